// fs_mesh.hip - the collider SDF of a closed triangle mesh on the device: manifold2SDF
// (FluidSim/src/cpu/util.cc:203-317), which cpu::FluidSimulator::setCollider calls with an exact
// band of 3 cells (FluidSim/include/FluidSim/cpu/fluid-simulator.h:62-71).
//
// The reference runs it serially: per triangle, the distance to every cell centre of its
// bounding box grown by the band (keeping the closest triangle), ray crossings along +z for the
// sign, then two rounds of eight fast sweeps that pass the closest triangle on to neighbours,
// and the sign from the crossing parity of each column. Here:
//   * band distances: one workgroup per triangle over its box, in two passes - the minimum
//     distance (atomic minimum on the bits of a non-negative double, which order like the
//     values), then the smallest index of a triangle at that distance (the serial loop keeps the
//     first triangle that reaches the minimum, since it replaces only on a strict decrease);
//   * crossings: integer atomic adds, order-free;
//   * sweeps: a sweep visits the cells in a fixed direction and each visit reads only the seven
//     neighbours the sweep has already visited, so it equals a wavefront over the planes
//     x' + y' + z' = L (x' the visiting order along x, ...): one launch per plane, every cell of a
//     plane independent. Same result as the serial sweep, bit for bit;
//   * sign: one thread per column.
// This is set-up work, outside the substep; its temporary arrays are freed before returning.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <vector>

#include "fs_ctx.h"

using namespace fsd;

namespace {

constexpr int TPB = 256;

struct D3 {
  double x, y, z;
};
__device__ inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 scl(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
// glm::dot for dvec3: (x*x' + y*y') + z*z'; glm::length = sqrt(dot)
__device__ inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double len3(D3 a) { return sqrt(dot3(a, a)); }

// distancePointTriangle (util.cc:86-125), the same case order and arithmetic. Where the
// closest point q is a + v ab (edge AB), a + w ac (edge AC) or a + v ab + w ac (inside), the
// reference measures |ap + v ab|, |ap + w ac|, |ap + v ab + w ac| (the length of p + q - 2a);
// sg = 1 reproduces that, sg = -1 gives |p - q| (FS_MESH_EXACT_DISTANCE).
__device__ double point_triangle(D3 p, D3 a, D3 b, D3 c, double sg) {
  const D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0 && d2 <= 0) return len3(ap);
  const D3 bp = sub(p, b);
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0 && d4 <= d3) return len3(bp);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) return len3(add(ap, scl(sg * (d1 / (d1 - d3)), ab)));
  const D3 cp = sub(p, c);
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0 && d5 <= d6) return len3(cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) return len3(add(ap, scl(sg * (d2 / (d2 - d6)), ac)));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return len3(add(bp, scl(w, sub(cp, bp))));
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  return len3(add(add(ap, scl(sg * v, ab)), scl(sg * w, ac)));
}

struct MeshArgs {
  const double* vtx;  // AoS xyz
  const uint32_t* idx;
  int ntri;
  int W, H, D;
  double h;
  double sg;  // +1: the reference's point-triangle distance; -1: the true one
};

__device__ inline D3 vertex(const MeshArgs& m, uint32_t i) { return {m.vtx[3 * i], m.vtx[3 * i + 1], m.vtx[3 * i + 2]}; }
__device__ inline D3 centre(const MeshArgs& m, int x, int y, int z) {
  return {coord1(x, 0.5, m.h), coord1(y, 0.5, m.h), coord1(z, 0.5, m.h)};
}
__device__ inline int box_lo(double a, double b, double c, double h, int band, int n) {
  return fclamp((int)floor(smin(smin(a, b), c) / h - 0.5) - band, 0, n - 1);
}
__device__ inline int box_hi(double a, double b, double c, double h, int band, int n) {
  return fclamp((int)ceil(smax(smax(a, b), c) / h - 0.5) + band, 0, n - 1);
}
struct Box {
  int x0, x1, y0, y1, z0, z1;
};
__device__ inline Box tri_box(const MeshArgs& m, D3 a, D3 b, D3 c, int band) {
  return {box_lo(a.x, b.x, c.x, m.h, band, m.W), box_hi(a.x, b.x, c.x, m.h, band, m.W),
          box_lo(a.y, b.y, c.y, m.h, band, m.H), box_hi(a.y, b.y, c.y, m.h, band, m.H),
          box_lo(a.z, b.z, c.z, m.h, band, m.D), box_hi(a.z, b.z, c.z, m.h, band, m.D)};
}

__global__ __launch_bounds__(TPB) void k_mesh_init(int64_t n, unsigned long long* __restrict__ dist,
                                                   int* __restrict__ closest, int* __restrict__ cnt) {
  const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (i >= n) return;
  dist[i] = (unsigned long long)__double_as_longlong(1e9);
  closest[i] = INT_MAX;
  cnt[i] = 0;
}

// Band distances, pass 1 (PICK = false): the minimum over triangles; pass 2 (PICK = true): the
// smallest index of a triangle at exactly that minimum. Workgroup t handles triangle t.
template <bool PICK>
__global__ __launch_bounds__(TPB) void k_mesh_band(MeshArgs m, int band, unsigned long long* __restrict__ dist,
                                                   int* __restrict__ closest) {
  for (int t = blockIdx.x; t < m.ntri; t += gridDim.x) {
    const D3 a = vertex(m, m.idx[3 * t]), b = vertex(m, m.idx[3 * t + 1]), c = vertex(m, m.idx[3 * t + 2]);
    const Box bx = tri_box(m, a, b, c, band);
    const int ex = bx.x1 - bx.x0 + 1, ey = bx.y1 - bx.y0 + 1, ez = bx.z1 - bx.z0 + 1;
    const int ncell = ex * ey * ez;
    for (int q = threadIdx.x; q < ncell; q += TPB) {
      const int z = bx.z0 + q % ez, y = bx.y0 + (q / ez) % ey, x = bx.x0 + q / (ez * ey);
      const double d = point_triangle(centre(m, x, y, z), a, b, c, m.sg);
      const int64_t id = lin3(x, y, z, m.H, m.D);
      const unsigned long long bits = (unsigned long long)__double_as_longlong(d);
      if (!PICK) {
        if (bits < dist[id]) atomicMin(dist + id, bits);
      } else if (bits == dist[id]) {
        atomicMin(closest + id, t);
      }
    }
  }
}

// Crossings of the +z rays through the cell centres of each triangle's xy box (util.cc:279-306).
__global__ __launch_bounds__(TPB) void k_mesh_crossings(MeshArgs m, int* __restrict__ cnt) {
  for (int t = blockIdx.x; t < m.ntri; t += gridDim.x) {
    const D3 a = vertex(m, m.idx[3 * t]), b = vertex(m, m.idx[3 * t + 1]), c = vertex(m, m.idx[3 * t + 2]);
    const Box bx = tri_box(m, a, b, c, 0);
    const int ey = bx.y1 - bx.y0 + 1;
    const int ncol = (bx.x1 - bx.x0 + 1) * ey;
    for (int q = threadIdx.x; q < ncol; q += TPB) {
      const int j = bx.x0 + q / ey, k = bx.y0 + q % ey;
      const double px = coord1(j, 0.5, m.h), py = coord1(k, 0.5, m.h);
      // insideTraingle2D (util.cc:131-143)
      const double c0 = (px - a.x) * (b.y - a.y) - (b.x - a.x) * (py - a.y);
      const double c1 = (px - b.x) * (c.y - b.y) - (c.x - b.x) * (py - b.y);
      const double c2 = (px - c.x) * (a.y - c.y) - (a.x - c.x) * (py - c.y);
      const int pos = (c0 >= 0.0) + (c1 >= 0.0) + (c2 >= 0.0);
      if (pos != 3 && pos != 0) continue;
      // calcBiocentricCoord / solveEquations (util.cc:145-157)
      const double a1 = a.x - c.x, b1 = b.x - c.x, e1 = px - c.x;
      const double a2 = a.y - c.y, b2 = b.y - c.y, e2 = py - c.y;
      const double det = a1 * b2 - a2 * b1;
      if (det == 0.0) continue;  // a triangle seen edge-on: the reference divides by zero here
      const double u = (e1 * b2 - e2 * b1) / det, v = (a1 * e2 - a2 * e1) / det;
      const double w = 1.0 - u - v;
      const double iz = a.z * u + b.z * v + c.z * w;
      const int l = (int)ceil(iz / m.h - 0.5);
      if (l < 0) atomicAdd(cnt + lin3(j, k, 0, m.H, m.D), 1);
      else if (l < m.D) atomicAdd(cnt + lin3(j, k, l, m.H, m.D), 1);
    }
  }
}

__global__ __launch_bounds__(TPB) void k_mesh_finish_band(int64_t n, const unsigned long long* __restrict__ bits,
                                                          double* __restrict__ sdf, int* __restrict__ closest) {
  const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (i >= n) return;
  sdf[i] = __longlong_as_double((long long)bits[i]);
  if (closest[i] == INT_MAX) closest[i] = -1;
}

// One plane x' + y' + z' = L of sweep (dx, dy, dz) (util.cc:180-201); x' = W-1-x for dx = 1 (the
// sweep starts at the high end) and x for dx = -1, likewise y', z'. checkNeighbour
// (util.cc:159-178) in the reference's order.
__global__ __launch_bounds__(TPB) void k_mesh_sweep_plane(MeshArgs m, int dx, int dy, int dz, int L,
                                                          double* __restrict__ sdf, int* __restrict__ closest) {
  const int64_t q = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (q >= (int64_t)m.W * m.H) return;
  const int xo = (int)(q / m.H), yo = (int)(q % m.H);
  const int zo = L - xo - yo;
  if (zo < 0 || zo >= m.D) return;
  const int x = dx > 0 ? m.W - 1 - xo : xo, y = dy > 0 ? m.H - 1 - yo : yo, z = dz > 0 ? m.D - 1 - zo : zo;
  const int64_t id = lin3(x, y, z, m.H, m.D);
  const D3 p = centre(m, x, y, z);
  double d = sdf[id];
  int cl = closest[id];
  const int off[7][3] = {{0, dy, dz}, {dx, 0, dz}, {dx, dy, 0}, {0, 0, dz}, {0, dy, 0}, {dx, 0, 0}, {dx, dy, dz}};
#pragma unroll
  for (int o = 0; o < 7; o++) {
    const int xn = x + off[o][0], yn = y + off[o][1], zn = z + off[o][2];
    if (xn < 0 || xn >= m.W || yn < 0 || yn >= m.H || zn < 0 || zn >= m.D) continue;
    const int t = closest[lin3(xn, yn, zn, m.H, m.D)];
    if (t == -1) continue;
    const double d1 = point_triangle(p, vertex(m, m.idx[3 * t]), vertex(m, m.idx[3 * t + 1]), vertex(m, m.idx[3 * t + 2]),
                                     m.sg);
    if (fabs(d1) < fabs(d)) {
      d = d1;
      cl = t;
    }
  }
  sdf[id] = d;
  closest[id] = cl;
}

// The sign: a cell is inside when the crossings at or below it in its column are odd.
template <typename T>
__global__ __launch_bounds__(TPB) void k_mesh_sign(int W, int H, int D, const double* __restrict__ sdf,
                                                   const int* __restrict__ cnt, T* __restrict__ out) {
  const int64_t q = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (q >= (int64_t)W * H) return;
  const int i = (int)(q / H), j = (int)(q % H);
  int tot = 0;
  for (int k = 0; k < D; k++) {
    const int64_t id = lin3(i, j, k, H, D);
    tot += cnt[id];
    out[id] = (T)((tot & 1) ? -sdf[id] : sdf[id]);
  }
}

int blocks_of(int64_t n) { return (int)((n + TPB - 1) / TPB); }

template <typename T>
int mesh_sdf_t(fs_ctx* c, const double* vtx, int64_t nv, const uint32_t* idx, int64_t ntri, int band, bool exact) {
  const int W = c->nx, H = c->ny, D = c->nz;
  const int64_t n = (int64_t)W * H * D;
  DBuf dv, di, dbits, dsdf, dclosest, dcnt;
  int st = FS_OK;
  auto alloc = [&](DBuf& b, size_t bytes) {
    if (st == FS_OK) st = fs_alloc(c, b, bytes ? bytes : 1);
  };
  alloc(dv, (size_t)nv * 3 * sizeof(double));
  alloc(di, (size_t)ntri * 3 * sizeof(uint32_t));
  alloc(dbits, (size_t)n * sizeof(unsigned long long));
  alloc(dsdf, (size_t)n * sizeof(double));
  alloc(dclosest, (size_t)n * sizeof(int));
  alloc(dcnt, (size_t)n * sizeof(int));
  auto run = [&]() -> int {
    FS_TRY(st);
    if (nv) FS_HIP(c, hipMemcpyAsync(dv.p, vtx, (size_t)nv * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (ntri) FS_HIP(c, hipMemcpyAsync(di.p, idx, (size_t)ntri * 3 * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    MeshArgs m{dv.as<double>(), di.as<uint32_t>(), (int)ntri, W, H, D, c->h[0], exact ? -1.0 : 1.0};
    k_mesh_init<<<blocks_of(n), TPB, 0, c->stream>>>(n, dbits.as<unsigned long long>(), dclosest.as<int>(),
                                                     dcnt.as<int>());
    const int tb = (int)(ntri < 65536 ? (ntri > 0 ? ntri : 1) : 65536);
    if (ntri > 0) {
      k_mesh_band<false><<<tb, TPB, 0, c->stream>>>(m, band, dbits.as<unsigned long long>(), dclosest.as<int>());
      k_mesh_band<true><<<tb, TPB, 0, c->stream>>>(m, band, dbits.as<unsigned long long>(), dclosest.as<int>());
      k_mesh_crossings<<<tb, TPB, 0, c->stream>>>(m, dcnt.as<int>());
    }
    k_mesh_finish_band<<<blocks_of(n), TPB, 0, c->stream>>>(n, dbits.as<unsigned long long>(), dsdf.as<double>(),
                                                           dclosest.as<int>());
    FS_HIP(c, hipGetLastError());
    const int dirs[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                            {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
    const int planes = W + H + D - 2;
    for (int round = 0; round < 2; round++)
      for (auto& d : dirs)
        for (int L = 0; L < planes; L++)
          k_mesh_sweep_plane<<<blocks_of((int64_t)W * H), TPB, 0, c->stream>>>(m, d[0], d[1], d[2], L,
                                                                               dsdf.as<double>(), dclosest.as<int>());
    FS_HIP(c, hipGetLastError());
    k_mesh_sign<T><<<blocks_of((int64_t)W * H), TPB, 0, c->stream>>>(W, H, D, dsdf.as<double>(), dcnt.as<int>(),
                                                                      c->collider.as<T>());
    FS_HIP(c, hipGetLastError());
    FS_HIP(c, hipStreamSynchronize(c->stream));
    return FS_OK;
  };
  st = run();
  for (DBuf* b : {&dv, &di, &dbits, &dsdf, &dclosest, &dcnt}) fs_free(*b);
  return st;
}

}  // namespace

extern "C" int fs_set_collider_mesh(fs_ctx* c, const double* vertices, int64_t nverts, const uint32_t* indices,
                                    int64_t ntris, int32_t exact_band, int32_t flags) {
  if (!c) return FS_BAD_ARG;
  if (nverts < 0 || ntris < 0 || ntris > INT_MAX / 3 || exact_band < 0 || (nverts > 0 && !vertices) ||
      (ntris > 0 && !indices) || (flags & ~FS_MESH_EXACT_DISTANCE)) {
    c->set_error("bad mesh arguments");
    return FS_BAD_ARG;
  }
  for (int64_t q = 0; q < 3 * ntris; q++)
    if ((int64_t)indices[q] >= nverts) {
      c->set_error("triangle index " + std::to_string(q / 3) + " names a missing vertex");
      return FS_BAD_ARG;
    }
  for (int64_t q = 0; q < 3 * nverts; q++)
    if (!std::isfinite(vertices[q])) {
      c->set_error("non-finite mesh vertex");
      return FS_BAD_ARG;
    }
  const bool exact = (flags & FS_MESH_EXACT_DISTANCE) != 0;
  return c->mixed ? mesh_sdf_t<float>(c, vertices, nverts, indices, ntris, exact_band, exact)
                  : mesh_sdf_t<double>(c, vertices, nverts, indices, ntris, exact_band, exact);
}
