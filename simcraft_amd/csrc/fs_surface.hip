// fs_surface.hip - surface output: rebuildSurface (FluidSim/src/cpu/rebuild-surface.cc:688-741),
// marching cubes over the fluid SDF on the device.
//
// Samples are the cell centres with indices -1..n per axis (a sample outside the grid reads 1000,
// :655-661); cubes have corner indices -1..n. The surface has one vertex per sign-changing
// lattice edge, at lerp(pa, pb, va/(va - vb)) with normal normalize(lerp(grad a, grad b, t))
// (:700-729, grad :666-687), and the triangles of each cube come from a case table. Differences
// from the reference, where it takes a table from elsewhere (fluid-engine-dev's
// triangleConnectionTable3D, :114-641): the table is built here by a rule (mc_table, below), and
// vertices are numbered in the order of their edges, triangles in the order of their cubes.
//
// Device pipeline, all deterministic: flag the sign-changing edges of each axis, prefix-sum the
// flags into vertex numbers (written per edge), write the vertices; count each cube's triangles,
// prefix-sum into offsets, write the triangles (prefix sums: fs_scan.h). The surface stays on the
// device until fs_get_surface copies it out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "fs_ctx.h"
#include "fs_scan.h"

using namespace fsd;

namespace {

constexpr int TPB = 256;

// ---- the case table ----------------------------------------------------------------------
// Cube corners c = x + 2y + 4z. Edge q = 4a + m of axis a starts at the corner whose other two
// coordinates are the bits of m (lower axis first) and 0 on a. For each of the 256 inside
// patterns (bit c: sample c < 0): on every face, the crossings of its sign-changing edges are
// joined by segments that separate the inside corners (two diagonal inside corners are each cut
// off, so the two cubes of a face agree); a segment runs with the inside corners on its right seen
// from outside the cube; the segments close into loops (each crossing edge lies on two faces);
// each loop is a fan from its first vertex none of whose chords lies on a cube face. Triangles
// face the positive (outside) side.
struct McTable {
  int8_t ntri[256];
  int8_t tri[256][36];
};

int edge_corner(int q, int end) {
  const int a = q >> 2, m = q & 3, o1 = (a + 1) % 3, o2 = (a + 2) % 3;
  const int lo = std::min(o1, o2), hi = std::max(o1, o2);
  return (end << a) | ((m & 1) << lo) | (((m >> 1) & 1) << hi);
}
int edge_between(int c0, int c1) {
  const int d = c0 ^ c1, ax = d == 1 ? 0 : (d == 2 ? 1 : 2), base = c0 & c1;
  const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
  const int lo = std::min(o1, o2), hi = std::max(o1, o2);
  return 4 * ax + ((base >> lo) & 1) + 2 * ((base >> hi) & 1);
}
bool edges_share_face(int q1, int q2) {
  auto faces = [](int q, int f[2][2]) {
    const int a = q >> 2, m = q & 3, o1 = (a + 1) % 3, o2 = (a + 2) % 3;
    f[0][0] = std::min(o1, o2);
    f[0][1] = m & 1;
    f[1][0] = std::max(o1, o2);
    f[1][1] = (m >> 1) & 1;
  };
  int f1[2][2], f2[2][2];
  faces(q1, f1);
  faces(q2, f2);
  for (auto& x : f1)
    for (auto& y : f2)
      if (x[0] == y[0] && x[1] == y[1]) return true;
  return false;
}

McTable build_table() {
  McTable T{};
  for (int cfg = 0; cfg < 256; cfg++) {
    auto in = [&](int c) { return (cfg >> c) & 1; };
    auto corner_pos = [](int c, double v[3]) {
      v[0] = c & 1;
      v[1] = (c >> 1) & 1;
      v[2] = (c >> 2) & 1;
    };
    int next[12];
    std::fill(next, next + 12, -1);
    for (int a = 0; a < 3; a++)
      for (int side = 0; side < 2; side++) {
        const int u = (a + 1) % 3, w = (a + 2) % 3;
        const int quad[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        int cyc[4];
        for (int k = 0; k < 4; k++) cyc[k] = (side << a) | (quad[k][0] << u) | (quad[k][1] << w);
        int nin = 0;
        for (int k = 0; k < 4; k++) nin += in(cyc[k]);
        if (nin == 0 || nin == 4) continue;
        int seg[2][3], ns = 0;  // (edge, edge, an inside corner that must lie on the right)
        if (nin == 2 && in(cyc[0]) == in(cyc[2])) {
          for (int k = 0; k < 4; k++)
            if (in(cyc[k])) {
              seg[ns][0] = edge_between(cyc[(k + 3) % 4], cyc[k]);
              seg[ns][1] = edge_between(cyc[k], cyc[(k + 1) % 4]);
              seg[ns][2] = cyc[k];
              ns++;
            }
        } else {
          int cross[2], nc = 0, keep = -1;
          for (int k = 0; k < 4; k++) {
            if (in(cyc[k]) != in(cyc[(k + 1) % 4])) cross[nc++] = edge_between(cyc[k], cyc[(k + 1) % 4]);
            if (in(cyc[k]) && keep < 0) keep = cyc[k];
          }
          if (nin == 3)
            for (int k = 0; k < 4; k++)
              if (!in(cyc[k])) keep = cyc[(k + 2) % 4];
          seg[0][0] = cross[0];
          seg[0][1] = cross[1];
          seg[0][2] = keep;
          ns = 1;
        }
        const double nrm[3] = {a == 0 ? (side ? 1.0 : -1.0) : 0.0, a == 1 ? (side ? 1.0 : -1.0) : 0.0,
                               a == 2 ? (side ? 1.0 : -1.0) : 0.0};
        for (int s = 0; s < ns; s++) {
          double A0[3], A1[3], B0[3], B1[3], C[3], P[3], Q[3];
          corner_pos(edge_corner(seg[s][0], 0), A0);
          corner_pos(edge_corner(seg[s][0], 1), A1);
          corner_pos(edge_corner(seg[s][1], 0), B0);
          corner_pos(edge_corner(seg[s][1], 1), B1);
          corner_pos(seg[s][2], C);
          for (int d = 0; d < 3; d++) {
            P[d] = 0.5 * (A0[d] + A1[d]);
            Q[d] = 0.5 * (B0[d] + B1[d]);
          }
          const double e[3] = {Q[0] - P[0], Q[1] - P[1], Q[2] - P[2]};
          const double f[3] = {C[0] - P[0], C[1] - P[1], C[2] - P[2]};
          const double side_of = (e[1] * f[2] - e[2] * f[1]) * nrm[0] + (e[2] * f[0] - e[0] * f[2]) * nrm[1] +
                                 (e[0] * f[1] - e[1] * f[0]) * nrm[2];
          if (side_of > 0) next[seg[s][1]] = seg[s][0];
          else next[seg[s][0]] = seg[s][1];
        }
      }
    int nt = 0;
    bool used[12] = {};
    for (int start = 0; start < 12; start++) {
      if (next[start] < 0 || used[start]) continue;
      int loop[12], m = 0;
      for (int q = start; q >= 0 && !used[q]; q = next[q]) {
        used[q] = true;
        loop[m++] = q;
      }
      int apex = 0;
      for (int s = 0; s < m; s++) {
        bool ok = true;
        for (int k = 0; k < m && ok; k++)
          if (k != s && k != (s + 1) % m && k != (s + m - 1) % m && edges_share_face(loop[s], loop[k])) ok = false;
        if (ok) {
          apex = s;
          break;
        }
      }
      for (int k = 1; k + 1 < m; k++) {
        T.tri[cfg][3 * nt] = (int8_t)loop[apex];
        T.tri[cfg][3 * nt + 1] = (int8_t)loop[(apex + k) % m];
        T.tri[cfg][3 * nt + 2] = (int8_t)loop[(apex + k + 1) % m];
        nt++;
      }
    }
    T.ntri[cfg] = (int8_t)nt;
  }
  return T;
}

__constant__ int8_t c_ntri[256];
__constant__ int8_t c_tri[256][36];

// ---- lattice -------------------------------------------------------------------------------
struct Lat {
  int nx, ny, nz;
  double h[3];
  // edges of axis a from sample (x, y, z), coordinates -1..n: id = ((a (nx+2) + x+1) (ny+2) + y+1)
  // (nz+2) + z+1; an edge whose far end lies past sample n does not exist
  __device__ int64_t nsamp() const { return (int64_t)(nx + 2) * (ny + 2) * (nz + 2); }
};

template <typename T>
__device__ inline double sample(const T* __restrict__ g, const Lat& L, int x, int y, int z) {
  if (x < 0 || x >= L.nx || y < 0 || y >= L.ny || z < 0 || z >= L.nz) return 1000;
  return (double)g[lin3(x, y, z, L.ny, L.nz)];
}

template <typename T>
__device__ inline void grad_at(const T* __restrict__ g, const Lat& L, int x, int y, int z, double out[3]) {
  const int xn = x < L.nx ? x + 1 : x, xp = x > -1 ? x - 1 : x;
  const int yn = y < L.ny ? y + 1 : y, yp = y > -1 ? y - 1 : y;
  const int zn = z < L.nz ? z + 1 : z, zp = z > -1 ? z - 1 : z;
  const double dx = x > -1 && x < L.nx - 1 ? 2 * L.h[0] : L.h[0];
  const double dy = y > -1 && y < L.ny - 1 ? 2 * L.h[1] : L.h[1];
  const double dz = z > -1 && z < L.nz - 1 ? 2 * L.h[2] : L.h[2];
  out[0] = (sample(g, L, xn, y, z) - sample(g, L, xp, y, z)) / dx;
  out[1] = (sample(g, L, x, yn, z) - sample(g, L, x, yp, z)) / dy;
  out[2] = (sample(g, L, x, y, zn) - sample(g, L, x, y, zp)) / dz;
}

__device__ inline void unlin_sample(const Lat& L, int64_t s, int& x, int& y, int& z) {
  z = (int)(s % (L.nz + 2)) - 1;
  const int64_t t = s / (L.nz + 2);
  y = (int)(t % (L.ny + 2)) - 1;
  x = (int)(t / (L.ny + 2)) - 1;
}

// 1 where edge e (of 3 * nsamp) changes sign
template <typename T>
__global__ __launch_bounds__(TPB) void k_mc_edge_flags(Lat L, const T* __restrict__ g, int* __restrict__ flag) {
  const int64_t e = blockIdx.x * (int64_t)TPB + threadIdx.x;
  const int64_t ns = L.nsamp();
  if (e >= 3 * ns) return;
  const int a = (int)(e / ns);
  int p[3];
  unlin_sample(L, e - a * ns, p[0], p[1], p[2]);
  const int lim[3] = {L.nx, L.ny, L.nz};
  int f = 0;
  if (p[a] < lim[a]) {
    int q[3] = {p[0], p[1], p[2]};
    q[a]++;
    f = (sample(g, L, p[0], p[1], p[2]) < 0.0) != (sample(g, L, q[0], q[1], q[2]) < 0.0);
  }
  flag[e] = f;
}

// vertex of every flagged edge, at its exclusive prefix index
template <typename T>
__global__ __launch_bounds__(TPB) void k_mc_vertices(Lat L, const T* __restrict__ g, const int* __restrict__ flag,
                                                     const int* __restrict__ vid, double* __restrict__ V,
                                                     double* __restrict__ N) {
  const int64_t e = blockIdx.x * (int64_t)TPB + threadIdx.x;
  const int64_t ns = L.nsamp();
  if (e >= 3 * ns || !flag[e]) return;
  const int a = (int)(e / ns);
  int p[3];
  unlin_sample(L, e - a * ns, p[0], p[1], p[2]);
  int q[3] = {p[0], p[1], p[2]};
  q[a]++;
  const double va = sample(g, L, p[0], p[1], p[2]), vb = sample(g, L, q[0], q[1], q[2]);
  const double t = va / (va - vb);
  double ga[3], gb[3], n[3];
  grad_at(g, L, p[0], p[1], p[2], ga);
  grad_at(g, L, q[0], q[1], q[2], gb);
  const int64_t v = vid[e];
  for (int d = 0; d < 3; d++) {
    const double pa = coord1(p[d], 0.5, L.h[d]), pb = coord1(q[d], 0.5, L.h[d]);
    V[3 * v + d] = pa * (1 - t) + pb * t;  // fluid::cpu lerp (cpu/util.h:27-29)
    n[d] = ga[d] * (1 - t) + gb[d] * t;
  }
  const double s = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);  // glm::normalize
  for (int d = 0; d < 3; d++) N[3 * v + d] = n[d] * s;
}

__device__ inline int cube_config(const double* s8) {
  int cfg = 0;
  for (int c = 0; c < 8; c++)
    if (s8[c] < 0.0) cfg |= 1 << c;
  return cfg;
}

// cubes: origins -1..n-1 per axis, (nx+1)(ny+1)(nz+1) of them in (x, y, z) order
template <typename T>
__global__ __launch_bounds__(TPB) void k_mc_cube_count(Lat L, const T* __restrict__ g, int* __restrict__ cnt) {
  const int64_t c = blockIdx.x * (int64_t)TPB + threadIdx.x;
  const int64_t nc = (int64_t)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
  if (c >= nc) return;
  const int z = (int)(c % (L.nz + 1)) - 1, y = (int)((c / (L.nz + 1)) % (L.ny + 1)) - 1,
            x = (int)(c / ((int64_t)(L.nz + 1) * (L.ny + 1))) - 1;
  double s8[8];
  for (int k = 0; k < 8; k++) s8[k] = sample(g, L, x + (k & 1), y + ((k >> 1) & 1), z + ((k >> 2) & 1));
  cnt[c] = c_ntri[cube_config(s8)];
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_mc_triangles(Lat L, const T* __restrict__ g, const int* __restrict__ toff,
                                                      const int* __restrict__ vid, uint32_t* __restrict__ I) {
  const int64_t c = blockIdx.x * (int64_t)TPB + threadIdx.x;
  const int64_t nc = (int64_t)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
  if (c >= nc) return;
  const int z = (int)(c % (L.nz + 1)) - 1, y = (int)((c / (L.nz + 1)) % (L.ny + 1)) - 1,
            x = (int)(c / ((int64_t)(L.nz + 1) * (L.ny + 1))) - 1;
  double s8[8];
  for (int k = 0; k < 8; k++) s8[k] = sample(g, L, x + (k & 1), y + ((k >> 1) & 1), z + ((k >> 2) & 1));
  const int cfg = cube_config(s8);
  const int nt = c_ntri[cfg];
  const int64_t ns = L.nsamp();
  const int64_t base = (int64_t)toff[c];
  for (int k = 0; k < 3 * nt; k++) {
    const int q = c_tri[cfg][k], a = q >> 2, m = q & 3, o1 = (a + 1) % 3, o2 = (a + 2) % 3;
    const int lo = o1 < o2 ? o1 : o2, hi = o1 < o2 ? o2 : o1;
    int p[3] = {x, y, z};
    p[lo] += m & 1;
    p[hi] += (m >> 1) & 1;
    const int64_t e = a * ns + (((int64_t)(p[0] + 1) * (L.ny + 2) + (p[1] + 1)) * (L.nz + 2) + (p[2] + 1));
    I[3 * base + k] = (uint32_t)vid[e];
  }
}

int blocks_of(int64_t n) { return (int)((n + TPB - 1) / TPB); }

// out = exclusive prefix sum of in (n ints); *total_host = the sum
int exclusive_scan(fs_ctx* c, int64_t n, const int* in, int* out, DBuf& tiles, int64_t* total_host) {
  const int64_t nt = fsscan::tiles_for(n);
  FS_TRY(fs_alloc(c, tiles, (size_t)(nt + 1) * sizeof(int)));
  FS_HIP(c, fsscan::exclusive_scan(c->stream, n, in, out, tiles.as<int>()));
  int tot = 0;
  FS_HIP(c, hipMemcpyAsync(&tot, tiles.as<int>() + nt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  *total_host = tot;
  fs_free(tiles);
  return FS_OK;
}

std::once_flag g_table_once;
McTable g_table;

template <typename T>
int rebuild_t(fs_ctx* c, int64_t* nverts, int64_t* ntris) {
  std::call_once(g_table_once, [] { g_table = build_table(); });
  FS_HIP(c, hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ntri), g_table.ntri, sizeof(g_table.ntri), 0, hipMemcpyHostToDevice,
                                   c->stream));
  FS_HIP(c, hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tri), g_table.tri, sizeof(g_table.tri), 0, hipMemcpyHostToDevice,
                                   c->stream));
  Lat L{c->nx, c->ny, c->nz, {c->h[0], c->h[1], c->h[2]}};
  const int64_t ns = (int64_t)(c->nx + 2) * (c->ny + 2) * (c->nz + 2);
  const int64_t ne = 3 * ns, nc = (int64_t)(c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  if (ne > INT32_MAX / 2) {
    c->set_error("grid too large for the surface's 32-bit vertex numbers");
    return FS_UNSUPPORTED;
  }
  const T* g = c->fsdf.as<T>();
  DBuf flag, vid, cnt, toff, tiles;
  int st = FS_OK;
  auto body = [&]() -> int {
    FS_TRY(fs_alloc(c, flag, (size_t)ne * sizeof(int)));
    FS_TRY(fs_alloc(c, vid, (size_t)ne * sizeof(int)));
    FS_TRY(fs_alloc(c, cnt, (size_t)nc * sizeof(int)));
    FS_TRY(fs_alloc(c, toff, (size_t)nc * sizeof(int)));
    k_mc_edge_flags<T><<<blocks_of(ne), TPB, 0, c->stream>>>(L, g, flag.as<int>());
    k_mc_cube_count<T><<<blocks_of(nc), TPB, 0, c->stream>>>(L, g, cnt.as<int>());
    FS_HIP(c, hipGetLastError());
    int64_t nv = 0, ntri = 0;
    FS_TRY(exclusive_scan(c, ne, flag.as<int>(), vid.as<int>(), tiles, &nv));
    FS_TRY(exclusive_scan(c, nc, cnt.as<int>(), toff.as<int>(), tiles, &ntri));
    fs_free(c->surf_v);
    fs_free(c->surf_n);
    fs_free(c->surf_i);
    FS_TRY(fs_alloc(c, c->surf_v, (size_t)std::max<int64_t>(nv, 1) * 3 * sizeof(double)));
    FS_TRY(fs_alloc(c, c->surf_n, (size_t)std::max<int64_t>(nv, 1) * 3 * sizeof(double)));
    FS_TRY(fs_alloc(c, c->surf_i, (size_t)std::max<int64_t>(ntri, 1) * 3 * sizeof(uint32_t)));
    k_mc_vertices<T><<<blocks_of(ne), TPB, 0, c->stream>>>(L, g, flag.as<int>(), vid.as<int>(), c->surf_v.as<double>(),
                                                           c->surf_n.as<double>());
    k_mc_triangles<T><<<blocks_of(nc), TPB, 0, c->stream>>>(L, g, toff.as<int>(), vid.as<int>(),
                                                            c->surf_i.as<uint32_t>());
    FS_HIP(c, hipGetLastError());
    FS_HIP(c, hipStreamSynchronize(c->stream));
    c->surf_nv = nv;
    c->surf_nt = ntri;
    *nverts = nv;
    *ntris = ntri;
    return FS_OK;
  };
  st = body();
  for (DBuf* b : {&flag, &vid, &cnt, &toff, &tiles}) fs_free(*b);
  return st;
}

}  // namespace

extern "C" {

int fs_rebuild_surface(fs_ctx* c, int64_t* nverts, int64_t* ntris) {
  if (!c || !nverts || !ntris) return FS_BAD_ARG;
  return c->mixed ? rebuild_t<float>(c, nverts, ntris) : rebuild_t<double>(c, nverts, ntris);
}

int fs_get_surface(fs_ctx* c, double* vertices, double* normals, uint32_t* indices) {
  if (!c) return FS_BAD_ARG;
  if (vertices && c->surf_nv)
    FS_HIP(c, hipMemcpyAsync(vertices, c->surf_v.p, (size_t)c->surf_nv * 3 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  if (normals && c->surf_nv)
    FS_HIP(c, hipMemcpyAsync(normals, c->surf_n.p, (size_t)c->surf_nv * 3 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  if (indices && c->surf_nt)
    FS_HIP(c, hipMemcpyAsync(indices, c->surf_i.p, (size_t)c->surf_nt * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  return FS_OK;
}

}  // extern "C"
