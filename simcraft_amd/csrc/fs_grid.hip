// fs_grid.hip - HIP kernels for every non-PCG stage of the FLIP substep, gfx950.
//
// Layout: dense arrays, k fastest (Spatify LinearHashXYZ, hash.h:25); particles SoA.
// Each kernel restates one SimCraft function (cited) with the same per-element operation
// order, so the FP64 build is bit-identical to a correctly rounded CPU evaluation.
// P2G is a deterministic gather over particles binned by their nearest node (counting sort,
// particle index ascending inside a bin) instead of the reference's racy TBB scatter.

#include <cfloat>
#include <cmath>
#include <cstring>

#include "fs_ctx.h"
#include "fs_scan.h"

using namespace fsd;

namespace {

constexpr int TPB = 256;

template <typename T>
struct GridArgs {
  int nx, ny, nz;
  T h[3];
  T ext[3];
};

template <typename T>
GridArgs<T> grid_args(const fs_ctx* c) {
  GridArgs<T> a;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  for (int d = 0; d < 3; d++) a.h[d] = (T)c->h[d];
  // HybridAdvectionSolver3D width/height/depth = resolution * h (cpu/fluid-simulator.h:88-95)
  a.ext[0] = (T)(c->nx * c->h[0]);
  a.ext[1] = (T)(c->ny * c->h[1]);
  a.ext[2] = (T)(c->nz * c->h[2]);
  return a;
}

// One thread per node of an nx*ny*nz grid without index division: blocks of 64 x 4 threads,
// x -> k (64 per block), y -> j (4 per block), block z -> i. Returns false past the grid's edge.
constexpr int G3X = 64, G3Y = 4;
static_assert(G3X * G3Y == TPB, "3D stencil blocks hold TPB threads");
inline dim3 g3_grid(int nx, int ny, int nz) { return dim3((nz + G3X - 1) / G3X, (ny + G3Y - 1) / G3Y, nx); }
inline dim3 g3_block() { return dim3(G3X, G3Y, 1); }
__device__ inline bool g3_index(int ny, int nz, int& i, int& j, int& k, int& id) {
  k = (int)blockIdx.x * G3X + (int)threadIdx.x;
  j = (int)blockIdx.y * G3Y + (int)threadIdx.y;
  i = (int)blockIdx.z;
  id = (i * ny + j) * nz + k;
  return k < nz && j < ny;
}

__device__ inline void unlin3(int id, int ny, int nz, int& i, int& j, int& k) {
  k = id % nz;
  int t = id / nz;
  j = t % ny;
  i = t / ny;
}

// 4x4x4 stencil of one face grid around TGrid::nearest: first node index and 1D kernel
// factors. The node weight CubicKernel::weight(h, p - pn) is (wx*wy)*wz.
template <typename T>
struct Stencil {
  int b[3];
  T w[3][4];
};

template <typename T>
__device__ inline void make_stencil(const T p[3], int axis, const T h[3], const T hs[3], Stencil<T>& st) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T off = (c == axis) ? T(0) : T(0.5);
    int nb = nearest1(p[c], off, h[c]);
    st.b[c] = nb - 1;
#pragma unroll
    for (int o = 0; o < 4; o++) {
      T pn = coord1(nb - 1 + o, off, h[c]);
      st.w[c][o] = bspline(wdiv(fabs(p[c] - pn), hs[c]));
    }
  }
}

// One grid of sampleVelocity (FluidSim/src/cpu/util.cc:40-85): i outer, k inner,
// out-of-range nodes skipped; two grids sampled with the same weights (FLIP: current, last).
template <typename T, bool TWO>
__device__ inline void sample_grid(const Stencil<T>& st, const T* __restrict__ G, const T* __restrict__ G2, int gx,
                                   int gy, int gz, T& acc, T& acc2, T& ws) {
  acc = T(0);
  acc2 = T(0);
  ws = T(0);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int x = st.b[0] + i;
    if (x < 0 || x >= gx) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int y = st.b[1] + j;
      if (y < 0 || y >= gy) continue;
      T wxy = st.w[0][i] * st.w[1][j];
      int base = gz * (y + gy * x);
      if (st.b[2] >= 0 && st.b[2] + 3 < gz) {  // the four z-nodes in range: one row, read at once
        T g4[4], h4[4];
        __builtin_memcpy(g4, G + base + st.b[2], sizeof(g4));
        if (TWO) __builtin_memcpy(h4, G2 + base + st.b[2], sizeof(h4));
#pragma unroll
        for (int k = 0; k < 4; k++) {
          T w = wxy * st.w[2][k];
          acc = acc + g4[k] * w;
          if (TWO) acc2 = acc2 + h4[k] * w;
          ws = ws + w;
        }
        continue;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int z = st.b[2] + k;
        if (z < 0 || z >= gz) continue;
        T w = wxy * st.w[2][k];
        acc = acc + G[base + z] * w;
        if (TWO) acc2 = acc2 + G2[base + z] * w;
        ws = ws + w;
      }
    }
  }
}

template <typename T>
struct Grids3 {
  const T* g[3];
  int d[3][3];
};

// sampleVelocity over the three face grids.
template <typename T>
__device__ inline V3<T> sample_velocity(const GridArgs<T>& ga, const Grids3<T>& G, T px, T py, T pz) {
  T p[3] = {px, py, pz};
  T hs[3] = {ga.h[0], ga.h[0], ga.h[0]};  // scalar h.x (util.cc:47)
  T out[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    Stencil<T> st;
    make_stencil(p, a, ga.h, hs, st);
    T acc, acc2, ws;
    sample_grid<T, false>(st, G.g[a], nullptr, G.d[a][0], G.d[a][1], G.d[a][2], acc, acc2, ws);
    out[a] = ws > T(0) ? acc / ws : T(0);
  }
  return {out[0], out[1], out[2]};
}

// HybridAdvectionSolver3D::handleCollision (FluidSim/src/cpu/advect-solver.cc:100-137)
template <typename T>
__device__ inline void handle_collision(const GridArgs<T>& ga, const T* __restrict__ col, V3<T>& p, V3<T>& v) {
  T d = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p.x, p.y, p.z);
  if (d < T(0)) {
    V3<T> n = gnormalize(sdf_grad(col, ga.nx, ga.ny, ga.nz, ga.h, p.x, p.y, p.z));
    p.x = p.x - d * n.x;
    p.y = p.y - d * n.y;
    p.z = p.z - d * n.z;
    T dv = v.x * n.x + v.y * n.y + v.z * n.z;
    T m = smin(dv, T(0));
    v.x = v.x - m * n.x;
    v.y = v.y - m * n.y;
    v.z = v.z - m * n.z;
  }
  if (p.x < T(0)) {
    p.x = T(0);
    if (v.x < T(0)) v.x = T(0);
  }
  if (p.x > ga.ext[0]) {
    p.x = ga.ext[0];
    if (v.x > T(0)) v.x = T(0);
  }
  if (p.y < T(0)) {
    p.y = T(0);
    if (v.y < T(0)) v.y = T(0);
  }
  if (p.y > ga.ext[1]) {
    p.y = ga.ext[1];
    if (v.y > T(0)) v.y = T(0);
  }
  if (p.z < T(0)) {
    p.z = T(0);
    if (v.z < T(0)) v.z = T(0);
  }
  if (p.z > ga.ext[2]) {
    p.z = ga.ext[2];
    if (v.z > T(0)) v.z = T(0);
  }
}

// FlipAdvectionSolver3D::advect (advect-solver.cc:170-187): Ralston RK3 + collision.
// *flag = 1 if any face velocity of the three grids is nonzero or NaN (+0 and -0 count as zero).
// 16-byte loads over each grid (its last n % (16 / sizeof(T)) values one by one).
template <typename T>
__global__ __launch_bounds__(TPB) void k_any_nonzero(int64_t n0, int64_t n1, int64_t n2, const T* __restrict__ g0,
                                                     const T* __restrict__ g1, const T* __restrict__ g2,
                                                     int* __restrict__ flag) {
  constexpr int V = 16 / (int)sizeof(T);
  typedef T TV __attribute__((ext_vector_type(V)));
  bool nz = false;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  const T* gs[3] = {g0, g1, g2};
  const int64_t ns[3] = {n0, n1, n2};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int64_t nv = ns[a] / V;
    const TV* gv = reinterpret_cast<const TV*>(gs[a]);
    for (int64_t e = tid; e < nv; e += stride) {
      const TV v = gv[e];
#pragma unroll
      for (int q = 0; q < V; q++) nz |= !(v[q] == T(0));
    }
    for (int64_t e = nv * V + tid; e < ns[a]; e += stride) nz |= !(gs[a][e] == T(0));
  }
  if (__builtin_amdgcn_ballot_w64(nz) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_advect(int n, GridArgs<T> ga, Grids3<T> G, const T* __restrict__ col, T* __restrict__ px,
                         T* __restrict__ py, T* __restrict__ pz, T* __restrict__ vx, T* __restrict__ vy,
                         T* __restrict__ vz, T dt, const int* __restrict__ nonzero) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // a grid that holds only zeros (the reference's order clears it just before advecting) samples
  // to exactly +0 at every point: the stage velocities are then taken as +0 without sampling
  const bool sample = *nonzero != 0;
  V3<T> p = {px[i], py[i], pz[i]};
  V3<T> k1 = {vx[i], vy[i], vz[i]};
  T c1 = T(0.5) * dt;
  V3<T> p1 = {p.x + c1 * k1.x, p.y + c1 * k1.y, p.z + c1 * k1.z};
  V3<T> k2 = sample ? sample_velocity(ga, G, p1.x, p1.y, p1.z) : V3<T>{T(0), T(0), T(0)};
  T c2 = T(0.75) * dt;
  V3<T> p2 = {p.x + c2 * k2.x, p.y + c2 * k2.y, p.z + c2 * k2.z};
  V3<T> k3 = sample ? sample_velocity(ga, G, p2.x, p2.y, p2.z) : V3<T>{T(0), T(0), T(0)};
  T a1 = T(2.0 / 9.0) * dt, a2 = T(1.0 / 3.0) * dt, a3 = T(4.0 / 9.0) * dt;
  p.x = p.x + a1 * k1.x + a2 * k2.x + a3 * k3.x;
  p.y = p.y + a1 * k1.y + a2 * k2.y + a3 * k3.y;
  p.z = p.z + a1 * k1.z + a2 * k2.z + a3 * k3.z;
  V3<T> v = k1;
  handle_collision(ga, col, p, v);
  px[i] = p.x;
  py[i] = p.y;
  pz[i] = p.z;
  vx[i] = v.x;
  vy[i] = v.y;
  vz[i] = v.z;
}

// FlipAdvectionSolver3D::solveG2P / PicAdvector3D::solveG2P (advect-solver.cc:7-17, 157-169).
template <typename T, bool FLIP>
__global__ __launch_bounds__(TPB) void k_g2p(int n, GridArgs<T> ga, Grids3<T> G, Grids3<T> L, const T* __restrict__ px,
                      const T* __restrict__ py, const T* __restrict__ pz, T* __restrict__ vx, T* __restrict__ vy,
                      T* __restrict__ vz, T alpha) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T p[3] = {px[i], py[i], pz[i]};
  T hs[3] = {ga.h[0], ga.h[0], ga.h[0]};
  T* vout[3] = {vx, vy, vz};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    Stencil<T> st;
    make_stencil(p, a, ga.h, hs, st);
    T acc, acc2, ws;
    sample_grid<T, FLIP>(st, G.g[a], L.g[a], G.d[a][0], G.d[a][1], G.d[a][2], acc, acc2, ws);
    T pic = ws > T(0) ? acc / ws : T(0);
    if (FLIP) {
      T old = ws > T(0) ? acc2 / ws : T(0);
      T fl = vout[a][i] + pic - old;
      vout[a][i] = alpha * fl + (T(1) - alpha) * pic;
    } else {
      vout[a][i] = pic;
    }
  }
}

// ---- counting sort of particles into bins (stable: index ascending inside a bin) ----
template <typename T>
__global__ __launch_bounds__(TPB) void k_keys_nearest(int n, int axis, GridArgs<T> ga, const T* __restrict__ px, const T* __restrict__ py,
                               const T* __restrict__ pz, int* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // bin grid covers nearest in [0, n_c] on every axis for positions inside the domain
  int b0 = nearest1(px[i], axis == 0 ? T(0) : T(0.5), ga.h[0]);
  int b1 = nearest1(py[i], axis == 1 ? T(0) : T(0.5), ga.h[1]);
  int b2 = nearest1(pz[i], axis == 2 ? T(0) : T(0.5), ga.h[2]);
  b0 = fclamp(b0, 0, ga.nx);
  b1 = fclamp(b1, 0, ga.ny);
  b2 = fclamp(b2, 0, ga.nz);
  keys[i] = lin3(b0, b1, b2, ga.ny + 1, ga.nz + 1);
}

// ParticleNeighbourSearcher::update key (Spatify/include/Spatify/ns-util.h:50-62)
template <typename T>
__global__ __launch_bounds__(TPB) void k_keys_morton(int n, GridArgs<T> ga, const T* __restrict__ px, const T* __restrict__ py,
                              const T* __restrict__ pz, int* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int ix = (int)floor(px[i] / ga.h[0]);
  if (ix < 0) ix = 0;
  if (ix >= ga.nx) ix = ga.nx - 1;
  int iy = (int)floor(py[i] / ga.h[1]);
  if (iy < 0) iy = 0;
  if (iy >= ga.ny) iy = ga.ny - 1;
  int iz = (int)floor(pz[i] / ga.h[2]);
  if (iz < 0) iz = 0;
  if (iz >= ga.nz) iz = ga.nz - 1;
  keys[i] = morton3D(ix, iy, iz);
}

// Particles are stored in a spatially coherent order, so the 64 particles of a wave fall into
// few bins. Lanes with equal keys are served by one atomic from the group's lowest lane, which
// takes the whole group's count; each lane's slot is the returned base plus its rank in the group.
// (Slots only seed the bin's positions: k_sort_bins restores index order inside every bin.)
__global__ __launch_bounds__(TPB) void k_count(int n, const int* __restrict__ keys, int* __restrict__ counts, int* __restrict__ slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63;
  const bool ok = i < n;
  const int key = ok ? keys[i] : -1;
  // lanes with equal keys form a group; the group's first lane adds the group size, all leaders
  // in one atomic instruction, and every lane takes its leader's old count plus its rank
  uint64_t todo = __builtin_amdgcn_ballot_w64(ok);
  int myleader = lane, myoff = 0, cnt = 0;
  while (todo != 0) {
    const int leader = __builtin_ctzll(todo);
    const int lk = __builtin_amdgcn_readlane(key, leader);
    const uint64_t grp = __builtin_amdgcn_ballot_w64(key == lk) & todo;
    if (lane == leader) cnt = __builtin_popcountll(grp);
    if ((grp >> lane) & 1) {
      myleader = leader;
      myoff = __builtin_popcountll(grp & ((1ull << lane) - 1));
    }
    todo &= ~grp;
  }
  int base = 0;
  if (cnt > 0) base = atomicAdd(&counts[key], cnt);
  base = __shfl(base, myleader, 64);
  if (ok) slots[i] = base + myoff;
}

__global__ __launch_bounds__(TPB) void k_scatter(int n, const int* __restrict__ keys, const int* __restrict__ slots,
                          const int* __restrict__ offs, int* __restrict__ perm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[offs[keys[i]] + slots[i]] = i;
}

__global__ __launch_bounds__(TPB) void k_sort_bins(int nb, const int* __restrict__ offs, int* __restrict__ perm) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int s = offs[b], e = offs[b + 1];
  if (e - s <= 1) return;
  if (e - s <= 16) {  // in registers: load, sorting network (INT_MAX padding), store
    int v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = s + q < e ? perm[s + q] : INT_MAX;
#pragma unroll
    for (int size = 2; size <= 16; size <<= 1)
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int r = q ^ stride;
          if (r > q) {
            const bool up = (q & size) == 0;
            const int lo = min(v[q], v[r]), hi = max(v[q], v[r]);
            v[q] = up ? lo : hi;
            v[r] = up ? hi : lo;
          }
        }
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (s + q < e) perm[s + q] = v[q];
    return;
  }
  for (int i = s + 1; i < e; i++) {
    int v = perm[i];
    int j = i - 1;
    while (j >= s && perm[j] > v) {
      perm[j + 1] = perm[j];
      j--;
    }
    perm[j + 1] = v;
  }
}

// Bin-ordered particle records for the P2G gather of face grid `axis`: the velocity component
// and the 12 one-dimensional kernel factors N(|p_c - x_node|/h_c) of the 4 nodes per axis
// around TGrid::nearest. A particle in bin b (its nearest node) meets node n at offset
// n - b + 1 on each axis, so the node weight CubicKernel::weight(h, p - pn) is
// (W[ox] * W[4+oy]) * W[8+oz], the same product the reference forms for that pair.
template <typename T>
__global__ __launch_bounds__(TPB) void k_p2g_records(int n, int axis, GridArgs<T> ga, const int* __restrict__ perm, const T* __restrict__ px,
                              const T* __restrict__ py, const T* __restrict__ pz, const T* __restrict__ vc,
                              T* __restrict__ W, T* __restrict__ sv) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int p = perm[t];
  const T pp[3] = {px[p], py[p], pz[p]};
  sv[t] = vc[p];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T off = (c == axis) ? T(0) : T(0.5);
    const int nb = nearest1(pp[c], off, ga.h[c]);
#pragma unroll
    for (int o = 0; o < 4; o++) {
      const T pn = coord1(nb - 1 + o, off, ga.h[c]);
      W[(int64_t)(4 * c + o) * n + t] = bspline(wdiv(fabs(pp[c] - pn), ga.h[c]));
    }
  }
}

// P2G gather for one face grid (FlipAdvectionSolver3D::solveP2G, advect-solver.cc:188-271).
// Node (x, y, z) sums v*w and w over the particles whose nearest node b has x-b0, y-b1, z-b2
// in {-1..2}, bins in lexicographic (b0, b1, b2) order and particles by index inside a bin;
// then ug = sum(v w) / sum(w) where the weight sum is positive, else 0 and not valid.
//
// A workgroup owns a tile of up to P2G_TZ consecutive z-nodes (one per lane) by up to P2G_TY
// y-nodes over a segment [xs, xe) of x-planes. It walks the bin slabs b0 = xs-2 .. xe and,
// inside a slab, the bin rows b1 = y0-2 .. y1 in ascending order. The bins z0-2 .. z1 of a row
// hold one contiguous run of bin-ordered particle records, which the workgroup copies to LDS
// with coalesced loads. Wave k owns the x-planes x with (x - xs) % 2 == k % 2 and half of the
// y-nodes (k / 2): at slab b0 two of the four planes b0-1 .. b0+2 belong to each wave, a wave
// reads the shared factors of a particle once for both, and skips a row that reaches none of
// its y-nodes; plane b0-1 is complete (its wave writes it) once slab b0 is done. A lane walks
// the particles of bins z-2 .. z+1 of the row in order, so every node still sees its
// contributions in (b0, b1, b2, index) order.
constexpr int P2G_TZ = 64;    // z-nodes per workgroup, one per lane of a wave
constexpr int P2G_TY = 8;     // y-nodes per tile (4 per wave)
constexpr int P2G_WAVES = 4;  // two x-parities by two y-halves

// Particles per staged chunk (13 fields each); a row with more is staged in several chunks.
template <typename T>
struct P2GCap {
  static constexpr int value = sizeof(T) == 8 ? 640 : 768;
  static constexpr int padded = value + value / 8;
};

// Staged chunk layout: fields 0-3 (x-factors), 8-11 (z-factors) and 12 (value) as arrays of CP
// slots; the four y-factors of a slot together at [4 CP + 4 slot] (one 16-byte read per particle).
template <typename T>
__device__ __forceinline__ T& p2g_field(T* L, int f, int s) {
  constexpr int CP = P2GCap<T>::padded;
  return (f >= 4 && f < 8) ? L[4 * CP + 4 * s + (f - 4)] : L[f * CP + s];
}

// LDS slot of particle i of a chunk: one pad word after every 8, so that the lanes of a wave,
// whose particles lie about one bin (about 8 particles) apart, fall into different banks.
__device__ __forceinline__ int p2g_slot(int i) { return i + (i >> 3); }

// One lane over its particles [t0, t1) of a staged row, for the wave's two x-planes (offsets oN
// and oN + 2 from b0 - 1) and its P2G_YH y-nodes, whose y-offset is C = y0 + YH h - b1 + 1: y-node
// j takes the y-factor W[4 + j + C] and is reached when 0 <= j + C <= 3. The z-factor index is
// 3 - (bins of the lane's four passed, from the bin bounds B1 .. B3 in registers). FP64 forms
// (wx * wy) * wz, v * w and the sums in the reference's order; FP32 (mixed precision)
// reassociates to (v wx wz) * wy and (wx wz) * wy, two fused multiply-adds per node. The loop is
// bound by LDS reads (at the irregular slot spacing of the bins): five per particle and lane for
// two planes, the four y-factors in one 16-byte read.
constexpr int P2G_YH = P2G_TY / 2;  // y-nodes per wave: the waves split the tile's y-nodes in two
template <typename T, bool NORM, int C>
__device__ __forceinline__ void p2g_row(const T* __restrict__ L, int oN, int t0, int t1, int c0, int B1, int B2,
                                        int B3, T (&acc)[2][P2G_YH], T (&ws)[2][P2G_YH]) {
  constexpr int CP = P2GCap<T>::padded;
  typedef T T4 __attribute__((ext_vector_type(4)));
  if (t0 >= t1) return;
  // The operands of the next particle are read while the current one is summed (up to two slots
  // past the chunk are read after the last one: the LDS array has room for them), so a particle
  // does not wait for its own LDS reads.
  struct Op {
    T wx0, wx1, wz, v;
    T4 wy4;
  };
  auto read = [&](int t) __attribute__((always_inline)) {
    const int s = p2g_slot(t - c0);
    const int iz = 3 - (int)(t >= B1) - (int)(t >= B2) - (int)(t >= B3);
    return Op{L[oN * CP + s], L[(oN + 2) * CP + s], L[(8 + iz) * CP + s], L[12 * CP + s],
              *reinterpret_cast<const T4*>(L + 4 * CP + 4 * s)};
  };
  auto sum = [&](const Op& op) __attribute__((always_inline)) {
    const T wx[2] = {op.wx0, op.wx1};
    const T wz = op.wz, v = op.v;
    const T4 wy4 = op.wy4;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if constexpr (sizeof(T) == 4) {
        const T xz = wx[q] * wz, vxz = v * xz;
#pragma unroll
        for (int j = 0; j < P2G_YH; j++) {
          if (j + C < 0 || j + C > 3) continue;
          acc[q][j] = fmaf(vxz, wy4[j + C], acc[q][j]);
          if (NORM) ws[q][j] = fmaf(xz, wy4[j + C], ws[q][j]);
        }
      } else {
#pragma unroll
        for (int j = 0; j < P2G_YH; j++) {
          if (j + C < 0 || j + C > 3) continue;
          const T w = wx[q] * wy4[j + C] * wz;
          acc[q][j] = acc[q][j] + v * w;
          if (NORM) ws[q][j] = ws[q][j] + w;
        }
      }
    }
  };
  // two particles per trip, operands in two register sets, so the rotation needs no copies
  Op a = read(t0);
  int t = t0;
  for (; t + 1 < t1; t += 2) {
    Op b = read(t + 1);
    __builtin_amdgcn_sched_barrier(0);  // the reads are issued before the arithmetic
    sum(a);
    asm volatile("" ::"v"(b.wx0), "v"(b.wx1), "v"(b.wz), "v"(b.v), "v"(b.wy4));
    a = read(t + 2);
    __builtin_amdgcn_sched_barrier(0);
    sum(b);
    asm volatile("" ::"v"(a.wx0), "v"(a.wx1), "v"(a.wz), "v"(a.v), "v"(a.wy4));
  }
  if (t < t1) sum(a);
}

// Tile `tile` of a face grid (gx, gy, gz) with bins (bx, by, bz): y-tiles fastest, then
// z-tiles, then x-segments. Nodes [xs, xe) x [y0, y1) x [z0, z1); bin rows b1 = b1s ..
// b1s + nb1 - 1 of the slabs b0 = xs - 2 .. xe (nrow rows in all), bins zlo .. zhi of a row.
struct P2GTile {
  int y0, y1, z0, z1, xs, xe, zlo, zhi, b1s, nb1, nrow;
  __device__ P2GTile(int tile, int gx, int gy, int gz, int nzt, int nyt, int nseg, int by, int bz) {
    const int yt = tile % nyt, zt = (tile / nyt) % nzt, seg = tile / (nyt * nzt);
    y0 = yt * gy / nyt;
    y1 = (yt + 1) * gy / nyt;
    z0 = zt * gz / nzt;
    z1 = (zt + 1) * gz / nzt;
    xs = seg * gx / nseg;
    xe = (seg + 1) * gx / nseg;
    zlo = max(z0 - 2, 0);
    zhi = min(z1, bz - 1);
    b1s = max(y0 - 2, 0);
    nb1 = min(y1, by - 1) - b1s + 1;
    nrow = (xe - xs + 3) * nb1;
  }
};

// Tile order for the gather, heaviest first. Tiles differ in work by orders of magnitude (a
// tile in air stages no particle), and the hardware hands workgroups to the XCDs round-robin in
// index order, so the gather takes tile order[blockIdx.x]: a longest-first list schedule that
// spreads the heavy tiles over all XCDs. The work of a tile is the number of particle records
// it stages; key = (~work << 32) | tile sorts ascending into that order.
constexpr int P2G_MAX_TILES = 8192;

__global__ __launch_bounds__(TPB) void k_p2g_tile_work(int gx, int gy, int gz, int nzt, int nyt, int nseg, int bx,
                                                       int by, int bz, const int* __restrict__ offs,
                                                       unsigned long long* __restrict__ keys) {
  __shared__ unsigned part[TPB / 64];
  const int tile = blockIdx.x;
  const P2GTile g(tile, gx, gy, gz, nzt, nyt, nseg, by, bz);
  unsigned cnt = 0;
  for (int ri = threadIdx.x; ri < g.nrow; ri += TPB) {
    const int b0 = g.xs - 2 + ri / g.nb1, b1 = g.b1s + ri % g.nb1;
    if (b0 < 0 || b0 >= bx) continue;
    const int row = bz * (b1 + by * b0);
    cnt += (unsigned)(offs[row + g.zhi + 1] - offs[row + g.zlo]);
  }
  for (int m = 32; m > 0; m >>= 1) cnt += __shfl_xor(cnt, m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned w = 0;
    for (int k = 0; k < TPB / 64; k++) w += part[k];
    keys[tile] = ((unsigned long long)(0xFFFFFFFFu - w) << 32) | (unsigned)tile;
  }
}

// Bitonic sort of n <= P2G_MAX_TILES keys in one workgroup; writes the tile of each rank.
__global__ __launch_bounds__(1024) void k_p2g_tile_order(int n, const unsigned long long* __restrict__ keys,
                                                         int* __restrict__ order) {
  __shared__ unsigned long long k[P2G_MAX_TILES];
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = threadIdx.x; i < P; i += 1024) k[i] = i < n ? keys[i] : ~0ull;
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < P / 2; i += 1024) {
        const int a = 2 * i - (i & (stride - 1)), b = a + stride;  // pair i of this stage
        const bool up = (a & size) == 0;
        const unsigned long long x = k[a], y = k[b];
        if ((x > y) == up) {
          k[a] = y;
          k[b] = x;
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < n; i += 1024) order[i] = (int)(unsigned)(k[i] & 0xFFFFFFFFull);
}

template <typename T, bool NORM = true>
__global__ __launch_bounds__(P2G_TZ* P2G_WAVES) void k_p2g_gather(GridArgs<T> ga, int gx, int gy, int gz, int n,
                                                                   int nzt, int nyt, int nseg,
                                                                   const int* __restrict__ order,
                                                                   const int* __restrict__ offs,
                                                                   const T* __restrict__ W, const T* __restrict__ vc,
                                                                   T* __restrict__ G, T* __restrict__ Wt,
                                                                   uint8_t* __restrict__ valid) {
  constexpr int CAP = P2GCap<T>::value, CP = P2GCap<T>::padded;
  __shared__ __align__(16) T L[13 * CP + 8];  // + the slots read past a full chunk

  // Without an order, workgroups go round-robin over the 8 XCDs (id = 8 * k + xcd) and each
  // XCD takes a contiguous run of tiles.
  int tile;
  if (order) {
    tile = order[blockIdx.x];
  } else {
    const int nwg = nzt * nyt * nseg;
    const int xcd = blockIdx.x & 7, q = nwg >> 3, r = nwg & 7;
    tile = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + (int)(blockIdx.x >> 3);
  }
  const int bx = ga.nx + 1, by = ga.ny + 1, bz = ga.nz + 1;
  const P2GTile g(tile, gx, gy, gz, nzt, nyt, nseg, by, bz);
  const int y0 = g.y0, y1 = g.y1, z0 = g.z0, z1 = g.z1, xs = g.xs, xe = g.xe;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int z = z0 + lane;
  const bool zok = z < z1;
  const int zlo = g.zlo, zhi = g.zhi;
  // Wave w owns the x-planes x with (x - xs) % 2 == w % 2 and the y-nodes y0 + YH (w / 2) + j.
  // At slab b0 it holds two planes, b0 - 1 + oN (near) and b0 + 1 + oN (far), oN = 0 or 1; when
  // oN = 0 the near plane has seen its last slab after this one and the far plane moves up.
  const int par = wave & 1, yh = wave >> 1, ya = y0 + P2G_YH * yh;
  T acc[2][P2G_YH], ws[2][P2G_YH];
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int j = 0; j < P2G_YH; j++) acc[q][j] = ws[q][j] = T(0);
  // Rows in walk order: ri = (b0 - xs + 2) * nb1 + (b1 - b1s). While row ri is summed, the
  // first chunk of row ri + 1 is loaded into registers and the bin offsets of row ri + 2 (the
  // row's record range and the lane's bin bounds) are read, so no load waits on another.
  const int b1s = g.b1s, nb1 = g.nb1, nrow = g.nrow;
  constexpr int NT = P2G_TZ * P2G_WAVES, PF = (CAP + NT - 1) / NT;
  T pre[PF][13];
  struct RowOff {
    int r0, r1, B[5];  // the row's records [r0, r1); the lane's bins z-2 .. z+1 start at B[0 .. 3], end at B[4]
  };
  auto offsets = [&](int ri) {
    RowOff q{0, 0, {0, 0, 0, 0, 0}};
    const int b0 = xs - 2 + ri / nb1, b1 = b1s + ri % nb1;
    if (ri >= nrow || b0 < 0 || b0 >= bx) return q;
    const int row = bz * (b1 + by * b0);
    q.r0 = offs[row + zlo];
    q.r1 = offs[row + zhi + 1];
#pragma unroll
    for (int m = 0; m < 5; m++) q.B[m] = zok ? offs[row + min(max(z - 2 + m, 0), bz)] : q.r0;
    return q;
  };
  auto fetch = [&](const RowOff& q) {
    const int pr0 = q.r0, cnt = min(CAP, q.r1 - q.r0);
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < cnt) {
#pragma unroll
        for (int f = 0; f < 12; f++) pre[k][f] = W[(int64_t)f * n + pr0 + i];
        pre[k][12] = vc[pr0 + i];
      }
    }
  };
  auto sum_chunk = [&](int oN, int cy, int c0, int cnt, const int (&B)[5]) {
    const int t0 = max(B[0], c0), t1 = min(B[4], c0 + cnt);
    switch (cy) {  // a row with |cy| > 3 reaches none of the wave's y-nodes
#define FS_P2G_ROW(CC) \
  case CC:             \
    p2g_row<T, NORM, CC>(L, oN, t0, t1, c0, B[1], B[2], B[3], acc, ws); \
    break;
      FS_P2G_ROW(3)
      FS_P2G_ROW(2)
      FS_P2G_ROW(1)
      FS_P2G_ROW(0)
      FS_P2G_ROW(-1)
      FS_P2G_ROW(-2)
      FS_P2G_ROW(-3)
#undef FS_P2G_ROW
      default:
        break;
    }
  };
  RowOff oc = offsets(0), on = offsets(1);
  fetch(oc);
  for (int ri = 0; ri < nrow; ri++) {
    const int b0 = xs - 2 + ri / nb1, b1 = b1s + ri % nb1;
    const int oN = (par + 1 + xs - b0) & 1;  // near plane x = b0 - 1 + oN, x - xs = par (mod 2)
    const int x = b0 - 1 + oN;
    const bool live = x + 2 >= xs && x < xe;  // either plane in the segment
    const int r0 = oc.r0, r1 = oc.r1, cy = ya - b1 + 1;
    const int B[5] = {oc.B[0], oc.B[1], oc.B[2], oc.B[3], oc.B[4]};
    if (r0 != r1) {
      const int cnt = min(CAP, r1 - r0);
      __syncthreads();  // the previous row has been read
#pragma unroll
      for (int k = 0; k < PF; k++) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < cnt) {
          const int s = p2g_slot(i);
#pragma unroll
          for (int f = 0; f < 13; f++) p2g_field(L, f, s) = pre[k][f];
        }
      }
      __syncthreads();
    }
    oc = on;
    if (ri + 1 < nrow) fetch(oc);
    on = offsets(ri + 2);
    if (r0 != r1) {
      if (live) sum_chunk(oN, cy, r0, min(CAP, r1 - r0), B);
      for (int c0 = r0 + CAP; c0 < r1; c0 += CAP) {  // a row beyond one chunk: stage the rest in place
        const int cnt = min(CAP, r1 - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < cnt; i += NT) {
          const int s = p2g_slot(i);
#pragma unroll
          for (int f = 0; f < 12; f++) p2g_field(L, f, s) = W[(int64_t)f * n + c0 + i];
          L[12 * CP + s] = vc[c0 + i];
        }
        __syncthreads();
        if (live) sum_chunk(oN, cy, c0, cnt, B);
      }
    }
    if (ri % nb1 == nb1 - 1 && oN == 0) {  // near plane x = b0 - 1 has seen every slab x-2 .. x+1
      if (x >= xs && x < xe && zok) {
#pragma unroll
        for (int j = 0; j < P2G_YH; j++) {
          if (ya + j >= y1) break;
          const int64_t nid = ((int64_t)x * gy + (ya + j)) * gz + z;
          const T a = acc[0][j], s = ws[0][j];
          if (NORM) {
            G[nid] = s > T(0) ? a / s : T(0);
            valid[nid] = s > T(0) ? 1 : 0;
            Wt[nid] = s;
          } else {
            G[nid] = a;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < P2G_YH; j++) {  // the far plane (x + 2) becomes the near one
        acc[0][j] = acc[1][j];
        ws[0][j] = ws[1][j];
        acc[1][j] = ws[1][j] = T(0);
      }
    }
  }
}
static_assert(P2G_TY == 8 && P2G_WAVES == 4, "four waves: two x-parities by two halves of the tile's 8 y-nodes");

// Launch of k_p2g_gather over face grid dims (gx, gy, gz). x is cut into segments so that
// there are about 2048 workgroups (several per CU), but not below 8 planes per segment:
// every segment walks three slabs more than it has planes.
template <typename T, bool NORM>
void launch_p2g_gather(hipStream_t st, GridArgs<T> ga, int gx, int gy, int gz, int n, const int* offs, const T* W,
                       const T* vc, T* G, T* Wt, uint8_t* valid, DBuf& tiles) {
  const int nzt = (gz + P2G_TZ - 1) / P2G_TZ, nyt = (gy + P2G_TY - 1) / P2G_TY;
  const int nseg = std::max(1, std::min((2048 + nzt * nyt - 1) / (nzt * nyt), gx / 8));
  const int nwg = nzt * nyt * nseg;
  const int* order = nullptr;
  if (nwg <= P2G_MAX_TILES) {
    unsigned long long* keys = tiles.as<unsigned long long>();
    int* ord = reinterpret_cast<int*>(keys + P2G_MAX_TILES);
    k_p2g_tile_work<<<nwg, TPB, 0, st>>>(gx, gy, gz, nzt, nyt, nseg, ga.nx + 1, ga.ny + 1, ga.nz + 1, offs, keys);
    k_p2g_tile_order<<<1, 1024, 0, st>>>(nwg, keys, ord);
    order = ord;
  }
  k_p2g_gather<T, NORM><<<nwg, P2G_TZ * P2G_WAVES, 0, st>>>(ga, gx, gy, gz, n, nzt, nyt, nseg, order, offs, W, vc, G,
                                                            Wt, valid);
}

// NaiveReconstructor::reconstruct (FluidSim/include/FluidSim/cpu/sdf.h:210-225) as a particle
// splat. The reference visits, for each cell, the particles of the Morton bins in
// [x_min,x_max]x[y_min,y_max]x[z_min,z_max] (forNeighbours, ns-util.h:80-109) except the last
// particle of each bin, keeps those within 2r and sets sdf = min(|x_c - x_p| - r). The minimum
// is order-free and both sqrt and x - r are monotone, so every particle instead lowers a
// per-cell minimum of the squared distance (atomicMin on the bit pattern, which orders like
// the value for non-negative floats); every inclusion test is the reference's, evaluated with
// the reference's arithmetic: the bin ranges of the cell, and |d| <= 2r as d^2 <= s_max where
// s_max is the largest float whose rounded sqrt is <= 2r.
template <typename T>
struct Bits;
template <>
struct Bits<double> {
  using U = unsigned long long;
  static constexpr U inf = 0x7FF0000000000000ull;
  __device__ static U of(double v) { return (U)__double_as_longlong(v); }
  __device__ static double val(U b) { return __longlong_as_double((long long)b); }
};
template <>
struct Bits<float> {
  using U = unsigned int;
  static constexpr U inf = 0x7F800000u;
  __device__ static U of(float v) { return (U)__float_as_uint(v); }
  __device__ static float val(U b) { return __uint_as_float(b); }
};

// The last particle of a Morton bin in the stable order is the one with the largest index.
// Lanes with equal keys are served by one atomic from the group's highest lane (largest index).
__global__ __launch_bounds__(TPB) void k_bin_last(int n, const int* __restrict__ keys, int* __restrict__ binmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63;
  const bool ok = i < n;
  const int key = ok ? keys[i] : -1;
  uint64_t todo = __builtin_amdgcn_ballot_w64(ok);
  bool last = false;  // the highest lane of its key group; all of them update in one instruction
  while (todo != 0) {
    const int lk = __builtin_amdgcn_readlane(key, __builtin_ctzll(todo));
    const uint64_t grp = __builtin_amdgcn_ballot_w64(key == lk) & todo;
    if (lane == 63 - __builtin_clzll(grp)) last = true;
    todo &= ~grp;
  }
  if (last) atomicMax(&binmax[key], i);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_fill_bits(int n, typename Bits<T>::U* __restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = Bits<T>::inf;
}

// A workgroup takes 256 consecutive particles. Particles are stored in a spatially coherent
// order, so their candidate cells fill a small box: the workgroup lowers the minima of that box
// in LDS (atomic minimum in LDS, no memory traffic), then lowers each per-cell minimum in memory
// once per box cell that received a candidate (an agent-scope atomic that skips cells whose
// minimum is already lower). When the box holds more cells than the LDS table, the workgroup
// lowers the minima in memory directly.
constexpr int RC_LDS_BYTES = 16384;

template <typename T>
struct ReconCand {  // one particle's candidate cells (empty box when it is skipped)
  int lo[3], hi[3], bin[3];
  int2 in[3];  // per axis, the cells whose bin range holds this particle's bin
  T p[3];
};

// Cell i's bin range [mn, mx] on axis d: cell c keeps bin b when
// floor((x_c - 2r)/h) <= b <= ceil((x_c + 2r)/h), clamped to the grid (ns-util.h:80-109).
template <typename T>
__global__ void k_recon_bounds(GridArgs<T> ga, T r2, int dmax, int2* __restrict__ bnd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * dmax) return;
  const int d = t / dmax, i = t % dmax;
  const int dims[3] = {ga.nx, ga.ny, ga.nz};
  if (i >= dims[d]) return;
  const T xc = coord1(i, T(0.5), ga.h[d]);
  bnd[t] = make_int2(max(0, (int)floor((xc - r2) / ga.h[d])), min(dims[d] - 1, (int)ceil((xc + r2) / ga.h[d])));
}

// Both bounds of a cell's bin range are non-decreasing in the cell index (each step - centre,
// shift, quotient, floor or ceiling, clamp - is monotone under rounding to nearest), so the cells
// that hold bin b form one interval [imin, imax]: imin is the first cell with mx >= b, imax the
// last with mn <= b, both found by bisection over the bounds (empty: imin > imax).
__global__ void k_recon_bins(int nx, int ny, int nz, int dmax, const int2* __restrict__ bnd, int2* __restrict__ tab) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * dmax) return;
  const int d = t / dmax, b = t % dmax;
  const int n = d == 0 ? nx : (d == 1 ? ny : nz);
  const int2* bd = bnd + d * dmax;
  int lo = n, hi = -1;
  if (b < n) {
    int l = 0, r = n;  // first i in [0, n) with mx(i) >= b
    while (l < r) {
      const int m = (l + r) >> 1;
      if (bd[m].y >= b) r = m;
      else l = m + 1;
    }
    lo = l;
    l = 0;
    r = n;  // first i with mn(i) > b, minus one
    while (l < r) {
      const int m = (l + r) >> 1;
      if (bd[m].x > b) r = m;
      else l = m + 1;
    }
    hi = l - 1;
  }
  tab[t] = make_int2(lo, hi);
}

template <typename T>
__device__ inline ReconCand<T> recon_candidates(int id, int n, const GridArgs<T>& ga, const T* px, const T* py,
                                                const T* pz, const int* keys, const int* binmax, T r2,
                                                const int2* bins, int dmax) {
  ReconCand<T> c;
  const bool skip = id >= n || binmax[keys[id]] == id;  // forNeighbours stops one short of the inclusive bin end
  const int dims[3] = {ga.nx, ga.ny, ga.nz};
  c.p[0] = skip ? T(0) : px[id];
  c.p[1] = skip ? T(0) : py[id];
  c.p[2] = skip ? T(0) : pz[id];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int bb = (int)floor(c.p[d] / ga.h[d]);  // ParticleNeighbourSearcher::update (ns-util.h:52-60)
    bb = bb < 0 ? 0 : bb;
    c.bin[d] = bb >= dims[d] ? dims[d] - 1 : bb;
    c.in[d] = bins[d * dmax + c.bin[d]];
    // cells whose centre (i + 1/2) h can lie within r2 of p: (p - r2)/h - 1/2 <= i <= (p + r2)/h - 1/2,
    // widened by a hundredth of a cell so that rounding of these bounds (and the reciprocal in
    // place of the quotient) never drops a cell the exact distance test below would keep (that
    // test alone decides)
    const T ih = T(1) / ga.h[d];
    c.lo[d] = max(0, (int)floor((c.p[d] - r2) * ih - T(0.5) - T(0.01)) + 1);
    c.hi[d] = min(dims[d] - 1, (int)ceil((c.p[d] + r2) * ih - T(0.5) + T(0.01)) - 1);
    c.lo[d] = max(c.lo[d], c.in[d].x);  // and whose bin range holds the particle's bin
    c.hi[d] = min(c.hi[d], c.in[d].y);
    if (skip) {
      c.lo[d] = INT_MAX / 4;
      c.hi[d] = -1;
    }
  }
  return c;
}

// Visits the candidate cells (i, j, k) of one particle with their squared distance: cell c keeps
// the particle's bin on every axis (the cell index lies in the particle's interval c.in) and
// |x_c - x_p| <= 2r as d^2 <= s_max, with d^2 = (dx^2 + dy^2) + dz^2. For each (i, j) the k range
// is narrowed to the cells whose dz^2 can fit under s_max - (dx^2 + dy^2), widened by a hundredth
// of a cell so that rounding never drops a cell the exact test keeps. The x rows are walked from a
// lane-dependent start, so particles of one cell (neighbouring lanes, same box) do not update the
// same LDS minimum at the same step. The set of visited cells and each distance do not depend on
// any of this.
template <typename T, class F>
__device__ inline void recon_visit(const ReconCand<T>& c, const GridArgs<T>& ga, T r2, T s_max, int rot, F&& fn) {
  const int ilo = max(c.lo[0], c.in[0].x), ihi = min(c.hi[0], c.in[0].y);
  const int jlo = max(c.lo[1], c.in[1].x), jhi = min(c.hi[1], c.in[1].y);
  const int klo = max(c.lo[2], c.in[2].x), khi = min(c.hi[2], c.in[2].y);
  const int ni = ihi - ilo + 1;
  if (ni <= 0 || jhi < jlo || khi < klo) return;
  const float ihz = 1.0f / (float)ga.h[2];
  int oi = rot % ni;
  for (int t = 0; t < ni; t++, oi = (oi + 1 == ni) ? 0 : oi + 1) {
    const int i = ilo + oi;
    const T xc = coord1(i, T(0.5), ga.h[0]);
    const T dx = xc - c.p[0];
    const T dx2 = dx * dx;
    if (dx2 > s_max) continue;
    for (int j = jlo; j <= jhi; j++) {
      const T yc = coord1(j, T(0.5), ga.h[1]);
      const T dy = yc - c.p[1];
      const T dxy = dx2 + dy * dy;
      if (dxy > s_max) continue;
      // bounds only (the margin absorbs the approximate square root and reciprocal)
      const float R = __builtin_amdgcn_sqrtf((float)(s_max - dxy));
      const float pz = (float)c.p[2] * ihz;
      const int kl = max(klo, (int)floorf(pz - R * ihz - 0.51f) + 1);
      const int kh = min(khi, (int)ceilf(pz + R * ihz - 0.49f) - 1);
      for (int k = kl; k <= kh; k++) {
        const T zc = coord1(k, T(0.5), ga.h[2]);
        const T dz = zc - c.p[2];
        const T d2 = dxy + dz * dz;
        if (d2 > s_max) continue;
        fn(i, j, k, d2);
      }
    }
  }
}

__device__ inline void lower_global(unsigned int* a, unsigned int v) {
  if (v < __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    __hip_atomic_fetch_min(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lower_global(unsigned long long* a, unsigned long long v) {
  if (v < __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    __hip_atomic_fetch_min(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_recon_splat(int n, GridArgs<T> ga, const T* __restrict__ px, const T* __restrict__ py,
                              const T* __restrict__ pz, const int* __restrict__ keys, const int* __restrict__ binmax,
                              T r2, T s_max, const int2* __restrict__ bins, int dmax,
                              typename Bits<T>::U* __restrict__ dmin) {
  using UB = typename Bits<T>::U;
  constexpr int CAP = RC_LDS_BYTES / (int)sizeof(UB);
  __shared__ UB table[CAP];
  __shared__ int box[6];
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const ReconCand<T> c = recon_candidates(id, n, ga, px, py, pz, keys, binmax, r2, bins, dmax);
  if (threadIdx.x < 6) box[threadIdx.x] = threadIdx.x < 3 ? INT_MAX : -1;
  __syncthreads();
  // the workgroup's box: reduced within each wave first (a skipped particle's empty box, lo =
  // INT_MAX / 4 and hi = -1, changes neither bound), then one LDS atomic per wave and bound
  int blo[3], bhi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    blo[d] = c.lo[d];
    bhi[d] = c.hi[d];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      blo[d] = min(blo[d], __shfl_xor(blo[d], off, 64));
      bhi[d] = max(bhi[d], __shfl_xor(bhi[d], off, 64));
    }
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (bhi[d] >= 0) {
        atomicMin(&box[d], blo[d]);
        atomicMax(&box[3 + d], bhi[d]);
      }
    }
  }
  __syncthreads();
  const int L0 = box[0], L1 = box[1], L2 = box[2];
  const int e0 = box[3] - L0 + 1, e1 = box[4] - L1 + 1, e2 = box[5] - L2 + 1;
  if (e0 <= 0 || e1 <= 0 || e2 <= 0) return;  // uniform: no particle of the workgroup is kept
  const long long vol = (long long)e0 * e1 * e2;
  if (vol > CAP) {  // uniform: lower the minima in memory directly
    recon_visit(c, ga, r2, s_max, threadIdx.x & 7, [&](int i, int j, int k, T d2) { lower_global(dmin + lin3(i, j, k, ga.ny, ga.nz), Bits<T>::of(d2)); });
    return;
  }
  for (int q = threadIdx.x; q < (int)vol; q += TPB) table[q] = Bits<T>::inf;
  __syncthreads();
  recon_visit(c, ga, r2, s_max, threadIdx.x & 7, [&](int i, int j, int k, T d2) {
    atomicMin(&table[((i - L0) * e1 + (j - L1)) * e2 + (k - L2)], Bits<T>::of(d2));
  });
  __syncthreads();
  for (int q = threadIdx.x; q < (int)vol; q += TPB) {
    const UB v = table[q];
    if (v == Bits<T>::inf) continue;
    const int k = L2 + q % e2, j = L1 + (q / e2) % e1, i = L0 + q / (e2 * e1);
    lower_global(dmin + lin3(i, j, k, ga.ny, ga.nz), v);
  }
}

// sdf.grid.fill(1e9); sdfValid.fill(false); then min(sdf, |d| - r) for the kept particles.
template <typename T>
__global__ __launch_bounds__(TPB) void k_recon_final(int ncell, const typename Bits<T>::U* __restrict__ dmin, T radius,
                                                     T* __restrict__ sdf, uint8_t* __restrict__ valid) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const typename Bits<T>::U b = dmin[c];
  T out = T(1e9);
  uint8_t v = 0;
  if (b != Bits<T>::inf) {
    const T dis = sqrt(Bits<T>::val(b)) - radius;
    out = smin(out, dis);
    v = 1;
  }
  sdf[c] = out;
  valid[c] = v;
}

// extrapolateFluidSdf iteration (cpu/fluid-simulator.h:124-164): valid cells are not copied.
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_sdf(int nx, int ny, int nz, const T* __restrict__ cur, T* __restrict__ buf,
                             const uint8_t* __restrict__ val, uint8_t* __restrict__ vbuf) {
  int i, j, k, id;
  if (!g3_index(ny, nz, i, j, k, id)) return;
  if (val[id]) {
    vbuf[id] = 1;
    return;
  }
  T sum = T(0);
  int count = 0;
  const int sx = ny * nz, sy = nz;
  if (i > 0 && val[id - sx]) { sum += cur[id - sx]; count++; }
  if (i < nx - 1 && val[id + sx]) { sum += cur[id + sx]; count++; }
  if (j > 0 && val[id - sy]) { sum += cur[id - sy]; count++; }
  if (j < ny - 1 && val[id + sy]) { sum += cur[id + sy]; count++; }
  if (k > 0 && val[id - 1]) { sum += cur[id - 1]; count++; }
  if (k < nz - 1 && val[id + 1]) { sum += cur[id + 1]; count++; }
  if (count > 0) {
    buf[id] = sum / (T)count;
    vbuf[id] = 1;
  } else {
    vbuf[id] = 0;
  }
}

// Four consecutive cells per thread (nz a multiple of 4): the flags of four cells come in one
// 4-byte load, the values of each x / y neighbour row in one 16-byte (FP32) or 32-byte (FP64)
// load; per cell the same terms in the same order as k_extrap_sdf, and the same writes (a valid
// cell and a cell with no valid neighbour leave the value buffer as it was).
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_sdf4(int nx, int ny, int nz, const T* __restrict__ cur,
                                                     T* __restrict__ buf, const uint8_t* __restrict__ val,
                                                     uint8_t* __restrict__ vbuf) {
  const int k0 = 4 * ((int)blockIdx.x * G3X + (int)threadIdx.x);
  const int j = (int)blockIdx.y * G3Y + (int)threadIdx.y;
  const int i = (int)blockIdx.z;
  if (k0 >= nz || j >= ny) return;
  const int id0 = (i * ny + j) * nz + k0;
  const int sx = ny * nz, sy = nz;
  auto flags = [&](int id) {
    const uchar4 f = *reinterpret_cast<const uchar4*>(val + id);
    return (unsigned)f.x | ((unsigned)f.y << 8) | ((unsigned)f.z << 16) | ((unsigned)f.w << 24);
  };
  auto values = [&](int id, T* v) {
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = cur[id + q];
  };
  const unsigned fc = flags(id0);
  if ((fc & 0x01010101u) == 0x01010101u) {  // four valid cells: flags only
    *reinterpret_cast<uchar4*>(vbuf + id0) = make_uchar4(1, 1, 1, 1);
    return;
  }
  const unsigned fxm = i > 0 ? flags(id0 - sx) : 0u, fxp = i < nx - 1 ? flags(id0 + sx) : 0u;
  const unsigned fym = j > 0 ? flags(id0 - sy) : 0u, fyp = j < ny - 1 ? flags(id0 + sy) : 0u;
  const bool km = k0 > 0, kp = k0 + 4 < nz;
  const uint8_t fkm = km ? val[id0 - 1] : 0, fkp = kp ? val[id0 + 4] : 0;
  T vxm[4] = {}, vxp[4] = {}, vym[4] = {}, vyp[4] = {}, vc[4] = {};
  if (fxm) values(id0 - sx, vxm);
  if (fxp) values(id0 + sx, vxp);
  if (fym) values(id0 - sy, vym);
  if (fyp) values(id0 + sy, vyp);
  if (fc) values(id0, vc);
  const T ckm = fkm ? cur[id0 - 1] : T(0), ckp = fkp ? cur[id0 + 4] : T(0);
  unsigned fo = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const unsigned bit = 0xFFu << (8 * q);
    if (fc & bit) {
      fo |= 1u << (8 * q);
      continue;
    }
    T sum = T(0);
    int count = 0;
    if (fxm & bit) { sum += vxm[q]; count++; }
    if (fxp & bit) { sum += vxp[q]; count++; }
    if (fym & bit) { sum += vym[q]; count++; }
    if (fyp & bit) { sum += vyp[q]; count++; }
    if (q > 0 ? (fc & (bit >> 8)) != 0 : fkm != 0) { sum += q > 0 ? vc[q - 1] : ckm; count++; }
    if (q < 3 ? (fc & (bit << 8)) != 0 : fkp != 0) { sum += q < 3 ? vc[q + 1] : ckp; count++; }
    if (count > 0) {
      buf[id0 + q] = sum / (T)count;
      fo |= 1u << (8 * q);
    }
  }
  *reinterpret_cast<uchar4*>(vbuf + id0) =
      make_uchar4((uint8_t)(fo & 1u), (uint8_t)((fo >> 8) & 1u), (uint8_t)((fo >> 16) & 1u), (uint8_t)(fo >> 24));
}

// smoothFluidSurface iteration (FluidSim/src/cpu/fluid-simulator.cc:40-71): min(avg6, self).
// FLAG also records whether the cell took the average (the reverse pass needs the branch).
template <typename T, bool FLAG = false>
__global__ __launch_bounds__(TPB) void k_smooth(int nx, int ny, int nz, const T* __restrict__ cur, T* __restrict__ buf,
                         uint8_t* __restrict__ took_avg = nullptr) {
  int i, j, k, id;
  if (!g3_index(ny, nz, i, j, k, id)) return;
  T sum = T(0);
  int count = 0;
  const int sx = ny * nz, sy = nz;
  if (i > 0) { sum += cur[id - sx]; count++; }
  if (i < nx - 1) { sum += cur[id + sx]; count++; }
  if (j > 0) { sum += cur[id - sy]; count++; }
  if (j < ny - 1) { sum += cur[id + sy]; count++; }
  if (k > 0) { sum += cur[id - 1]; count++; }
  if (k < nz - 1) { sum += cur[id + 1]; count++; }
  T b = sum / (T)count;
  T c0 = cur[id];
  if (FLAG) took_avg[id] = !(c0 < b);
  if (c0 < b) b = c0;
  buf[id] = b;
}

// k_smooth for four consecutive cells per thread (nz a multiple of 4), the same terms in the
// same order per cell.
template <typename T, bool FLAG = false>
__global__ __launch_bounds__(TPB) void k_smooth4(int nx, int ny, int nz, const T* __restrict__ cur,
                                                 T* __restrict__ buf, uint8_t* __restrict__ took_avg = nullptr) {
  const int k0 = 4 * ((int)blockIdx.x * G3X + (int)threadIdx.x);
  const int j = (int)blockIdx.y * G3Y + (int)threadIdx.y;
  const int i = (int)blockIdx.z;
  if (k0 >= nz || j >= ny) return;
  const int id0 = (i * ny + j) * nz + k0;
  const int sx = ny * nz, sy = nz;
  auto values = [&](int id, T* v) {
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = cur[id + q];
  };
  T vxm[4] = {}, vxp[4] = {}, vym[4] = {}, vyp[4] = {}, vc[4];
  values(id0, vc);
  if (i > 0) values(id0 - sx, vxm);
  if (i < nx - 1) values(id0 + sx, vxp);
  if (j > 0) values(id0 - sy, vym);
  if (j < ny - 1) values(id0 + sy, vyp);
  const T ckm = k0 > 0 ? cur[id0 - 1] : T(0), ckp = k0 + 4 < nz ? cur[id0 + 4] : T(0);
  uint8_t fl[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    T sum = T(0);
    int count = 0;
    if (i > 0) { sum += vxm[q]; count++; }
    if (i < nx - 1) { sum += vxp[q]; count++; }
    if (j > 0) { sum += vym[q]; count++; }
    if (j < ny - 1) { sum += vyp[q]; count++; }
    if (k0 + q > 0) { sum += q > 0 ? vc[q - 1] : ckm; count++; }
    if (k0 + q < nz - 1) { sum += q < 3 ? vc[q + 1] : ckp; count++; }
    T b = sum / (T)count;
    const T c0 = vc[q];
    fl[q] = !(c0 < b);
    if (c0 < b) b = c0;
    buf[id0 + q] = b;
  }
  if (FLAG) *reinterpret_cast<uchar4*>(took_avg + id0) = make_uchar4(fl[0], fl[1], fl[2], fl[3]);
}

// applyDirichletBoundary (fluid-simulator.cc:113-138) for face grid `axis`.
template <typename T>
__global__ __launch_bounds__(TPB) void k_dirichlet(int axis, int gx, int gy, int gz, T* __restrict__ G, uint8_t* __restrict__ valid) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= gx * gy * gz) return;
  int i, j, k;
  unlin3(id, gy, gz, i, j, k);
  int c = axis == 0 ? i : (axis == 1 ? j : k);
  int d = axis == 0 ? gx : (axis == 1 ? gy : gz);
  if (c == 0 || c == d - 1) {
    G[id] = T(0);
    valid[id] = 1;
  }
}

// extrapolate<GridType> iteration (cpu/fluid-simulator.h:209-251).
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_vel(int gx, int gy, int gz, const T* __restrict__ g, T* __restrict__ gbuf,
                             const uint8_t* __restrict__ val, uint8_t* __restrict__ vbuf) {
  int i, j, k, id;
  if (!g3_index(gy, gz, i, j, k, id)) return;
  if (val[id]) {
    vbuf[id] = 1;
    return;
  }
  T sum = T(0);
  int count = 0;
  const int sx = gy * gz, sy = gz;
  if (i > 0 && val[id - sx]) { sum += g[id - sx]; count++; }
  if (i < gx - 1 && val[id + sx]) { sum += g[id + sx]; count++; }
  if (j > 0 && val[id - sy]) { sum += g[id - sy]; count++; }
  if (j < gy - 1 && val[id + sy]) { sum += g[id + sy]; count++; }
  if (k > 0 && val[id - 1]) { sum += g[id - 1]; count++; }
  if (k < gz - 1 && val[id + 1]) { sum += g[id + 1]; count++; }
  if (count > 0) {
    gbuf[id] = sum / (T)count;
    vbuf[id] = 1;
  } else {
    gbuf[id] = T(0);
    vbuf[id] = 0;
  }
}

// The same iteration for four consecutive z-faces per thread, for face grids whose z extent is a
// multiple of 4 (rows stay 4-byte aligned): the flags of a face and of its x and y neighbours come
// as one 4-byte load per direction, the values as four loads per direction where a flag asks for
// them. Every face sums its valid neighbours in the order above and divides as above; valid faces
// write only their flag (the value buffer keeps its stale entry). Other grids take
// k_extrap_vel4f below (a byte-by-byte form of this kernel for them ran slower than the one-face
// kernel).
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_vel4(int gx, int gy, int gz, const T* __restrict__ g,
                                                     T* __restrict__ gbuf, const uint8_t* __restrict__ val,
                                                     uint8_t* __restrict__ vbuf) {
  const int k0 = 4 * ((int)blockIdx.x * G3X + (int)threadIdx.x);
  const int j = (int)blockIdx.y * G3Y + (int)threadIdx.y;
  const int i = (int)blockIdx.z;
  if (k0 >= gz || j >= gy) return;
  const int id0 = (i * gy + j) * gz + k0;
  const int sx = gy * gz, sy = gz;
  auto flags = [&](int id) {
    const uchar4 f = *reinterpret_cast<const uchar4*>(val + id);
    return (unsigned)f.x | ((unsigned)f.y << 8) | ((unsigned)f.z << 16) | ((unsigned)f.w << 24);
  };
  auto values = [&](int id, T* v) {
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = g[id + q];
  };
  const unsigned fc = flags(id0);
  if ((fc & 0x01010101u) == 0x01010101u) {  // four valid faces: flags only
    *reinterpret_cast<uchar4*>(vbuf + id0) = make_uchar4(1, 1, 1, 1);
    return;
  }
  const unsigned fxm = i > 0 ? flags(id0 - sx) : 0u, fxp = i < gx - 1 ? flags(id0 + sx) : 0u;
  const unsigned fym = j > 0 ? flags(id0 - sy) : 0u, fyp = j < gy - 1 ? flags(id0 + sy) : 0u;
  const bool km = k0 > 0, kp = k0 + 4 < gz;
  const uint8_t fkm = km ? val[id0 - 1] : 0, fkp = kp ? val[id0 + 4] : 0;
  T vxm[4] = {}, vxp[4] = {}, vym[4] = {}, vyp[4] = {}, vc[4] = {};
  if (fxm) values(id0 - sx, vxm);
  if (fxp) values(id0 + sx, vxp);
  if (fym) values(id0 - sy, vym);
  if (fyp) values(id0 + sy, vyp);
  if (fc) values(id0, vc);
  const T gkm = fkm ? g[id0 - 1] : T(0), gkp = fkp ? g[id0 + 4] : T(0);
  unsigned fo = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const unsigned bit = 0xFFu << (8 * q);
    if (fc & bit) {
      fo |= 1u << (8 * q);
      continue;
    }
    T sum = T(0);
    int count = 0;
    if (fxm & bit) { sum += vxm[q]; count++; }
    if (fxp & bit) { sum += vxp[q]; count++; }
    if (fym & bit) { sum += vym[q]; count++; }
    if (fyp & bit) { sum += vyp[q]; count++; }
    if (q > 0 ? (fc & (bit >> 8)) != 0 : fkm != 0) { sum += q > 0 ? vc[q - 1] : gkm; count++; }
    if (q < 3 ? (fc & (bit << 8)) != 0 : fkp != 0) { sum += q < 3 ? vc[q + 1] : gkp; count++; }
    if (count > 0) {
      gbuf[id0 + q] = sum / (T)count;
      fo |= 1u << (8 * q);
    } else {
      gbuf[id0 + q] = T(0);
    }
  }
  *reinterpret_cast<uchar4*>(vbuf + id0) =
      make_uchar4((uint8_t)(fo & 1u), (uint8_t)((fo >> 8) & 1u), (uint8_t)((fo >> 16) & 1u), (uint8_t)(fo >> 24));
}

// The same iteration for face grids whose z extent is not a multiple of 4 (the w grid, nz + 1
// faces per row), over the flat array: a thread takes four consecutive faces e0 .. e0+3, which
// may run from the end of one row into the next. Needs gz >= 4 and gy*gz a multiple of 4, so the
// flags of a group and of its x neighbours are aligned 4-byte loads and a group never spans two
// x-slabs. The y neighbours' flags are 4 bytes at e0 -/+ gz, read as two aligned words shifted
// into place; each face masks every neighbour that is not on the grid (row and slab ends), so it
// sums the same neighbours in the same order as the one-face kernel and writes what it writes.
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_vel4f(int gx, int gy, int gz, const T* __restrict__ g,
                                                      T* __restrict__ gbuf, const uint8_t* __restrict__ val,
                                                      uint8_t* __restrict__ vbuf) {
  const int n = gx * gy * gz, sx = gy * gz, sy = gz;
  const int e0 = 4 * (int)(blockIdx.x * TPB + threadIdx.x);
  if (e0 >= n) return;
  auto word = [&](int p) { return *reinterpret_cast<const unsigned*>(val + p); };
  // flag bytes p .. p+3 (n is a multiple of 4: both words lie below n when p + 4 <= n)
  auto flags = [&](int p) -> unsigned {
    if (p >= 0 && p + 4 <= n) {
      const int base = p & ~3, sh = p & 3;
      const unsigned w0 = word(base);
      if (sh == 0) return w0;
      const unsigned long long w = ((unsigned long long)word(base + 4) << 32) | w0;
      return (unsigned)(w >> (8 * sh));
    }
    unsigned f = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (p + q >= 0 && p + q < n) f |= (unsigned)val[p + q] << (8 * q);
    return f;
  };
  const unsigned fc = word(e0);
  if ((fc & 0x01010101u) == 0x01010101u) {  // four valid faces: flags only
    *reinterpret_cast<unsigned*>(vbuf + e0) = 0x01010101u;
    return;
  }
  // row of face e0 (j, i) and of the row after it, which faces past the row's end belong to
  const int r0 = e0 / gz, k0 = e0 - r0 * gz;
  const int ja = r0 % gy, ia = r0 / gy;  // a group stays in slab ia
  const int jb = ja + 1 < gy ? ja + 1 : 0;
  const unsigned fxm = ia > 0 ? word(e0 - sx) : 0u, fxp = ia < gx - 1 ? word(e0 + sx) : 0u;
  const unsigned fym = flags(e0 - sy), fyp = flags(e0 + sy);
  const uint8_t fkm = k0 > 0 ? val[e0 - 1] : 0;
  const uint8_t fkp = e0 + 4 < n ? val[e0 + 4] : 0;
  bool hxm[4], hxp[4], hym[4], hyp[4], hkm[4], hkp[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const unsigned bit = 0xFFu << (8 * q);
    const bool wrap = k0 + q >= gz;
    const int k = wrap ? k0 + q - gz : k0 + q, j = wrap ? jb : ja;
    hxm[q] = (fxm & bit) != 0;
    hxp[q] = (fxp & bit) != 0;
    hym[q] = j > 0 && (fym & bit) != 0;
    hyp[q] = j < gy - 1 && (fyp & bit) != 0;
    hkm[q] = k > 0 && (q > 0 ? (fc & (bit >> 8)) != 0 : fkm != 0);
    hkp[q] = k < gz - 1 && (q < 3 ? (fc & (bit << 8)) != 0 : fkp != 0);
  }
  // the four values at p .. p+3 where a face asks for one: all four loads at once when they lie
  // on the array, else one load per asking face
  auto values = [&](int p, const bool* h, T* v) {
    if (!(h[0] || h[1] || h[2] || h[3])) return;
    if (p >= 0 && p + 4 <= n) {
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = g[p + q];
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (h[q]) v[q] = g[p + q];
    }
  };
  T vxm[4] = {}, vxp[4] = {}, vym[4] = {}, vyp[4] = {}, vc[4] = {};
  values(e0 - sx, hxm, vxm);
  values(e0 + sx, hxp, vxp);
  values(e0 - sy, hym, vym);
  values(e0 + sy, hyp, vyp);
  if (fc) {  // the group's own values (z neighbours inside the group), always on the array
#pragma unroll
    for (int q = 0; q < 4; q++) vc[q] = g[e0 + q];
  }
  const T gkm = hkm[0] ? g[e0 - 1] : T(0), gkp = hkp[3] ? g[e0 + 4] : T(0);
  T vkm[4], vkp[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    vkm[q] = q > 0 ? vc[q - 1] : gkm;
    vkp[q] = q < 3 ? vc[q + 1] : gkp;
  }
  unsigned fo = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (fc & (0xFFu << (8 * q))) {
      fo |= 1u << (8 * q);
      continue;
    }
    T sum = T(0);
    int count = 0;
    if (hxm[q]) { sum += vxm[q]; count++; }
    if (hxp[q]) { sum += vxp[q]; count++; }
    if (hym[q]) { sum += vym[q]; count++; }
    if (hyp[q]) { sum += vyp[q]; count++; }
    if (hkm[q]) { sum += vkm[q]; count++; }
    if (hkp[q]) { sum += vkp[q]; count++; }
    if (count > 0) {
      gbuf[e0 + q] = sum / (T)count;
      fo |= 1u << (8 * q);
    } else {
      gbuf[e0 + q] = T(0);
    }
  }
  *reinterpret_cast<unsigned*>(vbuf + e0) = fo;
}

// applyForce (fluid-simulator.cc:13-18)
template <typename T>
__global__ __launch_bounds__(TPB) void k_force(int gx, int gy, int gz, T* __restrict__ V, const uint8_t* __restrict__ valid, T gdt) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= gx * gy * gz) return;
  int i, j, k;
  unlin3(id, gy, gz, i, j, k);
  if (valid[id] && j > 0 && j < gy - 1) V[id] = V[id] - gdt;
}

// Face solid weights of FvmSolver3D::buildSystem (project-solver.cc:77-127): faces between
// two air cells keep 0; otherwise 1 - fractionInside of the collider SDF at the 4 corners.
template <typename T>
__global__ __launch_bounds__(TPB) void k_face_weights(int axis, GridArgs<T> ga, int gx, int gy, int gz, const T* __restrict__ fs,
                               const T* __restrict__ col, T* __restrict__ W) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= gx * gy * gz) return;
  int i, j, k;
  unlin3(id, gy, gz, i, j, k);
  int c = axis == 0 ? i : (axis == 1 ? j : k);
  int d = axis == 0 ? gx : (axis == 1 ? gy : gz);
  if (c == 0 || c == d - 1) {
    W[id] = T(0);
    return;
  }
  int ny = ga.ny, nz = ga.nz;
  int im = i - (axis == 0), jm = j - (axis == 1), km = k - (axis == 2);
  if (fs[lin3(im, jm, km, ny, nz)] > T(0) && fs[lin3(i, j, k, ny, nz)] > T(0)) {
    W[id] = T(0);
    return;
  }
  T hh = ga.h[0];
  T p0 = coord1(i, axis == 0 ? T(0) : T(0.5), ga.h[0]);
  T p1 = coord1(j, axis == 1 ? T(0) : T(0.5), ga.h[1]);
  T p2 = coord1(k, axis == 2 ? T(0) : T(0.5), ga.h[2]);
  // corner offsets Vec3d(a,b,c) * 0.5 * h, in the argument order of each call
  T s0, s1, s2, s3;
  const T hp = T(1) * T(0.5) * hh, hm = T(-1) * T(0.5) * hh, z0 = T(0) * T(0.5) * hh;
  if (axis == 0) {  // bu(0,1,-1) bd(0,-1,-1) fd(0,-1,1) fu(0,1,1)
    s0 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + z0, p1 + hp, p2 + hm);
    s1 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + z0, p1 + hm, p2 + hm);
    s2 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + z0, p1 + hm, p2 + hp);
    s3 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + z0, p1 + hp, p2 + hp);
  } else if (axis == 1) {  // lb(-1,0,-1) rb(1,0,-1) rf(1,0,1) lf(-1,0,1)
    s0 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hm, p1 + z0, p2 + hm);
    s1 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hp, p1 + z0, p2 + hm);
    s2 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hp, p1 + z0, p2 + hp);
    s3 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hm, p1 + z0, p2 + hp);
  } else {  // ld(-1,-1,0) lu(-1,1,0) ru(1,1,0) rd(1,-1,0)
    s0 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hm, p1 + hm, p2 + z0);
    s1 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hm, p1 + hp, p2 + z0);
    s2 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hp, p1 + hp, p2 + z0);
    s3 = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0 + hp, p1 + hm, p2 + z0);
  }
  W[id] = T(1) - fraction_inside(s0, s1, s2, s3);
}

// Cell loop of FvmSolver3D::buildSystem (project-solver.cc:129-226): active marking and the
// 7-point ghost-fluid Poisson row. Neighbour order left, right, down, up, back, front.
template <typename T>
__global__ __launch_bounds__(TPB) void k_assemble(GridArgs<T> ga, T dt, const T* __restrict__ fs, const T* __restrict__ uW,
                           const T* __restrict__ vW, const T* __restrict__ wW, const T* __restrict__ U,
                           const T* __restrict__ V, const T* __restrict__ Wv, T* __restrict__ Adiag,
                           T* __restrict__ aL, T* __restrict__ aR, T* __restrict__ aU, T* __restrict__ aD,
                           T* __restrict__ aF, T* __restrict__ aB, T* __restrict__ rhs, uint8_t* __restrict__ active) {
  const int nx = ga.nx, ny = ga.ny, nz = ga.nz;
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nx * ny * nz) return;
  int i, j, k;
  unlin3(id, ny, nz, i, j, k);
  const int uL = lin3(i, j, k, ny, nz), uRi = lin3(i + 1, j, k, ny, nz);
  const int vD = lin3(i, j, k, ny + 1, nz), vUi = lin3(i, j + 1, k, ny + 1, nz);
  const int wB = lin3(i, j, k, ny, nz + 1), wFi = lin3(i, j, k + 1, ny, nz + 1);
  T ad = T(0), rh = T(0), l = T(0), r = T(0), u = T(0), d = T(0), f = T(0), b = T(0);
  bool act = false;
  if (!(uW[uL] == T(0) && uW[uRi] == T(0) && vW[vD] == T(0) && vW[vUi] == T(0) && wW[wB] == T(0) && wW[wFi] == T(0)) &&
      !(fs[id] > T(0))) {
    act = true;
    T sd = fs[id];
    T factor = dt / ga.h[0];
    const int sx = ny * nz, sy = nz;
    if (i > 0) {
      T fn = fs[id - sx];
      if (fn > T(0)) {
        T theta = smin(fn / (fn - sd), T(0.99));
        ad += uW[uL] * factor / (T(1) - theta);
      } else {
        ad += uW[uL] * factor;
        l -= uW[uL] * factor;
      }
      rh += uW[uL] * U[uL];
    }
    if (i < nx - 1) {
      T fn = fs[id + sx];
      if (fn > T(0)) {
        T theta = smax(sd / (sd - fn), T(0.01));
        ad += uW[uRi] * factor / theta;
      } else {
        ad += uW[uRi] * factor;
        r -= uW[uRi] * factor;
      }
      rh -= uW[uRi] * U[uRi];
    }
    if (j > 0) {
      T fn = fs[id - sy];
      if (fn > T(0)) {
        T theta = smin(fn / (fn - sd), T(0.99));
        ad += vW[vD] * factor / (T(1) - theta);
      } else {
        ad += vW[vD] * factor;
        d -= vW[vD] * factor;
      }
      rh += vW[vD] * V[vD];
    }
    if (j < nx - 1) {  // the reference bounds "up" by fluidSdf.width() (project-solver.cc:186)
      T fn = fs[id + sy];
      if (fn > T(0)) {
        T theta = smax(sd / (sd - fn), T(0.01));
        ad += vW[vUi] * factor / theta;
      } else {
        ad += vW[vUi] * factor;
        u -= vW[vUi] * factor;
      }
      rh -= vW[vUi] * V[vUi];
    }
    if (k > 0) {
      T fn = fs[id - 1];
      if (fn > T(0)) {
        T theta = smin(fn / (fn - sd), T(0.99));
        ad += wW[wB] * factor / (T(1) - theta);
      } else {
        ad += wW[wB] * factor;
        b -= wW[wB] * factor;
      }
      rh += wW[wB] * Wv[wB];
    }
    if (k < nz - 1) {
      T fn = fs[id + 1];
      if (fn > T(0)) {
        T theta = smax(sd / (sd - fn), T(0.01));
        ad += wW[wFi] * factor / theta;
      } else {
        ad += wW[wFi] * factor;
        f -= wW[wFi] * factor;
      }
      rh -= wW[wFi] * Wv[wFi];
    }
  }
  Adiag[id] = ad;
  aL[id] = l;
  aR[id] = r;
  aU[id] = u;
  aD[id] = d;
  aF[id] = f;
  aB[id] = b;
  rhs[id] = rh;
  active[id] = act ? 1 : 0;
}

// FvmSolver3D::project (project-solver.cc:348-421) for face grid `axis`.
template <typename T>
__global__ __launch_bounds__(TPB) void k_project(int axis, GridArgs<T> ga, int gx, int gy, int gz, T dt, const T* __restrict__ fs,
                          const T* __restrict__ pg, const T* __restrict__ Wt, T* __restrict__ G) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= gx * gy * gz) return;
  if (Wt[id] <= T(0)) return;
  int i, j, k;
  unlin3(id, gy, gz, i, j, k);
  int c = axis == 0 ? i : (axis == 1 ? j : k);
  int d = axis == 0 ? gx : (axis == 1 ? gy : gz);
  if (c == 0 || c == d - 1) {
    G[id] = T(0);
    return;
  }
  int ny = ga.ny, nz = ga.nz;
  int cm = lin3(i - (axis == 0), j - (axis == 1), k - (axis == 2), ny, nz);
  int cp = lin3(i, j, k, ny, nz);
  T hh = ga.h[0];
  T s0 = fs[cm], s1 = fs[cp];
  if (s0 >= T(0) && s1 >= T(0)) return;
  if (s0 < T(0) && s1 < T(0)) {
    G[id] = G[id] - (pg[cp] - pg[cm]) * dt / hh;
    return;
  }
  if (s0 < T(0)) {
    T theta = smax(s0 / (s0 - s1), T(0.01));
    G[id] = G[id] + pg[cm] * dt / hh / theta;
  } else {
    T theta = smin(s0 / (s0 - s1), T(0.99));
    G[id] = G[id] - pg[cp] * dt / hh / (T(1) - theta);
  }
}

// applyCollider (fluid-simulator.cc:76-111) for face grid `axis`.
template <typename T>
__global__ __launch_bounds__(TPB) void k_collider(int axis, GridArgs<T> ga, int gx, int gy, int gz, const T* __restrict__ col,
                           T* __restrict__ G) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= gx * gy * gz) return;
  int i, j, k;
  unlin3(id, gy, gz, i, j, k);
  int c = axis == 0 ? i : (axis == 1 ? j : k);
  int d = axis == 0 ? gx : (axis == 1 ? gy : gz);
  if (c == 0 || c == d - 1) {
    G[id] = T(0);
    return;
  }
  T p0 = coord1(i, axis == 0 ? T(0) : T(0.5), ga.h[0]);
  T p1 = coord1(j, axis == 1 ? T(0) : T(0.5), ga.h[1]);
  T p2 = coord1(k, axis == 2 ? T(0) : T(0.5), ga.h[2]);
  if (sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p0, p1, p2) < T(0)) {
    V3<T> n = gnormalize(sdf_grad(col, ga.nx, ga.ny, ga.nz, ga.h, p0, p1, p2));
    T nc = axis == 0 ? n.x : (axis == 1 ? n.y : n.z);
    G[id] = G[id] - G[id] * nc;
  }
}

// max |u| over a face grid (CFL, fluid-simulator.cc:184-203): fl(h/|u|) is monotone in |u|,
// so min over faces of h/|u| is h / max|u|.
template <typename T>
__global__ __launch_bounds__(TPB) void k_maxabs(int n, const T* __restrict__ g, unsigned long long* __restrict__ out) {
  __shared__ double sm[TPB];
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double a = fabs((double)g[i]);
    m = a > m ? a : m;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = sm[threadIdx.x + s] > sm[threadIdx.x] ? sm[threadIdx.x + s] : sm[threadIdx.x];
    __syncthreads();
  }
  // non-negative doubles order like their bit patterns
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(sm[0]));
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_to_double(int64_t n, const T* __restrict__ a, double* __restrict__ b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_from_double(int64_t n, const double* __restrict__ a, T* __restrict__ b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_split_aos(int64_t n, const double* __restrict__ aos, T* __restrict__ x, T* __restrict__ y,
                            T* __restrict__ z) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = (T)aos[3 * i];
  y[i] = (T)aos[3 * i + 1];
  z[i] = (T)aos[3 * i + 2];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_join_aos(int64_t n, const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
                           double* __restrict__ aos) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  aos[3 * i] = (double)x[i];
  aos[3 * i + 1] = (double)y[i];
  aos[3 * i + 2] = (double)z[i];
}

// ---------------------------------------------------------------------------------------------
// host side

template <typename T>
Grids3<T> grids_of(const fs_ctx* c, const DBuf* bufs) {
  Grids3<T> G;
  for (int a = 0; a < 3; a++) {
    G.g[a] = bufs[a].as<T>();
    for (int d = 0; d < 3; d++) G.d[a][d] = c->fdim[a][d];
  }
  return G;
}

int ensure_bins(fs_ctx* c, size_t nbins) {
  if (nbins <= c->bins_cap && c->counts.p) return FS_OK;
  fs_free(c->counts);
  fs_free(c->offs);
  fs_free(c->offs2);
  fs_free(c->scan_tmp);
  FS_TRY(fs_alloc(c, c->counts, (nbins + 1) * sizeof(int)));
  FS_TRY(fs_alloc(c, c->offs, (nbins + 1) * sizeof(int)));
  FS_TRY(fs_alloc(c, c->offs2, (nbins + 1) * sizeof(int)));
  const size_t tmp = (size_t)(fsscan::tiles_for((int64_t)nbins + 1) + 1) * sizeof(int);
  FS_TRY(fs_alloc(c, c->scan_tmp, tmp));
  c->scan_tmp_bytes = tmp;
  c->bins_cap = nbins;
  return FS_OK;
}

// keys must already be in c->keys; leaves offs[nbins+1] and perm.
int counting_sort(fs_ctx* c, int nbins, DBuf* offs_out = nullptr) {
  FS_TRY(ensure_bins(c, (size_t)nbins));
  DBuf& offs = offs_out ? *offs_out : c->offs;
  int n = (int)c->np;
  FS_HIP(c, hipMemsetAsync(c->counts.p, 0, (size_t)(nbins + 1) * sizeof(int), c->stream));
  if (n > 0) k_count<<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, c->keys.as<int>(), c->counts.as<int>(), c->slots.as<int>());
  FS_HIP(c, fsscan::exclusive_scan(c->stream, (int64_t)nbins + 1, c->counts.as<int>(), offs.as<int>(),
                                   c->scan_tmp.as<int>()));
  if (n > 0) {
    k_scatter<<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, c->keys.as<int>(), c->slots.as<int>(), offs.as<int>(),
                                                       c->perm.as<int>());
    k_sort_bins<<<fs_blocks(nbins, TPB), TPB, 0, c->stream>>>(nbins, offs.as<int>(), c->perm.as<int>());
  }
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

double smax_for(double r2) {
  double s = r2 * r2;
  while (std::sqrt(std::nextafter(s, INFINITY)) <= r2) s = std::nextafter(s, INFINITY);
  while (s > 0 && std::sqrt(s) > r2) s = std::nextafter(s, -INFINITY);
  return s;
}
float smax_for_f(float r2) {
  float s = r2 * r2;
  while (std::sqrt(std::nextafter(s, INFINITY)) <= r2) s = std::nextafter(s, INFINITY);
  while (s > 0 && std::sqrt(s) > r2) s = std::nextafter(s, -INFINITY);
  return s;
}

template <typename T>
int advect_t(fs_ctx* c, double dt) {
  int n = (int)c->np;
  if (n == 0) return FS_OK;
  GridArgs<T> ga = grid_args<T>(c);
  const int pf = fs_prof_begin(c, FS_KERNEL_ADVECT,
                               (double)n * 12 * sizeof(T) + (double)(c->n_face[0] + c->n_face[1] + c->n_face[2]) * sizeof(T));
  int* nonzero = reinterpret_cast<int*>(c->red.as<char>() + 16);
  FS_HIP(c, hipMemsetAsync(nonzero, 0, sizeof(int), c->stream));
  k_any_nonzero<T><<<2048, TPB, 0, c->stream>>>(c->n_face[0], c->n_face[1], c->n_face[2], c->g[0].as<T>(),
                                                  c->g[1].as<T>(), c->g[2].as<T>(), nonzero);
  k_advect<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, grids_of<T>(c, c->g), c->collider.as<T>(),
                                                        c->pos[0].as<T>(), c->pos[1].as<T>(), c->pos[2].as<T>(),
                                                        c->vel[0].as<T>(), c->vel[1].as<T>(), c->vel[2].as<T>(), (T)dt,
                                                        nonzero);
  fs_prof_end(c, pf);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int reconstruct_t(fs_ctx* c, double radius) {
  using UB = typename Bits<T>::U;
  GridArgs<T> ga = grid_args<T>(c);
  int n = (int)c->np;
  int nb = morton3D(c->nx - 1, c->ny - 1, c->nz - 1) + 1;
  // own keys and per-bin last particle (rkeys, rbinmax), not the P2G binning buffers
  if (!c->rbinmax.p) FS_TRY(fs_alloc(c, c->rbinmax, (size_t)nb * sizeof(int)));
  if (!c->recon_tmp.p) FS_TRY(fs_alloc(c, c->recon_tmp, (size_t)c->n_cell * sizeof(double)));
  const int dmax = std::max(c->nx, std::max(c->ny, c->nz));
  if (!c->rbins.p) FS_TRY(fs_alloc(c, c->rbins, (size_t)6 * dmax * sizeof(int2)));  // bounds, then intervals
  UB* dmin = c->recon_tmp.as<UB>();
  k_fill_bits<T><<<fs_blocks(c->n_cell, TPB), TPB, 0, c->stream>>>(c->n_cell, dmin);
  T rad = (T)radius;
  T r2 = T(2) * rad;
  T smax;
  if (sizeof(T) == 8) smax = (T)smax_for((double)r2);
  else smax = (T)smax_for_f((float)r2);
  const int pf = fs_prof_begin(c, FS_KERNEL_RECONSTRUCT, (double)n * 3 * sizeof(T) + (double)c->n_cell * (sizeof(T) + 1));
  if (n > 0) {
    FS_HIP(c, hipMemsetAsync(c->rbinmax.p, 0xFF, (size_t)nb * sizeof(int), c->stream));
    k_keys_morton<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, c->pos[0].as<T>(), c->pos[1].as<T>(),
                                                               c->pos[2].as<T>(), c->rkeys.as<int>());
    k_bin_last<<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, c->rkeys.as<int>(), c->rbinmax.as<int>());
    k_recon_bounds<T><<<fs_blocks(3 * dmax, 64), 64, 0, c->stream>>>(ga, r2, dmax, c->rbins.as<int2>());
    k_recon_bins<<<fs_blocks(3 * dmax, 64), 64, 0, c->stream>>>(c->nx, c->ny, c->nz, dmax, c->rbins.as<int2>(),
                                                                c->rbins.as<int2>() + 3 * dmax);
    k_recon_splat<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, c->pos[0].as<T>(), c->pos[1].as<T>(),
                                                               c->pos[2].as<T>(), c->rkeys.as<int>(),
                                                               c->rbinmax.as<int>(), r2, smax,
                                                               c->rbins.as<int2>() + 3 * dmax, dmax, dmin);
  }
  k_recon_final<T><<<fs_blocks(c->n_cell, TPB), TPB, 0, c->stream>>>(c->n_cell, dmin, rad, c->fsdf.as<T>(),
                                                                     c->sdfvalid.as<uint8_t>());
  fs_prof_end(c, pf);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int extrap_sdf_t(fs_ctx* c, int iters) {
  for (int it = 0; it < iters; it++) {
    if (c->nz % 4 == 0)
      k_extrap_sdf4<T><<<g3_grid(c->nx, c->ny, c->nz / 4), g3_block(), 0, c->stream>>>(
          c->nx, c->ny, c->nz, c->fsdf.as<T>(), c->fsdfbuf.as<T>(), c->sdfvalid.as<uint8_t>(),
          c->sdfvalidbuf.as<uint8_t>());
    else
      k_extrap_sdf<T><<<g3_grid(c->nx, c->ny, c->nz), g3_block(), 0, c->stream>>>(
          c->nx, c->ny, c->nz, c->fsdf.as<T>(), c->fsdfbuf.as<T>(), c->sdfvalid.as<uint8_t>(),
          c->sdfvalidbuf.as<uint8_t>());
    std::swap(c->fsdf, c->fsdfbuf);
    std::swap(c->sdfvalid, c->sdfvalidbuf);
  }
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int smooth_t(fs_ctx* c, int iters) {
  for (int it = 0; it < iters; it++) {
    if (c->nz % 4 == 0)
      k_smooth4<T><<<g3_grid(c->nx, c->ny, c->nz / 4), g3_block(), 0, c->stream>>>(c->nx, c->ny, c->nz,
                                                                                   c->fsdf.as<T>(), c->fsdfbuf.as<T>());
    else
      k_smooth<T><<<g3_grid(c->nx, c->ny, c->nz), g3_block(), 0, c->stream>>>(c->nx, c->ny, c->nz, c->fsdf.as<T>(),
                                                                              c->fsdfbuf.as<T>());
    std::swap(c->fsdf, c->fsdfbuf);
  }
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

// The side stream and its events, created on first use.
int ensure_side(fs_ctx* c) {
  if (!c->side) FS_HIP(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  if (!c->p2g_ev_ok) {
    for (int q = 0; q < 6; q++) FS_HIP(c, hipEventCreateWithFlags(&c->p2g_ev[q], hipEventDisableTiming));
    c->p2g_ev_ok = true;
  }
  return FS_OK;
}

// P2G of the three face grids. The gather of face grid a runs on the side stream while face grid
// a + 1 is binned on the main stream (two buffer sets, alternating); the main stream waits for the
// gathers before the stage ends. With per-launch event timing on, everything stays on the main
// stream, so that the timed intervals hold the whole stage.
// ---- APIC (affine particle-in-cell, Jiang et al. 2015) ----------------------------------------
// The reference declares AdvectionMethod::APIC without an implementation (setAdvector reaches
// ERROR, cpu/fluid-simulator.h:96-97). Here each particle carries, per velocity component a, the
// affine row c_a (the component's gradient near the particle); P2G transfers v_a + c_a.(x_n - x_p)
// to every node, and G2P sets v_a to the sampled velocity and c_a = D^-1 b over the same nodes, with
// b = sum w g (x_n - x_p) and D = sum w (x_n - x_p)(x_n - x_p)^T.

// P2G, one thread per node of face grid a: the particles of the bins that reach the node (nearest
// node b - 1 .. b + 2 around the node, so b in [i-2, i+1] per axis), bins in key order and
// particles in index order inside a bin - every node sums in (key, index) order, the order of
// FLIP's P2G. Writes the normalised velocity, the weight sum and the valid flag.
template <typename T>
__global__ __launch_bounds__(TPB) void k_p2g_apic(int a, GridArgs<T> ga, int gx, int gy, int gz,
                                                  const int* __restrict__ offs, const int* __restrict__ perm,
                                                  const T* __restrict__ px, const T* __restrict__ py,
                                                  const T* __restrict__ pz, const T* __restrict__ vc,
                                                  const T* __restrict__ c0, const T* __restrict__ c1,
                                                  const T* __restrict__ c2, T* __restrict__ G, T* __restrict__ Wt,
                                                  uint8_t* __restrict__ valid) {
  const int64_t id = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (id >= (int64_t)gx * gy * gz) return;
  const int k = (int)(id % gz), j = (int)((id / gz) % gy), i = (int)(id / ((int64_t)gz * gy));
  const T off[3] = {a == 0 ? T(0) : T(0.5), a == 1 ? T(0) : T(0.5), a == 2 ? T(0) : T(0.5)};
  const T pn[3] = {coord1(i, off[0], ga.h[0]), coord1(j, off[1], ga.h[1]), coord1(k, off[2], ga.h[2])};
  const int nbk[3] = {ga.nx + 1, ga.ny + 1, ga.nz + 1};  // bin index range of the nearest-node keys
  const int node[3] = {i, j, k};
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = max(node[d] - 2, 0);
    hi[d] = min(node[d] + 1, nbk[d] - 1);
  }
  T g = T(0), w_sum = T(0);
  for (int b0 = lo[0]; b0 <= hi[0]; b0++)
    for (int b1 = lo[1]; b1 <= hi[1]; b1++)
      for (int b2 = lo[2]; b2 <= hi[2]; b2++) {
        const int bin = lin3(b0, b1, b2, nbk[1], nbk[2]);
        for (int q = offs[bin]; q < offs[bin + 1]; q++) {
          const int pi = perm[q];
          const T d0 = px[pi] - pn[0], d1 = py[pi] - pn[1], d2 = pz[pi] - pn[2];  // p - x_n
          const T w = bspline(wdiv(fabs(d0), ga.h[0])) * bspline(wdiv(fabs(d1), ga.h[1])) *
                      bspline(wdiv(fabs(d2), ga.h[2]));
          // x_n - p = -(p - x_n), exactly
          const T val = vc[pi] + (c0[pi] * (-d0) + c1[pi] * (-d1) + c2[pi] * (-d2));
          g = g + val * w;
          w_sum = w_sum + w;
        }
      }
  const bool ok = w_sum > T(0);
  G[id] = ok ? g / w_sum : T(0);
  Wt[id] = w_sum;
  valid[id] = ok ? 1 : 0;
}

// G2P: per particle and component, the sampled velocity and the affine row (same nodes, weights
// and order as sampleVelocity, util.cc:40-85).
template <typename T>
__global__ __launch_bounds__(TPB) void k_g2p_apic(int n, GridArgs<T> ga, Grids3<T> G, const T* __restrict__ px,
                                                  const T* __restrict__ py, const T* __restrict__ pz,
                                                  T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                                                  T* __restrict__ cm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T p[3] = {px[i], py[i], pz[i]};
  const T hs[3] = {ga.h[0], ga.h[0], ga.h[0]};
  T* vout[3] = {vx, vy, vz};
#pragma unroll 1
  for (int a = 0; a < 3; a++) {
    const T off[3] = {a == 0 ? T(0) : T(0.5), a == 1 ? T(0) : T(0.5), a == 2 ? T(0) : T(0.5)};
    int b[3];
    T dn[3][4], wf[3][4];  // x_n - p per axis node, kernel factors
#pragma unroll
    for (int c = 0; c < 3; c++) {
      b[c] = nearest1(p[c], off[c], ga.h[c]) - 1;
#pragma unroll
      for (int o = 0; o < 4; o++) {
        const T xn = coord1(b[c] + o, off[c], ga.h[c]);
        dn[c][o] = xn - p[c];
        wf[c][o] = bspline(wdiv(fabs(p[c] - xn), hs[c]));
      }
    }
    const int gx = G.d[a][0], gy = G.d[a][1], gz = G.d[a][2];
    const T* __restrict__ g = G.g[a];
    T acc = T(0), ws = T(0), b0 = T(0), b1 = T(0), b2 = T(0);
    T D00 = T(0), D01 = T(0), D02 = T(0), D11 = T(0), D12 = T(0), D22 = T(0);
    for (int ii = 0; ii < 4; ii++) {
      const int x = b[0] + ii;
      if (x < 0 || x >= gx) continue;
      for (int jj = 0; jj < 4; jj++) {
        const int y = b[1] + jj;
        if (y < 0 || y >= gy) continue;
        for (int kk = 0; kk < 4; kk++) {
          const int z = b[2] + kk;
          if (z < 0 || z >= gz) continue;
          const T gn = g[lin3(x, y, z, gy, gz)];
          const T w = wf[0][ii] * wf[1][jj] * wf[2][kk];
          acc = acc + gn * w;
          ws = ws + w;
          const T dx = dn[0][ii], dy = dn[1][jj], dz = dn[2][kk];
          const T gw = gn * w;
          b0 = b0 + gw * dx;
          b1 = b1 + gw * dy;
          b2 = b2 + gw * dz;
          const T wx = w * dx, wy = w * dy, wz = w * dz;
          D00 = D00 + wx * dx;
          D01 = D01 + wx * dy;
          D02 = D02 + wx * dz;
          D11 = D11 + wy * dy;
          D12 = D12 + wy * dz;
          D22 = D22 + wz * dz;
        }
      }
    }
    const bool ok = ws > T(0);
    vout[a][i] = ok ? acc / ws : T(0);
    T r0 = T(0), r1 = T(0), r2 = T(0);
    if (ok) {
      const T i00 = D11 * D22 - D12 * D12, i01 = D02 * D12 - D01 * D22, i02 = D01 * D12 - D02 * D11;
      const T i11 = D00 * D22 - D02 * D02, i12 = D01 * D02 - D00 * D12, i22 = D00 * D11 - D01 * D01;
      const T det = D00 * i00 + D01 * i01 + D02 * i02;
      if (det != T(0)) {
        r0 = (i00 * b0 + i01 * b1 + i02 * b2) / det;
        r1 = (i01 * b0 + i11 * b1 + i12 * b2) / det;
        r2 = (i02 * b0 + i12 * b1 + i22 * b2) / det;
      }
    }
    cm[(int64_t)(3 * a) * n + i] = r0;
    cm[(int64_t)(3 * a + 1) * n + i] = r1;
    cm[(int64_t)(3 * a + 2) * n + i] = r2;
  }
}

template <typename T>
int p2g_apic_t(fs_ctx* c) {
  GridArgs<T> ga = grid_args<T>(c);
  const int n = (int)c->np;
  const int nb = (c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  FS_TRY(fs_ensure_apic(c));
  const T* cm = c->apicc.as<T>();
  for (int a = 0; a < 3; a++) {
    if (n > 0)
      k_keys_nearest<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, a, ga, c->pos[0].as<T>(), c->pos[1].as<T>(),
                                                                  c->pos[2].as<T>(), c->keys.as<int>());
    FS_TRY(counting_sort(c, nb, &c->offs));
    const int64_t nf = c->n_face[a];
    k_p2g_apic<T><<<fs_blocks(nf, TPB), TPB, 0, c->stream>>>(
        a, ga, c->fdim[a][0], c->fdim[a][1], c->fdim[a][2], c->offs.as<int>(), c->perm.as<int>(), c->pos[0].as<T>(),
        c->pos[1].as<T>(), c->pos[2].as<T>(), c->vel[a].as<T>(), cm + (int64_t)(3 * a) * n,
        cm + (int64_t)(3 * a + 1) * n, cm + (int64_t)(3 * a + 2) * n, c->g[a].as<T>(), c->gw[a].as<T>(),
        c->gvalid[a].as<uint8_t>());
    FS_HIP(c, hipGetLastError());
  }
  return FS_OK;
}

template <typename T>
int p2g_t(fs_ctx* c) {
  if (c->cfg.advection == FS_ADVECT_APIC) return p2g_apic_t<T>(c);
  GridArgs<T> ga = grid_args<T>(c);
  int n = (int)c->np;
  int nb = (c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  const bool overlap = !(c->cfg.profile_stages >= 2 && c->ktime_on);
  if (overlap) FS_TRY(ensure_side(c));
  if (!c->p2g_tiles.p) FS_TRY(fs_alloc(c, c->p2g_tiles, (size_t)P2G_MAX_TILES * 12));
  if (overlap && !c->p2g_tiles2.p) FS_TRY(fs_alloc(c, c->p2g_tiles2, (size_t)P2G_MAX_TILES * 12));
  for (int a = 0; a < 3; a++) {
    const int set = overlap ? (a & 1) : 0;
    DBuf& offs = set ? c->offs2 : c->offs;
    DBuf& W = set ? c->p2g_w2 : c->p2g_w;
    DBuf& sv = c->sorted[set];
    DBuf& tiles = set ? c->p2g_tiles2 : c->p2g_tiles;
    if (overlap && a >= 2) FS_HIP(c, hipStreamWaitEvent(c->stream, c->p2g_ev[a - 2], 0));  // set free again
    if (n > 0)
      k_keys_nearest<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, a, ga, c->pos[0].as<T>(), c->pos[1].as<T>(),
                                                                  c->pos[2].as<T>(), c->keys.as<int>());
    const int pf = fs_prof_begin(c, FS_KERNEL_P2G, (double)n * 4 * sizeof(T) + (double)c->n_face[a] * 2 * sizeof(T));
    FS_TRY(counting_sort(c, nb, &offs));
    if (n > 0)
      k_p2g_records<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, a, ga, c->perm.as<int>(), c->pos[0].as<T>(),
                                                                 c->pos[1].as<T>(), c->pos[2].as<T>(),
                                                                 c->vel[a].as<T>(), W.as<T>(), sv.as<T>());
    hipStream_t gs = c->stream;
    if (overlap) {
      FS_HIP(c, hipEventRecord(c->p2g_ev[3 + a], c->stream));
      FS_HIP(c, hipStreamWaitEvent(c->side, c->p2g_ev[3 + a], 0));
      gs = c->side;
    }
    int gx = c->fdim[a][0], gy = c->fdim[a][1], gz = c->fdim[a][2];
    launch_p2g_gather<T, true>(gs, ga, gx, gy, gz, n, offs.as<int>(), W.as<T>(), sv.as<T>(), c->g[a].as<T>(),
                               c->gw[a].as<T>(), c->gvalid[a].as<uint8_t>(), tiles);
    fs_prof_end(c, pf);
    FS_HIP(c, hipGetLastError());
    if (c->cfg.advection == FS_ADVECT_FLIP)
      FS_HIP(c, hipMemcpyAsync(c->glast[a].p, c->g[a].p, (size_t)c->n_face[a] * sizeof(T), hipMemcpyDeviceToDevice,
                               gs));
    if (overlap) FS_HIP(c, hipEventRecord(c->p2g_ev[a], c->side));
  }
  if (overlap) {
    FS_HIP(c, hipStreamWaitEvent(c->stream, c->p2g_ev[1], 0));
    FS_HIP(c, hipStreamWaitEvent(c->stream, c->p2g_ev[2], 0));
  }
  return FS_OK;
}

template <typename T>
int dirichlet_t(fs_ctx* c) {
  for (int a = 0; a < 3; a++)
    k_dirichlet<T><<<fs_blocks(c->n_face[a], TPB), TPB, 0, c->stream>>>(a, c->fdim[a][0], c->fdim[a][1], c->fdim[a][2],
                                                                        c->g[a].as<T>(), c->gvalid[a].as<uint8_t>());
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int extrap_vel_t(fs_ctx* c, int iters) {
  for (int a = 0; a < 3; a++)
    for (int it = 0; it < iters; it++) {
      const int gx = c->fdim[a][0], gy = c->fdim[a][1], gz = c->fdim[a][2];
      if (gz % 4 == 0)
        k_extrap_vel4<T><<<g3_grid(gx, gy, gz / 4), g3_block(), 0, c->stream>>>(
            gx, gy, gz, c->g[a].as<T>(), c->gbuf[a].as<T>(), c->gvalid[a].as<uint8_t>(), c->gvalidbuf[a].as<uint8_t>());
      else if (gz >= 4 && (gy * gz) % 4 == 0)
        k_extrap_vel4f<T><<<fs_blocks(c->n_face[a] / 4, TPB), TPB, 0, c->stream>>>(
            gx, gy, gz, c->g[a].as<T>(), c->gbuf[a].as<T>(), c->gvalid[a].as<uint8_t>(), c->gvalidbuf[a].as<uint8_t>());
      else
        k_extrap_vel<T><<<g3_grid(gx, gy, gz), g3_block(), 0, c->stream>>>(
            gx, gy, gz, c->g[a].as<T>(), c->gbuf[a].as<T>(), c->gvalid[a].as<uint8_t>(), c->gvalidbuf[a].as<uint8_t>());
      std::swap(c->g[a], c->gbuf[a]);
      std::swap(c->gvalid[a], c->gvalidbuf[a]);
    }
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int force_t(fs_ctx* c, double dt) {
  T gdt = (T)(c->cfg.gravity * dt);
  k_force<T><<<fs_blocks(c->n_face[1], TPB), TPB, 0, c->stream>>>(c->fdim[1][0], c->fdim[1][1], c->fdim[1][2],
                                                                  c->g[1].as<T>(), c->gvalid[1].as<uint8_t>(), gdt);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int build_t(fs_ctx* c, double dt) {
  GridArgs<T> ga = grid_args<T>(c);
  for (int a = 0; a < 3; a++)
    k_face_weights<T><<<fs_blocks(c->n_face[a], TPB), TPB, 0, c->stream>>>(
        a, ga, c->fdim[a][0], c->fdim[a][1], c->fdim[a][2], c->fsdf.as<T>(), c->collider.as<T>(),
        c->gweights[a].as<T>());
  // Aneighbour order Left, Right, Up, Down, Front, Back (project-solver.cc:7-14)
  k_assemble<T><<<fs_blocks(c->n_cell, TPB), TPB, 0, c->stream>>>(
      ga, (T)dt, c->fsdf.as<T>(), c->gweights[0].as<T>(), c->gweights[1].as<T>(), c->gweights[2].as<T>(),
      c->g[0].as<T>(), c->g[1].as<T>(), c->g[2].as<T>(), c->adiag.as<T>(), c->an[0].as<T>(), c->an[1].as<T>(),
      c->an[2].as<T>(), c->an[3].as<T>(), c->an[4].as<T>(), c->an[5].as<T>(), c->rhs.as<T>(),
      c->active.as<uint8_t>());
  FS_HIP(c, hipGetLastError());
  c->pcg.built = false;
  return FS_OK;
}

template <typename T>
int project_t(fs_ctx* c, double dt) {
  GridArgs<T> ga = grid_args<T>(c);
  for (int a = 0; a < 3; a++)
    k_project<T><<<fs_blocks(c->n_face[a], TPB), TPB, 0, c->stream>>>(a, ga, c->fdim[a][0], c->fdim[a][1],
                                                                      c->fdim[a][2], (T)dt, c->fsdf.as<T>(),
                                                                      c->pg.as<T>(), c->gweights[a].as<T>(),
                                                                      c->g[a].as<T>());
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int collider_t(fs_ctx* c) {
  GridArgs<T> ga = grid_args<T>(c);
  for (int a = 0; a < 3; a++)
    k_collider<T><<<fs_blocks(c->n_face[a], TPB), TPB, 0, c->stream>>>(a, ga, c->fdim[a][0], c->fdim[a][1],
                                                                       c->fdim[a][2], c->collider.as<T>(),
                                                                       c->g[a].as<T>());
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int g2p_t(fs_ctx* c) {
  int n = (int)c->np;
  if (n == 0) return FS_OK;
  GridArgs<T> ga = grid_args<T>(c);
  if (c->cfg.advection == FS_ADVECT_APIC) {
    FS_TRY(fs_ensure_apic(c));
    k_g2p_apic<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, grids_of<T>(c, c->g), c->pos[0].as<T>(),
                                                            c->pos[1].as<T>(), c->pos[2].as<T>(), c->vel[0].as<T>(),
                                                            c->vel[1].as<T>(), c->vel[2].as<T>(), c->apicc.as<T>());
    FS_HIP(c, hipGetLastError());
    return FS_OK;
  }
  const bool flip = c->cfg.advection == FS_ADVECT_FLIP;
  const int pf = fs_prof_begin(c, FS_KERNEL_G2P,
                               (double)n * 9 * sizeof(T) +
                                   (double)(c->n_face[0] + c->n_face[1] + c->n_face[2]) * (flip ? 2 : 1) * sizeof(T));
  if (flip)
    k_g2p<T, true><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(
        n, ga, grids_of<T>(c, c->g), grids_of<T>(c, c->glast), c->pos[0].as<T>(), c->pos[1].as<T>(),
        c->pos[2].as<T>(), c->vel[0].as<T>(), c->vel[1].as<T>(), c->vel[2].as<T>(), (T)c->cfg.flip_alpha);
  else
    k_g2p<T, false><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(
        n, ga, grids_of<T>(c, c->g), grids_of<T>(c, c->g), c->pos[0].as<T>(), c->pos[1].as<T>(), c->pos[2].as<T>(),
        c->vel[0].as<T>(), c->vel[1].as<T>(), c->vel[2].as<T>(), (T)c->cfg.flip_alpha);
  fs_prof_end(c, pf);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int cfl_t(fs_ctx* c, double* out) {
  FS_HIP(c, hipMemsetAsync(c->red.p, 0, sizeof(unsigned long long), c->stream));
  for (int a = 0; a < 3; a++)
    k_maxabs<T><<<256, TPB, 0, c->stream>>>(c->n_face[a], c->g[a].as<T>(), c->red.as<unsigned long long>());
  unsigned long long bits = 0;
  FS_HIP(c, hipMemcpyAsync(&bits, c->red.p, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  double m;
  std::memcpy(&m, &bits, sizeof(m));
  double hh = c->h[0];
  double cfl = hh / 1e-6;
  if (m > 0.0) {
    double v = hh / m;
    cfl = v < cfl ? v : cfl;
  }
  *out = c->cfg.cfl_factor * (1e-3 < cfl ? cfl : 1e-3);
  return FS_OK;
}



// ---- reverse mode: the fluid-surface chain ---------------------------------------------------
template <typename T>
__global__ __launch_bounds__(TPB) void k_accumulate(int64_t n, const T* __restrict__ a, T* __restrict__ b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = b[i] + a[i];
}

__device__ inline int n_count(int i, int j, int k, int nx, int ny, int nz) {
  return (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
}
__device__ inline int v_count(const uint8_t* V, int id, int i, int j, int k, int nx, int ny, int nz) {
  const int sx = ny * nz, sy = nz;
  return (i > 0 && V[id - sx]) + (i < nx - 1 && V[id + sx]) + (j > 0 && V[id - sy]) + (j < ny - 1 && V[id + sy]) +
         (k > 0 && V[id - 1]) + (k < nz - 1 && V[id + 1]);
}

// One smoothing iteration reversed: out(c) = cur(c), or the mean of c's neighbours when c took
// the average. curbar(n) = [n kept itself] outbar(n) + sum over neighbours c that averaged of
// outbar(c) / count(c).
template <typename T>
__global__ __launch_bounds__(TPB) void k_smooth_adj(int nx, int ny, int nz, const uint8_t* __restrict__ F, const T* __restrict__ ob,
                             T* __restrict__ cb) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nx * ny * nz) return;
  int i, j, k;
  unlin3(id, ny, nz, i, j, k);
  const int sx = ny * nz, sy = nz;
  T g = F[id] ? T(0) : ob[id];
  if (i > 0 && F[id - sx]) g += ob[id - sx] / (T)n_count(i - 1, j, k, nx, ny, nz);
  if (i < nx - 1 && F[id + sx]) g += ob[id + sx] / (T)n_count(i + 1, j, k, nx, ny, nz);
  if (j > 0 && F[id - sy]) g += ob[id - sy] / (T)n_count(i, j - 1, k, nx, ny, nz);
  if (j < ny - 1 && F[id + sy]) g += ob[id + sy] / (T)n_count(i, j + 1, k, nx, ny, nz);
  if (k > 0 && F[id - 1]) g += ob[id - 1] / (T)n_count(i, j, k - 1, nx, ny, nz);
  if (k < nz - 1 && F[id + 1]) g += ob[id + 1] / (T)n_count(i, j, k + 1, nx, ny, nz);
  cb[id] = g;
}

// One extrapolateFluidSdf iteration reversed. Forward with mask V: buf(c) = mean of the valid
// neighbours' cur for an invalid c that has some, else buf(c) is left as it was (valid and far
// cells); then (cur, buf) -> (buf, cur). From the adjoints (A, B) of the new pair:
//   curbar(n) = B(n) + [n valid] sum over invalid neighbours c of A(c) / count(c)
//   bufbar(c) = [c valid or count(c) = 0] A(c)
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_sdf_adj(int nx, int ny, int nz, const uint8_t* __restrict__ V, const T* __restrict__ A,
                                 const T* __restrict__ B, T* __restrict__ cb, T* __restrict__ bb) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nx * ny * nz) return;
  int i, j, k;
  unlin3(id, ny, nz, i, j, k);
  const int sx = ny * nz, sy = nz;
  T g = B[id];
  if (V[id]) {
    if (i > 0 && !V[id - sx]) g += A[id - sx] / (T)v_count(V, id - sx, i - 1, j, k, nx, ny, nz);
    if (i < nx - 1 && !V[id + sx]) g += A[id + sx] / (T)v_count(V, id + sx, i + 1, j, k, nx, ny, nz);
    if (j > 0 && !V[id - sy]) g += A[id - sy] / (T)v_count(V, id - sy, i, j - 1, k, nx, ny, nz);
    if (j < ny - 1 && !V[id + sy]) g += A[id + sy] / (T)v_count(V, id + sy, i, j + 1, k, nx, ny, nz);
    if (k > 0 && !V[id - 1]) g += A[id - 1] / (T)v_count(V, id - 1, i, j, k - 1, nx, ny, nz);
    if (k < nz - 1 && !V[id + 1]) g += A[id + 1] / (T)v_count(V, id + 1, i, j, k + 1, nx, ny, nz);
  }
  cb[id] = g;
  bb[id] = (V[id] || v_count(V, id, i, j, k, nx, ny, nz) == 0) ? A[id] : T(0);
}

// Reverse of the reconstruction: a valid cell's value is |x_c - x_p| - r for its nearest kept
// particle p (the one whose squared distance equals the cell's minimum), so p receives
// abar(c) (x_p - x_c) / |x_c - x_p|. Same candidate cells and inclusion tests as k_recon_splat.
template <typename T>
__global__ __launch_bounds__(TPB) void k_recon_adj(int n, GridArgs<T> ga, const T* __restrict__ px, const T* __restrict__ py,
                            const T* __restrict__ pz, const int* __restrict__ keys, const int* __restrict__ binmax,
                            T r2, T s_max, const typename Bits<T>::U* __restrict__ dmin, const T* __restrict__ abar,
                            T* __restrict__ gx, T* __restrict__ gy, T* __restrict__ gz) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  if (binmax[keys[id]] == id) return;
  const T p[3] = {px[id], py[id], pz[id]};
  const int dims[3] = {ga.nx, ga.ny, ga.nz};
  int bin[3], lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int bb = (int)floor(p[d] / ga.h[d]);
    bb = bb < 0 ? 0 : bb;
    bin[d] = bb >= dims[d] ? dims[d] - 1 : bb;
    lo[d] = max(0, (int)floor((p[d] - r2) / ga.h[d] - T(0.5)) - 1);
    hi[d] = min(dims[d] - 1, (int)floor((p[d] + r2) / ga.h[d] - T(0.5)) + 1);
  }
  auto in_range = [&](int d, T xc) {
    int mn = max(0, (int)floor((xc - r2) / ga.h[d]));
    int mx = min(dims[d] - 1, (int)ceil((xc + r2) / ga.h[d]));
    return bin[d] >= mn && bin[d] <= mx;
  };
  T g[3] = {T(0), T(0), T(0)};
  for (int i = lo[0]; i <= hi[0]; i++) {
    const T xc = coord1(i, T(0.5), ga.h[0]);
    const T dx = xc - p[0];
    const T dx2 = dx * dx;
    if (dx2 > s_max || !in_range(0, xc)) continue;
    for (int j = lo[1]; j <= hi[1]; j++) {
      const T yc = coord1(j, T(0.5), ga.h[1]);
      const T dy = yc - p[1];
      const T dxy = dx2 + dy * dy;
      if (dxy > s_max || !in_range(1, yc)) continue;
      for (int k = lo[2]; k <= hi[2]; k++) {
        const T zc = coord1(k, T(0.5), ga.h[2]);
        const T dz = zc - p[2];
        const T d2 = dxy + dz * dz;
        if (d2 > s_max || !in_range(2, zc)) continue;
        const int cell = lin3(i, j, k, ga.ny, ga.nz);
        if (Bits<T>::of(d2) != dmin[cell] || !(d2 > T(0))) continue;
        const T s = abar[cell] / sqrt(d2);
        g[0] -= s * dx;
        g[1] -= s * dy;
        g[2] -= s * dz;
      }
    }
  }
  gx[id] += g[0];
  gy[id] += g[1];
  gz[id] += g[2];
}

int ensure_cells(fs_ctx* c, DBuf& b, size_t bytes) {
  if (b.bytes != bytes) {
    fs_free(b);
    FS_TRY(fs_alloc(c, b, bytes));
  }
  return FS_OK;
}

template <typename T>
int adj_sdf_t(fs_ctx* c, AdjStep& t, DBuf& phibar, DBuf& psibar, DBuf* xbar) {
  using UB = typename Bits<T>::U;
  AdjTape& tp = c->tape;
  const int nc = c->n_cell, n = (int)c->np;
  const int ie = c->cfg.sdf_extrapolate_iters, is = c->cfg.sdf_smooth_iters;
  if (is < 1) {
    c->set_error("the reverse pass needs at least one SDF smoothing iteration");
    return FS_UNSUPPORTED;
  }
  const size_t cb = (size_t)nc * sizeof(T);
  for (DBuf* b : {&tp.sA, &tp.sB, &tp.sAb, &tp.sBb, &tp.sAb2, &tp.sBb2}) FS_TRY(ensure_cells(c, *b, cb));
  while ((int)tp.sV.size() < ie + 1) {
    tp.sV.emplace_back();
    FS_TRY(fs_alloc(c, tp.sV.back(), (size_t)nc));
  }
  while ((int)tp.sF.size() < is) {
    tp.sF.emplace_back();
    FS_TRY(fs_alloc(c, tp.sF.back(), (size_t)nc));
  }
  // forward: reconstruction of the recorded positions (values, mask, per-cell minima)
  GridArgs<T> ga = grid_args<T>(c);
  const int nb = morton3D(c->nx - 1, c->ny - 1, c->nz - 1) + 1;
  FS_TRY(ensure_bins(c, (size_t)nb));
  if (!c->recon_tmp.p) FS_TRY(fs_alloc(c, c->recon_tmp, (size_t)nc * sizeof(double)));
  const int dmax = std::max(c->nx, std::max(c->ny, c->nz));
  if (!c->rbins.p) FS_TRY(fs_alloc(c, c->rbins, (size_t)6 * dmax * sizeof(int2)));
  UB* dmin = c->recon_tmp.as<UB>();
  const T rad = (T)(c->cfg.recon_radius_factor * c->h[0]), r2 = T(2) * rad;
  const T smx = sizeof(T) == 8 ? (T)smax_for((double)r2) : (T)smax_for_f((float)r2);
  k_fill_bits<T><<<fs_blocks(nc, TPB), TPB, 0, c->stream>>>(nc, dmin);
  if (n > 0) {
    FS_HIP(c, hipMemsetAsync(c->counts.p, 0xFF, (size_t)nb * sizeof(int), c->stream));
    k_keys_morton<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, t.x1[0].as<T>(), t.x1[1].as<T>(),
                                                               t.x1[2].as<T>(), c->keys.as<int>());
    k_bin_last<<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, c->keys.as<int>(), c->counts.as<int>());
    k_recon_bounds<T><<<fs_blocks(3 * dmax, 64), 64, 0, c->stream>>>(ga, r2, dmax, c->rbins.as<int2>());
    k_recon_bins<<<fs_blocks(3 * dmax, 64), 64, 0, c->stream>>>(c->nx, c->ny, c->nz, dmax, c->rbins.as<int2>(),
                                                                c->rbins.as<int2>() + 3 * dmax);
    k_recon_splat<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, t.x1[0].as<T>(), t.x1[1].as<T>(),
                                                               t.x1[2].as<T>(), c->keys.as<int>(),
                                                               c->counts.as<int>(), r2, smx,
                                                               c->rbins.as<int2>() + 3 * dmax, dmax, dmin);
  }
  k_recon_final<T><<<fs_blocks(nc, TPB), TPB, 0, c->stream>>>(nc, dmin, rad, tp.sA.as<T>(), tp.sV[0].as<uint8_t>());
  FS_HIP(c, hipMemcpyAsync(tp.sB.p, t.sdfbuf0.p, cb, hipMemcpyDeviceToDevice, c->stream));
  // forward: extrapolation (values and masks), then smoothing with its branch flags
  DBuf *cur = &tp.sA, *buf = &tp.sB;
  for (int it = 0; it < ie; it++) {
    if (c->nz % 4 == 0)
      k_extrap_sdf4<T><<<g3_grid(c->nx, c->ny, c->nz / 4), g3_block(), 0, c->stream>>>(
          c->nx, c->ny, c->nz, cur->as<T>(), buf->as<T>(), tp.sV[it].as<uint8_t>(), tp.sV[it + 1].as<uint8_t>());
    else
      k_extrap_sdf<T><<<g3_grid(c->nx, c->ny, c->nz), g3_block(), 0, c->stream>>>(
          c->nx, c->ny, c->nz, cur->as<T>(), buf->as<T>(), tp.sV[it].as<uint8_t>(), tp.sV[it + 1].as<uint8_t>());
    std::swap(cur, buf);
  }
  for (int it = 0; it < is; it++) {
    if (c->nz % 4 == 0)
      k_smooth4<T, true><<<g3_grid(c->nx, c->ny, c->nz / 4), g3_block(), 0, c->stream>>>(
          c->nx, c->ny, c->nz, cur->as<T>(), buf->as<T>(), tp.sF[it].as<uint8_t>());
    else
      k_smooth<T, true><<<g3_grid(c->nx, c->ny, c->nz), g3_block(), 0, c->stream>>>(
          c->nx, c->ny, c->nz, cur->as<T>(), buf->as<T>(), tp.sF[it].as<uint8_t>());
    std::swap(cur, buf);
  }
  // reverse: smoothing (the buffer the step leaves behind is the last smoothing input)
  DBuf *ob = &phibar, *ib = &tp.sAb;
  for (int it = is - 1; it >= 0; it--) {
    k_smooth_adj<T><<<fs_blocks(nc, TPB), TPB, 0, c->stream>>>(c->nx, c->ny, c->nz, tp.sF[it].as<uint8_t>(),
                                                               ob->as<T>(), ib->as<T>());
    if (it == is - 1) {
      k_accumulate<T><<<fs_blocks(nc, TPB), TPB, 0, c->stream>>>(nc, psibar.as<T>(), ib->as<T>());
      ob = &tp.sAb2;
    }
    std::swap(ob, ib);
  }
  // reverse: extrapolation, from (adjoint of the smoothing input, 0 for the dead buffer)
  DBuf *A = ob, *B = &tp.sBb, *A2 = ib == &phibar ? &tp.sAb2 : ib, *B2 = &tp.sBb2;
  FS_HIP(c, hipMemsetAsync(B->p, 0, cb, c->stream));
  for (int it = ie - 1; it >= 0; it--) {
    k_extrap_sdf_adj<T><<<fs_blocks(nc, TPB), TPB, 0, c->stream>>>(c->nx, c->ny, c->nz, tp.sV[it].as<uint8_t>(),
                                                                   A->as<T>(), B->as<T>(), A2->as<T>(), B2->as<T>());
    std::swap(A, A2);
    std::swap(B, B2);
  }
  // reconstruction reversed; the start buffer's adjoint is carried to the previous step
  if (n > 0)
    k_recon_adj<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, t.x1[0].as<T>(), t.x1[1].as<T>(),
                                                             t.x1[2].as<T>(), c->keys.as<int>(), c->counts.as<int>(),
                                                             r2, smx, dmin, A->as<T>(), xbar[0].as<T>(),
                                                             xbar[1].as<T>(), xbar[2].as<T>());
  FS_HIP(c, hipMemcpyAsync(psibar.p, B->p, cb, hipMemcpyDeviceToDevice, c->stream));
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

// ---- reverse mode: particle side ------------------------------------------------------------
// Kernel factor of one stencil node and its derivative along the axis:
// w_c = N(|p_c - n_c| / hs_c), dw_c/dp_c = N'(r_c) sign(p_c - n_c) / hs_c.
template <typename T>
struct StencilD {
  int b[3];
  T w[3][4], d[3][4];
};

template <typename T>
__device__ inline void make_stencil_d(const T p[3], int axis, const T h[3], const T hs[3], StencilD<T>& st) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T off = (c == axis) ? T(0) : T(0.5);
    const int nb = nearest1(p[c], off, h[c]);
    st.b[c] = nb - 1;
#pragma unroll
    for (int o = 0; o < 4; o++) {
      const T dl = p[c] - coord1(nb - 1 + o, off, h[c]);
      const T r = wdiv(fabs(dl), hs[c]);
      st.w[c][o] = bspline(r);
      st.d[c][o] = wdiv(bspline_d(r) * (dl < T(0) ? T(-1) : T(1)), hs[c]);
    }
  }
}

// Sums over the in-grid nodes of one stencil: w, w*A, w*B and the position gradients of the
// same sums (node weight (wx*wy)*wz as in sample_grid).
template <typename T, bool TWO>
struct NodeSums {
  T s, a, b, ds[3], da[3], db[3];
};

template <typename T, bool TWO>
__device__ inline NodeSums<T, TWO> node_sums(const StencilD<T>& st, const T* __restrict__ A, const T* __restrict__ Bv,
                                             int gx, int gy, int gz) {
  NodeSums<T, TWO> r{};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int x = st.b[0] + i;
    if (x < 0 || x >= gx) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int y = st.b[1] + j;
      if (y < 0 || y >= gy) continue;
      const int base = gz * (y + gy * x);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int z = st.b[2] + k;
        if (z < 0 || z >= gz) continue;
        const T wx = st.w[0][i], wy = st.w[1][j], wz = st.w[2][k];
        const T w = wx * wy * wz;
        const T dw[3] = {st.d[0][i] * wy * wz, wx * st.d[1][j] * wz, wx * wy * st.d[2][k]};
        const T av = A[base + z];
        const T bv = TWO ? Bv[base + z] : T(0);
        r.s += w;
        r.a += w * av;
        if (TWO) r.b += w * bv;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          r.ds[d] += dw[d];
          r.da[d] += dw[d] * av;
          if (TWO) r.db[d] += dw[d] * bv;
        }
      }
    }
  }
  return r;
}

template <typename T>
struct PArr {
  T* v[3];
};

// Reverse of solveG2P (advect-solver.cc:7-17, 157-169). Per component a (its own face grid),
// pic = sum w g / S and old = sum w g_last / S with S = sum w; the new velocity is
// alpha*v + pic - alpha*old (FLIP) or pic (PIC). Writes the adjoint of the particle's own
// velocity (alpha * vbar), the transfer coefficients vbar / S of the transposed gather, and
// adds the position derivative to xbar.
template <typename T, bool FLIP>
__global__ __launch_bounds__(TPB) void k_g2p_adj(int n, GridArgs<T> ga, Grids3<T> G, Grids3<T> L, PArr<T> pos, PArr<T> vbar, PArr<T> xbar,
                          PArr<T> vbar1, PArr<T> coef, T alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T p[3] = {pos.v[0][i], pos.v[1][i], pos.v[2][i]};
  const T hs[3] = {ga.h[0], ga.h[0], ga.h[0]};  // sampleVelocity's scalar spacing (util.cc:47)
  T gp[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    StencilD<T> st;
    make_stencil_d(p, a, ga.h, hs, st);
    const NodeSums<T, FLIP> r = node_sums<T, FLIP>(st, G.g[a], L.g[a], G.d[a][0], G.d[a][1], G.d[a][2]);
    const T vb = vbar.v[a][i];
    T cf = T(0);
    if (r.s > T(0)) {
      const T pic = r.a / r.s, old = r.b / r.s;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        T g = r.da[d] - pic * r.ds[d];
        if (FLIP) g = g - alpha * (r.db[d] - old * r.ds[d]);
        gp[d] += vb * g / r.s;
      }
      cf = vb / r.s;
    }
    coef.v[a][i] = cf;
    vbar1.v[a][i] = FLIP ? alpha * vb : T(0);
  }
#pragma unroll
  for (int d = 0; d < 3; d++) xbar.v[d][i] += gp[d];
}

// Reverse of solveP2G (advect-solver.cc:188-271): g = sum_p w v_p / W where W = sum_p w > 0.
// With H = gbar / W (0 where W = 0) and HG = H * g per node:
// vbar_a += sum w H, xbar += v_a * sum dw H - sum dw HG.
template <typename T>
__global__ __launch_bounds__(TPB) void k_p2g_adj(int n, GridArgs<T> ga, Grids3<T> H, Grids3<T> HG, PArr<T> pos, PArr<T> vel, PArr<T> vbar,
                          PArr<T> xbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T p[3] = {pos.v[0][i], pos.v[1][i], pos.v[2][i]};
  T gp[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    StencilD<T> st;
    make_stencil_d(p, a, ga.h, ga.h, st);  // the transfer divides by each axis' spacing
    const NodeSums<T, true> r = node_sums<T, true>(st, H.g[a], HG.g[a], H.d[a][0], H.d[a][1], H.d[a][2]);
    vbar.v[a][i] += r.a;
    const T va = vel.v[a][i];
#pragma unroll
    for (int d = 0; d < 3; d++) gp[d] += va * r.da[d] - r.db[d];
  }
#pragma unroll
  for (int d = 0; d < 3; d++) xbar.v[d][i] += gp[d];
}

// Reverse of advect + handleCollision (advect-solver.cc:100-137, 170-187) on a cleared grid,
// where k2 = k3 = 0 and p = x + (2/9) dt v. The collider push-out p -= d(p) n is
// differentiated with its normal n held fixed (dd/dp taken as SDF::grad), the velocity
// projection v -= min(v.n, 0) n likewise; a clamped coordinate and a velocity component the
// wall zeroes carry no derivative. In: xbar, vbar of the outputs; out: of the inputs.
template <typename T>
__global__ __launch_bounds__(TPB) void k_advect_adj(int n, GridArgs<T> ga, const T* __restrict__ col, PArr<T> x0, PArr<T> v0, PArr<T> xbar,
                             PArr<T> vbar, T dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T a1 = T(2.0 / 9.0) * dt, a2 = T(1.0 / 3.0) * dt, a3 = T(4.0 / 9.0) * dt;
  V3<T> v = {v0.v[0][i], v0.v[1][i], v0.v[2][i]};
  V3<T> p = {x0.v[0][i] + a1 * v.x + a2 * T(0) + a3 * T(0), x0.v[1][i] + a1 * v.y + a2 * T(0) + a3 * T(0),
             x0.v[2][i] + a1 * v.z + a2 * T(0) + a3 * T(0)};
  const T d = sdf_eval(col, ga.nx, ga.ny, ga.nz, ga.h, p.x, p.y, p.z);
  const bool inside = d < T(0);
  V3<T> g = {T(0), T(0), T(0)}, nr = {T(0), T(0), T(0)};
  bool vproj = false;
  if (inside) {
    g = sdf_grad(col, ga.nx, ga.ny, ga.nz, ga.h, p.x, p.y, p.z);
    nr = gnormalize(g);
    p.x = p.x - d * nr.x;
    p.y = p.y - d * nr.y;
    p.z = p.z - d * nr.z;
    const T dv = v.x * nr.x + v.y * nr.y + v.z * nr.z;
    vproj = dv < T(0);
    const T m = smin(dv, T(0));
    v.x = v.x - m * nr.x;
    v.y = v.y - m * nr.y;
    v.z = v.z - m * nr.z;
  }
  T pb[3] = {xbar.v[0][i], xbar.v[1][i], xbar.v[2][i]};
  T vb[3] = {vbar.v[0][i], vbar.v[1][i], vbar.v[2][i]};
  const T pc[3] = {p.x, p.y, p.z}, vc[3] = {v.x, v.y, v.z};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const bool lo = pc[c] < T(0), hi = pc[c] > ga.ext[c];
    if (lo || hi) pb[c] = T(0);
    if ((lo && vc[c] < T(0)) || (hi && vc[c] > T(0))) vb[c] = T(0);
  }
  if (inside) {
    if (vproj) {
      const T t = nr.x * vb[0] + nr.y * vb[1] + nr.z * vb[2];
      vb[0] -= nr.x * t;
      vb[1] -= nr.y * t;
      vb[2] -= nr.z * t;
    }
    const T t = nr.x * pb[0] + nr.y * pb[1] + nr.z * pb[2];
    pb[0] -= g.x * t;
    pb[1] -= g.y * t;
    pb[2] -= g.z * t;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    xbar.v[c][i] = pb[c];
    vbar.v[c][i] = vb[c] + a1 * pb[c];
  }
}

template <typename T>
PArr<T> parr(DBuf* b) {
  return {{b[0].as<T>(), b[1].as<T>(), b[2].as<T>()}};
}
template <typename T>
Grids3<T> grids3(const fs_ctx* c, DBuf* b) {
  Grids3<T> G;
  for (int a = 0; a < 3; a++) {
    G.g[a] = b[a].as<T>();
    for (int d = 0; d < 3; d++) G.d[a][d] = c->fdim[a][d];
  }
  return G;
}

// out_a = sum_p w_pa(node) coef_a[p] for every node of face grid a (the transposed
// transfer), binned like P2G so the sums are deterministic.
template <typename T>
int p2g_sum_t(fs_ctx* c, DBuf* pos, DBuf* coef, DBuf* out) {
  GridArgs<T> ga = grid_args<T>(c);
  const int n = (int)c->np;
  const int nb = (c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  for (int a = 0; a < 3; a++) {
    if (n > 0)
      k_keys_nearest<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, a, ga, pos[0].as<T>(), pos[1].as<T>(),
                                                                  pos[2].as<T>(), c->keys.as<int>());
    FS_TRY(counting_sort(c, nb));
    if (n > 0)
      k_p2g_records<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, a, ga, c->perm.as<int>(), pos[0].as<T>(),
                                                                 pos[1].as<T>(), pos[2].as<T>(), coef[a].as<T>(),
                                                                 c->p2g_w.as<T>(), c->sorted[0].as<T>());
    const int gx = c->fdim[a][0], gy = c->fdim[a][1], gz = c->fdim[a][2];
    if (!c->p2g_tiles.p) FS_TRY(fs_alloc(c, c->p2g_tiles, (size_t)P2G_MAX_TILES * 12));
    launch_p2g_gather<T, false>(c->stream, ga, gx, gy, gz, n, c->offs.as<int>(), c->p2g_w.as<T>(),
                                c->sorted[0].as<T>(), out[a].as<T>(), nullptr, nullptr, c->p2g_tiles);
    FS_HIP(c, hipGetLastError());
  }
  return FS_OK;
}

template <typename T>
int adj_g2p_t(fs_ctx* c, DBuf* x1, DBuf* g5, DBuf* g0, DBuf* vbar, DBuf* xbar, DBuf* vbar1, DBuf* coef) {
  const int n = (int)c->np;
  if (n == 0) return FS_OK;
  GridArgs<T> ga = grid_args<T>(c);
  const bool flip = c->cfg.advection == FS_ADVECT_FLIP;
  const T alpha = (T)c->cfg.flip_alpha;
  if (flip)
    k_g2p_adj<T, true><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, grids3<T>(c, g5), grids3<T>(c, g0), parr<T>(x1),
                                                                 parr<T>(vbar), parr<T>(xbar), parr<T>(vbar1),
                                                                 parr<T>(coef), alpha);
  else
    k_g2p_adj<T, false><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, ga, grids3<T>(c, g5), grids3<T>(c, g5),
                                                                  parr<T>(x1), parr<T>(vbar), parr<T>(xbar),
                                                                  parr<T>(vbar1), parr<T>(coef), alpha);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int adj_p2g_t(fs_ctx* c, DBuf* x1, DBuf* v1, DBuf* H, DBuf* HG, DBuf* vbar, DBuf* xbar) {
  const int n = (int)c->np;
  if (n == 0) return FS_OK;
  k_p2g_adj<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, grid_args<T>(c), grids3<T>(c, H), grids3<T>(c, HG),
                                                         parr<T>(x1), parr<T>(v1), parr<T>(vbar), parr<T>(xbar));
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename T>
int adj_advect_t(fs_ctx* c, DBuf* x0, DBuf* v0, DBuf* xbar, DBuf* vbar, double dt) {
  const int n = (int)c->np;
  if (n == 0) return FS_OK;
  k_advect_adj<T><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, grid_args<T>(c), c->collider.as<T>(), parr<T>(x0),
                                                            parr<T>(v0), parr<T>(xbar), parr<T>(vbar), (T)dt);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
#define FS_DISPATCH(c, fn, ...) ((c)->mixed ? fn<float>(c, ##__VA_ARGS__) : fn<double>(c, ##__VA_ARGS__))

int fs_stage_clear(fs_ctx* c) {
  // clear() (fluid-simulator.cc:20-37)
  for (int a = 0; a < 3; a++) {
    FS_TRY(fs_zero(c, c->gvalid[a]));
    FS_TRY(fs_zero(c, c->gvalidbuf[a]));
    FS_TRY(fs_zero(c, c->g[a]));
    FS_TRY(fs_zero(c, c->gbuf[a]));
    FS_TRY(fs_zero(c, c->gw[a]));
  }
  FS_TRY(fs_zero(c, c->pg));
  return FS_OK;
}
int fs_stage_advect(fs_ctx* c, double dt) { return FS_DISPATCH(c, advect_t, dt); }
int fs_stage_reconstruct(fs_ctx* c, double radius) { return FS_DISPATCH(c, reconstruct_t, radius); }
int fs_stage_extrapolate_sdf(fs_ctx* c, int iters) { return FS_DISPATCH(c, extrap_sdf_t, iters); }
int fs_stage_smooth(fs_ctx* c, int iters) { return FS_DISPATCH(c, smooth_t, iters); }
int fs_stage_p2g(fs_ctx* c) { return FS_DISPATCH(c, p2g_t); }
int fs_stage_dirichlet(fs_ctx* c) { return FS_DISPATCH(c, dirichlet_t); }
int fs_stage_extrapolate_velocity(fs_ctx* c, int iters) { return FS_DISPATCH(c, extrap_vel_t, iters); }
int fs_stage_force(fs_ctx* c, double dt) { return FS_DISPATCH(c, force_t, dt); }
int fs_stage_build(fs_ctx* c, double dt) { return FS_DISPATCH(c, build_t, dt); }
int fs_stage_project(fs_ctx* c, double dt) { return FS_DISPATCH(c, project_t, dt); }
int fs_stage_collider(fs_ctx* c) { return FS_DISPATCH(c, collider_t); }
int fs_stage_g2p(fs_ctx* c) { return FS_DISPATCH(c, g2p_t); }
int fs_stage_cfl(fs_ctx* c, double* out) { return FS_DISPATCH(c, cfl_t, out); }
int fs_adj_p2g_sum(fs_ctx* c, DBuf* pos, DBuf* coef, DBuf* out) { return FS_DISPATCH(c, p2g_sum_t, pos, coef, out); }
int fs_adj_g2p(fs_ctx* c, DBuf* x1, DBuf* g5, DBuf* g0, DBuf* vbar, DBuf* xbar, DBuf* vbar1, DBuf* coef) {
  return FS_DISPATCH(c, adj_g2p_t, x1, g5, g0, vbar, xbar, vbar1, coef);
}
int fs_adj_p2g(fs_ctx* c, DBuf* x1, DBuf* v1, DBuf* H, DBuf* HG, DBuf* vbar, DBuf* xbar) {
  return FS_DISPATCH(c, adj_p2g_t, x1, v1, H, HG, vbar, xbar);
}
int fs_adj_advect(fs_ctx* c, DBuf* x0, DBuf* v0, DBuf* xbar, DBuf* vbar, double dt) {
  return FS_DISPATCH(c, adj_advect_t, x0, v0, xbar, vbar, dt);
}
int fs_adj_sdf(fs_ctx* c, AdjStep& t, DBuf& phibar, DBuf& psibar, DBuf* xbar) {
  return FS_DISPATCH(c, adj_sdf_t, t, phibar, psibar, xbar);
}

int fs_upload_real(fs_ctx* c, DBuf& dst, const double* src, size_t n) {
  if (!c->mixed) {
    FS_HIP(c, hipMemcpyAsync(dst.p, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return FS_OK;
  }
  DBuf tmp;
  FS_TRY(fs_alloc(c, tmp, n * sizeof(double)));
  FS_HIP(c, hipMemcpyAsync(tmp.p, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_from_double<float><<<fs_blocks((int64_t)n, TPB), TPB, 0, c->stream>>>((int64_t)n, tmp.as<double>(), dst.as<float>());
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_free(tmp);
  return FS_OK;
}

int fs_download_real(fs_ctx* c, const DBuf& src, double* dst, size_t n) {
  if (!c->mixed) {
    FS_HIP(c, hipMemcpyAsync(dst, src.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    return FS_OK;
  }
  DBuf tmp;
  FS_TRY(fs_alloc(c, tmp, n * sizeof(double)));
  k_to_double<float><<<fs_blocks((int64_t)n, TPB), TPB, 0, c->stream>>>((int64_t)n, src.as<float>(), tmp.as<double>());
  FS_HIP(c, hipMemcpyAsync(dst, tmp.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_free(tmp);
  return FS_OK;
}

int fs_particles_upload(fs_ctx* c, const double* pos, const double* vel) {
  int64_t n = c->np;
  if (n == 0) return FS_OK;
  for (int which = 0; which < 2; which++) {
    const double* src = which == 0 ? pos : vel;
    DBuf* dst = which == 0 ? c->pos : c->vel;
    if (!src) {
      for (int d = 0; d < 3; d++) FS_TRY(fs_zero(c, dst[d]));
      continue;
    }
    FS_HIP(c, hipMemcpyAsync(c->staging.p, src, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->mixed)
      k_split_aos<float><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, c->staging.as<double>(), dst[0].as<float>(),
                                                                   dst[1].as<float>(), dst[2].as<float>());
    else
      k_split_aos<double><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, c->staging.as<double>(), dst[0].as<double>(),
                                                                    dst[1].as<double>(), dst[2].as<double>());
    FS_HIP(c, hipGetLastError());
    FS_HIP(c, hipStreamSynchronize(c->stream));
  }
  return FS_OK;
}

int fs_particles_download(fs_ctx* c, double* pos, double* vel) {
  int64_t n = c->np;
  if (n == 0) return FS_OK;
  for (int which = 0; which < 2; which++) {
    double* dst = which == 0 ? pos : vel;
    DBuf* src = which == 0 ? c->pos : c->vel;
    if (!dst) continue;
    if (c->mixed)
      k_join_aos<float><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, src[0].as<float>(), src[1].as<float>(),
                                                                  src[2].as<float>(), c->staging.as<double>());
    else
      k_join_aos<double><<<fs_blocks(n, TPB), TPB, 0, c->stream>>>(n, src[0].as<double>(), src[1].as<double>(),
                                                                   src[2].as<double>(), c->staging.as<double>());
    FS_HIP(c, hipGetLastError());
    FS_HIP(c, hipMemcpyAsync(dst, c->staging.p, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
  }
  return FS_OK;
}
