// fs_adjoint.hip - reverse mode of fs_substep over a recorded rollout.
//
// The reference has no derivative code. The step is differentiated with its discrete structure
// held fixed: the marking, the active set, the face weights, the valid masks of the
// extrapolations, the branches of min/clamp operations and the Poisson matrix's sparsity are
// constants of the step (piecewise constant in the particle positions). Everything else is
// exact reverse mode: G2P (positions through the kernel weights, grids through the transposed
// transfer), the collider and projection as linear maps, the pressure solve through the same
// MIC-PCG on the symmetric matrix, the fluid SDF (reconstruction -> extrapolation -> smoothing,
// including the SDF buffer a step hands to the next) through the ghost-fluid theta of the
// matrix and the projection, the velocity extrapolation with its stale ping-pong buffers,
// gravity, P2G (velocities and, through the weights, positions) and advection with its clamps.
// Advection must run on the cleared grid (the reference's order, advect_on_cleared_grid = 1),
// where the RK3 stages k2 and k3 vanish.
//
// Rollouts: fs_rollout_begin turns recording on; every `every` substeps a checkpoint of the
// whole state (particles and the SDF ping-pong buffers, the only state a substep carries over)
// is stored. fs_rollout_vjp walks the segments backwards: restore the segment's checkpoint,
// recompute its substeps recording each one, then reverse them.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fs_ctx.h"

using namespace fsd;

namespace {

constexpr int TPB = 256;
inline int blocks(int64_t n) { return (int)((n + TPB - 1) / TPB); }

int copy_buf(fs_ctx* c, DBuf& dst, const DBuf& src) {
  if (dst.bytes != src.bytes) {
    fs_free(dst);
    FS_TRY(fs_alloc(c, dst, src.bytes));
  }
  if (src.bytes) FS_HIP(c, hipMemcpyAsync(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice, c->stream));
  return FS_OK;
}

// face index (i, j, k) of grid a with dims d
__device__ inline void unlin(int64_t id, const int d[3], int& i, int& j, int& k) {
  k = (int)(id % d[2]);
  int64_t t = id / d[2];
  j = (int)(t % d[1]);
  i = (int)(t / d[1]);
}

struct Dims {
  int d[3];
};

// applyCollider (fluid-simulator.cc:76-111) is u *= 0 on the walls and u *= (1 - n_a) inside
// the collider: the reverse multiplies by the same factor.
template <typename T>
__global__ __launch_bounds__(TPB) void k_collider_adj(int a, Dims D, int nx, int ny, int nz, T h0, T h1, T h2, const T* __restrict__ col,
                               T* __restrict__ G) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D.d[0] * D.d[1] * D.d[2]) return;
  int ix[3];
  unlin(id, D.d, ix[0], ix[1], ix[2]);
  if (ix[a] == 0 || ix[a] == D.d[a] - 1) {
    G[id] = T(0);
    return;
  }
  const T h[3] = {h0, h1, h2};
  T p[3];
  for (int c = 0; c < 3; c++) p[c] = coord1(ix[c], c == a ? T(0) : T(0.5), h[c]);
  if (sdf_eval(col, nx, ny, nz, h, p[0], p[1], p[2]) < T(0)) {
    const V3<T> n = gnormalize(sdf_grad(col, nx, ny, nz, h, p[0], p[1], p[2]));
    const T na = a == 0 ? n.x : (a == 1 ? n.y : n.z);
    G[id] = G[id] - G[id] * na;
  }
}

// project (project-solver.cc:348-420), reverse on the faces: a face the projection sets to 0
// (weight > 0 on a domain wall) passes nothing back; every other face passes its adjoint.
template <typename T>
__global__ __launch_bounds__(TPB) void k_project_faces_adj(int a, Dims D, const T* __restrict__ wts, T* __restrict__ G) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D.d[0] * D.d[1] * D.d[2]) return;
  int ix[3];
  unlin(id, D.d, ix[0], ix[1], ix[2]);
  if (wts[id] > T(0) && (ix[a] == 0 || ix[a] == D.d[a] - 1)) G[id] = T(0);
}

// project, reverse into the pressure: each face the projection updates adds its coefficient
// of p(c) times the face adjoint. Faces: u between cells (i-1) and (i): both fluid
// u -= (p_i - p_{i-1}) dt/h; only the lower fluid u += p_{i-1} dt/(h theta); only the upper
// u -= p_i dt/(h (1 - theta)). Only active cells carry a pressure.
template <typename T>
__global__ __launch_bounds__(TPB) void k_project_cells_adj(int nx, int ny, int nz, T dth, const T* __restrict__ fsd,
                                    const uint8_t* __restrict__ act, const T* __restrict__ uw,
                                    const T* __restrict__ vw, const T* __restrict__ ww, const T* __restrict__ gu,
                                    const T* __restrict__ gv, const T* __restrict__ gw, T* __restrict__ pbar) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)nx * ny * nz) return;
  const int k = (int)(id % nz), j = (int)((id / nz) % ny), i = (int)(id / ((int64_t)nz * ny));
  if (!act[id]) {
    pbar[id] = T(0);
    return;
  }
  const T* wts[3] = {uw, vw, ww};
  const T* gs[3] = {gu, gv, gw};
  const T sc = fsd[id];
  T acc = T(0);
  for (int a = 0; a < 3; a++) {
    int fd[3] = {nx, ny, nz};
    fd[a] += 1;
    for (int side = 0; side < 2; side++) {  // 0: the face below the cell, 1: the face above
      int f[3] = {i, j, k};
      f[a] += side;
      if (f[a] == 0 || f[a] == fd[a] - 1) continue;  // walls are set to 0, not projected
      const int64_t fid = ((int64_t)f[0] * fd[1] + f[1]) * fd[2] + f[2];
      if (!(wts[a][fid] > T(0))) continue;
      int o[3] = {i, j, k};
      o[a] += side == 0 ? -1 : 1;
      const T so = fsd[((int64_t)o[0] * ny + o[1]) * nz + o[2]];
      const T slo = side == 0 ? so : sc, shi = side == 0 ? sc : so;  // face's lower / upper cell
      if (slo >= T(0) && shi >= T(0)) continue;
      const T gb = gs[a][fid];
      if (slo < T(0) && shi < T(0)) {
        acc += side == 0 ? -dth * gb : dth * gb;  // cell is the upper / lower one
      } else if (slo < T(0)) {                    // only the lower cell is fluid
        if (side == 1) acc += dth / smax(slo / (slo - shi), T(0.01)) * gb;
      } else {  // only the upper cell is fluid
        if (side == 0) acc -= dth / (T(1) - smin(slo / (slo - shi), T(0.99))) * gb;
      }
    }
  }
  pbar[id] = acc;
}


// The ghost-fluid fraction theta of buildSystem (project-solver.cc:143-223) and project
// (:348-420) depends on the fluid SDF, which depends on the particles. For every cell c this
// gathers d(loss)/d(phi_c) from
//   (a) Adiag of c and of its neighbours (Adiag += w f / (1 - theta) on a lower, w f / theta on
//       an upper air neighbour; f = dt/h), with dL/dAdiag = -lambda p (p = A^-1 rhs);
//   (b) the projection of faces with exactly one fluid side (u += p_l dt/(h theta) or
//       u -= p_u dt/(h (1 - theta))), weighted by the face adjoint.
// theta = a/(a - b) with the clamp inactive; a clamped theta carries no derivative.
template <typename T>
__device__ inline void theta_grad(T a, T b, T& da, T& db) {  // theta = a / (a - b)
  const T den = (a - b) * (a - b);
  da = -b / den;
  db = a / den;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_theta_adj(int nx, int ny, int nz, T f, const T* __restrict__ phi, const uint8_t* __restrict__ act,
                            const T* __restrict__ uw, const T* __restrict__ vw, const T* __restrict__ ww,
                            const T* __restrict__ lam, const T* __restrict__ p, const T* __restrict__ gu,
                            const T* __restrict__ gv, const T* __restrict__ gw, T* __restrict__ phibar) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)nx * ny * nz) return;
  const int k = (int)(id % nz), j = (int)((id / nz) % ny), i = (int)(id / ((int64_t)nz * ny));
  const int ix[3] = {i, j, k};
  const int cd[3] = {nx, ny, nz};
  const T* wts[3] = {uw, vw, ww};
  const T* gs[3] = {gu, gv, gw};
  const T pc = phi[id];
  const T adc = act[id] ? -lam[id] * p[id] : T(0);  // dL/dAdiag(c)
  T acc = T(0);
  for (int a = 0; a < 3; a++) {
    int fd[3] = {nx, ny, nz};
    fd[a] += 1;
    // the "up" neighbour test of buildSystem uses the width (Q8 of the reference)
    const int upper_bound = a == 1 ? nx : cd[a];
    for (int side = 0; side < 2; side++) {  // 0: neighbour below, 1: neighbour above
      const int ni = ix[a] + (side == 0 ? -1 : 1);
      if (side == 0 ? ix[a] <= 0 : ix[a] >= upper_bound - 1) continue;
      if (ni < 0 || ni >= cd[a]) continue;
      int o[3] = {i, j, k};
      o[a] = ni;
      const int64_t nid = ((int64_t)o[0] * ny + o[1]) * nz + o[2];
      int fc[3] = {i, j, k};
      fc[a] += side;  // the face between c and the neighbour
      const int64_t fid = ((int64_t)fc[0] * fd[1] + fc[1]) * fd[2] + fc[2];
      const T w = wts[a][fid];
      const T pn = phi[nid];
      const T adn = act[nid] ? -lam[nid] * p[nid] : T(0);
      T da, db;
      // (a) Adiag of c, neighbour is air
      if (act[id] && pn > T(0)) {
        if (side == 0) {  // theta = min(pn/(pn - pc), 0.99); Adiag += w f/(1 - theta)
          const T th = pn / (pn - pc);
          if (th < T(0.99)) {
            theta_grad(pn, pc, db, da);
            acc += adc * w * f / ((T(1) - th) * (T(1) - th)) * da;
          }
        } else {  // theta = max(pc/(pc - pn), 0.01); Adiag += w f/theta
          const T th = pc / (pc - pn);
          if (th > T(0.01)) {
            theta_grad(pc, pn, da, db);
            acc += adc * (-w * f / (th * th)) * da;
          }
        }
      }
      // (a) Adiag of the neighbour, c is its air neighbour (c lies on its opposite side)
      if (act[nid] && pc > T(0)) {
        if (side == 0) {  // c is the neighbour's upper cell: theta = max(pn/(pn - pc), 0.01)
          const T th = pn / (pn - pc);
          if (th > T(0.01)) {
            theta_grad(pn, pc, da, db);
            acc += adn * (-w * f / (th * th)) * db;
          }
        } else {  // c is the neighbour's lower cell: theta = min(pc/(pc - pn), 0.99)
          const T th = pc / (pc - pn);
          if (th < T(0.99)) {
            theta_grad(pc, pn, da, db);
            acc += adn * w * f / ((T(1) - th) * (T(1) - th)) * da;
          }
        }
      }
      // (b) projection of the face (walls are zeroed, not projected)
      if (fc[a] == 0 || fc[a] == fd[a] - 1 || !(w > T(0))) continue;
      const T pl = side == 0 ? pn : pc, pu = side == 0 ? pc : pn;  // SDF below / above the face
      const bool c_low = side == 1;
      const T gb = gs[a][fid];
      if (pl < T(0) && pu >= T(0)) {  // lower fluid: u += p_l dt/(h theta), theta = max(., 0.01)
        const T th = pl / (pl - pu);
        if (th > T(0.01)) {
          theta_grad(pl, pu, da, db);
          const T pres = c_low ? p[id] : p[nid];
          acc += gb * (-pres * f / (th * th)) * (c_low ? da : db);
        }
      } else if (pl >= T(0) && pu < T(0)) {  // upper fluid: u -= p_u dt/(h (1 - theta))
        const T th = pl / (pl - pu);
        if (th < T(0.99)) {
          theta_grad(pl, pu, da, db);
          const T pres = c_low ? p[nid] : p[id];
          acc += gb * (-pres * f / ((T(1) - th) * (T(1) - th))) * (c_low ? da : db);
        }
      }
    }
  }
  phibar[id] = acc;
}

// buildSystem's right-hand side (project-solver.cc:143-223): rhs(c) = sum over c's faces of
// +w u (lower face) and -w u (upper face), interior faces only. Reverse:
// ubar(f) += w(f) (lambda(upper cell) - lambda(lower cell)), active cells only.
template <typename T>
__global__ __launch_bounds__(TPB) void k_rhs_adj(int a, Dims D, int nx, int ny, int nz, const T* __restrict__ wts,
                          const uint8_t* __restrict__ act, const T* __restrict__ lam, T* __restrict__ G) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D.d[0] * D.d[1] * D.d[2]) return;
  int ix[3];
  unlin(id, D.d, ix[0], ix[1], ix[2]);
  if (ix[a] == 0 || ix[a] == D.d[a] - 1) return;
  const int64_t up = ((int64_t)ix[0] * ny + ix[1]) * nz + ix[2];
  int lo3[3] = {ix[0], ix[1], ix[2]};
  lo3[a] -= 1;
  const int64_t lo = ((int64_t)lo3[0] * ny + lo3[1]) * nz + lo3[2];
  const T lu = act[up] ? lam[up] : T(0), ll = act[lo] ? lam[lo] : T(0);
  (void)nx;
  G[id] = G[id] + wts[id] * (lu - ll);
}

// applyForce (fluid-simulator.cc:13-18) adds gravity*dt on valid faces off the walls; the
// gravity derivative is dt times the sum of the adjoint over those faces (per block, summed on
// the host in block order).
template <typename T>
__global__ __launch_bounds__(TPB) void k_force_adj(int a, Dims D, const uint8_t* __restrict__ valid,
                                                   const T* __restrict__ G, double* __restrict__ part) {
  __shared__ double sm[TPB];
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double v = 0.0;
  if (id < (int64_t)D.d[0] * D.d[1] * D.d[2]) {
    int ix[3];
    unlin(id, D.d, ix[0], ix[1], ix[2]);
    if (valid[id] && ix[a] > 0 && ix[a] < D.d[a] - 1) v = (double)G[id];
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// extrapolate (cpu/fluid-simulator.h:204-253): the valid mask after one iteration.
__global__ __launch_bounds__(TPB) void k_extrap_mask(Dims D, const uint8_t* __restrict__ V, uint8_t* __restrict__ VN) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D.d[0] * D.d[1] * D.d[2]) return;
  int i, j, k;
  unlin(id, D.d, i, j, k);
  const int W = D.d[0], H = D.d[1], Z = D.d[2];
  const int64_t sx = (int64_t)H * Z, sy = Z;
  uint8_t r = V[id];
  if (!r) {
    r = (i > 0 && V[id - sx]) || (i < W - 1 && V[id + sx]) || (j > 0 && V[id - sy]) || (j < H - 1 && V[id + sy]) ||
        (k > 0 && V[id - 1]) || (k < Z - 1 && V[id + 1]);
  }
  VN[id] = r;
}

__device__ inline int valid_count(const uint8_t* V, int64_t id, int i, int j, int k, int W, int H, int Z) {
  const int64_t sx = (int64_t)H * Z, sy = Z;
  return (i > 0 && V[id - sx]) + (i < W - 1 && V[id + sx]) + (j > 0 && V[id - sy]) + (j < H - 1 && V[id + sy]) +
         (k > 0 && V[id - 1]) + (k < Z - 1 && V[id + 1]);
}

// One extrapolation iteration reversed. Forward, with mask V: GB'(c) = GB(c) for valid c,
// mean of the valid neighbours' G for invalid c with some, 0 otherwise; then the arrays swap,
// so (G, GB) -> (GB', G). Reverse, from the adjoints (A, B) of (G_new, GB_new):
//   Gbar(n)  = B(n) + [n valid] sum over invalid neighbours c of A(c) / count(c)
//   GBbar(c) = [c valid] A(c)
template <typename T>
__global__ __launch_bounds__(TPB) void k_extrap_adj(Dims D, const uint8_t* __restrict__ V, const T* __restrict__ A, const T* __restrict__ B,
                             T* __restrict__ Gb, T* __restrict__ GBb) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D.d[0] * D.d[1] * D.d[2]) return;
  int i, j, k;
  unlin(id, D.d, i, j, k);
  const int W = D.d[0], H = D.d[1], Z = D.d[2];
  const int64_t sx = (int64_t)H * Z, sy = Z;
  T g = B[id];
  if (V[id]) {
    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int q = 0; q < 6; q++) {
      const int ci = i + nb[q][0], cj = j + nb[q][1], ck = k + nb[q][2];
      if (ci < 0 || ci >= W || cj < 0 || cj >= H || ck < 0 || ck >= Z) continue;
      const int64_t cid = id + nb[q][0] * sx + nb[q][1] * sy + nb[q][2];
      if (V[cid]) continue;
      g += A[cid] / T(valid_count(V, cid, ci, cj, ck, W, H, Z));
    }
  }
  Gb[id] = g;
  GBb[id] = V[id] ? A[id] : T(0);
}

// applyDirichletBoundary (fluid-simulator.cc:113-138) sets the wall faces to 0.
template <typename T>
__global__ __launch_bounds__(TPB) void k_zero_walls(int a, Dims D, T* __restrict__ G) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D.d[0] * D.d[1] * D.d[2]) return;
  int ix[3];
  unlin(id, D.d, ix[0], ix[1], ix[2]);
  if (ix[a] == 0 || ix[a] == D.d[a] - 1) G[id] = T(0);
}

// P2G reverse node fields: H = gbar / W where W > 0 (else 0), HG = H * g.
template <typename T>
__global__ __launch_bounds__(TPB) void k_p2g_fields(int64_t n, const T* __restrict__ gb, const T* __restrict__ W, const T* __restrict__ g,
                             T* __restrict__ H, T* __restrict__ HG) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= n) return;
  const T h = W[id] > T(0) ? gb[id] / W[id] : T(0);
  H[id] = h;
  HG[id] = h * g[id];
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_scale(int64_t n, T s, const T* __restrict__ a, T* __restrict__ b) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id < n) b[id] = s * a[id];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_add(int64_t n, const T* __restrict__ a, T* __restrict__ b) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id < n) b[id] = b[id] + a[id];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_split3(int64_t n, const double* __restrict__ aos, T* __restrict__ x, T* __restrict__ y,
                         T* __restrict__ z) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = (T)aos[3 * i];
  y[i] = (T)aos[3 * i + 1];
  z[i] = (T)aos[3 * i + 2];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_join3(int64_t n, const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
                        double* __restrict__ aos) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  aos[3 * i] = (double)x[i];
  aos[3 * i + 1] = (double)y[i];
  aos[3 * i + 2] = (double)z[i];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_from_f64(int64_t n, const double* __restrict__ a, T* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_to_f64(int64_t n, const T* __restrict__ a, double* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

Dims dims_of(const fs_ctx* c, int a) { return {{c->fdim[a][0], c->fdim[a][1], c->fdim[a][2]}}; }

// Scratch of one reverse pass (context precision T; sizes in elements).
struct AdjWork {
  DBuf xb[3], vb[3], vb1[3], coef[3];  // particles
  DBuf gb5[3], gb0[3], A[3], B[3], A2[3], B2[3], H[3], HG[3];  // faces
  std::vector<DBuf> masks[3];  // valid masks of the extrapolation iterations
  DBuf part;                   // force partial sums
  DBuf phibar, psibar;         // adjoints of the smoothed SDF and of the carried SDF buffer
  void release() {
    for (int a = 0; a < 3; a++) {
      for (DBuf* b : {&xb[a], &vb[a], &vb1[a], &coef[a], &gb5[a], &gb0[a], &A[a], &B[a], &A2[a], &B2[a], &H[a],
                      &HG[a]})
        fs_free(*b);
      for (DBuf& m : masks[a]) fs_free(m);
      masks[a].clear();
    }
    fs_free(part);
    fs_free(phibar);
    fs_free(psibar);
  }
};

template <typename T>
int adj_step_t(fs_ctx* c, AdjStep& t, AdjWork& w, double grav[3]) {
  const int64_t np = c->np;
  const size_t R = sizeof(T);
  const double dt = t.dt;
  // 1. G2P: position terms into xbar, alpha*vbar into vbar1, transfer coefficients
  FS_TRY(fs_adj_g2p(c, t.x1, t.g5, t.g0, w.vb, w.xb, w.vb1, w.coef));
  // 2. transposed transfer to the grids: gbar5 = P^T coef; FLIP's u_last gets -alpha of it
  FS_TRY(fs_adj_p2g_sum(c, t.x1, w.coef, w.gb5));
  const bool flip = c->cfg.advection == FS_ADVECT_FLIP;
  for (int a = 0; a < 3; a++) {
    const int64_t nf = c->n_face[a];
    k_scale<T><<<blocks(nf), TPB, 0, c->stream>>>(nf, flip ? (T)-c->cfg.flip_alpha : T(0), w.gb5[a].as<T>(),
                                                  w.gb0[a].as<T>());
    // 3. collider
    const T h[3] = {(T)c->h[0], (T)c->h[1], (T)c->h[2]};
    k_collider_adj<T><<<blocks(nf), TPB, 0, c->stream>>>(a, dims_of(c, a), c->nx, c->ny, c->nz, h[0], h[1], h[2],
                                                         c->collider.as<T>(), w.gb5[a].as<T>());
  }
  FS_HIP(c, hipGetLastError());
  // 4. projection into the pressure, then the faces themselves
  k_project_cells_adj<T><<<blocks(c->n_cell), TPB, 0, c->stream>>>(
      c->nx, c->ny, c->nz, (T)(dt / c->h[0]), t.fsdf.as<T>(), t.active.as<uint8_t>(), t.weights[0].as<T>(),
      t.weights[1].as<T>(), t.weights[2].as<T>(), w.gb5[0].as<T>(), w.gb5[1].as<T>(), w.gb5[2].as<T>(),
      c->rhs.as<T>());
  for (int a = 0; a < 3; a++)
    k_project_faces_adj<T><<<blocks(c->n_face[a]), TPB, 0, c->stream>>>(a, dims_of(c, a), t.weights[a].as<T>(),
                                                                       w.gb5[a].as<T>());
  FS_HIP(c, hipGetLastError());
  // 5. lambda = A^-1 pbar with the forward's preconditioned CG (A is symmetric); the system of this step
  FS_TRY(copy_buf(c, c->active, t.active));
  FS_TRY(copy_buf(c, c->adiag, t.adiag));
  for (int q = 0; q < 6; q++) FS_TRY(copy_buf(c, c->an[q], t.an[q]));
  const fs_config saved = c->cfg;
  c->cfg.tol_mode = FS_TOL_REL_L2;
  c->cfg.cg_tolerance = c->mixed ? 1e-8 : 1e-10;
  c->cfg.cg_max_iterations = 4 * std::max(saved.cg_max_iterations, 100);
  fs_stats S{};
  int st = fs_stage_solve(c, &S);
  c->cfg = saved;
  if (st == FS_DEVICE_ERROR || st == FS_UNSUPPORTED || st == FS_BAD_ARG) return st;
  // 5b. the fluid SDF through theta (lambda = pg now, the forward pressure was recorded)
  k_theta_adj<T><<<blocks(c->n_cell), TPB, 0, c->stream>>>(
      c->nx, c->ny, c->nz, (T)(dt / c->h[0]), t.fsdf.as<T>(), t.active.as<uint8_t>(), t.weights[0].as<T>(),
      t.weights[1].as<T>(), t.weights[2].as<T>(), c->pg.as<T>(), t.pg.as<T>(), w.gb5[0].as<T>(), w.gb5[1].as<T>(),
      w.gb5[2].as<T>(), w.phibar.as<T>());
  FS_HIP(c, hipGetLastError());
  // 6. rhs: gbar3 = (faces) + weighted difference of lambda
  for (int a = 0; a < 3; a++)
    k_rhs_adj<T><<<blocks(c->n_face[a]), TPB, 0, c->stream>>>(a, dims_of(c, a), c->nx, c->ny, c->nz,
                                                             t.weights[a].as<T>(), t.active.as<uint8_t>(),
                                                             c->pg.as<T>(), w.gb5[a].as<T>());
  FS_HIP(c, hipGetLastError());
  // 7-8. extrapolation masks; gravity; extrapolation reversed (A: current array, B: buffer)
  const int iters = c->cfg.vel_extrapolate_iters;
  for (int a = 0; a < 3; a++) {
    const int64_t nf = c->n_face[a];
    Dims D = dims_of(c, a);
    while ((int)w.masks[a].size() < iters + 1) {
      w.masks[a].emplace_back();
      FS_TRY(fs_alloc(c, w.masks[a].back(), (size_t)nf));
    }
    FS_TRY(copy_buf(c, w.masks[a][0], t.valid[a]));
    for (int it = 0; it < iters; it++)
      k_extrap_mask<<<blocks(nf), TPB, 0, c->stream>>>(D, w.masks[a][it].as<uint8_t>(), w.masks[a][it + 1].as<uint8_t>());
    const int nbk = blocks(nf);
    if (w.part.bytes < (size_t)nbk * sizeof(double)) {
      fs_free(w.part);
      FS_TRY(fs_alloc(c, w.part, (size_t)nbk * sizeof(double)));
    }
    k_force_adj<T><<<nbk, TPB, 0, c->stream>>>(a, D, w.masks[a][iters].as<uint8_t>(), w.gb5[a].as<T>(),
                                               w.part.as<double>());
    std::vector<double> part(nbk);
    FS_HIP(c, hipMemcpyAsync(part.data(), w.part.p, nbk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : part) s += v;
    grav[a] += dt * s;
    DBuf *A = &w.gb5[a], *B = &w.B[a], *A2 = &w.A2[a], *B2 = &w.B2[a];
    FS_HIP(c, hipMemsetAsync(B->p, 0, nf * R, c->stream));
    for (int it = iters - 1; it >= 0; it--) {
      k_extrap_adj<T><<<blocks(nf), TPB, 0, c->stream>>>(D, w.masks[a][it].as<uint8_t>(), A->as<T>(), B->as<T>(),
                                                         A2->as<T>(), B2->as<T>());
      std::swap(A, A2);
      std::swap(B, B2);
    }
    // 9. Dirichlet walls, plus FLIP's u_last; 10. P2G node fields
    k_zero_walls<T><<<blocks(nf), TPB, 0, c->stream>>>(a, D, A->as<T>());
    k_add<T><<<blocks(nf), TPB, 0, c->stream>>>(nf, A->as<T>(), w.gb0[a].as<T>());
    k_p2g_fields<T><<<blocks(nf), TPB, 0, c->stream>>>(nf, w.gb0[a].as<T>(), t.w0[a].as<T>(), t.g0[a].as<T>(),
                                                       w.H[a].as<T>(), w.HG[a].as<T>());
    FS_HIP(c, hipGetLastError());
  }
  FS_TRY(fs_adj_p2g(c, t.x1, t.v1, w.H, w.HG, w.vb1, w.xb));
  // 10b. the fluid SDF chain back to the particle positions (and the carried SDF buffer)
  FS_TRY(fs_adj_sdf(c, t, w.phibar, w.psibar, w.xb));
  // 11. advection: xbar, vbar1 -> adjoints of the substep's inputs
  FS_TRY(fs_adj_advect(c, t.x0, t.v0, w.xb, w.vb1, dt));
  for (int d = 0; d < 3; d++) std::swap(w.vb[d], w.vb1[d]);
  (void)np;
  return FS_OK;
}

int alloc_work(fs_ctx* c, AdjWork& w) {
  const size_t R = c->rsz, np = (size_t)std::max<int64_t>(c->np, 1);
  for (int a = 0; a < 3; a++) {
    for (DBuf* b : {&w.xb[a], &w.vb[a], &w.vb1[a], &w.coef[a]}) FS_TRY(fs_alloc(c, *b, np * R));
    const size_t nf = (size_t)c->n_face[a];
    for (DBuf* b : {&w.gb5[a], &w.gb0[a], &w.A[a], &w.B[a], &w.A2[a], &w.B2[a], &w.H[a], &w.HG[a]})
      FS_TRY(fs_alloc(c, *b, nf * R));
  }
  FS_TRY(fs_alloc(c, w.phibar, (size_t)c->n_cell * R));
  FS_TRY(fs_alloc(c, w.psibar, (size_t)c->n_cell * R));
  FS_HIP(c, hipMemsetAsync(w.psibar.p, 0, (size_t)c->n_cell * R, c->stream));
  return FS_OK;
}

// with_grids: also the face grids and the pressure, so that the state a caller can read back
// after fs_rollout_vjp is the one it left (the recomputation overwrites them).
int save_checkpoint(fs_ctx* c, AdjCheckpoint& k, bool with_grids = false) {
  for (int d = 0; d < 3; d++) {
    FS_TRY(copy_buf(c, k.x[d], c->pos[d]));
    FS_TRY(copy_buf(c, k.v[d], c->vel[d]));
  }
  if (with_grids) {
    for (int a = 0; a < 3; a++) FS_TRY(copy_buf(c, k.g[a], c->g[a]));
    FS_TRY(copy_buf(c, k.pg, c->pg));
  }
  FS_TRY(copy_buf(c, k.fsdf, c->fsdf));
  FS_TRY(copy_buf(c, k.fsdfbuf, c->fsdfbuf));
  FS_TRY(copy_buf(c, k.sdfvalid, c->sdfvalid));
  FS_TRY(copy_buf(c, k.sdfvalidbuf, c->sdfvalidbuf));
  return FS_OK;
}

int load_checkpoint(fs_ctx* c, const AdjCheckpoint& k) {
  for (int d = 0; d < 3; d++) {
    FS_TRY(copy_buf(c, c->pos[d], k.x[d]));
    FS_TRY(copy_buf(c, c->vel[d], k.v[d]));
  }
  FS_TRY(copy_buf(c, c->fsdf, k.fsdf));
  FS_TRY(copy_buf(c, c->fsdfbuf, k.fsdfbuf));
  FS_TRY(copy_buf(c, c->sdfvalid, k.sdfvalid));
  FS_TRY(copy_buf(c, c->sdfvalidbuf, k.sdfvalidbuf));
  if (k.pg.p) {
    for (int a = 0; a < 3; a++) FS_TRY(copy_buf(c, c->g[a], k.g[a]));
    FS_TRY(copy_buf(c, c->pg, k.pg));
  }
  return FS_OK;
}

void free_checkpoint(AdjCheckpoint& k) {
  for (int d = 0; d < 3; d++) {
    fs_free(k.x[d]);
    fs_free(k.v[d]);
    fs_free(k.g[d]);
  }
  for (DBuf* b : {&k.fsdf, &k.fsdfbuf, &k.sdfvalid, &k.sdfvalidbuf, &k.pg}) fs_free(*b);
}

void free_step(AdjStep& s) {
  for (int a = 0; a < 3; a++)
    for (DBuf* b : {&s.x0[a], &s.v0[a], &s.x1[a], &s.v1[a], &s.g0[a], &s.w0[a], &s.valid[a], &s.g5[a],
                    &s.weights[a]})
      fs_free(*b);
  for (DBuf* b : {&s.fsdf, &s.active, &s.adiag, &s.sdfbuf0, &s.pg}) fs_free(*b);
  for (int q = 0; q < 6; q++) fs_free(s.an[q]);
}

template <typename T>
int upload_particles(fs_ctx* c, const double* aos, DBuf* dst) {
  const int64_t n = c->np;
  if (!aos) {
    for (int d = 0; d < 3; d++) FS_HIP(c, hipMemsetAsync(dst[d].p, 0, (size_t)n * sizeof(T), c->stream));
    return FS_OK;
  }
  FS_HIP(c, hipMemcpyAsync(c->staging.p, aos, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_split3<T><<<blocks(n), TPB, 0, c->stream>>>(n, c->staging.as<double>(), dst[0].as<T>(), dst[1].as<T>(),
                                                dst[2].as<T>());
  FS_HIP(c, hipGetLastError());
  FS_HIP(c, hipStreamSynchronize(c->stream));
  return FS_OK;
}

template <typename T>
int download_particles(fs_ctx* c, DBuf* src, double* aos) {
  const int64_t n = c->np;
  if (!aos) return FS_OK;
  k_join3<T><<<blocks(n), TPB, 0, c->stream>>>(n, src[0].as<T>(), src[1].as<T>(), src[2].as<T>(),
                                               c->staging.as<double>());
  FS_HIP(c, hipGetLastError());
  FS_HIP(c, hipMemcpyAsync(aos, c->staging.p, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  return FS_OK;
}

template <typename T>
int rollout_vjp_t(fs_ctx* c, const double* dpos, const double* dvel, double* dpos0, double* dvel0, double* dgrav) {
  AdjTape& tp = c->tape;
  const int nsteps = (int)tp.dts.size();
  AdjWork w;
  int st = alloc_work(c, w);
  double grav[3] = {0, 0, 0};
  if (st == FS_OK) st = save_checkpoint(c, tp.end, true);
  if (st == FS_OK) st = upload_particles<T>(c, dpos, w.xb);
  if (st == FS_OK) st = upload_particles<T>(c, dvel, w.vb);
  const int every = std::max(1, tp.every);
  const int nseg = (nsteps + every - 1) / every;
  if ((int)tp.seg.size() < every) tp.seg.resize(every);
  for (int seg = nseg - 1; seg >= 0 && st == FS_OK; seg--) {
    const int s0 = seg * every, s1 = std::min(s0 + every, nsteps);
    st = load_checkpoint(c, tp.ck[seg]);
    for (int j = s0; j < s1 && st == FS_OK; j++) {
      fs_stats S{};
      const int r = fs_substep_core(c, tp.dts[j], &S, &tp.seg[j - s0]);
      if (r == FS_DEVICE_ERROR || r == FS_UNSUPPORTED || r == FS_BAD_ARG) st = r;
    }
    for (int j = s1 - 1; j >= s0 && st == FS_OK; j--) st = adj_step_t<T>(c, tp.seg[j - s0], w, grav);
  }
  if (st == FS_OK) st = download_particles<T>(c, w.xb, dpos0);
  if (st == FS_OK) st = download_particles<T>(c, w.vb, dvel0);
  if (st == FS_OK) st = load_checkpoint(c, tp.end);
  if (st == FS_OK) FS_HIP(c, hipStreamSynchronize(c->stream));
  if (dgrav)
    for (int a = 0; a < 3; a++) dgrav[a] = grav[a];
  w.release();
  return st;
}

template <typename T>
int sample_vjp_t(fs_ctx* c, const double* pos, const double* dvel, double* dgrid[3]) {
  // v_a(p) = sum_n w g_a(n) / S: the transposed transfer of dvel / S, as in the G2P reverse
  AdjWork w;
  FS_TRY(alloc_work(c, w));
  DBuf* P = c->pos;
  int st = FS_OK;
  if (pos) {  // explicit positions: the G2P reverse's scratch holds them
    st = upload_particles<T>(c, pos, w.vb1);
    P = w.vb1;
  }
  if (st == FS_OK) st = upload_particles<T>(c, dvel, w.vb);
  // with FLIP off the G2P reverse gives coef = dvel / S; the grids only enter the position terms
  const int32_t adv = c->cfg.advection;
  c->cfg.advection = FS_ADVECT_PIC;
  if (st == FS_OK) st = fs_adj_g2p(c, P, c->g, c->g, w.vb, w.xb, w.xb, w.coef);
  c->cfg.advection = adv;
  if (st == FS_OK) st = fs_adj_p2g_sum(c, P, w.coef, w.gb5);
  for (int a = 0; a < 3 && st == FS_OK; a++) {
    const int64_t nf = c->n_face[a];
    DBuf tmp;
    st = fs_alloc(c, tmp, nf * sizeof(double));
    if (st != FS_OK) break;
    k_to_f64<T><<<blocks(nf), TPB, 0, c->stream>>>(nf, w.gb5[a].as<T>(), tmp.as<double>());
    hipError_t e = hipMemcpyAsync(dgrid[a], tmp.p, nf * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    fs_free(tmp);
    if (e != hipSuccess) {
      c->set_error(hipGetErrorString(e));
      st = FS_DEVICE_ERROR;
    }
  }
  w.release();
  return st;
}

}  // namespace

int fs_adj_record(fs_ctx* c, AdjStep* r, AdjPoint at) {
  switch (at) {
    case ADJ_INPUT:
      for (int d = 0; d < 3; d++) {
        FS_TRY(copy_buf(c, r->x0[d], c->pos[d]));
        FS_TRY(copy_buf(c, r->v0[d], c->vel[d]));
      }
      FS_TRY(copy_buf(c, r->sdfbuf0, c->fsdfbuf));
      break;
    case ADJ_ADVECTED:
      for (int d = 0; d < 3; d++) {
        FS_TRY(copy_buf(c, r->x1[d], c->pos[d]));
        FS_TRY(copy_buf(c, r->v1[d], c->vel[d]));
      }
      break;
    case ADJ_P2G:
      for (int a = 0; a < 3; a++) {
        FS_TRY(copy_buf(c, r->g0[a], c->g[a]));
        FS_TRY(copy_buf(c, r->w0[a], c->gw[a]));
      }
      break;
    case ADJ_DIRICHLET:
      for (int a = 0; a < 3; a++) FS_TRY(copy_buf(c, r->valid[a], c->gvalid[a]));
      break;
    case ADJ_BUILT:
      for (int a = 0; a < 3; a++) FS_TRY(copy_buf(c, r->weights[a], c->gweights[a]));
      FS_TRY(copy_buf(c, r->fsdf, c->fsdf));
      FS_TRY(copy_buf(c, r->active, c->active));
      FS_TRY(copy_buf(c, r->adiag, c->adiag));
      for (int q = 0; q < 6; q++) FS_TRY(copy_buf(c, r->an[q], c->an[q]));
      break;
    case ADJ_FINAL:
      for (int a = 0; a < 3; a++) FS_TRY(copy_buf(c, r->g5[a], c->g[a]));
      FS_TRY(copy_buf(c, r->pg, c->pg));
      break;
  }
  return FS_OK;
}

int fs_adj_on_substep(fs_ctx* c, double dt) {
  AdjTape& tp = c->tape;
  const int k = (int)tp.dts.size();
  if (k % std::max(1, tp.every) == 0) {
    tp.ck.emplace_back();
    FS_TRY(save_checkpoint(c, tp.ck.back()));
  }
  tp.dts.push_back(dt);
  return FS_OK;
}

void fs_adj_release(fs_ctx* c) {
  AdjTape& tp = c->tape;
  for (AdjCheckpoint& k : tp.ck) free_checkpoint(k);
  for (AdjStep& s : tp.seg) free_step(s);
  free_checkpoint(tp.end);
  for (DBuf* b : {&tp.sA, &tp.sB, &tp.sAb, &tp.sBb, &tp.sAb2, &tp.sBb2}) fs_free(*b);
  for (DBuf& b : tp.sV) fs_free(b);
  for (DBuf& b : tp.sF) fs_free(b);
  tp.sV.clear();
  tp.sF.clear();
  tp.ck.clear();
  tp.seg.clear();
  tp.dts.clear();
  tp.on = false;
  tp.recording = false;
}

extern "C" {

int fs_rollout_begin(fs_ctx* c, int32_t checkpoint_every) {
  if (!c) return FS_BAD_ARG;
  if (checkpoint_every < 1) {
    c->set_error("checkpoint interval must be >= 1");
    return FS_BAD_ARG;
  }
  if (!c->cfg.advect_on_cleared_grid) {
    c->set_error("the reverse pass needs advection on the cleared grid (advect_on_cleared_grid = 1)");
    return FS_UNSUPPORTED;
  }
  if (c->cfg.advection == FS_ADVECT_APIC) {
    c->set_error("the reverse pass covers FLIP and PIC, not APIC");
    return FS_UNSUPPORTED;
  }
  fs_adj_release(c);
  c->tape.on = true;
  c->tape.recording = true;
  c->tape.every = checkpoint_every;
  return FS_OK;
}

int fs_rollout_stop(fs_ctx* c) {
  if (!c) return FS_BAD_ARG;
  c->tape.recording = false;
  return FS_OK;
}

int fs_rollout_end(fs_ctx* c) {
  if (!c) return FS_BAD_ARG;
  fs_adj_release(c);
  return FS_OK;
}

int fs_rollout_steps(const fs_ctx* c) { return c ? (int)c->tape.dts.size() : -1; }

int fs_rollout_vjp(fs_ctx* c, const double* dpos, const double* dvel, double* dpos0, double* dvel0,
                   double* dgravity) {
  if (!c) return FS_BAD_ARG;
  if (!c->tape.on) {
    c->set_error("no rollout has been recorded (fs_rollout_begin)");
    return FS_BAD_ARG;
  }
  return c->mixed ? rollout_vjp_t<float>(c, dpos, dvel, dpos0, dvel0, dgravity)
                  : rollout_vjp_t<double>(c, dpos, dvel, dpos0, dvel0, dgravity);
}

int fs_sample_grid(fs_ctx* c, const double* pos, const double* u, const double* v, const double* w, double* vel) {
  if (!c || !u || !v || !w || !vel) return FS_BAD_ARG;
  // PIC G2P of the given grids at the given (or the current) positions; the simulation's
  // particles are left as they were (the face grids are scratch until the next substep)
  const double* src[3] = {u, v, w};
  for (int a = 0; a < 3; a++) FS_TRY(fs_upload_real(c, c->g[a], src[a], (size_t)c->n_face[a]));
  DBuf tp[3], tv[3];
  int st = FS_OK;
  for (int d = 0; d < 3 && st == FS_OK; d++) {
    st = fs_alloc(c, tv[d], (size_t)std::max<int64_t>(c->np, 1) * c->rsz);
    if (st == FS_OK && pos) st = fs_alloc(c, tp[d], (size_t)std::max<int64_t>(c->np, 1) * c->rsz);
  }
  if (st == FS_OK && pos) st = c->mixed ? upload_particles<float>(c, pos, tp) : upload_particles<double>(c, pos, tp);
  if (st == FS_OK) {
    for (int d = 0; d < 3; d++) {
      std::swap(c->vel[d], tv[d]);
      if (pos) std::swap(c->pos[d], tp[d]);
    }
    const int32_t adv = c->cfg.advection;
    c->cfg.advection = FS_ADVECT_PIC;
    st = fs_stage_g2p(c);
    c->cfg.advection = adv;
    if (st == FS_OK) st = fs_particles_download(c, nullptr, vel);
    for (int d = 0; d < 3; d++) {
      std::swap(c->vel[d], tv[d]);
      if (pos) std::swap(c->pos[d], tp[d]);
    }
  }
  for (int d = 0; d < 3; d++) {
    fs_free(tp[d]);
    fs_free(tv[d]);
  }
  return st;
}

int fs_sample_grid_vjp(fs_ctx* c, const double* pos, const double* dvel, double* du, double* dv, double* dw) {
  if (!c || !dvel || !du || !dv || !dw) return FS_BAD_ARG;
  double* out[3] = {du, dv, dw};
  return c->mixed ? sample_vjp_t<float>(c, pos, dvel, out) : sample_vjp_t<double>(c, pos, dvel, out);
}

}  // extern "C"
