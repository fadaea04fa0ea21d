// fs_mg.hip - multigrid-preconditioned CG for the pressure system.
//
// PreconditionerMethod::Multigrid is declared in the reference without an implementation
// (FluidSim/include/FluidSim/cpu/project-solver.h:117-118; setCompressedSolver reaches ERROR).
// This is the library's own multigrid; the CG loop around it is CgSolver3D::solve
// (FluidSim/src/cpu/project-solver.cc:308-346) with the same stopping rules and status codes.
//
// Hierarchy, rebuilt for every solve (the system changes every substep):
//   level 0  the assembled system on the grid: D = Adiag on active cells, Cx/Cy/Cz = the coupling
//            to the -x/-y/-z neighbour (An[Left], An[Down], An[Back]) where both cells are active,
//            else 0; a cell's +x coupling is Cx of its +x neighbour (the assembly is symmetric,
//            An[Right](i) == An[Left](i+1) bit for bit). W = omega / D where D > 0, else 0.
//   level l+1  2x2x2 aggregates of level l: the Galerkin product P^T A P of piecewise-constant
//            prolongation. D sums the children's D and twice their couplings inside the
//            aggregate; Cx sums the -x couplings of the children on the aggregate's -x face.
//            Coarsening stops at one cell.
// V-cycle on b (symmetric, so CG keeps its guarantees): two weighted-Jacobi sweeps from zero,
// the residual restricted by summation over the aggregate, the coarse cycle, the coarse solution
// added with weight alpha (over-correction, for the piecewise-constant prolongation), two
// sweeps; the coarsest level runs four sweeps.
//
// Layout: every level is a dense grid, k fastest, in the context's storage type S (FP32 in mixed
// precision); the CG vectors x, r, s, t are FP64 in both modes, so the stopping test is always
// made on an FP64 recurrence of the FP64 system. Kernels that reduce (dot products, norms) give a
// block of 256 threads one chunk of 4096 consecutive cells (thread t: cells t + 256 m, m = 0..15),
// the fixed tree; the others take one cell per thread. A cell with W = 0 (inactive) reads no
// neighbour. Sweeps that need the previous iterate of a neighbour recompute it from what is
// stored (W b for the first sweep, x + alpha x_coarse after the coarse correction), so each
// level costs four passes: two sweeps, restriction, correction with a sweep, a sweep. Dot
// products use the fixed tree of fs_reduce.h over the chunks, and the level-0 sweep that ends
// the cycle also forms z.r.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fs_ctx.h"
#include "fs_reduce.h"

namespace {

constexpr double MG_OMEGA = 0.9;  // Jacobi weight
constexpr double MG_ALPHA = 1.8;  // coarse-correction weight
constexpr int MCH = 4096;         // cells per block

template <typename S>
struct Lv {
  int nx, ny, nz;
  const S *D, *Cx, *Cy, *Cz, *W;
};

__device__ __forceinline__ void decode3(int64_t c, int ny, int nz, int& i, int& j, int& k) {
  const unsigned uc = (unsigned)c;  // levels hold fewer than 2^31 cells (checked at create)
  const unsigned q = uc / (unsigned)nz;
  k = (int)(uc - q * (unsigned)nz);
  i = (int)(q / (unsigned)ny);
  j = (int)(q - (unsigned)i * (unsigned)ny);
}

// Row c of the level's operator applied to xo: diagonal, -x, -y, -z, +x, +y, +z (the order of
// applyCompressedMatrix, project-solver.cc:23-35); terms outside the grid are dropped. xo(q, i, j, k)
// gives the operand at cell q = (i, j, k).
template <typename R, typename S, typename XO>
__device__ __forceinline__ R row(const Lv<S>& L, int64_t c, int i, int j, int k, XO xo) {
  const int64_t sx = (int64_t)L.ny * L.nz, sy = L.nz;
  R t = (R)L.D[c] * xo(c, i, j, k);
  if (i > 0) t = t + (R)L.Cx[c] * xo(c - sx, i - 1, j, k);
  if (j > 0) t = t + (R)L.Cy[c] * xo(c - sy, i, j - 1, k);
  if (k > 0) t = t + (R)L.Cz[c] * xo(c - 1, i, j, k - 1);
  if (i < L.nx - 1) t = t + (R)L.Cx[c + sx] * xo(c + sx, i + 1, j, k);
  if (j < L.ny - 1) t = t + (R)L.Cy[c + sy] * xo(c + sy, i, j + 1, k);
  if (k < L.nz - 1) t = t + (R)L.Cz[c + 1] * xo(c + 1, i, j, k + 1);
  return t;
}

// ---- hierarchy ----------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_mg_level0(int nx, int ny, int nz, const S* __restrict__ adiag,
                                                   const S* __restrict__ aL, const S* __restrict__ aD,
                                                   const S* __restrict__ aB, const uint8_t* __restrict__ act,
                                                   S* __restrict__ D, S* __restrict__ Cx, S* __restrict__ Cy,
                                                   S* __restrict__ Cz, S* __restrict__ W) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= (int64_t)nx * ny * nz) return;
  int i, j, k;
  decode3(c, ny, nz, i, j, k);
  const int64_t sx = (int64_t)ny * nz;
  S d = S(0), cx = S(0), cy = S(0), cz = S(0), w = S(0);
  if (act[c]) {
    d = adiag[c];
    if (i > 0 && act[c - sx]) cx = aL[c];
    if (j > 0 && act[c - nz]) cy = aD[c];
    if (k > 0 && act[c - 1]) cz = aB[c];
    w = d > S(0) ? (S)(MG_OMEGA / (double)d) : S(0);
  }
  D[c] = d;
  Cx[c] = cx;
  Cy[c] = cy;
  Cz[c] = cz;
  W[c] = w;
}

template <typename S>
__global__ __launch_bounds__(256) void k_mg_coarsen(Lv<S> F, int cx_, int cy_, int cz_, S* __restrict__ D,
                                                    S* __restrict__ Cx, S* __restrict__ Cy, S* __restrict__ Cz,
                                                    S* __restrict__ W) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= (int64_t)cx_ * cy_ * cz_) return;
  int I, J, K;
  decode3(c, cy_, cz_, I, J, K);
  S d = S(0), sx = S(0), sy = S(0), sz = S(0);
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++)
      for (int dk = 0; dk < 2; dk++) {
        const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
        if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
        const int64_t f = ((int64_t)i * F.ny + j) * F.nz + k;
        d = d + F.D[f];
        if (di) d = d + S(2) * F.Cx[f];
        if (dj) d = d + S(2) * F.Cy[f];
        if (dk) d = d + S(2) * F.Cz[f];
        if (!di) sx = sx + F.Cx[f];
        if (!dj) sy = sy + F.Cy[f];
        if (!dk) sz = sz + F.Cz[f];
      }
  D[c] = d;
  Cx[c] = sx;
  Cy[c] = sy;
  Cz[c] = sz;
  W[c] = d > S(0) ? (S)(MG_OMEGA / (double)d) : S(0);
}

// ---- V-cycle ------------------------------------------------------------------------------------
// B is the type of the level's right-hand side: FP64 on level 0 (the CG residual), S below.
// sc (may be null): the kernels of a CG iteration return at once after the solve has stopped.
// A cell with W = 0 is never written: the level's x and xs are zeroed when the hierarchy is
// built, and stay 0 there for the whole solve.

// x = two Jacobi sweeps from zero at a live cell: x1 = W b, x = x1 + W (b - A x1).
template <typename S, typename B>
__device__ __forceinline__ S pre_cell(const Lv<S>& L, const B* __restrict__ b, int64_t c, int i, int j, int k) {
  auto xo = [&](int64_t q, int, int, int) { return L.W[q] * (S)b[q]; };
  return xo(c, i, j, k) + L.W[c] * ((S)b[c] - row<S>(L, c, i, j, k, xo));
}
// b(f) - (A x)(f) summed over the live children f of aggregate (I, J, K), in lexicographic order.
template <typename S, typename B>
__device__ __forceinline__ S restrict_cell(const Lv<S>& L, const B* __restrict__ b, const S* __restrict__ x, int I,
                                           int J, int K) {
  auto xo = [&](int64_t q, int, int, int) { return x[q]; };
  S acc = S(0);
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++)
      for (int dk = 0; dk < 2; dk++) {
        const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
        if (i >= L.nx || j >= L.ny || k >= L.nz) continue;
        const int64_t f = ((int64_t)i * L.ny + j) * L.nz + k;
        if (L.W[f] > S(0)) acc = acc + ((S)b[f] - row<S>(L, f, i, j, k, xo));
      }
  return acc;
}
// The coarse correction x' = x + alpha xc(parent) on live cells, then one sweep at a live cell.
template <typename S, typename B>
__device__ __forceinline__ S post1_cell(const Lv<S>& L, const B* __restrict__ b, const S* __restrict__ x,
                                        const S* __restrict__ xc, int cny, int cnz, S alpha, int64_t c, int i, int j,
                                        int k) {
  auto xo = [&](int64_t q, int qi, int qj, int qk) {
    return L.W[q] > S(0) ? x[q] + alpha * xc[((int64_t)(qi >> 1) * cny + (qj >> 1)) * cnz + (qk >> 1)] : x[q];
  };
  return xo(c, i, j, k) + L.W[c] * ((S)b[c] - row<S>(L, c, i, j, k, xo));
}
// One sweep x = xs + W (b - A xs) at a live cell.
template <typename S, typename B>
__device__ __forceinline__ S sweep_cell(const Lv<S>& L, const B* __restrict__ b, const S* __restrict__ xs, int64_t c,
                                        int i, int j, int k) {
  auto xo = [&](int64_t q, int, int, int) { return xs[q]; };
  return xs[c] + L.W[c] * ((S)b[c] - row<S>(L, c, i, j, k, xo));
}

// The sweeps without a dot product take one cell per thread: the coarse levels are small, and
// sixteen cells per thread would leave most of the chip idle while each thread waits on its
// loads in turn.
template <typename S, typename B>
__global__ __launch_bounds__(256) void k_mg_pre(Lv<S> L, const B* __restrict__ b, S* __restrict__ x,
                                                const PcgScalars* sc) {
  if (sc && sc->converged != STOP_RUN) return;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= (int64_t)L.nx * L.ny * L.nz || !(L.W[c] > S(0))) return;
  int i, j, k;
  decode3(c, L.ny, L.nz, i, j, k);
  x[c] = pre_cell<S, B>(L, b, c, i, j, k);
}

template <typename S, typename B>
__global__ __launch_bounds__(256) void k_mg_restrict(Lv<S> L, const B* __restrict__ b, const S* __restrict__ x,
                                                     int cnx, int cny, int cnz, S* __restrict__ bc,
                                                     const PcgScalars* sc) {
  if (sc && sc->converged != STOP_RUN) return;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= (int64_t)cnx * cny * cnz) return;
  int I, J, K;
  decode3(c, cny, cnz, I, J, K);
  bc[c] = restrict_cell<S, B>(L, b, x, I, J, K);
}

template <typename S, typename B>
__global__ __launch_bounds__(256) void k_mg_post1(Lv<S> L, const B* __restrict__ b, const S* __restrict__ x,
                                                  const S* __restrict__ xc, int cny, int cnz, S alpha,
                                                  S* __restrict__ xs, const PcgScalars* sc) {
  if (sc && sc->converged != STOP_RUN) return;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= (int64_t)L.nx * L.ny * L.nz || !(L.W[c] > S(0))) return;
  int i, j, k;
  decode3(c, L.ny, L.nz, i, j, k);
  xs[c] = post1_cell<S, B>(L, b, x, xc, cny, cnz, alpha, c, i, j, k);
}

// One sweep from xs into x.
template <typename S, typename B>
__global__ __launch_bounds__(256) void k_mg_sweep(Lv<S> L, const B* __restrict__ b, const S* __restrict__ xs,
                                                  S* __restrict__ x, const PcgScalars* sc) {
  if (sc && sc->converged != STOP_RUN) return;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= (int64_t)L.nx * L.ny * L.nz || !(L.W[c] > S(0))) return;
  int i, j, k;
  decode3(c, L.ny, L.nz, i, j, k);
  x[c] = sweep_cell<S, B>(L, b, xs, c, i, j, k);
}

// The sweep that ends a CG cycle on level 0, from xs into x = z, with z.r over the chunk, reduced by
// the last block into sigma (FIRST: the initial sigma; else beta = sigma_new / sigma,
// sigma = sigma_new). Chunks of 4096 cells per block: the fixed reduction tree.
template <typename S, typename B, bool DOT, bool FIRST>
__global__ __launch_bounds__(256) void k_mg_post2(Lv<S> L, const B* __restrict__ b, const S* __restrict__ xs,
                                                  S* __restrict__ x, double* __restrict__ part, int nchunks,
                                                  PcgScalars* sc) {
  __shared__ double sm[4];
  if (sc && sc->converged != STOP_RUN) return;
  const int64_t n = (int64_t)L.nx * L.ny * L.nz;
  double acc = 0.0;
#pragma unroll 1
  for (int m = 0; m < 16; m++) {
    const int64_t c = (int64_t)blockIdx.x * MCH + m * 256 + threadIdx.x;
    if (c >= n) break;
    if (!(L.W[c] > S(0))) continue;  // z = 0 there, and so is its product with r
    int i, j, k;
    decode3(c, L.ny, L.nz, i, j, k);
    const S out = sweep_cell<S, B>(L, b, xs, c, i, j, k);
    x[c] = out;
    if (DOT) acc += (double)out * (double)b[c];
  }
  if (!DOT) return;
  const double bs = block_sum256(acc, sm);
  if (threadIdx.x == 0) xstore(part + blockIdx.x, bs);
  if (blockIdx.x == gridDim.x - 1) {
    gather_wait256(part, nchunks, sc);
    const double sig = gather_tree256(part, nchunks, sm);
    if (threadIdx.x == 0) {
      if (!FIRST) sc->beta = sig / sc->sigma;
      sc->sigma = sig;
    }
  }
}

// The levels from l0 down to the single cell and back up, in one workgroup (the small levels,
// where separate launches would cost more than the work): the same cell functions, the passes
// separated by workgroup barriers.
constexpr int MG_TAIL_CELLS = 4096;  // levels at or below this size run in the tail kernel
constexpr int MG_MAX_LEVELS = 16;
template <typename S>
struct TailLevels {
  Lv<S> L[MG_MAX_LEVELS];
  S* b[MG_MAX_LEVELS];
  S* x[MG_MAX_LEVELS];
  S* xs[MG_MAX_LEVELS];
};
template <typename S>
__global__ __launch_bounds__(1024) void k_mg_tail(TailLevels<S> T, int l0, int nl, S alpha, const PcgScalars* sc) {
  if (sc && sc->converged != STOP_RUN) return;
  const int tid = threadIdx.x;
  for (int l = l0; l < nl; l++) {
    const Lv<S>& L = T.L[l];
    const int n = L.nx * L.ny * L.nz;
    for (int c = tid; c < n; c += 1024) {
      if (!(L.W[c] > S(0))) continue;
      int i, j, k;
      decode3(c, L.ny, L.nz, i, j, k);
      T.x[l][c] = pre_cell<S, S>(L, T.b[l], c, i, j, k);
    }
    __syncthreads();
    if (l + 1 < nl) {
      const Lv<S>& C = T.L[l + 1];
      const int nc = C.nx * C.ny * C.nz;
      for (int c = tid; c < nc; c += 1024) {
        int I, J, K;
        decode3(c, C.ny, C.nz, I, J, K);
        T.b[l + 1][c] = restrict_cell<S, S>(L, T.b[l], T.x[l], I, J, K);
      }
      __syncthreads();
    }
  }
  auto sweep = [&](int l, const S* from, S* to) {
    const Lv<S>& L = T.L[l];
    const int n = L.nx * L.ny * L.nz;
    for (int c = tid; c < n; c += 1024) {
      if (!(L.W[c] > S(0))) continue;
      int i, j, k;
      decode3(c, L.ny, L.nz, i, j, k);
      to[c] = sweep_cell<S, S>(L, T.b[l], from, c, i, j, k);
    }
    __syncthreads();
  };
  sweep(nl - 1, T.x[nl - 1], T.xs[nl - 1]);
  sweep(nl - 1, T.xs[nl - 1], T.x[nl - 1]);
  for (int l = nl - 2; l >= l0; l--) {
    const Lv<S>& L = T.L[l];
    const Lv<S>& C = T.L[l + 1];
    const int n = L.nx * L.ny * L.nz;
    for (int c = tid; c < n; c += 1024) {
      if (!(L.W[c] > S(0))) continue;
      int i, j, k;
      decode3(c, L.ny, L.nz, i, j, k);
      T.xs[l][c] = post1_cell<S, S>(L, T.b[l], T.x[l], T.x[l + 1], C.ny, C.nz, alpha, c, i, j, k);
    }
    __syncthreads();
    sweep(l, T.xs[l], T.x[l]);
  }
}

// ---- CG kernels ---------------------------------------------------------------------------------
// r = rhs on active cells (0 elsewhere), x = 0; ||r||_inf, ||r||_2^2 and the stopping test at
// iteration 0.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_init(int64_t n, const uint8_t* __restrict__ act,
                                                 const T* __restrict__ rhs, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ part,
                                                 double* __restrict__ part2, int nchunks, PcgScalars* sc,
                                                 int tol_mode, double tol) {
  __shared__ double sm[4];
  double mx = 0.0, ss = 0.0;
  int na = 0;
  for (int m = 0; m < 16; m++) {
    const int64_t c = (int64_t)blockIdx.x * MCH + m * 256 + threadIdx.x;
    if (c >= n) break;
    const double v = act[c] ? (double)rhs[c] : 0.0;
    r[c] = v;
    x[c] = 0.0;
    ss += v * v;
    const double av = act[c] ? fabs(v) : 0.0;
    mx = (mx < av) ? av : mx;
    na += act[c] ? 1 : 0;
  }
  const double bm = block_max256(mx, sm);
  const double bs = block_sum256(ss, sm);
  const double bn = block_sum256((double)na, sm);
  if (threadIdx.x == 0) {
    xstore(part + blockIdx.x, bs);
    xstore(part2 + blockIdx.x, bm);
    atomicAdd(&sc->n_active, (unsigned long long)bn);
  }
  if (blockIdx.x == gridDim.x - 1) {
    gather_wait256(part, nchunks, sc);
    gather_wait256(part2, nchunks, sc);
    const double tot = gather_tree256(part, nchunks, sm);
    const double mxx = gather_max256(part2, nchunks, sm);
    if (threadIdx.x == 0) {
      sc->bnorm2sq = tot;
      sc->rnorm2sq = tot;
      sc->residual = mxx;
      sc->status = FS_OK;
      sc->conv_iter = 0;
      sc->sigma = sc->alpha = sc->beta = sc->sdotz = 0.0;
      const bool conv = tol_mode == FS_TOL_REL_L2 ? (sqrt(tot) <= tol * sqrt(tot)) : (mxx < tol);
      sc->converged = conv ? STOP_CONVERGED : STOP_RUN;
    }
  }
}

// s = z (iteration 1) or z + beta s_old, recomputed for the six neighbours; t = A s (the assembled
// system in FP64 on active cells); s.t reduced by the last block into alpha. Inactive cells are not
// written: s and t are zeroed when the solve starts and stay 0 there.
template <typename S>
__global__ __launch_bounds__(256) void k_mg_spmv(Lv<S> L, const uint8_t* __restrict__ act, const S* __restrict__ z,
                                                 const double* __restrict__ s_old, double* __restrict__ s_new,
                                                 double* __restrict__ t, double* __restrict__ part, int nchunks,
                                                 PcgScalars* sc, int iter) {
  __shared__ double sm[4];
  if (sc->converged != STOP_RUN) return;
  const int64_t n = (int64_t)L.nx * L.ny * L.nz;
  const bool first = iter == 1;
  const double beta = sc->beta;
  auto so = [&](int64_t q, int, int, int) { return first ? (double)z[q] : (double)z[q] + beta * s_old[q]; };
  double acc = 0.0;
#pragma unroll 1
  for (int m = 0; m < 16; m++) {
    const int64_t c = (int64_t)blockIdx.x * MCH + m * 256 + threadIdx.x;
    if (c >= n) break;
    if (!act[c]) continue;
    int i, j, k;
    decode3(c, L.ny, L.nz, i, j, k);
    const double sv = so(c, i, j, k);
    const double tv = row<double>(L, c, i, j, k, so);
    s_new[c] = sv;
    t[c] = tv;
    acc += sv * tv;
  }
  const double bs = block_sum256(acc, sm);
  if (threadIdx.x == 0) xstore(part + blockIdx.x, bs);
  if (blockIdx.x == gridDim.x - 1) {
    gather_wait256(part, nchunks, sc);
    const double sdott = gather_tree256(part, nchunks, sm);
    if (threadIdx.x == 0) {
      sc->sdotz = sdott;
      if (sdott == 0.0) {
        sc->converged = STOP_ABORT;
        sc->status = FS_BREAKDOWN;
        sc->conv_iter = iter;
      } else {
        sc->alpha = sc->sigma / sdott;
      }
    }
  }
}

// x += alpha s, r += -alpha t on active cells (saxpy, util.cc:394-402); ||r||_inf, ||r||_2^2 and the
// stopping test of iteration iter.
__global__ __launch_bounds__(256) void k_mg_update(int64_t n, const uint8_t* __restrict__ act,
                                                   double* __restrict__ x, const double* __restrict__ s,
                                                   double* __restrict__ r, const double* __restrict__ t,
                                                   double* __restrict__ part, double* __restrict__ part2,
                                                   int nchunks, PcgScalars* sc, int iter, int tol_mode, double tol) {
  __shared__ double sm[4];
  if (sc->converged != STOP_RUN) return;
  const double alpha = sc->alpha;
  double mx = 0.0, ss = 0.0;
  for (int m = 0; m < 16; m++) {
    const int64_t c = (int64_t)blockIdx.x * MCH + m * 256 + threadIdx.x;
    if (c >= n) break;
    if (!act[c]) continue;  // r = 0 there
    x[c] = x[c] + alpha * s[c];
    const double rv = r[c] + (-alpha) * t[c];
    r[c] = rv;
    const double av = fabs(rv);
    mx = (mx < av) ? av : mx;
    ss += rv * rv;
  }
  const double bm = block_max256(mx, sm);
  const double bs = block_sum256(ss, sm);
  if (threadIdx.x == 0) {
    xstore(part + blockIdx.x, bs);
    xstore(part2 + blockIdx.x, bm);
  }
  if (blockIdx.x == gridDim.x - 1) {
    gather_wait256(part, nchunks, sc);
    gather_wait256(part2, nchunks, sc);
    const double tot = gather_tree256(part, nchunks, sm);
    const double mxx = gather_max256(part2, nchunks, sm);
    if (threadIdx.x == 0) {
      sc->residual = mxx;
      sc->rnorm2sq = tot;
      const bool conv = tol_mode == FS_TOL_REL_L2 ? (sqrt(tot) <= tol * sqrt(sc->bnorm2sq)) : (mxx < tol);
      if (!(mxx <= 1.7976931348623157e308)) {
        sc->converged = STOP_ABORT;
        sc->status = FS_NONFINITE;
        sc->conv_iter = iter;
      } else if (conv) {
        sc->converged = STOP_CONVERGED;
        sc->conv_iter = iter;
      }
    }
  }
}

template <typename T>
__global__ void k_mg_store(int64_t n, const double* __restrict__ x, T* __restrict__ out) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < n) out[c] = (T)x[c];
}

// ---- host ---------------------------------------------------------------------------------------
int ensure(fs_ctx* c, DBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return FS_OK;
  fs_free(b);
  return fs_alloc(c, b, bytes);
}

template <typename S>
Lv<S> lv_of(const MgLevelBuf& B) {
  return {B.nx, B.ny, B.nz, B.D.as<S>(), B.Cx.as<S>(), B.Cy.as<S>(), B.Cz.as<S>(), B.W.as<S>()};
}

unsigned chunks_of(int64_t n) { return (unsigned)((n + MCH - 1) / MCH); }

// Levels and the hierarchy of the current assembled system.
template <typename S>
int mg_build(fs_ctx* c) {
  MgState& M = c->mg;
  std::vector<int3> dims{make_int3(c->nx, c->ny, c->nz)};
  while (dims.back().x > 1 || dims.back().y > 1 || dims.back().z > 1) {
    const int3 d = dims.back();
    dims.push_back(make_int3((d.x + 1) / 2, (d.y + 1) / 2, (d.z + 1) / 2));
  }
  M.lv.resize(dims.size());
  for (size_t l = 0; l < dims.size(); l++) {
    MgLevelBuf& B = M.lv[l];
    B.nx = dims[l].x;
    B.ny = dims[l].y;
    B.nz = dims[l].z;
    B.n = (int64_t)B.nx * B.ny * B.nz;
    const size_t bytes = (size_t)B.n * sizeof(S);
    for (DBuf* d : {&B.D, &B.Cx, &B.Cy, &B.Cz, &B.W, &B.x, &B.xs}) FS_TRY(ensure(c, *d, bytes));
    if (l > 0) FS_TRY(ensure(c, B.b, bytes));
    // the cycle never writes a cell with W = 0: its iterates stay 0 there for the solve
    FS_HIP(c, hipMemsetAsync(B.x.p, 0, bytes, c->stream));
    FS_HIP(c, hipMemsetAsync(B.xs.p, 0, bytes, c->stream));
  }
  if ((int)M.lv.size() > MG_MAX_LEVELS) {
    c->set_error("multigrid: too many levels");
    return FS_BAD_ARG;
  }
  MgLevelBuf& L0 = M.lv[0];
  k_mg_level0<S><<<fs_blocks(L0.n, 256), 256, 0, c->stream>>>(
      c->nx, c->ny, c->nz, c->adiag.as<S>(), c->an[0].as<S>(), c->an[3].as<S>(), c->an[5].as<S>(),
      c->active.as<uint8_t>(), L0.D.as<S>(), L0.Cx.as<S>(), L0.Cy.as<S>(), L0.Cz.as<S>(), L0.W.as<S>());
  for (size_t l = 1; l < M.lv.size(); l++) {
    MgLevelBuf& B = M.lv[l];
    k_mg_coarsen<S><<<fs_blocks(B.n, 256), 256, 0, c->stream>>>(lv_of<S>(M.lv[l - 1]), B.nx, B.ny, B.nz,
                                                                  B.D.as<S>(), B.Cx.as<S>(), B.Cy.as<S>(),
                                                                  B.Cz.as<S>(), B.W.as<S>());
  }
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

// One V-cycle on level 0 with right-hand side r0 (FP64). In a CG solve, sc is the solve's scalars
// and the final sweep forms sigma = z.r (FIRST: the initial sigma); z is level 0's x.
template <typename S>
int mg_cycle(fs_ctx* c, const double* r0, PcgScalars* sc, bool first) {
  MgState& M = c->mg;
  const int nl = (int)M.lv.size();
  const double na = (double)c->n_active_last;
  auto pre = [&](int l) {
    MgLevelBuf& B = M.lv[l];
    if (l == 0) k_mg_pre<S, double><<<fs_blocks(B.n, 256), 256, 0, c->stream>>>(lv_of<S>(B), r0, B.x.as<S>(), sc);
    else k_mg_pre<S, S><<<fs_blocks(B.n, 256), 256, 0, c->stream>>>(lv_of<S>(B), B.b.as<S>(), B.x.as<S>(), sc);
  };
  auto restrict_ = [&](int l) {
    MgLevelBuf &B = M.lv[l], &C = M.lv[l + 1];
    if (l == 0)
      k_mg_restrict<S, double><<<fs_blocks(C.n, 256), 256, 0, c->stream>>>(lv_of<S>(B), r0, B.x.as<S>(), C.nx, C.ny,
                                                                           C.nz, C.b.as<S>(), sc);
    else
      k_mg_restrict<S, S><<<fs_blocks(C.n, 256), 256, 0, c->stream>>>(lv_of<S>(B), B.b.as<S>(), B.x.as<S>(), C.nx,
                                                                      C.ny, C.nz, C.b.as<S>(), sc);
  };
  auto post1 = [&](int l) {
    MgLevelBuf &B = M.lv[l], &C = M.lv[l + 1];
    if (l == 0)
      k_mg_post1<S, double><<<fs_blocks(B.n, 256), 256, 0, c->stream>>>(lv_of<S>(B), r0, B.x.as<S>(), C.x.as<S>(),
                                                                         C.ny, C.nz, (S)MG_ALPHA, B.xs.as<S>(), sc);
    else
      k_mg_post1<S, S><<<fs_blocks(B.n, 256), 256, 0, c->stream>>>(lv_of<S>(B), B.b.as<S>(), B.x.as<S>(),
                                                                    C.x.as<S>(), C.ny, C.nz, (S)MG_ALPHA, B.xs.as<S>(),
                                                                    sc);
  };
  // one sweep from `from` into `to`; on level 0 of a CG cycle the sweep that ends the cycle
  // (into x) also forms sigma
  auto sweep = [&](int l, DBuf& from, DBuf& to, bool last) {
    MgLevelBuf& B = M.lv[l];
    const unsigned nb = chunks_of(B.n);
    const Lv<S> L = lv_of<S>(B);
    const unsigned nc1 = fs_blocks(B.n, 256);
    if (l > 0) {
      k_mg_sweep<S, S><<<nc1, 256, 0, c->stream>>>(L, B.b.as<S>(), from.as<S>(), to.as<S>(), sc);
    } else if (!last || !sc) {
      k_mg_sweep<S, double><<<nc1, 256, 0, c->stream>>>(L, r0, from.as<S>(), to.as<S>(), sc);
    } else if (first) {
      k_mg_post2<S, double, true, true><<<nb, 256, 0, c->stream>>>(L, r0, from.as<S>(), to.as<S>(),
                                                                    M.partials.as<double>(), (int)nb, sc);
    } else {
      k_mg_post2<S, double, true, false><<<nb, 256, 0, c->stream>>>(L, r0, from.as<S>(), to.as<S>(),
                                                                     M.partials.as<double>(), (int)nb, sc);
    }
  };
  // the levels from lt on run in one workgroup (the first level at or below MG_TAIL_CELLS cells,
  // never level 0, whose right-hand side is the FP64 residual)
  int lt = nl;
  for (int l = 1; l < nl; l++)
    if (M.lv[l].n <= MG_TAIL_CELLS) {
      lt = l;
      break;
    }
  // down
  int pf = fs_prof_begin(c, FS_KERNEL_MG_LEVEL0, (16.0 + 12.0 * sizeof(S)) * na);
  pre(0);
  if (nl > 1) restrict_(0);
  fs_prof_end(c, pf);
  // coarse levels hold about 1/7 of the active cells; 31 values per cell over the four passes
  pf = fs_prof_begin(c, FS_KERNEL_MG_COARSE, 31.0 * sizeof(S) * na / 7.0);
  for (int l = 1; l < lt; l++) {
    pre(l);
    restrict_(l);
  }
  if (lt < nl) {
    TailLevels<S> T{};
    for (int l = lt; l < nl; l++) {
      MgLevelBuf& B = M.lv[l];
      T.L[l] = lv_of<S>(B);
      T.b[l] = B.b.as<S>();
      T.x[l] = B.x.as<S>();
      T.xs[l] = B.xs.as<S>();
    }
    k_mg_tail<S><<<1, 1024, 0, c->stream>>>(T, lt, nl, (S)MG_ALPHA, sc);
  }
  // up: correction and post-smoothing
  for (int l = lt - 1; l >= 1; l--) {
    post1(l);
    sweep(l, M.lv[l].xs, M.lv[l].x, false);
  }
  fs_prof_end(c, pf);
  pf = fs_prof_begin(c, FS_KERNEL_MG_LEVEL0, (16.0 + 14.0 * sizeof(S)) * na);
  if (nl > 1) {
    post1(0);
  } else {  // a single-cell grid is its own coarsest level
    sweep(0, M.lv[0].x, M.lv[0].xs, false);
  }
  sweep(0, M.lv[0].xs, M.lv[0].x, true);
  fs_prof_end(c, pf);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

int mg_vectors(fs_ctx* c) {
  MgState& M = c->mg;
  const int64_t n = c->n_cell;
  for (DBuf* d : {&M.x, &M.r, &M.s0, &M.s1, &M.t}) FS_TRY(ensure(c, *d, (size_t)n * sizeof(double)));
  const int nc = (int)chunks_of(n);
  if (!M.scal.p) {
    FS_TRY(fs_alloc(c, M.scal, sizeof(PcgScalars)));
    FS_HIP(c, hipMemsetAsync(M.scal.p, 0, sizeof(PcgScalars), c->stream));
    FS_HIP(c, hipHostMalloc((void**)&M.host_scal, sizeof(PcgScalars), hipHostMallocDefault));
  }
  if (M.nchunks < nc || !M.partials.p) {
    FS_TRY(ensure(c, M.partials, (size_t)nc * sizeof(double)));
    FS_TRY(ensure(c, M.partials2, (size_t)nc * sizeof(double)));
    // reduction entries start unset; every reduction returns the entries it used to the sentinel
    k_fill_sentinel<<<fs_blocks(nc, 256), 256, 0, c->stream>>>(nc, M.partials.as<double>());
    k_fill_sentinel<<<fs_blocks(nc, 256), 256, 0, c->stream>>>(nc, M.partials2.as<double>());
    FS_HIP(c, hipGetLastError());
    M.nchunks = nc;
  }
  return FS_OK;
}

int mg_read_scalars(fs_ctx* c) {
  MgState& M = c->mg;
  FS_HIP(c, hipMemcpyAsync(M.host_scal, M.scal.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  return FS_OK;
}

template <typename S>
int mg_solve_t(fs_ctx* c, fs_stats* st) {
  MgState& M = c->mg;
  const int64_t n = c->n_cell;
  FS_TRY(mg_vectors(c));
  FS_TRY(mg_build<S>(c));
  const unsigned nb = chunks_of(n);
  const int nc = (int)nb;
  PcgScalars* sc = M.scal.as<PcgScalars>();
  PcgScalars* hs = M.host_scal;
  const int tmode = c->cfg.tol_mode;
  const double tol = c->cfg.cg_tolerance;
  const uint8_t* act = c->active.as<uint8_t>();
  FS_HIP(c, hipMemsetAsync(M.scal.p, 0, sizeof(PcgScalars), c->stream));
  // s and t are written on active cells only
  for (DBuf* d : {&M.s0, &M.s1, &M.t}) FS_HIP(c, hipMemsetAsync(d->p, 0, (size_t)n * sizeof(double), c->stream));
  k_mg_init<S><<<nb, 256, 0, c->stream>>>(n, act, c->rhs.as<S>(), M.x.as<double>(), M.r.as<double>(),
                                           M.partials.as<double>(), M.partials2.as<double>(), nc, sc, tmode, tol);
  FS_HIP(c, hipGetLastError());
  FS_TRY(mg_read_scalars(c));
  c->n_active_last = (int64_t)hs->n_active;
  int iters = 0, status = FS_OK;
  if (hs->converged == STOP_RUN) {
    MgLevelBuf& L0 = M.lv[0];
    const Lv<S> L = lv_of<S>(L0);
    FS_TRY(mg_cycle<S>(c, M.r.as<double>(), sc, true));
    const int maxit = c->cfg.cg_max_iterations;
    const int chk = std::max(1, c->cfg.check_interval);
    DBuf* sb[2] = {&M.s0, &M.s1};
    int it = 1;
    bool stopped = false;
    while (it < maxit && !stopped) {
      for (int m = 0; m < chk && it < maxit; m++, it++) {
        const double* s_old = sb[(it + 1) & 1]->as<double>();
        double* s_new = sb[it & 1]->as<double>();
        int pf = fs_prof_begin(c, FS_KERNEL_SPMV, (24.0 + 5.0 * sizeof(S)) * c->n_active_last);
        k_mg_spmv<S><<<nb, 256, 0, c->stream>>>(L, act, L0.x.as<S>(), s_old, s_new, M.t.as<double>(),
                                                M.partials.as<double>(), nc, sc, it);
        fs_prof_end(c, pf);
        pf = fs_prof_begin(c, FS_KERNEL_PCG_UPDATE, 48.0 * c->n_active_last);
        k_mg_update<<<nb, 256, 0, c->stream>>>(n, act, M.x.as<double>(), s_new, M.r.as<double>(), M.t.as<double>(),
                                              M.partials.as<double>(), M.partials2.as<double>(), nc, sc, it, tmode,
                                              tol);
        fs_prof_end(c, pf);
        FS_TRY(mg_cycle<S>(c, M.r.as<double>(), sc, false));
      }
      FS_HIP(c, hipGetLastError());
      FS_TRY(mg_read_scalars(c));
      fs_prof_flush(c);
      if (hs->wait_timeouts != 0) break;
      if (hs->converged != STOP_RUN) stopped = true;
    }
    if (hs->wait_timeouts != 0) {
      c->set_error("multigrid CG: a reduction wait exceeded its bound");
      return FS_DEVICE_ERROR;
    }
    if (hs->converged == STOP_CONVERGED || hs->converged == STOP_ABORT) iters = hs->conv_iter;
    else iters = maxit;
    if (hs->converged == STOP_ABORT) status = hs->status;
    else if (hs->converged == STOP_RUN) status = FS_NOT_CONVERGED;
  }
  if (c->mixed)
    k_mg_store<float><<<fs_blocks(n, 256), 256, 0, c->stream>>>(n, M.x.as<double>(), c->pg.as<float>());
  else
    k_mg_store<double><<<fs_blocks(n, 256), 256, 0, c->stream>>>(n, M.x.as<double>(), c->pg.as<double>());
  FS_HIP(c, hipGetLastError());
  FS_HIP(c, hipStreamSynchronize(c->stream));
  if (st) {
    st->cg_iterations = iters;
    st->residual_inf = hs->residual;
    st->rhs_norm2 = std::sqrt(hs->bnorm2sq);
    st->residual_rel = hs->bnorm2sq > 0 ? std::sqrt(hs->rnorm2sq / hs->bnorm2sq) : 0.0;
    st->n_active = (int32_t)c->n_active_last;
    st->refinements = 0;
  }
  return status;
}

template <typename S>
int mg_apply_t(fs_ctx* c, const double* r, double* z) {
  MgState& M = c->mg;
  const int64_t n = c->n_cell;
  FS_TRY(mg_vectors(c));
  FS_TRY(mg_build<S>(c));
  // the cycle's right-hand side: r on active cells, 0 elsewhere
  std::vector<uint8_t> act(n);
  FS_HIP(c, hipMemcpyAsync(act.data(), c->active.p, n, hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> host(n);
  for (int64_t q = 0; q < n; q++) host[q] = act[q] ? r[q] : 0.0;
  FS_HIP(c, hipMemcpyAsync(M.r.p, host.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FS_TRY(mg_cycle<S>(c, M.r.as<double>(), nullptr, false));
  return fs_download_real(c, M.lv[0].x, z, (size_t)n);
}

}  // namespace

int fs_mg_solve(fs_ctx* c, fs_stats* st) { return c->mixed ? mg_solve_t<float>(c, st) : mg_solve_t<double>(c, st); }

int fs_mg_apply_dense(fs_ctx* c, const double* r, double* z) {
  return c->mixed ? mg_apply_t<float>(c, r, z) : mg_apply_t<double>(c, r, z);
}

void fs_mg_release(fs_ctx* c) {
  MgState& M = c->mg;
  for (MgLevelBuf& B : M.lv)
    for (DBuf* d : {&B.D, &B.Cx, &B.Cy, &B.Cz, &B.W, &B.b, &B.x, &B.xs}) fs_free(*d);
  M.lv.clear();
  for (DBuf* d : {&M.x, &M.r, &M.s0, &M.s1, &M.t, &M.partials, &M.partials2, &M.scal}) fs_free(*d);
  if (M.host_scal) (void)hipHostFree(M.host_scal);
  M.host_scal = nullptr;
  M.nchunks = 0;
}
