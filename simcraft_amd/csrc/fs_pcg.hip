// fs_pcg.hip - MIC(0)-preconditioned CG pressure solve for gfx950
// (CgSolver3D::solve, FluidSim/src/cpu/project-solver.cc:308-346, with
// ModifiedIncompleteCholesky3D::precond :230-306 and applyCompressedMatrix :16-39).
//
// Layout ("column bricks"): the active cells are covered by 8x8 columns in (i,j) that span
// all k. A live brick (one with an active cell) owns S = nz+14 slots of 64 lanes; cell
// (i,j,k) sits in lane a*8+b (a=i%8, b=j%8) of slot k+a+b. The skew a+b turns the triangular
// sweeps into one wavefront per brick: at step s every lane of the wave handles slot s, and
// its (i-1) and (j-1) neighbours were finished by lanes a-1 / b-1 at step s-1, so they arrive
// by a lane shuffle. Bricks hand their boundary rows to the next brick through small exchange
// arrays whose entries double as ready flags (see the boundary exchange below); bricks are
// scheduled by ticket in dependency order, so every wait is on a brick that is already
// running. Each slot reads/writes one contiguous 512-byte row per array, and a brick streams
// its slots in chunks of U with the next chunk's loads in flight.
//
// Exactness: every cell does the reference's arithmetic in the reference's order. The MIC
// sweep coefficients are the products the reference forms first (AL(c)*E(i-1) and
// AL(i+1)*E(c)); inactive neighbours contribute +0 * +0. Dot products use one fixed tree.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "fs_ctx.h"
#include "fs_reduce.h"

namespace {

constexpr int CH = 4096;  // elements per chunk = 256 threads x 16
constexpr int NB_SELF = 1, NB_L = 2, NB_D = 4, NB_B = 8, NB_R = 16, NB_U = 32, NB_F = 64;
// factor-sweep flags
constexpr int TF_DO = 1, TF_GL = 2, TF_L1 = 4, TF_L2 = 8, TF_GD = 16, TF_D1 = 32, TF_D2 = 64, TF_GB = 128,
              TF_B1 = 256, TF_B2 = 512;

struct Lay {
  int nx, ny, nz, nbi, nbj, S, nlive;
  int64_t ne;
  const int* live;   // brick (I*nbj+J) -> live index or -1
  const int2* bij;   // live index -> (I, J)
};

__device__ inline int64_t eidx(const Lay& L, int B, int s, int lane) { return ((int64_t)B * L.S + s) * 64 + lane; }

// neighbour element of a valid cell; dir 0:-x 1:+x 2:-y 3:+y. -1 when outside the grid or in a
// brick that holds no active cell.
__device__ inline int64_t nb_elem(const Lay& L, int B, int s, int a, int b, int dir) {
  int2 ij = L.bij[B];
  if (dir == 0) {
    if (a > 0) return eidx(L, B, s - 1, (a - 1) * 8 + b);
    if (ij.x == 0) return -1;
    int B2 = L.live[(ij.x - 1) * L.nbj + ij.y];
    return B2 < 0 ? -1 : eidx(L, B2, s + 7, 56 + b);
  } else if (dir == 1) {
    if (8 * ij.x + a + 1 >= L.nx) return -1;
    if (a < 7) return eidx(L, B, s + 1, (a + 1) * 8 + b);
    int B2 = L.live[(ij.x + 1) * L.nbj + ij.y];
    return B2 < 0 ? -1 : eidx(L, B2, s - 7, b);
  } else if (dir == 2) {
    if (b > 0) return eidx(L, B, s - 1, a * 8 + b - 1);
    if (ij.y == 0) return -1;
    int B2 = L.live[ij.x * L.nbj + ij.y - 1];
    return B2 < 0 ? -1 : eidx(L, B2, s + 7, a * 8 + 7);
  } else {
    if (8 * ij.y + b + 1 >= L.ny) return -1;
    if (b < 7) return eidx(L, B, s + 1, a * 8 + b + 1);
    int B2 = L.live[ij.x * L.nbj + ij.y + 1];
    return B2 < 0 ? -1 : eidx(L, B2, s - 7, a * 8);
  }
}

// Element -> (brick, slot, lane, cell). Element counts stay below 2^31 (checked on the host),
// so the division is 32-bit.
__device__ inline bool decode(const Lay& L, int64_t e, int& B, int& s, int& a, int& b, int& i, int& j, int& k) {
  const unsigned ue = (unsigned)e;
  int lane = (int)(ue & 63u);
  const unsigned t = ue >> 6;
  B = (int)(t / (unsigned)L.S);
  s = (int)(t - (unsigned)B * (unsigned)L.S);
  a = lane >> 3;
  b = lane & 7;
  int2 ij = L.bij[B];
  i = 8 * ij.x + a;
  j = 8 * ij.y + b;
  k = s - a - b;
  return i < L.nx && j < L.ny && k >= 0 && k < L.nz;
}

constexpr int U = 8;  // slots per chunk; S is a multiple of U

__device__ inline int take_ticket(uint64_t* ctr, uint64_t base, const int* order, int n) {
  __shared__ int sB;
  if (threadIdx.x == 0) {
    uint64_t t = atomicAdd((unsigned long long*)ctr, 1ull) - base;
    // a ticket outside [0, n) can only follow a failed launch; clamp it so every access
    // stays in bounds (the exchange polls then time out and the solve reports an error)
    sB = order[t < (uint64_t)n ? t : (uint64_t)(n - 1)];
  }
  __syncthreads();
  return sB;
}

// Exchange entry of brick B, face f (0: i-face, 1: j-face), slot s, index w (0..7).
__device__ inline double* xent(double* X, const Lay& L, int f, int B, int s, int w) {
  return X + (((size_t)f * L.nlive + B) * L.S + s) * 8 + w;
}

// ---- setup kernels ------------------------------------------------------------------------------
__global__ void k_brick_flags(int nx, int ny, int nz, const uint8_t* __restrict__ active, int* __restrict__ flag,
                              int nbj, int* __restrict__ nact) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nx * ny) return;
  int i = row / ny, j = row % ny;
  const uint8_t* a = active + (size_t)row * nz;
  int cnt = 0;
  for (int k = 0; k < nz; k++) cnt += a[k] != 0;
  if (cnt) {
    atomicOr(&flag[(i / 8) * nbj + (j / 8)], 1);
    atomicAdd(nact, cnt);
  }
}

// Pack the dense system into the brick layout; neighbour-activity masks; factor flags and
// tau numerators (tau*A(c))*A(diagonal neighbour) of ModifiedIncompleteCholesky3D::precond.
template <typename TI>
__global__ void k_pack(Lay L, double tau, const TI* __restrict__ Adiag, const TI* __restrict__ AL,
                       const TI* __restrict__ AD, const TI* __restrict__ AB, const TI* __restrict__ rhs,
                       const uint8_t* __restrict__ act, double* __restrict__ padiag, double* __restrict__ pal,
                       double* __restrict__ pad, double* __restrict__ pab, double* __restrict__ prhs,
                       uint8_t* __restrict__ nbm, uint16_t* __restrict__ tfl, double* __restrict__ tnum,
                       double* __restrict__ x, double* __restrict__ q, double* __restrict__ z) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= L.ne) return;
  int B, s, a, b, i, j, k;
  bool ok = decode(L, e, B, s, a, b, i, j, k);
  double vd = 0, vl = 0, vdn = 0, vb = 0, vr = 0;
  uint8_t m = 0;
  uint16_t tf = 0;
  double t6[6] = {0, 0, 0, 0, 0, 0};
  if (ok) {
    const int ny = L.ny, nz = L.nz, nx = L.nx;
    const int sx = ny * nz, sy = nz;
    int c = k + nz * (j + ny * i);
    vd = Adiag[c];
    vl = AL[c];
    vdn = AD[c];
    vb = AB[c];
    vr = rhs[c];
    if (act[c]) {
      m = NB_SELF;
      if (i > 0 && act[c - sx]) m |= NB_L;
      if (j > 0 && act[c - sy]) m |= NB_D;
      if (k > 0 && act[c - 1]) m |= NB_B;
      if (i < nx - 1 && act[c + sx]) m |= NB_R;
      if (j < ny - 1 && act[c + sy]) m |= NB_U;
      if (k < nz - 1 && act[c + 1]) m |= NB_F;
      if (vd != 0.0) {
        tf = TF_DO;
        if (i > 0 && act[c - sx] && Adiag[c - sx] != 0.0) {
          tf |= TF_GL;
          if (j < ny - 1 && act[c - sx + sy]) {
            tf |= TF_L1;
            t6[0] = tau * vl * (double)AL[c - sx + sy];
          }
          if (k < nz - 1 && act[c - sx + 1]) {
            tf |= TF_L2;
            t6[1] = tau * vl * (double)AL[c - sx + 1];
          }
        }
        if (j > 0 && act[c - sy] && Adiag[c - sy] != 0.0) {
          tf |= TF_GD;
          if (i < nx - 1 && act[c + sx - sy]) {
            tf |= TF_D1;
            t6[2] = tau * vdn * (double)AD[c + sx - sy];
          }
          if (k < nz - 1 && act[c - sy + 1]) {
            tf |= TF_D2;
            t6[3] = tau * vdn * (double)AD[c - sy + 1];
          }
        }
        if (k > 0 && act[c - 1] && Adiag[c - 1] != 0.0) {
          tf |= TF_GB;
          if (i < nx - 1 && act[c + sx - 1]) {
            tf |= TF_B1;
            t6[4] = tau * vb * (double)AB[c + sx - 1];
          }
          if (j < ny - 1 && act[c + sy - 1]) {
            tf |= TF_B2;
            t6[5] = tau * vb * (double)AB[c + sy - 1];
          }
        }
      }
    }
  }
  padiag[e] = vd;
  pal[e] = vl;
  pad[e] = vdn;
  pab[e] = vb;
  prhs[e] = vr;
  nbm[e] = m;
  tfl[e] = tf;
#pragma unroll
  for (int u = 0; u < 6; u++) tnum[u * L.ne + e] = t6[u];
  x[e] = 0.0;
  q[e] = 0.0;
  z[e] = 0.0;
}

// E factor sweep (project-solver.cc:237-274), one wave per brick, tickets in dependency order.
// Chunks of UF slots: the inputs of chunk c + 1 (flags, diagonal, three off-diagonals and the
// six tau terms) are loaded while chunk c runs, so the per-slot chain (two divisions per
// neighbour term, a square root and a reciprocal) never waits on memory.
__global__ __launch_bounds__(64) void k_factor_sweep(Lay L, double sigma, const double* __restrict__ adiag,
                                                     const double* __restrict__ al, const double* __restrict__ ad,
                                                     const double* __restrict__ ab, const uint16_t* __restrict__ tfl,
                                                     const double* __restrict__ tnum, double* __restrict__ E,
                                                     double* X, uint64_t* ticket, uint64_t tbase,
                                                     const int* __restrict__ order, PcgScalars* sc) {
  constexpr int UF = 4;
  static_assert(U % UF == 0, "the slot count is a multiple of U");
  const int lane = threadIdx.x, a = lane >> 3, b = lane & 7;
  const int B = take_ticket(ticket, tbase, order, L.nlive);
  const int2 ij = L.bij[B];
  const int BL = ij.x > 0 ? L.live[(ij.x - 1) * L.nbj + ij.y] : -1;
  const int BD = ij.y > 0 ? L.live[ij.x * L.nbj + ij.y - 1] : -1;
  const int S = L.S;
  const bool cL = a == 0 && BL >= 0, cD = b == 0 && BD >= 0;
  const int64_t base = (int64_t)B * S * 64 + lane;
  struct In {
    int tf[UF];
    double dg[UF], vl[UF], vd[UF], vb[UF], tn[6][UF];
  };
  auto load = [&](In& in, int s0) {
#pragma unroll
    for (int u = 0; u < UF; u++) {
      const int64_t e = base + (int64_t)(s0 + u) * 64;
      in.tf[u] = tfl[e];
      in.dg[u] = adiag[e];
      in.vl[u] = al[e];
      in.vd[u] = ad[e];
      in.vb[u] = ab[e];
#pragma unroll
      for (int q = 0; q < 6; q++) in.tn[q][u] = tnum[q * L.ne + e];
    }
  };
  In cur, nxt;
  load(cur, 0);
  double eprev = 0.0;
  for (int s0 = 0; s0 < S; s0 += UF) {
    if (s0 + UF < S) load(nxt, s0 + UF);
    double xl[UF], xd[UF];
#pragma unroll
    for (int u = 0; u < UF; u++) {
      const int s = s0 + u;
      xl[u] = (cL && s + 7 < S) ? xload(xent(X, L, 0, BL, s + 7, b)) : 0.0;
      xd[u] = (cD && s + 7 < S) ? xload(xent(X, L, 1, BD, s + 7, a)) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UF; u++) {
      const int s = s0 + u;
      if (cL && s + 7 < S) xl[u] = xpoll(xent(X, L, 0, BL, s + 7, b), xl[u], sc);
      if (cD && s + 7 < S) xd[u] = xpoll(xent(X, L, 1, BD, s + 7, a), xd[u], sc);
    }
#pragma unroll
    for (int u = 0; u < UF; u++) {
      const int s = s0 + u;
      const int64_t e = base + (int64_t)s * 64;
      double EL = __shfl_up(eprev, 8, 64);
      double ED = __shfl_up(eprev, 1, 64);
      if (a == 0) EL = xl[u];
      if (b == 0) ED = xd[u];
      double ev = 0.0;
      const int f = cur.tf[u];
      if (f & TF_DO) {
        double x = cur.dg[u];
        if (f & TF_GL) {
          double t = cur.vl[u] * EL;
          x -= t * t;
          if (f & TF_L1) x -= cur.tn[0][u] / (EL * EL);
          if (f & TF_L2) x -= cur.tn[1][u] / (EL * EL);
        }
        if (f & TF_GD) {
          double t = cur.vd[u] * ED;
          x -= t * t;
          if (f & TF_D1) x -= cur.tn[2][u] / (ED * ED);
          if (f & TF_D2) x -= cur.tn[3][u] / (ED * ED);
        }
        if (f & TF_GB) {
          double t = cur.vb[u] * eprev;
          x -= t * t;
          if (f & TF_B1) x -= cur.tn[4][u] / (eprev * eprev);
          if (f & TF_B2) x -= cur.tn[5][u] / (eprev * eprev);
        }
        if (x < sigma * cur.dg[u]) x = cur.dg[u];
        ev = 1.0 / sqrt(x);
      }
      E[e] = ev;
      if (a == 7) xstore(xent(X, L, 0, B, s, b), ev);
      if (b == 7) xstore(xent(X, L, 1, B, s, a), ev);
      eprev = ev;
    }
#pragma unroll
    for (int u = 0; u < UF; u++) {
      const int s = s0 + u;
      if (cL && s + 7 < S) xstore(xent(X, L, 0, BL, s + 7, b), xsent());
      if (cD && s + 7 < S) xstore(xent(X, L, 1, BD, s + 7, a), xsent());
    }
    cur = nxt;
  }
}

// Sweep coefficients: forward cL = AL(c)*E(i-1) etc., backward dR = AL(i+1)*E(c) etc.
// Es (optional) receives E in the storage type.
template <typename S>
__global__ void k_coef(Lay L, const double* __restrict__ al, const double* __restrict__ ad,
                       const double* __restrict__ ab, const uint8_t* __restrict__ nbm, const double* __restrict__ E,
                       S* __restrict__ cl, S* __restrict__ cd, S* __restrict__ cb, S* __restrict__ dr,
                       S* __restrict__ du, S* __restrict__ df, S* __restrict__ Es) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= L.ne) return;
  int B, s, a, b, i, j, k;
  decode(L, e, B, s, a, b, i, j, k);
  const int m = nbm[e];
  double vl = 0, vd = 0, vb = 0, vr = 0, vu = 0, vf = 0;
  if (m & NB_SELF) {
    const double ec = E[e];
    if (m & NB_L) vl = al[e] * E[nb_elem(L, B, s, a, b, 0)];
    if (m & NB_D) vd = ad[e] * E[nb_elem(L, B, s, a, b, 2)];
    if (m & NB_B) vb = ab[e] * E[e - 64];
    if (m & NB_R) vr = al[nb_elem(L, B, s, a, b, 1)] * ec;
    if (m & NB_U) vu = ad[nb_elem(L, B, s, a, b, 3)] * ec;
    if (m & NB_F) vf = ab[e + 64] * ec;
  }
  cl[e] = (S)vl;
  cd[e] = (S)vd;
  cb[e] = (S)vb;
  dr[e] = (S)vr;
  du[e] = (S)vu;
  df[e] = (S)vf;
  if (Es) Es[e] = (S)E[e];
}

// r = rhs, ||r||_inf and ||r||_2^2 (fixed tree); the last block decides "naturally converged".
template <typename S>
__global__ __launch_bounds__(256) void k_init(Lay L, const double* __restrict__ rhs, S* __restrict__ r,
                                              double* __restrict__ part, double* __restrict__ part2, int nchunks,
                                              PcgScalars* sc, int tol_mode, double tol) {
  __shared__ double sm[4];
  double mx = 0.0, ss = 0.0;
  for (int m = 0; m < 16; m++) {
    int64_t e = (int64_t)blockIdx.x * CH + m * 256 + threadIdx.x;
    double v = 0.0;
    if (e < L.ne) {
      v = rhs[e];
      r[e] = (S)v;
    }
    double av = fabs(v);
    mx = (mx < av) ? av : mx;
    ss += v * v;
  }
  double bm = block_max256(mx, sm);
  double bs = block_sum256(ss, sm);
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
      sc->bnorm2sq = tot;
      sc->rnorm2sq = tot;
      sc->residual = mxx;
      bool conv = tol_mode == FS_TOL_REL_L2 ? (sqrt(tot) <= tol * sqrt(tot)) : (mxx < tol);
      sc->converged = conv ? STOP_CONVERGED : STOP_RUN;
      sc->conv_iter = 0;
      sc->status = FS_OK;
      sc->wait_timeouts = 0;
    }
  }
}

// A block's 4096 elements are 16 groups of 256; a group is 4 slots of one brick (S*64 is a
// multiple of 512), so the brick, its neighbour bricks and the group's first slot are uniform
// and every neighbour element is a fixed offset inside the brick or a uniform base plus the
// lane's index outside it.
struct Group {
  int B, s0, BL, BR, BD, BU;
};
__device__ inline Group group_of(const Lay& L, int base) {
  Group g;
  const int SL = L.S * 64;
  g.B = base / SL;
  g.s0 = (base - g.B * SL) >> 6;
  const int2 ij = L.bij[g.B];
  g.BL = ij.x > 0 ? L.live[(ij.x - 1) * L.nbj + ij.y] : -1;
  g.BR = ij.x + 1 < L.nbi ? L.live[(ij.x + 1) * L.nbj + ij.y] : -1;
  g.BD = ij.y > 0 ? L.live[ij.x * L.nbj + ij.y - 1] : -1;
  g.BU = ij.y + 1 < L.nbj ? L.live[ij.x * L.nbj + ij.y + 1] : -1;
  return g;
}
// neighbour elements of element e = (B*S + s)*64 + a*8 + b (the neighbour must be active)
__device__ inline int nb_left(const Lay& L, const Group& g, int e, int s, int a, int b) {
  return a > 0 ? e - 72 : (g.BL * L.S + s + 7) * 64 + 56 + b;
}
__device__ inline int nb_right(const Lay& L, const Group& g, int e, int s, int a, int b) {
  return a < 7 ? e + 72 : (g.BR * L.S + s - 7) * 64 + b;
}
__device__ inline int nb_down(const Lay& L, const Group& g, int e, int s, int a, int b) {
  return b > 0 ? e - 65 : (g.BD * L.S + s + 7) * 64 + a * 8 + 7;
}
__device__ inline int nb_up(const Lay& L, const Group& g, int e, int s, int a, int b) {
  return b < 7 ? e + 65 : (g.BU * L.S + s - 7) * 64 + a * 8;
}

// Byte offsets of the four cross-brick neighbours of a brick's face lanes, relative to the
// lane's own element (see nb_left .. nb_up): a = 0 lanes read the -x brick at slot s+7, lane 56+b;
// a = 7 lanes the +x brick at slot s-7, lane b; b = 0 / 7 lanes the -y / +y brick at slot s+7 /
// s-7, lanes 8a+7 / 8a. An absent brick gives 0 (its flags are never set).
struct FaceOff {
  uint32_t dL, dR, dD, dU;
};
template <typename S>
__device__ __forceinline__ FaceOff face_off(const Lay& L, int B) {
  const int SL = L.S * 64, sz = (int)sizeof(S);
  const int2 ij = L.bij[B];
  const int BL = ij.x > 0 ? L.live[(ij.x - 1) * L.nbj + ij.y] : -1;
  const int BR = ij.x + 1 < L.nbi ? L.live[(ij.x + 1) * L.nbj + ij.y] : -1;
  const int BD = ij.y > 0 ? L.live[ij.x * L.nbj + ij.y - 1] : -1;
  const int BU = ij.y + 1 < L.nbj ? L.live[ij.x * L.nbj + ij.y + 1] : -1;
  // byte offsets modulo 2^32: added to the lane's own offset they wrap to the neighbour's
  auto off = [&](int NB, int d) { return NB >= 0 ? (uint32_t)(((int64_t)(NB - B) * SL + d) * sz) : 0u; };
  FaceOff f;
  f.dL = off(BL, 7 * 64 + 56);
  f.dR = off(BR, -7 * 64 - 56);
  f.dD = off(BD, 7 * 64 + 7);
  f.dU = off(BU, -7 * 64 - 7);
  return f;
}
// element at byte offset `off` of a brick array (an unsigned 32-bit offset from a scalar base,
// so the address needs no 64-bit arithmetic; layouts are limited to 2^32 bytes per array)
template <typename T>
__device__ __forceinline__ T ld_at(const T* __restrict__ base, uint32_t off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

// z = A s with the fixed-tree s.z (applyCompressedMatrix + dotProduct(s, z)); last block
// forms alpha = sigma / s.z.
// S is the storage and arithmetic type of the vectors and the stencil; the dot product is
// accumulated in FP64 either way.
template <bool PCG, typename S>
__global__ __launch_bounds__(256) void k_spmv(Lay L, const S* __restrict__ x, const S* __restrict__ adiag,
                                              const S* __restrict__ al, const S* __restrict__ ad,
                                              const S* __restrict__ ab, const uint8_t* __restrict__ nbm,
                                              S* __restrict__ y, double* __restrict__ part, int nchunks,
                                              PcgScalars* sc, int iter) {
  __shared__ double sm[4];
  if (PCG && sc->converged != STOP_RUN) return;
  // A thread's 16 elements (one per group of 256) are independent until the dot product, which
  // adds them in order. Every neighbour is a byte offset from the element's own: a constant
  // inside the brick, a per-brick constant (FaceOff) across a face, zero when the flag is clear
  // (the element itself is read and the term dropped by a select, so the sums keep their
  // order). The operands of MB elements are loaded before any of their arithmetic, so the loads
  // overlap. Groups past the end reread the last group and are dropped.
  constexpr int MB = 4, sz = (int)sizeof(S);
  const int ne = (int)L.ne, chunk = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, a = lane >> 3, b = lane & 7;
  uint8_t fl[16];
#pragma unroll
  for (int m = 0; m < 16; m++) fl[m] = nbm[min(chunk * CH + m * 256, ne - 256) + tid];
  // with S >= 64 a chunk (64 slots) spans at most two bricks, whose faces are looked up once;
  // smaller layouts look them up per group
  const int SL = L.S * 64;
  const bool two = SL >= CH;
  const int B0 = min(chunk * CH, ne - 256) / SL;
  const FaceOff f0 = face_off<S>(L, B0), f1 = face_off<S>(L, min(B0 + 1, L.nlive - 1));
  double acc = 0.0;
#pragma unroll
  for (int m0 = 0; m0 < 16; m0 += MB) {
    uint32_t e4[MB], oL[MB], oD[MB], oB[MB], oR[MB], oU[MB], oF[MB];
    bool in[MB];
#pragma unroll
    for (int q = 0; q < MB; q++) {
      const int raw = chunk * CH + (m0 + q) * 256;
      in[q] = raw < ne;
      const int base = in[q] ? raw : ne - 256;
      FaceOff g;
      if (two) {
        g = base < (B0 + 1) * SL ? f0 : f1;
      } else {
        g = face_off<S>(L, base / SL);
      }
      const int f = fl[m0 + q];
      const uint32_t e = (uint32_t)(base + tid) * sz;
      e4[q] = e;
      oL[q] = (f & NB_L) ? e + (a > 0 ? (uint32_t)(-72 * sz) : g.dL) : e;
      oD[q] = (f & NB_D) ? e + (b > 0 ? (uint32_t)(-65 * sz) : g.dD) : e;
      oB[q] = (f & NB_B) ? e - 64u * sz : e;
      oR[q] = (f & NB_R) ? e + (a < 7 ? 72u * sz : g.dR) : e;
      oU[q] = (f & NB_U) ? e + (b < 7 ? 65u * sz : g.dU) : e;
      oF[q] = (f & NB_F) ? e + 64u * sz : e;
    }
    S xc[MB], dg[MB], cL[MB], cD[MB], cB[MB], cR[MB], cU[MB], cF[MB], xL[MB], xD[MB], xB[MB], xR[MB], xU[MB],
        xF[MB], v[MB];
#pragma unroll
    for (int q = 0; q < MB; q++) {
      xc[q] = ld_at(x, e4[q]);
      dg[q] = ld_at(adiag, e4[q]);
      cL[q] = ld_at(al, e4[q]);
      cD[q] = ld_at(ad, e4[q]);
      cB[q] = ld_at(ab, e4[q]);
      cR[q] = ld_at(al, oR[q]);
      cU[q] = ld_at(ad, oU[q]);
      cF[q] = ld_at(ab, oF[q]);
      xL[q] = ld_at(x, oL[q]);
      xD[q] = ld_at(x, oD[q]);
      xB[q] = ld_at(x, oB[q]);
      xR[q] = ld_at(x, oR[q]);
      xU[q] = ld_at(x, oU[q]);
      xF[q] = ld_at(x, oF[q]);
    }
#pragma unroll
    for (int q = 0; q < MB; q++) {
      const int f = fl[m0 + q];
      S w = dg[q] * xc[q];
      w = (f & NB_L) ? w + cL[q] * xL[q] : w;
      w = (f & NB_D) ? w + cD[q] * xD[q] : w;
      w = (f & NB_B) ? w + cB[q] * xB[q] : w;
      w = (f & NB_R) ? w + cR[q] * xR[q] : w;
      w = (f & NB_U) ? w + cU[q] * xU[q] : w;
      w = (f & NB_F) ? w + cF[q] * xF[q] : w;
      v[q] = (f & NB_SELF) ? w : S(0);
    }
#pragma unroll
    for (int q = 0; q < MB; q++) {
      if (in[q]) {
        *reinterpret_cast<S*>(reinterpret_cast<char*>(y) + e4[q]) = v[q];
        acc += (double)xc[q] * (double)v[q];
      }
    }
  }
  if (!PCG) return;
  double bs = block_sum256(acc, sm);
  if (threadIdx.x == 0) xstore(part + chunk, bs);
  if (chunk == (int)gridDim.x - 1) {
    gather_wait256(part, nchunks, sc);
    const double sdotz = gather_tree256(part, nchunks, sm);
    if (threadIdx.x == 0) {
      sc->sdotz = sdotz;
      if (sdotz == 0.0) {
        sc->converged = STOP_ABORT;
        sc->status = FS_BREAKDOWN;
        sc->conv_iter = iter;
      } else {
        sc->alpha = sc->sigma / sdotz;
      }
    }
  }
}

// Mixed precision: the FP64 residual b - A x of the FP64 iterate (A is the FP32 stencil, held
// exactly in FP64; the products follow applyCompressedMatrix's order, project-solver.cc:23-35),
// written into the FP32 recurrence residual. Its FP64 norms replace the recurrence's and
// decide convergence (tolerance as in k_fwd).
__global__ __launch_bounds__(256) void k_resid(Lay L, const double* __restrict__ rhs, const double* __restrict__ x,
                                               const double* __restrict__ adiag, const double* __restrict__ al,
                                               const double* __restrict__ ad, const double* __restrict__ ab,
                                               const uint8_t* __restrict__ nbm, float* __restrict__ r,
                                               double* __restrict__ part, double* __restrict__ part2, int nchunks,
                                               PcgScalars* sc, int iter, int tol_mode, double tol) {
  __shared__ double sm[4];
  double mx = 0.0, ss = 0.0;
  for (int m = 0; m < 16; m++) {
    const int base = blockIdx.x * CH + m * 256;
    if (base >= L.ne) break;
    const Group g = group_of(L, base);
    const int e = base + threadIdx.x, lane = threadIdx.x & 63, s = g.s0 + (threadIdx.x >> 6);
    const int a = lane >> 3, b = lane & 7;
    const int f = nbm[e];
    double res = 0.0;
    if (f & NB_SELF) {
      double v = adiag[e] * x[e];
      if (f & NB_L) v = v + al[e] * x[nb_left(L, g, e, s, a, b)];
      if (f & NB_D) v = v + ad[e] * x[nb_down(L, g, e, s, a, b)];
      if (f & NB_B) v = v + ab[e] * x[e - 64];
      if (f & NB_R) {
        const int n = nb_right(L, g, e, s, a, b);
        v = v + al[n] * x[n];
      }
      if (f & NB_U) {
        const int n = nb_up(L, g, e, s, a, b);
        v = v + ad[n] * x[n];
      }
      if (f & NB_F) v = v + ab[e + 64] * x[e + 64];
      res = rhs[e] - v;
    }
    r[e] = (float)res;
    const double av = fabs(res);
    mx = (mx < av) ? av : mx;
    ss += res * res;
  }
  double bm = block_max256(mx, sm);
  double bs = block_sum256(ss, sm);
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
      } else {
        sc->converged = STOP_RUN;
      }
    }
  }
}

// ---- MIC sweeps ---------------------------------------------------------------------------------
// One workgroup of four waves per live brick. Lane (a, b) = (lane >> 3, lane & 7) owns column
// (8I+a, 8J+b) and at slot s handles cell k = s - a - b, so its -x and -y neighbours at the same
// k were finished one slot earlier by lanes a-1 and b-1 (lane shuffles) or, on the brick face,
// 7 slots earlier by the neighbouring brick (exchange entries X[f][B][s][w]).
//
// Chunk c (U slots) of a brick array is one contiguous run of U*64 elements, and the entries a
// chunk consumes from one face are 64 doubles, one per lane. The dependent chain of a sweep is
// short arithmetic per slot; everything that waits on memory is kept off it:
//   Wave 3 (array loads) keeps SwCfg::R chunks of 16-byte-per-lane loads in flight and moves
//   each loaded chunk to an LDS ring (to the lanes that own the elements).
//   Wave 2 (exchange) reads each chunk's entries ahead, polling until the neighbours have
//   written them.
//   Wave 0 (arithmetic) runs the chain from LDS and writes each finished chunk to an outbox.
//   Wave 1 (stores) hands the outbox's boundary values to the next bricks (one 8-byte store per
//   lane and face), returns the consumed entries to the sentinel and writes the outputs with
//   16-byte stores.
// Waves 3 and 2 issue loads only: loads complete in issue order, so the waits the compiler
// places before using loaded values stay exact (with stores in flight it has to wait for
// everything, which would drain the pipeline). The waves meet at LDS counters; every wait is
// bounded by SPIN_LIMIT, after which the wave gives up, counts a timeout and the host reports
// FS_DEVICE_ERROR.
//
// Exchange entries hold a sentinel until the producer stores the value (agent-scope atomic
// stores and loads); the consumer uses the value and stores the sentinel back.



// v of lane-1 / lane+1 inside 16-lane rows (DPP row_shr:1 / row_shl:1); a lane whose source is
// outside its row gets 0. Used for the b-1 / b+1 neighbour, which lanes b = 0 / 7 never take.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const unsigned long long x = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHL1 = 0x101;

constexpr int SW_NA = 6;       // input arrays per chunk
constexpr int SW_CE = U * 64;  // elements per chunk and array
constexpr int SW_K = 2;        // outbox slots
constexpr int SW_KX = 4;       // exchange-entry slots (wave 2 reads up to this many chunks ahead)
typedef float sw_f4 __attribute__((ext_vector_type(4)));
typedef double sw_d2 __attribute__((ext_vector_type(2)));
template <typename V>
using Vec16 = typename std::conditional<sizeof(V) == 8, sw_d2, sw_f4>::type;  // 16 bytes, a register tuple
template <typename V>
struct SwCfg {
  static constexpr int R = sizeof(V) == 8 ? 3 : 4;           // chunks of loads in flight (loader register sets)
  static constexpr int RD = sizeof(V) == 8 ? 2 : 4;          // chunks loaded into LDS ahead of the compute wave
  static constexpr int OPS = SW_CE * (int)sizeof(V) / 1024;  // 16-byte loads per lane, array and chunk
  static constexpr int PER16 = 16 / (int)sizeof(V);          // elements per 16-byte access
};

// Loads of one chunk, in wave 0's registers.
template <typename V>
struct SwSet {
  Vec16<V> d[SW_NA][SwCfg<V>::OPS];
};

// Shared state of a sweep workgroup.
template <typename V>
struct SwLds {
  __align__(16) V tb[SwCfg<V>::RD][SW_NA][SW_CE];  // chunks c % RD, one element per lane and slot
  __align__(16) V ob[SW_K][SW_CE];        // outbox: outputs of finished chunks
  double xr[SW_KX][2][64];                // exchange entries of chunks c % SW_KX, from wave 2
  __align__(8) int loaded;  // chunks loaded by wave 3 and (next word) released by wave 1: one
  int freed;                // 8-byte word, so that wave 0 reads both at once
  int prog, xready, B, stop;       // slots finished by wave 0 / slots whose entries wave 2 has read
  int last;                                  // this brick holds the last ticket (it reduces)
};

static_assert(offsetof(SwLds<float>, freed) == offsetof(SwLds<float>, loaded) + 4 &&
                  offsetof(SwLds<double>, freed) == offsetof(SwLds<double>, loaded) + 4 &&
                  offsetof(SwLds<float>, loaded) % 8 == 0 && offsetof(SwLds<double>, loaded) % 8 == 0,
              "loaded and freed share one 8-byte word");

__device__ __forceinline__ int lds_get(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_put(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Wait until *p >= v; false if the wait gave up.
__device__ __forceinline__ bool lds_wait(const int* p, int v, PcgScalars* sc) {
  long spins = 0;
  while (lds_get(p) < v) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255) == 0) {
      const int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (to != 0 || spins > SPIN_LIMIT) {
        if ((threadIdx.x & 63) == 0) atomicAdd(&sc->wait_timeouts, 1);
        __builtin_amdgcn_s_waitcnt(0);  // leave nothing in flight on this path
        return false;
      }
    }
  }
  return true;
}

// Wait until p[0] >= v0 and p[1] >= v1 (two counters of one 8-byte word, read together).
__device__ __forceinline__ void lds_wait2(const int* p, int v0, int v1, PcgScalars* sc) {
  long spins = 0;
  while (true) {
    const unsigned long long w = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_ACQUIRE,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    if (((int)(unsigned)w >= v0) & ((int)(unsigned)(w >> 32) >= v1)) return;
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255) == 0) {
      const int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (to != 0 || spins > SPIN_LIMIT) {
        if ((threadIdx.x & 63) == 0) atomicAdd(&sc->wait_timeouts, 1);
        __builtin_amdgcn_s_waitcnt(0);  // leave nothing in flight on this path
        return;
      }
    }
  }
}

// Slot counters (prog, xready) are posted once per slot, so they use relaxed accesses: the LDS
// performs one wave's operations in issue order, so a counter stored after plain LDS stores is
// seen by another wave only once those stores are performed, and that wave's later reads (issued
// after it has read the counter) see them. A compiler-only fence keeps the program order, and
// no wait for the stores to complete is needed.
__device__ __forceinline__ int lds_peek(const int* p) {
  const int v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return v;
}
__device__ __forceinline__ void lds_post(int* p, int v) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Wait until *p >= v (bounded like lds_wait); returns the value read, or v if the wait gave up.
__device__ __forceinline__ int lds_until(const int* p, int v, PcgScalars* sc) {
  long spins = 0;
  int x;
  while ((x = __builtin_amdgcn_readfirstlane(lds_peek(p))) < v) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255) == 0) {
      const int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (to != 0 || spins > SPIN_LIMIT) {
        if ((threadIdx.x & 63) == 0) atomicAdd(&sc->wait_timeouts, 1);
        __builtin_amdgcn_s_waitcnt(0);  // leave nothing in flight on this path
        return v;
      }
    }
  }
  return x;
}

// Element offset of chunk c in a brick array: forward chunks run up from slot 0, backward
// chunks down from slot S-1 (chunk c then starts at slot S - (c+1)U).
struct RunFwd {
  __device__ size_t operator()(int c, int) const { return (size_t)c * SW_CE; }
};
struct RunBwd {
  __device__ size_t operator()(int c, int S) const { return (size_t)(S - (c + 1) * U) * 64; }
};

// Wave 3: issue the loads of chunk c (past the end: the last chunk again, unused). The loads are
// non-temporal: each element is read once per sweep, and keeping these streams out of the
// caches leaves them to the data that is reused soon (the SpMV's neighbour rows, the vectors
// handed from one kernel to the next): SpMV 45 -> 40 us, backward sweep 122 -> 118 us, substep
// 112 -> 108 ms at 256^3 mixed.
template <int NARR, typename V, typename RUN>
__device__ __forceinline__ void sw_load(SwSet<V>& st, const V* const* src, int nc, int c, int lane, RUN run, int S) {
  const int cs = min(c, nc - 1);
  const size_t o = run(cs, S);
#pragma unroll
  for (int k = 0; k < NARR; k++) {
#pragma unroll
    for (int h = 0; h < SwCfg<V>::OPS; h++)
      st.d[k][h] = __builtin_nontemporal_load(
          reinterpret_cast<const Vec16<V>*>(src[k] + o + h * (1024 / (int)sizeof(V)) + lane * SwCfg<V>::PER16));
  }
}

// Wave 3: move a chunk's array loads to an LDS ring slot, element e of array k at tb[k][e].
template <int NARR, typename V>
__device__ __forceinline__ void sw_spill(V (*tb)[SW_CE], const SwSet<V>& st, int lane) {
#pragma unroll
  for (int k = 0; k < NARR; k++) {
#pragma unroll
    for (int h = 0; h < SwCfg<V>::OPS; h++)
      *reinterpret_cast<Vec16<V>*>(&tb[k][h * (1024 / (int)sizeof(V)) + lane * SwCfg<V>::PER16]) = st.d[k][h];
  }
}

// One chunk of an LDS array to global memory, 16 bytes per lane.
template <typename V>
__device__ __forceinline__ void sw_store(V* dst, const V* lds, int lane) {
#pragma unroll
  for (int h = 0; h < SwCfg<V>::OPS; h++) {
    const int o = h * (1024 / (int)sizeof(V)) + lane * SwCfg<V>::PER16;
    *reinterpret_cast<Vec16<V>*>(dst + o) = *reinterpret_cast<const Vec16<V>*>(lds + o);
  }
}


// Wave 3 (loads): keeps SwCfg::R chunks of array loads in flight (register sets named rather
// than indexed so that all stay in registers), applies PREP to each loaded chunk (elementwise
// work that the chain's operands need) and moves its first NSPILL arrays to LDS ring slot
// c % RD once wave 0 (GUARD: and wave 1) is done with the chunk that used it. Wave 3 issues loads
// only, so the waits before the moves are exact. Whole groups of R chunks run as straight-line
// code (an exit inside would merge paths and cost exact waits).
template <int NARR, int NSPILL, bool GUARD, typename V, typename RUN, typename PREP>
__device__ __forceinline__ void sw_wave3(SwLds<V>& sh, const V* const* src, int nc, int S, RUN run, PcgScalars* sc,
                                         int lane, PREP prep) {
  constexpr int R = SwCfg<V>::R, RD = SwCfg<V>::RD;
  SwSet<V> s0, s1, s2, s3;
  sw_load<NARR>(s0, src, nc, 0, lane, run, S);
  sw_load<NARR>(s1, src, nc, 1, lane, run, S);
  if (R >= 3) sw_load<NARR>(s2, src, nc, 2, lane, run, S);
  if (R >= 4) sw_load<NARR>(s3, src, nc, 3, lane, run, S);
  auto chunk = [&](int c, SwSet<V>& cur) __attribute__((always_inline)) {
    if (c >= RD) lds_until(&sh.prog, (c - RD + 1) * U, sc);
    if (GUARD && c >= RD) lds_wait(&sh.freed, c - RD + 1, sc);  // wave 1 has read the ring slot too
    prep(cur);
    sw_spill<NSPILL>(sh.tb[c % RD], cur, lane);
    lds_put(&sh.loaded, c + 1);
    sw_load<NARR>(cur, src, nc, c + R, lane, run, S);
  };
  int c = 0;
  for (; c + R <= nc; c += R) {
    chunk(c, s0);
    chunk(c + 1, s1);
    if (R >= 3) chunk(c + 2, s2);
    if (R >= 4) chunk(c + 3, s3);
  }
  if (c < nc) chunk(c, s0);
  if (R >= 3 && c + 1 < nc) chunk(c + 1, s1);
  if (R >= 4 && c + 2 < nc) chunk(c + 2, s2);
}

// Wave 0 (arithmetic): for each chunk, once wave 3 has loaded it and wave 1 has released the
// outbox slot, BODY(c, slot, tb, xr, xw) computes it from LDS ring slot tb and entries xr into
// sh.ob[slot], posting prog after every slot. The exchange entries are tracked per slot through
// xw: a batch of slots whose entries are all read runs straight through, otherwise each slot
// waits for its own. Wave 0 issues no global memory access.
template <typename V>
struct SwXw {
  SwLds<V>& sh;
  PcgScalars* sc;
  int known;  // slots whose entries wave 2 has read, as last seen
  // true if the entries of every slot below `need` are in xr (one counter read at most)
  __device__ __forceinline__ bool all(int need) {
    if (known < need) known = __builtin_amdgcn_readfirstlane(lds_peek(&sh.xready));
    return known >= need;
  }
  __device__ __forceinline__ void wait(int need) {
    if (known < need) known = lds_until(&sh.xready, need, sc);
  }
  __device__ __forceinline__ void post(int v) { lds_post(&sh.prog, v); }
};
template <typename V, typename BODY>
__device__ __forceinline__ void sw_wave0(SwLds<V>& sh, int nc, PcgScalars* sc, BODY body) {
  SwXw<V> xw{sh, sc, 0};
  for (int c = 0; c < nc; c++) {
    const int slot = c % SW_K;
    lds_wait2(&sh.loaded, c + 1, c - SW_K + 1, sc);  // chunk loaded, outbox slot released
    body(c, slot, sh.tb[c % SwCfg<V>::RD], &sh.xr[c % SW_KX][0][0], xw);
  }
}

// Wave 2: read the exchange entries of each chunk (lane e: entry e of both faces, slot e / 8 of
// the chunk) into LDS slot c % SW_KX, up to SW_KX chunks ahead of wave 0, polling until the
// producers have written them (bounded by SPIN_LIMIT); xready counts the leading slots whose
// sixteen entries are all in. Wave 2 issues loads only.
template <typename V, typename FX>
__device__ __forceinline__ void sw_wave2(SwLds<V>& sh, int nc, FX ent, PcgScalars* sc, int lane) {
  bool give_up = false;
  for (int c = 0; c < nc; c++) {
    if (c >= SW_KX && !lds_wait(&sh.prog, (c - SW_KX + 1) * U, sc)) give_up = true;
    const double* p0 = ent(0, c);
    const double* p1 = ent(1, c);
    double v0 = xload(p0), v1 = xload(p1);
    int pub = 0;
    long spins = 0;
    while (true) {
      const uint64_t miss = __builtin_amdgcn_ballot_w64(xunset(v0) | xunset(v1));
      const int nv = (give_up | (miss == 0)) ? U : (int)(__builtin_ctzll(miss) >> 3);
      if (nv > pub) {
        sh.xr[c % SW_KX][0][lane] = v0;
        sh.xr[c % SW_KX][1][lane] = v1;
        lds_post(&sh.xready, c * U + nv);
        pub = nv;
      }
      if (pub == U) break;
      __builtin_amdgcn_s_sleep(1);
      if (xunset(v0)) v0 = xload(p0);
      if (xunset(v1)) v1 = xload(p1);
      if ((++spins & 255) == 0) {
        const int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (to != 0 || spins > SPIN_LIMIT) {
          if (lane == 0) atomicAdd(&sc->wait_timeouts, 1);
          give_up = true;
        }
      }
    }
  }
}

// Wave 1: hands each finished slot's boundary values to the next bricks as soon as wave 0 posts
// it (PUB(c, v, lo, hi): lanes of slots [lo, hi) of chunk c store v), then, once chunk c is
// complete, reads its outbox slot into registers, folds the chunk into the sweep's reductions
// (FOLD(c, outbox), off the arithmetic wave's chain), releases the slot and issues the chunk's
// stores via STORE(c, regs). Two register sets alternate so that reading chunk c+1 never has to
// wait for chunk c's stores to have read their data.
template <typename V>
struct SwOut {
  Vec16<V> d[2][SwCfg<V>::OPS];  // the chunk's outputs, 16 bytes per lane: [0] from the outbox,
                                 // [1] (if any) read from the ring by FOLD
};
template <typename V, typename BND, typename PUB, typename FOLD, typename STORE>
__device__ __forceinline__ void sw_wave1(SwLds<V>& sh, int nc, PcgScalars* sc, int lane, BND bnd, PUB pub,
                                         FOLD fold, STORE store) {
  SwOut<V> o0, o1;
  int seen = 0;  // slots finished by wave 0, as last read
  auto one = [&](int c, SwOut<V>& o) __attribute__((always_inline)) {
    const int slot = c % SW_K;
    for (int lo = 0; lo < U;) {
      if (seen <= c * U + lo) seen = lds_until(&sh.prog, c * U + lo + 1, sc);
      const int hi = min(seen - c * U, U);
      V v[2];
      bnd(&sh.ob[slot][0], v);
      pub(c, v, lo, hi);
      lo = hi;
    }
#pragma unroll
    for (int h = 0; h < SwCfg<V>::OPS; h++)
      o.d[0][h] = *reinterpret_cast<const Vec16<V>*>(&sh.ob[slot][h * (1024 / (int)sizeof(V)) +
                                                                 lane * SwCfg<V>::PER16]);
    fold(c, &sh.ob[slot][0], o);  // per-chunk reductions and ring reads, before the slot is released
    lds_put(&sh.freed, c + 1);
    store(c, o);
  };
  int c = 0;
  for (; c + 2 <= nc; c += 2) {
    one(c, o0);
    one(c + 1, o1);
  }
  if (c < nc) one(c, o0);
}

// Forward sweep L q = r (project-solver.cc:275-288). UPDATE: first r += (-alpha) t and
// ||r||_inf (saxpy + LinfNorm, :333-335); the last brick decides convergence.
// T: storage and arithmetic type of vectors and coefficients; the exchange entries stay FP64.
template <bool UPDATE, typename T>
__global__ __launch_bounds__(256) void k_fwd(Lay L, T* __restrict__ r, const T* __restrict__ t,
                                             const T* __restrict__ cl, const T* __restrict__ cd,
                                             const T* __restrict__ cb, const T* __restrict__ E,
                                             T* __restrict__ q, double* X, double* __restrict__ part2,
                                             double* __restrict__ part, uint64_t* ticket, uint64_t tbase,
                                             const int* __restrict__ order, PcgScalars* sc, int iter, int tol_mode,
                                             double tol) {
  enum { AR = 0, AL = 1, AD = 2, AB = 3, AE = 4, AT = 5 };  // t last: the plain sweep loads five arrays
  constexpr int NARR = UPDATE ? 6 : 5;
  __shared__ SwLds<T> sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, a = lane >> 3, b = lane & 7;
  if (threadIdx.x == 0) {
    const uint64_t tk = atomicAdd((unsigned long long*)ticket, 1ull) - tbase;
    // a ticket outside [0, n) can only follow a failed launch; clamp it so every access
    // stays in bounds (the exchange polls then time out and the solve reports an error)
    sh.B = order[tk < (uint64_t)L.nlive ? tk : (uint64_t)(L.nlive - 1)];
    sh.last = tk == (uint64_t)(L.nlive - 1);
    sh.stop = UPDATE && sc->converged != STOP_RUN;
    sh.prog = 0;
    sh.freed = 0;
    sh.xready = 0;
    sh.loaded = 0;
  }
  __syncthreads();
  if (sh.stop) return;
  const int B = __builtin_amdgcn_readfirstlane(sh.B);
  const int S = __builtin_amdgcn_readfirstlane(L.S), NC = S / U;
  const size_t brick = (size_t)B * S * 64, face = (size_t)__builtin_amdgcn_readfirstlane(L.nlive) * S * 8;
  const int2 ij = L.bij[B];
  const int BL = __builtin_amdgcn_readfirstlane(ij.x > 0 ? L.live[(ij.x - 1) * L.nbj + ij.y] : -1);
  const int BD = __builtin_amdgcn_readfirstlane(ij.y > 0 ? L.live[ij.x * L.nbj + ij.y - 1] : -1);
  const double* xn[2] = {BL >= 0 ? X + (size_t)BL * S * 8 : nullptr,  // f 0 entries of the -x neighbour
                         BD >= 0 ? X + face + (size_t)BD * S * 8 : nullptr};
  const double* Z = X + 2 * face;  // zero row: no neighbour, or past the last slot
  // this lane's entry of chunk c on face f: slot cU + lane/8 + 7 of the neighbour, index lane % 8
  const int es = (lane >> 3) + 7, ew = lane & 7;
  auto live = [&](int f, int c) __attribute__((always_inline)) { return xn[f] != nullptr && c * U + es < S; };
  auto ent = [&](int f, int c) __attribute__((always_inline)) -> const double* {
    return live(f, c) ? xn[f] + (size_t)(c * U + es) * 8 + ew : Z + ew;
  };
  const double nalpha = UPDATE ? -sc->alpha : 0.0;
  if (wave == 1) {  // stores
    double* o0 = X + (size_t)B * S * 8;  // this brick's f 0 entries (read by its +x neighbour)
    double* o1 = X + face + (size_t)B * S * 8;
    // lane e hands over slot cU + e/8, index e % 8 of each face
    auto bnd = [&](const T* qc, T* v) __attribute__((always_inline)) {
      v[0] = qc[(lane >> 3) * 64 + 56 + (lane & 7)];
      v[1] = qc[(lane >> 3) * 64 + (lane & 7) * 8 + 7];
    };
    auto pub = [&](int c, const T* v, int lo, int hi) __attribute__((always_inline)) {
      if ((lane >> 3) >= lo && (lane >> 3) < hi) {
        xstore(o0 + (size_t)c * U * 8 + lane, (double)v[0]);
        xstore(o1 + (size_t)c * U * 8 + lane, (double)v[1]);
      }
    };
    // UPDATE: the updated residual is in the chunk's ring slot (formed by wave 3): ||r||_inf and
    // r.r per lane in slot order, and its 16-byte-per-lane store values
    double mx = 0.0, ss = 0.0;
    auto fold = [&](int c, const T*, SwOut<T>& o) __attribute__((always_inline)) {
      if (UPDATE) {
        const T* rc = sh.tb[c % SwCfg<T>::RD][AR];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const double rr = (double)rc[u * 64 + lane];
          const double av = fabs(rr);
          mx = (mx < av) ? av : mx;
          ss += rr * rr;
        }
#pragma unroll
        for (int h = 0; h < SwCfg<T>::OPS; h++)
          o.d[1][h] = *reinterpret_cast<const Vec16<T>*>(&rc[h * (1024 / (int)sizeof(T)) + lane * SwCfg<T>::PER16]);
      }
    };
    sw_wave1(sh, NC, sc, lane, bnd, pub, fold, [&](int c, const SwOut<T>& o) __attribute__((always_inline)) {
      if (live(0, c)) xstore(const_cast<double*>(ent(0, c)), xsent());  // consumed entries back
      if (live(1, c)) xstore(const_cast<double*>(ent(1, c)), xsent());
#pragma unroll
      for (int k = 0; k < (UPDATE ? 2 : 1); k++)
#pragma unroll
        for (int h = 0; h < SwCfg<T>::OPS; h++)
          *reinterpret_cast<Vec16<T>*>((k == 0 ? q : r) + brick + RunFwd{}(c, S) + h * (1024 / (int)sizeof(T)) +
                                       lane * SwCfg<T>::PER16) = o.d[k][h];
    });
    if (UPDATE) {
      const double wm = wave_max(mx);
      const double wsum = wave_sum(ss);
      if (lane == 0) {
        xstore(part2 + B, wm);
        xstore(part + B, wsum);
      }
      if (__builtin_amdgcn_readfirstlane(sh.last)) {
        gather_wait_wave(part2, L.nlive, sc);
        gather_wait_wave(part, L.nlive, sc);
        const double res = gather_max_wave(part2, L.nlive);
        const double rr2 = gather_tree256_wave(part, L.nlive);
        if (lane == 0) {
          sc->residual = res;
          sc->rnorm2sq = rr2;
          bool conv = tol_mode == FS_TOL_REL_L2 ? (sqrt(rr2) <= tol * sqrt(sc->bnorm2sq)) : (res < tol);
          if (!(res <= 1.7976931348623157e308)) {
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
    return;
  }
  if (wave == 2) {  // exchange entries
    sw_wave2(sh, NC, ent, sc, lane);
    return;
  }
  if (wave == 3) {  // array loads
    const T* src[SW_NA] = {r + brick, cl + brick, cd + brick, cb + brick, E + brick, t + brick};
    // UPDATE: r + (-alpha) t into the r slot as each chunk arrives (t is not staged)
    const T na = (T)nalpha;
    sw_wave3<NARR, 5, UPDATE>(sh, src, NC, S, RunFwd{}, sc, lane, [&](SwSet<T>& st) __attribute__((always_inline)) {
      if (UPDATE) {
#pragma unroll
        for (int h = 0; h < SwCfg<T>::OPS; h++) st.d[AR][h] = st.d[AR][h] + na * st.d[AT][h];
      }
    });
    return;
  }
  T qprev = T(0);
  sw_wave0(sh, NC, sc, [&](int c, int slot, const T (*tb)[SW_CE], const double* xr, SwXw<T>& xw)
                           __attribute__((always_inline)) {
    // operands of UB slots at a time into registers first, so the chain of the slots waits for
    // LDS only in its lane shuffle
    constexpr int UB = sizeof(T) == 8 ? 4 : 8;
#pragma unroll
    for (int u0 = 0; u0 < U; u0 += UB) {
      T rv[UB], lv[UB], dv[UB], bv[UB], ev[UB];
#pragma unroll
      for (int k = 0; k < UB; k++) {
        const int i = (u0 + k) * 64 + lane;
        rv[k] = tb[AR][i];  // UPDATE: already r + (-alpha) t
        lv[k] = tb[AL][i];
        dv[k] = tb[AD][i];
        bv[k] = tb[AB][i];
        ev[k] = tb[AE][i];
      }
      auto step = [&](int k, T x0, T x1) __attribute__((always_inline)) {
        const int u = u0 + k;
        T qL = __shfl_up(qprev, 8, 64);
        T qD = dpp_row<DPP_ROW_SHR1>(qprev);
        qL = a == 0 ? x0 : qL;
        qD = b == 0 ? x1 : qD;
        T qv;
        if constexpr (sizeof(T) == 4) {
          // FP32 (mixed precision): q = rE - (bE) q_prev - (dE) qD - (lE) qL with the products
          // by E formed off the chain, three fused multiply-adds, the lane shuffle's operand last
          T tt = fmaf(-bv[k] * ev[k], qprev, rv[k] * ev[k]);
          tt = fmaf(-dv[k] * ev[k], qD, tt);
          qv = fmaf(-lv[k] * ev[k], qL, tt);
        } else {
          T tt = rv[k] - lv[k] * qL;
          tt = tt - dv[k] * qD;
          tt = tt - bv[k] * qprev;
          qv = tt * ev[k];
        }
        sh.ob[slot][u * 64 + lane] = qv;
        qprev = qv;
        xw.post(c * U + u + 1);
      };
      if (xw.all(c * U + u0 + UB)) {
        T x0[UB], x1[UB];  // read by every lane: no branch
#pragma unroll
        for (int k = 0; k < UB; k++) {
          x0[k] = (T)xr[(u0 + k) * 8 + b];
          x1[k] = (T)xr[64 + (u0 + k) * 8 + a];
        }
#pragma unroll
        for (int k = 0; k < UB; k++) step(k, x0[k], x1[k]);
      } else {
#pragma unroll
        for (int k = 0; k < UB; k++) {
          xw.wait(c * U + u0 + k + 1);
          step(k, (T)xr[(u0 + k) * 8 + b], (T)xr[64 + (u0 + k) * 8 + a]);
        }
      }
    }
  });
}

// Backward sweep L^T z = q (project-solver.cc:289-305) and the fixed-tree z.r; the last brick
// forms beta = sigma_new / sigma (INIT: sigma = z.r). Slots run from S-1 down to 0: chunk c
// covers slots S-1-cU down to S-U-cU, and its u-th slot, S-1-cU-u, is element (U-1-u)*64 + lane
// of the chunk's run.
template <bool INIT, typename T>
__global__ __launch_bounds__(256) void k_bwd(Lay L, const T* __restrict__ qv, const T* __restrict__ dr,
                                             const T* __restrict__ du, const T* __restrict__ df,
                                             const T* __restrict__ E, const T* __restrict__ r,
                                             T* __restrict__ z, double* X, double* __restrict__ part,
                                             uint64_t* ticket, uint64_t tbase, const int* __restrict__ order,
                                             PcgScalars* sc) {
  enum { AQ = 0, A1 = 1, A2 = 2, A3 = 3, AE = 4, AR = 5 };
  __shared__ SwLds<T> sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, a = lane >> 3, b = lane & 7;
  if (threadIdx.x == 0) {
    const uint64_t tk = atomicAdd((unsigned long long*)ticket, 1ull) - tbase;
    sh.B = order[tk < (uint64_t)L.nlive ? tk : (uint64_t)(L.nlive - 1)];
    sh.last = tk == (uint64_t)(L.nlive - 1);
    sh.stop = !INIT && sc->converged != STOP_RUN;
    sh.prog = 0;
    sh.freed = 0;
    sh.xready = 0;
    sh.loaded = 0;
  }
  __syncthreads();
  if (sh.stop) return;
  const int B = __builtin_amdgcn_readfirstlane(sh.B);
  const int S = __builtin_amdgcn_readfirstlane(L.S), NC = S / U;
  const size_t brick = (size_t)B * S * 64, face = (size_t)__builtin_amdgcn_readfirstlane(L.nlive) * S * 8;
  const int2 ij = L.bij[B];
  const int BR = __builtin_amdgcn_readfirstlane(ij.x + 1 < L.nbi ? L.live[(ij.x + 1) * L.nbj + ij.y] : -1);
  const int BU = __builtin_amdgcn_readfirstlane(ij.y + 1 < L.nbj ? L.live[ij.x * L.nbj + ij.y + 1] : -1);
  const double* xn[2] = {BR >= 0 ? X + (size_t)BR * S * 8 : nullptr,  // f 0 entries of the +x neighbour
                         BU >= 0 ? X + face + (size_t)BU * S * 8 : nullptr};
  const double* Z = X + 2 * face;
  // this lane's entry of chunk c on face f: slot S-1-cU-lane/8 - 7 of the neighbour, index lane % 8
  auto slot_of = [&](int c) __attribute__((always_inline)) { return S - 1 - c * U - (lane >> 3); };
  const int ew = lane & 7;
  auto live = [&](int f, int c) __attribute__((always_inline)) { return xn[f] != nullptr && slot_of(c) >= 7; };
  auto ent = [&](int f, int c) __attribute__((always_inline)) -> const double* {
    return live(f, c) ? xn[f] + (size_t)(slot_of(c) - 7) * 8 + ew : Z + ew;
  };
  if (wave == 1) {  // stores
    double* o0 = X + (size_t)B * S * 8;  // this brick's f 0 entries (read by its -x neighbour)
    double* o1 = X + face + (size_t)B * S * 8;
    // lane e hands over slot S-1-cU-e/8, index e % 8 of each face
    const int e0 = (U - 1 - (lane >> 3)) * 64;
    auto bnd = [&](const T* zc, T* v) __attribute__((always_inline)) {
      v[0] = zc[e0 + ew];
      v[1] = zc[e0 + ew * 8];
    };
    auto pub = [&](int c, const T* v, int lo, int hi) __attribute__((always_inline)) {
      if ((lane >> 3) >= lo && (lane >> 3) < hi) {
        const int so = slot_of(c);
        xstore(o0 + (size_t)so * 8 + ew, (double)v[0]);
        xstore(o1 + (size_t)so * 8 + ew, (double)v[1]);
      }
    };
    // z.r per lane in slot order: z from the outbox, r from the chunk's ring slot (which wave 3
    // refills only after this wave has released the chunk)
    double acc = 0.0;
    auto fold = [&](int c, const T* ob, SwOut<T>&) __attribute__((always_inline)) {
      const T* rc = sh.tb[c % SwCfg<T>::RD][AR];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = (U - 1 - u) * 64 + lane;
        acc += (double)ob[i] * (double)rc[i];
      }
    };
    sw_wave1(sh, NC, sc, lane, bnd, pub, fold, [&](int c, const SwOut<T>& o) __attribute__((always_inline)) {
      if (live(0, c)) xstore(const_cast<double*>(ent(0, c)), xsent());  // consumed entries back
      if (live(1, c)) xstore(const_cast<double*>(ent(1, c)), xsent());
#pragma unroll
      for (int h = 0; h < SwCfg<T>::OPS; h++)
        *reinterpret_cast<Vec16<T>*>(z + brick + RunBwd{}(c, S) + h * (1024 / (int)sizeof(T)) +
                                     lane * SwCfg<T>::PER16) = o.d[0][h];
    });
    const double ws = wave_sum(acc);
    if (lane == 0) xstore(part + B, ws);
    if (__builtin_amdgcn_readfirstlane(sh.last)) {
      gather_wait_wave(part, L.nlive, sc);
      const double sig = gather_tree256_wave(part, L.nlive);
      if (lane == 0) {
        if (INIT) {
          sc->sigma = sig;
        } else {
          sc->beta = sig / sc->sigma;
          sc->sigma = sig;
        }
      }
    }
    return;
  }
  if (wave == 2) {  // exchange entries
    sw_wave2(sh, NC, ent, sc, lane);
    return;
  }
  if (wave == 3) {  // array loads
    const T* src[SW_NA] = {qv + brick, dr + brick, du + brick, df + brick, E + brick, r + brick};
    sw_wave3<SW_NA, SW_NA, true>(sh, src, NC, S, RunBwd{}, sc, lane, [](SwSet<T>&) {});
    return;
  }
  T zprev = T(0);
  sw_wave0(sh, NC, sc, [&](int c, int slot, const T (*tb)[SW_CE], const double* xr, SwXw<T>& xw)
                           __attribute__((always_inline)) {
    constexpr int UB = sizeof(T) == 8 ? 4 : 8;  // operands of UB slots into registers first
#pragma unroll
    for (int u0 = 0; u0 < U; u0 += UB) {
      T qv[UB], v1[UB], v2[UB], v3[UB], ev[UB];
#pragma unroll
      for (int k = 0; k < UB; k++) {
        const int i = (U - 1 - (u0 + k)) * 64 + lane;
        qv[k] = tb[AQ][i];
        v1[k] = tb[A1][i];
        v2[k] = tb[A2][i];
        v3[k] = tb[A3][i];
        ev[k] = tb[AE][i];
      }
      auto step = [&](int k, T x0, T x1) __attribute__((always_inline)) {
        const int u = u0 + k;
        T zR = __shfl_down(zprev, 8, 64);
        T zU = dpp_row<DPP_ROW_SHL1>(zprev);
        zR = a == 7 ? x0 : zR;
        zU = b == 7 ? x1 : zU;
        T zv;
        if constexpr (sizeof(T) == 4) {  // FP32: as the forward step, the shuffle's operand last
          T tt = fmaf(-v3[k] * ev[k], zprev, qv[k] * ev[k]);
          tt = fmaf(-v2[k] * ev[k], zU, tt);
          zv = fmaf(-v1[k] * ev[k], zR, tt);
        } else {
          T tt = qv[k] - v1[k] * zR;
          tt = tt - v2[k] * zU;
          tt = tt - v3[k] * zprev;
          zv = tt * ev[k];
        }
        sh.ob[slot][(U - 1 - u) * 64 + lane] = zv;
        zprev = zv;
        xw.post(c * U + u + 1);
      };
      if (xw.all(c * U + u0 + UB)) {
        T x0[UB], x1[UB];  // read by every lane: no branch
#pragma unroll
        for (int k = 0; k < UB; k++) {
          x0[k] = (T)xr[(u0 + k) * 8 + b];
          x1[k] = (T)xr[64 + (u0 + k) * 8 + a];
        }
#pragma unroll
        for (int k = 0; k < UB; k++) step(k, x0[k], x1[k]);
      } else {
#pragma unroll
        for (int k = 0; k < UB; k++) {
          xw.wait(c * U + u0 + k + 1);
          step(k, (T)xr[(u0 + k) * 8 + b], (T)xr[64 + (u0 + k) * 8 + a]);
        }
      }
    }
  });
}

// pg += alpha s (saxpy, :332); s = z + beta s (scaleAndAdd, :341) unless the solve stopped.
// The iterate x is FP64 in both precisions; s and z have the storage type. One thread per 4
// consecutive elements (ne is a multiple of 64), with 4-wide loads and stores.
template <typename S>
__global__ __launch_bounds__(256) void k_update(int64_t ne, double* __restrict__ x, S* __restrict__ s,
                                                const S* __restrict__ z, PcgScalars* sc, int iter) {
  typedef S S4 __attribute__((ext_vector_type(4)));
  typedef double D4 __attribute__((ext_vector_type(4)));
  const int64_t e = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * 4;
  if (e >= ne) return;
  // the three streams' loads go out before the stop state is read (a chain of scalar loads);
  // x is next read by the next iteration's update, so it bypasses the caches (non-temporal),
  // leaving them to the sweeps' and the SpMV's streams
  const S4 sv = *reinterpret_cast<const S4*>(s + e);
  D4 xv = __builtin_nontemporal_load(reinterpret_cast<const D4*>(x + e));
  const S4 zv = *reinterpret_cast<const S4*>(z + e);
  const int st = sc->converged;
  if (st == STOP_ABORT) return;
  if (st == STOP_CONVERGED && sc->conv_iter != iter) return;
  const double alpha = sc->alpha;
  const S beta = (S)sc->beta;
#pragma unroll
  for (int q = 0; q < 4; q++) xv[q] = xv[q] + alpha * (double)sv[q];
  __builtin_nontemporal_store(xv, reinterpret_cast<D4*>(x + e));
  if (st == STOP_RUN) {
    S4 nv;
#pragma unroll
    for (int q = 0; q < 4; q++) nv[q] = zv[q] + beta * sv[q];
    *reinterpret_cast<S4*>(s + e) = nv;
  }
}

template <typename S>
__global__ void k_copy(int64_t n, const S* __restrict__ a, S* __restrict__ b) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) b[e] = a[e];
}

// layout -> dense for active cells only (inactive dense entries untouched)
template <typename TO>
__global__ void k_unpack(Lay L, const double* __restrict__ v, const uint8_t* __restrict__ nbm, TO* __restrict__ dense) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= L.ne) return;
  if (!(nbm[e] & NB_SELF)) return;
  int B, s, a, b, i, j, k;
  decode(L, e, B, s, a, b, i, j, k);
  dense[k + L.nz * (j + L.ny * i)] = (TO)v[e];
}

__global__ void k_narrow(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) b[e] = (float)a[e];
}
// dense -> layout (zero where the cell is padding or inactive)
__global__ void k_pack_vec(Lay L, const double* __restrict__ dense, const uint8_t* __restrict__ nbm, double* __restrict__ v) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= L.ne) return;
  double x = 0.0;
  if (nbm[e] & NB_SELF) {
    int B, s, a, b, i, j, k;
    decode(L, e, B, s, a, b, i, j, k);
    x = dense[k + L.nz * (j + L.ny * i)];
  }
  v[e] = x;
}

// ---------------------------------------------------------------------------------------------
Lay lay_of(const fs_ctx* c) {
  const PcgState& P = c->pcg;
  Lay L;
  L.nx = c->nx;
  L.ny = c->ny;
  L.nz = c->nz;
  L.nbi = P.nbi;
  L.nbj = P.nbj;
  L.S = P.S;
  L.nlive = P.nlive;
  L.ne = P.ne;
  L.live = P.live_map.as<int>();
  L.bij = P.brick_ij.as<int2>();
  return L;
}

int realloc_if(fs_ctx* c, DBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return FS_OK;
  fs_free(b);
  return fs_alloc(c, b, bytes);
}

// Exchange entries live in uncached device memory (hipDeviceMallocUncached): the sweeps'
// agent-scope stores and polls then bypass the L2, which shortens the brick-to-brick handoff
// (forward sweep ~3 us, backward ~1-2 us faster at 256^3). Memory of this kind is never handed
// back to the runtime once kernels have used it (on this platform, allocations made after such a
// free have held wrong values), so released buffers stay in a per-device pool until the process
// ends and are reused by later solves and contexts.
struct XPool {
  struct Ent {
    int dev;
    void* p;
    size_t bytes;
  };
  std::mutex mu;
  std::vector<Ent> idle;
};
XPool& xpool() {
  static XPool* pool = new XPool;  // never destroyed: its buffers are never freed either
  return *pool;
}
void xbuf_put(fs_ctx* c, DBuf& b) {
  if (!b.p) return;
  (void)hipStreamSynchronize(c->stream);  // no queued kernel of this context uses it any more
  XPool& xp = xpool();
  std::lock_guard<std::mutex> lock(xp.mu);
  xp.idle.push_back({c->device, b.p, b.bytes});
  b.p = nullptr;
  b.bytes = 0;
}
int xbuf_get(fs_ctx* c, DBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return FS_OK;
  xbuf_put(c, b);
  XPool& xp = xpool();
  std::lock_guard<std::mutex> lock(xp.mu);
  int best = -1;
  for (int m = 0; m < (int)xp.idle.size(); m++)
    if (xp.idle[m].dev == c->device && xp.idle[m].bytes >= bytes &&
        (best < 0 || xp.idle[m].bytes < xp.idle[best].bytes))
      best = m;
  if (best >= 0) {
    b.p = xp.idle[best].p;
    b.bytes = xp.idle[best].bytes;
    xp.idle.erase(xp.idle.begin() + best);
    return FS_OK;
  }
  const hipError_t e = hipExtMallocWithFlags(&b.p, bytes, hipDeviceMallocUncached);
  if (e != hipSuccess) {
    b.p = nullptr;
    c->set_error(std::string("hipExtMallocWithFlags(") + std::to_string(bytes) + ", uncached): " +
                 hipGetErrorString(e));
    return FS_DEVICE_ERROR;
  }
  b.bytes = bytes;
  return FS_OK;
}

// Build the brick layout from the assembled (dense) system and pack it.
int pcg_build(fs_ctx* c) {
  PcgState& P = c->pcg;
  P.nbi = (c->nx + 7) / 8;
  P.nbj = (c->ny + 7) / 8;
  P.S = ((c->nz + 14 + U - 1) / U) * U;
  const int nb = P.nbi * P.nbj;
  FS_TRY(realloc_if(c, P.brick_flag, (nb + 1) * sizeof(int)));
  FS_TRY(realloc_if(c, P.live_map, nb * sizeof(int)));
  if (!P.scal.p) {
    FS_TRY(fs_alloc(c, P.scal, sizeof(PcgScalars)));
    FS_HIP(c, hipMemsetAsync(P.scal.p, 0, sizeof(PcgScalars), c->stream));
    FS_HIP(c, hipHostMalloc((void**)&P.host_scal, sizeof(PcgScalars), hipHostMallocDefault));
    FS_HIP(c, hipHostMalloc((void**)&P.snap, 2 * sizeof(PcgScalars), hipHostMallocDefault));
    for (int q = 0; q < 2; q++) FS_HIP(c, hipEventCreateWithFlags(&P.sev[q], hipEventDisableTiming));
    P.sev_ok = true;
    FS_TRY(fs_alloc(c, P.tickets, 4 * sizeof(uint64_t)));
    FS_HIP(c, hipMemsetAsync(P.tickets.p, 0, 4 * sizeof(uint64_t), c->stream));
    for (int q = 0; q < 3; q++) P.ticket_base[q] = 0;
  }
  FS_HIP(c, hipMemsetAsync(P.brick_flag.p, 0, (nb + 1) * sizeof(int), c->stream));
  k_brick_flags<<<fs_blocks((int64_t)c->nx * c->ny, 256), 256, 0, c->stream>>>(
      c->nx, c->ny, c->nz, c->active.as<uint8_t>(), P.brick_flag.as<int>(), P.nbj, P.brick_flag.as<int>() + nb);
  FS_HIP(c, hipGetLastError());
  std::vector<int> flags(nb + 1);
  FS_HIP(c, hipMemcpyAsync(flags.data(), P.brick_flag.p, (nb + 1) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  c->n_active_last = flags[nb];
  P.live_host.assign(nb, -1);
  std::vector<int2> bij;
  for (int I = 0; I < P.nbi; I++)
    for (int J = 0; J < P.nbj; J++)
      if (flags[I * P.nbj + J]) {
        P.live_host[I * P.nbj + J] = (int)bij.size();
        bij.push_back(make_int2(I, J));
      }
  P.nlive = (int)bij.size();
  P.ne = (int64_t)P.nlive * P.S * 64;
  if (P.ne * (int64_t)sizeof(double) >= ((int64_t)1 << 32)) {
    c->set_error("pressure system too large for 32-bit byte offsets");
    return FS_BAD_ARG;
  }
  P.nchunks = (int)((P.ne + CH - 1) / CH);
  // dependency order: anti-diagonals I+J ascending (then I); backward uses the reverse
  std::vector<int> ord(P.nlive);
  for (int q = 0; q < P.nlive; q++) ord[q] = q;
  std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) {
    int du = bij[u].x + bij[u].y, dv = bij[v].x + bij[v].y;
    return du < dv;
  });
  std::vector<int> rord(ord.rbegin(), ord.rend());
  size_t nl = std::max(P.nlive, 1);
  FS_TRY(realloc_if(c, P.brick_ij, nl * sizeof(int2)));
  FS_TRY(realloc_if(c, P.fwd_order, nl * sizeof(int)));
  FS_TRY(realloc_if(c, P.bwd_order, nl * sizeof(int)));
  const size_t nent = 2 * nl * (size_t)P.S * 8, nx_ = nent + 64;  // entries + zero row (padded)
  FS_TRY(xbuf_get(c, P.xf, nx_ * sizeof(double)));
  FS_TRY(xbuf_get(c, P.xb, nx_ * sizeof(double)));
  if (P.nlive > 0) {
    FS_HIP(c, hipMemcpyAsync(P.live_map.p, P.live_host.data(), nb * sizeof(int), hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipMemcpyAsync(P.brick_ij.p, bij.data(), P.nlive * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipMemcpyAsync(P.fwd_order.p, ord.data(), P.nlive * sizeof(int), hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipMemcpyAsync(P.bwd_order.p, rord.data(), P.nlive * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  // every exchange entry starts as the sentinel
  k_fill_sentinel<<<fs_blocks((int64_t)nent, 256), 256, 0, c->stream>>>((int64_t)nent, P.xf.as<double>());
  k_fill_sentinel<<<fs_blocks((int64_t)nent, 256), 256, 0, c->stream>>>((int64_t)nent, P.xb.as<double>());
  FS_HIP(c, hipMemsetAsync(P.xf.as<double>() + nent, 0, 64 * sizeof(double), c->stream));
  FS_HIP(c, hipMemsetAsync(P.xb.as<double>() + nent, 0, 64 * sizeof(double), c->stream));
  FS_HIP(c, hipGetLastError());
  const size_t ne = (size_t)std::max<int64_t>(P.ne, 1);
  for (DBuf* d : {&P.adiag, &P.al, &P.ad, &P.ab, &P.rhs, &P.e, &P.cl, &P.cd, &P.cb, &P.dr, &P.du, &P.df, &P.x, &P.r,
                  &P.s, &P.t, &P.q, &P.z})
    FS_TRY(realloc_if(c, *d, ne * sizeof(double)));
  if (c->mixed)
    for (DBuf* d : {&P.f_adiag, &P.f_al, &P.f_ad, &P.f_ab, &P.f_e, &P.f_cl, &P.f_cd, &P.f_cb, &P.f_dr, &P.f_du,
                    &P.f_df, &P.f_r, &P.f_s, &P.f_t, &P.f_q, &P.f_z})
      FS_TRY(realloc_if(c, *d, ne * sizeof(float)));
  FS_TRY(realloc_if(c, P.nbm, ne));
  FS_TRY(realloc_if(c, P.tflags, ne * sizeof(uint16_t)));
  FS_TRY(realloc_if(c, P.tnum, 6 * ne * sizeof(double)));
  size_t np_ = std::max<size_t>(std::max(P.nchunks, P.nlive), 1);
  FS_TRY(realloc_if(c, P.partials, np_ * sizeof(double)));
  FS_TRY(realloc_if(c, P.partials2, np_ * sizeof(double)));
  // reduction entries start unset (every reduction returns the entries it used to the sentinel)
  k_fill_sentinel<<<fs_blocks((int64_t)np_, 256), 256, 0, c->stream>>>((int64_t)np_, P.partials.as<double>());
  k_fill_sentinel<<<fs_blocks((int64_t)np_, 256), 256, 0, c->stream>>>((int64_t)np_, P.partials2.as<double>());
  if (P.nlive > 0) {
    Lay L = lay_of(c);
    // Aneighbour order Left, Right, Up, Down, Front, Back. The packed system is FP64 (exact
    // for an FP32 assembly); the mixed-precision solve also gets FP32 copies of the stencil.
    auto pack = [&](auto* tag) {
      using TI = std::remove_pointer_t<decltype(tag)>;
      k_pack<TI><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(
          L, c->cfg.mic_tau, c->adiag.as<TI>(), c->an[0].as<TI>(), c->an[3].as<TI>(), c->an[5].as<TI>(),
          c->rhs.as<TI>(), c->active.as<uint8_t>(), P.adiag.as<double>(), P.al.as<double>(), P.ad.as<double>(),
          P.ab.as<double>(), P.rhs.as<double>(), P.nbm.as<uint8_t>(), P.tflags.as<uint16_t>(), P.tnum.as<double>(),
          P.x.as<double>(), P.q.as<double>(), P.z.as<double>());
    };
    if (c->mixed) {
      pack((float*)nullptr);
      const DBuf* src[4] = {&P.adiag, &P.al, &P.ad, &P.ab};
      DBuf* dst[4] = {&P.f_adiag, &P.f_al, &P.f_ad, &P.f_ab};
      for (int q = 0; q < 4; q++)
        k_narrow<<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(P.ne, src[q]->as<double>(), dst[q]->as<float>());
    } else {
      pack((double*)nullptr);
    }
    FS_HIP(c, hipGetLastError());
  }
  P.built = true;
  return FS_OK;
}

// tickets: one counter per sweep kind.
enum { SW_FACTOR = 0, SW_FWD = 1, SW_BWD = 2 };

// Typed views of the packed system and PCG vectors: FP64 arrays, or their FP32 copies in the
// mixed-precision solve. The iterate x is FP64 in both.
template <typename S>
struct Sys {
  S *adiag, *al, *ad, *ab, *e, *cl, *cd, *cb, *dr, *du, *df, *r, *s, *t, *q, *z;
};
template <typename S>
Sys<S> sys_of(PcgState& P);
template <>
Sys<double> sys_of<double>(PcgState& P) {
  return {P.adiag.as<double>(), P.al.as<double>(), P.ad.as<double>(), P.ab.as<double>(), P.e.as<double>(),
          P.cl.as<double>(),    P.cd.as<double>(), P.cb.as<double>(), P.dr.as<double>(), P.du.as<double>(),
          P.df.as<double>(),    P.r.as<double>(),  P.s.as<double>(),  P.t.as<double>(),  P.q.as<double>(),
          P.z.as<double>()};
}
template <>
Sys<float> sys_of<float>(PcgState& P) {
  return {P.f_adiag.as<float>(), P.f_al.as<float>(), P.f_ad.as<float>(), P.f_ab.as<float>(), P.f_e.as<float>(),
          P.f_cl.as<float>(),    P.f_cd.as<float>(), P.f_cb.as<float>(), P.f_dr.as<float>(), P.f_du.as<float>(),
          P.f_df.as<float>(),    P.f_r.as<float>(),  P.f_s.as<float>(),  P.f_t.as<float>(),  P.f_q.as<float>(),
          P.f_z.as<float>()};
}

// MIC factor (FP64) and the sweep coefficients in the storage type S.
template <typename S>
int pcg_factor(fs_ctx* c) {
  PcgState& P = c->pcg;
  Lay L = lay_of(c);
  uint64_t* tk = P.tickets.as<uint64_t>();
  const int pf = fs_prof_begin(c, FS_KERNEL_MIC_FACTOR, 90.0 * c->n_active_last);
  k_factor_sweep<<<P.nlive, 64, 0, c->stream>>>(L, c->cfg.mic_sigma, P.adiag.as<double>(), P.al.as<double>(),
                                                P.ad.as<double>(), P.ab.as<double>(), P.tflags.as<uint16_t>(),
                                                P.tnum.as<double>(), P.e.as<double>(), P.xf.as<double>(),
                                                tk + SW_FACTOR, P.ticket_base[SW_FACTOR], P.fwd_order.as<int>(),
                                                P.scal.as<PcgScalars>());
  fs_prof_end(c, pf);
  P.ticket_base[SW_FACTOR] += P.nlive;
  FS_HIP(c, hipGetLastError());
  Sys<S> y = sys_of<S>(P);
  const bool narrow = !std::is_same<S, double>::value;
  k_coef<S><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, P.al.as<double>(), P.ad.as<double>(), P.ab.as<double>(),
                                                          P.nbm.as<uint8_t>(), P.e.as<double>(), y.cl, y.cd, y.cb,
                                                          y.dr, y.du, y.df, narrow ? y.e : nullptr);
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

// fwd sweep: UPDATE applies r -= alpha t first.
template <typename S>
int launch_fwd(fs_ctx* c, bool update, int iter) {
  PcgState& P = c->pcg;
  Lay L = lay_of(c);
  Sys<S> y = sys_of<S>(P);
  uint64_t* tk = P.tickets.as<uint64_t>();
  auto go = [&](auto kern) {
    kern<<<P.nlive, 256, 0, c->stream>>>(L, y.r, y.t, y.cl, y.cd, y.cb, y.e, y.q, P.xf.as<double>(),
                                        P.partials2.as<double>(), P.partials.as<double>(), tk + SW_FWD,
                                        P.ticket_base[SW_FWD], P.fwd_order.as<int>(), P.scal.as<PcgScalars>(), iter,
                                        c->cfg.tol_mode, c->cfg.cg_tolerance);
  };
  const double bytes = (update ? 8.0 : 6.0) * sizeof(S) * c->n_active_last;
  const int pf = fs_prof_begin(c, FS_KERNEL_MIC_FWD, bytes);
  if (update) go(k_fwd<true, S>);
  else go(k_fwd<false, S>);
  fs_prof_end(c, pf);
  P.ticket_base[SW_FWD] += P.nlive;
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

template <typename S>
int launch_bwd(fs_ctx* c, bool init) {
  PcgState& P = c->pcg;
  Lay L = lay_of(c);
  Sys<S> y = sys_of<S>(P);
  uint64_t* tk = P.tickets.as<uint64_t>();
  auto go = [&](auto kern) {
    kern<<<P.nlive, 256, 0, c->stream>>>(L, y.q, y.dr, y.du, y.df, y.e, y.r, y.z, P.xb.as<double>(),
                                        P.partials.as<double>(), tk + SW_BWD, P.ticket_base[SW_BWD],
                                        P.bwd_order.as<int>(), P.scal.as<PcgScalars>());
  };
  const int pf = fs_prof_begin(c, FS_KERNEL_MIC_BWD, 7.0 * sizeof(S) * c->n_active_last);
  if (init) go(k_bwd<true, S>);
  else go(k_bwd<false, S>);
  fs_prof_end(c, pf);
  P.ticket_base[SW_BWD] += P.nlive;
  FS_HIP(c, hipGetLastError());
  return FS_OK;
}

int read_scalars(fs_ctx* c) {
  PcgState& P = c->pcg;
  FS_HIP(c, hipMemcpyAsync(P.host_scal, P.scal.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  return FS_OK;
}

// CgSolver3D::solve (project-solver.cc:308-346) in storage type S.
//   S = double: the reference's PCG in FP64, fixed-tree dot products.
//   S = float:  vectors, stencil and sweeps in FP32 with FP64 dot products and norms; the
//               iterate x is accumulated in FP64, and the FP32 recurrence residual is replaced
//               by the FP64 residual b - A x whenever it has dropped tenfold since the last
//               replacement and whenever it claims convergence, so the stopping test is always
//               made on the FP64 residual ("reliable updates"; Clark et al., Comput. Phys.
//               Commun. 181 (2010) 1517). A false convergence restarts the search direction
//               from the replaced residual.
template <typename S>
int solve_t(fs_ctx* c, fs_stats* st) {
  constexpr bool MIXED = !std::is_same<S, double>::value;
  FS_TRY(pcg_build(c));
  PcgState& P = c->pcg;
  FS_HIP(c, hipMemsetAsync(c->pg.p, 0, (size_t)c->n_cell * c->rsz, c->stream));
  int iters = 0, refinements = 0;
  int status = FS_OK;
  PcgScalars* hs = P.host_scal;
  if (P.nlive == 0) {
    std::memset(hs, 0, sizeof(*hs));
  } else {
    Lay L = lay_of(c);
    Sys<S> y = sys_of<S>(P);
    const double tol = c->cfg.cg_tolerance;
    const int tmode = c->cfg.tol_mode;
    k_init<S><<<P.nchunks, 256, 0, c->stream>>>(L, P.rhs.as<double>(), y.r, P.partials.as<double>(),
                                                P.partials2.as<double>(), P.nchunks, P.scal.as<PcgScalars>(), tmode,
                                                tol);
    FS_HIP(c, hipGetLastError());
    FS_TRY(read_scalars(c));
    if (hs->converged == 0) {
      auto restart = [&]() -> int {  // z = M^-1 r, sigma = z.r, s = z
        FS_TRY(launch_fwd<S>(c, false, 0));
        FS_TRY(launch_bwd<S>(c, true));
        k_copy<S><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(P.ne, y.z, y.s);
        return FS_OK;
      };
      FS_TRY(pcg_factor<S>(c));
      FS_TRY(restart());
      const int maxit = c->cfg.cg_max_iterations;
      const int chk = std::max(1, c->cfg.check_interval);
      double last_norm = std::sqrt(hs->rnorm2sq);
      int it = 1;
      bool stopped = false;
      auto batch = [&]() -> int {  // chk iterations (fewer at the cap)
        for (int m = 0; m < chk && it < maxit; m++, it++) {
          int pf = fs_prof_begin(c, FS_KERNEL_SPMV, 6.0 * sizeof(S) * c->n_active_last);
          k_spmv<true, S><<<P.nchunks, 256, 0, c->stream>>>(L, y.s, y.adiag, y.al, y.ad, y.ab, P.nbm.as<uint8_t>(),
                                                            y.t, P.partials.as<double>(), P.nchunks,
                                                            P.scal.as<PcgScalars>(), it);
          fs_prof_end(c, pf);
          FS_TRY(launch_fwd<S>(c, true, it));
          FS_TRY(launch_bwd<S>(c, false));
          pf = fs_prof_begin(c, FS_KERNEL_PCG_UPDATE, (16.0 + 3.0 * sizeof(S)) * c->n_active_last);
          k_update<S><<<fs_blocks(P.ne / 4, 256), 256, 0, c->stream>>>(P.ne, P.x.as<double>(), y.s, y.z,
                                                                   P.scal.as<PcgScalars>(), it);
          fs_prof_end(c, pf);
        }
        FS_HIP(c, hipGetLastError());
        return FS_OK;
      };
      // The host checks the scalars after every batch. Unless kernels are timed one by one, the
      // next batch is queued before the host waits for the previous batch's snapshot, so the
      // device does not idle while the host reads and decides; once the device has stopped (a
      // converged or aborted solve), the queued iterations return at once.
      const bool ahead = !(c->cfg.profile_stages >= 2 && c->ktime_on);
      auto snapshot = [&](int b) -> int {
        FS_HIP(c, hipMemcpyAsync(&P.snap[b], P.scal.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipEventRecord(P.sev[b], c->stream));
        return FS_OK;
      };
      int pend = -1;  // snapshot slot of the queued batch checked next (-1: none queued)
      while (!stopped) {
        if (pend < 0) {  // nothing queued (first batch, after a synchronous step, or not ahead)
          if (it >= maxit) break;
          FS_TRY(batch());
          if (ahead) {
            FS_TRY(snapshot(0));
            pend = 0;
          } else {
            FS_TRY(read_scalars(c));
            fs_prof_flush(c);
          }
        }
        if (ahead) {
          const bool more = it < maxit;
          if (more) {
            FS_TRY(batch());
            FS_TRY(snapshot(1 - pend));
          }
          FS_HIP(c, hipEventSynchronize(P.sev[pend]));
          *hs = P.snap[pend];
          pend = more ? 1 - pend : -1;
        }
        // hs: the state after the checked batch (one more batch may be queued behind it)
        if (hs->wait_timeouts != 0) break;
        if (MIXED && hs->converged != STOP_ABORT &&
            (hs->converged == STOP_CONVERGED || std::sqrt(hs->rnorm2sq) < 0.1 * last_norm)) {
          if (pend >= 0) {  // decide on the state after the queued batch as well
            FS_TRY(read_scalars(c));
            pend = -1;
            if (hs->wait_timeouts != 0) break;
            if (hs->converged == STOP_ABORT) {
              stopped = true;
              break;
            }
          }
          const bool claimed = hs->converged == STOP_CONVERGED;
          const int at = claimed ? hs->conv_iter : it - 1;
          const int pf = fs_prof_begin(c, FS_KERNEL_PCG_UPDATE, (8.0 * 7 + 4.0) * c->n_active_last);
          k_resid<<<P.nchunks, 256, 0, c->stream>>>(L, P.rhs.as<double>(), P.x.as<double>(), P.adiag.as<double>(),
                                                    P.al.as<double>(), P.ad.as<double>(), P.ab.as<double>(),
                                                    P.nbm.as<uint8_t>(), P.f_r.as<float>(),
                                                    P.partials.as<double>(), P.partials2.as<double>(), P.nchunks,
                                                    P.scal.as<PcgScalars>(), at, tmode, tol);
          fs_prof_end(c, pf);
          FS_HIP(c, hipGetLastError());
          FS_TRY(read_scalars(c));
          refinements++;
          last_norm = std::sqrt(hs->rnorm2sq);
          if (hs->converged == STOP_CONVERGED) {
            stopped = true;
          } else if (claimed) {
            it = at + 1;
            FS_TRY(restart());
          }
          continue;
        }
        if (hs->converged != 0) {
          stopped = true;
          break;
        }
        if (ahead && pend < 0) break;  // the cap: the last batch has been checked
      }
      if (hs->wait_timeouts != 0) {
        (void)hipStreamSynchronize(c->stream);  // a queued batch gives up too
        c->set_error("PCG sweep: a brick dependency wait exceeded its bound");
        return FS_DEVICE_ERROR;
      }
      if (hs->converged == 1 || hs->converged == 2) iters = hs->conv_iter;
      else iters = maxit;
      if (hs->converged == 2) status = hs->status;
      else if (hs->converged == 0) status = FS_NOT_CONVERGED;
    }
    if (MIXED)
      k_unpack<float><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, P.x.as<double>(), P.nbm.as<uint8_t>(),
                                                                   c->pg.as<float>());
    else
      k_unpack<double><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, P.x.as<double>(), P.nbm.as<uint8_t>(),
                                                                    c->pg.as<double>());
    FS_HIP(c, hipGetLastError());
    FS_HIP(c, hipStreamSynchronize(c->stream));
  }
  if (st) {
    st->cg_iterations = iters;
    st->residual_inf = hs->residual;
    st->rhs_norm2 = std::sqrt(hs->bnorm2sq);
    st->residual_rel = hs->bnorm2sq > 0 ? std::sqrt(hs->rnorm2sq / hs->bnorm2sq) : 0.0;
    st->n_active = (int32_t)c->n_active_last;
    st->refinements = refinements;
  }
  return status;
}

}  // namespace

int fs_stage_solve(fs_ctx* c, fs_stats* st) {
  if (c->cfg.preconditioner == FS_PRECOND_MULTIGRID) return fs_mg_solve(c, st);
  return c->mixed ? solve_t<float>(c, st) : solve_t<double>(c, st);
}

// applyCompressedMatrix on the assembled system for host vectors (dense); inactive entries of
// y are left as given.
int fs_pcg_spmv_dense(fs_ctx* c, const double* x, double* y) {
  FS_TRY(pcg_build(c));
  PcgState& P = c->pcg;
  if (P.nlive == 0) return FS_OK;
  Lay L = lay_of(c);
  DBuf dx;
  FS_TRY(fs_alloc(c, dx, (size_t)c->n_cell * sizeof(double)));
  FS_HIP(c, hipMemcpyAsync(dx.p, x, (size_t)c->n_cell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_pack_vec<<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, dx.as<double>(), P.nbm.as<uint8_t>(), P.s.as<double>());
  k_spmv<false, double><<<P.nchunks, 256, 0, c->stream>>>(L, P.s.as<double>(), P.adiag.as<double>(), P.al.as<double>(),
                                                  P.ad.as<double>(), P.ab.as<double>(), P.nbm.as<uint8_t>(),
                                                  P.t.as<double>(), P.partials.as<double>(), P.nchunks,
                                                  P.scal.as<PcgScalars>(), 0);
  FS_HIP(c, hipMemcpyAsync(dx.p, y, (size_t)c->n_cell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_unpack<double><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, P.t.as<double>(), P.nbm.as<uint8_t>(), dx.as<double>());
  FS_HIP(c, hipGetLastError());
  FS_HIP(c, hipMemcpyAsync(y, dx.p, (size_t)c->n_cell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_free(dx);
  return FS_OK;
}

// ModifiedIncompleteCholesky3D::precond for a host vector r: z = M^-1 r on active cells.
int fs_pcg_mic_dense(fs_ctx* c, const double* r, double* z) {
  FS_TRY(pcg_build(c));
  PcgState& P = c->pcg;
  if (P.nlive == 0) return FS_OK;
  Lay L = lay_of(c);
  DBuf dx;
  FS_TRY(fs_alloc(c, dx, (size_t)c->n_cell * sizeof(double)));
  FS_HIP(c, hipMemcpyAsync(dx.p, r, (size_t)c->n_cell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_pack_vec<<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, dx.as<double>(), P.nbm.as<uint8_t>(), P.r.as<double>());
  FS_TRY(pcg_factor<double>(c));
  FS_TRY(launch_fwd<double>(c, false, 0));
  FS_TRY(launch_bwd<double>(c, true));
  FS_HIP(c, hipMemcpyAsync(dx.p, z, (size_t)c->n_cell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_unpack<double><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, P.z.as<double>(), P.nbm.as<uint8_t>(), dx.as<double>());
  FS_HIP(c, hipGetLastError());
  FS_HIP(c, hipMemcpyAsync(z, dx.p, (size_t)c->n_cell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  PcgScalars hs;
  FS_HIP(c, hipMemcpy(&hs, P.scal.p, sizeof(hs), hipMemcpyDeviceToHost));
  fs_free(dx);
  if (hs.wait_timeouts) {
    c->set_error("MIC sweep: a brick dependency wait exceeded its bound");
    return FS_DEVICE_ERROR;
  }
  return FS_OK;
}

// MIC factor E of the assembled system (dense; zero on inactive cells).
int fs_pcg_export_e(fs_ctx* c, double* out) {
  FS_TRY(pcg_build(c));
  PcgState& P = c->pcg;
  std::memset(out, 0, (size_t)c->n_cell * sizeof(double));
  if (P.nlive == 0) return FS_OK;
  Lay L = lay_of(c);
  FS_TRY(pcg_factor<double>(c));
  DBuf dx;
  FS_TRY(fs_alloc(c, dx, (size_t)c->n_cell * sizeof(double)));
  FS_HIP(c, hipMemsetAsync(dx.p, 0, (size_t)c->n_cell * sizeof(double), c->stream));
  k_unpack<double><<<fs_blocks(P.ne, 256), 256, 0, c->stream>>>(L, P.e.as<double>(), P.nbm.as<uint8_t>(), dx.as<double>());
  FS_HIP(c, hipGetLastError());
  FS_HIP(c, hipMemcpyAsync(out, dx.p, (size_t)c->n_cell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_free(dx);
  return FS_OK;
}

void fs_pcg_release(fs_ctx* c) {
  PcgState& P = c->pcg;
  for (DBuf* d : {&P.live_map, &P.brick_ij, &P.fwd_order, &P.bwd_order, &P.brick_flag, &P.adiag, &P.al, &P.ad, &P.ab,
                  &P.nbm, &P.rhs, &P.e, &P.cl, &P.cd, &P.cb, &P.dr, &P.du, &P.df, &P.tflags, &P.tnum, &P.x, &P.r,
                  &P.s, &P.t, &P.q, &P.z, &P.partials, &P.partials2, &P.scal, &P.tickets, &P.f_adiag,
                  &P.f_al, &P.f_ad, &P.f_ab, &P.f_e, &P.f_cl, &P.f_cd, &P.f_cb, &P.f_dr, &P.f_du, &P.f_df, &P.f_r,
                  &P.f_s, &P.f_t, &P.f_q, &P.f_z, &P.f_x})
    fs_free(*d);
  xbuf_put(c, P.xf);
  xbuf_put(c, P.xb);
  if (P.host_scal) (void)hipHostFree(P.host_scal);
  P.host_scal = nullptr;
  if (P.snap) (void)hipHostFree(P.snap);
  P.snap = nullptr;
  if (P.sev_ok)
    for (int q = 0; q < 2; q++) (void)hipEventDestroy(P.sev[q]);
  P.sev_ok = false;
}
