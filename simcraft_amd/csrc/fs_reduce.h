// fs_reduce.h - fixed-order reductions and ready-flag exchange shared by the pressure solvers
// (fs_pcg.hip, fs_mg.hip). Every reduction adds in one fixed tree, so results do not depend on
// scheduling and a CPU restatement of the tree reproduces them bit for bit.
#pragma once
#include "fs_ctx.h"

namespace {

constexpr long SPIN_LIMIT = 1000000;  // bounded wait (~1 s), then the solve reports a device error
enum { STOP_RUN = 0, STOP_CONVERGED = 1, STOP_ABORT = 2 };

// ---- fixed-tree reductions -------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = (v < o) ? o : v;
  }
  return v;
}
// 256-thread block: per-thread values -> (w0 + w1) + (w2 + w3) of wave butterflies.
__device__ inline double block_sum256(double v, double* sm) {
  double w = wave_sum(v);
  int wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[wid] = w;
  __syncthreads();
  double r = (sm[0] + sm[1]) + (sm[2] + sm[3]);
  __syncthreads();
  return r;
}
__device__ inline double block_max256(double v, double* sm) {
  double w = wave_max(v);
  int wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[wid] = w;
  __syncthreads();
  double r = sm[0];
  for (int q = 1; q < 4; q++) r = (r < sm[q]) ? sm[q] : r;
  __syncthreads();
  return r;
}
// ---- boundary exchange between bricks ---------------------------------------------------------
// A brick hands its boundary row to the next brick through a small exchange array. Every entry
// holds a sentinel until the producer stores the value with an agent-scope atomic store; the
// consumer polls the entry with agent-scope atomic loads until the sentinel is gone, uses the
// value and stores the sentinel back. The sentinel is a signalling-NaN bit pattern, which
// floating-point arithmetic never returns, so every exchanged value is its own ready flag and
// no fence or cache maintenance is needed. Polls are bounded: after SPIN_LIMIT the wave gives
// up, counts a timeout and the host reports FS_DEVICE_ERROR.
constexpr unsigned long long XSENT = 0x7FF4DEADBEEF0001ull;

__device__ inline double xload(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void xstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool xready(double v) { return (unsigned long long)__double_as_longlong(v) != XSENT; }
__device__ inline double xsent() { return __longlong_as_double((long long)XSENT); }

// Wait until *p holds a written value; false when the wait gave up (the solve then reports
// FS_DEVICE_ERROR and every other waiter gives up within 256 polls).
__device__ inline bool xwait(const double* p, PcgScalars* sc) {
  long spins = 0;
  while (!xready(xload(p))) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255) == 0) {
      int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (to != 0 || spins > SPIN_LIMIT) {
        atomicAdd(&sc->wait_timeouts, 1);
        return false;
      }
    }
  }
  return true;
}

__device__ inline double xpoll(const double* p, double v, PcgScalars* sc) {
  long spins = 0;
  while (!xready(v)) {
    __builtin_amdgcn_s_sleep(1);
    v = xload(p);
    if ((++spins & 255) == 0) {
      int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (to != 0 || spins > SPIN_LIMIT) {
        atomicAdd(&sc->wait_timeouts, 1);
        return 0.0;  // give up: results are invalid and the solve reports FS_DEVICE_ERROR
      }
    }
  }
  return v;
}

// High word of the sentinel. Arithmetic never returns a signalling NaN, so for computed values
// the high word alone identifies the sentinel.
constexpr uint32_t XSENT_HI = (uint32_t)(XSENT >> 32);
__device__ __forceinline__ bool xunset(double v) { return (uint32_t)__double2hiint(v) == XSENT_HI; }

// ---- cross-block reductions --------------------------------------------------------------------
// Every block (in the sweeps: every brick) stores its partials with agent-scope atomic stores into
// entries that otherwise hold the sentinel, so a stored partial is its own ready flag and no
// fence is needed: an agent-scope fence writes back and invalidates the L2 on this GPU, which,
// done by every block, costs more than the kernel's own work. The block dispatched last (the
// brick with the last ticket) waits until every entry is set, reduces them in the fixed order and
// stores the sentinel back, so the entries are unset again for the next kernel.

// Block-wide (256 threads) wait until p[0, n) are all set; false if the wait gave up.
__device__ inline bool gather_wait256(const double* p, int n, PcgScalars* sc) {
  long spins = 0;
  while (true) {
    int mine = 1;
    for (int m = threadIdx.x; m < n; m += 256) mine &= xunset(xload(p + m)) ? 0 : 1;
    if (__syncthreads_and(mine)) return true;
    __builtin_amdgcn_s_sleep(2);
    int give = 0;
    if ((++spins & 63) == 0 && threadIdx.x == 0) {
      const int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      give = (to != 0 || spins > SPIN_LIMIT) ? 1 : 0;
      if (give) atomicAdd(&sc->wait_timeouts, 1);
    }
    if (__syncthreads_or(give)) return false;
  }
}
// Fixed tree over the n gathered entries (thread t sums p[t + 256m] for m ascending, then
// block_sum256), which it returns to the sentinel.
__device__ inline double gather_tree256(double* p, int n, double* sm) {
  double a = 0.0;
  for (int m = threadIdx.x; m < n; m += 256) {
    a += xload(p + m);
    xstore(p + m, xsent());
  }
  return block_sum256(a, sm);
}
// max over gathered entries (order-free: the comparison never takes a NaN), returned to the sentinel
__device__ inline double gather_max256(double* p, int n, double* sm) {
  double mx = 0.0;
  for (int m = threadIdx.x; m < n; m += 256) {
    const double v = xload(p + m);
    mx = (mx < v) ? v : mx;
    xstore(p + m, xsent());
  }
  return block_max256(mx, sm);
}
// One wave's wait until p[0, n) are all set; false if the wait gave up.
__device__ inline bool gather_wait_wave(const double* p, int n, PcgScalars* sc) {
  const int lane = threadIdx.x & 63;
  long spins = 0;
  while (true) {
    bool mine = true;
    for (int m = lane; m < n; m += 64) mine &= !xunset(xload(p + m));
    if (__builtin_amdgcn_ballot_w64(!mine) == 0) return true;
    __builtin_amdgcn_s_sleep(2);
    if ((++spins & 63) == 0) {
      const int to = __hip_atomic_load(&sc->wait_timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (to != 0 || spins > SPIN_LIMIT) {
        if (lane == 0) atomicAdd(&sc->wait_timeouts, 1);
        return false;
      }
    }
  }
}
// The same tree evaluated by one 64-lane wave (lane l plays threads l, l+64, l+128, l+192).
__device__ inline double gather_tree256_wave(double* p, int n) {
  double w[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    double a = 0.0;
    for (int m = (threadIdx.x & 63) + 64 * v; m < n; m += 256) a += xload(p + m);
    w[v] = wave_sum(a);
  }
  for (int m = threadIdx.x & 63; m < n; m += 64) xstore(p + m, xsent());
  return (w[0] + w[1]) + (w[2] + w[3]);
}
__device__ inline double gather_max_wave(double* p, int n) {
  double mx = 0.0;
  for (int m = threadIdx.x & 63; m < n; m += 64) {
    const double v = xload(p + m);
    mx = (mx < v) ? v : mx;
    xstore(p + m, xsent());
  }
  return wave_max(mx);
}

__global__ void k_fill_sentinel(int64_t n, double* __restrict__ x) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = __longlong_as_double((long long)XSENT);
}

}  // namespace
