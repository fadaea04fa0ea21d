// fs_noalloc.hip - the CFL time step of caller-owned device face grids, computed without any
// allocation: the reduction's partial maxima live in a file-scope device array.
//
// fs_cfl (fluid-simulator.cc:184-203) reduces the context's face grids into a buffer that the
// context allocates. fs_cfl_device_noalloc runs the same reduction on face grids the caller
// already holds in device memory and writes the time step to a device address, so it can be
// called where hipMalloc cannot (under stream capture, in a serving loop that must not allocate).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "simcraft_fluid.h"

namespace {

constexpr int kTpb = 256;
constexpr int kMaxPartials = 1024;  // blocks of the first pass = partial maxima kept in scratch

// One slot per block of the first pass (the bit pattern of a non-negative double, so that the
// unsigned order is the numeric order). 8 KiB, held by every device the library runs on until
// the process ends; each call writes every slot it later reads, so nothing is cleared.
__device__ __align__(256) unsigned long long scratch[kMaxPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device. In host code
// the array's name is only a stand-in for the symbol; kernels must be given this address.
int scratch_on_current_device(unsigned long long** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FS_DEVICE_ERROR;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FS_DEVICE_ERROR;
    g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FS_OK;
}

struct FaceGrids {
  const double* g[3];
  int64_t n[3];
};

// Pass 1: block b lowers max |u| over its grid-stride share of the three face grids into
// partial[b].
__global__ __launch_bounds__(kTpb) void k_cfl_partials(FaceGrids f, unsigned long long* __restrict__ partial) {
  __shared__ double sm[kTpb];
  double m = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kTpb;
  for (int a = 0; a < 3; a++) {
    const double* __restrict__ g = f.g[a];
    for (int64_t i = (int64_t)blockIdx.x * kTpb + threadIdx.x; i < f.n[a]; i += stride) {
      const double v = fabs(g[i]);
      m = v > m ? v : m;
    }
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kTpb / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = sm[threadIdx.x + s] > sm[threadIdx.x] ? sm[threadIdx.x + s] : sm[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = (unsigned long long)__double_as_longlong(sm[0]);
}

// Pass 2 (one block): the maximum of the partials, then CFL() = factor * max(1e-3, min(h / 1e-6,
// h / max|u|)) in the operation order of fs_cfl.
__global__ __launch_bounds__(kTpb) void k_cfl_finish(const unsigned long long* __restrict__ partial, int nparts,
                                                     double h, double factor, double* __restrict__ dt_out) {
  __shared__ unsigned long long sm[kTpb];
  unsigned long long m = 0;
  for (int i = threadIdx.x; i < nparts; i += kTpb) m = partial[i] > m ? partial[i] : m;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kTpb / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = sm[threadIdx.x + s] > sm[threadIdx.x] ? sm[threadIdx.x + s] : sm[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mx = __longlong_as_double((long long)sm[0]);
    double cfl = h / 1e-6;
    if (mx > 0.0) {
      const double v = h / mx;
      cfl = v < cfl ? v : cfl;
    }
    *dt_out = factor * (1e-3 < cfl ? cfl : 1e-3);
  }
}

}  // namespace

extern "C" int fs_cfl_device_noalloc(const double* u, const double* v, const double* w, int32_t nx, int32_t ny,
                                     int32_t nz, double h, double cfl_factor, double* dt_out, void* stream) {
  if (!u || !v || !w || !dt_out || nx <= 0 || ny <= 0 || nz <= 0 || !(h > 0.0)) return FS_BAD_ARG;
  // 32-bit face counts per grid at most (the largest grid the library's contexts accept)
  const int64_t nxy = (int64_t)nx * ny, nyz = (int64_t)ny * nz;
  const int64_t n_u = (int64_t)(nx + 1) * nyz, n_v = (int64_t)nx * (ny + 1) * nz, n_w = nxy * (nz + 1);
  if (n_u > INT32_MAX || n_v > INT32_MAX || n_w > INT32_MAX) return FS_BAD_ARG;
  unsigned long long* partial = nullptr;
  const int st = scratch_on_current_device(&partial);
  if (st != FS_OK) return st;
  int64_t most = n_u > n_v ? n_u : n_v;
  most = most > n_w ? most : n_w;
  int64_t blocks = (most + kTpb - 1) / kTpb;
  blocks = blocks < kMaxPartials ? blocks : kMaxPartials;
  hipStream_t s = static_cast<hipStream_t>(stream);
  FaceGrids f{{u, v, w}, {n_u, n_v, n_w}};
  k_cfl_partials<<<(int)blocks, kTpb, 0, s>>>(f, partial);
  k_cfl_finish<<<1, kTpb, 0, s>>>(partial, (int)blocks, h, cfl_factor, dt_out);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_DEVICE_ERROR;
}
