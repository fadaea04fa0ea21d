// fs_device.h - device-side scalar routines of the FLIP step, written for gfx950.
//
// Every routine keeps the operation order of the SimCraft function it cites, so that FP64
// results are identical to a correctly rounded CPU evaluation (the library is compiled with
// -ffp-contract=off; FP64 '/' and sqrt are correctly rounded on gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsd {

struct Geo {
  int nx, ny, nz;     // cells
  double h[3];        // spacing = size / resolution (Spatify/include/Spatify/grids.h:244-252)
  double ext[3];      // nx*h etc. (HybridAdvectionSolver3D width/height/depth)
};

__host__ __device__ inline int lin3(int i, int j, int k, int ny, int nz) { return k + nz * (j + ny * i); }

// fluid::clamp (FluidSim/include/FluidSim/cpu/util.h:21-24)
template <typename T>
__device__ inline T fclamp(T x, T lo, T hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
// std::min / std::max semantics
template <typename T>
__device__ inline T smin(T a, T b) {
  return b < a ? b : a;
}
template <typename T>
__device__ inline T smax(T a, T b) {
  return a < b ? b : a;
}

// TGrid::nearest (grids.h:287-292), origin 0, truncation toward zero.
template <typename T>
__device__ inline int nearest1(T p, T off, T h) {
  return (int)(((p - T(0)) - off * h) / h + T(0.5));
}
// TGrid::indexToCoord (grids.h:300-308)
template <typename T>
__device__ inline T coord1(int i, T off, T h) {
  return ((T)i + off) * h + T(0);
}

// Quotient a / b for interpolation weights: IEEE division in FP64; in FP32 (the mixed-precision
// mode) a times the hardware reciprocal of b (v_rcp_f32, 1 ulp), within about 2 ulp of the
// quotient, in place of the ten-instruction correctly rounded division. Cell and bin indices
// keep the exact division in both precisions.
template <typename T>
__device__ inline T wdiv(T a, T b) {
  return a / b;
}
template <>
__device__ inline float wdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

// Cubic B-spline kernel assumed for the absent Core/transfer-stencil.h CubicKernel.
template <typename T>
__device__ inline T bspline(T r) {
  if (r < T(1)) return (T(2) / T(3) - r * r) + T(0.5) * r * r * r;
  if (r < T(2)) {
    T t = T(2) - r;
    return sizeof(T) == 4 ? t * t * t * T(1.0 / 6.0) : t * t * t / T(6);
  }
  return T(0);
}

// dN/dr of bspline (r >= 0); N is C2, so the branches agree at r = 1 and r = 2.
template <typename T>
__device__ inline T bspline_d(T r) {
  if (r < T(1)) return T(1.5) * r * r - T(2) * r;
  if (r < T(2)) {
    T t = T(2) - r;
    return T(-0.5) * t * t;
  }
  return T(0);
}

// Morton code (Spatify/include/Spatify/ns-util.h:36-45)
__host__ __device__ inline int morton10(int x) {
  x = (x | (x << 16)) & 0x030000FF;
  x = (x | (x << 8)) & 0x0300F00F;
  x = (x | (x << 4)) & 0x030C30C3;
  x = (x | (x << 2)) & 0x09249249;
  return x;
}
__host__ __device__ inline int morton3D(int x, int y, int z) {
  return morton10(x) | (morton10(y) << 1) | (morton10(z) << 2);
}

template <typename T>
struct V3 {
  T x, y, z;
};

// SDF::eval (FluidSim/include/FluidSim/cpu/sdf.h:35-61) on a cell-centred grid: trilinear
// weights with the lower node weighted by t (mirrored).
template <typename T>
__device__ inline T sdf_eval(const T* __restrict__ g, int nx, int ny, int nz, const T h[3], T px, T py, T pz) {
  int ix = (int)((px - T(0)) / h[0] - T(0.5));
  int iy = (int)((py - T(0)) / h[1] - T(0.5));
  int iz = (int)((pz - T(0)) / h[2] - T(0.5));
  int x = fclamp(ix, 0, nx - 1), y = fclamp(iy, 0, ny - 1), z = fclamp(iz, 0, nz - 1);
  int xn = fclamp(ix + 1, 0, nx - 1), yn = fclamp(iy + 1, 0, ny - 1), zn = fclamp(iz + 1, 0, nz - 1);
  T cx = coord1(x, T(0.5), h[0]), cy = coord1(y, T(0.5), h[1]), cz = coord1(z, T(0.5), h[2]);
  T tx = fclamp(wdiv(px - cx, h[0]), T(0), T(1));
  T ty = fclamp(wdiv(py - cy, h[1]), T(0), T(1));
  T tz = fclamp(wdiv(pz - cz, h[2]), T(0), T(1));
  T r = tx * ty * tz * g[lin3(x, y, z, ny, nz)];
  r = r + tx * ty * (1 - tz) * g[lin3(x, y, zn, ny, nz)];
  r = r + tx * (1 - ty) * tz * g[lin3(x, yn, z, ny, nz)];
  r = r + tx * (1 - ty) * (1 - tz) * g[lin3(x, yn, zn, ny, nz)];
  r = r + (1 - tx) * ty * tz * g[lin3(xn, y, z, ny, nz)];
  r = r + (1 - tx) * ty * (1 - tz) * g[lin3(xn, y, zn, ny, nz)];
  r = r + (1 - tx) * (1 - ty) * tz * g[lin3(xn, yn, z, ny, nz)];
  r = r + (1 - tx) * (1 - ty) * (1 - tz) * g[lin3(xn, yn, zn, ny, nz)];
  return r;
}

// SDF::gradient (sdf.h:62-85)
template <typename T>
__device__ inline V3<T> sdf_gradient(const T* __restrict__ g, int nx, int ny, int nz, const T h[3], int x, int y, int z) {
  int xn = x < nx - 1 ? x + 1 : x, xp = x > 0 ? x - 1 : x;
  int yn = y < ny - 1 ? y + 1 : y, yp = y > 0 ? y - 1 : y;
  int zn = z < nz - 1 ? z + 1 : z, zp = z > 0 ? z - 1 : z;
  T dx = x > 0 && x < nx - 1 ? 2 * h[0] : h[0];
  T dy = y > 0 && y < ny - 1 ? 2 * h[1] : h[1];
  T dz = z > 0 && z < nz - 1 ? 2 * h[2] : h[2];
  T gx = g[lin3(xn, y, z, ny, nz)] - g[lin3(xp, y, z, ny, nz)];
  T gy = g[lin3(x, yn, z, ny, nz)] - g[lin3(x, yp, z, ny, nz)];
  T gz = g[lin3(x, y, zn, ny, nz)] - g[lin3(x, y, zp, ny, nz)];
  return {wdiv(gx, dx), wdiv(gy, dy), wdiv(gz, dz)};
}

// SDF::grad (sdf.h:86-118): mirrored trilinear blend of node gradients, t unclamped.
template <typename T>
__device__ inline V3<T> sdf_grad(const T* __restrict__ g, int nx, int ny, int nz, const T h[3], T px, T py, T pz) {
  int ix = (int)((px - T(0)) / h[0] - T(0.5));
  int iy = (int)((py - T(0)) / h[1] - T(0.5));
  int iz = (int)((pz - T(0)) / h[2] - T(0.5));
  int x = fclamp(ix, 0, nx - 1), y = fclamp(iy, 0, ny - 1), z = fclamp(iz, 0, nz - 1);
  int xn = fclamp(ix + 1, 0, nx - 1), yn = fclamp(iy + 1, 0, ny - 1), zn = fclamp(iz + 1, 0, nz - 1);
  T tx = wdiv(px - coord1(x, T(0.5), h[0]), h[0]);
  T ty = wdiv(py - coord1(y, T(0.5), h[1]), h[1]);
  T tz = wdiv(pz - coord1(z, T(0.5), h[2]), h[2]);
  const int xs[8] = {x, x, x, x, xn, xn, xn, xn};
  const int ys[8] = {y, y, yn, yn, y, y, yn, yn};
  const int zs[8] = {z, zn, z, zn, z, zn, z, zn};
  V3<T> acc = {T(0), T(0), T(0)};
#pragma unroll
  for (int c = 0; c < 8; c++) {
    T wx = (c & 4) ? (1 - tx) : tx;
    T wy = (c & 2) ? (1 - ty) : ty;
    T wz = (c & 1) ? (1 - tz) : tz;
    T w = wx * wy * wz;
    V3<T> gg = sdf_gradient(g, nx, ny, nz, h, xs[c], ys[c], zs[c]);
    if (c == 0) {
      acc = {w * gg.x, w * gg.y, w * gg.z};
    } else {
      acc.x = acc.x + w * gg.x;
      acc.y = acc.y + w * gg.y;
      acc.z = acc.z + w * gg.z;
    }
  }
  return acc;
}

// glm::normalize: v * (1/sqrt(dot(v,v))), dot = (x*x + y*y) + z*z
template <typename T>
__device__ inline V3<T> gnormalize(V3<T> v) {
  T d = v.x * v.x + v.y * v.y + v.z * v.z;
  T s = T(1) / sqrt(d);
  return {v.x * s, v.y * s, v.z * s};
}

// cycle (FluidSim/src/cpu/util.cc:319-325)
template <typename T>
__device__ inline void cycle4(T& a0, T& a1, T& a2, T& a3) {
  T tmp = a0;
  a0 = a3;
  a3 = a2;
  a2 = a1;
  a1 = tmp;
}
// fractionInside (util.cc:333-378)
template <typename T>
__device__ inline T fraction_inside(T lu, T ru, T rd, T ld) {
  int n_negs = (lu <= T(0)) + (ru <= T(0)) + (rd <= T(0)) + (ld <= T(0));
  if (n_negs == 0) return T(0);
  if (n_negs == 4) return T(1);
  for (int c = 0; c < 4 && lu > T(0); c++) cycle4(lu, ru, rd, ld);
  if (n_negs == 1) {
    T fracu = ru / (ru - lu);
    T fracl = ld / (ld - lu);
    return fracu * fracl * T(0.5);
  }
  if (n_negs == 2) {
    if (ru <= T(0)) {
      T fracl = ld / (ld - lu);
      T fracr = rd / (rd - ru);
      return T(0.5) * (fracl + fracr);
    }
    if (ld <= T(0)) {
      T fracu = ru / (ru - lu);
      T fracd = rd / (rd - ld);
      return T(0.5) * (fracu + fracd);
    }
    if (rd <= T(0)) {
      T fracu = ru / (ru - lu);
      T fracd = ld / (ld - rd);
      T fracl = ld / (ld - lu);
      T fracr = ru / (ru - rd);
      if (lu + ru + rd + ld <= T(0))
        return T(1) - T(0.5) * ((T(1) - fracu) * (T(1) - fracl) + (T(1) - fracd) * (T(1) - fracr));
      return T(0.5) * (fracu * fracl + fracd * fracr);
    }
    return T(0);
  }
  for (int c = 0; c < 4 && lu <= T(0); c++) cycle4(lu, ru, rd, ld);
  T fracu = ru / (ru - lu);
  T fracl = ld / (ld - lu);
  return T(1) - T(0.5) * fracu * fracl;
}

}  // namespace fsd
