// fs_oracle.cc - CPU FP64 restatement of SimCraft FluidSim's FLIP/PIC substep (the oracle).
//
// TEST INFRASTRUCTURE ONLY. It is loaded by tests/, by __graft_entry__.smoke() as the checker
// and by bench.py's cpu_baseline leg; the product library never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the SimCraft
// tree). The reference does not compile (missing Core/transfer-stencil.h, undefined ERROR,
// core:: vs sim::core namespaces, absent glm/TBB), so this is a restatement written from
// reading the source, in "reference mode": the reference's quirks are reproduced, e.g.
// clear-before-advect, stale ping-pong buffers, the skip-last-particle neighbour loop,
// mirrored trilinear weights and the non-textbook MIC(0) factor.
//
// Choices where the reference is undefined (documented, not pinned by any reference test):
//   * CubicKernel::weight (header absent): tensor-product cubic B-spline N(|d|/h), with
//     N(r) = (2/3 - r^2) + r^3/2 for r<1, (2-r)^3/6 for r<2, else 0.
//   * std::sort tie order in ParticleNeighbourSearcher::update (unstable): stable, i.e. particle
//     index ascending inside each Morton bin.
//   * FLIP P2G is a racy TBB scatter (advect-solver.cc:201-237): each node is summed over
//     particles in (nearest-node key, particle index) order. p2g_order=1 selects plain
//     particle-index order (the serial PIC order, advect-solver.cc:31-67).
//   * PCG dot products: dot_mode=0 is the reference's serial lexicographic sum
//     (util.cc:380-392); dot_mode=1 is the fixed reduction tree of the device layout
//     (column bricks, see pcg_layout below), so device and oracle agree bit for bit.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off). The reference's TBB loops are
// OpenMP loops over the same outer index; its serial loops stay serial.

#include "../include/simcraft_fluid.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <omp.h>
#include <string>
#include <vector>

namespace fso {

// ---------------------------------------------------------------------------------------------
// Spatify data model: dense arrays, k fastest (Spatify/include/Spatify/hash.h:19-28,
// arrays.h:108-263).
template <typename T>
struct Arr {
  int nx = 0, ny = 0, nz = 0;
  std::vector<T> d;
  Arr() = default;
  Arr(int a, int b, int c) : nx(a), ny(b), nz(c), d((size_t)a * b * c, T(0)) {}
  size_t id(int i, int j, int k) const {
    return (size_t)k + (size_t)nz * ((size_t)j + (size_t)ny * (size_t)i);
  }
  T& operator()(int i, int j, int k) { return d[id(i, j, k)]; }
  const T& operator()(int i, int j, int k) const { return d[id(i, j, k)]; }
  void fill(T v) { std::fill(d.begin(), d.end(), v); }
  size_t size() const { return d.size(); }
};
using RArr = Arr<double>;
using BArr = Arr<uint8_t>;

// glm::dvec3 arithmetic as the reference uses it.
struct V3 {
  double x, y, z;
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
// glm compute_dot for vec3: tmp = a*b; (tmp.x + tmp.y) + tmp.z
static inline double gdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::length / glm::distance(p0,p1) = length(p1 - p0)
static inline double glen(V3 a) { return std::sqrt(gdot(a, a)); }
// glm::normalize: v * inversesqrt(dot(v,v)), inversesqrt(x) = 1/sqrt(x)
static inline V3 gnormalize(V3 v) {
  double s = 1.0 / std::sqrt(gdot(v, v));
  return {v.x * s, v.y * s, v.z * s};
}

// fluid::clamp (FluidSim/include/FluidSim/cpu/util.h:21-24)
template <typename T>
static inline T fclamp(T x, T lo, T hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// TGrid::nearest (Spatify/include/Spatify/grids.h:287-292), origin 0:
// ivec3((coord - origin - offset*h)/h + 0.5), truncation toward zero.
static inline int nearest1(double p, double off, double h) {
  return (int)(((p - 0.0) - off * h) / h + 0.5);
}
// TGrid::indexToCoord (grids.h:300-308): (index + offset)*h + origin
static inline double coord1(int i, double off, double h) { return ((double)i + off) * h + 0.0; }
// FaceCentreOffset<3,Axis> (grids.h:479-482): 0 on the face axis, 1/2 elsewhere.
static inline double face_off(int axis, int c) { return axis == c ? 0.0 : 0.5; }

// Q1: CubicKernel::weight (advect-solver.cc:214, util.cc:58) is undefined in the reference.
static inline double bspline(double r) {
  if (r < 1.0) return (2.0 / 3.0 - r * r) + 0.5 * r * r * r;
  if (r < 2.0) {
    double t = 2.0 - r;
    return t * t * t / 6.0;
  }
  return 0.0;
}
static inline double kweight(double hx, double hy, double hz, V3 d) {
  return bspline(std::fabs(d.x) / hx) * bspline(std::fabs(d.y) / hy) * bspline(std::fabs(d.z) / hz);
}

// ---------------------------------------------------------------------------------------------
// SDF<3> on a cell-centred grid (FluidSim/include/FluidSim/cpu/sdf.h:20-190).
struct SDF {
  RArr g;
  double h[3];
  // SDF::eval (sdf.h:35-61): trilinear with the lower node weighted by t (mirrored, Q5).
  double eval(V3 p) const {
    // coordToSampledCellIndex (grids.h:297-299): ivec3((coord - origin)/h - offset)
    int ix = (int)((p.x - 0.0) / h[0] - 0.5);
    int iy = (int)((p.y - 0.0) / h[1] - 0.5);
    int iz = (int)((p.z - 0.0) / h[2] - 0.5);
    int x = fclamp(ix, 0, g.nx - 1), y = fclamp(iy, 0, g.ny - 1), z = fclamp(iz, 0, g.nz - 1);
    int xn = fclamp(ix + 1, 0, g.nx - 1), yn = fclamp(iy + 1, 0, g.ny - 1), zn = fclamp(iz + 1, 0, g.nz - 1);
    double px = coord1(x, 0.5, h[0]), py = coord1(y, 0.5, h[1]), pz = coord1(z, 0.5, h[2]);
    double tx = fclamp((p.x - px) / h[0], 0.0, 1.0);
    double ty = fclamp((p.y - py) / h[1], 0.0, 1.0);
    double tz = fclamp((p.z - pz) / h[2], 0.0, 1.0);
    return tx * ty * tz * g(x, y, z) + tx * ty * (1 - tz) * g(x, y, zn) +
           tx * (1 - ty) * tz * g(x, yn, z) + tx * (1 - ty) * (1 - tz) * g(x, yn, zn) +
           (1 - tx) * ty * tz * g(xn, y, z) + (1 - tx) * ty * (1 - tz) * g(xn, y, zn) +
           (1 - tx) * (1 - ty) * tz * g(xn, yn, z) + (1 - tx) * (1 - ty) * (1 - tz) * g(xn, yn, zn);
  }
  // SDF::gradient (sdf.h:62-85): central differences, one-sided at the border.
  V3 gradient(int x, int y, int z) const {
    int W = g.nx, H = g.ny, D = g.nz;
    int xn = x < W - 1 ? x + 1 : x, xp = x > 0 ? x - 1 : x;
    int yn = y < H - 1 ? y + 1 : y, yp = y > 0 ? y - 1 : y;
    int zn = z < D - 1 ? z + 1 : z, zp = z > 0 ? z - 1 : z;
    double dx = x > 0 && x < W - 1 ? 2 * h[0] : h[0];
    double dy = y > 0 && y < H - 1 ? 2 * h[1] : h[1];
    double dz = z > 0 && z < D - 1 ? 2 * h[2] : h[2];
    double gx = g(xn, y, z) - g(xp, y, z);
    double gy = g(x, yn, z) - g(x, yp, z);
    double gz = g(x, y, zn) - g(x, y, zp);
    return {gx / dx, gy / dy, gz / dz};
  }
  // SDF::grad (sdf.h:86-118): mirrored trilinear blend of node gradients, t not clamped.
  V3 grad(V3 p) const {
    int ix = (int)((p.x - 0.0) / h[0] - 0.5);
    int iy = (int)((p.y - 0.0) / h[1] - 0.5);
    int iz = (int)((p.z - 0.0) / h[2] - 0.5);
    int x = fclamp(ix, 0, g.nx - 1), y = fclamp(iy, 0, g.ny - 1), z = fclamp(iz, 0, g.nz - 1);
    int xn = fclamp(ix + 1, 0, g.nx - 1), yn = fclamp(iy + 1, 0, g.ny - 1), zn = fclamp(iz + 1, 0, g.nz - 1);
    double tx = (p.x - coord1(x, 0.5, h[0])) / h[0];
    double ty = (p.y - coord1(y, 0.5, h[1])) / h[1];
    double tz = (p.z - coord1(z, 0.5, h[2])) / h[2];
    V3 g000 = gradient(x, y, z), g001 = gradient(x, y, zn), g010 = gradient(x, yn, z),
       g011 = gradient(x, yn, zn), g100 = gradient(xn, y, z), g101 = gradient(xn, y, zn),
       g110 = gradient(xn, yn, z), g111 = gradient(xn, yn, zn);
    return tx * ty * tz * g000 + tx * ty * (1 - tz) * g001 + tx * (1 - ty) * tz * g010 +
           tx * (1 - ty) * (1 - tz) * g011 + (1 - tx) * ty * tz * g100 + (1 - tx) * ty * (1 - tz) * g101 +
           (1 - tx) * (1 - ty) * tz * g110 + (1 - tx) * (1 - ty) * (1 - tz) * g111;
  }
};

// cycle (FluidSim/src/cpu/util.cc:319-325)
static inline void cycle(double& a0, double& a1, double& a2, double& a3) {
  double tmp = a0;
  a0 = a3;
  a3 = a2;
  a2 = a1;
  a1 = tmp;
}
// fractionInside (util.cc:333-378)
double fractionInside(double lu, double ru, double rd, double ld) {
  int n_negs = (lu <= 0.0) + (ru <= 0.0) + (rd <= 0.0) + (ld <= 0.0);
  if (n_negs == 0) return 0.0;
  if (n_negs == 4) return 1.0;
  while (lu > 0.0) cycle(lu, ru, rd, ld);
  if (n_negs == 1) {
    double fracu = ru / (ru - lu);
    double fracl = ld / (ld - lu);
    return fracu * fracl * 0.5;
  }
  if (n_negs == 2) {
    if (ru <= 0.0) {
      double fracl = ld / (ld - lu);
      double fracr = rd / (rd - ru);
      return 0.5 * (fracl + fracr);
    }
    if (ld <= 0.0) {
      double fracu = ru / (ru - lu);
      double fracd = rd / (rd - ld);
      return 0.5 * (fracu + fracd);
    }
    if (rd <= 0.0) {
      double fracu = ru / (ru - lu);
      double fracd = ld / (ld - rd);
      double fracl = ld / (ld - lu);
      double fracr = ru / (ru - rd);
      if (lu + ru + rd + ld <= 0.0)
        return 1.0 - 0.5 * ((1.0 - fracu) * (1.0 - fracl) + (1.0 - fracd) * (1.0 - fracr));
      return 0.5 * (fracu * fracl + fracd * fracr);
    }
    return 0.0;  // unreachable: two negatives with lu<=0 leave one of ru/ld/rd <= 0
  }
  // n_negs == 3
  while (lu <= 0.0) cycle(lu, ru, rd, ld);
  double fracu = ru / (ru - lu);
  double fracl = ld / (ld - lu);
  return 1.0 - 0.5 * fracu * fracl;
}

// Morton code of ParticleNeighbourSearcher (Spatify/include/Spatify/ns-util.h:36-45).
static inline int morton10(int x) {
  x = (x | (x << 16)) & 0x030000FF;
  x = (x | (x << 8)) & 0x0300F00F;
  x = (x | (x << 4)) & 0x030C30C3;
  x = (x | (x << 2)) & 0x09249249;
  return x;
}
static inline int morton3D(int x, int y, int z) { return morton10(x) | (morton10(y) << 1) | (morton10(z) << 2); }

// ---------------------------------------------------------------------------------------------
// Device PCG layout, restated for the fixed-tree dot mode: live 8x8 column bricks in (i,j),
// each holding S = nz+14 (rounded up to a multiple of 8) slots of 64 lanes; cell (i,j,k) of brick (i/8, j/8) sits in lane
// (i%8)*8 + (j%8), slot k + i%8 + j%8. Element e = (brick*S + slot)*64 + lane.
struct PcgLayout {
  int nbi = 0, nbj = 0, S = 0, nlive = 0;
  std::vector<int> live;  // brick -> live index or -1
  std::vector<int64_t> cell_of;  // element -> dense id of an active cell, or -1
  void build(const BArr& active) {
    int nx = active.nx, ny = active.ny, nz = active.nz;
    nbi = (nx + 7) / 8;
    nbj = (ny + 7) / 8;
    S = ((nz + 14 + 7) / 8) * 8;  // rounded up to whole 8-slot chunks
    live.assign((size_t)nbi * nbj, -1);
    nlive = 0;
    for (int I = 0; I < nbi; I++)
      for (int J = 0; J < nbj; J++) {
        bool any = false;
        for (int a = 0; a < 8 && !any; a++)
          for (int b = 0; b < 8 && !any; b++) {
            int i = 8 * I + a, j = 8 * J + b;
            if (i >= nx || j >= ny) continue;
            for (int k = 0; k < nz; k++)
              if (active(i, j, k)) {
                any = true;
                break;
              }
          }
        if (any) live[(size_t)I * nbj + J] = nlive++;
      }
    cell_of.assign((size_t)nlive * S * 64, -1);
    for (int I = 0; I < nbi; I++)
      for (int J = 0; J < nbj; J++) {
        int L = live[(size_t)I * nbj + J];
        if (L < 0) continue;
        for (int a = 0; a < 8; a++)
          for (int b = 0; b < 8; b++) {
            int i = 8 * I + a, j = 8 * J + b;
            if (i >= nx || j >= ny) continue;
            for (int k = 0; k < nz; k++) {
              // only active cells contribute to a dot product (dotProduct masks by active,
              // util.cc:385); the stale values the reference leaves in inactive entries of z
              // and s must not enter the sum
              if (!active(i, j, k)) continue;
              int s = k + a + b;
              cell_of[((size_t)L * S + s) * 64 + a * 8 + b] = (int64_t)active.id(i, j, k);
            }
          }
      }
  }
};

// Wave butterfly: every lane adds its xor-partner, offsets 32..1.
static double butterfly64(double v[64]) {
  for (int off = 32; off >= 1; off >>= 1) {
    double nv[64];
    for (int l = 0; l < 64; l++) nv[l] = v[l] + v[l ^ off];
    std::memcpy(v, nv, sizeof(nv));
  }
  return v[0];
}
// Block of 256 threads: thread t sums vals[t + 256*m] for m ascending, then per-wave butterfly,
// then (w0 + w1) + (w2 + w3).
static double block_tree(const std::vector<double>& vals) {
  double acc[256];
  for (int t = 0; t < 256; t++) {
    double a = 0.0;
    for (size_t m = t; m < vals.size(); m += 256) a += vals[m];
    acc[t] = a;
  }
  double w[4];
  for (int q = 0; q < 4; q++) w[q] = butterfly64(acc + 64 * q);
  return (w[0] + w[1]) + (w[2] + w[3]);
}
// Elementwise-kernel dot: chunks of 4096 elements reduced by block_tree, then block_tree over
// the chunk partials. Inactive and padding elements contribute +0.
static double tree_dot_elementwise(const PcgLayout& L, const RArr& a, const RArr& b) {
  size_t ne = (size_t)L.nlive * L.S * 64;
  size_t nc = (ne + 4095) / 4096;
  std::vector<double> parts(nc);
  std::vector<double> chunk(4096);
  for (size_t c = 0; c < nc; c++) {
    for (int m = 0; m < 4096; m++) {
      size_t e = c * 4096 + m;
      double v = 0.0;
      if (e < ne && L.cell_of[e] >= 0) v = a.d[L.cell_of[e]] * b.d[L.cell_of[e]];
      chunk[m] = v;
    }
    parts[c] = block_tree(chunk);
  }
  return block_tree(parts);
}
// Backward-sweep dot: per brick, lane l accumulates slots S-1..0, wave butterfly; block_tree
// over brick partials.
static double tree_dot_sweep(const PcgLayout& L, const RArr& a, const RArr& b) {
  std::vector<double> parts(L.nlive);
  for (int B = 0; B < L.nlive; B++) {
    double lane[64];
    for (int l = 0; l < 64; l++) {
      double acc = 0.0;
      for (int s = L.S - 1; s >= 0; s--) {
        int64_t c = L.cell_of[((size_t)B * L.S + s) * 64 + l];
        double v = c >= 0 ? a.d[c] * b.d[c] : 0.0;
        acc += v;
      }
      lane[l] = acc;
    }
    parts[B] = butterfly64(lane);
  }
  return block_tree(parts);
}

// ---------------------------------------------------------------------------------------------
struct Sim {
  fs_config cfg{};
  int nx, ny, nz;
  double h[3];
  double width, height, depth;  // HybridAdvectionSolver3D extents (cpu/fluid-simulator.h:88-95)
  std::vector<V3> pos, vel;
  // cpu::FluidSimulator members (cpu/fluid-simulator.h:261-271)
  std::unique_ptr<RArr> ug, vg, wg, ubuf, vbuf, wbuf;
  std::unique_ptr<BArr> uValid, vValid, wValid, uValidBuf, vValidBuf, wValidBuf, sdfValid, sdfValidBuf;
  RArr uw, vw, ww, pg;
  std::unique_ptr<SDF> fluidSurface, fluidSurfaceBuf;
  SDF collider;
  // FlipAdvectionSolver3D (advect-solver.h:90-130)
  RArr u_last, v_last, w_last;
  // FvmSolver3D (project-solver.h:133-182)
  RArr Adiag, rhs, uWeights, vWeights, wWeights;
  RArr An[6];
  BArr active;
  // CgSolver3D + ModifiedIncompleteCholesky3D (project-solver.h:81-131)
  RArr r, z, s, E, q;
  int dot_mode = 0;
  int p2g_order = 0;
  int last_iters = 0;
  double last_residual = 0.0;
  int last_status = FS_OK;
  PcgLayout layout;

  enum Dir { Left, Right, Up, Down, Front, Back };  // project-solver.cc:7-14

  explicit Sim(const fs_config& c) : cfg(c) {
    nx = c.nx;
    ny = c.ny;
    nz = c.nz;
    // TGrid spacing: size / resolution (grids.h:244-252)
    h[0] = c.size_x / nx;
    h[1] = c.size_y / ny;
    h[2] = c.size_z / nz;
    width = nx * h[0];
    height = ny * h[1];
    depth = nz * h[2];
    ug.reset(new RArr(nx + 1, ny, nz));
    vg.reset(new RArr(nx, ny + 1, nz));
    wg.reset(new RArr(nx, ny, nz + 1));
    ubuf.reset(new RArr(nx + 1, ny, nz));
    vbuf.reset(new RArr(nx, ny + 1, nz));
    wbuf.reset(new RArr(nx, ny, nz + 1));
    uValid.reset(new BArr(nx + 1, ny, nz));
    vValid.reset(new BArr(nx, ny + 1, nz));
    wValid.reset(new BArr(nx, ny, nz + 1));
    uValidBuf.reset(new BArr(nx + 1, ny, nz));
    vValidBuf.reset(new BArr(nx, ny + 1, nz));
    wValidBuf.reset(new BArr(nx, ny, nz + 1));
    sdfValid.reset(new BArr(nx, ny, nz));
    sdfValidBuf.reset(new BArr(nx, ny, nz));
    uw = RArr(nx + 1, ny, nz);
    vw = RArr(nx, ny + 1, nz);
    ww = RArr(nx, ny, nz + 1);
    pg = RArr(nx, ny, nz);
    fluidSurface.reset(new SDF{RArr(nx, ny, nz), {h[0], h[1], h[2]}});
    fluidSurfaceBuf.reset(new SDF{RArr(nx, ny, nz), {h[0], h[1], h[2]}});
    collider = SDF{RArr(nx, ny, nz), {h[0], h[1], h[2]}};
    u_last = RArr(nx + 1, ny, nz);
    v_last = RArr(nx, ny + 1, nz);
    w_last = RArr(nx, ny, nz + 1);
    Adiag = RArr(nx, ny, nz);
    rhs = RArr(nx, ny, nz);
    uWeights = RArr(nx + 1, ny, nz);
    vWeights = RArr(nx, ny + 1, nz);
    wWeights = RArr(nx, ny, nz + 1);
    for (auto& a : An) a = RArr(nx, ny, nz);
    active = BArr(nx, ny, nz);
    r = RArr(nx, ny, nz);
    z = RArr(nx, ny, nz);
    s = RArr(nx, ny, nz);
    E = RArr(nx, ny, nz);
    q = RArr(nx, ny, nz);
  }

  // clear() (FluidSim/src/cpu/fluid-simulator.cc:20-37)
  void clear() {
    uValid->fill(0);
    vValid->fill(0);
    wValid->fill(0);
    uValidBuf->fill(0);
    vValidBuf->fill(0);
    wValidBuf->fill(0);
    vg->fill(0);
    ug->fill(0);
    wg->fill(0);
    ubuf->fill(0);
    vbuf->fill(0);
    wbuf->fill(0);
    uw.fill(0);
    vw.fill(0);
    ww.fill(0);
    pg.fill(0);
  }

  // sampleVelocity (FluidSim/src/cpu/util.cc:40-85): 4x4x4 nodes around TGrid::nearest,
  // i outer / k inner, weights with the scalar spacing h.x, normalised by the weight sum.
  V3 sampleVelocity(V3 p, const RArr& U, const RArr& Vv, const RArr& Wv) const {
    int ub[3] = {nearest1(p.x, face_off(0, 0), h[0]), nearest1(p.y, face_off(0, 1), h[1]),
                 nearest1(p.z, face_off(0, 2), h[2])};
    int vb[3] = {nearest1(p.x, face_off(1, 0), h[0]), nearest1(p.y, face_off(1, 1), h[1]),
                 nearest1(p.z, face_off(1, 2), h[2])};
    int wb[3] = {nearest1(p.x, face_off(2, 0), h[0]), nearest1(p.y, face_off(2, 1), h[1]),
                 nearest1(p.z, face_off(2, 2), h[2])};
    double hs = h[0];
    double u = 0.0, v = 0.0, w = 0.0;
    double w_u = 0.0, w_v = 0.0, w_w = 0.0;
    for (int i = -1; i <= 2; i++)
      for (int j = -1; j <= 2; j++)
        for (int k = -1; k <= 2; k++) {
          int x = ub[0] + i, y = ub[1] + j, zz = ub[2] + k;
          if (x >= 0 && x < U.nx && y >= 0 && y < U.ny && zz >= 0 && zz < U.nz) {
            double un = U(x, y, zz);
            V3 pn{coord1(x, 0.0, h[0]), coord1(y, 0.5, h[1]), coord1(zz, 0.5, h[2])};
            double uwt = kweight(hs, hs, hs, p - pn);
            u += un * uwt;
            w_u += uwt;
          }
          x = vb[0] + i, y = vb[1] + j, zz = vb[2] + k;
          if (x >= 0 && x < Vv.nx && y >= 0 && y < Vv.ny && zz >= 0 && zz < Vv.nz) {
            double vn = Vv(x, y, zz);
            V3 pn{coord1(x, 0.5, h[0]), coord1(y, 0.0, h[1]), coord1(zz, 0.5, h[2])};
            double vwt = kweight(hs, hs, hs, p - pn);
            v += vn * vwt;
            w_v += vwt;
          }
          x = wb[0] + i, y = wb[1] + j, zz = wb[2] + k;
          if (x >= 0 && x < Wv.nx && y >= 0 && y < Wv.ny && zz >= 0 && zz < Wv.nz) {
            double wn = Wv(x, y, zz);
            V3 pn{coord1(x, 0.5, h[0]), coord1(y, 0.5, h[1]), coord1(zz, 0.0, h[2])};
            double wwt = kweight(hs, hs, hs, p - pn);
            w += wn * wwt;
            w_w += wwt;
          }
        }
    return {w_u > 0.0 ? u / w_u : 0.0, w_v > 0.0 ? v / w_v : 0.0, w_w > 0.0 ? w / w_w : 0.0};
  }

  // HybridAdvectionSolver3D::handleCollision (FluidSim/src/cpu/advect-solver.cc:100-137)
  void handleCollision(V3& p, V3& v) const {
    double d = collider.eval(p);
    if (d < 0.0) {
      V3 normal = gnormalize(collider.grad(p));
      p = p - d * normal;
      double m = std::min(gdot(v, normal), 0.0);
      v = v - m * normal;
    }
    if (p.x < 0.0) {
      p.x = 0.0;
      if (v.x < 0.0) v.x = 0.0;
    }
    if (p.x > width) {
      p.x = width;
      if (v.x > 0.0) v.x = 0.0;
    }
    if (p.y < 0.0) {
      p.y = 0.0;
      if (v.y < 0.0) v.y = 0.0;
    }
    if (p.y > height) {
      p.y = height;
      if (v.y > 0.0) v.y = 0.0;
    }
    if (p.z < 0.0) {
      p.z = 0.0;
      if (v.z < 0.0) v.z = 0.0;
    }
    if (p.z > depth) {
      p.z = depth;
      if (v.z > 0.0) v.z = 0.0;
    }
  }

  // FlipAdvectionSolver3D::advect / PicAdvector3D::advect (advect-solver.cc:138-155, 170-187):
  // Ralston RK3 with k2, k3 sampled from the grids passed in.
  void advect(double dt) {
    const RArr &U = *ug, &Vv = *vg, &Wv = *wg;
    int n = (int)pos.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      V3 p = pos[i];
      V3 k1 = vel[i];
      V3 p1 = p + 0.5 * dt * k1;
      V3 k2 = sampleVelocity(p1, U, Vv, Wv);
      V3 p2 = p + 0.75 * dt * k2;
      V3 k3 = sampleVelocity(p2, U, Vv, Wv);
      p = p + 2.0 / 9.0 * dt * k1 + 1.0 / 3.0 * dt * k2 + 4.0 / 9.0 * dt * k3;
      V3 v = vel[i];
      handleCollision(p, v);
      pos[i] = p;
      vel[i] = v;
    }
  }

  // NaiveReconstructor<T,3>::reconstruct (FluidSim/include/FluidSim/cpu/sdf.h:210-225) with
  // ParticleNeighbourSearcher::update / forNeighbours (Spatify/include/Spatify/ns-util.h:46-109).
  void reconstruct(double radius) {
    SDF& sdf = *fluidSurface;
    BArr& valid = *sdfValid;
    int W = nx, H = ny, D = nz;
    const double* sp = h;  // resetGrid(..., sdf.grid.gridSpacing())
    int n = (int)pos.size();
    std::vector<int> cellmap(n), perm(n);
    for (int i = 0; i < n; i++) {
      perm[i] = i;
      int ix = (int)std::floor(pos[i].x / sp[0]);
      if (ix < 0) ix = 0;
      if (ix >= W) ix = W - 1;
      int iy = (int)std::floor(pos[i].y / sp[1]);
      if (iy < 0) iy = 0;
      if (iy >= H) iy = H - 1;
      int iz = (int)std::floor(pos[i].z / sp[2]);
      if (iz < 0) iz = 0;
      if (iz >= D) iz = D - 1;
      cellmap[i] = morton3D(ix, iy, iz);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return cellmap[a] < cellmap[b]; });
    std::vector<int> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = cellmap[perm[i]];
    // The reference sizes the bin tables W*H*D, which holds every Morton code only for
    // power-of-two cubes; they are sized to the largest code so other sizes stay in range.
    size_t nb = (size_t)morton3D(W - 1, H - 1, D - 1) + 1;
    std::vector<int> cbeg(nb, -1), cend(nb, -1);
    for (int i = 0; i < n; i++) {
      int g = sorted[i];
      if (i == 0 || sorted[i - 1] != g) cbeg[g] = i;
      if (i == n - 1 || sorted[i + 1] != g) cend[g] = i;  // inclusive end (Q4)
    }
    sdf.g.fill(1e9);
    valid.fill(0);
    double r2 = 2 * radius;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < W; i++)
      for (int j = 0; j < H; j++)
        for (int k = 0; k < D; k++) {
          V3 p{coord1(i, 0.5, h[0]), coord1(j, 0.5, h[1]), coord1(k, 0.5, h[2])};
          int x_min = std::max(0, (int)std::floor((p.x - r2) / sp[0]));
          int x_max = std::min(W - 1, (int)std::ceil((p.x + r2) / sp[0]));
          int y_min = std::max(0, (int)std::floor((p.y - r2) / sp[1]));
          int y_max = std::min(H - 1, (int)std::ceil((p.y + r2) / sp[1]));
          int z_min = std::max(0, (int)std::floor((p.z - r2) / sp[2]));
          int z_max = std::min(D - 1, (int)std::ceil((p.z + r2) / sp[2]));
          for (int a = x_min; a <= x_max; a++)
            for (int b = y_min; b <= y_max; b++)
              for (int c = z_min; c <= z_max; c++) {
                int cell = morton3D(a, b, c);
                for (int l = cbeg[cell]; l < cend[cell]; l++) {  // skips the bin's last particle (Q4)
                  V3 q = pos[perm[l]];
                  if (glen(p - q) <= r2) {
                    double dis = glen(q - p);
                    sdf.g(i, j, k) = std::min(sdf.g(i, j, k), dis - radius);
                    valid(i, j, k) = 1;
                  }
                }
              }
        }
  }

  // extrapolateFluidSdf (FluidSim/include/FluidSim/cpu/fluid-simulator.h:123-166): valid cells
  // are not copied into the buffer that is swapped in (Q3).
  void extrapolateFluidSdf(int iters) {
    for (int iter = 0; iter < iters; iter++) {
      sdfValidBuf->fill(0);
      RArr& cur = fluidSurface->g;
      RArr& buf = fluidSurfaceBuf->g;
      BArr& val = *sdfValid;
      BArr& vbuf2 = *sdfValidBuf;
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++) {
            if (val(i, j, k)) {
              vbuf2(i, j, k) = 1;
              continue;
            }
            double sum = 0.0;
            int count = 0;
            if (i > 0 && val(i - 1, j, k)) { sum += cur(i - 1, j, k); count++; }
            if (i < nx - 1 && val(i + 1, j, k)) { sum += cur(i + 1, j, k); count++; }
            if (j > 0 && val(i, j - 1, k)) { sum += cur(i, j - 1, k); count++; }
            if (j < ny - 1 && val(i, j + 1, k)) { sum += cur(i, j + 1, k); count++; }
            if (k > 0 && val(i, j, k - 1)) { sum += cur(i, j, k - 1); count++; }
            if (k < nz - 1 && val(i, j, k + 1)) { sum += cur(i, j, k + 1); count++; }
            if (count > 0) {
              buf(i, j, k) = sum / count;
              vbuf2(i, j, k) = 1;
            } else
              vbuf2(i, j, k) = 0;
          }
      std::swap(fluidSurface, fluidSurfaceBuf);
      std::swap(sdfValid, sdfValidBuf);
    }
  }

  // smoothFluidSurface (FluidSim/src/cpu/fluid-simulator.cc:39-74): buf = min(avg6, cur).
  void smoothFluidSurface(int iters) {
    for (int iter = 0; iter < iters; iter++) {
      RArr& cur = fluidSurface->g;
      RArr& buf = fluidSurfaceBuf->g;
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++) {
            double sum = 0;
            int count = 0;
            if (i > 0) { sum += cur(i - 1, j, k); count++; }
            if (i < nx - 1) { sum += cur(i + 1, j, k); count++; }
            if (j > 0) { sum += cur(i, j - 1, k); count++; }
            if (j < ny - 1) { sum += cur(i, j + 1, k); count++; }
            if (k > 0) { sum += cur(i, j, k - 1); count++; }
            if (k < nz - 1) { sum += cur(i, j, k + 1); count++; }
            buf(i, j, k) = sum / count;
            if (cur(i, j, k) < buf(i, j, k)) buf(i, j, k) = cur(i, j, k);
          }
      std::swap(fluidSurface, fluidSurfaceBuf);
    }
  }

  // One face grid of FlipAdvectionSolver3D::solveP2G (advect-solver.cc:188-271).
  void p2g_grid(int axis, RArr& G, RArr& Wt) {
    int n = (int)pos.size();
    std::vector<std::array<int, 3>> nb(n);
    for (int i = 0; i < n; i++) {
      nb[i] = {nearest1(pos[i].x, face_off(axis, 0), h[0]), nearest1(pos[i].y, face_off(axis, 1), h[1]),
               nearest1(pos[i].z, face_off(axis, 2), h[2])};
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (p2g_order == 0)
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nb[a] < nb[b]; });
    double o0 = face_off(axis, 0), o1 = face_off(axis, 1), o2 = face_off(axis, 2);
    // Every node sums its particles in `order`. The node planes x are split into slabs, one
    // per thread; a thread walks the whole order and adds only into its own planes, so each
    // node still receives its terms in exactly that order (the serial loop's rounding).
    const int nslab = std::max(1, std::min(G.nx, 4 * omp_get_max_threads()));
#pragma omp parallel for schedule(dynamic, 1)
    for (int sl = 0; sl < nslab; sl++) {
      const int xlo = (int)((int64_t)G.nx * sl / nslab), xhi = (int)((int64_t)G.nx * (sl + 1) / nslab);
      for (int idx : order) {
        const int b0 = nb[idx][0];
        if (b0 + 2 < xlo || b0 - 1 >= xhi) continue;
        V3 p = pos[idx];
        double vc = axis == 0 ? vel[idx].x : (axis == 1 ? vel[idx].y : vel[idx].z);
        for (int i = -1; i <= 2; i++) {
          int x = b0 + i;
          if (x < xlo || x >= xhi) continue;
          for (int j = -1; j <= 2; j++)
            for (int k = -1; k <= 2; k++) {
              int y = nb[idx][1] + j, zz = nb[idx][2] + k;
              if (x >= 0 && x < G.nx && y >= 0 && y < G.ny && zz >= 0 && zz < G.nz) {
                V3 pn{coord1(x, o0, h[0]), coord1(y, o1, h[1]), coord1(zz, o2, h[2])};
                double w = kweight(h[0], h[1], h[2], p - pn);
                G(x, y, zz) += vc * w;
                Wt(x, y, zz) += w;
              }
            }
        }
      }
    }
  }
  // p2g_order 2: the reference's parallel structure (TBB parallel_for over particles,
  // advect-solver.cc:201), with atomic adds in place of its racy ones. Used only to time the
  // CPU path; the sums then depend on thread timing at the rounding level.
  void p2g_grid_parallel(int axis, RArr& G, RArr& Wt) {
    int n = (int)pos.size();
    double o0 = face_off(axis, 0), o1 = face_off(axis, 1), o2 = face_off(axis, 2);
#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < n; idx++) {
      V3 p = pos[idx];
      double vc = axis == 0 ? vel[idx].x : (axis == 1 ? vel[idx].y : vel[idx].z);
      int b0 = nearest1(p.x, o0, h[0]), b1 = nearest1(p.y, o1, h[1]), b2 = nearest1(p.z, o2, h[2]);
      for (int i = -1; i <= 2; i++)
        for (int j = -1; j <= 2; j++)
          for (int k = -1; k <= 2; k++) {
            int x = b0 + i, y = b1 + j, zz = b2 + k;
            if (x >= 0 && x < G.nx && y >= 0 && y < G.ny && zz >= 0 && zz < G.nz) {
              V3 pn{coord1(x, o0, h[0]), coord1(y, o1, h[1]), coord1(zz, o2, h[2])};
              double w = kweight(h[0], h[1], h[2], p - pn);
              double vw = vc * w;
              double& g = G(x, y, zz);
              double& wt = Wt(x, y, zz);
#pragma omp atomic
              g += vw;
#pragma omp atomic
              wt += w;
            }
          }
    }
  }
  void p2g() {
    if (p2g_order == 2) {
      p2g_grid_parallel(0, *ug, uw);
      p2g_grid_parallel(1, *vg, vw);
      p2g_grid_parallel(2, *wg, ww);
    } else {
      p2g_grid(0, *ug, uw);
      p2g_grid(1, *vg, vw);
      p2g_grid(2, *wg, ww);
    }
    auto norm = [](RArr& G, RArr& Wt, BArr& V) {
      V.fill(0);
      for (size_t e = 0; e < G.size(); e++) {
        if (Wt.d[e] > 0.0) {
          G.d[e] /= Wt.d[e];
          V.d[e] = 1;
        } else {
          G.d[e] = 0.0;
          V.d[e] = 0;
        }
      }
    };
    norm(*ug, uw, *uValid);
    norm(*vg, vw, *vValid);
    norm(*wg, ww, *wValid);
    if (cfg.advection == FS_ADVECT_FLIP) {
      u_last.d = ug->d;
      v_last.d = vg->d;
      w_last.d = wg->d;
    }
  }

  // applyDirichletBoundary (fluid-simulator.cc:113-138)
  void applyDirichletBoundary() {
    RArr &U = *ug, &Vv = *vg, &Wv = *wg;
    for (int j = 0; j < U.ny; j++)
      for (int k = 0; k < U.nz; k++) {
        U(0, j, k) = 0.0;
        (*uValid)(0, j, k) = 1;
        U(U.nx - 1, j, k) = 0.0;
        (*uValid)(U.nx - 1, j, k) = 1;
      }
    for (int i = 0; i < Vv.nx; i++)
      for (int k = 0; k < Vv.nz; k++) {
        Vv(i, 0, k) = 0.0;
        (*vValid)(i, 0, k) = 1;
        Vv(i, Vv.ny - 1, k) = 0.0;
        (*vValid)(i, Vv.ny - 1, k) = 1;
      }
    for (int i = 0; i < Wv.nx; i++)
      for (int j = 0; j < Wv.ny; j++) {
        Wv(i, j, 0) = 0.0;
        (*wValid)(i, j, 0) = 1;
        Wv(i, j, Wv.nz - 1) = 0.0;
        (*wValid)(i, j, Wv.nz - 1) = 1;
      }
  }

  // extrapolate<GridType> (cpu/fluid-simulator.h:204-253): invalid cells get the mean of valid
  // neighbours (else 0); valid cells are not copied (Q3). Pointers swap every iteration.
  static void extrapolate(std::unique_ptr<RArr>& g, std::unique_ptr<RArr>& gbuf, std::unique_ptr<BArr>& valid,
                          std::unique_ptr<BArr>& validBuf, int iters) {
    for (int iter = 0; iter < iters; iter++) {
      validBuf->fill(0);
      RArr& G = *g;
      RArr& GB = *gbuf;
      BArr& V = *valid;
      BArr& VB = *validBuf;
      int W = G.nx, H = G.ny, D = G.nz;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < W; i++)
        for (int j = 0; j < H; j++)
          for (int k = 0; k < D; k++) {
            if (V(i, j, k)) {
              VB(i, j, k) = 1;
              continue;
            }
            double sum = 0.0;
            int count = 0;
            if (i > 0 && V(i - 1, j, k)) { sum += G(i - 1, j, k); count++; }
            if (i < W - 1 && V(i + 1, j, k)) { sum += G(i + 1, j, k); count++; }
            if (j > 0 && V(i, j - 1, k)) { sum += G(i, j - 1, k); count++; }
            if (j < H - 1 && V(i, j + 1, k)) { sum += G(i, j + 1, k); count++; }
            if (k > 0 && V(i, j, k - 1)) { sum += G(i, j, k - 1); count++; }
            if (k < D - 1 && V(i, j, k + 1)) { sum += G(i, j, k + 1); count++; }
            if (count > 0) {
              GB(i, j, k) = sum / count;
              VB(i, j, k) = 1;
            } else {
              GB(i, j, k) = 0.0;
              VB(i, j, k) = 0;
            }
          }
      std::swap(g, gbuf);
      std::swap(valid, validBuf);
    }
  }
  void extrapolateVelocity(int iters) {
    extrapolate(ug, ubuf, uValid, uValidBuf, iters);
    extrapolate(vg, vbuf, vValid, vValidBuf, iters);
    extrapolate(wg, wbuf, wValid, wValidBuf, iters);
  }

  // applyForce (fluid-simulator.cc:13-18)
  void applyForce(double dt) {
    RArr& V = *vg;
    for (int i = 0; i < V.nx; i++)
      for (int j = 0; j < V.ny; j++)
        for (int k = 0; k < V.nz; k++)
          if ((*vValid)(i, j, k) && j > 0 && j < V.ny - 1) V(i, j, k) -= cfg.gravity * dt;
  }

  // Vec3d(a, b, c) * 0.5 * h as written at project-solver.cc:85-122.
  static V3 corner(double a, double b, double c, double hh) { return {a * 0.5 * hh, b * 0.5 * hh, c * 0.5 * hh}; }

  // FvmSolver3D::buildSystem (FluidSim/src/cpu/project-solver.cc:59-228)
  void buildSystem(double dt) {
    const RArr& fs = fluidSurface->g;
    const RArr &U = *ug, &Vv = *vg, &Wv = *wg;
    double hh = h[0];
    Adiag.fill(0);
    for (auto& a : An) a.fill(0);
    uWeights.fill(0);
    vWeights.fill(0);
    wWeights.fill(0);
    rhs.fill(0);
    active.fill(0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx + 1; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          if (i == 0 || i == nx) {
            uWeights(i, j, k) = 0.0;
            continue;
          }
          if (fs(i - 1, j, k) > 0.0 && fs(i, j, k) > 0.0) continue;
          V3 p{coord1(i, 0.0, h[0]), coord1(j, 0.5, h[1]), coord1(k, 0.5, h[2])};
          double bu = collider.eval(p + corner(0.0, 1, -1, hh));
          double bd = collider.eval(p + corner(0.0, -1, -1, hh));
          double fd = collider.eval(p + corner(0.0, -1, 1, hh));
          double fu = collider.eval(p + corner(0.0, 1, 1, hh));
          uWeights(i, j, k) = 1.0 - fractionInside(bu, bd, fd, fu);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny + 1; j++)
        for (int k = 0; k < nz; k++) {
          if (j == 0 || j == ny) {
            vWeights(i, j, k) = 0.0;
            continue;
          }
          if (fs(i, j - 1, k) > 0.0 && fs(i, j, k) > 0.0) continue;
          V3 p{coord1(i, 0.5, h[0]), coord1(j, 0.0, h[1]), coord1(k, 0.5, h[2])};
          double lb = collider.eval(p + corner(-1, 0.0, -1, hh));
          double rb = collider.eval(p + corner(1, 0.0, -1, hh));
          double rf = collider.eval(p + corner(1, 0.0, 1, hh));
          double lf = collider.eval(p + corner(-1, 0.0, 1, hh));
          vWeights(i, j, k) = 1.0 - fractionInside(lb, rb, rf, lf);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz + 1; k++) {
          if (k == 0 || k == nz) {
            wWeights(i, j, k) = 0.0;
            continue;
          }
          if (fs(i, j, k - 1) > 0.0 && fs(i, j, k) > 0.0) continue;
          V3 p{coord1(i, 0.5, h[0]), coord1(j, 0.5, h[1]), coord1(k, 0.0, h[2])};
          double ld = collider.eval(p + corner(-1, -1, 0.0, hh));
          double lu = collider.eval(p + corner(-1, 1, 0.0, hh));
          double ru = collider.eval(p + corner(1, 1, 0.0, hh));
          double rd = collider.eval(p + corner(1, -1, 0.0, hh));
          wWeights(i, j, k) = 1.0 - fractionInside(ld, lu, ru, rd);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          if (uWeights(i, j, k) == 0.0 && uWeights(i + 1, j, k) == 0.0 && vWeights(i, j, k) == 0.0 &&
              vWeights(i, j + 1, k) == 0.0 && wWeights(i, j, k) == 0.0 && wWeights(i, j, k + 1) == 0.0)
            continue;
          if (fs(i, j, k) > 0.0) continue;
          active(i, j, k) = 1;
          double sd = fs(i, j, k);
          double factor = dt / hh;
          double& ad = Adiag(i, j, k);
          double& rh = rhs(i, j, k);
          if (i > 0) {  // left
            if (fs(i - 1, j, k) > 0.0) {
              double theta = std::min(fs(i - 1, j, k) / (fs(i - 1, j, k) - sd), 0.99);
              ad += uWeights(i, j, k) * factor / (1.0 - theta);
            } else {
              ad += uWeights(i, j, k) * factor;
              An[Left](i, j, k) -= uWeights(i, j, k) * factor;
            }
            rh += uWeights(i, j, k) * U(i, j, k);
          }
          if (i < nx - 1) {  // right
            if (fs(i + 1, j, k) > 0.0) {
              double theta = std::max(sd / (sd - fs(i + 1, j, k)), 0.01);
              ad += uWeights(i + 1, j, k) * factor / theta;
            } else {
              ad += uWeights(i + 1, j, k) * factor;
              An[Right](i, j, k) -= uWeights(i + 1, j, k) * factor;
            }
            rh -= uWeights(i + 1, j, k) * U(i + 1, j, k);
          }
          if (j > 0) {  // down
            if (fs(i, j - 1, k) > 0.0) {
              double theta = std::min(fs(i, j - 1, k) / (fs(i, j - 1, k) - sd), 0.99);
              ad += vWeights(i, j, k) * factor / (1.0 - theta);
            } else {
              ad += vWeights(i, j, k) * factor;
              An[Down](i, j, k) -= vWeights(i, j, k) * factor;
            }
            rh += vWeights(i, j, k) * Vv(i, j, k);
          }
          if (j < nx - 1) {  // up: bound uses fluidSdf.width() (Q8)
            if (fs(i, j + 1, k) > 0.0) {
              double theta = std::max(sd / (sd - fs(i, j + 1, k)), 0.01);
              ad += vWeights(i, j + 1, k) * factor / theta;
            } else {
              ad += vWeights(i, j + 1, k) * factor;
              An[Up](i, j, k) -= vWeights(i, j + 1, k) * factor;
            }
            rh -= vWeights(i, j + 1, k) * Vv(i, j + 1, k);
          }
          if (k > 0) {  // back
            if (fs(i, j, k - 1) > 0.0) {
              double theta = std::min(fs(i, j, k - 1) / (fs(i, j, k - 1) - sd), 0.99);
              ad += wWeights(i, j, k) * factor / (1.0 - theta);
            } else {
              ad += wWeights(i, j, k) * factor;
              An[Back](i, j, k) -= wWeights(i, j, k) * factor;
            }
            rh += wWeights(i, j, k) * Wv(i, j, k);
          }
          if (k < nz - 1) {  // front
            if (fs(i, j, k + 1) > 0.0) {
              double theta = std::max(sd / (sd - fs(i, j, k + 1)), 0.01);
              ad += wWeights(i, j, k + 1) * factor / theta;
            } else {
              ad += wWeights(i, j, k + 1) * factor;
              An[Front](i, j, k) -= wWeights(i, j, k + 1) * factor;
            }
            rh -= wWeights(i, j, k + 1) * Wv(i, j, k + 1);
          }
        }
  }

  // applyCompressedMatrix (project-solver.cc:16-39): sum order diag, L, D, B, R, U, F.
  void applyCompressedMatrix(const RArr& x, RArr& b) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          if (!active(i, j, k)) continue;
          double t = Adiag(i, j, k) * x(i, j, k);
          if (i > 0 && active(i - 1, j, k)) t += An[Left](i, j, k) * x(i - 1, j, k);
          if (j > 0 && active(i, j - 1, k)) t += An[Down](i, j, k) * x(i, j - 1, k);
          if (k > 0 && active(i, j, k - 1)) t += An[Back](i, j, k) * x(i, j, k - 1);
          if (i < nx - 1 && active(i + 1, j, k)) t += An[Right](i, j, k) * x(i + 1, j, k);
          if (j < ny - 1 && active(i, j + 1, k)) t += An[Up](i, j, k) * x(i, j + 1, k);
          if (k < nz - 1 && active(i, j, k + 1)) t += An[Front](i, j, k) * x(i, j, k + 1);
          b(i, j, k) = t;
        }
  }

  // ModifiedIncompleteCholesky3D::precond (project-solver.cc:230-306). The factor E is
  // recomputed on every call, as in the reference (Q7).
  void micFactor() {
    const double tau = cfg.mic_tau, sigma = cfg.mic_sigma;
    const RArr &AL = An[Left], &AD = An[Down], &AB = An[Back];
    E.fill(0);
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          if (!active(i, j, k)) continue;
          double e = Adiag(i, j, k);
          if (e == 0.0) continue;
          if (i > 0 && active(i - 1, j, k) && Adiag(i - 1, j, k) != 0.0) {
            double El = E(i - 1, j, k);
            double t = AL(i, j, k) * El;
            e -= t * t;
            if (j < ny - 1 && active(i - 1, j + 1, k)) e -= tau * AL(i, j, k) * AL(i - 1, j + 1, k) / (El * El);
            if (k < nz - 1 && active(i - 1, j, k + 1)) e -= tau * AL(i, j, k) * AL(i - 1, j, k + 1) / (El * El);
          }
          if (j > 0 && active(i, j - 1, k) && Adiag(i, j - 1, k) != 0.0) {
            double Ed = E(i, j - 1, k);
            double t = AD(i, j, k) * Ed;
            e -= t * t;
            if (i < nx - 1 && active(i + 1, j - 1, k)) e -= tau * AD(i, j, k) * AD(i + 1, j - 1, k) / (Ed * Ed);
            if (k < nz - 1 && active(i, j - 1, k + 1)) e -= tau * AD(i, j, k) * AD(i, j - 1, k + 1) / (Ed * Ed);
          }
          if (k > 0 && active(i, j, k - 1) && Adiag(i, j, k - 1) != 0.0) {
            double Eb = E(i, j, k - 1);
            double t = AB(i, j, k) * Eb;
            e -= t * t;
            if (i < nx - 1 && active(i + 1, j, k - 1)) e -= tau * AB(i, j, k) * AB(i + 1, j, k - 1) / (Eb * Eb);
            if (j < ny - 1 && active(i, j + 1, k - 1)) e -= tau * AB(i, j, k) * AB(i, j + 1, k - 1) / (Eb * Eb);
          }
          if (e < sigma * Adiag(i, j, k)) e = Adiag(i, j, k);
          E(i, j, k) = 1.0 / std::sqrt(e);
        }
  }
  void micSolve(const RArr& rr, RArr& zz) {
    const RArr &AL = An[Left], &AD = An[Down], &AB = An[Back];
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          if (!active(i, j, k)) continue;
          double t = rr(i, j, k);
          if (i > 0 && active(i - 1, j, k)) t -= AL(i, j, k) * E(i - 1, j, k) * q(i - 1, j, k);
          if (j > 0 && active(i, j - 1, k)) t -= AD(i, j, k) * E(i, j - 1, k) * q(i, j - 1, k);
          if (k > 0 && active(i, j, k - 1)) t -= AB(i, j, k) * E(i, j, k - 1) * q(i, j, k - 1);
          q(i, j, k) = t * E(i, j, k);
        }
    for (int i = nx - 1; i >= 0; i--)  // forEachReversed (arrays.h:205-211)
      for (int j = ny - 1; j >= 0; j--)
        for (int k = nz - 1; k >= 0; k--) {
          if (!active(i, j, k)) continue;
          double t = q(i, j, k);
          if (i < nx - 1 && active(i + 1, j, k)) t -= AL(i + 1, j, k) * E(i, j, k) * zz(i + 1, j, k);
          if (j < ny - 1 && active(i, j + 1, k)) t -= AD(i, j + 1, k) * E(i, j, k) * zz(i, j + 1, k);
          if (k < nz - 1 && active(i, j, k + 1)) t -= AB(i, j, k + 1) * E(i, j, k) * zz(i, j, k + 1);
          zz(i, j, k) = t * E(i, j, k);
        }
  }
  void precond(const RArr& rr, RArr& zz) {
    micFactor();
    micSolve(rr, zz);
  }

  // dotProduct (util.cc:380-392): serial lexicographic sum over active cells.
  double dotSerial(const RArr& a, const RArr& b) const {
    double sum = 0.0;
    for (size_t e = 0; e < a.size(); e++)
      if (active.d[e]) sum += a.d[e] * b.d[e];
    return sum;
  }
  // LinfNorm (util.h:56-63)
  double linf(const RArr& a) const {
    double m = 0;
    for (size_t e = 0; e < a.size(); e++)
      if (active.d[e]) m = std::max(m, std::fabs(a.d[e]));
    return m;
  }
  double norm2(const RArr& a) const {
    double sum = 0.0;
    for (size_t e = 0; e < a.size(); e++)
      if (active.d[e]) sum += a.d[e] * a.d[e];
    return std::sqrt(sum);
  }

  // CgSolver3D::solve (project-solver.cc:308-346). Returns the final ||r||_inf.
  double solvePressure() {
    const double tol = cfg.cg_tolerance;
    const int max_iterations = cfg.cg_max_iterations;
    bool tree = dot_mode == 1;
    if (tree) layout.build(active);
    pg.fill(0);
    r.d = rhs.d;
    double bnorm = norm2(rhs);
    auto converged = [&](double res) {
      if (cfg.tol_mode == FS_TOL_REL_L2) return norm2(r) <= tol * bnorm;
      return res < tol;
    };
    double residual = linf(r);
    last_status = FS_OK;
    if (converged(residual)) {
      last_iters = 0;
      last_residual = residual;
      return residual;
    }
    precond(rhs, z);
    s.d = z.d;
    double sigma = tree ? tree_dot_sweep(layout, z, r) : dotSerial(z, r);
    int iter = 1;
    bool conv = false;
    for (; iter < max_iterations; iter++) {
      applyCompressedMatrix(s, z);
      double sdotz = tree ? tree_dot_elementwise(layout, s, z) : dotSerial(s, z);
      if (sdotz == 0.0) {
        last_status = FS_BREAKDOWN;
        break;
      }
      double alpha = sigma / sdotz;
      for (size_t e = 0; e < pg.size(); e++)  // saxpy (util.cc:394-402)
        if (active.d[e]) pg.d[e] += alpha * s.d[e];
      for (size_t e = 0; e < r.size(); e++)
        if (active.d[e]) r.d[e] += -alpha * z.d[e];
      residual = linf(r);
      if (!(residual <= 1.7976931348623157e308)) {  // inf: stop (NaN entries are ignored by LinfNorm)
        last_status = FS_NONFINITE;
        break;
      }
      if (converged(residual)) {
        conv = true;
        break;
      }
      precond(r, z);
      double sigma_new = tree ? tree_dot_sweep(layout, z, r) : dotSerial(z, r);
      double beta = sigma_new / sigma;
      for (size_t e = 0; e < s.size(); e++)  // scaleAndAdd (util.cc:404-411)
        if (active.d[e]) s.d[e] = z.d[e] + beta * s.d[e];
      sigma = sigma_new;
    }
    if (last_status == FS_OK && !conv) last_status = FS_NOT_CONVERGED;
    last_iters = iter;
    last_residual = residual;
    return residual;
  }

  // FvmSolver3D::project (project-solver.cc:348-421)
  void project(double dt) {
    const RArr& fsd = fluidSurface->g;
    double hh = h[0];
    RArr &U = *ug, &Vv = *vg, &Wv = *wg;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < U.nx; i++)
      for (int j = 0; j < U.ny; j++)
        for (int k = 0; k < U.nz; k++) {
          if (uWeights(i, j, k) <= 0.0) continue;
          if (i == 0 || i == U.nx - 1) {
            U(i, j, k) = 0.0;
            continue;
          }
          double sl = fsd(i - 1, j, k), sr = fsd(i, j, k);
          if (sl >= 0.0 && sr >= 0.0) continue;
          if (sl < 0.0 && sr < 0.0) {
            U(i, j, k) -= (pg(i, j, k) - pg(i - 1, j, k)) * dt / hh;
            continue;
          }
          if (sl < 0.0) {
            double theta = std::max(sl / (sl - sr), 0.01);
            U(i, j, k) += pg(i - 1, j, k) * dt / hh / theta;
          } else {
            double theta = std::min(sl / (sl - sr), 0.99);
            U(i, j, k) -= pg(i, j, k) * dt / hh / (1.0 - theta);
          }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Vv.nx; i++)
      for (int j = 0; j < Vv.ny; j++)
        for (int k = 0; k < Vv.nz; k++) {
          if (vWeights(i, j, k) <= 0.0) continue;
          if (j == 0 || j == Vv.ny - 1) {
            Vv(i, j, k) = 0.0;
            continue;
          }
          double sd = fsd(i, j - 1, k), su = fsd(i, j, k);
          if (sd >= 0.0 && su >= 0.0) continue;
          if (sd < 0.0 && su < 0.0) {
            Vv(i, j, k) -= (pg(i, j, k) - pg(i, j - 1, k)) * dt / hh;
            continue;
          }
          if (sd < 0.0) {
            double theta = std::max(sd / (sd - su), 0.01);
            Vv(i, j, k) += pg(i, j - 1, k) * dt / hh / theta;
          } else {
            double theta = std::min(sd / (sd - su), 0.99);
            Vv(i, j, k) -= pg(i, j, k) * dt / hh / (1.0 - theta);
          }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Wv.nx; i++)
      for (int j = 0; j < Wv.ny; j++)
        for (int k = 0; k < Wv.nz; k++) {
          if (wWeights(i, j, k) <= 0.0) continue;
          if (k == 0 || k == Wv.nz - 1) {
            Wv(i, j, k) = 0.0;
            continue;
          }
          double sb = fsd(i, j, k - 1), sf = fsd(i, j, k);
          if (sb >= 0.0 && sf >= 0.0) continue;
          if (sb < 0.0 && sf < 0.0) {
            Wv(i, j, k) -= (pg(i, j, k) - pg(i, j, k - 1)) * dt / hh;
            continue;
          }
          if (sb < 0.0) {
            double theta = std::max(sb / (sb - sf), 0.01);
            Wv(i, j, k) += pg(i, j, k - 1) * dt / hh / theta;
          } else {
            double theta = std::min(sb / (sb - sf), 0.99);
            Wv(i, j, k) -= pg(i, j, k) * dt / hh / (1.0 - theta);
          }
        }
  }

  // applyCollider (fluid-simulator.cc:76-111)
  void applyCollider() {
    RArr &U = *ug, &Vv = *vg, &Wv = *wg;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < U.nx; i++)
      for (int j = 0; j < U.ny; j++)
        for (int k = 0; k < U.nz; k++) {
          if (i == 0 || i == U.nx - 1) {
            U(i, j, k) = 0.0;
            continue;
          }
          V3 p{coord1(i, 0.0, h[0]), coord1(j, 0.5, h[1]), coord1(k, 0.5, h[2])};
          if (collider.eval(p) < 0.0) {
            V3 n = gnormalize(collider.grad(p));
            U(i, j, k) -= U(i, j, k) * n.x;
          }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Vv.nx; i++)
      for (int j = 0; j < Vv.ny; j++)
        for (int k = 0; k < Vv.nz; k++) {
          if (j == 0 || j == Vv.ny - 1) {
            Vv(i, j, k) = 0.0;
            continue;
          }
          V3 p{coord1(i, 0.5, h[0]), coord1(j, 0.0, h[1]), coord1(k, 0.5, h[2])};
          if (collider.eval(p) < 0.0) {
            V3 n = gnormalize(collider.grad(p));
            Vv(i, j, k) -= Vv(i, j, k) * n.y;
          }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Wv.nx; i++)
      for (int j = 0; j < Wv.ny; j++)
        for (int k = 0; k < Wv.nz; k++) {
          if (k == 0 || k == Wv.nz - 1) {
            Wv(i, j, k) = 0.0;
            continue;
          }
          V3 p{coord1(i, 0.5, h[0]), coord1(j, 0.5, h[1]), coord1(k, 0.0, h[2])};
          if (collider.eval(p) < 0.0) {
            V3 n = gnormalize(collider.grad(p));
            Wv(i, j, k) -= Wv(i, j, k) * n.z;
          }
        }
  }

  // FlipAdvectionSolver3D::solveG2P (advect-solver.cc:157-169), PicAdvector3D::solveG2P (:7-17)
  void g2p() {
    int n = (int)pos.size();
    double alpha = cfg.flip_alpha;
    bool flip = cfg.advection == FS_ADVECT_FLIP;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      V3 pic = sampleVelocity(pos[i], *ug, *vg, *wg);
      if (flip) {
        V3 old = sampleVelocity(pos[i], u_last, v_last, w_last);
        V3 fl = vel[i] + pic - old;
        vel[i] = alpha * fl + (1.0 - alpha) * pic;
      } else {
        vel[i] = pic;
      }
    }
  }

  // CFL (fluid-simulator.cc:184-203)
  double CFL() const {
    double hh = h[0];
    double cfl = hh / 1e-6;
    for (const RArr* g : {ug.get(), vg.get(), wg.get()})
      for (double u : g->d)
        if (u != 0.0) cfl = std::min(cfl, hh / std::fabs(u));
    return cfg.cfl_factor * std::max(1e-3, cfl);
  }

  // substep (fluid-simulator.cc:140-182)
  int substep(double dt, fs_stats* st) {
    if (cfg.advect_on_cleared_grid) {
      clear();
      advect(dt);
    } else {
      advect(dt);
      clear();
    }
    reconstruct(cfg.recon_radius_factor * h[0]);
    extrapolateFluidSdf(cfg.sdf_extrapolate_iters);
    smoothFluidSurface(cfg.sdf_smooth_iters);
    p2g();
    applyDirichletBoundary();
    extrapolateVelocity(cfg.vel_extrapolate_iters);
    applyForce(dt);
    buildSystem(dt);
    solvePressure();
    project(dt);
    applyCollider();
    g2p();
    if (st) fill_stats(st);
    return last_status;
  }

  void fill_stats(fs_stats* st) {
    std::memset(st, 0, sizeof(*st));
    st->status = last_status;
    st->cg_iterations = last_iters;
    st->residual_inf = last_residual;
    int na = 0;
    for (auto a : active.d) na += a;
    st->n_active = na;
    double bn = norm2(rhs);
    st->rhs_norm2 = bn;
    st->residual_rel = bn > 0 ? norm2(r) / bn : 0.0;
  }

  // step(Frame&) (fluid-simulator.cc:205-220)
  int stepFrame(double frame_dt, fs_stats* st) {
    double t = 0;
    int cnt = 0;
    int status = FS_OK;
    while (t < frame_dt) {
      double cfl = CFL();
      double dt = std::min(cfl, frame_dt - t);
      cnt++;
      status = substep(dt, st);
      if (status != FS_OK && status != FS_NOT_CONVERGED) break;
      t += dt;
    }
    if (st) st->substeps = cnt;
    return status;
  }

  // Named state, matching fs_get_array / fs_set_array of the device library.
  bool lookup(const std::string& n, RArr** ra, BArr** ba) {
    *ra = nullptr;
    *ba = nullptr;
    if (n == "u") *ra = ug.get();
    else if (n == "v") *ra = vg.get();
    else if (n == "w") *ra = wg.get();
    else if (n == "ubuf") *ra = ubuf.get();
    else if (n == "vbuf") *ra = vbuf.get();
    else if (n == "wbuf") *ra = wbuf.get();
    else if (n == "u_valid") *ba = uValid.get();
    else if (n == "v_valid") *ba = vValid.get();
    else if (n == "w_valid") *ba = wValid.get();
    else if (n == "u_valid_buf") *ba = uValidBuf.get();
    else if (n == "v_valid_buf") *ba = vValidBuf.get();
    else if (n == "w_valid_buf") *ba = wValidBuf.get();
    else if (n == "uw") *ra = &uw;
    else if (n == "vw") *ra = &vw;
    else if (n == "ww") *ra = &ww;
    else if (n == "u_last") *ra = &u_last;
    else if (n == "v_last") *ra = &v_last;
    else if (n == "w_last") *ra = &w_last;
    else if (n == "fluid_sdf") *ra = &fluidSurface->g;
    else if (n == "fluid_sdf_buf") *ra = &fluidSurfaceBuf->g;
    else if (n == "sdf_valid") *ba = sdfValid.get();
    else if (n == "sdf_valid_buf") *ba = sdfValidBuf.get();
    else if (n == "collider_sdf") *ra = &collider.g;
    else if (n == "u_weights") *ra = &uWeights;
    else if (n == "v_weights") *ra = &vWeights;
    else if (n == "w_weights") *ra = &wWeights;
    else if (n == "adiag") *ra = &Adiag;
    else if (n == "a_left") *ra = &An[Left];
    else if (n == "a_right") *ra = &An[Right];
    else if (n == "a_up") *ra = &An[Up];
    else if (n == "a_down") *ra = &An[Down];
    else if (n == "a_front") *ra = &An[Front];
    else if (n == "a_back") *ra = &An[Back];
    else if (n == "rhs") *ra = &rhs;
    else if (n == "active") *ba = &active;
    else if (n == "pg") *ra = &pg;
    else if (n == "mic_e") *ra = &E;
    return *ra || *ba;
  }
};

}  // namespace fso

// ---------------------------------------------------------------------------------------------
// C API for the tests (ctypes). Same names and meaning as the device library, prefix fso_.
using fso::Sim;
extern "C" {

// The reference's parameters: alpha 0.97 (advect-solver.h:129), tau 0.97 and sigma 0.25
// (project-solver.h:83-84), |r|_inf < 1e-6 within 300 iterations (project-solver.h:129-130,
// project-solver.cc:327), gravity 9.8 (fluid-simulator.cc:13-18), 10 SDF extrapolation and 5
// smoothing passes, 10 velocity extrapolation passes and radius 0.8h (fluid-simulator.cc:
// 146-160), CFL factor 5 (fluid-simulator.cc:202).
void fso_config_default(fs_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->nx = cfg->ny = cfg->nz = 32;
  cfg->size_x = cfg->size_y = cfg->size_z = 1.0;
  cfg->precision = FS_PRECISION_FP64;
  cfg->advection = FS_ADVECT_FLIP;
  cfg->flip_alpha = 0.97;
  cfg->mic_tau = 0.97;
  cfg->mic_sigma = 0.25;
  cfg->cg_tolerance = 1e-6;
  cfg->cg_max_iterations = 300;
  cfg->tol_mode = FS_TOL_ABS_LINF;
  cfg->gravity = 9.8;
  cfg->sdf_extrapolate_iters = 10;
  cfg->sdf_smooth_iters = 5;
  cfg->vel_extrapolate_iters = 10;
  cfg->recon_radius_factor = 0.8;
  cfg->advect_on_cleared_grid = 1;
  cfg->cfl_factor = 5.0;
  cfg->inner_tolerance = 1e-3;
  cfg->max_refinements = 8;
  cfg->check_interval = 8;
}

void* fso_create(const fs_config* cfg) {
  if (!cfg || cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0) return nullptr;
  return new Sim(*cfg);
}
void fso_destroy(void* s) { delete static_cast<Sim*>(s); }
int fso_set_dot_mode(void* s, int mode) {
  static_cast<Sim*>(s)->dot_mode = mode;
  return 0;
}
int fso_set_p2g_order(void* s, int order) {
  static_cast<Sim*>(s)->p2g_order = order;
  return 0;
}
int fso_set_particles(void* sp, const double* p, const double* v, int64_t n) {
  Sim* s = static_cast<Sim*>(sp);
  s->pos.resize(n);
  s->vel.resize(n);
  for (int64_t i = 0; i < n; i++) {
    s->pos[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    s->vel[i] = v ? fso::V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]} : fso::V3{0, 0, 0};
  }
  return 0;
}
int fso_get_particles(void* sp, double* p, double* v) {
  Sim* s = static_cast<Sim*>(sp);
  for (size_t i = 0; i < s->pos.size(); i++) {
    if (p) {
      p[3 * i] = s->pos[i].x;
      p[3 * i + 1] = s->pos[i].y;
      p[3 * i + 2] = s->pos[i].z;
    }
    if (v) {
      v[3 * i] = s->vel[i].x;
      v[3 * i + 1] = s->vel[i].y;
      v[3 * i + 2] = s->vel[i].z;
    }
  }
  return 0;
}
int64_t fso_num_particles(void* s) { return (int64_t) static_cast<Sim*>(s)->pos.size(); }
int fso_set_collider_sdf(void* s, const double* sdf) {
  Sim* S = static_cast<Sim*>(s);
  std::memcpy(S->collider.g.d.data(), sdf, S->collider.g.d.size() * sizeof(double));
  return 0;
}
int fso_clear(void* s) {
  static_cast<Sim*>(s)->clear();
  return 0;
}
int fso_advect(void* s, double dt) {
  static_cast<Sim*>(s)->advect(dt);
  return 0;
}
int fso_reconstruct(void* s, double radius) {
  static_cast<Sim*>(s)->reconstruct(radius);
  return 0;
}
int fso_extrapolate_fluid_sdf(void* s, int iters) {
  static_cast<Sim*>(s)->extrapolateFluidSdf(iters);
  return 0;
}
int fso_smooth_fluid_surface(void* s, int iters) {
  static_cast<Sim*>(s)->smoothFluidSurface(iters);
  return 0;
}
int fso_p2g(void* s, double) {
  static_cast<Sim*>(s)->p2g();
  return 0;
}
int fso_apply_dirichlet_boundary(void* s) {
  static_cast<Sim*>(s)->applyDirichletBoundary();
  return 0;
}
int fso_extrapolate_velocity(void* s, int iters) {
  static_cast<Sim*>(s)->extrapolateVelocity(iters);
  return 0;
}
int fso_apply_force(void* s, double dt) {
  static_cast<Sim*>(s)->applyForce(dt);
  return 0;
}
int fso_build_system(void* s, double dt) {
  static_cast<Sim*>(s)->buildSystem(dt);
  return 0;
}
int fso_solve_pressure(void* sp, fs_stats* st) {
  Sim* s = static_cast<Sim*>(sp);
  s->solvePressure();
  if (st) s->fill_stats(st);
  return s->last_status;
}
int fso_project(void* s, double dt) {
  static_cast<Sim*>(s)->project(dt);
  return 0;
}
int fso_apply_collider(void* s) {
  static_cast<Sim*>(s)->applyCollider();
  return 0;
}
int fso_g2p(void* s, double) {
  static_cast<Sim*>(s)->g2p();
  return 0;
}
int fso_substep(void* s, double dt, fs_stats* st) { return static_cast<Sim*>(s)->substep(dt, st); }
int fso_step_frame(void* s, double frame_dt, fs_stats* st) { return static_cast<Sim*>(s)->stepFrame(frame_dt, st); }
int fso_cfl(void* s, double* out) {
  *out = static_cast<Sim*>(s)->CFL();
  return 0;
}
int fso_spmv(void* sp, const double* x, double* y) {
  Sim* s = static_cast<Sim*>(sp);
  fso::RArr X(s->nx, s->ny, s->nz), Y(s->nx, s->ny, s->nz);
  std::memcpy(X.d.data(), x, X.size() * sizeof(double));
  std::memcpy(Y.d.data(), y, Y.size() * sizeof(double));
  s->applyCompressedMatrix(X, Y);
  std::memcpy(y, Y.d.data(), Y.size() * sizeof(double));
  return 0;
}
int fso_mic_apply(void* sp, const double* r, double* z) {
  Sim* s = static_cast<Sim*>(sp);
  fso::RArr R(s->nx, s->ny, s->nz), Z(s->nx, s->ny, s->nz);
  std::memcpy(R.d.data(), r, R.size() * sizeof(double));
  std::memcpy(Z.d.data(), z, Z.size() * sizeof(double));
  s->precond(R, Z);
  std::memcpy(z, Z.d.data(), Z.size() * sizeof(double));
  return 0;
}
int fso_array_info(void* sp, const char* name, int64_t dims[3], int32_t* dtype) {
  Sim* s = static_cast<Sim*>(sp);
  fso::RArr* ra;
  fso::BArr* ba;
  if (!s->lookup(name, &ra, &ba)) return FS_BAD_ARG;
  if (ra) {
    dims[0] = ra->nx;
    dims[1] = ra->ny;
    dims[2] = ra->nz;
    *dtype = 0;
  } else {
    dims[0] = ba->nx;
    dims[1] = ba->ny;
    dims[2] = ba->nz;
    *dtype = 1;
  }
  return 0;
}
int fso_get_array(void* sp, const char* name, void* out) {
  Sim* s = static_cast<Sim*>(sp);
  fso::RArr* ra;
  fso::BArr* ba;
  if (!s->lookup(name, &ra, &ba)) return FS_BAD_ARG;
  if (ra) std::memcpy(out, ra->d.data(), ra->size() * sizeof(double));
  else std::memcpy(out, ba->d.data(), ba->size());
  return 0;
}
int fso_set_array(void* sp, const char* name, const void* in) {
  Sim* s = static_cast<Sim*>(sp);
  fso::RArr* ra;
  fso::BArr* ba;
  if (!s->lookup(name, &ra, &ba)) return FS_BAD_ARG;
  if (ra) std::memcpy(ra->d.data(), in, ra->size() * sizeof(double));
  else std::memcpy(ba->d.data(), in, ba->size());
  return 0;
}
// Free functions exposed for known-answer tests.
double fso_fraction_inside(double lu, double ru, double rd, double ld) { return fso::fractionInside(lu, ru, rd, ld); }
int fso_morton3d(int x, int y, int z) { return fso::morton3D(x, y, z); }
double fso_bspline(double r) { return fso::bspline(r); }
double fso_sdf_eval(void* sp, const double* p) {
  Sim* s = static_cast<Sim*>(sp);
  return s->collider.eval({p[0], p[1], p[2]});
}
void fso_sdf_grad(void* sp, const double* p, double* g) {
  Sim* s = static_cast<Sim*>(sp);
  fso::V3 v = s->collider.grad({p[0], p[1], p[2]});
  g[0] = v.x;
  g[1] = v.y;
  g[2] = v.z;
}
void fso_sample_velocity(void* sp, const double* p, double* out) {
  Sim* s = static_cast<Sim*>(sp);
  fso::V3 v = s->sampleVelocity({p[0], p[1], p[2]}, *s->ug, *s->vg, *s->wg);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}
void fso_nearest(void* sp, int axis, const double* p, int* out) {
  Sim* s = static_cast<Sim*>(sp);
  for (int c = 0; c < 3; c++) out[c] = fso::nearest1(p[c], fso::face_off(axis, c), s->h[c]);
}
int fso_last_iterations(void* s) { return static_cast<Sim*>(s)->last_iters; }
}
