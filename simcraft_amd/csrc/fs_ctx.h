// fs_ctx.h - the context behind the C ABI: device state of one FLIP simulation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../include/simcraft_fluid.h"
#include "fs_device.h"

// Status-returning HIP check used by host code of the library.
#define FS_HIP(ctx, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      (ctx)->set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return FS_DEVICE_ERROR;                                                      \
    }                                                                              \
  } while (0)

#define FS_TRY(expr)          \
  do {                        \
    int s_ = (expr);          \
    if (s_ != FS_OK) return s_; \
  } while (0)

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Device scalars of one pressure solve. Written by the last block of a reduction kernel,
// read by the next kernels and copied to the host at convergence checks.
struct PcgScalars {
  double sigma;      // z.r of the current iteration
  double alpha;      // sigma / (s.As)
  double beta;
  double sdotz;
  double residual;   // ||r||_inf
  double rnorm2sq;   // ||r||_2^2 (fixed tree)
  double bnorm2sq;   // ||rhs||_2^2
  double tol_abs;    // absolute threshold applied to ||r||_inf or ||r||_2^2 depending on mode
  int32_t converged;
  int32_t conv_iter;
  int32_t status;
  int32_t wait_timeouts;
  unsigned long long n_active;  // active cells (multigrid solve; integer sum, order-free)
};

struct PcgState {
  int nbi = 0, nbj = 0, S = 0, nlive = 0;
  int64_t ne = 0;                // elements nlive*S*64
  int nchunks = 0;               // ceil(ne/4096)
  std::vector<int> live_host;    // brick -> live index or -1
  DBuf live_map;                 // int[nbi*nbj]
  DBuf brick_ij;                 // int2[nlive]
  DBuf fwd_order, bwd_order;     // int[nlive]: dependency order (anti-diagonal) and its reverse
  DBuf brick_flag;               // int[nbi*nbj]
  // packed system (element layout)
  DBuf adiag, al, ad, ab, nbm, rhs;
  DBuf e, cl, cd, cb, dr, du, df;  // MIC factor and sweep coefficients
  DBuf tflags, tnum;             // factor sweep: flag word and 6 tau numerators per element
  DBuf x, r, s, t, q, z;         // PCG vectors
  DBuf partials;                 // double[max(nchunks, nlive)]
  DBuf partials2;                // second partial array (max reductions)
  DBuf scal;                     // PcgScalars
  DBuf xf, xb;                   // boundary exchange (forward/factor, backward): [2][nlive][S][8] + zero row
  DBuf tickets;                  // uint64[3]: brick tickets of the factor, forward and backward sweeps
  uint64_t ticket_base[3] = {0, 0, 0};
  PcgScalars* host_scal = nullptr;  // pinned
  // snapshots of the scalars after a batch of iterations (pinned, two slots) and their events:
  // the solve queues the next batch before it waits for the previous batch's snapshot
  PcgScalars* snap = nullptr;
  hipEvent_t sev[2];
  bool sev_ok = false;
  bool built = false;
  // mixed-precision copies
  DBuf f_adiag, f_al, f_ad, f_ab, f_e, f_cl, f_cd, f_cb, f_dr, f_du, f_df;
  DBuf f_r, f_s, f_t, f_q, f_z, f_x;
};

// Multigrid-preconditioned CG (fs_mg.hip). Levels are dense grids, k fastest, storage type of the
// context; level 0 is the assembled system, level l+1 the 2x2x2 aggregation of level l.
struct MgLevelBuf {
  int nx = 0, ny = 0, nz = 0;
  int64_t n = 0;
  DBuf D, Cx, Cy, Cz, W;  // diagonal, couplings to the -x/-y/-z neighbour, Jacobi weight omega/D
  DBuf b, x, xs;          // right-hand side (levels >= 1), cycle solution and its ping-pong copy
};
struct MgState {
  std::vector<MgLevelBuf> lv;
  DBuf x, r, s0, s1, t;  // CG vectors (FP64); z is level 0's x
  DBuf partials, partials2, scal;
  PcgScalars* host_scal = nullptr;  // pinned
  int nchunks = 0;
};

// Reverse mode. One recorded substep (device copies, context precision) and the rollout tape:
// checkpoints of the full state every `every` substeps, from which a segment is recomputed
// with its steps recorded, then reversed.
struct AdjStep {
  double dt = 0;
  DBuf x0[3], v0[3], x1[3], v1[3];  // particles before and after advection
  DBuf g0[3], w0[3];                // P2G output (FLIP's u_last) and its weight sums
  DBuf valid[3];                    // valid flags after the Dirichlet walls
  DBuf g5[3];                       // grids after projection and collider (G2P input)
  DBuf weights[3];                  // face weights of the assembly
  DBuf fsdf, active, adiag, an[6];  // the pressure system's geometry
  DBuf sdfbuf0;                     // SDF ping-pong buffer at the start (carried over)
  DBuf pg;                          // pressure
};
struct AdjCheckpoint {
  DBuf x[3], v[3], fsdf, fsdfbuf, sdfvalid, sdfvalidbuf;
  DBuf g[3], pg;  // face grids and pressure: only in the end-of-rollout checkpoint
};
struct AdjTape {
  bool on = false;         // a tape exists (fs_rollout_begin .. fs_rollout_end)
  bool recording = false;  // substeps are appended to it (until fs_rollout_stop)
  int every = 8;
  std::vector<double> dts;
  std::vector<AdjCheckpoint> ck;
  std::vector<AdjStep> seg;
  AdjCheckpoint end;
  // reverse-pass scratch of the fluid-surface chain (cell arrays; masks and branch flags)
  DBuf sA, sB, sAb, sBb, sAb2, sBb2;
  std::vector<DBuf> sV, sF;
};
enum AdjPoint { ADJ_INPUT, ADJ_ADVECTED, ADJ_P2G, ADJ_DIRICHLET, ADJ_BUILT, ADJ_FINAL };

struct fs_ctx {
  fs_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nx = 0, ny = 0, nz = 0;
  double h[3] = {0, 0, 0};
  bool mixed = false;
  size_t rsz = 8;  // bytes per real of grids/particles (8 FP64, 4 mixed)
  int n_face[3] = {0, 0, 0};
  int fdim[3][3];  // face grid dims
  int n_cell = 0;

  // particles (SoA)
  int64_t np = 0;
  DBuf pos[3], vel[3];
  DBuf staging;  // AoS upload/download
  // face grids: current, ping-pong buffer, valid flags, P2G weight sums, FLIP copy, solid weights
  DBuf g[3], gbuf[3], gvalid[3], gvalidbuf[3], gw[3], glast[3], gweights[3];
  // cell-centred
  DBuf fsdf, fsdfbuf, sdfvalid, sdfvalidbuf, collider, pg;
  DBuf adiag, an[6], rhs, active;
  // binning
  DBuf keys, slots, perm, counts, offs, scan_tmp;
  DBuf sorted[4];   // bin-ordered velocity component (P2G); spare particle-sized scratch
  DBuf p2g_w;       // 12 kernel factors per particle in bin order (P2G)
  DBuf recon_tmp;   // per-cell squared-distance minima, one array per XCD and a spare (reconstruction)
  DBuf p2g_tiles;   // P2G gather tile keys and heaviest-first tile order
  // second P2G buffer set (bin offsets, records, tile order; the velocity records go to
  // sorted[1]): a face grid's gather runs on `side` while the next face grid is binned
  DBuf offs2, p2g_w2, p2g_tiles2;
  hipStream_t side = nullptr;
  hipEvent_t p2g_ev[6];  // gather done (face grid a), records ready (3 + a)
  bool p2g_ev_ok = false;
  // reconstruction's own keys and per-bin last particle, so that the fluid-SDF stages (on side2)
  // can run beside P2G and the velocity stages (on the main stream) within a substep
  DBuf rkeys, rbinmax;
  // reconstruction: per axis and bin b, the cells i whose bin range (ns-util.h forNeighbours)
  // holds b, as one interval [imin, imax] (int2[3][max(nx, ny, nz)])
  DBuf rbins;
  hipStream_t side2 = nullptr;
  hipEvent_t sdf_ev[2];  // fork (main), join (side2)
  bool sdf_ev_ok = false;
  size_t scan_tmp_bytes = 0;
  size_t bins_cap = 0;
  DBuf red;  // small reduction scratch
  // APIC affine rows, 9 per particle, component-major: [3a + d][particle] (allocated and zeroed
  // by fs_ensure_apic when APIC is in use; freed when the particles change, so they restart at 0)
  DBuf apicc;
  // the last surface built by fs_rebuild_surface (vertices and normals AoS xyz FP64, triangles)
  DBuf surf_v, surf_n, surf_i;
  int64_t surf_nv = 0, surf_nt = 0;
  PcgState pcg;
  MgState mg;
  hipEvent_t ev[12];
  bool ev_ok = false;
  // kernel timing (profile_stages >= 2): event pairs recorded around launches, resolved at the
  // next synchronisation point
  static constexpr int NPROF_EV = 256;
  hipEvent_t pev[NPROF_EV];
  int pev_kind[NPROF_EV / 2];
  double pev_bytes[NPROF_EV / 2];
  int pev_used = 0;
  bool pev_ok = false;
  bool ktime_on = true;  // fs_set_kernel_timing
  int64_t kprof_launches[FS_KERNEL_COUNT] = {};
  double kprof_ms[FS_KERNEL_COUNT] = {};
  double kprof_bytes[FS_KERNEL_COUNT] = {};
  int64_t n_active_last = 0;
  AdjTape tape;

  void set_error(const std::string& s) { err = s; }
};

// Allocation helpers
int fs_alloc(fs_ctx* c, DBuf& b, size_t bytes);
void fs_free(DBuf& b);
int fs_zero(fs_ctx* c, DBuf& b);
// APIC affine rows of the current particles (9 * np values at storage width), zeroed on first use
int fs_ensure_apic(fs_ctx* c);

// Stages implemented in the .hip files.
int fs_stage_clear(fs_ctx* c);
int fs_stage_advect(fs_ctx* c, double dt);
int fs_stage_reconstruct(fs_ctx* c, double radius);
int fs_stage_extrapolate_sdf(fs_ctx* c, int iters);
int fs_stage_smooth(fs_ctx* c, int iters);
int fs_stage_p2g(fs_ctx* c);
int fs_stage_dirichlet(fs_ctx* c);
int fs_stage_extrapolate_velocity(fs_ctx* c, int iters);
int fs_stage_force(fs_ctx* c, double dt);
int fs_stage_build(fs_ctx* c, double dt);
int fs_stage_solve(fs_ctx* c, fs_stats* st);
int fs_stage_project(fs_ctx* c, double dt);
int fs_stage_collider(fs_ctx* c);
int fs_stage_g2p(fs_ctx* c);
int fs_stage_cfl(fs_ctx* c, double* out);

// Substep with an optional recording of the state the reverse pass needs (fs_capi.hip).
int fs_substep_core(fs_ctx* c, double dt, fs_stats* S, AdjStep* rec);
// Reverse mode (fs_adjoint.hip; particle kernels in fs_grid.hip).
int fs_adj_record(fs_ctx* c, AdjStep* rec, AdjPoint at);
int fs_adj_on_substep(fs_ctx* c, double dt);  // checkpoints of a recorded rollout
void fs_adj_release(fs_ctx* c);
int fs_adj_p2g_sum(fs_ctx* c, DBuf* pos, DBuf* coef, DBuf* out);
int fs_adj_g2p(fs_ctx* c, DBuf* x1, DBuf* g5, DBuf* g0, DBuf* vbar, DBuf* xbar, DBuf* vbar1, DBuf* coef);
int fs_adj_p2g(fs_ctx* c, DBuf* x1, DBuf* v1, DBuf* H, DBuf* HG, DBuf* vbar, DBuf* xbar);
int fs_adj_advect(fs_ctx* c, DBuf* x0, DBuf* v0, DBuf* xbar, DBuf* vbar, double dt);
// Reverse of reconstruct -> extrapolateFluidSdf -> smoothFluidSurface for one recorded step:
// phibar (adjoint of the smoothed SDF) and psibar (in: adjoint of the SDF buffer the step leaves
// behind; out: adjoint of the buffer it started from) -> position adjoint added to xbar.
int fs_adj_sdf(fs_ctx* c, AdjStep& t, DBuf& phibar, DBuf& psibar, DBuf* xbar);
int fs_pcg_spmv_dense(fs_ctx* c, const double* x, double* y);
int fs_pcg_mic_dense(fs_ctx* c, const double* r, double* z);
int fs_pcg_export_e(fs_ctx* c, double* out);
void fs_pcg_release(fs_ctx* c);
// Multigrid-preconditioned CG of the assembled system (fs_mg.hip).
int fs_mg_solve(fs_ctx* c, fs_stats* st);
int fs_mg_apply_dense(fs_ctx* c, const double* r, double* z);
void fs_mg_release(fs_ctx* c);
int fs_particles_upload(fs_ctx* c, const double* pos, const double* vel);
int fs_particles_download(fs_ctx* c, double* pos, double* vel);

// Conversions between the FP64 ABI and device storage of a given real size.
int fs_upload_real(fs_ctx* c, DBuf& dst, const double* src, size_t n);
int fs_download_real(fs_ctx* c, const DBuf& src, double* dst, size_t n);

inline unsigned fs_blocks(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// Kernel timing helpers (active when cfg.profile_stages >= 2).
int fs_prof_begin(fs_ctx* c, int kind, double bytes);  // returns a slot or -1
void fs_prof_end(fs_ctx* c, int slot);
void fs_prof_flush(fs_ctx* c);  // resolves recorded pairs; call after a stream synchronisation
