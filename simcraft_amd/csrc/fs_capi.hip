// fs_capi.hip - extern "C" entry points of libsimcraft_fluid.so (declared in
// include/simcraft_fluid.h). Host orchestration only; the kernels live in fs_grid.hip and
// fs_pcg.hip.
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "fs_ctx.h"

int fs_alloc(fs_ctx* c, DBuf& b, size_t bytes) {
  b.bytes = bytes;
  if (bytes == 0) {
    b.p = nullptr;
    return FS_OK;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    b.bytes = 0;
    c->set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return FS_DEVICE_ERROR;
  }
  return FS_OK;
}

void fs_free(DBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int fs_zero(fs_ctx* c, DBuf& b) {
  if (!b.p || !b.bytes) return FS_OK;
  FS_HIP(c, hipMemsetAsync(b.p, 0, b.bytes, c->stream));
  return FS_OK;
}

int fs_ensure_apic(fs_ctx* c) {
  const size_t bytes = (size_t)std::max<int64_t>(c->np, 1) * 9 * c->rsz;
  if (c->apicc.p && c->apicc.bytes == bytes) return FS_OK;
  fs_free(c->apicc);
  FS_TRY(fs_alloc(c, c->apicc, bytes));
  return fs_zero(c, c->apicc);
}

int fs_prof_begin(fs_ctx* c, int kind, double bytes) {
  if (c->cfg.profile_stages < 2 || !c->pev_ok || !c->ktime_on) return -1;
  if (c->pev_used + 2 > fs_ctx::NPROF_EV) {
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    fs_prof_flush(c);
  }
  int slot = c->pev_used / 2;
  c->pev_used += 2;
  c->pev_kind[slot] = kind;
  c->pev_bytes[slot] = bytes;
  (void)hipEventRecord(c->pev[2 * slot], c->stream);
  return slot;
}

void fs_prof_end(fs_ctx* c, int slot) {
  if (slot < 0) return;
  (void)hipEventRecord(c->pev[2 * slot + 1], c->stream);
}

void fs_prof_flush(fs_ctx* c) {
  for (int q = 0; q < c->pev_used / 2; q++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->pev[2 * q], c->pev[2 * q + 1]) != hipSuccess) continue;
    int k = c->pev_kind[q];
    c->kprof_launches[k] += 1;
    c->kprof_ms[k] += ms;
    c->kprof_bytes[k] += c->pev_bytes[q];
  }
  c->pev_used = 0;
}

namespace {

thread_local std::string g_create_error;

int check_ctx(fs_ctx* c) { return c ? FS_OK : FS_BAD_ARG; }

void release_all(fs_ctx* c) {
  for (int d = 0; d < 3; d++) {
    fs_free(c->pos[d]);
    fs_free(c->vel[d]);
  }
  fs_free(c->staging);
  for (int a = 0; a < 3; a++)
    for (DBuf* b : {&c->g[a], &c->gbuf[a], &c->gvalid[a], &c->gvalidbuf[a], &c->gw[a], &c->glast[a], &c->gweights[a]})
      fs_free(*b);
  for (DBuf* b : {&c->surf_v, &c->surf_n, &c->surf_i, &c->apicc}) fs_free(*b);
  for (DBuf* b : {&c->fsdf, &c->fsdfbuf, &c->sdfvalid, &c->sdfvalidbuf, &c->collider, &c->pg, &c->adiag, &c->rhs,
                  &c->active, &c->keys, &c->slots, &c->perm, &c->counts, &c->offs, &c->scan_tmp, &c->red,
                  &c->recon_tmp, &c->sorted[0], &c->sorted[1], &c->sorted[2], &c->sorted[3], &c->p2g_w,
                  &c->p2g_tiles, &c->offs2, &c->p2g_w2, &c->p2g_tiles2, &c->rkeys, &c->rbinmax, &c->rbins})
    fs_free(*b);
  for (int q = 0; q < 6; q++) fs_free(c->an[q]);
  fs_pcg_release(c);
  fs_mg_release(c);
  fs_adj_release(c);
}

struct NamedArray {
  DBuf* buf;
  int dims[3];
  int u8;
};

bool lookup(fs_ctx* c, const std::string& n, NamedArray& out) {
  static const char* fnames[3] = {"u", "v", "w"};
  out.u8 = 0;
  for (int a = 0; a < 3; a++) {
    std::string f = fnames[a];
    DBuf* b = nullptr;
    int u8 = 0;
    if (n == f) b = &c->g[a];
    else if (n == f + "buf") b = &c->gbuf[a];
    else if (n == f + "_valid") { b = &c->gvalid[a]; u8 = 1; }
    else if (n == f + "_valid_buf") { b = &c->gvalidbuf[a]; u8 = 1; }
    else if (n == f + "w") b = &c->gw[a];
    else if (n == f + "_last") b = &c->glast[a];
    else if (n == f + "_weights") b = &c->gweights[a];
    if (b) {
      out.buf = b;
      out.u8 = u8;
      for (int d = 0; d < 3; d++) out.dims[d] = c->fdim[a][d];
      return true;
    }
  }
  DBuf* b = nullptr;
  int u8 = 0;
  if (n == "fluid_sdf") b = &c->fsdf;
  else if (n == "fluid_sdf_buf") b = &c->fsdfbuf;
  else if (n == "sdf_valid") { b = &c->sdfvalid; u8 = 1; }
  else if (n == "sdf_valid_buf") { b = &c->sdfvalidbuf; u8 = 1; }
  else if (n == "collider_sdf") b = &c->collider;
  else if (n == "pg") b = &c->pg;
  else if (n == "adiag") b = &c->adiag;
  else if (n == "a_left") b = &c->an[0];
  else if (n == "a_right") b = &c->an[1];
  else if (n == "a_up") b = &c->an[2];
  else if (n == "a_down") b = &c->an[3];
  else if (n == "a_front") b = &c->an[4];
  else if (n == "a_back") b = &c->an[5];
  else if (n == "rhs") b = &c->rhs;
  else if (n == "active") { b = &c->active; u8 = 1; }
  else if (n == "mic_e") b = nullptr;
  if (!b && n != "mic_e") return false;
  out.buf = b;
  out.u8 = u8;
  out.dims[0] = c->nx;
  out.dims[1] = c->ny;
  out.dims[2] = c->nz;
  return true;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

}  // namespace

// APIC affine rows: device layout [3a + d][particle] at storage width <-> AoS doubles
// [particle][a][d] across the ABI.
template <typename T>
__global__ void k_apic_to_aos(int64_t n, const T* __restrict__ cm, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < 9; q++) out[9 * i + q] = (double)cm[q * n + i];
}
template <typename T>
__global__ void k_apic_from_aos(int64_t n, const double* __restrict__ in, T* __restrict__ cm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < 9; q++) cm[q * n + i] = (T)in[9 * i + q];
}

extern "C" {

void fs_config_default(fs_config* cfg) {
  if (!cfg) return;
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
  cfg->profile_stages = 0;
  cfg->preconditioner = FS_PRECOND_MIC;
}

int fs_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* fs_last_error(const fs_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int fs_create(const fs_config* cfg, fs_ctx** out) {
  if (!cfg || !out) return FS_BAD_ARG;
  *out = nullptr;
  g_create_error.clear();
  if (cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0 || !(cfg->size_x > 0) || !(cfg->size_y > 0) || !(cfg->size_z > 0)) {
    g_create_error = "resolution and size must be positive";
    return FS_BAD_ARG;
  }
  if (cfg->nx != cfg->ny) {
    // FvmSolver3D::buildSystem bounds the "up" neighbour by fluidSdf.width()
    // (project-solver.cc:186): for nx != ny the reference system is asymmetric or reads out of
    // range, so only nx == ny is a defined configuration.
    g_create_error = "nx must equal ny (the reference assembly bounds the up-neighbour by the width)";
    return FS_UNSUPPORTED;
  }
  if ((int64_t)(cfg->nx + 1) * (cfg->ny + 1) * (cfg->nz + 1) >= (int64_t)1 << 31) {
    g_create_error = "grid too large for 32-bit indexing";
    return FS_BAD_ARG;
  }
  if (cfg->precision != FS_PRECISION_FP64 && cfg->precision != FS_PRECISION_MIXED) {
    g_create_error = "unknown precision";
    return FS_BAD_ARG;
  }
  if (cfg->preconditioner != FS_PRECOND_MIC && cfg->preconditioner != FS_PRECOND_MULTIGRID) {
    g_create_error = "unknown preconditioner";
    return FS_BAD_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_create_error = "no HIP device";
    return FS_NO_DEVICE;
  }
  fs_ctx* c = new (std::nothrow) fs_ctx();
  if (!c) return FS_DEVICE_ERROR;
  c->cfg = *cfg;
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  c->nx = cfg->nx;
  c->ny = cfg->ny;
  c->nz = cfg->nz;
  c->h[0] = cfg->size_x / cfg->nx;
  c->h[1] = cfg->size_y / cfg->ny;
  c->h[2] = cfg->size_z / cfg->nz;
  c->mixed = cfg->precision == FS_PRECISION_MIXED;
  c->rsz = c->mixed ? 4 : 8;
  c->n_cell = c->nx * c->ny * c->nz;
  int dims[3][3] = {{c->nx + 1, c->ny, c->nz}, {c->nx, c->ny + 1, c->nz}, {c->nx, c->ny, c->nz + 1}};
  for (int a = 0; a < 3; a++) {
    for (int d = 0; d < 3; d++) c->fdim[a][d] = dims[a][d];
    c->n_face[a] = dims[a][0] * dims[a][1] * dims[a][2];
  }
  auto fail = [&](int st) {
    g_create_error = c->err;
    release_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return st;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->err = "hipStreamCreate failed";
    return fail(FS_DEVICE_ERROR);
  }
  const size_t R = c->rsz;
  int st = FS_OK;
  for (int a = 0; a < 3 && st == FS_OK; a++) {
    size_t n = (size_t)c->n_face[a];
    for (DBuf* b : {&c->g[a], &c->gbuf[a], &c->gw[a], &c->glast[a], &c->gweights[a]})
      if (st == FS_OK) st = fs_alloc(c, *b, n * R);
    for (DBuf* b : {&c->gvalid[a], &c->gvalidbuf[a]})
      if (st == FS_OK) st = fs_alloc(c, *b, n);
  }
  size_t nc = (size_t)c->n_cell;
  for (DBuf* b : {&c->fsdf, &c->fsdfbuf, &c->collider, &c->pg})
    if (st == FS_OK) st = fs_alloc(c, *b, nc * R);
  // the assembled system, at storage width (the pressure solve widens it to FP64 where needed)
  for (DBuf* b : {&c->adiag, &c->rhs, &c->an[0], &c->an[1], &c->an[2], &c->an[3], &c->an[4], &c->an[5]})
    if (st == FS_OK) st = fs_alloc(c, *b, nc * R);
  for (DBuf* b : {&c->sdfvalid, &c->sdfvalidbuf, &c->active})
    if (st == FS_OK) st = fs_alloc(c, *b, nc);
  if (st == FS_OK) st = fs_alloc(c, c->red, 64);
  if (st != FS_OK) return fail(st);
  // every buffer starts zeroed, like the reference's value-initialised std::vectors
  for (int a = 0; a < 3 && st == FS_OK; a++)
    for (DBuf* b : {&c->g[a], &c->gbuf[a], &c->gw[a], &c->glast[a], &c->gweights[a], &c->gvalid[a], &c->gvalidbuf[a]})
      if (st == FS_OK) st = fs_zero(c, *b);
  for (DBuf* b : {&c->fsdf, &c->fsdfbuf, &c->collider, &c->pg, &c->adiag, &c->rhs, &c->an[0], &c->an[1], &c->an[2],
                  &c->an[3], &c->an[4], &c->an[5], &c->sdfvalid, &c->sdfvalidbuf, &c->active})
    if (st == FS_OK) st = fs_zero(c, *b);
  if (st == FS_OK && hipStreamSynchronize(c->stream) != hipSuccess) {
    c->err = "initial clear failed";
    st = FS_DEVICE_ERROR;
  }
  if (st != FS_OK) return fail(st);
  bool ok = true;
  for (int q = 0; q < 12 && ok; q++) ok = hipEventCreate(&c->ev[q]) == hipSuccess;
  c->ev_ok = ok;
  if (cfg->profile_stages >= 2) {
    ok = true;
    for (int q = 0; q < fs_ctx::NPROF_EV && ok; q++) ok = hipEventCreate(&c->pev[q]) == hipSuccess;
    c->pev_ok = ok;
  }
  *out = c;
  return FS_OK;
}

void fs_destroy(fs_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->side2) (void)hipStreamSynchronize(c->side2);
  release_all(c);
  if (c->p2g_ev_ok)
    for (int q = 0; q < 6; q++) (void)hipEventDestroy(c->p2g_ev[q]);
  if (c->sdf_ev_ok)
    for (int q = 0; q < 2; q++) (void)hipEventDestroy(c->sdf_ev[q]);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->side2) (void)hipStreamDestroy(c->side2);
  if (c->ev_ok)
    for (int q = 0; q < 12; q++) (void)hipEventDestroy(c->ev[q]);
  if (c->pev_ok)
    for (int q = 0; q < fs_ctx::NPROF_EV; q++) (void)hipEventDestroy(c->pev[q]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int fs_set_collider_sdf(fs_ctx* c, const double* sdf) {
  FS_TRY(check_ctx(c));
  if (!sdf) return FS_BAD_ARG;
  return fs_upload_real(c, c->collider, sdf, (size_t)c->n_cell);
}

int fs_set_particles(fs_ctx* c, const double* pos, const double* vel, int64_t n) {
  FS_TRY(check_ctx(c));
  if (n < 0 || (n > 0 && !pos) || n >= ((int64_t)1 << 31) / 3) {
    c->set_error("bad particle count or null positions");
    return FS_BAD_ARG;
  }
  const double ext[3] = {c->nx * c->h[0], c->ny * c->h[1], c->nz * c->h[2]};
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double p = pos[3 * i + d];
      if (!(p >= 0.0 && p <= ext[d])) {
        c->set_error("particle " + std::to_string(i) + " lies outside the domain");
        return FS_BAD_ARG;
      }
    }
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_adj_release(c);  // a recorded rollout refers to the previous particles
  fs_free(c->apicc);  // APIC affine rows restart at zero
  if (n != c->np) {
    for (int d = 0; d < 3; d++) {
      fs_free(c->pos[d]);
      fs_free(c->vel[d]);
    }
    fs_free(c->staging);
    fs_free(c->keys);
    fs_free(c->slots);
    fs_free(c->perm);
    for (int q = 0; q < 4; q++) fs_free(c->sorted[q]);
    fs_free(c->p2g_w);
    fs_free(c->p2g_w2);
    fs_free(c->rkeys);
    c->np = 0;
    for (int d = 0; d < 3; d++) {
      FS_TRY(fs_alloc(c, c->pos[d], (size_t)n * c->rsz));
      FS_TRY(fs_alloc(c, c->vel[d], (size_t)n * c->rsz));
    }
    FS_TRY(fs_alloc(c, c->staging, (size_t)n * 3 * sizeof(double)));
    FS_TRY(fs_alloc(c, c->keys, (size_t)n * sizeof(int)));
    FS_TRY(fs_alloc(c, c->slots, (size_t)n * sizeof(int)));
    FS_TRY(fs_alloc(c, c->perm, (size_t)n * sizeof(int)));
    for (int q = 0; q < 4; q++) FS_TRY(fs_alloc(c, c->sorted[q], (size_t)n * c->rsz));
    FS_TRY(fs_alloc(c, c->p2g_w, (size_t)n * 12 * c->rsz));
    FS_TRY(fs_alloc(c, c->p2g_w2, (size_t)n * 12 * c->rsz));
    FS_TRY(fs_alloc(c, c->rkeys, (size_t)n * sizeof(int)));
    c->np = n;
  }
  return fs_particles_upload(c, pos, vel);
}

int fs_get_particles(fs_ctx* c, double* pos, double* vel) {
  FS_TRY(check_ctx(c));
  return fs_particles_download(c, pos, vel);
}

int64_t fs_num_particles(const fs_ctx* c) { return c ? c->np : -1; }

int fs_get_particle_affine(fs_ctx* c, double* c9) {
  FS_TRY(check_ctx(c));
  if (!c9) return FS_BAD_ARG;
  if (c->np == 0) return FS_OK;
  FS_TRY(fs_ensure_apic(c));
  DBuf tmp;
  FS_TRY(fs_alloc(c, tmp, (size_t)c->np * 9 * sizeof(double)));
  const int nb = (int)((c->np + 255) / 256);
  if (c->mixed) k_apic_to_aos<float><<<nb, 256, 0, c->stream>>>(c->np, c->apicc.as<float>(), tmp.as<double>());
  else k_apic_to_aos<double><<<nb, 256, 0, c->stream>>>(c->np, c->apicc.as<double>(), tmp.as<double>());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(c9, tmp.p, (size_t)c->np * 9 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  fs_free(tmp);
  if (e != hipSuccess) {
    c->set_error(hipGetErrorString(e));
    return FS_DEVICE_ERROR;
  }
  return FS_OK;
}

int fs_set_particle_affine(fs_ctx* c, const double* c9) {
  FS_TRY(check_ctx(c));
  if (!c9) return FS_BAD_ARG;
  if (c->np == 0) return FS_OK;
  FS_TRY(fs_ensure_apic(c));
  DBuf tmp;
  FS_TRY(fs_alloc(c, tmp, (size_t)c->np * 9 * sizeof(double)));
  hipError_t e = hipMemcpyAsync(tmp.p, c9, (size_t)c->np * 9 * sizeof(double), hipMemcpyHostToDevice, c->stream);
  const int nb = (int)((c->np + 255) / 256);
  if (e == hipSuccess) {
    if (c->mixed) k_apic_from_aos<float><<<nb, 256, 0, c->stream>>>(c->np, tmp.as<double>(), c->apicc.as<float>());
    else k_apic_from_aos<double><<<nb, 256, 0, c->stream>>>(c->np, tmp.as<double>(), c->apicc.as<double>());
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  fs_free(tmp);
  if (e != hipSuccess) {
    c->set_error(hipGetErrorString(e));
    return FS_DEVICE_ERROR;
  }
  return FS_OK;
}

int fs_set_advector(fs_ctx* c, int32_t method) {
  FS_TRY(check_ctx(c));
  if (method == FS_ADVECT_PIC || method == FS_ADVECT_FLIP || method == FS_ADVECT_APIC) {
    c->cfg.advection = method;
    return FS_OK;
  }
  c->set_error("unknown advection method");
  return FS_BAD_ARG;
}

int fs_set_projector(fs_ctx* c, int32_t method) {
  FS_TRY(check_ctx(c));
  if (method == FS_PROJECT_FVM) return FS_OK;
  c->set_error("only the FVM projector exists in the reference");
  return FS_UNSUPPORTED;
}

int fs_set_compressed_solver(fs_ctx* c, int32_t solver, int32_t precond) {
  FS_TRY(check_ctx(c));
  if (solver != FS_SOLVER_CG) {
    c->set_error("unknown compressed solver");
    return FS_UNSUPPORTED;
  }
  if (precond != FS_PRECOND_MIC && precond != FS_PRECOND_MULTIGRID) {
    c->set_error("unknown preconditioner");
    return FS_BAD_ARG;
  }
  c->cfg.preconditioner = precond;
  return FS_OK;
}

int fs_set_reconstructor(fs_ctx* c, int32_t method) {
  FS_TRY(check_ctx(c));
  if (method == FS_RECON_NAIVE) return FS_OK;
  c->set_error("unknown reconstructor");
  return FS_UNSUPPORTED;
}

int fs_clear(fs_ctx* c) {
  FS_TRY(check_ctx(c));
  return fs_stage_clear(c);
}
int fs_advect(fs_ctx* c, double dt) {
  FS_TRY(check_ctx(c));
  return fs_stage_advect(c, dt);
}
int fs_reconstruct(fs_ctx* c, double radius) {
  FS_TRY(check_ctx(c));
  return fs_stage_reconstruct(c, radius);
}
int fs_extrapolate_fluid_sdf(fs_ctx* c, int32_t iters) {
  FS_TRY(check_ctx(c));
  return fs_stage_extrapolate_sdf(c, iters);
}
int fs_smooth_fluid_surface(fs_ctx* c, int32_t iters) {
  FS_TRY(check_ctx(c));
  return fs_stage_smooth(c, iters);
}
int fs_p2g(fs_ctx* c, double) {
  FS_TRY(check_ctx(c));
  return fs_stage_p2g(c);
}
int fs_apply_dirichlet_boundary(fs_ctx* c) {
  FS_TRY(check_ctx(c));
  return fs_stage_dirichlet(c);
}
int fs_extrapolate_velocity(fs_ctx* c, int32_t iters) {
  FS_TRY(check_ctx(c));
  return fs_stage_extrapolate_velocity(c, iters);
}
int fs_apply_force(fs_ctx* c, double dt) {
  FS_TRY(check_ctx(c));
  return fs_stage_force(c, dt);
}
int fs_build_system(fs_ctx* c, double dt) {
  FS_TRY(check_ctx(c));
  return fs_stage_build(c, dt);
}
int fs_solve_pressure(fs_ctx* c, fs_stats* st) {
  FS_TRY(check_ctx(c));
  if (st) std::memset(st, 0, sizeof(*st));
  int s = fs_stage_solve(c, st);
  if (st) st->status = s;
  return s;
}
int fs_project(fs_ctx* c, double dt) {
  FS_TRY(check_ctx(c));
  return fs_stage_project(c, dt);
}
int fs_apply_collider(fs_ctx* c) {
  FS_TRY(check_ctx(c));
  return fs_stage_collider(c);
}
int fs_g2p(fs_ctx* c, double) {
  FS_TRY(check_ctx(c));
  return fs_stage_g2p(c);
}
int fs_cfl(fs_ctx* c, double* out) {
  FS_TRY(check_ctx(c));
  if (!out) return FS_BAD_ARG;
  return fs_stage_cfl(c, out);
}

int fs_spmv(fs_ctx* c, const double* x, double* y) {
  FS_TRY(check_ctx(c));
  if (!x || !y) return FS_BAD_ARG;
  return fs_pcg_spmv_dense(c, x, y);
}

int fs_mic_apply(fs_ctx* c, const double* r, double* z) {
  FS_TRY(check_ctx(c));
  if (!r || !z) return FS_BAD_ARG;
  return fs_pcg_mic_dense(c, r, z);
}

int fs_mg_apply(fs_ctx* c, const double* r, double* z) {
  FS_TRY(check_ctx(c));
  if (!r || !z) return FS_BAD_ARG;
  return fs_mg_apply_dense(c, r, z);
}

// cpu::FluidSimulator::substep (FluidSim/src/cpu/fluid-simulator.cc:140-182).
int fs_substep(fs_ctx* c, double dt, fs_stats* st) {
  FS_TRY(check_ctx(c));
  fs_stats local;
  fs_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof(*S));
  if (c->tape.on && c->tape.recording) FS_TRY(fs_adj_on_substep(c, dt));
  return fs_substep_core(c, dt, S, nullptr);
}

}  // extern "C"

// cpu::FluidSimulator::substep (fluid-simulator.cc:140-182); `rec` records what the reverse
// pass needs.
int fs_substep_core(fs_ctx* c, double dt, fs_stats* S, AdjStep* rec) {
  const bool prof = c->cfg.profile_stages && c->ev_ok;
  if (rec) {
    rec->dt = dt;
    FS_TRY(fs_adj_record(c, rec, ADJ_INPUT));
  }
  auto mark = [&](int q) {
    if (prof) (void)hipEventRecord(c->ev[q], c->stream);
  };
  mark(0);
  if (c->cfg.advect_on_cleared_grid) {
    FS_TRY(fs_stage_clear(c));
    FS_TRY(fs_stage_advect(c, dt));
  } else {
    FS_TRY(fs_stage_advect(c, dt));
    FS_TRY(fs_stage_clear(c));
  }
  if (rec) FS_TRY(fs_adj_record(c, rec, ADJ_ADVECTED));
  mark(1);
  // The fluid-SDF stages and the velocity stages (P2G .. force) are independent until the system
  // is built: without a recording, without per-stage timing (profile_stages 1) and without
  // per-launch event timing, the SDF stages run on side2, forked from and joined back into the
  // context's stream. Every return after the fork joins side2 first (SideJoin), so no SDF work
  // is left running behind a later call that only synchronises the context's stream.
  const bool sdf_side =
      !rec && c->cfg.profile_stages != 1 && !(c->cfg.profile_stages >= 2 && c->ktime_on);
  struct SideJoin {
    fs_ctx* c = nullptr;
    ~SideJoin() {
      if (!c) return;
      (void)hipEventRecord(c->sdf_ev[1], c->side2);
      (void)hipStreamWaitEvent(c->stream, c->sdf_ev[1], 0);
    }
  } side_join;
  auto sdf_stages = [&]() -> int {
    FS_TRY(fs_stage_reconstruct(c, c->cfg.recon_radius_factor * c->h[0]));
    FS_TRY(fs_stage_extrapolate_sdf(c, c->cfg.sdf_extrapolate_iters));
    FS_TRY(fs_stage_smooth(c, c->cfg.sdf_smooth_iters));
    return FS_OK;
  };
  if (sdf_side) {
    if (!c->side2) FS_HIP(c, hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking));
    if (!c->sdf_ev_ok) {
      for (int q = 0; q < 2; q++) FS_HIP(c, hipEventCreateWithFlags(&c->sdf_ev[q], hipEventDisableTiming));
      c->sdf_ev_ok = true;
    }
    FS_HIP(c, hipEventRecord(c->sdf_ev[0], c->stream));
    FS_HIP(c, hipStreamWaitEvent(c->side2, c->sdf_ev[0], 0));
    side_join.c = c;
    hipStream_t main_stream = c->stream;
    c->stream = c->side2;  // the stage functions enqueue on c->stream
    const int st = sdf_stages();
    c->stream = main_stream;
    FS_TRY(st);
  } else {
    FS_TRY(sdf_stages());
  }
  mark(2);
  FS_TRY(fs_stage_p2g(c));
  if (rec) FS_TRY(fs_adj_record(c, rec, ADJ_P2G));
  FS_TRY(fs_stage_dirichlet(c));
  if (rec) FS_TRY(fs_adj_record(c, rec, ADJ_DIRICHLET));
  mark(3);
  FS_TRY(fs_stage_extrapolate_velocity(c, c->cfg.vel_extrapolate_iters));
  FS_TRY(fs_stage_force(c, dt));
  mark(4);
  if (sdf_side) {  // join the SDF stages
    side_join.c = nullptr;
    FS_HIP(c, hipEventRecord(c->sdf_ev[1], c->side2));
    FS_HIP(c, hipStreamWaitEvent(c->stream, c->sdf_ev[1], 0));
  }
  FS_TRY(fs_stage_build(c, dt));
  if (rec) FS_TRY(fs_adj_record(c, rec, ADJ_BUILT));
  mark(5);
  int solve = fs_stage_solve(c, S);
  if (solve == FS_DEVICE_ERROR || solve == FS_UNSUPPORTED || solve == FS_BAD_ARG) return solve;
  mark(6);
  FS_TRY(fs_stage_project(c, dt));
  FS_TRY(fs_stage_collider(c));
  if (rec) FS_TRY(fs_adj_record(c, rec, ADJ_FINAL));
  mark(7);
  FS_TRY(fs_stage_g2p(c));
  mark(8);
  FS_HIP(c, hipStreamSynchronize(c->stream));
  if (prof) {
    S->ms_advect = elapsed(c->ev[0], c->ev[1]);
    S->ms_reconstruct = elapsed(c->ev[1], c->ev[2]);
    S->ms_p2g = elapsed(c->ev[2], c->ev[3]);
    S->ms_extrapolate = elapsed(c->ev[3], c->ev[4]);
    S->ms_build = elapsed(c->ev[4], c->ev[5]);
    S->ms_pcg = elapsed(c->ev[5], c->ev[6]);
    S->ms_project = elapsed(c->ev[6], c->ev[7]);
    S->ms_g2p = elapsed(c->ev[7], c->ev[8]);
    S->ms_total = elapsed(c->ev[0], c->ev[8]);
  }
  S->status = solve;
  return solve;
}

extern "C" {

// cpu::FluidSimulator::step (fluid-simulator.cc:205-220)
int fs_step_frame(fs_ctx* c, double frame_dt, fs_stats* st) {
  FS_TRY(check_ctx(c));
  double t = 0;
  int cnt = 0;
  int status = FS_OK;
  fs_stats local;
  fs_stats* S = st ? st : &local;
  while (t < frame_dt) {
    double cfl;
    FS_TRY(fs_stage_cfl(c, &cfl));
    double dt = cfl < frame_dt - t ? cfl : frame_dt - t;
    cnt++;
    status = fs_substep(c, dt, S);
    if (status != FS_OK && status != FS_NOT_CONVERGED) break;
    t += dt;
  }
  S->substeps = cnt;
  return status;
}

int fs_array_info(const fs_ctx* cc, const char* name, int64_t dims[3], int32_t* dtype) {
  fs_ctx* c = const_cast<fs_ctx*>(cc);
  FS_TRY(check_ctx(c));
  if (!name || !dims || !dtype) return FS_BAD_ARG;
  NamedArray na;
  if (!lookup(c, name, na)) {
    c->set_error(std::string("unknown array ") + name);
    return FS_BAD_ARG;
  }
  for (int d = 0; d < 3; d++) dims[d] = na.dims[d];
  *dtype = na.u8;
  return FS_OK;
}

int fs_get_array(fs_ctx* c, const char* name, void* out) {
  FS_TRY(check_ctx(c));
  if (!name || !out) return FS_BAD_ARG;
  NamedArray na;
  if (!lookup(c, name, na)) {
    c->set_error(std::string("unknown array ") + name);
    return FS_BAD_ARG;
  }
  size_t n = (size_t)na.dims[0] * na.dims[1] * na.dims[2];
  if (std::string(name) == "mic_e") return fs_pcg_export_e(c, static_cast<double*>(out));
  if (na.u8) {
    FS_HIP(c, hipMemcpyAsync(out, na.buf->p, n, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    return FS_OK;
  }
  return fs_download_real(c, *na.buf, static_cast<double*>(out), n);
}

int fs_set_array(fs_ctx* c, const char* name, const void* in) {
  FS_TRY(check_ctx(c));
  if (!name || !in) return FS_BAD_ARG;
  NamedArray na;
  if (!lookup(c, name, na) || !na.buf) {
    c->set_error(std::string("unknown or read-only array ") + name);
    return FS_BAD_ARG;
  }
  size_t n = (size_t)na.dims[0] * na.dims[1] * na.dims[2];
  c->pcg.built = false;
  if (na.u8) {
    FS_HIP(c, hipMemcpyAsync(na.buf->p, in, n, hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    return FS_OK;
  }
  int s = fs_upload_real(c, *na.buf, static_cast<const double*>(in), n);
  if (s == FS_OK) FS_HIP(c, hipStreamSynchronize(c->stream));
  return s;
}

int fs_synchronize(fs_ctx* c) {
  FS_TRY(check_ctx(c));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_prof_flush(c);
  return FS_OK;
}

int fs_kernel_profile(fs_ctx* c, int32_t kind, int64_t* launches, double* total_ms, double* total_bytes) {
  FS_TRY(check_ctx(c));
  if (kind < 0 || kind >= FS_KERNEL_COUNT) return FS_BAD_ARG;
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_prof_flush(c);
  if (launches) *launches = c->kprof_launches[kind];
  if (total_ms) *total_ms = c->kprof_ms[kind];
  if (total_bytes) *total_bytes = c->kprof_bytes[kind];
  return FS_OK;
}

int fs_set_kernel_timing(fs_ctx* c, int32_t on) {
  FS_TRY(check_ctx(c));
  c->ktime_on = on != 0;
  return FS_OK;
}

int fs_kernel_profile_reset(fs_ctx* c) {
  FS_TRY(check_ctx(c));
  FS_HIP(c, hipStreamSynchronize(c->stream));
  fs_prof_flush(c);
  for (int k = 0; k < FS_KERNEL_COUNT; k++) {
    c->kprof_launches[k] = 0;
    c->kprof_ms[k] = 0;
    c->kprof_bytes[k] = 0;
  }
  return FS_OK;
}

}  // extern "C"
