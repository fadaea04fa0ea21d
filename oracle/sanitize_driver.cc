// sanitize_driver.cc - drives the oracle (fs_oracle.cc) through every entry point on a small
// scene, for a build with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile,
// target `sanitize`). TEST INFRASTRUCTURE ONLY; exits non-zero when a call fails, and the
// sanitizers abort on the first error they find.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../include/simcraft_fluid.h"

extern "C" {
void fso_config_default(fs_config* cfg);
void* fso_create(const fs_config* cfg);
void fso_destroy(void* s);
int fso_set_dot_mode(void* s, int mode);
int fso_set_p2g_order(void* s, int order);
int fso_set_particles(void* s, const double* p, const double* v, int64_t n);
int fso_get_particles(void* s, double* p, double* v);
int fso_set_collider_sdf(void* s, const double* sdf);
int fso_set_collider_mesh(void* s, const double* v, int64_t nv, const uint32_t* ix, int64_t nt, int band, int flags);
int fso_substep(void* s, double dt, fs_stats* st);
int fso_step_frame(void* s, double frame_dt, fs_stats* st);
int fso_cfl(void* s, double* out);
int fso_spmv(void* s, const double* x, double* y);
int fso_mic_apply(void* s, const double* r, double* z);
int fso_get_array(void* s, const char* name, void* out);
int fso_rollout_begin(void* s);
int fso_rollout_stop(void* s);
int fso_rollout_vjp(void* s, const double* dpos, const double* dvel, double* dpos0, double* dvel0, double* dgrav);
int fso_get_apic(void* s, double* c);
int fso_mg_apply(void* s, const double* r, double* z);
int fso_set_apic(void* s, const double* c);
}

static int check(int st, const char* what) {
  if (st != 0 && st != FS_NOT_CONVERGED) {
    std::fprintf(stderr, "%s failed with status %d\n", what, st);
    return 1;
  }
  return 0;
}

int main() {
  const int n = 12;
  const double h = 1.0 / n;
  // the dam break block, one particle per octant of each fluid cell (a fixed jitter)
  std::vector<double> pos, vel;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        if (!((i + 0.5) * h < 0.4 && (j + 0.5) * h < 0.6)) continue;
        for (int q = 0; q < 8; q++) {
          const double o[3] = {0.25 + 0.5 * ((q >> 2) & 1), 0.25 + 0.5 * ((q >> 1) & 1), 0.25 + 0.5 * (q & 1)};
          const int c[3] = {i, j, k};
          for (int d = 0; d < 3; d++) {
            pos.push_back((c[d] + o[d] + 0.03 * std::sin(7.0 * q + d + i)) * h);
            vel.push_back(0.2 * std::cos(3.0 * q + d + j));
          }
        }
      }
  const int64_t np = (int64_t)pos.size() / 3;
  // a closed mesh collider: the octahedron around (0.7, 0.3, 0.5)
  const double cx = 0.7, cy = 0.3, cz = 0.5, r = 0.15;
  const std::vector<double> verts = {cx + r, cy, cz, cx - r, cy, cz, cx, cy + r, cz,
                                     cx, cy - r, cz, cx, cy, cz + r, cx, cy, cz - r};
  const std::vector<uint32_t> tris = {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
  int bad = 0;
  for (int advection : {(int)FS_ADVECT_FLIP, (int)FS_ADVECT_PIC, (int)FS_ADVECT_APIC}) {
    const bool apic = advection == FS_ADVECT_APIC;
    fs_config cfg;
    fso_config_default(&cfg);
    cfg.nx = cfg.ny = cfg.nz = n;
    cfg.flip_alpha = 0.95;
    cfg.advection = advection;
    void* s = fso_create(&cfg);
    if (!s) return 1;
    fso_set_dot_mode(s, 1);
    for (int flags : {0, (int)FS_MESH_EXACT_DISTANCE})
      bad |= check(fso_set_collider_mesh(s, verts.data(), 6, tris.data(), 8, 3, flags), "set_collider_mesh");
    bad |= check(fso_set_particles(s, pos.data(), vel.data(), np), "set_particles");
    std::vector<double> cm(9 * np);
    for (size_t e = 0; e < cm.size(); e++) cm[e] = 0.5 * std::sin(0.37 * (double)e);
    if (apic) {
      bad |= check(fso_set_apic(s, cm.data()), "set_apic");
      if (fso_rollout_begin(s) != FS_UNSUPPORTED) bad = 1;  // APIC has no reverse pass
    } else {
      bad |= check(fso_rollout_begin(s), "rollout_begin");
    }
    fs_stats st;
    for (int step = 0; step < 3; step++) bad |= check(fso_substep(s, 0.004, &st), "substep");
    fso_rollout_stop(s);
    std::vector<double> dp(3 * np, 1.0), dv(3 * np, -0.5), dp0(3 * np), dv0(3 * np), dg(3);
    if (apic)
      bad |= check(fso_get_apic(s, cm.data()), "get_apic");
    else
      bad |= check(fso_rollout_vjp(s, dp.data(), dv.data(), dp0.data(), dv0.data(), dg.data()), "rollout_vjp");
    double dt = 0;
    bad |= check(fso_cfl(s, &dt), "cfl");
    bad |= check(fso_step_frame(s, 0.01, &st), "step_frame");
    std::vector<double> x(n * n * n), y(n * n * n, 0.0), z(n * n * n, 0.0);
    for (size_t e = 0; e < x.size(); e++) x[e] = std::sin(0.1 * (double)e);
    bad |= check(fso_spmv(s, x.data(), y.data()), "spmv");
    bad |= check(fso_mic_apply(s, x.data(), z.data()), "mic_apply");
    bad |= check(fso_mg_apply(s, x.data(), z.data()), "mg_apply");
    bad |= check(fso_get_array(s, "fluid_sdf", y.data()), "get_array");
    fso_set_p2g_order(s, 2);
    bad |= check(fso_substep(s, 0.004, &st), "substep (parallel P2G)");
    fso_destroy(s);
    cfg.preconditioner = FS_PRECOND_MULTIGRID;  // the multigrid-preconditioned CG
    s = fso_create(&cfg);
    if (!s) return 1;
    fso_set_dot_mode(s, 1);
    bad |= check(fso_set_particles(s, pos.data(), vel.data(), np), "set_particles (multigrid)");
    for (int step = 0; step < 2; step++) bad |= check(fso_substep(s, 0.004, &st), "substep (multigrid)");
    bad |= check(fso_get_particles(s, dp.data(), dv.data()), "get_particles");
    fso_destroy(s);
  }
  std::printf("sanitized oracle run: %s (%lld particles)\n", bad ? "FAILED" : "ok", (long long)np);
  return bad;
}
