/*
 * A plain C caller of libsimcraft_fluid.so, compiled against include/simcraft_fluid.h only.
 * It follows the reference application's call order (construct, set collider, set fluid,
 * choose solvers, step; FluidSim/disabled/fluid-sim.cc:186-227):
 *
 *   capi_caller IN OUT
 *
 * IN  (binary, native endianness): int32 n, int32 advection, int32 steps, int32 pad,
 *     double dt, double flip_alpha, int64 particles,
 *     collider SDF n*n*n doubles, positions particles*3 doubles, velocities particles*3 doubles.
 * OUT (binary): int32 steps, then per step int32 status, int32 cg_iterations, double residual_inf;
 *     positions particles*3 doubles, velocities particles*3 doubles, pressure n*n*n doubles.
 *
 * Exit status: 0 on success, 1 on bad input, 2 when the library reports an error (its message
 * is printed).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "simcraft_fluid.h"

static int read_all(FILE* f, void* p, size_t size, size_t count) {
  return fread(p, size, count, f) == count;
}

static int fail(fs_ctx* ctx, const char* what, int st) {
  fprintf(stderr, "%s failed with status %d: %s\n", what, st, fs_last_error(ctx));
  if (ctx) fs_destroy(ctx);
  return 2;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 1;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 1;
  int32_t head[4];
  double dt, alpha;
  int64_t np;
  if (!read_all(in, head, sizeof(int32_t), 4) || !read_all(in, &dt, sizeof dt, 1) ||
      !read_all(in, &alpha, sizeof alpha, 1) || !read_all(in, &np, sizeof np, 1))
    return 1;
  const int32_t n = head[0], advection = head[1], steps = head[2];
  const size_t ncell = (size_t)n * n * n;
  double* sdf = malloc(ncell * sizeof(double));
  double* pos = malloc((size_t)np * 3 * sizeof(double));
  double* vel = malloc((size_t)np * 3 * sizeof(double));
  double* pg = malloc(ncell * sizeof(double));
  if (!sdf || !pos || !vel || !pg) return 1;
  if (!read_all(in, sdf, sizeof(double), ncell) || !read_all(in, pos, sizeof(double), (size_t)np * 3) ||
      !read_all(in, vel, sizeof(double), (size_t)np * 3))
    return 1;
  fclose(in);

  fs_config cfg;
  fs_config_default(&cfg);
  cfg.nx = cfg.ny = cfg.nz = n;
  cfg.flip_alpha = alpha;
  fs_ctx* ctx = NULL;
  int st = fs_create(&cfg, &ctx);
  if (st != FS_OK) return fail(NULL, "fs_create", st);
  if ((st = fs_set_collider_sdf(ctx, sdf)) != FS_OK) return fail(ctx, "fs_set_collider_sdf", st);
  if ((st = fs_set_particles(ctx, pos, vel, np)) != FS_OK) return fail(ctx, "fs_set_particles", st);
  if ((st = fs_set_advector(ctx, advection)) != FS_OK) return fail(ctx, "fs_set_advector", st);
  if ((st = fs_set_projector(ctx, FS_PROJECT_FVM)) != FS_OK) return fail(ctx, "fs_set_projector", st);
  if ((st = fs_set_compressed_solver(ctx, FS_SOLVER_CG, FS_PRECOND_MIC)) != FS_OK)
    return fail(ctx, "fs_set_compressed_solver", st);

  FILE* out = fopen(argv[2], "wb");
  if (!out) return fail(ctx, "fopen", 0);
  fwrite(&steps, sizeof steps, 1, out);
  for (int s = 0; s < steps; s++) {
    fs_stats stats;
    st = fs_substep(ctx, dt, &stats);
    if (st != FS_OK && st != FS_NOT_CONVERGED) return fail(ctx, "fs_substep", st);
    printf("substep %d: status %d, %d PCG iterations, residual %.3e\n", s, st, stats.cg_iterations,
           stats.residual_inf);
    fwrite(&stats.status, sizeof(int32_t), 1, out);
    fwrite(&stats.cg_iterations, sizeof(int32_t), 1, out);
    fwrite(&stats.residual_inf, sizeof(double), 1, out);
  }
  if (fs_num_particles(ctx) != np) return fail(ctx, "fs_num_particles", -1);
  if ((st = fs_get_particles(ctx, pos, vel)) != FS_OK) return fail(ctx, "fs_get_particles", st);
  if ((st = fs_get_array(ctx, "pg", pg)) != FS_OK) return fail(ctx, "fs_get_array", st);
  fwrite(pos, sizeof(double), (size_t)np * 3, out);
  fwrite(vel, sizeof(double), (size_t)np * 3, out);
  fwrite(pg, sizeof(double), ncell, out);
  fclose(out);
  fs_destroy(ctx);
  free(sdf);
  free(pos);
  free(vel);
  free(pg);
  return 0;
}
