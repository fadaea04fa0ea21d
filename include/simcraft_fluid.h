/*
 * simcraft_fluid.h - C ABI of libsimcraft_fluid.so, an MI355X (gfx950) implementation of
 * SimCraft FluidSim's FLIP/PIC substep.
 *
 * Plain C: pointers, sizes and status codes only. All arrays crossing the ABI use the
 * reference's dense layout (Spatify LinearHashXYZ, Spatify/include/Spatify/hash.h:25):
 *   idx = k + nz' * (j + ny' * i)     (k fastest, x slowest)
 * where (nx', ny', nz') are the dimensions of the array in question:
 *   cell-centred arrays           nx     x ny     x nz
 *   u face grid (axis 0)          (nx+1) x ny     x nz
 *   v face grid (axis 1)          nx     x (ny+1) x nz
 *   w face grid (axis 2)          nx     x ny     x (nz+1)
 * Particles are AoS xyz doubles (std::vector<Vec3d> in the reference,
 * FluidSim/include/FluidSim/cpu/fluid-simulator.h:200-202).
 *
 * A context owns every device buffer and its HIP streams: the stream every entry point orders
 * its work on, plus two internal side streams that a substep forks from and joins back into that
 * stream before it returns (P2G gathers beside the next face grid's binning; the fluid-SDF stages
 * beside P2G and the velocity stages). Host arrays are copied at the call and never retained.
 * A context is not re-entrant; contexts are independent.
 * No entry point aborts the process: failures return a status and fs_last_error()
 * holds the message.
 */
#ifndef SIMCRAFT_FLUID_H_
#define SIMCRAFT_FLUID_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference's failure channels are an undefined ERROR() macro, asserts and
 * a std::cerr warning (FluidSim/src/cpu/fluid-simulator.cc:169-172); these replace them. */
typedef enum fs_status {
  FS_OK = 0,
  FS_NOT_CONVERGED = 1, /* PCG hit its iteration cap (CgSolver3D max_iterations) */
  FS_BREAKDOWN = 2,     /* s.z == 0 in PCG (asserted at project-solver.cc:330) */
  FS_NONFINITE = 3,     /* NaN/Inf residual */
  FS_BAD_ARG = 4,
  FS_DEVICE_ERROR = 5,  /* HIP error, or a device-side wait that exceeded its bound */
  FS_UNSUPPORTED = 6,   /* a configuration the library does not implement */
  FS_NO_DEVICE = 7
} fs_status;

/* Mirrors fluid::AdvectionMethod (FluidSim/include/FluidSim/fluid-simulator.h:14-18). */
enum { FS_ADVECT_PIC = 0, FS_ADVECT_FLIP = 1, FS_ADVECT_APIC = 2 };
/* Mirrors fluid::ProjectSolver (fluid-simulator.h:20-23). */
enum { FS_PROJECT_FVM = 0, FS_PROJECT_FDM = 1 };
/* Mirrors fluid::PreconditionerMethod (fluid-simulator.h:30-33). */
enum { FS_PRECOND_MULTIGRID = 0, FS_PRECOND_MIC = 1 };
/* Mirrors fluid::CompressedSolverMethod (fluid-simulator.h:35-37). */
enum { FS_SOLVER_CG = 0 };
/* Mirrors fluid::ReconstructorMethod (fluid-simulator.h:39-41). */
enum { FS_RECON_NAIVE = 0 };

/* Arithmetic of the whole step. */
enum {
  FS_PRECISION_FP64 = 0,  /* reference arithmetic: FP64 everywhere */
  FS_PRECISION_MIXED = 1  /* FP32 particles/grids/stencil/PCG vectors and sweeps; FP64 dot
                             products, norms and iterate, and FP64 residual replacement
                             (b - Ax) deciding convergence */
};

/* PCG stopping rule. */
enum {
  FS_TOL_ABS_LINF = 0, /* reference: ||r||_inf < tol (project-solver.cc:334-335, project-solver.h:129) */
  FS_TOL_REL_L2 = 1    /* ||r||_2 <= tol * ||rhs||_2 */
};

/* Dot products inside PCG on the device always use one fixed reduction tree, so results are
 * deterministic from run to run. */

typedef struct fs_config {
  int32_t nx, ny, nz;             /* resolution (cells) */
  double size_x, size_y, size_z;  /* domain extent; origin is (0,0,0) */
  int32_t precision;              /* FS_PRECISION_* */
  int32_t advection;              /* FS_ADVECT_* (setAdvector) */
  double flip_alpha;              /* FLIP blend, advect-solver.h:129 (0.97) */
  double mic_tau;                 /* MIC(0) tuning, project-solver.h:83 (0.97) */
  double mic_sigma;               /* MIC(0) safety, project-solver.h:84 (0.25) */
  double cg_tolerance;            /* project-solver.h:129 (1e-6) */
  int32_t cg_max_iterations;      /* project-solver.h:130 (300) */
  int32_t tol_mode;               /* FS_TOL_* */
  double gravity;                 /* applyForce, fluid-simulator.cc:16 (9.8, along -y) */
  int32_t sdf_extrapolate_iters;  /* fluid-simulator.cc:151 (10) */
  int32_t sdf_smooth_iters;       /* fluid-simulator.cc:152 (5) */
  int32_t vel_extrapolate_iters;  /* fluid-simulator.cc:160-162 (10) */
  double recon_radius_factor;     /* fluid-simulator.cc:147 (0.8 * h) */
  int32_t advect_on_cleared_grid; /* 1 = reference order: clear() before advect (fluid-simulator.cc:141-143) */
  double cfl_factor;              /* fluid-simulator.cc:202 (5.0) */
  double inner_tolerance;         /* unused, kept for layout: the mixed-precision "outer FP64
                                     refinement" is residual replacement inside one PCG run (the
                                     FP64 residual b - Ax replaces the FP32 recurrence whenever it
                                     has dropped tenfold and whenever it claims convergence), not
                                     nested solves with an inner tolerance */
  int32_t max_refinements;        /* unused, kept for layout (see inner_tolerance) */
  int32_t check_interval;         /* PCG iterations between host convergence reads */
  int32_t profile_stages;         /* 1 = record per-stage device times into fs_stats (the fluid-SDF
                                     stages then stay on the main stream, so ms_reconstruct is
                                     their time); 2 = also per-launch kernel timing (see
                                     fs_kernel_profile); while that timing is switched off, the SDF
                                     stages overlap P2G and ms_reconstruct is not separable */
  int32_t preconditioner;         /* FS_PRECOND_* (setCompressedSolver, cpu/fluid-simulator.h:109-113):
                                     FS_PRECOND_MIC (default) or FS_PRECOND_MULTIGRID, this
                                     library's aggregation multigrid (fs_mg_apply) */
  int32_t reserved[6];
} fs_config;

typedef struct fs_stats {
  int32_t status;         /* fs_status of the step */
  int32_t cg_iterations;  /* PCG iterations, counted as CgSolver3D::solve counts them (project-solver.cc:326-344) */
  int32_t refinements;    /* FP64 residual replacements (mixed precision), else 0 */
  int32_t n_active;       /* active cells in the last pressure solve */
  double residual_inf;    /* final ||r||_inf of the pressure system (FP64) */
  double residual_rel;    /* final ||r||_2 / ||rhs||_2 (FP64) */
  double rhs_norm2;       /* ||rhs||_2 */
  int32_t substeps;       /* substeps taken by fs_step_frame */
  int32_t pad0;
  double ms_advect, ms_reconstruct, ms_p2g, ms_extrapolate, ms_build, ms_pcg, ms_project, ms_g2p;
  double ms_total;
} fs_stats;

typedef struct fs_ctx fs_ctx;

/* Reference defaults (each field cites its source above). */
void fs_config_default(fs_config* cfg);

/* Construct / destroy. Replaces cpu::FluidSimulator(n, size, resolution)
 * (FluidSim/include/FluidSim/cpu/fluid-simulator.h:28-60) plus FluidSimulator::setBackend
 * (FluidSim/src/fluid-simulator.cc:4-9); the particle count is set by fs_set_particles. */
int fs_create(const fs_config* cfg, fs_ctx** out);
void fs_destroy(fs_ctx* ctx);
const char* fs_last_error(const fs_ctx* ctx);
/* Number of visible HIP devices (0 when none); never initialises more than the runtime. */
int fs_device_count(void);

/* ---- Backend level: FluidComputeBackend (fluid-simulator.h:43-53) and cpu::FluidSimulator ---- */
/* setCollider(mesh) (cpu/fluid-simulator.h:62-71) takes a mesh and builds the SDF with
 * manifold2SDF; this entry takes the cell-centred collider SDF (nx*ny*nz doubles) directly. */
int fs_set_collider_sdf(fs_ctx* ctx, const double* sdf);
/* setCollider(mesh) itself (cpu/fluid-simulator.h:62-71): the collider SDF of a closed triangle
 * mesh by manifold2SDF (FluidSim/src/cpu/util.cc:203-317), computed on the device. vertices: AoS
 * xyz doubles (core::Mesh::vertices, Core/include/Core/mesh.h:9-14); indices: three 0-based vertex
 * indices per triangle (core::Mesh::indices); exact_band: cells around each triangle's bounding
 * box that get exact distances before the sweeps (setCollider passes 3). Load an OBJ file with
 * the reference's core::myLoadObj (Core/src/obj-io.cc:13-64) and pass its arrays.
 * flags: 0 = the reference's point-triangle distance, which measures |p + q - 2a| instead of
 * |p - q| when the closest point q lies on edge AB, on edge AC or inside triangle abc
 * (util.cc:101-125); FS_MESH_EXACT_DISTANCE = |p - q| in every case. */
enum { FS_MESH_EXACT_DISTANCE = 1 };
int fs_set_collider_mesh(fs_ctx* ctx, const double* vertices, int64_t nverts, const uint32_t* indices,
                         int64_t ntris, int32_t exact_band, int32_t flags);
/* setInitialFluid (cpu/fluid-simulator.h:73-78) draws unseeded random particles; this entry
 * takes explicit positions (and velocities; NULL = zero). Positions must lie in the domain. */
int fs_set_particles(fs_ctx* ctx, const double* pos_xyz, const double* vel_xyz, int64_t n);
int fs_get_particles(fs_ctx* ctx, double* pos_xyz, double* vel_xyz);
int64_t fs_num_particles(const fs_ctx* ctx);
/* cpu/fluid-simulator.h:80-98. FS_ADVECT_APIC, which the reference declares without an
 * implementation (setAdvector reaches ERROR at :96-97), is affine particle-in-cell: each particle
 * carries per velocity component a the row c_a; P2G transfers v_a + c_a.(x_n - x_p) to every
 * node of its stencil, G2P sets v_a to the sampled velocity and c_a = D^-1 b (b = sum w g (x_n -
 * x_p), D = sum w (x_n - x_p)(x_n - x_p)^T, over the same nodes and weights). Not differentiated
 * by the reverse mode. */
int fs_set_advector(fs_ctx* ctx, int32_t method);
/* The APIC rows of every particle: 9 doubles per particle, [particle][component u, v, w][x, y, z];
 * zero after fs_set_particles. */
int fs_get_particle_affine(fs_ctx* ctx, double* c9);
int fs_set_particle_affine(fs_ctx* ctx, const double* c9);
int fs_set_projector(fs_ctx* ctx, int32_t method);                      /* cpu/fluid-simulator.h:100-107 */
int fs_set_compressed_solver(fs_ctx* ctx, int32_t solver, int32_t precond); /* :109-113 */
int fs_set_reconstructor(fs_ctx* ctx, int32_t method);                  /* :168-179 */
/* cpu::FluidSimulator::substep(dt) (FluidSim/src/cpu/fluid-simulator.cc:140-182). */
int fs_substep(fs_ctx* ctx, double dt, fs_stats* stats);
/* cpu::FluidSimulator::step(Frame&) (fluid-simulator.cc:205-220): CFL-limited substeps
 * until frame_dt is covered. */
int fs_step_frame(fs_ctx* ctx, double frame_dt, fs_stats* stats);
/* CFL() (fluid-simulator.cc:184-203). */
int fs_cfl(fs_ctx* ctx, double* dt_out);

/* rebuildSurface (FluidSim/src/cpu/rebuild-surface.cc:732-741): a marching-cubes triangle mesh of
 * the current fluid SDF, built on the device and kept there; returns its vertex and triangle
 * counts. Vertices lie on the sign-changing edges between cell centres (samples outside the grid
 * read 1000), with the reference's interpolated positions and normals; the triangles of a cube
 * follow this library's case table (the reference's comes from an external table), vertices are
 * numbered in the order of their edges and triangles in the order of their cubes. */
int fs_rebuild_surface(fs_ctx* ctx, int64_t* nverts, int64_t* ntris);
/* Copies the last rebuilt surface: vertices and normals AoS xyz (nverts*3 doubles each),
 * indices three per triangle (ntris*3). Any output may be NULL. */
int fs_get_surface(fs_ctx* ctx, double* vertices, double* normals, uint32_t* indices);

/* ---- Operator level: the stages of substep, each on the context's device state ---- */
int fs_clear(fs_ctx* ctx);                                   /* fluid-simulator.cc:20-37 */
int fs_advect(fs_ctx* ctx, double dt);                       /* advect-solver.cc:170-187 (+handleCollision :100-137) */
int fs_reconstruct(fs_ctx* ctx, double radius);              /* sdf.h:210-225, ns-util.h:46-109 */
int fs_extrapolate_fluid_sdf(fs_ctx* ctx, int32_t iters);    /* cpu/fluid-simulator.h:123-166 */
int fs_smooth_fluid_surface(fs_ctx* ctx, int32_t iters);     /* fluid-simulator.cc:39-74 */
int fs_p2g(fs_ctx* ctx, double dt);                          /* advect-solver.cc:188-271 (PIC :18-98) */
int fs_apply_dirichlet_boundary(fs_ctx* ctx);                /* fluid-simulator.cc:113-138 */
int fs_extrapolate_velocity(fs_ctx* ctx, int32_t iters);     /* cpu/fluid-simulator.h:204-253, 3 grids */
int fs_apply_force(fs_ctx* ctx, double dt);                  /* fluid-simulator.cc:13-18 */
int fs_build_system(fs_ctx* ctx, double dt);                 /* project-solver.cc:59-228 */
int fs_solve_pressure(fs_ctx* ctx, fs_stats* stats);         /* CgSolver3D::solve, project-solver.cc:308-346 */
int fs_project(fs_ctx* ctx, double dt);                      /* project-solver.cc:348-435 */
int fs_apply_collider(fs_ctx* ctx);                          /* fluid-simulator.cc:76-111 */
int fs_g2p(fs_ctx* ctx, double dt);                          /* advect-solver.cc:157-169 (PIC :7-17) */

/* applyCompressedMatrix (project-solver.cc:16-39) on the assembled system: y = A x over active
 * cells (inactive entries of y are left untouched). x, y: dense nx*ny*nz. */
int fs_spmv(fs_ctx* ctx, const double* x, double* y);
/* ModifiedIncompleteCholesky3D::precond (project-solver.cc:230-306): z = M^-1 r over active
 * cells of the assembled system. */
int fs_mic_apply(fs_ctx* ctx, const double* r, double* z);
/* The multigrid preconditioner of FS_PRECOND_MULTIGRID (declared without an implementation in the
 * reference, project-solver.h:117-118): one V-cycle of the assembled system, z = M^-1 r on every
 * cell (r read on active cells; z is 0 on inactive ones). */
int fs_mg_apply(fs_ctx* ctx, const double* r, double* z);

/* ---- State access (dense reference layout) ----
 * Names: "u","v","w","ubuf","vbuf","wbuf","u_valid","v_valid","w_valid","u_valid_buf",
 * "v_valid_buf","w_valid_buf","uw","vw","ww","u_last","v_last","w_last","fluid_sdf",
 * "fluid_sdf_buf","sdf_valid","sdf_valid_buf","collider_sdf","u_weights","v_weights",
 * "w_weights","adiag","a_left","a_right","a_up","a_down","a_front","a_back","rhs",
 * "active","pg","mic_e".
 * dtype: 0 = float64, 1 = uint8. Floating arrays are always exchanged as float64. */
int fs_array_info(const fs_ctx* ctx, const char* name, int64_t dims[3], int32_t* dtype);
int fs_get_array(fs_ctx* ctx, const char* name, void* out);
int fs_set_array(fs_ctx* ctx, const char* name, const void* in);
int fs_synchronize(fs_ctx* ctx);

/* ---- Kernel timing (no counterpart in the reference) ----
 * With fs_config.profile_stages >= 2 the library brackets every launch of the kernels below
 * with HIP events on its own stream and accumulates, per kernel class: launches, device
 * milliseconds and algorithmic bytes (the minimum HBM traffic of the launch: every array the
 * kernel must read or write, counted once per active cell / face / particle it covers; masks
 * and the P2G binning keys excluded). */
enum {
  FS_KERNEL_SPMV = 0,       /* applyCompressedMatrix + s.z          48 B / active cell (FP64) */
  FS_KERNEL_MIC_FWD = 1,    /* r -= alpha t, ||r||inf, L q = r      64 B / active cell */
  FS_KERNEL_MIC_BWD = 2,    /* L^T z = q, z.r                       56 B / active cell */
  FS_KERNEL_PCG_UPDATE = 3, /* pg += alpha s, s = z + beta s        40 B / active cell */
  FS_KERNEL_MIC_FACTOR = 4, /* E factor                             90 B / active cell */
  FS_KERNEL_P2G = 5,        /* records + gather, one face grid */
  FS_KERNEL_G2P = 6,
  FS_KERNEL_ADVECT = 7,
  FS_KERNEL_RECONSTRUCT = 8,
  FS_KERNEL_MG_LEVEL0 = 9,  /* multigrid cycle on the grid: sweeps, restriction, correction, z.r */
  FS_KERNEL_MG_COARSE = 10, /* multigrid cycle below the grid (levels of 2x2x2 aggregates) */
  FS_KERNEL_COUNT = 11
};
int fs_kernel_profile(fs_ctx* ctx, int32_t kind, int64_t* launches, double* total_ms, double* total_bytes);
int fs_kernel_profile_reset(fs_ctx* ctx);
/* Turns the bracketing on (default) or off at run time, so that a timed run can leave the
 * events (and the synchronisations that drain them) out; only for profile_stages >= 2. */
int fs_set_kernel_timing(fs_ctx* ctx, int32_t on);

/* ---- Reverse mode (no counterpart in the reference, which has no derivative code) ----
 * Vector-Jacobian products of a rollout of fs_substep calls, for differentiable simulation.
 * The discrete structure of each substep (fluid SDF, marking, face weights, active set,
 * extrapolation masks, Poisson matrix) is held fixed; the transfers (including their kernel
 * weights' dependence on positions), collider, projection (through the same preconditioned CG on the
 * symmetric matrix), velocity extrapolation, gravity and advection are differentiated exactly.
 * Requires advect_on_cleared_grid = 1 (the reference order). Arrays are AoS xyz doubles. */
int fs_rollout_begin(fs_ctx* ctx, int32_t checkpoint_every); /* record the following substeps */
int fs_rollout_stop(fs_ctx* ctx);  /* stop appending substeps; the tape stays for fs_rollout_vjp */
int fs_rollout_end(fs_ctx* ctx);                             /* stop recording, free the tape */
int fs_rollout_steps(const fs_ctx* ctx);                     /* substeps recorded so far */
/* Given dL/dpos and dL/dvel of the particles at the end of the recorded substeps, returns dL/dpos
 * and dL/dvel of the particles at fs_rollout_begin and dL/dG for the gravity vector
 * G = (0, -gravity, 0) (the generalised force u_a += G_a dt on valid interior faces). Recomputes
 * the rollout segment by segment from checkpoints taken every `checkpoint_every` substeps; then
 * restores the state found at the call (particles, fluid SDF buffers, face grids, pressure).
 * Any output may be NULL. */
int fs_rollout_vjp(fs_ctx* ctx, const double* dpos, const double* dvel, double* dpos0, double* dvel0,
                   double* dgravity);
/* PIC transfer of the given face grids (dense, reference layout) to particles at pos_xyz (NULL:
 * the simulation's current particles; the simulation state is not changed), and its transpose:
 * grid gradients from particle-velocity gradients. */
int fs_sample_grid(fs_ctx* ctx, const double* pos_xyz, const double* u, const double* v, const double* w,
                   double* vel_xyz);
int fs_sample_grid_vjp(fs_ctx* ctx, const double* pos_xyz, const double* dvel_xyz, double* du, double* dv,
                       double* dw);

/* ---- Allocation-free entry (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP stream/graph capture and
 * serving loops that forbid hipMalloc. CFL() (fluid-simulator.cc:184-203, as fs_cfl) of face
 * grids the caller holds in device memory (u, v, w: dense reference layout, FP64, for an
 * nx*ny*nz cell grid of spacing h); the time step cfl_factor * max(1e-3, min(h/1e-6, h/max|u|))
 * is written to the device address dt_out on `stream` (a hipStream_t; NULL = the null stream),
 * asynchronously. Nothing is allocated: the reduction's partial maxima live in a fixed 8 KiB
 * device array of the library (one per GPU, held until the process ends), which every call
 * rewrites before reading. Calls that overlap on one GPU race on that array. Each face grid
 * must have fewer than 2^31 faces (FS_BAD_ARG otherwise, before anything is launched). */
int fs_cfl_device_noalloc(const double* u, const double* v, const double* w, int32_t nx, int32_t ny, int32_t nz,
                          double h, double cfl_factor, double* dt_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SIMCRAFT_FLUID_H_ */
