/*
 * cfd2d_hip.h - C ABI of the MI355X (gfx950) implementation of the cfd2d
 * projection time step (kaeptnzursee1910/CFD-, package cfd2d).
 *
 * Plain C: integers, doubles and pointers only.  Host arrays use the
 * reference's NumPy layout (C order, index [i, j], j contiguous):
 *     u, F, hu, hu_old : (nx+1) x (ny+2)     v, G, hv, hv_old : (nx+2) x (ny+1)
 *     p, phi, fluid    : (nx+2) x (ny+2)     pressure cell arrays : nx x ny
 * Every entry point returns 0 on success, > 0 for "solver not converged after
 * k iterations" (as scipy's `info`), < 0 on error; cfd_last_error() gives the
 * message of the last error of the calling thread.
 *
 * Ownership: a context owns all device memory it allocates (hipMalloc) and
 * runs on one HIP stream (its own, or one set by cfd_ctx_set_stream).  Calls
 * that only enqueue work return before the GPU has finished it; calls that
 * return host data synchronise the stream.  A context is used by one host
 * thread at a time.
 */
#ifndef CFD2D_HIP_H
#define CFD2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_OK 0
#define CFD_ERR_ARG (-1)
#define CFD_ERR_HIP (-2)
#define CFD_ERR_STATE (-3)
#define CFD_ERR_LIMIT (-4)
#define CFD_ERR_UNSTABLE (-5)

/* boundary.py:19-25 */
enum { CFD_BC_NOSLIP = 0, CFD_BC_SLIP = 1, CFD_BC_INFLOW = 2, CFD_BC_OUTFLOW = 3,
       CFD_BC_PERIODIC = 4 };
/* kernels.py:33 */
enum { CFD_LIM_UPWIND = 0, CFD_LIM_VANLEER = 1, CFD_LIM_MINMOD = 2, CFD_LIM_SUPERBEE = 3,
       CFD_LIM_LINEAR = 4, CFD_LIM_MC = 5 };
/* fields of a context, reference shapes as listed above; X is the pressure
   solver's warm start (pressure.py:140, 177) as an nx x ny array */
enum { CFD_FIELD_U = 0, CFD_FIELD_V = 1, CFD_FIELD_P = 2, CFD_FIELD_F = 3, CFD_FIELD_G = 4,
       CFD_FIELD_HU = 5, CFD_FIELD_HV = 6, CFD_FIELD_HU_OLD = 7, CFD_FIELD_HV_OLD = 8,
       CFD_FIELD_X = 9,
       /* turbulence model outputs, (nx+2) x (ny+2), download only: nu_t and the
          SST diffusivities of k and omega */
       CFD_FIELD_NUT = 10, CFD_FIELD_DIFF_K = 11, CFD_FIELD_DIFF_OMEGA = 12,
       /* two-phase VOF, (nx+2) x (ny+2): the phase fraction (upload sets the
          properties from it) and the mixture density and dynamic viscosity
          (download only) */
       CFD_FIELD_ALPHA = 13, CFD_FIELD_RHO = 14, CFD_FIELD_MU = 15,
       /* the solution before the last one, nx x ny, kept for the extrapolated
          initial iterate (cfd_set_initial_iterate mode 1); an upload continues
          the time series of X (checkpoint restart) */
       CFD_FIELD_XPREV = 16 };
/* pressure solvers: Jacobi-PCG in FP64 (the reference's "cg", pressure.py:171),
   and flexible CG with FP64 vectors and dot products preconditioned by an FP32
   aggregation-multigrid V-cycle (set up with cfd_set_multigrid), the solve
   ending on the FP64 true residual */
enum { CFD_SOLVER_PCG64 = 0, CFD_SOLVER_MGCG = 1, CFD_SOLVER_LEARNED = 2 };

typedef struct cfd_ctx cfd_ctx;

/* Grid, boundary conditions and the assembled pressure-system description.
   Mirrors Grid (grid.py:26-152), DomainBoundary (boundary.py:92-141) and the
   setup half of PressureSolver (pressure.py:32-141). */
typedef struct cfd_grid_desc {
    int32_t nx, ny;
    double dx, dy;
    int32_t periodic;            /* periodic in x (grid.set_periodic_x applied to fluid) */
    const uint8_t* fluid;        /* (nx+2) x (ny+2), grid.fluid after finalize() */
    int32_t bc_kind[4];          /* left, right, bottom, top: CFD_BC_* */
    double bc_velocity[4];       /* wall tangential velocity (noslip) */
    const double* inflow[4];     /* signed inflow normal velocity per boundary face, or NULL */
    const uint8_t* open[4];      /* boundary face open (adjacent inner cell is fluid) */
    const uint8_t* pcode;        /* nx x ny: index 1..63 into diag_table (0 = not fluid)
                                    | 64 if the cell has a Dirichlet x-face
                                    | 128 if it has a Dirichlet y-face */
    const uint8_t* pcomp;        /* nx x ny: pure-Neumann component 1..ncomp, 0 = none */
    int32_t ncomp;               /* <= 8 */
    const double* comp_count;    /* cells per component */
    const double* diag_table;    /* 64 diagonal entries of A = -Laplacian */
    const double* inv_table;     /* 64 reciprocals (Jacobi preconditioner, pressure.py:139) */
} cfd_grid_desc;

/* One transported cell field (models.py:73-136 ScalarField.attach). */
typedef struct cfd_scalar_desc {
    double diffusivity;          /* D0: constant diffusivity field */
    int32_t limiter;             /* CFD_LIM_* */
    int32_t obstacle_fixed;      /* obstacle BC: 0 zero-gradient, 1 fixed value */
    double obstacle_value;
    const int64_t* bt[4];        /* per boundary face: 0 zero-gradient, 1 fixed, 2 flux */
    const double* bv[4];
    int32_t wall_flux;           /* accumulate wall fluxes for Nusselt monitors */
    int32_t boussinesq;          /* add the buoyancy source of this field (models.py:386-396) */
    double gx, gy, t_ref;
} cfd_scalar_desc;

/* Time-step settings (solver.py:95-173, 209-253, 448-492). */
typedef struct cfd_step_params {
    double nu, tau;
    int32_t scheme_ab2;          /* 1 AB2, 0 explicit Euler */
    int32_t gamma_auto;
    double gamma_fixed;
    double dt_fixed;             /* > 0: fixed dt */
    double dt_model;             /* min of the models' dt_limit(), or >= 1e300 */
    double t_end;                /* run(): shorten the last step to end at t_end */
    int32_t clamp_t_end;
    int32_t solver;              /* CFD_SOLVER_* */
    double fx, fy;               /* body force */
    double rtol;                 /* pressure solve relative tolerance */
    int32_t maxiter;
    int32_t check_t_end;         /* run(): stop once t >= t_end - 1e-12 */
    double reliable_delta;       /* mixed solver: residual-replacement trigger factor */
    double steady_tol;           /* run(): stop once n > 10 and residual < steady_tol (0: off) */
    int64_t max_steps;           /* run(): stop at this step count (0: off) */
    int32_t flow_control;        /* 1: hold the bulk velocity at flow_rate (solver.py:405-425) */
    int32_t pad_flow;
    double flow_rate;
} cfd_step_params;

/* Scalars of a running simulation (Simulation attributes t, n, dt, dt_old,
   flow_force and whether AB2 history exists). */
typedef struct cfd_state {
    double t, dt, dt_old, gamma, flow_force;
    int64_t n;
    int32_t has_old;
    int32_t done;                /* a run reached t_end / max_steps / steady_tol */
    int32_t error;               /* instability detected at step n (solver.py:395-399) */
    int32_t pad;
} cfd_state;

/* One history record (solver.py:427-443) plus solver statistics.  mon[] holds
   raw sums over the grid: [0] max |div u|, [1] sum of uc^2 + vc^2 over fluid
   cells, [2..5] sums of p over fluid cells with a solid east / west / north /
   south neighbour, [6..7] uc summed over the north / south ones, [8..9] vc
   over the east / west ones (obstacle forces, solver.py:542-567), then per
   scalar s (5 values from index 10+5s, s < 4): sum of phi over fluid cells and
   the wall-flux sums of the left, right, bottom, top sides (models.py:228-250);
   [30] the largest nu_t of the step (turbulence models, else 0), [31] the
   driving force after the step (flow_force), [32..34] VOF: the sum of alpha
   over fluid cells, the last cell of the bottom row and of the left column
   with alpha > 0.5 (1-based, -1 for none). */
typedef struct cfd_record {
    double t, dt, gamma, residual, divergence, kinetic_energy;
    double cg_relres;
    double mon[36];
    int64_t n;
    int32_t cg_iters, cg_info, cg_outer, nmon;
} cfd_record;

const char* cfd_last_error(void);
int cfd_device_count(int32_t* count);
/* End of use (no counterpart in the reference, which holds no device
   resources): waits for the device.  Callers destroy their contexts,
   ensembles and networks first and make no further call afterwards; the
   Python binding does both from an atexit handler, before the process's exit
   handlers tear the HIP runtime down. */
int cfd_shutdown(void);

/* Simulation.__init__ device half (solver.py:110-147): allocate and upload. */
int cfd_ctx_create(const cfd_grid_desc* desc, cfd_ctx** out);
int cfd_ctx_destroy(cfd_ctx* ctx);
/* Run on a caller-owned hipStream_t (NULL restores the context's own stream). */
int cfd_ctx_set_stream(cfd_ctx* ctx, void* stream);
/* Add a transported field (PassiveScalar / BoussinesqTemperature attach,
   models.py:102-136, 296-333, 339-384); *id receives its index. */
int cfd_add_scalar(cfd_ctx* ctx, const cfd_scalar_desc* desc, int32_t* id);

/* Host <-> device copies in the reference layout (checkpoint save/load,
   functions.py:326-379; probes and reports read fields this way). */
int cfd_upload(cfd_ctx* ctx, int32_t field, const double* host);
int cfd_download(cfd_ctx* ctx, int32_t field, double* host);
int cfd_upload_scalar(cfd_ctx* ctx, int32_t id, const double* host);   /* + fill_ghosts */
int cfd_download_scalar(cfd_ctx* ctx, int32_t id, double* host);
int cfd_set_state(cfd_ctx* ctx, const cfd_state* st);
int cfd_get_state(cfd_ctx* ctx, cfd_state* st);

/* Multigrid hierarchy for CFD_SOLVER_MGCG (no reference counterpart; the
   reference's solver choice is pressure.py:136-139).  Level 0 is the context's
   grid (only nx, ny given); level l >= 1 holds the 5-point Galerkin operator
   of 2x2 aggregation as nx x ny host arrays (diagonal, coupling magnitudes to
   the east and north neighbours, 0 where no fluid).  The last level is solved
   with the dense inverse cinv (ncoarse x ncoarse over its cells, row-major). */
typedef struct cfd_mg_level_desc {
    int32_t nx, ny;
    const float* diag;
    const float* ce;
    const float* cn;
} cfd_mg_level_desc;
typedef struct cfd_mg_desc {
    int32_t nlev;                /* 2..16 */
    float omega;                 /* damped-Jacobi weight */
    float wc;                    /* coarse-correction weight */
    int32_t ncoarse;             /* nx*ny of the last level, <= 4096 */
    const float* cinv;
    cfd_mg_level_desc lev[16];
} cfd_mg_desc;
int cfd_set_multigrid(cfd_ctx* ctx, const cfd_mg_desc* desc);

/* Simulation._apply_bc (solver.py:190-193): DomainBoundary.apply + obstacle_bc. */
int cfd_apply_bc(cfd_ctx* ctx);
/* Simulation._project_initial (solver.py:195-204). */
int cfd_project_initial(cfd_ctx* ctx, const cfd_step_params* prm);
/* Simulation.step (solver.py:326-403): enqueue `count` steps, no host sync. */
int cfd_step(cfd_ctx* ctx, const cfd_step_params* prm, int32_t count);
/* Wait for enqueued work.  Returns CFD_ERR_STATE when a device wait gave up
   (cfd_set_spin_limit), as does every call that reads the device state. */
int cfd_sync(cfd_ctx* ctx);
/* Bound of the device's waits on other work-groups (no counterpart in the
   reference, which has no device): the persistent pressure kernels' grid
   barriers and the one-pass reductions' gathers.  A wait longer than
   `seconds` (default 2) gives up, every kernel still ends, and the context
   reports CFD_ERR_STATE from then on (its fields are invalid); calling this
   again clears that error.  seconds <= 0 makes every wait give up at once
   (a test of the error path that hangs nothing). */
int cfd_set_spin_limit(cfd_ctx* ctx, double seconds);
/* Copy the records of steps first_n .. first_n+count-1 (1-based step numbers,
   at most the ring length back from the current step). */
int cfd_read_records(cfd_ctx* ctx, int64_t first_n, int32_t count, cfd_record* out);
/* Time of the last cfd_step batch on the device, in ms (HIP events). */
int cfd_last_elapsed_ms(cfd_ctx* ctx, double* ms);
/* Time every pressure-projection launch with HIP events on the context's
   stream (measurement aid, no reference counterpart): switch on, run steps,
   then read the summed kernel time and launch count (which resets them). */
int cfd_set_profiling(cfd_ctx* ctx, int32_t on);
int cfd_projection_times(cfd_ctx* ctx, double* total_ms, int64_t* launches);

/* PressureSolver.solve (pressure.py:143-185) on a host right-hand side
   (nx x ny) -> host pressure (nx x ny); warm start from the context's X. */
int cfd_pressure_solve(cfd_ctx* ctx, const double* rhs, double* p_out, int32_t solver,
                       double rtol, int32_t maxiter, int32_t* iters, double* relres);

/* ---- device memory for callers of the stateless kernels (no counterpart in
   the reference, which has no device; these replace framework tensors, since a
   second HIP runtime in the same process cannot open the GPU) ---- */
int cfd_dev_alloc(int64_t bytes, void** out);
int cfd_dev_free(void* ptr);
/* kind: 1 host->device, 2 device->host, 3 device->device; synchronous */
int cfd_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind);

/* ---- stateless kernels on caller device arrays in the reference layout ---- */
/* kernels.momentum_tendencies (kernels.py:44-155), constant viscosity nu0 */
int cfd_momentum_tendencies(const double* u, const double* v, const uint8_t* fluid, int32_t nx,
                            int32_t ny, double dx, double dy, double nu0, double gamma,
                            int32_t periodic, double* hu, double* hv, void* stream);
/* kernels.obstacle_bc (kernels.py:158-209) */
int cfd_obstacle_bc(double* u, double* v, const uint8_t* fluid, int32_t nx, int32_t ny,
                    int32_t periodic, void* stream);
/* kernels.scalar_transport (kernels.py:254-391); diff may be NULL for a
   constant diffusivity d0 */
int cfd_scalar_transport(const double* phi, const double* u, const double* v,
                         const uint8_t* fluid, int32_t nx, int32_t ny, double dx, double dy,
                         const double* diff, double d0, int32_t limiter, int32_t periodic,
                         const int64_t* bt_l, const double* bv_l, const int64_t* bt_r,
                         const double* bv_r, const int64_t* bt_b, const double* bv_b,
                         const int64_t* bt_t, const double* bv_t, int32_t obs_t, double obs_v,
                         double* out, void* stream);

/* ---- learned flexible CG (pressure solver CFD_SOLVER_LEARNED; no counterpart
   in the reference, whose solvers are splu and Jacobi-PCG, pressure.py:136-139).
   FP64 flexible CG, ending on the FP64 true residual, preconditioned by an FP32
   V-cycle over the multigrid hierarchy (cfd_set_multigrid) that stops at
   `level`, where the U-Net supplies the coarse correction after an exact
   correction in the constant vector.  Modes: 0 network, 1 no network (the
   constant-vector correction alone), 3 V-cycle (no network: the V-cycle goes on
   below `level` to the dense coarsest level, the complete multigrid
   preconditioner), 2 teacher (level solved by
   `teacher_cycles` iterations of flexible CG preconditioned by V-cycles down
   to the dense coarsest level; with harvest
   buffers, each iteration stores the pair (network input, normalised
   correction) as two H x W float images, H and W the level's sizes rounded up
   to multiples of 8).  A solve that does not converge within max_iter
   iterations is continued by the FP64 Jacobi-PCG when `fallback` is set.  A
   network that has not reduced ||r|| tenfold after 8 iterations of a solve is
   switched off for the context (the solver then runs as in mode 1).
   The context keeps `net` (and the harvest buffers) until cfd_set_learned is
   called again or the context is destroyed: the caller keeps them alive. ---- */
/* Initial iterate of the pressure solves of time steps, for the
   constant-density solvers (no counterpart in the reference, whose CG starts
   from the last solution, pressure.py:140, 171-177): mode 0 the last
   solution, mode 1 the linear extrapolation in time
   x + (dt/dt_old)(x - x_prev) of the last two solutions (the first solve
   after the initial projection, an upload of X or a change of mode starts
   from the last solution).  cfd_learned_desc.x0_extrapolate sets the same
   mode.  cfd_initial_iterate reports the mode and whether x_prev holds the
   solution before the last one. */
int cfd_set_initial_iterate(cfd_ctx* ctx, int32_t mode);
int cfd_initial_iterate(cfd_ctx* ctx, int32_t* mode, int32_t* has_history);

typedef struct cfd_unet cfd_unet;
typedef struct cfd_learned_desc {
    cfd_unet* net;
    int32_t level;
    int32_t mode;
    int32_t teacher_cycles;
    int32_t max_iter;
    int32_t fallback;
    int32_t use_graph;        /* replay each iteration as a captured HIP graph */
    float sigma;              /* network output scale */
    double one_a_one;         /* 1^T A_level 1 (0: no constant-vector correction) */
    float* harvest_in;        /* device, harvest_cap x H x W floats (mode 2) */
    float* harvest_out;
    int64_t harvest_cap;
    int32_t coarse_fused;     /* bit 0: levels with <= 64 x 64 cells, the network and the
                                 level-K correction in one work-group launch (modes 0, 1
                                 on a 32 x 32 level K; mode 3: those levels down to the
                                 dense solve; ignored otherwise); bit 1 (with bit 0): the
                                 levels above that coarse end in one launch per level
                                 and leg (the same FP32 arithmetic; the partial dot
                                 products of the fine up-leg in tile order); bit 2 (with
                                 bit 1): the fine level's up-leg still as two launches
                                 (row-band partials: bit-identical to bit 1 clear); bit 3
                                 (with bit 1, bit 2 clear, mode 3): the two levels just
                                 above the coarse end in one launch per leg (closed
                                 domains whose sizes halve exactly; bit-identical) */
    int32_t x0_extrapolate;   /* 1: start each solve of a time step from the linear
                                 extrapolation x + (dt/dt_old)(x - x_previous) of the last
                                 two solutions; 0: from the last solution (SciPy's x0) */
    int32_t sweeps;           /* damped-Jacobi sweeps per leg on the levels above the fused
                                 coarse end: 0 or 2 (the multigrid solver's V-cycle), 3 or
                                 4 (a stronger preconditioner; needs coarse_fused bits 0
                                 and 1 without bit 2), -1: 3 for a solve after one of at
                                 least 4 iterations, 2 again after 8 consecutive 3-sweep
                                 solves of at most 2 (2 throughout without those fused
                                 legs) */
} cfd_learned_desc;
typedef struct cfd_learned_stats_t {
    int64_t harvested, solves, iterations, fallbacks;
    int32_t last_iters, last_info;
    double last_relres;
    int32_t net_off;           /* 1: the network failed its check and is switched off */
    int32_t last_sweeps;       /* smoothing sweeps per leg of the last solve */
} cfd_learned_stats_t;
int cfd_set_learned(cfd_ctx* ctx, const cfd_learned_desc* desc);
int cfd_learned_stats(cfd_ctx* ctx, cfd_learned_stats_t* out);
/* zero the harvest counter and the solve statistics */
int cfd_learned_reset(cfd_ctx* ctx);

/* ---- function objects on the device (cfd2d/functions.py:59-240 evaluates
   Probes and FieldAverage on the host after every step,
   solver.py:402-403; here they run in the step's launch sequence, so a
   multi-step cfd_step needs no host synchronisation) ---- */
enum {
    CFD_PROBE_U = 0,          /* face u, grid (i dx, (j - 1/2) dy), host layout (nx+1, ny+2) */
    CFD_PROBE_V = 1,          /* face v, grid ((i - 1/2) dx, j dy), host layout (nx+2, ny+1) */
    CFD_PROBE_PRESSURE = 2,   /* cell fields from here on: cell centres, host layout (nx, ny), */
    CFD_PROBE_SPEED = 3,      /*   NaN in solid cells (solver.py:526-537, 569-578) */
    CFD_PROBE_VORTICITY = 4,
    CFD_PROBE_SCALAR = 5      /* + scalar id */
};
typedef struct cfd_probe_desc {
    int32_t nsample;          /* (field, point) pairs; 0 removes the probes */
    int32_t every;            /* record after the steps n with n % every == 0 */
    int32_t capacity;         /* records held until cfd_probes_read drains them */
    int32_t pad;
    const int32_t* src;       /* per sample: CFD_PROBE_* */
    const int32_t* corner;    /* per sample (i, j): lower corner of the interpolation cell in
                                 the field's host layout, 0 <= i <= rows-2, 0 <= j <= cols-2 */
    const double* frac;       /* per sample (a, b): normalised distances along x and y; value =
                                 f00 (1-a)(1-b) + f01 (1-a) b + f10 a (1-b) + f11 a b, evaluated
                                 left to right (scipy's linear RegularGridInterpolator) */
} cfd_probe_desc;
int cfd_probes_set(cfd_ctx* ctx, const cfd_probe_desc* desc);
/* the held records (times [n], values [n][nsample]) and the count dropped for
   lack of room; the device buffer is emptied */
int cfd_probes_read(cfd_ctx* ctx, int32_t max_records, double* times, double* values,
                    int32_t* nrecords, int64_t* lost);
/* FieldAverage (functions.py:180-210): from the first step with t >= t_start,
   sum += a dt and sum2 += a^2 dt of the cell-centred u, v and pressure (NaN in
   solid cells); enable = 0 stops it; enabling clears the sums */
int cfd_favg_set(cfd_ctx* ctx, int32_t enable, double t_start);
/* T and the sums [3][nx][ny] (u, v, p; either pointer may be NULL) */
int cfd_favg_read(cfd_ctx* ctx, double* T, double* sum, double* sum2);
/* Node fields for reports, computed on the device from the current velocity
   and copied to host (nx+1) x (ny+1): the vorticity dv/dx - du/dy
   (solver.py:569-573) or the streamfunction with psi(0, 0) = 0 and no flux
   through faces without a fluid cell beside them, the sums running as numpy's
   cumsum does (solver.py:580-596) */
enum { CFD_DERIVED_VORTICITY = 0, CFD_DERIVED_STREAMFUNCTION = 1 };
int cfd_download_derived(cfd_ctx* ctx, int32_t kind, double* host);
/* Allocation-free variant of cfd_download_derived.  It has no counterpart in
   the reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid it).
   The node field is staged in a static device array of
   CFD_DERIVED_NOALLOC_MAX doubles on each GPU (about 34 MB, held until the
   process ends) instead of a buffer allocated on the first call; grids with
   more nodes get CFD_ERR_LIMIT.  All calls on one GPU share that array:
   calls that overlap on one GPU race. */
#define CFD_DERIVED_NOALLOC_MAX (2049L * 2049L)
int cfd_download_derived_noalloc(cfd_ctx* ctx, int32_t kind, double* host);

/* ---- turbulence models (turbulence.py:130-325); no counterpart in a C
   interface of the reference, which runs them in NumPy ----
   Smagorinsky: nu_t = l2 |S| on fluid cells.  k-omega SST (Menter 2003, low
   Re): k and omega are two scalars added with cfd_add_scalar (diffusivity nu,
   the reference's boundary conditions, upwind or another limiter); the step
   then transports them with the SST diffusivities and sources instead of as
   passive scalars.  Arrays are nx x ny host layout.  Call once, after the
   scalars are added and uploaded; it evaluates nu_t for the current fields. */
enum { CFD_TURB_SMAGORINSKY = 1, CFD_TURB_SST = 2 };
typedef struct cfd_turb_desc {
    int32_t model;
    int32_t sid_k, sid_omega;        /* SST */
    int32_t pad;
    double nu;                       /* molecular viscosity */
    const double* l2;                /* Smagorinsky: min(C_s sqrt(dx dy), 0.41 d)^2 */
    const double* wall_distance;     /* SST: d >= 1e-12 (turbulence.py:42-69, 237-238) */
    const uint8_t* wall_adjacent;    /* SST: fluid cells next to a wall */
    const double* omega_wall;        /* SST: 6 nu / (beta1 y^2) there (turbulence.py:239-243) */
} cfd_turb_desc;
int cfd_set_turbulence(cfd_ctx* ctx, const cfd_turb_desc* desc);
/* Two-phase VOF (vof.py:136-306, cases.py:294-320); no counterpart in a C
   interface of the reference.  alpha0 is nx x ny (clipped to [0, 1], 0 off
   the fluid); the pressure system becomes div((1/rho) grad p) = div(u*)/dt,
   solved by CG to 1e-12 (the reference solves it directly) with
   the pure-Neumann components' pinned cells (i, j) from the pressure set-up;
   the CG is preconditioned by an aggregation multigrid rebuilt every solve.
   Gravity is the caller's body force; sigma > 0 adds CSF surface tension. */
typedef struct cfd_vof_desc {
    double rho1, rho2, mu1, mu2, sigma, c_alpha, cfl_alpha, alpha_inflow;
    const double* alpha0;
    int32_t npinned;
    int32_t jacobi_only;             /* 1: Jacobi-PCG instead of the multigrid-preconditioned CG */
    const int32_t* pinned;
} cfd_vof_desc;
int cfd_set_vof(cfd_ctx* ctx, const cfd_vof_desc* desc);
/* kernels.momentum_tendencies with variable = True (kernels.py:96-103,
   131-138): stress-form diffusion of the cell viscosity field nu
   ((nx+2) x (ny+2), nu_eff) */
int cfd_momentum_tendencies_var(const double* u, const double* v, const uint8_t* fluid,
                                int32_t nx, int32_t ny, double dx, double dy, const double* nu,
                                double gamma, int32_t periodic, double* hu, double* hv,
                                void* stream);

/* ---- learned preconditioner network (no counterpart in the reference: the
   U-Net of the learned flexible PCG, trained in-process on residual ->
   correction pairs harvested from the pressure solves) ----
   4 levels, 3x3 convolutions, channels 16-32-64-64, ReLU, 2x2 average pooling,
   nearest upsampling with skip concatenation; FP32 master parameters (flat
   vector, layer by layer: weights [tap][Cout][Cin], then biases), BF16 weights
   and activations on the matrix cores with FP32 accumulation.  Images are
   H x W fp32 on the device, H and W multiples of 8. */
int64_t cfd_unet_param_count(void);
int cfd_unet_create(int32_t H, int32_t W, cfd_unet** out);
int cfd_unet_destroy(cfd_unet* net);
/* host <-> device parameter copies; set_params also resets the Adam moments */
int cfd_unet_set_params(cfd_unet* net, const float* host);
int cfd_unet_get_params(cfd_unet* net, float* host);
int cfd_unet_get_grad(cfd_unet* net, float* host);
int cfd_unet_zero_grad(cfd_unet* net);
/* forward passes treat the H x W image as a batch of H / image_rows images
   stacked along H (each zero-padded at its own borders); training is for
   single images */
int cfd_unet_set_batch(cfd_unet* net, int32_t image_rows);
/* y = net(x) */
int cfd_unet_forward(cfd_unet* net, const float* x_dev, float* y_dev);
/* loss = sum mask (net(x) - t)^2 / tnorm2 (mask may be NULL); the gradient of
   the loss is added to the network's gradient vector */
int cfd_unet_loss_grad(cfd_unet* net, const float* x_dev, const float* t_dev,
                       const float* mask_dev, double tnorm2, double* loss);
/* loss of cfd_unet_loss_grad: kind 0 the masked squared error above; kind 1
   the energy norm (y - t).A(y - t) / tnorm2 on the nx x ny corner of the
   image, A = 5-point Laplacian with zero-flux sides and unit couplings + eps I
   (mask unused) */
int cfd_unet_set_loss(cfd_unet* net, int32_t kind, int32_t nx, int32_t ny, float eps);
/* one Adam step (beta 0.9/0.999) on the gradient times `scale`, then zero it */
int cfd_unet_adam(cfd_unet* net, float lr, float scale);
/* D = A B for 16 x 16 row-major matrices, through one bf16 matrix-core
   instruction (checks the operand lane layout the convolutions rely on) */
int cfd_mfma_layout_check(const float* A, const float* B, float* D);
/* D = A B for A 16 x 32 and B 32 x 16 (row-major) through one
   v_mfma_f32_16x16x32_bf16 fed with two 16-deep K steps in the lane layout
   above (checks the pairing the convolutions rely on).  No counterpart in the
   reference, which has no matrix-core code. */
int cfd_mfma_layout_check_k32(const float* A, const float* B, float* D);

/* ---- batched ensembles (no counterpart in the reference, which runs one case
   per process): many lid-driven / heated cavities on the unit square
   (cases.py:60-70, 207-223; AB2, gamma auto, no-slip walls) advanced in
   lockstep.  Instances of one side length n (a power of two >= 64) form a
   class; every kernel of a step serves all instances of a class at once.
   Bulk in FP32 (velocities, temperature, pressure, V-cycle), the pressure
   solve an FP64 flexible CG to a true relative residual of rtol,
   preconditioned by V-cycles down to each instance's 32 x 32 level where one
   batched pass of the shared U-Net (cfd_ens_set_net) gives the coarse
   corrections.  Per class: nu (1/Re or sqrt(Pr/Ra)), lid velocity (1 or 0),
   thermal diffusivity kappa (0 for a lid cavity, 1/sqrt(Ra Pr) for a heated
   one: T = 0.5 / -0.5 on the left / right walls, buoyancy (0, -T)); the
   pressure multigrid hierarchy of the n x n pinned Neumann cavity (levels >= 1
   as n_l x n_l float arrays, the coarsest one's dense inverse). ---- */
#define CFD_ENS_MAX_CLASSES 8
#define CFD_ENS_MAX_LEVELS 12
typedef struct cfd_ens cfd_ens;
typedef struct cfd_ens_class_desc {
    int32_t n, B, nlev;
    const double* nu;          /* B values each */
    const double* lid;
    const double* kappa;
    int32_t lev_n[CFD_ENS_MAX_LEVELS];
    const float* diag[CFD_ENS_MAX_LEVELS];
    const float* ce[CFD_ENS_MAX_LEVELS];
    const float* cn[CFD_ENS_MAX_LEVELS];
    const float* cinv;
    double one_a_one;          /* 1^T A_K 1 of the 32 x 32 level */
} cfd_ens_class_desc;
typedef struct cfd_ens_desc {
    int32_t nclass;
    cfd_ens_class_desc cls[CFD_ENS_MAX_CLASSES];
    double tau, rtol;
    int32_t maxiter;
    float omega, wc;           /* damped-Jacobi weight, coarse-correction weight */
    int32_t autotune;          /* 1: time work-group shapes per class (cfd_ens_tuning) */
    int32_t use_graph;         /* 1: replay each solver iteration as a captured HIP graph */
    int32_t x0_extrapolate;    /* 1: each solve starts from the linear extrapolation in time
                                  of the instance's last two solutions */
    int32_t pad;
} cfd_ens_desc;
typedef struct cfd_ens_stats_t {
    double t, dt, gamma, kinetic_energy, divergence, residual, relres;
    int64_t steps, sweeps_total;
    int32_t iters, info, pred_tx, pred_ty;
    float pred_us;
} cfd_ens_stats_t;
int cfd_ens_create(const cfd_ens_desc* desc, cfd_ens** out);
int cfd_ens_destroy(cfd_ens* ens);
/* net: a batch of (total instances) images of 32 x 32 (cfd_unet_set_batch(net,
   32)), or NULL for the constant-vector coarse correction alone */
int cfd_ens_set_net(cfd_ens* ens, cfd_unet* net, float sigma);
/* Coarse end of every instance's V-cycle: 0 the network (or the
   constant-vector correction, cfd_ens_set_net), 1 the V-cycle carried on to
   the dense coarsest solve, the levels of at most 64 x 64 cells held in LDS
   by one work-group per instance (the complete multigrid preconditioner) */
int cfd_ens_set_coarse(cfd_ens* ens, int32_t mode);
/* `count` steps of every instance.  With use_graph each step is one graph
   launch whose solver iterations (a cap adapted to the previous steps' count)
   are controlled on the device; a step whose solves need more continues in
   the next launch, and the call returns once the device has completed
   `count` steps.  Without it, the host polls the solves' progress. */
int cfd_ens_step(cfd_ens* ens, int32_t count);
/* field 0 u, 1 v, 2 p, 3 T of one instance, reference layout, float */
int cfd_ens_download(cfd_ens* ens, int32_t cls, int32_t inst, int32_t field, float* host);
int cfd_ens_stats(cfd_ens* ens, int32_t cls, int32_t inst, cfd_ens_stats_t* out);
/* What the autotuner chose for class `cls` (desc.autotune; the defaults
   otherwise): the predictor's work-group shape; per level the work-group
   shape (sweep_tx columns x sweep_ty rows) of the pointwise V-cycle kernels;
   the cells per thread and batch of the FP64 vector kernels; the time per
   launch of each choice and of the default (256 x 1, 8 cells).  The choices
   change no computed value.  No counterpart in the reference, which has no
   ensembles: it reports the tile shapes the autotuner picked. */
typedef struct cfd_ens_tuning_t {
    int32_t pred_tx, pred_ty, nlev, vec_batch;
    int32_t sweep_tx[CFD_ENS_MAX_LEVELS], sweep_ty[CFD_ENS_MAX_LEVELS];
    float pred_us, vec_us, vec_default_us, pad;
    float sweep_us[CFD_ENS_MAX_LEVELS], sweep_default_us[CFD_ENS_MAX_LEVELS];
} cfd_ens_tuning_t;
int cfd_ens_tuning(cfd_ens* ens, int32_t cls, cfd_ens_tuning_t* out);

#ifdef __cplusplus
}
#endif
#endif
