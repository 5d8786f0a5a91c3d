/*
 * qesmc.h -- C ABI of the MI355X mass-consistency solver (libqesmc.so).
 *
 * The library replaces the body of QES-Winds' Solver::solve for one time step:
 *   R = divergence of u0 (src/winds/Solver_CPU_RB.cpp:56-80),
 *   red-black SOR for lambda (Solver_CPU_RB.cpp:95-184),
 *   u = u0 + grad(lambda)/(2 alpha^2) and solid zeroing (Solver_CPU_RB.cpp:194-284),
 * on one AMD Instinct MI355X (gfx950).  Plain C types only; host pointers are
 * borrowed for the duration of a call.  Every entry point returns QESMC_OK (0)
 * or a negative status; qesmc_status_string() names it.  A context is not
 * thread-safe: use one per host thread.
 *
 * Layouts are the reference's (src/qes/Domain.h:143-175): nx = ncx+1,
 * ny = ncy+1, nz = ncz+2; cell arrays (nx-1)*(ny-1)*(nz-1) and face arrays
 * nx*ny*nz, x fastest; dz_array has nz-1 entries with [0] the ground ghost.
 */
#ifndef QESMC_H
#define QESMC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QESMC_ABI_VERSION 2

/* ---- status codes ---------------------------------------------------------
 * The reference exits the process on every error (Solver_GPU_GlobalMemory.cu:
 * 58-65 _cudaCheck, SolverFactory.cpp:57-58 QESout::error); the library
 * returns these codes instead and the C++ adapter maps them back to exits. */
#define QESMC_OK 0
#define QESMC_ERR_ARG (-1)          /* null pointer, bad size, dx != dy ...     */
#define QESMC_ERR_NO_DEVICE (-2)    /* no HIP device / bad ordinal              */
#define QESMC_ERR_OOM (-3)          /* device allocation failed                 */
#define QESMC_ERR_HIP (-4)          /* other HIP runtime error                  */
#define QESMC_ERR_ZERO_DIAG (-5)    /* an updated cell has e+f+g+h+m+n == 0     */
#define QESMC_ERR_NONFINITE (-6)    /* lambda became NaN/Inf during the solve   */
#define QESMC_ERR_STATE (-7)        /* resident call before its inputs were set */
#define QESMC_ERR_UNSUPPORTED (-8)  /* option not available for these inputs   */
#define QESMC_ERR_IO (-9)           /* file missing, unwritable or not a NetCDF
                                       classic (CDF-1/CDF-2) file               */
#define QESMC_ERR_LIMIT (-10)       /* beyond the fixed capacity of an entry
                                       point that allocates nothing             */

/* ---- grid / input / output descriptors ---------------------------------- */

/* qes::Domain sizes (Domain.cpp:7-31).  dy is accepted for completeness; the
 * reference's coefficients are normalised with dx in both horizontal
 * directions (Domain.h:104 getDomainSize quirk) and its R/correction use
 * domain.dy(); the library uses dx and dy exactly where the reference does. */
typedef struct qesmc_grid {
  int nx, ny, nz;          /* staggered sizes                                */
  float dx, dy;            /* horizontal spacing [m]                         */
  const float *dz_array;   /* nz-1 vertical spacings [m], [0] = ghost        */
} qesmc_grid;

/* WINDSGeneralData::e,f,g,h,m,n and icellflag (WINDSGeneralData.h:231-239). */
typedef struct qesmc_coeffs {
  const float *e, *f, *g, *h, *m, *n; /* cell arrays, already normalised     */
  const int *icellflag;               /* cell array: 0 building, 1 fluid, 2 terrain */
} qesmc_coeffs;

/* WINDSGeneralData::u0,v0,w0 (inputs) and u,v,w (outputs), face arrays. */
typedef struct qesmc_fields {
  const float *u0, *v0, *w0;
  float *u, *v, *w;
} qesmc_fields;

/* Kernel path selection (qesmc_params.path). */
#define QESMC_PATH_AUTO 0     /* fastest available: FUSED where the colour
                                 arrays exceed the 256 MB Infinity Cache and
                                 the run allows it (wall mask, below 2^31
                                 half-cells, its
                                 buffers already allocated in a
                                 qesmc_run_static call), SPLIT otherwise        */
#define QESMC_PATH_PLAIN 1    /* in-place red/black launches on the x-fastest
                                 layout, full fp32 coefficients: the classic
                                 GPU red-black SOR (baseline)                   */
#define QESMC_PATH_SPLIT 2    /* colour-split (checkerboard) layout, per-cell
                                 6-bit wall mask when the coefficients are
                                 stair-step, full coefficients otherwise        */
#define QESMC_PATH_FUSED 3    /* colour-split layout and wall mask, both colour
                                 passes of an iteration in one launch (two
                                 lambda buffers); stair-step coefficients only,
                                 else QESMC_ERR_UNSUPPORTED                     */

typedef struct qesmc_params {
  float omega;      /* over-relaxation, reference 1.78f (Solver.h:83)          */
  float tol;        /* stop when max|dlambda| <= tol (Solver_CPU_RB.cpp:95)    */
  int itermax;      /* iteration cap (SimulationParameters maxIterations)      */
  int alpha1;       /* Gaussian precision moduli, reference 1 (Solver.cpp:79)  */
  int alpha2;
  int path;         /* QESMC_PATH_*                                            */
  int check_every;  /* host polls convergence every this many iterations (>=1);
                       results never depend on it                             */
  int flags;        /* QESMC_RUN_* bits                                        */
  int tile_k;       /* k planes marched by one SOR wave tile (0 = library's
                       choice)                                                */
  int tile_v;       /* split path: half-cells per lane, 1, 2 or 4 (0 = choice) */
  int wg_per_cu;    /* split path: resident work-groups per CU (0 = choice)    */
  int slab_planes;  /* split path: k-slab of the interleaved red/black schedule
                       (0 = library's choice, -1 = two whole-domain passes)   */
  /* The tile fields and slab_planes only change speed, never results. */
  float stop_extend; /* stop rule.  0: the reference's (iterate while
                        max|dlambda| > tol, Solver_CPU_RB.cpp:95).  > 0: the
                        production rule for a field accuracy: the first
                        iteration t* that the reference's rule would not run
                        fixes the end at t* + ceil(stop_extend * t*); the
                        iterations up to it run whatever max|dlambda| does.
                        Decided by the device gate, like the reference's.   */
  int refine_every; /* 0: off.  K > 0 (split and fused paths): the run iterates on a
                       correction delta of lambda = base + delta, and every K
                       iterations folds delta into base and forms the
                       residual of base in FP64 (as qesmc_refine does), so
                       the fp32 rounding of lambda is not amplified by SOR
                       into a noise floor of the field, and max|dlambda| (the
                       stop rule's measure) keeps falling at omega near 2.
                       In exact arithmetic the iterates are the plain run's.
                       Not in qesmc_run_batch or qesmc_run_static
                       (QESMC_ERR_UNSUPPORTED). */
  double chebyshev_rho; /* 0: every colour pass at omega.  In (0, 1): the
                       Jacobi spectral radius rho_J (e.g.
                       from qesmc_estimate_omega), and omega varies by colour
                       pass as cyclic Chebyshev semi-iteration prescribes
                       (Golub & Varga 1961): pass h = 0, 1, 2, ... of the
                       schedule (red, black, red, ...) uses w_0 = 1,
                       w_1 = 1 / (1 - rho^2 / 2), w_h = 1 / (1 - rho^2 w_h-1 / 4),
                       which falls to Young's optimum from above; omega is then
                       unused.  Same fixed point; fewer iterations than a
                       fixed omega because the early passes are optimal for
                       the polynomial degree reached.  No counterpart in the
                       reference.  Not in qesmc_run_batch
                       (QESMC_ERR_UNSUPPORTED). */
  int chebyshev_start; /* iterations of the schedule run by earlier calls: a
                       solve split over several calls passes the running
                       count so the schedule continues (0: a fresh start). */
} qesmc_params;

/* qesmc_params.flags: rebuild the stencil encoding from the resident
 * coefficients inside qesmc_run (so its cost is part of the run's time). */
#define QESMC_RUN_REENCODE 1

typedef struct qesmc_result {
  int iterations;            /* SOR iterations executed (reference prints it, :192) */
  int path;                  /* path that ran (QESMC_PATH_PLAIN / _SPLIT / _FUSED) */
  int compressed;            /* 1 if the wall-mask stencil was used               */
  int status;                /* same as the return code                           */
  double final_max_dlambda;  /* max|dlambda| of the last iteration                */
  double ms_total;           /* device time of R + SOR + correction (HIP events)  */
  double ms_sor;             /* device time of the SOR loop alone                 */
  double glups;              /* (nx-3)(ny-3)(nz-3) * iterations / ms_sor          */
  int tile_k, tile_v, wg_per_cu; /* tiling that ran                               */
  int reserved;
} qesmc_result;

typedef struct qesmc_ctx qesmc_ctx;

/* ---- lifetime -------------------------------------------------------------- */

/* Create a context on HIP device `device` with its own stream.  Replaces the
 * device selection and per-solve allocation of Solver_GPU_GlobalMemory
 * (Solver_GPU_GlobalMemory.cu:201-268); buffers persist across solves. */
int qesmc_create(qesmc_ctx **out, int device);
int qesmc_destroy(qesmc_ctx *ctx);

/* Reference defaults: omega 1.78f, tol 1e-9f, itermax 500, alpha 1,
 * path AUTO, check_every 64, flags 0, tile fields 0, slab_planes 0,
 * stop_extend 0 (the reference's stop rule), refine_every 0. */
void qesmc_default_params(qesmc_params *p);

const char *qesmc_status_string(int status);

/* ---- one-call drop-in ------------------------------------------------------ */

/* Solver::solve(WINDSGeneralData*, const int& itermax) (Solver.h:112): copies
 * inputs to the device, runs R + SOR + correction, copies u, v, w back.
 * lambda_inout: NULL for a cold start (Solver::resetLambda, Solver.h:115-120),
 * else the (nx-1)(ny-1)(nz-1) initial lambda, overwritten with the final one
 * (keeps Solver::lambda in step, so the CPU solvers' warm start across time
 * steps in qes/qesMain.cpp:136-150 is preserved). */
int qesmc_solve(qesmc_ctx *ctx, const qesmc_grid *grid, const qesmc_coeffs *coeffs,
                qesmc_fields *fields, float *lambda_inout, const qesmc_params *params,
                qesmc_result *result);

/* ---- resident interface (inputs stay in HBM across solves) ----------------- */

/* Upload grid + coefficients + flags (the WINDSGeneralData constructor's
 * products, WINDSGeneralData.cpp:859-866) and build the stencil encoding. */
int qesmc_set_geometry(qesmc_ctx *ctx, const qesmc_grid *grid, const qesmc_coeffs *coeffs);
/* Upload u0, v0, w0 (applyWindProfile + applyParametrizations products). */
int qesmc_set_inflow(qesmc_ctx *ctx, const float *u0, const float *v0, const float *w0);
/* Initial lambda for the next run; NULL = zeros (resetLambda). */
int qesmc_set_lambda(qesmc_ctx *ctx, const float *lambda);
/* R + SOR + correction on the resident inputs; u, v, w and lambda stay on
 * the device.  The next run starts from this run's lambda unless
 * qesmc_set_lambda is called in between. */
int qesmc_run(qesmc_ctx *ctx, const qesmc_params *params, qesmc_result *result);
int qesmc_get_velocity(qesmc_ctx *ctx, float *u, float *v, float *w);
int qesmc_get_lambda(qesmc_ctx *ctx, float *lambda);

/* qesmc_run without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  qesmc_run allocates
 * its per-iteration convergence rows when a run needs more of them than any
 * earlier run of the context; this variant takes them from a fixed
 * per-device array of the library (16.8 MB, held until the process ends) and
 * calls no allocating routine, so params->itermax is limited to 65536
 * (QESMC_ERR_LIMIT beyond).  The fused path needs its second lambda buffers
 * from an earlier qesmc_run with QESMC_PATH_FUSED (QESMC_ERR_STATE
 * otherwise).  Results are qesmc_run's, bit for bit.  All calls on one
 * device share the array: calls that overlap on one device race. */
int qesmc_run_static(qesmc_ctx *ctx, const qesmc_params *params, qesmc_result *result);
/* Max |div| over interior fluid cells of the corrected field, FP64 on the
 * device: physical = 0 uses the reference operator R(u)/(-2 alpha1^2),
 * physical = 1 the dimensionally consistent one (vertical term / dz_k). */
int qesmc_divergence_linf(qesmc_ctx *ctx, int physical, double *out);

/* ---- batches and tuning --------------------------------------------------- */

/* Solve n resident problems (one context each, same device, any sizes)
 * together: every SOR colour pass of all of them is one kernel launch, each
 * problem keeping its own stop test (the reference solves one domain per
 * Solver::solve; this has no counterpart there).  With use_graph != 0 the
 * iterations are replayed as a captured HIP graph of check_every iterations.
 * Per problem the results equal qesmc_run's; results[i] (optional) receives
 * them (their timings span the whole batch).  ms_loop (optional): device
 * time of the batched iterations.  Split path only.  One tiling serves the
 * whole batch: tile_v 1 or 2 and tile_k as given, else the most work per
 * wave tile (at most 4 planes x 2 half-cells) that leaves >= 6 wave tiles
 * per CU over all members; wg_per_cu as given, else 5 from 48 wave tiles
 * per CU up and qesmc_run's untuned choice below that.  Tilings autotuned
 * per context are not used. */
int qesmc_run_batch(qesmc_ctx **ctxs, int n, const qesmc_params *params, qesmc_result *results,
                    int use_graph, double *ms_loop);

/* Time `iterations` SOR iterations of the resident problem for a set of
 * candidate tilings (tile_k, tile_v, wg_per_cu) and keep the fastest as the
 * context's default for tile fields left at 0.  lambda is restored.  best
 * (optional) receives base with the chosen tile fields. */
int qesmc_autotune(qesmc_ctx *ctx, const qesmc_params *base, int iterations, qesmc_params *best);

/* ---- input builder (flat stair-step cities) ------------------------------ */

/* One RectangularBuilding (RectangularBuilding.h:69-110). */
typedef struct qesmc_building {
  float x_start, y_start, length, width, height, base_height, rotation_deg;
} qesmc_building;

/* One sensor profile (WindProfilerSensorType.cpp:290-333):
 * kind 0 uniform, 1 log law, 2 power law (exponent in z0_or_p, as siteZ0). */
typedef struct qesmc_inflow {
  int kind;
  float U_ref, z_ref, z0_or_p, wind_dir_deg;
} qesmc_inflow;

/* Build the WINDSGeneralData arrays of a flat city: ground ghost flags,
 * stair-step building flags (PolyBuilding.cpp:130-236), wall coefficients
 * (Wall.cpp:44-349, 827-846), single-sensor u0/v0/w0 (WindProfilerSensorType.cpp
 * :571-621) and solid zeroing (Wall.cpp:805-824).  Outputs are caller-owned
 * arrays of the sizes above. */
int qesmc_city_build(int nx, int ny, int nz, float dx, const float *dz_array, int nbuildings,
                     const qesmc_building *buildings, const qesmc_inflow *inflow, float *e,
                     float *f, float *g, float *h, float *m, float *n, int *icellflag, float *u0,
                     float *v0, float *w0);

/* Seeded non-overlapping rectangles: footprints U(lmin,lmax) m, heights
 * U(hmin,hmax) m (or lognormal(median hmin, sigma hmax) capped at hcap when
 * lognormal != 0), one free cell between buildings and `margin` cells from
 * the lateral boundaries.  Returns the number placed (<= nbuildings). */
int qesmc_city_random_buildings(unsigned long long seed, int nx, int ny, float dx, int nbuildings,
                                float lmin, float lmax, float hmin, float hmax, int lognormal,
                                float hcap, int margin, qesmc_building *out);

/* The same city built on the device, straight into the context: replaces the
 * serial set-up of WINDSGeneralData's constructor for this case (flags
 * WINDSGeneralData.cpp:439-444 and PolyBuilding.cpp:130-236, walls and
 * coefficients Wall.cpp:44-349, 827-846, inflow WindProfilerSensorType.cpp
 * :571-621, zeroing Wall.cpp:805-824) followed by qesmc_set_geometry and
 * qesmc_set_inflow.  dy = dx.  The resident arrays are bitwise equal to
 * qesmc_city_build's.  Building fields must be finite (|value| <= 1e8). */
int qesmc_set_city(qesmc_ctx *ctx, int nx, int ny, int nz, float dx, const float *dz_array, int nbuildings,
                   const qesmc_building *buildings, const qesmc_inflow *inflow);

/* Replace u0/v0/w0 of the resident geometry by one sensor profile extruded
 * over the domain with solid zeroing (a new wind direction or speed on the
 * same city: WindProfilerSensorType.cpp:571-621, Wall.cpp:805-824). */
int qesmc_set_city_inflow(qesmc_ctx *ctx, const qesmc_inflow *inflow);

/* Copy the resident inputs back to the host; NULL pointers are skipped. */
int qesmc_get_inputs(qesmc_ctx *ctx, float *e, float *f, float *g, float *h, float *m, float *n,
                     int *icellflag, float *u0, float *v0, float *w0);

/* ---- after the solve: velocity gradient tensor (QES-Turb) --------------- */

/* TURBGeneralData::derivativeVelocity (TURBGeneralData.cpp:479-552; device
 * counterpart TURBGeneralData.cu:37-306) on the resident corrected field of
 * the last qesmc_run / qesmc_solve, without the wall corrections (:553-556,
 * qesmc_velocity_gradient_walls).  G: NULL (result stays on the device) or 9*(nx-1)(ny-1)(nz-1)
 * floats receiving Gxx, Gxy, Gxz, Gyx, Gyy, Gyz, Gzx, Gzy, Gzz (G_ab = d u_a /
 * d x_b) for the fluid cells with i in [1,nx-3], j in [1,ny-3], k in [1,nz-3],
 * 0 elsewhere.  uniform_z != 0 selects the reference's uniform-grid branch
 * (flagUniformZGrid) with spacing dz; otherwise the stretched-grid formulas
 * on the resident dz_array.  ms (optional): device time of the kernel. */
int qesmc_velocity_gradient(qesmc_ctx *ctx, int uniform_z, float dz, float *G, double *ms);

/* The wall step of TURBGeneralData::derivativeVelocity (TURBGeneralData.cpp:
 * 553-556) on the resident gradient of the last qesmc_velocity_gradient call
 * (QESMC_ERR_STATE without one), for stair-step walls: for buildings (cell
 * flag 0) and then terrain (flag 2), the velocity wall treatment of
 * turbParams buildingWallFlag / terrainWallFlag (TURBParams.h:67-89, default
 * 2; TURBWallBuilding.cpp:37-70, TURBWallTerrain.cpp:37-70):
 *   1  log law (set_loglaw_stairstep, TURBWallBuilding.cpp:87-169): the
 *      wall-normal derivatives of the wall-parallel components next to a
 *      wall are 0.5 (q_f + q_f') / (0.5 d log(0.5 d / z0)); z0 is `z0`
 *      (WGD->z0) for buildings and, for terrain, the mean of the four
 *      surrounding entries of z0_domain_u / z0_domain_v (nx*ny floats each,
 *      WGD->z0_domain_u / _v; read only in this case, else may be NULL);
 *   2  one-sided differences into the fluid
 *      (comp_velocity_deriv_finitediff_stairstep, TURBWall.cpp:206-300);
 *   0  none.
 * Same cells, pass order and arithmetic as the reference, including its
 * quirks (the log law's "left wall" pass walks the cells with a wall below;
 * dy is dx).  Cut cells (flags 7/8, the meshTypeFlag == 1 branch) are not
 * built: QESMC_ERR_UNSUPPORTED, gradient unchanged.  G, ms: as in
 * qesmc_velocity_gradient. */
int qesmc_velocity_gradient_walls(qesmc_ctx *ctx, int building_mode, int terrain_mode, float z0,
                                  const float *z0_domain_u, const float *z0_domain_v, float *G, double *ms);

/* Wall::wallLogBC (Wall.cpp:352-797) on the resident corrected field, as
 * qesWindsMain.cpp:157-165 applies it after a solve when logLawFlag == 1
 * (u0 = u; wallLogBC; u = u0): for every fluid cell next to a solid (flag 0
 * or 2) below, above, behind (i > 0), in front, right (j > 0) or left, in
 * that pass order (the lists of Wall.cpp:288-340), the face (i,j,k) gets a
 * zero wall-normal component and the 20-step log-law speed of the next cell
 * out along the wall-parallel direction of that cell.  z0: WGD->z0 (the
 * sensor's site z0, WINDSGeneralData.cpp:396-406); dz is the smallest
 * dz_array entry and dy = dx, as the reference's Domain reports them.  Cut
 * cells (flags 7/8, Wall.cpp:381-655) are not built: QESMC_ERR_UNSUPPORTED
 * and the field is left unchanged.  ms (optional): device time of the six
 * passes. */
int qesmc_wall_log_bc(qesmc_ctx *ctx, float z0, double *ms);

/* Over-relaxation factor for a field accuracy (no counterpart in the
 * reference, which fixes omega = 1.78f, Solver.h:83).  Estimates rho_J, the
 * spectral radius of the Jacobi iteration of the resident problem (fluid
 * cells the sweep relaxes, the k = 0 mirror included), by `steps` Lanczos
 * steps in FP64 on its symmetrised form from a smooth start, and returns
 * omega = 2 - 1.2 (2 - omega_opt) with Young's omega_opt = 2 / (1 +
 * sqrt(1 - rho_J^2)) for red-black SOR: slightly below the optimum, since
 * fp32 rounding noise is amplified as omega nears 2 and past the optimum the
 * slowest modes only oscillate.  Few steps underestimate rho_J (the Ritz
 * value approaches it from below), which errs toward a smaller, safer
 * omega.  rho or omega may be
 * NULL.  The converged solution does not depend on omega; results at the
 * reference's omega (parity) are unaffected. */
int qesmc_estimate_omega(qesmc_ctx *ctx, int steps, double *rho, float *omega);

/* One step of mixed-precision iterative refinement of the resident lambda
 * (no counterpart in the reference, whose fp32 SOR stops where its own
 * rounding noise does): qesmc_run with the resident lambda as the base of
 * refine_every (itself when params->refine_every > 0, else no refresh
 * inside the run).  The residual of the current lambda, r' = R + D lambda -
 * sum_nb c_nb lambda_nb, is formed in FP64 and replaces R; the same
 * red-black SOR (omega, tol, itermax, stop rule, tiling) solves for the
 * correction from zero; lambda += correction (one fp32 rounding); then the
 * velocity correction and zeroing as in qesmc_run.  Near 2, SOR amplifies
 * the fp32 rounding of lambda into a noise floor of the velocity field
 * (about 1e-4 m/s on a 1000 x 1000 x 150 grid at omega 1.99); refinement
 * takes the field below it.  From lambda = 0 the result is qesmc_run's,
 * bit for bit.  Split and fused paths, with the same values on both
 * (QESMC_PATH_PLAIN: QESMC_ERR_UNSUPPORTED).  In res, ms_total includes the
 * residual. */
int qesmc_refine(qesmc_ctx *ctx, const qesmc_params *params, qesmc_result *res);

/* ---- ensembles on one geometry (no counterpart in the reference) ----------
 *
 * The discrete problem is linear in (u0, v0, w0), and a single-sensor inflow
 * from direction theta is cos(theta) x the inflow along x plus sin(theta) x
 * the inflow along y.  Combining converged basis solutions gives an initial
 * lambda for any direction; SOR then runs to the same stop rule. */

/* lambda(ctx) = sum_t coefs[t] * lambda(srcs[t]), 1 <= n <= 8, all contexts on
 * one device with the same grid (a source may be ctx itself). */
int qesmc_lambda_combine(qesmc_ctx *ctx, int n, qesmc_ctx *const *srcs, const float *coefs);

/* *out = max over all faces of |u_a - u_b|, |v_a - v_b|, |w_a - w_b| (FP64;
 * a NaN difference gives +inf).  Contexts on one device, same grid. */
int qesmc_velocity_linf_diff(qesmc_ctx *a, qesmc_ctx *b, double *out);

/* Velocity probes: the resident corrected u, v, w at n face indices
 * (i + j*nx + k*nx*ny, each < nx*ny*nz), gathered on the device, so a field
 * can be compared with sampled reference values without copying it whole.
 * The index list stays in the context for later qesmc_get_probes calls
 * (until the grid changes).  No counterpart in the reference. */
int qesmc_set_probes(qesmc_ctx *ctx, long long n, const long long *faces);
/* u, v, w: n_probes floats each, in the order of the index list. */
int qesmc_get_probes(qesmc_ctx *ctx, float *u, float *v, float *w);

/* ---- QES workspace files (NetCDF classic, 64-bit offset) ------------------ */

/* Write the solver's arrays as a QES-Winds workspace file: the reference's
 * WINDSOutputWorkspace layout (WINDSOutputWorkspace.cpp:40-150) for u, v, w,
 * u0, v0, w0, icellflag, e..n, dz_array and the coordinates, one time
 * record.  The reference writes NetCDF-4; the netCDF library reads this
 * classic format the same way. */
int qesmc_workspace_write(const char *path, const qesmc_grid *grid, const qesmc_coeffs *coeffs,
                          const qesmc_fields *fields);

/* Staggered sizes of a workspace file (x_face, y_face, z_face dimensions,
 * as Domain(const std::string&) reads them, Domain.cpp:59-66). */
int qesmc_workspace_dims(const char *path, int *nx, int *ny, int *nz);

/* Read a workspace file (NetCDF classic or 64-bit offset; a NetCDF-4 file
 * converts with `nccopy -k 64-bit-offset`) into caller buffers of the given
 * staggered sizes, which must match the file; any pointer may be NULL.
 * dx = x[1] - x[0], dy = y[1] - y[0], dz_array from the file or uniform
 * z[1] - z[0] (Domain.cpp:59-109); e..n and icellflag as the coefficient
 * read of WINDSGeneralData.cpp:880-911; record 0 of the time dimension;
 * numeric types other than float / int are converted. */
int qesmc_workspace_read(const char *path, int nx, int ny, int nz, float *dx, float *dy, float *dz_array,
                         float *e, float *f, float *g, float *h, float *m, float *n, int *icellflag, float *u0,
                         float *v0, float *w0, float *u, float *v, float *w);

/* ---- learned warm start: 3-D U-Net predicting lambda0 --------------------
 *
 * No counterpart in the reference (whose solvers start from lambda = 0 or the
 * previous time step, Solver.h:115-120).  A U-Net maps the cell-type mask, R
 * and the inflow profile of a resident problem to an initial lambda; SOR then
 * runs to the same stopping rule, so accuracy is unchanged and only the
 * iteration count changes.  BF16 activations/weights on MFMA with FP32
 * accumulation; FP32 master weights and Adam.  Volumes are the ncx x ncy x ncz
 * physical cells (k = 1..ncz), padded to multiples of 2^(levels-1). */
typedef struct qesmc_unet qesmc_unet;

typedef struct qesmc_unet_config {
  int levels;          /* resolution levels (4 in the shipped network)       */
  int base_channels;   /* channels at level 0, doubled per level (16)        */
  int ncx, ncy, ncz;   /* problem size in cells (fixes the activation sizes) */
  unsigned long long seed; /* weight initialisation                           */
  float target_scale;  /* network output unit: lambda0 = output * target_scale *
                          rms(R); 0 = set from the data at the first training
                          call (RMS of lambda / rms(R) over the samples)     */
} qesmc_unet_config;

/* The output unit in use (see qesmc_unet_config.target_scale). */
int qesmc_unet_target_scale(const qesmc_unet *net, float *scale);

/* Number of input feature channels (mask, R, inflow u, inflow v, height, pad). */
#define QESMC_UNET_IN_CHANNELS 8

int qesmc_unet_create(qesmc_unet **out, int device, const qesmc_unet_config *cfg);
int qesmc_unet_destroy(qesmc_unet *net);
/* FP32 master parameters, layer by layer: for every 3x3x3 conv W[cout][27][cin]
 * (tap = dz*9 + dy*3 + dx) then b[cout]; last the 1x1x1 head W[c0], b[1]. */
int qesmc_unet_num_params(const qesmc_unet *net, long long *count);
int qesmc_unet_get_params(qesmc_unet *net, float *host, long long count);
int qesmc_unet_set_params(qesmc_unet *net, const float *host, long long count);
/* Padded volume dims (x, y, z) of level 0. */
int qesmc_unet_dims(const qesmc_unet *net, int *dx, int *dy, int *dz);

/* Raw network access (tests): input is the padded volume, channels last
 * [dz][dy][dx][QESMC_UNET_IN_CHANNELS] fp32 (stored as BF16 on the device);
 * output [dz][dy][dx] fp32.  backward takes dL/doutput after a forward and
 * returns dL/dparams in the parameter order above. */
int qesmc_unet_forward(qesmc_unet *net, const float *input, float *output);
int qesmc_unet_backward(qesmc_unet *net, const float *grad_output, float *grad_params);

/* Training data from a resident solver context: features of its geometry and
 * inflow, target = its current lambda (run it to convergence first). */
int qesmc_unet_add_sample(qesmc_unet *net, qesmc_ctx *ctx);
int qesmc_unet_num_samples(const qesmc_unet *net, int *count);
int qesmc_unet_clear_samples(qesmc_unet *net);
/* Adam steps on samples [first, last) drawn uniformly (batch of one sample),
 * learning rate lr; mean loss over the steps.  ms: device time. */
int qesmc_unet_train(qesmc_unet *net, int steps, float lr, int first, int last,
                     unsigned long long seed, float *mean_loss, double *ms);
/* Training loss: mean squared error of lambda0 (in target units) plus, when
 * weight > 0, weight x the mean over blocks of block^3 voxels of the squared
 * block-mean error -- smooth errors, which SOR removes slowest, then cost
 * (1 + weight) times more.  block: a power of two dividing the padded
 * volume (multiples of 2^(levels-1)); weight 0 turns the term off (default). */
int qesmc_unet_set_loss(qesmc_unet *net, int block, float weight);
/* Mean loss over samples [first, last) (no update). */
int qesmc_unet_eval(qesmc_unet *net, int first, int last, float *mean_loss);
/* Predict lambda0 for the resident problem of ctx and install it as the
 * initial lambda of the next qesmc_run (boundary ring and top kept at 0, so
 * the converged solution is unchanged).  ms: wall time of the whole call
 * (R and input features, network, installation of lambda0); tflops: achieved
 * BF16 MFMA rate of the network's convolutions (device time of the network). */
int qesmc_unet_predict(qesmc_unet *net, qesmc_ctx *ctx, double *ms, double *tflops);

#ifdef __cplusplus
}
#endif

#endif /* QESMC_H */
