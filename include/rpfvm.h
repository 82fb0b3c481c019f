/*
 * rpfvm.h -- C ABI of the MI355X (gfx950) finite-volume step for the 2-D
 * compressible Euler equations.  Drop-in for the FP64 CPU/OpenMP step of
 * RP_FVM (xujf520/OverheatingProblem).
 *
 * Everything here is plain C: pointers, sizes and ints, no C++ or torch types.
 * A caller owns every host buffer; device buffers belong to the context.
 *
 * Layout contract (the reference's, src/RP_FVM.c:124-142):
 *   a host state is double U[4][R][C], R = nx + 8 (x, first index),
 *   C = ny + 8 (y, contiguous index), 4 ghost layers per side, variables
 *   (rho, rho*u, rho*v, E).
 *
 * Return codes: 0 = ok; < 0 = invalid argument or HIP runtime error;
 * > 0 = numerical fault detected on the device (see RPFVM_FAULT_*).
 * The library never calls exit().
 *
 * One host thread per context.  All work of a context is ordered on the
 * context's own HIP stream.
 */
#ifndef RPFVM_H
#define RPFVM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPFVM_GHOST 4      /* GhostCell, include/Golbal.h:49 */
#define RPFVM_NVAR 4       /* var, include/Golbal.h:33 */

/* BC_Type, include/Boundary_Condition.h:22-28 (same numbering) */
enum {
    RPFVM_BC_OUTFLOW = 0,
    RPFVM_BC_REFLECTION = 1,
    RPFVM_BC_PERIODICITY = 2,
    RPFVM_BC_INFLOW = 3,
    RPFVM_BC_FIXED_VALUE = 4,
    /* Double-Mach-reflection sides (Woodward & Colella), the two non-trivial sides of
     * BC_DoubleMach_2D (include/Boundary_Condition.h:480-660), which the reference's
     * steppers leave commented out (include/time_advance.h:591-620,826-861).  Both are
     * evaluated at the time of the RK stage being filled.  Cell-centre position along
     * the side: x_i = ((i - 4 + 0.5) / nx) * lx. */
    RPFVM_BC_DMR_WALL = 5,   /* bottom only: x_i <= bc_par[2][0]: fixed_state[2];
                                beyond: mirrored reflection (v negated) */
    RPFVM_BC_DMR_SHOCK = 6   /* top only: x_i < x_s(t): fixed_state[3], else
                                bc_alt_state[3]; x_s(t) = p0 + (p1 / p2) + (p3 * t / p2)
                                with p = bc_par[3] (start x, height, slope, speed) */
};

/* Riemann solver ids as the "scheme" control key, include/CFD_convection.h:89-117.
 * Any other id except 5 selects LLF, as the reference's default branch.  The
 * heat-capacity solvers 11, 22, 33 (HLLHC_Flux, HLLCHC_Flux, RoeHC_Flux,
 * include/scheme.h:1860, 2153, 2793) run in both forms, in a context and in batches. */
enum {
    RPFVM_SCHEME_HLL = 1, RPFVM_SCHEME_HLLC = 2, RPFVM_SCHEME_ROE = 3, RPFVM_SCHEME_LLF = 10,
    RPFVM_SCHEME_HLLHC = 11, RPFVM_SCHEME_HLLCHC = 22, RPFVM_SCHEME_ROEHC = 33
};

/* Arithmetic form of the step. */
enum {
    RPFVM_PRECISION_FP64 = 0,  /* bit-compatible with the reference FP64 path */
    RPFVM_PRECISION_MIXED = 1  /* FP32 reconstruction/projection/flux, FP64 state and dt */
};

/* Fault bits reported by rpfvm_get_faults (also returned by rpfvm_sync). */
enum {
    RPFVM_FAULT_HLL_BAD_STATE = 1,   /* HLL face with rho <= 0, NaN rho or NaN wave speed,
                                        or RoeHC face with a NaN Roe density (the reference
                                        prints and skips: scheme.h:427-436,500-503,2936-2939) */
    RPFVM_FAULT_NONFINITE_DT = 2,    /* dt reduction produced a non-finite or non-positive dt */
    RPFVM_FAULT_MIXED_FALLBACK = 4,  /* mixed form recomputed at least one face in FP64
                                        (informational, not an error) */
    RPFVM_FAULT_POSITIVITY = 8,      /* the positivity safeguard left a cell with rho <= 0
                                        or p <= 0 */
    RPFVM_FAULT_GRID_SYNC = 16       /* reserved (not raised) */
};

/* rpfvm_config.debug_flags bits. */
enum {
    RPFVM_DEBUG_MIXED_ALL_FP64 = 1,  /* every mixed-form face through the FP64 recompute list */
    RPFVM_DEBUG_LARGE_ARRAY_FORMS = 2, /* the row forms of the sweeps (the fused y sweep one
                                        row per wavefront instead of the LDS-staged marching
                                        form, global instead of buffer addressing, the mixed
                                        form's two-sweep stage) at any size; they run by
                                        default where an array (4 planes) reaches 4 GB and on
                                        grids below 8 M cells (FP64) / 2 M cells (mixed),
                                        which they fill with more wavefronts; identical
                                        results */
    RPFVM_DEBUG_TWO_SWEEP = 4,       /* the mixed form's stage as an x sweep followed by the
                                        fused y sweep (the default for characteristic
                                        WENO5); identical results */
    RPFVM_DEBUG_ONE_PASS = 8,        /* the mixed form's stage as one pass (x faces, y faces
                                        and update in one kernel; the default for the other
                                        reconstructions) also for characteristic WENO5;
                                        identical results */
    RPFVM_DEBUG_MARCH_FORMS = 16     /* the marching forms on grids below the size where
                                        the row forms run by default (arrays below 4 GB);
                                        identical results */
};

/* Error codes (< 0). */
enum {
    RPFVM_OK = 0,
    RPFVM_ERR_INVALID = -1,      /* bad argument or configuration */
    RPFVM_ERR_UNSUPPORTED = -2,  /* configuration not implemented on the device */
    RPFVM_ERR_HIP = -3,          /* HIP runtime error (see rpfvm_last_error) */
    RPFVM_ERR_NOMEM = -4
};

/*
 * Step configuration.  The ControlFile.txt keys (src/RP_FVM.c:434-529) map as
 *   L_nx -> nx, L_ny -> ny, Time_ADM -> time_order, scheme -> scheme,
 *   Recon_Accur -> recon_order, Characteriz -> characteristic, Source -> source,
 *   M_gamma -> gamma, CFL -> cfl, Gravity -> gravity,
 * and the process globals bc_config / inflow_state / *_fixed_value_state
 * (src/RP_FVM.c:86, include/Boundary_Condition.h:49, include/initialize.h:1298-1301)
 * become bc[4] / inflow_state / fixed_state.
 */
typedef struct rpfvm_config {
    int32_t nx, ny;            /* interior cells in x (rows) and y (columns) */
    int32_t time_order;        /* 1, 2 or 3 (RK1_TimeAd, RK2_TimeAd, RK3_TimeAd) */
    int32_t scheme;            /* RPFVM_SCHEME_* */
    int32_t recon_order;       /* 1 Godunov, 2 TVD, 3 WENO3, 5 WENO5 */
    int32_t characteristic;    /* 0 component-wise, 1 characteristic-variable reconstruction */
    int32_t source;            /* gravity source term on/off */
    int32_t precision;         /* RPFVM_PRECISION_* */
    int32_t positivity;        /* 0: reference behaviour.  1: a face whose reconstructed
                                  rho or p is not positive (or not finite) falls back to
                                  first-order (piecewise-constant) states.  Not in the
                                  reference, which has no positivity safeguard. */
    int32_t debug_flags;       /* 0 normally; RPFVM_DEBUG_* bits select verification-only
                                  paths (RPFVM_DEBUG_MIXED_ALL_FP64: the mixed form routes
                                  every face through its FP64 recompute list;
                                  RPFVM_DEBUG_LARGE_ARRAY_FORMS: see above) */
    int32_t bc[4];             /* left, right, bottom, top: RPFVM_BC_* */
    double dx, dy;
    double gamma, cfl, gravity;
    double inflow_state[4];    /* conserved (rho, rho u, rho v, E) */
    double fixed_state[4][4];  /* [side][var], side order left, right, bottom, top */
    double lx;                 /* domain length in x for RPFVM_BC_DMR_* (0: nx * dx) */
    double bc_par[4][4];       /* per-side parameters of RPFVM_BC_DMR_* */
    double bc_alt_state[4][4]; /* per-side second state of RPFVM_BC_DMR_SHOCK */
} rpfvm_config;

typedef struct rpfvm_ctx rpfvm_ctx;

/* Fill a config with the reference's effective defaults: RK3, HLLC, WENO5
 * component-wise (Recon_Accur = 5, Characteriz = false, include/Golbal.h:58,60),
 * gamma 1.4, CFL 0.5, outflow on all sides, inflow/fixed states as the
 * reference's initialisers (include/Boundary_Condition.h:49). */
void rpfvm_config_default(rpfvm_config *cfg);

/* Context lifecycle.  Allocates device state and scratch on the current HIP
 * device.  Replaces the per-call malloc/free of RK*_TimeAd and
 * Flux_Reconstruction_RP (include/time_advance.h:791-821, CFD_convection.h:43-45). */
int rpfvm_create(const rpfvm_config *cfg, rpfvm_ctx **out);
int rpfvm_destroy(rpfvm_ctx *ctx);

/* Host <-> device state transfer, reference layout double[4][R][C]. */
int rpfvm_set_state(rpfvm_ctx *ctx, const double *U);
int rpfvm_get_state(rpfvm_ctx *ctx, double *U);

/* Get_Delta_T_2D, include/function.h:144-175: dt = CFL*min(dx,dy)/(max(|u|+a)+max(|v|+a))
 * over the interior (NaN speeds skipped).  After rpfvm_time_advance the maxima come from
 * the last stage's kernel (the state is not read again); after rpfvm_set_state or
 * rpfvm_device_state, from a pass over the state.  Both give the same dt, bit for bit.
 * Synchronises the stream and returns dt on the host. */
int rpfvm_get_delta_t(rpfvm_ctx *ctx, double *dt);

/* RK1_TimeAd / RK2_TimeAd / RK3_TimeAd (include/time_advance.h:643-882), chosen by
 * cfg.time_order: ghost fill of every stage input (including the state itself),
 * spatial operator, stage combine.  Asynchronous on the context stream. */
int rpfvm_time_advance(rpfvm_ctx *ctx, double dt);

/* Boundary_Conditions, include/Boundary_Condition.h:963-980: ghost fill of the
 * state in place, sides in the order left, right, bottom, top.  Asynchronous. */
int rpfvm_boundary_conditions(rpfvm_ctx *ctx);

/* Source_Gravity + Boundary_Conditions + Flux_Reconstruction_RP + Space_Discrete_Item
 * (include/time_advance.h:824-829) on the current state: ghost fill in place, then
 * L = -(F_i - F_{i-1})/dx - (G_j - G_{j-1})/dy + S.  Writes host arrays
 * L[4][R][C] (interior), F[4][R][C] (x faces i+1/2 at [i], i = 3..R-5) and
 * G[4][R][C] (y faces j+1/2 at [j]); any of them may be NULL.  Synchronous. */
int rpfvm_space_operator(rpfvm_ctx *ctx, double *L, double *F, double *G);

/* One Riemann solver alone, as the reference's per-solver flux calls
 * HLL_Flux / HLLC_Flux / Roe_Flux / LLF_Flux / HLLHC_Flux / HLLCHC_Flux / RoeHC_Flux
 * (dir, var, rows, cols, GC, L_state, R_state, flux) (include/scheme.h; dispatched at
 * include/CFD_convection.h:89-117): the context's scheme on host face states
 * Ls[4][R][C], Rs[4][R][C] for dir 1 (x faces i+1/2 at [i], i = 3..R-5, j = 4..C-5)
 * or 2 (y faces), into host F[4][R][C]; other entries of F are left as they are on
 * the device (the flux array of the context).  A face the reference skips gets NaN and
 * sets RPFVM_FAULT_HLL_BAD_STATE.  FP64 contexts only; overwrites the stage arrays,
 * not the state.  Synchronous. */
int rpfvm_riemann(rpfvm_ctx *ctx, int dir, const double *Ls, const double *Rs, double *F);

/* Time-controlled loop of the reference driver (src/RP_FVM.c:179-191), on the
 * device: repeat { dt = Get_Delta_T_2D; clip to t_stop; step } until *t reaches
 * t_stop or max_steps steps have run.  dt and t never leave the device inside the
 * loop; *t and *steps are updated on return.  Synchronous. */
int rpfvm_advance_to(rpfvm_ctx *ctx, double t_stop, int max_steps, double *t, int *steps);

/* Clock of the context, read by the time-dependent boundary types: rpfvm_time_advance
 * fills stage k's ghosts at t + c_k dt (c = 0, 1, 1/2 for RK3; 0, 1 for RK2) and then
 * advances the clock by dt; rpfvm_advance_to starts from *t.  Starts at 0. */
int rpfvm_set_time(rpfvm_ctx *ctx, double t);
int rpfvm_get_time(rpfvm_ctx *ctx, double *t);

/* Wait for the context stream and return the fault bits accumulated since the last
 * call (0 when clean); the bits are cleared. */
int rpfvm_sync(rpfvm_ctx *ctx);
int rpfvm_get_faults(rpfvm_ctx *ctx, int *faults);

/* Get_Delta_T_2D on a caller's device state, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same reduction and dt as
 * rpfvm_get_delta_t's pass over the state, bit for bit, without a context: U is the
 * device address of plane 0, row 0, column 0 of a state laid out as four planes of
 * nx + 8 rows with a row pitch of ld doubles (ld >= ny + 8) and plane v at
 * U + v * plane_stride; dt (device) receives dt = CFL*min(dx,dy)/(S_x+S_y), clipped like
 * the driver (src/RP_FVM.c:186-191) against clock[0] = t, clock[1] = t_stop when clock
 * (device, two doubles) is not NULL; faults (device, optional) is OR-ed with
 * RPFVM_FAULT_NONFINITE_DT when dt is not positive and finite (the caller clears it).
 * Only cfg's nx, ny, dx, dy, gamma and cfl are read.  The launches go on `stream` (a
 * hipStream_t, NULL for the default stream) and the call returns without waiting.
 * Scratch: a fixed device array of 2048 workgroup maxima pairs (32 KiB) per GPU, held
 * until the process ends; any grid the configuration admits fits in it.  Calls that
 * overlap on one GPU (other streams or host threads) share that array and race. */
int rpfvm_delta_t_noalloc(const rpfvm_config *cfg, const double *U, size_t ld,
                          size_t plane_stride, const double *clock, double *dt,
                          unsigned *faults, void *stream);

/* Device access for callers that keep data on the GPU.  The state lives in four
 * planes of R rows with a row pitch of *ld doubles (ld >= C); plane v starts at
 * base + v * plane_stride.  The stream is a hipStream_t.  The next dt is normally
 * reduced from wave-speed maxima the last step left behind; calling
 * rpfvm_device_state makes the next rpfvm_get_delta_t read the state itself, so call
 * it again after writing the state through the pointer. */
int rpfvm_device_state(rpfvm_ctx *ctx, double **base, size_t *ld, size_t *plane_stride);
int rpfvm_stream(rpfvm_ctx *ctx, void **stream);

/* Kernel timing (no counterpart in the reference, which only times whole steps with
 * omp_get_wtime, src/RP_FVM.c:181-216).  While enabled, every kernel launch of the
 * context is bracketed by HIP events on the context stream.  rpfvm_kernel_times
 * synchronises, returns total milliseconds and launch counts per kernel class
 * (RPFVM_KCLASS_*), and resets them. */
enum {
    RPFVM_KCLASS_BC = 0,       /* ghost fill */
    RPFVM_KCLASS_FLUX_X = 1,   /* x-face reconstruction + Riemann flux */
    RPFVM_KCLASS_FLUX_Y = 2,   /* y-face reconstruction + Riemann flux (fused with the
                                  residual + stage combine where the stage writes a
                                  separate output; the mixed form's one-pass stage, x faces
                                  included, is timed here too) */
    RPFVM_KCLASS_UPDATE = 3,   /* residual + RK stage combine */
    RPFVM_KCLASS_DT = 4,       /* CFL reduction */
    RPFVM_KCLASS_COUNT = 5
};
int rpfvm_set_kernel_timing(rpfvm_ctx *ctx, int enabled);
int rpfvm_kernel_times(rpfvm_ctx *ctx, double ms[RPFVM_KCLASS_COUNT], int counts[RPFVM_KCLASS_COUNT]);

/* ---------------------------------------------------------------------------
 * Batching form: many independent problems of uneven size advanced together
 * (no counterpart in the reference, whose driver runs one problem per process,
 * src/RP_FVM.c:92-372).  Each problem keeps its own configuration (grid, dx/dy,
 * gamma, CFL, Riemann solver, boundary conditions, source) and its own dt and clock;
 * recon_order (1, 2, 3 or 5), characteristic, time_order, precision and the positivity
 * safeguard (two-sweep schedule only) must be the same for every problem.  One RK
 * stage of all active problems is one persistent launch driven by a work queue
 * of strips; the whole step is captured once in a HIP graph and replayed.  Each
 * problem's interior evolves bit-identically to an rpfvm_ctx with the same
 * configuration driven by rpfvm_advance_to(t_end).  rpfvm_batch_get_state returns the
 * interior and leaves the ghost entries of U as rpfvm_batch_set_state wrote them (the
 * one-pass schedules apply the boundary conditions on load; the two-sweep schedule
 * fills device ghost frames that are not copied back).
 * ------------------------------------------------------------------------- */
typedef struct rpfvm_batch rpfvm_batch;

/* Schedule of a stage: one marching pass (stencils from cache, or through an LDS row
 * ring), or two sweeps (ghost fill, x fluxes to a buffer, y fluxes fused with the
 * update). */
enum { RPFVM_STAGING_DIRECT = 0, RPFVM_STAGING_LDS = 1, RPFVM_STAGING_SWEEPS = 2 };

/* Launch shape of the stage kernel.  rpfvm_batch_tune measures candidates on the
 * batch's current state and keeps the fastest. */
typedef struct rpfvm_batch_tuning {
    int staging;        /* RPFVM_STAGING_* */
    int rows;           /* rows a wavefront marches per strip (4..1024) */
    int waves_per_cu;   /* resident wavefronts of the persistent grid per CU (1..32) */
    int use_graph;      /* 1: replay the captured step graph; 0: launch eagerly */
    double ms_per_stage;/* measured stage time of this shape (0 when set by hand) */
} rpfvm_batch_tuning;

int rpfvm_batch_create(int count, const rpfvm_config *cfgs, rpfvm_batch **out);
int rpfvm_batch_destroy(rpfvm_batch *b);
/* Host U[4][nx+8][ny+8] of problem k; resets its clock to t = 0. */
int rpfvm_batch_set_state(rpfvm_batch *b, int k, const double *U);
int rpfvm_batch_get_state(rpfvm_batch *b, int k, double *U);
/* One step of every problem with t < t_end (dt clipped to reach t_end exactly, as
 * src/RP_FVM.c:186-191).  Asynchronous. */
int rpfvm_batch_step(rpfvm_batch *b, double t_end);
/* Steps until every problem has reached t_end or max_steps steps were launched;
 * *steps = steps launched.  Synchronous. */
int rpfvm_batch_advance_to(rpfvm_batch *b, double t_end, int max_steps, int *steps);
/* Per-problem time, step count and fault bits (arrays of count entries; any may
 * be NULL).  Synchronous. */
int rpfvm_batch_status(rpfvm_batch *b, double *t, int *steps, int *faults);
/* OR of the fault bits of all problems since the last call (cleared). */
int rpfvm_batch_sync(rpfvm_batch *b);
int rpfvm_batch_tune(rpfvm_batch *b, rpfvm_batch_tuning *chosen);
int rpfvm_batch_set_tuning(rpfvm_batch *b, const rpfvm_batch_tuning *t);
int rpfvm_batch_get_tuning(rpfvm_batch *b, rpfvm_batch_tuning *t);
int rpfvm_batch_stream(rpfvm_batch *b, void **stream);

/* Message of the last error seen on this thread. */
const char *rpfvm_last_error(void);

/* Library version string. */
const char *rpfvm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* RPFVM_H */
