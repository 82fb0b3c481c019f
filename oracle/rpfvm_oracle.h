/*
 * rpfvm_oracle.h -- CPU oracle for the 2-D Euler FV step (TEST INFRASTRUCTURE).
 *
 * A clean-room FP64 restatement of the reference solver's hot path
 * (xujf520/OverheatingProblem, RP_FVM).  Only the tests, the smoke check and the
 * CPU-baseline leg of bench.py load it, always as the checker; the product
 * (libs under overheatingproblem_amd/) never links it.
 *
 * Arrays use the reference layout double[4][R][C], R = nx + 8, C = ny + 8.
 */
#ifndef RPFVM_ORACLE_H
#define RPFVM_ORACLE_H

#include "../include/rpfvm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Get_Delta_T_2D (include/function.h:144-175). */
double orc_delta_t(const rpfvm_config *cfg, const double *U);

/* Boundary_Conditions (include/Boundary_Condition.h:963-980), in place. */
int orc_boundary_conditions(const rpfvm_config *cfg, double *U);
/* The same at clock t (time-dependent sides). */
int orc_boundary_conditions_t(const rpfvm_config *cfg, double *U, double t);

/* Source_Gravity + Boundary_Conditions (in place on U) + Flux_Reconstruction_RP +
 * Space_Discrete_Item.  L, F, G are full [4][R][C] arrays (any may be NULL). */
int orc_space_operator(const rpfvm_config *cfg, double *U, double *L, double *F, double *G);

/* RK1/2/3_TimeAd (include/time_advance.h:643-882), in place on U. */
int orc_time_advance(const rpfvm_config *cfg, double *U, double dt);
/* The same starting at clock t: stage k fills its ghosts at t + c_k dt. */
int orc_time_advance_t(const rpfvm_config *cfg, double *U, double dt, double t);

/* Driver time loop (src/RP_FVM.c:179-191) with one output time t_stop: steps until
 * t reaches t_stop or max_steps; returns steps taken, *t updated, dts[] optional. */
int orc_advance_to(const rpfvm_config *cfg, double *U, double t_stop, int max_steps,
                   double *t, double *dts);

/* The reconstruction alone on a state whose ghosts are filled: left/right face
 * states for dir 1 (x) or 2 (y), stored at the face's left cell as the reference does. */
int orc_reconstruct(const rpfvm_config *cfg, const double *U, int dir, double *CL, double *CR);

/* Riemann flux of cfg->scheme on face states Ls/Rs ([4][R][C]) for dir 1 or 2,
 * over the reference's face ranges; other entries of F are left untouched. */
int orc_riemann(const rpfvm_config *cfg, int dir, const double *Ls, const double *Rs, double *F);

/* Point reconstructions on n six-point stencils f[0..5] (cells i-2..i+3):
 * out[4*s + 0..3] = WENO5Z_L, WENO5Z_R, WENO5_L, WENO5_R of &f[2]. */
void orc_weno5_points(int n, const double *in6, double *out4);

/* Roe-averaged eigenvectors at every face of a direction (Con_to_Pri_2D +
 * Compute_Eigen_2DX/2DY); EL/ER are [4][4][R][C] (zero where no face). */
int orc_eigen(const rpfvm_config *cfg, const double *U, int dir, double *EL, double *ER);

/* Cells repaired by the positivity safeguard during the last orc_time_advance. */
long orc_last_repairs(void);

/* Number of OpenMP threads the oracle runs with. */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
