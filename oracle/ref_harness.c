/*
 * ref_harness.c -- test infrastructure, not product code.
 *
 * A driver for the reference solver's own FP64 step, used only to write the
 * golden fixtures under tests/golden/ and, as "cpu_baseline.kind = reference",
 * to time the reference CPU/OpenMP path.  It is compiled by oracle/Makefile
 * against the reference headers where they lie (-I <reference>/include); no
 * reference source is copied into this repository.  This file takes the place
 * of the reference's main() (src/RP_FVM.c:92-372): it sets the solver globals
 * the way set_material_parameters (src/RP_FVM.c:600-605) does, and it also
 * sets the file-scope statics Recon_Accur / Characteriz (include/Golbal.h:58,60),
 * which the reference program parses but never applies.
 *
 * Usage:  ref_harness <job.txt> <in.bin> <out.bin>
 *   job.txt  "key value" lines (see read_job)
 *   in.bin   raw little-endian doubles, layout depends on mode
 *   out.bin  raw doubles, layout depends on mode
 *
 * Modes
 *   step      in: U[4][R][C]; out: dt[steps], U[4][R][C]
 *             per step: dt = Get_Delta_T_2D (or fixed_dt), RK{1,2,3}_TimeAd
 *             (src/RP_FVM.c:182, 193-209)
 *   bc        in: U; out: U after Boundary_Conditions (include/Boundary_Condition.h:963)
 *   dmrbc     in: U; out: U after BC_DoubleMach_2D at `time` (include/Boundary_Condition.h:480)
 *   residual  in: U; out: U (after BC), L[4][R][C], F[4][R][C], G[4][R][C]
 *             Source_Gravity, Boundary_Conditions, Flux_Reconstruction_RP,
 *             Space_Discrete_Item (include/time_advance.h:824-829)
 *   recon     in: U; out: conserl, conserr for dir = 1 then dir = 2
 *             (Flux_Reconstruction_RP's Recon_Accur switch, include/CFD_convection.h:67-85)
 *   flux      in: Lstate[4][R][C], Rstate[4][R][C]; out: Flux(dir=1), Flux(dir=2)
 *             for the solver id in "scheme" (include/CFD_convection.h:89-117)
 *   eigen     in: U; out: EL(x)[4][4][R][C], ER(x), EL(y), ER(y)
 *             (Con_to_Pri_2D + Compute_Eigen_2DX/2DY)
 *   weno      in: n x 6 doubles; out: n x 4 doubles
 *             (WENO5Z_L, WENO5Z_R, WENO5_L, WENO5_R on &f[2])
 *   dt        in: U; out: 1 double (Get_Delta_T_2D)
 *   bench     in: U; runs "steps" steps like "step", prints timing to stdout,
 *             out: dt[steps], U
 *   run       in: U; the reference driver's time loop to "t_end" (src/RP_FVM.c:179-191
 *             with next_output_time = Tmax = t_end, at most "steps" steps when
 *             steps > 0), timing printed to stdout as in "bench" (dt + step, IO
 *             excluded, src/RP_FVM.c:181-216); out: n, t, dt[n], U
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include "Golbal.h"
#include "mesh.h"
#include "initialize.h"
#include "CFD_convection.h"
#include "Boundary_Condition.h"
#include "time_advance.h"
#include "scheme.h"

/* The reference program defines this global in src/RP_FVM.c:86. */
BoundaryConfig bc_config = {BC_OUTFLOW, BC_OUTFLOW, BC_OUTFLOW, BC_OUTFLOW};

typedef struct {
    char mode[32];
    int nx, ny, steps, time_order, scheme, recon, characteriz, source;
    int bc[4];
    int n; /* weno: number of stencils */
    double gamma, cfl, dx, dy, gravity, fixed_dt, time, t_end;
    double inflow[4];
    double fixed[4][4];
} job_t;

static void die(const char *msg) { fprintf(stderr, "ref_harness: %s\n", msg); exit(2); }

static void read_job(const char *path, job_t *j) {
    memset(j, 0, sizeof *j);
    strcpy(j->mode, "step");
    j->steps = 1; j->time_order = 3; j->scheme = 2; j->recon = 5;
    j->gamma = 1.4; j->cfl = 0.5; j->dx = j->dy = 1.0; j->fixed_dt = -1.0;
    for (int k = 0; k < 4; k++) j->inflow[k] = inflow_state[k];
    for (int k = 0; k < 4; k++) {
        j->fixed[0][k] = L_fixed_value_state[k];
        j->fixed[1][k] = R_fixed_value_state[k];
        j->fixed[2][k] = B_fixed_value_state[k];
        j->fixed[3][k] = T_fixed_value_state[k];
    }
    FILE *f = fopen(path, "r");
    if (!f) die("cannot open job file");
    char key[64], val[128];
    while (fscanf(f, "%63s %127s", key, val) == 2) {
        if (!strcmp(key, "mode")) strncpy(j->mode, val, sizeof j->mode - 1);
        else if (!strcmp(key, "nx")) j->nx = atoi(val);
        else if (!strcmp(key, "ny")) j->ny = atoi(val);
        else if (!strcmp(key, "n")) j->n = atoi(val);
        else if (!strcmp(key, "steps")) j->steps = atoi(val);
        else if (!strcmp(key, "time_order")) j->time_order = atoi(val);
        else if (!strcmp(key, "scheme")) j->scheme = atoi(val);
        else if (!strcmp(key, "recon")) j->recon = atoi(val);
        else if (!strcmp(key, "characteriz")) j->characteriz = atoi(val);
        else if (!strcmp(key, "source")) j->source = atoi(val);
        else if (!strcmp(key, "bc_left")) j->bc[0] = atoi(val);
        else if (!strcmp(key, "bc_right")) j->bc[1] = atoi(val);
        else if (!strcmp(key, "bc_bottom")) j->bc[2] = atoi(val);
        else if (!strcmp(key, "bc_top")) j->bc[3] = atoi(val);
        else if (!strcmp(key, "gamma")) j->gamma = strtod(val, NULL);
        else if (!strcmp(key, "cfl")) j->cfl = strtod(val, NULL);
        else if (!strcmp(key, "dx")) j->dx = strtod(val, NULL);
        else if (!strcmp(key, "dy")) j->dy = strtod(val, NULL);
        else if (!strcmp(key, "gravity")) j->gravity = strtod(val, NULL);
        else if (!strcmp(key, "fixed_dt")) j->fixed_dt = strtod(val, NULL);
        else if (!strcmp(key, "time")) j->time = strtod(val, NULL);
        else if (!strcmp(key, "t_end")) j->t_end = strtod(val, NULL);
        else if (!strncmp(key, "inflow", 6)) j->inflow[atoi(key + 6)] = strtod(val, NULL);
        else if (!strncmp(key, "fixed", 5)) {
            int side = key[5] - '0', comp = key[6] - '0';
            if (side < 0 || side > 3 || comp < 0 || comp > 3) die("bad fixedSC key");
            j->fixed[side][comp] = strtod(val, NULL);
        } else {
            fprintf(stderr, "ref_harness: unknown key %s\n", key);
            exit(2);
        }
    }
    fclose(f);
}

static double *read_doubles(const char *path, size_t n) {
    double *p = malloc(n * sizeof(double));
    if (!p) die("oom");
    FILE *f = fopen(path, "rb");
    if (!f) die("cannot open input");
    if (fread(p, sizeof(double), n, f) != n) die("short input");
    fclose(f);
    return p;
}

static void write_doubles(FILE *f, const double *p, size_t n) {
    if (fwrite(p, sizeof(double), n, f) != n) die("short write");
}

static void run_step(int scheme, int order, int rows, int cols, double (*U)[rows][cols],
                     double dt, double dx, double dy) {
    switch (order) {
        case 1: RK1_TimeAd(scheme, var, rows, cols, GhostCell, U, dt, dx, dy); break;
        case 2: RK2_TimeAd(scheme, var, rows, cols, GhostCell, U, dt, dx, dy); break;
        case 3: RK3_TimeAd(scheme, var, rows, cols, GhostCell, U, dt, dx, dy); break;
        default: die("bad time_order");
    }
}

int main(int argc, char **argv) {
    if (argc != 4) die("usage: ref_harness job.txt in.bin out.bin");
    job_t j;
    read_job(argv[1], &j);

    /* set_material_parameters (src/RP_FVM.c:600-605) plus the statics it leaves alone */
    M_gamma = j.gamma;
    Source = j.source != 0;
    Gravity = j.gravity;
    CFL = j.cfl;
    Recon_Accur = j.recon;
    Characteriz = j.characteriz != 0;
    bc_config.left = (BC_Type)j.bc[0];
    bc_config.right = (BC_Type)j.bc[1];
    bc_config.bottom = (BC_Type)j.bc[2];
    bc_config.top = (BC_Type)j.bc[3];
    for (int k = 0; k < 4; k++) {
        inflow_state[k] = j.inflow[k];
        L_fixed_value_state[k] = j.fixed[0][k];
        R_fixed_value_state[k] = j.fixed[1][k];
        B_fixed_value_state[k] = j.fixed[2][k];
        T_fixed_value_state[k] = j.fixed[3][k];
    }

    FILE *out = fopen(argv[3], "wb");
    if (!out) die("cannot open output");

    if (!strcmp(j.mode, "weno")) {
        double *in = read_doubles(argv[2], (size_t)j.n * 6);
        for (int s = 0; s < j.n; s++) {
            double *f = in + 6 * (size_t)s;
            double r[4] = {WENO5Z_L(&f[2]), WENO5Z_R(&f[2]), WENO5_L(&f[2]), WENO5_R(&f[2])};
            write_doubles(out, r, 4);
        }
        free(in);
        fclose(out);
        return 0;
    }

    const int rows = j.nx + 2 * GhostCell, cols = j.ny + 2 * GhostCell;
    const size_t plane = (size_t)rows * cols, nU = 4 * plane;
    typedef double grid_t[rows][cols];

    if (!strcmp(j.mode, "flux")) {
        double *in = read_doubles(argv[2], 2 * nU);
        grid_t *Ls = (grid_t *)in, *Rs = (grid_t *)(in + nU);
        grid_t *Fl = calloc(nU, sizeof(double));
        for (int dir = 1; dir <= 2; dir++) {
            memset(Fl, 0, nU * sizeof(double));
            switch (j.scheme) {
                case 1: HLL_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
                case 2: HLLC_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
                case 3: Roe_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
                case 11: HLLHC_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
                case 22: HLLCHC_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
                case 33: RoeHC_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
                default: LLF_Flux(dir, var, rows, cols, GhostCell, Ls, Rs, Fl); break;
            }
            write_doubles(out, (double *)Fl, nU);
        }
        free(Fl); free(in); fclose(out);
        return 0;
    }

    double *ubuf = read_doubles(argv[2], nU);
    grid_t *U = (grid_t *)ubuf;

    if (!strcmp(j.mode, "step") || !strcmp(j.mode, "bench")) {
        double *dts = malloc(sizeof(double) * (size_t)(j.steps > 0 ? j.steps : 1));
        double t0 = omp_get_wtime();
        for (int s = 0; s < j.steps; s++) {
            double dt = j.fixed_dt > 0 ? j.fixed_dt
                                       : Get_Delta_T_2D(var, rows, cols, U, j.dx, j.dy);
            dts[s] = dt;
            run_step(j.scheme, j.time_order, rows, cols, U, dt, j.dx, j.dy);
        }
        double t1 = omp_get_wtime();
        if (!strcmp(j.mode, "bench"))
            printf("{\"seconds\": %.9f, \"steps\": %d, \"cells\": %lld, \"threads\": %d}\n",
                   t1 - t0, j.steps, (long long)j.nx * j.ny, omp_get_max_threads());
        write_doubles(out, dts, (size_t)j.steps);
        write_doubles(out, ubuf, nU);
        free(dts);
    } else if (!strcmp(j.mode, "run")) {
        size_t cap = 1024, n = 0;
        double *dts = malloc(sizeof(double) * cap);
        double Time = 0.0, Delta_T = 0.0, wall = 0.0;
        for (Time = 0.0; Time < j.t_end && (j.steps <= 0 || (int)n < j.steps); Time = Time + Delta_T) {
            double t0 = omp_get_wtime();
            Delta_T = Get_Delta_T_2D(var, rows, cols, U, j.dx, j.dy);
            if (Time + Delta_T > j.t_end) Delta_T = j.t_end - Time;
            if (Time + Delta_T >= j.t_end) Delta_T = j.t_end - Time;
            run_step(j.scheme, j.time_order, rows, cols, U, Delta_T, j.dx, j.dy);
            wall += omp_get_wtime() - t0;
            if (n == cap) {
                cap *= 2;
                dts = realloc(dts, sizeof(double) * cap);
                if (!dts) die("oom");
            }
            dts[n++] = Delta_T;
            if (n % 500 == 0) fprintf(stderr, "ref_harness run: step %zu t %.6f\n", n, Time + Delta_T);
        }
        printf("{\"seconds\": %.9f, \"steps\": %zu, \"t\": %.17g, \"cells\": %lld, \"threads\": %d}\n",
               wall, n, Time, (long long)j.nx * j.ny, omp_get_max_threads());
        double hdr[2] = {(double)n, Time};
        write_doubles(out, hdr, 2);
        write_doubles(out, dts, n);
        write_doubles(out, ubuf, nU);
        free(dts);
    } else if (!strcmp(j.mode, "dt")) {
        double dt = Get_Delta_T_2D(var, rows, cols, U, j.dx, j.dy);
        write_doubles(out, &dt, 1);
    } else if (!strcmp(j.mode, "bc")) {
        Boundary_Conditions(var, rows, cols, U, GhostCell);
        write_doubles(out, ubuf, nU);
    } else if (!strcmp(j.mode, "dmrbc")) {
        BC_DoubleMach_2D(var, rows, cols, U, GhostCell, j.time);
        write_doubles(out, ubuf, nU);
    } else if (!strcmp(j.mode, "residual")) {
        grid_t *F = calloc(nU, sizeof(double)), *G = calloc(nU, sizeof(double));
        grid_t *S = calloc(nU, sizeof(double)), *Lr = calloc(nU, sizeof(double));
        if (Source) Source_Gravity(var, rows, cols, GhostCell, U, S);
        Boundary_Conditions(var, rows, cols, U, GhostCell);
        Flux_Reconstruction_RP(j.scheme, var, rows, cols, GhostCell, U, F, G, 0.0, j.dx, j.dy);
        Space_Discrete_Item(var, rows, cols, GhostCell, j.dx, j.dy, F, G, S, Lr);
        write_doubles(out, ubuf, nU);
        write_doubles(out, (double *)Lr, nU);
        write_doubles(out, (double *)F, nU);
        write_doubles(out, (double *)G, nU);
        free(F); free(G); free(S); free(Lr);
    } else if (!strcmp(j.mode, "recon")) {
        grid_t *cl = calloc(nU, sizeof(double)), *cr = calloc(nU, sizeof(double));
        for (int dir = 1; dir <= 2; dir++) {
            memset(cl, 0, nU * sizeof(double));
            memset(cr, 0, nU * sizeof(double));
            switch (Recon_Accur) {
                case 1: Reconstruction_Godunov(dir, var, rows, cols, GhostCell, U, cl, cr, j.dx); break;
                case 2: TVD_Reconstruction(dir, var, rows, cols, GhostCell, U, cl, cr, j.dx, j.dy); break;
                case 3: WENO3_Reconstruction(dir, var, rows, cols, GhostCell, U, cl, cr); break;
                case 5: WENO5_Reconstruction(dir, var, rows, cols, GhostCell, U, cl, cr); break;
                default: die("bad recon");
            }
            write_doubles(out, (double *)cl, nU);
            write_doubles(out, (double *)cr, nU);
        }
        free(cl); free(cr);
    } else if (!strcmp(j.mode, "eigen")) {
        grid_t *Pri = calloc(nU, sizeof(double));
        double *EL = calloc(4 * nU, sizeof(double)), *ER = calloc(4 * nU, sizeof(double));
        typedef double egrid_t[4][rows][cols];
        Con_to_Pri_2D(var, rows, cols, Pri, U);
        Compute_Eigen_2DX(var, rows, cols, Pri, (egrid_t *)EL, (egrid_t *)ER);
        write_doubles(out, EL, 4 * nU);
        write_doubles(out, ER, 4 * nU);
        memset(EL, 0, 4 * nU * sizeof(double));
        memset(ER, 0, 4 * nU * sizeof(double));
        Compute_Eigen_2DY(var, rows, cols, Pri, (egrid_t *)EL, (egrid_t *)ER);
        write_doubles(out, EL, 4 * nU);
        write_doubles(out, ER, 4 * nU);
        free(Pri); free(EL); free(ER);
    } else {
        die("unknown mode");
    }
    free(ubuf);
    fclose(out);
    return 0;
}
