/*
 * rp_fvm_gpu.c -- the RP_FVM program (src/RP_FVM.c) with its step on the gfx950
 * library (include/rpfvm.h).
 *
 *   rp_fvm_gpu [ControlFile.txt] [--mixed] [--positivity] [--control-recon]
 *
 * Same inputs and outputs as the reference program: the ControlFile.txt keys and
 * defaults (src/RP_FVM.c:382-589), the test-case initial data (include/initialize.h),
 * the time-controlled loop with output times Tmax / Control_output and the step-count
 * mode (src/RP_FVM.c:169-320), and the Tecplot .plt / 1-D .dat files
 * (src/RP_FVM.c:751-966) that Post/Blast_post.py reads.  The state lives on the GPU;
 * it comes back to the host only for output.
 *
 * Like the reference, the run uses the solver's effective reconstruction -- WENO5,
 * component-wise (include/Golbal.h:58,60) -- whatever Recon_Accur / Characteriz say
 * (the reference parses them but never applies them); --control-recon applies them.
 * --mixed selects the mixed-precision form, --positivity the positivity safeguard
 * (neither exists in the reference).  TestCase Precision_Test runs the reference's
 * convergence harness (Precision_test) with its steps on the GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <time.h>

#include "rpfvm.h"

#define GC RPFVM_GHOST

/* Test cases, in the order of the reference's TestCase2D (include/Golbal.h:8-30). */
typedef enum {
    TC_PRECISION,
    TC_1D_SHOCKTUBE,
    TC_1D_CONTACTWAVE,
    TC_1D_SHOCKIMPACT,
    TC_1D_IMPACTWALL,
    TC_1D_DOUBLERARE,
    TC_1D_NOH,
    TC_RIEMANN1,
    TC_RIEMANN2,
    TC_RIEMANN3,
    TC_RIEMANN4,
    TC_RIEMANN5,
    TC_TAYLOR_GREEN,
    TC_GAUSSIAN,
    TC_KELVIN_HELMHOLTZ,
    TC_RAYLEIGH_TAYLOR,
    TC_DOUBLE_MACH,
    TC_BACKWARD_STEP,
    TC_ODD_EVEN,
    TC_BLAST_WAVE,
    TC_NOH_2D,
    TC_COUNT
} test_case;

/* control-file name, short name (file names), name in the .plt title */
static const struct {
    const char *key, *shortname, *title;
} CASES[TC_COUNT] = {
    {"Precision_Test", "UNKNOWN", "Error"},
    {"1D_Sod_Shocktube", "SOD", "1D_SOD"},
    {"1D_Contact_Wave", "CONTACT", "1D_CONTACT"},
    {"1D_Shock_Impact", "SHOCKIMPACT", "1D_SHOCKIMPACT"},
    {"1D_Impact_Wall", "IMPACTWALL", "1D_IMPACTWALL"},
    {"1D_Double_Rarefaction", "DOUBLERARE", "1D_DOUBLERARE"},
    {"1D_Noh_Problem", "NOH1D", "1D_NOH"},
    {"2D_Riemann_Case1", "RIEMANN1", "2D_RIEMANN1"},
    {"2D_Riemann_Case2", "RIEMANN2", "2D_RIEMANN2"},
    {"2D_Riemann_Case3", "RIEMANN3", "2D_RIEMANN3"},
    {"2D_Riemann_Case4", "RIEMANN4", "2D_RIEMANN4"},
    {"2D_Riemann_Case5", "RIEMANN5", "2D_RIEMANN5"},
    {"Taylor_Green_Vortex", "TGV", "TGV"},
    {"Gaussian_Pulse", "GAUSSIAN", "GAUSSIAN"},
    {"Kelvin_Helmholtz", "KH", "KH"},
    {"Rayleigh_Taylor", "RT", "RT"},
    {"Double_Mach_Reflection", "DMR", "DMR"},
    {"Backward_Step", "BACKSTEP", "BACKSTEP"},
    {"Odd_Even_Decoupling", "OEDC", "OEDC"},
    {"Blast_Wave", "BLAST", "BLASTWAVE"},
    {"2D_Noh_Problem", "NOH2D", "2D_NOH"},
};

typedef struct {
    test_case tc;
    int nx, ny, time_adm, scheme;
    double gamma;
    int source;
    double gravity, cfl, tmax;
    int control_comp, control_out, recon, characteriz;
    char output_dir[256];
} control;

/* --------------------------------------------------------------- control file */
static void trim(char *s) {
    char *p = s;
    while (*p == ' ') p++;
    memmove(s, p, strlen(p) + 1);
    size_t n = strlen(s);
    while (n && s[n - 1] == ' ') s[--n] = 0;
}

static int truthy(const char *v) { return !strcmp(v, "true") || !strcmp(v, "1"); }

/* ReadControlFile (src/RP_FVM.c:382-589): "key: value" lines, '#' comments. */
static void read_control_file(const char *path, control *c) {
    c->tc = TC_RAYLEIGH_TAYLOR;
    c->nx = 100;
    c->ny = 400;
    c->time_adm = 3;
    c->scheme = 1;
    c->gamma = 1.4;
    c->source = 0;
    c->gravity = 1.0;
    c->cfl = 0.4;
    c->tmax = 1.0;
    c->control_comp = 0;
    c->control_out = 2;
    c->recon = 2;
    c->characteriz = 0;
    strcpy(c->output_dir, "/mnt/d/Desktop/RP_FVM/data");
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "Error: Cannot open control file %s\n", path);
        exit(1);
    }
    char line[256], key[100], val[150];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n' || line[0] == '\r') continue;
        line[strcspn(line, "\n")] = 0;
        line[strcspn(line, "\r")] = 0;
        if (sscanf(line, "%99[^:]: %149[^\n]", key, val) != 2) continue;
        trim(key);
        trim(val);
        if (!strcmp(key, "TestCase")) {
            int found = 0;
            for (int k = 0; k < TC_COUNT; k++)
                if (!strcmp(val, CASES[k].key)) {
                    c->tc = (test_case)k;
                    found = 1;
                }
            if (!found) fprintf(stderr, "Warning: Unknown test case '%s', using default Rayleigh_Taylor\n", val);
        } else if (!strcmp(key, "L_nx")) c->nx = atoi(val);
        else if (!strcmp(key, "L_ny")) c->ny = atoi(val);
        else if (!strcmp(key, "Time_ADM")) c->time_adm = atoi(val);
        else if (!strcmp(key, "scheme")) c->scheme = atoi(val);
        else if (!strcmp(key, "M_gamma")) c->gamma = atof(val);
        else if (!strcmp(key, "Source")) c->source = truthy(val);
        else if (!strcmp(key, "Gravity")) c->gravity = atof(val);
        else if (!strcmp(key, "CFL")) c->cfl = atof(val);
        else if (!strcmp(key, "Tmax")) c->tmax = atof(val);
        else if (!strcmp(key, "Control_Compution")) c->control_comp = atoi(val);
        else if (!strcmp(key, "Control_output")) c->control_out = atoi(val);
        else if (!strcmp(key, "output_dir")) snprintf(c->output_dir, sizeof c->output_dir, "%s", val);
        else if (!strcmp(key, "Recon_Accur")) {
            const int r = atoi(val);
            if (r == 1 || r == 2 || r == 3 || r == 5) c->recon = r;
            else {
                fprintf(stderr, "Warning: Invalid Recon_Accur value %d. Valid values are 1, 2, 3, 5. Using default 2.\n", r);
                c->recon = 2;
            }
        } else if (!strcmp(key, "Characteriz")) {
            if (truthy(val)) c->characteriz = 1;
            else if (!strcmp(val, "false") || !strcmp(val, "0")) c->characteriz = 0;
            else {
                fprintf(stderr, "Warning: Invalid Characteriz value '%s'. Using default false.\n", val);
                c->characteriz = 0;
            }
        } else {
            fprintf(stderr, "Warning: Unknown parameter '%s' in control file\n", key);
        }
    }
    fclose(f);
}

/* ------------------------------------------------------------- initial data */
/* Primitive fields P[4][R][C] (rho, u, v, p) of a test case, its domain, end time and
 * boundary sides, as the initialisers of include/initialize.h set them. */
typedef struct {
    int R, C;
    double *P;
} grid;

#define AT(g, v, i, j) ((g).P[((size_t)(v) * (g).R + (size_t)(i)) * (g).C + (size_t)(j)])

static void fill(grid g, int i, int j, double r, double u, double v, double p) {
    AT(g, 0, i, j) = r;
    AT(g, 1, i, j) = u;
    AT(g, 2, i, j) = v;
    AT(g, 3, i, j) = p;
}

/* two states split at rows/2 over the whole array (1-D tubes, initialize.h:70-350) */
static void tube(grid g, const double a[4], const double b[4]) {
    for (int i = 0; i < g.R; i++)
        for (int j = 0; j < g.C; j++) {
            const double *s = i < g.R / 2 ? a : b;
            fill(g, i, j, s[0], s[1], s[2], s[3]);
        }
}

/* four quadrants split at rows/2, cols/2 (initialize.h:351-625): q[0] i<, j<;
 * q[1] i<, j>=; q[2] i>=, j<; q[3] i>=, j>= */
static void quadrants(grid g, const double q[4][4]) {
    for (int i = 0; i < g.R; i++)
        for (int j = 0; j < g.C; j++) {
            const double *s = q[(i >= g.R / 2) * 2 + (j >= g.C / 2)];
            fill(g, i, j, s[0], s[1], s[2], s[3]);
        }
}

static void init_case(const control *c, grid g, double *Lx, double *Ly, double *T, int bc[4]) {
    const double nx = g.R - 2 * GC, ny = g.C - 2 * GC;
    for (int s = 0; s < 4; s++) bc[s] = RPFVM_BC_OUTFLOW;
    *Lx = *Ly = 1.0;
    switch (c->tc) {
        case TC_PRECISION: { /* initialize.h:36-60 */
            bc[0] = bc[1] = RPFVM_BC_PERIODICITY;
            *Lx = 2.0 * M_PI;
            const double dx = *Lx / (int)nx;
            for (int i = GC; i < g.R - GC; i++)
                fill(g, i, GC, 1.0 + 0.2 * sin((i - GC + 0.5) * dx), 1.0, 0.0, 1.0);
            *T = 2.0;
            break;
        }
        case TC_1D_SHOCKTUBE: {
            const double a[4] = {1.0, 0.0, 0.0, 1.0}, b[4] = {0.125, 0.0, 0.0, 0.1};
            tube(g, a, b);
            *T = 0.2;
            break;
        }
        case TC_1D_CONTACTWAVE: {
            const double a[4] = {1.0, 0.0, 0.0, 1.0}, b[4] = {0.125, 0.0, 0.0, 1.0};
            tube(g, a, b);
            *T = 0.2;
            break;
        }
        case TC_1D_SHOCKIMPACT: {
            const double a[4] = {1.0, 4.0, 0.0, 1.0}, b[4] = {1.0, -4.0, 0.0, 1.0};
            tube(g, a, b);
            *T = 0.2;
            break;
        }
        case TC_1D_DOUBLERARE: {
            const double a[4] = {1.0, -2.0, 0.0, 1.0}, b[4] = {1.0, 2.0, 0.0, 1.0};
            tube(g, a, b);
            *T = 0.15;
            break;
        }
        case TC_1D_IMPACTWALL: {
            const double a[4] = {1.0, 1.0, 0.0, 1.0};
            bc[1] = RPFVM_BC_REFLECTION;
            tube(g, a, a);
            *T = 0.6;
            break;
        }
        case TC_1D_NOH: {
            const double a[4] = {1.0, -1.0, 0.0, 1e-6};
            bc[0] = RPFVM_BC_REFLECTION;
            tube(g, a, a);
            *T = 0.6;
            break;
        }
        case TC_RIEMANN1: {
            const double q[4][4] = {{2, -0.75, 0.5, 1}, {1, -0.75, -0.5, 1}, {1, 0.75, 0.5, 1}, {3, 0.75, -0.5, 1}};
            quadrants(g, q);
            *T = 0.23;
            break;
        }
        case TC_RIEMANN2: {
            const double q[4][4] = {{2, 0, -0.3, 1}, {1, 0, -0.4, 1}, {1.0625, 0, 0.2145, 0.4}, {0.5197, 0, -1.125, 0.4}};
            quadrants(g, q);
            *T = 0.30;
            break;
        }
        case TC_RIEMANN3: {
            const double q[4][4] = {{0.5065, 0.8939, 0, 0.35}, {1.5, 0, 0, 1.5}, {1.1, 0.8939, 0.8939, 1.1},
                                    {0.5065, 0, 0.8939, 0.35}};
            quadrants(g, q);
            *T = 0.25;
            break;
        }
        case TC_RIEMANN4: {
            const double q[4][4] = {{2, 0.75, 0.5, 1}, {1, 0.75, -0.5, 1}, {1, -0.75, 0.5, 1}, {3, -0.75, -0.5, 1}};
            quadrants(g, q);
            *T = 0.30;
            break;
        }
        case TC_RIEMANN5: {
            const double q[4][4] = {{1, -0.6259, 0.1, 1}, {0.5197, 0.1, 0.1, 0.4}, {0.8, 0.1, 0.1, 1},
                                    {1, 0.1, -0.6259, 1}};
            quadrants(g, q);
            *T = 0.25;
            break;
        }
        case TC_TAYLOR_GREEN: { /* initialize.h:626-683: positions from the array index */
            for (int s = 0; s < 4; s++) bc[s] = RPFVM_BC_PERIODICITY;
            *Lx = *Ly = 2.0 * M_PI;
            const double inx = 1.0 / nx, iny = 1.0 / ny, pre = 1.0 * 1.0 * 1.0 / 4.0;
            for (int i = 0; i < g.R; i++)
                for (int j = 0; j < g.C; j++) {
                    const double x = (double)i * inx * (*Lx), y = (double)j * iny * (*Ly);
                    const double u = 1.0 * sin(x) * cos(y), v = -1.0 * cos(x) * sin(y);
                    fill(g, i, j, 1.0, u, v, 1.0 + pre * (cos(2.0 * x) + cos(2.0 * y)));
                }
            *T = 5.0;
            break;
        }
        case TC_GAUSSIAN: { /* initialize.h:684-736 */
            *Lx = *Ly = 10.0;
            const double x0 = *Lx / 2.0, y0 = *Ly / 2.0;
            for (int i = 0; i < g.R; i++)
                for (int j = 0; j < g.C; j++) {
                    const double x = (double)i / nx * (*Lx), y = (double)j / ny * (*Ly);
                    const double r2 = (x - x0) * (x - x0) + (y - y0) * (y - y0);
                    fill(g, i, j, 1.0 + 1.0 * exp(-r2 / (1.0 * 1.0)), 0.5, 0.3, 1.0);
                }
            *T = 4.0;
            break;
        }
        case TC_KELVIN_HELMHOLTZ: { /* initialize.h:737-835, interior only */
            for (int s = 0; s < 4; s++) bc[s] = RPFVM_BC_PERIODICITY;
            const int nxp = g.R - 2 * GC, nyp = g.C - 2 * GC;
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) {
                    const double x = (double)(i - GC + 0.5) / nxp * (*Lx);
                    const double y = (double)(j - GC + 0.5) / nyp * (*Ly);
                    const int slab = fabs(y - 0.5) < 0.25;
                    const double vy = 0.1 * sin(4 * M_PI * x);
                    const double g1 = exp(-pow(y - 0.25, 2) / (2.0 * 0.05 * 0.05));
                    const double g2 = exp(-pow(y - 0.75, 2) / (2.0 * 0.05 * 0.05));
                    fill(g, i, j, slab ? 2.0 : 1.0, slab ? 0.5 : -0.5, vy * (g1 + g2), 2.5);
                }
            *T = 2.0;
            break;
        }
        case TC_RAYLEIGH_TAYLOR: { /* initialize.h:836-884, interior only */
            bc[0] = bc[1] = RPFVM_BC_REFLECTION;
            bc[2] = bc[3] = RPFVM_BC_FIXED_VALUE;
            *Lx = 0.25;
            *T = 1.95;
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) {
                    const double x = (double)(i + 0.5 - GC) / nx * (*Lx);
                    const double y = (double)(j + 0.5 - GC) / ny * (*Ly);
                    if (y < 0.5)
                        fill(g, i, j, 2.0, 0.0,
                             -0.025 * sqrt((2.0 * y + 1.0) * c->gamma / 2.0) * cos(8.0 * M_PI * x),
                             2.0 * y + 1.0);
                    else
                        fill(g, i, j, 1.0, 0.0, -0.025 * sqrt((y + 1.5) * c->gamma / 1.0) * cos(8.0 * M_PI * x),
                             y + 1.5);
                }
            break;
        }
        case TC_DOUBLE_MACH: { /* initialize.h:890-966, interior only */
            bc[0] = RPFVM_BC_INFLOW;
            bc[2] = bc[3] = RPFVM_BC_REFLECTION;
            *Lx = 4.0;
            *T = 0.2;
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) {
                    const double x = (double)(i + 0.5 - GC) / nx * (*Lx);
                    const double y = (double)(j + 0.5 - GC) / ny * (*Ly);
                    if (y < sqrt(3.0) * (x - 1.0 / 6.0)) fill(g, i, j, 1.4, 0.0, 0.0, 1.0);
                    else fill(g, i, j, 8.0, 7.1447, -4.125, 116.5);
                }
            break;
        }
        case TC_BLAST_WAVE: { /* initialize.h:967-1060, interior only */
            for (int s = 0; s < 4; s++) bc[s] = RPFVM_BC_REFLECTION;
            *T = 0.01;
            const double dx = *Lx / nx, dy = *Ly / ny;
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) {
                    const double x = (i + 0.5 - GC) * dx - 0.5, y = (j + 0.5 - GC) * dy - 0.5;
                    const double r = sqrt((x - 0.0) * (x - 0.0) + (y - 0.0) * (y - 0.0));
                    fill(g, i, j, 1.0, 0.0, 0.0, r <= 0.1 ? 10.0 : 0.1);
                }
            break;
        }
        case TC_NOH_2D: { /* initialize.h:1061-1185: as there, conserved values are stored in
                             the primitive array and converted once more, and the
                             perturbation draws rand() (here in serial row-major order) */
            for (int s = 0; s < 4; s++) bc[s] = RPFVM_BC_REFLECTION;
            *T = 0.6;
            const double dx = *Lx / nx, dy = *Ly / ny;
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) {
                    const double x = (i + 0.5 - GC) * dx, y = (j + 0.5 - GC) * dy;
                    const double dist = sqrt(x * x + y * y);
                    double u = 0.0, v = 0.0;
                    if (dist > 1.0e-10) {
                        u = -x / dist;
                        v = -y / dist;
                    }
                    const double pert = 0.001 * (2.0 * ((double)rand() / RAND_MAX) - 1.0);
                    u += pert;
                    v += pert;
                    const double norm = sqrt(u * u + v * v);
                    if (norm > 1.0e-10) {
                        u = u * 1.0 / norm;
                        v = v * 1.0 / norm;
                    }
                    fill(g, i, j, 1.0, 1.0 * u, 1.0 * v,
                         1.0e-6 / (c->gamma - 1.0) + 0.5 * 1.0 * (u * u + v * v));
                    if (dist < 0.05) AT(g, 0, i, j) = 1.0 * (1.0 + 0.001 * ((double)rand() / RAND_MAX - 0.5));
                }
            break;
        }
        case TC_BACKWARD_STEP: { /* initialize.h:1302-1388, interior only */
            bc[0] = RPFVM_BC_FIXED_VALUE;
            bc[2] = bc[3] = RPFVM_BC_REFLECTION;
            *Lx = 3.0;
            *T = 4.0;
            const double u = 3.0 * sqrt(c->gamma * 1.0 / 1.4);
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) fill(g, i, j, 1.4, u, 0.0, 1.0);
            break;
        }
        case TC_ODD_EVEN: { /* initialize.h:1186-1297, interior only */
            bc[2] = bc[3] = RPFVM_BC_REFLECTION;
            *Lx = 20.0;
            *T = 2.0;
            const double gm = 1.4, Ms = 6.0, rho0 = 1.4, p0 = 1.0;
            const double a0 = sqrt(gm * p0 / rho0);
            const double rho1 = rho0 * ((gm + 1.0) * Ms * Ms) / ((gm - 1.0) * Ms * Ms + 2.0);
            const double p1 = p0 * (2.0 * gm * Ms * Ms - (gm - 1.0)) / (gm + 1.0);
            const double u1 = Ms * a0 * (1.0 - rho0 / rho1);
            for (int i = GC; i < g.R - GC; i++)
                for (int j = GC; j < g.C - GC; j++) {
                    const double x = (double)(i + 0.5 - GC) / nx * (*Lx);
                    if (x < (*Lx) * 0.25) fill(g, i, j, rho1, u1, 0.0, p1);
                    else fill(g, i, j, rho0, 0.0, 0.0, p0);
                }
            break;
        }
        default:
            break;
    }
}

/* initEulerconser2D (include/initialize.h:1389-1405), every cell */
static void prim_to_cons(grid g, double gamma, double *U) {
    for (int i = 0; i < g.R; i++)
        for (int j = 0; j < g.C; j++) {
            const double r = AT(g, 0, i, j), u = AT(g, 1, i, j), v = AT(g, 2, i, j), p = AT(g, 3, i, j);
            const size_t o = (size_t)i * g.C + j, pl = (size_t)g.R * g.C;
            U[o] = r;
            U[pl + o] = u * r;
            U[2 * pl + o] = v * r;
            U[3 * pl + o] = p / (gamma - 1) + 0.5 * r * (u * u + v * v);
        }
}

/* Con_to_Pri_2D (include/initialize.h:1437-1454) over [GC-1, R-GC] x [GC-1, C-GC] */
static void cons_to_prim(grid g, double gamma, const double *U) {
    const size_t pl = (size_t)g.R * g.C;
    for (int i = GC - 1; i <= g.R - GC; i++)
        for (int j = GC - 1; j <= g.C - GC; j++) {
            const size_t o = (size_t)i * g.C + j;
            const double r = U[o], u = U[pl + o] / U[o], v = U[2 * pl + o] / U[o];
            const double p = (U[3 * pl + o] - 0.5 * r * (u * u + v * v)) * (gamma - 1);
            fill(g, i, j, r, u, v, p);
        }
}

/* ---------------------------------------------------------------- output */
static const char *scheme_name(int s) {
    switch (s) {
        case 1: return "HLL";
        case 2: return "HLLC";
        case 3: return "Roe";
        case 11: return "HLLHC";
        case 22: return "HLLCHC";
        case 33: return "RoeHC";
        default: return "ExactRiemann";
    }
}

static void make_dirs(const char *dir) {
    char tmp[512];
    snprintf(tmp, sizeof tmp, "%s", dir);
    for (char *p = tmp + 1; *p; p++)
        if (*p == '/') {
            *p = 0;
            mkdir(tmp, 0777);
            *p = '/';
        }
    mkdir(tmp, 0777);
}

/* OutputData_file_2D (src/RP_FVM.c:751-966): cell centres from 0, x outer / y inner,
 * 12 tab-separated %.8f columns (+ IBLANK), or the 1-D profile when ny == 1. */
static int write_output(const control *c, grid pri, const double *U, const double *mx, const double *my,
                        double now) {
    const int nx = pri.R - 2 * GC, ny = pri.C - 2 * GC;
    const char *sn = scheme_name(c->scheme), *title = CASES[c->tc].title;
    char fn[768];
    if (ny == 1)
        snprintf(fn, sizeof fn, "%s/%s_%s_profile_%.6f.dat", c->output_dir, CASES[c->tc].shortname, sn, now);
    else
        snprintf(fn, sizeof fn, "%s/%s_%s_output_data_%.6f.plt", c->output_dir, CASES[c->tc].shortname, sn, now);
    make_dirs(c->output_dir);
    FILE *f = fopen(fn, "w");
    if (!f) {
        printf("Error: Failed to open file: %s\n", fn);
        return 1;
    }
    const size_t pl = (size_t)pri.R * pri.C;
    const double gm = c->gamma;
    if (ny == 1) {
        fprintf(f, "! 1D Profile Data - %s with %s\n", title, sn);
        fprintf(f, "! Time = %.6f\n", now);
        fprintf(f, "! Number of points = %d\n", nx);
        fprintf(f, "X\tDensity\tU\tV\tPressure\tTemperature\tEntropy\tMomentumX\tMomentumY\tEnergy\n");
        for (int i = 0; i < nx; i++) {
            const int a = i + GC, b = GC;
            const size_t o = (size_t)a * pri.C + b;
            const double r = AT(pri, 0, a, b), p = AT(pri, 3, a, b);
            fprintf(f, "%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\n", mx[i], r,
                    AT(pri, 1, a, b), AT(pri, 2, a, b), p, p / r, p / pow(r, gm), U[pl + o], U[2 * pl + o],
                    U[3 * pl + o]);
        }
    } else {
        fprintf(f, "TITLE = \"2D Fluid Dynamics Data - %s with %s\"\n", title, sn);
        fprintf(f, "VARIABLES = \"X\", \"Y\", \"rho\", \"u\", \"v\", \"p\", \"T\", \"Entropy\", \"rhou\", \"rhov\", \"rhoE\", \"IBLANK\"\n");
        const int step = c->tc == TC_BACKWARD_STEP;
        if (step)
            fprintf(f, "ZONE T=\"Backward Step: Time=%.6f\", I=%d, J=%d, DATAPACKING=POINT\n", now, ny, nx);
        else
            fprintf(f, "ZONE T=\"Time=%.6f\", I=%d, J=%d, DATAPACKING=POINT\n", now, ny, nx);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                const int a = i + GC, b = j + GC;
                const size_t o = (size_t)a * pri.C + b;
                const double r = AT(pri, 0, a, b), p = AT(pri, 3, a, b);
                const int iblank = step && mx[i] >= 0.6 && mx[i] <= 3.0 && my[j] >= 0.0 && my[j] <= 0.2 ? 0 : 1;
                fprintf(f, "%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%.8f\t%d\n", mx[i], my[j],
                        r, AT(pri, 1, a, b), AT(pri, 2, a, b), p, p / r, p / pow(r, gm), U[pl + o],
                        U[2 * pl + o], U[3 * pl + o], iblank);
            }
    }
    fclose(f);
    printf("File: %s\n", fn);
    return 0;
}

/* ------------------------------------------------------------------ main */
static void check(int rc, const char *what) {
    if (rc < 0) {
        fprintf(stderr, "%s: %s (%d)\n", what, rpfvm_last_error(), rc);
        exit(2);
    }
}

typedef struct {
    int mixed, positivity, control_recon;
} options;

/* The library configuration of a run: the control keys, the case's sides and the
 * effective reconstruction (WENO5 component-wise unless --control-recon). */
static void run_config(const control *c, const options *o, int nx, int ny, double dx, double dy,
                       rpfvm_config *cfg) {
    cfg->nx = nx;
    cfg->ny = ny;
    cfg->time_order = c->time_adm;
    cfg->scheme = (c->scheme == 1 || c->scheme == 2 || c->scheme == 3 || c->scheme == 11 ||
                   c->scheme == 22 || c->scheme == 33) ? c->scheme : RPFVM_SCHEME_LLF;
    cfg->recon_order = o->control_recon ? c->recon : 5;
    cfg->characteristic = o->control_recon ? c->characteriz : 0;
    cfg->source = c->source;
    cfg->gravity = c->gravity;
    cfg->gamma = c->gamma;
    cfg->cfl = c->cfl;
    cfg->dx = dx;
    cfg->dy = dy;
    cfg->precision = o->mixed ? RPFVM_PRECISION_MIXED : RPFVM_PRECISION_FP64;
    cfg->positivity = o->positivity;
}

/* ------------------------------------------------------- precision harness */
/* Precision_test (src/RP_FVM.c:974-1072): the advected sine wave of Init_Precision
 * (include/initialize.h:36-60, 1609-1629) on nx = 10 .. 640 with ny = L_ny, stepped on
 * the GPU with the fixed dt = 0.5 dx^(5/3) of Precision_Get_Delta_T_2D
 * (include/function.h:239-241) to T = 2, then the L1 / L2 / Linf density errors against
 * 1 + 0.2 sin(x - t) along column GC, the orders between successive grids and the error
 * table (include/Error.h:129-329), printed as the reference prints them. */
typedef struct {
    int n;
    double dx, l1, l2, linf, o1, o2, oinf;
} error_row;

static void density_errors(const double *U, int R, int C, double dx, double t, error_row *e) {
    double l1 = 0.0, l2 = 0.0, linf = 0.0;
    for (int i = GC; i < R - GC; i++) { /* TEST_Scheme_Error_L1, include/Error.h:129-160 */
        const double exact = 1.0 + 0.2 * sin((i - GC + 0.5) * dx - t);
        l1 += fabs(exact - U[(size_t)i * C + GC]) * dx;
    }
    for (int i = GC; i < R - GC; i++) { /* TEST_Scheme_Error_L2, :177-203 */
        const double err = 1.0 + 0.2 * sin((i - GC + 0.5) * dx - t) - U[(size_t)i * C + GC];
        l2 += err * err * dx;
    }
    l2 = sqrt(l2);
    for (int i = GC; i < R - GC; i++) { /* TEST_Scheme_Error_Linf, :220-242 */
        const double err = fabs(1.0 + 0.2 * sin((i - GC + 0.5) * dx - t) - U[(size_t)i * C + GC]);
        if (err > linf) linf = err;
    }
    e->l1 = l1;
    e->l2 = l2;
    e->linf = linf;
    printf("L1 Error = %.10e \n", l1);
    printf("L2 Error = %.10e \n", l2);
    printf("L\u221e Error = %.10e \n", linf);
}

static double order_of(double coarse, double fine, double h_ratio) {
    return coarse > 0 && fine > 0 ? log(coarse / fine) / log(h_ratio) : 0.0;
}

static int precision_test(const control *c, const options *o) {
    static const int NX[] = {10, 20, 40, 80, 160, 320, 640};
    const int count = (int)(sizeof NX / sizeof NX[0]);
    error_row rows[sizeof NX / sizeof NX[0]];
    for (int k = 0; k < count; k++) {
        const int nx = NX[k], R = nx + 2 * GC, C = c->ny + 2 * GC;
        const size_t n = (size_t)4 * R * C;
        grid pri = {R, C, calloc(n, sizeof(double))};
        double *U = calloc(n, sizeof(double));
        if (!pri.P || !U) {
            fprintf(stderr, "Memory allocation failed in Main\n");
            free(pri.P);
            free(U);
            return 1;
        }
        double Lx, Ly, Tmax = c->tmax;
        rpfvm_config cfg;
        rpfvm_config_default(&cfg);
        init_case(c, pri, &Lx, &Ly, &Tmax, cfg.bc);
        prim_to_cons(pri, c->gamma, U);
        const double dx = Lx / nx, dy = Ly / c->ny;
        run_config(c, o, nx, c->ny, dx, dy, &cfg);
        rpfvm_ctx *ctx = NULL;
        check(rpfvm_create(&cfg, &ctx), "rpfvm_create");
        check(rpfvm_set_state(ctx, U), "rpfvm_set_state");
        double time, dt = 0.0;
        for (time = 0.0; time < Tmax; time = time + dt) {
            dt = 0.5 * pow(dx, 5.0 / 3);
            if (time + dt >= Tmax) dt = Tmax - time;
            check(rpfvm_time_advance(ctx, dt), "rpfvm_time_advance");
        }
        const int faults = rpfvm_sync(ctx);
        check(rpfvm_get_state(ctx, U), "rpfvm_get_state");
        rpfvm_destroy(ctx);
        if (faults > 0) fprintf(stderr, "Nx = %d: device fault bits %d\n", nx, faults);
        rows[k].n = nx;
        rows[k].dx = dx;
        density_errors(U, R, C, dx, time, &rows[k]);
        printf("\n--- Completed calculation for Nx = %d, \u0394x = %.6f ---", nx, dx);
        free(pri.P);
        free(U);
    }
    for (int k = 1; k < count; k++) { /* calculate_orders, include/Error.h:259-294 */
        const double h = (double)rows[k].n / (double)rows[k - 1].n;
        rows[k].o1 = order_of(rows[k - 1].l1, rows[k].l1, h);
        rows[k].o2 = order_of(rows[k - 1].l2, rows[k].l2, h);
        rows[k].oinf = order_of(rows[k - 1].linf, rows[k].linf, h);
        printf("Order calculation for grid %d (N=%d) to %d (N=%d): h_ratio = %.4f, L1_order = %.4f\n", k - 1,
               rows[k - 1].n, k, rows[k].n, h, rows[k].o1);
    }
    /* print_error_table, include/Error.h:303-329 */
    const char *rule = "===============================================================================";
    printf("\n\n%s\n", rule);
    printf("Table 3.3: Errors and numerical orders of accuracy of the density\n");
    printf("%s\n", rule);
    printf("Grid points  \u0394x        L1 error    Order     L2 error    Order     L\u221e error    Order\n");
    printf("------------------------------------------------------------------------------\n");
    printf("%-11d  %-9.6f  %-10.2e  %-8s  %-10.2e  %-8s  %-10.2e  %-8s\n", rows[0].n, rows[0].dx, rows[0].l1,
           "---", rows[0].l2, "---", rows[0].linf, "---");
    for (int k = 1; k < count; k++)
        printf("%-11d  %-9.6f  %-10.2e  %-8.2f  %-10.2e  %-8.2f  %-10.2e  %-8.2f\n", rows[k].n, rows[k].dx,
               rows[k].l1, rows[k].o1, rows[k].l2, rows[k].o2, rows[k].linf, rows[k].oinf);
    printf("%s\n", rule);
    printf("The program has completed its execution.");
    return 0;
}

int main(int argc, char **argv) {
    control c;
    const char *path = NULL;
    options o = {0, 0, 0};
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--mixed")) o.mixed = 1;
        else if (!strcmp(argv[a], "--positivity")) o.positivity = 1;
        else if (!strcmp(argv[a], "--control-recon")) o.control_recon = 1;
        else path = argv[a];
    }
    if (path) {
        read_control_file(path, &c);
    } else { /* src/RP_FVM.c:97-113 */
        printf("Using default parameters (no control file specified)\n");
        memset(&c, 0, sizeof c);
        c.tc = TC_BACKWARD_STEP;
        c.nx = 300;
        c.ny = 100;
        c.time_adm = 1;
        c.scheme = 10;
        c.gamma = 1.4;
        c.gravity = 1.0;
        c.cfl = 0.4;
        c.tmax = 0.1;
        c.control_out = 4;
        c.recon = 1;
        c.characteriz = 1;
        strcpy(c.output_dir, "/mnt/d/Desktop/RP_FVM/data");
    }
    if (c.scheme == 5) {
        fprintf(stderr, "scheme 5 (the Marquina stub) is not implemented on the GPU\n");
        return 2;
    }
    if (c.tc == TC_PRECISION) return precision_test(&c, &o); /* src/RP_FVM.c:118-122 */
    printf("Test case: %s (%s), %d x %d, scheme %s\n", CASES[c.tc].key, CASES[c.tc].shortname, c.nx, c.ny,
           scheme_name(c.scheme));

    const int R = c.nx + 2 * GC, C = c.ny + 2 * GC;
    const size_t n = (size_t)4 * R * C;
    grid pri = {R, C, calloc(n, sizeof(double))};
    double *U = calloc(n, sizeof(double));
    double *mx = calloc((size_t)c.nx, sizeof(double)), *my = calloc((size_t)c.ny, sizeof(double));
    if (!pri.P || !U || !mx || !my) {
        fprintf(stderr, "Memory allocation failed in Main\n");
        return 1;
    }
    double Lx, Ly, Tmax = c.tmax;
    rpfvm_config cfg;
    rpfvm_config_default(&cfg);
    init_case(&c, pri, &Lx, &Ly, &Tmax, cfg.bc);
    prim_to_cons(pri, c.gamma, U);
    const double dx = Lx / c.nx, dy = Ly / c.ny;
    for (int j = 0; j < c.nx; j++) mx[j] = dx * j + 0.5 * dx; /* Mesh_2D, include/mesh.h:151-160 */
    for (int k = 0; k < c.ny; k++) my[k] = dy * k + 0.5 * dy;

    run_config(&c, &o, c.nx, c.ny, dx, dy, &cfg);
    rpfvm_ctx *ctx = NULL;
    check(rpfvm_create(&cfg, &ctx), "rpfvm_create");
    check(rpfvm_set_state(ctx, U), "rpfvm_set_state");

    int ite = 0;
    double time = 0.0, dt = 0.0;
    const double t0 = (double)clock() / CLOCKS_PER_SEC;
    if (c.control_comp == 0) {
        /* src/RP_FVM.c:169-267 */
        const double output_time = Tmax / c.control_out;
        double next_output = output_time;
        for (time = 0.0; time < Tmax; time = time + dt) {
            check(rpfvm_get_delta_t(ctx, &dt), "rpfvm_get_delta_t");
            if (time + dt > next_output) dt = next_output - time;
            if (time + dt >= Tmax) dt = Tmax - time;
            check(rpfvm_set_time(ctx, time), "rpfvm_set_time");
            check(rpfvm_time_advance(ctx, dt), "rpfvm_time_advance");
            const double now = time + dt;
            if (fabs(now - next_output) < 1e-10 * output_time) {
                if (ite > 0) {
                    check(rpfvm_get_state(ctx, U), "rpfvm_get_state");
                    cons_to_prim(pri, c.gamma, U);
                    printf("Step %d completed | Time: %.6f\n", ite, now);
                    write_output(&c, pri, U, mx, my, now);
                }
                next_output += output_time;
            }
            ite++;
        }
    } else if (c.control_comp == 1) {
        /* src/RP_FVM.c:269-317: 400 steps of dt = 0.1 dx, RK1 or RK3 */
        time = 0.0;
        for (int m = 0; m <= 399; m++) {
            dt = 0.1 * dx;
            if (c.time_adm == 1 || c.time_adm == 3) check(rpfvm_time_advance(ctx, dt), "rpfvm_time_advance");
            else fprintf(stderr, "Error: Invalid time advancement method\n");
            ite++;
            time = time + dt;
            if (ite % c.control_out == 0) printf("Step = %d     Time = %.6f\n", ite, time);
        }
    } else {
        fprintf(stderr, "Error: Invalid control computation mode\n");
    }
    const int faults = rpfvm_sync(ctx);
    const double t1 = (double)clock() / CLOCKS_PER_SEC;
    printf("Steps Completed: %d  Final Time: %.6f  Host CPU time: %.3f s  Fault bits: %d\n", ite, time,
           t1 - t0, faults);
    check(rpfvm_get_state(ctx, U), "rpfvm_get_state");
    cons_to_prim(pri, c.gamma, U);
    write_output(&c, pri, U, mx, my, Tmax); /* src/RP_FVM.c:356-360 */
    rpfvm_destroy(ctx);
    free(pri.P);
    free(U);
    free(mx);
    free(my);
    return 0;
}
