/*
 * rpfvm_oracle.c -- CPU oracle of the 2-D Euler FV step.  TEST INFRASTRUCTURE:
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * and only as the checker.  Nothing in the product links it.
 *
 * A clean-room FP64 restatement of the reference solver (xujf520/OverheatingProblem,
 * "RP_FVM").  It is organised by face (gather a 6-cell stencil, reconstruct,
 * solve the Riemann problem) instead of the reference's plane-by-plane passes,
 * but every floating-point expression is evaluated in the same order as the
 * reference's, so the results are bit-identical.  Build with -O3 and
 * -ffp-contract=off (no FMA), like the reference build (CmakeLists.txt:17-22).
 *
 * Pinning (see tests/test_oracle_golden.py and tests/test_oracle_math.py):
 * every function below is checked bit-for-bit against fixtures written by the
 * reference's own code (oracle/ref_harness.c, tests/golden/make_golden.py), and
 * against closed forms (WENO consistency, L.R = I, flux consistency F(U,U) = F(U),
 * conservation, free-stream preservation).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "rpfvm_oracle.h"

#define GC 4

typedef struct {
    int R, C;
} dims_t;

static dims_t dims_of(const rpfvm_config *c) {
    dims_t d = {c->nx + 2 * GC, c->ny + 2 * GC};
    return d;
}

#define AT(A, v, i, j) ((A)[((size_t)(v) * d.R + (size_t)(i)) * (size_t)d.C + (size_t)(j)])

int orc_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ helpers */
/* sgn, max_of_two, min_of_two: include/function.h:31-63 */
static inline int sgn_(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
static inline double max2_(double a, double b) { return a > b ? a : b; }
static inline double min2_(double a, double b) { return a > b ? b : a; }
/* min_mod: include/function.h:66-68 */
static inline double minmod_(double a, double b) {
    return 0.5 * (sgn_(a) + sgn_(b)) * min2_(fabs(a), fabs(b));
}
/* van_leer: include/function.h:71-74 */
static inline double vanleer_(double a, double b) {
    return ((sgn_(a) + sgn_(b)) * a * b) / (fabs(a) + fabs(b) + 1.0e-6);
}

/* ------------------------------------------------------- point reconstructions
 * Stencil s[0..5] holds cells i-2 .. i+3 of the face i+1/2. */

/* WENO5-Z (Borges et al., exponent 1): include/Reconstruction.h:1519-1624.
 * (v1..v5) are the stencil values in upwind order for the side being built. */
static inline double weno5z_(double v1, double v2, double v3, double v4, double v5) {
    const double eps = 1.0e-15;
    double b1 = 13.0 / 12.0 * (v1 - 2.0 * v2 + v3) * (v1 - 2.0 * v2 + v3)
              + 0.25 * (v1 - 4.0 * v2 + 3.0 * v3) * (v1 - 4.0 * v2 + 3.0 * v3);
    double b2 = 13.0 / 12.0 * (v2 - 2.0 * v3 + v4) * (v2 - 2.0 * v3 + v4)
              + 0.25 * (v2 - v4) * (v2 - v4);
    double b3 = 13.0 / 12.0 * (v3 - 2.0 * v4 + v5) * (v3 - 2.0 * v4 + v5)
              + 0.25 * (3.0 * v3 - 4.0 * v4 + v5) * (3.0 * v3 - 4.0 * v4 + v5);
    double tau = fabs(b1 - b3);
    double a1 = 0.1 * (1.0 + tau / (b1 + eps));
    double a2 = 0.6 * (1.0 + tau / (b2 + eps));
    double a3 = 0.3 * (1.0 + tau / (b3 + eps));
    double w1 = a1 / (a1 + a2 + a3);
    double w2 = a2 / (a1 + a2 + a3);
    double w3 = a3 / (a1 + a2 + a3);
    return w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3) / 6.0
         + w2 * (-v2 + 5.0 * v3 + 2.0 * v4) / 6.0
         + w3 * (2.0 * v3 + 5.0 * v4 - v5) / 6.0;
}

/* WENO5-JS (Jiang & Shu 1996): include/Reconstruction.h:1408-1508. */
static inline double weno5js_(double v1, double v2, double v3, double v4, double v5) {
    const double eps = 1.0e-15;
    double b1 = 13.0 / 12.0 * (v1 - 2.0 * v2 + v3) * (v1 - 2.0 * v2 + v3)
              + 0.25 * (v1 - 4.0 * v2 + 3.0 * v3) * (v1 - 4.0 * v2 + 3.0 * v3);
    double b2 = 13.0 / 12.0 * (v2 - 2.0 * v3 + v4) * (v2 - 2.0 * v3 + v4)
              + 0.25 * (v2 - v4) * (v2 - v4);
    double b3 = 13.0 / 12.0 * (v3 - 2.0 * v4 + v5) * (v3 - 2.0 * v4 + v5)
              + 0.25 * (3.0 * v3 - 4.0 * v4 + v5) * (3.0 * v3 - 4.0 * v4 + v5);
    double a1 = 0.1 / (eps + b1) / (eps + b1);
    double a2 = 0.6 / (eps + b2) / (eps + b2);
    double a3 = 0.3 / (eps + b3) / (eps + b3);
    double w1 = a1 / (a1 + a2 + a3);
    double w2 = a2 / (a1 + a2 + a3);
    double w3 = a3 / (a1 + a2 + a3);
    return w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3) / 6.0
         + w2 * (-v2 + 5.0 * v3 + 2.0 * v4) / 6.0
         + w3 * (2.0 * v3 + 5.0 * v4 - v5) / 6.0;
}

/* WENO3: include/Reconstruction.h:1326-1398 (pow(x,2) is x*x). */
static inline double weno3_(double v1, double v2, double v3) {
    const double eps = 1e-6;
    double s1 = (v2 - v1) * (v2 - v1);
    double s2 = (v3 - v2) * (v3 - v2);
    double a1 = (1.0 / 3.0) / ((eps + s1) * (eps + s1));
    double a2 = (2.0 / 3.0) / ((eps + s2) * (eps + s2));
    double w1 = a1 / (a1 + a2);
    double w2 = a2 / (a1 + a2);
    return w1 * (-0.5 * v1 + 1.5 * v2) + w2 * (0.5 * v2 + 0.5 * v3);
}

/* TVD slopes with minmod (lim 0) or van Leer (lim 1): include/Reconstruction.h:1088-1209. */
static inline double tvd_slope_(double v1, double v2, double v3, double delta, int lim) {
    double s1 = (v2 - v1) / delta;
    double s2 = (v3 - v2) / delta;
    return lim ? vanleer_(s1, s2) : minmod_(s1, s2);
}

/* Left/right face values of one scalar stencil s[0..5] for reconstruction order
 * `order` (1 Godunov, 2 TVD, 3 WENO3, 5 WENO5); `z` selects WENO5-Z over JS and the
 * van Leer limiter over minmod, as the characteristic branch does. */
static inline void recon_scalar_(const double s[6], int order, int z, double delta,
                                 double *left, double *right) {
    switch (order) {
        case 1:
            *left = s[2];
            *right = s[3];
            break;
        case 2: {
            double sl = tvd_slope_(s[1], s[2], s[3], delta, z);
            double sr = tvd_slope_(s[2], s[3], s[4], delta, z);
            *left = s[2] + 0.5 * sl * delta;
            *right = s[3] - 0.5 * sr * delta;
            break;
        }
        case 3:
            *left = weno3_(s[1], s[2], s[3]);
            *right = weno3_(s[4], s[3], s[2]);
            break;
        default:
            if (z) {
                *left = weno5z_(s[0], s[1], s[2], s[3], s[4]);
                *right = weno5z_(s[5], s[4], s[3], s[2], s[1]);
            } else {
                *left = weno5js_(s[0], s[1], s[2], s[3], s[4]);
                *right = weno5js_(s[5], s[4], s[3], s[2], s[1]);
            }
            break;
    }
}

void orc_weno5_points(int n, const double *in6, double *out4) {
    for (int s = 0; s < n; s++) {
        const double *f = in6 + 6 * (size_t)s;
        out4[4 * (size_t)s + 0] = weno5z_(f[0], f[1], f[2], f[3], f[4]);
        out4[4 * (size_t)s + 1] = weno5z_(f[5], f[4], f[3], f[2], f[1]);
        out4[4 * (size_t)s + 2] = weno5js_(f[0], f[1], f[2], f[3], f[4]);
        out4[4 * (size_t)s + 3] = weno5js_(f[5], f[4], f[3], f[2], f[1]);
    }
}

/* ----------------------------------------------------------- characteristics */
/* Con_to_Pri_2D, include/initialize.h:1437-1454, for one cell. */
static inline void prim_(double g, const double c[4], double p[4]) {
    double rho = c[0];
    double u = c[1] / c[0];
    double v = c[2] / c[0];
    p[0] = rho;
    p[1] = u;
    p[2] = v;
    p[3] = (c[3] - 0.5 * rho * (u * u + v * v)) * (g - 1);
}

/* Cell enthalpy as Compute_Eigen_2DX/2DY compute it (include/Characteriz.h:318, 420). */
static inline double enthalpy_(double g, const double p[4]) {
    return 0.5 * (p[1] * p[1] + p[2] * p[2]) + (g / (g - 1)) * (p[3] / p[0]);
}

/* Roe-averaged left (EL[wave][var]) and right (ER[var][wave]) eigenvectors of the
 * face between primitive states a and b.  dir 1: Compute_Eigen_2DX
 * (include/Characteriz.h:325-395); dir 2: Compute_Eigen_2DY (:427-497). */
static void face_eigen_(double g, int dir, const double a[4], const double b[4],
                        double EL[4][4], double ER[4][4]) {
    double Ha = enthalpy_(g, a), Hb = enthalpy_(g, b);
    double D = sqrt(b[0] / a[0]);
    double D1 = D + 1.0;
    double u = (a[1] + D * b[1]) / D1;
    double v = (a[2] + D * b[2]) / D1;
    double H = (Ha + D * Hb) / D1;
    double q2 = u * u + v * v;
    double c2 = (g - 1.0) * (H - 0.5 * q2);
    double c = sqrt(c2);
    double b1 = (g - 1.0) / c2;
    double b2 = 1.0 + b1 * q2 - b1 * H;
    if (dir == 1) {
        EL[0][0] = 0.5 * (b2 + u / c);
        EL[0][1] = -0.5 * (b1 * u + 1.0 / c);
        EL[0][2] = -0.5 * b1 * v;
        EL[0][3] = 0.5 * b1;
        EL[1][0] = -q2 + H;
        EL[1][1] = u;
        EL[1][2] = v;
        EL[1][3] = -1.0;
        EL[2][0] = v;
        EL[2][1] = 0.0;
        EL[2][2] = -1.0;
        EL[2][3] = 0.0;
        EL[3][0] = 0.5 * (b2 - u / c);
        EL[3][1] = 0.5 * (-b1 * u + 1.0 / c);
        EL[3][2] = 0.5 * -b1 * v;
        EL[3][3] = 0.5 * b1;
        /* ER[var][wave] = eigen_r[var][wave] */
        ER[0][0] = 1.0;
        ER[0][1] = b1;
        ER[0][2] = 0.0;
        ER[0][3] = 1.0;
        ER[1][0] = u - 1.0 * c;
        ER[1][1] = u * b1;
        ER[1][2] = 0.0;
        ER[1][3] = u + 1.0 * c;
        ER[2][0] = v - 0.0 * c;
        ER[2][1] = v * b1;
        ER[2][2] = -1.0;
        ER[2][3] = v + 0.0 * c;
        ER[3][0] = H - u * c;
        ER[3][1] = H * b1 - 1.0;
        ER[3][2] = -v;
        ER[3][3] = H + u * c;
    } else {
        EL[0][0] = 0.5 * (b2 + v / c);
        EL[0][1] = -0.5 * (b1 * u + 0.0 / c);
        EL[0][2] = -0.5 * (b1 * v + 1.0 / c);
        EL[0][3] = 0.5 * b1;
        EL[1][0] = -q2 + H;
        EL[1][1] = u;
        EL[1][2] = v;
        EL[1][3] = -1.0;
        EL[2][0] = -u;
        EL[2][1] = 1.0;
        EL[2][2] = -0.0;
        EL[2][3] = 0.0;
        EL[3][0] = 0.5 * (b2 - v / c);
        EL[3][1] = 0.5 * (-b1 * u + 0.0 / c);
        EL[3][2] = 0.5 * (-b1 * v + 1.0 / c);
        EL[3][3] = 0.5 * b1;
        ER[0][0] = 1.0;
        ER[0][1] = b1;
        ER[0][2] = 0.0;
        ER[0][3] = 1.0;
        ER[1][0] = u - 0.0 * c;
        ER[1][1] = u * b1;
        ER[1][2] = 1.0;
        ER[1][3] = u + 0.0 * c;
        ER[2][0] = v - 1.0 * c;
        ER[2][1] = v * b1;
        ER[2][2] = -0.0;
        ER[2][3] = v + 1.0 * c;
        ER[3][0] = H - v * c;
        ER[3][1] = H * b1 - 1.0;
        ER[3][2] = u;
        ER[3][3] = H + v * c;
    }
}

/* Face states from the 6-cell stencil st[n][var] (cells i-2..i+3).
 * Characteristic branch: include/Reconstruction.h:909-967 (and the Godunov /
 * TVD / WENO3 twins at :73-149, :285-380, :704-791); component branch: :1031-1073
 * (and :185-236, :392-444, :822-868). */
static void reconstruct_face_(const rpfvm_config *cfg, int dir, double st[6][4],
                              double UL[4], double UR[4]) {
    const double g = cfg->gamma;
    const int order = cfg->recon_order;
    const double delta = dir == 1 ? cfg->dx : cfg->dy;
    if (cfg->characteristic) {
        double pa[4], pb[4], EL[4][4], ER[4][4];
        prim_(g, st[2], pa);
        prim_(g, st[3], pb);
        face_eigen_(g, dir, pa, pb, EL, ER);
        double wl[4], wr[4];
        for (int k = 0; k < 4; k++) {
            if (order == 1) {
                /* Godunov projects only cells i and i+1 (:95-99) */
                double l = 0.0, r = 0.0;
                for (int m = 0; m < 4; m++) {
                    l += st[2][m] * EL[k][m];
                    r += st[3][m] * EL[k][m];
                }
                wl[k] = l;
                wr[k] = r;
            } else {
                double w[6];
                for (int n = 0; n < 6; n++) {
                    w[n] = 0.0;
                    for (int m = 0; m < 4; m++) w[n] += st[n][m] * EL[k][m];
                }
                recon_scalar_(w, order, 1, delta, &wl[k], &wr[k]);
            }
        }
        for (int v = 0; v < 4; v++) {
            double l = 0.0, r = 0.0;
            for (int k = 0; k < 4; k++) {
                l += wl[k] * ER[v][k];
                r += wr[k] * ER[v][k];
            }
            UL[v] = l;
            UR[v] = r;
        }
    } else {
        for (int v = 0; v < 4; v++) {
            double s[6];
            for (int n = 0; n < 6; n++) s[n] = st[n][v];
            recon_scalar_(s, order, 0, delta, &UL[v], &UR[v]);
        }
    }
}

/* Positivity test of a reconstructed face state: rho > 0 and p > 0 (false for NaN). */
static inline int admissible_(double g, const double c[4]) {
    double ir = 1.0 / c[0];
    double u = c[1] * ir, v = c[2] * ir;
    double p = (c[3] - 0.5 * c[0] * (u * u + v * v)) * (g - 1.0);
    return c[0] > 0.0 && p > 0.0 && isfinite(p) && isfinite(u) && isfinite(v);
}

/* Face states with the optional positivity safeguard (an extension; the reference
 * has none): when either reconstructed state has rho <= 0 or p <= 0, the face falls
 * back to the first-order states of its two cells. */
static void face_states_(const rpfvm_config *cfg, int dir, double st[6][4], double UL[4],
                         double UR[4]) {
    reconstruct_face_(cfg, dir, st, UL, UR);
    if (cfg->positivity && !(admissible_(cfg->gamma, UL) && admissible_(cfg->gamma, UR))) {
        for (int v = 0; v < 4; v++) {
            UL[v] = st[2][v];
            UR[v] = st[3][v];
        }
    }
}

/* -------------------------------------------------------------- Riemann flux */
typedef struct {
    double r, ru, rv, E;   /* conserved */
    double ir, u, v, u2, v2, p;
} side_t;

static inline void side_(double g, const double c[4], side_t *s) {
    s->r = c[0];
    s->ru = c[1];
    s->rv = c[2];
    s->E = c[3];
    s->ir = 1.0 / s->r;
    s->u = s->ru * s->ir;
    s->v = s->rv * s->ir;
    s->u2 = s->u * s->u;
    s->v2 = s->v * s->v;
    s->p = (s->E - 0.5 * s->r * (s->u2 + s->v2)) * (g - 1.0);
}

/* Physical flux of a side with enthalpy H, in direction dir. */
static inline void phys_flux_(int dir, const side_t *s, double H, double F[4]) {
    if (dir == 1) {
        F[0] = s->r * s->u;
        F[1] = s->r * s->u2 + s->p;
        F[2] = s->r * s->v * s->u;
        F[3] = s->r * H * s->u;
    } else {
        F[0] = s->r * s->v;
        F[1] = s->r * s->u * s->v;
        F[2] = s->r * s->v2 + s->p;
        F[3] = s->r * H * s->v;
    }
}

/* HLLC with Einfeldt speeds: include/scheme.h:690-1001.  Returns 0. */
static int hllc_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    const double gm1 = g - 1.0;
    const double gr = g / gm1;
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double aL, aR;
    if (dir == 1) { /* :745-746 */
        aL = sqrt((g - 1.0) * (HL - 0.5 * (L.u2 + L.v2)));
        aR = sqrt((g - 1.0) * (HR - 0.5 * (R.u2 + R.v2)));
    } else { /* :895-896 */
        aL = sqrt(g * L.p * L.ir);
        aR = sqrt(g * R.p * R.ir);
    }
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double sL = sqrt(L.r), sR = sqrt(R.r);
    double ss = sL + sR;
    double iss = 1.0 / ss;
    double ubar = (sL * L.u + sR * R.u) * iss;
    double vbar = (sL * L.v + sR * R.v) * iss;
    double eta = dir == 1 ? 0.5 * sL * sR * iss * iss      /* :777 */
                          : 0.5 * sL * sR / (ss * ss);     /* :929 */
    double qL = dir == 1 ? L.u : L.v, qR = dir == 1 ? R.u : R.v;
    double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    double nbar = dir == 1 ? ubar : vbar;
    double sl = nbar - d, sr = nbar + d;
    double denom = L.r * (sl - qL) - R.r * (sr - qR);
    double sstar = (R.p - L.p + L.r * qL * (sl - qL) - R.r * qR * (sr - qR)) / denom;
    double mL = L.r * (sl - qL) / (sl - sstar);
    double mR = R.r * (sr - qR) / (sr - sstar);
    double eL = L.E * L.ir + (sstar - qL) * (sstar + L.p / (L.r * (sl - qL)));
    double eR = R.E * R.ir + (sstar - qR) * (sstar + R.p / (R.r * (sr - qR)));
    if (sl >= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else if (sstar >= 0.0) {
        F[0] = FL[0] + sl * (mL - L.r);
        if (dir == 1) {
            F[1] = FL[1] + sl * (mL * sstar - L.ru);
            F[2] = FL[2] + sl * (mL * L.v - L.rv);
        } else {
            F[1] = FL[1] + sl * (mL * L.u - L.ru);
            F[2] = FL[2] + sl * (mL * sstar - L.rv);
        }
        F[3] = FL[3] + sl * (mL * eL - L.E);
    } else {
        F[0] = FR[0] + sr * (mR - R.r);
        if (dir == 1) {
            F[1] = FR[1] + sr * (mR * sstar - R.ru);
            F[2] = FR[2] + sr * (mR * R.v - R.rv);
        } else {
            F[1] = FR[1] + sr * (mR * R.u - R.ru);
            F[2] = FR[2] + sr * (mR * sstar - R.rv);
        }
        F[3] = FR[3] + sr * (mR * eR - R.E);
    }
    return 0;
}

/* HLL with Einfeldt speeds: include/scheme.h:400-674.  A face the reference skips
 * (rho <= 0, NaN rho, NaN speed: :427-436, :500-503) gets a NaN flux and returns 1. */
static int hll_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    if (cl[0] <= 0.0 || cr[0] <= 0.0 || isnan(cl[0]) || isnan(cr[0])) {
        for (int k = 0; k < 4; k++) F[k] = NAN;
        return 1;
    }
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    const double gr = g / (g - 1.0);
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double aL = sqrt((g - 1.0) * (HL - 0.5 * (L.u2 + L.v2)));
    double aR = sqrt((g - 1.0) * (HR - 0.5 * (R.u2 + R.v2)));
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double sL = sqrt(L.r), sR = sqrt(R.r);
    double ss = sL + sR;
    double iss = 1.0 / ss;
    double ubar = (sL * L.u + sR * R.u) * iss;
    double vbar = (sL * L.v + sR * R.v) * iss;
    double eta = 0.5 * sL * sR / (ss * ss);
    double qL = dir == 1 ? L.u : L.v, qR = dir == 1 ? R.u : R.v;
    double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    if (isnan(d)) {
        for (int k = 0; k < 4; k++) F[k] = NAN;
        return 1;
    }
    double nbar = dir == 1 ? ubar : vbar;
    double sl = nbar - d, sr = nbar + d;
    if (sl >= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else {
        double isd = 1.0 / (sr - sl);
        double slsr = sl * sr;
        for (int k = 0; k < 4; k++) F[k] = (sr * FL[k] - sl * FR[k] + slsr * (cr[k] - cl[k])) * isd;
    }
    return 0;
}

/* Positively conservative HLLE (Einfeldt, Munz, Roe & Sjogreen 1991): wave-speed
 * bounds s_L = min(q_L - a_L, q~ - c~), s_R = max(q_R + a_R, q~ + c~) with Roe averages.
 * Used only by the positivity safeguard (an extension; the reference's HLL uses the
 * unbounded speeds q~ -/+ d, which are not positivity preserving). */
static int hlle_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    if (!(cl[0] > 0.0) || !(cr[0] > 0.0)) {
        for (int k = 0; k < 4; k++) F[k] = NAN;
        return 1;
    }
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    const double gr = g / (g - 1.0);
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double aL = sqrt(fmax(g * L.p * L.ir, 0.0));
    double aR = sqrt(fmax(g * R.p * R.ir, 0.0));
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double sL = sqrt(L.r), sR = sqrt(R.r);
    double iss = 1.0 / (sL + sR);
    double ubar = (sL * L.u + sR * R.u) * iss;
    double vbar = (sL * L.v + sR * R.v) * iss;
    double Hbar = (sL * HL + sR * HR) * iss;
    double cbar = sqrt(fmax((g - 1.0) * (Hbar - 0.5 * (ubar * ubar + vbar * vbar)), 0.0));
    double qL = dir == 1 ? L.u : L.v, qR = dir == 1 ? R.u : R.v;
    double qbar = dir == 1 ? ubar : vbar;
    double sl = fmin(qL - aL, qbar - cbar);
    double sr = fmax(qR + aR, qbar + cbar);
    if (sl >= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else {
        double isd = 1.0 / (sr - sl);
        for (int k = 0; k < 4; k++) F[k] = (sr * FL[k] - sl * FR[k] + sl * sr * (cr[k] - cl[k])) * isd;
    }
    return 0;
}

/* Roe with a no-op entropy fix: include/scheme.h:1018-1322.  With hc set it is
 * RoeHC (include/scheme.h:2793-3110): the wave-strength reciprocals carry eps = 1e-12,
 * a face whose Roe density sqrt(rho_L) sqrt(rho_R) is NaN is skipped (:2936-2939; here:
 * NaN flux, return 1), and the energy flux loses 0.5 |q~ + c~| (e_R - e_L) rho~ with
 * e = p / (rho (gamma - 1)). */
static int roe_core_(double g, int dir, const double cl[4], const double cr[4], double F[4],
                     int hc);

static int roe_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    return roe_core_(g, dir, cl, cr, F, 0);
}

static int roe_core_(double g, int dir, const double cl[4], const double cr[4], double F[4],
                     int hc) {
    const double gm1 = g - 1.0;
    const double gr = g / gm1;
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double sL = sqrt(L.r), sR = sqrt(R.r);
    double ss = sL + sR;
    double iss = 1.0 / ss;
    double rbar = sL * sR;
    double ubar = (sL * L.u + sR * R.u) * iss;
    double vbar = (sL * L.v + sR * R.v) * iss;
    double Hbar = (sL * HL + sR * HR) * iss;
    double q2 = ubar * ubar + vbar * vbar;
    double cbar = sqrt(gm1 * (Hbar - 0.5 * q2));
    double nbar = dir == 1 ? ubar : vbar;
    double lam[4] = {fabs(nbar - cbar), fabs(nbar), fabs(nbar), fabs(nbar + cbar)};
    double dr = R.r - L.r, du = R.u - L.u, dv = R.v - L.v, dp = R.p - L.p;
    double cs = cbar * cbar;
    double i2cs = hc ? 1.0 / (2.0 * cs + 1e-12) : 1.0 / (2.0 * cs);
    double ics = hc ? 1.0 / (cs + 1e-12) : 1.0 / cs;
    double dn = dir == 1 ? du : dv, dt = dir == 1 ? dv : du;
    double al[4] = {(dp - rbar * cbar * dn) * i2cs, dr - dp * ics, rbar * dt,
                    (dp + rbar * cbar * dn) * i2cs};
    double ev[4][4];
    if (dir == 1) {
        double e[4][4] = {{1.0, ubar - cbar, vbar, Hbar - ubar * cbar},
                          {1.0, ubar, vbar, 0.5 * q2},
                          {0.0, 0.0, 1.0, vbar},
                          {1.0, ubar + cbar, vbar, Hbar + ubar * cbar}};
        memcpy(ev, e, sizeof e);
    } else {
        double e[4][4] = {{1.0, ubar, vbar - cbar, Hbar - vbar * cbar},
                          {1.0, ubar, vbar, 0.5 * q2},
                          {0.0, 1.0, 0.0, ubar},
                          {1.0, ubar, vbar + cbar, Hbar + vbar * cbar}};
        memcpy(ev, e, sizeof e);
    }
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]);
    for (int w = 0; w < 4; w++) {
        double term = 0.5 * lam[w] * al[w];
        for (int k = 0; k < 4; k++) F[k] -= term * ev[w][k];
    }
    if (hc) {
        if (isnan(rbar)) {
            for (int k = 0; k < 4; k++) F[k] = NAN;
            return 1;
        }
        const double igm1 = 1.0 / gm1;
        double eL = L.ir * L.p * igm1, eR = R.ir * R.p * igm1;
        F[3] = F[3] - 0.5 * lam[3] * (eR - eL) * rbar;
    }
    return 0;
}

/* HLL with the heat-capacity energy term (HLLHC): include/scheme.h:1860-2137.  Sound
 * speed sqrt(gamma p / rho), Einfeldt speeds with eta = 0.5 sqrt(rho_L rho_R) / (sum)^2
 * formed through the reciprocal of the sum; a NaN Einfeldt speed stays NaN on x faces
 * (the fallback at :1948-1953 is commented out) and falls back to the Roe sound speed
 * on y faces (:2073-2076).  The star-region energy flux adds
 * (s_L s_R / (s_R - s_L)) (e_R - e_L) rho_HLL, rho_HLL the HLL mean density. */
static int hllhc_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    const double gm1 = g - 1.0;
    const double gr = g / gm1;
    const double igm1 = 1.0 / gm1;
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    double aL = sqrt(g * L.p * L.ir);
    double aR = sqrt(g * R.p * R.ir);
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double sL = sqrt(L.r), sR = sqrt(R.r);
    double ss = sL + sR;
    double iss = 1.0 / ss;
    double ubar = (sL * L.u + sR * R.u) * iss;
    double vbar = (sL * L.v + sR * R.v) * iss;
    double Hbar = (sL * HL + sR * HR) * iss;
    double eta = 0.5 * sL * sR * iss * iss;
    double qL = dir == 1 ? L.u : L.v, qR = dir == 1 ? R.u : R.v;
    double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    if (dir == 2 && isnan(d)) d = sqrt(gm1 * (Hbar - 0.5 * (ubar * ubar + vbar * vbar)));
    double nbar = dir == 1 ? ubar : vbar;
    double sl = nbar - d, sr = nbar + d;
    double rhll = (sr * R.r - sl * L.r + FL[0] - FR[0]) / (sr - sl);
    double eL = L.p * L.ir * igm1, eR = R.p * R.ir * igm1;
    if (sl >= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else {
        double isd = 1.0 / (sr - sl);
        double slsr = sl * sr;
        for (int k = 0; k < 3; k++) F[k] = (sr * FL[k] - sl * FR[k] + slsr * (cr[k] - cl[k])) * isd;
        F[3] = (sr * FL[3] - sl * FR[3] + slsr * (cr[3] - cl[3])) * isd + (slsr * isd) * (eR - eL) * rhll;
    }
    return 0;
}

/* HLLC with the heat-capacity energy term (HLLCHC): include/scheme.h:2153-2478.  Sound
 * speed from the enthalpy, eta through (sum)^2 in both directions, no NaN fallback.
 * Branch order as the reference: s_L >= 0; s_L < 0 and s* >= 0; s* < 0 and s_R > 0;
 * otherwise the right flux (so NaN speeds take the right flux).  The star-state energy
 * fluxes add (s_L s_R / (s_R - s_L)) (e_R - e_L) times the star density. */
static int hllchc_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    const double gm1 = g - 1.0;
    const double gr = g / gm1;
    const double igm1 = 1.0 / gm1;
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double aL = sqrt((g - 1.0) * (HL - 0.5 * (L.u2 + L.v2)));
    double aR = sqrt((g - 1.0) * (HR - 0.5 * (R.u2 + R.v2)));
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double sL = sqrt(L.r), sR = sqrt(R.r);
    double ss = sL + sR;
    double iss = 1.0 / ss;
    double nbar = dir == 1 ? (sL * L.u + sR * R.u) * iss : (sL * L.v + sR * R.v) * iss;
    double eta = 0.5 * sL * sR / (ss * ss);
    double qL = dir == 1 ? L.u : L.v, qR = dir == 1 ? R.u : R.v;
    double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    double sl = nbar - d, sr = nbar + d;
    double denom = L.r * (sl - qL) - R.r * (sr - qR);
    double sstar = (R.p - L.p + L.r * qL * (sl - qL) - R.r * qR * (sr - qR)) / denom;
    double mL = L.r * (sl - qL) / (sl - sstar);
    double mR = R.r * (sr - qR) / (sr - sstar);
    double feL = L.E * L.ir + (sstar - qL) * (sstar + L.p / (L.r * (sl - qL)));
    double feR = R.E * R.ir + (sstar - qR) * (sstar + R.p / (R.r * (sr - qR)));
    double eL = L.ir * L.p * igm1, eR = R.ir * R.p * igm1;
    double isd = 1.0 / (sr - sl);
    double slsr = sl * sr;
    double ec = slsr * isd * (eR - eL);
    if (sl >= 0.0) {
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sl < 0.0 && sstar >= 0.0) {
        F[0] = FL[0] + sl * (mL - L.r);
        if (dir == 1) {
            F[1] = FL[1] + sl * (mL * sstar - L.ru);
            F[2] = FL[2] + sl * (mL * L.v - L.rv);
        } else {
            F[1] = FL[1] + sl * (mL * L.u - L.ru);
            F[2] = FL[2] + sl * (mL * sstar - L.rv);
        }
        F[3] = FL[3] + sl * (mL * feL - L.E) + ec * mL;
    } else if (sstar < 0.0 && sr > 0.0) {
        F[0] = FR[0] + sr * (mR - R.r);
        if (dir == 1) {
            F[1] = FR[1] + sr * (mR * sstar - R.ru);
            F[2] = FR[2] + sr * (mR * R.v - R.rv);
        } else {
            F[1] = FR[1] + sr * (mR * R.u - R.ru);
            F[2] = FR[2] + sr * (mR * sstar - R.rv);
        }
        F[3] = FR[3] + sr * (mR * feR - R.E) + ec * mR;
    } else {
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    }
    return 0;
}

/* Local Lax-Friedrichs (the reference's default branch): include/scheme.h:229-381. */
static int llf_(double g, int dir, const double cl[4], const double cr[4], double F[4]) {
    const double gm1 = g - 1.0;
    side_t L, R;
    side_(g, cl, &L);
    side_(g, cr, &R);
    double aL = sqrt(g * L.p * L.ir);
    double aR = sqrt(g * R.p * R.ir);
    const double gr = g / gm1;
    double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double FL[4], FR[4];
    phys_flux_(dir, &L, HL, FL);
    phys_flux_(dir, &R, HR, FR);
    double qL = dir == 1 ? L.u : L.v, qR = dir == 1 ? R.u : R.v;
    double s = max2_(fabs(qL) + aL, fabs(qR) + aR);
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]) - 0.5 * s * (cr[k] - cl[k]);
    return 0;
}

static int riemann_(const rpfvm_config *cfg, int dir, const double cl[4], const double cr[4],
                    double F[4]) {
    switch (cfg->scheme) {
        case 1: return hll_(cfg->gamma, dir, cl, cr, F);
        case 2: return hllc_(cfg->gamma, dir, cl, cr, F);
        case 3: return roe_(cfg->gamma, dir, cl, cr, F);
        case 11: return hllhc_(cfg->gamma, dir, cl, cr, F);
        case 22: return hllchc_(cfg->gamma, dir, cl, cr, F);
        case 33: return roe_core_(cfg->gamma, dir, cl, cr, F, 1);
        default: return llf_(cfg->gamma, dir, cl, cr, F);
    }
}

static int check_cfg_(const rpfvm_config *c) {
    if (c->nx < 1 || c->ny < 1) return RPFVM_ERR_INVALID;
    if (c->recon_order != 1 && c->recon_order != 2 && c->recon_order != 3 && c->recon_order != 5)
        return RPFVM_ERR_INVALID;
    if (c->time_order < 1 || c->time_order > 3) return RPFVM_ERR_INVALID;
    /* Scheme 5 (the RS_Marquina stub) is not restated here. */
    if (c->scheme == 5) return RPFVM_ERR_UNSUPPORTED;
    for (int s = 0; s < 4; s++)
        if (c->bc[s] < 0 || c->bc[s] > 6) return RPFVM_ERR_INVALID;
    /* the double-Mach-reflection sides exist on the bottom / top only */
    for (int s = 0; s < 4; s++)
        if ((c->bc[s] == RPFVM_BC_DMR_WALL && s != 2) || (c->bc[s] == RPFVM_BC_DMR_SHOCK && s != 3))
            return RPFVM_ERR_INVALID;
    return 0;
}

/* Cell-centre x of row i for the double-Mach-reflection sides, in the form of
 * BC_DoubleMach_2D (include/Boundary_Condition.h:564): ((i + 0.5 - GC) / nx) * Lx. */
static double dmr_x_(const rpfvm_config *cfg, int i) {
    const double lx = cfg->lx > 0.0 ? cfg->lx : cfg->nx * cfg->dx;
    return (double)(i + 0.5 - GC) / (double)cfg->nx * lx;
}

/* The bottom and top sides of BC_DoubleMach_2D (include/Boundary_Condition.h:558-660)
 * at time t: the bottom is the fixed state up to the wall start and a mirrored
 * reflection beyond it; the top is the fixed state left of the moving shock trace and
 * the second state right of it.  Rows i = GC .. R-GC-1 only, as the reference. */
static void bc_dmr_(const rpfvm_config *cfg, double *x, int side, double t) {
    const dims_t d = dims_of(cfg);
    const double *p = cfg->bc_par[side];
    for (int i = GC; i < d.R - GC; i++) {
        const double xp = dmr_x_(cfg, i);
        for (int v = 0; v < 4; v++) {
            for (int m = 0; m < 4; m++) {
                if (side == 2) {
                    if (xp > p[0]) {
                        const double s = AT(x, v, i, GC + m);
                        AT(x, v, i, GC - 1 - m) = v == 2 ? -s : s;
                    } else {
                        AT(x, v, i, GC - 1 - m) = cfg->fixed_state[2][v];
                    }
                } else {
                    const double xs = p[0] + (p[1] / p[2]) + (p[3] * t / p[2]);
                    AT(x, v, i, d.C - GC + m) = xp < xs ? cfg->fixed_state[3][v] : cfg->bc_alt_state[3][v];
                }
            }
        }
    }
}

/* ------------------------------------------------------------ ghost-cell fill */
/* One side of Boundary_Conditions; index maps of include/Boundary_Condition.h:
 * fixed 63-135, inflow 148-202, outflow 216-270, periodic 283-336, reflection
 * 350-465 (left/top mirrored, right/bottom copy the first interior cell). */
static void bc_side_(const rpfvm_config *cfg, double *x, int side, double time) {
    const dims_t d = dims_of(cfg);
    const int R = d.R, C = d.C;
    const int type = cfg->bc[side];
    if (type == RPFVM_BC_DMR_WALL || type == RPFVM_BC_DMR_SHOCK) {
        bc_dmr_(cfg, x, side, time);
        return;
    }
    const int lr = side < 2;                       /* left/right walk k over columns */
    const int lo = GC;
    const int hi = type == RPFVM_BC_OUTFLOW ? (lr ? C - GC : R - GC)   /* inclusive */
                                            : (lr ? C - GC - 1 : R - GC - 1);
    for (int v = 0; v < 4; v++) {
        for (int t = lo; t <= hi; t++) {
            for (int m = 0; m < 4; m++) {
                /* ghost index g and source index s along the normal direction */
                int gi, si = -1;
                double val = 0.0, sign = 1.0;
                int normal_var = lr ? 1 : 2;
                switch (side) {
                    case 0: gi = GC - 1 - m; break;
                    case 1: gi = R - GC + m; break;
                    case 2: gi = GC - 1 - m; break;
                    default: gi = C - GC + m; break;
                }
                switch (type) {
                    case RPFVM_BC_OUTFLOW:
                        si = side == 0 ? GC : side == 1 ? R - GC - 1 : side == 2 ? GC : C - GC - 1;
                        break;
                    case RPFVM_BC_PERIODICITY:
                        si = side == 0 ? R - GC - 1 - m : side == 1 ? GC + m
                           : side == 2 ? C - GC - 1 - m : GC + m;
                        break;
                    case RPFVM_BC_REFLECTION:
                        si = side == 0 ? GC + m : side == 1 ? R - GC - 1
                           : side == 2 ? GC : C - GC - 1 - m;
                        if (v == normal_var) sign = -1.0;
                        break;
                    case RPFVM_BC_INFLOW:
                        val = cfg->inflow_state[v];
                        break;
                    default:
                        val = cfg->fixed_state[side][v];
                        break;
                }
                if (si >= 0) {
                    double s = lr ? AT(x, v, si, t) : AT(x, v, t, si);
                    val = sign < 0 ? -s : s;
                }
                if (lr) AT(x, v, gi, t) = val;
                else AT(x, v, t, gi) = val;
            }
        }
    }
}

int orc_boundary_conditions_t(const rpfvm_config *cfg, double *U, double t) {
    int e = check_cfg_(cfg);
    if (e) return e;
    for (int side = 0; side < 4; side++) bc_side_(cfg, U, side, t);
    return 0;
}

int orc_boundary_conditions(const rpfvm_config *cfg, double *U) {
    return orc_boundary_conditions_t(cfg, U, 0.0);
}

/* ---------------------------------------------------------------- dt (CFL) */
double orc_delta_t(const rpfvm_config *cfg, const double *U) {
    const dims_t d = dims_of(cfg);
    const double g = cfg->gamma;
    double Sx = 0.0, Sy = 0.0;
#pragma omp parallel for collapse(2) reduction(max : Sx, Sy)
    for (int i = GC; i < d.R - GC; i++) {
        for (int j = GC; j < d.C - GC; j++) {
            double rho = AT(U, 0, i, j);
            double u = AT(U, 1, i, j) / rho;
            double v = AT(U, 2, i, j) / rho;
            double p = (AT(U, 3, i, j) - 0.5 * rho * (u * u + v * v)) * (g - 1);
            double a = sqrt(g * p / rho);
            double sx = fabs(u) + a, sy = fabs(v) + a;
            if (sx > Sx) Sx = sx;
            if (sy > Sy) Sy = sy;
        }
    }
    return cfg->cfl * min2_(cfg->dx, cfg->dy) / (Sx + Sy);
}

/* ------------------------------------------------------------ spatial operator */
static int fluxes_(const rpfvm_config *cfg, const double *U, double *F, double *G) {
    const dims_t d = dims_of(cfg);
    int bad = 0;
    /* x faces i+1/2, i = GC-1 .. R-GC-1 (include/Reconstruction.h:916-917) */
#pragma omp parallel for collapse(2) reduction(| : bad)
    for (int i = GC - 1; i < d.R - GC; i++) {
        for (int j = GC; j < d.C - GC; j++) {
            double st[6][4], UL[4], UR[4], f[4];
            for (int n = 0; n < 6; n++)
                for (int v = 0; v < 4; v++) st[n][v] = AT(U, v, i - 2 + n, j);
            face_states_(cfg, 1, st, UL, UR);
            bad |= riemann_(cfg, 1, UL, UR, f);
            for (int v = 0; v < 4; v++) AT(F, v, i, j) = f[v];
        }
    }
    /* y faces j+1/2, j = GC-1 .. C-GC-1 (include/Reconstruction.h:976-977) */
#pragma omp parallel for collapse(2) reduction(| : bad)
    for (int i = GC; i < d.R - GC; i++) {
        for (int j = GC - 1; j < d.C - GC; j++) {
            double st[6][4], UL[4], UR[4], f[4];
            for (int n = 0; n < 6; n++)
                for (int v = 0; v < 4; v++) st[n][v] = AT(U, v, i, j - 2 + n);
            face_states_(cfg, 2, st, UL, UR);
            bad |= riemann_(cfg, 2, UL, UR, f);
            for (int v = 0; v < 4; v++) AT(G, v, i, j) = f[v];
        }
    }
    return bad ? RPFVM_FAULT_HLL_BAD_STATE : 0;
}

/* Source_Gravity (include/scheme.h:3723-3751) on the interior of S (zeros elsewhere). */
static void source_(const rpfvm_config *cfg, const double *U, double *S) {
    const dims_t d = dims_of(cfg);
    for (int i = GC; i <= d.R - GC - 1; i++)
        for (int j = GC; j <= d.C - GC - 1; j++) {
            double rho = AT(U, 0, i, j);
            double v = AT(U, 2, i, j) / rho;
            AT(S, 0, i, j) = 0.0;
            AT(S, 1, i, j) = 0.0;
            AT(S, 2, i, j) = rho * cfg->gravity;
            AT(S, 3, i, j) = rho * cfg->gravity * v;
        }
}

/* Space_Discrete_Item (include/scheme.h:3758-3777) on the interior. */
static void residual_(const rpfvm_config *cfg, const double *F, const double *G,
                      const double *S, double *L) {
    const dims_t d = dims_of(cfg);
    const double idx = 1.0 / cfg->dx, idy = 1.0 / cfg->dy;
#pragma omp parallel for collapse(2)
    for (int i = GC; i <= d.R - GC - 1; i++)
        for (int j = GC; j <= d.C - GC - 1; j++)
            for (int k = 0; k < 4; k++)
                AT(L, k, i, j) = -(AT(F, k, i, j) - AT(F, k, i - 1, j)) * idx
                                 - (AT(G, k, i, j) - AT(G, k, i, j - 1)) * idy
                                 + AT(S, k, i, j);
}

typedef struct {
    size_t n;
    double *F, *G, *S, *L;
} work_t;

static int work_alloc_(const rpfvm_config *cfg, work_t *w) {
    const dims_t d = dims_of(cfg);
    w->n = 4 * (size_t)d.R * (size_t)d.C;
    w->F = calloc(w->n, sizeof(double));
    w->G = calloc(w->n, sizeof(double));
    w->S = calloc(w->n, sizeof(double));
    w->L = calloc(w->n, sizeof(double));
    if (!w->F || !w->G || !w->S || !w->L) {
        free(w->F); free(w->G); free(w->S); free(w->L);
        return RPFVM_ERR_NOMEM;
    }
    return 0;
}

static void work_free_(work_t *w) {
    free(w->F); free(w->G); free(w->S); free(w->L);
}

/* One stage operator on Y: source, ghost fill in place, fluxes, residual into w->L. */
static int stage_operator_(const rpfvm_config *cfg, double *Y, work_t *w, double t) {
    if (cfg->source) source_(cfg, Y, w->S);
    for (int side = 0; side < 4; side++) bc_side_(cfg, Y, side, t);
    int bad = fluxes_(cfg, Y, w->F, w->G);
    residual_(cfg, w->F, w->G, w->S, w->L);
    return bad;
}

int orc_space_operator(const rpfvm_config *cfg, double *U, double *L, double *F, double *G) {
    int e = check_cfg_(cfg);
    if (e) return e;
    work_t w;
    if (work_alloc_(cfg, &w)) return RPFVM_ERR_NOMEM;
    int bad = stage_operator_(cfg, U, &w, 0.0);
    if (L) memcpy(L, w.L, w.n * sizeof(double));
    if (F) memcpy(F, w.F, w.n * sizeof(double));
    if (G) memcpy(G, w.G, w.n * sizeof(double));
    work_free_(&w);
    return bad;
}

/* --------------------------------------------------------------- time advance */
enum { K_EULER = 1, K_RK2_2 = 2, K_RK3_2 = 3, K_RK3_3 = 4 };

/* Stage combine of RK1/RK2/RK3 (include/time_advance.h:697, 757-777, 832-872). */
static inline double combine_(int kind, double un, double y, double l, double dt) {
    switch (kind) {
        case K_EULER: return y + dt * l;
        case K_RK2_2: return 0.5 * un + 0.5 * (y + dt * l);
        case K_RK3_2: return (3.0 / 4.0) * un + (1.0 / 4.0) * (y + dt * l);
        default: return (1.0 / 3.0) * un + (2.0 / 3.0) * (y + dt * l);
    }
}

static long g_repairs = 0;  /* cells repaired by the positivity safeguard, last step */

long orc_last_repairs(void) { return g_repairs; }

/* Residual, forward-Euler state W = Y + dt L and the stage output of one cell. */
static int cell_update_(const rpfvm_config *cfg, const double *Un, const double *Y, double *out,
                        const work_t *w, int kind, double dt, int i, int j) {
    const dims_t d = dims_of(cfg);
    const double idx = 1.0 / cfg->dx, idy = 1.0 / cfg->dy;
    double W[4];
    for (int k = 0; k < 4; k++) {
        double l = -(AT(w->F, k, i, j) - AT(w->F, k, i - 1, j)) * idx
                   - (AT(w->G, k, i, j) - AT(w->G, k, i, j - 1)) * idy + AT(w->S, k, i, j);
        AT(w->L, k, i, j) = l;
        W[k] = AT(Y, k, i, j) + dt * l;
        AT(out, k, i, j) = combine_(kind, AT(Un, k, i, j), AT(Y, k, i, j), l, dt);
    }
    return admissible_(cfg->gamma, W);
}

/* Positivity safeguard, an extension the reference does not have (a posteriori, cell
 * based): a cell whose forward-Euler state W = Y + dt L has rho <= 0 or p <= 0 gets all
 * four of its face fluxes replaced by the first-order HLL flux of the adjacent cells,
 * and every cell touching a replaced face is updated again; cells that become
 * inadmissible are repaired in the next round (at most 8 rounds).  Every step here is
 * order-independent, so the device path reproduces it exactly. */
static int repair_(const rpfvm_config *cfg, const double *Un, const double *Y, double *out,
                   work_t *w, int kind, double dt, unsigned char *state) {
    const dims_t d = dims_of(cfg);
    const int nx = cfg->nx, ny = cfg->ny;
    const size_t ncell = (size_t)nx * ny;
    /* state bit 1: in current list, bit 2: repaired, bit 4: touched this round */
    int *list = malloc(sizeof(int) * ncell), *next = malloc(sizeof(int) * ncell);
    int *touched = malloc(sizeof(int) * 5 * ncell);
    int nlist = 0, faults = 0;
    for (int i = GC; i < d.R - GC; i++)
        for (int j = GC; j < d.C - GC; j++) {
            double W[4];
            for (int k = 0; k < 4; k++) W[k] = AT(Y, k, i, j) + dt * AT(w->L, k, i, j);
            if (!admissible_(cfg->gamma, W)) list[nlist++] = (i - GC) * ny + (j - GC);
        }
    int rounds = 0;
    while (nlist > 0 && rounds < 8) {
        for (int q = 0; q < nlist; q++) {
            const int i = GC + list[q] / ny, j = GC + list[q] % ny;
            state[list[q]] |= 2;
            double a[4], b[4], f[4];
            for (int side = 0; side < 4; side++) {
                /* faces [i-1][j], [i][j] (x) and [i][j-1], [i][j] (y) */
                const int dir = side < 2 ? 1 : 2;
                const int fi = side == 0 ? i - 1 : i, fj = side == 2 ? j - 1 : j;
                for (int k = 0; k < 4; k++) {
                    a[k] = AT(Y, k, fi, fj);
                    b[k] = dir == 1 ? AT(Y, k, fi + 1, fj) : AT(Y, k, fi, fj + 1);
                }
                faults |= hlle_(cfg->gamma, dir, a, b, f);
                for (int k = 0; k < 4; k++) {
                    if (dir == 1) AT(w->F, k, fi, fj) = f[k];
                    else AT(w->G, k, fi, fj) = f[k];
                }
            }
        }
        int nt = 0;
        for (int q = 0; q < nlist; q++) {
            const int ci = list[q] / ny, cj = list[q] % ny;
            const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
            for (int m = 0; m < 5; m++) {
                const int ti = ci + di[m], tj = cj + dj[m];
                if (ti < 0 || ti >= nx || tj < 0 || tj >= ny) continue;
                const int c = ti * ny + tj;
                if (!(state[c] & 4)) {
                    state[c] |= 4;
                    touched[nt++] = c;
                }
            }
        }
        int nnext = 0;
        for (int q = 0; q < nt; q++) {
            const int c = touched[q];
            state[c] &= ~4;
            const int ok = cell_update_(cfg, Un, Y, out, w, kind, dt, GC + c / ny, GC + c % ny);
            if (!ok && !(state[c] & 2)) next[nnext++] = c;
            else if (!ok) faults |= 8;
        }
        g_repairs += nlist;
        int *t = list;
        list = next;
        next = t;
        nlist = nnext;
        rounds++;
    }
    if (nlist > 0) faults |= 8;
    free(list);
    free(next);
    free(touched);
    return faults ? (faults & 1 ? RPFVM_FAULT_HLL_BAD_STATE : 0) | (faults & 8 ? RPFVM_FAULT_POSITIVITY : 0) : 0;
}

/* One RK stage: out = combine(kind, Un, Y, dt L(Y)).  out may alias Un. */
static int stage_(const rpfvm_config *cfg, const double *Un, double *Y, double *out, int kind,
                  double dt, double t, work_t *w, double *tmp, unsigned char *state) {
    const dims_t d = dims_of(cfg);
    int bad = stage_operator_(cfg, Y, w, t);
    double *dst = cfg->positivity ? tmp : out;
#pragma omp parallel for collapse(2)
    for (int i = GC; i <= d.R - GC - 1; i++)
        for (int j = GC; j <= d.C - GC - 1; j++)
            for (int k = 0; k < 4; k++)
                AT(dst, k, i, j) = combine_(kind, AT(Un, k, i, j), AT(Y, k, i, j), AT(w->L, k, i, j), dt);
    if (cfg->positivity) {
        memset(state, 0, (size_t)cfg->nx * cfg->ny);
        bad |= repair_(cfg, Un, Y, dst, w, kind, dt, state);
        for (int k = 0; k < 4; k++)
            for (int i = GC; i <= d.R - GC - 1; i++)
                memcpy(&AT(out, k, i, GC), &AT(dst, k, i, GC), sizeof(double) * cfg->ny);
    }
    return bad;
}

int orc_time_advance(const rpfvm_config *cfg, double *U, double dt) {
    return orc_time_advance_t(cfg, U, dt, 0.0);
}

/* Stage k fills its ghosts at t + c_k dt: SSP-RK3 abscissae c = (0, 1, 1/2), RK2 (0, 1). */
int orc_time_advance_t(const rpfvm_config *cfg, double *U, double dt, double t) {
    int e = check_cfg_(cfg);
    if (e) return e;
    work_t w;
    if (work_alloc_(cfg, &w)) return RPFVM_ERR_NOMEM;
    double *U1 = calloc(w.n, sizeof(double)), *U2 = calloc(w.n, sizeof(double));
    double *tmp = cfg->positivity ? calloc(w.n, sizeof(double)) : NULL;
    unsigned char *state = cfg->positivity ? calloc((size_t)cfg->nx * cfg->ny, 1) : NULL;
    if (!U1 || !U2 || (cfg->positivity && (!tmp || !state))) {
        free(U1); free(U2); free(tmp); free(state); work_free_(&w);
        return RPFVM_ERR_NOMEM;
    }
    int bad = 0;
    g_repairs = 0;
    /* RK1_TimeAd :643-708, RK2_TimeAd :711-785, RK3_TimeAd :788-882 */
    if (cfg->time_order == 1) {
        bad |= stage_(cfg, U, U, U, K_EULER, dt, t, &w, tmp, state);
    } else if (cfg->time_order == 2) {
        bad |= stage_(cfg, U, U, U1, K_EULER, dt, t, &w, tmp, state);
        bad |= stage_(cfg, U, U1, U, K_RK2_2, dt, t + dt, &w, tmp, state);
    } else {
        bad |= stage_(cfg, U, U, U1, K_EULER, dt, t, &w, tmp, state);
        bad |= stage_(cfg, U, U1, U2, K_RK3_2, dt, t + dt, &w, tmp, state);
        bad |= stage_(cfg, U, U2, U, K_RK3_3, dt, t + 0.5 * dt, &w, tmp, state);
    }
    free(U1);
    free(U2);
    free(tmp);
    free(state);
    work_free_(&w);
    return bad;
}

int orc_advance_to(const rpfvm_config *cfg, double *U, double t_stop, int max_steps,
                   double *t, double *dts) {
    int e = check_cfg_(cfg);
    if (e) return e;
    double time = *t;
    int n = 0;
    /* src/RP_FVM.c:179-191 with next_output_time == Tmax == t_stop */
    while (time < t_stop && n < max_steps) {
        double dt = orc_delta_t(cfg, U);
        if (time + dt > t_stop) dt = t_stop - time;
        if (time + dt >= t_stop) dt = t_stop - time;
        if (dts) dts[n] = dt;
        orc_time_advance_t(cfg, U, dt, time);
        time = time + dt;
        n++;
    }
    *t = time;
    return n;
}

/* ------------------------------------------------------------- unit entries */
int orc_reconstruct(const rpfvm_config *cfg, const double *U, int dir, double *CL, double *CR) {
    int e = check_cfg_(cfg);
    if (e) return e;
    const dims_t d = dims_of(cfg);
    int i0 = dir == 1 ? GC - 1 : GC, i1 = d.R - GC;
    int j0 = dir == 1 ? GC : GC - 1, j1 = d.C - GC;
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++) {
            double st[6][4], UL[4], UR[4];
            for (int n = 0; n < 6; n++)
                for (int v = 0; v < 4; v++)
                    st[n][v] = dir == 1 ? AT(U, v, i - 2 + n, j) : AT(U, v, i, j - 2 + n);
            reconstruct_face_(cfg, dir, st, UL, UR);
            for (int v = 0; v < 4; v++) {
                AT(CL, v, i, j) = UL[v];
                AT(CR, v, i, j) = UR[v];
            }
        }
    return 0;
}

int orc_riemann(const rpfvm_config *cfg, int dir, const double *Ls, const double *Rs, double *F) {
    int e = check_cfg_(cfg);
    if (e) return e;
    const dims_t d = dims_of(cfg);
    int i0 = dir == 1 ? GC - 1 : GC, i1 = d.R - GC;
    int j0 = dir == 1 ? GC : GC - 1, j1 = d.C - GC;
    int bad = 0;
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++) {
            double cl[4], cr[4], f[4];
            for (int v = 0; v < 4; v++) {
                cl[v] = AT(Ls, v, i, j);
                cr[v] = AT(Rs, v, i, j);
            }
            int b = riemann_(cfg, dir, cl, cr, f);
            bad |= b;
            if (!b)
                for (int v = 0; v < 4; v++) AT(F, v, i, j) = f[v];
        }
    return bad ? RPFVM_FAULT_HLL_BAD_STATE : 0;
}

int orc_eigen(const rpfvm_config *cfg, const double *U, int dir, double *EL, double *ER) {
    const dims_t d = dims_of(cfg);
    const size_t plane = (size_t)d.R * d.C;
    /* face ranges of Compute_Eigen_2DX (:325-326) and 2DY (:427-428, j inclusive of
     * C-GC; that last face reads a cell Con_to_Pri_2D never fills and is skipped here) */
    int i0 = dir == 1 ? GC - 1 : GC, i1 = d.R - GC;
    int j0 = dir == 1 ? GC : GC - 1, j1 = d.C - GC;
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++) {
            double ca[4], cb[4], pa[4], pb[4], L4[4][4], R4[4][4];
            for (int v = 0; v < 4; v++) {
                ca[v] = AT(U, v, i, j);
                cb[v] = dir == 1 ? AT(U, v, i + 1, j) : AT(U, v, i, j + 1);
            }
            prim_(cfg->gamma, ca, pa);
            prim_(cfg->gamma, cb, pb);
            face_eigen_(cfg->gamma, dir, pa, pb, L4, R4);
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) {
                    EL[((size_t)a * 4 + b) * plane + (size_t)i * d.C + j] = L4[a][b];
                    ER[((size_t)a * 4 + b) * plane + (size_t)i * d.C + j] = R4[a][b];
                }
        }
    return 0;
}
