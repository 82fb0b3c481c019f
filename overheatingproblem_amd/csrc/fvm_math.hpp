// fvm_math.hpp -- FP64 point math of the 2-D Euler FV step on gfx950.
//
// Every expression is written so that, compiled with -ffp-contract=off (no FMA
// contraction) and IEEE division/sqrt, it rounds exactly like the reference's
// x86-64 build (-O3, SSE2, no FMA).  Each function names the reference routine
// whose arithmetic it reproduces.
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_kernels.hpp"

namespace fvm {

// ------------------------------------------------------ shared-denominator division
// gfx950's IEEE FP64 division a/b is the sequence
//   d = div_scale(b), n = div_scale(a), y0 = rcp(d), y1 = y0 + y0(1 - d y0),
//   y2 = y1 + y1(1 - d y1), q = n y2, r = n - d q, div_fmas(r, y2, q), div_fixup.
// div_scale rescales only near the exponent limits (exponent difference >= 768,
// numerator within 2^53 of the denormals, denominator or quotient denormal), and
// div_fixup only changes special or out-of-range results.  For |a|, |b| in
// [2^-300, 2^300] neither applies, so a/b == fma(fma(-b, a*y2, a), y2, a*y2) exactly.
// Recip holds y2 of one denominator; quot() finishes each numerator with 3 FP64 ops
// (and falls back to a true division outside the safe range), so quotients sharing a
// denominator cost one reciprocal refinement instead of one full sequence each.
struct Recip {
    double b, y;
    bool ok;
};

__device__ __forceinline__ bool in_safe_range(double x) {
    const double ax = fabs(x);
    return ax >= 0x1p-300 && ax <= 0x1p300;
}

__device__ __forceinline__ Recip recip(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    return Recip{b, y, in_safe_range(b)};
}

__device__ __forceinline__ double quot(double a, const Recip& r) {
    if (r.ok && in_safe_range(a)) {
        const double q = a * r.y;
        const double rem = fma(-r.b, q, a);
        return fma(rem, r.y, q);
    }
    return a / r.b;
}

// IEEE square root of x in [2^-700, 2^700]: the refinement of the hardware rsq estimate
// (one Goldschmidt step, two Newton corrections of the root) that the compiler's own
// double sqrt applies to an operand it does not need to rescale (x >= 2^-767) and that is
// neither zero nor infinite; bit-identical to sqrt(x) there.
__device__ __forceinline__ double sqrt_mid(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// Quotients in groups.  The per-quotient test of quot() costs each quotient a divergent
// branch (an exec-mask save, a skip and a restore: about as many scalar instructions as
// the quotient has vector ones).  A group of quotients (one WENO weight combination, the
// primitives of a cell, the Roe-averaged basis of a face) instead runs the fast path of
// every quotient unconditionally with Div<true>, which records whether each was inside
// the exact range; one vote of the wavefront then keeps the group, or reruns it with
// Div<false> (quot() per quotient) when any lane left the range.  Either way every lane
// gets exactly the IEEE quotients.
template <bool FAST>
struct Div {
    bool ok = true;
    __device__ __forceinline__ double operator()(double a, const Recip& r) {
        if constexpr (FAST) {
            ok = ok && r.ok && in_safe_range(a);
            const double q = a * r.y;
            const double rem = fma(-r.b, q, a);
            return fma(rem, r.y, q);
        } else {
            return quot(a, r);
        }
    }
    // sqrt(x), on the fast path through sqrt_mid for x in [2^-700, 2^700].
    __device__ __forceinline__ double root(double x) {
        if constexpr (FAST) {
            ok = ok && x >= 0x1p-700 && x <= 0x1p700;
            return sqrt_mid(x);
        } else {
            return sqrt(x);
        }
    }
    // As operator(), except that the fast path also takes a = +-0 and then returns a zero
    // whose sign can be wrong (-0 / b gives +0 for b > 0).  Only for quotients whose
    // zeros cannot reach a result with their sign (see reconstruct_face).
    __device__ __forceinline__ double zq(double a, const Recip& r) {
        if constexpr (FAST) {
            ok = ok && r.ok && (in_safe_range(a) || a == 0.0);
            const double q = a * r.y;
            const double rem = fma(-r.b, q, a);
            return fma(rem, r.y, q);
        } else {
            return quot(a, r);
        }
    }
};

// The refined reciprocal and the finished quotient of recip() / quot() without their
// range tests, for groups whose ranges are established as a whole.
__device__ __forceinline__ double rcp2(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    return fma(y, e, y);
}
__device__ __forceinline__ double qfin(double a, double b, double y) {
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}

// True when `ok` holds in every active lane of the wavefront (a uniform value).
__device__ __forceinline__ bool all_lanes(bool ok) { return __ballot(!ok) == 0ull; }

// ------------------------------------------------------------------- helpers
// sgn / max_of_two / min_of_two / min_mod / van_leer: include/function.h:31-74
__device__ __forceinline__ int sgn(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
__device__ __forceinline__ double max2(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double min2(double a, double b) { return a > b ? b : a; }
__device__ __forceinline__ double minmod(double a, double b) {
    return 0.5 * (sgn(a) + sgn(b)) * min2(fabs(a), fabs(b));
}
__device__ __forceinline__ double vanleer(double a, double b) {
    return ((sgn(a) + sgn(b)) * a * b) / (fabs(a) + fabs(b) + 1.0e-6);
}

// ---------------------------------------------------- scalar reconstructions
// Smoothness indicators shared by WENO5-JS and WENO5-Z (Jiang & Shu 1996),
// include/Reconstruction.h:1426-1433 and :1537-1544.
struct Beta5 {
    double b1, b2, b3;
};
__device__ __forceinline__ Beta5 beta5(double v1, double v2, double v3, double v4, double v5) {
    Beta5 r;
    r.b1 = 13.0 / 12.0 * (v1 - 2.0 * v2 + v3) * (v1 - 2.0 * v2 + v3)
         + 0.25 * (v1 - 4.0 * v2 + 3.0 * v3) * (v1 - 4.0 * v2 + 3.0 * v3);
    r.b2 = 13.0 / 12.0 * (v2 - 2.0 * v3 + v4) * (v2 - 2.0 * v3 + v4)
         + 0.25 * (v2 - v4) * (v2 - v4);
    r.b3 = 13.0 / 12.0 * (v3 - 2.0 * v4 + v5) * (v3 - 2.0 * v4 + v5)
         + 0.25 * (3.0 * v3 - 4.0 * v4 + v5) * (3.0 * v3 - 4.0 * v4 + v5);
    return r;
}

// Weighted combination of the three third-order candidates, :1567-1569.  r6 is the
// Recip of 6.0 (built from a runtime value so the compiler cannot fold the hardware
// reciprocal estimate).
template <bool FAST>
__device__ __forceinline__ double weno5_combine_t(double a1, double a2, double a3, double v1,
                                                  double v2, double v3, double v4, double v5,
                                                  const Recip& r6, Div<FAST>& q) {
    const Recip rs = recip(a1 + a2 + a3);
    const double w1 = q(a1, rs);
    const double w2 = q(a2, rs);
    const double w3 = q(a3, rs);
    return q(w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3), r6)
         + q(w2 * (-v2 + 5.0 * v3 + 2.0 * v4), r6)
         + q(w3 * (2.0 * v3 + 5.0 * v4 - v5), r6);
}

__device__ __forceinline__ double weno5_combine(double a1, double a2, double a3, double v1,
                                                double v2, double v3, double v4, double v5,
                                                const Recip& r6) {
    Div<true> f;
    const double r = weno5_combine_t(a1, a2, a3, v1, v2, v3, v4, v5, r6, f);
    if (all_lanes(f.ok)) return r;
    Div<false> c;
    return weno5_combine_t(a1, a2, a3, v1, v2, v3, v4, v5, r6, c);
}

// WENO5-Z (Borges et al. 2008, exponent 1, eps 1e-15): WENO5Z_L/R,
// include/Reconstruction.h:1519-1624.  (v1..v5) in upwind order.
//
// FAST = false is the reference's arithmetic (IEEE quotients).  FAST = true gives the same
// value whenever it leaves q.ok set, and the caller reruns FAST = false when any lane
// cleared it.  It takes the caller's word (q.ok on entry) that |v| <= 2^120, from which
// every range it needs follows:
//  * 2 v and 4 v are exact, so fma(-2, v2, v1) rounds like v1 - 2.0 * v2 (one operation
//    instead of two, 8 per value);
//  * beta is a sum of squares, 0 <= beta <= 34 (2^120)^2 < 2^246, so d = beta + eps lies
//    in [1e-15, 2^246];
//  * tau = |beta1 - beta3| <= 2^246: tau / d is exact for tau >= 2^-300, and below that
//    any value under 2^-54 gives the reference's 1.0 + tau / d == 1.0, as the fast
//    quotient does (it is at most 2^-249);
//  * each alpha is at least 0.1 (0.1 (1 + x), x >= 0) and the sum s at most
//    1 + 2^246 / 1e-15 < 2^296, so the weights alpha / s are exact;
//  * each weight is at least 0.1 / 2^296 and each candidate polynomial at most 20 2^120
//    in magnitude, so a sixth t / 6 is exact for t = 0 (up to the sign, Div::zq: WENO5-Z
//    runs in the characteristic branch only) or |t| >= 2^-300: the one test left.
template <bool FAST>
__device__ __forceinline__ double weno5z_t(double v1, double v2, double v3, double v4, double v5,
                                           const Recip& r6, Div<FAST>& q) {
    const double eps = 1.0e-15;
    if constexpr (!FAST) {
        const Beta5 b = beta5(v1, v2, v3, v4, v5);
        const double tau = fabs(b.b1 - b.b3);
        const double a1 = 0.1 * (1.0 + tau / (b.b1 + eps));
        const double a2 = 0.6 * (1.0 + tau / (b.b2 + eps));
        const double a3 = 0.3 * (1.0 + tau / (b.b3 + eps));
        return weno5_combine_t(a1, a2, a3, v1, v2, v3, v4, v5, r6, q);
    } else {
        const double t1 = fma(-2.0, v2, v1) + v3;
        const double s1 = fma(-4.0, v2, v1) + 3.0 * v3;
        const double t2 = fma(-2.0, v3, v2) + v4;
        const double t3 = fma(-2.0, v4, v3) + v5;
        const double s3 = fma(-4.0, v4, 3.0 * v3) + v5;
        const double b1 = 13.0 / 12.0 * t1 * t1 + 0.25 * s1 * s1;
        const double b2 = 13.0 / 12.0 * t2 * t2 + 0.25 * (v2 - v4) * (v2 - v4);
        const double b3 = 13.0 / 12.0 * t3 * t3 + 0.25 * s3 * s3;
        const double tau = fabs(b1 - b3);
        const double d1 = b1 + eps, d2 = b2 + eps, d3 = b3 + eps;
        const double a1 = 0.1 * (1.0 + qfin(tau, d1, rcp2(d1)));
        const double a2 = 0.6 * (1.0 + qfin(tau, d2, rcp2(d2)));
        const double a3 = 0.3 * (1.0 + qfin(tau, d3, rcp2(d3)));
        const double sum = a1 + a2 + a3;
        const double ys = rcp2(sum);
        const double p1 = fma(2.0, v1, -(7.0 * v2)) + 11.0 * v3;
        const double p2 = fma(2.0, v4, -v2 + 5.0 * v3);
        const double p3 = fma(2.0, v3, 5.0 * v4) - v5;
        const double x1 = qfin(a1, sum, ys) * p1;
        const double x2 = qfin(a2, sum, ys) * p2;
        const double x3 = qfin(a3, sum, ys) * p3;
        q.ok = q.ok && (fabs(x1) >= 0x1p-300 || x1 == 0.0) && (fabs(x2) >= 0x1p-300 || x2 == 0.0) &&
               (fabs(x3) >= 0x1p-300 || x3 == 0.0);
        return qfin(x1, 6.0, r6.y) + qfin(x2, 6.0, r6.y) + qfin(x3, 6.0, r6.y);
    }
}

// The two WENO5-Z values of a characteristic field's six stencil values, each under its
// own vote.
__device__ __forceinline__ void weno5z_pair(const double s[6], const Recip& r6, double& left,
                                            double& right) {
    // |v| <= 2^120 for all six (the sum of magnitudes bounds each; a NaN or an infinity
    // fails the test)
    double m = fabs(s[0]);
#pragma unroll
    for (int n = 1; n < 6; n++) m += fabs(s[n]);
    const bool bounded = m <= 0x1p120;
    {
        Div<true> f;
        f.ok = bounded;
        left = weno5z_t(s[0], s[1], s[2], s[3], s[4], r6, f);
        if (!all_lanes(f.ok)) {
            Div<false> c;
            left = weno5z_t(s[0], s[1], s[2], s[3], s[4], r6, c);
        }
    }
    {
        Div<true> f;
        f.ok = bounded;
        right = weno5z_t(s[5], s[4], s[3], s[2], s[1], r6, f);
        if (!all_lanes(f.ok)) {
            Div<false> c;
            right = weno5z_t(s[5], s[4], s[3], s[2], s[1], r6, c);
        }
    }
}

// WENO5-JS: WENO5_L/R, include/Reconstruction.h:1408-1508.
__device__ __forceinline__ double weno5js(double v1, double v2, double v3, double v4, double v5,
                                          const Recip& r6) {
    const double eps = 1.0e-15;
    const Beta5 b = beta5(v1, v2, v3, v4, v5);
    const Recip r1 = recip(eps + b.b1), r2 = recip(eps + b.b2), r3 = recip(eps + b.b3);
    Div<true> f;
    double a1 = f(f(0.1, r1), r1);
    double a2 = f(f(0.6, r2), r2);
    double a3 = f(f(0.3, r3), r3);
    if (!all_lanes(f.ok)) {
        a1 = quot(quot(0.1, r1), r1);
        a2 = quot(quot(0.6, r2), r2);
        a3 = quot(quot(0.3, r3), r3);
    }
    return weno5_combine(a1, a2, a3, v1, v2, v3, v4, v5, r6);
}

// WENO3_L/R, include/Reconstruction.h:1326-1398 (pow(x,2) folded to x*x).
__device__ __forceinline__ double weno3(double v1, double v2, double v3) {
    const double eps = 1e-6;
    const double s1 = (v2 - v1) * (v2 - v1);
    const double s2 = (v3 - v2) * (v3 - v2);
    const double a1 = (1.0 / 3.0) / ((eps + s1) * (eps + s1));
    const double a2 = (2.0 / 3.0) / ((eps + s2) * (eps + s2));
    const double w1 = a1 / (a1 + a2);
    const double w2 = a2 / (a1 + a2);
    return w1 * (-0.5 * v1 + 1.5 * v2) + w2 * (0.5 * v2 + 0.5 * v3);
}

// TVD slopes, include/Reconstruction.h:1088-1209: minmod for the component branch,
// van Leer for the characteristic branch.
template <bool VANLEER>
__device__ __forceinline__ double tvd_slope(double v1, double v2, double v3, double delta) {
    const double s1 = (v2 - v1) / delta;
    const double s2 = (v3 - v2) / delta;
    return VANLEER ? vanleer(s1, s2) : minmod(s1, s2);
}

// Face values of one scalar stencil s[0..5] (cells i-2..i+3) for reconstruction
// order ORDER; CHAR selects WENO5-Z / van Leer (characteristic branch) over
// WENO5-JS / minmod (component branch).
template <int ORDER, bool CHAR>
__device__ __forceinline__ void recon_scalar(const double s[6], double delta, const Recip& r6,
                                             double& left, double& right) {
    if constexpr (ORDER == 1) {
        left = s[2];
        right = s[3];
    } else if constexpr (ORDER == 2) {
        const double sl = tvd_slope<CHAR>(s[1], s[2], s[3], delta);
        const double sr = tvd_slope<CHAR>(s[2], s[3], s[4], delta);
        left = s[2] + 0.5 * sl * delta;
        right = s[3] - 0.5 * sr * delta;
    } else if constexpr (ORDER == 3) {
        left = weno3(s[1], s[2], s[3]);
        right = weno3(s[4], s[3], s[2]);
    } else {
        if constexpr (CHAR) {
            weno5z_pair(s, r6, left, right);
        } else {
            left = weno5js(s[0], s[1], s[2], s[3], s[4], r6);
            right = weno5js(s[5], s[4], s[3], s[2], s[1], r6);
        }
    }
}

// ------------------------------------------------------------ eigenvectors
// Uniform constants of a kernel: gm1 = gamma - 1 and gr = gamma / (gamma - 1) come from
// the host (IEEE division rounds identically there), r6 is the Recip of 6.0.
struct Consts {
    double g, gm1, gr;
    Recip r6;
};

__device__ __forceinline__ Consts make_consts(const Phys& ph) {
    return Consts{ph.gamma, ph.gm1, ph.gr, recip(ph.six)};
}

// Con_to_Pri_2D for one cell (include/initialize.h:1443-1451) and the enthalpy
// Compute_Eigen_2DX/2DY derive from it (include/Characteriz.h:318,420); the three
// quotients by rho share one reciprocal.
template <bool FAST>
__device__ __forceinline__ void primitive_h(const Consts& k, const double c[4], double p[4],
                                            double& H, Recip& rr, Div<FAST>& q) {
    rr = recip(c[0]);
    const double rho = c[0];
    const double u = q.zq(c[1], rr);
    const double v = q.zq(c[2], rr);
    p[0] = rho;
    p[1] = u;
    p[2] = v;
    p[3] = (c[3] - 0.5 * rho * (u * u + v * v)) * k.gm1;
    H = 0.5 * (u * u + v * v) + k.gr * q(p[3], rr);
}

// The left (EL[wave][var]) and right (ER[var][wave]) eigenvectors of the face between
// cells a and b (conserved states).  DIR 1: Compute_Eigen_2DX (include/Characteriz.h:
// 300-398); DIR 2: Compute_Eigen_2DY (:402-500), with its literal 0.0/c and -0.0 terms
// (0.0/c == 0.0 * (1/c) for every c = sqrt(c2) >= +0 or NaN).
//
// The 32 entries are simple products of nine face scalars, so a face keeps the scalars
// (FaceBasis) and forms each row where it is used: a row of EL per projected wave, a row
// of ER per back-projected variable.  Every entry is the same expression as in the
// reference and is formed once per face, as when the matrices are held; holding nine
// scalars instead of 32 entries removed the fused y sweep's register spill.
//
// The velocity quotients (cell and Roe-averaged u, v and u/c or v/c) take Div::zq: a
// velocity component enters the entries only through sums with nonzero terms, squares
// and products, so the sign of a zero velocity can only change the sign of a zero entry,
// which reconstruct_face's sums cannot pass on.
struct FaceBasis {
    double u, v, H, q2, c, b1, b2, ic, wc;  // wc = u/c (DIR 1) or v/c (DIR 2)
};

template <int DIR, bool FAST>
__device__ __forceinline__ FaceBasis face_basis_t(const Consts& k, const double ca[4],
                                                  const double cb[4], Div<FAST>& quot) {
    double a[4], b[4], Ha, Hb;
    Recip ra, rb;
    primitive_h(k, ca, a, Ha, ra, quot);
    primitive_h(k, cb, b, Hb, rb, quot);
    const double D = quot.root(quot(b[0], ra));
    const double D1 = D + 1.0;
    const Recip rd = recip(D1);
    FaceBasis e;
    e.u = quot.zq(a[1] + D * b[1], rd);
    e.v = quot.zq(a[2] + D * b[2], rd);
    e.H = quot(Ha + D * Hb, rd);
    e.q2 = e.u * e.u + e.v * e.v;
    const double c2 = k.gm1 * (e.H - 0.5 * e.q2);
    e.c = quot.root(c2);
    e.b1 = k.gm1 / c2;
    e.b2 = 1.0 + e.b1 * e.q2 - e.b1 * e.H;
    const Recip rc = recip(e.c);
    e.ic = quot(1.0, rc);
    e.wc = quot.zq(DIR == 1 ? e.u : e.v, rc);
    return e;
}

template <int DIR>
__device__ __forceinline__ FaceBasis face_basis(const Consts& k, const double ca[4],
                                                const double cb[4]) {
    Div<true> f;
    const FaceBasis e = face_basis_t<DIR>(k, ca, cb, f);
    if (all_lanes(f.ok)) return e;
    Div<false> c;
    return face_basis_t<DIR>(k, ca, cb, c);
}

// Row `w` of EL (the left eigenvector of wave w); w is a constant after unrolling.
template <int DIR>
__device__ __forceinline__ void el_row(const FaceBasis& e, int w, double r[4]) {
    const double u = e.u, v = e.v, b1 = e.b1, b2 = e.b2, ic = e.ic, wc = e.wc;
    if constexpr (DIR == 1) {
        if (w == 0) {
            r[0] = 0.5 * (b2 + wc);
            r[1] = -0.5 * (b1 * u + ic);
            r[2] = -0.5 * b1 * v;
            r[3] = 0.5 * b1;
        } else if (w == 1) {
            r[0] = -e.q2 + e.H;
            r[1] = u;
            r[2] = v;
            r[3] = -1.0;
        } else if (w == 2) {
            r[0] = v;
            r[1] = 0.0;
            r[2] = -1.0;
            r[3] = 0.0;
        } else {
            r[0] = 0.5 * (b2 - wc);
            r[1] = 0.5 * (-b1 * u + ic);
            r[2] = 0.5 * -b1 * v;
            r[3] = 0.5 * b1;
        }
    } else {
        const double zc = 0.0 * ic;
        if (w == 0) {
            r[0] = 0.5 * (b2 + wc);
            r[1] = -0.5 * (b1 * u + zc);
            r[2] = -0.5 * (b1 * v + ic);
            r[3] = 0.5 * b1;
        } else if (w == 1) {
            r[0] = -e.q2 + e.H;
            r[1] = u;
            r[2] = v;
            r[3] = -1.0;
        } else if (w == 2) {
            r[0] = -u;
            r[1] = 1.0;
            r[2] = -0.0;
            r[3] = 0.0;
        } else {
            r[0] = 0.5 * (b2 - wc);
            r[1] = 0.5 * (-b1 * u + zc);
            r[2] = 0.5 * (-b1 * v + ic);
            r[3] = 0.5 * b1;
        }
    }
}

// Row `n` of ER (variable n over the four waves); n is a constant after unrolling.
template <int DIR>
__device__ __forceinline__ void er_row(const FaceBasis& e, int n, double r[4]) {
    const double u = e.u, v = e.v, H = e.H, c = e.c, b1 = e.b1;
    if (n == 0) {
        r[0] = 1.0;
        r[1] = b1;
        r[2] = 0.0;
        r[3] = 1.0;
    } else if constexpr (DIR == 1) {
        if (n == 1) {
            r[0] = u - 1.0 * c;
            r[1] = u * b1;
            r[2] = 0.0;
            r[3] = u + 1.0 * c;
        } else if (n == 2) {
            r[0] = v - 0.0 * c;
            r[1] = v * b1;
            r[2] = -1.0;
            r[3] = v + 0.0 * c;
        } else {
            r[0] = H - u * c;
            r[1] = H * b1 - 1.0;
            r[2] = -v;
            r[3] = H + u * c;
        }
    } else {
        if (n == 1) {
            r[0] = u - 0.0 * c;
            r[1] = u * b1;
            r[2] = 1.0;
            r[3] = u + 0.0 * c;
        } else if (n == 2) {
            r[0] = v - 1.0 * c;
            r[1] = v * b1;
            r[2] = -0.0;
            r[3] = v + 1.0 * c;
        } else {
            r[0] = H - v * c;
            r[1] = H * b1 - 1.0;
            r[2] = u;
            r[3] = H + v * c;
        }
    }
}

// Face states from the six-cell stencil st[n][var].  Characteristic branch:
// project w = L.U per wave (sum from 0.0 in var order), reconstruct, back-project
// U = R.w (sum from 0.0 in wave order): include/Reconstruction.h:909-967 and its
// Godunov/TVD/WENO3 twins; component branch: :1031-1073 and twins.
//
// Signs of zeros.  The back-projection's sums start from +0.0, and +0.0 + (-0) == +0.0, so
// they give the same result for terms that differ only in the signs of zeros.  Every
// reconstruction here maps values that differ only in the signs of zeros to values that
// differ only in the signs of zeros (sums, products, fabs, comparisons, and quotients by
// denominators that are never zero: beta + eps, the weight sum, 6, delta, (eps + s)^2,
// |a| + |b| + 1e-6).  Neither the zero signs of the entries of EL and ER nor those of
// the projected values w can therefore change UL or UR, which lets the projection start
// from its first product instead of 0.0 + that product, and the WENO5-Z sixths and the
// basis's velocity quotients take Div::zq.
template <int DIR, int ORDER, bool CHAR>
__device__ __forceinline__ void reconstruct_face(const Consts& kc, double delta,
                                                 const double st[6][4], double UL[4],
                                                 double UR[4]) {
    if constexpr (CHAR) {
        const FaceBasis e = face_basis<DIR>(kc, st[2], st[3]);
        double wl[4], wr[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double EL[4];
            el_row<DIR>(e, k, EL);
            if constexpr (ORDER == 1) {
                double l = st[2][0] * EL[0], r = st[3][0] * EL[0];
#pragma unroll
                for (int m = 1; m < 4; m++) {
                    l += st[2][m] * EL[m];
                    r += st[3][m] * EL[m];
                }
                wl[k] = l;
                wr[k] = r;
            } else {
                double w[6];
#pragma unroll
                for (int n = 0; n < 6; n++) {
                    double acc = st[n][0] * EL[0];
#pragma unroll
                    for (int m = 1; m < 4; m++) acc += st[n][m] * EL[m];
                    w[n] = acc;
                }
                recon_scalar<ORDER, true>(w, delta, kc.r6, wl[k], wr[k]);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double ER[4];
            er_row<DIR>(e, v, ER);
            double l = 0.0, r = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                l += wl[k] * ER[k];
                r += wr[k] * ER[k];
            }
            UL[v] = l;
            UR[v] = r;
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double s[6] = {st[0][v], st[1][v], st[2][v], st[3][v], st[4][v], st[5][v]};
            recon_scalar<ORDER, false>(s, delta, kc.r6, UL[v], UR[v]);
        }
    }
}

// ------------------------------------------------------------ Riemann solvers
struct Side {
    double r, ru, rv, E, ir, u, v, u2, v2, p;
};

__device__ __forceinline__ Side make_side(const Consts& k, const double c[4]) {
    Side s;
    s.r = c[0];
    s.ru = c[1];
    s.rv = c[2];
    s.E = c[3];
    s.ir = 1.0 / s.r;
    s.u = s.ru * s.ir;
    s.v = s.rv * s.ir;
    s.u2 = s.u * s.u;
    s.v2 = s.v * s.v;
    s.p = (s.E - 0.5 * s.r * (s.u2 + s.v2)) * k.gm1;
    return s;
}

// make_side with 1 / rho from rcp2 (exact for rho in [2^-300, 2^300]).
__device__ __forceinline__ Side make_side_fast(const Consts& k, const double c[4]) {
    Side s;
    s.r = c[0];
    s.ru = c[1];
    s.rv = c[2];
    s.E = c[3];
    s.ir = qfin(1.0, s.r, rcp2(s.r));
    s.u = s.ru * s.ir;
    s.v = s.rv * s.ir;
    s.u2 = s.u * s.u;
    s.v2 = s.v * s.v;
    s.p = (s.E - 0.5 * s.r * (s.u2 + s.v2)) * k.gm1;
    return s;
}

template <int DIR>
__device__ __forceinline__ void phys_flux(const Side& s, double H, double F[4]) {
    if constexpr (DIR == 1) {
        F[0] = s.r * s.u;
        F[1] = s.r * s.u2 + s.p;
        F[2] = s.r * s.v * s.u;
        F[3] = s.r * H * s.u;
    } else {
        F[0] = s.r * s.v;
        F[1] = s.r * s.u * s.v;
        F[2] = s.r * s.v2 + s.p;
        F[3] = s.r * H * s.v;
    }
}

// HLLC with Einfeldt speeds, include/scheme.h:690-1001 (per-direction round-off:
// sound speed :745-746 vs :895-896, Einfeldt eta :777 vs :929).
//
// FAST: the reciprocals of rho and of sqrt(rho_L) + sqrt(rho_R), the Einfeldt quotient and
// the five square roots come from rcp2 / sqrt_mid, which give the IEEE results when both
// densities lie in [2^-290, 2^290] (then every reciprocal and the quotient have their
// operands in [2^-300, 2^300]) and the sound speeds' and the Einfeldt speed's squares in
// [2^-700, 2^700]; `ok` says whether they did, and hllc() reruns FAST = false otherwise.
template <int DIR, bool FAST>
__device__ __forceinline__ int hllc_t(const Consts& k, const double cl[4], const double cr[4],
                                      double F[4], bool& ok) {
    const double g = k.g, gm1 = k.gm1, gr = k.gr;
    auto root = [&](double x) {
        if constexpr (FAST) {
            ok = ok && x >= 0x1p-700 && x <= 0x1p700;
            return sqrt_mid(x);
        } else {
            return sqrt(x);
        }
    };
    if constexpr (FAST)
        ok = cl[0] >= 0x1p-290 && cl[0] <= 0x1p290 && cr[0] >= 0x1p-290 && cr[0] <= 0x1p290;
    const Side L = FAST ? make_side_fast(k, cl) : make_side(k, cl);
    const Side R = FAST ? make_side_fast(k, cr) : make_side(k, cr);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double aL, aR;
    if constexpr (DIR == 1) {
        aL = root((g - 1.0) * (HL - 0.5 * (L.u2 + L.v2)));
        aR = root((g - 1.0) * (HR - 0.5 * (R.u2 + R.v2)));
    } else {
        aL = root(g * L.p * L.ir);
        aR = root(g * R.p * R.ir);
    }
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    // rho in [2^-290, 2^290]: sL, sR in [2^-145, 2^145], ss in [2^-145, 2^146]
    const double sL = FAST ? sqrt_mid(L.r) : sqrt(L.r), sR = FAST ? sqrt_mid(R.r) : sqrt(R.r);
    const double ss = sL + sR;
    const double iss = FAST ? qfin(1.0, ss, rcp2(ss)) : 1.0 / ss;
    double eta;
    if constexpr (DIR == 1) eta = 0.5 * sL * sR * iss * iss;
    else {
        const double s2 = ss * ss;  // in [2^-290, 2^292]
        eta = FAST ? qfin(0.5 * sL * sR, s2, rcp2(s2)) : 0.5 * sL * sR / s2;
    }
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const double nbar = ((DIR == 1 ? sL * L.u + sR * R.u : sL * L.v + sR * R.v)) * iss;
    const double d = root((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    const double sl = nbar - d, sr = nbar + d;
    if (sl >= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FL[k];
        return 0;
    }
    if (sr <= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FR[k];
        return 0;
    }
    const double denom = L.r * (sl - qL) - R.r * (sr - qR);
    const double sstar = (R.p - L.p + L.r * qL * (sl - qL) - R.r * qR * (sr - qR)) / denom;
    if (sstar >= 0.0) {
        const double m = L.r * (sl - qL) / (sl - sstar);
        const double e = L.E * L.ir + (sstar - qL) * (sstar + L.p / (L.r * (sl - qL)));
        F[0] = FL[0] + sl * (m - L.r);
        if constexpr (DIR == 1) {
            F[1] = FL[1] + sl * (m * sstar - L.ru);
            F[2] = FL[2] + sl * (m * L.v - L.rv);
        } else {
            F[1] = FL[1] + sl * (m * L.u - L.ru);
            F[2] = FL[2] + sl * (m * sstar - L.rv);
        }
        F[3] = FL[3] + sl * (m * e - L.E);
    } else {
        const double m = R.r * (sr - qR) / (sr - sstar);
        const double e = R.E * R.ir + (sstar - qR) * (sstar + R.p / (R.r * (sr - qR)));
        F[0] = FR[0] + sr * (m - R.r);
        if constexpr (DIR == 1) {
            F[1] = FR[1] + sr * (m * sstar - R.ru);
            F[2] = FR[2] + sr * (m * R.v - R.rv);
        } else {
            F[1] = FR[1] + sr * (m * R.u - R.ru);
            F[2] = FR[2] + sr * (m * sstar - R.rv);
        }
        F[3] = FR[3] + sr * (m * e - R.E);
    }
    return 0;
}

template <int DIR>
__device__ __forceinline__ int hllc(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    bool ok = true;
    const int r = hllc_t<DIR, true>(k, cl, cr, F, ok);
    if (all_lanes(ok)) return r;
    return hllc_t<DIR, false>(k, cl, cr, F, ok);
}

// HLL with Einfeldt speeds, include/scheme.h:400-674.  The reference prints and
// skips a face with rho <= 0, NaN rho (:427-436) or a NaN speed (:500-503), which
// leaves stale memory in its flux array; here such a face gets NaN and a fault bit.
template <int DIR>
__device__ __forceinline__ int hll(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    if (cl[0] <= 0.0 || cr[0] <= 0.0 || isnan(cl[0]) || isnan(cr[0])) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = __builtin_nan("");
        return 1;
    }
    const double g = k.g, gr = k.gr;
    const Side L = make_side(k, cl), R = make_side(k, cr);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    const double aL = sqrt((g - 1.0) * (HL - 0.5 * (L.u2 + L.v2)));
    const double aR = sqrt((g - 1.0) * (HR - 0.5 * (R.u2 + R.v2)));
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double ss = sL + sR;
    const double iss = 1.0 / ss;
    const double eta = 0.5 * sL * sR / (ss * ss);
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const double nbar = ((DIR == 1 ? sL * L.u + sR * R.u : sL * L.v + sR * R.v)) * iss;
    const double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    if (isnan(d)) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = __builtin_nan("");
        return 1;
    }
    const double sl = nbar - d, sr = nbar + d;
    if (sl >= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else {
        const double isd = 1.0 / (sr - sl);
        const double slsr = sl * sr;
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = (sr * FL[k] - sl * FR[k] + slsr * (cr[k] - cl[k])) * isd;
    }
    return 0;
}

// Positively conservative HLLE (Einfeldt, Munz, Roe & Sjogreen 1991): wave-speed
// bounds s_L = min(q_L - a_L, q~ - c~), s_R = max(q_R + a_R, q~ + c~) with Roe
// averages.  Used only by the positivity safeguard (an extension): the reference's HLL
// uses the unbounded speeds q~ -/+ d, which do not preserve positivity.
template <int DIR>
__device__ __forceinline__ int hlle(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    if (!(cl[0] > 0.0) || !(cr[0] > 0.0)) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = __builtin_nan("");
        return 1;
    }
    const double g = k.g, gr = k.gr;
    const Side L = make_side(k, cl), R = make_side(k, cr);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    const double aL = sqrt(fmax(g * L.p * L.ir, 0.0));
    const double aR = sqrt(fmax(g * R.p * R.ir, 0.0));
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double iss = 1.0 / (sL + sR);
    const double ubar = (sL * L.u + sR * R.u) * iss;
    const double vbar = (sL * L.v + sR * R.v) * iss;
    const double Hbar = (sL * HL + sR * HR) * iss;
    const double cbar = sqrt(fmax((g - 1.0) * (Hbar - 0.5 * (ubar * ubar + vbar * vbar)), 0.0));
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const double qbar = DIR == 1 ? ubar : vbar;
    const double sl = fmin(qL - aL, qbar - cbar);
    const double sr = fmax(qR + aR, qbar + cbar);
    if (sl >= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else {
        const double isd = 1.0 / (sr - sl);
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = (sr * FL[k] - sl * FR[k] + sl * sr * (cr[k] - cl[k])) * isd;
    }
    return 0;
}

// Roe with the reference's no-op entropy fix, include/scheme.h:1018-1322.  With HC it
// is RoeHC (include/scheme.h:2793-3110): eps = 1e-12 in the wave-strength reciprocals,
// a NaN Roe density skips the face (:2936-2939; here NaN and a fault bit), and the
// energy flux loses 0.5 |q~ + c~| (e_R - e_L) rho~, e = p / (rho (gamma - 1)).
template <int DIR, bool HC = false>
__device__ __forceinline__ int roe(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    const double g = k.g, gm1 = k.gm1, gr = k.gr;
    const Side L = make_side(k, cl), R = make_side(k, cr);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double ss = sL + sR;
    const double iss = 1.0 / ss;
    const double rbar = sL * sR;
    const double ubar = (sL * L.u + sR * R.u) * iss;
    const double vbar = (sL * L.v + sR * R.v) * iss;
    const double Hbar = (sL * HL + sR * HR) * iss;
    const double q2 = ubar * ubar + vbar * vbar;
    const double cbar = sqrt(gm1 * (Hbar - 0.5 * q2));
    const double nbar = DIR == 1 ? ubar : vbar;
    const double lam[4] = {fabs(nbar - cbar), fabs(nbar), fabs(nbar), fabs(nbar + cbar)};
    const double dr = R.r - L.r, du = R.u - L.u, dv = R.v - L.v, dp = R.p - L.p;
    const double cs = cbar * cbar;
    const double i2cs = HC ? 1.0 / (2.0 * cs + 1e-12) : 1.0 / (2.0 * cs);
    const double ics = HC ? 1.0 / (cs + 1e-12) : 1.0 / cs;
    const double dn = DIR == 1 ? du : dv, dtg = DIR == 1 ? dv : du;
    const double al[4] = {(dp - rbar * cbar * dn) * i2cs, dr - dp * ics, rbar * dtg,
                          (dp + rbar * cbar * dn) * i2cs};
    double ev[4][4];
    if constexpr (DIR == 1) {
        ev[0][0] = 1.0; ev[0][1] = ubar - cbar; ev[0][2] = vbar; ev[0][3] = Hbar - ubar * cbar;
        ev[1][0] = 1.0; ev[1][1] = ubar;        ev[1][2] = vbar; ev[1][3] = 0.5 * q2;
        ev[2][0] = 0.0; ev[2][1] = 0.0;         ev[2][2] = 1.0;  ev[2][3] = vbar;
        ev[3][0] = 1.0; ev[3][1] = ubar + cbar; ev[3][2] = vbar; ev[3][3] = Hbar + ubar * cbar;
    } else {
        ev[0][0] = 1.0; ev[0][1] = ubar; ev[0][2] = vbar - cbar; ev[0][3] = Hbar - vbar * cbar;
        ev[1][0] = 1.0; ev[1][1] = ubar; ev[1][2] = vbar;        ev[1][3] = 0.5 * q2;
        ev[2][0] = 0.0; ev[2][1] = 1.0;  ev[2][2] = 0.0;         ev[2][3] = ubar;
        ev[3][0] = 1.0; ev[3][1] = ubar; ev[3][2] = vbar + cbar; ev[3][3] = Hbar + vbar * cbar;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]);
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const double term = 0.5 * lam[w] * al[w];
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] -= term * ev[w][k];
    }
    if constexpr (HC) {
        if (isnan(rbar)) {
#pragma unroll
            for (int k = 0; k < 4; k++) F[k] = __builtin_nan("");
            return 1;
        }
        const double igm1 = 1.0 / gm1;
        const double eL = L.ir * L.p * igm1, eR = R.ir * R.p * igm1;
        F[3] = F[3] - 0.5 * lam[3] * (eR - eL) * rbar;
    }
    return 0;
}

// HLL with the heat-capacity energy term (HLLHC), include/scheme.h:1860-2137: sound
// speed sqrt(gamma p / rho), eta through the reciprocal of sqrt(rho_L) + sqrt(rho_R)
// squared, a NaN Einfeldt speed kept on x faces (the fallback is commented out there,
// :1948-1953) and replaced by the Roe sound speed on y faces (:2073-2076); the
// star-region energy flux adds (s_L s_R / (s_R - s_L)) (e_R - e_L) rho_HLL.
template <int DIR>
__device__ __forceinline__ int hllhc(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    const double g = k.g, gm1 = k.gm1, gr = k.gr;
    const Side L = make_side(k, cl), R = make_side(k, cr);
    const double aL = sqrt(g * L.p * L.ir);
    const double aR = sqrt(g * R.p * R.ir);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double ss = sL + sR;
    const double iss = 1.0 / ss;
    const double ubar = (sL * L.u + sR * R.u) * iss;
    const double vbar = (sL * L.v + sR * R.v) * iss;
    const double eta = 0.5 * sL * sR * iss * iss;
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    if constexpr (DIR == 2) {
        if (isnan(d)) {
            const double Hbar = (sL * HL + sR * HR) * iss;
            d = sqrt(gm1 * (Hbar - 0.5 * (ubar * ubar + vbar * vbar)));
        }
    }
    const double nbar = DIR == 1 ? ubar : vbar;
    const double sl = nbar - d, sr = nbar + d;
    if (sl >= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FL[k];
    } else if (sr <= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    } else {
        const double rhll = (sr * R.r - sl * L.r + FL[0] - FR[0]) / (sr - sl);
        const double igm1 = 1.0 / gm1;
        const double eL = L.p * L.ir * igm1, eR = R.p * R.ir * igm1;
        const double isd = 1.0 / (sr - sl);
        const double slsr = sl * sr;
#pragma unroll
        for (int k = 0; k < 3; k++) F[k] = (sr * FL[k] - sl * FR[k] + slsr * (cr[k] - cl[k])) * isd;
        F[3] = (sr * FL[3] - sl * FR[3] + slsr * (cr[3] - cl[3])) * isd + (slsr * isd) * (eR - eL) * rhll;
    }
    return 0;
}

// HLLC with the heat-capacity energy term (HLLCHC), include/scheme.h:2153-2478: sound
// speed from the enthalpy, eta through (sqrt(rho_L) + sqrt(rho_R))^2 in both
// directions, no NaN fallback, and the reference's branch order (s_L >= 0; s_L < 0 and
// s* >= 0; s* < 0 and s_R > 0; else the right flux, which NaN speeds also take).  The
// star-state energy fluxes add (s_L s_R / (s_R - s_L)) (e_R - e_L) times the star
// density.
template <int DIR>
__device__ __forceinline__ int hllchc(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    const double g = k.g, gm1 = k.gm1, gr = k.gr;
    const Side L = make_side(k, cl), R = make_side(k, cr);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    const double aL = sqrt((g - 1.0) * (HL - 0.5 * (L.u2 + L.v2)));
    const double aR = sqrt((g - 1.0) * (HR - 0.5 * (R.u2 + R.v2)));
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double ss = sL + sR;
    const double iss = 1.0 / ss;
    const double nbar = ((DIR == 1 ? sL * L.u + sR * R.u : sL * L.v + sR * R.v)) * iss;
    const double eta = 0.5 * sL * sR / (ss * ss);
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const double d = sqrt((sL * aL * aL + sR * aR * aR) * iss + eta * (qR - qL) * (qR - qL));
    const double sl = nbar - d, sr = nbar + d;
    if (sl >= 0.0) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FL[k];
        return 0;
    }
    const double denom = L.r * (sl - qL) - R.r * (sr - qR);
    const double sstar = (R.p - L.p + L.r * qL * (sl - qL) - R.r * qR * (sr - qR)) / denom;
    const double igm1 = 1.0 / gm1;
    const double eL = L.ir * L.p * igm1, eR = R.ir * R.p * igm1;
    const double isd = 1.0 / (sr - sl);
    const double slsr = sl * sr;
    const double ec = slsr * isd * (eR - eL);
    if (sl < 0.0 && sstar >= 0.0) {
        const double m = L.r * (sl - qL) / (sl - sstar);
        const double e = L.E * L.ir + (sstar - qL) * (sstar + L.p / (L.r * (sl - qL)));
        F[0] = FL[0] + sl * (m - L.r);
        if constexpr (DIR == 1) {
            F[1] = FL[1] + sl * (m * sstar - L.ru);
            F[2] = FL[2] + sl * (m * L.v - L.rv);
        } else {
            F[1] = FL[1] + sl * (m * L.u - L.ru);
            F[2] = FL[2] + sl * (m * sstar - L.rv);
        }
        F[3] = FL[3] + sl * (m * e - L.E) + ec * m;
    } else if (sstar < 0.0 && sr > 0.0) {
        const double m = R.r * (sr - qR) / (sr - sstar);
        const double e = R.E * R.ir + (sstar - qR) * (sstar + R.p / (R.r * (sr - qR)));
        F[0] = FR[0] + sr * (m - R.r);
        if constexpr (DIR == 1) {
            F[1] = FR[1] + sr * (m * sstar - R.ru);
            F[2] = FR[2] + sr * (m * R.v - R.rv);
        } else {
            F[1] = FR[1] + sr * (m * R.u - R.ru);
            F[2] = FR[2] + sr * (m * sstar - R.rv);
        }
        F[3] = FR[3] + sr * (m * e - R.E) + ec * m;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = FR[k];
    }
    return 0;
}

// Local Lax-Friedrichs (the reference's default solver branch), include/scheme.h:229-381.
template <int DIR>
__device__ __forceinline__ int llf(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    const double g = k.g, gr = k.gr;
    const Side L = make_side(k, cl), R = make_side(k, cr);
    const double aL = sqrt(g * L.p * L.ir);
    const double aR = sqrt(g * R.p * R.ir);
    const double HL = 0.5 * (L.u2 + L.v2) + gr * (L.p * L.ir);
    const double HR = 0.5 * (R.u2 + R.v2) + gr * (R.p * R.ir);
    double FL[4], FR[4];
    phys_flux<DIR>(L, HL, FL);
    phys_flux<DIR>(R, HR, FR);
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const double s = max2(fabs(qL) + aL, fabs(qR) + aR);
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]) - 0.5 * s * (cr[k] - cl[k]);
    return 0;
}

enum : int {
    SOLVER_HLL = 1, SOLVER_HLLC = 2, SOLVER_ROE = 3, SOLVER_LLF = 10,
    SOLVER_HLLHC = 11, SOLVER_HLLCHC = 22, SOLVER_ROEHC = 33
};

template <int DIR, int SOLVER>
__device__ __forceinline__ int riemann(const Consts& k, const double cl[4], const double cr[4], double F[4]) {
    if constexpr (SOLVER == SOLVER_HLL) return hll<DIR>(k, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLC) return hllc<DIR>(k, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_ROE) return roe<DIR>(k, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLHC) return hllhc<DIR>(k, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLCHC) return hllchc<DIR>(k, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_ROEHC) return roe<DIR, true>(k, cl, cr, F);
    else return llf<DIR>(k, cl, cr, F);
}

// Positivity test of a face state (rho > 0, p > 0, finite); used only by the
// optional positivity safeguard, which the reference does not have.
__device__ __forceinline__ bool admissible(const Consts& k, const double c[4]) {
    const double ir = 1.0 / c[0];
    const double u = c[1] * ir, v = c[2] * ir;
    const double p = (c[3] - 0.5 * c[0] * (u * u + v * v)) * k.gm1;
    return c[0] > 0.0 && p > 0.0 && isfinite(p) && isfinite(u) && isfinite(v);
}

}  // namespace fvm
