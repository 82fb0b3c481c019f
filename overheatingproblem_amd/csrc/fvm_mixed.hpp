// fvm_mixed.hpp -- the mixed-precision face computation: FP32 WENO/TVD (smoothness
// indicators, nonlinear weights, interpolation) and FP32 Riemann wave algebra, with
// FP64 state, characteristic basis, pressure recovery and physical fluxes.
//
// What stays FP64 and why:
//  * the state and the face states handed to the Riemann solver: each face state is
//    U_ref + dU, with U_ref the mean of the face's two cells (FP64) and dU the
//    reconstruction of the stencil's deviations from U_ref, so FP32 round-off scales with
//    the local jump, not with |U| (Noh has E = 0.5 + 1.5e-6);
//  * the characteristic basis, the projection w = L dU and the back-projection R w.  The
//    basis has condition number ~ Mach^2 (b1 = (gamma-1)/c^2 ~ 4e5 in Noh's p = 1e-6
//    inflow, Mach ~ 800): in FP32, R L = I fails at the 1e-1 level.  The divides and
//    square roots of the basis use the hardware FP64 reciprocal / reciprocal square root
//    with one Newton step (relative error ~1e-14), not IEEE sequences;
//  * pressure recovery p = (gamma-1)(E - (rho u^2 + rho v^2)/2) and the physical fluxes;
//  * every FP32 quantity of the Riemann solvers is a jump or a speed relative to the
//    face's Roe velocity (HLLC: s - q = (q~ - q) -/+ d with q~ - q from dq; s* - q from
//    dp and dq), so hypersonic faces (c << |q|) do not cancel in FP32.
// A face whose reconstructed states have rho <= 0 or p <= 0 is recomputed with the exact
// FP64 path (fvm_math.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_math.hpp"

namespace fvm {
namespace mx {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// 1/x to ~1e-14 relative: hardware FP64 reciprocal estimate + one Newton step.
__device__ __forceinline__ double rcp64(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// sqrt(x) to ~1e-14 relative: hardware FP64 reciprocal square root + one Newton step.
__device__ __forceinline__ double sqrt64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double s = x * y;
    return fma(0.5 * s, fma(-s, y, 1.0), s);
}

// ---------------------------------------------------------------- FP32 WENO etc.
// FP32 WENO5 of both face values of one field from its six-cell stencil s[0..5] (cells
// i-2..i+3): the left value from (v1..v5) = (s0..s4), the right one from the mirrored
// (s5..s1).  Per side: Jiang-Shu smoothness indicators
// b = 13/12 (second difference)^2 + 1/4 (first difference)^2 + eps of the three
// sub-stencils, weights d_k (1 + tau / b_k) with tau = |b_1 - b_3| (WENO5-Z, exponent 1;
// z) or d_k / b_k^2 (WENO5-JS), d = (0.1, 0.6, 0.3), normalised, and the candidates
// (2 v1 - 7 v2 + 11 v3, -v2 + 5 v3 + 2 v4, 2 v3 + 5 v4 - v5) / 6.  The two sides share
// what their sub-stencils share: the second differences of (s1, s2, s3) and (s2, s3, s4)
// and the candidates -s1 + 5 s2 + 2 s3 and 2 s2 + 5 s3 - s4 (the right side's third and
// second).  eps rides in the first FMA of each indicator.  84 FP32 operations per field;
// the two sides computed apart take 102.
__device__ __forceinline__ void weno5_pair_f(const float s[6], bool z, float& left, float& right) {
    const float eps = 1.0e-15f, c13 = 13.0f / 12.0f;
    // second differences of (s0,s1,s2), (s1,s2,s3), (s2,s3,s4), (s3,s4,s5)
    const float d0 = fmaf(-2.0f, s[1], s[0]) + s[2];
    const float d1 = fmaf(-2.0f, s[2], s[1]) + s[3];
    const float d2 = fmaf(-2.0f, s[3], s[2]) + s[4];
    const float d3 = fmaf(-2.0f, s[4], s[3]) + s[5];
    const float q0 = fmaf(c13 * d0, d0, eps), q1 = fmaf(c13 * d1, d1, eps);
    const float q2 = fmaf(c13 * d2, d2, eps), q3 = fmaf(c13 * d3, d3, eps);
    // one-sided first differences: left stencils 1..3, right stencils 1..3
    const float e1 = fmaf(3.0f, s[2], fmaf(-4.0f, s[1], s[0])), e2 = s[1] - s[3];
    const float e3 = fmaf(3.0f, s[2], fmaf(-4.0f, s[3], s[4]));
    const float f1 = fmaf(3.0f, s[3], fmaf(-4.0f, s[4], s[5])), f2 = s[4] - s[2];
    const float f3 = fmaf(3.0f, s[3], fmaf(-4.0f, s[2], s[1]));
    // left: (s0,s1,s2), (s1,s2,s3), (s2,s3,s4); right: (s5,s4,s3), (s4,s3,s2), (s3,s2,s1)
    const float bl1 = fmaf(0.25f * e1, e1, q0), bl2 = fmaf(0.25f * e2, e2, q1);
    const float bl3 = fmaf(0.25f * e3, e3, q2);
    const float br1 = fmaf(0.25f * f1, f1, q3), br2 = fmaf(0.25f * f2, f2, q2);
    const float br3 = fmaf(0.25f * f3, f3, q1);
    // candidates: left p1 p2 p3, right p1' (p2' = p3, p3' = p2)
    const float pl1 = fmaf(2.0f, s[0], fmaf(-7.0f, s[1], 11.0f * s[2]));
    const float p2 = fmaf(5.0f, s[2], fmaf(2.0f, s[3], -s[1]));
    const float p3 = fmaf(2.0f, s[2], fmaf(5.0f, s[3], -s[4]));
    const float pr1 = fmaf(2.0f, s[5], fmaf(-7.0f, s[4], 11.0f * s[3]));
    float al1, al2, al3, ar1, ar2, ar3;
    if (z) {  // WENO5-Z, exponent 1
        const float tl = fabsf(bl1 - bl3), tr = fabsf(br1 - br3);
        al1 = fmaf(0.1f * tl, rcp(bl1), 0.1f);
        al2 = fmaf(0.6f * tl, rcp(bl2), 0.6f);
        al3 = fmaf(0.3f * tl, rcp(bl3), 0.3f);
        ar1 = fmaf(0.1f * tr, rcp(br1), 0.1f);
        ar2 = fmaf(0.6f * tr, rcp(br2), 0.6f);
        ar3 = fmaf(0.3f * tr, rcp(br3), 0.3f);
    } else {  // WENO5-JS
        const float il1 = rcp(bl1), il2 = rcp(bl2), il3 = rcp(bl3);
        const float ir1 = rcp(br1), ir2 = rcp(br2), ir3 = rcp(br3);
        al1 = (0.1f * il1) * il1;
        al2 = (0.6f * il2) * il2;
        al3 = (0.3f * il3) * il3;
        ar1 = (0.1f * ir1) * ir1;
        ar2 = (0.6f * ir2) * ir2;
        ar3 = (0.3f * ir3) * ir3;
    }
    const float isl = rcp(al1 + al2 + al3) * (1.0f / 6.0f);
    const float isr = rcp(ar1 + ar2 + ar3) * (1.0f / 6.0f);
    left = fmaf(al1, pl1, fmaf(al2, p2, al3 * p3)) * isl;
    right = fmaf(ar1, pr1, fmaf(ar2, p3, ar3 * p2)) * isr;
}

__device__ __forceinline__ float weno3_f(float v1, float v2, float v3) {
    const float eps = 1e-6f;
    const float s1 = (v2 - v1) * (v2 - v1), s2 = (v3 - v2) * (v3 - v2);
    const float i1 = rcp(eps + s1), i2 = rcp(eps + s2);
    const float a1 = (1.0f / 3.0f) * i1 * i1, a2 = (2.0f / 3.0f) * i2 * i2;
    return (a1 * (-0.5f * v1 + 1.5f * v2) + a2 * (0.5f * v2 + 0.5f * v3)) * rcp(a1 + a2);
}

__device__ __forceinline__ float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// TVD slope on an undivided stencil (the reference divides by delta and multiplies back;
// minmod is scale-free, van Leer's 1e-6 guard is applied to the divided slopes).
__device__ __forceinline__ float tvd_half_slope(float v1, float v2, float v3, float delta, bool vanleer) {
    const float id = rcp(delta);
    const float a = (v2 - v1) * id, b = (v3 - v2) * id;
    float s;
    if (vanleer) s = ((sgnf(a) + sgnf(b)) * a * b) * rcp(fabsf(a) + fabsf(b) + 1.0e-6f);
    else s = 0.5f * (sgnf(a) + sgnf(b)) * fminf(fabsf(a), fabsf(b));
    return 0.5f * s * delta;
}

template <int ORDER, bool CHAR>
__device__ __forceinline__ void recon_scalar_f(const float s[6], float delta, float& l, float& r) {
    if constexpr (ORDER == 1) {
        l = s[2];
        r = s[3];
    } else if constexpr (ORDER == 2) {
        l = s[2] + tvd_half_slope(s[1], s[2], s[3], delta, CHAR);
        r = s[3] - tvd_half_slope(s[2], s[3], s[4], delta, CHAR);
    } else if constexpr (ORDER == 3) {
        l = weno3_f(s[1], s[2], s[3]);
        r = weno3_f(s[4], s[3], s[2]);
    } else {
        weno5_pair_f(s, CHAR, l, r);
    }
}

// Primitive state of a cell with FP64 pressure recovery.
struct Prim64 {
    double r, ir, u, v, p;
};
__device__ __forceinline__ Prim64 prim64(double g, const double c[4]) {
    Prim64 s;
    s.r = c[0];
    s.ir = rcp64(c[0]);
    s.u = c[1] * s.ir;
    s.v = c[2] * s.ir;
    s.p = fma(-0.5, fma(c[1], s.u, c[2] * s.v), c[3]) * (g - 1.0);
    return s;
}

__device__ __forceinline__ bool admissible64(double g, const double c[4]) {
    const Prim64 s = prim64(g, c);
    return c[0] > 0.0 && s.p > 0.0 && isfinite(s.p) && isfinite(s.u) && isfinite(s.v);
}

// The Roe-averaged characteristic basis in FP64 (the eigenvectors of
// Compute_Eigen_2DX/2DY, include/Characteriz.h:346-392,448-494), kept as the scalars
// its rows are built from.  The Roe sound speed is formed without cancellation,
// c~^2 = (sqrt(rho_a) c_a^2 + sqrt(rho_b) c_b^2)/(sqrt(rho_a)+sqrt(rho_b))
//   + (gamma-1)/2 sqrt(rho_a rho_b)/(sqrt(rho_a)+sqrt(rho_b))^2 |u_b - u_a|^2,
// which equals (gamma-1)(H~ - |u~|^2/2) for an ideal gas; then b1 = (gamma-1)/c~^2,
// b2 = b1 |u~|^2/2, H~ = c~^2/(gamma-1) + |u~|^2/2.  With n the face-normal velocity
// component and t the tangential one, the left eigenvectors (rows) are, for x faces,
//   l0 = ((b2 + n/c)/2, -(b1 u + 1/c)/2, -b1 v/2, b1/2),  l1 = (H~ - |u~|^2, u, v, -1),
//   l2 = (v, 0, -1, 0),  l3 = ((b2 - n/c)/2, (-b1 u + 1/c)/2, -b1 v/2, b1/2)
// (y faces: u and v exchange roles in l0, l3 and l2 = (-u, 1, 0, 0)), and the right
// eigenvectors (columns) r0 = (1, u - c, v, H~ - u c), r1 = b1 (1, u, v, |u~|^2/2),
// r2 = (0, 0, -1, -v), r3 = (1, u + c, v, H~ + u c) for x faces (y faces:
// r0 = (1, u, v - c, H~ - v c), r2 = (0, 1, 0, u), r3 = (1, u, v + c, H~ + v c)).
struct Basis64 {
    double u, v, c, nc;    // Roe velocity, sound speed, normal velocity x c
    double b1, b2, H;      // right-eigenvector scalars
    double hb1, hb2, hic;  // b1/2, b2/2, 1/(2c)
    double Hmq2;           // H~ - |u~|^2
};

template <int DIR>
__device__ __forceinline__ Basis64 face_basis64(double g, double igm1, const Prim64& A,
                                                const Prim64& B) {
    const double gm1 = g - 1.0;
    const double sa = sqrt64(A.r), sb = sqrt64(B.r);
    const double iss = rcp64(sa + sb);
    Basis64 e;
    e.u = fma(sa, A.u, sb * B.u) * iss;
    e.v = fma(sa, A.v, sb * B.v) * iss;
    const double du = B.u - A.u, dv = B.v - A.v;
    const double a2A = g * (A.p * A.ir), a2B = g * (B.p * B.ir);
    const double siss = sa * iss, tiss = sb * iss;
    double c2 = fma(siss, a2A, tiss * a2B) + (0.5 * gm1) * (siss * tiss) * fma(du, du, dv * dv);
    c2 = fmax(c2, 1.0e-300);
    const double q2 = fma(e.u, e.u, e.v * e.v);
    const double ic2 = rcp64(c2);
    e.c = sqrt64(c2);
    e.nc = (DIR == 1 ? e.u : e.v) * e.c;
    e.b1 = gm1 * ic2;
    e.b2 = 0.5 * e.b1 * q2;
    e.H = fma(c2, igm1, 0.5 * q2);
    e.Hmq2 = fma(c2, igm1, -0.5 * q2);
    e.hb1 = 0.5 * e.b1;
    e.hb2 = 0.5 * e.b2;
    e.hic = 0.5 * (e.c * ic2);
    return e;
}

// w = L d for one stencil cell, from the row structure: with P = u d1 + v d2 - d3,
// w1 = (H~ - |u~|^2) d0 + P, A = (b2 d0 - b1 P)/2, B = (n d0 - d_n)/(2c),
// w0 = A + B, w3 = A - B, w2 = v d0 - d2 (x) or d1 - u d0 (y).
template <int DIR>
__device__ __forceinline__ void project64(const Basis64& e, const double d[4], float w[4]) {
    const double P = fma(e.u, d[1], fma(e.v, d[2], -d[3]));
    const double A = fma(e.hb2, d[0], -e.hb1 * P);
    const double B = DIR == 1 ? e.hic * fma(e.u, d[0], -d[1]) : e.hic * fma(e.v, d[0], -d[2]);
    w[0] = (float)(A + B);
    w[1] = (float)fma(e.Hmq2, d[0], P);
    w[2] = (float)(DIR == 1 ? fma(e.v, d[0], -d[2]) : fma(-e.u, d[0], d[1]));
    w[3] = (float)(A - B);
}

// dU = R w: with S = w0 + w3 and D = w3 - w0, dU0 = S + b1 w1 and
// x: dU1 = u dU0 + c D, dU2 = v dU0 - w2, dU3 = H~ S + b2 w1 - v w2 + u c D;
// y: dU1 = u dU0 + w2, dU2 = v dU0 + c D, dU3 = H~ S + b2 w1 + u w2 + v c D.
template <int DIR>
__device__ __forceinline__ void unproject64(const Basis64& e, const float w[4], double dU[4]) {
    const double w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    const double S = w0 + w3, D = w3 - w0;
    dU[0] = fma(e.b1, w1, S);
    if constexpr (DIR == 1) {
        dU[1] = fma(e.u, dU[0], e.c * D);
        dU[2] = fma(e.v, dU[0], -w2);
        dU[3] = fma(e.H, S, fma(e.b2, w1, fma(-e.v, w2, e.nc * D)));
    } else {
        dU[1] = fma(e.u, dU[0], w2);
        dU[2] = fma(e.v, dU[0], e.c * D);
        dU[3] = fma(e.H, S, fma(e.b2, w1, fma(e.u, w2, e.nc * D)));
    }
}

// Face states U_ref + dU from the six-cell stencil (FP64 in and out), U_ref the mean of
// the face's two cells.  Their deviations are taken as d and -d (d = U_2 - U_ref; the
// right cell's own difference U_3 - U_ref differs from -d by at most one rounding of
// U_ref, 1e-16 relative, far below the FP32 reconstruction's error): one subtraction,
// projection and conversion fewer per face, and w_3 = -w_2 exactly, since the projection
// is odd under rounding to nearest.  (The left cell as U_ref, d_2 = 0, saves more, but
// on Noh 256^2 over 150 steps the HLL-family runs then ended up to 40x further from FP64:
// HLL 1.4e-5 against 3.0e-7 relative L1(rho), FP64's own sensitivity to a 1e-12
// perturbation being 3.4e-7.  The likely cause is the reflecting walls, where the mean
// keeps the round-off of a face's two mirror-image states alike.)  Characteristic
// branch: w = L dU in FP64, WENO/TVD on w in FP32, dU = R w in FP64.
template <int DIR, int ORDER, bool CHAR>
__device__ __forceinline__ void reconstruct_face(double g, double igm1, double delta,
                                                 const double st[6][4],
                                                 double UL[4], double UR[4]) {
    double ref[4];
#pragma unroll
    for (int v = 0; v < 4; v++) ref[v] = 0.5 * (st[2][v] + st[3][v]);
    double d[6][4];
#pragma unroll
    for (int n = 0; n < 6; n++)
#pragma unroll
        for (int v = 0; v < 4; v++) d[n][v] = n == 3 ? -d[2][v] : st[n][v] - ref[v];
    const float df = (float)delta;
    if constexpr (CHAR) {
        const Basis64 e = face_basis64<DIR>(g, igm1, prim64(g, st[2]), prim64(g, st[3]));
        float w[6][4];
#pragma unroll
        for (int n = 0; n < 6; n++) {
            if (n == 3) {
#pragma unroll
                for (int k = 0; k < 4; k++) w[n][k] = -w[2][k];
            } else {
                project64<DIR>(e, d[n], w[n]);
            }
        }
        float wl[4], wr[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float s[6] = {w[0][k], w[1][k], w[2][k], w[3][k], w[4][k], w[5][k]};
            recon_scalar_f<ORDER, true>(s, df, wl[k], wr[k]);
        }
        double dl[4], dr[4];
        unproject64<DIR>(e, wl, dl);
        unproject64<DIR>(e, wr, dr);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            UL[v] = ref[v] + dl[v];
            UR[v] = ref[v] + dr[v];
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float d2 = (float)d[2][v];
            const float s[6] = {(float)d[0][v], (float)d[1][v], d2,
                                -d2, (float)d[4][v], (float)d[5][v]};
            float l, r;
            recon_scalar_f<ORDER, false>(s, df, l, r);
            UL[v] = ref[v] + (double)l;
            UR[v] = ref[v] + (double)r;
        }
    }
}

// ------------------------------------------------------------- Riemann solvers
// Physical flux of a side in FP64.
template <int DIR>
__device__ __forceinline__ void phys64(const double c[4], const Prim64& s, double F[4]) {
    const double q = DIR == 1 ? s.u : s.v;
    F[0] = c[0] * q;
    F[1] = c[1] * q + (DIR == 1 ? s.p : 0.0);
    F[2] = c[2] * q + (DIR == 2 ? s.p : 0.0);
    F[3] = (c[3] + s.p) * q;
}

// The jump of the specific internal energy e = p / (rho (gamma - 1)) across the face, in
// FP32 from the FP64 primitives (the heat-capacity term of HLLHC / HLLCHC / RoeHC).
__device__ __forceinline__ float energy_jump(double g, const Prim64& L, const Prim64& R) {
    return (float)(R.p * R.ir - L.p * L.ir) * rcp((float)(g - 1.0));
}

// Einfeldt speeds (the reference's HLL/HLLC estimate, include/scheme.h:765-790) relative
// to the face: with the Roe velocity q~, s_L - q_L = (q~ - q_L) - d and
// s_R - q_R = (q~ - q_R) + d, where q~ - q_L = sqrt(rho_R) dq / (sqrt(rho_L)+sqrt(rho_R))
// and q~ - q_R = -sqrt(rho_L) dq / (...): each from the FP64 jump dq.
template <int DIR>
struct Speeds {
    float aL, aR;   // s_L - q_L, s_R - q_R
    float sl, sr;   // s_L, s_R
};

template <int DIR>
__device__ __forceinline__ Speeds<DIR> einfeldt(double g, const Prim64& L, const Prim64& R) {
    const float gf = (float)g;
    const float aL2 = gf * (float)(L.p * L.ir), aR2 = gf * (float)(R.p * R.ir);
    const float sqL = fsqrt((float)L.r), sqR = fsqrt((float)R.r);
    const float iss = rcp(sqL + sqR);
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const float dq = (float)(qR - qL);
    const float eta = 0.5f * sqL * sqR * iss * iss;
    const float d = fsqrt(fmaxf(fmaf(fmaf(sqL, aL2, sqR * aR2), iss, eta * dq * dq), 0.0f));
    const float nL = sqR * dq * iss, nR = -sqL * dq * iss;  // q~ - q_L, q~ - q_R
    Speeds<DIR> s;
    s.aL = nL - d;
    s.aR = nR + d;
    s.sl = (float)qL + s.aL;
    s.sr = (float)qR + s.aR;
    return s;
}

template <int DIR>
__device__ __forceinline__ void hllc(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    const Speeds<DIR> w = einfeldt<DIR>(g, L, R);
    if (w.sl >= 0.0f) {
        phys64<DIR>(cl, L, F);
        return;
    }
    if (w.sr <= 0.0f) {
        phys64<DIR>(cr, R, F);
        return;
    }
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const float rL = (float)L.r, rR = (float)R.r;
    const float dq = (float)(qR - qL), dp = (float)(R.p - L.p);
    const float mL = rL * w.aL, mR = rR * w.aR;
    const float im = rcp(mL - mR);
    // s* - q_L = (dp - m_R dq)/(m_L - m_R), s* - q_R = (dp - m_L dq)/(m_L - m_R)
    const float sqstL = (dp - mR * dq) * im, sqstR = (dp - mL * dq) * im;
    const float sstar = (float)qL + sqstL;
    const bool left = sstar >= 0.0f;
    const double* c = left ? cl : cr;
    const Prim64& S = left ? L : R;
    const float s = left ? w.sl : w.sr, r = left ? rL : rR;
    const float sq = left ? w.aL : w.aR;           // s - q
    const float sqst = left ? sqstL : sqstR;       // s* - q
    phys64<DIR>(c, S, F);
    // U*_K - U_K = chi (rho, rho s (normal), rho v_t, E + p + rho (s - q) s*),
    // chi = (s* - q)/(s - s*), s - s* = (s - q) - (s* - q)
    const float chi = sqst * rcp(sq - sqst);
    const float sc = s * chi;
    const float tang = (float)(DIR == 1 ? S.v : S.u);
    const float dE = (float)(c[3] + S.p) + r * sq * sstar;
    F[0] += (double)(sc * r);
    const double dn = (double)(sc * r * s), dt = (double)(sc * r * tang);
    F[1] += DIR == 1 ? dn : dt;
    F[2] += DIR == 1 ? dt : dn;
    F[3] += (double)(sc * dE);
}

template <int DIR>
__device__ __forceinline__ void hll(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    const Speeds<DIR> w = einfeldt<DIR>(g, L, R);
    if (w.sl >= 0.0f) {
        phys64<DIR>(cl, L, F);
        return;
    }
    if (w.sr <= 0.0f) {
        phys64<DIR>(cr, R, F);
        return;
    }
    // F = F_L + s_L/(s_R - s_L) [s_R dU - dF] with, from F(U) = q U + P(U),
    // s_R dU - dF = (s_R - q_R) dU - dq U_L - dP: every term is a jump
    phys64<DIR>(cl, L, F);
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const float dq = (float)(qR - qL);
    const float k = w.sl * rcp(w.sr - w.sl);
    float dP[4] = {0.0f, 0.0f, 0.0f, (float)(R.p * qR - L.p * qL)};
    dP[DIR] = (float)(R.p - L.p);
#pragma unroll
    for (int m = 0; m < 4; m++)
        F[m] += (double)(k * (w.aR * (float)(cr[m] - cl[m]) - dq * (float)cl[m] - dP[m]));
}

// Roe; with HC RoeHC (include/scheme.h:2793-3110): eps = 1e-12 in the wave-strength
// reciprocals and the energy flux less 0.5 |q~ + c~| (e_R - e_L) rho~.  Unlike the HLL
// family, the wave algebra runs in FP64 (hardware reciprocal / square root with one
// Newton step): the acoustic wave strengths grow like the Mach number
// (a0, a3 ~ -/+ rho~ dn / (2 c~)) and largely cancel in the dissipation sum, and in a
// cold hypersonic flow (Noh's inflow: internal energy ~1e-6 of the total) FP32 cannot
// resolve the energy dissipation.  With FP32 wave algebra the mixed Noh state left the
// FP64 one by 3e-5 (relative L1) within 10 steps; the jump forms of HLL / HLLC stay at
// 1e-8.  The reconstruction stays FP32.
template <int DIR, bool HC = false>
__device__ __forceinline__ void roe(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    double FL[4], FR[4];
    phys64<DIR>(cl, L, FL);
    phys64<DIR>(cr, R, FR);
    const double gm1 = g - 1.0;
    const double sL = sqrt64(L.r), sR = sqrt64(R.r);
    const double iss = rcp64(sL + sR);
    const double rbar = sL * sR;
    const double ubar = fma(sL, L.u, sR * R.u) * iss, vbar = fma(sL, L.v, sR * R.v) * iss;
    const double aL2 = g * (L.p * L.ir), aR2 = g * (R.p * R.ir);
    const double du = R.u - L.u, dv = R.v - L.v;
    const double c2 = fmax(fma(fma(sL, aL2, sR * aR2), iss, (0.5 * gm1) * rbar * iss * iss *
                                                               fma(du, du, dv * dv)),
                           1.0e-300);
    const double cbar = sqrt64(c2);
    const double q2 = fma(ubar, ubar, vbar * vbar);
    const double Hbar = fma(c2, rcp64(gm1), 0.5 * q2);
    const double nbar = DIR == 1 ? ubar : vbar;
    const double lam0 = fabs(nbar - cbar), lam1 = fabs(nbar), lam3 = fabs(nbar + cbar);
    const double dr = cr[0] - cl[0], dp = R.p - L.p;
    const double i2cs = HC ? rcp64(fma(2.0, c2, 1.0e-12)) : 0.5 * rcp64(c2);
    const double ics = HC ? rcp64(c2 + 1.0e-12) : rcp64(c2);
    const double dn = DIR == 1 ? du : dv, dtg = DIR == 1 ? dv : du;
    const double rcdn = rbar * cbar * dn;
    const double a0 = (dp - rcdn) * i2cs, a1 = fma(-dp, ics, dr), a2 = rbar * dtg,
                 a3 = (dp + rcdn) * i2cs;
    const double t0 = 0.5 * lam0 * a0, t1 = 0.5 * lam1 * a1, t2 = 0.5 * lam1 * a2,
                 t3 = 0.5 * lam3 * a3;
    const double s03 = t0 + t3, d03 = t3 - t0;  // the acoustic pair: sum and difference
    double D[4];
    D[0] = s03 + t1;
    if constexpr (DIR == 1) {
        D[1] = fma(ubar, D[0], cbar * d03);
        D[2] = fma(vbar, D[0], t2);
        D[3] = fma(Hbar, s03, fma(ubar * cbar, d03, fma(t1, 0.5 * q2, t2 * vbar)));
    } else {
        D[1] = fma(ubar, D[0], t2);
        D[2] = fma(vbar, D[0], cbar * d03);
        D[3] = fma(Hbar, s03, fma(vbar * cbar, d03, fma(t1, 0.5 * q2, t2 * ubar)));
    }
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]) - D[k];
    if constexpr (HC) F[3] -= 0.5 * lam3 * ((R.p * R.ir - L.p * L.ir) * rcp64(gm1)) * rbar;
}

// LLF: the dissipation 0.5 s (U_R - U_L) multiplies the FP64 jumps (the FP32 speed s
// only scales it; rounding the jumps themselves would, as for Roe, swamp the energy
// dissipation of a cold hypersonic flow).
template <int DIR>
__device__ __forceinline__ void llf(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    double FL[4], FR[4];
    phys64<DIR>(cl, L, FL);
    phys64<DIR>(cr, R, FR);
    const float gf = (float)g;
    const float aL = fsqrt(fmaxf(gf * (float)(L.p * L.ir), 0.0f)), aR = fsqrt(fmaxf(gf * (float)(R.p * R.ir), 0.0f));
    const float qL = (float)(DIR == 1 ? L.u : L.v), qR = (float)(DIR == 1 ? R.u : R.v);
    const double hs = (double)(0.5f * fmaxf(fabsf(qL) + aL, fabsf(qR) + aR));
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = fma(-hs, cr[k] - cl[k], 0.5 * (FL[k] + FR[k]));
}

// ------------------------------------------------- heat-capacity solvers (mixed)
// HLLHC (include/scheme.h:1860-2137): the HLL flux of hll() and, between the waves, the
// energy term (s_L s_R / (s_R - s_L)) (e_R - e_L) rho_HLL with the HLL density
// rho_HLL = (s_R rho_R - s_L rho_L + F_L0 - F_R0) / (s_R - s_L), which in speeds relative
// to each side is (rho_R (s_R - q_R) - rho_L (s_L - q_L)) / (s_R - s_L).  (The reference's
// NaN fallback of the y-face wave speed cannot arise here: the mixed form hands the
// solver admissible states only.)
template <int DIR>
__device__ __forceinline__ void hllhc(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    const Speeds<DIR> w = einfeldt<DIR>(g, L, R);
    if (w.sl >= 0.0f) {
        phys64<DIR>(cl, L, F);
        return;
    }
    if (w.sr <= 0.0f) {
        phys64<DIR>(cr, R, F);
        return;
    }
    phys64<DIR>(cl, L, F);
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const float dq = (float)(qR - qL);
    const float isd = rcp(w.sr - w.sl);
    const float k = w.sl * isd;
    float dP[4] = {0.0f, 0.0f, 0.0f, (float)(R.p * qR - L.p * qL)};
    dP[DIR] = (float)(R.p - L.p);
#pragma unroll
    for (int m = 0; m < 4; m++)
        F[m] += (double)(k * (w.aR * (float)(cr[m] - cl[m]) - dq * (float)cl[m] - dP[m]));
    const float rhll = ((float)R.r * w.aR - (float)L.r * w.aL) * isd;
    F[3] += (double)(k * w.sr * energy_jump(g, L, R) * rhll);
}

// HLLCHC (include/scheme.h:2153-2478): the HLLC flux of hllc() and, in either star region,
// the energy term (s_L s_R / (s_R - s_L)) (e_R - e_L) rho*_K with the star density
// rho*_K = rho_K (s_K - q_K) / (s_K - s*).
template <int DIR>
__device__ __forceinline__ void hllchc(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    const Speeds<DIR> w = einfeldt<DIR>(g, L, R);
    if (w.sl >= 0.0f) {
        phys64<DIR>(cl, L, F);
        return;
    }
    if (w.sr <= 0.0f) {
        phys64<DIR>(cr, R, F);
        return;
    }
    const double qL = DIR == 1 ? L.u : L.v, qR = DIR == 1 ? R.u : R.v;
    const float rL = (float)L.r, rR = (float)R.r;
    const float dq = (float)(qR - qL), dp = (float)(R.p - L.p);
    const float mL = rL * w.aL, mR = rR * w.aR;
    const float im = rcp(mL - mR);
    const float sqstL = (dp - mR * dq) * im, sqstR = (dp - mL * dq) * im;
    const float sstar = (float)qL + sqstL;
    const bool left = sstar >= 0.0f;
    const double* c = left ? cl : cr;
    const Prim64& S = left ? L : R;
    const float s = left ? w.sl : w.sr, r = left ? rL : rR;
    const float sq = left ? w.aL : w.aR;
    const float sqst = left ? sqstL : sqstR;
    phys64<DIR>(c, S, F);
    const float iden = rcp(sq - sqst);  // 1 / (s - s*)
    const float chi = sqst * iden;
    const float sc = s * chi;
    const float tang = (float)(DIR == 1 ? S.v : S.u);
    const float dE = (float)(c[3] + S.p) + r * sq * sstar;
    const float ec = w.sl * w.sr * rcp(w.sr - w.sl) * energy_jump(g, L, R);
    F[0] += (double)(sc * r);
    const double dn = (double)(sc * r * s), dt = (double)(sc * r * tang);
    F[1] += DIR == 1 ? dn : dt;
    F[2] += DIR == 1 ? dt : dn;
    F[3] += (double)(fmaf(sc, dE, ec * (r * sq * iden)));
}

template <int DIR, int SOLVER>
__device__ __forceinline__ void riemann(double g, const double cl[4], const double cr[4], double F[4]) {
    if constexpr (SOLVER == SOLVER_HLL) hll<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLC) hllc<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_ROE) roe<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLHC) hllhc<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLCHC) hllchc<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_ROEHC) roe<DIR, true>(g, cl, cr, F);
    else llf<DIR>(g, cl, cr, F);
}

}  // namespace mx
}  // namespace fvm
