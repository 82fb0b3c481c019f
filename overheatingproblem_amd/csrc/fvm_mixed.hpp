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
__device__ __forceinline__ float weno5_f(float v1, float v2, float v3, float v4, float v5, bool z) {
    const float eps = 1.0e-15f;
    // second differences and one-sided first differences of the three stencils
    const float t1 = fmaf(-2.0f, v2, v1) + v3, t2 = fmaf(3.0f, v3, fmaf(-4.0f, v2, v1));
    const float t3 = fmaf(-2.0f, v3, v2) + v4, t4 = v2 - v4;
    const float t5 = fmaf(-2.0f, v4, v3) + v5, t6 = fmaf(3.0f, v3, fmaf(-4.0f, v4, v5));
    const float b1 = fmaf((13.0f / 12.0f) * t1, t1, (0.25f * t2) * t2);
    const float b2 = fmaf((13.0f / 12.0f) * t3, t3, (0.25f * t4) * t4);
    const float b3 = fmaf((13.0f / 12.0f) * t5, t5, (0.25f * t6) * t6);
    float a1, a2, a3;
    if (z) {  // WENO5-Z, exponent 1
        const float tau = fabsf(b1 - b3);
        a1 = fmaf(0.1f * tau, rcp(b1 + eps), 0.1f);
        a2 = fmaf(0.6f * tau, rcp(b2 + eps), 0.6f);
        a3 = fmaf(0.3f * tau, rcp(b3 + eps), 0.3f);
    } else {  // WENO5-JS
        const float i1 = rcp(eps + b1), i2 = rcp(eps + b2), i3 = rcp(eps + b3);
        a1 = (0.1f * i1) * i1;
        a2 = (0.6f * i2) * i2;
        a3 = (0.3f * i3) * i3;
    }
    const float is = rcp(a1 + a2 + a3) * (1.0f / 6.0f);
    // candidate stencils 2 v1 - 7 v2 + 11 v3, -v2 + 5 v3 + 2 v4, 2 v3 + 5 v4 - v5
    const float p1 = fmaf(2.0f, v1, fmaf(-7.0f, v2, 11.0f * v3));
    const float p2 = fmaf(5.0f, v3, fmaf(2.0f, v4, -v2));
    const float p3 = fmaf(2.0f, v3, fmaf(5.0f, v4, -v5));
    return fmaf(a1, p1, fmaf(a2, p2, a3 * p3)) * is;
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
        l = weno5_f(s[0], s[1], s[2], s[3], s[4], CHAR);
        r = weno5_f(s[5], s[4], s[3], s[2], s[1], CHAR);
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
__device__ __forceinline__ Basis64 face_basis64(double g, const Prim64& A, const Prim64& B) {
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
    const double igm1 = 1.0 / gm1;
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

// Face states U_ref + dU from the six-cell stencil (FP64 in and out).  Characteristic
// branch: w = L dU in FP64, WENO/TVD on w in FP32, dU = R w in FP64.
template <int DIR, int ORDER, bool CHAR>
__device__ __forceinline__ void reconstruct_face(double g, double delta, const double st[6][4],
                                                 double UL[4], double UR[4]) {
    double ref[4];
#pragma unroll
    for (int v = 0; v < 4; v++) ref[v] = 0.5 * (st[2][v] + st[3][v]);
    double d[6][4];
#pragma unroll
    for (int n = 0; n < 6; n++)
#pragma unroll
        for (int v = 0; v < 4; v++) d[n][v] = st[n][v] - ref[v];
    const float df = (float)delta;
    if constexpr (CHAR) {
        const Basis64 e = face_basis64<DIR>(g, prim64(g, st[2]), prim64(g, st[3]));
        float w[6][4];
#pragma unroll
        for (int n = 0; n < 6; n++) project64<DIR>(e, d[n], w[n]);
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
            const float s[6] = {(float)d[0][v], (float)d[1][v], (float)d[2][v],
                                (float)d[3][v], (float)d[4][v], (float)d[5][v]};
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

template <int DIR>
__device__ __forceinline__ void roe(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    double FL[4], FR[4];
    phys64<DIR>(cl, L, FL);
    phys64<DIR>(cr, R, FR);
    const float gm1 = (float)(g - 1.0), gf = (float)g;
    const float sL = fsqrt((float)L.r), sR = fsqrt((float)R.r);
    const float iss = rcp(sL + sR);
    const float rbar = sL * sR;
    const float uL = (float)L.u, uR = (float)R.u, vL = (float)L.v, vR = (float)R.v;
    const float ubar = (sL * uL + sR * uR) * iss, vbar = (sL * vL + sR * vR) * iss;
    const float aL2 = gf * (float)(L.p * L.ir), aR2 = gf * (float)(R.p * R.ir);
    const float du = uR - uL, dv = vR - vL;
    const float c2 = fmaxf((sL * aL2 + sR * aR2) * iss + 0.5f * gm1 * rbar * iss * iss * (du * du + dv * dv),
                           1.0e-30f);
    const float cbar = fsqrt(c2);
    const float q2 = ubar * ubar + vbar * vbar;
    const float Hbar = c2 * rcp(gm1) + 0.5f * q2;
    const float nbar = DIR == 1 ? ubar : vbar;
    const float lam0 = fabsf(nbar - cbar), lam1 = fabsf(nbar), lam3 = fabsf(nbar + cbar);
    const float dr = (float)(cr[0] - cl[0]), dp = (float)(R.p - L.p);
    const float i2cs = 0.5f * rcp(c2), ics = rcp(c2);
    const float dn = DIR == 1 ? du : dv, dtg = DIR == 1 ? dv : du;
    const float a0 = (dp - rbar * cbar * dn) * i2cs, a1 = dr - dp * ics, a2 = rbar * dtg,
                a3 = (dp + rbar * cbar * dn) * i2cs;
    const float t0 = 0.5f * lam0 * a0, t1 = 0.5f * lam1 * a1, t2 = 0.5f * lam1 * a2, t3 = 0.5f * lam3 * a3;
    float D[4];
    if constexpr (DIR == 1) {
        D[0] = t0 + t1 + t3;
        D[1] = t0 * (ubar - cbar) + t1 * ubar + t3 * (ubar + cbar);
        D[2] = (t0 + t1 + t3) * vbar + t2;
        D[3] = t0 * (Hbar - ubar * cbar) + t1 * 0.5f * q2 + t2 * vbar + t3 * (Hbar + ubar * cbar);
    } else {
        D[0] = t0 + t1 + t3;
        D[1] = (t0 + t1 + t3) * ubar + t2;
        D[2] = t0 * (vbar - cbar) + t1 * vbar + t3 * (vbar + cbar);
        D[3] = t0 * (Hbar - vbar * cbar) + t1 * 0.5f * q2 + t2 * ubar + t3 * (Hbar + vbar * cbar);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]) - (double)D[k];
}

template <int DIR>
__device__ __forceinline__ void llf(double g, const double cl[4], const double cr[4], double F[4]) {
    const Prim64 L = prim64(g, cl), R = prim64(g, cr);
    double FL[4], FR[4];
    phys64<DIR>(cl, L, FL);
    phys64<DIR>(cr, R, FR);
    const float gf = (float)g;
    const float aL = fsqrt(fmaxf(gf * (float)(L.p * L.ir), 0.0f)), aR = fsqrt(fmaxf(gf * (float)(R.p * R.ir), 0.0f));
    const float qL = (float)(DIR == 1 ? L.u : L.v), qR = (float)(DIR == 1 ? R.u : R.v);
    const float s = fmaxf(fabsf(qL) + aL, fabsf(qR) + aR);
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]) - (double)(0.5f * s * (float)(cr[k] - cl[k]));
}

template <int DIR, int SOLVER>
__device__ __forceinline__ void riemann(double g, const double cl[4], const double cr[4], double F[4]) {
    if constexpr (SOLVER == SOLVER_HLL) hll<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_HLLC) hllc<DIR>(g, cl, cr, F);
    else if constexpr (SOLVER == SOLVER_ROE) roe<DIR>(g, cl, cr, F);
    else llf<DIR>(g, cl, cr, F);
}

}  // namespace mx
}  // namespace fvm
