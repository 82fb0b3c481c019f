// fvm_fused.hpp -- y-face flux fused with the residual and RK stage combine.
//
// A 64-lane row of a block computes the y faces f = jb .. jb+63 of one grid row and
// updates the 63 cells c = f + 1 between them: lane t owns face f = jb + t (the cell's
// lower face, G_{c-1/2}) and receives the upper face G_{c+1/2} from lane t + 1 with a
// wavefront shuffle, so the face fluxes never round-trip through memory before use (they
// are still stored, for the positivity safeguard and the mixed form's FP64 recompute).
// The cell arithmetic is exactly cell_update's (Space_Discrete_Item + stage combine), so
// the fused path is bit-identical to k_flux<2> + k_update.
#pragma once

#include "fvm_buffer.hpp"
#include "fvm_kernels.hpp"
#include "fvm_math.hpp"

namespace fvm {


// Block (x, y) of a 2-D grid, remapped so that each XCD works on one contiguous range of
// the grid.  Workgroups are dispatched to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2: with the identity map, blocks that share stencil rows and cache
// lines (the neighbours in x and y) land on different XCDs and fetch the shared lines
// from HBM once each.  Here linear id L runs as logical block (L % 8) * (N / 8) + L / 8,
// so neighbouring logical blocks run on the same XCD at nearly the same time (a tail of
// N % 8 blocks keeps the identity map).
struct Tile2 {
    int x, y;
};
__device__ __forceinline__ Tile2 xcd_tile() {
    const unsigned nx = gridDim.x, n = gridDim.x * gridDim.y;
    const unsigned lin = blockIdx.y * nx + blockIdx.x;
    const unsigned per = n / 8;
    const unsigned l = lin < per * 8 ? (lin % 8) * per + lin / 8 : lin;
    return Tile2{(int)(l % nx), (int)(l / nx)};
}

// |u| + a and |v| + a of one cell, formed as Get_Delta_T_2D does (include/function.h:
// 144-175); the CFL reduction's max ignores NaN like the reference's `if (s > S) S = s`.
__device__ __forceinline__ void cell_speeds(double gam, double rho, double ru, double rv,
                                            double E, double& lx, double& ly) {
    const double u = ru / rho;
    const double v = rv / rho;
    const double p = (E - 0.5 * rho * (u * u + v * v)) * (gam - 1.0);
    const double a = sqrt(gam * p / rho);
    lx = fabs(u) + a;
    ly = fabs(v) + a;
}
__device__ __forceinline__ double nanmax(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = nanmax(x, __shfl_xor(x, o, 64));
    return x;
}

// Space_Discrete_Item (include/scheme.h:3758-3777) with Source_Gravity (:3723-3751)
// fused with the RK stage combine (include/time_advance.h:697, 757-777, 832-872) for
// cell (i, j).  `out` may alias `Un` (the last stage writes the state in place): each
// thread reads its own cell before writing it.  Returns whether the forward-Euler
// state W = Y + dt L has rho > 0 and p > 0 (used by the positivity safeguard).
__device__ __forceinline__ bool cell_update(const double* Un, const double* Y, double* out,
                                            const double* __restrict__ F,
                                            const double* __restrict__ G, const Geom& g,
                                            const Phys& ph, double dt, int kind, int i, int j) {
    const size_t c = (size_t)i * g.ld + j;
    double S[4] = {0.0, 0.0, 0.0, 0.0};
    if (ph.source) {
        const double rho = Y[c];
        const double v = Y[2 * g.plane + c] / rho;
        S[2] = rho * ph.gravity;
        S[3] = rho * ph.gravity * v;
    }
    const double idx = ph.idx, idy = ph.idy;
    double W[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t o = (size_t)k * g.plane + c;
        const double L = -(F[o] - F[o - g.ld]) * idx - (G[o] - G[o - 1]) * idy + S[k];
        const double y = Y[o];
        W[k] = y + dt * L;
        double r;
        switch (kind) {
            case UPD_RESIDUAL: r = L; break;                                   // L only
            case UPD_EULER: r = y + dt * L; break;                             // U1 = Y + dt L
            case UPD_RK2_2: r = 0.5 * Un[o] + 0.5 * (y + dt * L); break;
            case UPD_RK3_2: r = (3.0 / 4.0) * Un[o] + (1.0 / 4.0) * (y + dt * L); break;
            default: r = (1.0 / 3.0) * Un[o] + (2.0 / 3.0) * (y + dt * L); break;  // UPD_RK3_3
        }
        out[o] = r;
    }
    const Consts kc{ph.gamma, ph.gm1, ph.gr, Recip{}};
    return admissible(kc, W);
}


// Residual + stage combine of cell (i, c) with the y-face fluxes given in registers
// (gl = G_{c-1/2}, gu = G_{c+1/2}) and the cell's own state y[] (already loaded).
// Leaves the forward-Euler state Y + dt L in W (for the positivity safeguard's test).
__device__ __forceinline__ void cell_update_regs(const double* Un, double* out,
                                                 const double* __restrict__ F, const Geom& g,
                                                 const Phys& ph, double dt, int kind, int i,
                                                 int c, const double y[4], const double gl[4],
                                                 const double gu[4], double W[4], double R[4]) {
    const size_t o0 = (size_t)i * g.ld + c;
    double S[4] = {0.0, 0.0, 0.0, 0.0};
    if (ph.source) {
        const double rho = y[0];
        const double v = y[2] / rho;
        S[2] = rho * ph.gravity;
        S[3] = rho * ph.gravity * v;
    }
    const double idx = ph.idx, idy = ph.idy;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t o = (size_t)k * g.plane + o0;
        const double L = -(F[o] - F[o - g.ld]) * idx - (gu[k] - gl[k]) * idy + S[k];
        W[k] = y[k] + dt * L;
        double r;
        switch (kind) {
            case UPD_EULER: r = y[k] + dt * L; break;
            case UPD_RK2_2: r = 0.5 * Un[o] + 0.5 * (y[k] + dt * L); break;
            case UPD_RK3_2: r = (3.0 / 4.0) * Un[o] + (1.0 / 4.0) * (y[k] + dt * L); break;
            default: r = (1.0 / 3.0) * Un[o] + (2.0 / 3.0) * (y[k] + dt * L); break;
        }
        out[o] = r;
        R[k] = r;
    }
}

// cell_update_regs with the x fluxes (fl = F_{i-1/2}, fu = F_{i+1/2}) and U^n already
// in registers (un is not read by UPD_EULER); the new state is left in R, not stored.
__device__ __forceinline__ void cell_update_vals(const Phys& ph, double dt, int kind,
                                                 const double y[4], const double un[4],
                                                 const double fl[4], const double fu[4],
                                                 const double gl[4], const double gu[4],
                                                 double W[4], double R[4]) {
    double S[4] = {0.0, 0.0, 0.0, 0.0};
    if (ph.source) {
        const double rho = y[0];
        const double v = y[2] / rho;
        S[2] = rho * ph.gravity;
        S[3] = rho * ph.gravity * v;
    }
    const double idx = ph.idx, idy = ph.idy;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double L = -(fu[k] - fl[k]) * idx - (gu[k] - gl[k]) * idy + S[k];
        W[k] = y[k] + dt * L;
        double r;
        switch (kind) {
            case UPD_EULER: r = y[k] + dt * L; break;
            case UPD_RK2_2: r = 0.5 * un[k] + 0.5 * (y[k] + dt * L); break;
            case UPD_RK3_2: r = (3.0 / 4.0) * un[k] + (1.0 / 4.0) * (y[k] + dt * L); break;
            default: r = (1.0 / 3.0) * un[k] + (2.0 / 3.0) * (y[k] + dt * L); break;
        }
        R[k] = r;
    }
}

// cell_update_vals, storing the new state of cell (i, c) to out.
__device__ __forceinline__ void cell_update_pre(double* out, const Geom& g, const Phys& ph,
                                                double dt, int kind, int i, int c,
                                                const double y[4], const double un[4],
                                                const double fl[4], const double fu[4],
                                                const double gl[4], const double gu[4],
                                                double W[4], double R[4]) {
    const size_t o0 = (size_t)i * g.ld + c;
    cell_update_vals(ph, dt, kind, y, un, fl, fu, gl, gu, W, R);
#pragma unroll
    for (int k = 0; k < 4; k++) out[(size_t)k * g.plane + o0] = R[k];
}

// FaceOp::compute(st, f) returns false when the face must be recomputed later (mixed
// form); FaceOp::defer(c0) records such a face.  FaceOp::kXDeferred: the x sweep may have
// deferred x faces too, writing NaN into F for them (the mixed form); a cell next to such
// a face is left to the fix-up like one next to a deferred y face.  With PREFETCH the cell's x fluxes and
// U^n are loaded before the face arithmetic, so that their latency overlaps it (24 more
// VGPRs: it pays where the kernel keeps its occupancy, the mixed form at 4 waves/SIMD,
// and loses where it does not, the FP64 form).
template <bool PREFETCH, typename FaceOp>
__device__ __forceinline__ void yflux_update_body(const double* Un, const double* Y, double* out,
                                                  const double* __restrict__ F, double* G,
                                                  const Geom& g, const Phys& ph, double dt,
                                                  int kind, const Repair& rp, const DtSlots& ds,
                                                  FaceOp& op) {
    static_assert(!FaceOp::kXDeferred || PREFETCH, "the x-face test reads the prefetched fluxes");
    const int t = threadIdx.x;
    const Tile2 b = xcd_tile();
    const int i = GC + b.y * blockDim.y + threadIdx.y;
    const int f = GC - 1 + b.x * kStrip + t;
    const bool face_ok = i < g.R - GC && f <= g.C - GC - 1;
    const int c = f + 1;
    const bool cell_ok = t < kStrip && face_ok && c <= g.C - GC - 1;
    double fl[4] = {0.0, 0.0, 0.0, 0.0}, fu[4] = {0.0, 0.0, 0.0, 0.0};
    double un[4] = {0.0, 0.0, 0.0, 0.0};
    if (PREFETCH && cell_ok) {
        const size_t o0 = (size_t)i * g.ld + c;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            fu[k] = F[(size_t)k * g.plane + o0];
            fl[k] = F[(size_t)k * g.plane + o0 - g.ld];
            if (kind != UPD_EULER) un[k] = Un[(size_t)k * g.plane + o0];
        }
    }
    double gf[4] = {0.0, 0.0, 0.0, 0.0};
    int pending = 0;
    if (face_ok) {
        const size_t c0 = (size_t)i * g.ld + f;
        double st[6][4];
#pragma unroll
        for (int n = 0; n < 6; n++)
#pragma unroll
            for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * g.plane + c0 + (n - 2)];
        if (op.compute(st, gf)) {
#pragma unroll
            for (int v = 0; v < 4; v++) G[(size_t)v * g.plane + c0] = gf[v];
        } else {
            op.defer(c0);
            pending = 1;
        }
    }
    double gu[4];
#pragma unroll
    for (int v = 0; v < 4; v++) gu[v] = __shfl_down(gf[v], 1, 64);
    const int pending_up = __shfl_down(pending, 1, 64);
    // last stage: the cell's wave speeds for the next step's dt, "if (s > S) S = s" from
    // S = 0 in each lane (a NaN speed is skipped, as in the full pass and Get_Delta_T_2D,
    // instead of hiding the row's other maxima)
    double lx = 0.0, ly = 0.0;
    if (cell_ok) {
        // the cell's stage input, read again (a cache hit) rather than kept live in
        // registers across the face arithmetic
        double y[4];
#pragma unroll
        for (int v = 0; v < 4; v++) y[v] = Y[(size_t)v * g.plane + (size_t)i * g.ld + c];
        double W[4], R[4];
        if constexpr (PREFETCH) cell_update_pre(out, g, ph, dt, kind, i, c, y, un, fl, fu, gf, gu, W, R);
        else cell_update_regs(Un, out, F, g, ph, dt, kind, i, c, y, gf, gu, W, R);
        if (ds.slot) {
            double sx, sy;
            cell_speeds(ph.gamma, R[0], R[1], R[2], R[3], sx, sy);
            if (sx > lx) lx = sx;
            if (sy > ly) ly = sy;
        }
        bool cpend = pending || pending_up;
        if constexpr (FaceOp::kXDeferred) cpend = cpend || isnan(fl[0]) || isnan(fu[0]);
        if (rp.cnt && !cpend) {
            const Consts kc{ph.gamma, ph.gm1, ph.gr, Recip{}};
            if (!admissible(kc, W)) {
                const unsigned k = atomicAdd(&rp.cnt[0], 1u);
                rp.list[0][k] = (i - GC) * rp.ny + (c - GC);
            }
        }
    }
    if (ds.slot) {
        lx = wave_max(lx);
        ly = wave_max(ly);
        if (t == 0 && i < g.R - GC) {
            // one row per wavefront: ds.rows == 1
            const size_t w = (size_t)(i - GC) * ds.nstrips + (size_t)((f + 1 - GC) / kStrip);
            ds.slot[2 * w] = (unsigned long long)__double_as_longlong(lx);
            ds.slot[2 * w + 1] = (unsigned long long)__double_as_longlong(ly);
        }
    }
}

constexpr int kFusedRows = RPFVM_FUSED_ROWS;  // grid rows per workgroup of the fused kernels

// ------------------------------------------------------------------ marching form
// A wavefront takes kMarchRows consecutive rows of one 63-cell strip and walks down them.
// While it computes row i, the stencil window of row i + 1 (cells f0-2 .. f0+66, 4
// variables) is on its way into LDS by direct loads (buffer_load_dword ... lds: each lane
// moves one dword and no VGPR holds it), and the x-flux and U^n rows of i + 1
// follow as soon as row i's update has read its own.  A cell's lower x-flux row is the
// upper row of the cell above, carried in registers.  The wait for those loads sits after
// the face arithmetic and before the row's first store: by then they have arrived, and
// no store of the row is outstanding yet.  The arithmetic and its inputs are
// yflux_update_body's, so the results are identical.
#ifndef RPFVM_MARCH_ROWS
#define RPFVM_MARCH_ROWS 4
#endif
#ifndef RPFVM_MARCH_WAVES
#define RPFVM_MARCH_WAVES 2
#endif
constexpr int kMarchRows = RPFVM_MARCH_ROWS;    // rows per wavefront
constexpr int kMarchWaves = RPFVM_MARCH_WAVES;  // wavefronts (adjacent strips) per workgroup
constexpr int kWinCells = kStrip + 6;  // stencil window of a row: cells f0-2 .. f0+66
constexpr int kWinPitch = 96;          // doubles per variable of a window in LDS (3 x 64 dwords)
constexpr int kRowPitch = 64;          // doubles per variable of an F or U^n row (2 x 64 dwords)
// doubles of LDS per wavefront: two windows, one F row, one U^n row (10 KB: 16 wavefronts
// fill the CU's 160 KB, 4 per SIMD)
constexpr int kMarchLds = 2 * 4 * kWinPitch + 2 * 4 * kRowPitch;

// Compiler barriers around the hand-placed waits: no memory access moves across them.
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void wait_lds() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// NI x 64 dwords into dst in LDS: dword 64k + t of dst gets the dword at byte
// soffset + voff[k] of the buffer, where voff[k] = 4 min(64k + t, valid dwords - 1) (every
// lane loads; the copies past the valid range are never read).
template <int NI>
__device__ __forceinline__ void lds_fetch(i32x4 rsrc, int soffset, double* dst, const int (&voff)[NI]) {
    unsigned* d = reinterpret_cast<unsigned*>(dst);
#pragma unroll
    for (int k = 0; k < NI; k++) buffer_load_lds_dword(rsrc, (lds_void*)(d + 64 * k), 4, voff[k], soffset, 0, 0);
}

// The marching form addresses every array through a buffer descriptor (buffer_fits).

template <typename FaceOp>
__device__ __forceinline__ void yflux_update_march(const double* Un, const double* Y, double* out,
                                                   const double* __restrict__ F, double* G,
                                                   const Geom& g, const Phys& ph, double dt,
                                                   int kind, const Repair& rp, const DtSlots& ds,
                                                   FaceOp& op, double* lds) {
    const int t = threadIdx.x;
    const Tile2 b = xcd_tile();
    // wave-uniform values in SGPRs (xcd_tile divides in the VALU): every address below is
    // then an SGPR base plus a lane offset
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    // the wavefronts of a workgroup take adjacent strips of the same rows, so the cache
    // lines at their common edge are fetched once for both
    const int strip = __builtin_amdgcn_readfirstlane(b.x * kMarchWaves + wv);
    const int i0 = __builtin_amdgcn_readfirstlane(GC + b.y * kMarchRows);
    if (strip * kStrip >= g.C - 2 * GC) return;  // whole wavefront
    const int i1 = min(i0 + kMarchRows, g.R - GC);
    if (i0 >= i1) return;  // whole wavefront
    double* const win = lds + wv * kMarchLds;      // [2][4][kWinPitch]
    double* const frow = win + 2 * 4 * kWinPitch;  // [4][kRowPitch]
    double* const urow = frow + 4 * kRowPitch;     // [4][kRowPitch]
    const int f0 = __builtin_amdgcn_readfirstlane(GC - 1 + strip * kStrip);
    const int f_lane = f0 + t;
    const bool face_ok = f_lane <= g.C - GC - 1;
    const int c_lane = f_lane + 1;
    const bool cell_ok = t < kStrip && face_ok && c_lane <= g.C - GC - 1;
    // window cells f0-2 .. f0-2+ncell-1 (<= C-1); row cells f0+1 .. f0+nrow (<= C-GC-1)
    const int win_words = 2 * min(kWinCells, g.C - (f0 - 2));
    const int row_words = 2 * min(kStrip, g.C - GC - 1 - f0);
    int woff[3], roff[2];
#pragma unroll
    for (int k = 0; k < 3; k++) woff[k] = 4 * min(64 * k + t, win_words - 1);
#pragma unroll
    for (int k = 0; k < 2; k++) roff[k] = 4 * min(64 * k + t, row_words - 1);
    const bool need_un = kind != UPD_EULER;
    const unsigned bytes = (unsigned)array_bytes(g);
    const i32x4 ry = buffer_rsrc(Y, bytes), rf = buffer_rsrc(F, bytes), ru = buffer_rsrc(Un, bytes);
    const i32x4 rg = buffer_rsrc(G, bytes), ro = buffer_rsrc(out, bytes);
    const int fb = f_lane * (int)sizeof(double);  // byte offset of the lane's face in a row
    // byte offsets below 4 GB (buffer_fits), formed in unsigned arithmetic
    const unsigned pl8 = (unsigned)(g.plane * sizeof(double)), ld8 = (unsigned)g.ld * 8u;
    auto fetch_window = [&](int i, int buf) {
        const unsigned o = (unsigned)i * ld8 + (unsigned)(f0 - 2) * 8u;
#pragma unroll
        for (int v = 0; v < 4; v++)
            lds_fetch<3>(ry, (int)(o + v * pl8), win + (buf * 4 + v) * kWinPitch, woff);
    };
    auto fetch_rows = [&](int i) {
        const unsigned o = (unsigned)i * ld8 + (unsigned)(f0 + 1) * 8u;
#pragma unroll
        for (int v = 0; v < 4; v++) lds_fetch<2>(rf, (int)(o + v * pl8), frow + v * kRowPitch, roff);
        if (need_un) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                lds_fetch<2>(ru, (int)(o + v * pl8), urow + v * kRowPitch, roff);
        }
    };
    double fl[4] = {0.0, 0.0, 0.0, 0.0};
    if (cell_ok) {
#pragma unroll
        for (int k = 0; k < 4; k++) fl[k] = F[(size_t)k * g.plane + (size_t)(i0 - 1) * g.ld + c_lane];
    }
    fetch_window(i0, 0);
    fetch_rows(i0);
    wait_vmem();
    double lx = 0.0, ly = 0.0;  // the lane's wave-speed maxima over its rows (last stage)
    for (int row = i0; row < i1; row++) {
        // Row and lane indices pass through empty asm statements, so that the compiler
        // forms each store address afresh instead of carrying ~20 per-lane 64-bit
        // pointers across the loop (40 VGPRs).
        int i = row, f = f_lane, c = c_lane;
        asm volatile("" : "+s"(i));
        asm volatile("" : "+v"(f), "+v"(c));
        const int cur = (i - i0) & 1;
        const double* w = win + cur * 4 * kWinPitch;
        if (i + 1 < i1) fetch_window(i + 1, cur ^ 1);  // window cur ^ 1 was last read in row i - 1
        double gf[4] = {0.0, 0.0, 0.0, 0.0};
        bool ok = true;
        if (face_ok) {
            double st[6][4];
#pragma unroll
            for (int n = 0; n < 6; n++)
#pragma unroll
                for (int v = 0; v < 4; v++) st[n][v] = w[v * kWinPitch + t + n];
            ok = op.compute(st, gf);
        }
        wait_vmem();  // window i + 1, the F / U^n rows of i, the stores of row i - 1
        const unsigned rowb = (unsigned)i * ld8;
        const int pending = face_ok && !ok ? 1 : 0;
        if (pending) op.defer((size_t)i * g.ld + f);
        double gu[4];
#pragma unroll
        for (int v = 0; v < 4; v++) gu[v] = __shfl_down(gf[v], 1, 64);
        const int pending_up = __shfl_down(pending, 1, 64);    // face f + 1 deferred
        // the cell is left to the fix-up when one of its faces was deferred: its y faces f,
        // f + 1, or (FaceOp::kXDeferred) an x face, which the x sweep marked with a NaN
        int cpend = cell_ok && (pending | pending_up) ? 1 : 0;
        if constexpr (FaceOp::kXDeferred) {
            if (cell_ok && (isnan(fl[0]) || isnan(frow[t]))) cpend = 1;
        }
        const int cpend_dn = __shfl_up(cpend, 1, 64);  // the cell below face f (lane t - 1's)
        // G: every face (FaceOp::kStoreAllG), or only those the fix-up of a deferred cell
        // reads (the y faces of the cells on either side; the face shared by two strips is
        // computed by both, each seeing its own neighbour cell)
        if (face_ok && ok && (FaceOp::kStoreAllG || cpend || cpend_dn)) {
#pragma unroll
            for (int v = 0; v < 4; v++) buffer_store_f64(gf[v], rg, fb, (int)(rowb + v * pl8), 0);
        }
        if (cell_ok) {
            double y[4], un[4], fu[4], W[4], R[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                y[v] = w[v * kWinPitch + t + 3];
                fu[v] = frow[v * kRowPitch + t];
                un[v] = need_un ? urow[v * kRowPitch + t] : 0.0;
            }
            cell_update_vals(ph, dt, kind, y, un, fl, fu, gf, gu, W, R);
#pragma unroll
            for (int v = 0; v < 4; v++)  // cell c = f + 1: 8 bytes past the face
                buffer_store_f64(R[v], ro, fb + (int)sizeof(double), (int)(rowb + v * pl8), 0);
            if (ds.slot) {  // see yflux_update_body
                double sx, sy;
                cell_speeds(ph.gamma, R[0], R[1], R[2], R[3], sx, sy);
                if (sx > lx) lx = sx;
                if (sy > ly) ly = sy;
            }
            if (rp.cnt && !cpend) {
                const Consts kc{ph.gamma, ph.gm1, ph.gr, Recip{}};
                if (!admissible(kc, W)) {
                    const unsigned k = atomicAdd(&rp.cnt[0], 1u);
                    rp.list[0][k] = (i - GC) * rp.ny + (c - GC);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) fl[v] = fu[v];
        }
        wait_lds();  // this row's reads of the window and the F / U^n rows are complete
        if (i + 1 < i1) fetch_rows(i + 1);
    }
    if (ds.slot) {  // one slot for the wavefront's rows (ds.rows == kMarchRows)
        lx = wave_max(lx);
        ly = wave_max(ly);
        if (t == 0) {
            const size_t sw = (size_t)((i0 - GC) / ds.rows) * ds.nstrips + (size_t)strip;
            ds.slot[2 * sw] = (unsigned long long)__double_as_longlong(lx);
            ds.slot[2 * sw + 1] = (unsigned long long)__double_as_longlong(ly);
        }
    }
}

inline dim3 yflux_march_grid(const Geom& g) {
    const int nj = g.C - 2 * GC, ni = g.R - 2 * GC;
    const int ns = (nj + kStrip - 1) / kStrip;
    return dim3((ns + kMarchWaves - 1) / kMarchWaves, (ni + kMarchRows - 1) / kMarchRows);
}
inline dim3 yflux_march_block() { return dim3(64, kMarchWaves); }

inline dim3 yflux_update_grid(const Geom& g) {
    const int nj = g.C - 2 * GC, ni = g.R - 2 * GC;
    return dim3((nj + kStrip - 1) / kStrip, (ni + kFusedRows - 1) / kFusedRows);
}
inline dim3 yflux_update_block() { return dim3(64, kFusedRows); }

}  // namespace fvm
