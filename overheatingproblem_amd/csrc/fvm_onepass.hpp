// fvm_onepass.hpp -- one RK stage in one pass: x faces, y faces, residual and stage
// combine in a single marching kernel, so that no face flux round-trips through HBM.
//
// A wavefront takes kOpRows consecutive cell rows of one 63-cell strip: lane t owns the y
// face f0 + t and, for t < 63, the cell c = f0 + 1 + t and the x faces of its column.  It
// keeps the six cell rows i-2 .. i+3 that the x faces of row i need in a ring of LDS slots
// (row r in slot r mod 6); a slot holds the row's window of cells f0-2 .. f0+66 (4
// variables), which is also the y faces' stencil of that row.  Per row i:
//   * y face f0 + t of row i from the ring's row-i window;
//   * wait for row i+3 (requested one row earlier), read the x-face stencil (rows
//     i-2 .. i+3, the lane's column) from the ring, then request row i+4 into the slot
//     just read for the last time (direct buffer loads into LDS: no VGPR holds the row)
//     and the lane's U^n cell (into registers);
//   * x face i (between rows i and i+1) is the cell's upper x flux; its lower one is the
//     previous row's, carried in registers (the first row's is computed before the loop,
//     so a wavefront computes kOpRows + 1 x-face rows);
//   * the upper y flux by shuffle; update the cell, store it.
// Computing the y face before the x face keeps the x-face stencil out of the registers
// while the y face is in flight, and puts the y face between a row request and its wait.
// A face that FaceOp::compute declines (the mixed form's faces whose FP32 states are not
// admissible) is deferred with FaceOp::defer: its two cells are left to the fix-up kernels,
// and the fluxes those kernels read -- every other face of a cell next to a deferred face
// -- are stored in F and G.  No other flux is stored.  The face and cell arithmetic is the
// two-sweep path's (x sweep, then yflux_update_march), so the results are identical.
#pragma once

#include "fvm_buffer.hpp"
#include "fvm_fused.hpp"
#include "fvm_kernels.hpp"

namespace fvm {

#ifndef RPFVM_OP_ROWS
#define RPFVM_OP_ROWS 16
#endif
#ifndef RPFVM_OP_WAVES
#define RPFVM_OP_WAVES 2
#endif
constexpr int kOpRows = RPFVM_OP_ROWS;    // cell rows per wavefront
constexpr int kOpWaves = RPFVM_OP_WAVES;  // wavefronts (adjacent strips) per workgroup
constexpr int kOpPitch = kStrip + 6;      // doubles per variable of a ring slot: cells f0-2 .. f0+66
constexpr int kOpTail = 2 * kOpPitch - 128;  // dwords of a slot row past two 64-dword loads
// doubles of LDS per wavefront: 6 slots x 4 variables (13,248 B: 12 wavefronts, 3 per SIMD,
// fill the CU's 160 KB)
constexpr int kOpLds = 6 * 4 * kOpPitch;
// Minimum waves/SIMD the one-pass kernels are compiled for: the LDS ring allows 3.  With
// characteristic WENO5 faces in both directions a 3-wave register budget (168 VGPRs)
// spills 20-30 registers and the kernel runs 23% slower than the two-sweep stage; at 2
// waves (256 VGPRs) it does not spill.  RPFVM_OP_OCC overrides both (tuning builds).
template <int ORDER, bool CHAR>
constexpr int op_occupancy() {
#ifdef RPFVM_OP_OCC
    return RPFVM_OP_OCC;
#else
    return ORDER == 5 && CHAR ? 2 : 3;
#endif
}

template <typename OpX, typename OpY>
__device__ __forceinline__ void stage_march(const double* Un, const double* Y, double* out, double* F,
                                            double* G, const Geom& g, const Phys& ph, double dt,
                                            int kind, const Repair& rp, const DtSlots& ds, OpX& opx,
                                            OpY& opy, double* lds) {
    const int t = threadIdx.x;
    const Tile2 b = xcd_tile();
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int strip = __builtin_amdgcn_readfirstlane(b.x * kOpWaves + wv);
    const int i0 = __builtin_amdgcn_readfirstlane(GC + b.y * kOpRows);
    if (strip * kStrip >= g.C - 2 * GC) return;  // whole wavefront
    const int i1 = min(i0 + kOpRows, g.R - GC);
    if (i0 >= i1) return;  // whole wavefront
    double* const ring = lds + wv * kOpLds;  // [6][4][kOpPitch]
    const int f0 = __builtin_amdgcn_readfirstlane(GC - 1 + strip * kStrip);
    const int f_lane = f0 + t;
    const bool face_ok = f_lane <= g.C - GC - 1;
    const int c_lane = f_lane + 1;
    const bool cell_ok = t < kStrip && face_ok && c_lane <= g.C - GC - 1;
    // window cells f0-2 .. f0-2+ncell-1 (<= C-1): lanes past the last valid dword load it
    // again (never read)
    const int win_words = 2 * min(kOpPitch, g.C - (f0 - 2));
    int woff[3];
#pragma unroll
    for (int k = 0; k < 3; k++) woff[k] = 4 * min(64 * k + t, win_words - 1);
    const bool need_un = kind != UPD_EULER;
    const unsigned bytes = (unsigned)array_bytes(g);
    const i32x4 ry = buffer_rsrc(Y, bytes), ru = buffer_rsrc(Un, bytes), ro = buffer_rsrc(out, bytes);
    const i32x4 rf = buffer_rsrc(F, bytes), rg = buffer_rsrc(G, bytes);
    const int fb = f_lane * (int)sizeof(double);  // byte offset of the lane's y face in a row
    const int cb = fb + (int)sizeof(double);      // ... and of its cell (and x faces)
    const unsigned pl8 = (unsigned)(g.plane * sizeof(double)), ld8 = (unsigned)g.ld * 8u;
    // rows r >= i0 - 3 >= 1 and r <= i1 + 3 <= R - 1 only: the offsets stay in the array
    auto fetch_row = [&](int r) {
        const unsigned o = (unsigned)r * ld8 + (unsigned)(f0 - 2) * 8u;
        unsigned* slot = reinterpret_cast<unsigned*>(ring + (r % 6) * 4 * kOpPitch);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            unsigned* d = slot + v * 2 * kOpPitch;
            buffer_load_lds_dword(ry, (lds_void*)d, 4, woff[0], (int)(o + v * pl8), 0, 0);
            buffer_load_lds_dword(ry, (lds_void*)(d + 64), 4, woff[1], (int)(o + v * pl8), 0, 0);
        }
        if (t < kOpTail) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                buffer_load_lds_dword(ry, (lds_void*)(slot + v * 2 * kOpPitch + 128), 4, woff[2],
                                      (int)(o + v * pl8), 0, 0);
        }
    };
    // the x-face stencil of face row i: rows i-2 .. i+3 in the lane's column (window index
    // t + 3)
    auto read_xstencil = [&](int i, double st[6][4]) {
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const double* slot = ring + ((i - 2 + n) % 6) * 4 * kOpPitch;
#pragma unroll
            for (int v = 0; v < 4; v++) st[n][v] = slot[v * kOpPitch + t + 3];
        }
    };

    // prologue: rows i0-3 .. i0+2, the x face i0 - 1 (the first row's lower face), then row
    // i0 + 3 into the slot of row i0 - 3
    for (int r = i0 - 3; r <= i0 + 2; r++) fetch_row(r);
    wait_vmem();
    double fl[4] = {0.0, 0.0, 0.0, 0.0};
    int xlo = 0;  // the lower x face was deferred
    {
        double st[6][4];
        read_xstencil(i0 - 1, st);
        wait_lds();
        fetch_row(i0 + 3);
        if (cell_ok && !opx.compute(st, fl)) {
            xlo = 1;
            opx.defer((size_t)(i0 - 1) * g.ld + c_lane);
        }
    }
    for (int row = i0; row < i1; row++) {
        // the row and lane indices pass through empty asm statements so that the compiler
        // forms each address afresh instead of carrying per-lane pointers across the loop
        int i = row, f = f_lane, c = c_lane;
        asm volatile("" : "+s"(i));
        asm volatile("" : "+v"(f), "+v"(c));
        const unsigned rowb = (unsigned)i * ld8;
        // y face f of row i first: its window (slot i mod 6) arrived three rows ago, and
        // the x-face stencil does not occupy registers meanwhile
        const double* w = ring + (i % 6) * 4 * kOpPitch;
        double gf[4] = {0.0, 0.0, 0.0, 0.0};
        int yd = 0;
        if (face_ok) {
            double st[6][4];
#pragma unroll
            for (int n = 0; n < 6; n++)
#pragma unroll
                for (int v = 0; v < 4; v++) st[n][v] = w[v * kOpPitch + t + n];
            if (!opy.compute(st, gf)) {
                yd = 1;
                opy.defer((size_t)i * g.ld + f);
            }
        }
        // then x face i: row i + 3 (requested a row ago) and the stores of row i - 1 are
        // complete after this wait, which the y face has covered
        wait_vmem();
        double fu[4] = {0.0, 0.0, 0.0, 0.0};
        int xhi = 0;
        double un[4] = {0.0, 0.0, 0.0, 0.0};
        {
            double st[6][4];
            read_xstencil(i, st);
            wait_lds();  // slot (i - 2) mod 6 is read for the last time
            if (i + 1 < i1) fetch_row(i + 4);
            if (cell_ok && need_un) {
#pragma unroll
                for (int v = 0; v < 4; v++) un[v] = buffer_load_f64(ru, cb, (int)(rowb + v * pl8), 0);
            }
            if (cell_ok && !opx.compute(st, fu)) {
                xhi = 1;
                opx.defer((size_t)i * g.ld + c);
            }
        }
        double gu[4];
#pragma unroll
        for (int v = 0; v < 4; v++) gu[v] = __shfl_down(gf[v], 1, 64);
        const int yd_up = __shfl_down(yd, 1, 64);  // face f + 1 deferred
        // the cell is left to the fix-up when one of its four faces was deferred
        const int pend = cell_ok && (xlo | xhi | yd | yd_up) ? 1 : 0;
        const int pend_dn = __shfl_up(pend, 1, 64);  // the cell below face f (lane t - 1's)
        // the fix-up of a pending cell reads its other faces from F and G
        if (face_ok && !yd && (pend || pend_dn)) {
#pragma unroll
            for (int v = 0; v < 4; v++) buffer_store_f64(gf[v], rg, fb, (int)(rowb + v * pl8), 0);
        }
        double lx = 0.0, ly = 0.0;
        if (cell_ok) {
            double y[4], W[4], R[4];
#pragma unroll
            for (int v = 0; v < 4; v++) y[v] = w[v * kOpPitch + t + 3];
            cell_update_vals(ph, dt, kind, y, un, fl, fu, gf, gu, W, R);
#pragma unroll
            for (int v = 0; v < 4; v++) buffer_store_f64(R[v], ro, cb, (int)(rowb + v * pl8), 0);
            if (pend) {
                if (!xlo) {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        buffer_store_f64(fl[v], rf, cb, (int)(rowb - ld8 + v * pl8), 0);
                }
                if (!xhi) {
#pragma unroll
                    for (int v = 0; v < 4; v++) buffer_store_f64(fu[v], rf, cb, (int)(rowb + v * pl8), 0);
                }
            }
            if (ds.slot) {  // see yflux_update_body
                double sx, sy;
                cell_speeds(ph.gamma, R[0], R[1], R[2], R[3], sx, sy);
                if (sx > lx) lx = sx;
                if (sy > ly) ly = sy;
            }
            if (rp.cnt && !pend) {
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
            if (t == 0) {
                const size_t sw = (size_t)(i - GC) * ds.nstrips + (size_t)strip;  // ds.rows == 1
                ds.slot[2 * sw] = (unsigned long long)__double_as_longlong(lx);
                ds.slot[2 * sw + 1] = (unsigned long long)__double_as_longlong(ly);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) fl[v] = fu[v];
        xlo = xhi;
    }
}

inline dim3 stage_march_grid(const Geom& g) {
    const int nj = g.C - 2 * GC, ni = g.R - 2 * GC;
    const int ns = (nj + kStrip - 1) / kStrip;
    return dim3((ns + kOpWaves - 1) / kOpWaves, (ni + kOpRows - 1) / kOpRows);
}
inline dim3 stage_march_block() { return dim3(64, kOpWaves); }

}  // namespace fvm
