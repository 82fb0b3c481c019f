// fvm_kernels.hip -- gfx950 kernels of the FP64 finite-volume step.
//
// Per RK stage (include/time_advance.h:823-872):
//   k_bc_pair  x2   ghost fill of the stage input: sides L + R, then B + T (k_bc_side x4,
//                   one side per launch in order, on grids under 4 interior cells a side)
//   k_flux<1>       x-face fluxes  F[i+1/2][j]: 6-cell stencil -> reconstruction -> Riemann
//   k_flux<2>       y-face fluxes  G[i][j+1/2]
//   k_update        L = -(dF)/dx - (dG)/dy + S and the SSP-RK stage combine
// where a stage writes a separate output, k_flux<2> + k_update run fused as k_yflux_march
// (inputs staged through LDS; k_yflux_update, one row per wavefront, for arrays of 4 GB
// and more), and per step the CFL reduction: k_dt_slots + k_dt_final over the maxima the
// last stage left, or k_dt_partial + k_dt_final over the state (include/function.h:144-175).
//
// Layout in HBM: four planes (rho, rho u, rho v, E) of R rows by `ld` doubles, ld a
// multiple of 16 and the first interior column 128-byte aligned; a wavefront's 64
// lanes walk the contiguous (y) index, so every load and store is coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fvm_bc.hpp"
#include "fvm_buffer.hpp"
#include "fvm_repair.hpp"
#include "fvm_fused.hpp"
#include "fvm_kernels.hpp"
#include "fvm_math.hpp"

namespace fvm {

// --------------------------------------------------------------- ghost fill
// The stage clock as the driver forms it: t, t + dt, t + 0.5 dt.
__device__ __forceinline__ double stage_time(const BcSide& s, const StepCtl* step) {
    return s.c == 0.0 ? step->t0 : step->t0 + s.c * step->dt;
}

__global__ void k_bc_side(double* __restrict__ U, Geom g, BcSide s, const StepCtl* __restrict__ step) {
    if (!step->active) return;
    const int t = s.lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (t > s.hi) return;
    bc_side_point(U, g, s, t, blockIdx.y, stage_time(s, step));
}

// Two opposite sides in one launch (blockIdx.y: variable, then side): left + right touch
// disjoint rows, bottom + top disjoint columns, once the grid has at least 4 interior
// cells each way; launched in the reference's order L + R, then B + T (which read the
// corner cells L and R write).
__global__ void k_bc_pair(double* __restrict__ U, Geom g, BcSide a, BcSide b,
                          const StepCtl* __restrict__ step) {
    if (!step->active) return;
    const bool first = blockIdx.y < 4;
    const int lo = first ? a.lo : b.lo, hi = first ? a.hi : b.hi;
    const int t = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (t > hi) return;
    if (first) bc_side_point(U, g, a, t, blockIdx.y, stage_time(a, step));
    else bc_side_point(U, g, b, t, blockIdx.y - 4, stage_time(b, step));
}

// ------------------------------------------------------------ face fluxes
// One thread per face.  DIR 1: faces (i+1/2, j), i = GC-1..R-GC-1, j = GC..C-GC-1;
// DIR 2: faces (i, j+1/2), i = GC..R-GC-1, j = GC-1..C-GC-1.  The flux is stored at
// the face's left cell, [i][j], as the reference's Flux array.
template <int DIR, int ORDER, bool CHAR, int SOLVER>
__global__ FLUX_BOUNDS void k_flux(const double* __restrict__ Y, double* __restrict__ F,
                                              Geom g, Phys ph, unsigned* __restrict__ faults) {
    if (!step_on(ph)) return;
    const int j = (DIR == 1 ? GC : GC - 1) + blockIdx.x * 64 + threadIdx.x;
    const int i = (DIR == 1 ? GC - 1 : GC) + blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.R - GC || j >= g.C - GC) return;
    const size_t c0 = (size_t)i * g.ld + j;
    const size_t step = DIR == 1 ? (size_t)g.ld : 1;
    double st[6][4];
#pragma unroll
    for (int n = 0; n < 6; n++)
#pragma unroll
        for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * g.plane + c0 + (n - 2) * step];
    const Consts kc = make_consts(ph);
    double UL[4], UR[4];
    reconstruct_face<DIR, ORDER, CHAR>(kc, DIR == 1 ? ph.dx : ph.dy, st, UL, UR);
    if (ph.positivity && !(admissible(kc, UL) && admissible(kc, UR))) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            UL[v] = st[2][v];
            UR[v] = st[3][v];
        }
    }
    double f[4];
    if (riemann<DIR, SOLVER>(kc, UL, UR, f)) atomicOr(faults, 1u);
#pragma unroll
    for (int v = 0; v < 4; v++) F[(size_t)v * g.plane + c0] = f[v];
}

// The x faces in a marching form: a wavefront takes kRingRows face rows of a 64-column
// strip and keeps the 6 cell rows of its stencil in a ring of LDS slots (row r in slot
// r mod 6).  Per face row it reads the stencil from LDS, then requests the one new cell
// row (direct buffer loads into the slot just read for the last time), computes the face
// while that row arrives, and waits for it before the flux store.  Same arithmetic as
// k_flux; used when the arrays are below 4 GB (buffer offsets).
#ifndef RPFVM_RING_ROWS
#define RPFVM_RING_ROWS 8
#endif
#ifndef RPFVM_RING_WAVES
#define RPFVM_RING_WAVES 1
#endif
#ifndef RPFVM_RING_OCC
#define RPFVM_RING_OCC 3
#endif
#ifndef RPFVM_MARCH_OCC
#define RPFVM_MARCH_OCC 3
#endif
constexpr int kRingRows = RPFVM_RING_ROWS;    // face rows per wavefront
constexpr int kRingWaves = RPFVM_RING_WAVES;  // wavefronts per workgroup
constexpr int kRingLds = 6 * 4 * 64;          // doubles per wavefront (12 KB: 3 waves/SIMD)

template <int ORDER, bool CHAR, int SOLVER>
__global__ __launch_bounds__(64 * kRingWaves, RPFVM_RING_OCC) void k_flux_ring(const double* __restrict__ Y,
                                                                 double* __restrict__ F, Geom g,
                                                                 Phys ph, unsigned* __restrict__ faults) {
    __shared__ double lds[kRingWaves * kRingLds];
    if (!step_on(ph)) return;
    const Tile2 b = xcd_tile();
    const int t = threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    // face rows ia .. ib-1 (face row i lies between cell rows i and i + 1)
    const int ia = __builtin_amdgcn_readfirstlane(GC - 1 + (b.y * kRingWaves + wv) * kRingRows);
    const int ib = min(ia + kRingRows, g.R - GC);
    if (ia >= ib) return;  // whole wavefront
    const int j0 = __builtin_amdgcn_readfirstlane(GC + b.x * 64);
    const bool ok = j0 + t <= g.C - GC - 1;
    double* const ring = lds + wv * kRingLds;  // [6][4][64]
    const unsigned bytes = (unsigned)array_bytes(g);
    const i32x4 ry = buffer_rsrc(Y, bytes), rf = buffer_rsrc(F, bytes);
    const unsigned pl8 = (unsigned)g.plane * 8u, ld8 = (unsigned)g.ld * 8u;
    const int row_words = 2 * min(64, g.C - GC - j0);  // the strip's cells of a row
    int roff[2];
#pragma unroll
    for (int k = 0; k < 2; k++) roff[k] = 4 * min(64 * k + t, row_words - 1);
    auto fetch_row = [&](int r) {
        const unsigned o = (unsigned)r * ld8 + (unsigned)j0 * 8u;
        double* slot = ring + (r % 6) * 4 * 64;
#pragma unroll
        for (int v = 0; v < 4; v++) lds_fetch<2>(ry, (int)(o + v * pl8), slot + v * 64, roff);
    };
    for (int r = ia - 2; r <= ia + 3; r++) fetch_row(r);
    wait_vmem();
    const Consts kc = make_consts(ph);
    int fault = 0;
    for (int row = ia; row < ib; row++) {
        int i = row;
        asm volatile("" : "+s"(i));
        double st[6][4];
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const double* slot = ring + ((i - 2 + n) % 6) * 4 * 64;
#pragma unroll
            for (int v = 0; v < 4; v++) st[n][v] = slot[v * 64 + t];
        }
        wait_lds();  // slot (i - 2) mod 6 is read for the last time
        if (i + 1 < ib) fetch_row(i + 4);
        double f[4] = {0.0, 0.0, 0.0, 0.0};
        if (ok) {
            double UL[4], UR[4];
            reconstruct_face<1, ORDER, CHAR>(kc, ph.dx, st, UL, UR);
            if (ph.positivity && !(admissible(kc, UL) && admissible(kc, UR))) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    UL[v] = st[2][v];
                    UR[v] = st[3][v];
                }
            }
            fault |= riemann<1, SOLVER>(kc, UL, UR, f);
        }
        wait_vmem();  // row i + 4, the stores of face row i - 1
        if (ok) {
            const unsigned o = (unsigned)i * ld8 + (unsigned)j0 * 8u;
#pragma unroll
            for (int v = 0; v < 4; v++) buffer_store_f64(f[v], rf, t * 8, (int)(o + v * pl8), 0);
        }
    }
    if (__ballot(fault != 0) && t == 0) atomicOr(faults, 1u);
}

// Solver alone on given face states, as the reference's per-solver calls
// *_Flux(dir, var, rows, cols, GC, L_state, R_state, flux) (include/scheme.h), over the
// face ranges of k_flux.
template <int DIR, int SOLVER>
__global__ __launch_bounds__(256) void k_riemann(const double* __restrict__ Ls,
                                                 const double* __restrict__ Rs,
                                                 double* __restrict__ F, Geom g, Phys ph,
                                                 unsigned* __restrict__ faults) {
    const int j = (DIR == 1 ? GC : GC - 1) + blockIdx.x * 64 + threadIdx.x;
    const int i = (DIR == 1 ? GC - 1 : GC) + blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.R - GC || j >= g.C - GC) return;
    const size_t c0 = (size_t)i * g.ld + j;
    double cl[4], cr[4], f[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cl[v] = Ls[(size_t)v * g.plane + c0];
        cr[v] = Rs[(size_t)v * g.plane + c0];
    }
    if (riemann<DIR, SOLVER>(make_consts(ph), cl, cr, f)) atomicOr(faults, 1u);
#pragma unroll
    for (int v = 0; v < 4; v++) F[(size_t)v * g.plane + c0] = f[v];
}

// FP64 y faces fused with the cell update (fvm_fused.hpp).
template <int ORDER, bool CHAR, int SOLVER>
struct FaceOp64 {
    static constexpr bool kStoreAllG = true;  // the safeguard reads every y face from G
    static constexpr bool kXDeferred = false;  // every x face is in F
    const Phys& ph;
    Consts kc;
    unsigned* faults;
    __device__ __forceinline__ bool compute(const double st[6][4], double f[4]) {
        double UL[4], UR[4];
        reconstruct_face<2, ORDER, CHAR>(kc, ph.dy, st, UL, UR);
        if (ph.positivity && !(admissible(kc, UL) && admissible(kc, UR))) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                UL[v] = st[2][v];
                UR[v] = st[3][v];
            }
        }
        if (riemann<2, SOLVER>(kc, UL, UR, f)) atomicOr(faults, 1u);
        return true;
    }
    __device__ __forceinline__ void defer(size_t) {}
};

template <int ORDER, bool CHAR, int SOLVER>
__global__ FUSED_BOUNDS void k_yflux_update(const double* Un, const double* Y,
                                                      double* out, const double* __restrict__ F,
                                                      double* G, Geom g, Phys ph,
                                                      int kind, Repair rp, DtSlots ds,
                                                      unsigned* __restrict__ faults) {
    if (!step_on(ph)) return;
    FaceOp64<ORDER, CHAR, SOLVER> op{ph, make_consts(ph), faults};
    yflux_update_body<false>(Un, Y, out, F, G, g, ph, ph.step->dt, kind, rp, ds, op);
}

// The marching form (yflux_update_march) of k_yflux_update: identical results, 8% faster
// at 4096^2 (WENO5-char HLLC); used whenever an array's 4 planes fit in 4 GB.  Bounded
// to 3 waves/SIMD like k_flux_ring (WENO5-char: 167 VGPRs, no spill).  A 4-wave bound
// (RPFVM_MARCH_OCC / RPFVM_RING_OCC = 4) spills 120 and 84 bytes a lane and makes the
// WENO5-char HLLC step at 4096^2 17% slower.
template <int ORDER, bool CHAR, int SOLVER>
__global__ __launch_bounds__(64 * kMarchWaves, RPFVM_MARCH_OCC) void k_yflux_march(
    const double* Un, const double* Y, double* out, const double* __restrict__ F, double* G, Geom g,
    Phys ph, int kind, Repair rp, DtSlots ds, unsigned* __restrict__ faults) {
    __shared__ double lds[kMarchWaves * kMarchLds];
    if (!step_on(ph)) return;
    FaceOp64<ORDER, CHAR, SOLVER> op{ph, make_consts(ph), faults};
    yflux_update_march(Un, Y, out, F, G, g, ph, ph.step->dt, kind, rp, ds, op, lds);
}

// ------------------------------------------------------------------ update
// k_update: cell_update (fvm_fused.hpp) over the interior.
__global__ __launch_bounds__(256) void k_update(const double* Un, const double* Y, double* out,
                                                const double* __restrict__ F,
                                                const double* __restrict__ G, Geom g, Phys ph,
                                                int kind, Repair rp) {
    if (!step_on(ph)) return;
    const int j = GC + blockIdx.x * 64 + threadIdx.x;
    const int i = GC + blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.R - GC || j >= g.C - GC) return;
    const bool ok = cell_update(Un, Y, out, F, G, g, ph, ph.step->dt, kind, i, j);
    if (rp.cnt && !ok) {
        const unsigned k = atomicAdd(&rp.cnt[0], 1u);
        rp.list[0][k] = (i - GC) * rp.ny + (j - GC);
    }
}

// Copies the ghost frame of src into dst (all four variables).
__global__ void k_copy_frame(const double* __restrict__ src, double* __restrict__ dst, Geom g,
                             const StepCtl* __restrict__ step) {
    if (!step->active) return;
    const int nframe = 2 * GC * g.C + 2 * GC * (g.R - 2 * GC);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nframe) return;
    int i, j;
    if (t < 2 * GC * g.C) {
        const int r = t / g.C;
        i = r < GC ? r : g.R - 2 * GC + r;
        j = t % g.C;
    } else {
        const int u = t - 2 * GC * g.C;
        i = GC + u / (2 * GC);
        const int k = u % (2 * GC);
        j = k < GC ? k : g.C - 2 * GC + k;
    }
    const size_t c = (size_t)i * g.ld + j;
#pragma unroll
    for (int v = 0; v < 4; v++) dst[(size_t)v * g.plane + c] = src[(size_t)v * g.plane + c];
}

// ------------------------------------------------------------------- CFL dt
// Get_Delta_T_2D (include/function.h:144-175): S_x = max(|u|+a), S_y = max(|v|+a)
// over the interior; "if (s > S) S = s" ignores NaN exactly as the reference.  The
// max is exact, so the block order does not matter.

__global__ __launch_bounds__(256) void k_dt_partial(const double* __restrict__ U, Geom g, Phys ph,
                                                    double* __restrict__ partial) {
    const double gam = ph.gamma;
    double sx = 0.0, sy = 0.0;
    const int nj = g.C - 2 * GC;
    const long total = (long)(g.R - 2 * GC) * nj;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int i = GC + (int)(t / nj), j = GC + (int)(t % nj);
        const size_t c = (size_t)i * g.ld + j;
        double lx, ly;
        cell_speeds(gam, U[c], U[g.plane + c], U[2 * g.plane + c], U[3 * g.plane + c], lx, ly);
        if (lx > sx) sx = lx;
        if (ly > sy) sy = ly;
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    __shared__ double red[2][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[0][w] = sx;
        red[1][w] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0][0], b = red[1][0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
            a = nanmax(a, red[0][k]);
            b = nanmax(b, red[1][k]);
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// The same partial maxima from the slots of the last stage.  Each wavefront takes 64
// consecutive slots; a stale one (kSlotStale) has its cells (up to 63 columns of ds.rows
// rows) recomputed from U by the whole wavefront, one column per lane.
__global__ __launch_bounds__(256) void k_dt_slots(const double* __restrict__ U, Geom g, Phys ph,
                                                  DtSlots ds, double* __restrict__ partial) {
    const double gam = ph.gamma;
    double sx = 0.0, sy = 0.0;
    const int nj = g.C - 2 * GC;
    const int lane = threadIdx.x & 63;
    const long total = (long)dt_slot_count(g, ds.rows);
    const long waves = (long)gridDim.x * (blockDim.x >> 6);
    for (long base = ((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; base < total;
         base += waves * 64) {
        const long w = base + lane;
        bool stale = false;
        if (w < total) {
            const unsigned long long bx = ds.slot[2 * w], by = ds.slot[2 * w + 1];
            stale = bx == kSlotStale;
            if (!stale) {  // slots hold maxima from 0 with NaN skipped: never NaN
                sx = nanmax(sx, __longlong_as_double((long long)bx));
                sy = nanmax(sy, __longlong_as_double((long long)by));
            }
        }
        for (unsigned long long m = __ballot(stale); m; m &= m - 1) {
            const long ws = base + (__ffsll((long long)m) - 1);
            const int ia = GC + (int)(ws / ds.nstrips) * ds.rows;
            const int ib = min(ia + ds.rows, g.R - GC);
            const int j = (int)(ws % ds.nstrips) * kStrip + lane;
            if (lane < kStrip && j < nj) {
                for (int i = ia; i < ib; i++) {
                    const size_t c = (size_t)i * g.ld + GC + j;
                    double lx, ly;
                    cell_speeds(gam, U[c], U[g.plane + c], U[2 * g.plane + c], U[3 * g.plane + c], lx, ly);
                    if (lx > sx) sx = lx;
                    if (ly > sy) sy = ly;
                }
            }
        }
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    __shared__ double red[2][4];
    const int wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = sx;
        red[1][wv] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0][0], b = red[1][0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
            a = nanmax(a, red[0][k]);
            b = nanmax(b, red[1][k]);
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// Reduces the per-block maxima and forms dt = CFL*min(dx,dy)/(S_x+S_y); optionally
// clips it against the device clock as the driver does (src/RP_FVM.c:186-191):
// clock[0] = t, clock[1] = t_stop.
__global__ void k_dt_final(const double* __restrict__ partial, int nblocks, Phys ph,
                           double* __restrict__ dt_out, const double* __restrict__ clock,
                           unsigned* __restrict__ faults) {
    double sx = 0.0, sy = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) {
        sx = nanmax(sx, partial[2 * k]);
        sy = nanmax(sy, partial[2 * k + 1]);
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    if (threadIdx.x == 0) {
        double dt = ph.cfl * min2(ph.dx, ph.dy) / (sx + sy);
        if (clock) {
            const double t = clock[0], ts = clock[1];
            if (t + dt > ts) dt = ts - t;
            if (t + dt >= ts) dt = ts - t;
        }
        if (!(dt > 0.0) || !isfinite(dt)) atomicOr(faults, 2u);
        *dt_out = dt;
    }
}

// ----------------------------------------------------------------- launchers
template <int DIR, int ORDER, bool CHAR, int SOLVER>
static void launch_flux(const double* Y, double* F, const Geom& g, const Phys& ph,
                        unsigned* faults, const FaceFix&, hipStream_t s) {
    const int nj = DIR == 1 ? g.C - 2 * GC : g.C - 2 * GC + 1;
    const int ni = DIR == 1 ? g.R - 2 * GC + 1 : g.R - 2 * GC;
    dim3 block(64, 4);
    dim3 grid((nj + 63) / 64, (ni + 3) / 4);
    if (DIR == 1 && use_march_forms(g, ph)) {
        const int rows = g.R - 2 * GC + 1, per = kRingRows * kRingWaves;
        hipLaunchKernelGGL((k_flux_ring<ORDER, CHAR, SOLVER>), dim3((nj + 63) / 64, (rows + per - 1) / per),
                           dim3(64, kRingWaves), 0, s, Y, F, g, ph, faults);
        return;
    }
    hipLaunchKernelGGL((k_flux<DIR, ORDER, CHAR, SOLVER>), grid, block, 0, s, Y, F, g, ph, faults);
}

template <int ORDER, bool CHAR, int SOLVER>
static void launch_yflux_update(const double* Un, const double* Y, double* out, const double* F,
                                double* G, const Geom& g, const Phys& ph, int kind,
                                const Repair& rp, const FaceFix&, unsigned, const DtSlots& ds,
                                unsigned* faults, hipStream_t s) {
    if (use_march_forms(g, ph)) {
        hipLaunchKernelGGL((k_yflux_march<ORDER, CHAR, SOLVER>), yflux_march_grid(g), yflux_march_block(), 0,
                           s, Un, Y, out, F, G, g, ph, kind, rp, ds, faults);
        return;
    }
    hipLaunchKernelGGL((k_yflux_update<ORDER, CHAR, SOLVER>), yflux_update_grid(g), yflux_update_block(), 0,
                       s, Un, Y, out, F, G, g, ph, kind, rp, ds, faults);
}

template <int ORDER, bool CHAR, int SOLVER>
static FluxPair pick3() {
    return FluxPair{&launch_flux<1, ORDER, CHAR, SOLVER>, &launch_flux<2, ORDER, CHAR, SOLVER>,
                    &launch_flux<1, ORDER, CHAR, SOLVER>, &launch_yflux_update<ORDER, CHAR, SOLVER>,
                    &launch_repair, nullptr, nullptr, nullptr};
}

template <int ORDER, bool CHAR>
static FluxPair pick2(int solver) {
    switch (solver) {
        case SOLVER_HLL: return pick3<ORDER, CHAR, SOLVER_HLL>();
        case SOLVER_HLLC: return pick3<ORDER, CHAR, SOLVER_HLLC>();
        case SOLVER_ROE: return pick3<ORDER, CHAR, SOLVER_ROE>();
        case SOLVER_HLLHC: return pick3<ORDER, CHAR, SOLVER_HLLHC>();
        case SOLVER_HLLCHC: return pick3<ORDER, CHAR, SOLVER_HLLCHC>();
        case SOLVER_ROEHC: return pick3<ORDER, CHAR, SOLVER_ROEHC>();
        default: return pick3<ORDER, CHAR, SOLVER_LLF>();
    }
}

template <int ORDER>
static FluxPair pick1(bool ch, int solver) {
    return ch ? pick2<ORDER, true>(solver) : pick2<ORDER, false>(solver);
}

FluxPair select_flux(int order, bool characteristic, int solver) {
    switch (order) {
        case 1: return pick1<1>(characteristic, solver);
        case 2: return pick1<2>(characteristic, solver);
        case 3: return pick1<3>(characteristic, solver);
        default: return pick1<5>(characteristic, solver);
    }
}

template <int DIR, int SOLVER>
static void launch_riemann_t(const double* Ls, const double* Rs, double* F, const Geom& g,
                             const Phys& ph, unsigned* faults, hipStream_t s) {
    const int nj = DIR == 1 ? g.C - 2 * GC : g.C - 2 * GC + 1;
    const int ni = DIR == 1 ? g.R - 2 * GC + 1 : g.R - 2 * GC;
    hipLaunchKernelGGL((k_riemann<DIR, SOLVER>), dim3((nj + 63) / 64, (ni + 3) / 4), dim3(64, 4),
                       0, s, Ls, Rs, F, g, ph, faults);
}

template <int DIR>
static void launch_riemann_d(int solver, const double* Ls, const double* Rs, double* F,
                             const Geom& g, const Phys& ph, unsigned* faults, hipStream_t s) {
    switch (solver) {
        case SOLVER_HLL: return launch_riemann_t<DIR, SOLVER_HLL>(Ls, Rs, F, g, ph, faults, s);
        case SOLVER_HLLC: return launch_riemann_t<DIR, SOLVER_HLLC>(Ls, Rs, F, g, ph, faults, s);
        case SOLVER_ROE: return launch_riemann_t<DIR, SOLVER_ROE>(Ls, Rs, F, g, ph, faults, s);
        case SOLVER_HLLHC: return launch_riemann_t<DIR, SOLVER_HLLHC>(Ls, Rs, F, g, ph, faults, s);
        case SOLVER_HLLCHC: return launch_riemann_t<DIR, SOLVER_HLLCHC>(Ls, Rs, F, g, ph, faults, s);
        case SOLVER_ROEHC: return launch_riemann_t<DIR, SOLVER_ROEHC>(Ls, Rs, F, g, ph, faults, s);
        default: return launch_riemann_t<DIR, SOLVER_LLF>(Ls, Rs, F, g, ph, faults, s);
    }
}

void launch_riemann(int dir, int solver, const double* Ls, const double* Rs, double* F,
                    const Geom& g, const Phys& ph, unsigned* faults, hipStream_t s) {
    if (dir == 1) launch_riemann_d<1>(solver, Ls, Rs, F, g, ph, faults, s);
    else launch_riemann_d<2>(solver, Ls, Rs, F, g, ph, faults, s);
}

void launch_bc_side(double* U, const Geom& g, const BcSide& s, const StepCtl* step, hipStream_t st) {
    const int n = s.hi - s.lo + 1;
    if (n <= 0) return;
    dim3 block(64);
    dim3 grid((n + 63) / 64, 4);
    hipLaunchKernelGGL(k_bc_side, grid, block, 0, st, U, g, s, step);
}

void launch_bc(double* U, const Geom& g, const BcSide side[4], const StepCtl* step, hipStream_t st) {
    if (g.R - 2 * GC < GC || g.C - 2 * GC < GC) {  // fewer than 4 interior cells: in order
        for (int k = 0; k < 4; k++) launch_bc_side(U, g, side[k], step, st);
        return;
    }
    for (int k = 0; k < 4; k += 2) {
        const int n = std::max(side[k].hi - side[k].lo + 1, side[k + 1].hi - side[k + 1].lo + 1);
        if (n <= 0) continue;
        hipLaunchKernelGGL(k_bc_pair, dim3((n + 63) / 64, 8), dim3(64), 0, st, U, g, side[k],
                           side[k + 1], step);
    }
}

void launch_update(const double* Un, const double* Y, double* out, const double* F,
                   const double* G, const Geom& g, const Phys& ph, int kind,
                   const Repair& rp, hipStream_t s) {
    const int nj = g.C - 2 * GC, ni = g.R - 2 * GC;
    dim3 block(64, 4);
    dim3 grid((nj + 63) / 64, (ni + 3) / 4);
    hipLaunchKernelGGL(k_update, grid, block, 0, s, Un, Y, out, F, G, g, ph, kind, rp);
}

void launch_repair(const double* Un, const double* Y, double* out, double* F, double* G,
                   const Geom& g, const Phys& ph, int kind, const Repair& rp,
                   unsigned stamp, const DtSlots& ds, unsigned* faults, hipStream_t s) {
    launch_repair_t<YFacesStored>(Un, Y, out, F, G, g, ph, kind, rp, stamp, ds, faults, s);
}

void launch_copy_frame(const double* src, double* dst, const Geom& g, const StepCtl* step, hipStream_t s) {
    const int nframe = 2 * GC * g.C + 2 * GC * (g.R - 2 * GC);
    hipLaunchKernelGGL(k_copy_frame, dim3((nframe + 255) / 256), dim3(256), 0, s, src, dst, g, step);
}

// ------------------------------------------------------------ step control
__global__ void k_step_set(StepCtl* step, double dt, double t0) {
    step->dt = dt;
    step->t0 = t0;
    step->active = 1u;
}

__global__ void k_loop_init(StepCtl* step, double t, double t_stop) {
    step->t = t;
    step->t_stop = t_stop;
    step->steps = 0u;
    step->stop = 0u;
    step->active = 0u;
}

// The device loop's step decision, on the per-block maxima: the driver loop of
// src/RP_FVM.c:179-191 (and rpfvm_advance_to's host loop) -- no step once t >= t_stop;
// dt = CFL*min(dx,dy)/(S_x+S_y) (fault bit 2 when not positive and finite, as
// k_dt_final), clipped to t_stop; a dt that does not advance the clock ends the loop
// with fault bit 2; otherwise the step runs from t and the clock becomes t + dt.  Once
// ended, the loop stays ended.
__global__ void k_dt_loop(const double* __restrict__ partial, int nblocks, Phys ph,
                          StepCtl* __restrict__ step, unsigned* __restrict__ faults) {
    double sx = 0.0, sy = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) {
        sx = nanmax(sx, partial[2 * k]);
        sy = nanmax(sy, partial[2 * k + 1]);
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    if (threadIdx.x != 0) return;
    const double t = step->t, ts = step->t_stop;
    if (step->stop || !(t < ts)) {
        step->stop = 1u;
        step->active = 0u;
        return;
    }
    double dt = ph.cfl * min2(ph.dx, ph.dy) / (sx + sy);
    if (!(dt > 0.0) || !isfinite(dt)) atomicOr(faults, 2u);
    if (t + dt > ts) dt = ts - t;
    if (t + dt >= ts) dt = ts - t;
    if (!(t + dt > t)) {
        atomicOr(faults, 2u);
        step->stop = 1u;
        step->active = 0u;
        return;
    }
    step->dt = dt;
    step->t0 = t;
    step->t = t + dt;
    step->steps += 1u;
    step->active = 1u;
}

void launch_step_set(StepCtl* step, double dt, double t0, hipStream_t s) {
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, s, step, dt, t0);
}

void launch_loop_init(StepCtl* step, double t, double t_stop, hipStream_t s) {
    hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(1), 0, s, step, t, t_stop);
}

int dt_blocks(const Geom& g) {
    const long cells = (long)(g.R - 2 * GC) * (g.C - 2 * GC);
    long b = (cells + 255) / 256;
    return (int)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

int march_rows() { return kMarchRows; }

// A pitched 2-D copy on the device (host arrays in the reference layout <-> device rows).
__global__ __launch_bounds__(256) void k_repitch(const double* __restrict__ src, size_t spitch,
                                                 double* __restrict__ dst, size_t dpitch,
                                                 size_t rows, size_t cols) {
    const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= cols) return;
    for (size_t r = blockIdx.y; r < rows; r += gridDim.y) dst[r * dpitch + j] = src[r * spitch + j];
}

void launch_repitch(const double* src, size_t spitch, double* dst, size_t dpitch, size_t rows,
                    size_t cols, hipStream_t s) {
    const unsigned gy = (unsigned)std::min<size_t>(rows, 16384);
    hipLaunchKernelGGL(k_repitch, dim3((unsigned)((cols + 255) / 256), gy), dim3(256), 0, s, src,
                       spitch, dst, dpitch, rows, cols);
}

void launch_dt_slots(const double* U, const Geom& g, const Phys& ph, const DtSlots& ds,
                     double* partial, double* dt_out, const double* clock, unsigned* faults,
                     hipStream_t s) {
    const long slots = dt_slot_count(g, ds.rows);
    const int nb = (int)std::min<long>(dt_blocks(g), std::max<long>(1, (slots + 255) / 256));
    hipLaunchKernelGGL(k_dt_slots, dim3(nb), dim3(256), 0, s, U, g, ph, ds, partial);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(64), 0, s, partial, nb, ph, dt_out, clock, faults);
}

void launch_dt_loop(const double* U, const Geom& g, const Phys& ph, const DtSlots& ds,
                    double* partial, StepCtl* step, unsigned* faults, hipStream_t s) {
    if (ds.slot) {
        const long slots = dt_slot_count(g, ds.rows);
        const int nb = (int)std::min<long>(dt_blocks(g), std::max<long>(1, (slots + 255) / 256));
        hipLaunchKernelGGL(k_dt_slots, dim3(nb), dim3(256), 0, s, U, g, ph, ds, partial);
        hipLaunchKernelGGL(k_dt_loop, dim3(1), dim3(64), 0, s, partial, nb, ph, step, faults);
    } else {
        const int nb = dt_blocks(g);
        hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(256), 0, s, U, g, ph, partial);
        hipLaunchKernelGGL(k_dt_loop, dim3(1), dim3(64), 0, s, partial, nb, ph, step, faults);
    }
}

void launch_dt(const double* U, const Geom& g, const Phys& ph, double* partial, double* dt_out,
               const double* clock, unsigned* faults, hipStream_t s) {
    const int nb = dt_blocks(g);
    hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(256), 0, s, U, g, ph, partial);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(64), 0, s, partial, nb, ph, dt_out, clock, faults);
}

}  // namespace fvm
