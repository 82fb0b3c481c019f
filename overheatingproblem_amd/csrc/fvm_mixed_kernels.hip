// fvm_mixed_kernels.hip -- face-flux kernels of the mixed-precision form.
//
// The face arithmetic is fvm::mx (fvm_mixed.hpp).  A face whose FP32-reconstructed states
// are not admissible (rho <= 0 or p <= 0 or not finite) is appended to a list instead of
// computed, and its cells are left to the fix-up after the stage's sweeps: k_fixfaces_mx
// recomputes the listed faces with the FP64 path (with the positivity safeguard on, a face
// still inadmissible there falls back to first order), k_fixcells_mx updates their cells
// again, and the safeguard's rounds follow.  Keeping the rare FP64 path out of the hot
// kernels keeps their registers and occupancy.
// A stage is either two sweeps -- k_flux_mixed (x faces into F; a deferred face gets NaN
// there) and the fused y sweep k_yflux_march_mixed, which leaves the cells next to a
// deferred face of either direction to the fix-up -- or one pass, k_stage_mixed
// (fvm_onepass.hpp).
#include <hip/hip_runtime.h>

#include "fvm_buffer.hpp"
#include "fvm_fused.hpp"
#include "fvm_onepass.hpp"
#include "fvm_repair.hpp"
#include "fvm_kernels.hpp"
#include "fvm_math.hpp"
#include "fvm_mixed.hpp"

namespace fvm {

// Appends face c0 to the FP64 recompute list of direction d (0: x, 1: y).  The lists hold
// every face of the grid plus those two wavefronts both compute (FaceFix::cap), so the
// capacity test never fails; it keeps a stray entry from writing past the list.
__device__ __forceinline__ void defer_face(const FaceFix& fix, int d, size_t c0) {
    const unsigned k = atomicAdd(&fix.cnt[d], 1u);
    if (k < fix.cap) fix.list[d][k] = (int)c0;
}

template <int DIR, int ORDER, bool CHAR, int SOLVER>
__device__ __forceinline__ int face_fp64(const Phys& ph, const double st[6][4], double f[4]) {
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
    return riemann<DIR, SOLVER>(kc, UL, UR, f);
}

// BUF: loads and stores through buffer descriptors (fvm_buffer.hpp), for arrays below 4 GB.
// Minimum waves/SIMD of the x sweep (0: the compiler's choice; WENO5-char: 111 VGPRs,
// 4 waves; a bound of 5 holds 96 VGPRs and spills 17 dwords).
#ifndef RPFVM_FLUX_MX_WAVES
#define RPFVM_FLUX_MX_WAVES 0
#endif
#if RPFVM_FLUX_MX_WAVES > 0
#define FLUX_MX_BOUNDS __launch_bounds__(256, RPFVM_FLUX_MX_WAVES)
#else
#define FLUX_MX_BOUNDS __launch_bounds__(256)
#endif
template <int DIR, int ORDER, bool CHAR, int SOLVER, bool BUF>
__global__ FLUX_MX_BOUNDS void k_flux_mixed(const double* __restrict__ Y,
                                                    double* __restrict__ F, Geom g, Phys ph,
                                                    FaceFix fix) {
    if (!step_on(ph)) return;
    const int j0 = (DIR == 1 ? GC : GC - 1) + blockIdx.x * 64;
    const int j = j0 + threadIdx.x;
    const int i = (DIR == 1 ? GC - 1 : GC) + blockIdx.y * blockDim.y +
                  (BUF ? __builtin_amdgcn_readfirstlane((int)threadIdx.y) : (int)threadIdx.y);
    if (i >= g.R - GC || j >= g.C - GC) return;
    const size_t c0 = (size_t)i * g.ld + j;
    const size_t step = DIR == 1 ? (size_t)g.ld : 1;
    double st[6][4];
    FaceIO<BUF> io(Y, F, g, i, j0, DIR == 1 ? g.ld : 1);
#pragma unroll
    for (int n = 0; n < 6; n++)
#pragma unroll
        for (int v = 0; v < 4; v++) st[n][v] = io.load(Y, v, n - 2, c0, step);
    double UL[4], UR[4], f[4];
    mx::reconstruct_face<DIR, ORDER, CHAR>(ph.gamma, ph.igm1, DIR == 1 ? ph.dx : ph.dy, st, UL, UR);
    if (ph.force_fix || !(mx::admissible64(ph.gamma, UL) && mx::admissible64(ph.gamma, UR))) {
        defer_face(fix, DIR - 1, c0);
        // the fused y sweep leaves the cells of a NaN x face to the fix-up
#pragma unroll
        for (int v = 0; v < 4; v++) io.store(F, v, c0, __builtin_nan(""));
        return;
    }
    mx::riemann<DIR, SOLVER>(ph.gamma, UL, UR, f);
#pragma unroll
    for (int v = 0; v < 4; v++) io.store(F, v, c0, f[v]);
}

// FP64 recomputation of the listed faces of one direction, for the complete sweeps of
// rpfvm_space_operator; also resets the other direction's list, which the next sweep (the
// other direction) refills.
template <int DIR, int ORDER, bool CHAR, int SOLVER>
__global__ __launch_bounds__(256) void k_flux_fix(const double* __restrict__ Y,
                                                  double* __restrict__ F, Geom g, Phys ph,
                                                  FaceFix fix, unsigned* __restrict__ faults) {
    if (!step_on(ph)) return;
    const unsigned n = min(fix.cnt[DIR - 1], fix.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        fix.cnt[2 - DIR] = 0;
        if (n) atomicOr(faults, 4u);
    }
    const size_t step = DIR == 1 ? (size_t)g.ld : 1;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const size_t c0 = (size_t)fix.list[DIR - 1][q];
        double st[6][4], f[4];
#pragma unroll
        for (int m = 0; m < 6; m++)
#pragma unroll
            for (int v = 0; v < 4; v++) st[m][v] = Y[(size_t)v * g.plane + c0 + (m - 2) * step];
        if (face_fp64<DIR, ORDER, CHAR, SOLVER>(ph, st, f)) atomicOr(faults, 1u);
#pragma unroll
        for (int v = 0; v < 4; v++) F[(size_t)v * g.plane + c0] = f[v];
    }
}

// Mixed-form y faces fused with the cell update (fvm_fused.hpp).  A face that needs the
// FP64 path is deferred to fix.list[1]; the cells next to it, and those next to an x face
// the x sweep deferred, are updated again by the fix-up.
template <int ORDER, bool CHAR, int SOLVER>
struct FaceOpMx {
    // The fused sweep stores a face flux in G only where the FP64 fix-up reads it (next to
    // a deferred cell); the safeguard recomputes the others (YFacesMx).
    static constexpr bool kStoreAllG = false;
    static constexpr bool kXDeferred = true;
    const Phys& ph;
    FaceFix fix;
    __device__ __forceinline__ bool compute(const double st[6][4], double f[4]) {
        double UL[4], UR[4];
        mx::reconstruct_face<2, ORDER, CHAR>(ph.gamma, ph.igm1, ph.dy, st, UL, UR);
        if (ph.force_fix || !(mx::admissible64(ph.gamma, UL) && mx::admissible64(ph.gamma, UR)))
            return false;
        mx::riemann<2, SOLVER>(ph.gamma, UL, UR, f);
        return true;
    }
    __device__ __forceinline__ void defer(size_t c0) { defer_face(fix, 1, c0); }
};

// The x faces of the one-pass stage (k_stage_mixed): k_flux_mixed's face arithmetic.
template <int ORDER, bool CHAR, int SOLVER>
struct FaceOpMxX {
    const Phys& ph;
    FaceFix fix;
    __device__ __forceinline__ bool compute(const double st[6][4], double f[4]) {
        double UL[4], UR[4];
        mx::reconstruct_face<1, ORDER, CHAR>(ph.gamma, ph.igm1, ph.dx, st, UL, UR);
        if (ph.force_fix || !(mx::admissible64(ph.gamma, UL) && mx::admissible64(ph.gamma, UR)))
            return false;
        mx::riemann<1, SOLVER>(ph.gamma, UL, UR, f);
        return true;
    }
    __device__ __forceinline__ void defer(size_t c0) { defer_face(fix, 0, c0); }
};

// SLOTS: the last stage, which also leaves the next step's wave-speed maxima in `ds`.
// A kernel of its own, so that the earlier stages keep the register budget of 4
// waves/SIMD (the speeds' divisions and square root add 7 VGPRs).
template <int ORDER, bool CHAR, int SOLVER, bool SLOTS>
__global__ FUSED_MX_BOUNDS void k_yflux_update_mixed(const double* Un, const double* Y,
                                                            double* out,
                                                            const double* __restrict__ F,
                                                            double* G, Geom g, Phys ph,
                                                            int kind, Repair rp, FaceFix fix,
                                                            DtSlots ds) {
    if (!step_on(ph)) return;
    FaceOpMx<ORDER, CHAR, SOLVER> op{ph, fix};
    yflux_update_body<true>(Un, Y, out, F, G, g, ph, ph.step->dt, kind, rp, SLOTS ? ds : DtSlots{}, op);
}

// The marching form of the same kernel (yflux_update_march): identical results.  Used
// whenever an array's 4 planes fit in 4 GB (32-bit buffer offsets); measured 8% faster
// than the row kernel at 4096^2, from staging the stencil rows through LDS with
// scalar-offset buffer loads (the rows per wavefront hardly matter: 1, 2 and 4 within
// 0.5%).
// 4 waves/SIMD: the compiler's own choice is 131-135 VGPRs (3 waves); bounded, it needs
// 121-123 and spills nothing.
template <int ORDER, bool CHAR, int SOLVER, bool SLOTS>
__global__ __launch_bounds__(64 * kMarchWaves, 4) void k_yflux_march_mixed(const double* Un, const double* Y,
                                                           double* out,
                                                           const double* __restrict__ F,
                                                           double* G, Geom g, Phys ph,
                                                           int kind, Repair rp, FaceFix fix,
                                                           DtSlots ds) {
    __shared__ double lds[kMarchWaves * kMarchLds];
    if (!step_on(ph)) return;
    FaceOpMx<ORDER, CHAR, SOLVER> op{ph, fix};
    yflux_update_march(Un, Y, out, F, G, g, ph, ph.step->dt, kind, rp, SLOTS ? ds : DtSlots{}, op, lds);
}

template <int ORDER, bool CHAR, int SOLVER>
static void launch_yflux_update_mixed(const double* Un, const double* Y, double* out,
                                      const double* F, double* G, const Geom& g, const Phys& ph,
                                      int kind, const Repair& rp, const FaceFix& fix,
                                      unsigned stamp, const DtSlots& ds, unsigned* faults,
                                      hipStream_t s) {
    if (use_march_forms(g, ph)) {
        if (ds.slot)
            hipLaunchKernelGGL((k_yflux_march_mixed<ORDER, CHAR, SOLVER, true>), yflux_march_grid(g),
                               yflux_march_block(), 0, s, Un, Y, out, F, G, g, ph, kind, rp, fix, ds);
        else
            hipLaunchKernelGGL((k_yflux_march_mixed<ORDER, CHAR, SOLVER, false>), yflux_march_grid(g),
                               yflux_march_block(), 0, s, Un, Y, out, F, G, g, ph, kind, rp, fix, ds);
    } else if (ds.slot) {
        hipLaunchKernelGGL((k_yflux_update_mixed<ORDER, CHAR, SOLVER, true>), yflux_update_grid(g),
                           yflux_update_block(), 0, s, Un, Y, out, F, G, g, ph, kind, rp, fix, ds);
    } else {
        hipLaunchKernelGGL((k_yflux_update_mixed<ORDER, CHAR, SOLVER, false>), yflux_update_grid(g),
                           yflux_update_block(), 0, s, Un, Y, out, F, G, g, ph, kind, rp, fix, ds);
    }
}

// The x (DIR 1) or y (DIR 2) faces into F; FIX: their FP64 recomputation right after
// (complete fluxes, rpfvm_space_operator), else left to the stage's fix-up.
template <int DIR, int ORDER, bool CHAR, int SOLVER, bool FIX>
static void launch_flux_mixed(const double* Y, double* F, const Geom& g, const Phys& ph,
                              unsigned* faults, const FaceFix& fix, hipStream_t s) {
    const int nj = DIR == 1 ? g.C - 2 * GC : g.C - 2 * GC + 1;
    const int ni = DIR == 1 ? g.R - 2 * GC + 1 : g.R - 2 * GC;
    dim3 block(64, 4);
    dim3 grid((nj + 63) / 64, (ni + 3) / 4);
    if (use_march_forms(g, ph))
        hipLaunchKernelGGL((k_flux_mixed<DIR, ORDER, CHAR, SOLVER, true>), grid, block, 0, s, Y, F, g,
                           ph, fix);
    else
        hipLaunchKernelGGL((k_flux_mixed<DIR, ORDER, CHAR, SOLVER, false>), grid, block, 0, s, Y, F, g,
                           ph, fix);
    if (FIX)
        hipLaunchKernelGGL((k_flux_fix<DIR, ORDER, CHAR, SOLVER>), dim3(64), dim3(256), 0, s, Y, F, g,
                           ph, fix, faults);
}

__global__ void k_fix_reset(FaceFix fix) {
    fix.cnt[0] = 0;
    fix.cnt[1] = 0;
    fix.cnt[2] = 0;
}

void launch_fix_reset(const FaceFix& fix, hipStream_t s) {
    hipLaunchKernelGGL(k_fix_reset, dim3(1), dim3(1), 0, s, fix);
}

// The y-face fluxes of the safeguard's re-updated cells in the mixed form, which does not
// store them all: face (i, f) between cells (i, f) and (i, f + 1) comes from G when the
// safeguard replaced it (a neighbour repaired in this stage) or when the FP32 path
// deferred it (the fix-up wrote its FP64 flux there); otherwise it is recomputed with
// the fused sweep's own arithmetic, which gives the value the sweep had.
template <int ORDER, bool CHAR, int SOLVER>
struct YFacesMx {
    static constexpr bool kStored = false;
    __device__ static void face(const double* Y, const double* G, const Geom& g, const Phys& ph,
                                const Repair& rp, unsigned stamp, int i, int f, double out[4]) {
        const int row = (i - GC) * rp.ny;
        const bool replaced = (f >= GC && rp.rep[row + f - GC] == stamp) ||
                              (f + 1 <= g.C - GC - 1 && rp.rep[row + f + 1 - GC] == stamp);
        const size_t c0 = (size_t)i * g.ld + f;
        if (!replaced) {
            double st[6][4];
#pragma unroll
            for (int n = 0; n < 6; n++)
#pragma unroll
                for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * g.plane + c0 + (n - 2)];
            FaceOpMx<ORDER, CHAR, SOLVER> op{ph, FaceFix{}};
            if (op.compute(st, out)) return;
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = G[(size_t)v * g.plane + c0];
    }
};

// ------------------------------------------------------------- one-pass stage
// x faces, y faces, residual and stage combine in one marching kernel (fvm_onepass.hpp):
// the x-face fluxes never reach HBM.  Faces whose FP32 states are not admissible go to the
// two FP64 lists for the fix-up.  The 13 KB LDS ring of each wavefront allows 12 per CU
// (3 per SIMD); op_occupancy sets the register budget.
template <int ORDER, bool CHAR, int SOLVER, bool SLOTS>
__global__ __launch_bounds__(64 * kOpWaves, (op_occupancy<ORDER, CHAR>())) void k_stage_mixed(const double* Un, const double* Y,
                                                                  double* out, double* F, double* G,
                                                                  Geom g, Phys ph, int kind,
                                                                  Repair rp, FaceFix fix, DtSlots ds) {
    __shared__ double lds[kOpWaves * kOpLds];
    if (!step_on(ph)) return;
    FaceOpMxX<ORDER, CHAR, SOLVER> opx{ph, fix};
    FaceOpMx<ORDER, CHAR, SOLVER> opy{ph, fix};
    stage_march(Un, Y, out, F, G, g, ph, ph.step->dt, kind, rp, SLOTS ? ds : DtSlots{}, opx, opy, lds);
}

template <int ORDER, bool CHAR, int SOLVER>
static void launch_stage_mixed(const double* Un, const double* Y, double* out, double* F,
                               double* G, const Geom& g, const Phys& ph, int kind,
                               const Repair& rp, const FaceFix& fix, unsigned stamp,
                               const DtSlots& ds, unsigned* faults, hipStream_t s) {
    if (ds.slot)
        hipLaunchKernelGGL((k_stage_mixed<ORDER, CHAR, SOLVER, true>), stage_march_grid(g),
                           stage_march_block(), 0, s, Un, Y, out, F, G, g, ph, kind, rp, fix, ds);
    else
        hipLaunchKernelGGL((k_stage_mixed<ORDER, CHAR, SOLVER, false>), stage_march_grid(g),
                           stage_march_block(), 0, s, Un, Y, out, F, G, g, ph, kind, rp, fix, ds);
}

// The x-face fluxes of the safeguard's re-updated cells after a one-pass stage, which stores
// only those its own fix-up reads: face (i, j) between cells (i, j) and (i + 1, j) comes
// from F when the safeguard replaced it or the FP32 path deferred it, and is otherwise
// recomputed with the stage's own arithmetic (the value the stage had).
template <int ORDER, bool CHAR, int SOLVER>
struct XFacesMx {
    static constexpr bool kStored = false;
    __device__ static void face(const double* Y, const double* F, const Geom& g, const Phys& ph,
                                const Repair& rp, unsigned stamp, int i, int j, double out[4]) {
        const int col = j - GC;
        const bool replaced = (i >= GC && rp.rep[(i - GC) * rp.ny + col] == stamp) ||
                              (i + 1 <= g.R - GC - 1 && rp.rep[(i + 1 - GC) * rp.ny + col] == stamp);
        const size_t c0 = (size_t)i * g.ld + j;
        if (!replaced) {
            double st[6][4];
#pragma unroll
            for (int n = 0; n < 6; n++)
#pragma unroll
                for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * g.plane + c0 + (size_t)(n - 2) * g.ld];
            FaceOpMxX<ORDER, CHAR, SOLVER> op{ph, FaceFix{}};
            if (op.compute(st, out)) return;
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = F[(size_t)v * g.plane + c0];
    }
};

// ------------------------------------------------------------------ fix-up
// After a stage's sweeps.  k_fixfaces_mx recomputes the listed faces with the FP64 path
// (grid-stride: x faces into F, y faces into G).  k_fixcells_mx updates the two cells of
// every listed face again from F and G (cell_update; a cell next to two listed faces is
// computed twice with identical results and enters the safeguard's list once) and clears
// the lists.  Then the safeguard's rounds (k_repair_block); YF / XF: where they take the
// faces of a re-updated cell (fvm_repair.hpp).
constexpr int kFixFaceBlocks = 256, kFixCellBlocks = 64;

template <int ORDER, bool CHAR, int SOLVER>
__global__ __launch_bounds__(256) void k_fixfaces_mx(const double* __restrict__ Y, double* F, double* G,
                                                     Geom g, Phys ph, FaceFix fix,
                                                     unsigned* __restrict__ faults) {
    if (!step_on(ph)) return;
    const unsigned n0 = min(fix.cnt[0], fix.cap), n1 = min(fix.cnt[1], fix.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n0 | n1)) atomicOr(faults, 4u);
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n0 + n1; q += gridDim.x * blockDim.x) {
        const bool xf = q < n0;
        const size_t c0 = (size_t)(xf ? fix.list[0][q] : fix.list[1][q - n0]);
        const size_t step = xf ? (size_t)g.ld : 1;
        double st[6][4], f[4];
#pragma unroll
        for (int m = 0; m < 6; m++)
#pragma unroll
            for (int v = 0; v < 4; v++) st[m][v] = Y[(size_t)v * g.plane + c0 + (m - 2) * step];
        const int bad = xf ? face_fp64<1, ORDER, CHAR, SOLVER>(ph, st, f)
                           : face_fp64<2, ORDER, CHAR, SOLVER>(ph, st, f);
        if (bad) atomicOr(faults, 1u);
        double* D = xf ? F : G;
#pragma unroll
        for (int v = 0; v < 4; v++) D[(size_t)v * g.plane + c0] = f[v];
    }
}

__global__ __launch_bounds__(1024) void k_fixcells_mx(const double* Un, const double* Y, double* out,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ G, Geom g, Phys ph,
                                                      int kind, Repair rp, FaceFix fix,
                                                      unsigned stamp, DtSlots ds) {
    if (!step_on(ph)) return;
    const double dt = ph.step->dt;
    const unsigned n0 = min(fix.cnt[0], fix.cap), n1 = min(fix.cnt[1], fix.cap);
    const unsigned tstamp = stamp * 16u + 15u;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * (n0 + n1); q += gridDim.x * blockDim.x) {
        int i, j;
        if (q < 2 * n0) {  // x face (i, j): cells (i, j) and (i + 1, j)
            const size_t c0 = (size_t)fix.list[0][q >> 1];
            i = (int)(c0 / g.ld) + (int)(q & 1);
            j = (int)(c0 % g.ld);
            if (i < GC || i > g.R - GC - 1) continue;
        } else {  // y face (i, j): cells (i, j) and (i, j + 1)
            const unsigned r = q - 2 * n0;
            const size_t c0 = (size_t)fix.list[1][r >> 1];
            i = (int)(c0 / g.ld);
            j = (int)(c0 % g.ld) + (int)(r & 1);
            if (j < GC || j > g.C - GC - 1) continue;
        }
        const bool ok = cell_update(Un, Y, out, F, G, g, ph, dt, kind, i, j);
        if (ds.slot) {  // the last stage's wave-speed slot of this cell is stale now
            const size_t w = (size_t)((i - GC) / ds.rows) * ds.nstrips + (size_t)((j - GC) / kStrip);
            ds.slot[2 * w] = kSlotStale;
        }
        if (rp.cnt && !ok) {
            const int cell = (i - GC) * rp.ny + (j - GC);
            if (atomicExch(&rp.tch[cell], tstamp) != tstamp) {
                const unsigned k = atomicAdd(&rp.cnt[0], 1u);
                rp.list[0][k] = cell;
            }
        }
    }
    // The last workgroup to count clears the lists: every workgroup read both counts before
    // counting itself, and only the counters are shared (no data to publish, so no fence,
    // which on this multi-XCD part would write an L2 back).
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&fix.cnt[2], 1u) == gridDim.x - 1) {
        atomicExch(&fix.cnt[0], 0u);
        atomicExch(&fix.cnt[1], 0u);
        atomicExch(&fix.cnt[2], 0u);
    }
}

template <int ORDER, bool CHAR, int SOLVER, typename YF, typename XF>
static void launch_fixup(const double* Un, const double* Y, double* out, double* F, double* G,
                         const Geom& g, const Phys& ph, int kind, const Repair& rp,
                         const FaceFix& fix, unsigned stamp, const DtSlots& ds, unsigned* faults,
                         hipStream_t s) {
    hipLaunchKernelGGL((k_fixfaces_mx<ORDER, CHAR, SOLVER>), dim3(kFixFaceBlocks), dim3(256), 0, s, Y, F,
                       G, g, ph, fix, faults);
    hipLaunchKernelGGL(k_fixcells_mx, dim3(kFixCellBlocks), dim3(1024), 0, s, Un, Y, out, F, G, g, ph,
                       kind, rp, fix, stamp, ds);
    if (rp.cnt) launch_repair_t<YF, XF>(Un, Y, out, F, G, g, ph, kind, rp, stamp, ds, faults, s);
}

// After the two-sweep stage: every x face is in F (the fix-up writes the deferred ones);
// the marching y sweep stores only the y faces the fix-up reads, the row form (arrays of
// 4 GB and more) every one.
template <int ORDER, bool CHAR, int SOLVER>
static void launch_fixup2(const double* Un, const double* Y, double* out, double* F, double* G,
                          const Geom& g, const Phys& ph, int kind, const Repair& rp,
                          const FaceFix& fix, unsigned stamp, const DtSlots& ds, unsigned* faults,
                          hipStream_t s) {
    if (use_march_forms(g, ph))
        launch_fixup<ORDER, CHAR, SOLVER, YFacesMx<ORDER, CHAR, SOLVER>, XFacesStored>(
            Un, Y, out, F, G, g, ph, kind, rp, fix, stamp, ds, faults, s);
    else
        launch_fixup<ORDER, CHAR, SOLVER, YFacesStored, XFacesStored>(Un, Y, out, F, G, g, ph, kind,
                                                                      rp, fix, stamp, ds, faults, s);
}

// After the one-pass stage, which stores only the faces of either direction the fix-up reads.
template <int ORDER, bool CHAR, int SOLVER>
static void launch_fixup1(const double* Un, const double* Y, double* out, double* F, double* G,
                          const Geom& g, const Phys& ph, int kind, const Repair& rp,
                          const FaceFix& fix, unsigned stamp, const DtSlots& ds, unsigned* faults,
                          hipStream_t s) {
    launch_fixup<ORDER, CHAR, SOLVER, YFacesMx<ORDER, CHAR, SOLVER>, XFacesMx<ORDER, CHAR, SOLVER>>(
        Un, Y, out, F, G, g, ph, kind, rp, fix, stamp, ds, faults, s);
}

template <int ORDER, bool CHAR, int SOLVER>
static FluxPair mpick3() {
    return FluxPair{&launch_flux_mixed<1, ORDER, CHAR, SOLVER, true>,
                    &launch_flux_mixed<2, ORDER, CHAR, SOLVER, true>,
                    &launch_flux_mixed<1, ORDER, CHAR, SOLVER, false>,
                    &launch_yflux_update_mixed<ORDER, CHAR, SOLVER>,
                    nullptr,
                    &launch_stage_mixed<ORDER, CHAR, SOLVER>,
                    &launch_fixup2<ORDER, CHAR, SOLVER>,
                    &launch_fixup1<ORDER, CHAR, SOLVER>};
}

template <int ORDER, bool CHAR>
static FluxPair mpick2(int solver) {
    switch (solver) {
        case SOLVER_HLL: return mpick3<ORDER, CHAR, SOLVER_HLL>();
        case SOLVER_HLLC: return mpick3<ORDER, CHAR, SOLVER_HLLC>();
        case SOLVER_ROE: return mpick3<ORDER, CHAR, SOLVER_ROE>();
        case SOLVER_HLLHC: return mpick3<ORDER, CHAR, SOLVER_HLLHC>();
        case SOLVER_HLLCHC: return mpick3<ORDER, CHAR, SOLVER_HLLCHC>();
        case SOLVER_ROEHC: return mpick3<ORDER, CHAR, SOLVER_ROEHC>();
        default: return mpick3<ORDER, CHAR, SOLVER_LLF>();
    }
}

template <int ORDER>
static FluxPair mpick1(bool ch, int solver) {
    return ch ? mpick2<ORDER, true>(solver) : mpick2<ORDER, false>(solver);
}

FluxPair select_flux_mixed(int order, bool characteristic, int solver) {
    switch (order) {
        case 1: return mpick1<1>(characteristic, solver);
        case 2: return mpick1<2>(characteristic, solver);
        case 3: return mpick1<3>(characteristic, solver);
        default: return mpick1<5>(characteristic, solver);
    }
}

}  // namespace fvm
