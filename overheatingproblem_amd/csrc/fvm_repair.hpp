// fvm_repair.hpp -- the positivity safeguard's rounds: device phases, kernels and their
// launcher, templated on where a re-updated cell's face fluxes come from (YF for y faces,
// XF for x faces): YFacesStored / XFacesStored read them from G / F (every face stored, as
// the FP64 form does); the mixed form, whose sweeps store only the faces its own fix-up
// reads, recomputes the others (fvm_mixed_kernels.hip): the y faces after its fused y
// sweep, and the x faces too after its one-pass stage.
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_fused.hpp"
#include "fvm_kernels.hpp"
#include "fvm_math.hpp"

namespace fvm {

// Every y face is in G.
struct YFacesStored {
    static constexpr bool kStored = true;
};
// Every x face is in F.
struct XFacesStored {
    static constexpr bool kStored = true;
};

// ------------------------------------------------------ positivity safeguard
// An extension the reference does not have.  A cell whose forward-Euler state is
// inadmissible gets its four face fluxes replaced by the first-order HLLE flux of the
// adjacent cells; every cell touching a replaced face is updated again, and cells that
// become inadmissible are repaired in the next round.  Each kernel works on a list
// (grid-stride), so a round with an empty list costs one small launch.
// The three phases of a round, grid-stride over their lists (block b of nb).
__device__ __forceinline__ void repair_faces_phase(const double* __restrict__ Y, double* F,
                                                   double* G, const Geom& g, const Phys& ph,
                                                   const Repair& rp, int round, unsigned stamp,
                                                   unsigned* __restrict__ faults, unsigned b,
                                                   unsigned nb) {
    const int cur = round & 1;
    const unsigned n = rp.cnt[cur];
    if (b == 0 && threadIdx.x == 0) {
        rp.cnt[cur ^ 1] = 0;
        rp.cnt[2] = 0;
        if (n) atomicAdd(&rp.cnt[3], n);
    }
    for (unsigned q = b * blockDim.x + threadIdx.x; q < n; q += nb * blockDim.x) {
        const int cell = rp.list[cur][q];
        rp.rep[cell] = stamp;
        const int i = GC + cell / rp.ny, j = GC + cell % rp.ny;
#pragma unroll
        for (int side = 0; side < 4; side++) {
            const int fi = side == 0 ? i - 1 : i, fj = side == 2 ? j - 1 : j;
            const size_t c = (size_t)fi * g.ld + fj;
            const size_t nbr = side < 2 ? c + g.ld : c + 1;
            double a[4], bb[4], f[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                a[k] = Y[(size_t)k * g.plane + c];
                bb[k] = Y[(size_t)k * g.plane + nbr];
            }
            const Consts kc{ph.gamma, ph.gm1, ph.gr, Recip{}};
            const int bad = side < 2 ? hlle<1>(kc, a, bb, f) : hlle<2>(kc, a, bb, f);
            if (bad) atomicOr(faults, 1u);
            double* D = side < 2 ? F : G;
#pragma unroll
            for (int k = 0; k < 4; k++) D[(size_t)k * g.plane + c] = f[k];
        }
    }
}

__device__ __forceinline__ void repair_touch_phase(const Repair& rp, int round, unsigned tstamp,
                                                   unsigned b, unsigned nb) {
    const int cur = round & 1;
    const unsigned n = rp.cnt[cur];
    for (unsigned q = b * blockDim.x + threadIdx.x; q < 5 * n; q += nb * blockDim.x) {
        const int cell = rp.list[cur][q / 5];
        const int m = q % 5;
        int ti = cell / rp.ny, tj = cell % rp.ny;
        ti += m == 1 ? -1 : m == 2 ? 1 : 0;
        tj += m == 3 ? -1 : m == 4 ? 1 : 0;
        if (ti < 0 || ti >= rp.nx || tj < 0 || tj >= rp.ny) continue;
        const int c = ti * rp.ny + tj;
        if (atomicExch(&rp.tch[c], tstamp) != tstamp) {
            const unsigned k = atomicAdd(&rp.cnt[2], 1u);
            rp.touched[k] = c;
        }
    }
}

template <typename YF, typename XF>
__device__ __forceinline__ void repair_cells_phase(const double* Un, const double* Y, double* out,
                                                   const double* __restrict__ F,
                                                   const double* __restrict__ G, const Geom& g,
                                                   const Phys& ph, double dt, int kind,
                                                   const Repair& rp, int round, unsigned stamp,
                                                   const DtSlots& ds, unsigned* __restrict__ faults,
                                                   unsigned b, unsigned nb) {
    const int nxt = (round & 1) ^ 1;
    const unsigned n = rp.cnt[2];
    for (unsigned q = b * blockDim.x + threadIdx.x; q < n; q += nb * blockDim.x) {
        const int c = rp.touched[q];
        const int i = GC + c / rp.ny, j = GC + c % rp.ny;
        bool ok;
        if constexpr (YF::kStored && XF::kStored) {
            ok = cell_update(Un, Y, out, F, G, g, ph, dt, kind, i, j);
        } else if constexpr (!XF::kStored) {  // every face from XF / YF, then cell_update's arithmetic
            const size_t o = (size_t)i * g.ld + j;
            double fl[4], fu[4], gl[4], gu[4], y[4], un[4], W[4], R[4];
            XF::face(Y, F, g, ph, rp, stamp, i - 1, j, fl);
            XF::face(Y, F, g, ph, rp, stamp, i, j, fu);
            YF::face(Y, G, g, ph, rp, stamp, i, j - 1, gl);
            YF::face(Y, G, g, ph, rp, stamp, i, j, gu);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                y[v] = Y[(size_t)v * g.plane + o];
                un[v] = kind != UPD_EULER ? Un[(size_t)v * g.plane + o] : 0.0;
            }
            cell_update_pre(out, g, ph, dt, kind, i, j, y, un, fl, fu, gl, gu, W, R);
            ok = admissible(Consts{ph.gamma, ph.gm1, ph.gr, Recip{}}, W);
        } else {  // the same update with the y faces from YF (cell_update's arithmetic)
            const size_t o = (size_t)i * g.ld + j;
            double gl[4], gu[4], y[4], W[4], R[4];
            YF::face(Y, G, g, ph, rp, stamp, i, j - 1, gl);
            YF::face(Y, G, g, ph, rp, stamp, i, j, gu);
#pragma unroll
            for (int v = 0; v < 4; v++) y[v] = Y[(size_t)v * g.plane + o];
            cell_update_regs(Un, out, F, g, ph, dt, kind, i, j, y, gl, gu, W, R);
            ok = admissible(Consts{ph.gamma, ph.gm1, ph.gr, Recip{}}, W);
        }
        if (ds.slot) {  // the last stage's wave-speed slot of this cell is stale now
            const size_t w = (size_t)((c / rp.ny) / ds.rows) * ds.nstrips + (size_t)((c % rp.ny) / kStrip);
            ds.slot[2 * w] = kSlotStale;
        }
        if (!ok) {
            if (rp.rep[c] != stamp) {
                const unsigned k = atomicAdd(&rp.cnt[nxt], 1u);
                rp.list[nxt][k] = c;
            } else {
                atomicOr(faults, 8u);
            }
        }
    }
}

// All rounds in one workgroup (1024 threads), __syncthreads() between the phases (the
// phases are order-independent, so the result does not depend on how many threads share
// a list).  One ordinary launch per stage: with the usual empty first list it returns at
// once, and a list of a few hundred cells takes a few microseconds.  (A cooperative
// multi-workgroup launch with grid barriers measured 0.1 ms of stream stall per step.)
// The rounds, run by every thread of one workgroup (any size).
template <typename YF, typename XF>
__device__ __forceinline__ void repair_rounds(const double* Un, const double* Y, double* out, double* F,
                                              double* G, const Geom& g, const Phys& ph, double dt,
                                              int kind, const Repair& rp, unsigned stamp,
                                              const DtSlots& ds, unsigned* __restrict__ faults) {
    if (__hip_atomic_load(&rp.cnt[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
    int r = 0;
    for (; r < kRepairRounds; r++) {
        if (__hip_atomic_load(&rp.cnt[r & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) break;
        repair_faces_phase(Y, F, G, g, ph, rp, r, stamp, faults, 0u, 1u);
        __syncthreads();
        repair_touch_phase(rp, r, stamp * 16u + (unsigned)r, 0u, 1u);
        __syncthreads();
        repair_cells_phase<YF, XF>(Un, Y, out, F, G, g, ph, dt, kind, rp, r, stamp, ds, faults, 0u, 1u);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (r == kRepairRounds && rp.cnt[kRepairRounds & 1]) atomicOr(faults, 8u);
        rp.cnt[0] = 0;
        rp.cnt[1] = 0;
        rp.cnt[2] = 0;
    }
}

template <typename YF, typename XF>
__global__ __launch_bounds__(1024) void k_repair_block(const double* Un, const double* Y,
                                                       double* out, double* F, double* G, Geom g,
                                                       Phys ph, int kind, Repair rp,
                                                       unsigned stamp, DtSlots ds,
                                                       unsigned* __restrict__ faults) {
    if (!step_on(ph)) return;
    repair_rounds<YF, XF>(Un, Y, out, F, G, g, ph, ph.step->dt, kind, rp, stamp, ds, faults);
}

template <typename YF, typename XF = XFacesStored>
void launch_repair_t(const double* Un, const double* Y, double* out, double* F, double* G,
                     const Geom& g, const Phys& ph, int kind, const Repair& rp,
                     unsigned stamp, const DtSlots& ds, unsigned* faults, hipStream_t s) {
    hipLaunchKernelGGL((k_repair_block<YF, XF>), dim3(1), dim3(1024), 0, s, Un, Y, out, F, G, g, ph, kind,
                       rp, stamp, ds, faults);
}

}  // namespace fvm
