// fvm_repair.hpp -- the positivity safeguard's rounds: device phases, kernels and their
// launcher, templated on where a re-updated cell's y-face fluxes come from (YF):
// YFacesStored reads them from G (every face stored, as the FP64 form does); the mixed
// form, whose fused sweep stores only the faces its own fix-up reads, recomputes them
// (fvm_mixed_kernels.hip).  x-face fluxes are always read from F.
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

template <typename YF>
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
        if constexpr (YF::kStored) {
            ok = cell_update(Un, Y, out, F, G, g, ph, dt, kind, i, j);
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
            const size_t w = (size_t)(c / rp.ny) * ds.nstrips + (size_t)((c % rp.ny) / kStrip);
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

template <typename YF>
__global__ __launch_bounds__(256) void k_repair_faces(const double* __restrict__ Y, double* F,
                                                      double* G, Geom g, Phys ph, Repair rp,
                                                      int round, unsigned stamp,
                                                      unsigned* __restrict__ faults) {
    repair_faces_phase(Y, F, G, g, ph, rp, round, stamp, faults, blockIdx.x, gridDim.x);
}

template <typename YF>
__global__ __launch_bounds__(256) void k_repair_touch(Repair rp, int round, unsigned tstamp) {
    repair_touch_phase(rp, round, tstamp, blockIdx.x, gridDim.x);
}

template <typename YF>
__global__ __launch_bounds__(256) void k_repair_cells(const double* Un, const double* Y, double* out,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ G, Geom g,
                                                      Phys ph, double dt, int kind, Repair rp,
                                                      int round, unsigned stamp, DtSlots ds,
                                                      unsigned* __restrict__ faults) {
    repair_cells_phase<YF>(Un, Y, out, F, G, g, ph, dt, kind, rp, round, stamp, ds, faults,
                           blockIdx.x, gridDim.x);
}

// After the last round: cells still listed were never repaired; reset the counters.
template <typename YF>
__global__ void k_repair_finish(Repair rp, int last_list, unsigned* __restrict__ faults) {
    if (threadIdx.x == 0) {
        if (rp.cnt[last_list]) atomicOr(faults, 8u);
        rp.cnt[0] = 0;
        rp.cnt[1] = 0;
        rp.cnt[2] = 0;
    }
}

// Grid barrier of the cooperative repair kernel (every block resident): block leaders
// arrive on bar[0] and poll it, through atomics, until it reaches `target`.  The poll is
// bounded; a barrier that does not complete sets fault bit 16 and lets the block go on
// (never expected with all blocks resident).
__device__ __forceinline__ void repair_grid_sync(unsigned* bar, unsigned target,
                                                 unsigned* __restrict__ faults) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        unsigned polls = 0;
        while (atomicAdd(bar, 0u) < target) {
            if (++polls > (1u << 22)) {
                atomicOr(faults, 16u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
    }
    __syncthreads();
}

// All rounds of the safeguard in one cooperative launch: with an empty first list (the
// usual case) it returns at once instead of running 3 x kRepairRounds + 1 launches.
// rp.cnt[8] is the barrier counter, rp.cnt[9] counts finished blocks; the last block
// out resets both and the lists, as k_repair_finish.
template <typename YF>
__global__ __launch_bounds__(256) void k_repair_all(const double* Un, const double* Y, double* out,
                                                    double* F, double* G, Geom g, Phys ph,
                                                    double dt, int kind, Repair rp, unsigned stamp,
                                                    DtSlots ds, unsigned* __restrict__ faults) {
    if (__hip_atomic_load(&rp.cnt[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
    const unsigned b = blockIdx.x, nb = gridDim.x;
    unsigned* bar = &rp.cnt[8];
    unsigned passed = 0;
    int r = 0;
    for (; r < kRepairRounds; r++) {
        if (__hip_atomic_load(&rp.cnt[r & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) break;
        repair_faces_phase(Y, F, G, g, ph, rp, r, stamp, faults, b, nb);
        repair_grid_sync(bar, ++passed * nb, faults);
        repair_touch_phase(rp, r, stamp * 16u + (unsigned)r, b, nb);
        repair_grid_sync(bar, ++passed * nb, faults);
        repair_cells_phase<YF>(Un, Y, out, F, G, g, ph, dt, kind, rp, r, stamp, ds, faults, b, nb);
        repair_grid_sync(bar, ++passed * nb, faults);
    }
    if (threadIdx.x == 0 && atomicAdd(&rp.cnt[9], 1u) == nb - 1) {
        // every block has left the loop: no list is read any more
        if (r == kRepairRounds && rp.cnt[kRepairRounds & 1]) atomicOr(faults, 8u);
        atomicExch(&rp.cnt[0], 0u);
        atomicExch(&rp.cnt[1], 0u);
        atomicExch(&rp.cnt[2], 0u);
        atomicExch(&rp.cnt[8], 0u);
        atomicExch(&rp.cnt[9], 0u);
    }
}

template <typename YF>
void launch_repair_t(const double* Un, const double* Y, double* out, double* F, double* G,
                     const Geom& g, const Phys& ph, double dt, int kind, const Repair& rp,
                     unsigned stamp, const DtSlots& ds, unsigned* faults, hipStream_t s) {
    const dim3 block(256), grid(64);
    {
        // one cooperative launch (all 64 blocks resident); the separate launches below
        // remain for a device or runtime that refuses it
        Geom gg = g;
        Phys pp = ph;
        Repair rr = rp;
        DtSlots dd = ds;
        void* args[] = {(void*)&Un, (void*)&Y, (void*)&out, (void*)&F, (void*)&G, (void*)&gg,
                        (void*)&pp, (void*)&dt, (void*)&kind, (void*)&rr, (void*)&stamp,
                        (void*)&dd, (void*)&faults};
        if (hipLaunchCooperativeKernel((const void*)k_repair_all<YF>, grid, block, args, 0, s) ==
            hipSuccess)
            return;
        (void)hipGetLastError();
    }
    for (int r = 0; r < kRepairRounds; r++) {
        hipLaunchKernelGGL(k_repair_faces<YF>, grid, block, 0, s, Y, F, G, g, ph, rp, r, stamp, faults);
        hipLaunchKernelGGL(k_repair_touch<YF>, grid, block, 0, s, rp, r, stamp * 16u + (unsigned)r);
        hipLaunchKernelGGL(k_repair_cells<YF>, grid, block, 0, s, Un, Y, out, F, G, g, ph, dt, kind, rp,
                           r, stamp, ds, faults);
    }
    hipLaunchKernelGGL(k_repair_finish<YF>, dim3(1), dim3(64), 0, s, rp, kRepairRounds & 1, faults);
}

}  // namespace fvm
