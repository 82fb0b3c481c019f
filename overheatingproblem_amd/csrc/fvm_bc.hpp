// fvm_bc.hpp -- the ghost fill of one side at one (tangential index, variable).
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_kernels.hpp"

namespace fvm {

// One side of Boundary_Conditions (include/Boundary_Condition.h:963-980).  A thread
// owns one (variable, tangential index) pair and writes the four ghost layers in the
// reference's order, so a side that reads its own ghosts (periodic or reflecting
// with fewer than four interior cells) sees the same values as the reference loop.
// `time` is the stage clock (read by the double-Mach-reflection shock side only).
__device__ __forceinline__ void bc_side_point(double* __restrict__ U, const Geom& g,
                                              const BcSide& s, int t, int v, double time) {
    double* P = U + (size_t)v * g.plane;
    if (s.type == BC_DMR_WALL || s.type == BC_DMR_SHOCK) {
        // bottom / top of BC_DoubleMach_2D (include/Boundary_Condition.h:558-660), row t
        const double xp = (double)(t + 0.5 - GC) / s.nx * s.lx;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (s.type == BC_DMR_WALL) {
                const double x = P[(size_t)t * g.ld + GC + m];
                P[(size_t)t * g.ld + GC - 1 - m] = xp > s.par[0] ? (v == 2 ? -x : x) : s.value[v];
            } else {
                const double xs = s.par[0] + (s.par[1] / s.par[2]) + (s.par[3] * time / s.par[2]);
                P[(size_t)t * g.ld + g.C - GC + m] = xp < xs ? s.value[v] : s.alt[v];
            }
        }
        return;
    }
    const bool lr = s.side < 2;
    const double sign = (s.type == BC_REFLECT && v == (lr ? 1 : 2)) ? -1.0 : 1.0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        int gi, si = -1;
        switch (s.side) {
            case 0: gi = GC - 1 - m; break;
            case 1: gi = g.R - GC + m; break;
            case 2: gi = GC - 1 - m; break;
            default: gi = g.C - GC + m; break;
        }
        double val = s.value[v];
        switch (s.type) {
            case BC_OUT:
                si = s.side == 0 ? GC : s.side == 1 ? g.R - GC - 1 : s.side == 2 ? GC : g.C - GC - 1;
                break;
            case BC_PERIODIC:
                si = s.side == 0 ? g.R - GC - 1 - m : s.side == 1 ? GC + m
                   : s.side == 2 ? g.C - GC - 1 - m : GC + m;
                break;
            case BC_REFLECT:
                si = s.side == 0 ? GC + m : s.side == 1 ? g.R - GC - 1
                   : s.side == 2 ? GC : g.C - GC - 1 - m;
                break;
            default:
                break;  // inflow / fixed value: s.value
        }
        if (si >= 0) {
            const double x = lr ? P[(size_t)si * g.ld + t] : P[(size_t)t * g.ld + si];
            val = sign < 0.0 ? -x : x;
        }
        if (lr) P[(size_t)gi * g.ld + t] = val;
        else P[(size_t)t * g.ld + gi] = val;
    }
}

}  // namespace fvm
