// fvm_batch_kernels.hip -- kernels of the batching form (layout and schedule in
// fvm_batch.hpp).
//
//   k_batch_control   one workgroup: per-problem dt / clock / activity from the
//                     previous step's wave-speed maxima (Get_Delta_T_2D,
//                     include/function.h:144-175, clipped as src/RP_FVM.c:186-191), the
//                     prefix table of active strips, and the reset of the work counters
//   k_batch_stage     persistent: one RK stage of every active problem, strip by strip
//   k_batch_fix       mixed form: the cells next to faces whose FP32 states were not
//                     admissible, with those faces recomputed in FP64
//   k_batch_copy_back RK1 only: U1 -> U over the active interiors
//   two-sweep schedule (per stage):
//   k_batch_ghosts    ghost frame of the stage input, one side of every active problem
//   k_batch_xsweep    persistent: x-face fluxes of every active strip into a.flux
//   k_batch_xfix      mixed form: x faces whose FP32 states were not admissible, in FP64
//   k_batch_ysweep    persistent: y faces fused with the residual and stage combine
#include <hip/hip_runtime.h>

#include "fvm_batch.hpp"
#include "fvm_bc.hpp"
#include "fvm_buffer.hpp"
#include "fvm_fused.hpp"
#include "fvm_kernels.hpp"
#include "fvm_math.hpp"
#include "fvm_mixed.hpp"

namespace fvm {
namespace {

// Stage input value at (i, j) of one problem with the ghost layers replaced by the
// boundary maps of Boundary_Conditions (include/Boundary_Condition.h:63-465, as
// k_bc_side applies them): outflow copies the first interior cell, reflection mirrors
// on the left and top and copies the first interior cell on the right and bottom, with
// the normal momentum negated, periodic wraps by n, inflow / fixed return the side's
// state.  At most one of i, j lies outside the interior.
__device__ __forceinline__ double bload(const double* P, const BProb& d, int i, int j, int v) {
    const Geom& g = d.g;
    bool neg = false;
    if (i < GC || i >= g.R - GC) {
        const int s = i < GC ? 0 : 1;
        switch (d.bc[s]) {
            case BC_OUT: i = s == 0 ? GC : g.R - GC - 1; break;
            case BC_REFLECT:
                i = s == 0 ? 2 * GC - 1 - i : g.R - GC - 1;
                neg = v == 1;
                break;
            case BC_PERIODIC: i = s == 0 ? i + d.nx : i - d.nx; break;
            default: return d.bcv[s][v];
        }
    } else if (j < GC || j >= g.C - GC) {
        const int s = j < GC ? 2 : 3;
        switch (d.bc[s]) {
            case BC_OUT: j = s == 2 ? GC : g.C - GC - 1; break;
            case BC_REFLECT:
                j = s == 2 ? GC : 2 * (g.C - GC) - 1 - j;
                neg = v == 2;
                break;
            case BC_PERIODIC: j = s == 2 ? j + d.ny : j - d.ny; break;
            default: return d.bcv[s][v];
        }
    }
    const double x = P[(size_t)v * g.plane + (size_t)i * g.ld + j];
    return neg ? -x : x;
}

// FP64 face (the single-problem k_flux arithmetic) with the solver chosen at run time.
// HC adds the heat-capacity solvers (every FP64 face of a batch but the mixed form's own
// sweep kernels, which compute FP64 faces through their fix-ups only).
template <int DIR, int ORDER, bool CHAR, bool HC = false>
__device__ __forceinline__ int face64(const Consts& kc, const Phys& ph, int solver,
                                      const double st[6][4], double f[4]) {
    double UL[4], UR[4];
    reconstruct_face<DIR, ORDER, CHAR>(kc, DIR == 1 ? ph.dx : ph.dy, st, UL, UR);
    if (ph.positivity && !(admissible(kc, UL) && admissible(kc, UR))) {
        // the safeguard's face level: first-order states (k_flux)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            UL[v] = st[2][v];
            UR[v] = st[3][v];
        }
    }
    if constexpr (HC) {
        switch (solver) {
            case SOLVER_HLLHC: return riemann<DIR, SOLVER_HLLHC>(kc, UL, UR, f);
            case SOLVER_HLLCHC: return riemann<DIR, SOLVER_HLLCHC>(kc, UL, UR, f);
            case SOLVER_ROEHC: return riemann<DIR, SOLVER_ROEHC>(kc, UL, UR, f);
            default: break;
        }
    }
    switch (solver) {
        case SOLVER_HLL: return riemann<DIR, SOLVER_HLL>(kc, UL, UR, f);
        case SOLVER_HLLC: return riemann<DIR, SOLVER_HLLC>(kc, UL, UR, f);
        case SOLVER_ROE: return riemann<DIR, SOLVER_ROE>(kc, UL, UR, f);
        default: return riemann<DIR, SOLVER_LLF>(kc, UL, UR, f);
    }
}

// Mixed-form face (k_flux_mixed arithmetic, every solver incl. the heat-capacity ones);
// false when the FP32-reconstructed states are not admissible and the face must be
// recomputed in FP64.
template <int DIR, int ORDER, bool CHAR>
__device__ __forceinline__ bool face_mx(const Phys& ph, int solver, const double st[6][4],
                                        double f[4]) {
    double UL[4], UR[4];
    mx::reconstruct_face<DIR, ORDER, CHAR>(ph.gamma, ph.igm1, DIR == 1 ? ph.dx : ph.dy, st, UL, UR);
    if (ph.force_fix || !(mx::admissible64(ph.gamma, UL) && mx::admissible64(ph.gamma, UR)))
        return false;
    switch (solver) {
        case SOLVER_HLL: mx::riemann<DIR, SOLVER_HLL>(ph.gamma, UL, UR, f); break;
        case SOLVER_HLLC: mx::riemann<DIR, SOLVER_HLLC>(ph.gamma, UL, UR, f); break;
        case SOLVER_ROE: mx::riemann<DIR, SOLVER_ROE>(ph.gamma, UL, UR, f); break;
        case SOLVER_HLLHC: mx::riemann<DIR, SOLVER_HLLHC>(ph.gamma, UL, UR, f); break;
        case SOLVER_HLLCHC: mx::riemann<DIR, SOLVER_HLLCHC>(ph.gamma, UL, UR, f); break;
        case SOLVER_ROEHC: mx::riemann<DIR, SOLVER_ROEHC>(ph.gamma, UL, UR, f); break;
        default: mx::riemann<DIR, SOLVER_LLF>(ph.gamma, UL, UR, f); break;
    }
    return true;
}

// Space_Discrete_Item (include/scheme.h:3758-3777) with Source_Gravity (:3723-3751) and
// the RK stage combine (include/time_advance.h:697, 757-777, 832-872) of one cell, with
// the four face fluxes in registers: fl/fu the lower/upper x faces, gl/gu the y faces.
// The expressions are cell_update's, term for term; W receives the forward-Euler state
// Y + dt L that the positivity safeguard tests.
__device__ __forceinline__ void combine(const Phys& ph, double dt, int kind, const double y[4],
                                        const double un[4], const double fl[4],
                                        const double fu[4], const double gl[4],
                                        const double gu[4], double r[4], double W[4]) {
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
        switch (kind) {
            case UPD_EULER: r[k] = y[k] + dt * L; break;
            case UPD_RK2_2: r[k] = 0.5 * un[k] + 0.5 * (y[k] + dt * L); break;
            case UPD_RK3_2: r[k] = (3.0 / 4.0) * un[k] + (1.0 / 4.0) * (y[k] + dt * L); break;
            default: r[k] = (1.0 / 3.0) * un[k] + (2.0 / 3.0) * (y[k] + dt * L); break;
        }
    }
}

// |u| + a and |v| + a of a cell, as k_dt_partial.
__device__ __forceinline__ void speeds(double gam, const double c[4], double& lx, double& ly) {
    const double rho = c[0];
    const double u = c[1] / rho;
    const double v = c[2] / rho;
    const double p = (c[3] - 0.5 * rho * (u * u + v * v)) * (gam - 1.0);
    const double a = sqrt(gam * p / rho);
    lx = fabs(u) + a;
    ly = fabs(v) + a;
}

__device__ __forceinline__ double nanmax_b(double a, double b) { return b > a ? b : a; }

__device__ __forceinline__ double wave_max_b(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = nanmax_b(x, __shfl_xor(x, o, 64));
    return x;
}

// Non-negative doubles order as their bit patterns.
__device__ __forceinline__ void speed_max(BStat& s, double sx, double sy) {
    atomicMax(&s.smax[0], (unsigned long long)__double_as_longlong(sx));
    atomicMax(&s.smax[1], (unsigned long long)__double_as_longlong(sy));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// A wave-uniform row index the optimiser cannot derive from the loop counter.
__device__ __forceinline__ int opaque_row(int r) {
    asm volatile("" : "+s"(r));
    return r;
}

// Which active problem owns strip `item` (binary search of the prefix table).
__device__ __forceinline__ int owner(const BatchArgs& a, int nact, int item) {
    int lo = 0, hi = nact - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.act_start[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr int kRingCols = 72;  // 69 columns of a strip row (+3 padding)

// The positivity safeguard's first list: cell (i, c) of problem k, whose forward-Euler
// state W is not admissible.
__device__ __forceinline__ void repair_candidate(const BatchArgs& a, const BProb& d, const Phys& ph,
                                                 int k, int i, int c, const double W[4]) {
    const Consts kc{ph.gamma, ph.gm1, ph.gr, Recip{}};
    if (admissible(kc, W)) return;
    const unsigned q = atomicAdd(&a.rep.cnt[0], 1u);
    a.rep.list[0][q] = ((unsigned long long)k << 40) |
                       (unsigned long long)((long)(i - GC) * d.ny + (c - GC));
}

}  // namespace

// ----------------------------------------------------------------- control
// mode 0: advance the clocks (one step); mode 1: schedule only (autotuner trials).
// A problem stops at t_end or on a dt fault; an HLL bad-state fault (bit 1) does not
// stop it, as the reference's HLL_Flux prints and carries on (include/scheme.h:427-436).
__global__ __launch_bounds__(256) void k_batch_control(BatchArgs a, int mode) {
    BSched& sc = *a.sched;
    const double t_end = sc.t_end;
    for (int k = threadIdx.x; k < a.nprob; k += blockDim.x) {
        BStat& s = a.stat[k];
        const BProb& d = a.probs[k];
        if (mode == 0) {
            bool act = s.t < t_end && !(s.faults & 2u);
            if (act) {
                const double sx = __longlong_as_double((long long)s.smax[0]);
                const double sy = __longlong_as_double((long long)s.smax[1]);
                double dt = d.ph.cfl * min2(d.ph.dx, d.ph.dy) / (sx + sy);
                if (s.t + dt > t_end) dt = t_end - s.t;
                if (s.t + dt >= t_end) dt = t_end - s.t;
                // a dt that no longer advances the clock is a fault too (the loop
                // would never reach t_end)
                if (!(dt > 0.0) || !isfinite(dt) || !(s.t + dt > s.t)) {
                    s.faults |= 2u;
                    atomicOr(a.faults, 2u);
                    act = false;
                } else {
                    s.dt = dt;
                    s.t = s.t + dt;
                    s.steps += 1;
                    s.smax[0] = 0ull;
                    s.smax[1] = 0ull;
                }
            }
            s.active = act ? 1 : 0;
            if (a.swap_pools) s.mirror = act ? 0 : ((s.mirror & 1) ? 1 : 3);
        } else {
            s.active = (s.t < t_end && !(s.faults & 2u)) ? 1 : 0;
        }
    }
    // The table of active problems (ascending k) and their first work items: an exclusive
    // scan over the problems in chunks of the workgroup (a serial loop in one thread
    // waited on each problem's loads in turn, 42 us for 64 problems).
    __shared__ int any_sh, left_sh, take_sh[256], items_sh[256];
    if (threadIdx.x == 0) {
        any_sh = 0;
        left_sh = 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < a.nprob; k += blockDim.x) {
        const BStat& s = a.stat[k];
        if (s.active) atomicOr(&any_sh, 1);
        if (s.t < t_end && !(s.faults & 2u)) atomicAdd(&left_sh, 1);
    }
    __syncthreads();
    const bool any = any_sh != 0;
    int nact = 0, total = 0;  // the same in every thread
    for (int k0 = 0; k0 < a.nprob; k0 += blockDim.x) {
        const int k = k0 + (int)threadIdx.x;
        bool take = false;
        int items = 0;
        if (k < a.nprob) {
            const BStat& s = a.stat[k];
            take = mode == 0 ? s.active != 0 : (s.active != 0 || !any);
            if (take) {
                const BProb& d = a.probs[k];
                items = ((d.nx + a.rows - 1) / a.rows) * d.strips_y;
            }
        }
        // inclusive scans of the counts (Hillis-Steele over the workgroup)
        take_sh[threadIdx.x] = take ? 1 : 0;
        items_sh[threadIdx.x] = items;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            const int tv = threadIdx.x >= (unsigned)o ? take_sh[threadIdx.x - o] : 0;
            const int iv = threadIdx.x >= (unsigned)o ? items_sh[threadIdx.x - o] : 0;
            __syncthreads();
            take_sh[threadIdx.x] += tv;
            items_sh[threadIdx.x] += iv;
            __syncthreads();
        }
        if (take) {
            const int q = nact + take_sh[threadIdx.x] - 1;
            a.act_prob[q] = k;
            a.act_start[q] = total + items_sh[threadIdx.x] - items;
        }
        nact += take_sh[blockDim.x - 1];
        total += items_sh[blockDim.x - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.act_start[nact] = total;
        sc.nact = nact;
        sc.total = total;
        sc.active_left = left_sh;
        sc.stamp += 4u;  // stage stamps of this step: stamp + 0, 1, 2 (the first is 4)
        for (int q = 0; q < 4; q++) {
            sc.queue[q] = 0u;
            sc.defer[q] = 0u;
            sc.xqueue[q] = 0u;
            sc.xdefer[q] = 0u;
        }
    }
}

// ------------------------------------------------------------------- stage
// One RK stage: Y = pool[yi] -> out = pool[oi] (U^n = pool[0]); `last` accumulates the
// wave-speed maxima of the new state for the next step's dt.
template <int ORDER, bool CHAR, bool MIXED, int STAGING>
__global__ __launch_bounds__(64) void k_batch_stage(BatchArgs a, int stage, int kind, int yi,
                                                    int oi, int last) {
    const int lane = threadIdx.x;
    const int nact = a.sched->nact, total = a.sched->total;
    unsigned* queue = &a.sched->queue[stage];
    __shared__ double ring[STAGING == STAGING_LDS ? 6 : 1][4][kRingCols];
    // Work fetch without a lane-dependent branch: every lane takes part in the atomic,
    // lane 0 adding 1 and the others 0, and the wavefront takes lane 0's ticket.  (An
    // `if (lane == 0) atomicAdd` + shuffle let the compiler lower this loop as a
    // divergent one whose latch re-entered the strip body without a new fetch.)  The
    // trip count is bounded by the number of work items as well.
    for (int trip = 0; trip <= total; trip++) {
        const unsigned ticket = atomicAdd(queue, lane == 0 ? 1u : 0u);
        const int item = uniform((int)ticket);
        if (item >= total) break;
        const int ai = uniform(owner(a, nact, item));
        const int k = uniform(a.act_prob[ai]);
        const int local = item - a.act_start[ai];
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const Phys ph = d.ph;
        const int solver = d.solver;
        const double dt = a.stat[k].dt;
        const double* Y = a.pool[yi] + d.off;
        const double* Un = a.pool[0] + d.off;
        double* out = a.pool[oi] + d.off;
        const int sx = local / d.strips_y, sy = local - sx * d.strips_y;
        const int i0 = GC + sx * a.rows;
        const int i1 = min(i0 + a.rows, g.R - GC);
        const int j0 = GC + sy * kStripCols;
        const int jf = j0 - 1 + lane;                 // this lane's y face G[i][jf]
        const int c = jf + 1;                         // this lane's cell column
        const bool face_ok = jf <= g.C - GC - 1;
        const bool cell_ok = lane < kStripCols && c <= g.C - GC - 1;
        int fault = 0;
        double smx = 0.0, smy = 0.0;

        // x-face stencil rows r-2..r+3 at column c, y-face stencil columns jf-2..jf+3
        // at row r
        // Row indices pass through opaque_row before addressing: left visible, loop
        // strength reduction turns the 24 stencil addresses into 24 live 64-bit
        // induction pointers (about 50 VGPRs).
        auto load_x = [&](int r, double st[6][4]) {
            r = opaque_row(r);
            if constexpr (STAGING == STAGING_LDS) {
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = ring[(r - 2 + n + 6) % 6][v][lane + 3];
            } else {
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = bload(Y, d, r - 2 + n, c, v);
            }
        };
        auto load_y = [&](int r, double st[6][4]) {
            r = opaque_row(r);
            if constexpr (STAGING == STAGING_LDS) {
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = ring[r % 6][v][lane + n];
            } else {
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = bload(Y, d, r, jf - 2 + n, v);
            }
        };
        // LDS ring: row r, columns j0-3 .. j0+65 (clamped into the array) at slot r % 6
        auto stage_row = [&](int r) {
            if constexpr (STAGING == STAGING_LDS) {
                r = opaque_row(r);
                const int ja = min(j0 - 3 + lane, g.C - 1);
                const int jb = min(j0 + 61 + lane, g.C - 1);
                const bool interior_row = r >= GC && r < g.R - GC;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    // a ghost row is read only at interior columns; a ghost column only
                    // in interior rows
                    const double xa = (interior_row || (ja >= GC && ja < g.C - GC))
                                          ? bload(Y, d, r, ja, v) : 0.0;
                    ring[r % 6][v][lane] = xa;
                    if (lane < 5) {
                        const double xb = (interior_row || (jb >= GC && jb < g.C - GC))
                                              ? bload(Y, d, r, jb, v) : 0.0;
                        ring[r % 6][v][64 + lane] = xb;
                    }
                }
            }
        };
        // The face arithmetic reads the problem's physics through a pointer the
        // compiler cannot see through, once per row: otherwise it hoists every
        // gamma-derived constant out of the row loop and keeps them all live, which
        // costs 40-60 VGPRs (and occupancy).
        auto row_phys = [&]() {
            const BProb* dq = a.probs + k;
            asm volatile("" : "+s"(dq));
            return dq->ph;
        };
        auto x_face = [&](int r, double f[4]) -> bool {
            double st[6][4];
            load_x(r, st);
            const Phys pr = row_phys();
            if constexpr (MIXED) {
                return face_mx<1, ORDER, CHAR>(pr, solver, st, f);
            } else {
                fault |= face64<1, ORDER, CHAR, !MIXED>(make_consts(pr), pr, solver, st, f);
                return true;
            }
        };

        if constexpr (STAGING == STAGING_LDS) {
            __syncthreads();  // the previous strip's reads are done
            for (int r = i0 - 3; r < i0 + 2; r++) stage_row(r);
        }
        // Row r computes the x face F[r] (between rows r and r + 1); from r = i0 on it
        // also computes the y faces of row r and updates the row's cells with F[r - 1]
        // carried from the previous iteration.
        double fp[4] = {0.0, 0.0, 0.0, 0.0};
        bool bad_p = false;
        for (int i = i0 - 1; i < i1; i++) {
            if constexpr (STAGING == STAGING_LDS) {
                __syncthreads();     // slot (i + 3) % 6 held row i - 3, last read for face i - 1
                stage_row(i + 3);
                __syncthreads();
            }
            double fc[4] = {0.0, 0.0, 0.0, 0.0};
            bool bad_c = false;
            if (cell_ok) bad_c = !x_face(i, fc);
            if (i >= i0) {
                double gf[4] = {0.0, 0.0, 0.0, 0.0}, y[4] = {0.0, 0.0, 0.0, 0.0};
                int bad_g = 0;
                if (face_ok) {
                    double st[6][4];
                    load_y(i, st);
#pragma unroll
                    for (int v = 0; v < 4; v++) y[v] = st[3][v];
                    const Phys pr = row_phys();
                    if constexpr (MIXED) {
                        bad_g = face_mx<2, ORDER, CHAR>(pr, solver, st, gf) ? 0 : 1;
                    } else {
                        fault |= face64<2, ORDER, CHAR, !MIXED>(make_consts(pr), pr, solver, st, gf);
                    }
                }
                double gu[4];
#pragma unroll
                for (int v = 0; v < 4; v++) gu[v] = __shfl_down(gf[v], 1, 64);
                const int bad_gu = __shfl_down(bad_g, 1, 64);
                if (cell_ok) {
                    const size_t o0 = (size_t)opaque_row(i) * g.ld + c;
                    if (MIXED && (bad_p || bad_c || bad_g || bad_gu)) {
                        const unsigned q = atomicAdd(&a.sched->defer[stage], 1u);
                        a.defer_list[q] = ((unsigned long long)k << 40) |
                                          ((unsigned long long)i << 20) | (unsigned long long)c;
                    } else {
                        double un[4] = {0.0, 0.0, 0.0, 0.0}, r[4];
                        if (kind != UPD_EULER) {
#pragma unroll
                            for (int v = 0; v < 4; v++) un[v] = Un[(size_t)v * g.plane + o0];
                        }
                        double W[4];
                        combine(ph, dt, kind, y, un, fp, fc, gf, gu, r, W);
#pragma unroll
                        for (int v = 0; v < 4; v++) out[(size_t)v * g.plane + o0] = r[v];
                        if (last) {
                            double lx, ly;
                            speeds(ph.gamma, r, lx, ly);
                            if (lx > smx) smx = lx;
                            if (ly > smy) smy = ly;
                        }
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) fp[v] = fc[v];
            bad_p = bad_c;
        }
        if (last) {
            smx = wave_max_b(smx);
            smy = wave_max_b(smy);
            if (lane == 0) speed_max(a.stat[k], smx, smy);
        }
        if (!MIXED) {
            const unsigned long long any = __ballot(fault != 0);
            if (any && lane == 0) {
                atomicOr(a.faults, 1u);
                atomicOr(&a.stat[k].faults, 1u);
            }
        }
    }
}

// -------------------------------------------------------- two-sweep schedule
// Ghost frame of side `side` of every active problem's stage input (pool[yi]), as the
// single path's k_bc_side: grid (tangential blocks, 4 variables, problems).
__global__ __launch_bounds__(64) void k_batch_ghosts(BatchArgs a, int side, int yi) {
    for (int k = blockIdx.z; k < a.nprob; k += gridDim.z) {
        if (!a.stat[k].active) continue;
        const BProb& d = a.probs[k];
        const BcSide s = d.side[side];
        const int t = s.lo + blockIdx.x * blockDim.x + threadIdx.x;
        if (t > s.hi) continue;
        bc_side_point(a.pool[yi] + d.off, d.g, s, t, blockIdx.y, s.t);  // no time-dependent sides in batches
    }
}

// The x faces F[r][c] (between rows r and r + 1) of every active strip: lane t owns
// column c = j0 + t (t < 63) and marches the strip's rows, as one row of k_flux<1> per
// step (the first strip row of a problem also computes its lower boundary face).
// Mixed form: a face whose FP32 states are not admissible goes to xdefer_list.  BUF: the
// stencil rows staged through an LDS ring by direct buffer loads, one row ahead of the face
// arithmetic, and the fluxes stored through a buffer descriptor (row offsets in SGPRs, the
// lane's column in one VGPR; every problem's arrays below 4 GB); otherwise global loads of
// the 24 stencil values per face.  A row walk that loads each face's stencil just before
// its arithmetic left 30% of the VALU issue idle (0.68 busy at 3.9 waves/SIMD).
template <int ORDER, bool CHAR, bool MIXED, bool BUF>
__global__ __launch_bounds__(64, (MIXED || !BUF ? 1 : 3)) void k_batch_xsweep(BatchArgs a, int stage, int yi) {
    __shared__ double xring[BUF ? 6 * 4 * 64 : 1];  // 12 KB: six stencil rows of the strip
    const int lane = threadIdx.x;
    const int nact = a.sched->nact, total = a.sched->total;
    unsigned* queue = &a.sched->xqueue[stage];
    for (int trip = 0; trip <= total; trip++) {
        const unsigned ticket = atomicAdd(queue, lane == 0 ? 1u : 0u);
        const int item = uniform((int)ticket);
        if (item >= total) break;
        const int ai = uniform(owner(a, nact, item));
        const int k = uniform(a.act_prob[ai]);
        const int local = item - a.act_start[ai];
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const int solver = d.solver;
        const double* Y = a.pool[yi] + d.off;
        double* F = a.flux + d.off;
        const int sx = local / d.strips_y, sy = local - sx * d.strips_y;
        const int i0 = GC + sx * a.rows;
        const int i1 = min(i0 + a.rows, g.R - GC);
        const int c = GC + sy * kStripCols + lane;
        const bool col_ok = lane < kStripCols && c <= g.C - GC - 1;
        int fault = 0;
        auto row_phys = [&]() {
            const BProb* dq = a.probs + k;
            asm volatile("" : "+s"(dq));
            return dq->ph;
        };
        i32x4 ry, rf;
        unsigned pl8 = 0, ld8 = 0, cb0 = 0;
        if constexpr (BUF) {
            const unsigned bytes = (unsigned)array_bytes(g);
            ry = buffer_rsrc(Y, bytes);
            rf = buffer_rsrc(F, bytes);
            pl8 = (unsigned)g.plane * 8u;
            ld8 = (unsigned)g.ld * 8u;
            cb0 = (unsigned)(GC + sy * kStripCols) * 8u;  // the strip's first column
        }
        const int lb = lane * 8;
        const int r_first = i0 == GC ? GC - 1 : i0;
        // BUF: the stencil rows r_first-2 .. i1+2 of the strip (64 columns, 4 variables each)
        // pass through a ring of 6 LDS slots, row q in slot q mod 6, loaded straight into
        // LDS (no VGPR holds them); row r + 4 is requested as soon as face row r has read
        // the slot of row r - 2, so it arrives during face row r's arithmetic.  Lanes past
        // the row's last column load its last dword again (never read).  Rows stay within
        // 1 .. R-2 (r_first >= GC - 1, i1 <= R - GC).
        int off0 = 0, off1 = 0;
        if constexpr (BUF) {
            const int cw = 2 * min(64, g.C - (GC + sy * kStripCols));  // valid dwords of a row
            off0 = 4 * min(lane, cw - 1);
            off1 = 4 * min(64 + lane, cw - 1);
        }
        auto fetch_row = [&](int q) {
            const unsigned o = (unsigned)q * ld8 + cb0;
            unsigned* slot = reinterpret_cast<unsigned*>(xring + (q % 6) * 4 * 64);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                buffer_load_lds_dword(ry, (lds_void*)(slot + v * 128), 4, off0, (int)(o + v * pl8), 0, 0);
                buffer_load_lds_dword(ry, (lds_void*)(slot + v * 128 + 64), 4, off1, (int)(o + v * pl8), 0, 0);
            }
        };
        if constexpr (BUF) {
            for (int q = r_first - 2; q <= r_first + 3; q++) fetch_row(q);
        }
        for (int r = r_first; r < i1; r++) {
            double st[6][4], f[4];
            size_t c0 = 0, pl = 0;
            unsigned rb = 0;
            if constexpr (BUF) {
                wait_vmem();  // rows up to r + 3 (and the stores of face row r - 1)
#pragma unroll
                for (int n = 0; n < 6; n++) {
                    const double* slot = xring + ((r - 2 + n) % 6) * 4 * 64;
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = slot[v * 64 + lane];
                }
                wait_lds();  // the slot of row r - 2 is read for the last time
                if (r + 1 < i1) fetch_row(r + 4);
                if (!col_ok) continue;
                rb = (unsigned)opaque_row(r) * ld8 + cb0;
            } else {
                if (!col_ok) continue;
                // strides through an opaque copy, so that the 24 stencil offsets are formed
                // at their loads instead of hoisted out of the row loop into live registers
                size_t ld = g.ld;
                pl = g.plane;
                asm volatile("" : "+v"(ld), "+v"(pl));
                c0 = (size_t)opaque_row(r) * ld + c;
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * pl + c0 + (size_t)(n - 2) * ld];
            }
            const Phys pr = row_phys();
            if constexpr (MIXED) {
                if (!face_mx<1, ORDER, CHAR>(pr, solver, st, f)) {
                    const unsigned q = atomicAdd(&a.sched->xdefer[stage], 1u);
                    a.xdefer_list[q] = ((unsigned long long)k << 40) |
                                       ((unsigned long long)r << 20) | (unsigned long long)c;
                    continue;
                }
            } else {
                fault |= face64<1, ORDER, CHAR, !MIXED>(make_consts(pr), pr, solver, st, f);
            }
            if constexpr (BUF) {
#pragma unroll
                for (int v = 0; v < 4; v++) buffer_store_f64(f[v], rf, lb, (int)(rb + v * pl8), 0);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) F[(size_t)v * pl + c0] = f[v];
            }
        }
        if (!MIXED) {
            const unsigned long long any = __ballot(fault != 0);
            if (any && lane == 0) {
                atomicOr(a.faults, 1u);
                atomicOr(&a.stat[k].faults, 1u);
            }
        }
    }
}

// The deferred x faces recomputed with the FP64 path (k_flux_fix's arithmetic).
template <int ORDER, bool CHAR>
__global__ __launch_bounds__(256) void k_batch_xfix(BatchArgs a, int stage, int yi) {
    const unsigned n = a.sched->xdefer[stage];
    if (n && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(a.faults, 4u);
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const unsigned long long e = a.xdefer_list[q];
        const int k = (int)(e >> 40), r = (int)((e >> 20) & 0xFFFFF), c = (int)(e & 0xFFFFF);
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const double* Y = a.pool[yi] + d.off;
        double* F = a.flux + d.off;
        const size_t c0 = (size_t)r * g.ld + c;
        double st[6][4], f[4];
#pragma unroll
        for (int m = 0; m < 6; m++)
#pragma unroll
            for (int v = 0; v < 4; v++) st[m][v] = Y[(size_t)v * g.plane + c0 + (size_t)(m - 2) * g.ld];
        if (face64<1, ORDER, CHAR, true>(make_consts(d.ph), d.ph, d.solver, st, f)) {
            atomicOr(a.faults, 1u);
            atomicOr(&a.stat[k].faults, 1u);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) F[(size_t)v * g.plane + c0] = f[v];
    }
}

// The y faces of every active strip fused with the residual and stage combine (the
// single path's k_yflux_update, one row per step).  Mixed form: a cell next to a y face
// whose FP32 states are not admissible is left to k_batch_fix (which recomputes its four
// faces).  The next step's wave-speed maxima come from k_batch_speeds_all afterwards:
// folding them in here (and loading the x fluxes ahead of the face arithmetic, as the
// single path does) would cost this kernel 18-22 VGPRs each and its fourth wave/SIMD.
template <int ORDER, bool CHAR, bool MIXED>
__global__ __launch_bounds__(64) void k_batch_ysweep(BatchArgs a, int stage, int kind, int yi,
                                                     int oi) {
    const int lane = threadIdx.x;
    const int nact = a.sched->nact, total = a.sched->total;
    unsigned* queue = &a.sched->queue[stage];
    for (int trip = 0; trip <= total; trip++) {
        const unsigned ticket = atomicAdd(queue, lane == 0 ? 1u : 0u);
        const int item = uniform((int)ticket);
        if (item >= total) break;
        const int ai = uniform(owner(a, nact, item));
        const int k = uniform(a.act_prob[ai]);
        const int local = item - a.act_start[ai];
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const int solver = d.solver;
        const double dt = a.stat[k].dt;
        const double* Y = a.pool[yi] + d.off;
        const double* Un = a.pool[0] + d.off;
        double* out = a.pool[oi] + d.off;
        const double* F = a.flux + d.off;
        const int sx = local / d.strips_y, sy = local - sx * d.strips_y;
        const int i0 = GC + sx * a.rows;
        const int i1 = min(i0 + a.rows, g.R - GC);
        const int jf = GC + sy * kStripCols - 1 + lane;   // this lane's y face G[i][jf]
        const int c = jf + 1;                             // this lane's cell column
        const bool face_ok = jf <= g.C - GC - 1;
        const bool cell_ok = lane < kStripCols && c <= g.C - GC - 1;
        int fault = 0;
        auto row_phys = [&]() {
            const BProb* dq = a.probs + k;
            asm volatile("" : "+s"(dq));
            return dq->ph;
        };
        for (int i = i0; i < i1; i++) {
            size_t ld = g.ld, pl = g.plane;
            asm volatile("" : "+v"(ld), "+v"(pl));
            const size_t row = (size_t)opaque_row(i) * ld;
            double gf[4] = {0.0, 0.0, 0.0, 0.0};
            int bad_g = 0;
            if (face_ok) {
                double st[6][4];
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * pl + row + jf - 2 + n];
                const Phys pr = row_phys();
                if constexpr (MIXED) {
                    bad_g = face_mx<2, ORDER, CHAR>(pr, solver, st, gf) ? 0 : 1;
                } else {
                    fault |= face64<2, ORDER, CHAR, !MIXED>(make_consts(pr), pr, solver, st, gf);
                }
            }
            double gu[4];
#pragma unroll
            for (int v = 0; v < 4; v++) gu[v] = __shfl_down(gf[v], 1, 64);
            const int bad_gu = __shfl_down(bad_g, 1, 64);
            if (cell_ok) {
                const size_t o0 = row + c;
                if (MIXED && (bad_g || bad_gu)) {
                    const unsigned q = atomicAdd(&a.sched->defer[stage], 1u);
                    a.defer_list[q] = ((unsigned long long)k << 40) |
                                      ((unsigned long long)i << 20) | (unsigned long long)c;
                } else {
                    double y[4], un[4] = {0.0, 0.0, 0.0, 0.0}, fl[4], fu[4], r[4];
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        y[v] = Y[(size_t)v * pl + o0];
                        fu[v] = F[(size_t)v * pl + o0];
                        fl[v] = F[(size_t)v * pl + o0 - ld];
                        if (kind != UPD_EULER) un[v] = Un[(size_t)v * pl + o0];
                    }
                    const Phys pc = row_phys();
                    double W[4];
                    combine(pc, dt, kind, y, un, fl, fu, gf, gu, r, W);
#pragma unroll
                    for (int v = 0; v < 4; v++) out[(size_t)v * pl + o0] = r[v];
                    if (a.positivity) repair_candidate(a, d, pc, k, i, c, W);
                }
            }
        }
        if (!MIXED) {
            const unsigned long long any = __ballot(fault != 0);
            if (any && lane == 0) {
                atomicOr(a.faults, 1u);
                atomicOr(&a.stat[k].faults, 1u);
            }
        }
    }
}

// k_batch_ysweep with its inputs staged through LDS, as the single path's marching form
// (yflux_update_march): for each row of a strip, the stencil window, the x-flux row and
// the U^n row reach LDS by direct buffer loads (no VGPRs), the next row's window requested
// before the current row's face arithmetic and waited for after it; the lower x-flux row
// is carried in registers.  Same arithmetic and results as k_batch_ysweep.  Used when
// every problem's arrays are below 4 GB (32-bit buffer offsets, BatchArgs::buf_ok).
// Waves/SIMD: mixed 4 (128 VGPRs, 3 dwords spilled; bounded to 3 it does not spill and the
// batch runs 2% slower); FP64 3 (168 VGPRs, 4 spilled; unbounded it takes 179).
#ifndef RPFVM_BATCH_MIXED_WAVES
#define RPFVM_BATCH_MIXED_WAVES 4
#endif
#ifndef RPFVM_BATCH_FP64_WAVES
#define RPFVM_BATCH_FP64_WAVES 3
#endif
template <int ORDER, bool CHAR, bool MIXED>
__global__ __launch_bounds__(64, MIXED ? RPFVM_BATCH_MIXED_WAVES : RPFVM_BATCH_FP64_WAVES) void k_batch_ysweep_lds(BatchArgs a, int stage, int kind,
                                                            int yi, int oi) {
    __shared__ double lds[kMarchLds];
    double* const win = lds;                      // [2][4][kWinPitch]
    double* const frow = lds + 2 * 4 * kWinPitch;  // [4][kRowPitch]
    double* const urow = frow + 4 * kRowPitch;     // [4][kRowPitch]
    const int lane = threadIdx.x;
    const int nact = a.sched->nact, total = a.sched->total;
    unsigned* queue = &a.sched->queue[stage];
    const bool need_un = kind != UPD_EULER;
    for (int trip = 0; trip <= total; trip++) {
        const unsigned ticket = atomicAdd(queue, lane == 0 ? 1u : 0u);
        const int item = uniform((int)ticket);
        if (item >= total) break;
        const int ai = uniform(owner(a, nact, item));
        const int k = uniform(a.act_prob[ai]);
        const int local = item - a.act_start[ai];
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const int solver = d.solver;
        const double dt = a.stat[k].dt;
        const int sx = local / d.strips_y, sy = local - sx * d.strips_y;
        const int i0 = GC + sx * a.rows;
        const int i1 = min(i0 + a.rows, g.R - GC);
        const int f0 = GC - 1 + sy * kStripCols;
        const int jf = f0 + lane;  // this lane's y face G[i][jf]
        const int c = jf + 1;      // this lane's cell column
        const bool face_ok = jf <= g.C - GC - 1;
        const bool cell_ok = lane < kStripCols && c <= g.C - GC - 1;
        const unsigned bytes = (unsigned)array_bytes(g);
        const i32x4 ry = buffer_rsrc(a.pool[yi] + d.off, bytes);
        const i32x4 ru = buffer_rsrc(a.pool[0] + d.off, bytes);
        const i32x4 rf = buffer_rsrc(a.flux + d.off, bytes);
        const i32x4 ro = buffer_rsrc(a.pool[oi] + d.off, bytes);
        const unsigned pl8 = (unsigned)g.plane * 8u, ld8 = (unsigned)g.ld * 8u;
        // window cells f0-2 .. (<= C-1); row cells f0+1 .. (<= C-GC-1)
        const int win_words = 2 * min(kWinCells, g.C - (f0 - 2));
        const int row_words = 2 * min(kStripCols, g.C - GC - 1 - f0);
        int woff[3], roff[2];
#pragma unroll
        for (int q = 0; q < 3; q++) woff[q] = 4 * min(64 * q + lane, win_words - 1);
#pragma unroll
        for (int q = 0; q < 2; q++) roff[q] = 4 * min(64 * q + lane, row_words - 1);
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
        auto row_phys = [&]() {
            const BProb* dq = a.probs + k;
            asm volatile("" : "+s"(dq));
            return dq->ph;
        };
        const int cb = c * 8;  // byte offset of the lane's cell in a row
        double fl[4] = {0.0, 0.0, 0.0, 0.0};
        if (cell_ok) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                fl[v] = buffer_load_f64(rf, cb, (int)((unsigned)(i0 - 1) * ld8 + v * pl8), 0);
        }
        fetch_window(i0, 0);
        fetch_rows(i0);
        wait_vmem();
        int fault = 0;
        for (int row = i0; row < i1; row++) {
            const int i = opaque_row(row);
            const int cur = (i - i0) & 1;
            const double* w = win + cur * 4 * kWinPitch;
            if (i + 1 < i1) fetch_window(i + 1, cur ^ 1);  // window cur ^ 1 was read in row i - 1
            double gf[4] = {0.0, 0.0, 0.0, 0.0};
            int bad_g = 0;
            if (face_ok) {
                double st[6][4];
#pragma unroll
                for (int n = 0; n < 6; n++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[n][v] = w[v * kWinPitch + lane + n];
                const Phys pr = row_phys();
                if constexpr (MIXED) {
                    bad_g = face_mx<2, ORDER, CHAR>(pr, solver, st, gf) ? 0 : 1;
                } else {
                    fault |= face64<2, ORDER, CHAR, !MIXED>(make_consts(pr), pr, solver, st, gf);
                }
            }
            wait_vmem();  // window i + 1, the F / U^n rows of i, the stores of row i - 1
            double gu[4];
#pragma unroll
            for (int v = 0; v < 4; v++) gu[v] = __shfl_down(gf[v], 1, 64);
            const int bad_gu = __shfl_down(bad_g, 1, 64);
            if (cell_ok) {
                double fu[4];
#pragma unroll
                for (int v = 0; v < 4; v++) fu[v] = frow[v * kRowPitch + lane];
                if (MIXED && (bad_g || bad_gu)) {
                    const unsigned q = atomicAdd(&a.sched->defer[stage], 1u);
                    a.defer_list[q] = ((unsigned long long)k << 40) |
                                      ((unsigned long long)i << 20) | (unsigned long long)c;
                } else {
                    double y[4], un[4] = {0.0, 0.0, 0.0, 0.0}, r[4];
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        y[v] = w[v * kWinPitch + lane + 3];
                        if (need_un) un[v] = urow[v * kRowPitch + lane];
                    }
                    const Phys pc = row_phys();
                    double W[4];
                    combine(pc, dt, kind, y, un, fl, fu, gf, gu, r, W);
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        buffer_store_f64(r[v], ro, cb, (int)((unsigned)i * ld8 + v * pl8), 0);
                    if (a.positivity) repair_candidate(a, d, pc, k, i, c, W);
                }
#pragma unroll
                for (int v = 0; v < 4; v++) fl[v] = fu[v];
            }
            wait_lds();  // this row's reads of the window and the F / U^n rows are complete
            if (i + 1 < i1) fetch_rows(i + 1);
        }
        if (!MIXED) {
            const unsigned long long any = __ballot(fault != 0);
            if (any && lane == 0) {
                atomicOr(a.faults, 1u);
                atomicOr(&a.stat[k].faults, 1u);
            }
        }
    }
}

// Wave-speed maxima of every active problem's new state (the two-sweep schedule's last
// stage): grid (blocks, problems); the same per-cell arithmetic and NaN-ignoring max as
// k_dt_partial.
__global__ __launch_bounds__(256) void k_batch_speeds_all(BatchArgs a, int pi) {
    for (int k = blockIdx.y; k < a.nprob; k += gridDim.y) {
        if (!a.stat[k].active) continue;
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const double* U = a.pool[pi] + d.off;
        double sx = 0.0, sy = 0.0;
        const long total = (long)d.nx * d.ny;
        for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
             t += (long)gridDim.x * blockDim.x) {
            const int i = GC + (int)(t / d.ny), j = GC + (int)(t % d.ny);
            const size_t o = (size_t)i * g.ld + j;
            double c[4];
#pragma unroll
            for (int v = 0; v < 4; v++) c[v] = U[(size_t)v * g.plane + o];
            double lx, ly;
            speeds(d.ph.gamma, c, lx, ly);
            if (lx > sx) sx = lx;
            if (ly > sy) sy = ly;
        }
        sx = wave_max_b(sx);
        sy = wave_max_b(sy);
        if ((threadIdx.x & 63) == 0) speed_max(a.stat[k], sx, sy);
    }
}

// ------------------------------------------------------------------ fix-up
// Mixed form: each deferred cell with its four faces recomputed (mixed arithmetic,
// FP64 for a face whose FP32 states are not admissible) -- the single-problem mixed
// path's fix-up (k_fixfaces_mx + k_fixcells_mx), per cell.
template <int ORDER, bool CHAR>
__global__ __launch_bounds__(256) void k_batch_fix(BatchArgs a, int stage, int kind, int yi,
                                                   int oi, int last) {
    const unsigned n = a.sched->defer[stage];
    if (n && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(a.faults, 4u);
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const unsigned long long e = a.defer_list[q];
        const int k = (int)(e >> 40), i = (int)((e >> 20) & 0xFFFFF), c = (int)(e & 0xFFFFF);
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const Phys ph = d.ph;
        const double* Y = a.pool[yi] + d.off;
        const double* Un = a.pool[0] + d.off;
        double* out = a.pool[oi] + d.off;
        const Consts kc = make_consts(ph);
        double fl[4], fu[4], gl[4], gu[4];
        double* faces[4] = {fl, fu, gl, gu};
        int fault = 0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            double st[6][4];
            if (m < 2) {
                const int r = i - 1 + m;  // x face between rows r and r + 1
#pragma unroll
                for (int s = 0; s < 6; s++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[s][v] = bload(Y, d, r - 2 + s, c, v);
                if (!face_mx<1, ORDER, CHAR>(ph, d.solver, st, faces[m]))
                    fault |= face64<1, ORDER, CHAR, true>(kc, ph, d.solver, st, faces[m]);
            } else {
                const int jf = c - 1 + (m - 2);  // y face between columns jf and jf + 1
#pragma unroll
                for (int s = 0; s < 6; s++)
#pragma unroll
                    for (int v = 0; v < 4; v++) st[s][v] = bload(Y, d, i, jf - 2 + s, v);
                if (!face_mx<2, ORDER, CHAR>(ph, d.solver, st, faces[m]))
                    fault |= face64<2, ORDER, CHAR, true>(kc, ph, d.solver, st, faces[m]);
            }
        }
        const size_t o0 = (size_t)i * g.ld + c;
        double y[4], un[4] = {0.0, 0.0, 0.0, 0.0}, r[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            y[v] = Y[(size_t)v * g.plane + o0];
            if (kind != UPD_EULER) un[v] = Un[(size_t)v * g.plane + o0];
        }
        double W[4];
        combine(ph, a.stat[k].dt, kind, y, un, fl, fu, gl, gu, r, W);
#pragma unroll
        for (int v = 0; v < 4; v++) out[(size_t)v * g.plane + o0] = r[v];
        if (a.positivity) repair_candidate(a, d, ph, k, i, c, W);
        if (last) {
            double lx, ly;
            speeds(ph.gamma, r, lx, ly);
            speed_max(a.stat[k], lx > 0.0 ? lx : 0.0, ly > 0.0 ? ly : 0.0);
        }
        if (fault) {
            atomicOr(a.faults, 1u);
            atomicOr(&a.stat[k].faults, 1u);
        }
    }
}

// ------------------------------------------------------- positivity safeguard
// The single-problem path's rounds (fvm_repair.hpp) over the batch-wide lists, in one
// workgroup with __syncthreads() between the phases: the cells of list r are marked
// repaired in this stage, every cell beside them (and they themselves) is touched, and
// each touched cell is updated again; one whose forward-Euler state is still not
// admissible goes to list r + 1, or raises RPFVM_FAULT_POSITIVITY if it was repaired
// already.  Problems do not interact, so each problem's rounds are the single path's.
// A face is replaced once a cell beside it is repaired (rep == stamp).  The single path
// stores the replacement HLLE fluxes; here the cells phase evaluates them where it needs
// them, and takes every other face from the stage's own results: x faces from the flux
// pool, y faces recomputed with the sweep's arithmetic (the mixed form's FP32 face, or
// its FP64 face where the sweep deferred it) -- the same values, so the same cells.
template <int ORDER, bool CHAR, bool MIXED>
__device__ __forceinline__ int sweep_yface(const BProb& d, const Phys& ph, const double* Y, int i,
                                           int f, double out[4]) {
    const Geom& g = d.g;
    const size_t c0 = (size_t)i * g.ld + f;
    double st[6][4];
#pragma unroll
    for (int n = 0; n < 6; n++)
#pragma unroll
        for (int v = 0; v < 4; v++) st[n][v] = Y[(size_t)v * g.plane + c0 + (n - 2)];
    if constexpr (MIXED) {
        if (face_mx<2, ORDER, CHAR>(ph, d.solver, st, out)) return 0;
        return face64<2, ORDER, CHAR, true>(make_consts(ph), ph, d.solver, st, out);
    } else {
        return face64<2, ORDER, CHAR, true>(make_consts(ph), ph, d.solver, st, out);
    }
}

// HLLE flux of the face between cells a and b = a + step (offsets in a plane).
template <int DIR>
__device__ __forceinline__ int pair_hlle(const Consts& kc, const double* Y, const Geom& g, size_t a,
                                         size_t step, double f[4]) {
    double l[4], r[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        l[v] = Y[(size_t)v * g.plane + a];
        r[v] = Y[(size_t)v * g.plane + a + step];
    }
    return hlle<DIR>(kc, l, r, f);
}

constexpr unsigned long long kCellMask = (1ull << 40) - 1;

__device__ __forceinline__ unsigned cnt_load(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int ORDER, bool CHAR, bool MIXED>
__global__ __launch_bounds__(256) void k_batch_repair(BatchArgs a, int stage, int kind, int yi,
                                                      int oi) {
    const BRepair& rp = a.rep;
    if (cnt_load(&rp.cnt[0]) == 0u) return;
    const unsigned stamp = a.sched->stamp + (unsigned)stage;
    const unsigned tid = threadIdx.x, nth = blockDim.x;
    int r = 0;
    for (; r < kRepairRounds; r++) {
        const int cur = r & 1, nxt = cur ^ 1;
        const unsigned n = cnt_load(&rp.cnt[cur]);
        if (n == 0u) break;
        __syncthreads();  // every thread has read n before the counters are reset
        if (tid == 0) {
            rp.cnt[nxt] = 0u;
            rp.cnt[2] = 0u;
            atomicAdd(&rp.cnt[3], n);
        }
        // faces: the listed cells are repaired in this stage
        for (unsigned q = tid; q < n; q += nth) {
            const unsigned long long e = rp.list[cur][q];
            rp.rep[a.probs[(int)(e >> 40)].cbase + (e & kCellMask)] = stamp;
        }
        __syncthreads();
        // touch: each listed cell and its four neighbours, once per round
        const unsigned tstamp = stamp * 16u + (unsigned)r;
        for (unsigned q = tid; q < 5u * n; q += nth) {
            const unsigned long long e = rp.list[cur][q / 5u];
            const int k = (int)(e >> 40);
            const BProb& d = a.probs[k];
            const long c = (long)(e & kCellMask);
            const unsigned m = q % 5u;
            int ti = (int)(c / d.ny), tj = (int)(c % d.ny);
            ti += m == 1 ? -1 : m == 2 ? 1 : 0;
            tj += m == 3 ? -1 : m == 4 ? 1 : 0;
            if (ti < 0 || ti >= d.nx || tj < 0 || tj >= d.ny) continue;
            const long cc = (long)ti * d.ny + tj;
            if (atomicExch(&rp.tch[d.cbase + cc], tstamp) != tstamp) {
                const unsigned w = atomicAdd(&rp.cnt[2], 1u);
                rp.touched[w] = ((unsigned long long)k << 40) | (unsigned long long)cc;
            }
        }
        __syncthreads();
        // cells: update every touched cell again with its faces as they stand now
        const unsigned nt = cnt_load(&rp.cnt[2]);
        for (unsigned q = tid; q < nt; q += nth) {
            const unsigned long long e = rp.touched[q];
            const int k = (int)(e >> 40);
            const BProb& d = a.probs[k];
            const Geom g = d.g;
            const Phys ph = d.ph;
            const Consts kc = make_consts(ph);
            const long c = (long)(e & kCellMask);
            const int i = GC + (int)(c / d.ny), j = GC + (int)(c % d.ny);
            const unsigned* rep = rp.rep + d.cbase;
            const double* Y = a.pool[yi] + d.off;
            const double* F = a.flux + d.off;
            const size_t o = (size_t)i * g.ld + j;
            const bool here = rep[c] == stamp;
            int fault = 0;
            double fl[4], fu[4], gl[4], gu[4];
            if (here || (i - 1 >= GC && rep[c - d.ny] == stamp)) {
                fault |= pair_hlle<1>(kc, Y, g, o - g.ld, g.ld, fl);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) fl[v] = F[(size_t)v * g.plane + o - g.ld];
            }
            if (here || (i + 1 <= g.R - GC - 1 && rep[c + d.ny] == stamp)) {
                fault |= pair_hlle<1>(kc, Y, g, o, g.ld, fu);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) fu[v] = F[(size_t)v * g.plane + o];
            }
            if (here || (j - 1 >= GC && rep[c - 1] == stamp))
                fault |= pair_hlle<2>(kc, Y, g, o - 1, 1, gl);
            else
                fault |= sweep_yface<ORDER, CHAR, MIXED>(d, ph, Y, i, j - 1, gl);
            if (here || (j + 1 <= g.C - GC - 1 && rep[c + 1] == stamp))
                fault |= pair_hlle<2>(kc, Y, g, o, 1, gu);
            else
                fault |= sweep_yface<ORDER, CHAR, MIXED>(d, ph, Y, i, j, gu);
            double y[4], un[4] = {0.0, 0.0, 0.0, 0.0}, res[4], W[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                y[v] = Y[(size_t)v * g.plane + o];
                if (kind != UPD_EULER) un[v] = a.pool[0][d.off + (size_t)v * g.plane + o];
            }
            combine(ph, a.stat[k].dt, kind, y, un, fl, fu, gl, gu, res, W);
            double* out = a.pool[oi] + d.off;
#pragma unroll
            for (int v = 0; v < 4; v++) out[(size_t)v * g.plane + o] = res[v];
            if (fault) {
                atomicOr(a.faults, 1u);
                atomicOr(&a.stat[k].faults, 1u);
            }
            if (!admissible(kc, W)) {
                if (!here) {
                    const unsigned w = atomicAdd(&rp.cnt[nxt], 1u);
                    rp.list[nxt][w] = e;
                } else {
                    atomicOr(a.faults, 8u);
                    atomicOr(&a.stat[k].faults, 8u);
                }
            }
        }
        __syncthreads();
    }
    if (r == kRepairRounds) {  // cells still listed after the last round
        const int cur = kRepairRounds & 1;
        const unsigned n = cnt_load(&rp.cnt[cur]);
        for (unsigned q = tid; q < n; q += nth) {
            const int k = (int)(rp.list[cur][q] >> 40);
            atomicOr(a.faults, 8u);
            atomicOr(&a.stat[k].faults, 8u);
        }
        __syncthreads();
    }
    if (tid == 0) {
        rp.cnt[0] = 0u;
        rp.cnt[1] = 0u;
        rp.cnt[2] = 0u;
    }
}

// ------------------------------------------------------------- RK1, speeds
__global__ __launch_bounds__(256) void k_batch_copy_back(BatchArgs a) {
    const int nact = a.sched->nact;
    for (int ai = blockIdx.y; ai < nact; ai += gridDim.y) {
        const int k = a.act_prob[ai];
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const long total = (long)d.nx * d.ny;
        for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
             t += (long)gridDim.x * blockDim.x) {
            const int i = GC + (int)(t / d.ny), j = GC + (int)(t % d.ny);
            const size_t o = d.off + (size_t)i * g.ld + j;
#pragma unroll
            for (int v = 0; v < 4; v++) a.pool[0][(size_t)v * g.plane + o] = a.pool[1][(size_t)v * g.plane + o];
        }
    }
}

// Pools swapped every step: the interior of each problem that sits out this step and
// is not mirrored yet, pool[0] -> pool[3] (k_batch_control set bit 1 of its flags).
__global__ __launch_bounds__(256) void k_batch_mirror(BatchArgs a) {
    for (int k = blockIdx.y; k < a.nprob; k += gridDim.y) {
        if (!(a.stat[k].mirror & 2)) continue;
        const BProb& d = a.probs[k];
        const Geom g = d.g;
        const long total = (long)d.nx * d.ny;
        for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
             t += (long)gridDim.x * blockDim.x) {
            const int i = GC + (int)(t / d.ny), j = GC + (int)(t % d.ny);
            const size_t o = d.off + (size_t)i * g.ld + j;
#pragma unroll
            for (int v = 0; v < 4; v++) a.pool[3][(size_t)v * g.plane + o] = a.pool[0][(size_t)v * g.plane + o];
        }
    }
}

// Clock reset and wave-speed maxima of problem k's current state (after set_state).
__global__ void k_batch_reset(BatchArgs a, int k) {
    if (threadIdx.x == 0) {
        BStat& s = a.stat[k];
        s.t = 0.0;
        s.dt = 0.0;
        s.smax[0] = 0ull;
        s.smax[1] = 0ull;
        s.steps = 0;
        s.active = 0;
        s.faults = 0u;
        s.mirror = 0;
    }
}

__global__ __launch_bounds__(256) void k_batch_speeds(BatchArgs a, int k) {
    const BProb& d = a.probs[k];
    const Geom g = d.g;
    const double* U = a.pool[0] + d.off;
    double sx = 0.0, sy = 0.0;
    const long total = (long)d.nx * d.ny;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int i = GC + (int)(t / d.ny), j = GC + (int)(t % d.ny);
        const size_t o = (size_t)i * g.ld + j;
        double c[4];
#pragma unroll
        for (int v = 0; v < 4; v++) c[v] = U[(size_t)v * g.plane + o];
        double lx, ly;
        speeds(d.ph.gamma, c, lx, ly);
        if (lx > sx) sx = lx;
        if (ly > sy) sy = ly;
    }
    sx = wave_max_b(sx);
    sy = wave_max_b(sy);
    if ((threadIdx.x & 63) == 0) speed_max(a.stat[k], sx, sy);
}

// --------------------------------------------------------------- launchers
namespace {

template <int ORDER, bool CHAR, bool MIXED>
void stage_launch(const BatchArgs& a, int time_order, int stage, const BatchStageLaunch& l,
                  hipStream_t s) {
    // (Y, out, kind, last) per stage, include/time_advance.h:643-882
    int yi = 0, oi = 1, kind = UPD_EULER, last = 0;
    if (time_order == 1) {
        last = 1;
    } else if (time_order == 2) {
        if (stage == 1) {
            yi = 1;
            oi = 0;
            kind = UPD_RK2_2;
            last = 1;
        }
    } else {
        if (stage == 1) {
            yi = 1;
            oi = 2;
            kind = UPD_RK3_2;
        } else if (stage == 2) {
            yi = 2;
            oi = 0;
            kind = UPD_RK3_3;
            last = 1;
        }
    }
    // with the safeguard the last stage of RK2 / RK3 writes pool[3] (a repair recomputes
    // cells from U^n, which an in-place update would have overwritten)
    if (a.positivity && time_order > 1 && last) oi = 3;
    if (l.staging == STAGING_SWEEPS) {
        // blocks cover the longest side of any problem; each side checks its own range
        const dim3 gg((a.max_side + 63) / 64, 4, a.nprob < 65535 ? a.nprob : 65535);
        for (int side = 0; side < 4; side++)
            hipLaunchKernelGGL(k_batch_ghosts, gg, dim3(64), 0, s, a, side, yi);
        if (a.buf_ok)
            hipLaunchKernelGGL((k_batch_xsweep<ORDER, CHAR, MIXED, true>), dim3(l.grid), dim3(64), 0,
                               s, a, stage, yi);
        else
            hipLaunchKernelGGL((k_batch_xsweep<ORDER, CHAR, MIXED, false>), dim3(l.grid), dim3(64), 0,
                               s, a, stage, yi);
        if constexpr (MIXED)
            hipLaunchKernelGGL((k_batch_xfix<ORDER, CHAR>), dim3(256), dim3(256), 0, s, a, stage, yi);
        if (a.buf_ok)
            hipLaunchKernelGGL((k_batch_ysweep_lds<ORDER, CHAR, MIXED>), dim3(l.grid), dim3(64), 0, s,
                               a, stage, kind, yi, oi);
        else
            hipLaunchKernelGGL((k_batch_ysweep<ORDER, CHAR, MIXED>), dim3(l.grid), dim3(64), 0, s, a,
                               stage, kind, yi, oi);
        if constexpr (MIXED)
            hipLaunchKernelGGL((k_batch_fix<ORDER, CHAR>), dim3(1024), dim3(256), 0, s, a, stage,
                               kind, yi, oi, 0);
        if (a.positivity)
            hipLaunchKernelGGL((k_batch_repair<ORDER, CHAR, MIXED>), dim3(1), dim3(256), 0, s, a,
                               stage, kind, yi, oi);
        // RK1 writes U1 and copies back afterwards: its maxima come after the copy
        if (last && time_order != 1)
            launch_batch_speeds_all(a, oi, s);
        return;
    } else if (l.staging == STAGING_LDS)
        hipLaunchKernelGGL((k_batch_stage<ORDER, CHAR, MIXED, STAGING_LDS>), dim3(l.grid), dim3(64),
                           0, s, a, stage, kind, yi, oi, last);
    else
        hipLaunchKernelGGL((k_batch_stage<ORDER, CHAR, MIXED, STAGING_DIRECT>), dim3(l.grid),
                           dim3(64), 0, s, a, stage, kind, yi, oi, last);
    if constexpr (MIXED)
        // 1024 workgroups: at one wavefront per SIMD (its FP64 fallback path needs 320
        // registers) that covers every SIMD of the chip
        hipLaunchKernelGGL((k_batch_fix<ORDER, CHAR>), dim3(1024), dim3(256), 0, s, a, stage, kind,
                           yi, oi, last);
}

template <int ORDER>
BatchKernels select_order(bool characteristic, bool mixed) {
    if (mixed)
        return {characteristic ? &stage_launch<ORDER, true, true> : &stage_launch<ORDER, false, true>};
    return {characteristic ? &stage_launch<ORDER, true, false> : &stage_launch<ORDER, false, false>};
}

}  // namespace

BatchKernels select_batch(int order, bool characteristic, bool mixed) {
    switch (order) {  // the API admits 1, 2, 3 and 5
        case 1: return select_order<1>(characteristic, mixed);
        case 2: return select_order<2>(characteristic, mixed);
        case 3: return select_order<3>(characteristic, mixed);
        default: return select_order<5>(characteristic, mixed);
    }
}

void launch_batch_control(const BatchArgs& a, int mode, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_control, dim3(1), dim3(256), 0, s, a, mode);
}

void launch_batch_copy_back(const BatchArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_copy_back, dim3(64, 64), dim3(256), 0, s, a);
}

void launch_batch_reset_speeds(const BatchArgs& a, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a, k);
}

void launch_batch_speeds_all(const BatchArgs& a, int pi, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_speeds_all, dim3(64, a.nprob < 65535 ? a.nprob : 65535), dim3(256),
                       0, s, a, pi);
}

void launch_batch_mirror(const BatchArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_mirror, dim3(64, a.nprob < 65535 ? a.nprob : 65535), dim3(256), 0, s, a);
}

void launch_batch_speeds(const BatchArgs& a, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_speeds, dim3(256), dim3(256), 0, s, a, k);
}

}  // namespace fvm
