// fvm_kernels.hpp -- launch interface between the C ABI (fvm_api.cpp) and the
// gfx950 kernels (fvm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace fvm {

constexpr int GC = 4;  // GhostCell, include/Golbal.h:49

// Minimum resident wavefronts per SIMD the face-flux kernels are compiled for (0: the
// compiler's choice).  4 caps them at 128 VGPRs (with spills); build-time tuning knob.
#ifndef RPFVM_FLUX_WAVES
#define RPFVM_FLUX_WAVES 0
#endif
#if RPFVM_FLUX_WAVES > 0
#define FLUX_BOUNDS __launch_bounds__(256, RPFVM_FLUX_WAVES)
#else
#define FLUX_BOUNDS __launch_bounds__(256)
#endif
// Grid rows per workgroup of the fused y-flux + update kernels (64 lanes per row).
#ifndef RPFVM_FUSED_ROWS
#define RPFVM_FUSED_ROWS 2
#endif
#if RPFVM_FLUX_WAVES > 0
#define FUSED_BOUNDS __launch_bounds__(64 * RPFVM_FUSED_ROWS, RPFVM_FLUX_WAVES)
#else
#define FUSED_BOUNDS __launch_bounds__(64 * RPFVM_FUSED_ROWS)
#endif
// The waves knob for the mixed form's fused kernel alone.
#ifndef RPFVM_FUSED_MX_WAVES
#define RPFVM_FUSED_MX_WAVES RPFVM_FLUX_WAVES
#endif
#if RPFVM_FUSED_MX_WAVES > 0
#define FUSED_MX_BOUNDS __launch_bounds__(64 * RPFVM_FUSED_ROWS, RPFVM_FUSED_MX_WAVES)
#else
#define FUSED_MX_BOUNDS __launch_bounds__(64 * RPFVM_FUSED_ROWS)
#endif

// Device grid geometry: planes of R rows x ld doubles; plane v at base + v*plane.
struct Geom {
    int R, C;       // rows (x, with ghosts) and logical columns (y, with ghosts)
    int ld;         // row pitch in doubles (>= C)
    size_t plane;   // plane stride in doubles
};

// The step's control block in device memory.  Every kernel of a step reads dt and the
// stage clock from it and does nothing when `active` is 0, so a step can be enqueued
// before its dt is known: rpfvm_time_advance writes a host dt into it (k_step_set), the
// device-side loop of rpfvm_advance_to has the CFL reduction write it (k_dt_loop).
struct StepCtl {
    double dt;        // the step's dt
    double t0;        // clock at the step's start: stage c runs its ghost fill at t0 + c dt
    double t;         // device loop: the clock after the last step taken
    double t_stop;    // device loop: the end time
    unsigned active;  // 1: the step's kernels run
    unsigned steps;   // device loop: steps taken
    unsigned stop;    // device loop: ended (t_stop reached, or dt did not advance the clock)
    unsigned pad;
};

// Physical and scheme parameters read by the kernels.
struct Phys {
    double gamma, dx, dy, cfl, gravity;
    double gm1, gr, six;  // gamma - 1, gamma / (gamma - 1), 6.0 (host-computed)
    double igm1;          // 1 / (gamma - 1) (host-computed; IEEE division rounds alike)
    double idx, idy;      // 1 / dx, 1 / dy (host-computed; IEEE division rounds alike)
    int source, positivity;
    int force_fix;   // mixed form: send every face to the FP64 recompute list
    int row_forms;    // the row forms of the sweeps (one row per wavefront, global addressing):
                      // arrays of 4 GB and more, grids too small to fill the GPU with
                      // marching wavefronts, or on request
    const StepCtl* step;  // the step's control block (null where no step runs)
};

// Kernel entry of a step: false when the step is not taken (device loop past its end).
__device__ __forceinline__ bool step_on(const Phys& ph) { return ph.step->active != 0u; }

enum : int { BC_OUT = 0, BC_REFLECT = 1, BC_PERIODIC = 2, BC_INFLOW = 3, BC_FIXED = 4,
             BC_DMR_WALL = 5, BC_DMR_SHOCK = 6 };

// One side of the ghost fill: side 0..3 = left, right, bottom, top; tangential index
// range [lo, hi] (inclusive); value[] is the inflow or fixed state when used.  The
// double-Mach-reflection sides also use alt[] (state right of the shock), par[] (wall
// start, or shock start / height / slope / speed), the cell-centre map x_i =
// ((i + 0.5 - GC) / nx) * lx and the stage time t (k_bc_side sets it from the step's
// clock: t0 + c dt).
struct BcSide {
    int side, type, lo, hi;
    double value[4];
    double alt[4];
    double par[4];
    double nx, lx, t;
    double c;  // the stage's time coefficient (SSP-RK3: 0, 1, 1/2)
};

enum : int { UPD_RESIDUAL = 0, UPD_EULER = 1, UPD_RK2_2 = 2, UPD_RK3_2 = 3, UPD_RK3_3 = 4 };

// Device lists of the positivity safeguard (all null when it is off).  Cells are
// numbered over the interior, c = (i - GC) * ny + (j - GC).
struct Repair {
    unsigned* cnt;       // [0], [1]: list lengths, [2]: touched length, [3]: repairs total
    int* list[2];        // cells to repair in the current / next round
    int* touched;        // cells to update again in the current round
    unsigned* rep;       // rep[c] == stage stamp: c was repaired in this stage
    unsigned* tch;       // tch[c] == round stamp: c is already in the touched list
    int nx, ny;
};
constexpr int kRepairRounds = 8;

// Faces of the mixed form that need the FP64 path.  The two-sweep stage resets the lists
// alternately (each direction's fix-up clears the other's count); the one-pass stage
// fills both and its cell fix-up clears both (cnt[2] counts its finished workgroups).
struct FaceFix {
    unsigned* cnt;  // [0] x-face count, [1] y-face count, [2] workgroups done
    int* list[2];   // face offsets i*ld + j
    unsigned cap;   // entries per list (every face, plus the faces two wavefronts share)
};

// Wave-speed maxima of the last stage's new state, one slot per wavefront of the fused
// y-flux kernel: slot w = ((i - GC) / rows) * nstrips + (j - GC) / 63 holds the bit
// patterns of max(|u|+a) and max(|v|+a) over the cells of strip s = w % nstrips (columns
// GC + 63 s .. GC + 63 s + 62) in the `rows` rows from GC + rows (w / nstrips) (non-negative
// doubles order as unsigned integers).  rows: the rows a wavefront of the marching fused y
// sweep walks (which then reduces its lanes' maxima once), 1 for the row forms and the
// one-pass stage.  A kernel that changes such a cell afterwards
// marks its slot stale with the all-ones pattern; the reduction recomputes stale slots
// from the state.  slot == nullptr: not the last stage.
constexpr int kStrip = 63;  // cells updated per 64-lane row of the fused y-flux kernels

struct DtSlots {
    unsigned long long* slot;
    int nstrips;
    int rows;  // grid rows per slot
};
constexpr unsigned long long kSlotStale = ~0ull;

using FluxLaunch = void (*)(const double* Y, double* F, const Geom& g, const Phys& ph,
                            unsigned* faults, const FaceFix& fix, hipStream_t s);
// y faces fused with the residual + stage combine (out must not alias Y); `stamp` is
// the stage's unique stamp for de-duplicating safeguard list entries.
using FusedLaunch = void (*)(const double* Un, const double* Y, double* out, const double* F,
                             double* G, const Geom& g, const Phys& ph, int kind,
                             const Repair& rp, const FaceFix& fix, unsigned stamp,
                             const DtSlots& ds, unsigned* faults, hipStream_t s);
// The positivity safeguard's rounds after a stage (fvm_repair.hpp).
using RepairLaunch = void (*)(const double* Un, const double* Y, double* out, double* F, double* G,
                              const Geom& g, const Phys& ph, int kind, const Repair& rp,
                              unsigned stamp, const DtSlots& ds, unsigned* faults, hipStream_t s);

// The whole stage in one pass (fvm_onepass.hpp): x faces, y faces, residual and stage
// combine (out must not alias Y; buffer forms only).
using StageLaunch = void (*)(const double* Un, const double* Y, double* out, double* F, double* G,
                             const Geom& g, const Phys& ph, int kind, const Repair& rp,
                             const FaceFix& fix, unsigned stamp, const DtSlots& ds,
                             unsigned* faults, hipStream_t s);
// After a stage of the mixed form: the FP64 recomputation of the deferred faces, the update
// of their cells, the positivity safeguard's rounds; clears the lists.
using FixupLaunch = StageLaunch;

struct FluxPair {
    FluxLaunch x, y;      // complete face fluxes into F / G (space operator, unfused stages)
    FluxLaunch xs;        // the two-sweep stage's x sweep (the mixed form defers faces to fixup2)
    FusedLaunch yu;       // y faces fused with the residual and stage combine
    RepairLaunch repair;  // the safeguard's rounds after a stage of a form without fix-ups
    StageLaunch stage;    // the one-pass stage (null: the form has none)
    FixupLaunch fixup2;   // after xs + yu (null: none)
    FixupLaunch fixup1;   // after the one-pass stage
};

FluxPair select_flux(int order, bool characteristic, int solver);
FluxPair select_flux_mixed(int order, bool characteristic, int solver);
// FP64 solver `solver` on face states Ls/Rs into F, faces of direction dir (1 or 2).
void launch_riemann(int dir, int solver, const double* Ls, const double* Rs, double* F,
                    const Geom& g, const Phys& ph, unsigned* faults, hipStream_t s);
void launch_bc_side(double* U, const Geom& g, const BcSide& s, const StepCtl* step, hipStream_t st);
// The four sides in the reference's order L, R, B, T (two launches: L + R, then B + T).
void launch_bc(double* U, const Geom& g, const BcSide side[4], const StepCtl* step, hipStream_t st);
void launch_update(const double* Un, const double* Y, double* out, const double* F,
                   const double* G, const Geom& g, const Phys& ph, int kind,
                   const Repair& rp, hipStream_t s);
void launch_repair(const double* Un, const double* Y, double* out, double* F, double* G,
                   const Geom& g, const Phys& ph, int kind, const Repair& rp,
                   unsigned stamp, const DtSlots& ds, unsigned* faults, hipStream_t s);
// Clears the mixed form's FP64 recompute lists (after the space operator's sweeps).
void launch_fix_reset(const FaceFix& fix, hipStream_t s);
void launch_copy_frame(const double* src, double* dst, const Geom& g, const StepCtl* step, hipStream_t s);
// Writes a host dt and clock into the step's control block and marks the step active.
void launch_step_set(StepCtl* step, double dt, double t0, hipStream_t s);
// Starts the device loop: clock t, end t_stop, no steps taken.
void launch_loop_init(StepCtl* step, double t, double t_stop, hipStream_t s);
// The device loop's CFL reduction: dt from the slots (ds.slot) or a full pass, clipped to
// the loop's end as the driver does (src/RP_FVM.c:186-191); sets the next step's control.
void launch_dt_loop(const double* U, const Geom& g, const Phys& ph, const DtSlots& ds,
                    double* partial, StepCtl* step, unsigned* faults, hipStream_t s);
int dt_blocks(const Geom& g);
void launch_dt(const double* U, const Geom& g, const Phys& ph, double* partial, double* dt_out,
               const double* clock, unsigned* faults, hipStream_t s);
// The same dt from the slots the last stage left (stale slots recomputed from U).
void launch_dt_slots(const double* U, const Geom& g, const Phys& ph, const DtSlots& ds,
                     double* partial, double* dt_out, const double* clock, unsigned* faults,
                     hipStream_t s);
__host__ __device__ inline int dt_slot_count(const Geom& g, int rows) {
    return ((g.R - 2 * GC + rows - 1) / rows) * ((g.C - 2 * GC + kStrip - 1) / kStrip);
}
int march_rows();  // rows a wavefront of the marching fused y sweep walks
// rows x cols doubles from src (row pitch spitch) to dst (row pitch dpitch)
void launch_repitch(const double* src, size_t spitch, double* dst, size_t dpitch, size_t rows,
                    size_t cols, hipStream_t s);

}  // namespace fvm
