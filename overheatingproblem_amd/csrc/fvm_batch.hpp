// fvm_batch.hpp -- many independent problems of uneven size, advanced together.
//
// Every problem keeps its own grid, physics, Riemann solver, boundary conditions, dt
// and clock; reconstruction, time order and precision are shared by the batch.  One
// RK stage of every active problem is ONE persistent launch (k_batch_stage) whose
// wavefronts pull work items ("strips") from a device counter:
//
//   strip = (problem, TX consecutive rows, 63 consecutive columns)
//
// A wavefront marches down its strip one row at a time: lane t owns the y face
// G[i][jf] with jf = j0 - 1 + t and the cell (i, jf + 1); the x-face flux F[i][c] of the
// row is carried in registers to the next row as F[i-1][c], and the cell's upper y face
// comes from lane t + 1 by shuffle.  Face fluxes therefore never reach HBM: per stage
// the kernel reads the stage input (plus U^n) and writes the stage output.  Ghost
// cells are not stored: a halo load outside the interior goes through the side's
// boundary map (the index map and sign of Boundary_Conditions), so there is no
// ghost-fill pass either.
//
// The face and cell arithmetic is the single-problem path's (fvm_math.hpp,
// fvm_mixed.hpp), so each problem's interior is bit-identical to advancing it alone.
// The per-problem CFL reduction of the next step is folded into the last stage
// (atomic max of bit patterns of non-negative doubles, which order as unsigned
// integers), and k_batch_control turns it into each problem's dt, clock, activity and
// the schedule of active strips before stage 1.  The step is a fixed sequence of
// launches on one stream, captured once into a hipGraph and replayed.
//
// The two-sweep schedule (STAGING_SWEEPS) trades the fluxes-in-registers property for
// the single-problem kernels' register profile: per stage, the ghost frames of the
// stage input are filled (the four sides in order, as the single path), a persistent
// sweep computes the x faces of every active strip into a flux buffer, and a second
// persistent sweep computes the y faces and updates the cells with the x fluxes read
// back (the marching kernel keeps both sweeps' state live and runs at 2 waves/SIMD).
// Both schedules use the same strips, work counters and face arithmetic, so their
// results are identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "fvm_kernels.hpp"

namespace fvm {

constexpr int kStripCols = 63;  // cells per strip row (64 y faces)

// How a stage gets its stencils: the one-pass marching kernel reading the stage input
// directly or through an LDS ring, or the two-sweep schedule (ghost fill, x-flux sweep to
// a flux buffer, fused y-flux + update sweep).
enum : int { STAGING_DIRECT = 0, STAGING_LDS = 1, STAGING_SWEEPS = 2 };

// One problem of the batch (device-resident, read with scalar loads).
struct BProb {
    Geom g;
    Phys ph;
    int nx, ny;
    int solver;
    int bc[4];            // L, R, B, T
    double bcv[4][4];     // inflow / fixed state per side
    size_t off;           // offset in doubles of the problem's planes in each pool buffer
    size_t cbase;         // first entry of the problem's interior cells in the repair arrays
    int strips_y;         // ceil(ny / 63)
    int pad;
    BcSide side[4];       // ghost fill of the two-sweep schedule (as the single path's)
};

// Per-problem state of the time loop.
struct BStat {
    double t;                      // time reached after the step in flight
    double dt;                     // dt of the step in flight
    unsigned long long smax[2];    // bits of max(|u|+a), max(|v|+a) over the interior
    int steps;                     // steps taken
    int active;                    // taking part in the step in flight
    unsigned faults;               // RPFVM_FAULT_* bits
    int mirror;                    // pools swapped every step (safeguard): bit 0, both
                                   // pool[0] and pool[3] hold the state; bit 1, copy it
                                   // across in the step in flight (the problem sits out)
};

// Schedule of the step in flight, written by k_batch_control.
struct BSched {
    unsigned queue[4];    // work counters of stages 1..3 (+1 spare)
    unsigned defer[4];    // deferred-cell counts of stages 1..3 (mixed form)
    unsigned xqueue[4];   // two-sweep schedule: x-sweep work counters
    unsigned xdefer[4];   // two-sweep schedule: deferred x faces (mixed form)
    int nact;             // active problems
    int total;            // active strips
    int active_left;      // problems with t < t_end after this step's dt
    unsigned stamp;       // positivity safeguard: stamp of stage s is stamp + s (never 0)
    double t_end;         // stop time (set by the host)
};

// Lists of the positivity safeguard over the whole batch (all null when it is off).
// Entries are (problem << 40) | c with c = (i - GC) * ny + (j - GC) the problem's
// interior cell; rep and tch are indexed by cbase + c.
struct BRepair {
    unsigned* cnt;                  // [0], [1]: list lengths, [2]: touched length, [3]: repairs
    unsigned long long* list[2];    // cells to repair in the current / next round
    unsigned long long* touched;    // cells to update again in the current round
    unsigned* rep;                  // == stage stamp: repaired in this stage
    unsigned* tch;                  // == round stamp: already in the touched list
};

struct BatchArgs {
    const BProb* probs;
    BStat* stat;
    BSched* sched;
    int* act_prob;             // [nact] problem index of each active problem
    int* act_start;            // [nact + 1] first strip of each active problem
    double* pool[4];           // U, U1, U2 of every problem (problem k at + probs[k].off),
                               // and with the safeguard U3, the last stage's output (the
                               // host swaps pool[0] and pool[3] after every step)
    double* flux;              // x-face fluxes of the two-sweep schedule (same layout)
    unsigned long long* defer_list;  // (problem << 40) | (i << 20) | j of deferred cells
    unsigned long long* xdefer_list; // (problem << 40) | (i << 20) | j of deferred x faces
    unsigned* faults;          // batch-wide OR of fault bits
    int nprob;
    int rows;                  // TX: rows per strip
    int max_side;              // largest R or C of any problem (ghost-fill launch shape)
    int buf_ok;                // every problem's arrays below 4 GB: buffer-addressed kernels
    int positivity;            // the positivity safeguard (two-sweep schedule only)
    int swap_pools;            // the last stage writes pool[3] and the host swaps it in
    BRepair rep;
};

struct BatchStageLaunch {
    int staging;      // STAGING_*
    int grid;         // persistent workgroups
    int block;        // threads per workgroup (multiple of 64)
};

// Launchers (fvm_batch_kernels.hip).  `stage` is 0-based; the kinds and buffers of
// each stage follow time_order (1: U -> U1 (+copy back), 2: U -> U1 -> U,
// 3: U -> U1 -> U2 -> U).
struct BatchKernels {
    void (*stage)(const BatchArgs& a, int time_order, int stage, const BatchStageLaunch& l,
                  hipStream_t s);
};
BatchKernels select_batch(int order, bool characteristic, bool mixed);
void launch_batch_control(const BatchArgs& a, int mode, hipStream_t s);
void launch_batch_copy_back(const BatchArgs& a, hipStream_t s);
void launch_batch_speeds(const BatchArgs& a, int k, hipStream_t s);
// Two-sweep schedule: wave-speed maxima of every active problem's state in pool[pi]
// (after RK1's copy back; the other orders launch it from their last stage).
void launch_batch_speeds_all(const BatchArgs& a, int pi, hipStream_t s);
void launch_batch_reset_speeds(const BatchArgs& a, int k, hipStream_t s);
// Pools swapped every step: a problem that sits out a step has its state copied from
// pool[0] to pool[3] once (k_batch_control marks it), so the swap leaves it in place.
void launch_batch_mirror(const BatchArgs& a, hipStream_t s);

}  // namespace fvm
