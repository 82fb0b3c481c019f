// fvm_batch_api.cpp -- the batching form's C ABI (include/rpfvm.h, rpfvm_batch_*).
//
// Device memory of a batch, allocated once:
//   three state pools (U, U1, U2), an x-flux pool (two-sweep schedule) and, with the
//     positivity safeguard, a fourth state pool for the last stage's output: every
//     problem's four planes of R rows x ld doubles, problem k at offset probs[k].off
//     (interior column 128-byte aligned)
//   probs[count] (BProb), stat[count] (BStat), one BSched, the active-problem tables,
//   the deferred-cell and deferred-x-face lists of the mixed form (one entry per interior
//   cell / x face at most), the safeguard's repair lists and stamps
// A step is k_batch_control, then per RK stage the launches of the schedule the tuning
// selects (fvm_batch_kernels.hip, stage_launch):
//   two-sweep (the tuner's choice on the seeded workloads, the only one with the
//     safeguard): ghost fill of the stage input, persistent x sweep into the flux pool
//     (+ its FP64 fix-up in the mixed form), persistent y sweep fused with the stage
//     update (+ the mixed form's cell fix-up), the safeguard's repair rounds, and after the
//     last stage the pass that takes the next dt's wave-speed maxima;
//   one-pass: one persistent k_batch_stage launch per stage (+ k_batch_fix in the mixed
//     form), the maxima folded into the last stage;
// plus k_batch_copy_back for RK1.  It contains no memset and no host synchronisation, so
// it is captured into a hipGraph (re-captured when the tuning changes; one graph per pool
// assignment when the safeguard alternates the state between pool[0] and pool[3]) and
// replayed by rpfvm_batch_step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rpfvm.h"
#include "fvm_api_util.hpp"
#include "fvm_batch.hpp"
#include "fvm_buffer.hpp"

using namespace fvm;
using fvm::api::fail;

struct rpfvm_batch {
    int count = 0;
    int device = 0;
    int ncu = 256;
    int time_order = 3;
    bool mixed = false;
    hipStream_t stream = nullptr;
    std::vector<rpfvm_config> cfg;
    std::vector<BProb> probs;
    void* pool = nullptr;
    BatchArgs args{};
    BatchKernels kern{};
    rpfvm_batch_tuning tuning{};
    // With the safeguard the last stage writes pool[3], which becomes the state: the
    // step's launches alternate between two pool assignments (parity), each captured in
    // its own graph.
    bool swap_pools = false;
    int parity = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    BSched* host_sched = nullptr;  // pinned
    BStat* host_stat = nullptr;    // pinned
    unsigned* host_faults = nullptr;
    double t_end = -1.0;           // value last written to the device
};

namespace {

int stages_of(int time_order) { return time_order; }

BatchStageLaunch launch_shape(const rpfvm_batch* b) {
    BatchStageLaunch l;
    l.staging = b->tuning.staging;
    l.grid = std::max(1, b->ncu * b->tuning.waves_per_cu);
    l.block = 64;
    return l;
}

// The step's launches, in stream order (also what the graph captures).
void enqueue_step(rpfvm_batch* b) {
    launch_batch_control(b->args, 0, b->stream);
    if (b->swap_pools) launch_batch_mirror(b->args, b->stream);
    const BatchStageLaunch l = launch_shape(b);
    for (int s = 0; s < stages_of(b->time_order); s++)
        b->kern.stage(b->args, b->time_order, s, l, b->stream);
    if (b->time_order == 1) {
        launch_batch_copy_back(b->args, b->stream);
        if (l.staging == STAGING_SWEEPS) launch_batch_speeds_all(b->args, 0, b->stream);
    }
}

void drop_graph(rpfvm_batch* b) {
    for (int q = 0; q < 2; q++) {
        if (b->exec[q]) (void)hipGraphExecDestroy(b->exec[q]);
        if (b->graph[q]) (void)hipGraphDestroy(b->graph[q]);
        b->exec[q] = nullptr;
        b->graph[q] = nullptr;
    }
}

// Graph of the step under the current pool assignment.
int capture(rpfvm_batch* b) {
    const int q = b->parity;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(b);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(b->stream, &g));
    b->graph[q] = g;
    HIPCHK(hipGraphInstantiate(&b->exec[q], b->graph[q], nullptr, nullptr, 0));
    return 0;
}

// After a step whose last stage wrote pool[3]: that buffer is the state now.
void swap_state(rpfvm_batch* b) {
    if (!b->swap_pools) return;
    std::swap(b->args.pool[0], b->args.pool[3]);
    b->parity ^= 1;
}

int set_t_end(rpfvm_batch* b, double t_end) {
    if (t_end == b->t_end) return 0;
    HIPCHK(hipStreamSynchronize(b->stream));  // the pinned staging slot may be in flight
    b->host_sched->t_end = t_end;
    HIPCHK(hipMemcpyAsync(&b->args.sched->t_end, &b->host_sched->t_end, sizeof(double),
                          hipMemcpyHostToDevice, b->stream));
    b->t_end = t_end;
    return 0;
}

int run_step(rpfvm_batch* b) {
    if (b->tuning.use_graph) {
        if (!b->exec[b->parity]) {
            int e = capture(b);
            if (e) return e;
        }
        HIPCHK(hipGraphLaunch(b->exec[b->parity], b->stream));
    } else {
        enqueue_step(b);
        HIPCHK(hipGetLastError());
    }
    swap_state(b);
    return 0;
}

int check_tuning(const rpfvm_batch* b, const rpfvm_batch_tuning* t) {
    if (t->staging != RPFVM_STAGING_DIRECT && t->staging != RPFVM_STAGING_LDS &&
        t->staging != RPFVM_STAGING_SWEEPS)
        return fail(RPFVM_ERR_INVALID, "staging must be 0, 1 or 2");
    if (b->args.positivity && t->staging != RPFVM_STAGING_SWEEPS)
        return fail(RPFVM_ERR_UNSUPPORTED,
                    "the positivity safeguard runs in the two-sweep schedule (staging 2) only");
    if (t->rows < 4 || t->rows > 1024) return fail(RPFVM_ERR_INVALID, "rows must be in 4..1024");
    if (t->waves_per_cu < 1 || t->waves_per_cu > 32)
        return fail(RPFVM_ERR_INVALID, "waves_per_cu must be in 1..32");
    return 0;
}

}  // namespace

extern "C" {

int rpfvm_batch_destroy(rpfvm_batch* b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->pool) (void)hipFree(b->pool);
    if (b->host_sched) (void)hipHostFree(b->host_sched);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int rpfvm_batch_create(int count, const rpfvm_config* cfgs, rpfvm_batch** out) {
    if (!out) return fail(RPFVM_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (count < 1 || !cfgs) return fail(RPFVM_ERR_INVALID, "count must be >= 1 with configs");
    for (int k = 0; k < count; k++) {
        const rpfvm_config& c = cfgs[k];
        int e = fvm::api::validate(&c);
        if (e) {
            char msg[512];
            std::snprintf(msg, sizeof msg, "%s", rpfvm_last_error());
            return fail(e, "problem %d: %s", k, msg);
        }
        for (int s = 0; s < 4; s++)
            if (c.bc[s] > RPFVM_BC_FIXED_VALUE)
                return fail(RPFVM_ERR_UNSUPPORTED, "problem %d: time-dependent sides are not batched", k);
        if (c.nx < 2 * GC || c.ny < 2 * GC || c.nx >= (1 << 20) - 2 * GC || c.ny >= (1 << 20) - 2 * GC)
            return fail(RPFVM_ERR_INVALID, "problem %d: nx, ny must be in 8 .. 2^20 - 9", k);
        if ((c.characteristic != 0) != (cfgs[0].characteristic != 0) ||
            c.recon_order != cfgs[0].recon_order || c.time_order != cfgs[0].time_order ||
            c.precision != cfgs[0].precision || (c.positivity != 0) != (cfgs[0].positivity != 0))
            return fail(RPFVM_ERR_INVALID,
                        "problem %d: recon_order, characteristic, time_order, precision and "
                        "positivity must match problem 0", k);
    }
    if (count >= (1 << 23)) return fail(RPFVM_ERR_INVALID, "too many problems");
    rpfvm_batch* b = new (std::nothrow) rpfvm_batch();
    if (!b) return fail(RPFVM_ERR_NOMEM, "out of host memory");
    auto cleanup = [&](int code) {
        rpfvm_batch_destroy(b);
        return code;
    };
    b->count = count;
    b->cfg.assign(cfgs, cfgs + count);
    b->time_order = cfgs[0].time_order;
    b->mixed = cfgs[0].precision == RPFVM_PRECISION_MIXED;
    b->kern = select_batch(cfgs[0].recon_order, cfgs[0].characteristic != 0, b->mixed);
    const bool pos = cfgs[0].positivity != 0;
    b->args.positivity = pos ? 1 : 0;
    b->swap_pools = pos && b->time_order > 1;
    b->args.swap_pools = b->swap_pools ? 1 : 0;
    b->tuning = {pos ? RPFVM_STAGING_SWEEPS : RPFVM_STAGING_DIRECT, 32, 8, 1, 0.0};

    // problem layout in the pools
    size_t off = 0, cells = 0;
    bool large_forms = false;  // some problem needs the global-address kernels
    b->probs.resize(count);
    for (int k = 0; k < count; k++) {
        const rpfvm_config& c = cfgs[k];
        BProb& d = b->probs[k];
        std::memset(&d, 0, sizeof d);
        d.g.R = c.nx + 2 * GC;
        d.g.C = c.ny + 2 * GC;
        d.g.ld = ((d.g.C + 15) / 16) * 16;
        d.g.plane = (size_t)d.g.R * d.g.ld;
        d.ph.gamma = c.gamma;
        d.ph.dx = c.dx;
        d.ph.dy = c.dy;
        d.ph.cfl = c.cfl;
        d.ph.gravity = c.gravity;
        d.ph.gm1 = c.gamma - 1.0;
        d.ph.gr = c.gamma / (c.gamma - 1.0);
        d.ph.six = 6.0;
        d.ph.igm1 = 1.0 / (c.gamma - 1.0);
        d.ph.idx = 1.0 / c.dx;
        d.ph.idy = 1.0 / c.dy;
        d.ph.source = c.source ? 1 : 0;
        d.ph.positivity = c.positivity ? 1 : 0;
        d.ph.force_fix = (c.debug_flags & RPFVM_DEBUG_MIXED_ALL_FP64) ? 1 : 0;
        d.nx = c.nx;
        d.ny = c.ny;
        d.solver = (c.scheme == RPFVM_SCHEME_HLL || c.scheme == RPFVM_SCHEME_HLLC ||
                    c.scheme == RPFVM_SCHEME_ROE || c.scheme == RPFVM_SCHEME_HLLHC ||
                    c.scheme == RPFVM_SCHEME_HLLCHC || c.scheme == RPFVM_SCHEME_ROEHC)
                       ? c.scheme
                       : RPFVM_SCHEME_LLF;
        for (int s = 0; s < 4; s++) {
            d.bc[s] = c.bc[s];
            for (int v = 0; v < 4; v++)
                d.bcv[s][v] = c.bc[s] == RPFVM_BC_INFLOW ? c.inflow_state[v] : c.fixed_state[s][v];
        }
        d.strips_y = (c.ny + kStripCols - 1) / kStripCols;
        for (int s = 0; s < 4; s++) {  // the single path's sides (fvm_api.cpp, rpfvm_create)
            BcSide& bs = d.side[s];
            bs.side = s;
            bs.type = c.bc[s];
            bs.lo = GC;
            const int last = s < 2 ? d.g.C - GC : d.g.R - GC;
            bs.hi = bs.type == RPFVM_BC_OUTFLOW ? last : last - 1;
            for (int v = 0; v < 4; v++) bs.value[v] = d.bcv[s][v];
            bs.nx = (double)c.nx;
            bs.lx = c.lx > 0.0 ? c.lx : c.nx * c.dx;
        }
        b->args.max_side = std::max(b->args.max_side, std::max(d.g.R, d.g.C));
        if (!buffer_fits(d.g) || (c.debug_flags & RPFVM_DEBUG_LARGE_ARRAY_FORMS)) large_forms = true;
        off = (off + 31) / 32 * 32;  // 256-byte aligned region
        d.off = off + fvm::api::kFrontPad;
        off = d.off + 4 * d.g.plane + 32;
        d.cbase = cells;
        cells += (size_t)c.nx * c.ny;
    }
    b->args.buf_ok = large_forms ? 0 : 1;
    const size_t pool_doubles = (off + 31) / 32 * 32;
    auto a256 = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t probs_bytes = a256(sizeof(BProb) * count);
    const size_t stat_bytes = a256(sizeof(BStat) * count);
    const size_t sched_bytes = a256(sizeof(BSched));
    const size_t act_bytes = a256(sizeof(int) * (2 * (size_t)count + 1));
    const size_t defer_bytes = b->mixed ? a256(sizeof(unsigned long long) * cells) : 0;
    size_t xfaces = 0;  // x faces of every problem: (nx + 1) ny
    for (int k = 0; k < count; k++) xfaces += (size_t)(cfgs[k].nx + 1) * cfgs[k].ny;
    const size_t xdefer_bytes = b->mixed ? a256(sizeof(unsigned long long) * xfaces) : 0;
    // pools: U, U1, U2, the two-sweep schedule's x-flux buffer and, with the safeguard,
    // U3; the safeguard's lists (3 x 8 B per cell) and stamps (2 x 4 B per cell)
    const int npools = pos ? 5 : 4;
    const size_t rep_bytes = pos ? 256 + 3 * a256(sizeof(unsigned long long) * cells) +
                                       2 * a256(sizeof(unsigned) * cells)
                                 : 0;
    const size_t total = npools * pool_doubles * sizeof(double) + probs_bytes + stat_bytes +
                         sched_bytes + act_bytes + 256 + defer_bytes + xdefer_bytes + rep_bytes;

    hipError_t he = hipGetDevice(&b->device);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(he)));
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess &&
        ncu > 0)
        b->ncu = ncu;
    he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he)));
    he = hipMalloc(&b->pool, total);
    if (he != hipSuccess)
        return cleanup(fail(RPFVM_ERR_NOMEM, "hipMalloc(%zu): %s", total, hipGetErrorString(he)));
    char* p = static_cast<char*>(b->pool);
    for (int q = 0; q < 3; q++) b->args.pool[q] = reinterpret_cast<double*>(p) + q * pool_doubles;
    b->args.flux = reinterpret_cast<double*>(p) + 3 * pool_doubles;
    b->args.pool[3] = pos ? reinterpret_cast<double*>(p) + 4 * pool_doubles : nullptr;
    char* r = p + npools * pool_doubles * sizeof(double);
    b->args.probs = reinterpret_cast<const BProb*>(r);
    BProb* dprobs = reinterpret_cast<BProb*>(r);
    r += probs_bytes;
    b->args.stat = reinterpret_cast<BStat*>(r);
    r += stat_bytes;
    b->args.sched = reinterpret_cast<BSched*>(r);
    r += sched_bytes;
    b->args.act_prob = reinterpret_cast<int*>(r);
    b->args.act_start = b->args.act_prob + count;
    r += act_bytes;
    b->args.faults = reinterpret_cast<unsigned*>(r);
    r += 256;
    b->args.defer_list = b->mixed ? reinterpret_cast<unsigned long long*>(r) : nullptr;
    r += defer_bytes;
    b->args.xdefer_list = b->mixed ? reinterpret_cast<unsigned long long*>(r) : nullptr;
    r += xdefer_bytes;
    if (pos) {
        BRepair& rp = b->args.rep;
        rp.cnt = reinterpret_cast<unsigned*>(r);
        r += 256;
        for (int q = 0; q < 2; q++) {
            rp.list[q] = reinterpret_cast<unsigned long long*>(r);
            r += a256(sizeof(unsigned long long) * cells);
        }
        rp.touched = reinterpret_cast<unsigned long long*>(r);
        r += a256(sizeof(unsigned long long) * cells);
        rp.rep = reinterpret_cast<unsigned*>(r);
        r += a256(sizeof(unsigned) * cells);
        rp.tch = reinterpret_cast<unsigned*>(r);
    }
    b->args.nprob = count;
    b->args.rows = b->tuning.rows;

    he = hipMemsetAsync(b->pool, 0, total, b->stream);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipMemset: %s", hipGetErrorString(he)));
    he = hipMemcpyAsync(dprobs, b->probs.data(), sizeof(BProb) * count, hipMemcpyHostToDevice, b->stream);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(he)));
    const size_t host_bytes = a256(sizeof(BSched)) + a256(sizeof(BStat) * count) + 256;
    he = hipHostMalloc(reinterpret_cast<void**>(&b->host_sched), host_bytes, hipHostMallocDefault);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(he)));
    b->host_stat = reinterpret_cast<BStat*>(reinterpret_cast<char*>(b->host_sched) + a256(sizeof(BSched)));
    b->host_faults = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(b->host_stat) +
                                                 a256(sizeof(BStat) * count));
    he = hipStreamSynchronize(b->stream);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "sync: %s", hipGetErrorString(he)));
    *out = b;
    return 0;
}

int rpfvm_batch_set_state(rpfvm_batch* b, int k, const double* U) {
    if (!b || !U) return fail(RPFVM_ERR_INVALID, "null argument");
    if (k < 0 || k >= b->count) return fail(RPFVM_ERR_INVALID, "problem %d out of range", k);
    const BProb& d = b->probs[k];
    HIPCHK(hipMemcpy2DAsync(b->args.pool[0] + d.off, d.g.ld * sizeof(double), U,
                            d.g.C * sizeof(double), d.g.C * sizeof(double), (size_t)4 * d.g.R,
                            hipMemcpyHostToDevice, b->stream));
    launch_batch_reset_speeds(b->args, k, b->stream);
    launch_batch_speeds(b->args, k, b->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int rpfvm_batch_get_state(rpfvm_batch* b, int k, double* U) {
    if (!b || !U) return fail(RPFVM_ERR_INVALID, "null argument");
    if (k < 0 || k >= b->count) return fail(RPFVM_ERR_INVALID, "problem %d out of range", k);
    const BProb& d = b->probs[k];
    // interior rows only: the ghost layers of U are left as they are
    const size_t row0 = (size_t)GC * d.g.ld + GC;
    for (int v = 0; v < 4; v++) {
        HIPCHK(hipMemcpy2DAsync(U + (size_t)v * d.g.R * d.g.C + (size_t)GC * d.g.C + GC,
                                d.g.C * sizeof(double),
                                b->args.pool[0] + d.off + (size_t)v * d.g.plane + row0,
                                d.g.ld * sizeof(double), (size_t)d.ny * sizeof(double), (size_t)d.nx,
                                hipMemcpyDeviceToHost, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int rpfvm_batch_step(rpfvm_batch* b, double t_end) {
    if (!b) return fail(RPFVM_ERR_INVALID, "null batch");
    int e = set_t_end(b, t_end);
    return e ? e : run_step(b);
}

int rpfvm_batch_advance_to(rpfvm_batch* b, double t_end, int max_steps, int* steps) {
    if (!b) return fail(RPFVM_ERR_INVALID, "null batch");
    int e = set_t_end(b, t_end);
    if (e) return e;
    int n = 0;
    const int chunk = 16;  // steps launched between checks of the active count
    while (n < max_steps) {
        const int m = std::min(chunk, max_steps - n);
        for (int q = 0; q < m; q++)
            if ((e = run_step(b))) return e;
        n += m;
        HIPCHK(hipMemcpyAsync(&b->host_sched->active_left, &b->args.sched->active_left, sizeof(int),
                              hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        if (b->host_sched->active_left == 0) break;
    }
    if (steps) *steps = n;
    return 0;
}

int rpfvm_batch_status(rpfvm_batch* b, double* t, int* steps, int* faults) {
    if (!b) return fail(RPFVM_ERR_INVALID, "null batch");
    HIPCHK(hipMemcpyAsync(b->host_stat, b->args.stat, sizeof(BStat) * b->count, hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int k = 0; k < b->count; k++) {
        if (t) t[k] = b->host_stat[k].t;
        if (steps) steps[k] = b->host_stat[k].steps;
        if (faults) faults[k] = (int)b->host_stat[k].faults;
    }
    return 0;
}

int rpfvm_batch_sync(rpfvm_batch* b) {
    if (!b) return fail(RPFVM_ERR_INVALID, "null batch");
    HIPCHK(hipMemcpyAsync(b->host_faults, b->args.faults, sizeof(unsigned), hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipMemsetAsync(b->args.faults, 0, sizeof(unsigned), b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return (int)*b->host_faults;
}

int rpfvm_batch_get_tuning(rpfvm_batch* b, rpfvm_batch_tuning* t) {
    if (!b || !t) return fail(RPFVM_ERR_INVALID, "null argument");
    *t = b->tuning;
    return 0;
}

int rpfvm_batch_set_tuning(rpfvm_batch* b, const rpfvm_batch_tuning* t) {
    if (!b || !t) return fail(RPFVM_ERR_INVALID, "null argument");
    int e = check_tuning(b, t);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tuning = *t;
    b->args.rows = t->rows;
    drop_graph(b);  // the launch shape is baked into the graph
    return 0;
}

// Autotuner: times stage 1 (U -> U1, which leaves U, the clocks and the dt maxima
// alone) for every candidate shape on the current state and keeps the fastest.
int rpfvm_batch_tune(rpfvm_batch* b, rpfvm_batch_tuning* chosen) {
    if (!b) return fail(RPFVM_ERR_INVALID, "null batch");
    const int rows_c[] = {8, 16, 32, 64, 128};
    const int waves_c[] = {4, 8, 12, 16};
    const int all_stagings[] = {RPFVM_STAGING_DIRECT, RPFVM_STAGING_LDS, RPFVM_STAGING_SWEEPS};
    const int sweeps_only[] = {RPFVM_STAGING_SWEEPS};
    const bool pos = b->args.positivity != 0;  // the safeguard runs in the two-sweep schedule
    const int* const stagings = pos ? sweeps_only : all_stagings;
    const int nstagings = pos ? 1 : 3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const rpfvm_batch_tuning before = b->tuning;
    rpfvm_batch_tuning best = b->tuning;
    double best_ms = 1e300;
    int err = 0;
    for (int si = 0; si < nstagings; si++) {
        const int st = stagings[si];
        for (int rows : rows_c) {
            for (int w : waves_c) {
                rpfvm_batch_tuning t{st, rows, w, b->tuning.use_graph, 0.0};
                b->tuning = t;
                b->args.rows = rows;
                const BatchStageLaunch l = launch_shape(b);
                float best_t = 1e30f;
                for (int rep = 0; rep < 3 && !err; rep++) {
                    launch_batch_control(b->args, 1, b->stream);
                    if (hipEventRecord(e0, b->stream) != hipSuccess) err = 1;
                    b->kern.stage(b->args, 3, 0, l, b->stream);
                    if (hipEventRecord(e1, b->stream) != hipSuccess) err = 1;
                    float ms = 0.0f;
                    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                        err = 1;
                    if (rep > 0) best_t = std::min(best_t, ms);  // first run warms the caches
                }
                if (err) break;
                if (best_t < best_ms) {
                    best_ms = best_t;
                    best = t;
                    best.ms_per_stage = best_t;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err) {
        // back to the shape the batch had before the sweep, and no graph of a trial shape
        b->tuning = before;
        b->args.rows = before.rows;
        drop_graph(b);
        return fail(RPFVM_ERR_HIP, "tuning launch failed: %s", hipGetErrorString(hipGetLastError()));
    }
    b->tuning = best;
    b->args.rows = best.rows;
    drop_graph(b);
    if (chosen) *chosen = best;
    return 0;
}

int rpfvm_batch_stream(rpfvm_batch* b, void** stream) {
    if (!b || !stream) return fail(RPFVM_ERR_INVALID, "null argument");
    *stream = b->stream;
    return 0;
}

}  // extern "C"
