// fvm_api.cpp -- the C ABI of include/rpfvm.h on top of the gfx950 kernels.
//
// A context owns its device buffers (state, two stage arrays, two face-flux arrays,
// a residual array and reduction scratch), allocated once at rpfvm_create: the
// reference instead mallocs and zeroes six state-sized arrays per RK call
// (include/time_advance.h:791-821) and three per spatial operator
// (include/CFD_convection.h:43-45).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rpfvm.h"
#include "fvm_api_util.hpp"
#include "fvm_buffer.hpp"
#include "fvm_kernels.hpp"

using namespace fvm;

namespace fvm {
namespace api {

namespace {
thread_local char g_err[512] = "";
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char* last_error() { return g_err; }

}  // namespace api
}  // namespace fvm

using fvm::api::fail;
using fvm::api::kFrontPad;

struct rpfvm_ctx {
    rpfvm_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    Geom g{};
    Phys ph{};
    FluxPair flux{};
    BcSide bc[4]{};
    void* pool = nullptr;
    double *U = nullptr, *U1 = nullptr, *U2 = nullptr, *F = nullptr, *G = nullptr, *L = nullptr;
    double* U3 = nullptr;          // stage output when the positivity safeguard is on
    Repair rp{};                   // safeguard lists (null pointers when off)
    FaceFix fix{};                 // mixed form: faces recomputed in FP64
    unsigned stamp = 0;            // stage stamp of the safeguard
    double* partial = nullptr;
    DtSlots slots{};               // last stage's wave-speed maxima (folded CFL reduction)
    bool slots_valid = false;      // the slots describe the current state U
    double* scal = nullptr;        // [0] dt, [1] t, [2] t_stop
    StepCtl* ctl = nullptr;        // the step's control block (dt, stage clock, active)
    StepCtl* host_ctl = nullptr;   // pinned mirror of ctl (the device loop's result)
    unsigned* faults = nullptr;
    double* host_scal = nullptr;   // pinned mirror of scal
    unsigned* host_faults = nullptr;
    unsigned host_pending_faults = 0;  // fault bits found on the host (advance_to)
    double time = 0.0;                 // clock of the time-dependent boundary sides
    // kernel timing (rpfvm_set_kernel_timing)
    bool timing = false;
    struct Pending {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> spare;
    double ms[RPFVM_KCLASS_COUNT] = {};
    int counts[RPFVM_KCLASS_COUNT] = {};
};

namespace {

hipEvent_t take_event(rpfvm_ctx* x) {
    if (!x->spare.empty()) {
        hipEvent_t e = x->spare.back();
        x->spare.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the launches issued during its lifetime with a pair of events.
struct Timed {
    rpfvm_ctx* x;
    int cls;
    hipEvent_t a = nullptr;
    Timed(rpfvm_ctx* ctx, int c) : x(ctx), cls(c) {
        if (x->timing && (a = take_event(x))) (void)hipEventRecord(a, x->stream);
    }
    ~Timed() {
        if (!a) return;
        hipEvent_t b = take_event(x);
        if (!b) return;
        (void)hipEventRecord(b, x->stream);
        x->pending.push_back({cls, a, b});
    }
};

}  // namespace

// Checks one step configuration (shared with the batch entry points).
int fvm::api::validate(const rpfvm_config* c) {
    if (!c) return fail(RPFVM_ERR_INVALID, "null config");
    if (c->nx < 1 || c->ny < 1) return fail(RPFVM_ERR_INVALID, "nx, ny must be >= 1");
    if ((long)(c->nx + 2 * GC) * (c->ny + 2 * GC) > (1L << 31))
        return fail(RPFVM_ERR_INVALID, "grid too large");
    if (c->time_order < 1 || c->time_order > 3)
        return fail(RPFVM_ERR_INVALID, "time_order must be 1, 2 or 3");
    if (c->recon_order != 1 && c->recon_order != 2 && c->recon_order != 3 && c->recon_order != 5)
        return fail(RPFVM_ERR_INVALID, "recon_order must be 1, 2, 3 or 5");
    if (c->scheme == 5)
        return fail(RPFVM_ERR_UNSUPPORTED, "scheme 5 (the Marquina stub) is not implemented");
    if (c->precision != RPFVM_PRECISION_FP64 && c->precision != RPFVM_PRECISION_MIXED)
        return fail(RPFVM_ERR_INVALID, "precision must be 0 (FP64) or 1 (mixed)");
    for (int s = 0; s < 4; s++)
        if (c->bc[s] < 0 || c->bc[s] > 6) return fail(RPFVM_ERR_INVALID, "bad bc[%d]", s);
    if (c->bc[0] >= RPFVM_BC_DMR_WALL || c->bc[1] >= RPFVM_BC_DMR_WALL ||
        c->bc[2] == RPFVM_BC_DMR_SHOCK || c->bc[3] == RPFVM_BC_DMR_WALL)
        return fail(RPFVM_ERR_INVALID, "the double-Mach-reflection wall is a bottom side, the shock a top side");
    if (!(c->dx > 0) || !(c->dy > 0) || !(c->gamma > 1.0))
        return fail(RPFVM_ERR_INVALID, "dx, dy must be > 0 and gamma > 1");
    return 0;
}

extern "C" {

const char* rpfvm_last_error(void) { return fvm::api::last_error(); }

const char* rpfvm_version(void) { return "rpfvm-gfx950 0.1.0"; }

void rpfvm_config_default(rpfvm_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->nx = c->ny = 100;
    c->time_order = 3;
    c->scheme = RPFVM_SCHEME_HLLC;
    c->recon_order = 5;
    c->characteristic = 0;
    c->dx = c->dy = 0.01;
    c->gamma = 1.4;
    c->cfl = 0.5;
    c->gravity = 1.0;
    // include/Boundary_Condition.h:49 and include/initialize.h:1298-1301
    const double mach3[4] = {1.4, 1.4 * 3.0, 0.0, 1.0 / 0.4 + 0.5 * 1.4 * 3.0 * 3.0};
    for (int k = 0; k < 4; k++) {
        c->inflow_state[k] = mach3[k];
        for (int s = 0; s < 4; s++) c->fixed_state[s][k] = mach3[k];
    }
}

static bool one_pass(const rpfvm_ctx* x);

// The row forms of the sweeps where the marching forms cannot run (arrays of 4 GB and
// more: 32-bit buffer offsets) or lose: a marching wavefront walks 4-8 rows one after
// the other, so a small grid leaves most SIMDs idle.  Measured on Noh (ms per step through
// the device loop, marching / row): FP64 200^2 0.267 / 0.141, 1024^2 0.593 / 0.500,
// 2048^2 1.64 / 1.62, 3072^2 3.25 / 3.40, 4096^2 5.60 / 5.83; mixed 200^2 0.247 / 0.211,
// 1024^2 0.369 / 0.337, 1280^2 0.474 / 0.457, 1536^2 0.600 / 0.635, 2048^2 0.895 / 0.969.
static bool row_forms(const rpfvm_config& c, const Geom& g) {
    if (!buffer_fits(g) || (c.debug_flags & RPFVM_DEBUG_LARGE_ARRAY_FORMS)) return true;
    if (c.debug_flags & RPFVM_DEBUG_MARCH_FORMS) return false;
    const double cells = (double)c.nx * (double)c.ny;
    return cells < (c.precision == RPFVM_PRECISION_MIXED ? 2.0e6 : 8.0e6);
}

int rpfvm_create(const rpfvm_config* cfg, rpfvm_ctx** out) {
    if (!out) return fail(RPFVM_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int e = fvm::api::validate(cfg);
    if (e) return e;
    rpfvm_ctx* x = new (std::nothrow) rpfvm_ctx();
    if (!x) return fail(RPFVM_ERR_NOMEM, "out of host memory");
    x->cfg = *cfg;
    const rpfvm_config& c = x->cfg;
    x->g.R = c.nx + 2 * GC;
    x->g.C = c.ny + 2 * GC;
    x->g.ld = ((x->g.C + 15) / 16) * 16;
    x->g.plane = (size_t)x->g.R * x->g.ld;
    x->ph.gamma = c.gamma;
    x->ph.dx = c.dx;
    x->ph.dy = c.dy;
    x->ph.cfl = c.cfl;
    x->ph.gravity = c.gravity;
    x->ph.gm1 = c.gamma - 1.0;
    x->ph.gr = c.gamma / (c.gamma - 1.0);
    x->ph.six = 6.0;
    x->ph.igm1 = 1.0 / (c.gamma - 1.0);
    x->ph.idx = 1.0 / c.dx;
    x->ph.idy = 1.0 / c.dy;
    x->ph.source = c.source ? 1 : 0;
    x->ph.positivity = c.positivity ? 1 : 0;
    x->ph.force_fix = (c.debug_flags & RPFVM_DEBUG_MIXED_ALL_FP64) ? 1 : 0;
    x->ph.row_forms = row_forms(c, x->g) ? 1 : 0;
    x->flux = c.precision == RPFVM_PRECISION_MIXED
                  ? select_flux_mixed(c.recon_order, c.characteristic != 0, c.scheme)
                  : select_flux(c.recon_order, c.characteristic != 0, c.scheme);
    for (int s = 0; s < 4; s++) {
        BcSide& b = x->bc[s];
        b.side = s;
        b.type = c.bc[s];
        b.lo = GC;
        const bool lr = s < 2;
        const int last = lr ? x->g.C - GC : x->g.R - GC;  // outflow loops are inclusive
        b.hi = b.type == BC_OUT ? last : last - 1;
        for (int k = 0; k < 4; k++) {
            b.value[k] = b.type == BC_INFLOW ? c.inflow_state[k] : c.fixed_state[s][k];
            b.alt[k] = c.bc_alt_state[s][k];
            b.par[k] = c.bc_par[s][k];
        }
        b.nx = (double)c.nx;
        b.lx = c.lx > 0.0 ? c.lx : c.nx * c.dx;
        b.t = 0.0;
    }

    auto cleanup = [&](int code) {
        rpfvm_destroy(x);
        return code;
    };
    hipError_t he = hipGetDevice(&x->device);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(he)));
    he = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he)));

    const size_t arr = 4 * x->g.plane + 32;             // doubles per state-sized array
    const size_t arr_bytes = ((arr * sizeof(double) + 255) / 256) * 256;
    const bool separate_out = c.positivity || c.precision == RPFVM_PRECISION_MIXED;
    const int narr = separate_out ? 7 : 6;
    const int nb = dt_blocks(x->g);
    const size_t small = 256 * 4 + (size_t)nb * 2 * sizeof(double) + 256;
    const size_t ncell = (size_t)c.nx * c.ny;
    const size_t cell_bytes = ((ncell * 4 + 255) / 256) * 256;
    const size_t rep_bytes = c.positivity ? 5 * cell_bytes + 256 : 0;
    // FP64 recompute lists: every face, plus the faces two wavefronts both compute (the
    // first x-face row of each block of >= 4 rows in the one-pass stage, the strip-edge y
    // faces of the fused y sweeps)
    const size_t fix_cap = (size_t)(c.nx + 2 + (c.nx + 3) / 4) * (size_t)(c.ny + 2 + (c.ny + kStrip - 1) / kStrip);
    const size_t face_bytes = ((fix_cap * 4 + 255) / 256) * 256;
    const size_t fix_bytes = c.precision == RPFVM_PRECISION_MIXED ? 2 * face_bytes + 256 : 0;
    // slots for one row per slot, the most any form uses
    const size_t slot_bytes = (((size_t)dt_slot_count(x->g, 1) * 2 * sizeof(unsigned long long) + 255) / 256) * 256;
    const size_t ctl_bytes = 256;
    const size_t total = narr * arr_bytes + small + rep_bytes + fix_bytes + slot_bytes + ctl_bytes;
    he = hipMalloc(&x->pool, total);
    if (he != hipSuccess)
        return cleanup(fail(RPFVM_ERR_NOMEM, "hipMalloc(%zu): %s", total, hipGetErrorString(he)));
    char* p = static_cast<char*>(x->pool);
    double** arrays[7] = {&x->U, &x->U1, &x->U2, &x->F, &x->G, &x->L, &x->U3};
    for (int k = 0; k < narr; k++) *arrays[k] = reinterpret_cast<double*>(p + k * arr_bytes) + kFrontPad;
    char* q = p + narr * arr_bytes;
    x->scal = reinterpret_cast<double*>(q);
    x->faults = reinterpret_cast<unsigned*>(q + 256);
    x->partial = reinterpret_cast<double*>(q + 512);
    if (c.positivity) {
        char* r = q + small;
        x->rp.cnt = reinterpret_cast<unsigned*>(r);
        r += 256;
        x->rp.list[0] = reinterpret_cast<int*>(r);
        x->rp.list[1] = reinterpret_cast<int*>(r + cell_bytes);
        x->rp.touched = reinterpret_cast<int*>(r + 2 * cell_bytes);
        x->rp.rep = reinterpret_cast<unsigned*>(r + 3 * cell_bytes);
        x->rp.tch = reinterpret_cast<unsigned*>(r + 4 * cell_bytes);
        x->rp.nx = c.nx;
        x->rp.ny = c.ny;
    }
    if (fix_bytes) {
        char* r = q + small + rep_bytes;
        x->fix.cnt = reinterpret_cast<unsigned*>(r);
        x->fix.list[0] = reinterpret_cast<int*>(r + 256);
        x->fix.list[1] = reinterpret_cast<int*>(r + 256 + face_bytes);
        x->fix.cap = (unsigned)fix_cap;
    }
    x->slots.slot = reinterpret_cast<unsigned long long*>(q + small + rep_bytes + fix_bytes);
    x->ctl = reinterpret_cast<StepCtl*>(q + small + rep_bytes + fix_bytes + slot_bytes);
    x->ph.step = x->ctl;
    x->slots.nstrips = (c.ny + kStrip - 1) / kStrip;
    // rows per slot: those a wavefront of the marching fused y sweep walks, which reduces its
    // lanes' maxima once; one row in the row forms and in the one-pass stage (whose 16-row
    // wavefronts reduce per row: carrying the maxima across its rows spills registers)
    x->slots.rows = use_march_forms(x->g, x->ph) && !one_pass(x) ? march_rows() : 1;
    he = hipMemsetAsync(x->pool, 0, total, x->stream);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipMemset: %s", hipGetErrorString(he)));
    he = hipHostMalloc(reinterpret_cast<void**>(&x->host_scal), 256, hipHostMallocDefault);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(he)));
    x->host_faults = reinterpret_cast<unsigned*>(x->host_scal + 8);
    x->host_ctl = reinterpret_cast<StepCtl*>(x->host_scal + 16);
    he = hipStreamSynchronize(x->stream);
    if (he != hipSuccess) return cleanup(fail(RPFVM_ERR_HIP, "sync: %s", hipGetErrorString(he)));
    *out = x;
    return 0;
}

int rpfvm_destroy(rpfvm_ctx* x) {
    if (!x) return 0;
    if (x->stream) (void)hipStreamSynchronize(x->stream);
    for (auto& p : x->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : x->spare) (void)hipEventDestroy(e);
    if (x->pool) (void)hipFree(x->pool);
    if (x->host_scal) (void)hipHostFree(x->host_scal);
    if (x->stream) (void)hipStreamDestroy(x->stream);
    delete x;
    return 0;
}

// Host arrays in the reference layout [4][R][C] cross PCIe as one contiguous copy into (or
// out of) a state-sized scratch array, re-pitched to the device rows of ld doubles by a
// kernel: a pitched host copy moves row by row (32 KB rows at 4096^2) at 3.3 GB/s.  The
// scratch is U1 (L where U1 or U2 is the copy's own end): stage arrays whose contents no
// caller reads between steps.
static double* copy_stage(rpfvm_ctx* x, const double* dev) {
    return dev == x->U1 || dev == x->U2 ? x->L : x->U1;
}

static int copy_in(rpfvm_ctx* x, double* dev, const double* host) {
    const size_t C = (size_t)x->g.C, rows = 4 * (size_t)x->g.R;
    double* stage = copy_stage(x, dev);
    HIPCHK(hipMemcpyAsync(stage, host, rows * C * sizeof(double), hipMemcpyHostToDevice, x->stream));
    launch_repitch(stage, C, dev, (size_t)x->g.ld, rows, C, x->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

static int copy_out(rpfvm_ctx* x, double* host, const double* dev) {
    const size_t C = (size_t)x->g.C, rows = 4 * (size_t)x->g.R;
    double* stage = copy_stage(x, dev);
    launch_repitch(dev, (size_t)x->g.ld, stage, C, rows, C, x->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, stage, rows * C * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int rpfvm_set_state(rpfvm_ctx* x, const double* U) {
    if (!x || !U) return fail(RPFVM_ERR_INVALID, "null argument");
    x->slots_valid = false;
    return copy_in(x, x->U, U);
}

int rpfvm_get_state(rpfvm_ctx* x, double* U) {
    if (!x || !U) return fail(RPFVM_ERR_INVALID, "null argument");
    return copy_out(x, U, x->U);
}

// Ghost fill of Y at the stage clock t0 + c dt of the step's control block (read only by
// the double-Mach-reflection sides).
static void enqueue_bc(rpfvm_ctx* x, double* Y, double c) {
    Timed tm(x, RPFVM_KCLASS_BC);
    for (int s = 0; s < 4; s++) x->bc[s].c = c;
    launch_bc(Y, x->g, x->bc, x->ctl, x->stream);
}

static void enqueue_fluxes(rpfvm_ctx* x, const double* Y) {
    {
        Timed t(x, RPFVM_KCLASS_FLUX_X);
        x->flux.x(Y, x->F, x->g, x->ph, x->faults, x->fix, x->stream);
    }
    {
        Timed t(x, RPFVM_KCLASS_FLUX_Y);
        x->flux.y(Y, x->G, x->g, x->ph, x->faults, x->fix, x->stream);
    }
}

static void enqueue_update(rpfvm_ctx* x, const double* Un, const double* Y, double* out,
                           int kind, const Repair& rp) {
    Timed t(x, RPFVM_KCLASS_UPDATE);
    launch_update(Un, Y, out, x->F, x->G, x->g, x->ph, kind, rp, x->stream);
}

static void enqueue_dt(rpfvm_ctx* x, const double* clock) {
    Timed t(x, RPFVM_KCLASS_DT);
    if (x->slots_valid)
        launch_dt_slots(x->U, x->g, x->ph, x->slots, x->partial, x->scal, clock, x->faults,
                        x->stream);
    else
        launch_dt(x->U, x->g, x->ph, x->partial, x->scal, clock, x->faults, x->stream);
}

// Whether a stage writing a separate output runs as one pass: the form has it (mixed),
// the arrays take buffer offsets, and the schedule is the faster one for the
// reconstruction.  Measured at Noh 4096^2 (ms per step, one-pass / two-sweep): Godunov
// 1.63 / 2.07, TVD 2.27 / 2.37, WENO3 2.16 / 2.42, characteristic WENO5 2.88 / 2.84 (its
// two face computations need a 2-wave register budget in one kernel).  The debug flags
// force either schedule.
static bool one_pass(const rpfvm_ctx* x) {
    if (!x->flux.stage || !use_march_forms(x->g, x->ph)) return false;
    const int f = x->cfg.debug_flags;
    if (f & RPFVM_DEBUG_TWO_SWEEP) return false;
    if (f & RPFVM_DEBUG_ONE_PASS) return true;
    return !(x->cfg.recon_order == 5 && x->cfg.characteristic);
}

// One stage: ghost fill of Y, then either the one-pass stage (x faces, y faces, residual
// and stage combine in one kernel) or the x-face fluxes followed by the y faces fused with
// the residual and stage combine into out (separately when out aliases Y, which the fused
// kernel's neighbour reads forbid), then the mixed form's fix-up or the positivity
// safeguard's rounds.
// `last`: the step's last stage, which leaves the next step's wave-speed slots (fused
// path only; returns whether it did).
static bool enqueue_stage(rpfvm_ctx* x, const double* Un, double* Y, double* out, int kind,
                          double c, bool last) {
    const unsigned stamp = ++x->stamp;
    enqueue_bc(x, Y, c);
    const bool fused = out != Y;
    const DtSlots ds{last && fused ? x->slots.slot : nullptr, x->slots.nstrips, x->slots.rows};
    const bool op = fused && one_pass(x);
    if (op) {
        Timed t(x, RPFVM_KCLASS_FLUX_Y);
        x->flux.stage(Un, Y, out, x->F, x->G, x->g, x->ph, kind, x->rp, x->fix, stamp, ds, x->faults,
                      x->stream);
    } else if (fused) {
        {
            Timed t(x, RPFVM_KCLASS_FLUX_X);
            x->flux.xs(Y, x->F, x->g, x->ph, x->faults, x->fix, x->stream);
        }
        Timed t(x, RPFVM_KCLASS_FLUX_Y);
        x->flux.yu(Un, Y, out, x->F, x->G, x->g, x->ph, kind, x->rp, x->fix, stamp, ds, x->faults,
                   x->stream);
    } else {
        enqueue_fluxes(x, Y);
        enqueue_update(x, Un, Y, out, kind, x->rp);
    }
    // the mixed form's fix-up (deferred faces, their cells, the safeguard's rounds), or the
    // safeguard's rounds alone
    const FixupLaunch fixup = op ? x->flux.fixup1 : fused ? x->flux.fixup2 : nullptr;
    if (fixup) {
        Timed t(x, RPFVM_KCLASS_UPDATE);
        fixup(Un, Y, out, x->F, x->G, x->g, x->ph, kind, x->rp, x->fix, stamp, ds, x->faults,
              x->stream);
    } else if (x->rp.cnt) {
        Timed t(x, RPFVM_KCLASS_UPDATE);
        const RepairLaunch repair = x->flux.repair ? x->flux.repair : &launch_repair;
        repair(Un, Y, out, x->F, x->G, x->g, x->ph, kind, x->rp, stamp, ds, x->faults, x->stream);
    }
    return ds.slot != nullptr;
}

// The stages of one step, with dt and the stage clock from the step's control block.
static void enqueue_stages(rpfvm_ctx* x) {
    // With the positivity safeguard or the mixed form the last stage cannot update U in
    // place (a repaired or FP64-recomputed cell is recomputed from U^n): it writes U3,
    // which takes over U's ghost frame (the reference leaves U's ghosts as filled at
    // stage 1) and becomes the state.
    double* last = x->U3 ? x->U3 : x->U;
    // stage k fills its ghosts at t + c_k dt (SSP-RK3: c = 0, 1, 1/2; RK2: 0, 1)
    bool slots = false;
    switch (x->cfg.time_order) {
        case 1:
            slots = enqueue_stage(x, x->U, x->U, last, UPD_EULER, 0.0, true);
            break;
        case 2:
            enqueue_stage(x, x->U, x->U, x->U1, UPD_EULER, 0.0, false);
            slots = enqueue_stage(x, x->U, x->U1, last, UPD_RK2_2, 1.0, true);
            break;
        default:
            enqueue_stage(x, x->U, x->U, x->U1, UPD_EULER, 0.0, false);
            enqueue_stage(x, x->U, x->U1, x->U2, UPD_RK3_2, 1.0, false);
            slots = enqueue_stage(x, x->U, x->U2, last, UPD_RK3_3, 0.5, true);
            break;
    }
    x->slots_valid = slots;
    if (last != x->U) {
        launch_copy_frame(x->U, last, x->g, x->ctl, x->stream);
        x->U3 = x->U;
        x->U = last;
    }
}

static int enqueue_step(rpfvm_ctx* x, double dt) {
    launch_step_set(x->ctl, dt, x->time, x->stream);
    enqueue_stages(x);
    x->time = x->time + dt;
    HIPCHK(hipGetLastError());
    return 0;
}

int rpfvm_time_advance(rpfvm_ctx* x, double dt) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    return enqueue_step(x, dt);
}

int rpfvm_boundary_conditions(rpfvm_ctx* x) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    launch_step_set(x->ctl, 0.0, x->time, x->stream);
    enqueue_bc(x, x->U, 0.0);
    HIPCHK(hipGetLastError());
    return 0;
}

int rpfvm_set_time(rpfvm_ctx* x, double t) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    x->time = t;
    return 0;
}

int rpfvm_get_time(rpfvm_ctx* x, double* t) {
    if (!x || !t) return fail(RPFVM_ERR_INVALID, "null argument");
    *t = x->time;
    return 0;
}

int rpfvm_get_delta_t(rpfvm_ctx* x, double* dt) {
    if (!x || !dt) return fail(RPFVM_ERR_INVALID, "null argument");
    enqueue_dt(x, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(x->host_scal, x->scal, sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    *dt = x->host_scal[0];
    return 0;
}

int rpfvm_riemann(rpfvm_ctx* x, int dir, const double* Ls, const double* Rs, double* F) {
    if (!x || !Ls || !Rs || !F) return fail(RPFVM_ERR_INVALID, "null argument");
    if (dir != 1 && dir != 2) return fail(RPFVM_ERR_INVALID, "dir must be 1 (x) or 2 (y)");
    if (x->cfg.precision != RPFVM_PRECISION_FP64)
        return fail(RPFVM_ERR_UNSUPPORTED, "rpfvm_riemann evaluates the FP64 solvers: FP64 contexts only");
    // the face states travel through the stage arrays U1 / U2 (the state U is untouched)
    int e = copy_in(x, x->U1, Ls);
    if (!e) e = copy_in(x, x->U2, Rs);
    if (e) return e;
    const int solver = x->cfg.scheme;
    launch_riemann(dir, solver, x->U1, x->U2, x->F, x->g, x->ph, x->faults, x->stream);
    HIPCHK(hipGetLastError());
    if ((e = copy_out(x, F, x->F))) return e;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int rpfvm_space_operator(rpfvm_ctx* x, double* L, double* F, double* G) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    launch_step_set(x->ctl, 0.0, x->time, x->stream);
    enqueue_bc(x, x->U, 0.0);
    enqueue_fluxes(x, x->U);
    enqueue_update(x, x->U, x->U, x->L, UPD_RESIDUAL, Repair{});
    if (x->fix.cnt) launch_fix_reset(x->fix, x->stream);  // the y sweep's list is left filled
    HIPCHK(hipGetLastError());
    int e = 0;
    if (L && (e = copy_out(x, L, x->L))) return e;
    if (F && (e = copy_out(x, F, x->F))) return e;
    if (G && (e = copy_out(x, G, x->G))) return e;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

// The driver loop of src/RP_FVM.c:179-191 (next_output_time == Tmax == t_stop) on the
// device: each step is the CFL reduction deciding the step (k_dt_loop: stop at t_stop,
// clip, stop when dt does not advance the clock) followed by the step's kernels, which do
// nothing once the loop has stopped.  No dt crosses to the host: steps are enqueued in
// chunks and the host reads the device clock once per chunk.  The chunk grows from 4 as
// the remaining time allows, so at most a few steps past the end are enqueued as no-ops.
int rpfvm_advance_to(rpfvm_ctx* x, double t_stop, int max_steps, double* t, int* steps) {
    if (!x || !t) return fail(RPFVM_ERR_INVALID, "null argument");
    x->time = *t;
    int n = 0;
    if (*t < t_stop && max_steps > 0) {
        launch_loop_init(x->ctl, *t, t_stop, x->stream);
        int enq = 0, chunk = 4;
        // U and U3 trade places after each step the host enqueues; a step the device does
        // not take leaves them, so the host undoes the swaps of the steps not taken
        for (;;) {
            const int k = std::min(chunk, max_steps - enq);
            for (int q = 0; q < k; q++) {
                {
                    Timed tm(x, RPFVM_KCLASS_DT);
                    launch_dt_loop(x->U, x->g, x->ph, x->slots_valid ? x->slots : DtSlots{},
                                   x->partial, x->ctl, x->faults, x->stream);
                }
                enqueue_stages(x);
            }
            enq += k;
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(x->host_ctl, x->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, x->stream));
            HIPCHK(hipStreamSynchronize(x->stream));
            const StepCtl c = *x->host_ctl;
            n = (int)c.steps;
            if (c.stop || enq >= max_steps) break;
            // the steps left at the last step's dt, with 5% slack (more rounds if dt shrinks)
            const double left = c.dt > 0.0 ? (t_stop - c.t) / c.dt : 1.0;
            chunk = (int)std::min<double>(256.0, std::max<double>(1.0, std::ceil(1.05 * left) + 1.0));
        }
        if (x->U3 && ((enq - n) & 1)) std::swap(x->U, x->U3);
        x->time = x->host_ctl->t;
    }
    *t = x->time;
    if (steps) *steps = n;
    return 0;
}

int rpfvm_get_faults(rpfvm_ctx* x, int* faults) {
    if (!x || !faults) return fail(RPFVM_ERR_INVALID, "null argument");
    HIPCHK(hipMemcpyAsync(x->host_faults, x->faults, sizeof(unsigned), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemsetAsync(x->faults, 0, sizeof(unsigned), x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    *faults = (int)(*x->host_faults | x->host_pending_faults);
    x->host_pending_faults = 0;
    return 0;
}

int rpfvm_sync(rpfvm_ctx* x) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    int f = 0;
    int e = rpfvm_get_faults(x, &f);
    return e ? e : f;
}

int rpfvm_set_kernel_timing(rpfvm_ctx* x, int enabled) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    x->timing = enabled != 0;
    return 0;
}

int rpfvm_kernel_times(rpfvm_ctx* x, double* ms, int* counts) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    HIPCHK(hipStreamSynchronize(x->stream));
    for (auto& p : x->pending) {
        float f = 0.0f;
        HIPCHK(hipEventElapsedTime(&f, p.a, p.b));
        x->ms[p.cls] += f;
        x->counts[p.cls] += 1;
        x->spare.push_back(p.a);
        x->spare.push_back(p.b);
    }
    x->pending.clear();
    for (int k = 0; k < RPFVM_KCLASS_COUNT; k++) {
        if (ms) ms[k] = x->ms[k];
        if (counts) counts[k] = x->counts[k];
        x->ms[k] = 0.0;
        x->counts[k] = 0;
    }
    return 0;
}

int rpfvm_device_state(rpfvm_ctx* x, double** base, size_t* ld, size_t* plane_stride) {
    if (!x) return fail(RPFVM_ERR_INVALID, "null context");
    x->slots_valid = false;  // the caller may write the state: next dt from the state itself
    if (base) *base = x->U;
    if (ld) *ld = (size_t)x->g.ld;
    if (plane_stride) *plane_stride = x->g.plane;
    return 0;
}

int rpfvm_stream(rpfvm_ctx* x, void** stream) {
    if (!x || !stream) return fail(RPFVM_ERR_INVALID, "null argument");
    *stream = (void*)x->stream;
    return 0;
}

}  // extern "C"
