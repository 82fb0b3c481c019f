// fvm_noalloc.cpp -- rpfvm_delta_t_noalloc: the CFL reduction of rpfvm_get_delta_t on a
// caller's device state, with no allocation of any kind.
//
// rpfvm_get_delta_t reduces the wave speeds in two kernels (k_dt_partial: one pair of
// maxima per workgroup; k_dt_final: their maximum and dt) through a scratch array of
// per-workgroup maxima that the context allocates.  Here that scratch is a file-scope
// device array, `scratch`, of fixed size: at most kScratchBlocks workgroup pairs (the
// reduction never launches more: dt_blocks caps it) plus one fault word.  Every GPU the
// library runs on holds its own copy until the process ends; its device address is taken
// once per device with hipGetSymbolAddress (the array's name in host code is only a
// host-side stand-in) and kept in `g_scratch`.
//
// Calls on one GPU share that GPU's array, so two calls whose launches overlap on that
// GPU (different streams, or different host threads) race on it.  Each call overwrites
// every entry it reads before reading it, so nothing carries over from an earlier call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rpfvm.h"
#include "fvm_api_util.hpp"
#include "fvm_kernels.hpp"

using fvm::api::fail;

namespace {

constexpr int kScratchBlocks = 2048;                    // workgroup maxima (pairs)
constexpr int kScratchDoubles = 2 * kScratchBlocks + 32;  // + the fault word (and padding)
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

// Device address of the current device's scratch array (taken on its first use).
int scratch_of_current_device(double** out) {
    int dev = 0;
    hipError_t he = hipGetDevice(&dev);
    if (he != hipSuccess) return fail(RPFVM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(he));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(RPFVM_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> hold(g_lock);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !p)
            return fail(RPFVM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(he));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" int rpfvm_delta_t_noalloc(const rpfvm_config* cfg, const double* U, size_t ld,
                                     size_t plane_stride, const double* clock, double* dt,
                                     unsigned* faults, void* stream) {
    int e = fvm::api::validate(cfg);
    if (e) return e;
    if (!U || !dt) return fail(RPFVM_ERR_INVALID, "null state or dt pointer");
    fvm::Geom g{};
    g.R = cfg->nx + 2 * fvm::GC;
    g.C = cfg->ny + 2 * fvm::GC;
    if (ld < (size_t)g.C || plane_stride < ld * (size_t)g.R)
        return fail(RPFVM_ERR_INVALID, "ld must be >= ny + 8 and plane_stride >= ld * (nx + 8)");
    if (ld > (size_t)0x7fffffff) return fail(RPFVM_ERR_INVALID, "ld too large");
    g.ld = (int)ld;
    g.plane = plane_stride;
    const int nb = fvm::dt_blocks(g);
    if (nb > kScratchBlocks)  // dt_blocks caps the launch at kScratchBlocks workgroups
        return fail(RPFVM_ERR_UNSUPPORTED, "reduction of %d workgroups exceeds the %d of the scratch",
                    nb, kScratchBlocks);
    double* sp = nullptr;
    if ((e = scratch_of_current_device(&sp))) return e;
    fvm::Phys ph{};
    ph.gamma = cfg->gamma;
    ph.dx = cfg->dx;
    ph.dy = cfg->dy;
    ph.cfl = cfg->cfl;
    unsigned* fw = faults ? faults : reinterpret_cast<unsigned*>(sp + 2 * kScratchBlocks);
    fvm::launch_dt(U, g, ph, sp, dt, clock, fw, static_cast<hipStream_t>(stream));
    HIPCHK(hipGetLastError());
    return 0;
}
