// fvm_buffer.hpp -- buffer-descriptor addressing of the state and flux arrays.
//
// A buffer instruction forms its address as descriptor base + SGPR offset + VGPR offset.
// With one descriptor per array (4 planes), the row / plane / stencil-row part of every
// offset is wave-uniform and lives in an SGPR, and a lane contributes one fixed VGPR (its
// column), so the kernels do no per-lane 64-bit address arithmetic.  Offsets are 32-bit:
// the buffer forms are used only when an array's 4 planes fit in 4 GB (buffer_fits), and
// the global-address forms otherwise.
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_kernels.hpp"

namespace fvm {

typedef __attribute__((address_space(3))) void lds_void;
typedef int i32x4 __attribute__((ext_vector_type(4)));

// buffer_load_dword ... lds: LDS[M0 + 4 * lane] = dword at base + voffset + soffset.
__device__ void buffer_load_lds_dword(i32x4 rsrc, lds_void* lds, int size, int voffset, int soffset,
                                      int offset, int aux) __asm("llvm.amdgcn.raw.buffer.load.lds");
__device__ double buffer_load_f64(i32x4 rsrc, int voffset, int soffset,
                                  int aux) __asm("llvm.amdgcn.raw.buffer.load.f64");
__device__ void buffer_store_f64(double v, i32x4 rsrc, int voffset, int soffset,
                                 int aux) __asm("llvm.amdgcn.raw.buffer.store.f64");

// Raw descriptor (stride 0; gfx9 word 3) over `bytes` bytes from `base`: loads past
// `bytes` return zero, stores past it are dropped.
__device__ __forceinline__ i32x4 buffer_rsrc(const void* base, unsigned bytes) {
    const unsigned long long a = reinterpret_cast<unsigned long long>(base);
    i32x4 r;
    r.x = __builtin_amdgcn_readfirstlane((int)(unsigned)a);
    r.y = __builtin_amdgcn_readfirstlane((int)(unsigned)(a >> 32));
    r.z = __builtin_amdgcn_readfirstlane((int)bytes);
    r.w = 0x00020000;
    return r;
}

// Bytes of one array (4 planes); fits the 32-bit offsets of the buffer forms?
__host__ __device__ inline unsigned long long array_bytes(const Geom& g) {
    return (unsigned long long)g.plane * 4 * sizeof(double);
}
__host__ __device__ inline bool buffer_fits(const Geom& g) { return array_bytes(g) < (1ull << 32); }
// The marching, buffer-addressed kernel forms (LDS-staged sweeps, the one-pass stage)
// unless the context runs the row forms (Phys::row_forms).
inline bool use_march_forms(const Geom&, const Phys& ph) { return !ph.row_forms; }

// Stencil loads and face stores of the flux kernels for a wavefront on row i, columns
// j0 + lane: element (v, row/column offset n) of the stencil and the face (v) itself.
// FaceIO<true> addresses through descriptors (one SGPR offset per (v, n), one lane VGPR);
// FaceIO<false> is the plain global form, for arrays of 4 GB and more.
template <bool BUF>
struct FaceIO;

template <>
struct FaceIO<true> {
    i32x4 ry, rf;
    unsigned base, pl8, st8;  // bytes: (i, j0) in a plane, a plane, one stencil step
    int vo;                   // bytes: the lane's column
    __device__ __forceinline__ FaceIO(const double* Y, const double* F, const Geom& g, int i, int j0,
                                      int step) {
        const unsigned bytes = (unsigned)array_bytes(g);
        ry = buffer_rsrc(Y, bytes);
        rf = buffer_rsrc(F, bytes);
        base = ((unsigned)i * (unsigned)g.ld + (unsigned)j0) * 8u;
        pl8 = (unsigned)g.plane * 8u;
        st8 = (unsigned)step * 8u;
        vo = (int)threadIdx.x * 8;
    }
    // n >= -2 and the stencil never leaves the array, so the modular sum is the offset
    __device__ __forceinline__ double load(const double*, int v, int n, size_t, size_t) const {
        return buffer_load_f64(ry, vo, (int)(base + (unsigned)v * pl8 + (unsigned)n * st8), 0);
    }
    __device__ __forceinline__ void store(double*, int v, size_t, double x) const {
        buffer_store_f64(x, rf, vo, (int)(base + (unsigned)v * pl8), 0);
    }
};

template <>
struct FaceIO<false> {
    size_t plane;
    __device__ __forceinline__ FaceIO(const double*, const double*, const Geom& g, int, int, int)
        : plane(g.plane) {}
    __device__ __forceinline__ double load(const double* Y, int v, int n, size_t c0, size_t step) const {
        return Y[(size_t)v * plane + c0 + n * step];
    }
    __device__ __forceinline__ void store(double* F, int v, size_t c0, double x) const {
        F[(size_t)v * plane + c0] = x;
    }
};

}  // namespace fvm
