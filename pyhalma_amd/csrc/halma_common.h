// Shared device/host helpers for libhalma (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/halma.h"

namespace halma {

// ---- error state (thread-local message; entry points return the code) --------------
int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

constexpr int kWave = 64;

// Number of compute units of the current device (cached per device).
int device_cu_count();

// Round up to a multiple of 256 bytes (workspace carving).
inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Carves sub-buffers out of a caller-provided workspace.  When base is null the
// carver only measures (used by the *_workspace_size queries).
struct Carver {
    char* base;
    size_t used = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += align256(count * sizeof(T));
        return p;
    }
};

}  // namespace halma

#define HALMA_HIP(expr)                                                      \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess) return halma::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define HALMA_LAUNCH_CHECK() HALMA_HIP(hipGetLastError())

// ---- device helpers ----------------------------------------------------------------
namespace halma {

// Uniform (wave-invariant) loads through the scalar path: address space 4 is the
// constant address space, so a wave-uniform address compiles to a scalar load.
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "uniform_load works on whole dwords");
    typedef const __attribute__((address_space(4))) int* cptr;
    cptr q = (cptr)(const void*)p;  // C-style cast performs the address-space conversion
    T r;
    int* rw = reinterpret_cast<int*>(&r);
#pragma unroll
    for (int i = 0; i < static_cast<int>(sizeof(T) / 4); ++i) rw[i] = q[i];
    return r;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_bcast0(T v) {
    return __shfl(v, 0, kWave);
}

__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_down(v, s, kWave);
    return v;
}

__device__ __forceinline__ double wave_max_f64(double v) {
    for (int s = 32; s >= 1; s >>= 1) v = fmax(v, __shfl_down(v, s, kWave));
    return v;
}

__device__ __forceinline__ int wave_sum_i32(int v) {
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_down(v, s, kWave);
    return v;
}

// Element i of an FP32 or FP64 array (dtype HALMA_F32 or HALMA_F64) as a double.  The raw
// bits are loaded inside the branch on the type and converted after it: a conversion
// inside the branch would wait for its load at once, and the loads of several calls would
// no longer overlap.
__device__ __forceinline__ double ldf(const void* p, int dtype, int64_t i) {
    uint64_t bits;
    if (dtype == HALMA_F64)
        bits = static_cast<const uint64_t*>(p)[i];
    else
        bits = static_cast<const uint32_t*>(p)[i];
    return dtype == HALMA_F64 ? __longlong_as_double(static_cast<long long>(bits))
                              : static_cast<double>(__uint_as_float(static_cast<uint32_t>(bits)));
}

// The same with the type fixed at compile time where every array of a call has one type
// (DT 1: all FP32, 2: all FP64; 0: each array's type read at run time).
template <int DT>
__device__ __forceinline__ double ldt(const void* p, int dtype, int64_t i) {
    if constexpr (DT == 1) return static_cast<double>(static_cast<const float*>(p)[i]);
    else if constexpr (DT == 2) return static_cast<const double*>(p)[i];
    else return ldf(p, dtype, i);
}

// Error-free transformation a + b = s + e (Knuth TwoSum), FP64.  Requires no
// contraction or reassociation (the library is compiled with -ffp-contract=off).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Double-double accumulator add: (hi, lo) += x.
__device__ __forceinline__ void dd_add(double& hi, double& lo, double x) {
    double s, e;
    two_sum(hi, x, s, e);
    lo += e;
    hi = s;
}

// (hi, lo) += (h2, l2)
__device__ __forceinline__ void dd_add_dd(double& hi, double& lo, double h2, double l2) {
    double s, e;
    two_sum(hi, h2, s, e);
    e += lo + l2;
    double s2, e2;
    two_sum(s, e, s2, e2);
    hi = s2;
    lo = e2;
}

}  // namespace halma
