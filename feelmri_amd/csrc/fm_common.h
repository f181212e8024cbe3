// Shared helpers for the libfeelmri_hip kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/feelmri_hip.h"

namespace fm {

// Per-thread message behind fm_last_error().
void set_error(const char *fmt, ...);

#define FM_CHECK_ARG(cond, ...)          \
    do {                                 \
        if (!(cond)) {                   \
            ::fm::set_error(__VA_ARGS__); \
            return FM_ERR_ARG;           \
        }                                \
    } while (0)

#define FM_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t fm_e_ = (call);                                                       \
        if (fm_e_ != hipSuccess) {                                                       \
            ::fm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(fm_e_),   \
                            __FILE__, __LINE__);                                         \
            return FM_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

// Launch-status check after a <<<>>> launch.
#define FM_LAUNCH_CHECK(name)                                                           \
    do {                                                                                \
        hipError_t fm_e_ = hipGetLastError();                                           \
        if (fm_e_ != hipSuccess) {                                                      \
            ::fm::set_error("launch of %s failed: %s", name, hipGetErrorString(fm_e_)); \
            return FM_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A phase u in cycles reduced for the FP32 hardware sine and cosine, which
// take revolutions and want an argument in [0, 1]: one fract (u - floor(u),
// exact) and, from FP64, one conversion.  The result can round to 1.0f, which
// the periodic functions take as 0.
__device__ inline float turns_f32(double u) { return (float)__builtin_amdgcn_fract(u); }
__device__ inline float turns_f32(float u) { return __builtin_amdgcn_fractf(u); }

// (double)(float)pi / pi: the reference forms 2*pi from a float constant
// (cpp/feelmri/MRI.h:13), so every k.x phase carries this factor.
constexpr double kPiFOverPi = static_cast<double>(static_cast<float>(3.14159265358979323846)) /
                              3.14159265358979323846;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// Number of CUs on the current device (cached per device).
int device_cu_count();

// Workgroups of `kernel` (a __global__ function) that one CU holds at once
// with `threads` threads and `dyn` bytes of dynamic LDS; 1 when the runtime
// cannot tell.
template <typename F>
int blocks_per_cu(F kernel, int threads, size_t dyn)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, dyn) != hipSuccess || n < 1) n = 1;
    return n;
}

// Node splits of a grid of `base` workgroups per split, each split a whole
// number of `chunk`-node chunks: the count in [s_min, 2 s_min] (and at most
// s_max, 0 = one chunk per split) that minimises (rounds of `capacity`
// concurrent workgroups) x (nodes per split), the fewest on a tie.
// Workgroups of one grid take the same time, so a last round that is only
// partly filled leaves the rest of the device idle.
int64_t balanced_splits(int64_t N, int64_t chunk, int64_t base, int64_t s_min, int64_t capacity,
                        int64_t s_max = 0);

// Zero n doubles on the stream with a kernel (not a memset node, so the
// call can be captured into a HIP graph and replayed).
int zero_f64_async(double *x, int64_t n, hipStream_t st);

}  // namespace fm
