// POD motion: mode contraction and node positions.
//
// reference: cpp/feelmri/POD.cpp:5-38 (tensordot_modes_weights) and the
// trajectory evaluation in python/feelmri/Motion.py:240-254 (POD),
// :272-290 (PODVelocity), :303-309 (PODSum), :58-68 (RespiratoryMotion).
// Memory-bound: each output row reads its M mode values once.
#include <cstdint>

#include "fm_common.h"

namespace fm {

template <typename T>
__global__ void k_tensordot(int64_t rows, int64_t M, const T *__restrict__ modes, const T *__restrict__ w,
                            T *__restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const T *a = modes + r * M;
    T acc = 0;
    for (int64_t m = 0; m < M; ++m) acc += a[m] * w[m];
    out[r] = acc;
}

// Same contraction for M <= kTdMaxM: a workgroup stages its 256 rows (256*M
// contiguous values) through LDS with 16-byte loads, so the HBM reads are
// fully coalesced, then each lane sums its row in mode order.
constexpr int kTdMaxM = 32;

template <typename T>
__global__ __launch_bounds__(256) void k_tensordot_tiled(int64_t rows, int M, const T *__restrict__ modes,
                                                        const T *__restrict__ w, T *__restrict__ out)
{
    __shared__ __align__(16) T tile[256 * kTdMaxM];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t nrows = (rows - r0 < 256) ? rows - r0 : 256;
    const int64_t nval = nrows * M;
    const T *src = modes + r0 * M;
    constexpr int kVec = 16 / sizeof(T);
    // 16-byte loads where the tile start is 16-byte aligned, scalar otherwise
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const int64_t nvec = nval / kVec;
        for (int64_t i = threadIdx.x; i < nvec; i += 256) {
            const uint4 v = reinterpret_cast<const uint4 *>(src)[i];
            *reinterpret_cast<uint4 *>(&tile[i * kVec]) = v;
        }
        for (int64_t i = nvec * kVec + threadIdx.x; i < nval; i += 256) tile[i] = src[i];
    } else {
        for (int64_t i = threadIdx.x; i < nval; i += 256) tile[i] = src[i];
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
        const T *a = &tile[threadIdx.x * M];
        T acc = 0;
        for (int m = 0; m < M; ++m) acc += a[m] * w[m];
        out[r0 + threadIdx.x] = acc;
    }
}

__global__ void k_positions(int64_t N, const double *__restrict__ xyz, const float *__restrict__ modes, int R,
                            const double *__restrict__ coef, double *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double d = coef[R + c];
        for (int r = 0; r < R; ++r) d += (double)modes[(int64_t)(c * R + r) * N + n] * coef[r];
        out[3 * n + c] = xyz[(int64_t)c * N + n] + d;
    }
}

__global__ void k_modes_pack(int64_t P, int R, const float *__restrict__ in, float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // over P*3*R, input order
    if (i >= P * 3 * R) return;
    const int64_t p = i / (3 * R);
    const int64_t rem = i % (3 * R);  // c*R + r
    out[rem * P + p] = in[i];
}

}  // namespace fm

using namespace fm;

extern "C" int fm_pod_tensordot(int dtype, int64_t P, int64_t C, int64_t M, const void *modes,
                                const void *weights, void *out, void *stream)
{
    FM_CHECK_ARG(dtype == FM_DTYPE_F32 || dtype == FM_DTYPE_F64, "unknown dtype %d", dtype);
    FM_CHECK_ARG(P >= 0 && C >= 0 && M >= 0, "negative size");
    const int64_t rows = P * C;
    if (rows == 0) return FM_OK;
    FM_CHECK_ARG(out && (M == 0 || (modes && weights)), "NULL array");
    hipStream_t st = as_stream(stream);
    const unsigned blocks = (unsigned)cdiv(rows, 256);
    if (M > 0 && M <= kTdMaxM) {
        if (dtype == FM_DTYPE_F64)
            k_tensordot_tiled<double><<<blocks, 256, 0, st>>>(rows, (int)M, (const double *)modes,
                                                            (const double *)weights, (double *)out);
        else
            k_tensordot_tiled<float><<<blocks, 256, 0, st>>>(rows, (int)M, (const float *)modes,
                                                           (const float *)weights, (float *)out);
        FM_LAUNCH_CHECK("k_tensordot_tiled");
        return FM_OK;
    }
    if (dtype == FM_DTYPE_F64)
        k_tensordot<double><<<blocks, 256, 0, st>>>(rows, M, (const double *)modes, (const double *)weights,
                                                   (double *)out);
    else
        k_tensordot<float><<<blocks, 256, 0, st>>>(rows, M, (const float *)modes, (const float *)weights,
                                                  (float *)out);
    FM_LAUNCH_CHECK("k_tensordot");
    return FM_OK;
}

extern "C" int fm_pod_positions(int64_t n_nodes, const double *packed_xyz, const float *modes, int n_modes,
                                const double *coef, double *out, void *stream)
{
    FM_CHECK_ARG(n_nodes >= 0 && n_modes >= 0, "negative size");
    if (n_nodes == 0) return FM_OK;
    FM_CHECK_ARG(packed_xyz && coef && out && (n_modes == 0 || modes), "NULL array");
    k_positions<<<(unsigned)cdiv(n_nodes, 256), 256, 0, as_stream(stream)>>>(n_nodes, packed_xyz, modes, n_modes,
                                                                          coef, out);
    FM_LAUNCH_CHECK("k_positions");
    return FM_OK;
}

extern "C" int fm_modes_pack(int64_t P, int R, const float *modes_pcr, float *modes_soa, void *stream)
{
    FM_CHECK_ARG(P >= 0 && R >= 0, "negative size");
    if (P == 0 || R == 0) return FM_OK;
    FM_CHECK_ARG(modes_pcr && modes_soa, "NULL array");
    const int64_t total = P * 3 * R;
    k_modes_pack<<<(unsigned)cdiv(total, 256), 256, 0, as_stream(stream)>>>(P, R, modes_pcr, modes_soa);
    FM_LAUNCH_CHECK("k_modes_pack");
    return FM_OK;
}
