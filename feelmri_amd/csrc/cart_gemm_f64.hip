// FP64 Cartesian GEMM on v_mfma_f64_16x16x4_f64 (the plain FP64 form of
// the factorised signal; see signal_gemm.hip for the factorisation).
#include "cart_gemm.h"

namespace fm {

// ---------------------------------------------------------------------------
// FP64 GEMM: v_mfma_f64_16x16x4_f64
// ---------------------------------------------------------------------------
// Workgroup tile: RB*16 rows x 64 columns; wave w owns columns [16w, 16w+16).
// A 16-node step is 4 MFMA k-steps of 4 nodes; lane l supplies row/column
// l%16 and node 4*st + l/16 of k-step st.  Within a k-step the MFMAs are
// issued term by term across the RB row blocks, so an accumulator is
// updated every RB-th MFMA.  RB <= 4 keeps the kernel within 256 registers
// (two waves per SIMD).  The E tiles of two 16-node steps form one LDS stage
// (double-buffered), copied by direct global -> LDS DMA a stage ahead, so
// the workgroup synchronises once per 32 nodes and stages through no
// registers; the B operands (Ey * Ez) come from global memory one step ahead.
template <int RB>
__device__ inline void stage_E64_to_lds(const double *__restrict__ src, double *dst, int wave, int lane)
{
    // RB row-block tiles of 4 KB: RB * 4 chunks of 1 KB (64 lanes x 16 B)
#pragma unroll
    for (int e = 0; e < RB; ++e) {
        const int chunk = wave * RB + e;
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void *)(reinterpret_cast<const uint4 *>(src) + chunk * 64 + lane),
            (__attribute__((address_space(3))) void *)(reinterpret_cast<uint4 *>(dst) + chunk * 64), 16, 0, 0);
    }
}

template <int RB>
__global__ __launch_bounds__(kGemmThreads) void k_cart_gemm_f64(
    const double *__restrict__ E, const double2 *__restrict__ Ey, const double2 *__restrict__ Ez, int64_t Np,
    int64_t Mp, int64_t M, int64_t NC, int64_t n_kzV, int64_t n_rt, int64_t n_ct, int64_t n_splits,
    int64_t steps_per_split, int64_t ksteps, double2 *__restrict__ part)
{
    constexpr int kStep = RB * 2 * 4 * 64;             // doubles of one step's E tiles
    __shared__ double sE[2][2 * kStep];
    int64_t split, rt, ct;
    gemm_tile_ids(n_rt, n_ct, n_splits, split, rt, ct);
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int64_t nrb = Mp / 16;
    const int64_t k0 = split * steps_per_split;
    const int64_t k1 = (k0 + steps_per_split < ksteps) ? k0 + steps_per_split : ksteps;

    const int64_t col = ct * 64 + wave * 16 + (lane % 16);
    const int64_t colc = col < NC ? col : NC - 1;
    const int64_t j = colc / n_kzV, kv = colc % n_kzV;
    const double2 *eyrow = Ey + j * Np + lane / 16;
    const double2 *ezrow = Ez + kv * Np + lane / 16;

    f64x4 acc_re[RB], acc_im[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc_re[r][q] = acc_im[r][q] = 0.0;

    // E tiles of steps ks and ks + 1 (the second only below k1) into stage
    // buffer b
    auto stage = [&](int64_t ks, int b) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (ks + h < k1)
                stage_E64_to_lds<RB>(E + ((ks + h) * nrb + rt * RB) * 2 * 4 * 64, sE[b] + h * kStep, wave, lane);
    };

    double2 cy[4], cz[4], ny[4], nz[4];
    if (k0 < k1) {
        stage(k0, 0);
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            cy[st] = eyrow[k0 * 16 + 4 * st];
            cz[st] = ezrow[k0 * 16 + 4 * st];
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // LDS copies are tracked by vmcnt
    __syncthreads();
    int buf = 0;
    for (int64_t ks0 = k0; ks0 < k1; ks0 += 2) {
        if (ks0 + 2 < k1) stage(ks0 + 2, buf ^ 1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t ks = ks0 + h;
            if (ks >= k1) break;
            double br[4], bi[4];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                br[st] = cy[st].x * cz[st].x - cy[st].y * cz[st].y;
                bi[st] = cy[st].x * cz[st].y + cy[st].y * cz[st].x;
            }
            if (ks + 1 < k1) {
#pragma unroll
                for (int st = 0; st < 4; ++st) {
                    ny[st] = eyrow[(ks + 1) * 16 + 4 * st];
                    nz[st] = ezrow[(ks + 1) * 16 + 4 * st];
                }
            }
            const double *se = sE[buf] + h * kStep;
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                double er[RB], ei[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    er[r] = se[(r * 2 * 4 + 0 * 4 + st) * 64 + lane];
                    ei[r] = se[(r * 2 * 4 + 1 * 4 + st) * 64 + lane];
                }
#pragma unroll
                for (int r = 0; r < RB; ++r) acc_re[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(er[r], br[st], acc_re[r], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < RB; ++r) acc_im[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(er[r], bi[st], acc_im[r], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < RB; ++r) acc_re[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ei[r], bi[st], acc_re[r], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < RB; ++r) acc_im[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei[r], br[st], acc_im[r], 0, 0, 0);
            }
            if (ks + 1 < k1) {
#pragma unroll
                for (int st = 0; st < 4; ++st) {
                    cy[st] = ny[st];
                    cz[st] = nz[st];
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);  // the next stage has landed
        __syncthreads();
        buf ^= 1;
    }
    // acc register q of lane l holds row l/16 + 4q, column l%16
    // (v_mfma_f64_16x16x4_f64 result layout, checked on gfx950)
    if (col < NC) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t row = (rt * RB + r) * 16 + (lane / 16) + 4 * q;
                if (row < M) part[(split * M + row) * NC + col] = make_double2(acc_re[r][q], acc_im[r][q]);
            }
    }
}

template <int RB>
static void launch_f64(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    const unsigned grid = (unsigned)(s.n_rt * s.n_ct * s.n_splits);
    k_cart_gemm_f64<RB><<<grid, kGemmThreads, 0, st>>>(
        reinterpret_cast<const double *>(ws + s.off_E), reinterpret_cast<const double2 *>(ws + s.off_Ey),
        reinterpret_cast<const double2 *>(ws + s.off_Ez), s.Np, s.Mp, s.M, s.NC, n_kzV, s.n_rt, s.n_ct, s.n_splits,
        s.steps_per_split, s.ksteps, reinterpret_cast<double2 *>(ws + s.off_part));
}

void launch_cart_gemm_f64(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    switch (s.rb_per_wave) {
    case 1: launch_f64<1>(s, ws, n_kzV, st); break;
    case 2: launch_f64<2>(s, ws, n_kzV, st); break;
    case 3: launch_f64<3>(s, ws, n_kzV, st); break;
    case 4: launch_f64<4>(s, ws, n_kzV, st); break;
    case 5: launch_f64<5>(s, ws, n_kzV, st); break;
    case 6: launch_f64<6>(s, ws, n_kzV, st); break;
    case 7: launch_f64<7>(s, ws, n_kzV, st); break;
    default: launch_f64<8>(s, ws, n_kzV, st); break;
    }
}

}  // namespace fm
