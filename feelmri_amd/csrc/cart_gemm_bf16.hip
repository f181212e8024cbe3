// Mixed-precision Cartesian GEMM: bf16 hi+lo operands on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation (the factorisation is
// described in signal_gemm.hip).
//
// Workgroup tile: RB*32 readout rows x 128 columns.  Wave w owns columns
// [32w, 32w+32) of the tile and all RB row blocks, i.e. RB complex 32x32
// accumulator pairs (RB*32 AGPRs).  Per 16-node step:
//   * the E tiles of the step (RB row blocks x 4 bf16 planes, 1 KB each)
//     are copied global -> LDS with global_load_lds_dwordx4 one step ahead
//     into the other half of a double buffer (no staging registers);
//   * the A operand of the lane's column is formed in registers from the
//     Ey/Ezb tables (complex product, split into bf16 hi and lo), one step
//     ahead, after the current step's MFMAs have been issued;
//   * 12 MFMAs per row block: Re += Er*Ar - Ei*Ai, Im += Er*Ai + Ei*Ar, each
//     real product as hh + hl + lh; -Ai comes from a sign flip of the A
//     operand, and the MFMAs are issued term by term across the row blocks so
//     that consecutive MFMAs never update the same accumulator.
// One barrier per step.
#include "cart_gemm.h"

namespace fm {

struct BFrag {
    bf16x8 rh, rl, ih, il;  // Re/Im of the A operand, hi and lo parts
};

// A operand for 8 nodes of one column: product of Ey and Ezb (4 float4 =
// 8 complex each), split into bf16 hi + lo.
__device__ inline BFrag make_bfrag(const float4 (&y)[4], const float4 (&z)[4])
{
    BFrag f;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const float4 a = y[h], b = z[h];
        const float r0 = a.x * b.x - a.y * b.y, i0 = a.x * b.y + a.y * b.x;
        const float r1 = a.z * b.z - a.w * b.w, i1 = a.z * b.w + a.w * b.z;
        const __bf16 r0h = (__bf16)r0, r1h = (__bf16)r1, i0h = (__bf16)i0, i1h = (__bf16)i1;
        f.rh[2 * h] = r0h;
        f.rh[2 * h + 1] = r1h;
        f.ih[2 * h] = i0h;
        f.ih[2 * h + 1] = i1h;
        f.rl[2 * h] = (__bf16)(r0 - (float)r0h);
        f.rl[2 * h + 1] = (__bf16)(r1 - (float)r1h);
        f.il[2 * h] = (__bf16)(i0 - (float)i0h);
        f.il[2 * h + 1] = (__bf16)(i1 - (float)i1h);
    }
    return f;
}

__device__ inline bf16x8 neg_bf16x8(bf16x8 v)
{
    uint4 u = __builtin_bit_cast(uint4, v);
    const unsigned s = 0x80008000u;
    u.x ^= s;
    u.y ^= s;
    u.z ^= s;
    u.w ^= s;
    return __builtin_bit_cast(bf16x8, u);
}

// Copy one E stage (RB tiles of 4 KB) into LDS: each wave moves RB KB,
// 1 KB (64 lanes x 16 B) per instruction.
template <int RB>
__device__ inline void stage_E_to_lds(const uint4 *__restrict__ src, uint4 *dst, int wave, int lane)
{
#pragma unroll
    for (int e = 0; e < RB; ++e) {
        const int chunk = wave * RB + e;  // 1 KB chunks, RB*4 per stage
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void *)(src + chunk * 64 + lane),
            (__attribute__((address_space(3))) void *)(dst + chunk * 64), 16, 0, 0);
    }
}

template <int RB>
__global__ __launch_bounds__(kGemmThreads) void k_cart_gemm_bf16(
    const uint4 *__restrict__ E, const float2 *__restrict__ Ey, const float2 *__restrict__ Ez, int64_t Np,
    int64_t Mp, int64_t M, int64_t NC, int64_t n_kzV, int64_t n_rt, int64_t n_ct, int64_t n_splits,
    int64_t steps_per_split, int64_t ksteps, float2 *__restrict__ part)
{
    constexpr int kStage = RB * 4 * 64;  // uint4 per stage
    __shared__ uint4 sE[2][kStage];
    int64_t split, rt, ct;
    gemm_tile_ids(n_rt, n_ct, n_splits, split, rt, ct);
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int64_t nrb = Mp / 32;
    const int64_t k0 = split * steps_per_split;
    const int64_t k1 = (k0 + steps_per_split < ksteps) ? k0 + steps_per_split : ksteps;

    // this lane's output column (clamped for padding lanes; results discarded)
    const int64_t col = ct * 128 + wave * 32 + (lane % 32);
    const int64_t colc = col < NC ? col : NC - 1;
    const int64_t j = colc / n_kzV, kv = colc % n_kzV;
    // nodes 16*ks + 8*(lane/32) .. +7 of this column: 4 float4 per table
    const float4 *ey4 = reinterpret_cast<const float4 *>(Ey + j * Np) + 4 * (lane / 32);
    const float4 *ez4 = reinterpret_cast<const float4 *>(Ez + kv * Np) + 4 * (lane / 32);

    f32x16 acc_re[RB], acc_im[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc_re[r][q] = acc_im[r][q] = 0.0f;

    BFrag cur;
    if (k0 < k1) {
        stage_E_to_lds<RB>(E + (k0 * nrb + rt * RB) * 4 * 64, sE[0], wave, lane);
        float4 y[4], z[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            y[h] = ey4[k0 * 8 + h];
            z[h] = ez4[k0 * 8 + h];
        }
        cur = make_bfrag(y, z);
    }
    __builtin_amdgcn_s_waitcnt(0);  // LDS copies are tracked by vmcnt
    __syncthreads();
    int buf = 0;
    for (int64_t ks = k0; ks < k1; ++ks) {
        // Prefetch step ks+1 (the last step re-fetches itself into the idle
        // buffer): unconditional, so the loaded registers are used in place
        // and nothing waits on them before the MFMAs below.
        const int64_t kn = (ks + 1 < k1) ? ks + 1 : ks;
        stage_E_to_lds<RB>(E + (kn * nrb + rt * RB) * 4 * 64, sE[buf ^ 1], wave, lane);
        float4 ny[4], nz[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            ny[h] = ey4[kn * 8 + h];
            nz[h] = ez4[kn * 8 + h];
        }
        __builtin_amdgcn_sched_barrier(0);  // issue the prefetch before the MFMAs
        const bf16x8 nih = neg_bf16x8(cur.ih), nil = neg_bf16x8(cur.il);
        // E operand planes of every row block for this step
        bf16x8 erh[RB], erl[RB], eih[RB], eil[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const uint4 *t = &sE[buf][r * 4 * 64];
            erh[r] = __builtin_bit_cast(bf16x8, t[0 * 64 + lane]);
            erl[r] = __builtin_bit_cast(bf16x8, t[1 * 64 + lane]);
            eih[r] = __builtin_bit_cast(bf16x8, t[2 * 64 + lane]);
            eil[r] = __builtin_bit_cast(bf16x8, t[3 * 64 + lane]);
        }
#define FM_MFMA(ACC, A, B)                                                                   \
    _Pragma("unroll") for (int r = 0; r < RB; ++r) ACC[r] =                                  \
        __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[r], B, ACC[r], 0, 0, 0);
        FM_MFMA(acc_re, erh, cur.rh)
        FM_MFMA(acc_im, erh, cur.ih)
        FM_MFMA(acc_re, erh, cur.rl)
        FM_MFMA(acc_im, erh, cur.il)
        FM_MFMA(acc_re, erl, cur.rh)
        FM_MFMA(acc_im, erl, cur.ih)
        FM_MFMA(acc_re, eih, nih)
        FM_MFMA(acc_im, eih, cur.rh)
        FM_MFMA(acc_re, eih, nil)
        FM_MFMA(acc_im, eih, cur.rl)
        FM_MFMA(acc_re, eil, nih)
        FM_MFMA(acc_im, eil, cur.rh)
#undef FM_MFMA
        // keep the next step's operand build after the MFMA issue, so the
        // wait for its global loads does not stall the MFMAs
        __builtin_amdgcn_sched_barrier(0);
        cur = make_bfrag(ny, nz);
        __builtin_amdgcn_s_waitcnt(0);  // next stage landed in LDS
        __syncthreads();
        buf ^= 1;
    }

    // acc register q of lane l holds row 8*(q/4) + 4*(l/32) + q%4 and
    // column l%32 of its 32x32 block
    if (col < NC) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int64_t row = (rt * RB + r) * 32 + 8 * (q / 4) + 4 * (lane / 32) + (q % 4);
                if (row < M) part[(split * M + row) * NC + col] = make_float2(acc_re[r][q], acc_im[r][q]);
            }
        }
    }
}

template <int RB>
static void launch_bf16(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    const unsigned grid = (unsigned)(s.n_rt * s.n_ct * s.n_splits);
    k_cart_gemm_bf16<RB><<<grid, kGemmThreads, 0, st>>>(
        reinterpret_cast<const uint4 *>(ws + s.off_E), reinterpret_cast<const float2 *>(ws + s.off_Ey),
        reinterpret_cast<const float2 *>(ws + s.off_Ez), s.Np, s.Mp, s.M, s.NC, n_kzV, s.n_rt, s.n_ct, s.n_splits,
        s.steps_per_split, s.ksteps, reinterpret_cast<float2 *>(ws + s.off_part));
}

void launch_cart_gemm_bf16(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    switch (s.rb_per_wave) {
    case 1: launch_bf16<1>(s, ws, n_kzV, st); break;
    case 2: launch_bf16<2>(s, ws, n_kzV, st); break;
    case 3: launch_bf16<3>(s, ws, n_kzV, st); break;
    default: launch_bf16<4>(s, ws, n_kzV, st); break;
    }
}

}  // namespace fm
