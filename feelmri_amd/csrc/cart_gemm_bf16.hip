// Mixed-precision Cartesian GEMM: bf16 hi+lo operands on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation (the factorisation is
// described in signal_gemm.hip).
//
// Workgroup tile: RB*32 readout rows x 128 columns.  Wave w owns columns
// [32w, 32w+32) of the tile and all RB row blocks, i.e. RB complex 32x32
// accumulator pairs (RB*32 accumulator registers).  Per 16-node step, 12
// MFMAs per row block: Re += Er*Ar - Ei*Ai, Im += Er*Ai + Ei*Ar, each real
// product as hh + hl + lh; -Ai comes from a sign flip of the A operand, and
// the MFMAs are issued term by term across the row blocks so that
// consecutive MFMAs never update the same accumulator.  The E tiles of a step
// (RB row blocks x 4 bf16 planes, 1 KB each) reach LDS by direct
// global -> LDS DMA (global_load_lds_dwordx4), one step ahead.  One barrier
// per step.
//
// Two kernels differ in where the A operand (Ey * Ezb per column, split into
// bf16 hi and lo) comes from:
//   * k_cart_gemm_bf16_lds: 8 waves (256 columns) per workgroup when RB is
//     even, else 4; the distinct Ey and Ezb rows of the workgroup's
//     columns are DMA'd into LDS two steps ahead (a 3-deep ring), and the
//     A operand of step ks+1 is built from LDS in four chunks interleaved
//     with the four MFMA phases of step ks.  Used when those rows fit the
//     ring (kYZRows), e.g. 3-D stacks with up to ~64 kz x encodings.
//   * k_cart_gemm_bf16: every lane loads its column's Ey/Ezb values from
//     global memory one step ahead and builds the A operand after the
//     step's MFMAs; any shape.
#include "cart_gemm.h"

#include <type_traits>

namespace fm {

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

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

// Chunk h (nodes 2h, 2h+1 of the lane's 8) of the A operand: the complex
// products of a (Ey) and b (Ezb), split into bf16 hi and lo, converted in
// pairs (one packed convert per pair).
__device__ inline void build_chunk(BFrag &f, int h, const float4 a, const float4 b)
{
    const f32x2 re = {a.x * b.x - a.y * b.y, a.z * b.z - a.w * b.w};
    const f32x2 im = {a.x * b.y + a.y * b.x, a.z * b.w + a.w * b.z};
    const bf16x2 reh = __builtin_convertvector(re, bf16x2);
    const bf16x2 imh = __builtin_convertvector(im, bf16x2);
    const bf16x2 rel = __builtin_convertvector(re - __builtin_convertvector(reh, f32x2), bf16x2);
    const bf16x2 iml = __builtin_convertvector(im - __builtin_convertvector(imh, f32x2), bf16x2);
    f.rh[2 * h] = reh[0];
    f.rh[2 * h + 1] = reh[1];
    f.ih[2 * h] = imh[0];
    f.ih[2 * h + 1] = imh[1];
    f.rl[2 * h] = rel[0];
    f.rl[2 * h + 1] = rel[1];
    f.il[2 * h] = iml[0];
    f.il[2 * h + 1] = iml[1];
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


// Staged kernel.  Per step the workgroup's 128 columns need nJ distinct Ey
// rows and nK = min(128, kz*V) distinct Ezb rows (16 nodes, 128 B each).
constexpr int kYZRows = 72;
// Waves per workgroup of the staged kernel: 8 (256 columns, one E stage
// shared by twice the columns, halving E staging per MFMA) when the E stage's
// RB*4 chunks divide over 8 waves, else 4.
template <int RB>
constexpr int lds_waves() { return (RB % 2 == 0) ? 8 : 4; }
constexpr int kYZStride = 9;  // 16-byte pieces per staged row: 8 + 1 pad

// Copy one E stage (RB tiles of 4 KB) into LDS with WAVES waves: each wave
// moves RB*4/WAVES chunks of 1 KB (64 lanes x 16 B per instruction).
template <int RB, int WAVES>
__device__ inline void stage_E_waves(const uint4 *__restrict__ src, uint4 *dst, int wave, int lane)
{
    static_assert((RB * 4) % WAVES == 0, "E stage chunks must divide over the waves");
    constexpr int kPer = RB * 4 / WAVES;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int chunk = wave * kPer + e;
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void *)(src + chunk * 64 + lane),
            (__attribute__((address_space(3))) void *)(dst + chunk * 64), 16, 0, 0);
    }
}

// WAVES waves per workgroup, 32 columns each (the workgroup tile is
// RB*32 rows x 32*WAVES columns); the E stage is shared by all of them.
template <int RB, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 2) void k_cart_gemm_bf16_lds(
    const uint4 *__restrict__ E, const float2 *__restrict__ Ey, const float2 *__restrict__ Ez, int64_t Np,
    int64_t Mp, int64_t M, int64_t NC, int64_t n_kzV, int64_t n_rt, int64_t n_ct, int64_t n_splits,
    int64_t steps_per_split, int64_t ksteps, float2 *__restrict__ part)
{
    constexpr int kThreads = 64 * WAVES;
    constexpr int kCols = 32 * WAVES;
    constexpr int kStage = RB * 4 * 64;
    constexpr int kYZ = kYZRows * kYZStride;
    constexpr int kPieces = (kYZ + kThreads - 1) / kThreads;
    // Every ring slot is its own LDS variable and the step loop is unrolled
    // by the ring depth, so each LDS read names a different variable than
    // the DMA writes in flight and needs no wait on them.
    __shared__ uint4 sE0[kStage], sE1[kStage], sE2[kStage];
    __shared__ uint4 sY0[kYZ], sY1[kYZ], sY2[kYZ];
    int64_t split, rt, ct;
    gemm_tile_ids(n_rt, n_ct, n_splits, split, rt, ct);
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int64_t nrb = Mp / 32;
    const int64_t k0 = split * steps_per_split;
    const int64_t k1 = (k0 + steps_per_split < ksteps) ? k0 + steps_per_split : ksteps;

    const int64_t c0 = ct * kCols;
    const int64_t clast = (c0 + kCols - 1 < NC) ? c0 + kCols - 1 : NC - 1;
    const int64_t jf = c0 / n_kzV;
    const int nJ = (int)(clast / n_kzV - jf) + 1;
    const int nK = (int)(n_kzV < kCols ? n_kzV : kCols);
    const int T = kYZStride * (nJ + nK);  // LDS slots per step; nJ + nK <= kYZRows

    // staging: LDS slot s = tid + kThreads p holds piece s%9 of row s/9 (the
    // pad slot 8 repeats piece 7); rows are 144 B apart, so lanes reading
    // the same piece of consecutive rows hit distinct banks
    const float2 *stg[kPieces];
#pragma unroll
    for (int p = 0; p < kPieces; ++p) {
        const int s = tid + kThreads * p;
        const int row = s / kYZStride;
        const int prt = (s % kYZStride) < 8 ? s % kYZStride : 7;
        const float2 *base = Ey;
        if (s < T) base = row < nJ ? Ey + (jf + row) * Np : Ez + ((c0 + row - nJ) % n_kzV) * Np;
        stg[p] = base + 2 * prt;
    }

    // this lane's output column (clamped for padding lanes; results discarded)
    const int64_t col = c0 + wave * 32 + (lane % 32);
    const int64_t colc = col < NC ? col : NC - 1;
    const int ys = (int)(colc / n_kzV - jf);
    const int zs = nJ + (int)((colc - c0) % nK);
    const int hb = lane / 32;
    const int yoff = ys * kYZStride + 4 * hb, zoff = zs * kYZStride + 4 * hb;
    const uint4 *const E_rt = E + rt * RB * 4 * 64;

    f32x16 acc_re[RB], acc_im[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc_re[r][q] = acc_im[r][q] = 0.0f;

    auto stage_yz = [&](int64_t ks, uint4 *dst) {
#pragma unroll
        for (int p = 0; p < kPieces; ++p) {
            if (wave * 64 + kThreads * p < T) {
                if (tid + kThreads * p < T)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(stg[p] + ks * 16),
                        (__attribute__((address_space(3))) void *)(dst + wave * 64 + kThreads * p), 16, 0, 0);
            }
        }
    };
    auto ybuf = [&](auto b) -> uint4 * {
        if constexpr (decltype(b)::value == 0) return sY0;
        else if constexpr (decltype(b)::value == 1) return sY1;
        else return sY2;
    };
    auto ebuf = [&](auto b) -> uint4 * {
        if constexpr (decltype(b)::value == 0) return sE0;
        else if constexpr (decltype(b)::value == 1) return sE1;
        else return sE2;
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;

    BFrag cur;
    if (k0 < k1) {
        stage_E_waves<RB, WAVES>(E_rt + k0 * nrb * 4 * 64, sE0, wave, lane);
        stage_yz(k0, sY0);
        stage_yz(k0 + 1 < k1 ? k0 + 1 : k0, sY1);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        float4 y[4], z[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            y[h] = __builtin_bit_cast(float4, sY0[yoff + h]);
            z[h] = __builtin_bit_cast(float4, sY0[zoff + h]);
        }
        cur = make_bfrag(y, z);
    }

    // One 16-node step with ring slot B: MFMAs on the E tile in slot B and
    // the A operand `cur`; meanwhile DMA E(ks+1) into slot B+1 and the Ey/Ezb
    // rows of ks+2 into slot B+2, and build the A operand of ks+1 from slot
    // B+1 in four chunks interleaved with the MFMA phases.
    auto step = [&](int64_t ks, auto B) {
        constexpr int b = decltype(B)::value;
        using N1 = std::integral_constant<int, (b + 1) % 3>;
        using N2 = std::integral_constant<int, (b + 2) % 3>;
        const int64_t kn1 = (ks + 1 < k1) ? ks + 1 : ks;
        const int64_t kn2 = (ks + 2 < k1) ? ks + 2 : kn1;
        stage_E_waves<RB, WAVES>(E_rt + kn1 * nrb * 4 * 64, ebuf(N1{}), wave, lane);
        stage_yz(kn2, ybuf(N2{}));
        __builtin_amdgcn_sched_barrier(0);

        const uint4 *tE = ebuf(B) + lane;
        const uint4 *tY = ybuf(N1{});
        BFrag nxt;
        bf16x8 ea[RB], eb[RB];
        float4 ya, za, yb, zb;
        // phase 0: E plane Re-hi, chunk 0 of the next A operand
#pragma unroll
        for (int r = 0; r < RB; ++r) ea[r] = __builtin_bit_cast(bf16x8, tE[(r * 4 + 0) * 64]);
        ya = __builtin_bit_cast(float4, tY[yoff + 0]);
        za = __builtin_bit_cast(float4, tY[zoff + 0]);
        __builtin_amdgcn_sched_barrier(0);
        // phase 1: Re-hi MFMAs; load plane Re-lo and chunk 1; build chunk 0
#pragma unroll
        for (int r = 0; r < RB; ++r) eb[r] = __builtin_bit_cast(bf16x8, tE[(r * 4 + 1) * 64]);
        yb = __builtin_bit_cast(float4, tY[yoff + 1]);
        zb = __builtin_bit_cast(float4, tY[zoff + 1]);
#define FM_MFMA(ACC, A, Bv)                                                                  \
    _Pragma("unroll") for (int r = 0; r < RB; ++r) ACC[r] =                                  \
        __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[r], Bv, ACC[r], 0, 0, 0);
        FM_MFMA(acc_re, ea, cur.rh)
        FM_MFMA(acc_im, ea, cur.ih)
        FM_MFMA(acc_re, ea, cur.rl)
        FM_MFMA(acc_im, ea, cur.il)
        build_chunk(nxt, 0, ya, za);
#pragma unroll
        for (int i = 0; i < 4 * RB; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 1);
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 1);
        }
        __builtin_amdgcn_sched_barrier(0);
        // phase 2: Re-lo MFMAs; load plane Im-hi and chunk 2; build chunk 1
#pragma unroll
        for (int r = 0; r < RB; ++r) ea[r] = __builtin_bit_cast(bf16x8, tE[(r * 4 + 2) * 64]);
        ya = __builtin_bit_cast(float4, tY[yoff + 2]);
        za = __builtin_bit_cast(float4, tY[zoff + 2]);
        FM_MFMA(acc_re, eb, cur.rh)
        FM_MFMA(acc_im, eb, cur.ih)
        build_chunk(nxt, 1, yb, zb);
#pragma unroll
        for (int i = 0; i < 2 * RB; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 2);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 2);
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 2);
        }
        __builtin_amdgcn_sched_barrier(0);
        // phase 3: Im-hi MFMAs; load plane Im-lo and chunk 3; build chunk 2
        const bf16x8 nih = neg_bf16x8(cur.ih), nil = neg_bf16x8(cur.il);
#pragma unroll
        for (int r = 0; r < RB; ++r) eb[r] = __builtin_bit_cast(bf16x8, tE[(r * 4 + 3) * 64]);
        yb = __builtin_bit_cast(float4, tY[yoff + 3]);
        zb = __builtin_bit_cast(float4, tY[zoff + 3]);
        FM_MFMA(acc_re, ea, nih)
        FM_MFMA(acc_im, ea, cur.rh)
        FM_MFMA(acc_re, ea, nil)
        FM_MFMA(acc_im, ea, cur.rl)
        build_chunk(nxt, 2, ya, za);
#pragma unroll
        for (int i = 0; i < 4 * RB; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 3);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 3);
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 3);
        }
        __builtin_amdgcn_sched_barrier(0);
        // phase 4: Im-lo MFMAs; build chunk 3
        FM_MFMA(acc_re, eb, nih)
        FM_MFMA(acc_im, eb, cur.rh)
        build_chunk(nxt, 3, yb, zb);
#pragma unroll
        for (int i = 0; i < 2 * RB; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 4);
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 4);
        }
#undef FM_MFMA
        __builtin_amdgcn_sched_barrier(0);
        cur = nxt;
        __builtin_amdgcn_s_waitcnt(0);  // E(ks+1) and the rows of ks+2 landed
        __syncthreads();
    };

    int64_t ks = k0;
    for (; ks + 3 <= k1; ks += 3) {
        step(ks, I0{});
        step(ks + 1, I1{});
        step(ks + 2, I2{});
    }
    if (ks < k1) step(ks, I0{});
    if (ks + 1 < k1) step(ks + 1, I1{});

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

// Largest number of distinct Ey + Ezb rows any workgroup's `cols` columns use.
static int max_yz_rows(int64_t NC, int64_t n_kzV, int64_t cols)
{
    int64_t best = 0;
    const int64_t nK = n_kzV < cols ? n_kzV : cols;
    for (int64_t ct = 0; ct < cdiv(NC, cols); ++ct) {
        const int64_t c0 = ct * cols, cl = std::min<int64_t>(c0 + cols - 1, NC - 1);
        best = std::max<int64_t>(best, cl / n_kzV - c0 / n_kzV + 1 + nK);
    }
    return (int)best;
}

// Staged kernel with WAVES waves when the column tiles' rows fit the ring.
template <int RB, int WAVES>
static bool launch_bf16_lds(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    constexpr int kCols = 32 * WAVES;
    if (max_yz_rows(s.NC, n_kzV, kCols) > kYZRows) return false;
    const int64_t n_ct = cdiv(s.NC, kCols);
    const unsigned grid = (unsigned)(s.n_rt * n_ct * s.n_splits);
    k_cart_gemm_bf16_lds<RB, WAVES><<<grid, 64 * WAVES, 0, st>>>(
        reinterpret_cast<const uint4 *>(ws + s.off_E), reinterpret_cast<const float2 *>(ws + s.off_Ey),
        reinterpret_cast<const float2 *>(ws + s.off_Ez), s.Np, s.Mp, s.M, s.NC, n_kzV, s.n_rt, n_ct, s.n_splits,
        s.steps_per_split, s.ksteps, reinterpret_cast<float2 *>(ws + s.off_part));
    return true;
}

template <int RB>
static void launch_bf16(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    const unsigned grid = (unsigned)(s.n_rt * s.n_ct * s.n_splits);
    if (launch_bf16_lds<RB, lds_waves<RB>()>(s, ws, n_kzV, st)) return;
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
