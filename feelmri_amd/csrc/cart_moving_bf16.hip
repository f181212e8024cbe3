// Moving nodes on a Cartesian stack: one complex GEMM per readout index on
// the matrix cores, with the operands generated inside the kernel (mixed
// form).
//
// With POD motion, positions x_n(t) follow t, and on a Cartesian stack t is a
// function of the readout index i only, so every sample of readout i sees the
// same positions x_n(t_i) (the reference re-evaluates them whenever t changes
// between consecutive samples, MRI.cpp:109-134).  For fixed i the sum over
// nodes of MRI.cpp:138-152 is the product
//
//   S[i, j, (k,v)] = sum_n Ey_i[j,n] B_i[n,(k,v)]
//   Ey_i[j,n]      = exp(-2 pi i G_j . x_n(t_i))
//   B_i[n,(k,v)]   = exp(-2 pi i H_k . x_n(t_i)) c_i[n,v]
//   c_i[n,v]       = b_v[n] exp(-t_i/T2_n) exp(i phi_n t_i) exp(-2 pi i F_i . x_n(t_i))
//
// of a (lines x nodes) and a (nodes x kz*V) complex matrix, k(i,j,k) = F_i +
// G_j + H_k being the plan's additive split.  The operands depend on i through
// the positions, so they are built per workgroup and never stored:
//   * per chunk of 64 W nodes, each thread computes one node record into LDS:
//     the position (FP64, initial position plus the POD modes weighted by the
//     coefficients of t_i) and c_i[n, 0..V) (FP64 phase with range reduction,
//     FP32 hardware sine/cosine and exp2);
//   * per 16-node step, lane l of wave w builds the A fragment it feeds to the
//     MFMA itself, in registers: row 32 (W rt + w) + l%32, nodes 8 (l/32) ..
//     +7, FP64 phase G_j . x, range reduction, FP32 sine/cosine, bf16 hi + lo;
//   * exp(-2 pi i H_k . x) of the step's nodes, one per kz of the column tile,
//     goes to LDS one step ahead; the B fragments (column 32 cb + l%32) are
//     formed from it and the records, split into bf16 hi + lo planes in LDS
//     (one buffer: a barrier after the build and one after the MFMA phase's
//     reads; the smaller footprint lets 4 workgroups share a CU);
//   * 12 v_mfma_f32_32x32x16_bf16 per (row block, column block) and step:
//     Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br, each real product as hh + hl +
//     lh, with -Ai formed once per step.
// Nodes are split across workgroups, at most ~16 384 per split: the FP32
// accumulators are the dominant error term (at 131 072 nodes per split the
// 4-D flow case is 1.9e-5 from FP64, at 16 384 4.5e-6, at 10 176 3.1e-6),
// and k_cart_reduce sums the per-split partial tiles in FP64.  The split
// count fills whole rounds of the workgroups the device holds.  The workgroups of one node
// split are dealt onto one XCD, so the split's node data is fetched from HBM
// about once and re-read from L2 by the readout indices.
#include <algorithm>

#include "cart_gemm.h"

namespace fm {

typedef __attribute__((ext_vector_type(2))) float mv_f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 mv_bf16x2;

// bf16 hi + lo planes of 8 complex values.
__device__ inline void split8(const float (&re)[8], const float (&im)[8], bf16x8 &rh, bf16x8 &rl, bf16x8 &ih,
                              bf16x8 &il)
{
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const mv_f32x2 r = {re[2 * h], re[2 * h + 1]};
        const mv_f32x2 m = {im[2 * h], im[2 * h + 1]};
        const mv_bf16x2 r_h = __builtin_convertvector(r, mv_bf16x2);
        const mv_bf16x2 m_h = __builtin_convertvector(m, mv_bf16x2);
        const mv_bf16x2 r_l = __builtin_convertvector(r - __builtin_convertvector(r_h, mv_f32x2), mv_bf16x2);
        const mv_bf16x2 m_l = __builtin_convertvector(m - __builtin_convertvector(m_h, mv_f32x2), mv_bf16x2);
        rh[2 * h] = r_h[0];
        rh[2 * h + 1] = r_h[1];
        ih[2 * h] = m_h[0];
        ih[2 * h + 1] = m_h[1];
        rl[2 * h] = r_l[0];
        rl[2 * h + 1] = r_l[1];
        il[2 * h] = m_l[0];
        il[2 * h + 1] = m_l[1];
    }
}

__device__ inline bf16x8 mv_neg(bf16x8 v)
{
    uint4 u = __builtin_bit_cast(uint4, v);
    const unsigned s = 0x80008000u;
    u.x ^= s;
    u.y ^= s;
    u.z ^= s;
    u.w ^= s;
    return __builtin_bit_cast(bf16x8, u);
}

// exp(-2 pi i u), u in cycles: FP64 range reduction, FP32 hardware sine and
// cosine (which take revolutions).
__device__ inline void cis_neg_f32(double u, float &c, float &s)
{
    const float rf = turns_f32(u);
    c = __builtin_amdgcn_cosf(rf);
    s = -__builtin_amdgcn_sinf(rf);
}

// kz rows of the exp(-2 pi i H_k . x) buffer: 16 nodes + 2 entries of pad
constexpr int kEzStride = 18;

// Up to 3 column blocks the kernel fits 168 VGPRs (3 waves per SIMD);
// 4 column blocks keep 8 accumulator tiles per wave and take 2.
// STD: a standard stack (lines along y, kz along z: the other components of
// G and H are exactly 0) whose column tiles are all full.  A line or kz phase
// is then one FP64 product, which rounds as the general three-term form
// does (fma(0, x, fma(G_y, y, 0 z)) = G_y y), and no column needs masking:
// the same bits with ~16 % fewer operand instructions per step.
template <int W, int CB, bool STD = false>
__global__ __launch_bounds__(64 * W, (CB <= 3 ? 3 : 2)) void k_cart_moving_bf16(
    const double *__restrict__ packed, int64_t N, int V, const float *__restrict__ modes, int R,
    const double *__restrict__ mcoef, const int32_t *__restrict__ ro_group, const double *__restrict__ F,
    const double *__restrict__ G, const double *__restrict__ H, const double *__restrict__ tr, int64_t M,
    int64_t L, int64_t K, int64_t n_rt, int64_t n_ct, int64_t n_splits, int64_t nodes_per_split, int ez_rows,
    float2 *__restrict__ part)
{
    constexpr int kThreads = 64 * W;
    constexpr int kChunk = kThreads;       // nodes per record chunk
    constexpr int kSteps = kChunk / 16;    // 16-node steps per chunk
    constexpr int kCols = 32 * CB;         // columns of the workgroup tile
    constexpr int kBF = (CB * 64 + kThreads - 1) / kThreads;  // B fragments per thread
    __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
    // rows padded by 2 entries (16 B): the B build reads rows of different
    // encodings / kz in one instruction, and unpadded strides (a multiple of
    // 128 B) put them on the same LDS banks
    __shared__ float2 sc[kMovMaxV][kChunk + 2];
    __shared__ uint4 sB[CB * 4 * 64];
    __shared__ double sH[3 * (kCols + 1)];
    __shared__ double scoef[kMaxModes + 3];
    extern __shared__ float2 sEz[];  // 2 buffers of ez_rows rows of kEzStride (kz row, node)

    // workgroup -> (split, readout index, row tile, column tile); the ids of
    // one split are contiguous and dealt onto one XCD
    const int64_t total = n_splits * M * n_rt * n_ct;
    const int64_t h = blockIdx.x;
    int64_t lg = h;
    if (total % kXcds == 0) lg = (h % kXcds) * (total / kXcds) + h / kXcds;
    const int64_t ct = lg % n_ct;
    int64_t rest = lg / n_ct;
    const int64_t rt = rest % n_rt;
    rest /= n_rt;
    const int64_t i = rest % M;
    const int64_t split = rest / M;

    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int64_t KV = K * V;
    const int64_t n_begin = split * nodes_per_split;
    const int64_t n_end = (n_begin + nodes_per_split < N) ? n_begin + nodes_per_split : N;

    const double ti = tr[i];
    const double Fx = F[3 * i], Fy = F[3 * i + 1], Fz = F[3 * i + 2];
    const int64_t grp = ro_group[i];
    for (int q = tid; q < R + 3; q += kThreads) scoef[q] = mcoef[grp * (R + 3) + q];

    // this lane's A row
    const int64_t row = (rt * W + wave) * 32 + (lane % 32);
    const bool row_ok = row < L;
    const double Gx = row_ok ? G[3 * row] : 0.0, Gy = row_ok ? G[3 * row + 1] : 0.0,
                 Gz = row_ok ? G[3 * row + 2] : 0.0;

    // the column tile's kz rows
    const int64_t c0 = ct * kCols;
    const int64_t kmin = c0 / V;
    const int64_t kmax = std::min<int64_t>(K - 1, (c0 + kCols - 1) / V);
    const int nk = (int)(kmax - kmin + 1);  // <= ez_rows
    for (int q = tid; q < 3 * nk; q += kThreads) sH[q] = H[3 * kmin + q];

    // this thread's B fragments: fragment f = tid + r kThreads feeds lane
    // f % 64 (= lane) of column block f / 64
    int bkk[kBF], bv[kBF];
    bool bok[kBF];
#pragma unroll
    for (int r = 0; r < kBF; ++r) {
        const int f = tid + r * kThreads;
        const int64_t c = c0 + (f / 64) * 32 + (lane % 32);
        bok[r] = f < CB * 64 && c < KV;
        bkk[r] = bok[r] ? (int)(c / V - kmin) : 0;
        bv[r] = bok[r] ? (int)(c % V) : 0;
    }

    f32x16 acc_re[CB], acc_im[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc_re[cb][q] = acc_im[cb][q] = 0.0f;

    const int ez_buf = ez_rows * kEzStride;
    int pz = 0;
    const int half = lane / 32;

    // exp(-2 pi i H_k . x) of step s's 16 nodes into buffer `buf`
    auto ez_gen = [&](int s, int buf) {
        float2 *dst = sEz + buf * ez_buf;
        for (int e = tid; e < nk * 16; e += kThreads) {
            const int kk = e / 16, m = s * 16 + e % 16;
            const double u = STD ? sH[3 * kk + 2] * sz[m]
                                 : fma(sH[3 * kk], sx[m], fma(sH[3 * kk + 1], sy[m], sH[3 * kk + 2] * sz[m]));
            float c, sn;
            cis_neg_f32(u, c, sn);
            dst[kk * kEzStride + e % 16] = make_float2(c, sn);
        }
    };

    for (int64_t base = n_begin; base < n_end; base += kChunk) {
        __syncthreads();  // previous chunk consumed; scoef and sH visible
        {
            const int64_t n = base + tid;
            double x = 0.0, y = 0.0, z = 0.0;
            float2 cv[kMovMaxV];
#pragma unroll
            for (int v = 0; v < kMovMaxV; ++v) cv[v] = make_float2(0.0f, 0.0f);
            if (n < n_end) {
                double dx = scoef[R], dy = scoef[R + 1], dz = scoef[R + 2];
                for (int r = 0; r < R; ++r) {
                    const double w = scoef[r];
                    dx += (double)modes[(int64_t)(0 * R + r) * N + n] * w;
                    dy += (double)modes[(int64_t)(1 * R + r) * N + n] * w;
                    dz += (double)modes[(int64_t)(2 * R + r) * N + n] * w;
                }
                x = packed[n] + dx;
                y = packed[N + n] + dy;
                z = packed[2 * N + n] + dz;
                const double phi = packed[3 * N + n], r2 = packed[4 * N + n];
                // c = b exp(-t/T2) exp(i phi t) exp(-2 pi i F.x)   (MRI.cpp:124-146)
                const double u = fma(Fx, x, fma(Fy, y, Fz * z)) - phi * ti * (1.0 / kTwoPi);
                float ec, es;
                cis_neg_f32(u, ec, es);
                const float d = __builtin_amdgcn_exp2f((float)(-ti * r2 * 1.4426950408889634));
                const float er = d * ec, ei = d * es;
#pragma unroll
                for (int v = 0; v < kMovMaxV; ++v) {
                    if (v < V) {
                        const float br = (float)packed[(int64_t)(5 + 2 * v) * N + n];
                        const float bi = (float)packed[(int64_t)(6 + 2 * v) * N + n];
                        cv[v] = make_float2(br * er - bi * ei, br * ei + bi * er);
                    }
                }
            }
            sx[tid] = x;
            sy[tid] = y;
            sz[tid] = z;
#pragma unroll
            for (int v = 0; v < kMovMaxV; ++v) sc[v][tid] = cv[v];
        }
        __syncthreads();
        const int nsteps = (int)std::min<int64_t>(kSteps, (n_end - base + 15) / 16);
        ez_gen(0, pz);
        __syncthreads();
        for (int s = 0; s < nsteps; ++s) {
            if (s + 1 < nsteps) ez_gen(s + 1, pz ^ 1);
            // B fragments of step s
            {
                const float2 *ez = sEz + pz * ez_buf;
#pragma unroll
                for (int r = 0; r < kBF; ++r) {
                    const int f = tid + r * kThreads;
                    if (f < CB * 64) {
                        float br_[8], bi_[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int m = 8 * half + q;
                            const float2 a = ez[bkk[r] * kEzStride + m];
                            const float2 b = sc[bv[r]][s * 16 + m];
                            br_[q] = (STD || bok[r]) ? a.x * b.x - a.y * b.y : 0.0f;
                            bi_[q] = (STD || bok[r]) ? a.x * b.y + a.y * b.x : 0.0f;
                        }
                        bf16x8 rh, rl, ih, il;
                        split8(br_, bi_, rh, rl, ih, il);
                        uint4 *dst = &sB[(f / 64) * 4 * 64 + lane];
                        dst[0 * 64] = __builtin_bit_cast(uint4, rh);
                        dst[1 * 64] = __builtin_bit_cast(uint4, rl);
                        dst[2 * 64] = __builtin_bit_cast(uint4, ih);
                        dst[3 * 64] = __builtin_bit_cast(uint4, il);
                    }
                }
            }
            // A fragment of step s (registers)
            bf16x8 arh, arl, aih, ail;
            {
                float ar_[8], ai_[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int m = s * 16 + 8 * half + q;
                    const double u = STD ? Gy * sy[m] : fma(Gx, sx[m], fma(Gy, sy[m], Gz * sz[m]));
                    cis_neg_f32(u, ar_[q], ai_[q]);
                }
                split8(ar_, ai_, arh, arl, aih, ail);
            }
            const bf16x8 naih = mv_neg(aih), nail = mv_neg(ail);
            __syncthreads();
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const uint4 *src = &sB[cb * 4 * 64 + lane];
                const bf16x8 brh = __builtin_bit_cast(bf16x8, src[0 * 64]);
                const bf16x8 brl = __builtin_bit_cast(bf16x8, src[1 * 64]);
                const bf16x8 bih = __builtin_bit_cast(bf16x8, src[2 * 64]);
                const bf16x8 bil = __builtin_bit_cast(bf16x8, src[3 * 64]);
                f32x16 re = acc_re[cb], im = acc_im[cb];
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bih, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(naih, bil, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nail, bih, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, im, 0, 0, 0);
                acc_re[cb] = re;
                acc_im[cb] = im;
            }
            pz ^= 1;
            __syncthreads();  // B planes read before the next step rewrites them
        }
    }

    // acc register q of lane l holds row 8 (q/4) + 4 (l/32) + q%4 and column
    // l%32 of its 32x32 block
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        const int64_t col = c0 + cb * 32 + (lane % 32);
        if (col >= KV) continue;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t r = (rt * W + wave) * 32 + 8 * (q / 4) + 4 * half + (q % 4);
            if (r < L) part[((split * M + i) * L + r) * KV + col] = make_float2(acc_re[cb][q], acc_im[cb][q]);
        }
    }
}

// The instantiation a shape launches: a standard stack whose column tiles
// are all full takes the STD form.
static bool moving_std(const MovingShape &s, const CartesianTables &c)
{
    return c.axis_aligned && s.KV == s.n_ct * 32 * s.CB;
}

static size_t moving_dyn_lds(const MovingShape &s) { return 2 * (size_t)s.ez_rows * kEzStride * sizeof(float2); }

template <int W, int CB>
static int moving_blocks_wc(bool std_stack, size_t dyn)
{
    return std_stack ? blocks_per_cu(k_cart_moving_bf16<W, CB, true>, 64 * W, dyn)
                     : blocks_per_cu(k_cart_moving_bf16<W, CB, false>, 64 * W, dyn);
}

template <int W>
static int moving_blocks_w(int CB, bool std_stack, size_t dyn)
{
    switch (CB) {
    case 1: return moving_blocks_wc<W, 1>(std_stack, dyn);
    case 2: return moving_blocks_wc<W, 2>(std_stack, dyn);
    case 3: return moving_blocks_wc<W, 3>(std_stack, dyn);
    default: return moving_blocks_wc<W, 4>(std_stack, dyn);
    }
}

// Workgroups of the shape's instantiation that one CU holds at once.
static int moving_blocks_per_cu(const MovingShape &s, const CartesianTables &c)
{
    const bool st = moving_std(s, c);
    const size_t dyn = moving_dyn_lds(s);
    switch (s.W) {
    case 1: return moving_blocks_w<1>(s.CB, st, dyn);
    case 2: return moving_blocks_w<2>(s.CB, st, dyn);
    case 3: return moving_blocks_w<3>(s.CB, st, dyn);
    default: return moving_blocks_w<4>(s.CB, st, dyn);
    }
}

MovingShape moving_shape(const CartesianTables &c, int64_t N, int V)
{
    MovingShape s{};
    s.M = c.n_ro;
    s.L = c.n_ln;
    s.K = c.n_kz;
    s.V = V;
    s.KV = s.K * V;
    const int64_t nrb = cdiv(s.L, 32), ncb = cdiv(s.KV, 32);
    s.W = (int)std::min<int64_t>(4, nrb);
    s.CB = (int)std::min<int64_t>(4, ncb);
    s.n_rt = cdiv(nrb, s.W);
    s.n_ct = cdiv(ncb, s.CB);
    const int64_t chunk = 64 * s.W;
    const int64_t base = s.M * s.n_rt * s.n_ct;
    const int64_t cols = 32 * (int64_t)s.CB;
    s.ez_rows = (int)std::min<int64_t>(cols, (cols - 1) / V + 2);
    // at most ~16 384 nodes per split (FP32 accumulation), four workgroups
    // per CU at least, and whole rounds of the workgroups the device holds
    const int64_t s_min =
        std::max(cdiv(4 * (int64_t)device_cu_count(), base), cdiv(std::max<int64_t>(N, 1), 16384));
    const int64_t capacity = (int64_t)moving_blocks_per_cu(s, c) * device_cu_count();
    const int64_t splits = balanced_splits(N, chunk, base, s_min, capacity);
    s.nodes_per_split = cdiv(cdiv(std::max<int64_t>(N, 1), splits), chunk) * chunk;
    s.n_splits = cdiv(std::max<int64_t>(N, 1), s.nodes_per_split);
    s.part_bytes = (size_t)s.n_splits * (size_t)s.M * (size_t)s.L * (size_t)s.KV * sizeof(float2);
    return s;
}

template <int W, int CB>
static void launch_moving_wc(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                             const float *modes, int R, const double *mcoef, float2 *part, hipStream_t st)
{
    const int64_t grid = s.n_splits * s.M * s.n_rt * s.n_ct;
    const size_t dyn = moving_dyn_lds(s);
    const bool std_stack = moving_std(s, c);
#define FM_MV(STDV)                                                                                           \
    k_cart_moving_bf16<W, CB, STDV><<<(unsigned)grid, 64 * W, dyn, st>>>(                                    \
        packed, N, (int)s.V, modes, R, mcoef, c.d_ro_group, c.d_F, c.d_G, c.d_H, c.d_t, s.M, s.L, s.K, s.n_rt, \
        s.n_ct, s.n_splits, s.nodes_per_split, s.ez_rows, part)
    if (std_stack) FM_MV(true);
    else FM_MV(false);
#undef FM_MV
}

template <int W>
static void launch_moving_w(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                            const float *modes, int R, const double *mcoef, float2 *part, hipStream_t st)
{
    switch (s.CB) {
    case 1: launch_moving_wc<W, 1>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 2: launch_moving_wc<W, 2>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 3: launch_moving_wc<W, 3>(s, c, packed, N, modes, R, mcoef, part, st); break;
    default: launch_moving_wc<W, 4>(s, c, packed, N, modes, R, mcoef, part, st); break;
    }
}

void launch_cart_moving_bf16(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                             const float *modes, int R, const double *mcoef, float2 *part, hipStream_t st)
{
    switch (s.W) {
    case 1: launch_moving_w<1>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 2: launch_moving_w<2>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 3: launch_moving_w<3>(s, c, packed, N, modes, R, mcoef, part, st); break;
    default: launch_moving_w<4>(s, c, packed, N, modes, R, mcoef, part, st); break;
    }
}

}  // namespace fm
