// Moving nodes on a Cartesian stack, FP64 form: the per-readout product of
// cart_moving_bf16.hip on v_mfma_f64_16x16x4_f64, operands generated in the
// kernel in FP64.
//
//   S[i, j, (k,v)] = sum_n Ey_i[j,n] B_i[n,(k,v)]          (per readout index i)
//   Ey_i[j,n] = exp(-2 pi i G_j . x_n(t_i)),  B_i[n,(k,v)] = exp(-2 pi i H_k . x_n(t_i)) c_i[n,v]
//
// Layout (v_mfma_f64_16x16x4_f64, 4 nodes per MFMA): lane l supplies row /
// column l%16 and node l/16 of the k-step; accumulator register q of lane l
// holds row l/16 + 4q, column l%16.
//   * per chunk of 64 W nodes each thread forms one node record in LDS: the
//     position x0 + sum_r modes_r w_r(t_i) + translation (FP64) and
//     c_i[n, 0..V) = b_v exp(-t_i/T2) exp(i phi t_i) exp(-2 pi i F_i . x)
//     (FP64 phase, range reduction, sincospi, exp);
//   * per 16-node step (4 k-steps), wave w builds its A fragments (16 lines
//     of row block W rt + w) in registers, one FP64 sincospi per (line, node);
//     the kz factors exp(-2 pi i H_k . x) of the step's nodes are in LDS
//     (one buffer, written during the previous step's products) and the
//     workgroup forms the B fragments (16 (k,v) columns per column block,
//     FP64 products) in one LDS buffer.  The kz factors of the next step
//     are formed between the step's barrier and its MFMAs;
//   * per k-step and column block, 4 MFMAs: Re += Ar Br - Ai Bi,
//     Im += Ar Bi + Ai Br.
// FP64 accumulators need no bound on the split length; node splits exist for
// parallelism only and are summed in FP64 by k_cart_reduce.
#include <algorithm>
#include <cmath>

#include "cart_gemm.h"

namespace fm {

// exp(-2 pi i u), u in cycles: exact range reduction, FP64 sincospi.
__device__ inline void cis_neg_f64(double u, double &c, double &s)
{
    const double r = u - __builtin_rint(u);
    double sn, cs;
    sincospi(2.0 * r, &sn, &cs);
    c = cs;
    s = -sn;
}

// kz rows of the kz-factor buffer: 16 nodes + 1 entry of pad
constexpr int kEz64Stride = 17;

template <int W, int CB>
__global__ __launch_bounds__(64 * W, 2) void k_cart_moving_f64(
    const double *__restrict__ packed, int64_t N, int V, const float *__restrict__ modes, int R,
    const double *__restrict__ mcoef, const int32_t *__restrict__ ro_group, const double *__restrict__ F,
    const double *__restrict__ G, const double *__restrict__ H, const double *__restrict__ tr, int64_t M,
    int64_t L, int64_t K, int64_t n_rt, int64_t n_ct, int64_t n_splits, int64_t nodes_per_split, int ez_rows,
    double2 *__restrict__ part)
{
    constexpr int kThreads = 64 * W;
    constexpr int kChunk = kThreads;             // nodes per record chunk
    constexpr int kSteps = kChunk / 16;          // 16-node steps per chunk
    constexpr int kCols = 16 * CB;               // columns of the workgroup tile
    constexpr int kFrag = CB * 4;                // B fragments per step: (column block, k-step)
    constexpr int kBF = (kFrag * 64 + kThreads - 1) / kThreads;
    __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
    __shared__ double2 sc[kMovMaxV][kChunk + 1];
    __shared__ double sB[kFrag * 2 * 64];        // (fragment, Re/Im plane, lane)
    __shared__ double sH[3 * (kCols + 1)];
    __shared__ double scoef[kMaxModes + 3];
    extern __shared__ double2 sEz64[];           // ez_rows rows of kEz64Stride

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
    const int64_t row = (rt * W + wave) * 16 + (lane % 16);
    const bool row_ok = row < L;
    const double Gx = row_ok ? G[3 * row] : 0.0, Gy = row_ok ? G[3 * row + 1] : 0.0,
                 Gz = row_ok ? G[3 * row + 2] : 0.0;

    // the column tile's kz rows
    const int64_t c0 = ct * kCols;
    const int64_t kmin = c0 / V;
    const int64_t kmax = std::min<int64_t>(K - 1, (c0 + kCols - 1) / V);
    const int nk = (int)(kmax - kmin + 1);  // <= ez_rows
    for (int q = tid; q < 3 * nk; q += kThreads) sH[q] = H[3 * kmin + q];

    // this thread's B fragments: fragment f = (tid + r kThreads) / 64 is
    // (column block f / 4, k-step f % 4) and feeds lane tid % 64 (= lane)
    int bkk[kBF], bv[kBF];
    bool bok[kBF];
#pragma unroll
    for (int r = 0; r < kBF; ++r) {
        const int f = (tid + r * kThreads) / 64;
        const int64_t c = c0 + (f / 4) * 16 + (lane % 16);
        bok[r] = f < kFrag && c < KV;
        bkk[r] = bok[r] ? (int)(c / V - kmin) : 0;
        bv[r] = bok[r] ? (int)(c % V) : 0;
    }

    f64x4 acc_re[CB], acc_im[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc_re[cb][q] = acc_im[cb][q] = 0.0;

    // kz factors of step s's 16 nodes
    auto ez_gen = [&](int s) {
        double2 *dst = sEz64;
        for (int e = tid; e < nk * 16; e += kThreads) {
            const int kk = e / 16, m = s * 16 + e % 16;
            const double u = fma(sH[3 * kk], sx[m], fma(sH[3 * kk + 1], sy[m], sH[3 * kk + 2] * sz[m]));
            double c, sn;
            cis_neg_f64(u, c, sn);
            dst[kk * kEz64Stride + e % 16] = make_double2(c, sn);
        }
    };

    for (int64_t base = n_begin; base < n_end; base += kChunk) {
        __syncthreads();  // previous chunk consumed; scoef and sH visible
        {
            const int64_t n = base + tid;
            double x = 0.0, y = 0.0, z = 0.0;
            double2 cv[kMovMaxV];
#pragma unroll
            for (int v = 0; v < kMovMaxV; ++v) cv[v] = make_double2(0.0, 0.0);
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
                double ec, es;
                cis_neg_f64(u, ec, es);
                const double d = exp(-ti * r2);
                const double er = d * ec, ei = d * es;
#pragma unroll
                for (int v = 0; v < kMovMaxV; ++v) {
                    if (v < V) {
                        const double br = packed[(int64_t)(5 + 2 * v) * N + n];
                        const double bi = packed[(int64_t)(6 + 2 * v) * N + n];
                        cv[v] = make_double2(br * er - bi * ei, br * ei + bi * er);
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
        ez_gen(0);
        __syncthreads();
        for (int s = 0; s < nsteps; ++s) {
            // B fragments of step s: node 16 s + 4 st + lane/16, column c
            {
                const double2 *ez = sEz64;
#pragma unroll
                for (int r = 0; r < kBF; ++r) {
                    const int f = (tid + r * kThreads) / 64;
                    if (f < kFrag) {
                        const int st = f % 4;
                        const int m = 4 * st + lane / 16;
                        const double2 a = ez[bkk[r] * kEz64Stride + m];
                        const double2 b = sc[bv[r]][s * 16 + m];
                        const double br = bok[r] ? a.x * b.x - a.y * b.y : 0.0;
                        const double bi = bok[r] ? a.x * b.y + a.y * b.x : 0.0;
                        sB[(f * 2 + 0) * 64 + lane] = br;
                        sB[(f * 2 + 1) * 64 + lane] = bi;
                    }
                }
            }
            // A fragments of step s (registers): row `row`, node 16 s + 4 st + lane/16
            double ar[4], ai[4];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const int m = s * 16 + 4 * st + lane / 16;
                const double u = fma(Gx, sx[m], fma(Gy, sy[m], Gz * sz[m]));
                cis_neg_f64(u, ar[st], ai[st]);
            }
            __syncthreads();  // B planes written, kz factors of step s read
            if (s + 1 < nsteps) ez_gen(s + 1);
#pragma unroll
            for (int st = 0; st < 4; ++st) {
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) {
                    const int f = cb * 4 + st;
                    const double br = sB[(f * 2 + 0) * 64 + lane], bi = sB[(f * 2 + 1) * 64 + lane];
                    acc_re[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[st], br, acc_re[cb], 0, 0, 0);
                    acc_im[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[st], bi, acc_im[cb], 0, 0, 0);
                }
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) {
                    const int f = cb * 4 + st;
                    const double br = sB[(f * 2 + 0) * 64 + lane], bi = sB[(f * 2 + 1) * 64 + lane];
                    acc_re[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[st], bi, acc_re[cb], 0, 0, 0);
                    acc_im[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[st], br, acc_im[cb], 0, 0, 0);
                }
            }
            __syncthreads();  // B planes read and the next kz factors written
        }
    }

    // accumulator register q of lane l holds row l/16 + 4q, column l%16
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        const int64_t col = c0 + cb * 16 + (lane % 16);
        if (col >= KV) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t r = (rt * W + wave) * 16 + (lane / 16) + 4 * q;
            if (r < L) part[((split * M + i) * L + r) * KV + col] = make_double2(acc_re[cb][q], acc_im[cb][q]);
        }
    }
}

// Blocks per workgroup tile: the largest count in `opts` (descending, at most
// `cap`) whose tiles use at least 80 % of their blocks over n blocks.
static int pick_blocks(int64_t n, std::initializer_list<int> opts, int cap)
{
    for (int o : opts) {
        if (o > cap || (o > n && o != 1)) continue;
        if (5 * n >= 4 * cdiv(n, o) * o) return o;
    }
    return 1;
}

// LDS bytes of one workgroup (static arrays + the kz-factor buffer).
static size_t lds_bytes_f64(int W, int CB, int ez_rows)
{
    const size_t chunk = 64 * (size_t)W;
    return 3 * chunk * 8 + kMovMaxV * (chunk + 1) * 16 + (size_t)CB * 4 * 2 * 64 * 8 +
           3 * (16 * (size_t)CB + 1) * 8 + (kMaxModes + 3) * 8 + (size_t)ez_rows * kEz64Stride * 16;
}

template <int W>
static int mf64_blocks_w(int CB, size_t dyn)
{
    switch (CB) {
    case 1: return blocks_per_cu(k_cart_moving_f64<W, 1>, 64 * W, dyn);
    case 2: return blocks_per_cu(k_cart_moving_f64<W, 2>, 64 * W, dyn);
    case 3: if constexpr (W >= 2) return blocks_per_cu(k_cart_moving_f64<W, 3>, 64 * W, dyn); break;
    case 4: if constexpr (W >= 2) return blocks_per_cu(k_cart_moving_f64<W, 4>, 64 * W, dyn); break;
    case 6: if constexpr (W >= 3) return blocks_per_cu(k_cart_moving_f64<W, 6>, 64 * W, dyn); break;
    default: break;
    }
    return 1;
}

// Workgroups of the shape's instantiation that one CU holds at once.
static int mf64_blocks_per_cu(int W, int CB, int ez_rows)
{
    const size_t dyn = (size_t)ez_rows * kEz64Stride * sizeof(double2);
    switch (W) {
    case 1: return mf64_blocks_w<1>(CB, dyn);
    case 2: return mf64_blocks_w<2>(CB, dyn);
    case 3: return mf64_blocks_w<3>(CB, dyn);
    default: return mf64_blocks_w<4>(CB, dyn);
    }
}

MovingShape moving_shape_f64(const CartesianTables &c, int64_t N, int V)
{
    MovingShape s{};
    s.M = c.n_ro;
    s.L = c.n_ln;
    s.K = c.n_kz;
    s.V = V;
    s.KV = s.K * V;
    const int64_t nrb = cdiv(s.L, 16), ncb = cdiv(s.KV, 16);
    s.W = pick_blocks(nrb, {4, 3, 2, 1}, 4);
    // at most 2 W column blocks (8 B fragments per thread) and 6 (8 would
    // exceed 256 registers), within the LDS of a CU
    int cap = std::min(2 * s.W, 6);
    auto ez_rows_of = [&](int cb) {
        const int64_t cols = 16 * (int64_t)cb;
        return (int)std::min<int64_t>(cols, (cols - 1) / V + 2);
    };
    s.CB = pick_blocks(ncb, {6, 4, 3, 2, 1}, cap);
    while (s.CB > 1 && lds_bytes_f64(s.W, s.CB, ez_rows_of(s.CB)) > 96 * 1024) {
        cap = s.CB - 1;
        s.CB = pick_blocks(ncb, {6, 4, 3, 2, 1}, cap);
    }
    s.n_rt = cdiv(nrb, s.W);
    s.n_ct = cdiv(ncb, s.CB);
    const int64_t chunk = 64 * s.W;
    const int64_t base = s.M * s.n_rt * s.n_ct;
    s.ez_rows = ez_rows_of(s.CB);
    // four workgroups per CU at least, in whole rounds of the workgroups the
    // device holds; FP64 accumulators need no bound on the split length
    const int64_t capacity = (int64_t)mf64_blocks_per_cu(s.W, s.CB, s.ez_rows) * device_cu_count();
    const int64_t splits = balanced_splits(N, chunk, base, cdiv(4 * (int64_t)device_cu_count(), base), capacity);
    s.nodes_per_split = cdiv(cdiv(std::max<int64_t>(N, 1), splits), chunk) * chunk;
    s.n_splits = cdiv(std::max<int64_t>(N, 1), s.nodes_per_split);
    s.part_bytes = (size_t)s.n_splits * (size_t)s.M * (size_t)s.L * (size_t)s.KV * sizeof(double2);
    return s;
}

template <int W, int CB>
static void launch_mf64_wc(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                           const float *modes, int R, const double *mcoef, double2 *part, hipStream_t st)
{
    const int64_t grid = s.n_splits * s.M * s.n_rt * s.n_ct;
    const size_t dyn = (size_t)s.ez_rows * kEz64Stride * sizeof(double2);
    k_cart_moving_f64<W, CB><<<(unsigned)grid, 64 * W, dyn, st>>>(
        packed, N, (int)s.V, modes, R, mcoef, c.d_ro_group, c.d_F, c.d_G, c.d_H, c.d_t, s.M, s.L, s.K, s.n_rt,
        s.n_ct, s.n_splits, s.nodes_per_split, s.ez_rows, part);
}

// Column-block counts instantiated per W: those <= min(2 W, 6) (the shape
// never picks more).
template <int W>
static void launch_mf64_w(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                          const float *modes, int R, const double *mcoef, double2 *part, hipStream_t st)
{
    switch (s.CB) {
    case 1: launch_mf64_wc<W, 1>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 2: launch_mf64_wc<W, 2>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 3: if constexpr (W >= 2) launch_mf64_wc<W, 3>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 4: if constexpr (W >= 2) launch_mf64_wc<W, 4>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 6: if constexpr (W >= 3) launch_mf64_wc<W, 6>(s, c, packed, N, modes, R, mcoef, part, st); break;
    default: break;
    }
}

void launch_cart_moving_f64(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                            const float *modes, int R, const double *mcoef, double2 *part, hipStream_t st)
{
    switch (s.W) {
    case 1: launch_mf64_w<1>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 2: launch_mf64_w<2>(s, c, packed, N, modes, R, mcoef, part, st); break;
    case 3: launch_mf64_w<3>(s, c, packed, N, modes, R, mcoef, part, st); break;
    default: launch_mf64_w<4>(s, c, packed, N, modes, R, mcoef, part, st); break;
    }
}

}  // namespace fm
