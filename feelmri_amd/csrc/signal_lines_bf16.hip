// Static nodes, lines linear in k and t (radial spokes, any straight
// readout): one complex GEMM per line on the matrix cores, operands built in
// the kernel (mixed form).
//
// Along a line k(m) = k0 + m dk and t(m) = t0 + m dt (m = 0 .. L-1), so the
// sum over nodes of MRI.cpp:196-297 is
//
//   S[m] = sum_n c_n e_n z_n^m,   e_n = exp(-t0/T2) exp(-2 pi i (k0.x - phi t0 / 2 pi))
//                                 z_n = exp(-dt/T2) exp(-2 pi i (dk.x - phi dt / 2 pi))
//
// With m = Q p + q (q < Q, p < P, P Q >= L) this is the product
//
//   S[p, q] = sum_n A[p, n] B[n, q],   A[p, n] = c_n e_n Z_n^p,  Z_n = z_n^Q
//                                      B[n, q] = z_n^q
//
// of a (P x nodes) and a (nodes x Q) complex matrix: P + Q operands per node
// and line produce P Q samples (48 for 512, against 512 sample updates of a
// recurrence).  Per 32-node build, lane l of the wave owns node l % 32 and
// half h = l / 32 of its operands: B columns q = 2 j + h and A rows
// p = 2 j + h, each chain advanced by z^2 or Z^2 from its start (1 or z, c e
// or c e Z).  The four complex factors z, Z, c e, c e Z come from FP64
// phases, range-reduced to FP32 turns for the hardware sine and cosine (two
// per lane; the halves swap z and Z).  Values are split into bf16 hi + lo
// planes in LDS, (re, im) pairs adjacent along K, so that
//   Re S = sum_k A'[p,k] B[k,q],   A' = (Ar, -Ai) per node
//   Im S = sum_k A''[p,k] B[k,q],  A'' = (Ai, Ar)  (A' with halves swapped)
// each as hh + hl + lh on v_mfma_f32_16x16x32_bf16.  FP32 accumulators over
// at most 16 384 nodes per split; k_lines_reduce adds the splits in FP64 and
// writes the samples in the output order.
#include <algorithm>

#include "cart_gemm.h"

namespace fm {

typedef __attribute__((ext_vector_type(4))) float ln_f32x4;
typedef __attribute__((ext_vector_type(2))) float ln_f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 ln_bf16x2;

constexpr int kLineNB = 32;                  // nodes per operand build (two K steps of 16)
constexpr int kLineRow = 2 * kLineNB + 4;    // LDS words of an operand row: hi words, lo words, pad
constexpr int64_t kLineNodesPerSplit = 16384;
constexpr int kLineMaxRT = 4;                // row tiles of 16 (P <= 64)
constexpr int kLineValuPerMfma = 12;         // scheduling: VALU slots between two products

// exp(-r2 tt) exp(-2 pi i u): FP64 phase reduced to FP32 turns.
__device__ inline ln_f32x2 ln_cis_decay(double u, double tt, double r2)
{
    const float rf = turns_f32(u);
    const float d = __builtin_amdgcn_exp2f((float)(-tt * r2) * 1.44269504f);
    return ln_f32x2{d * __builtin_amdgcn_cosf(rf), -d * __builtin_amdgcn_sinf(rf)};
}

__device__ inline ln_f32x2 ln_cmul(ln_f32x2 a, ln_f32x2 b)
{
    return ln_f32x2{fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)};
}

// bf16 hi and lo words of the pair (a, b): a in the low half.
__device__ inline void ln_split(float a, float b, uint32_t &hw, uint32_t &lw)
{
    const ln_f32x2 v = {a, b};
    const ln_bf16x2 h = __builtin_convertvector(v, ln_bf16x2);
    const ln_bf16x2 l = __builtin_convertvector(v - __builtin_convertvector(h, ln_f32x2), ln_bf16x2);
    hw = __builtin_bit_cast(uint32_t, h);
    lw = __builtin_bit_cast(uint32_t, l);
}

// (Ar, -Ai) -> (Ai, Ar) in each 32-bit word: rotate by 16, flip the new low
// half's sign.
__device__ inline bf16x8 ln_swap_neg(bf16x8 v)
{
    uint4 u = __builtin_bit_cast(uint4, v);
    u.x = __builtin_amdgcn_alignbit(u.x, u.x, 16) ^ 0x8000u;
    u.y = __builtin_amdgcn_alignbit(u.y, u.y, 16) ^ 0x8000u;
    u.z = __builtin_amdgcn_alignbit(u.z, u.z, 16) ^ 0x8000u;
    u.w = __builtin_amdgcn_alignbit(u.w, u.w, 16) ^ 0x8000u;
    return __builtin_bit_cast(bf16x8, u);
}

// LDS operand of R rows (32-bit words): row r holds the bf16 hi words of its
// 32 nodes, then the lo words (kLineRow words a row).  Even rows form one
// block and odd rows a second block placed 32 banks away, since the two wave
// halves write rows 2 j and 2 j + 1 at once (64 banks of 4 bytes); the
// 16-byte fragment reads of 16 rows then also fall on distinct banks.
template <int R>
struct LinePlane {
    static constexpr int kOdd = (R / 2) * kLineRow + (32 - ((R / 2) * kLineRow) % 64 + 64) % 64;
    static constexpr int kWords = kOdd + (R / 2) * kLineRow;
    __device__ static constexpr int row(int r) { return (r & 1) * kOdd + (r >> 1) * kLineRow; }
};

// Node data of one build, loaded a chunk ahead of its use.
struct LineNodeRaw {
    double x, y, z, phi, r2, br, bi;
};

// One wave per workgroup: line `line` of the (split, line) pair the
// workgroup id maps to; P = 16 NRT rows, Q = 16 NCT columns.  Node offsets
// are 32-bit (N < kLineMaxNodes).
template <int NRT, int NCT>
__global__ __launch_bounds__(64) void k_signal_lines_bf16(const double *__restrict__ packed, int64_t N, int v,
                                                          const RunDesc *__restrict__ lines, int64_t n_lines,
                                                          int64_t n_splits, int64_t nodes_per_split,
                                                          float2 *__restrict__ part)
{
    constexpr int P = 16 * NRT, Q = 16 * NCT, PQ = P * Q;
    using PB = LinePlane<Q>;
    using PA = LinePlane<P>;
    __shared__ uint32_t sB[PB::kWords], sA[PA::kWords];

    // workgroup -> (split, line); the lines of one split are contiguous and
    // dealt onto one XCD, which then reads the split's nodes from its L2
    const int64_t total = n_splits * n_lines;
    const int64_t h = blockIdx.x;
    int64_t lg = h;
    if (total % kXcds == 0) lg = (h % kXcds) * (total / kXcds) + h / kXcds;
    const int64_t line = lg % n_lines;
    const int64_t split = lg / n_lines;

    const int lane = threadIdx.x;
    const int nl = lane & 31, half = lane >> 5;
    const int r16 = lane & 15, g16 = lane >> 4;
    const RunDesc ld = lines[line];
    const uint32_t n_begin = (uint32_t)(split * nodes_per_split);
    const uint32_t n_end = (uint32_t)std::min<int64_t>((int64_t)n_begin + nodes_per_split, N);

    // per half: the factor this lane forms (z or Z) and its chain start's
    // phase and time (c e or c e Z)
    const double dkx = ld.dkx, dky = ld.dky, dkz = ld.dkz, dt = ld.dt;
    const double sa = half ? (double)Q : 1.0;          // z (half 0) or Z = z^Q (half 1)
    const double tb = half ? fma((double)Q, dt, ld.t0) : ld.t0;

    ln_f32x4 acc_re[NRT][NCT], acc_im[NRT][NCT];
#pragma unroll
    for (int a = 0; a < NRT; ++a)
#pragma unroll
        for (int b = 0; b < NCT; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc_re[a][b][q] = acc_im[a][b][q] = 0.0f;

    const double *px = packed, *py = packed + N, *pz = packed + 2 * N, *pphi = packed + 3 * N,
                 *pr2 = packed + 4 * N, *pbr = packed + (5 + 2 * (int64_t)v) * N,
                 *pbi = packed + (6 + 2 * (int64_t)v) * N;

    // this lane's node of the chunk at `base` (clamped into the split: the
    // build zeroes c past its end)
    auto load = [&](uint32_t base) {
        const uint32_t n = base + nl;
        const uint32_t nn = n < n_end ? n : n_end - 1;
        return LineNodeRaw{px[nn], py[nn], pz[nn], pphi[nn], pr2[nn], pbr[nn], pbi[nn]};
    };

    // operand build of the chunk at `base` into the LDS planes; nodes past
    // the split's end get c = 0 (so A = 0) and are never read back
    auto build = [&](uint32_t base, const LineNodeRaw &nd) {
        const bool live = base + nl < n_end;
        const double dwc = nd.phi * (1.0 / kTwoPi);
        const float cr = live ? (float)nd.br : 0.0f, ci = live ? (float)nd.bi : 0.0f;
        // phases in cycles (MRI.cpp:138-146, with k' = k pi_f / pi)
        const double uz = fma(dkx, nd.x, fma(dky, nd.y, fma(dkz, nd.z, -dwc * dt)));
        const double ue = fma(ld.k0x, nd.x, fma(ld.k0y, nd.y, fma(ld.k0z, nd.z, -dwc * ld.t0)));
        const ln_f32x2 own = ln_cis_decay(sa * uz, sa * dt, nd.r2);                    // z | Z
        const ln_f32x2 eb = ln_cis_decay(half ? fma((double)Q, uz, ue) : ue, tb, nd.r2);  // e | e Z
        const ln_f32x2 oth = {__shfl_xor(own.x, 32), __shfl_xor(own.y, 32)};          // Z | z
        const ln_f32x2 zf = half ? oth : own, Zf = half ? own : oth;
        const ln_f32x2 z2 = ln_cmul(zf, zf);
        // B: q = 2 j + half, from 1 (half 0) or z (half 1)
        ln_f32x2 b = half ? zf : ln_f32x2{1.0f, 0.0f};
        uint32_t *wb = sB + half * PB::kOdd + nl;
#pragma unroll
        for (int j = 0; j < Q / 2; ++j) {
            uint32_t hw, lw;
            ln_split(b.x, b.y, hw, lw);
            wb[j * kLineRow] = hw;
            wb[j * kLineRow + kLineNB] = lw;
            b = ln_cmul(b, z2);
        }
        // A' = conj(A): p = 2 j + half, from conj(c e) (half 0) or conj(c e Z)
        // (half 1), advanced by conj(Z^2)
        const ln_f32x2 Zc = {Zf.x, -Zf.y};
        const ln_f32x2 Z2c = ln_cmul(Zc, Zc);
        ln_f32x2 a = ln_cmul(ln_f32x2{cr, -ci}, ln_f32x2{eb.x, -eb.y});
        uint32_t *wa = sA + half * PA::kOdd + nl;
#pragma unroll
        for (int j = 0; j < P / 2; ++j) {
            uint32_t hw, lw;
            ln_split(a.x, a.y, hw, lw);
            wa[j * kLineRow] = hw;
            wa[j * kLineRow + kLineNB] = lw;
            a = ln_cmul(a, Z2c);
        }
    };

    // Software pipeline: the fragments of chunk k go to registers, then the
    // wave builds chunk k + 1 into the planes (its node data loaded one chunk
    // earlier) while the matrix cores run chunk k's products.
    LineNodeRaw nxt = load(n_begin);
    build(n_begin, nxt);
    nxt = load(n_begin + kLineNB);
    __syncthreads();
    for (uint32_t base = n_begin; base < n_end; base += kLineNB) {
        bf16x8 bh[2][NCT], bl[2][NCT], ah[2][NRT], al[2][NRT];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int col = 16 * ks + 4 * g16;  // this lane's 4 nodes (words) of the K step
#pragma unroll
            for (int b = 0; b < NCT; ++b) {
                const uint32_t *rb = sB + PB::row(16 * b + r16) + col;
                bh[ks][b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(rb));
                bl[ks][b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(rb + kLineNB));
            }
#pragma unroll
            for (int a = 0; a < NRT; ++a) {
                const uint32_t *ra = sA + PA::row(16 * a + r16) + col;
                ah[ks][a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(ra));
                al[ks][a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(ra + kLineNB));
            }
        }
        __syncthreads();  // the planes are read before the next build rewrites them
        const LineNodeRaw cur = nxt;
        nxt = load(base + 2 * kLineNB);
        build(base + kLineNB, cur);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int a = 0; a < NRT; ++a) {
                const bf16x8 ah2 = ln_swap_neg(ah[ks][a]), al2 = ln_swap_neg(al[ks][a]);
#pragma unroll
                for (int b = 0; b < NCT; ++b) {
                    ln_f32x4 re = acc_re[a][b], im = acc_im[a][b];
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks][a], bh[ks][b], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah2, bh[ks][b], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks][a], bl[ks][b], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah2, bl[ks][b], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks][a], bh[ks][b], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al2, bh[ks][b], im, 0, 0, 0);
                    acc_re[a][b] = re;
                    acc_im[a][b] = im;
                }
            }
        // interleave: one product, then a share of the build's VALU work
#pragma unroll
        for (int i = 0; i < 12 * NRT * NCT; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, kLineValuPerMfma, 0);
        }
        __syncthreads();  // the next chunk's planes are written
    }

    // acc register r of lane l holds row 4 (l / 16) + r and column l % 16 of
    // its 16 x 16 tile: sample m = Q p + q
    float2 *dst = part + (split * n_lines + line) * (int64_t)PQ;
#pragma unroll
    for (int a = 0; a < NRT; ++a)
#pragma unroll
        for (int b = 0; b < NCT; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = 16 * a + 4 * g16 + r, q = 16 * b + r16;
                dst[p * Q + q] = make_float2(acc_re[a][b][r], acc_im[a][b][r]);
            }
}

// out[(m stride + line) V + v] = sum over splits of the partial samples, FP64.
__global__ void k_lines_reduce(const float2 *__restrict__ part, int64_t n_splits, int64_t n_lines, int64_t PQ,
                               int64_t L, const RunDesc *__restrict__ lines, int64_t stride, int V, int v,
                               double2 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_lines * L) return;
    const int64_t line = i / L, m = i % L;
    double re = 0.0, im = 0.0;
    for (int64_t s = 0; s < n_splits; ++s) {
        const float2 p = part[(s * n_lines + line) * PQ + m];
        re += (double)p.x;
        im += (double)p.y;
    }
    const int64_t sid = (int64_t)lines[line].s0 + m * stride;
    out[sid * V + v] = make_double2(re, im);
}

struct LineShape {
    int nrt, nct;          // row tiles (P = 16 nrt), column tiles (Q = 16 nct)
    int64_t PQ;
    int64_t n_splits, nodes_per_split;
    size_t part_bytes;
};

// Workgroups of the (NRT, NCT) instantiation that one CU holds at once.
static int line_blocks_per_cu(int nrt, int nct)
{
    if (nct == 1) return blocks_per_cu(k_signal_lines_bf16<1, 1>, 64, 0);
    switch (nrt) {
    case 1: return blocks_per_cu(k_signal_lines_bf16<1, 2>, 64, 0);
    case 2: return blocks_per_cu(k_signal_lines_bf16<2, 2>, 64, 0);
    case 3: return blocks_per_cu(k_signal_lines_bf16<3, 2>, 64, 0);
    default: return blocks_per_cu(k_signal_lines_bf16<4, 2>, 64, 0);
    }
}

static LineShape line_shape(const fm_signal_plan *p, int64_t N)
{
    LineShape s{};
    const int64_t L = p->n_meas;
    s.nct = L <= 256 ? 1 : 2;
    const int64_t Q = 16 * s.nct;
    s.nrt = (int)cdiv(cdiv(L, Q), 16);
    s.PQ = 16 * (int64_t)s.nrt * Q;
    const int64_t nl = std::max<int64_t>(p->n_line_tabs, 1);
    // ~8 workgroups per CU at least, at most 16 384 nodes per split (FP32
    // accumulation), at least 256 nodes per split, in whole rounds of the
    // workgroups the device holds
    const int64_t s_min = std::max(cdiv(8 * (int64_t)device_cu_count(), nl),
                                   cdiv(std::max<int64_t>(N, 1), kLineNodesPerSplit));
    const int64_t capacity = (int64_t)line_blocks_per_cu(s.nrt, s.nct) * device_cu_count();
    const int64_t splits = balanced_splits(N, kLineNB, nl, s_min, capacity, cdiv(std::max<int64_t>(N, 1), 256));
    s.nodes_per_split = cdiv(cdiv(std::max<int64_t>(N, 1), splits), kLineNB) * kLineNB;
    s.n_splits = cdiv(std::max<int64_t>(N, 1), s.nodes_per_split);
    s.part_bytes = (size_t)s.n_splits * (size_t)nl * (size_t)s.PQ * sizeof(float2);
    return s;
}

bool lines_usable(const fm_signal_plan *p, int64_t N)
{
    return p->n_line_tabs > 0 && p->n_meas >= kLineMinLen && p->n_meas <= 16 * kLineMaxRT * 32 &&
           N < kLineMaxNodes;
}

size_t lines_workspace(const fm_signal_plan *p, int64_t N)
{
    if (!lines_usable(p, N) || N <= 0) return 0;
    return line_shape(p, N).part_bytes;
}

template <int NRT, int NCT>
static void launch_lines_t(const LineShape &s, const fm_signal_plan *p, const double *packed, int64_t N, int v,
                           float2 *part, hipStream_t st)
{
    const int64_t grid = s.n_splits * p->n_line_tabs;
    k_signal_lines_bf16<NRT, NCT><<<(unsigned)grid, 64, 0, st>>>(packed, N, v, p->d_lines, p->n_line_tabs,
                                                                 s.n_splits, s.nodes_per_split, part);
}

int lines_execute(const fm_signal_plan *p, int64_t N, int V, const double *packed, void *ws, size_t ws_bytes,
                  double *out, hipStream_t st)
{
    if (!lines_usable(p, N)) {
        set_error("plan has no linear lines for the line GEMM");
        return FM_ERR_UNSUPPORTED;
    }
    const LineShape s = line_shape(p, N);
    if (s.part_bytes > ws_bytes) {
        set_error("signal workspace too small: need %zu bytes, got %zu", s.part_bytes, ws_bytes);
        return FM_ERR_WORKSPACE;
    }
    const int64_t grid = s.n_splits * p->n_line_tabs;
    if (grid > (int64_t)0x7fffffff) {
        set_error("line GEMM exceeds the launch grid (%lld workgroups)", (long long)grid);
        return FM_ERR_UNSUPPORTED;
    }
    float2 *part = reinterpret_cast<float2 *>(ws);
    const int64_t L = p->n_meas;
    for (int v = 0; v < V; ++v) {
        plan_timing_start(p, st);
#define FM_LN(R, C) launch_lines_t<R, C>(s, p, packed, N, v, part, st)
        if (s.nct == 1) {
            FM_LN(1, 1);
        } else {
            switch (s.nrt) {
            case 1: FM_LN(1, 2); break;
            case 2: FM_LN(2, 2); break;
            case 3: FM_LN(3, 2); break;
            default: FM_LN(4, 2); break;
            }
        }
#undef FM_LN
        plan_timing_stop(p, st);
        FM_LAUNCH_CHECK("k_signal_lines_bf16");
        const int64_t n_out = p->n_line_tabs * L;
        k_lines_reduce<<<(unsigned)cdiv(n_out, 256), 256, 0, st>>>(part, s.n_splits, p->n_line_tabs, s.PQ, L,
                                                                   p->d_lines, p->n_lines * p->n_kz, V, v,
                                                                   reinterpret_cast<double2 *>(out));
        FM_LAUNCH_CHECK("k_lines_reduce");
    }
    return FM_OK;
}

}  // namespace fm
