// Direct non-uniform k-space sum (any trajectory), FP64 and mixed forms.
//
//   S[s, v] = sum_n exp(i(-2 pi_f k_s . r_n + phi_n t_s)) exp(-t_s / T2_n) b[n, v]
//
// reference: cpp/feelmri/MRI.cpp:196-297 (static nodes), :38-159 (moving).
//
// Work decomposition.  Samples of equal t are grouped (the reference's t
// cache) and cut into slots of up to 256 samples; a workgroup owns a tile of
// two slots (2 samples per lane, the slots possibly of different t) and one
// contiguous split of the nodes.  Nodes stream
// through LDS in chunks of 256: each lane of the workgroup stages one node,
// folding the t-dependent factors (T2* decay, off-resonance, and the POD
// position update for moving nodes) into the staged coefficient, so the
// inner loop per (node, sample) is just the k.x phase, one sincos and a
// complex MAC.  All lanes read the same LDS node at the same time
// (broadcast).  Node splits write partial sums that a second kernel adds in
// a fixed order, so results do not depend on scheduling.
//
// Phase handling.  u = k'.x is accumulated in FP64 in cycles (k' carries the
// reference's float-pi factor).  FP64 form: r = u - rint(u) in [-1/2, 1/2],
// sincospi(2r).  Mixed form: fract(u) in [0, 1) (one instruction, exact),
// rounded to FP32 and fed to the hardware v_sin_f32 / v_cos_f32, which take
// revolutions; partial sums over a 256-node chunk are kept in FP32 and added
// into FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <unordered_map>

#include "signal_plan.h"

namespace fm {

template <bool MIXED>
struct CoefT;
template <>
struct CoefT<true> {
    using T = float;
};
template <>
struct CoefT<false> {
    using T = double;
};

// A node staged in LDS.  Shared-t tiles carry one coefficient set per slot
// (the slots of a tile may belong to different t-groups) or, in packed slots
// of GP t-groups, one per (slot, group); for moving nodes one position per
// slot.
template <int VB, bool MIXED, bool SHARED_T, bool MOTION, int GP = 1, bool CONJ = false>
struct alignas(16) StagedNode {
    using CT = typename CoefT<MIXED>::T;
    static constexpr int kPos = MOTION ? kDirectSPL : 1;
    // conjugate-pair tiles: both slots hold the same GP t-groups
    static constexpr int kCoef = SHARED_T ? (CONJ ? GP : kDirectSPL * GP) : 1;
    static constexpr bool kPair = MIXED && GP > 1 && !CONJ;  // second slot's phase by FP32 offset
    double x[kPos], y[kPos], z[kPos];
    double dwc;                  // phi/(2 pi), cycles/ms   (per-sample-t kernel only)
    double r2;                   // 1/T2, 1/ms              (per-sample-t kernel only)
    CT c[kCoef][2 * VB];
    float xf[kPair ? 4 : 1];     // packed mixed slots: x, y, z in FP32 and max |x|, |y|, |z|
};

// Tiles.  Shared-t ("grouped") tiles are kDirectSPL slot records
// tiles[kDirectSPL * b + j]: slot j holds up to 256 samples of one t-group
// (lane l takes sample perm[offset + l]).  Packed slots (GP > 1, static
// nodes) hold GP consecutive t-groups of 256 / GP samples each, starting at
// group `group`: lane l belongs to group + l / (256 / GP), and the
// t-dependent factor exp(-t/T2) exp(i phi t) is staged per (node, slot,
// group) instead of being formed per (node, sample).  Per-sample-t
// ("consecutive") tiles are one record of up to 512 consecutive samples
// (lane l takes offset + l + 256 j).
//
// In packed slots of the mixed form the two samples of a lane lie on one
// line, GP readout indices apart.  When |k'1 - k'0|_1 max(|x|,|y|,|z|) <= 4
// cycles for every node of a chunk (pair_dk is the plan's bound on the
// first factor), the second phase is formed as the first one's FP64-reduced
// phase plus (k'1 - k'0).x in FP32: an FP32 error of ~1e-6 cycles on top of
// the FP64 range reduction.  PLANAR drops the kz.z term (every kz = 0).
//
// Conjugate-pair tiles (CONJ, mixed form): both slots hold the same GP
// t-groups, and lane l's slot-1 sample has k'_b = -k'_a of its slot-0
// sample (to 1e-12 of the largest |k'|, the plan checks; e.g. spiral
// interleaves half a turn apart).  One phase and one sine/cosine then serve
// both: exp(-2 pi i k'_b.x) = conj(exp(-2 pi i k'_a.x)), and the two share
// the t-group's coefficient c, so the lane adds c e and c conj(e).
template <int VB, bool MIXED, bool SHARED_T, bool MOTION, int GP = 1, bool PLANAR = false, bool CONJ = false>
__global__ __launch_bounds__(kDirectThreads) void k_signal_direct(
    const double *__restrict__ packed, int64_t N, int V, int v0,
    const double4 *__restrict__ ks, const int32_t *__restrict__ perm,
    const SignalTile *__restrict__ tiles, const double *__restrict__ tg_t,
    const float *__restrict__ modes, int R, const double *__restrict__ mcoef,
    int64_t nodes_per_split, double2 *__restrict__ dst, int64_t S, int write_final, double pair_dk)
{
    using CT = typename CoefT<MIXED>::T;
    using Node = StagedNode<VB, MIXED, SHARED_T, MOTION, GP, CONJ>;
    static_assert(GP == 1 || (SHARED_T && !MOTION), "packed t-groups: shared-t tiles of static nodes");
    static_assert(!CONJ || (MIXED && GP > 1 && kDirectSPL == 2), "conjugate pairs: mixed packed slots");
    static_assert(!PLANAR || GP > 1, "planar phases: packed slots only");
    constexpr bool kPair = Node::kPair && kDirectSPL == 2;
    constexpr int kSlots = SHARED_T ? kDirectSPL : 1;
    constexpr int kGSize = kDirectThreads / GP;  // samples per group of a packed slot
    __shared__ Node snode[kDirectThreads];
    __shared__ double scoef[MOTION ? kDirectSPL : 1][kMaxModes + 3];

    const int64_t split = blockIdx.y;
    const int64_t n_begin = split * nodes_per_split;
    const int64_t n_end = (n_begin + nodes_per_split < N) ? n_begin + nodes_per_split : N;
    const int tid = threadIdx.x;

    const double *px = packed;
    const double *py = packed + N;
    const double *pz = packed + 2 * N;
    const double *pphi = packed + 3 * N;
    const double *pr2 = packed + 4 * N;
    const double *pb = packed + 5 * N;  // rows 2v, 2v+1

    SignalTile tl[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) tl[j] = tiles[(int64_t)blockIdx.x * kSlots + j];
    // slot (and group) times: uniform over the workgroup, kept in LDS
    __shared__ double tg[kDirectSPL * GP];
    bool same_group = GP == 1;  // both slots share one t-group (stage once)
    if (SHARED_T) {
        if (tid < kDirectSPL * GP) {
            const int j = tid / GP, q = tid % GP;
            tg[tid] = tl[j].count > q * kGSize ? tg_t[tl[j].group + q] : 0.0;
        }
#pragma unroll
        for (int j = 1; j < kDirectSPL; ++j) same_group = same_group && (tl[j].group == tl[0].group || tl[j].count == 0);
    }
    const int gq = GP > 1 ? threadIdx.x / kGSize : 0;  // this lane's group in a packed slot
    if (MOTION) {
#pragma unroll
        for (int j = 0; j < kDirectSPL; ++j)
            for (int q = tid; q < R + 3; q += kDirectThreads)
                scoef[j][q] = tl[j].count > 0 ? mcoef[(int64_t)tl[j].group * (R + 3) + q] : 0.0;
    }

    int sid[kDirectSPL];
    bool valid[kDirectSPL];
    double kxs[kDirectSPL], kys[kDirectSPL], kzs[kDirectSPL], ts[kDirectSPL];
#pragma unroll
    for (int j = 0; j < kDirectSPL; ++j) {
        if (SHARED_T) {
            valid[j] = tid < tl[j].count;
            sid[j] = valid[j] ? perm[tl[j].offset + tid] : 0;
        } else {
            const int idx = tid + j * kDirectThreads;
            valid[j] = idx < tl[0].count;
            sid[j] = valid[j] ? tl[0].offset + idx : 0;
        }
        const double4 k = ks[sid[j]];
        kxs[j] = k.x;
        kys[j] = k.y;
        kzs[j] = k.z;
        ts[j] = k.w;
    }
    const float dkx = (float)(kxs[kSlots - 1] - kxs[0]), dky = (float)(kys[kSlots - 1] - kys[0]),
                dkz = (float)(kzs[kSlots - 1] - kzs[0]);

    double tot_re[kDirectSPL][VB], tot_im[kDirectSPL][VB];
#pragma unroll
    for (int j = 0; j < kDirectSPL; ++j)
#pragma unroll
        for (int v = 0; v < VB; ++v) tot_re[j][v] = tot_im[j][v] = 0.0;

    for (int64_t base = n_begin; base < n_end; base += kDirectThreads) {
        __syncthreads();  // previous chunk fully consumed (and scoef visible)
        int far_node = 0;  // this thread's node too far out for the FP32 offset
        {
            const int64_t n = base + tid;
            // staged in place: packed slots carry 2 GP coefficient sets per
            // node, too many to assemble in registers first
            Node &nd = snode[tid];
            if (n < n_end) {
                const double x0 = px[n], y0 = py[n], z0 = pz[n];
#pragma unroll
                for (int j = 0; j < Node::kPos; ++j) {
                    double x = x0, y = y0, z = z0;
                    if (MOTION) {
                        if (j > 0 && same_group) {
                            x = nd.x[0];
                            y = nd.y[0];
                            z = nd.z[0];
                        } else {
                            double dx = scoef[j][R], dy = scoef[j][R + 1], dz = scoef[j][R + 2];
                            for (int r = 0; r < R; ++r) {
                                const double w = scoef[j][r];
                                dx += (double)modes[(int64_t)(0 * R + r) * N + n] * w;
                                dy += (double)modes[(int64_t)(1 * R + r) * N + n] * w;
                                dz += (double)modes[(int64_t)(2 * R + r) * N + n] * w;
                            }
                            x += dx;
                            y += dy;
                            z += dz;
                        }
                    }
                    nd.x[j] = x;
                    nd.y[j] = y;
                    nd.z[j] = z;
                }
                const double phi = pphi[n], r2 = pr2[n];
                nd.dwc = phi * (1.0 / kTwoPi);
                nd.r2 = r2;
                if constexpr (kPair) {
                    nd.xf[0] = (float)x0;
                    nd.xf[1] = (float)y0;
                    nd.xf[2] = PLANAR ? 0.0f : (float)z0;
                    nd.xf[3] = (float)fmax(fabs(x0), fmax(fabs(y0), PLANAR ? 0.0 : fabs(z0)));
                    far_node = fmax(fabs(x0), fmax(fabs(y0), PLANAR ? 0.0 : fabs(z0))) * pair_dk > 4.0;
                }
                for (int j = 0; j < Node::kCoef; ++j) {
                    if (SHARED_T) {
                        if (j > 0 && same_group) {
#pragma unroll
                            for (int v = 0; v < 2 * VB; ++v) nd.c[j][v] = nd.c[0][v];
                            continue;
                        }
                        // c = b exp(-t/T2) exp(i phi t)   (MRI.cpp:264-271); the
                        // mixed form reduces phi t / 2 pi in FP64 and takes the
                        // FP32 hardware sine, cosine and exp2
                        double fr, fi;
                        if constexpr (MIXED) {
                            const double u = phi * tg[j] * (1.0 / kTwoPi);
                            const float r = turns_f32(u);
                            const float d = __builtin_amdgcn_exp2f((float)(-tg[j] * r2 * 1.4426950408889634));
                            fr = (double)(d * __builtin_amdgcn_cosf(r));
                            fi = (double)(d * __builtin_amdgcn_sinf(r));
                        } else {
                            const double d = exp(-tg[j] * r2);
                            double so, co;
                            sincos(phi * tg[j], &so, &co);
                            fr = d * co;
                            fi = d * so;
                        }
#pragma unroll
                        for (int v = 0; v < VB; ++v) {
                            const double br = pb[(int64_t)(2 * (v0 + v)) * N + n];
                            const double bi = pb[(int64_t)(2 * (v0 + v) + 1) * N + n];
                            nd.c[j][2 * v] = (CT)(br * fr - bi * fi);
                            nd.c[j][2 * v + 1] = (CT)(br * fi + bi * fr);
                        }
                    } else {
#pragma unroll
                        for (int v = 0; v < VB; ++v) {
                            nd.c[j][2 * v] = (CT)pb[(int64_t)(2 * (v0 + v)) * N + n];
                            nd.c[j][2 * v + 1] = (CT)pb[(int64_t)(2 * (v0 + v) + 1) * N + n];
                        }
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < Node::kPos; ++j) nd.x[j] = nd.y[j] = nd.z[j] = 0.0;
                nd.dwc = nd.r2 = 0.0;
                if constexpr (kPair) nd.xf[0] = nd.xf[1] = nd.xf[2] = nd.xf[3] = 0.0f;
#pragma unroll
                for (int j = 0; j < Node::kCoef; ++j)
#pragma unroll
                    for (int v = 0; v < 2 * VB; ++v) nd.c[j][v] = (CT)0;
            }
        }
        bool pair_path = false;
        if constexpr (kPair) pair_path = !__syncthreads_or(far_node);
        else __syncthreads();
        const int cnt = (n_end - base < kDirectThreads) ? (int)(n_end - base) : kDirectThreads;

        CT acc_re[kDirectSPL][VB], acc_im[kDirectSPL][VB];
#pragma unroll
        for (int j = 0; j < kDirectSPL; ++j)
#pragma unroll
            for (int v = 0; v < VB; ++v) acc_re[j][v] = acc_im[j][v] = (CT)0;

        if constexpr (CONJ) {
            // one phase per pair; slot 1 takes the conjugate factor
#pragma unroll 2
            for (int m = 0; m < cnt; ++m) {
                const Node &nd = snode[m];
                const double u = PLANAR ? fma(kxs[0], nd.x[0], kys[0] * nd.y[0])
                                        : fma(kxs[0], nd.x[0], fma(kys[0], nd.y[0], kzs[0] * nd.z[0]));
                const float r = turns_f32(u);
                const CT sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);
#pragma unroll
                for (int v = 0; v < VB; ++v) {
                    const CT cr = nd.c[gq][2 * v], ci = nd.c[gq][2 * v + 1];
                    acc_re[0][v] = fma(cs, cr, fma(sn, ci, acc_re[0][v]));
                    acc_im[0][v] = fma(cs, ci, fma(-sn, cr, acc_im[0][v]));
                    acc_re[1][v] = fma(cs, cr, fma(-sn, ci, acc_re[1][v]));
                    acc_im[1][v] = fma(cs, ci, fma(sn, cr, acc_im[1][v]));
                }
            }
        } else if (pair_path) {
            // first slot: FP64 phase, range-reduced; second: plus dk.x in FP32
            if constexpr (kPair) {
#pragma unroll 2
            for (int m = 0; m < cnt; ++m) {
                const Node &nd = snode[m];
                const double u = PLANAR ? fma(kxs[0], nd.x[0], kys[0] * nd.y[0])
                                        : fma(kxs[0], nd.x[0], fma(kys[0], nd.y[0], kzs[0] * nd.z[0]));
                float r[2];
                r[0] = turns_f32(u);
                r[1] = turns_f32(r[0] + (PLANAR ? fmaf(dkx, nd.xf[0], dky * nd.xf[1])
                                                : fmaf(dkx, nd.xf[0], fmaf(dky, nd.xf[1], dkz * nd.xf[2]))));
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const CT sn = __builtin_amdgcn_sinf(r[j]), cs = __builtin_amdgcn_cosf(r[j]);
                    const int jc = j * GP + gq;
#pragma unroll
                    for (int v = 0; v < VB; ++v) {
                        const CT cr = nd.c[jc][2 * v], ci = nd.c[jc][2 * v + 1];
                        acc_re[j][v] = fma(cs, cr, fma(sn, ci, acc_re[j][v]));
                        acc_im[j][v] = fma(cs, ci, fma(-sn, cr, acc_im[j][v]));
                    }
                }
            }
            }
        } else {
#pragma unroll 2
        for (int m = 0; m < cnt; ++m) {
            const Node &nd = snode[m];
#pragma unroll
            for (int j = 0; j < kDirectSPL; ++j) {
                const int jp = MOTION ? j : 0;
                double u = PLANAR ? fma(kxs[j], nd.x[jp], kys[j] * nd.y[jp])
                                  : fma(kxs[j], nd.x[jp], fma(kys[j], nd.y[jp], kzs[j] * nd.z[jp]));
                if (!SHARED_T) u = fma(-nd.dwc, ts[j], u);  // phase - phi t, in cycles
                CT cs, sn;
                if (MIXED) {
                    const float rf = turns_f32(u);
                    sn = __builtin_amdgcn_sinf(rf);
                    cs = __builtin_amdgcn_cosf(rf);
                } else {
                    const double rr = u - __builtin_rint(u);
                    double s_, c_;
                    sincospi(2.0 * rr, &s_, &c_);
                    sn = (CT)s_;
                    cs = (CT)c_;
                }
                if (!SHARED_T) {
                    // exp(-t/T2) per (node, sample)
                    CT d;
                    if (MIXED) {
                        d = __builtin_amdgcn_exp2f((float)(-ts[j] * nd.r2 * 1.4426950408889634));
                    } else {
                        d = (CT)exp(-ts[j] * nd.r2);
                    }
                    cs *= d;
                    sn *= d;
                }
                // e = cs - i sn ; acc += e * c
                const int jc = SHARED_T ? j * GP + gq : 0;
#pragma unroll
                for (int v = 0; v < VB; ++v) {
                    const CT cr = nd.c[jc][2 * v], ci = nd.c[jc][2 * v + 1];
                    acc_re[j][v] = fma(cs, cr, fma(sn, ci, acc_re[j][v]));
                    acc_im[j][v] = fma(cs, ci, fma(-sn, cr, acc_im[j][v]));
                }
            }
        }
        }
#pragma unroll
        for (int j = 0; j < kDirectSPL; ++j)
#pragma unroll
            for (int v = 0; v < VB; ++v) {
                tot_re[j][v] += (double)acc_re[j][v];
                tot_im[j][v] += (double)acc_im[j][v];
            }
    }

#pragma unroll
    for (int j = 0; j < kDirectSPL; ++j) {
        if (!valid[j]) continue;
        const int64_t row = write_final ? (int64_t)sid[j] : split * S + sid[j];
#pragma unroll
        for (int v = 0; v < VB; ++v) dst[row * V + v0 + v] = make_double2(tot_re[j][v], tot_im[j][v]);
    }
}

// Mixed form, static nodes, readout runs.  Lane l of a workgroup owns run
// blockIdx.x * 256 + l (kRunLen samples of one readout line).  Per staged
// node the lane forms
//   e0 = exp(-t0/T2) exp(-2 pi i (k0.x - phi t0 / 2 pi))      (FP64 phase,
//        range-reduced, FP32 hardware sine/cosine)
//   q  = exp(-dt/T2) exp(-2 pi i (dk.x - phi dt / 2 pi))       (FP32)
// and accumulates e0 q^m c for m = 0 .. kRunLen-1, advancing e (one
// encoding: y = e c) by one complex multiply per sample.  FP32 partial sums over a 256-node chunk are
// added into FP64, as in k_signal_direct.
template <int VB>
struct alignas(16) RunNode {
    double x, y, z, dwc, r2;
    float xf, yf, zf, dwcf, r2f;
    float c[2 * VB];
};

// SHQ: the tile's runs share at most kRunSteps steps (dk, dt); the staging
// threads form each node's step factor once per step into LDS, and the lanes
// read it instead of forming their own.
constexpr int kSqStride = kDirectThreads + 1;  // LDS row of step factors (+1 entry)

template <int VB, bool SHQ = false>
__global__ __launch_bounds__(kDirectThreads) void k_signal_runs(
    const double *__restrict__ packed, int64_t N, int V, int v0, const RunDesc *__restrict__ runs,
    int64_t n_runs, int64_t stride, int64_t nodes_per_split, double2 *__restrict__ dst, int64_t S,
    int write_final, const double *__restrict__ run_steps = nullptr, const int32_t *__restrict__ run_nsteps = nullptr)
{
    __shared__ RunNode<VB> snode[kDirectThreads];
    extern __shared__ float2 sq32[];  // SHQ: (step, node) factors, rows of kSqStride
    const int nst = SHQ ? run_nsteps[blockIdx.x] : 0;
    const int64_t split = blockIdx.y;
    const int64_t n_begin = split * nodes_per_split;
    const int64_t n_end = (n_begin + nodes_per_split < N) ? n_begin + nodes_per_split : N;
    const int tid = threadIdx.x;
    const int64_t rid = (int64_t)blockIdx.x * kDirectThreads + tid;
    const bool valid = rid < n_runs;
    const RunDesc rd = runs[valid ? rid : 0];
    const float dkx = (float)rd.dkx, dky = (float)rd.dky, dkz = (float)rd.dkz, dtf = (float)rd.dt;

    const double *px = packed;
    const double *py = packed + N;
    const double *pz = packed + 2 * N;
    const double *pphi = packed + 3 * N;
    const double *pr2 = packed + 4 * N;
    const double *pb = packed + 5 * N;

    double tot_re[kRunLen][VB], tot_im[kRunLen][VB];
#pragma unroll
    for (int m = 0; m < kRunLen; ++m)
#pragma unroll
        for (int v = 0; v < VB; ++v) tot_re[m][v] = tot_im[m][v] = 0.0;

    for (int64_t base = n_begin; base < n_end; base += kDirectThreads) {
        __syncthreads();
        {
            const int64_t n = base + tid;
            RunNode<VB> nd;
            if (n < n_end) {
                nd.x = px[n];
                nd.y = py[n];
                nd.z = pz[n];
                nd.dwc = pphi[n] * (1.0 / kTwoPi);
                nd.r2 = pr2[n];
#pragma unroll
                for (int v = 0; v < VB; ++v) {
                    nd.c[2 * v] = (float)pb[(int64_t)(2 * (v0 + v)) * N + n];
                    nd.c[2 * v + 1] = (float)pb[(int64_t)(2 * (v0 + v) + 1) * N + n];
                }
            } else {
                nd.x = nd.y = nd.z = nd.dwc = nd.r2 = 0.0;
#pragma unroll
                for (int v = 0; v < 2 * VB; ++v) nd.c[v] = 0.0f;
            }
            nd.xf = (float)nd.x;
            nd.yf = (float)nd.y;
            nd.zf = (float)nd.z;
            nd.dwcf = (float)nd.dwc;
            nd.r2f = (float)nd.r2;
            snode[tid] = nd;
            if constexpr (SHQ) {
                for (int j = 0; j < nst; ++j) {
                    const double *st = &run_steps[((int64_t)blockIdx.x * kRunSteps + j) * 4];
                    const float sdkx = (float)st[0], sdky = (float)st[1], sdkz = (float)st[2], sdt = (float)st[3];
                    const float du = turns_f32(fmaf(sdkx, nd.xf, fmaf(sdky, nd.yf, fmaf(sdkz, nd.zf, -nd.dwcf * sdt))));
                    const float ds = __builtin_amdgcn_exp2f(-sdt * nd.r2f * 1.44269504f);
                    sq32[j * kSqStride + tid] = make_float2(ds * __builtin_amdgcn_cosf(du), -ds * __builtin_amdgcn_sinf(du));
                }
            }
        }
        __syncthreads();
        const int cnt = (n_end - base < kDirectThreads) ? (int)(n_end - base) : kDirectThreads;

        float acc_re[kRunLen][VB], acc_im[kRunLen][VB];
#pragma unroll
        for (int m = 0; m < kRunLen; ++m)
#pragma unroll
            for (int v = 0; v < VB; ++v) acc_re[m][v] = acc_im[m][v] = 0.0f;

        for (int i = 0; i < cnt; ++i) {
            const RunNode<VB> &nd = snode[i];
            // anchor: FP64 phase in cycles, range-reduced
            const double u = fma(rd.k0x, nd.x, fma(rd.k0y, nd.y, fma(rd.k0z, nd.z, -nd.dwc * rd.t0)));
            const float r = turns_f32(u);
            const float d0 = __builtin_amdgcn_exp2f((float)(-rd.t0 * nd.r2) * 1.44269504f);
            float er = d0 * __builtin_amdgcn_cosf(r);
            float ei = -d0 * __builtin_amdgcn_sinf(r);
            // step factor
            float qr, qi;
            if constexpr (SHQ) {
                const float2 q = sq32[rd.step * kSqStride + i];
                qr = q.x;
                qi = q.y;
            } else {
                const float du = turns_f32(fmaf(dkx, nd.xf, fmaf(dky, nd.yf, fmaf(dkz, nd.zf, -nd.dwcf * dtf))));
                const float ds = __builtin_amdgcn_exp2f(-dtf * nd.r2f * 1.44269504f);
                qr = ds * __builtin_amdgcn_cosf(du);
                qi = -ds * __builtin_amdgcn_sinf(du);
            }
            if constexpr (VB == 1) {
                // y = e c advances by q: one complex multiply and one complex
                // add per sample
                float yr = fmaf(er, nd.c[0], -ei * nd.c[1]);
                float yi = fmaf(er, nd.c[1], ei * nd.c[0]);
#pragma unroll
                for (int m = 0; m < kRunLen; ++m) {
                    acc_re[m][0] += yr;
                    acc_im[m][0] += yi;
                    const float nr = fmaf(yr, qr, -yi * qi);
                    yi = fmaf(yr, qi, yi * qr);
                    yr = nr;
                }
            } else {
#pragma unroll
                for (int m = 0; m < kRunLen; ++m) {
#pragma unroll
                    for (int v = 0; v < VB; ++v) {
                        const float cr = nd.c[2 * v], ci = nd.c[2 * v + 1];
                        acc_re[m][v] = fmaf(er, cr, fmaf(-ei, ci, acc_re[m][v]));
                        acc_im[m][v] = fmaf(er, ci, fmaf(ei, cr, acc_im[m][v]));
                    }
                    const float nr = er * qr - ei * qi;
                    ei = fmaf(er, qi, ei * qr);
                    er = nr;
                }
            }
        }
#pragma unroll
        for (int m = 0; m < kRunLen; ++m)
#pragma unroll
            for (int v = 0; v < VB; ++v) {
                tot_re[m][v] += (double)acc_re[m][v];
                tot_im[m][v] += (double)acc_im[m][v];
            }
    }

    if (!valid) return;
#pragma unroll
    for (int m = 0; m < kRunLen; ++m) {
        if (m >= rd.len) continue;
        const int64_t sid = (int64_t)rd.s0 + m * stride;
        const int64_t row = write_final ? sid : split * S + sid;
#pragma unroll
        for (int v = 0; v < VB; ++v) dst[row * V + v0 + v] = make_double2(tot_re[m][v], tot_im[m][v]);
    }
}

// FP64 form, static nodes, readout runs: the same recurrence in FP64
// (FP64 phase and sincospi for the anchor and the step, FP64 products and
// accumulators).  Taken only when every run is linear to 1e-14 of the
// largest |k'| and |t|; 15 FP64 complex products then leave the result
// ~1e-15 from the per-sample sum of MRI.cpp:260-288.
template <int VB>
struct alignas(16) RunNode64 {
    double x, y, z, dwc, r2, pad;
    double c[2 * VB];
};

template <int VB, bool SHQ = false>
__global__ __launch_bounds__(kDirectThreads) void k_signal_runs_f64(
    const double *__restrict__ packed, int64_t N, int V, int v0, const RunDesc *__restrict__ runs,
    int64_t n_runs, int64_t stride, int64_t nodes_per_split, double2 *__restrict__ dst, int64_t S,
    int write_final, const double *__restrict__ run_steps = nullptr, const int32_t *__restrict__ run_nsteps = nullptr)
{
    __shared__ RunNode64<VB> snode[kDirectThreads];
    extern __shared__ double2 sq64[];  // SHQ: (step, node) factors, rows of kSqStride
    const int nst = SHQ ? run_nsteps[blockIdx.x] : 0;
    const int64_t split = blockIdx.y;
    const int64_t n_begin = split * nodes_per_split;
    const int64_t n_end = (n_begin + nodes_per_split < N) ? n_begin + nodes_per_split : N;
    const int tid = threadIdx.x;
    const int64_t rid = (int64_t)blockIdx.x * kDirectThreads + tid;
    const bool valid = rid < n_runs;
    const RunDesc rd = runs[valid ? rid : 0];

    double tot_re[kRunLen][VB], tot_im[kRunLen][VB];
#pragma unroll
    for (int m = 0; m < kRunLen; ++m)
#pragma unroll
        for (int v = 0; v < VB; ++v) tot_re[m][v] = tot_im[m][v] = 0.0;

    for (int64_t base = n_begin; base < n_end; base += kDirectThreads) {
        __syncthreads();
        {
            const int64_t n = base + tid;
            RunNode64<VB> nd;
            nd.pad = 0.0;
            if (n < n_end) {
                nd.x = packed[n];
                nd.y = packed[N + n];
                nd.z = packed[2 * N + n];
                nd.dwc = packed[3 * N + n] * (1.0 / kTwoPi);
                nd.r2 = packed[4 * N + n];
#pragma unroll
                for (int v = 0; v < VB; ++v) {
                    nd.c[2 * v] = packed[(int64_t)(5 + 2 * (v0 + v)) * N + n];
                    nd.c[2 * v + 1] = packed[(int64_t)(6 + 2 * (v0 + v)) * N + n];
                }
            } else {
                nd.x = nd.y = nd.z = nd.dwc = nd.r2 = 0.0;
#pragma unroll
                for (int v = 0; v < 2 * VB; ++v) nd.c[v] = 0.0;
            }
            snode[tid] = nd;
            if constexpr (SHQ) {
                for (int j = 0; j < nst; ++j) {
                    const double *st = &run_steps[((int64_t)blockIdx.x * kRunSteps + j) * 4];
                    double du = fma(st[0], nd.x, fma(st[1], nd.y, fma(st[2], nd.z, -nd.dwc * st[3])));
                    du -= __builtin_rint(du);
                    double qs, qc;
                    sincospi(2.0 * du, &qs, &qc);
                    const double ds = exp(-st[3] * nd.r2);
                    sq64[j * kSqStride + tid] = make_double2(ds * qc, -ds * qs);
                }
            }
        }
        __syncthreads();
        const int cnt = (n_end - base < kDirectThreads) ? (int)(n_end - base) : kDirectThreads;
        for (int i = 0; i < cnt; ++i) {
            const RunNode64<VB> &nd = snode[i];
            // anchor: the first sample's term, as the per-sample sum forms it
            const double u = fma(rd.k0x, nd.x, fma(rd.k0y, nd.y, fma(rd.k0z, nd.z, -nd.dwc * rd.t0)));
            const double r = u - __builtin_rint(u);
            double sn, cs;
            sincospi(2.0 * r, &sn, &cs);
            const double d0 = exp(-rd.t0 * nd.r2);
            double er = d0 * cs, ei = -d0 * sn;
            // step factor
            double qr, qi;
            if constexpr (SHQ) {
                const double2 q = sq64[rd.step * kSqStride + i];
                qr = q.x;
                qi = q.y;
            } else {
                double du = fma(rd.dkx, nd.x, fma(rd.dky, nd.y, fma(rd.dkz, nd.z, -nd.dwc * rd.dt)));
                du -= __builtin_rint(du);
                double qs, qc;
                sincospi(2.0 * du, &qs, &qc);
                const double ds = exp(-rd.dt * nd.r2);
                qr = ds * qc;
                qi = -ds * qs;
            }
            if constexpr (VB == 1) {
                // y = e c advances by q, as in k_signal_runs
                double yr = fma(er, nd.c[0], -ei * nd.c[1]);
                double yi = fma(er, nd.c[1], ei * nd.c[0]);
#pragma unroll
                for (int m = 0; m < kRunLen; ++m) {
                    tot_re[m][0] += yr;
                    tot_im[m][0] += yi;
                    const double nr = fma(yr, qr, -yi * qi);
                    yi = fma(yr, qi, yi * qr);
                    yr = nr;
                }
            } else {
#pragma unroll
                for (int m = 0; m < kRunLen; ++m) {
#pragma unroll
                    for (int v = 0; v < VB; ++v) {
                        const double cr = nd.c[2 * v], ci = nd.c[2 * v + 1];
                        tot_re[m][v] = fma(er, cr, fma(-ei, ci, tot_re[m][v]));
                        tot_im[m][v] = fma(er, ci, fma(ei, cr, tot_im[m][v]));
                    }
                    const double nr = er * qr - ei * qi;
                    ei = fma(er, qi, ei * qr);
                    er = nr;
                }
            }
        }
    }

    if (!valid) return;
#pragma unroll
    for (int m = 0; m < kRunLen; ++m) {
        if (m >= rd.len) continue;
        const int64_t sid = (int64_t)rd.s0 + m * stride;
        const int64_t row = write_final ? sid : split * S + sid;
#pragma unroll
        for (int v = 0; v < VB; ++v) dst[row * V + v0 + v] = make_double2(tot_re[m][v], tot_im[m][v]);
    }
}

// out[s, v] = sum over splits, in split order.
__global__ void k_reduce_splits(const double2 *__restrict__ part, int64_t n_splits, int64_t SV,
                                double2 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= SV) return;
    double re = 0.0, im = 0.0;
    for (int64_t s = 0; s < n_splits; ++s) {
        const double2 p = part[s * SV + i];
        re += p.x;
        im += p.y;
    }
    out[i] = make_double2(re, im);
}

template <typename T>
__global__ void k_nodes_pack(int64_t N, int V, const T *__restrict__ r0, const T *__restrict__ phi,
                             const T *__restrict__ T2, const T *__restrict__ mxy,
                             const double *__restrict__ w, double *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = (double)r0[3 * n];
    out[N + n] = (double)r0[3 * n + 1];
    out[2 * N + n] = (double)r0[3 * n + 2];
    out[3 * N + n] = (double)phi[n];
    out[4 * N + n] = 1.0 / (double)T2[n];
    const double ww = w ? w[n] : 1.0;
    for (int v = 0; v < V; ++v) {
        out[(int64_t)(5 + 2 * v) * N + n] = ww * (double)mxy[2 * (n * V + v)];
        out[(int64_t)(6 + 2 * v) * N + n] = ww * (double)mxy[2 * (n * V + v) + 1];
    }
}

// CSR SpMV, G lanes per row (G a power of two, at most 64): lane j of a
// row's group adds the row's terms j, j + G, ... for up to 4 encodings at a
// time (the row's values and column indices read once per pass, coalesced
// across the group), then the group sums its lanes in a fixed butterfly, so
// a result does not depend on the launch.  Groups past the last row work on
// the last row and write nothing.
template <int G>
__global__ __launch_bounds__(256) void k_csr_spmv_c128(int64_t n_rows, int V, const int64_t *__restrict__ indptr,
                                                        const int64_t *__restrict__ indices,
                                                        const double *__restrict__ data,
                                                        const double2 *__restrict__ x, double2 *__restrict__ y)
{
    const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int j = threadIdx.x % G;
    const bool live = g < n_rows;
    const int64_t r = live ? g : n_rows - 1;
    const int64_t k0 = indptr[r], k1 = indptr[r + 1];
    for (int v0 = 0; v0 < V; v0 += 4) {
        double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
        for (int64_t k = k0 + j; k < k1; k += G) {
            const double a = data[k];
            const double2 *xr = x + indices[k] * V + v0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (v0 + u < V) {
                    const double2 xv = xr[u];
                    re[u] = fma(a, xv.x, re[u]);
                    im[u] = fma(a, xv.y, im[u]);
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                re[u] += __shfl_xor(re[u], off, G);
                im[u] += __shfl_xor(im[u], off, G);
            }
        if (live && j == 0) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (v0 + u < V) y[r * V + v0 + u] = make_double2(re[u], im[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
struct DirectLaunch {
    int64_t n_splits;
    int64_t nodes_per_split;
};

// Tilings of the direct path: 0 consecutive, 1 t-grouped, 2 readout runs,
// 3 packed t-groups, 4 conjugate pairs in packed t-groups (mixed form).
static int64_t direct_tiles(const fm_signal_plan *p, int tiling)
{
    if (tiling == 4) return p->n_tiles_conj;
    if (tiling == 3) return p->n_tiles_packed;
    if (tiling == 2) return cdiv(p->n_runs, kDirectThreads);
    return tiling == 1 ? p->n_tiles_grouped : p->n_tiles_consec;
}

static DirectLaunch direct_launch_shape(const fm_signal_plan *p, int64_t N, int tiling)
{
    // Aim for ~8 workgroups per CU over the whole grid; keep >= 2048 nodes
    // per split so the staging cost stays amortised.
    const int64_t target = 8 * (int64_t)device_cu_count();
    const int64_t nt = direct_tiles(p, tiling);
    int64_t splits = cdiv(target, nt > 0 ? nt : 1);
    const int64_t max_splits = cdiv(N, 2048);
    if (splits > max_splits) splits = max_splits;
    if (splits < 1) splits = 1;
    int64_t nps = cdiv(N, splits);
    nps = cdiv(nps, kDirectThreads) * kDirectThreads;
    splits = cdiv(N, nps);
    return {splits, nps};
}

template <int VB, bool MIXED>
static void launch_direct_vb(const fm_signal_plan *p, bool grouped, bool motion, dim3 grid,
                             const double *packed, int64_t N, int V, int v0, const float *modes,
                             int R, const double *mcoef, int64_t nps, double2 *dst, int write_final,
                             hipStream_t st)
{
    if (grouped) {
        if (motion)
            k_signal_direct<VB, MIXED, true, true><<<grid, kDirectThreads, 0, st>>>(
                packed, N, V, v0, p->d_ks, p->d_perm, p->d_tiles_grouped, p->d_tg_t, modes, R, mcoef,
                nps, dst, p->n_samples, write_final, 0.0);
        else
            k_signal_direct<VB, MIXED, true, false><<<grid, kDirectThreads, 0, st>>>(
                packed, N, V, v0, p->d_ks, p->d_perm, p->d_tiles_grouped, p->d_tg_t, modes, R, mcoef,
                nps, dst, p->n_samples, write_final, 0.0);
    } else {
        k_signal_direct<VB, MIXED, false, false><<<grid, kDirectThreads, 0, st>>>(
            packed, N, V, v0, p->d_ks, p->d_perm, p->d_tiles_consec, p->d_tg_t, modes, R, mcoef, nps,
            dst, p->n_samples, write_final, 0.0);
    }
}

template <bool MIXED>
static void launch_direct(const fm_signal_plan *p, bool grouped, bool motion, dim3 grid,
                          const double *packed, int64_t N, int V, int v0, int vb, const float *modes,
                          int R, const double *mcoef, int64_t nps, double2 *dst, int write_final,
                          hipStream_t st)
{
#define FM_LD(VBV) launch_direct_vb<VBV, MIXED>(p, grouped, motion, grid, packed, N, V, v0, modes, R, mcoef, nps, dst, write_final, st)
    switch (vb) {
    case 1: FM_LD(1); break;
    case 2: FM_LD(2); break;
    case 3: FM_LD(3); break;
    default: FM_LD(4); break;
    }
#undef FM_LD
}

template <bool MIXED>
static void launch_packed(const fm_signal_plan *p, dim3 grid, const double *packed, int64_t N, int V, int v0,
                          int64_t nps, double2 *dst, int write_final, hipStream_t st)
{
#define FM_LP(GPV, PL)                                                                                      \
    k_signal_direct<1, MIXED, true, false, GPV, PL><<<grid, kDirectThreads, 0, st>>>(                      \
        packed, N, V, v0, p->d_ks, p->d_perm, p->d_tiles_packed, p->d_tg_t, nullptr, 0, nullptr, nps, dst, \
        p->n_samples, write_final, p->packed_dk)
    if (p->planar) {
        switch (p->group_pack) {
        case 2: FM_LP(2, true); break;
        case 4: FM_LP(4, true); break;
        default: FM_LP(8, true); break;
        }
    } else {
        switch (p->group_pack) {
        case 2: FM_LP(2, false); break;
        case 4: FM_LP(4, false); break;
        default: FM_LP(8, false); break;
        }
    }
#undef FM_LP
}

static void launch_conj(const fm_signal_plan *p, dim3 grid, const double *packed, int64_t N, int V, int v0,
                        int64_t nps, double2 *dst, int write_final, hipStream_t st)
{
#define FM_LQ(GPV, PL)                                                                                          \
    k_signal_direct<1, true, true, false, GPV, PL, true><<<grid, kDirectThreads, 0, st>>>(                     \
        packed, N, V, v0, p->d_ks, p->d_perm_conj, p->d_tiles_conj, p->d_tg_t, nullptr, 0, nullptr, nps, dst, \
        p->n_samples, write_final, 0.0)
    if (p->planar) {
        switch (p->conj_pack) {
        case 4: FM_LQ(4, true); break;
        case 8: FM_LQ(8, true); break;
        default: FM_LQ(16, true); break;
        }
    } else {
        switch (p->conj_pack) {
        case 4: FM_LQ(4, false); break;
        case 8: FM_LQ(8, false); break;
        default: FM_LQ(16, false); break;
        }
    }
#undef FM_LQ
}

static size_t direct_workspace(const fm_signal_plan *p, int64_t N, int V)
{
    if (N <= 0 || p->n_samples == 0) return 0;
    // the largest of the tilings' needs (motion forces the grouped one) and
    // of the per-line GEMM's partial samples (one encoding at a time)
    size_t need = lines_workspace(p, N);
    for (int g = 0; g < 5; ++g) {
        if (g == 2 && p->n_runs == 0) continue;
        if (g == 3 && p->group_pack <= 1) continue;
        if (g == 4 && p->conj_pack <= 1) continue;
        const DirectLaunch L = direct_launch_shape(p, N, g);
        if (L.n_splits > 1)
            need = std::max(need, (size_t)L.n_splits * (size_t)p->n_samples * (size_t)V * sizeof(double2));
    }
    return need;
}

static int direct_execute(const fm_signal_plan *p, int precision, int64_t N, int V,
                          const double *packed, const float *modes, int R, const double *mcoef,
                          void *ws, size_t ws_bytes, double *out, hipStream_t st)
{
    const int64_t S = p->n_samples;
    if (S == 0) return FM_OK;
    if (N == 0) {
        return zero_f64_async(out, 2 * S * V, st);
    }
    const bool motion = modes != nullptr && R > 0;
    // mixed form, static nodes, whole linear lines: the per-line GEMM
    if (!motion && precision != FM_PREC_F64 && lines_usable(p, N)) return lines_execute(p, N, V, packed, ws, ws_bytes, out, st);
    // positions depend on t, so moving nodes always take the t-grouped tiles;
    // static nodes take the readout runs when the plan has them (the FP64
    // form only when they are linear to 1e-14)
    const bool use_runs = !motion && p->n_runs > 0 && (precision != FM_PREC_F64 || p->runs_exact);
    // static nodes with small equal-size t-groups: packed slots
    const bool use_packed = !motion && !use_runs && p->group_pack > 1;
    // mixed form: conjugate pairs when every packed group pairs up
    const bool use_conj = use_packed && precision != FM_PREC_F64 && p->conj_pack > 1;
    const bool grouped = motion || p->prefer_grouped;
    const int tiling = use_runs ? 2 : (use_conj ? 4 : (use_packed ? 3 : (grouped ? 1 : 0)));
    const DirectLaunch L = direct_launch_shape(p, N, tiling);
    const size_t need = L.n_splits > 1 ? (size_t)L.n_splits * (size_t)S * (size_t)V * sizeof(double2) : 0;
    if (need > ws_bytes) {
        set_error("signal workspace too small: need %zu bytes, got %zu", need, ws_bytes);
        return FM_ERR_WORKSPACE;
    }
    double2 *dst = L.n_splits > 1 ? reinterpret_cast<double2 *>(ws) : reinterpret_cast<double2 *>(out);
    const int write_final = L.n_splits > 1 ? 0 : 1;
    dim3 grid((unsigned)direct_tiles(p, tiling), (unsigned)L.n_splits);
    plan_timing_start(p, st);
    if (use_runs) {
        // one encoding per launch: 16 samples x 1 encoding of accumulators
        // keeps the kernel at ~130 VGPRs (3 waves per SIMD)
        for (int v0 = 0; v0 < V; ++v0) {
            if (precision == FM_PREC_F64)
                if (p->run_steps > 0)
                    k_signal_runs_f64<1, true><<<grid, kDirectThreads, sizeof(double2) * kSqStride * p->run_steps, st>>>(
                        packed, N, V, v0, p->d_runs, p->n_runs, p->run_stride, L.nodes_per_split, dst, S, write_final,
                        p->d_run_steps, p->d_run_nsteps);
                else
                    k_signal_runs_f64<1><<<grid, kDirectThreads, 0, st>>>(packed, N, V, v0, p->d_runs, p->n_runs,
                                                                          p->run_stride, L.nodes_per_split, dst, S,
                                                                          write_final);
            else if (p->run_steps > 0)
                k_signal_runs<1, true><<<grid, kDirectThreads, sizeof(float2) * kSqStride * p->run_steps, st>>>(
                    packed, N, V, v0, p->d_runs, p->n_runs, p->run_stride, L.nodes_per_split, dst, S, write_final,
                    p->d_run_steps, p->d_run_nsteps);
            else
                k_signal_runs<1><<<grid, kDirectThreads, 0, st>>>(packed, N, V, v0, p->d_runs, p->n_runs,
                                                                  p->run_stride, L.nodes_per_split, dst, S,
                                                                  write_final);
            FM_LAUNCH_CHECK("k_signal_runs");
        }
    }
    if (use_packed) {
        // one encoding per launch: the per-(slot, group) coefficients of a
        // staged node take 2 * group_pack complex values per encoding
        for (int v0 = 0; v0 < V; ++v0) {
            if (precision == FM_PREC_F64)
                launch_packed<false>(p, grid, packed, N, V, v0, L.nodes_per_split, dst, write_final, st);
            else if (use_conj)
                launch_conj(p, grid, packed, N, V, v0, L.nodes_per_split, dst, write_final, st);
            else
                launch_packed<true>(p, grid, packed, N, V, v0, L.nodes_per_split, dst, write_final, st);
            FM_LAUNCH_CHECK("k_signal_direct (packed t-groups)");
        }
    }
    for (int v0 = 0; v0 < V && !use_runs && !use_packed; v0 += kMaxVB) {
        const int vb = (V - v0 < kMaxVB) ? V - v0 : kMaxVB;
        if (precision == FM_PREC_F64)
            launch_direct<false>(p, grouped, motion, grid, packed, N, V, v0, vb, modes, R, mcoef, L.nodes_per_split, dst, write_final, st);
        else
            launch_direct<true>(p, grouped, motion, grid, packed, N, V, v0, vb, modes, R, mcoef, L.nodes_per_split, dst, write_final, st);
        FM_LAUNCH_CHECK("k_signal_direct");
    }
    plan_timing_stop(p, st);
    if (L.n_splits > 1) {
        const int64_t SV = S * V;
        k_reduce_splits<<<(unsigned)cdiv(SV, 256), 256, 0, st>>>(dst, L.n_splits, SV,
                                                               reinterpret_cast<double2 *>(out));
        FM_LAUNCH_CHECK("k_reduce_splits");
    }
    return FM_OK;
}

}  // namespace fm

using namespace fm;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace fm {
// Host tables of a plan before upload.
struct PlanHost {
    std::vector<int32_t> perm;
    std::vector<SignalTile> tiles_g, tiles_c, tiles_p, tiles_q;
    std::vector<int32_t> perm_conj;
    std::vector<double4> ks;
    std::vector<RunDesc> runs;
    std::vector<RunDesc> lines;       // whole linear lines (per-line GEMM)
    std::vector<double> run_steps;    // (tiles, kRunSteps, 4)
    std::vector<int32_t> run_nsteps;  // (tiles)
};

// Host analysis of a trajectory: distinct-t groups, Cartesian
// factorisation, tilings, readout runs.  No device calls.
static int plan_analyse(fm_signal_plan *p, const double *kx, const double *ky, const double *kz,
                        const double *t, int method, PlanHost &h)
{
    const int64_t S = p->n_samples, n_meas = p->n_meas, n_lines = p->n_lines, n_kz = p->n_kz;
    // distinct t in first-appearance order (equal values share decay/phase)
    std::unordered_map<uint64_t, int32_t> gid;
    p->sample_group.resize((size_t)S);
    for (int64_t s = 0; s < S; ++s) {
        double tv = t[s];
        if (tv == 0.0) tv = 0.0;  // -0.0 and 0.0 compare equal in the reference's cache
        uint64_t key;
        std::memcpy(&key, &tv, sizeof key);
        auto it = gid.find(key);
        int32_t g;
        if (it == gid.end()) {
            g = (int32_t)p->tg_t.size();
            gid.emplace(key, g);
            p->tg_t.push_back(t[s]);
        } else {
            g = it->second;
        }
        p->sample_group[(size_t)s] = g;
    }
    p->n_tgroups = (int64_t)p->tg_t.size();

    if (method != FM_METHOD_DIRECT && S > 0) {
        const bool ok = cartesian_analyse(p, kx, ky, kz, t);
        if (method == FM_METHOD_CARTESIAN && !ok) {
            set_error("k-space sampling does not factorise as a Cartesian grid "
                      "(k not additively separable per axis, or t not a function of the readout index)");
            return FM_ERR_UNSUPPORTED;
        }
        p->method = ok ? FM_METHOD_CARTESIAN : FM_METHOD_DIRECT;
    }

    // Tilings.  Grouped: samples ordered by t-group, each tile within one
    // group, so the t-dependent node factors are staged once per tile.
    // Consecutive: plain runs of samples, per-sample t.  Static nodes prefer
    // the grouped tiles when groups fill most of a 512-sample tile.
    std::vector<int32_t> &perm = h.perm;
    perm.resize((size_t)S);
    std::vector<SignalTile> &tiles_g = h.tiles_g, &tiles_c = h.tiles_c;
    {
        std::vector<int64_t> count((size_t)p->n_tgroups + 1, 0);
        for (int64_t s = 0; s < S; ++s) count[(size_t)p->sample_group[(size_t)s] + 1]++;
        for (int64_t g = 0; g < p->n_tgroups; ++g) count[(size_t)g + 1] += count[(size_t)g];
        std::vector<int64_t> fill(count.begin(), count.end() - 1);
        for (int64_t s = 0; s < S; ++s) perm[(size_t)fill[(size_t)p->sample_group[(size_t)s]]++] = (int32_t)s;
        // grouped: slot records of up to 256 samples of one group, packed
        // kDirectSPL to a tile (slots of one tile may hold different groups)
        for (int64_t g = 0; g < p->n_tgroups; ++g) {
            for (int64_t o = count[(size_t)g]; o < count[(size_t)g + 1]; o += kDirectThreads) {
                const int64_t c = std::min<int64_t>(kDirectThreads, count[(size_t)g + 1] - o);
                tiles_g.push_back({(int32_t)g, (int32_t)o, (int32_t)c, 0});
            }
        }
        while (tiles_g.size() % kDirectSPL) tiles_g.push_back({0, 0, 0, 0});
        // packed slots: every t-group has the same size g in {32, 64, 128}
        // (e.g. the interleaves of a spiral at one readout index), so a slot
        // takes 256 / g whole groups
        const int64_t g0 = p->n_tgroups > 0 ? count[1] - count[0] : 0;
        bool uniform = g0 == 32 || g0 == 64 || g0 == 128;
        for (int64_t g = 0; g < p->n_tgroups && uniform; ++g) uniform = count[(size_t)g + 1] - count[(size_t)g] == g0;
        if (uniform) {
            const int64_t gp = kDirectThreads / g0;
            for (int64_t g = 0; g < p->n_tgroups; g += gp) {
                const int64_t ng = std::min<int64_t>(gp, p->n_tgroups - g);
                h.tiles_p.push_back({(int32_t)g, (int32_t)count[(size_t)g], (int32_t)(ng * g0), 0});
            }
            while (h.tiles_p.size() % kDirectSPL) h.tiles_p.push_back({0, 0, 0, 0});
            p->group_pack = (int)gp;
            p->n_tiles_packed = (int64_t)h.tiles_p.size() / kDirectSPL;
        }
        for (int64_t o = 0; o < S; o += kDirectTile)
            tiles_c.push_back({-1, (int32_t)o, (int32_t)std::min<int64_t>(kDirectTile, S - o), 0});
    }
    const double avg_group = p->n_tgroups ? (double)S / (double)p->n_tgroups : 0.0;
    p->prefer_grouped = avg_group >= 64.0;
    p->n_tiles_grouped = (int64_t)tiles_g.size() / kDirectSPL;
    p->n_tiles_consec = (int64_t)tiles_c.size();

    std::vector<double4> &ks = h.ks;
    ks.resize((size_t)S);
    for (int64_t s = 0; s < S; ++s)
        ks[(size_t)s] = make_double4(kx[s] * kPiFOverPi, ky[s] * kPiFOverPi, kz[s] * kPiFOverPi, t[s]);
    p->planar = 1;
    for (int64_t s = 0; s < S && p->planar; ++s) p->planar = ks[(size_t)s].z == 0.0;
    // the two samples of a lane in a packed tile (same line, group_pack
    // readout indices apart): the kernel may form the second one's phase as
    // the first one's reduced phase plus dk.x in FP32 when |dk|_1 |x|_inf
    // stays small; record the largest |dk|_1
    p->packed_dk = 0.0;
    for (size_t b = 0; b + 1 < h.tiles_p.size(); b += kDirectSPL) {
        const SignalTile &s0 = h.tiles_p[b], &s1 = h.tiles_p[b + 1];
        const int32_t c = std::min(s0.count, s1.count);
        for (int32_t l = 0; l < c; ++l) {
            const double4 a = ks[(size_t)perm[(size_t)(s0.offset + l)]];
            const double4 q = ks[(size_t)perm[(size_t)(s1.offset + l)]];
            p->packed_dk = std::max(p->packed_dk, std::fabs(q.x - a.x) + std::fabs(q.y - a.y) + std::fabs(q.z - a.z));
        }
    }

    // Conjugate pairs (mixed form): in every packed t-group pair each sample
    // with one whose k' is its negative, to 1e-12 of the largest |k'| (an
    // O(g0^2) search per group of g0 <= 128 samples; k' = 0 pairs with
    // another k' = 0).  All groups or none.
    p->conj_pack = 1;
    p->n_tiles_conj = 0;
    p->conj_dk = 0.0;
    if (p->group_pack > 1 && S > 0) {
        const int64_t g0 = kDirectThreads / p->group_pack;
        double kmax = 0.0;
        for (int64_t s = 0; s < S; ++s)
            kmax = std::max({kmax, std::fabs(ks[(size_t)s].x), std::fabs(ks[(size_t)s].y), std::fabs(ks[(size_t)s].z)});
        const double tol = 1e-12 * (kmax > 0.0 ? kmax : 1.0);
        std::vector<int32_t> first((size_t)S / 2), second((size_t)S / 2);  // per group: g0/2 pairs
        bool ok = g0 % 2 == 0;
        double worst = 0.0;
        std::vector<char> used((size_t)g0);
        for (int64_t g = 0; g < p->n_tgroups && ok; ++g) {
            const int32_t *mem = &perm[(size_t)(g * g0)];  // uniform groups: group g at g * g0
            std::fill(used.begin(), used.end(), 0);
            int64_t np = 0;
            for (int64_t a = 0; a < g0 && ok; ++a) {
                if (used[(size_t)a]) continue;
                const double4 ka = ks[(size_t)mem[a]];
                int64_t best = -1;
                double bd = 0.0;
                for (int64_t b = a + 1; b < g0; ++b) {
                    if (used[(size_t)b]) continue;
                    const double4 kb = ks[(size_t)mem[b]];
                    const double d = std::fabs(ka.x + kb.x) + std::fabs(ka.y + kb.y) + std::fabs(ka.z + kb.z);
                    if (d <= tol && (best < 0 || d < bd)) {
                        best = b;
                        bd = d;
                    }
                }
                if (best < 0) {
                    ok = false;
                    break;
                }
                used[(size_t)a] = used[(size_t)best] = 1;
                first[(size_t)(g * g0 / 2 + np)] = mem[a];
                second[(size_t)(g * g0 / 2 + np)] = mem[best];
                worst = std::max(worst, bd);
                ++np;
            }
        }
        if (ok) {
            const int64_t half = g0 / 2, gpc = kDirectThreads / half;
            std::vector<int32_t> &pc = h.perm_conj;
            pc.clear();
            for (int64_t g = 0; g < p->n_tgroups; g += gpc) {
                const int64_t ng = std::min<int64_t>(gpc, p->n_tgroups - g);
                const int32_t o0 = (int32_t)pc.size();
                for (int64_t q = g * half; q < (g + ng) * half; ++q) pc.push_back(first[(size_t)q]);
                const int32_t o1 = (int32_t)pc.size();
                for (int64_t q = g * half; q < (g + ng) * half; ++q) pc.push_back(second[(size_t)q]);
                h.tiles_q.push_back({(int32_t)g, o0, (int32_t)(ng * half), 0});
                h.tiles_q.push_back({(int32_t)g, o1, (int32_t)(ng * half), 0});
            }
            p->conj_pack = (int)gpc;
            p->n_tiles_conj = (int64_t)h.tiles_q.size() / kDirectSPL;
            p->conj_dk = worst;
        }
    }

    // Readout runs: kRunLen consecutive readout samples of a line (stride
    // lines * kz in the sample index) whose k' and t are linear in the
    // readout index to 1e-12 of the largest |k'| / |t|.  All runs or none.
    std::vector<RunDesc> &runs = h.runs;
    if (n_meas >= 2 && S > 0) {
        const int64_t stride = n_lines * n_kz;
        double kmax = 0.0, tmax = 0.0;
        for (int64_t s = 0; s < S; ++s) {
            kmax = std::max({kmax, std::fabs(ks[(size_t)s].x), std::fabs(ks[(size_t)s].y), std::fabs(ks[(size_t)s].z)});
            tmax = std::max(tmax, std::fabs(ks[(size_t)s].w));
        }
        const double tolk = 1e-12 * (kmax > 0.0 ? kmax : 1.0), tolt = 1e-12 * (tmax > 0.0 ? tmax : 1.0);
        bool linear = true;
        double devk = 0.0, devt = 0.0;  // largest departure from the linear model
        runs.reserve((size_t)(stride * cdiv(n_meas, kRunLen)));
        for (int64_t line = 0; line < stride && linear; ++line) {
            // one slope per line (first to last sample), so that the runs of
            // a line share their step factor
            const double4 f = ks[(size_t)line], l = ks[(size_t)((n_meas - 1) * stride + line)];
            const double inv = 1.0 / (double)(n_meas - 1);
            const double4 d = make_double4((l.x - f.x) * inv, (l.y - f.y) * inv, (l.z - f.z) * inv, (l.w - f.w) * inv);
            for (int64_t i0 = 0; i0 < n_meas && linear; i0 += kRunLen) {
                const int64_t len = std::min<int64_t>(kRunLen, n_meas - i0);
                const int64_t s0 = i0 * stride + line;
                const double4 a = ks[(size_t)s0];
                for (int64_t m = 1; m < len && linear; ++m) {
                    const double4 c = ks[(size_t)(s0 + m * stride)];
                    const double ek = std::max({std::fabs(c.x - a.x - m * d.x), std::fabs(c.y - a.y - m * d.y),
                                                std::fabs(c.z - a.z - m * d.z)});
                    const double et = std::fabs(c.w - a.w - m * d.w);
                    devk = std::max(devk, ek);
                    devt = std::max(devt, et);
                    linear = ek <= tolk && et <= tolt;
                }
                RunDesc r{};
                r.k0x = a.x;
                r.k0y = a.y;
                r.k0z = a.z;
                r.t0 = a.w;
                r.dkx = d.x;
                r.dky = d.y;
                r.dkz = d.z;
                r.dt = d.w;
                r.s0 = (int32_t)s0;
                r.len = (int32_t)len;
                runs.push_back(r);
            }
        }
        if (linear) {
            p->n_runs = (int64_t)runs.size();
            p->run_stride = stride;
            p->runs_exact = devk <= 1e-2 * tolk && devt <= 1e-2 * tolt;
            // the distinct steps of each 256-run tile (bitwise)
            const int64_t n_tiles = cdiv(p->n_runs, kDirectThreads);
            h.run_steps.assign((size_t)(n_tiles * kRunSteps * 4), 0.0);
            h.run_nsteps.assign((size_t)n_tiles, 0);
            int most = 0;
            for (int64_t b = 0; b < n_tiles && most <= kRunSteps; ++b) {
                int cnt = 0;
                for (int64_t r = b * kDirectThreads; r < std::min<int64_t>(p->n_runs, (b + 1) * kDirectThreads); ++r) {
                    RunDesc &rd = runs[(size_t)r];
                    const double v[4] = {rd.dkx, rd.dky, rd.dkz, rd.dt};
                    int j = 0;
                    for (; j < cnt; ++j)
                        if (std::memcmp(&h.run_steps[(size_t)((b * kRunSteps + j) * 4)], v, sizeof v) == 0) break;
                    if (j == cnt) {
                        if (cnt == kRunSteps) {
                            most = kRunSteps + 1;
                            break;
                        }
                        std::memcpy(&h.run_steps[(size_t)((b * kRunSteps + j) * 4)], v, sizeof v);
                        ++cnt;
                    }
                    rd.step = j;
                }
                h.run_nsteps[(size_t)b] = cnt;
                most = std::max(most, cnt);
            }
            p->run_steps = most <= kRunSteps ? most : 0;
            if (p->run_steps == 0) {
                h.run_steps.clear();
                h.run_nsteps.clear();
            }
            // whole lines: every sample within 1e-12 of the line's first
            // sample plus m times its slope (the per-line GEMM's model)
            bool whole = true;
            std::vector<RunDesc> &lines = h.lines;
            lines.reserve((size_t)stride);
            for (int64_t line = 0; line < stride && whole; ++line) {
                const double4 f = ks[(size_t)line], l = ks[(size_t)((n_meas - 1) * stride + line)];
                const double inv = 1.0 / (double)(n_meas - 1);
                const double4 d = make_double4((l.x - f.x) * inv, (l.y - f.y) * inv, (l.z - f.z) * inv, (l.w - f.w) * inv);
                for (int64_t m = 1; m < n_meas && whole; ++m) {
                    const double4 c = ks[(size_t)(m * stride + line)];
                    const double ek = std::max({std::fabs(c.x - f.x - m * d.x), std::fabs(c.y - f.y - m * d.y),
                                                std::fabs(c.z - f.z - m * d.z)});
                    whole = ek <= tolk && std::fabs(c.w - f.w - m * d.w) <= tolt;
                }
                RunDesc r{};
                r.k0x = f.x;
                r.k0y = f.y;
                r.k0z = f.z;
                r.t0 = f.w;
                r.dkx = d.x;
                r.dky = d.y;
                r.dkz = d.z;
                r.dt = d.w;
                r.s0 = (int32_t)line;
                r.len = (int32_t)n_meas;
                lines.push_back(r);
            }
            if (whole) p->n_line_tabs = stride;
            else lines.clear();
        } else {
            runs.clear();
        }
    }

    return FM_OK;
}
}  // namespace fm

extern "C" int fm_signal_plan_create(fm_signal_plan **plan, int64_t n_meas, int64_t n_lines,
                                     int64_t n_kz, const double *kx, const double *ky,
                                     const double *kz, const double *t, int method)
{
    FM_CHECK_ARG(plan != nullptr, "plan pointer is NULL");
    *plan = nullptr;
    FM_CHECK_ARG(n_meas >= 0 && n_lines >= 0 && n_kz >= 0, "negative k-space dimension");
    FM_CHECK_ARG(method == FM_METHOD_AUTO || method == FM_METHOD_DIRECT || method == FM_METHOD_CARTESIAN,
                 "unknown method %d", method);
    const int64_t S = n_meas * n_lines * n_kz;
    FM_CHECK_ARG(S < (int64_t(1) << 31), "too many k-space samples (%lld)", (long long)S);
    FM_CHECK_ARG(S == 0 || (kx && ky && kz && t), "NULL k-space array");

    auto *p = new fm_signal_plan();
    p->n_meas = n_meas;
    p->n_lines = n_lines;
    p->n_kz = n_kz;
    p->n_samples = S;
    {
        const hipError_t e = hipGetDevice(&p->device);
        if (e != hipSuccess) {
            delete p;
            set_error("hipGetDevice failed: %s", hipGetErrorString(e));
            return FM_ERR_HIP;
        }
    }
    PlanHost h;
    {
        const int rc = plan_analyse(p, kx, ky, kz, t, method, h);
        if (rc != FM_OK) {
            delete p;
            return rc;
        }
    }
    const std::vector<int32_t> &perm = h.perm;
    const std::vector<SignalTile> &tiles_g = h.tiles_g, &tiles_c = h.tiles_c;
    const std::vector<double4> &ks = h.ks;
    const std::vector<RunDesc> &runs = h.runs;

    auto fail = [&](hipError_t e, const char *what) {
        set_error("%s failed: %s", what, hipGetErrorString(e));
        fm_signal_plan_destroy(p);
        return FM_ERR_HIP;
    };
    hipError_t e;
    if (S > 0) {
        if ((e = hipMalloc(&p->d_ks, sizeof(double4) * (size_t)S)) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipMalloc(&p->d_perm, sizeof(int32_t) * (size_t)S)) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipMalloc(&p->d_tiles_grouped, sizeof(SignalTile) * tiles_g.size())) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipMalloc(&p->d_tiles_consec, sizeof(SignalTile) * tiles_c.size())) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipMalloc(&p->d_tg_t, sizeof(double) * (size_t)p->n_tgroups)) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipMemcpy(p->d_ks, ks.data(), sizeof(double4) * (size_t)S, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        if ((e = hipMemcpy(p->d_perm, perm.data(), sizeof(int32_t) * (size_t)S, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        if ((e = hipMemcpy(p->d_tiles_grouped, tiles_g.data(), sizeof(SignalTile) * tiles_g.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        if ((e = hipMemcpy(p->d_tiles_consec, tiles_c.data(), sizeof(SignalTile) * tiles_c.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        if (!h.tiles_p.empty()) {
            if ((e = hipMalloc(&p->d_tiles_packed, sizeof(SignalTile) * h.tiles_p.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_tiles_packed, h.tiles_p.data(), sizeof(SignalTile) * h.tiles_p.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        }
        if (!h.tiles_q.empty()) {
            if ((e = hipMalloc(&p->d_tiles_conj, sizeof(SignalTile) * h.tiles_q.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_tiles_conj, h.tiles_q.data(), sizeof(SignalTile) * h.tiles_q.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
            if ((e = hipMalloc(&p->d_perm_conj, sizeof(int32_t) * h.perm_conj.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_perm_conj, h.perm_conj.data(), sizeof(int32_t) * h.perm_conj.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        }
        if ((e = hipMemcpy(p->d_tg_t, p->tg_t.data(), sizeof(double) * (size_t)p->n_tgroups, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        if (p->n_runs > 0) {
            if ((e = hipMalloc(&p->d_runs, sizeof(RunDesc) * runs.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_runs, runs.data(), sizeof(RunDesc) * runs.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        }
        if (p->n_line_tabs > 0) {
            if ((e = hipMalloc(&p->d_lines, sizeof(RunDesc) * h.lines.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_lines, h.lines.data(), sizeof(RunDesc) * h.lines.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        }
        if (p->run_steps > 0) {
            if ((e = hipMalloc(&p->d_run_steps, sizeof(double) * h.run_steps.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_run_steps, h.run_steps.data(), sizeof(double) * h.run_steps.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
            if ((e = hipMalloc(&p->d_run_nsteps, sizeof(int32_t) * h.run_nsteps.size())) != hipSuccess) return fail(e, "hipMalloc");
            if ((e = hipMemcpy(p->d_run_nsteps, h.run_nsteps.data(), sizeof(int32_t) * h.run_nsteps.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
        }
        if (p->method == FM_METHOD_CARTESIAN) {
            const int rc = cartesian_upload(p);
            if (rc != FM_OK) {
                fm_signal_plan_destroy(p);
                return rc;
            }
        }
    }
    *plan = p;
    return FM_OK;
}

extern "C" int fm_signal_plan_analyse(int64_t n_meas, int64_t n_lines, int64_t n_kz, const double *kx,
                                      const double *ky, const double *kz, const double *t, int method,
                                      int *method_out, int64_t *n_tgroups, int64_t *n_runs,
                                      int64_t *n_tiles_grouped, int *runs_exact, int *group_pack,
                                      int *conj_pack)
{
    FM_CHECK_ARG(n_meas >= 0 && n_lines >= 0 && n_kz >= 0, "negative k-space dimension");
    FM_CHECK_ARG(method == FM_METHOD_AUTO || method == FM_METHOD_DIRECT || method == FM_METHOD_CARTESIAN,
                 "unknown method %d", method);
    const int64_t S = n_meas * n_lines * n_kz;
    FM_CHECK_ARG(S < (int64_t(1) << 31), "too many k-space samples (%lld)", (long long)S);
    FM_CHECK_ARG(S == 0 || (kx && ky && kz && t), "NULL k-space array");
    fm_signal_plan p;
    p.n_meas = n_meas;
    p.n_lines = n_lines;
    p.n_kz = n_kz;
    p.n_samples = S;
    PlanHost h;
    const int rc = plan_analyse(&p, kx, ky, kz, t, method, h);
    if (rc != FM_OK) return rc;
    if (method_out) *method_out = p.method;
    if (n_tgroups) *n_tgroups = p.n_tgroups;
    if (n_runs) *n_runs = p.n_runs;
    if (n_tiles_grouped) *n_tiles_grouped = p.n_tiles_grouped;
    if (runs_exact) *runs_exact = p.runs_exact ? 1 : 0;
    if (group_pack) *group_pack = p.group_pack;
    if (conj_pack) *conj_pack = p.conj_pack;
    return FM_OK;
}

extern "C" int fm_signal_plan_set_timing(fm_signal_plan *p, int max_launches)
{
    FM_CHECK_ARG(p != nullptr, "plan is NULL");
    FM_CHECK_ARG(max_launches >= 0 && max_launches <= 1 << 16, "max_launches out of range");
    for (hipEvent_t e : p->ev_start) (void)hipEventDestroy(e);
    for (hipEvent_t e : p->ev_stop) (void)hipEventDestroy(e);
    p->ev_start.clear();
    p->ev_stop.clear();
    p->ev_used = 0;
    for (int i = 0; i < max_launches; ++i) {
        hipEvent_t a, b;
        FM_HIP(hipEventCreate(&a));
        FM_HIP(hipEventCreate(&b));
        p->ev_start.push_back(a);
        p->ev_stop.push_back(b);
    }
    return FM_OK;
}

extern "C" int fm_signal_plan_kernel_ms(fm_signal_plan *p, double *total_ms, int *n_timed)
{
    FM_CHECK_ARG(p != nullptr && total_ms != nullptr && n_timed != nullptr, "NULL argument");
    double sum = 0.0;
    const int n = p->ev_used;
    if (n > 0) FM_HIP(hipEventSynchronize(p->ev_stop[(size_t)n - 1]));
    for (int i = 0; i < n; ++i) {
        float ms = 0.0f;
        FM_HIP(hipEventElapsedTime(&ms, p->ev_start[(size_t)i], p->ev_stop[(size_t)i]));
        sum += ms;
    }
    p->ev_used = 0;
    *total_ms = sum;
    *n_timed = n;
    return FM_OK;
}

extern "C" int fm_signal_plan_destroy(fm_signal_plan *p)
{
    if (!p) return FM_OK;
    // executes still in flight on any stream read the plan's tables: wait
    // for the device before they are freed
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != p->device) (void)hipSetDevice(p->device);
        (void)hipDeviceSynchronize();
        if (cur != p->device) (void)hipSetDevice(cur);
    }
    for (hipEvent_t e : p->ev_start) (void)hipEventDestroy(e);
    for (hipEvent_t e : p->ev_stop) (void)hipEventDestroy(e);
    (void)hipFree(p->d_ks);
    (void)hipFree(p->d_perm);
    (void)hipFree(p->d_tiles_grouped);
    (void)hipFree(p->d_runs);
    (void)hipFree(p->d_lines);
    (void)hipFree(p->d_run_steps);
    (void)hipFree(p->d_run_nsteps);
    (void)hipFree(p->d_tiles_consec);
    (void)hipFree(p->d_tiles_packed);
    (void)hipFree(p->d_tiles_conj);
    (void)hipFree(p->d_perm_conj);
    (void)hipFree(p->d_tg_t);
    cartesian_free(p);
    delete p;
    return FM_OK;
}

extern "C" int fm_signal_plan_info(const fm_signal_plan *p, int64_t *n_samples, int64_t *n_tgroups,
                                   int *method, double *tgroup_t)
{
    FM_CHECK_ARG(p != nullptr, "plan is NULL");
    if (n_samples) *n_samples = p->n_samples;
    if (n_tgroups) *n_tgroups = p->n_tgroups;
    if (method) *method = p->method;
    if (tgroup_t && p->n_tgroups) std::memcpy(tgroup_t, p->tg_t.data(), sizeof(double) * (size_t)p->n_tgroups);
    return FM_OK;
}

extern "C" int fm_signal_plan_runs(const fm_signal_plan *p, int64_t *n_runs, int *run_len, int *runs_exact)
{
    FM_CHECK_ARG(p != nullptr, "plan is NULL");
    if (n_runs) *n_runs = p->n_runs;
    if (run_len) *run_len = kRunLen;
    if (runs_exact) *runs_exact = p->runs_exact ? 1 : 0;
    return FM_OK;
}

extern "C" int fm_signal_plan_lines(const fm_signal_plan *p, int64_t *n_lines, int *line_gemm)
{
    FM_CHECK_ARG(p != nullptr, "plan is NULL");
    if (n_lines) *n_lines = p->n_line_tabs;
    if (line_gemm) *line_gemm = lines_usable(p, 0) ? 1 : 0;
    return FM_OK;
}

extern "C" int fm_signal_workspace_size(const fm_signal_plan *p, int precision, int64_t n_nodes,
                                        int n_enc, int moving, size_t *bytes)
{
    FM_CHECK_ARG(p != nullptr && bytes != nullptr, "NULL argument");
    FM_CHECK_ARG(n_nodes >= 0 && n_enc >= 1, "bad sizes");
    // the path fm_signal_execute takes: a Cartesian plan runs static nodes
    // on its GEMM and moving nodes on the per-readout moving GEMM (either
    // form, up to kMovMaxV encodings) or the direct kernels
    if (p->method == FM_METHOD_CARTESIAN && !moving)
        *bytes = cartesian_workspace(p, precision, n_nodes, n_enc);
    else if (p->method == FM_METHOD_CARTESIAN && n_enc <= kMovMaxV)
        *bytes = cartesian_moving_workspace(p, precision, n_nodes, n_enc);
    else
        *bytes = direct_workspace(p, n_nodes, n_enc);
    return FM_OK;
}

extern "C" int fm_signal_nodes_pack(int64_t n_nodes, int n_enc, int dtype, const void *r0,
                                    const void *phi_dB0, const void *T2, const void *mxy,
                                    const double *weights, double *packed, void *stream)
{
    FM_CHECK_ARG(n_nodes >= 0 && n_enc >= 1, "bad sizes (n_nodes=%lld, n_enc=%d)", (long long)n_nodes, n_enc);
    FM_CHECK_ARG(dtype == FM_DTYPE_F32 || dtype == FM_DTYPE_F64, "unknown dtype %d", dtype);
    if (n_nodes == 0) return FM_OK;
    FM_CHECK_ARG(r0 && phi_dB0 && T2 && mxy && packed, "NULL array");
    const unsigned blocks = (unsigned)cdiv(n_nodes, 256);
    hipStream_t st = as_stream(stream);
    if (dtype == FM_DTYPE_F64)
        k_nodes_pack<double><<<blocks, 256, 0, st>>>(n_nodes, n_enc, (const double *)r0, (const double *)phi_dB0,
                                                    (const double *)T2, (const double *)mxy, weights, packed);
    else
        k_nodes_pack<float><<<blocks, 256, 0, st>>>(n_nodes, n_enc, (const float *)r0, (const float *)phi_dB0,
                                                   (const float *)T2, (const float *)mxy, weights, packed);
    FM_LAUNCH_CHECK("k_nodes_pack");
    return FM_OK;
}

extern "C" int fm_signal_execute(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc,
                                 const double *packed, const float *modes, int n_modes,
                                 const double *motion_coef, void *workspace, size_t workspace_bytes,
                                 double *out, void *stream)
{
    FM_CHECK_ARG(p != nullptr, "plan is NULL");
    FM_CHECK_ARG(precision == FM_PREC_F64 || precision == FM_PREC_MIXED,
                 "signal precision must be FM_PREC_F64 or FM_PREC_MIXED (got %d)", precision);
    FM_CHECK_ARG(n_nodes >= 0 && n_enc >= 1, "bad sizes (n_nodes=%lld, n_enc=%d)", (long long)n_nodes, n_enc);
    FM_CHECK_ARG(p->n_samples == 0 || out != nullptr, "out is NULL");
    FM_CHECK_ARG(n_nodes == 0 || packed != nullptr, "packed is NULL");
    FM_CHECK_ARG(n_modes >= 0 && n_modes <= kMaxModes, "n_modes must be in [0, %d]", kMaxModes);
    FM_CHECK_ARG(!(modes != nullptr && n_modes > 0) || motion_coef != nullptr, "motion_coef is NULL");
    int dev = 0;
    FM_HIP(hipGetDevice(&dev));
    FM_CHECK_ARG(dev == p->device, "plan was created on device %d, current device is %d", p->device, dev);
    hipStream_t st = as_stream(stream);
    const bool motion = modes != nullptr && n_modes > 0;
    if (p->method == FM_METHOD_CARTESIAN && !motion)
        return cartesian_execute(p, precision, n_nodes, n_enc, packed, workspace, workspace_bytes, out, st);
    // moving nodes on a Cartesian plan: the per-readout GEMM in either form
    // (up to kMovMaxV encodings), otherwise the direct kernels
    if (p->method == FM_METHOD_CARTESIAN && motion && n_enc <= kMovMaxV)
        return cartesian_moving_execute(p, precision, n_nodes, n_enc, packed, modes, n_modes, motion_coef,
                                        workspace, workspace_bytes, out, st);
    return direct_execute(p, precision, n_nodes, n_enc, packed, modes, n_modes, motion_coef, workspace,
                          workspace_bytes, out, st);
}

extern "C" int fm_csr_spmv_c128(int64_t n_rows, int n_enc, const int64_t *indptr, const int64_t *indices,
                                const double *data, const double *x, double *y, void *stream)
{
    FM_CHECK_ARG(n_rows >= 0 && n_enc >= 1, "bad sizes");
    if (n_rows == 0) return FM_OK;
    FM_CHECK_ARG(indptr && indices && data && x && y, "NULL array");
    // 16 lanes per row: P1 mass rows hold ~15 entries on tetrahedral meshes
    // (longer rows take more passes)
    constexpr int G = 16;
    const int64_t threads = n_rows * G;
    FM_CHECK_ARG(cdiv(threads, 256) <= 0x7fffffff, "too many rows for one launch");
    k_csr_spmv_c128<G><<<(unsigned)cdiv(threads, 256), 256, 0, as_stream(stream)>>>(
        n_rows, n_enc, indptr, indices, data, reinterpret_cast<const double2 *>(x), reinterpret_cast<double2 *>(y));
    FM_LAUNCH_CHECK("k_csr_spmv_c128");
    return FM_OK;
}
