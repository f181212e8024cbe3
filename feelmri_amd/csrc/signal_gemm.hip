// Cartesian factorised signal: a complex GEMM on the matrix cores.
//
// When every k-space coordinate splits additively over the three sample
// indices, k(i,j,k) = F(i) + G(j) + H(k) per axis, and t depends on the
// readout index i only (Cartesian stacks with one line per shot,
// python/feelmri/KSpaceTraj.py:183-220), the direct sum of
// cpp/feelmri/MRI.cpp:279-381 factorises:
//
//   S[i, (j,k,v)] = sum_n E[i,n] A[n,(j,k,v)]
//   E[i,n]        = exp(-t_i/T2_n) exp(i phi_n t_i) exp(-2 pi i F_i . x_n)
//   A[n,(j,k,v)]  = Ey[j,n] Ezb[(k,v),n]
//   Ey[j,n]       = exp(-2 pi i G_j . x_n),  Ezb[(k,v),n] = b[n,v] exp(-2 pi i H_k . x_n)
//
// (2 pi carries the reference's float-pi factor, MRI.h:13).  Per execute the
// E, Ey and Ezb tables are computed once (O((n_ro + n_ln + n_kz V) N)
// transcendentals); the GEMM then streams E through LDS and builds the A
// operand in registers from Ey and Ezb, so the (N x lines*kz*V) matrix A is
// never stored.
//
// Mixed form: E and A are split into bf16 hi + lo parts and each real product
// x*y is formed as xh*yh + xh*yl + xl*yh on v_mfma_f32_32x32x16_bf16 with
// FP32 accumulation (12 MFMAs per 16 nodes per 32x32 complex tile).  FP64
// form: v_mfma_f64_16x16x4_f64 (4 MFMAs per 4 nodes per 16x16 complex tile).
// Nodes are split across workgroups; per-split partial tiles are summed in
// FP64 in a fixed order by a second kernel.  Workgroups that share a node
// split (and so the same E tiles) are mapped onto the same XCD so the tiles
// are fetched from HBM once per split and served from that XCD's L2.
#include <cmath>
#include <vector>

#include "signal_plan.h"

namespace fm {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) double f64x4;

constexpr int kGemmThreads = 256;  // 4 waves
constexpr int kXcds = 8;

// ---------------------------------------------------------------------------
// Sampling analysis (host)
// ---------------------------------------------------------------------------
bool cartesian_analyse(fm_signal_plan *p, const double *kx, const double *ky, const double *kz,
                       const double *t)
{
    const int64_t R = p->n_meas, L = p->n_lines, K = p->n_kz;
    auto at = [&](const double *a, int64_t i, int64_t j, int64_t k) { return a[(i * L + j) * K + k]; };
    // t must be a function of the readout index only
    for (int64_t i = 0; i < R; ++i)
        for (int64_t j = 0; j < L; ++j)
            for (int64_t k = 0; k < K; ++k)
                if (!(at(t, i, j, k) == at(t, i, 0, 0))) return false;
    CartesianTables &c = p->cart;
    c.n_ro = R;
    c.n_ln = L;
    c.n_kz = K;
    c.F.assign((size_t)R * 3, 0.0);
    c.G.assign((size_t)L * 3, 0.0);
    c.H.assign((size_t)K * 3, 0.0);
    c.t.resize((size_t)R);
    for (int64_t i = 0; i < R; ++i) c.t[(size_t)i] = at(t, i, 0, 0);
    const double *ax[3] = {kx, ky, kz};
    for (int a = 0; a < 3; ++a) {
        const double *k3 = ax[a];
        double kmax = 0.0;
        bool only_i = true, only_j = true, only_k = true;
        for (int64_t i = 0; i < R; ++i)
            for (int64_t j = 0; j < L; ++j)
                for (int64_t k = 0; k < K; ++k) {
                    const double v = at(k3, i, j, k);
                    kmax = std::max(kmax, std::fabs(v));
                    only_i &= v == at(k3, i, 0, 0);
                    only_j &= v == at(k3, 0, j, 0);
                    only_k &= v == at(k3, 0, 0, k);
                }
        if (only_i) {
            for (int64_t i = 0; i < R; ++i) c.F[(size_t)(3 * i + a)] = at(k3, i, 0, 0);
        } else if (only_j) {
            for (int64_t j = 0; j < L; ++j) c.G[(size_t)(3 * j + a)] = at(k3, 0, j, 0);
        } else if (only_k) {
            for (int64_t k = 0; k < K; ++k) c.H[(size_t)(3 * k + a)] = at(k3, 0, 0, k);
        } else {
            // general additive split (e.g. a rotated Cartesian grid)
            const double k0 = at(k3, 0, 0, 0);
            for (int64_t i = 0; i < R; ++i) c.F[(size_t)(3 * i + a)] = at(k3, i, 0, 0);
            for (int64_t j = 0; j < L; ++j) c.G[(size_t)(3 * j + a)] = at(k3, 0, j, 0) - k0;
            for (int64_t k = 0; k < K; ++k) c.H[(size_t)(3 * k + a)] = at(k3, 0, 0, k) - k0;
            const double tol = 1e-12 * std::max(kmax, 1.0);
            for (int64_t i = 0; i < R; ++i)
                for (int64_t j = 0; j < L; ++j)
                    for (int64_t k = 0; k < K; ++k) {
                        const double s = c.F[(size_t)(3 * i + a)] + c.G[(size_t)(3 * j + a)] + c.H[(size_t)(3 * k + a)];
                        if (!(std::fabs(at(k3, i, j, k) - s) <= tol)) return false;
                    }
        }
    }
    for (auto *v : {&c.F, &c.G, &c.H})
        for (double &x : *v) x *= kPiFOverPi;
    c.ok = true;
    return true;
}

int cartesian_upload(fm_signal_plan *p)
{
    CartesianTables &c = p->cart;
    FM_HIP(hipMalloc(&c.d_F, sizeof(double) * c.F.size()));
    FM_HIP(hipMalloc(&c.d_G, sizeof(double) * c.G.size()));
    FM_HIP(hipMalloc(&c.d_H, sizeof(double) * c.H.size()));
    FM_HIP(hipMalloc(&c.d_t, sizeof(double) * c.t.size()));
    FM_HIP(hipMemcpy(c.d_F, c.F.data(), sizeof(double) * c.F.size(), hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(c.d_G, c.G.data(), sizeof(double) * c.G.size(), hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(c.d_H, c.H.data(), sizeof(double) * c.H.size(), hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(c.d_t, c.t.data(), sizeof(double) * c.t.size(), hipMemcpyHostToDevice));
    return FM_OK;
}

void cartesian_free(fm_signal_plan *p)
{
    CartesianTables &c = p->cart;
    (void)hipFree(c.d_F);
    (void)hipFree(c.d_G);
    (void)hipFree(c.d_H);
    (void)hipFree(c.d_t);
    c.d_F = c.d_G = c.d_H = c.d_t = nullptr;
}

// ---------------------------------------------------------------------------
// Geometry of one execute
// ---------------------------------------------------------------------------
struct GemmShape {
    bool mixed;
    int64_t N, Np;         // nodes, padded to the K step of a stage
    int64_t M, Mp;         // readout rows, padded to the row block
    int64_t NC, NCp;       // output columns lines*kz*V, padded to the column block
    int rb_per_wave;       // row blocks per wave
    int64_t n_rt, n_ct;    // row tiles, column tiles (per workgroup)
    int64_t ksteps;        // K stages in total
    int64_t n_splits, steps_per_split;
    // workspace offsets (bytes)
    size_t off_E, off_Ey, off_Ez, off_part, total;
};

// mixed: row block 32, K stage 16 nodes, column block 32 per wave (4 per WG)
// f64:   row block 16, K stage 16 nodes (4 MFMA k-steps), column block 16 per wave
static GemmShape gemm_shape(const fm_signal_plan *p, int precision, int64_t N, int V)
{
    GemmShape s{};
    const CartesianTables &c = p->cart;
    s.mixed = precision != FM_PREC_F64;
    const int rowblk = s.mixed ? 32 : 16;
    const int colblk = s.mixed ? 32 : 16;
    const int max_rb = s.mixed ? 4 : 8;
    s.N = N;
    s.Np = cdiv(std::max<int64_t>(N, 1), 16) * 16;
    s.M = c.n_ro;
    const int64_t nrb = cdiv(s.M, rowblk);
    s.rb_per_wave = (int)std::min<int64_t>(max_rb, nrb);
    s.n_rt = cdiv(nrb, s.rb_per_wave);
    s.Mp = s.n_rt * s.rb_per_wave * rowblk;
    s.NC = c.n_ln * c.n_kz * V;
    s.n_ct = cdiv(s.NC, 4 * colblk);
    s.NCp = s.n_ct * 4 * colblk;
    s.ksteps = s.Np / 16;
    const int64_t base = s.n_rt * s.n_ct;
    const int64_t target = 4 * (int64_t)device_cu_count();
    int64_t splits = cdiv(target, base);
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, cdiv(s.ksteps, 8)));
    s.steps_per_split = cdiv(s.ksteps, splits);
    s.n_splits = cdiv(s.ksteps, s.steps_per_split);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    if (s.mixed) {
        s.off_E = take((size_t)s.Mp * s.Np * 4 * sizeof(uint16_t));          // 4 bf16 planes
        s.off_Ey = take((size_t)c.n_ln * s.Np * sizeof(float2));
        s.off_Ez = take((size_t)c.n_kz * V * s.Np * sizeof(float2));
        s.off_part = take((size_t)s.n_splits * s.M * s.NC * sizeof(float2));
    } else {
        s.off_E = take((size_t)s.Mp * s.Np * 2 * sizeof(double));            // Re, Im planes
        s.off_Ey = take((size_t)c.n_ln * s.Np * sizeof(double2));
        s.off_Ez = take((size_t)c.n_kz * V * s.Np * sizeof(double2));
        s.off_part = take((size_t)s.n_splits * s.M * s.NC * sizeof(double2));
    }
    s.total = off;
    return s;
}

size_t cartesian_workspace(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc)
{
    if (p->n_samples == 0 || n_nodes == 0) return 0;
    return gemm_shape(p, precision, n_nodes, n_enc).total;
}

// ---------------------------------------------------------------------------
// Operand tables
// ---------------------------------------------------------------------------
// exp(-2 pi i u) for u in cycles: FP64 range reduction, then FP64 sincospi
// (FP64 form) or the hardware FP32 sine/cosine, which take revolutions.
template <bool MIXED>
__device__ inline void cis_neg_cycles(double u, double &c, double &s)
{
    const double r = u - __builtin_rint(u);
    if (MIXED) {
        const float rf = (float)r;
        c = (double)__builtin_amdgcn_cosf(rf);
        s = -(double)__builtin_amdgcn_sinf(rf);
    } else {
        double sn, cs;
        sincospi(2.0 * r, &sn, &cs);
        c = cs;
        s = -sn;
    }
}

__device__ inline void split_bf16(float x, uint16_t &hi, uint16_t &lo)
{
    const __bf16 h = (__bf16)x;
    const float hf = (float)h;
    const __bf16 l = (__bf16)(x - hf);
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, l);
}

// E table.  Mixed: tiles of (32 rows x 16 nodes) per 4 planes (Er_h, Er_l,
// Ei_h, Ei_l), each plane in MFMA lane order: lane l holds row l%32 and
// nodes 8*(l/32) .. +7 (8 bf16 = 16 bytes).  Tile index (kstep, rowblock).
// FP64: tiles of (16 rows x 16 nodes) per 2 planes (Er, Ei), plane layout
// [kstep4 (4)][lane (64)] with lane l holding row l%16, node l/16 of that
// 4-node step.  Padding rows/nodes are zero.
template <bool MIXED>
__global__ void k_cart_E(const double *__restrict__ packed, int64_t N, int64_t Np, int64_t M, int64_t Mp,
                         const double *__restrict__ F, const double *__restrict__ tr, void *__restrict__ Eout)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // over Mp * Np
    if (idx >= Mp * Np) return;
    const int64_t n = idx % Np, i = idx / Np;
    double er = 0.0, ei = 0.0;
    if (n < N && i < M) {
        const double x = packed[n], y = packed[N + n], z = packed[2 * N + n];
        const double phi = packed[3 * N + n], r2 = packed[4 * N + n];
        const double ti = tr[i];
        // total phase in cycles: -(F.x) + phi t / (2 pi)
        const double u = fma(F[3 * i], x, fma(F[3 * i + 1], y, F[3 * i + 2] * z)) - phi * ti * (1.0 / kTwoPi);
        double c, s;
        cis_neg_cycles<MIXED>(u, c, s);
        const double d = exp(-ti * r2);
        er = d * c;
        ei = d * s;
    }
    if (MIXED) {
        const int64_t kb = n / 16, q16 = n % 16;
        const int64_t rb = i / 32, row = i % 32;
        const int lane = (int)(row + 32 * (q16 / 8));
        const int q = (int)(q16 % 8);
        const int64_t nrb = Mp / 32;
        uint16_t *E = reinterpret_cast<uint16_t *>(Eout);
        const int64_t tile = (kb * nrb + rb) * 4;  // 4 planes per tile
        uint16_t h, l;
        split_bf16((float)er, h, l);
        E[((tile + 0) * 64 + lane) * 8 + q] = h;
        E[((tile + 1) * 64 + lane) * 8 + q] = l;
        split_bf16((float)ei, h, l);
        E[((tile + 2) * 64 + lane) * 8 + q] = h;
        E[((tile + 3) * 64 + lane) * 8 + q] = l;
    } else {
        const int64_t kb = n / 16, q16 = n % 16;
        const int64_t rb = i / 16, row = i % 16;
        const int64_t step = q16 / 4;
        const int lane = (int)(row + 16 * (q16 % 4));
        const int64_t nrb = Mp / 16;
        double *E = reinterpret_cast<double *>(Eout);
        const int64_t tile = (kb * nrb + rb) * 2;
        E[((tile + 0) * 4 + step) * 64 + lane] = er;
        E[((tile + 1) * 4 + step) * 64 + lane] = ei;
    }
}

// Ey[j][n] = exp(-2 pi i G_j . x_n) ; Ezb[k*V+v][n] = b[n,v] exp(-2 pi i H_k . x_n)
template <bool MIXED>
__global__ void k_cart_yz(const double *__restrict__ packed, int64_t N, int64_t Np, int V, int64_t n_ln,
                          int64_t n_kz, const double *__restrict__ G, const double *__restrict__ H,
                          void *__restrict__ Ey, void *__restrict__ Ez)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // over (n_ln + n_kz) * Np
    if (idx >= (n_ln + n_kz) * Np) return;
    const int64_t n = idx % Np, row = idx / Np;
    const bool is_y = row < n_ln;
    const double *kv = is_y ? &G[3 * row] : &H[3 * (row - n_ln)];
    double c = 0.0, s = 0.0;
    if (n < N) {
        const double x = packed[n], y = packed[N + n], z = packed[2 * N + n];
        cis_neg_cycles<MIXED>(fma(kv[0], x, fma(kv[1], y, kv[2] * z)), c, s);
    }
    if (is_y) {
        if (MIXED)
            reinterpret_cast<float2 *>(Ey)[row * Np + n] = make_float2((float)c, (float)s);
        else
            reinterpret_cast<double2 *>(Ey)[row * Np + n] = make_double2(c, s);
    } else {
        const int64_t k = row - n_ln;
        for (int v = 0; v < V; ++v) {
            double br = 0.0, bi = 0.0;
            if (n < N) {
                br = packed[(int64_t)(5 + 2 * v) * N + n];
                bi = packed[(int64_t)(6 + 2 * v) * N + n];
            }
            const double ar = br * c - bi * s, ai = br * s + bi * c;
            if (MIXED)
                reinterpret_cast<float2 *>(Ez)[(k * V + v) * Np + n] = make_float2((float)ar, (float)ai);
            else
                reinterpret_cast<double2 *>(Ez)[(k * V + v) * Np + n] = make_double2(ar, ai);
        }
    }
}

// Map the hardware workgroup id to (split, row tile, column tile) so that
// the workgroups of one split run on one XCD (ids are dealt round-robin over
// the 8 XCDs).
__device__ inline void gemm_tile_ids(int64_t n_rt, int64_t n_ct, int64_t n_splits, int64_t &split, int64_t &rt,
                                     int64_t &ct)
{
    const int64_t total = n_rt * n_ct * n_splits;
    const int64_t h = blockIdx.x;
    int64_t L = h;
    if (total % kXcds == 0) L = (h % kXcds) * (total / kXcds) + h / kXcds;
    const int64_t per_split = n_rt * n_ct;
    split = L / per_split;
    const int64_t rem = L % per_split;
    rt = rem / n_ct;
    ct = rem % n_ct;
}

// ---------------------------------------------------------------------------
// Mixed GEMM: bf16 hi+lo, v_mfma_f32_32x32x16_bf16
// ---------------------------------------------------------------------------
// Workgroup tile: RB*32 readout rows x 128 columns; wave w owns columns
// [32w, 32w+32) of the tile and all RB row blocks.  Per 16-node step the E
// tiles (RB x 4 planes, 64 lanes x 16 B each) are double-buffered in LDS;
// the next step's E tiles and Ey/Ezb values are loaded into registers while
// the current step's MFMAs run (one barrier per step).
template <int RB>
__global__ __launch_bounds__(kGemmThreads) void k_cart_gemm_bf16(
    const uint4 *__restrict__ E, const float2 *__restrict__ Ey, const float2 *__restrict__ Ez, int64_t Np,
    int64_t Mp, int64_t M, int64_t NC, int64_t n_kzV, int64_t n_rt, int64_t n_ct, int64_t n_splits,
    int64_t steps_per_split, int64_t ksteps, float2 *__restrict__ part)
{
    constexpr int kStage = RB * 4 * 64;                // uint4 per stage
    constexpr int kPer = kStage / kGemmThreads;        // = RB
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
    const float4 *ey4 = reinterpret_cast<const float4 *>(Ey + j * Np) + 4 * (lane / 32);
    const float4 *ez4 = reinterpret_cast<const float4 *>(Ez + kv * Np) + 4 * (lane / 32);

    f32x16 acc_re[RB], acc_im[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc_re[r][q] = acc_im[r][q] = 0.0f;

    uint4 pre[kPer];
    float4 cy[4], cz[4], ny[4], nz[4];
    if (k0 < k1) {
        const uint4 *src = E + (k0 * nrb + rt * RB) * 4 * 64;
#pragma unroll
        for (int e = 0; e < kPer; ++e) sE[0][tid + e * kGemmThreads] = src[tid + e * kGemmThreads];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            cy[h] = ey4[k0 * 8 + h];   // 16 nodes = 8 float4 per step
            cz[h] = ez4[k0 * 8 + h];
        }
    }
    __syncthreads();
    int buf = 0;
    for (int64_t ks = k0; ks < k1; ++ks) {
        // A operand for this lane's column and nodes ks*16 + 8*(lane/32) + q
        bf16x8 brh, brl, bih, bil;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const float4 a = cy[h], b = cz[h];
            const float r0 = a.x * b.x - a.y * b.y, i0 = a.x * b.y + a.y * b.x;
            const float r1 = a.z * b.z - a.w * b.w, i1 = a.z * b.w + a.w * b.z;
            const __bf16 r0h = (__bf16)r0, r1h = (__bf16)r1, i0h = (__bf16)i0, i1h = (__bf16)i1;
            brh[2 * h] = r0h;
            brh[2 * h + 1] = r1h;
            bih[2 * h] = i0h;
            bih[2 * h + 1] = i1h;
            brl[2 * h] = (__bf16)(r0 - (float)r0h);
            brl[2 * h + 1] = (__bf16)(r1 - (float)r1h);
            bil[2 * h] = (__bf16)(i0 - (float)i0h);
            bil[2 * h + 1] = (__bf16)(i1 - (float)i1h);
        }
        const bool has_next = ks + 1 < k1;
        if (has_next) {
            const uint4 *src = E + ((ks + 1) * nrb + rt * RB) * 4 * 64;
#pragma unroll
            for (int e = 0; e < kPer; ++e) pre[e] = src[tid + e * kGemmThreads];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                ny[h] = ey4[(ks + 1) * 8 + h];
                nz[h] = ez4[(ks + 1) * 8 + h];
            }
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const uint4 *t = &sE[buf][r * 4 * 64];
            const uint4 urh = t[0 * 64 + lane], url = t[1 * 64 + lane];
            const uint4 uih = t[2 * 64 + lane], uil = t[3 * 64 + lane];
            const unsigned sg = 0x80008000u;
            const uint4 unh = make_uint4(uih.x ^ sg, uih.y ^ sg, uih.z ^ sg, uih.w ^ sg);
            const uint4 unl = make_uint4(uil.x ^ sg, uil.y ^ sg, uil.z ^ sg, uil.w ^ sg);
            const bf16x8 erh = __builtin_bit_cast(bf16x8, urh), erl = __builtin_bit_cast(bf16x8, url);
            const bf16x8 eih = __builtin_bit_cast(bf16x8, uih), eil = __builtin_bit_cast(bf16x8, uil);
            const bf16x8 enh = __builtin_bit_cast(bf16x8, unh), enl = __builtin_bit_cast(bf16x8, unl);
            // Re += Er*Ar - Ei*Ai ; Im += Er*Ai + Ei*Ar  (each as hh + hl + lh)
            acc_re[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(erh, brh, acc_re[r], 0, 0, 0);
            acc_im[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(erh, bih, acc_im[r], 0, 0, 0);
            acc_re[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(erh, brl, acc_re[r], 0, 0, 0);
            acc_im[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(erh, bil, acc_im[r], 0, 0, 0);
            acc_re[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(erl, brh, acc_re[r], 0, 0, 0);
            acc_im[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(erl, bih, acc_im[r], 0, 0, 0);
            acc_re[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(enh, bih, acc_re[r], 0, 0, 0);
            acc_im[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(eih, brh, acc_im[r], 0, 0, 0);
            acc_re[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(enh, bil, acc_re[r], 0, 0, 0);
            acc_im[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(eih, brl, acc_im[r], 0, 0, 0);
            acc_re[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(enl, bih, acc_re[r], 0, 0, 0);
            acc_im[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(eil, brh, acc_im[r], 0, 0, 0);
        }
        if (has_next) {
#pragma unroll
            for (int e = 0; e < kPer; ++e) sE[buf ^ 1][tid + e * kGemmThreads] = pre[e];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                cy[h] = ny[h];
                cz[h] = nz[h];
            }
        }
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

// ---------------------------------------------------------------------------
// FP64 GEMM: v_mfma_f64_16x16x4_f64
// ---------------------------------------------------------------------------
// Workgroup tile: RB*16 rows x 64 columns; wave w owns columns [16w, 16w+16).
// A 16-node step is 4 MFMA k-steps of 4 nodes; lane l supplies row/column
// l%16 and node 4*st + l/16 of k-step st.
template <int RB>
__global__ __launch_bounds__(kGemmThreads) void k_cart_gemm_f64(
    const double *__restrict__ E, const double2 *__restrict__ Ey, const double2 *__restrict__ Ez, int64_t Np,
    int64_t Mp, int64_t M, int64_t NC, int64_t n_kzV, int64_t n_rt, int64_t n_ct, int64_t n_splits,
    int64_t steps_per_split, int64_t ksteps, double2 *__restrict__ part)
{
    constexpr int kStage = RB * 2 * 4 * 64;            // doubles per stage
    constexpr int kPer = kStage / kGemmThreads;        // = 2 RB
    __shared__ double sE[2][kStage];
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

    double pre[kPer];
    double2 cy[4], cz[4], ny[4], nz[4];
    if (k0 < k1) {
        const double *src = E + (k0 * nrb + rt * RB) * 2 * 4 * 64;
#pragma unroll
        for (int e = 0; e < kPer; ++e) sE[0][tid + e * kGemmThreads] = src[tid + e * kGemmThreads];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            cy[st] = eyrow[k0 * 16 + 4 * st];
            cz[st] = ezrow[k0 * 16 + 4 * st];
        }
    }
    __syncthreads();
    int buf = 0;
    for (int64_t ks = k0; ks < k1; ++ks) {
        double br[4], bi[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            br[st] = cy[st].x * cz[st].x - cy[st].y * cz[st].y;
            bi[st] = cy[st].x * cz[st].y + cy[st].y * cz[st].x;
        }
        const bool has_next = ks + 1 < k1;
        if (has_next) {
            const double *src = E + ((ks + 1) * nrb + rt * RB) * 2 * 4 * 64;
#pragma unroll
            for (int e = 0; e < kPer; ++e) pre[e] = src[tid + e * kGemmThreads];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                ny[st] = eyrow[(ks + 1) * 16 + 4 * st];
                nz[st] = ezrow[(ks + 1) * 16 + 4 * st];
            }
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const double *t = &sE[buf][r * 2 * 4 * 64];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const double er = t[(0 * 4 + st) * 64 + lane];
                const double ei = t[(1 * 4 + st) * 64 + lane];
                acc_re[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, br[st], acc_re[r], 0, 0, 0);
                acc_im[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, bi[st], acc_im[r], 0, 0, 0);
                acc_re[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ei, bi[st], acc_re[r], 0, 0, 0);
                acc_im[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei, br[st], acc_im[r], 0, 0, 0);
            }
        }
        if (has_next) {
#pragma unroll
            for (int e = 0; e < kPer; ++e) sE[buf ^ 1][tid + e * kGemmThreads] = pre[e];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                cy[st] = ny[st];
                cz[st] = nz[st];
            }
        }
        __syncthreads();
        buf ^= 1;
    }
    // acc register q of lane l holds row 4*(l/16) + q, column l%16
    if (col < NC) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t row = (rt * RB + r) * 16 + 4 * (lane / 16) + q;
                if (row < M) part[(split * M + row) * NC + col] = make_double2(acc_re[r][q], acc_im[r][q]);
            }
    }
}

template <typename P>
__global__ void k_cart_reduce(const P *__restrict__ part, int64_t n_splits, int64_t MNC, double2 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MNC) return;
    double re = 0.0, im = 0.0;
    for (int64_t s = 0; s < n_splits; ++s) {
        const P v = part[s * MNC + i];
        re += (double)v.x;
        im += (double)v.y;
    }
    out[i] = make_double2(re, im);
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

template <int RB>
static void launch_f64(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st)
{
    const unsigned grid = (unsigned)(s.n_rt * s.n_ct * s.n_splits);
    k_cart_gemm_f64<RB><<<grid, kGemmThreads, 0, st>>>(
        reinterpret_cast<const double *>(ws + s.off_E), reinterpret_cast<const double2 *>(ws + s.off_Ey),
        reinterpret_cast<const double2 *>(ws + s.off_Ez), s.Np, s.Mp, s.M, s.NC, n_kzV, s.n_rt, s.n_ct, s.n_splits,
        s.steps_per_split, s.ksteps, reinterpret_cast<double2 *>(ws + s.off_part));
}

int cartesian_execute(const fm_signal_plan *p, int precision, int64_t N, int V, const double *packed, void *wsv,
                      size_t ws_bytes, double *out, hipStream_t st)
{
    const CartesianTables &c = p->cart;
    const int64_t S = p->n_samples;
    if (S == 0) return FM_OK;
    if (N == 0) {
        FM_HIP(hipMemsetAsync(out, 0, (size_t)S * V * sizeof(double2), st));
        return FM_OK;
    }
    const GemmShape s = gemm_shape(p, precision, N, V);
    if (s.total > ws_bytes) {
        set_error("signal workspace too small: need %zu bytes, got %zu", s.total, ws_bytes);
        return FM_ERR_WORKSPACE;
    }
    char *ws = reinterpret_cast<char *>(wsv);
    const int64_t n_kzV = c.n_kz * V;
    {
        const int64_t tot = s.Mp * s.Np;
        if (s.mixed)
            k_cart_E<true><<<(unsigned)cdiv(tot, 256), 256, 0, st>>>(packed, N, s.Np, s.M, s.Mp, c.d_F, c.d_t,
                                                                    ws + s.off_E);
        else
            k_cart_E<false><<<(unsigned)cdiv(tot, 256), 256, 0, st>>>(packed, N, s.Np, s.M, s.Mp, c.d_F, c.d_t,
                                                                     ws + s.off_E);
        FM_LAUNCH_CHECK("k_cart_E");
        const int64_t tot2 = (c.n_ln + c.n_kz) * s.Np;
        if (s.mixed)
            k_cart_yz<true><<<(unsigned)cdiv(tot2, 256), 256, 0, st>>>(packed, N, s.Np, V, c.n_ln, c.n_kz, c.d_G,
                                                                      c.d_H, ws + s.off_Ey, ws + s.off_Ez);
        else
            k_cart_yz<false><<<(unsigned)cdiv(tot2, 256), 256, 0, st>>>(packed, N, s.Np, V, c.n_ln, c.n_kz, c.d_G,
                                                                       c.d_H, ws + s.off_Ey, ws + s.off_Ez);
        FM_LAUNCH_CHECK("k_cart_yz");
    }
    if (s.mixed) {
        switch (s.rb_per_wave) {
        case 1: launch_bf16<1>(s, ws, n_kzV, st); break;
        case 2: launch_bf16<2>(s, ws, n_kzV, st); break;
        case 3: launch_bf16<3>(s, ws, n_kzV, st); break;
        default: launch_bf16<4>(s, ws, n_kzV, st); break;
        }
        FM_LAUNCH_CHECK("k_cart_gemm_bf16");
        const int64_t MNC = s.M * s.NC;
        k_cart_reduce<float2><<<(unsigned)cdiv(MNC, 256), 256, 0, st>>>(
            reinterpret_cast<const float2 *>(ws + s.off_part), s.n_splits, MNC, reinterpret_cast<double2 *>(out));
    } else {
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
        FM_LAUNCH_CHECK("k_cart_gemm_f64");
        const int64_t MNC = s.M * s.NC;
        k_cart_reduce<double2><<<(unsigned)cdiv(MNC, 256), 256, 0, st>>>(
            reinterpret_cast<const double2 *>(ws + s.off_part), s.n_splits, MNC, reinterpret_cast<double2 *>(out));
    }
    FM_LAUNCH_CHECK("k_cart_reduce");
    return FM_OK;
}

}  // namespace fm
