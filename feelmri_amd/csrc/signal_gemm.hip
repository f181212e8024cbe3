// Cartesian factorised signal: a complex GEMM on the matrix cores.
//
// When every k-space coordinate splits additively over the three sample
// indices, k(i,j,k) = F(i) + G(j) + H(k) per axis, and t depends on the
// readout index i only (Cartesian stacks with one line per shot,
// python/feelmri/KSpaceTraj.py:183-220), the direct sum of
// cpp/feelmri/MRI.cpp:196-297 factorises:
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

#include "cart_gemm.h"

namespace fm {


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
    c.ro_group.resize((size_t)R);
    for (int64_t i = 0; i < R; ++i) c.ro_group[(size_t)i] = p->sample_group[(size_t)(i * L * K)];
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
    c.axis_aligned = true;
    for (int64_t j = 0; j < L; ++j) c.axis_aligned &= c.G[(size_t)(3 * j)] == 0.0 && c.G[(size_t)(3 * j + 2)] == 0.0;
    for (int64_t k = 0; k < K; ++k) c.axis_aligned &= c.H[(size_t)(3 * k)] == 0.0 && c.H[(size_t)(3 * k + 1)] == 0.0;
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
    FM_HIP(hipMalloc(&c.d_ro_group, sizeof(int32_t) * c.ro_group.size()));
    FM_HIP(hipMemcpy(c.d_ro_group, c.ro_group.data(), sizeof(int32_t) * c.ro_group.size(), hipMemcpyHostToDevice));
    return FM_OK;
}

void cartesian_free(fm_signal_plan *p)
{
    CartesianTables &c = p->cart;
    (void)hipFree(c.d_F);
    (void)hipFree(c.d_G);
    (void)hipFree(c.d_H);
    (void)hipFree(c.d_t);
    (void)hipFree(c.d_ro_group);
    c.d_F = c.d_G = c.d_H = c.d_t = nullptr;
    c.d_ro_group = nullptr;
}


// mixed: row block 32, K stage 16 nodes, column block 32 per wave (4 per WG)
// f64:   row block 16, K stage 16 nodes (4 MFMA k-steps), column block 16 per wave
static GemmShape gemm_shape(const fm_signal_plan *p, int precision, int64_t N, int V)
{
    GemmShape s{};
    const CartesianTables &c = p->cart;
    s.mixed = precision != FM_PREC_F64;
    const int rowblk = s.mixed ? 32 : 16;
    const int colblk = s.mixed ? 32 : 16;
    const int max_rb = s.mixed ? 4 : 4;
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
    // mixed: one round of workgroups (one 8-wave or two 4-wave workgroups
    // per CU); FP64: two rounds (two 4-wave workgroups per CU, measured faster)
    const int64_t target = (s.mixed ? 2 : 4) * (int64_t)device_cu_count();
    int64_t splits = cdiv(target, base);
    // mixed: FP32 accumulators sum at most 8192 steps (131 072 nodes) per
    // split before the FP64 split reduction, bounding the accumulation error
    // for large meshes
    if (s.mixed) splits = std::max<int64_t>(splits, cdiv(s.ksteps, 8192));
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

// E tables.  Mixed: tiles of (32 rows x 16 nodes) per 4 planes (Er_h, Er_l,
// Ei_h, Ei_l), each plane in MFMA lane order: lane l holds row l%32 and
// nodes 8*(l/32) .. +7 (8 bf16 = 16 bytes).  Tile index (kstep, rowblock).
// FP64: tiles of (16 rows x 16 nodes) per 2 planes (Er, Ei), plane layout
// [kstep4 (4)][lane (64)] with lane l holding row l%16, node l/16 of that
// 4-node step.  Padding rows/nodes are zero.
//
// FP64 E table, one thread per (k-step, 4-node step, lane): the thread owns
// node n = 16 kb + 4 st + l/16 and walks the row blocks (row 16 rb + l%16),
// so a wave writes 512 contiguous bytes per plane.
__device__ __forceinline__ void cart_E_f64(int64_t t, const double *__restrict__ packed, int64_t N, int64_t Np,
                                           int64_t M, int64_t Mp, const double *__restrict__ F,
                                           const double *__restrict__ tr, double *__restrict__ E)
{
    // t over (Np / 16) * 4 * 64
    if (t >= (Np / 16) * 256) return;
    const int lane = (int)(t % 64);
    const int st = (int)((t / 64) % 4);
    const int64_t kb = t / 256;
    const int64_t nrb = Mp / 16;
    const int64_t n = kb * 16 + 4 * st + lane / 16;
    const bool ok = n < N;
    const double x = ok ? packed[n] : 0.0, y = ok ? packed[N + n] : 0.0, z = ok ? packed[2 * N + n] : 0.0;
    const double phi = ok ? packed[3 * N + n] : 0.0, r2 = ok ? packed[4 * N + n] : 0.0;
    for (int64_t rb = 0; rb < nrb; ++rb) {
        const int64_t i = rb * 16 + (lane % 16);
        double er = 0.0, ei = 0.0;
        if (ok && i < M) {
            const double ti = tr[i];
            // total phase in cycles: -(F.x) + phi t / (2 pi)
            const double u = fma(F[3 * i], x, fma(F[3 * i + 1], y, F[3 * i + 2] * z)) - phi * ti * (1.0 / kTwoPi);
            double c, s;
            cis_neg_cycles<false>(u, c, s);
            const double d = exp(-ti * r2);
            er = d * c;
            ei = d * s;
        }
        const int64_t tile = (kb * nrb + rb) * 2;
        E[((tile + 0) * 4 + st) * 64 + lane] = er;
        E[((tile + 1) * 4 + st) * 64 + lane] = ei;
    }
}

// Mixed E table, one thread per lane fragment column: the thread owns the 8
// nodes n0..n0+7 that lane l (row l%32, node group l/32) of k-step kb feeds
// to the MFMA, keeps their attributes in registers and walks the row blocks,
// writing each of the 4 planes of a tile as one 16-byte store (a wave writes
// 1 KB contiguous per plane).  Phase in FP64 with range reduction, then the
// hardware FP32 sine/cosine and exp2.
__device__ __forceinline__ void cart_E_bf16(int64_t t, const double *__restrict__ packed, int64_t N, int64_t Np,
                                            int64_t M, int64_t Mp, const double *__restrict__ F,
                                            const double *__restrict__ tr, uint4 *__restrict__ E)
{
    // t over (Np / 16) * 64
    if (t >= (Np / 16) * 64) return;
    const int lane = (int)(t % 64);
    const int64_t kb = t / 64;
    const int64_t nrb = Mp / 32;
    const int64_t n0 = kb * 16 + 8 * (lane / 32);
    double x[8], y[8], z[8], ph[8], r2[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int64_t n = n0 + q;
        const bool ok = n < N;
        x[q] = ok ? packed[n] : 0.0;
        y[q] = ok ? packed[N + n] : 0.0;
        z[q] = ok ? packed[2 * N + n] : 0.0;
        ph[q] = ok ? packed[3 * N + n] : 0.0;
        r2[q] = ok ? packed[4 * N + n] : 0.0;
    }
    const int nvalid = (int)(N - n0 < 8 ? (N - n0 > 0 ? N - n0 : 0) : 8);
    for (int64_t rb = 0; rb < nrb; ++rb) {
        const int64_t i = rb * 32 + (lane % 32);
        const bool row_ok = i < M;
        const double fx = row_ok ? F[3 * i] : 0.0, fy = row_ok ? F[3 * i + 1] : 0.0,
                     fz = row_ok ? F[3 * i + 2] : 0.0;
        const double ti = row_ok ? tr[i] : 0.0;
        unsigned short h[4][8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            float er = 0.0f, ei = 0.0f;
            if (row_ok && q < nvalid) {
                const double u = fma(fx, x[q], fma(fy, y[q], fz * z[q])) - ph[q] * ti * (1.0 / kTwoPi);
                const float rf = turns_f32(u);
                const float d = __builtin_amdgcn_exp2f((float)(-ti * r2[q] * 1.4426950408889634));
                er = d * __builtin_amdgcn_cosf(rf);
                ei = -d * __builtin_amdgcn_sinf(rf);
            }
            const __bf16 erh = (__bf16)er, eih = (__bf16)ei;
            const __bf16 erl = (__bf16)(er - (float)erh), eil = (__bf16)(ei - (float)eih);
            h[0][q] = __builtin_bit_cast(unsigned short, erh);
            h[1][q] = __builtin_bit_cast(unsigned short, erl);
            h[2][q] = __builtin_bit_cast(unsigned short, eih);
            h[3][q] = __builtin_bit_cast(unsigned short, eil);
        }
        const int64_t tile = kb * nrb + rb;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
            uint4 v;
            v.x = (unsigned)h[pl][0] | ((unsigned)h[pl][1] << 16);
            v.y = (unsigned)h[pl][2] | ((unsigned)h[pl][3] << 16);
            v.z = (unsigned)h[pl][4] | ((unsigned)h[pl][5] << 16);
            v.w = (unsigned)h[pl][6] | ((unsigned)h[pl][7] << 16);
            E[(tile * 4 + pl) * 64 + lane] = v;
        }
    }
}

// Ey rows exp(-2 pi i (G_j . x_n)) and Ezb rows exp(-2 pi i (H_k . x_n)) b_v[n]
// (the encoding's weighted magnetisation folded into the kz factor).  One
// thread per node keeps its position in registers and walks a block of
// kYZRowsPerBlock rows (Ey rows first, then Ezb).
constexpr int kYZRowsPerBlock = 16;

template <bool MIXED>
__device__ __forceinline__ void cart_yz(int64_t n, int64_t rblk, const double *__restrict__ packed, int64_t N,
                                        int64_t Np, int V, int64_t n_ln, int64_t n_kz,
                                        const double *__restrict__ G, const double *__restrict__ H,
                                        void *__restrict__ Ey, void *__restrict__ Ez)
{
    if (n >= Np) return;
    const int64_t r0 = rblk * kYZRowsPerBlock;
    const int64_t r1 = (r0 + kYZRowsPerBlock < n_ln + n_kz) ? r0 + kYZRowsPerBlock : n_ln + n_kz;
    const bool ok = n < N;
    const double x = ok ? packed[n] : 0.0, y = ok ? packed[N + n] : 0.0, z = ok ? packed[2 * N + n] : 0.0;
    for (int64_t row = r0; row < r1; ++row) {
        const bool is_y = row < n_ln;
        const double *kv = is_y ? &G[3 * row] : &H[3 * (row - n_ln)];
        double c = 0.0, s = 0.0;
        if (ok) cis_neg_cycles<MIXED>(fma(kv[0], x, fma(kv[1], y, kv[2] * z)), c, s);
        if (is_y) {
            if (MIXED)
                reinterpret_cast<float2 *>(Ey)[row * Np + n] = make_float2((float)c, (float)s);
            else
                reinterpret_cast<double2 *>(Ey)[row * Np + n] = make_double2(c, s);
        } else {
            const int64_t k = row - n_ln;
            for (int v = 0; v < V; ++v) {
                double br = 0.0, bi = 0.0;
                if (ok) {
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
}

// Both operand tables in one launch: blocks [0, nb_E) build the E table,
// the rest the Ey/Ezb rows (nb_n node blocks x row blocks), so the two
// memory-bound jobs overlap instead of running back to back.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_cart_tables(const double *__restrict__ packed, int64_t N, int64_t Np,
                                                     int64_t M, int64_t Mp, const double *__restrict__ F,
                                                     const double *__restrict__ tr, void *__restrict__ E,
                                                     int64_t nb_E, int V, int64_t n_ln, int64_t n_kz,
                                                     const double *__restrict__ G, const double *__restrict__ H,
                                                     void *__restrict__ Ey, void *__restrict__ Ez, int64_t nb_n)
{
    const int64_t b = blockIdx.x;
    if (b < nb_E) {
        const int64_t t = b * 256 + threadIdx.x;
        if (MIXED)
            cart_E_bf16(t, packed, N, Np, M, Mp, F, tr, reinterpret_cast<uint4 *>(E));
        else
            cart_E_f64(t, packed, N, Np, M, Mp, F, tr, reinterpret_cast<double *>(E));
        return;
    }
    const int64_t c = b - nb_E;
    cart_yz<MIXED>((c % nb_n) * 256 + threadIdx.x, c / nb_n, packed, N, Np, V, n_ln, n_kz, G, H, Ey, Ez);
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

int cartesian_execute(const fm_signal_plan *p, int precision, int64_t N, int V, const double *packed, void *wsv,
                      size_t ws_bytes, double *out, hipStream_t st)
{
    const CartesianTables &c = p->cart;
    const int64_t S = p->n_samples;
    if (S == 0) return FM_OK;
    if (N == 0) {
        return zero_f64_async(out, 2 * S * V, st);
    }
    const GemmShape s = gemm_shape(p, precision, N, V);
    if (s.total > ws_bytes) {
        set_error("signal workspace too small: need %zu bytes, got %zu", s.total, ws_bytes);
        return FM_ERR_WORKSPACE;
    }
    char *ws = reinterpret_cast<char *>(wsv);
    const int64_t n_kzV = c.n_kz * V;
    {
        const int64_t nb_E = s.mixed ? cdiv(s.Np / 16 * 64, 256) : cdiv(s.Np / 16 * 256, 256);
        const int64_t nb_n = cdiv(s.Np, 256);
        const int64_t nb = nb_E + nb_n * cdiv(c.n_ln + c.n_kz, kYZRowsPerBlock);
        if (nb > (int64_t)0x7fffffff) {
            set_error("Cartesian operand tables exceed the launch grid (%lld blocks)", (long long)nb);
            return FM_ERR_UNSUPPORTED;
        }
        if (s.mixed)
            k_cart_tables<true><<<(unsigned)nb, 256, 0, st>>>(packed, N, s.Np, s.M, s.Mp, c.d_F, c.d_t,
                                                             ws + s.off_E, nb_E, V, c.n_ln, c.n_kz, c.d_G, c.d_H,
                                                             ws + s.off_Ey, ws + s.off_Ez, nb_n);
        else
            k_cart_tables<false><<<(unsigned)nb, 256, 0, st>>>(packed, N, s.Np, s.M, s.Mp, c.d_F, c.d_t,
                                                              ws + s.off_E, nb_E, V, c.n_ln, c.n_kz, c.d_G, c.d_H,
                                                              ws + s.off_Ey, ws + s.off_Ez, nb_n);
        FM_LAUNCH_CHECK("k_cart_tables");
    }
    if (s.mixed) {
        plan_timing_start(p, st);
        launch_cart_gemm_bf16(s, ws, n_kzV, st);
        plan_timing_stop(p, st);
        FM_LAUNCH_CHECK("k_cart_gemm_bf16");
        const int64_t MNC = s.M * s.NC;
        k_cart_reduce<float2><<<(unsigned)cdiv(MNC, 256), 256, 0, st>>>(
            reinterpret_cast<const float2 *>(ws + s.off_part), s.n_splits, MNC, reinterpret_cast<double2 *>(out));
    } else {
        plan_timing_start(p, st);
        launch_cart_gemm_f64(s, ws, n_kzV, st);
        plan_timing_stop(p, st);
        FM_LAUNCH_CHECK("k_cart_gemm_f64");
        const int64_t MNC = s.M * s.NC;
        k_cart_reduce<double2><<<(unsigned)cdiv(MNC, 256), 256, 0, st>>>(
            reinterpret_cast<const double2 *>(ws + s.off_part), s.n_splits, MNC, reinterpret_cast<double2 *>(out));
    }
    FM_LAUNCH_CHECK("k_cart_reduce");
    return FM_OK;
}

size_t cartesian_moving_workspace(const fm_signal_plan *p, int precision, int64_t N, int V)
{
    if (p->n_samples == 0 || N == 0 || V > kMovMaxV) return 0;
    return (precision == FM_PREC_F64 ? moving_shape_f64(p->cart, N, V) : moving_shape(p->cart, N, V)).part_bytes;
}

int cartesian_moving_execute(const fm_signal_plan *p, int precision, int64_t N, int V, const double *packed,
                             const float *modes, int R, const double *mcoef, void *ws, size_t ws_bytes, double *out,
                             hipStream_t st)
{
    const CartesianTables &c = p->cart;
    const int64_t S = p->n_samples;
    if (S == 0) return FM_OK;
    if (N == 0) {
        return zero_f64_async(out, 2 * S * V, st);
    }
    if (V > kMovMaxV) {
        set_error("moving Cartesian GEMM takes at most %d encodings per call (got %d)", kMovMaxV, V);
        return FM_ERR_UNSUPPORTED;
    }
    const bool f64 = precision == FM_PREC_F64;
    const MovingShape s = f64 ? moving_shape_f64(c, N, V) : moving_shape(c, N, V);
    if (s.part_bytes > ws_bytes) {
        set_error("signal workspace too small: need %zu bytes, got %zu", s.part_bytes, ws_bytes);
        return FM_ERR_WORKSPACE;
    }
    const int64_t grid = s.n_splits * s.M * s.n_rt * s.n_ct;
    if (grid > (int64_t)0x7fffffff) {
        set_error("moving Cartesian GEMM exceeds the launch grid (%lld workgroups)", (long long)grid);
        return FM_ERR_UNSUPPORTED;
    }
    const int64_t MNC = s.M * s.L * s.KV;
    plan_timing_start(p, st);
    if (f64) {
        double2 *part = reinterpret_cast<double2 *>(ws);
        launch_cart_moving_f64(s, c, packed, N, modes, R, mcoef, part, st);
        plan_timing_stop(p, st);
        FM_LAUNCH_CHECK("k_cart_moving_f64");
        k_cart_reduce<double2><<<(unsigned)cdiv(MNC, 256), 256, 0, st>>>(part, s.n_splits, MNC,
                                                                          reinterpret_cast<double2 *>(out));
    } else {
        float2 *part = reinterpret_cast<float2 *>(ws);
        launch_cart_moving_bf16(s, c, packed, N, modes, R, mcoef, part, st);
        plan_timing_stop(p, st);
        FM_LAUNCH_CHECK("k_cart_moving_bf16");
        k_cart_reduce<float2><<<(unsigned)cdiv(MNC, 256), 256, 0, st>>>(part, s.n_splits, MNC,
                                                                         reinterpret_cast<double2 *>(out));
    }
    FM_LAUNCH_CHECK("k_cart_reduce");
    return FM_OK;
}

}  // namespace fm
