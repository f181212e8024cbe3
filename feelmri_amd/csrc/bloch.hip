// Per-node Bloch integrator (Cayley-Klein rotation, then T1/T2 relaxation).
//
// reference: cpp/feelmri/BlochSimulator.cpp:151-276 (static nodes) and
// :12-148 (moving nodes), solve_mri<T>.
//
// One lane per node; node attributes come from an SoA table (coalesced).
// The event table is walked in chunks of 256 events: each lane of the
// workgroup stages one event (rf, |rf|^2, -gamma dt / 2, G, dt and, for
// moving nodes, the POD weights of that step) into LDS, then every lane
// integrates its node through the chunk reading the events by broadcast.
// Step i -> i+1 uses event i+1; e1/e2 are refreshed only when dt changes;
// |B| is floored at 1e-12; rotation first, then relaxation toward M0.
//
// Events without RF (rf = 0, the same for every lane) take the precession
// the rotation reduces to: beta = 0 and conj(alpha)^2 = exp(-i gamma dt Bz),
// so Mxy turns by -gamma dt Bz and the rotation leaves Mz alone.  This skips
// |B|, its reciprocal and the Cayley-Klein products; the result departs from
// the general step by rounding only (|alpha|^2 is 1 to an ulp).
#include <algorithm>
#include <type_traits>

#include "fm_common.h"

namespace fm {

constexpr int kBlochThreads = 256;
constexpr int kBlochChunk = 256;
constexpr int kBlochMaxModes = 16;
constexpr int kChainMaxBlocks = 128;  // blocks per chained launch

// A chain of blocks in one launch (static nodes, final state): the blocks'
// event tables are concatenated; offset[b] is block b's first event.  That
// event is not a step (step i of a block uses its event i + 1): it ends
// block b - 1 instead, applying code[b - 1] = spoil | (store column + 1) << 1
// -- the store of the state into column `col` of the (n_store, N) stores,
// then the artificial spoiling Mxy = 0 (Bloch.py:355-368).  code[n - 1] is
// applied after the last event.
struct ChainDesc {
    int64_t offset[kChainMaxBlocks + 1];
    int32_t code[kChainMaxBlocks];
    int32_t n;
};
struct NoChain {};

template <typename T>
struct alignas(16) Event {
    T rf_re, rf_im, rf2, kappa;
    T gx, gy, gz, dt;
};

template <typename T>
__device__ inline void dev_sincos(T x, T *s, T *c);
// FP64 form: sincospi of x / pi.  Its range reduction is exact (x/pi -
// rint), where sincos of a many-radian argument takes the general reduction;
// the one extra rounding of x / pi is 1.1e-16 relative.  7 % faster on the
// 1000-event, 512k-node integrator.
template <>
__device__ inline void dev_sincos<double>(double x, double *s, double *c)
{
    sincospi(x * 0.31830988618379067154, s, c);
}
// FP32 form: the hardware sine/cosine, which take revolutions, after
// reducing the argument to [-1/2, 1/2] revolution.  Their error is absolute
// (a few 1e-7), which is large relative to the small per-event rotations of
// an RF pulse (e.g. 0.013 rad per 10 us step of a 15 deg, 0.2 ms pulse) and
// accumulates over the pulse; below 0.25 rad the Taylor series to x^7 / x^6
// (truncation < 1e-10) is used instead.  The choice is per lane; the wave
// evaluates only the form(s) its lanes need (series only, hardware only, or
// both and a select).
template <>
__device__ inline void dev_sincos<float>(float x, float *s, float *c)
{
    float r = x * 0.159154943091895336f;
    r -= __builtin_rintf(r);
    const float a = r * 6.28318530717958648f;
    const bool small = __builtin_fabsf(a) < 0.25f;
    auto series = [&](float &ps, float &pc) {
        const float a2 = a * a;
        ps = a * (1.0f - a2 * (1.0f / 6.0f) * (1.0f - a2 * (1.0f / 20.0f) * (1.0f - a2 * (1.0f / 42.0f))));
        pc = 1.0f - a2 * 0.5f * (1.0f - a2 * (1.0f / 12.0f) * (1.0f - a2 * (1.0f / 30.0f)));
    };
    if (__all(small)) {
        series(*s, *c);
    } else if (!__any(small)) {
        *s = __builtin_amdgcn_sinf(r);
        *c = __builtin_amdgcn_cosf(r);
    } else {
        float ps, pc;
        series(ps, pc);
        const float hs = __builtin_amdgcn_sinf(r), hc = __builtin_amdgcn_cosf(r);
        *s = small ? ps : hs;
        *c = small ? pc : hc;
    }
}

template <typename T>
__device__ inline T dev_exp(T x);
template <>
__device__ inline double dev_exp<double>(double x) { return exp(x); }
template <>
__device__ inline float dev_exp<float>(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// T1 recovery Mz' = Mz e1 + (1 - e1) M0.  FP64 form: exactly the reference's
// expression (BlochSimulator.cpp:271).  FP32 form: Mz - om1 (Mz - M0) with
// om1 = 1 - e1 = -expm1(x) evaluated in FP64 and rounded once.  The spoiled
// steady state (1 - E1) / (1 - E1 cos a) magnifies a relative error in E1 by
// E1 / (1 - E1) (~200 for TR 5 ms, T1 1 s), so e1 rounded to FP32 (or 1 - e1
// formed in FP32) would bias Mz by ~1e-5; om1 carries the recovery with full
// FP32 relative precision.  Evaluated only when dt changes.
template <typename T>
__device__ inline T one_minus_exp(T x, T e);
template <>
__device__ inline double one_minus_exp<double>(double, double e) { return 1.0 - e; }
template <>
__device__ inline float one_minus_exp<float>(float x, float) { return (float)(-expm1((double)x)); }

template <typename T>
__device__ inline T relax_mz(T mz, T e1, T om1, T M0);
template <>
__device__ inline double relax_mz<double>(double mz, double e1, double om1, double M0) { return fma(mz, e1, om1 * M0); }
template <>
__device__ inline float relax_mz<float>(float mz, float, float om1, float M0) { return mz - om1 * (mz - M0); }

// Mz coefficient |alpha|^2 - |beta|^2 of the rotation.  FP64 form: the
// reference's expression (BlochSimulator.cpp:262, :133).  FP32 form: 1 - 2
// |beta|^2, equal to it for a unitary (alpha, beta).  In FP32, c^2 + s^2
// misses 1 by up to an ulp, the same way on every event of a readout (same
// node, same gradient), so a2 - b2 scales Mz systematically during free
// precession; 1 - 2 b2 is exactly 1 there (beta = 0).
template <typename T>
__device__ inline T mz_coef(T a2, T b2);
template <>
__device__ inline double mz_coef<double>(double a2, double b2) { return a2 - b2; }
template <>
__device__ inline float mz_coef<float>(float, float b2) { return fmaf(-2.0f, b2, 1.0f); }

template <typename T>
__device__ inline T dev_sqrt(T x);
template <>
__device__ inline double dev_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ inline float dev_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T>
__device__ inline T dev_rcp(T x);
template <>
__device__ inline double dev_rcp<double>(double x) { return 1.0 / x; }
template <>
__device__ inline float dev_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// a != b for values that are equal across the wave, decided on the scalar
// unit (first lane's bits).
__device__ inline bool uniform_ne(double a, double b)
{
    const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
    const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
    const unsigned alo = __builtin_amdgcn_readfirstlane((unsigned)ua);
    const unsigned ahi = __builtin_amdgcn_readfirstlane((unsigned)(ua >> 32));
    const unsigned blo = __builtin_amdgcn_readfirstlane((unsigned)ub);
    const unsigned bhi = __builtin_amdgcn_readfirstlane((unsigned)(ub >> 32));
    return alo != blo || ahi != bhi;
}
__device__ inline bool uniform_ne(float a, float b)
{
    return __builtin_amdgcn_readfirstlane(__builtin_bit_cast(unsigned, a)) !=
           __builtin_amdgcn_readfirstlane(__builtin_bit_cast(unsigned, b));
}
// a == 0 for a value that is equal across the wave (first lane's bits,
// either sign of zero)
__device__ inline bool uniform_zero(double a)
{
    const unsigned long long ua = __builtin_bit_cast(unsigned long long, a) << 1;
    return (__builtin_amdgcn_readfirstlane((unsigned)ua) | __builtin_amdgcn_readfirstlane((unsigned)(ua >> 32))) == 0;
}
__device__ inline bool uniform_zero(float a)
{
    return (__builtin_amdgcn_readfirstlane(__builtin_bit_cast(unsigned, a)) << 1) == 0;
}

template <typename T>
struct Cplx2;
template <>
struct Cplx2<double> {
    using type = double2;
};
template <>
struct Cplx2<float> {
    using type = float2;
};

// Chained blocks: kChainHost takes the block boundaries from a ChainDesc
// passed by value, kChainMarks from per-event marks in device memory
// (marks[e] = 1 + code of the block ending before event e, 0 inside a
// block; marks[n_time] = 1 + code of the last block) with the event count
// read from device memory as well.
constexpr int kChainNone = 0, kChainHost = 1, kChainMarks = 2;

template <typename T, bool MOTION, bool HISTORY, int CHAIN = kChainNone>
__global__ __launch_bounds__(kBlochThreads) void k_bloch(
    const double *__restrict__ packed, int64_t N, int64_t n_time, double M0d, double gammad,
    const double *__restrict__ rf, const double *__restrict__ G, const double *__restrict__ dt,
    const double *mxy_in, const double *mz_in,  // may alias mxy_fin / mz_fin
    const float *__restrict__ modes, int R, const double *__restrict__ mcoef,
    typename Cplx2<T>::type *__restrict__ mxy_out, T *__restrict__ mz_out,
    double2 *mxy_fin, double *mz_fin, const std::conditional_t<CHAIN == kChainHost, ChainDesc, NoChain> chain = {},
    float2 *__restrict__ store_xy = nullptr, float *__restrict__ store_z = nullptr,
    const int32_t *__restrict__ marks = nullptr, const int64_t *__restrict__ n_events = nullptr,
    int64_t n_store = 0)
{
    using C2 = typename Cplx2<T>::type;
    static_assert(CHAIN == kChainNone || !HISTORY, "chained blocks keep the final state only");
    __shared__ Event<T> sev[kBlochChunk];
    __shared__ T scoef[MOTION ? kBlochChunk * (kBlochMaxModes + 3) : 1];
    __shared__ int32_t smark[CHAIN != kChainNone ? kBlochChunk : 1];  // 0: a step; else 1 + block-end code
    if constexpr (CHAIN == kChainMarks) {
        // the event count is produced on the device; n_time is the capacity
        const int64_t ne = *n_events;
        n_time = ne < n_time ? ne : n_time;
    }

    const int tid = threadIdx.x;
    const int64_t p = (int64_t)blockIdx.x * kBlochThreads + tid;
    const bool active = p < N;
    const T M0 = (T)M0d, gamma = (T)gammad;

    T x = 0, y = 0, z = 0, invT1 = 0, invT2 = 0, dB = 0;
    T mxr = 0, mxi = 0, mz = 0;
    float md[MOTION ? 3 * kBlochMaxModes : 1];
    if (active) {
        x = (T)packed[p];
        y = (T)packed[N + p];
        z = (T)packed[2 * N + p];
        invT1 = (T)packed[3 * N + p];
        invT2 = (T)packed[4 * N + p];
        dB = (T)packed[5 * N + p];
        mxr = (T)mxy_in[2 * p];
        mxi = (T)mxy_in[2 * p + 1];
        mz = (T)mz_in[p];
        if (MOTION) {
#pragma unroll
            for (int r = 0; r < kBlochMaxModes; ++r) {
                if (r < R) {
                    md[r] = modes[(int64_t)r * N + p];
                    md[kBlochMaxModes + r] = modes[(int64_t)(R + r) * N + p];
                    md[2 * kBlochMaxModes + r] = modes[(int64_t)(2 * R + r) * N + p];
                }
            }
        }
        if (HISTORY) {
            C2 o;
            o.x = mxr;
            o.y = mxi;
            mxy_out[p] = o;
            mz_out[p] = mz;
        }
    }
    T e1 = 0, e2 = 0, om1 = 0, dt_prev = (T)-1;
    const int64_t n_steps = n_time > 0 ? n_time - 1 : 0;
    const int RC = R + 3;

    for (int64_t base = 0; base < n_steps; base += kBlochChunk) {
        __syncthreads();
        {
            const int64_t i = base + tid;  // step i uses event i + 1
            Event<T> ev;
            if (i < n_steps) {
                const double rr = rf[2 * (i + 1)], ri = rf[2 * (i + 1) + 1];
                ev.rf_re = (T)rr;
                ev.rf_im = (T)ri;
                ev.rf2 = ev.rf_re * ev.rf_re + ev.rf_im * ev.rf_im;
                ev.dt = (T)dt[i + 1];
                ev.kappa = -(T)0.5 * gamma * ev.dt;
                ev.gx = (T)G[3 * (i + 1)];
                ev.gy = (T)G[3 * (i + 1) + 1];
                ev.gz = (T)G[3 * (i + 1) + 2];
            } else {
                ev.rf_re = ev.rf_im = ev.rf2 = ev.kappa = ev.gx = ev.gy = ev.gz = ev.dt = 0;
            }
            sev[tid] = ev;
            if constexpr (CHAIN == kChainMarks) {
                smark[tid] = i < n_steps ? marks[i + 1] : 0;
            }
            if constexpr (CHAIN == kChainHost) {
                // event i + 1 starts block b > 0: it ends block b - 1
                int mk = 0;
                if (i < n_steps) {
                    int lo = 1, hi = chain.n - 1;
                    while (lo <= hi) {
                        const int mid = (lo + hi) / 2;
                        if (chain.offset[mid] == i + 1) {
                            mk = 1 + chain.code[mid - 1];
                            break;
                        }
                        if (chain.offset[mid] < i + 1) lo = mid + 1;
                        else hi = mid - 1;
                    }
                }
                smark[tid] = mk;
            }
            if (MOTION) {
                for (int q = 0; q < RC; ++q)
                    scoef[tid * RC + q] = (i < n_steps) ? (T)mcoef[i * RC + q] : (T)0;
            }
        }
        __syncthreads();
        const int cnt = (n_steps - base < kBlochChunk) ? (int)(n_steps - base) : kBlochChunk;
        if (active) {
            for (int e = 0; e < cnt; ++e) {
                if constexpr (CHAIN != kChainNone) {
                    const int mk = smark[e];
                    if (mk != 0) {  // the same for every lane: a block ends here
                        const int col = ((mk - 1) >> 1) - 1;
                        if (col >= 0 && (CHAIN == kChainHost || col < n_store)) {
                            store_xy[p + col * N] = make_float2((float)mxr, (float)mxi);
                            store_z[p + col * N] = (float)mz;
                        }
                        if ((mk - 1) & 1) mxr = mxi = 0;
                        continue;
                    }
                }
                const Event<T> ev = sev[e];
                // BlochSimulator.cpp:207-214: e1/e2 only when dt changes.  dt
                // is the same for every lane (LDS broadcast), so test it on
                // the scalar unit: a real branch, not two exp per event
                // under a lane mask.
                if (uniform_ne(ev.dt, dt_prev)) {
                    e1 = dev_exp<T>(-ev.dt * invT1);
                    e2 = dev_exp<T>(-ev.dt * invT2);
                    om1 = one_minus_exp<T>(-ev.dt * invT1, e1);
                    dt_prev = ev.dt;
                }
                T px = x, py = y, pz = z;
                if (MOTION) {
                    const T *cf = &scoef[e * RC];
                    T dx = cf[R], dy = cf[R + 1], dz = cf[R + 2];
#pragma unroll
                    for (int r = 0; r < kBlochMaxModes; ++r) {
                        if (r < R) {
                            dx = fma((T)md[r], cf[r], dx);
                            dy = fma((T)md[kBlochMaxModes + r], cf[r], dy);
                            dz = fma((T)md[2 * kBlochMaxModes + r], cf[r], dz);
                        }
                    }
                    px += dx;
                    py += dy;
                    pz += dz;
                }
                const T Bz = fma(pz, ev.gz, fma(py, ev.gy, px * ev.gx)) + dB;
                if (uniform_zero(ev.rf2)) {
                    // free precession: Mxy exp(-i gamma dt Bz); kappa = -gamma dt / 2
                    T s, c;
                    dev_sincos<T>((ev.kappa + ev.kappa) * Bz, &s, &c);
                    const T nmr = fma(c, mxr, -(s * mxi)), nmi = fma(c, mxi, s * mxr);
                    mxr = nmr * e2;
                    mxi = nmi * e2;
                    mz = relax_mz<T>(mz, e1, om1, M0);
                    if (HISTORY) {
                        const int64_t o = p + (base + e + 1) * N;
                        C2 v;
                        v.x = mxr;
                        v.y = mxi;
                        mxy_out[o] = v;
                        mz_out[o] = mz;
                    }
                    continue;
                }
                T Bn = dev_sqrt<T>(fma(Bz, Bz, ev.rf2));
                if (Bn < (T)1e-12) Bn = (T)1e-12;
                const T inv = dev_rcp<T>(Bn);
                const T nz = Bz * inv;
                T s, c;
                dev_sincos<T>(Bn * ev.kappa, &s, &c);
                // alpha = c - i nz s ; beta = -i (rf/|B|) s
                const T ar = c, ai = -nz * s;
                const T br = ev.rf_im * inv * s, bi = -ev.rf_re * inv * s;
                // every sum of two products is written as one fma and one
                // product, so that each instantiation of this kernel rounds
                // alike (a free contraction may fuse either product)
                const T a2 = fma(ar, ar, ai * ai), b2 = fma(br, br, bi * bi);
                const T car = ar, cai = -ai;
                const T cab_r = fma(car, br, -(cai * bi)), cab_i = fma(car, bi, cai * br);
                const T ca2_r = fma(car, car, -(cai * cai)), ca2_i = ((T)2 * car) * cai;
                const T bb_r = fma(br, br, -(bi * bi)), bb_i = ((T)2 * br) * bi;
                const T cmr = mxr, cmi = -mxi;
                const T nmr = fma((T)2 * cab_r, mz, fma(ca2_r, mxr, -(ca2_i * mxi)) - fma(bb_r, cmr, -(bb_i * cmi)));
                const T nmi = fma((T)2 * cab_i, mz, fma(ca2_r, mxi, ca2_i * mxr) - fma(bb_r, cmi, bb_i * cmr));
                const T ab_r = fma(ar, br, -(ai * bi)), ab_i = fma(ar, bi, ai * br);
                const T nmz = fma(mz_coef<T>(a2, b2), mz, -((T)2 * fma(ab_r, cmr, -(ab_i * cmi))));
                mxr = nmr * e2;
                mxi = nmi * e2;
                mz = relax_mz<T>(nmz, e1, om1, M0);
                if (HISTORY) {
                    const int64_t o = p + (base + e + 1) * N;
                    C2 v;
                    v.x = mxr;
                    v.y = mxi;
                    mxy_out[o] = v;
                    mz_out[o] = mz;
                }
            }
        }
    }
    if constexpr (CHAIN != kChainNone) {
        if (active) {  // the last block's end
            int code;
            if constexpr (CHAIN == kChainHost) code = chain.code[chain.n - 1];
            else code = marks[n_time] - 1;
            const int col = (code >> 1) - 1;
            if (col >= 0 && (CHAIN == kChainHost || col < n_store)) {
                store_xy[p + col * N] = make_float2((float)mxr, (float)mxi);
                store_z[p + col * N] = (float)mz;
            }
            if (code & 1) mxr = mxi = 0;
        }
    }
    if (!HISTORY && active) {
        mxy_fin[p] = make_double2((double)mxr, (double)mxi);
        mz_fin[p] = (double)mz;
    }
}

template <typename T>
__global__ void k_bloch_pack(int64_t N, const T *__restrict__ r0, const T *__restrict__ T1,
                             const T *__restrict__ T2, const T *__restrict__ dB, double *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = (double)r0[3 * n];
    out[N + n] = (double)r0[3 * n + 1];
    out[2 * N + n] = (double)r0[3 * n + 2];
    out[3 * N + n] = 1.0 / (double)T1[n];
    out[4 * N + n] = 1.0 / (double)T2[n];
    out[5 * N + n] = (double)dB[n];
}

__global__ void k_store_spoil(int64_t N, double2 *__restrict__ mxy, const double *__restrict__ mz,
                              float2 *__restrict__ mxy_store, float *__restrict__ mz_store, int64_t col,
                              int spoil)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double2 v = mxy[n];
    if (col >= 0) {
        mxy_store[n + col * N] = make_float2((float)v.x, (float)v.y);
        mz_store[n + col * N] = (float)mz[n];
    }
    if (spoil) mxy[n] = make_double2(0.0, 0.0);
}

template <typename T>
static int launch_bloch(bool motion, bool history, int64_t N, int64_t n_time, const double *packed,
                        double M0, double gamma, const double *rf, const double *G, const double *dt,
                        const double *mxy_in, const double *mz_in, const float *modes, int R,
                        const double *mcoef, void *mxy_out, void *mz_out, hipStream_t st)
{
    using C2 = typename Cplx2<T>::type;
    const unsigned blocks = (unsigned)cdiv(N, kBlochThreads);
    C2 *mo = reinterpret_cast<C2 *>(mxy_out);
    T *zo = reinterpret_cast<T *>(mz_out);
    double2 *mf = reinterpret_cast<double2 *>(mxy_out);
    double *zf = reinterpret_cast<double *>(mz_out);
#define FM_BLOCH_ARGS packed, N, n_time, M0, gamma, rf, G, dt, mxy_in, mz_in, modes, R, mcoef, mo, zo, mf, zf
    if (motion) {
        if (history)
            k_bloch<T, true, true><<<blocks, kBlochThreads, 0, st>>>(FM_BLOCH_ARGS);
        else
            k_bloch<T, true, false><<<blocks, kBlochThreads, 0, st>>>(FM_BLOCH_ARGS);
    } else {
        if (history)
            k_bloch<T, false, true><<<blocks, kBlochThreads, 0, st>>>(FM_BLOCH_ARGS);
        else
            k_bloch<T, false, false><<<blocks, kBlochThreads, 0, st>>>(FM_BLOCH_ARGS);
    }
#undef FM_BLOCH_ARGS
    FM_LAUNCH_CHECK("k_bloch");
    return FM_OK;
}

template <typename T>
static void launch_bloch_chain(int64_t N, int64_t n_time, const double *packed, double M0, double gamma,
                               const double *rf, const double *G, const double *dt, const double *mxy_in,
                               const double *mz_in, double *mxy_out, double *mz_out, const ChainDesc &cd,
                               float2 *store_xy, float *store_z, hipStream_t st)
{
    const unsigned blocks = (unsigned)cdiv(N, kBlochThreads);
    k_bloch<T, false, false, kChainHost><<<blocks, kBlochThreads, 0, st>>>(
        packed, N, n_time, M0, gamma, rf, G, dt, mxy_in, mz_in, nullptr, 0, nullptr, nullptr, nullptr,
        reinterpret_cast<double2 *>(mxy_out), mz_out, cd, store_xy, store_z);
}

template <typename T>
static void launch_bloch_marked(bool motion, int64_t N, int64_t capacity, const int64_t *n_events,
                                const double *packed, double M0, double gamma, const double *rf, const double *G,
                                const double *dt, const int32_t *marks, const float *modes, int R,
                                const double *mcoef, const double *mxy_in, const double *mz_in, double *mxy_out,
                                double *mz_out, float2 *store_xy, float *store_z, int64_t n_store, hipStream_t st)
{
    const unsigned blocks = (unsigned)cdiv(N, kBlochThreads);
    double2 *mf = reinterpret_cast<double2 *>(mxy_out);
    if (motion)
        k_bloch<T, true, false, kChainMarks><<<blocks, kBlochThreads, 0, st>>>(
            packed, N, capacity, M0, gamma, rf, G, dt, mxy_in, mz_in, modes, R, mcoef, nullptr, nullptr, mf, mz_out,
            NoChain{}, store_xy, store_z, marks, n_events, n_store);
    else
        k_bloch<T, false, false, kChainMarks><<<blocks, kBlochThreads, 0, st>>>(
            packed, N, capacity, M0, gamma, rf, G, dt, mxy_in, mz_in, nullptr, 0, nullptr, nullptr, nullptr, mf,
            mz_out, NoChain{}, store_xy, store_z, marks, n_events, n_store);
}

}  // namespace fm

using namespace fm;

extern "C" int fm_bloch_nodes_pack(int64_t n_pos, int dtype, const void *r0, const void *T1,
                                   const void *T2, const void *delta_B, double *packed, void *stream)
{
    FM_CHECK_ARG(n_pos >= 0, "negative n_pos");
    FM_CHECK_ARG(dtype == FM_DTYPE_F32 || dtype == FM_DTYPE_F64, "unknown dtype %d", dtype);
    if (n_pos == 0) return FM_OK;
    FM_CHECK_ARG(r0 && T1 && T2 && delta_B && packed, "NULL array");
    const unsigned blocks = (unsigned)cdiv(n_pos, 256);
    hipStream_t st = as_stream(stream);
    if (dtype == FM_DTYPE_F64)
        k_bloch_pack<double><<<blocks, 256, 0, st>>>(n_pos, (const double *)r0, (const double *)T1,
                                                    (const double *)T2, (const double *)delta_B, packed);
    else
        k_bloch_pack<float><<<blocks, 256, 0, st>>>(n_pos, (const float *)r0, (const float *)T1,
                                                   (const float *)T2, (const float *)delta_B, packed);
    FM_LAUNCH_CHECK("k_bloch_pack");
    return FM_OK;
}

extern "C" int fm_bloch_execute(int precision, int64_t n_pos, int64_t n_time, const double *packed,
                                double M0, double gamma, const double *rf, const double *G,
                                const double *dt, const double *mxy_in, const double *mz_in,
                                const float *modes, int n_modes, const double *motion_coef,
                                void *mxy_out, void *mz_out, int history, void *stream)
{
    FM_CHECK_ARG(precision == FM_PREC_F64 || precision == FM_PREC_F32,
                 "Bloch precision must be FM_PREC_F64 or FM_PREC_F32 (got %d)", precision);
    FM_CHECK_ARG(n_pos >= 0 && n_time >= 0, "negative size");
    FM_CHECK_ARG(n_modes >= 0 && n_modes <= kBlochMaxModes, "n_modes must be in [0, %d]", kBlochMaxModes);
    if (n_pos == 0) return FM_OK;
    FM_CHECK_ARG(packed && mxy_in && mz_in && mxy_out && mz_out, "NULL array");
    FM_CHECK_ARG(n_time == 0 || (rf && G && dt), "NULL event table");
    const bool motion = modes != nullptr && n_modes > 0;
    FM_CHECK_ARG(!motion || motion_coef != nullptr, "motion_coef is NULL");
    if (n_time == 0) {
        set_error("empty event table (n_time = 0)");
        return FM_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    if (precision == FM_PREC_F64)
        return launch_bloch<double>(motion, history != 0, n_pos, n_time, packed, M0, gamma, rf, G, dt, mxy_in,
                                    mz_in, modes, n_modes, motion_coef, mxy_out, mz_out, st);
    return launch_bloch<float>(motion, history != 0, n_pos, n_time, packed, M0, gamma, rf, G, dt, mxy_in,
                               mz_in, modes, n_modes, motion_coef, mxy_out, mz_out, st);
}

extern "C" int fm_bloch_store_spoil(int64_t n_pos, double *mxy_state, const double *mz_state,
                                    void *mxy_store, void *mz_store, int64_t store_col, int spoil,
                                    void *stream)
{
    FM_CHECK_ARG(n_pos >= 0, "negative n_pos");
    if (n_pos == 0) return FM_OK;
    FM_CHECK_ARG(mxy_state && mz_state, "NULL state");
    FM_CHECK_ARG(store_col < 0 || (mxy_store && mz_store), "NULL store");
    k_store_spoil<<<(unsigned)cdiv(n_pos, 256), 256, 0, as_stream(stream)>>>(
        n_pos, reinterpret_cast<double2 *>(mxy_state), mz_state, reinterpret_cast<float2 *>(mxy_store),
        reinterpret_cast<float *>(mz_store), store_col, spoil);
    FM_LAUNCH_CHECK("k_store_spoil");
    return FM_OK;
}

extern "C" int fm_bloch_chain(int precision, int64_t n_pos, const double *packed, double M0, double gamma,
                              int64_t n_blocks, const int64_t *block_len, const int32_t *store_col,
                              const int32_t *spoil, const double *rf, const double *G, const double *dt,
                              const double *mxy_in, const double *mz_in, void *mxy_store, void *mz_store,
                              int64_t n_store, double *mxy_out, double *mz_out, void *stream)
{
    FM_CHECK_ARG(precision == FM_PREC_F64 || precision == FM_PREC_F32,
                 "Bloch precision must be FM_PREC_F64 or FM_PREC_F32 (got %d)", precision);
    FM_CHECK_ARG(n_pos >= 0 && n_blocks >= 0 && n_store >= 0, "negative size");
    if (n_pos == 0) return FM_OK;
    FM_CHECK_ARG(mxy_in && mz_in && mxy_out && mz_out, "NULL state");
    hipStream_t st = as_stream(stream);
    if (n_blocks == 0) {  // no block: the state passes through
        hipError_t e = hipSuccess;
        if (mxy_out != mxy_in)
            e = hipMemcpyAsync(mxy_out, mxy_in, 2 * sizeof(double) * (size_t)n_pos, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && mz_out != mz_in)
            e = hipMemcpyAsync(mz_out, mz_in, sizeof(double) * (size_t)n_pos, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) {
            set_error("fm_bloch_chain: %s", hipGetErrorString(e));
            return FM_ERR_HIP;
        }
        return FM_OK;
    }
    FM_CHECK_ARG(block_len && store_col && spoil, "NULL block description");
    FM_CHECK_ARG(packed && rf && G && dt, "NULL array");
    for (int64_t b = 0; b < n_blocks; ++b) {
        FM_CHECK_ARG(block_len[b] >= 1, "block %lld has no event", (long long)b);
        FM_CHECK_ARG(store_col[b] >= -1 && store_col[b] < n_store, "block %lld: store column %d outside [-1, %lld)",
                     (long long)b, store_col[b], (long long)n_store);
    }
    FM_CHECK_ARG(n_store == 0 || (mxy_store && mz_store), "NULL store");
    int64_t ev0 = 0;
    for (int64_t b0 = 0; b0 < n_blocks; b0 += kChainMaxBlocks) {
        const int nb = (int)std::min<int64_t>(kChainMaxBlocks, n_blocks - b0);
        ChainDesc cd;
        cd.n = nb;
        int64_t o = 0;
        for (int k = 0; k < nb; ++k) {
            cd.offset[k] = o;
            cd.code[k] = (spoil[b0 + k] ? 1 : 0) | ((store_col[b0 + k] + 1) << 1);
            o += block_len[b0 + k];
        }
        cd.offset[nb] = o;
        // later launches continue from the state the previous one left
        const double *xin = b0 == 0 ? mxy_in : mxy_out;
        const double *zin = b0 == 0 ? mz_in : mz_out;
        if (precision == FM_PREC_F64)
            launch_bloch_chain<double>(n_pos, o, packed, M0, gamma, rf + 2 * ev0, G + 3 * ev0, dt + ev0, xin, zin,
                                       mxy_out, mz_out, cd, reinterpret_cast<float2 *>(mxy_store),
                                       reinterpret_cast<float *>(mz_store), st);
        else
            launch_bloch_chain<float>(n_pos, o, packed, M0, gamma, rf + 2 * ev0, G + 3 * ev0, dt + ev0, xin, zin,
                                      mxy_out, mz_out, cd, reinterpret_cast<float2 *>(mxy_store),
                                      reinterpret_cast<float *>(mz_store), st);
        FM_LAUNCH_CHECK("k_bloch (chained blocks)");
        ev0 += o;
    }
    return FM_OK;
}

extern "C" int fm_bloch_chain_marked(int precision, int64_t n_pos, const double *packed, double M0, double gamma,
                                     const int64_t *n_events, int64_t capacity, const double *rf, const double *G,
                                     const double *dt, const int32_t *marks, const float *modes, int n_modes,
                                     const double *motion_coef, const double *mxy_in, const double *mz_in,
                                     void *mxy_store, void *mz_store, int64_t n_store, double *mxy_out,
                                     double *mz_out, void *stream)
{
    FM_CHECK_ARG(precision == FM_PREC_F64 || precision == FM_PREC_F32,
                 "Bloch precision must be FM_PREC_F64 or FM_PREC_F32 (got %d)", precision);
    FM_CHECK_ARG(n_pos >= 0 && capacity >= 1 && n_store >= 0, "negative size or capacity < 1");
    FM_CHECK_ARG(n_modes >= 0 && n_modes <= kBlochMaxModes, "n_modes must be in [0, %d]", kBlochMaxModes);
    if (n_pos == 0) return FM_OK;
    FM_CHECK_ARG(packed && n_events && rf && G && dt && marks, "NULL array");
    FM_CHECK_ARG(mxy_in && mz_in && mxy_out && mz_out, "NULL state");
    FM_CHECK_ARG(n_store == 0 || (mxy_store && mz_store), "NULL store");
    const bool motion = modes != nullptr && n_modes > 0;
    FM_CHECK_ARG(!motion || motion_coef != nullptr, "motion_coef is NULL");
    hipStream_t st = as_stream(stream);
    float2 *sx = reinterpret_cast<float2 *>(mxy_store);
    float *sz = reinterpret_cast<float *>(mz_store);
    if (precision == FM_PREC_F64)
        launch_bloch_marked<double>(motion, n_pos, capacity, n_events, packed, M0, gamma, rf, G, dt, marks, modes,
                                    n_modes, motion_coef, mxy_in, mz_in, mxy_out, mz_out, sx, sz, n_store, st);
    else
        launch_bloch_marked<float>(motion, n_pos, capacity, n_events, packed, M0, gamma, rf, G, dt, marks, modes,
                                   n_modes, motion_coef, mxy_in, mz_in, mxy_out, mz_out, sx, sz, n_store, st);
    FM_LAUNCH_CHECK("k_bloch (marked chain)");
    return FM_OK;
}
