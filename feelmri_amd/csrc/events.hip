// Sequence event tables built on the device from block descriptions.
//
// reference: python/feelmri/Bloch.py:105-129 (SequenceBlock._discretization:
// the union of the dt grid over the block extent, the RF start/end times and
// dt_rf grids, the gradient knots and dt_gr grids, sorted and
// de-duplicated), :44-49 (__call__: RF and gradients summed over the block's
// pulses and gradients at each time) and :340 (dt = diff(times, prepend=0)).
// Pulses are hard (the reference's RF table of one constant value, read by
// linear interpolation with 0 outside the closed interval [start, end],
// python/feelmri/MRObjects.py:591-614), gradients trapezoids evaluated as
// numpy's interp over their four knots.
//
// Three launches, no host round trip: (1) one workgroup per block generates
// its candidate times into LDS, sorts them (bitonic) and counts the distinct
// ones; (2) one workgroup scans the counts into block offsets; (3) each
// block's workgroup regenerates and sorts its candidates and writes the
// distinct times, dt, RF, gradients and the block-boundary marks at its
// offset.  Regenerating costs less than a scratch table of every block's
// sorted candidates.
//
// Every value is formed as the host rule forms it, operation by operation,
// with contraction off, so the tables are bitwise those of the host:
// arange as numpy fills it (start, start + step, start + i * ((start + step)
// - start)), interp with numpy's branches (left/right zero, knot hits, slope
// times offset plus value), sums in the pulses' and gradients' order.
#include <cmath>

#include "fm_common.h"

namespace fm {

constexpr int kEvThreads = 256;
constexpr int kEvMaxSegs = 4 * FM_EVENTS_MAX_PARTS + 1;

// A run of candidate times: an arange (kind 0: a = start, b = step), the
// four knots of a trapezoid (kind 1: idx), or a pulse's start and end
// (kind 2: idx).
struct EvSeg {
    double a, b;
    int64_t n;
    int kind, idx;
};

__host__ __device__ inline int64_t arange_len(double start, double stop, double step)
{
#pragma clang fp contract(off)
    const double n = ceil((stop - start) / step);
    return n > 0.0 ? (n < 9.0e15 ? (int64_t)n : (int64_t)9e15) : 0;
}

__host__ __device__ inline double arange_at(double start, double step, int64_t i)
{
#pragma clang fp contract(off)
    if (i == 0) return start;
    const double second = start + step;
    if (i == 1) return second;
    const double delta = second - start;
    return start + (double)i * delta;
}

__host__ __device__ inline void trap_knots(const fm_trapezoid &g, double x[4])
{
#pragma clang fp contract(off)
    x[0] = g.start;
    x[1] = g.start + g.ramp;
    x[2] = x[1] + g.flat;
    x[3] = (g.start + 2.0 * g.ramp) + g.flat;
}

// The candidate runs of one block, in the host rule's order (the order does
// not change the sorted result).  Returns the run count, or -1 when the
// block has more parts than FM_EVENTS_MAX_PARTS.
__host__ __device__ inline int block_segments(const fm_block_desc &bd, const fm_rf_pulse *pulses,
                                              const fm_trapezoid *traps, EvSeg *seg, int64_t *n_cand)
{
#pragma clang fp contract(off)
    if (bd.n_pulses < 0 || bd.n_traps < 0 || bd.n_pulses > FM_EVENTS_MAX_PARTS || bd.n_traps > FM_EVENTS_MAX_PARTS)
        return -1;
    int ns = 0;
    int64_t total = 0;
    double ext0 = bd.t0, ext1 = bd.t0 + bd.dur;
    for (int k = 0; k < bd.n_traps; ++k) {
        const fm_trapezoid &g = traps[bd.trap0 + k];
        double x[4];
        trap_knots(g, x);
        seg[ns++] = EvSeg{0.0, 0.0, 4, 1, bd.trap0 + k};
        total += 4;
        ext0 = x[0] < ext0 ? x[0] : ext0;
        ext1 = x[3] > ext1 ? x[3] : ext1;
        if (bd.dt_gr > 0.0) {
            const int64_t n = arange_len(x[0], x[3], bd.dt_gr);
            seg[ns++] = EvSeg{x[0], bd.dt_gr, n, 0, 0};
            total += n;
        }
    }
    for (int k = 0; k < bd.n_pulses; ++k) {
        const fm_rf_pulse &p = pulses[bd.pulse0 + k];
        const double end = p.start + p.dur;
        seg[ns++] = EvSeg{0.0, 0.0, 2, 2, bd.pulse0 + k};
        total += 2;
        ext0 = p.start < ext0 ? p.start : ext0;
        ext1 = end > ext1 ? end : ext1;
        if (bd.dt_rf > 0.0) {
            const int64_t n = arange_len(p.start, end, bd.dt_rf);
            seg[ns++] = EvSeg{p.start, bd.dt_rf, n, 0, 0};
            total += n;
        }
    }
    const int64_t n = arange_len(ext0, ext1, bd.dt);
    seg[ns++] = EvSeg{ext0, bd.dt, n, 0, 0};
    total += n;
    *n_cand = total;
    return ns;
}

__device__ inline double seg_value(const EvSeg &s, int64_t j, const fm_rf_pulse *pulses,
                                   const fm_trapezoid *traps)
{
#pragma clang fp contract(off)
    if (s.kind == 0) return arange_at(s.a, s.b, j);
    if (s.kind == 1) {
        double x[4];
        trap_knots(traps[s.idx], x);
        return x[j];
    }
    const fm_rf_pulse &p = pulses[s.idx];
    return j == 0 ? p.start : p.start + p.dur;
}

// A pulse's value at t: its constant amp on [start, start + dur], or numpy's
// interp(t, linspace(start, start + dur, n), table, left=0, right=0) -- the
// grid j * step + start with step = (stop - start) / (n - 1), the last point
// stop, the knot hits, slope times offset plus value.
__device__ inline void pulse_value(const fm_rf_pulse &p, const double *__restrict__ tables, double t, double &re,
                                   double &im)
{
#pragma clang fp contract(off)
    const double stop = p.start + p.dur;
    re = im = 0.0;
    if (t < p.start || t > stop) return;
    if (p.n_table == 0) {
        re = p.amp_re;
        im = p.amp_im;
        return;
    }
    const int n = p.n_table;
    const double *f = tables + 2 * p.table0;
    if (t == stop) {
        re = f[2 * (n - 1)];
        im = f[2 * (n - 1) + 1];
        return;
    }
    const double step = (stop - p.start) / (double)(n - 1);
    auto X = [&](int j) { return j == n - 1 ? stop : (double)j * step + p.start; };
    int lo = 0, hi = n - 1;  // X(lo) <= t < X(hi)
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (X(mid) <= t) lo = mid;
        else hi = mid;
    }
    const double x0 = X(lo), x1 = X(lo + 1);
    if (x0 == t) {
        re = f[2 * lo];
        im = f[2 * lo + 1];
        return;
    }
    const double c[2][2] = {{f[2 * lo], f[2 * lo + 2]}, {f[2 * lo + 1], f[2 * lo + 3]}};
    double out[2];
    for (int k = 0; k < 2; ++k) {
        const double slope = (c[k][1] - c[k][0]) / (x1 - x0);
        double r = slope * (t - x0) + c[k][0];
        if (isnan(r)) {
            r = slope * (t - x1) + c[k][1];
            if (isnan(r) && c[k][0] == c[k][1]) r = c[k][0];
        }
        out[k] = r;
    }
    re = out[0];
    im = out[1];
}

// numpy's interp(t, knots, [0, amp, amp, 0], left=0, right=0)
__device__ inline double trap_value(const fm_trapezoid &g, double t)
{
#pragma clang fp contract(off)
    double x[4];
    trap_knots(g, x);
    const double y[4] = {0.0, g.amp, g.amp, 0.0};
    if (t > x[3] || t < x[0]) return 0.0;
    int i = 1;
    while (i < 4 && t >= x[i]) ++i;
    const int j = i - 1;
    if (j == 3) return y[3];
    if (x[j] == t) return y[j];
    const double slope = (y[j + 1] - y[j]) / (x[j + 1] - x[j]);
    double r = slope * (t - x[j]) + y[j];
    if (isnan(r)) {
        r = slope * (t - x[j + 1]) + y[j + 1];
        if (isnan(r) && y[j] == y[j + 1]) r = y[j];
    }
    return r;
}

struct EvShared {
    double v[FM_EVENTS_MAX_CANDIDATES];
    EvSeg seg[kEvMaxSegs];
    int64_t seg_end[kEvMaxSegs];
    int32_t count[kEvThreads];
    int32_t nseg, n_cand, P;
};

// Candidates of block b into sh.v[0, n_cand), sorted, padded to a power of
// two with +inf.  Returns false (for the whole workgroup) when the block
// does not fit.
__device__ bool block_sorted_times(EvShared &sh, const fm_block_desc &bd, const fm_rf_pulse *pulses,
                                   const fm_trapezoid *traps, int64_t max_cand)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        int64_t n = 0;
        const int ns = block_segments(bd, pulses, traps, sh.seg, &n);
        int64_t acc = 0;
        for (int k = 0; k < (ns > 0 ? ns : 0); ++k) {
            acc += sh.seg[k].n;
            sh.seg_end[k] = acc;
        }
        const bool ok = ns > 0 && n <= max_cand && n <= FM_EVENTS_MAX_CANDIDATES;
        sh.nseg = ok ? ns : -1;
        sh.n_cand = ok ? (int32_t)n : 0;
        int P = 1;
        while (P < sh.n_cand) P <<= 1;
        sh.P = P;
    }
    __syncthreads();
    if (sh.nseg < 0) return false;
    const int n = sh.n_cand, P = sh.P;
    for (int i = tid; i < P; i += kEvThreads) {
        double val = INFINITY;
        if (i < n) {
            int k = 0;
            while (sh.seg_end[k] <= i) ++k;
            const int64_t j = i - (sh.seg_end[k] - sh.seg[k].n);
            val = seg_value(sh.seg[k], j, pulses, traps);
        }
        sh.v[i] = val;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P; i += kEvThreads) {
                const int l = i ^ j;
                if (l > i) {
                    const double a = sh.v[i], c = sh.v[l];
                    if ((a > c) == ((i & k) == 0)) {
                        sh.v[i] = c;
                        sh.v[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    return true;
}

// distinct values among this thread's slice [lo, hi) of the sorted times
__device__ inline int32_t count_distinct(const EvShared &sh, int lo, int hi)
{
    int32_t c = 0;
    for (int i = lo; i < hi; ++i) c += (i == 0 || sh.v[i] != sh.v[i - 1]) ? 1 : 0;
    return c;
}

__global__ __launch_bounds__(kEvThreads) void k_events_count(int64_t n_blocks, const fm_block_desc *blocks,
                                                             const fm_rf_pulse *pulses,
                                                             const fm_trapezoid *traps, int64_t max_cand,
                                                             int64_t *offsets)
{
    __shared__ EvShared sh;
    const int64_t b = blockIdx.x;
    const fm_block_desc bd = blocks[b];
    if (!block_sorted_times(sh, bd, pulses, traps, max_cand)) {
        if (threadIdx.x == 0) offsets[b + 1] = -1;  // does not fit: no events, flagged by the scan
        return;
    }
    const int n = sh.n_cand;
    const int chunk = (n + kEvThreads - 1) / kEvThreads;
    const int lo = min(n, (int)threadIdx.x * chunk), hi = min(n, lo + chunk);
    sh.count[threadIdx.x] = count_distinct(sh, lo, hi);
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int k = 0; k < kEvThreads; ++k) t += sh.count[k];
        offsets[b + 1] = t;
    }
}

// offsets[1..n] hold the counts (-1: the block did not fit): turn them
// into starts, offsets[n] the total; status = 1 (a block did not fit) | 2
// (the events exceed capacity)
__global__ __launch_bounds__(kEvThreads) void k_events_scan(int64_t n_blocks, int64_t capacity, int64_t *offsets,
                                                            int32_t *status)
{
    __shared__ int64_t part[kEvThreads];
    __shared__ int32_t bad[kEvThreads];
    const int tid = threadIdx.x;
    const int64_t chunk = (n_blocks + kEvThreads - 1) / kEvThreads;
    const int64_t lo = min<int64_t>(n_blocks, tid * chunk), hi = min<int64_t>(n_blocks, lo + chunk);
    int64_t s = 0;
    int32_t nf = 0;
    for (int64_t i = lo; i < hi; ++i) {
        const int64_t c = offsets[i + 1];
        nf |= c < 0;
        s += c < 0 ? 0 : c;
    }
    part[tid] = s;
    bad[tid] = nf;
    __syncthreads();
    if (tid == 0) {
        int64_t acc = 0;
        int32_t any = 0;
        for (int k = 0; k < kEvThreads; ++k) {
            const int64_t v = part[k];
            part[k] = acc;
            acc += v;
            any |= bad[k];
        }
        *status = (any ? 1 : 0) | (acc > capacity ? 2 : 0);
        offsets[0] = 0;
    }
    __syncthreads();
    int64_t acc = part[tid];
    for (int64_t i = lo; i < hi; ++i) {
        const int64_t c = offsets[i + 1];
        acc += c < 0 ? 0 : c;
        offsets[i + 1] = acc;
    }
}

__device__ inline int32_t block_code(const fm_block_desc &bd)
{
    return (bd.spoil ? 1 : 0) | ((bd.store_col + 1) << 1);
}

__global__ __launch_bounds__(kEvThreads) void k_events_write(int64_t n_blocks, const fm_block_desc *blocks,
                                                             const fm_rf_pulse *pulses, const double *rf_tables,
                                                             const fm_trapezoid *traps, int64_t max_cand,
                                                             int64_t capacity, const int64_t *offsets,
                                                             double *times, double *rf, double *G, double *dt,
                                                             int32_t *marks)
{
#pragma clang fp contract(off)
    __shared__ EvShared sh;
    const int64_t b = blockIdx.x;
    const fm_block_desc bd = blocks[b];
    const int64_t off = offsets[b];
    if (b == n_blocks - 1 && threadIdx.x == 0) {
        const int64_t total = offsets[n_blocks];
        if (total <= capacity) marks[total] = 1 + block_code(bd);
    }
    if (!block_sorted_times(sh, bd, pulses, traps, max_cand)) return;
    const int n = sh.n_cand;
    const int chunk = (n + kEvThreads - 1) / kEvThreads;
    const int lo = min(n, (int)threadIdx.x * chunk), hi = min(n, lo + chunk);
    sh.count[threadIdx.x] = count_distinct(sh, lo, hi);
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (int k = 0; k < kEvThreads; ++k) {
            const int32_t v = sh.count[k];
            sh.count[k] = acc;
            acc += v;
        }
    }
    __syncthreads();
    const int32_t code_prev = b > 0 ? block_code(blocks[b - 1]) : 0;
    int64_t r = off + sh.count[threadIdx.x];
    for (int i = lo; i < hi; ++i) {
        if (i > 0 && sh.v[i] == sh.v[i - 1]) continue;
        if (r < capacity) {
            const double t = sh.v[i];
            times[r] = t;
            dt[r] = t - (i == 0 ? 0.0 : sh.v[i - 1]);
            double re = 0.0, im = 0.0;
            for (int k = 0; k < bd.n_pulses; ++k) {
                double pr, pi;
                pulse_value(pulses[bd.pulse0 + k], rf_tables, t, pr, pi);
                re = re + pr;
                im = im + pi;
            }
            rf[2 * r] = re;
            rf[2 * r + 1] = im;
            double g[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < bd.n_traps; ++k) {
                const fm_trapezoid &tr = traps[bd.trap0 + k];
                const int ax = tr.axis;
                if (ax >= 0 && ax < 3) g[ax] = g[ax] + trap_value(tr, t);
            }
            G[3 * r] = g[0];
            G[3 * r + 1] = g[1];
            G[3 * r + 2] = g[2];
            marks[r] = (i == 0 && b > 0) ? 1 + code_prev : 0;
        }
        ++r;
    }
}

}  // namespace fm

using namespace fm;

extern "C" int fm_events_bound(int64_t n_blocks, const fm_block_desc *blocks, const fm_rf_pulse *pulses,
                               int64_t n_table_points, const fm_trapezoid *traps, int64_t *max_candidates,
                               int64_t *total_bound)
{
    FM_CHECK_ARG(n_blocks >= 0, "negative n_blocks");
    FM_CHECK_ARG(max_candidates && total_bound && (n_blocks == 0 || blocks), "NULL argument");
    FM_CHECK_ARG(n_table_points >= 0, "negative n_table_points");
    int64_t mx = 0, tot = 0;
    EvSeg seg[kEvMaxSegs];
    for (int64_t b = 0; b < n_blocks; ++b) {
        const fm_block_desc &bd = blocks[b];
        FM_CHECK_ARG(bd.dt > 0.0, "block %lld: dt must be > 0", (long long)b);
        FM_CHECK_ARG(bd.n_pulses == 0 || pulses, "block %lld has pulses but pulses is NULL", (long long)b);
        FM_CHECK_ARG(bd.n_traps == 0 || traps, "block %lld has trapezoids but traps is NULL", (long long)b);
        int64_t n = 0;
        const int ns = block_segments(bd, pulses, traps, seg, &n);
        FM_CHECK_ARG(ns > 0, "block %lld: at most %d pulses and %d trapezoids", (long long)b, FM_EVENTS_MAX_PARTS,
                     FM_EVENTS_MAX_PARTS);
        for (int k = 0; k < bd.n_pulses; ++k) {
            const fm_rf_pulse &p = pulses[bd.pulse0 + k];
            FM_CHECK_ARG(p.n_table == 0 || (p.n_table >= 2 && p.table0 >= 0 && p.table0 + p.n_table <= n_table_points),
                         "block %lld pulse %d: table [%lld, +%d) outside the %lld table points", (long long)b, k,
                         (long long)p.table0, p.n_table, (long long)n_table_points);
            FM_CHECK_ARG(p.dur > 0.0 || p.n_table == 0, "block %lld pulse %d: a table needs dur > 0", (long long)b, k);
        }
        FM_CHECK_ARG(n >= 1, "block %lld has no event (empty extent)", (long long)b);
        mx = n > mx ? n : mx;
        tot += n;
    }
    *max_candidates = mx;
    *total_bound = tot;
    return FM_OK;
}

extern "C" int fm_events_build(int64_t n_blocks, const fm_block_desc *blocks, const fm_rf_pulse *pulses,
                               const double *rf_tables, const fm_trapezoid *traps, int64_t max_candidates,
                               int64_t capacity, int64_t *offsets, double *times, double *rf, double *G, double *dt, int32_t *marks,
                               int32_t *status, void *stream)
{
    FM_CHECK_ARG(n_blocks >= 1, "need at least one block");
    FM_CHECK_ARG(n_blocks <= 0x7fffffff, "too many blocks for one launch");
    FM_CHECK_ARG(max_candidates >= 1 && max_candidates <= FM_EVENTS_MAX_CANDIDATES,
                 "max_candidates must be in [1, %d]", FM_EVENTS_MAX_CANDIDATES);
    FM_CHECK_ARG(capacity >= 1, "capacity must be >= 1");
    FM_CHECK_ARG(blocks && offsets && times && rf && G && dt && marks && status, "NULL array");
    hipStream_t st = as_stream(stream);
    const unsigned nb = (unsigned)n_blocks;
    k_events_count<<<nb, kEvThreads, 0, st>>>(n_blocks, blocks, pulses, traps, max_candidates, offsets);
    FM_LAUNCH_CHECK("k_events_count");
    k_events_scan<<<1, kEvThreads, 0, st>>>(n_blocks, capacity, offsets, status);
    FM_LAUNCH_CHECK("k_events_scan");
    k_events_write<<<nb, kEvThreads, 0, st>>>(n_blocks, blocks, pulses, rf_tables, traps, max_candidates, capacity,
                                              offsets,
                                              times, rf, G, dt, marks);
    FM_LAUNCH_CHECK("k_events_write");
    return FM_OK;
}
