// Internal layout of fm_signal_plan (shared by the direct and Cartesian paths).
#pragma once
#include <vector>

#include "fm_common.h"

namespace fm {

// Samples handled by one workgroup of the direct kernels (256 lanes x 2).
constexpr int kDirectThreads = 256;
constexpr int kDirectSPL = 2;
constexpr int kDirectTile = kDirectThreads * kDirectSPL;
// Encodings handled per direct launch (register-resident accumulators).
constexpr int kMaxVB = 4;
// Upper bound on POD modes fused into the loaders.
constexpr int kMaxModes = 32;
// Encodings per launch of the moving-node Cartesian GEMM (node records in LDS).
constexpr int kMovMaxV = 4;

// Readout runs (mixed-form direct kernel for static nodes): kRunLen
// consecutive readout samples of one line whose k and t are linear in the
// readout index, so exp(-2 pi i k.x) exp((i phi - 1/T2) t) advances by one
// complex factor per sample.
constexpr int kRunLen = 16;

struct RunDesc {
    double k0x, k0y, k0z, t0;  // first sample (k' in cycles/m, t in ms)
    double dkx, dky, dkz, dt;  // step per sample (the line's slope)
    int32_t s0;                // first sample index; sample m is s0 + m * run_stride
    int32_t len;               // samples in the run (<= kRunLen)
    int32_t step;              // index of (dk, dt) in its tile's step table
    int32_t pad;
};
// distinct (dk, dt) steps a 256-run tile may share through LDS
constexpr int kRunSteps = 8;
// Lines linear in k and t over their whole length take the per-line GEMM
// (mixed form, static nodes) from this many samples per line.
constexpr int kLineMinLen = 128;
// ... and up to this many nodes (32-bit byte offsets into the node arrays).
constexpr int64_t kLineMaxNodes = int64_t(1) << 28;

struct SignalTile {
    int32_t group;   // t-group index (shared-t kernels) or -1
    int32_t offset;  // first entry in perm
    int32_t count;   // samples in the tile
    int32_t pad;
};

struct CartesianTables {
    bool ok = false;
    int64_t n_ro = 0, n_ln = 0, n_kz = 0;
    // Additive split k(i,j,k) = F(i) + G(j) + H(k) per axis, cycles/m,
    // already multiplied by (double)(float)pi / pi.
    std::vector<double> F, G, H;  // (n_ro,3), (n_ln,3), (n_kz,3)
    std::vector<double> t;        // (n_ro)
    std::vector<int32_t> ro_group;  // (n_ro) t-group of each readout index
    // a standard stack: G varies along y only and H along z only (their
    // other components are exactly 0), so a line or kz phase is one product
    bool axis_aligned = false;
    double *d_F = nullptr, *d_G = nullptr, *d_H = nullptr, *d_t = nullptr;
    int32_t *d_ro_group = nullptr;
};

}  // namespace fm

struct fm_signal_plan {
    int device = 0;
    int64_t n_meas = 0, n_lines = 0, n_kz = 0, n_samples = 0;
    int method = FM_METHOD_DIRECT;
    bool prefer_grouped = true;   // static nodes: use the t-grouped tiles
    int64_t n_tgroups = 0;
    std::vector<double> tg_t;     // distinct t, first-appearance order
    std::vector<int32_t> sample_group;
    // Two tilings of the samples: t-grouped tiles (each tile one distinct t;
    // always used for moving nodes) and consecutive tiles (any t per sample).
    int64_t n_tiles_grouped = 0, n_tiles_consec = 0;
    // packed slots of group_pack equal-size t-groups (1: none), static nodes
    int group_pack = 1;
    int64_t n_tiles_packed = 0;
    // largest |k'(slot 1) - k'(slot 0)|_1 over the lanes of the packed tiles
    // (cycles per unit length), and whether every sample has kz = 0
    double packed_dk = 0.0;
    int planar = 0;
    // conjugate pairs (mixed form): within every packed t-group the samples
    // pair up as k'_b = -k'_a (to 1e-12 of the largest |k'|); a conj tile
    // holds one member of each pair in slot 0 and its partner in slot 1,
    // lane for lane, conj_pack t-groups per tile (1: none)
    int conj_pack = 1;
    int64_t n_tiles_conj = 0;
    double conj_dk = 0.0;  // largest |k'_a + k'_b|_1 over the pairs
    int32_t *d_perm_conj = nullptr;
    fm::SignalTile *d_tiles_conj = nullptr;
    // device tables
    double4 *d_ks = nullptr;      // per sample: kx', ky', kz' (cycles/m, pi_f-scaled), t
    int32_t *d_perm = nullptr;    // samples ordered by t-group (grouped tiles index it)
    fm::SignalTile *d_tiles_grouped = nullptr;
    fm::SignalTile *d_tiles_consec = nullptr;  // offsets are sample indices
    fm::SignalTile *d_tiles_packed = nullptr;
    double *d_tg_t = nullptr;
    // readout runs (empty when some run is not linear in k and t)
    int64_t n_runs = 0, run_stride = 0;
    // every run linear to 1e-14 of the largest |k'| and |t| (not only the
    // 1e-12 of run detection): the FP64 form may then take the runs too
    bool runs_exact = false;
    fm::RunDesc *d_runs = nullptr;
    // whole lines linear in k and t (to 1e-12 of the largest |k'| / |t|
    // from the line's first sample): one RunDesc per line, len = n_meas,
    // s0 = the line's first sample; 0 when some line is not
    int64_t n_line_tabs = 0;
    fm::RunDesc *d_lines = nullptr;
    // per 256-run tile, its distinct steps (kRunSteps x (dkx, dky, dkz, dt))
    // and their count; run_steps = the largest count, 0 when some tile has
    // more than kRunSteps (each lane then forms its own step factor)
    int run_steps = 0;
    double *d_run_steps = nullptr;
    int32_t *d_run_nsteps = nullptr;
    fm::CartesianTables cart;
    // optional timing of the dominant kernel (fm_signal_plan_set_timing)
    std::vector<hipEvent_t> ev_start, ev_stop;
    mutable int ev_used = 0;
};

namespace fm {
// Record the start / stop event of the dominant kernel when timing is on.
inline void plan_timing_start(const fm_signal_plan *p, hipStream_t st)
{
    if (p->ev_used < (int)p->ev_start.size()) (void)hipEventRecord(p->ev_start[(size_t)p->ev_used], st);
}
inline void plan_timing_stop(const fm_signal_plan *p, hipStream_t st)
{
    if (p->ev_used < (int)p->ev_stop.size()) (void)hipEventRecord(p->ev_stop[(size_t)p->ev_used++], st);
}
}  // namespace fm

namespace fm {
// Cartesian path (signal_gemm.hip).
bool cartesian_analyse(fm_signal_plan *p, const double *kx, const double *ky, const double *kz,
                       const double *t);
int cartesian_upload(fm_signal_plan *p);
void cartesian_free(fm_signal_plan *p);
size_t cartesian_workspace(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc);
int cartesian_execute(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc,
                      const double *packed, void *ws, size_t ws_bytes, double *out,
                      hipStream_t stream);
// Static nodes on linear lines, per-line GEMM, mixed form
// (signal_lines_bf16.hip).
bool lines_usable(const fm_signal_plan *p, int64_t n_nodes);
size_t lines_workspace(const fm_signal_plan *p, int64_t n_nodes);
int lines_execute(const fm_signal_plan *p, int64_t n_nodes, int n_enc, const double *packed, void *ws,
                  size_t ws_bytes, double *out, hipStream_t stream);
// Moving nodes, per-readout GEMM in either form (n_enc <= kMovMaxV).
size_t cartesian_moving_workspace(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc);
int cartesian_moving_execute(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc,
                             const double *packed, const float *modes, int n_modes, const double *motion_coef,
                             void *ws, size_t ws_bytes, double *out, hipStream_t stream);
}  // namespace fm
