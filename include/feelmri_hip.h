/*
 * feelmri_hip.h -- C ABI of libfeelmri_hip.so, the MI355X (gfx950) signal-
 * generation path of FEelMRI.
 *
 * Every entry point returns an int status (FM_OK = 0).  On failure
 * fm_last_error() returns a message for the calling thread.  Device
 * pointers are plain HIP device addresses (e.g. a torch tensor's
 * data_ptr()); `stream` is a hipStream_t passed as void*.  Kernels run
 * asynchronously on that stream; buffers are caller-owned.  Host pointers
 * are marked "host".  Complex arrays are interleaved (re, im).
 *
 * Units follow the reference: positions m, k cycles/m, t/dt/T1/T2 ms,
 * phi_dB0 rad/ms, delta_B mT, G mT/m, rf mT, gamma rad/(ms*mT).
 *
 * The reference interface each entry point replaces is cited as
 * path:line in the reference repository (hmella/feelmri).
 */
#ifndef FEELMRI_HIP_H
#define FEELMRI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FM_OK = 0,
    FM_ERR_ARG = 1,          /* invalid argument or shape */
    FM_ERR_HIP = 2,          /* HIP runtime error */
    FM_ERR_UNSUPPORTED = 3,  /* combination not supported by the chosen method */
    FM_ERR_WORKSPACE = 4     /* workspace too small */
};

/* Arithmetic form. */
enum {
    FM_PREC_F64 = 0,    /* plain FP64 (the drop-in form) */
    FM_PREC_MIXED = 1,  /* FP64 range reduction + FP32 sincos (direct);
                           BF16 hi+lo MFMA with FP32 accumulation (Cartesian) */
    FM_PREC_F32 = 2     /* FP32 throughout (Bloch) */
};

/* Element type of reference-layout inputs handed to the pack routines. */
enum { FM_DTYPE_F32 = 0, FM_DTYPE_F64 = 1 };

/* Signal evaluation method. */
enum {
    FM_METHOD_AUTO = 0,       /* Cartesian GEMM when the sampling factorises, else direct */
    FM_METHOD_DIRECT = 1,     /* direct non-uniform sum (any trajectory) */
    FM_METHOD_CARTESIAN = 2   /* factorised complex GEMM on the matrix cores */
};

const char *fm_last_error(void);
int fm_version(void);

/* ------------------------------------------------------------------------
 * Signal -- replaces MRI.Signal (cpp/feelmri/MRI.h:43-82 bindings;
 * cpp/feelmri/MRI.cpp:196-297 static nodes, :38-159 moving nodes):
 *
 *   S[s, v] = sum_n exp(i(-2 pi_f k_s . r_n(t_s) + phi_dB0_n t_s))
 *                   exp(-t_s / T2_n) (M Mxy)[n, v]
 *
 * with pi_f = (double)(float)pi (MRI.h:13).  Split into a plan (host
 * analysis of the sampling, made once per trajectory) and an execute call
 * (device work only, no allocation).
 * ---------------------------------------------------------------------- */
typedef struct fm_signal_plan fm_signal_plan;

/* kx, ky, kz, t: host arrays of shape (n_meas, n_lines, n_kz), C order --
 * the kloc and t arguments of MRI.Signal (MRI.h:19-20).  Samples with equal
 * t are grouped (the reference's t cache, MRI.cpp:260-276); the plan keeps
 * the sampling tables resident in device memory (allocated here, once). */
int fm_signal_plan_create(fm_signal_plan **plan, int64_t n_meas, int64_t n_lines, int64_t n_kz,
                          const double *kx, const double *ky, const double *kz,
                          const double *t, int method);
/* Waits for the plan's device (executes in flight read its tables), then
 * frees it. */
int fm_signal_plan_destroy(fm_signal_plan *plan);
/* Host-only analysis of a trajectory, as fm_signal_plan_create performs it,
 * without a device: the method it would choose, the number of distinct t,
 * of 16-sample readout runs and of t-grouped tiles, whether the runs are
 * linear to 1e-14 (runs_exact: the FP64 form takes them too), and how many
 * equal-size t-groups a packed slot holds (group_pack, 1 when the groups do
 * not pack), and how many a conjugate-pair tile holds (conj_pack, 1 when
 * some packed group's samples do not pair up as k and -k: the mixed form of
 * static nodes then takes the packed slots).  Out-pointers may be NULL.  No
 * counterpart in the reference (plan introspection; usable without a
 * GPU). */
int fm_signal_plan_analyse(int64_t n_meas, int64_t n_lines, int64_t n_kz, const double *kx,
                           const double *ky, const double *kz, const double *t, int method,
                           int *method_out, int64_t *n_tgroups, int64_t *n_runs,
                           int64_t *n_tiles_grouped, int *runs_exact, int *group_pack, int *conj_pack);
/* n_tgroups: number of distinct t; method: the method chosen
 * (FM_METHOD_DIRECT or FM_METHOD_CARTESIAN).  tgroup_t (host, may be
 * NULL) receives the n_tgroups distinct times in the order the motion
 * coefficient table of fm_signal_execute must follow. */
int fm_signal_plan_info(const fm_signal_plan *plan, int64_t *n_samples, int64_t *n_tgroups,
                        int *method, double *tgroup_t);
/* Readout runs of the plan: 0 unless k and t are linear in the readout
 * index along every line to 1e-12 (then the mixed form of static nodes takes
 * the run kernel); runs_exact: linear to 1e-14 (then the FP64 form takes its
 * FP64 run kernel too).  No counterpart in the reference (plan
 * introspection). */
int fm_signal_plan_runs(const fm_signal_plan *plan, int64_t *n_runs, int *run_len, int *runs_exact);
/* Whole lines of the plan linear in k and t (k(m) = k0 + m dk, t(m) = t0 +
 * m dt from each line's first sample, to 1e-12 of the largest |k| / |t|):
 * n_lines = lines * kz when every line is, else 0.  line_gemm = 1 when the
 * mixed form of static nodes takes the per-line GEMM (n_meas from 128 to
 * 2048 samples, fewer than 2^28 nodes): S[Q p + q] = sum_n (c e Z^p)[n]
 * (z^q)[n] on the matrix cores.  No counterpart in the reference (plan
 * introspection). */
int fm_signal_plan_lines(const fm_signal_plan *plan, int64_t *n_lines, int *line_gemm);
/* Workspace (device bytes) fm_signal_execute needs for this plan, with
 * static nodes (moving = 0) or moving nodes (moving = 1: modes given).  The
 * GEMM paths size their node splits by the current device's occupancy, so
 * query on the device that will execute. */
int fm_signal_workspace_size(const fm_signal_plan *plan, int precision, int64_t n_nodes,
                             int n_enc, int moving, size_t *bytes);

/* Pack reference-layout node inputs into the SoA table fm_signal_execute
 * reads: rows x, y, z, phi_dB0 (rad/ms), 1/T2, then Re/Im of (w .* Mxy)[:, v]
 * for each encoding v; (5 + 2 n_enc) rows of n_nodes doubles.
 *   r0 (n_nodes, 3) row-major, phi_dB0 (n_nodes), T2 (n_nodes): dtype `dtype`
 *   mxy (n_nodes, n_enc) complex of dtype `dtype` (M * Mxy already formed,
 *       or Mxy when `weights` holds the lumped diagonal of M)
 *   weights (n_nodes) float64 lumped diagonal of M, or NULL for identity. */
int fm_signal_nodes_pack(int64_t n_nodes, int n_enc, int dtype, const void *r0,
                         const void *phi_dB0, const void *T2, const void *mxy,
                         const double *weights, double *packed, void *stream);

/* Evaluate the signal.  out: (n_samples, n_enc) complex128, sample order
 * (i, j, k) of the plan's arrays.  On a Cartesian plan, moving nodes with
 * n_enc <= 4 take the per-readout GEMM (bf16 or FP64 MFMA by precision),
 * otherwise the direct kernels.  Motion (MRI.cpp:109-121): modes (NULL
 * for static nodes) is (3 n_modes) rows of n_nodes float32, row c*n_modes+r
 * holding component c of mode r; motion_coef (device) is (n_tgroups,
 * n_modes + 3) doubles: per distinct t the mode weights followed by a
 * uniform translation, so r_n(t) = r0_n + sum_r modes[:, r][n] w_r + trans. */
int fm_signal_execute(const fm_signal_plan *plan, int precision, int64_t n_nodes, int n_enc,
                      const double *packed, const float *modes, int n_modes,
                      const double *motion_coef, void *workspace, size_t workspace_bytes,
                      double *out, void *stream);

/* Size of the library's static signal workspace (bytes per device). */
#define FM_SIGNAL_SCRATCH_BYTES (128u << 20)

/* fm_signal_execute with no workspace argument and no allocation: the
 * workspace is a file-scope device array of FM_SIGNAL_SCRATCH_BYTES that each
 * device holds until the process ends (its address is looked up on the
 * device's first call).  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Evaluations whose
 * fm_signal_workspace_size exceeds the array return FM_ERR_WORKSPACE.  All
 * calls on one device share the array, so calls that overlap on one device
 * (other streams, other host threads) race; serialise them. */
int fm_signal_execute_noalloc(const fm_signal_plan *plan, int precision, int64_t n_nodes, int n_enc,
                              const double *packed, const float *modes, int n_modes,
                              const double *motion_coef, double *out, void *stream);

/* Measurement support (no reference counterpart): when enabled, each
 * fm_signal_execute records a pair of HIP events on its stream around the
 * dominant kernel (the GEMM, or the direct-sum launches), for up to
 * max_launches executes.  fm_signal_plan_kernel_ms waits for the last
 * recorded event and returns the summed milliseconds and the number of
 * timed executes since the previous query.  max_launches = 0 disables. */
int fm_signal_plan_set_timing(fm_signal_plan *plan, int max_launches);
int fm_signal_plan_kernel_ms(fm_signal_plan *plan, double *total_ms, int *n_timed);

/* y = M x for a CSR matrix M (int64 indptr/indices, float64 data) and
 * x (n_cols, n_enc) complex128: the M * Mxy product of MRI.cpp:243 when M
 * is the consistent mass matrix (examples/free_running.py:177). */
int fm_csr_spmv_c128(int64_t n_rows, int n_enc, const int64_t *indptr, const int64_t *indices,
                     const double *data, const double *x, double *y, void *stream);

/* ------------------------------------------------------------------------
 * Bloch -- replaces BlochSimulator.solve_mri (cpp/feelmri/BlochSimulator.h:
 * 50-80 bindings; BlochSimulator.cpp:151-276 static, :12-148 moving).
 * One lane per node; the event table is staged through LDS.
 * ---------------------------------------------------------------------- */
/* Pack r0 (n,3) row-major, T1, T2, delta_B (n) of dtype `dtype` into the
 * SoA table fm_bloch_execute reads: rows x, y, z, 1/T1, 1/T2, delta_B. */
int fm_bloch_nodes_pack(int64_t n_pos, int dtype, const void *r0, const void *T1,
                        const void *T2, const void *delta_B, double *packed, void *stream);

/* precision FM_PREC_F64 or FM_PREC_F32.
 * rf (n_time) complex128, G (n_time, 3) row-major, dt (n_time): device.
 * mxy_in/mz_in: initial state, complex128 / float64 (n_pos).
 * modes/motion_coef: as fm_signal_execute, with motion_coef holding one
 * row per step i (n_time - 1 rows), evaluated at the end of the step.
 * history != 0: mxy_out (n_pos, n_time) complex and mz_out (n_pos, n_time)
 * column-major (element (p, i) at p + i n_pos), including column 0;
 * history output element type is complex128/float64 for F64 and
 * complex64/float32 for F32.  history == 0: final state only, (n_pos)
 * complex128/float64 for both precisions, so it can seed the next call.
 * mxy_out/mz_out may alias mxy_in/mz_in only when history == 0. */
int fm_bloch_execute(int precision, int64_t n_pos, int64_t n_time, const double *packed,
                     double M0, double gamma, const double *rf, const double *G,
                     const double *dt, const double *mxy_in, const double *mz_in,
                     const float *modes, int n_modes, const double *motion_coef,
                     void *mxy_out, void *mz_out, int history, void *stream);

/* Artificial spoiling and storage between blocks (python/feelmri/
 * Bloch.py:357-368): copies the state into column `store_col` of the
 * (n_pos, n_store) column-major stores when store_col >= 0, then zeroes
 * mxy_state when `spoil` is set.  State: complex128/float64; stores:
 * complex64/float32 (the reference's BlochSolver arrays, Bloch.py:321-322). */
int fm_bloch_store_spoil(int64_t n_pos, double *mxy_state, const double *mz_state,
                         void *mxy_store, void *mz_store, int64_t store_col, int spoil,
                         void *stream);

/* A whole block chain of BlochSolver.solve (python/feelmri/Bloch.py:
 * 331-368: per block one solve_mri call from the carried state, the last
 * column stored when the block stores, Mxy zeroed after a non-empty block,
 * Mz carried) for static nodes, with the state kept in registers across
 * blocks.  Equal, bit for bit, to fm_bloch_execute (history == 0) followed by
 * fm_bloch_store_spoil for every block.
 * block_len, store_col, spoil: host arrays of n_blocks entries; block b has
 * block_len[b] >= 1 events, stores into column store_col[b] (-1: none) of
 * the (n_pos, n_store) column-major complex64/float32 stores and spoils when
 * spoil[b] != 0.  rf, G, dt: device, the blocks' event tables concatenated
 * (sum of block_len events, layouts as fm_bloch_execute).  mxy_in/mz_in:
 * initial state; mxy_out/mz_out: state after the last block (may alias the
 * inputs), complex128/float64; with n_blocks == 0 they receive the input
 * state.  One launch per 128 blocks; nothing is allocated, so the call can
 * be recorded into a HIP graph. */
int fm_bloch_chain(int precision, int64_t n_pos, const double *packed, double M0, double gamma,
                   int64_t n_blocks, const int64_t *block_len, const int32_t *store_col,
                   const int32_t *spoil, const double *rf, const double *G, const double *dt,
                   const double *mxy_in, const double *mz_in, void *mxy_store, void *mz_store,
                   int64_t n_store, double *mxy_out, double *mz_out, void *stream);

/* ------------------------------------------------------------------------
 * Sequence event tables built on the device -- replaces
 * SequenceBlock._discretization and __call__ (python/feelmri/Bloch.py:105-129,
 * :44-49) with dt = diff(times, prepend=0) (Bloch.py:340), for RF pulses
 * (constant or tabulated) and trapezoidal gradients; with fm_bloch_chain_marked a whole block chain
 * (BlochSolver.solve, Bloch.py:298-380) runs from device descriptions with
 * nothing read back to the host.
 * ---------------------------------------------------------------------- */
/* One block: the union of a dt grid over its extent, each pulse's start,
 * end and dt_rf grid, each trapezoid's 4 knots and dt_gr grid, sorted and
 * de-duplicated (grids as numpy's arange: start, start + step, then
 * start + i * ((start + step) - start); dt_rf, dt_gr <= 0: no grid).  The
 * extent runs from min(t0, pulse starts, trapezoid starts) to
 * max(t0 + dur, pulse ends, trapezoid ends).  RF at time t: the sum, in
 * order, of each pulse's value (fm_rf_pulse); gradient on
 * an axis: the sum, in order, of numpy's interp of each trapezoid of that
 * axis over its knots (zero outside).  dt must be > 0. */
typedef struct fm_block_desc {
    double t0, dur, dt, dt_rf, dt_gr;   /* ms */
    int32_t pulse0, n_pulses;           /* pulses [pulse0, pulse0 + n_pulses) */
    int32_t trap0, n_traps;             /* trapezoids [trap0, trap0 + n_traps) */
    int32_t spoil;                      /* != 0: Mxy zeroed after the block (a non-empty block) */
    int32_t store_col;                  /* store column, -1: none */
} fm_block_desc;

/* One RF pulse on [start, start + dur], as the reference's RF reads its
 * table (MRObjects.py:591-614: linear interpolation, 0 outside the closed
 * interval).  n_table == 0: the constant amp (the hard shape);
 * n_table >= 2: the (re, im) table rf_tables[table0 .. table0 + n_table) on
 * the grid linspace(start, start + dur, n_table) (j * step + start, the last
 * point start + dur), read as numpy's interp -- any shape (apodized sinc,
 * sinc, hard with a frequency offset). */
typedef struct fm_rf_pulse {
    double start, dur;                  /* ms */
    double amp_re, amp_im;              /* mT, when n_table == 0 */
    int64_t table0;                     /* first table point, when n_table >= 2 */
    int32_t n_table, reserved;
} fm_rf_pulse;

typedef struct fm_trapezoid {
    double start, ramp, flat, amp;      /* ms, mT/m: knots start, +ramp, +flat, +ramp */
    int32_t axis, reserved;             /* 0, 1, 2 */
} fm_trapezoid;

#define FM_EVENTS_MAX_CANDIDATES 8192   /* grid candidates per block */
#define FM_EVENTS_MAX_PARTS 16          /* pulses and trapezoids per block, each */

/* Host: per-block candidate counts (grid points before de-duplication):
 * max_candidates receives the largest (the builder's per-block bound) and
 * total_bound their sum (a bound on the event count).  Same arithmetic as
 * the device builder.  Also checks every pulse table against the
 * n_table_points points of rf_tables: fm_events_build reads the
 * descriptions as given, so run this check on them first.  Host arrays. */
int fm_events_bound(int64_t n_blocks, const fm_block_desc *blocks, const fm_rf_pulse *pulses,
                    int64_t n_table_points, const fm_trapezoid *traps, int64_t *max_candidates,
                    int64_t *total_bound);

/* Device: the blocks' event tables, concatenated.  blocks, pulses,
 * rf_tables (2 x float64 per point; NULL when no pulse has a table), traps:
 * device arrays; max_candidates: a bound from fm_events_bound (at most
 * FM_EVENTS_MAX_CANDIDATES); capacity: entries of times, dt (float64), rf
 * (2 x float64), G (3 x float64) and, plus one, of marks (int32).
 * offsets (device, n_blocks + 1 int64) receives each block's first event and
 * the total; marks[e] is 0, or 1 + code of the block that ends before event
 * e when e starts block b > 0, and marks[total] = 1 + code of the last block
 * (code = spoil | (store_col + 1) << 1).  status (device int32) receives 0,
 * or 1 when a block has more candidates than max_candidates, or 2 when the
 * events exceed capacity (the tables are then incomplete; nothing is
 * written out of bounds).  Three launches, nothing allocated: the call can
 * be recorded into a HIP graph. */
int fm_events_build(int64_t n_blocks, const fm_block_desc *blocks, const fm_rf_pulse *pulses,
                    const double *rf_tables, const fm_trapezoid *traps, int64_t max_candidates, int64_t capacity,
                    int64_t *offsets,
                    double *times, double *rf, double *G, double *dt, int32_t *marks, int32_t *status,
                    void *stream);

/* A block chain over concatenated event tables whose block boundaries are
 * given per event (marks, as fm_events_build writes them) and whose event
 * count is read on the device (*n_events, at most capacity): static nodes,
 * or moving nodes with modes / n_modes / motion_coef as fm_bloch_execute,
 * motion_coef holding one row per step of the concatenated table (step i
 * uses event i + 1; rows of the steps that start a block are not read).
 * Same results, bit for bit, as fm_bloch_execute (history == 0) and
 * fm_bloch_store_spoil per block.  Store columns at or beyond n_store are
 * not written.  One launch, any number of blocks, nothing allocated. */
int fm_bloch_chain_marked(int precision, int64_t n_pos, const double *packed, double M0, double gamma,
                          const int64_t *n_events, int64_t capacity, const double *rf, const double *G,
                          const double *dt, const int32_t *marks, const float *modes, int n_modes,
                          const double *motion_coef, const double *mxy_in, const double *mz_in, void *mxy_store,
                          void *mz_store, int64_t n_store, double *mxy_out, double *mz_out, void *stream);

/* ------------------------------------------------------------------------
 * P1 mass -- replaces Assemble.MassAssemble (cpp/feelmri/Assemble.h:21-39;
 * Assemble.cpp:78-116) followed by lumping (python/feelmri/
 * Phantom.py:379-399).  P1 tetrahedra only.
 * ---------------------------------------------------------------------- */
/* Host: greedy element colouring so that no two elements of a colour share
 * a node.  elems (n_elems, 4) int32, host.  perm (host, n_elems) receives
 * the elements ordered by colour; color_offsets (host, max_colors + 1)
 * the start of each colour in perm.  At most 64 colours. */
int fm_mass_p1_color(int64_t n_elems, int64_t n_nodes, const int32_t *elems, int32_t *perm,
                     int64_t *color_offsets, int max_colors, int *n_colors);

/* Lumped weights w_i = sum_{e ni i} |T_e| / 4 by an atomic-free scatter:
 * one launch per colour.  elems (n_elems, 4) int32, nodes (n_nodes, 3)
 * row-major of dtype `dtype`, perm: device; color_offsets: host. */
int fm_mass_p1_lumped(int64_t n_elems, int64_t n_nodes, const int32_t *elems, int dtype,
                      const void *nodes, const int32_t *perm, const int64_t *color_offsets,
                      int n_colors, double *w, void *stream);

/* Host: each node's incident elements in the order of perm (colour order,
 * or element order for perm = identity), as their positions k in perm
 * (perm[k] = e) in the slots of a sliced-ELL table: node n's j-th element is
 * stored in slot chunk_offsets[n / 64] + 64 j + n % 64; padding slots hold
 * -1.  Outputs (host): chunk_offsets (ceil(n_nodes / 64) + 1), degree
 * (n_nodes), n_slots (the table length, chunk_offsets' last entry) and,
 * unless incident is NULL (a size query), incident (max_slots >= n_slots
 * int32 entries).  elems and perm as for fm_mass_p1_color (host). */
int fm_mass_p1_incidence(int64_t n_elems, int64_t n_nodes, const int32_t *elems, const int32_t *perm,
                         int64_t *chunk_offsets, int32_t *degree, int32_t *incident, int64_t max_slots,
                         int64_t *n_slots);

/* Lumped weights w_i = sum_{e ni i} |T_e| / 4 in two launches: element e
 * writes |det J| = 6 |T| into det_scratch[pos[e]], then node i adds the
 * determinants of its degree[i] incident elements -- the additions of
 * fm_mass_p1_lumped, in the same order and with the same rounding, so the
 * result is bitwise the same when perm is a colouring.  elems: the
 * connectivity (device); pos: the inverse of the perm the incidence table
 * was built with (pos[perm[k]] = k, device int32; NULL for perm = identity);
 * chunk_offsets, degree, incident from fm_mass_p1_incidence (copied to the
 * device); n_slots = the slot count it returned; det_scratch (n_elems
 * float64, overwritten); nodes, w: device.  With perm = identity the tables
 * may belong to a larger mesh whose elements are ordered by their largest
 * node id: positions >= n_elems are skipped, so its first n_nodes nodes and
 * first n_elems elements (those whose nodes all lie below n_nodes) are
 * weighted as a mesh of their own. */
int fm_mass_p1_lumped_gather(int64_t n_elems, int64_t n_nodes, const int32_t *elems, const int32_t *pos, int dtype,
                             const void *nodes, const int64_t *chunk_offsets, const int32_t *degree,
                             const int32_t *incident, int64_t n_slots, double *det_scratch,
                             double *w, void *stream);

/* Per-element 4x4 consistent mass matrices |T_e|/20 (1 + delta_ij),
 * (n_elems, 16) float64 -- the triplet values of Assemble.cpp:109-111. */
int fm_mass_p1_element_matrices(int64_t n_elems, const int32_t *elems, int dtype,
                                const void *nodes, double *vals, void *stream);

/* ------------------------------------------------------------------------
 * POD -- replaces POD.tensordot_modes_weights (cpp/feelmri/POD.h:17-23;
 * POD.cpp:5-38): out (P, C) = modes (P, C, M) . weights (M), dtype f32/f64.
 * ---------------------------------------------------------------------- */
int fm_pod_tensordot(int dtype, int64_t P, int64_t C, int64_t M, const void *modes,
                     const void *weights, void *out, void *stream);

/* Node positions under POD motion (python/feelmri/Motion.py:240-254,
 * :272-290, :303-309, :58-68): out (n_nodes, 3) row-major float64 =
 * r0 + sum_r modes[:, r] w_r + trans for one coefficient row
 * (n_modes + 3 doubles, device). */
int fm_pod_positions(int64_t n_nodes, const double *packed_xyz, const float *modes,
                     int n_modes, const double *coef, double *out, void *stream);

/* Convert reference (P, 3, R) float32 modes to the (3R, P) SoA rows the
 * signal and Bloch kernels read. */
int fm_modes_pack(int64_t P, int R, const float *modes_pcr, float *modes_soa, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FEELMRI_HIP_H */
