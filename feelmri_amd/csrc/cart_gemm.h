// Shared pieces of the Cartesian GEMM kernels (signal_gemm.hip drives them).
#pragma once
#include "signal_plan.h"

namespace fm {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) double f64x4;

constexpr int kGemmThreads = 256;  // 4 waves
constexpr int kXcds = 8;

// Geometry of one execute.
struct GemmShape {
    bool mixed;
    int64_t N, Np;         // nodes, padded to the 16-node K step
    int64_t M, Mp;         // readout rows, padded to the row block
    int64_t NC, NCp;       // output columns lines*kz*V, padded to the column block
    int rb_per_wave;       // row blocks per wave
    int64_t n_rt, n_ct;    // row tiles, column tiles (per workgroup)
    int64_t ksteps;        // K steps in total
    int64_t n_splits, steps_per_split;
    // workspace offsets (bytes)
    size_t off_E, off_Ey, off_Ez, off_part, total;
};

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

// Launchers (cart_gemm_bf16.hip, cart_gemm_f64.hip): run the GEMM over the
// operand tables in the workspace and write the per-split partial tiles.
void launch_cart_gemm_bf16(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st);
void launch_cart_gemm_f64(const GemmShape &s, char *ws, int64_t n_kzV, hipStream_t st);

// Moving nodes on a Cartesian stack (cart_moving_bf16.hip): per readout index
// a (lines x nodes) by (nodes x kz*V) product with in-kernel operands.
struct MovingShape {
    int W, CB;                        // row blocks (waves) and column blocks per workgroup
    int64_t M, L, K, V, KV;           // readouts, lines, kz, encodings, kz*V
    int64_t n_rt, n_ct;               // row tiles, column tiles
    int64_t n_splits, nodes_per_split;
    int ez_rows;                      // kz rows of one column tile (LDS buffer rows)
    size_t part_bytes;                // (n_splits, M, L, KV) float2 partial tiles
};
MovingShape moving_shape(const CartesianTables &c, int64_t N, int V);
void launch_cart_moving_bf16(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                             const float *modes, int R, const double *mcoef, float2 *part, hipStream_t st);
// FP64 form (cart_moving_f64.hip): row and column blocks of 16, double2 partials.
MovingShape moving_shape_f64(const CartesianTables &c, int64_t N, int V);
void launch_cart_moving_f64(const MovingShape &s, const CartesianTables &c, const double *packed, int64_t N,
                            const float *modes, int R, const double *mcoef, double2 *part, hipStream_t st);

}  // namespace fm
