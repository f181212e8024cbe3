// P1 tetrahedron mass: lumped weights by an atomic-free coloured scatter,
// and the per-element consistent matrices.
//
// reference: cpp/feelmri/Assemble.cpp:15-116 (IsoMapping, LocalMassAssemble,
// MassAssemble) and the lumping in python/feelmri/Phantom.py:379-399.  For
// P1 tetrahedra the quadrature (exact at order 2) reduces to closed forms:
// M^e_ij = |T_e|/20 (1 + delta_ij) and the row sums w_i = sum_{e ni i} |T_e|/4.
//
// Elements are greedily coloured on the host so that no two elements of a
// colour share a node; one launch per colour then adds |T_e|/4 into the
// weights of its four nodes with plain loads and stores.  The order in which
// a node receives its contributions (colour order) is fixed, so results are
// reproducible.
//
// The gather form does the same sums in two launches over a host-built
// incidence table in sliced-ELL layout: node n's j-th incident element in
// colour order is stored, as its position k in the colour order, in slot
// chunk_offsets[n/64] + 64 j + n%64.  Launch 1 walks the elements in their
// own order (coalesced connectivity, coordinates re-read from L2) and writes
// |det J| = 6 |T| once per element into dets[pos[e]], pos the inverse of the
// colour permutation.  Launch 2, per node, reads its deg(n) positions -- a
// wave reads 64 consecutive slots per step -- and adds their determinants.
// The j-th elements of neighbouring nodes mostly share a colour, and within
// a colour positions follow element order, so the determinants gathered
// for a chunk of nodes lie close together.  Both launches deal contiguous
// ranges of elements / nodes onto each XCD, so that the lines of one range
// (coordinates, determinant writes of one colour, determinant reads) stay in
// that XCD's L2.  A node has at most one element per colour, so it performs
// the scatter's additions in the scatter's order; both forms add with an
// explicit fma(|det J_e|, 1/24, w), so they give the same bits.
#include <algorithm>
#include <vector>

#include "fm_common.h"

namespace fm {

template <typename T>
__device__ inline double tet_abs_det(const int32_t *__restrict__ el, const T *__restrict__ nodes, int32_t *ids)
{
    double X[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t n = el[i];
        ids[i] = (int32_t)n;
#pragma unroll
        for (int c = 0; c < 3; ++c) X[i][c] = (double)nodes[3 * n + c];
    }
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = X[r + 1][c] - X[0][c];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    return fabs(det);
}

template <typename T>
__global__ void k_lumped_colour(int64_t begin, int64_t end, const int32_t *__restrict__ perm,
                                const int32_t *__restrict__ elems, const T *__restrict__ nodes,
                                double *__restrict__ w)
{
    const int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    const int64_t e = perm[k];
    int32_t ids[4];
    const double det = tet_abs_det<T>(elems + 4 * e, nodes, ids);
#pragma unroll
    for (int i = 0; i < 4; ++i) w[ids[i]] = fma(det, 1.0 / 24.0, w[ids[i]]);  // += |T|/4
}

// Workgroup b of a launch of ceil(n / 8) * 8 workgroups takes range
// (b % 8) * ceil(n / 8) + b / 8 of n ranges: workgroups are dealt onto the
// 8 XCDs round-robin, so each XCD runs one contiguous eighth of the ranges.
__device__ inline int64_t xcd_range(int64_t n_ranges)
{
    const int64_t per = (n_ranges + 7) / 8;
    return (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

inline unsigned xcd_grid(int64_t n_ranges) { return (unsigned)(((n_ranges + 7) / 8) * 8); }

template <typename T>
__global__ __launch_bounds__(256) void k_abs_dets(int64_t E, const int32_t *__restrict__ elems,
                                                  const int32_t *__restrict__ pos, const T *__restrict__ nodes,
                                                  double *__restrict__ dets)
{
    const int64_t e = xcd_range(cdiv(E, 256)) * 256 + threadIdx.x;
    if (e >= E) return;
    int32_t ids[4];
    dets[pos ? (int64_t)pos[e] : e] = tet_abs_det<T>(elems + 4 * e, nodes, ids);
}

// Positions >= n_elems are skipped: with the identity order on a mesh whose
// elements are ordered by their largest node id, the first n_nodes nodes and
// the first n_elems elements form a sub-mesh weighted without a table of its
// own.
__global__ __launch_bounds__(256) void k_lumped_gather(int64_t n_nodes, int64_t n_elems,
                                                       const int64_t *__restrict__ chunk_offsets,
                                                       const int32_t *__restrict__ degree,
                                                       const int32_t *__restrict__ incident,
                                                       const double *__restrict__ dets, double *__restrict__ w)
{
    const int64_t n = xcd_range(cdiv(n_nodes, 256)) * 256 + threadIdx.x;
    if (n >= n_nodes) return;
    const int32_t *ids = incident + chunk_offsets[n / 64] + n % 64;
    const int d = degree[n];
    double acc = 0.0;
    for (int j = 0; j < d; ++j) {
        const int32_t k = ids[64 * j];
        if (k < n_elems) acc = fma(dets[k], 1.0 / 24.0, acc);  // += |T|/4
    }
    w[n] = acc;
}

template <typename T>
__global__ void k_element_matrices(int64_t E, const int32_t *__restrict__ elems, const T *__restrict__ nodes,
                                   double *__restrict__ vals)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int32_t ids[4];
    const double m = tet_abs_det<T>(elems + 4 * e, nodes, ids) * (1.0 / 120.0);  // |T|/20
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) vals[16 * e + 4 * i + j] = (i == j) ? 2.0 * m : m;
}

__global__ void k_fill_f64(int64_t n, double v, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

}  // namespace fm

using namespace fm;

extern "C" int fm_mass_p1_color(int64_t n_elems, int64_t n_nodes, const int32_t *elems, int32_t *perm,
                                int64_t *color_offsets, int max_colors, int *n_colors)
{
    FM_CHECK_ARG(n_elems >= 0 && n_nodes >= 0, "negative size");
    FM_CHECK_ARG(max_colors >= 1 && max_colors <= 64, "max_colors must be in [1, 64]");
    FM_CHECK_ARG(n_colors && color_offsets && (n_elems == 0 || (elems && perm)), "NULL argument");
    std::vector<uint64_t> used((size_t)n_nodes, 0);
    std::vector<uint8_t> col((size_t)n_elems);
    int nc = 0;
    for (int64_t e = 0; e < n_elems; ++e) {
        uint64_t m = 0;
        for (int i = 0; i < 4; ++i) {
            const int64_t n = elems[4 * e + i];
            if (n < 0 || n >= n_nodes) {
                set_error("element %lld references node %lld outside [0, %lld)", (long long)e, (long long)n,
                          (long long)n_nodes);
                return FM_ERR_ARG;
            }
            m |= used[(size_t)n];
        }
        if (m == ~uint64_t(0)) {
            set_error("element colouring needs more than 64 colours");
            return FM_ERR_UNSUPPORTED;
        }
        const int c = __builtin_ctzll(~m);
        if (c >= max_colors) {
            set_error("element colouring needs more than max_colors=%d colours", max_colors);
            return FM_ERR_UNSUPPORTED;
        }
        col[(size_t)e] = (uint8_t)c;
        for (int i = 0; i < 4; ++i) used[(size_t)elems[4 * e + i]] |= uint64_t(1) << c;
        nc = std::max(nc, c + 1);
    }
    std::vector<int64_t> off((size_t)nc + 1, 0);
    for (int64_t e = 0; e < n_elems; ++e) off[(size_t)col[(size_t)e] + 1]++;
    for (int c = 0; c < nc; ++c) off[(size_t)c + 1] += off[(size_t)c];
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < n_elems; ++e) perm[fill[col[(size_t)e]]++] = (int32_t)e;
    for (int c = 0; c <= nc; ++c) color_offsets[c] = off[(size_t)c];
    *n_colors = nc;
    return FM_OK;
}

extern "C" int fm_mass_p1_lumped(int64_t n_elems, int64_t n_nodes, const int32_t *elems, int dtype,
                                 const void *nodes, const int32_t *perm, const int64_t *color_offsets,
                                 int n_colors, double *w, void *stream)
{
    FM_CHECK_ARG(n_elems >= 0 && n_nodes >= 0 && n_colors >= 0, "negative size");
    FM_CHECK_ARG(dtype == FM_DTYPE_F32 || dtype == FM_DTYPE_F64, "unknown dtype %d", dtype);
    if (n_nodes == 0) return FM_OK;
    FM_CHECK_ARG(w != nullptr, "w is NULL");
    hipStream_t st = as_stream(stream);
    k_fill_f64<<<(unsigned)cdiv(n_nodes, 256), 256, 0, st>>>(n_nodes, 0.0, w);
    FM_LAUNCH_CHECK("k_fill_f64");
    if (n_elems == 0) return FM_OK;
    FM_CHECK_ARG(elems && nodes && perm && color_offsets, "NULL array");
    FM_CHECK_ARG(color_offsets[0] == 0 && color_offsets[n_colors] == n_elems,
                 "color_offsets must span [0, n_elems]");
    for (int c = 0; c < n_colors; ++c) {
        const int64_t b = color_offsets[c], e = color_offsets[c + 1];
        FM_CHECK_ARG(e >= b, "color_offsets not monotonic");
        if (e == b) continue;
        const unsigned blocks = (unsigned)cdiv(e - b, 256);
        if (dtype == FM_DTYPE_F64)
            k_lumped_colour<double><<<blocks, 256, 0, st>>>(b, e, perm, elems, (const double *)nodes, w);
        else
            k_lumped_colour<float><<<blocks, 256, 0, st>>>(b, e, perm, elems, (const float *)nodes, w);
        FM_LAUNCH_CHECK("k_lumped_colour");
    }
    return FM_OK;
}

extern "C" int fm_mass_p1_incidence(int64_t n_elems, int64_t n_nodes, const int32_t *elems, const int32_t *perm,
                                    int64_t *chunk_offsets, int32_t *degree, int32_t *incident,
                                    int64_t max_slots, int64_t *n_slots)
{
    FM_CHECK_ARG(n_elems >= 0 && n_nodes >= 0, "negative size");
    FM_CHECK_ARG(chunk_offsets && n_slots && (n_nodes == 0 || degree) && (n_elems == 0 || (elems && perm)),
                 "NULL argument");
    for (int64_t n = 0; n < n_nodes; ++n) degree[n] = 0;
    for (int64_t e = 0; e < n_elems; ++e)
        for (int i = 0; i < 4; ++i) {
            const int64_t n = elems[4 * e + i];
            if (n < 0 || n >= n_nodes) {
                set_error("element %lld references node %lld outside [0, %lld)", (long long)e, (long long)n,
                          (long long)n_nodes);
                return FM_ERR_ARG;
            }
            degree[n]++;
        }
    const int64_t n_chunks = cdiv(n_nodes, 64);
    chunk_offsets[0] = 0;
    for (int64_t c = 0; c < n_chunks; ++c) {
        int32_t wmax = 0;
        for (int64_t n = 64 * c; n < std::min<int64_t>(n_nodes, 64 * c + 64); ++n) wmax = std::max(wmax, degree[n]);
        chunk_offsets[c + 1] = chunk_offsets[c] + 64 * (int64_t)wmax;
    }
    *n_slots = chunk_offsets[n_chunks];
    if (incident == nullptr) return FM_OK;  // size query: chunk_offsets, degree and n_slots only
    if (chunk_offsets[n_chunks] > max_slots) {
        set_error("incidence needs %lld slots, the table holds %lld", (long long)chunk_offsets[n_chunks],
                  (long long)max_slots);
        return FM_ERR_ARG;
    }
    // slots in perm (colour) order: node n's j-th element takes slot j;
    // padding slots (j >= degree) hold -1
    for (int64_t q = 0; q < chunk_offsets[n_chunks]; ++q) incident[q] = -1;
    std::vector<int32_t> next((size_t)n_nodes, 0);
    std::vector<uint8_t> seen((size_t)n_elems, 0);
    for (int64_t k = 0; k < n_elems; ++k) {
        const int64_t e = perm[k];
        if (e < 0 || e >= n_elems || seen[(size_t)e]) {
            set_error("perm is not a permutation of [0, %lld) (entry %lld)", (long long)n_elems, (long long)k);
            return FM_ERR_ARG;
        }
        seen[(size_t)e] = 1;
        for (int i = 0; i < 4; ++i) {
            const int64_t n = elems[4 * e + i];
            incident[chunk_offsets[n / 64] + 64 * (int64_t)next[(size_t)n]++ + n % 64] = (int32_t)k;
        }
    }
    return FM_OK;
}

extern "C" int fm_mass_p1_lumped_gather(int64_t n_elems, int64_t n_nodes, const int32_t *elems, const int32_t *pos,
                                        int dtype, const void *nodes, const int64_t *chunk_offsets,
                                        const int32_t *degree, const int32_t *incident, int64_t n_slots,
                                        double *det_scratch, double *w, void *stream)
{
    FM_CHECK_ARG(n_elems >= 0 && n_nodes >= 0 && n_slots >= 0, "negative size");
    FM_CHECK_ARG(dtype == FM_DTYPE_F32 || dtype == FM_DTYPE_F64, "unknown dtype %d", dtype);
    // every element has 4 incidence slots, so a table built for at least
    // these elements by fm_mass_p1_incidence holds at least 4 n_elems
    FM_CHECK_ARG(n_slots >= 4 * n_elems,
                 "n_slots (%lld) below 4 n_elems: pass the slot count fm_mass_p1_incidence returned for this mesh",
                 (long long)n_slots);
    if (n_nodes == 0) return FM_OK;
    FM_CHECK_ARG(w && chunk_offsets && degree, "NULL array");
    FM_CHECK_ARG(n_elems <= 0x7fffffffLL && n_nodes <= 0x7fffffffLL * 32, "mesh too large for one launch");
    hipStream_t st = as_stream(stream);
    if (n_elems > 0) {
        FM_CHECK_ARG(elems && nodes && incident && det_scratch, "NULL array");
        const unsigned eblocks = xcd_grid(cdiv(n_elems, 256));
        if (dtype == FM_DTYPE_F64)
            k_abs_dets<double><<<eblocks, 256, 0, st>>>(n_elems, elems, pos, (const double *)nodes, det_scratch);
        else
            k_abs_dets<float><<<eblocks, 256, 0, st>>>(n_elems, elems, pos, (const float *)nodes, det_scratch);
        FM_LAUNCH_CHECK("k_abs_dets");
    }
    k_lumped_gather<<<xcd_grid(cdiv(n_nodes, 256)), 256, 0, st>>>(n_nodes, n_elems, chunk_offsets, degree,
                                                                   incident, det_scratch, w);
    FM_LAUNCH_CHECK("k_lumped_gather");
    return FM_OK;
}

extern "C" int fm_mass_p1_element_matrices(int64_t n_elems, const int32_t *elems, int dtype, const void *nodes,
                                           double *vals, void *stream)
{
    FM_CHECK_ARG(n_elems >= 0, "negative size");
    FM_CHECK_ARG(dtype == FM_DTYPE_F32 || dtype == FM_DTYPE_F64, "unknown dtype %d", dtype);
    if (n_elems == 0) return FM_OK;
    FM_CHECK_ARG(elems && nodes && vals, "NULL array");
    const unsigned blocks = (unsigned)cdiv(n_elems, 256);
    hipStream_t st = as_stream(stream);
    if (dtype == FM_DTYPE_F64)
        k_element_matrices<double><<<blocks, 256, 0, st>>>(n_elems, elems, (const double *)nodes, vals);
    else
        k_element_matrices<float><<<blocks, 256, 0, st>>>(n_elems, elems, (const float *)nodes, vals);
    FM_LAUNCH_CHECK("k_element_matrices");
    return FM_OK;
}
