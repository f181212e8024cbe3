/*
 * oracle.c -- FP64 CPU restatement of FEelMRI's signal-generation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/_build/liboracle.so (strict IEEE flags, used for parity) and
 * oracle/_build/liboracle_fast.so (the reference's -O3 -ffast-math flags,
 * used only as the timed CPU baseline).  Nothing in the product package
 * (feelmri_amd/) may load or call it; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg do.
 *
 * Every function follows one reference routine loop for loop; the citation
 * is given on each.  Paths are relative to the reference repository root.
 *
 * Units (as in the reference): positions m, k cycles/m, t/dt/T1/T2 ms,
 * phi_dB0 rad/ms, delta_B mT, G mT/m, rf mT, gamma rad/(ms*mT).
 * Complex arrays are interleaved (re, im) doubles.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* cpp/feelmri/MRI.h:13 declares `const float PI = M_PI;`, so every
 * two_pi = 2.0 * PI in MRI.cpp (:73, :231) is 2*(double)(float)pi even in
 * Signal<double>. */
static const float ORC_PI_F = (float)M_PI;

int orc_version(void) { return 3; }

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * Signal, static nodes.  cpp/feelmri/MRI.cpp:196-297 (Signal<T>(..., py::none)).
 *
 *   S[s, v] = sum_n exp(i(-2pi_f (kx x_n + ky y_n + kz z_n) + phi_dB0_n t_s))
 *                   * exp(-t_s / T2_n) * MMxy[n, v]
 *
 * MMxy = M * Mxy is formed by the caller (MRI.cpp:243).  Samples s run in the
 * reference's (i, j, k) row-major order; the per-node decay and off-resonance
 * terms are recomputed only when t changes (MRI.cpp:260-276), exactly as in
 * the reference.  With OpenMP each thread owns a contiguous block of samples,
 * which leaves every per-sample sum in the same node order as the serial loop.
 *
 * r is (n_nodes, 3) row-major; mmxy is (n_nodes, n_enc) complex row-major;
 * out is (n_samples, n_enc) complex row-major.
 * ---------------------------------------------------------------------- */
void orc_signal(int64_t n_nodes, int64_t n_samples, int64_t n_enc,
                const double *kx, const double *ky, const double *kz,
                const double *t, const double *r, const double *phi_dB0,
                const double *T2, const double *mmxy, double *out,
                int n_threads)
{
    const double two_pi = 2.0 * (double)ORC_PI_F;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        const int64_t s0 = n_samples * tid / nth;
        const int64_t s1 = n_samples * (tid + 1) / nth;
        double *phi_off = (double *)malloc(sizeof(double) * (size_t)(n_nodes > 0 ? n_nodes : 1));
        double *decay = (double *)malloc(sizeof(double) * (size_t)(n_nodes > 0 ? n_nodes : 1));
        double *acc = (double *)malloc(sizeof(double) * 2 * (size_t)(n_enc > 0 ? n_enc : 1));
        double t_old = -1.0;
        int have_t = 0;
        for (int64_t s = s0; s < s1; ++s) {
            if (!have_t || t_old != t[s]) {
                /* MRI.cpp:264-274 */
                for (int64_t n = 0; n < n_nodes; ++n) {
                    phi_off[n] = phi_dB0[n] * t[s];
                    decay[n] = exp(-t[s] * (1.0 / T2[n]));
                }
                t_old = t[s];
                have_t = 1;
            }
            /* MRI.cpp:279 */
            const double kx_ = two_pi * kx[s], ky_ = two_pi * ky[s], kz_ = two_pi * kz[s];
            for (int64_t v = 0; v < n_enc; ++v) acc[2 * v] = acc[2 * v + 1] = 0.0;
            for (int64_t n = 0; n < n_nodes; ++n) {
                /* MRI.cpp:282-288 */
                const double phase = -kx_ * r[3 * n] - ky_ * r[3 * n + 1] - kz_ * r[3 * n + 2] + phi_off[n];
                const double c = cos(phase), sn = sin(phase);
                /* MRI.cpp:291: fourier^T * (MMxy .* decay) */
                for (int64_t v = 0; v < n_enc; ++v) {
                    const double br = mmxy[2 * (n * n_enc + v)] * decay[n];
                    const double bi = mmxy[2 * (n * n_enc + v) + 1] * decay[n];
                    acc[2 * v] += c * br - sn * bi;
                    acc[2 * v + 1] += c * bi + sn * br;
                }
            }
            for (int64_t v = 0; v < n_enc; ++v) {
                out[2 * (s * n_enc + v)] = acc[2 * v];
                out[2 * (s * n_enc + v) + 1] = acc[2 * v + 1];
            }
        }
        free(phi_off);
        free(decay);
        free(acc);
    }
}

/* ------------------------------------------------------------------------
 * Bloch integrator.  cpp/feelmri/BlochSimulator.cpp:151-276 (static nodes,
 * solve_mri<T>(..., py::none)) and :12-148 (moving nodes), instantiated at
 * T = double (the reference binds only T = float, BlochSimulator.h:50-80).
 *
 * Step i -> i+1 uses rf[i+1], G[i+1,:], dt[i+1]; column 0 is the initial
 * condition.  e1/e2 are refreshed only when dt changes (:207-214); |B| is
 * floored at 1e-12 (:238-239); rotation first, then relaxation toward the
 * scalar M0 (:270-271).  regime_idx is unused by the reference (:253-267).
 *
 * disp: NULL for static nodes, else (n_time-1, n_pos, 3) row-major: the
 * displacement pod_trajectory(ti) evaluated at the end of step i
 * (ti = dt[1] + ... + dt[i+1], BlochSimulator.cpp:81-89).
 *
 * history != 0: mxy_out is (n_pos, n_time) complex and mz_out (n_pos,
 * n_time), column-major like the reference's Eigen matrices
 * (element (p, i) at p + i*n_pos).  history == 0: only the last column
 * (n_pos values) is written.
 * ---------------------------------------------------------------------- */
void orc_bloch(int64_t n_pos, int64_t n_time,
               const double *r0, const double *T1, const double *T2,
               const double *delta_B, double M0, double gamma,
               const double *rf, const double *G, const double *dt,
               const double *mxy0, const double *mz0, const double *disp,
               double *mxy_out, double *mz_out, int history, int n_threads)
{
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(n_threads)
#endif
    for (int64_t p = 0; p < n_pos; ++p) {
        /* Loop interchange relative to the reference (event loop outer,
         * node loop inner): every node evolves independently, so each node
         * sees exactly the same sequence of floating-point operations. */
        double mxr = mxy0[2 * p], mxi = mxy0[2 * p + 1], mz = mz0[p];
        if (history) {
            mxy_out[2 * p] = mxr;
            mxy_out[2 * p + 1] = mxi;
            mz_out[p] = mz;
        }
        const double invT1 = 1.0 / T1[p], invT2 = 1.0 / T2[p];
        double e1 = 0.0, e2 = 0.0, dt_i = -1.0;
        for (int64_t i = 0; i + 1 < n_time; ++i) {
            if (dt_i != dt[i + 1]) {
                e1 = exp(-dt[i + 1] * invT1);
                e2 = exp(-dt[i + 1] * invT2);
            }
            dt_i = dt[i + 1];
            const double rfr = rf[2 * (i + 1)], rfi = rf[2 * (i + 1) + 1];
            const double rf2 = rfr * rfr + rfi * rfi;
            const double kappa = -0.5 * gamma * dt_i;
            const double Gx = G[3 * (i + 1)], Gy = G[3 * (i + 1) + 1], Gz = G[3 * (i + 1) + 2];
            double x = r0[3 * p], y = r0[3 * p + 1], z = r0[3 * p + 2];
            if (disp) {
                const double *d = disp + 3 * ((size_t)i * (size_t)n_pos + (size_t)p);
                x += d[0];
                y += d[1];
                z += d[2];
            }
            const double Bz = x * Gx + y * Gy + z * Gz + delta_B[p];
            double Bn = sqrt(Bz * Bz + rf2);
            if (Bn < 1e-12) Bn = 1e-12;
            const double nz = Bz / Bn;
            const double hphi = Bn * kappa;
            const double c = cos(hphi), s = sin(hphi);
            /* alpha = c - i nz s ; beta = -i (rf/|B|) s */
            const double ar = c, ai = -nz * s;
            const double nxr = rfr / Bn, nxi = rfi / Bn;
            const double br = nxi * s, bi = -nxr * s;
            const double a2 = ar * ar + ai * ai;
            const double b2 = br * br + bi * bi;
            /* conj(alpha) */
            const double car = ar, cai = -ai;
            /* 2 conj(a) b Mz */
            const double cab_r = car * br - cai * bi, cab_i = car * bi + cai * br;
            /* conj(a)^2 */
            const double ca2_r = car * car - cai * cai, ca2_i = 2.0 * car * cai;
            /* beta^2 */
            const double bb_r = br * br - bi * bi, bb_i = 2.0 * br * bi;
            /* conj(Mxy) */
            const double cmr = mxr, cmi = -mxi;
            const double nmr = 2.0 * cab_r * mz + (ca2_r * mxr - ca2_i * mxi) - (bb_r * cmr - bb_i * cmi);
            const double nmi = 2.0 * cab_i * mz + (ca2_r * mxi + ca2_i * mxr) - (bb_r * cmi + bb_i * cmr);
            /* Mz' = (|a|^2 - |b|^2) Mz - 2 Re(a b conj(Mxy)) */
            const double ab_r = ar * br - ai * bi, ab_i = ar * bi + ai * br;
            const double nmz = (a2 - b2) * mz - 2.0 * (ab_r * cmr - ab_i * cmi);
            mxr = nmr * e2;
            mxi = nmi * e2;
            mz = nmz * e1 + (1.0 - e1) * M0;
            if (history) {
                const size_t o = (size_t)p + (size_t)(i + 1) * (size_t)n_pos;
                mxy_out[2 * o] = mxr;
                mxy_out[2 * o + 1] = mxi;
                mz_out[o] = mz;
            }
        }
        if (!history) {
            mxy_out[2 * p] = mxr;
            mxy_out[2 * p + 1] = mxi;
            mz_out[p] = mz;
        }
    }
}

/* ------------------------------------------------------------------------
 * P1 tetrahedron mass matrix.  cpp/feelmri/Assemble.cpp:15-116
 * (IsoMapping, LocalMassAssemble, MassAssemble) for the P1 Lagrange element
 * on the reference tetrahedron with an order-2 quadrature rule (the call in
 * python/feelmri/Phantom.py:379-392 uses quadrature_order=2).
 *
 * P1 basis: S = (1-r-s-t, r, s, t); dS/dr = (-1,1,0,0), dS/ds = (-1,0,1,0),
 * dS/dt = (-1,0,0,1).  J(a,b) = sum_i dS_a(i) x_i(b) (Assemble.cpp:36-45),
 * Me = sum_q S S^T |det J| w_q (Assemble.cpp:69-73).
 *
 * The quadrature is the symmetric 4-point rule exact for degree 2
 * (points (a,b,b), (b,a,b), (b,b,a), (b,b,b) with b = (5-sqrt5)/20,
 * a = 1-3b, weights 1/24).  S_i S_j is quadratic, so any degree-2-exact rule
 * (basix's default included) yields the same element matrix.
 *
 * vals: (n_elems, 4, 4) element matrices, row-major, in element order -- the
 * triplets the reference emits at Assemble.cpp:109-111.
 * ---------------------------------------------------------------------- */
void orc_mass_p1_element_matrices(int64_t n_elems, const int32_t *elems,
                                  const double *nodes, double *vals)
{
    const double b = (5.0 - sqrt(5.0)) / 20.0, a = 1.0 - 3.0 * b;
    const double qp[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
    const double qw = 1.0 / 24.0;
    for (int64_t e = 0; e < n_elems; ++e) {
        const int32_t *el = elems + 4 * e;
        double X[4][3];
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 3; ++c) X[i][c] = nodes[3 * (int64_t)el[i] + c];
        double J[3][3];
        for (int rr = 0; rr < 3; ++rr)
            for (int c = 0; c < 3; ++c) J[rr][c] = X[rr + 1][c] - X[0][c];
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                         - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                         + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        const double adet = fabs(det);
        double *Me = vals + 16 * e;
        for (int k = 0; k < 16; ++k) Me[k] = 0.0;
        for (int q = 0; q < 4; ++q) {
            const double S[4] = {1.0 - qp[q][0] - qp[q][1] - qp[q][2], qp[q][0], qp[q][1], qp[q][2]};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) Me[4 * i + j] += S[i] * S[j] * adet * qw;
        }
    }
}

/* Lumped weights: w_i = sum_j M_ij (python/feelmri/Phantom.py:395-398),
 * accumulated from the element matrices in element order, the order in
 * which Eigen's setFromTriplets sums duplicates (Assemble.cpp:114). */
void orc_mass_p1_lumped(int64_t n_elems, int64_t n_nodes, const int32_t *elems,
                        const double *nodes, double *w)
{
    for (int64_t n = 0; n < n_nodes; ++n) w[n] = 0.0;
    double Me[16];
    for (int64_t e = 0; e < n_elems; ++e) {
        orc_mass_p1_element_matrices(1, elems + 4 * e, nodes, Me);
        for (int i = 0; i < 4; ++i) {
            double row = 0.0;
            for (int j = 0; j < 4; ++j) row += Me[4 * i + j];
            w[elems[4 * e + i]] += row;
        }
    }
}

/* ------------------------------------------------------------------------
 * POD mode contraction.  cpp/feelmri/POD.cpp:5-38: (P,C,M) . (M,) -> (P,C),
 * the modes viewed as a (P*C) x M row-major matrix times the weight vector.
 * ---------------------------------------------------------------------- */
void orc_pod_tensordot(int64_t P, int64_t C, int64_t M, const double *modes,
                       const double *weights, double *out)
{
    for (int64_t r = 0; r < P * C; ++r) {
        double acc = 0.0;
        for (int64_t m = 0; m < M; ++m) acc += modes[r * M + m] * weights[m];
        out[r] = acc;
    }
}

/* Sparse (CSR) complex SpMV y = M x, x (n, n_enc) complex: the M * Mxy
 * product of MRI.cpp:243 for a general (consistent) mass matrix. */
void orc_csr_spmv_c128(int64_t n_rows, int64_t n_enc, const int64_t *indptr,
                       const int64_t *indices, const double *data,
                       const double *x, double *y)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        for (int64_t v = 0; v < n_enc; ++v) {
            double ar = 0.0, ai = 0.0;
            for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
                const int64_t c = indices[k];
                ar += data[k] * x[2 * (c * n_enc + v)];
                ai += data[k] * x[2 * (c * n_enc + v) + 1];
            }
            y[2 * (r * n_enc + v)] = ar;
            y[2 * (r * n_enc + v) + 1] = ai;
        }
    }
}
