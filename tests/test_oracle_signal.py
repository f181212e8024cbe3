"""Pins for the FP64 Signal oracle (oracle.orc_signal / oracle.Signal).

Each check is fixed by mathematics independent of the oracle's loop:
closed forms (one node, Dirichlet kernel of a lattice), the shift theorem,
linearity/additivity, and the reference's float-pi constant
(cpp/feelmri/MRI.h:13) and +phi t sign (MRI.cpp:282-285).
"""
import numpy as np
import pytest

PI_F = float(np.float32(np.pi))


def _rand_nodes(rng, n, v=1):
    r = rng.uniform(-0.05, 0.05, (n, 3))
    phi = rng.normal(0, 0.3, n)
    T2 = rng.uniform(40, 100, n)
    mxy = rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))
    return r, phi, T2, mxy


def test_single_node_closed_form(oracle):
    rng = np.random.default_rng(1)
    r, phi, T2, mxy = _rand_nodes(rng, 1, 2)
    S = 37
    kx, ky, kz = rng.uniform(-300, 300, (3, S))
    t = rng.uniform(0, 5, S)
    out = oracle.signal_samples(kx, ky, kz, t, r, phi, T2, mxy)
    phase = -2 * PI_F * (kx * r[0, 0] + ky * r[0, 1] + kz * r[0, 2]) + phi[0] * t
    ref = (np.exp(-t / T2[0]) * np.exp(1j * phase))[:, None] * mxy[0][None, :]
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_float_pi_constant(oracle):
    # one node far from the origin, large k: float-pi and exact-pi phases
    # differ by ~1e-5 rad; the oracle must follow the float constant
    r = np.array([[0.1, 0.0, 0.0]])
    kx = np.array([800.0])
    z = np.zeros(1)
    out = oracle.signal_samples(kx, z, z, z, r, np.zeros(1), np.full(1, 1e300), np.ones((1, 1)))[0, 0]
    want_f = np.exp(-1j * 2 * PI_F * 80.0)
    want_exact = np.exp(-1j * 2 * np.pi * 80.0)
    assert abs(out - want_f) < 1e-11
    assert abs(out - want_exact) > 1e-6


def _dirichlet(theta, M):
    # sum_{m<M} exp(-i theta m)
    theta = np.asarray(theta, dtype=np.float64)
    small = np.abs(np.sin(theta / 2)) < 1e-14
    num = 1 - np.exp(-1j * theta * M)
    den = 1 - np.exp(-1j * theta)
    return np.where(small, M * np.exp(-1j * theta * 0), num / np.where(small, 1, den))


def test_lattice_dirichlet_kernel(oracle):
    # uniform 3-D lattice with equal weights: S factorises into three
    # geometric series (Dirichlet kernels)
    Mx, My, Mz = 7, 5, 3
    h = np.array([0.011, 0.013, 0.017])
    x0 = np.array([-0.03, -0.02, 0.005])
    g = np.stack(np.meshgrid(np.arange(Mx), np.arange(My), np.arange(Mz), indexing="ij"), -1).reshape(-1, 3)
    r = x0 + g * h
    N = r.shape[0]
    rng = np.random.default_rng(2)
    S = 50
    k = rng.uniform(-120, 120, (3, S))
    out = oracle.signal_samples(*k, np.zeros(S), r, np.zeros(N), np.full(N, 1e300), np.full((N, 1), 0.7 + 0.2j))
    th = 2 * PI_F * k * h[:, None]
    ref = (0.7 + 0.2j) * np.exp(-1j * 2 * PI_F * (k * x0[:, None]).sum(0))
    ref = ref * _dirichlet(th[0], Mx) * _dirichlet(th[1], My) * _dirichlet(th[2], Mz)
    np.testing.assert_allclose(out[:, 0], ref, rtol=1e-10, atol=1e-10 * N)


def test_uniform_offresonance_and_decay_factor_out(oracle):
    rng = np.random.default_rng(3)
    r, _, _, mxy = _rand_nodes(rng, 40)
    N = r.shape[0]
    S = 20
    k = rng.uniform(-200, 200, (3, S))
    t = rng.uniform(0, 4, S)
    base = oracle.signal_samples(*k, np.zeros(S), r, np.zeros(N), np.full(N, 1e300), mxy)
    dw, T2 = 0.9, 55.0
    out = oracle.signal_samples(*k, t, r, np.full(N, dw), np.full(N, T2), mxy)
    # reference sign: phase = -2 pi k.x + phi t  (exp(+i dw t))
    np.testing.assert_allclose(out, base * (np.exp(-t / T2) * np.exp(1j * dw * t))[:, None], rtol=1e-11,
                               atol=1e-12)


def test_k0_t0_is_weighted_sum(oracle):
    rng = np.random.default_rng(4)
    r, phi, T2, mxy = _rand_nodes(rng, 100, 3)
    z = np.zeros(1)
    out = oracle.signal_samples(z, z, z, z, r, phi, T2, mxy)
    np.testing.assert_allclose(out[0], mxy.sum(0), rtol=1e-13)


def test_linearity_and_partition_additivity(oracle):
    rng = np.random.default_rng(5)
    r, phi, T2, a = _rand_nodes(rng, 90, 2)
    b = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    S = 25
    k = rng.uniform(-250, 250, (3, S))
    t = np.repeat(rng.uniform(0, 3, 5), 5)
    f = lambda m, sl=slice(None): oracle.signal_samples(*k, t, r[sl], phi[sl], T2[sl], m[sl])
    c = 0.3 - 1.7j
    np.testing.assert_allclose(f(a + c * b), f(a) + c * f(b), rtol=1e-10, atol=1e-10)
    # the MPI invariant: sum over node partitions equals the global sum
    np.testing.assert_allclose(f(a), f(a, slice(0, 33)) + f(a, slice(33, None)), rtol=1e-10, atol=1e-10)


def test_shift_theorem_for_motion_callable(oracle):
    rng = np.random.default_rng(6)
    r, phi, T2, mxy = _rand_nodes(rng, 30)
    N = r.shape[0]
    kx, ky, kz = rng.uniform(-200, 200, (3, 4, 3, 2))
    t = np.broadcast_to(np.linspace(0.5, 2.0, 4)[:, None, None], (4, 3, 2)).copy()
    d = np.array([0.004, -0.002, 0.001])
    static = oracle.Signal(0, None, (kx, ky, kz), t, r, phi, T2, mxy)
    moved = oracle.Signal(0, None, (kx, ky, kz), t, r, phi, T2, mxy, lambda tt: np.tile(d, (N, 1)))
    ramp = np.exp(-1j * 2 * PI_F * (kx * d[0] + ky * d[1] + kz * d[2]))
    np.testing.assert_allclose(moved, static * ramp[..., None], rtol=1e-10, atol=1e-12)


def test_motion_evaluated_at_sample_time(oracle):
    # linear motion r(t) = r0 + v t, single node: closed form per sample
    r = np.array([[0.01, -0.02, 0.003]])
    v = np.array([0.002, 0.001, -0.0005])
    rng = np.random.default_rng(7)
    kx, ky, kz = rng.uniform(-300, 300, (3, 6, 2, 1))
    t = np.broadcast_to(np.linspace(0.1, 3.0, 6)[:, None, None], (6, 2, 1)).copy()
    out = oracle.Signal(0, None, (kx, ky, kz), t, r, np.zeros(1), np.full(1, 1e300), np.ones((1, 1)),
                        lambda tt: (v * tt)[None, :])
    pos = r[0][None, None, None, :] + v * t[..., None]
    ref = np.exp(-1j * 2 * PI_F * (kx * pos[..., 0] + ky * pos[..., 1] + kz * pos[..., 2]))
    np.testing.assert_allclose(out[..., 0], ref, rtol=1e-11, atol=1e-12)


def test_sparse_mass_product(oracle):
    import scipy.sparse as sp

    rng = np.random.default_rng(8)
    M = sp.random(60, 60, density=0.1, random_state=1, format="csc") + sp.eye(60)
    mxy = rng.normal(size=(60, 2)) + 1j * rng.normal(size=(60, 2))
    np.testing.assert_allclose(oracle.weighted_mxy(M, mxy), M @ mxy, rtol=1e-13)
    w = rng.uniform(1, 2, 60)
    np.testing.assert_allclose(oracle.weighted_mxy(sp.diags(w), mxy), w[:, None] * mxy, rtol=1e-14)


def test_threads_do_not_change_results(oracle):
    rng = np.random.default_rng(9)
    r, phi, T2, mxy = _rand_nodes(rng, 70)
    S = 41
    k = rng.uniform(-250, 250, (3, S))
    t = rng.uniform(0, 3, S)
    a = oracle.signal_samples(*k, t, r, phi, T2, mxy, n_threads=1)
    b = oracle.signal_samples(*k, t, r, phi, T2, mxy, n_threads=4)
    assert np.array_equal(a, b)
