"""Pins for the FP64 Bloch oracle (oracle.orc_bloch, oracle.bloch_solver_solve).

Closed forms: on-resonance hard pulse, free precession with T2 decay and
off-resonance, T1 recovery, the Ernst/FLASH steady state under the
reference's ideal spoiling (python/feelmri/Bloch.py:363-368), rotation norm
preservation, and the step semantics of BlochSimulator.cpp:204-271
(column 0 = initial state; rf[0], G[0], dt[0] unused).
"""
import numpy as np

GAMMA = 42.58e6 * 2 * np.pi / 1e6
BIG = 1e300


def _nodes(rng, n):
    r = rng.uniform(-0.05, 0.05, (n, 3))
    return r


def test_hard_pulse_on_resonance(oracle):
    n, steps, dt = 5, 40, 0.005
    theta = np.deg2rad(37.0)
    B1 = theta / (GAMMA * steps * dt)
    phase = 0.4
    rf = np.full(steps + 1, B1 * np.exp(1j * phase))
    G = np.zeros((steps + 1, 3))
    dts = np.full(steps + 1, dt)
    r = _nodes(np.random.default_rng(0), n)
    mxy, mz = oracle.solve_mri(r, np.full(n, BIG), np.full(n, BIG), np.zeros(n), 1.0, GAMMA, rf, G, dts, None,
                               np.zeros(n), np.ones(n), history=False)
    np.testing.assert_allclose(mxy, 1j * np.exp(1j * phase) * np.sin(theta), rtol=1e-12)
    np.testing.assert_allclose(mz, np.cos(theta), rtol=1e-12)


def test_free_precession_decay_and_recovery(oracle):
    rng = np.random.default_rng(1)
    n, steps = 6, 25
    r = _nodes(rng, n)
    G = np.tile([3.0, -2.0, 1.5], (steps + 1, 1))
    dts = np.full(steps + 1, 0.02)
    dB = rng.normal(0, 1e-4, n)
    T1 = rng.uniform(500, 1500, n)
    T2 = rng.uniform(40, 100, n)
    mxy0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    mz0 = rng.uniform(-1, 1, n)
    M0 = 0.9
    mxy, mz = oracle.solve_mri(r, T1, T2, dB, M0, GAMMA, np.zeros(steps + 1), G, dts, None, mxy0, mz0,
                               history=False)
    T = steps * 0.02
    Bz = r @ G[0] + dB
    np.testing.assert_allclose(mxy, mxy0 * np.exp(-1j * GAMMA * Bz * T) * np.exp(-T / T2), rtol=1e-10)
    np.testing.assert_allclose(mz, M0 + (mz0 - M0) * np.exp(-T / T1), rtol=1e-12)


def test_ernst_steady_state_with_ideal_spoiling(oracle):
    n = 4
    r = _nodes(np.random.default_rng(2), n)
    T1 = np.array([200.0, 350.0, 500.0, 800.0])
    T2 = np.array([50.0, 60.0, 70.0, 80.0])
    theta = np.deg2rad(15.0)
    tau = 0.01
    B1 = theta / (GAMMA * tau)
    pulse = dict(rf=np.array([0.0, B1]), G=np.zeros((2, 3)), dt=np.array([0.0, tau]), empty=False,
                 store=True, t0=0.0)
    delay_dt = np.array([0.0] + [1.0] * 4 + [0.99])
    delay = dict(rf=np.zeros(6), G=np.zeros((6, 3)), dt=delay_dt, empty=True, store=False, t0=0.0)
    blocks = [pulse, delay] * 1200
    mxy, _, _, _ = oracle.bloch_solver_solve(blocks, r, T1, T2, np.zeros(n), 1.0, GAMMA, np.zeros(n),
                                             np.ones(n))
    E1 = np.exp(-(tau + delay_dt[1:].sum()) / T1)
    mz_minus = (1 - E1) / (1 - E1 * np.cos(theta))
    want = 1j * np.sin(theta) * mz_minus * np.exp(-tau / T2)
    np.testing.assert_allclose(mxy[:, -1], want, rtol=1e-10)


def test_norm_preserved_without_relaxation(oracle):
    rng = np.random.default_rng(3)
    n, T = 8, 60
    r = _nodes(rng, n)
    rf = (rng.normal(size=T) + 1j * rng.normal(size=T)) * 5e-3
    G = rng.normal(size=(T, 3)) * 10
    dts = rng.uniform(0.001, 0.02, T)
    mxy0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    mz0 = rng.normal(size=n)
    mxy, mz = oracle.solve_mri(r, np.full(n, BIG), np.full(n, BIG), rng.normal(0, 1e-3, n), 1.0, GAMMA, rf, G,
                               dts, None, mxy0, mz0, history=True)
    norm = np.abs(mxy) ** 2 + mz ** 2
    np.testing.assert_allclose(norm, (np.abs(mxy0) ** 2 + mz0 ** 2)[:, None] * np.ones((1, T)), rtol=1e-12)


def test_step_semantics(oracle):
    rng = np.random.default_rng(4)
    n, T = 5, 12
    r = _nodes(rng, n)
    rf = (rng.normal(size=T) + 1j * rng.normal(size=T)) * 1e-3
    G = rng.normal(size=(T, 3))
    dts = rng.uniform(0.01, 0.1, T)
    args = (np.full(n, 900.0), np.full(n, 80.0), np.zeros(n), 1.0, GAMMA)
    mxy0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    mz0 = rng.normal(size=n)
    a = oracle.solve_mri(r, *args, rf, G, dts, None, mxy0, mz0, history=True)
    rf2, G2, dt2 = rf.copy(), G.copy(), dts.copy()
    rf2[0], G2[0], dt2[0] = 99.0, 99.0, 99.0
    b = oracle.solve_mri(r, *args, rf2, G2, dt2, None, mxy0, mz0, history=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[0][:, 0], mxy0)
    np.testing.assert_array_equal(a[1][:, 0], mz0)
    # final-only mode returns the last history column
    c = oracle.solve_mri(r, *args, rf, G, dts, None, mxy0, mz0, history=False)
    np.testing.assert_array_equal(c[0], a[0][:, -1])


def test_motion_positions_at_step_end(oracle):
    # free precession of a node moving with constant velocity: phase is
    # gamma G . (r0 t + v t^2/2) accumulated with the position at each
    # step's end (the reference evaluates motion after ti += dt)
    n, steps, dt = 1, 30, 0.01
    r = np.array([[0.01, 0.0, 0.0]])
    v = np.array([0.05, 0.0, 0.0])
    G = np.tile([20.0, 0, 0], (steps + 1, 1))
    mxy, _ = oracle.solve_mri(r, np.full(n, BIG), np.full(n, BIG), np.zeros(n), 1.0, GAMMA,
                              np.zeros(steps + 1), G, np.full(steps + 1, dt), None, np.ones(n), np.zeros(n),
                              lambda t: (v * t)[None, :], history=False)
    ti = dt * np.arange(1, steps + 1)
    phase = -GAMMA * 20.0 * (r[0, 0] + v[0] * ti).sum() * dt
    np.testing.assert_allclose(mxy, np.exp(1j * phase), rtol=1e-11)
