"""Pins for the host producers that feed the kernels: the event grid of a
sequence block (python/feelmri/Bloch.py:105-129, :340), the Cartesian and
golden-angle radial k-space extents (python/feelmri/KSpaceTraj.py:40-44,
:183-220, :335-379), the spiral's rotation of interleaves and the
gradient-limited spiral stack (KSpaceTraj.py:386-562): its time law against
closed forms and quadrature in the amplitude- and slew-limited regimes, the
ADC grid, the end points, the turn count and the kz planes.

Each is checked by what the definitions fix, not by retyping them: grid
membership of every knot, the half-open dt grid, dt = diff(prepend 0), the
integrated RF area against the nominal flip; the centre sample at k = 0, the
sample spacing 1/(oversampling FOV), the golden-angle increment
2 pi / phi^2 between spokes and the spoke radii.
"""
import numpy as np
import pytest

from feelmri_amd import sequences, trajectories


def test_block_grid_contains_knots_and_half_open_dt_grid():
    p = sequences.HardPulse(start=2.0, dur=0.3, flip=np.deg2rad(20))
    g = sequences.Trapezoid(axis=1, start=2.5, amp=10.0, ramp=0.1, flat=0.4)
    b = sequences.make_block(2.0, 3.0, pulses=[p], gradients=[g], dt_rf=0.05, dt_gr=0.1, dt=1.0)
    t = b.times
    assert np.all(np.diff(t) > 0)                                   # sorted, unique
    for knot in [2.0, 2.3, *g.timings]:                              # RF start/end, gradient knots
        assert np.min(np.abs(t - knot)) < 1e-12, knot
    # the block dt grid starts at the extent and excludes its end (arange)
    assert np.min(np.abs(t - 3.0)) < 1e-12 and np.min(np.abs(t - 4.0)) < 1e-12
    assert np.min(np.abs(t - 5.0)) > 0.5                            # extent end 5.0 is not on the grid
    # dt = diff(times, prepend=0): the first step is the absolute start time
    assert b.dt[0] == t[0] == 2.0
    np.testing.assert_allclose(np.cumsum(b.dt), t, rtol=0, atol=1e-12)
    # grid spacing inside the pulse and the gradient
    inside = (t > 2.0) & (t <= 2.3)
    assert np.all(np.diff(t[(t >= 2.0) & (t <= 2.3)]) <= 0.05 + 1e-12) and inside.sum() >= 6
    assert np.all(np.diff(t[(t >= 2.5) & (t <= g.timings[-1])]) <= 0.1 + 1e-12)


def test_block_rf_area_is_the_flip_and_gradient_follows_trapezoid():
    flip = np.deg2rad(15.0)
    p = sequences.HardPulse(start=0.0, dur=0.2, flip=flip)
    b = sequences.make_block(0.0, 0.2, pulses=[p], dt_rf=0.01, dt=1.0)
    # step i integrates over (t_{i-1}, t_i] with rf(t_i): the hard pulse's
    # B1 times its duration, so gamma * sum |rf| dt is the flip angle
    np.testing.assert_allclose(sequences.GAMMA * np.sum(np.abs(b.rf[1:]) * b.dt[1:]), flip, rtol=1e-12)
    # [start, end]: on at the start knot too (the reference's RF table is
    # read with 0 outside the closed interval); the first event's value is
    # never integrated
    assert b.rf[0] == p.b1 and b.rf[-1] == p.b1
    g = sequences.Trapezoid(axis=2, start=1.0, amp=-7.0, ramp=0.2, flat=0.6)
    b2 = sequences.make_block(1.0, 1.0, gradients=[g], dt_gr=0.05, dt=1.0)
    area = np.sum(b2.G[1:, 2] * b2.dt[1:])                           # right-point rule on a piecewise-linear G
    np.testing.assert_allclose(area, -7.0 * (0.6 + 0.2), rtol=0.02)
    assert b2.G[:, 0].max() == 0 and b2.G[:, 1].max() == 0


def test_delay_block():
    b = sequences.delay_block(5.0, 3.0, dt=1.0)
    assert b.empty and not b.store
    np.testing.assert_allclose(b.times, [5.0, 6.0, 7.0])            # end excluded
    assert not np.any(b.rf) and not np.any(b.G)


def test_spoiled_gre_blocks():
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=3, n_lines=2, rf_dur=0.2)
    assert len(blocks) == 10
    assert [b.store for b in blocks[::2]] == [False, False, False, True, True]
    assert all(b.empty for b in blocks[1::2]) and not any(b.empty for b in blocks[::2])
    starts = [b.t0 for b in blocks[::2]]
    np.testing.assert_allclose(np.diff(starts), 5.0)
    # each TR (imaging + delay block) ends where the next starts
    for img, dly, nxt in zip(blocks[::2], blocks[1::2], blocks[2::2]):
        assert abs(dly.info["extent"][1] - nxt.t0) < 1e-9


@pytest.mark.parametrize("res,ofac", [((64, 48, 8), 1), ((40, 30, 5), 2), ((33, 17, 3), 1)])
def test_cartesian_extents(res, ofac):
    fov = (0.24, 0.20, 0.06)
    (kx, ky, kz), t = trajectories.cartesian(fov, res, oversampling=ofac, t_first=1.0, t_readout=2.0)
    ro = ofac * res[0]
    assert kx.shape == (ro, res[1], res[2])
    dk = 1.0 / (np.array([ofac, 1, 1]) * np.array(fov))
    np.testing.assert_allclose(np.diff(kx[:, 0, 0]), dk[0], rtol=1e-12)
    np.testing.assert_allclose(np.diff(ky[0, :, 0]), dk[1], rtol=1e-12)
    if res[2] > 1:
        np.testing.assert_allclose(np.diff(kz[0, 0, :]), dk[2], rtol=1e-12)
    # centre sample: k = 0 at index n // 2 (readout: plus one step for an odd resolution)
    assert abs(kx[ro // 2, 0, 0] - (dk[0] if res[0] % 2 else 0.0)) < 1e-9 * dk[0]
    assert abs(ky[0, res[1] // 2, 0]) < 1e-9 * dk[1]
    assert abs(kz[0, 0, res[2] // 2]) < 1e-9 * dk[2]
    # every line and slice has the same readout times, t_first .. t_first + t_readout
    np.testing.assert_allclose(t, np.broadcast_to(t[:, :1, :1], t.shape))
    np.testing.assert_allclose(t[[0, -1], 0, 0], [1.0, 3.0])


@pytest.mark.parametrize("n,ofac", [(512, 2), (256, 2), (129, 1)])
def test_radial_golden_angle_and_extent(n, ofac):
    fov = 0.24
    (kx, ky, kz), t = trajectories.radial(fov, n, 13, oversampling=ofac)
    kx, ky = kx[..., 0], ky[..., 0]
    assert not np.any(kz)
    r = np.hypot(kx, ky)
    res0 = n // ofac
    kmax = ((n - 1 - n // 2) + (res0 % 2)) / (ofac * fov)
    np.testing.assert_allclose(r[:, 0], np.linspace(0, kmax, n), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(r, np.broadcast_to(r[:, :1], r.shape), rtol=1e-12)
    # spoke angle: k = kr (cos theta, -sin theta), theta_n = n * 2 pi / phi^2 (mod 2 pi),
    # the golden angle 137.5078 deg
    theta = np.mod(-np.arctan2(ky[-1], kx[-1]), 2 * np.pi)
    golden = 2 * np.pi / ((1 + np.sqrt(5)) / 2) ** 2
    assert abs(np.rad2deg(golden) - 137.50776) < 1e-4
    want = np.mod(np.arange(13) * golden, 2 * np.pi)
    d = np.angle(np.exp(1j * (theta - want)))
    assert np.abs(d).max() < 1e-9


def test_spiral_interleaves_are_rotations():
    (kx, ky, _), t = trajectories.spiral(0.24, 256, 1024, 8, density_exponent=2.0)
    K = kx[..., 0] + 1j * ky[..., 0]
    for m in range(1, 8):
        np.testing.assert_allclose(K[:, m], K[:, 0] * np.exp(2j * np.pi * m / 8), rtol=1e-12, atol=1e-9)
    # variable density: radius kmax u^2 reaches 0.5 res / FOV at the last sample
    np.testing.assert_allclose(np.abs(K[-1, 0]), 0.5 * 256 / 0.24, rtol=1e-12)
    u = (len(K) // 2) / (len(K) - 1)
    np.testing.assert_allclose(np.abs(K[len(K) // 2, 0]) / np.abs(K[-1, 0]), u ** 2, rtol=1e-9)


def _spiral_p1_law(n0, turns, kmax, g_max, slew, n=256):
    """T of the p = 1 spiral K(u) = kmax u exp(2 pi i turns u) through
    spiral_time_law on n points."""
    u = np.linspace(0.0, 1.0, n)
    K = kmax * u * np.exp(2j * np.pi * turns * u)
    return trajectories.spiral_time_law(u, K, g_max=g_max, slew=slew)[1]


def test_spiral_time_law_amplitude_limited_is_arc_length():
    """With the slew limit out of reach the time is the arc length of K over
    gamma G: kmax / (gamma G) * (a sqrt(1 + a^2) + asinh a) / (2 a), a = 2 pi
    turns, gamma = 2 pi gammabar, G = 0.95 g_max."""
    kmax, turns = 500.0, 6.0
    T = _spiral_p1_law(256, turns, kmax, g_max=33.0, slew=1e9)
    a = 2 * np.pi * turns
    arc = kmax * (a * np.sqrt(1 + a * a) + np.arcsinh(a)) / (2 * a)
    want = arc / (2 * np.pi * trajectories.GAMMABAR * 33e-3 * 0.95)
    np.testing.assert_allclose(T, want, rtol=5e-3)


def test_spiral_time_law_slew_limited_matches_quadrature():
    """With the amplitude limit out of reach the time is the integral of
    sqrt(|K''(u)| / (gamma S)), |K''| = kmax a sqrt((a u)^2 + 4) for p = 1."""
    from scipy.integrate import quad

    kmax, turns = 500.0, 6.0
    T = _spiral_p1_law(256, turns, kmax, g_max=1e9, slew=180.0)
    a = 2 * np.pi * turns
    gS = 2 * np.pi * trajectories.GAMMABAR * 180.0 * 0.95
    want = quad(lambda u: np.sqrt(kmax * a * np.sqrt((a * u) ** 2 + 4) / gS), 0.0, 1.0)[0]
    np.testing.assert_allclose(T, want, rtol=1e-2)


def test_spiral_time_law_respects_both_limits():
    """On the default scanner the discrete gradient |dK/dt| / gamma of every
    interval stays under 0.95 G_max."""
    u = np.linspace(0.0, 1.0, 200)
    K = 400.0 * u ** 2 * np.exp(2j * np.pi * 6.0 * u ** 0.5)
    t, T = trajectories.spiral_time_law(u, K)
    assert np.all(np.diff(t) > 0) and t[0] == 0.0 and T == t[-1]
    G = np.abs(np.diff(K)) / np.diff(t) / (2 * np.pi * trajectories.GAMMABAR)
    assert G.max() <= 33e-3 * 0.95 * 1.01


@pytest.mark.parametrize("res,p,bw", [((64, 8, 1), 2.0, 128e3), ((100, 12, 3), 1.0, 256e3), ((16, 4, 2), 1.5, 64e3)])
def test_spiral_stack_adc_grid_ends_and_planes(res, p, bw):
    fov = (0.24, 0.22, 0.08)
    (kx, ky, kz), t, T_ms = trajectories.spiral_stack(fov=fov, res=res, oversampling=2, t_start=1.5,
                                                      receiver_bw=bw, density_exponent=p, dtype=np.float64)
    n0, n_il, n_sl = res
    n = 2 * int(np.round(T_ms * 1e-3 * bw))
    assert kx.shape == ky.shape == kz.shape == t.shape == (n, n_il, n_sl)
    # ADC times: uniform over [0, T], the same for every interleave and plane
    np.testing.assert_allclose(t[:, 0, 0], 1.5 + np.linspace(0.0, T_ms, n), rtol=0, atol=1e-12)
    assert np.all(t == t[:, :1, :1])
    # ends: the centre, then kmax at 2 pi turns, turns = min(max(1, res/2), max(1, res/32))
    K = kx[:, :, 0] + 1j * ky[:, :, 0]
    kmax = 0.5 * n0 / fov[0]
    turns = min(max(1.0, 0.5 * n0), max(1.0, n0 / 32))
    np.testing.assert_allclose(K[0], 0.0, atol=1e-9)
    np.testing.assert_allclose(K[-1, 0], kmax * np.exp(2j * np.pi * turns), rtol=1e-9, atol=1e-9 * kmax)
    # interleaves are rotations by 2 pi m / M; radius kmax u^p along the base spiral
    for m in range(n_il):
        np.testing.assert_allclose(K[:, m], K[:, 0] * np.exp(2j * np.pi * m / n_il), atol=1e-9 * kmax)
    # kz planes at (s - n // 2) / FOV_z
    np.testing.assert_allclose(kz[0, 0, :], (np.arange(n_sl) - n_sl // 2) / fov[2], rtol=1e-12)
    assert np.all(kz == kz[:1, :1, :])


def test_hard_pulse_on_closed_interval_adds_the_preceding_step():
    """A hard pulse that starts after its block's first event is on at its
    start knot, so the step that ends there carries it too: gamma * sum
    |rf| dt over the steps is the flip times (dur + that step) / dur."""
    flip = np.deg2rad(20.0)
    p = sequences.HardPulse(start=0.3, dur=0.2, flip=flip)
    b = sequences.make_block(0.0, 1.0, pulses=[p], dt_rf=0.01, dt=0.25)
    k = int(np.flatnonzero(np.abs(b.times - 0.3) < 1e-12)[0])
    assert b.times[k - 1] == 0.25 and b.rf[k] == p.b1 and b.rf[k - 1] == 0
    area = sequences.GAMMA * np.sum(np.abs(b.rf[1:]) * b.dt[1:])
    np.testing.assert_allclose(area, flip * (0.2 + 0.05) / 0.2, rtol=1e-12)


def test_rf_pulse_shapes():
    """The reference's RF: gamma times the trapezoid area of the table is
    the flip angle; the sinc's zero crossings sit on its ends (bw (n_l + 1)
    lobes from the centre), the main lobe peaks at start + (n_l + 1) / bw;
    the hard shape equals HardPulse; a phase offset rotates the table."""
    flip = np.deg2rad(30.0)
    for shape in ("apodized_sinc", "sinc", "hard"):
        p = sequences.RFPulse(start=1.0, dur=2.0, flip=flip, shape=shape, n_lobes=(2, 3))
        tt = np.linspace(1.0, 3.0, 200)
        v = p(tt)
        area = np.sum((v[1:] + v[:-1]) / 2) * (tt[1] - tt[0]) * sequences.GAMMA
        np.testing.assert_allclose(abs(area), flip, rtol=1e-12)
        assert p(np.array([0.999, 3.001])).tolist() == [0, 0]
        if shape != "hard":
            bw = 7 / 2.0
            np.testing.assert_allclose(abs(v[0]), 0.0, atol=1e-12 * abs(v).max())
            np.testing.assert_allclose(abs(p(np.array([3.0]))[0]), 0.0, atol=1e-12 * abs(v).max())
            peak = 1.0 + 3 / bw
            fine = np.linspace(1.0, 3.0, 20001)
            assert abs(fine[np.argmax(np.abs(p(fine)))] - peak) < 2.0 / 199
    hard = sequences.RFPulse(start=1.0, dur=2.0, flip=flip, shape="hard", phase=0.3)
    hp = sequences.HardPulse(start=1.0, dur=2.0, flip=flip, phase=0.3)
    t = np.array([0.5, 1.0, 1.7, 3.0, 3.5])
    np.testing.assert_allclose(hard(t), hp(t), rtol=1e-15, atol=0)
    s0 = sequences.RFPulse(start=0.0, dur=1.0, flip=flip)
    s1 = sequences.RFPulse(start=0.0, dur=1.0, flip=flip, phase=0.7)
    np.testing.assert_allclose(s1(t), s0(t) * np.exp(0.7j), rtol=1e-12, atol=1e-15)



def test_sinc_pulse_excitation_through_the_oracle():
    """An on-resonance node (no gradient, no off-resonance, no relaxation)
    under a real sinc pulse rotates about one axis: Mz = cos(gamma sum rf dt)
    exactly over the block's steps, and that sum is the nominal flip to the
    event grid's quadrature error."""
    import oracle

    flip = np.deg2rad(40.0)
    p = sequences.RFPulse(start=0.0, dur=2.0, flip=flip, n_lobes=(2, 2))
    b = sequences.make_block(0.0, 2.0, pulses=[p], dt_rf=0.005, dt=1.0)
    ang = sequences.GAMMA * np.sum(b.rf[1:].real * b.dt[1:])
    np.testing.assert_allclose(ang, flip, rtol=1e-3)
    r0 = np.zeros((1, 3))
    mxy, mz = oracle.solve_mri(r0, np.array([1e12]), np.array([1e12]), np.zeros(1), 1.0, sequences.GAMMA, b.rf, b.G,
                               b.dt, None, np.zeros(1, complex), np.ones(1), None)
    np.testing.assert_allclose(mz[0, -1], np.cos(ang), rtol=0, atol=1e-12)
    np.testing.assert_allclose(abs(mxy[0, -1]), np.sin(ang), rtol=0, atol=1e-12)
