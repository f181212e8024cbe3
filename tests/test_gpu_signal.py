"""Signal parity on the GPU: HIP direct sum vs the FP64 oracle.

Tolerances (relative L2 over the whole output):
  FP64 form   <= 1e-10  (the bar is 1e-6; same arithmetic up to rounding)
  mixed form  <= 1e-4   (FP64 range reduction + FP32 sincos / accumulation)
"""
import numpy as np
import pytest
import scipy.sparse as sp

import motion_fixtures
from feelmri_amd import Motion, trajectories

pytestmark = pytest.mark.gpu

TOL_F64 = 1e-10
TOL_MIXED = 1e-4


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _nodes(n, v=1, seed=0):
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(-0.06, 0.06, (n, 3))
    phi = rng.normal(0, 2 * np.pi * 20 / 1000, n)
    T2 = rng.uniform(40, 100, n)
    mxy = (rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))) * 0.1
    w = rng.uniform(0.5, 1.5, n) * 1e-6
    return r0, phi, T2, mxy, w


def descriptor_callable(obj):
    """t -> displacement formed in FP64 from the device tables of a motion
    object (float32 modes and coefficients, exactly what the kernels read)."""
    modes = Motion.motion_modes(obj).astype(np.float64)
    R = Motion.n_modes(obj)

    def f(t):
        c = Motion.motion_coefficients(obj, [t])[0]
        d = np.tensordot(modes, c[:R], axes=([2], [0])) if R else 0.0
        return d + c[R:]

    return f


def _check(kloc, t, n=3000, v=1, seed=0, M="diag", motion=None, precisions=("f64", "mixed"), gpu=None,
           oracle_motion=None, method="auto"):
    from feelmri_amd import MRI
    import oracle

    r0, phi, T2, mxy, w = _nodes(n, v, seed)
    if M == "diag":
        Mm = sp.diags(w)
    elif M == "csr":
        Mm = (sp.random(n, n, density=3.0 / max(n, 1), random_state=seed) + sp.diags(w * 1e6)) * 1e-6
        Mm = sp.csc_matrix(Mm)
    else:
        Mm = None
    ref = oracle.Signal(0, Mm, kloc, t, r0, phi, T2, mxy, oracle_motion if oracle_motion else motion)
    errs = {}
    for prec in precisions:
        got = MRI.Signal(0, Mm, kloc, t, r0, phi, T2, mxy, motion, precision=prec, method=method)
        assert got.shape == ref.shape
        errs[prec] = rel_l2(got, ref)
    print("relL2", {k: f"{e:.2e}" for k, e in errs.items()})
    if "f64" in errs:
        assert errs["f64"] <= TOL_F64, errs
    if "mixed" in errs:
        assert errs["mixed"] <= TOL_MIXED, errs
    return errs


def test_cartesian_small_diag(gpu):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (16, 12, 4), t_first=1.2, t_readout=2.0)
    _check(kloc, t, n=3000, v=1)


@pytest.mark.parametrize("v", [2, 3, 5])
def test_multiple_encodings(gpu, v):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (12, 8, 2), t_first=0.5)
    _check(kloc, t, n=1500, v=v, seed=v)


def test_consistent_mass_csr(gpu):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (10, 10, 1))
    _check(kloc, t, n=2000, M="csr")


def test_identity_mass(gpu):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (10, 10, 1))
    _check(kloc, t, n=700, M=None)


def test_radial(gpu):
    kloc, t = trajectories.radial(0.24, 128, 16, t_first=0.3)
    _check(kloc, t, n=5000, seed=3)


# Readout runs: the mixed direct form advances exp(-2 pi i k.x) along each
# readout by one complex factor per sample when k and t are linear in the
# readout index (Cartesian lines, radial spokes); readouts that are not a
# multiple of the run length end in a short run.
@pytest.mark.parametrize("shape,v", [((21, 6, 3), 3), ((16, 5, 1), 1), ((40, 3, 2), 2)])
def test_direct_readout_runs_cartesian(gpu, shape, v):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), shape, t_first=0.7, t_readout=2.5)
    _check(kloc, t, n=2500, v=v, seed=sum(shape), method="direct")


def test_readout_run_detection(gpu):
    from feelmri_amd import device

    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (21, 6, 3), t_first=0.7)
    p = device.SignalPlan(kloc, t, method="direct")
    assert p.n_runs == 6 * 3 * 2 and p.run_len == 16   # ceil(21 / 16) runs per line
    kloc, t = trajectories.radial(0.24, 100, 7, t_first=0.3)
    assert device.SignalPlan(kloc, t).n_runs == 7 * 7
    kloc, t = trajectories.spiral(0.24, 64, 200, 8, t_first=0.2)
    assert device.SignalPlan(kloc, t).n_runs == 0     # k not linear along the readout
    rng = np.random.default_rng(5)
    kl = tuple(rng.uniform(-200, 200, (30, 7, 2)) for _ in range(3))
    assert device.SignalPlan(kl, rng.uniform(0.0, 4.0, (30, 7, 2))).n_runs == 0


def test_direct_readout_runs_radial_short_last_run(gpu):
    kloc, t = trajectories.radial(0.24, 100, 7, t_first=0.3)
    _check(kloc, t, n=6000, seed=11)


def test_spiral_small_groups_general_kernel(gpu):
    kloc, t = trajectories.spiral(0.24, 64, 200, 8, t_first=0.2)
    _check(kloc, t, n=4000, seed=4)


@pytest.mark.parametrize("n_il,v", [(32, 1), (64, 3), (128, 2)])
def test_spiral_packed_t_groups(gpu, n_il, v):
    """Interleaves sharing t per readout index: 256 / n_il equal-size groups
    per packed slot; 90 readout indices leave a partial last slot."""
    from feelmri_amd.device import analyse_trajectory

    kloc, t = trajectories.spiral(0.24, 64, 90, n_il, t_first=0.2)
    assert analyse_trajectory(kloc, t)["group_pack"] == 256 // n_il
    _check(kloc, t, n=3001, v=v, seed=n_il)


def _irregular_interleaves(n_il=32, seed=3):
    """Spiral interleaves at irregular angles: equal-size t-groups that pack,
    with no sample at -k of another."""
    kloc, t = trajectories.spiral(0.24, 64, 90, 1, t_first=0.2)
    base = kloc[0][:, 0, 0] + 1j * kloc[1][:, 0, 0]
    ang = 2 * np.pi * np.sort(np.random.default_rng(seed).uniform(size=n_il))
    K = base[:, None] * np.exp(1j * ang)[None, :]
    kx, ky = np.ascontiguousarray(K.real[..., None]), np.ascontiguousarray(K.imag[..., None])
    return (kx, ky, np.zeros_like(kx)), np.ascontiguousarray(np.broadcast_to(t[:, :1, :], kx.shape))


@pytest.mark.parametrize("kind", ["conj", "pair"])
@pytest.mark.parametrize("far", [False, True])
def test_spiral_packed_mixed_paths(gpu, kind, far):
    """The mixed form's two packed paths.  ``conj``: interleaves half a turn
    apart pair up as k and -k at equal t, so one phase and one sine/cosine
    serve both samples of a lane.  ``pair``: irregular interleave angles; the
    second sample of a lane (same interleave, 8 readout indices on) is the
    first one's FP64-reduced phase plus dk.x in FP32, for 256-node chunks
    whose nodes all have |dk|_1 max|x| <= 4 cycles.  With ``far``, nodes
    700..1299 sit 25x further out (up to 1.5 m): pair chunks take the FP64
    phase for both samples.  Either way the error stays far inside the mixed
    bar: <= 2e-6."""
    from feelmri_amd import MRI
    from feelmri_amd.device import analyse_trajectory
    import oracle

    if kind == "conj":
        kloc, t = trajectories.spiral(0.24, 64, 90, 32, t_first=0.2)
    else:
        kloc, t = _irregular_interleaves()
    a = analyse_trajectory(kloc, t)
    assert a["group_pack"] == 8 and a["conj_pack"] == (16 if kind == "conj" else 1)
    r0, phi, T2, mxy, w = _nodes(3001, 1, 21)
    if far:
        r0[700:1300] *= 25.0
    ref = oracle.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy, None)
    errs = {}
    for prec in ("mixed", "f64"):
        got = MRI.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy, None, precision=prec)
        errs[prec] = rel_l2(got, ref)
    print("relL2", {k: f"{e:.2e}" for k, e in errs.items()})
    assert errs["mixed"] <= 2e-6 and errs["f64"] <= TOL_F64


def test_random_times_per_sample(gpu):
    rng = np.random.default_rng(5)
    kloc = tuple(rng.uniform(-200, 200, (30, 7, 2)) for _ in range(3))
    t = rng.uniform(0.0, 4.0, (30, 7, 2))
    _check(kloc, t, n=2600, v=2, seed=5)


@pytest.mark.parametrize("n", [1, 255, 257, 4099])
def test_ragged_node_counts(gpu, n):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (8, 6, 1))
    _check(kloc, t, n=n, seed=n)


def test_single_sample(gpu):
    kloc = tuple(np.array([[[v]]]) for v in (13.0, -7.0, 2.0))
    t = np.array([[[1.5]]])
    _check(kloc, t, n=1000)


def test_zero_nodes_gives_zeros(gpu):
    from feelmri_amd import MRI

    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (4, 4, 1))
    z = np.zeros((0, 3))
    out = MRI.Signal(0, None, kloc, t, z, np.zeros(0), np.zeros(0), np.zeros((0, 1), np.complex128))
    assert out.shape == (4, 4, 1, 1) and not np.any(out)


def test_float32_inputs_use_float_overload(gpu):
    from feelmri_amd import MRI
    import oracle

    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (8, 8, 1), dtype=np.float32)
    r0, phi, T2, mxy, w = _nodes(2000)
    args = (r0.astype(np.float32), phi.astype(np.float32), T2.astype(np.float32), mxy.astype(np.complex64))
    out = MRI.Signal(0, sp.diags(w.astype(np.float32)), kloc, t, *args)
    assert out.dtype == np.complex64
    ref = oracle.Signal(0, sp.diags(w.astype(np.float32)), kloc, t, *args)
    assert rel_l2(out, ref) <= TOL_MIXED


def _pod(n, velocity=False, seed=11, family="own"):
    return motion_fixtures.pod(n, seed=seed, velocity=velocity, family=family)


@pytest.mark.parametrize("kind", ["pod", "velocity", "resp", "sum", "callable", "ref-pod", "ref-velocity",
                                  "ref-sum"])
def test_moving_nodes(gpu, kind):
    """Own motion classes, reference-shaped objects (duck-typed) and a generic
    callable (sampled at the distinct t on the host)."""
    n = 2500
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (12, 6, 2), t_first=0.4)
    family = "ref" if kind.startswith("ref-") else "own"
    base = kind.split("-")[-1]
    resp = motion_fixtures.resp(direction=(0, 1, 0), family=family)
    if base == "pod":
        m = _pod(n, family=family)
    elif base == "velocity":
        m = _pod(n, velocity=True, family=family)
    elif base == "resp":
        m = resp
    elif base == "sum":
        m = resp + _pod(n, family=family)
    else:
        pod = _pod(n)
        m = lambda tt: pod(tt) * 1.5  # noqa: E731
    if hasattr(m, "update_timeshift"):
        m.update_timeshift(3.0)
    # kernel parity: the oracle moves the nodes with the same float32 tables
    exact = descriptor_callable(m) if kind != "callable" else None
    _check(kloc, t, n=n, motion=m, seed=12, oracle_motion=exact)
    if kind != "callable":
        # semantics: against the object's own __call__ (float32 trajectories,
        # so the FP64 bar is the float32 displacement resolution)
        from feelmri_amd import MRI
        import oracle

        r0, phi, T2, mxy, w = _nodes(n, 1, 12)
        ref = oracle.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy, m)
        got = MRI.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy, m, precision="f64")
        assert rel_l2(got, ref) <= 1e-5


def test_c0_slab_full_kspace(gpu):
    """61 206-node slab, 64 x 64 Cartesian, FOV 0.24 m, 2 ms readout."""
    from feelmri_amd import phantoms

    m = phantoms.slab_mesh()
    tis = phantoms.draw_tissue(m.n_nodes, seed=1000)
    import oracle

    w = oracle.mass_p1_lumped(m.elems, m.nodes)
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (64, 64, 1), t_first=1.25, t_readout=2.0)
    rng = np.random.default_rng(1000)
    mxy = (1j * np.sin(np.deg2rad(15)) * tis.M0 * rng.uniform(0.05, 0.2, m.n_nodes))[:, None]
    from feelmri_amd import MRI

    ref = oracle.Signal(0, sp.diags(w), kloc, t, m.nodes, tis.delta_omega, tis.T2star, mxy)
    errs = {}
    for prec in ("f64", "mixed"):
        got = MRI.Signal(0, sp.diags(w), kloc, t, m.nodes, tis.delta_omega, tis.T2star, mxy, precision=prec)
        errs[prec] = rel_l2(got, ref)
    print("c0 relL2", errs)
    assert errs["f64"] <= TOL_F64 and errs["mixed"] <= TOL_MIXED
