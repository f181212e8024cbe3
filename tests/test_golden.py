"""The oracle against its committed golden vectors (tests/golden/*.npz,
written by tests/golden/make_golden.py): a change to oracle.c, its build
flags or the host producers that moves any output shows here."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

GOLDEN = Path(__file__).resolve().parent / "golden"
SIGNALS = ["signal_cartesian", "signal_radial", "signal_spiral", "signal_linear_motion"]


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("name", SIGNALS)
def test_oracle_signal_reproduces_golden(oracle, name):
    g = load(name)
    motion = None
    if "vel" in g:
        vel, n = g["vel"], g["r0"].shape[0]
        motion = lambda tt: np.broadcast_to(vel * tt, (n, 3))  # noqa: E731
    out = oracle.Signal(0, sp.diags(g["w"]), (g["kx"], g["ky"], g["kz"]), g["t"], g["r0"], g["phi"], g["T2"],
                        g["mxy"], motion)
    assert out.shape == g["out"].shape
    assert rel(out, g["out"]) <= 1e-13


def test_oracle_bloch_reproduces_golden(oracle):
    g = load("bloch_hard_pulse_gradients")
    mxy, mz = oracle.solve_mri(g["r0"], g["T1"], g["T2"], g["dB"], 1.0, float(g["gamma"]), g["rf"], g["G"], g["dt"],
                               None, g["mxy0"], g["mz0"], None)
    assert rel(mxy, g["mxy"]) <= 1e-13 and rel(mz, g["mz"]) <= 1e-13


def test_oracle_bloch_sinc_slice_reproduces_golden(oracle):
    """A slice-selective apodized-sinc excitation: the oracle reproduces the
    fixture; the fixture's tables are what the host producers build today;
    the slice centre ends near a 60 deg tip and the nodes outside the slice
    (|z| > 8.5 mm for a 3 kHz pulse on 6 mT/m: 11.7 mm thick) stay near
    the z axis."""
    from feelmri_amd import sequences

    g = load("bloch_sinc_slice")
    mxy, mz = oracle.solve_mri(g["r0"], g["T1"], g["T2"], g["dB"], 1.0, float(g["gamma"]), g["rf"], g["G"], g["dt"],
                               None, g["mxy0"], g["mz0"], None)
    assert rel(mxy, g["mxy"]) <= 1e-13 and rel(mz, g["mz"]) <= 1e-13
    blk = sequences.make_block(0.0, 3.5, pulses=[sequences.RFPulse(0.1, 2.0, np.deg2rad(60), n_lobes=(2, 2), phase=0.2)],
                               gradients=[sequences.Trapezoid(2, 0.0, 6.0, 0.1, 2.0),
                                          sequences.Trapezoid(2, 2.2, -6.0, 0.1, 0.95)],
                               dt_rf=0.02, dt_gr=0.05, dt=0.5)
    assert np.array_equal(blk.rf, g["rf"]) and np.array_equal(blk.G, g["G"]) and np.array_equal(blk.dt, g["dt"])
    z = g["r0"][:, 2]
    centre = np.abs(z) < 1e-3
    # centre: Mz at cos 60 deg; |Mxy| at sin 60 deg less ~5 % of T2 decay and
    # imperfect refocusing over the 3.5 ms block
    assert np.all(np.abs(g["mz"][centre, -1] - 0.5) < 0.05)
    assert np.all(np.abs(np.abs(g["mxy"][centre, -1]) - np.sin(np.deg2rad(60))) < 0.08)
    assert np.all(np.abs(g["mxy"][np.abs(z) > 8.5e-3, -1]) < 0.1)


def test_oracle_mass_and_pod_reproduce_golden(oracle):
    g = load("mass_p1_ellipsoid")
    assert rel(oracle.mass_p1_lumped(g["elems"], g["nodes"]), g["lumped"]) <= 1e-15
    assert rel(oracle.mass_p1_element_matrices(g["elems"], g["nodes"]), g["element_matrices"]) <= 1e-15
    p = load("pod_tensordot")
    assert rel(oracle.tensordot_modes_weights(p["modes"], p["weights"]), p["out"]) <= 1e-15


def test_golden_inputs_follow_the_host_producers():
    """The trajectories in the fixtures are what the producers make today."""
    from feelmri_amd import trajectories

    g = load("signal_spiral")
    kloc, t = trajectories.spiral(0.24, 32, 40, 32, t_first=0.2)
    assert np.array_equal(g["kx"], kloc[0]) and np.array_equal(g["t"], t)
    g = load("signal_radial")
    kloc, t = trajectories.radial(0.24, 48, 20, t_first=0.3)
    assert np.array_equal(g["ky"], kloc[1])
