"""Writes the golden vectors of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Each case is a small seeded input of one hot-path call with the oracle's
output: Signal (Cartesian, radial, spiral; static and with linear motion),
solve_mri (hard pulse with gradients, full history), the P1 lumped and
consistent masses, and the POD contraction.  The fixtures pin the oracle
against silent change (tests/test_golden.py recomputes them) and give the
GPU tests a reference that does not need the oracle (tests/test_gpu_golden.py).
Stored as .npz (numpy arrays, no pickles).
"""
from pathlib import Path
import sys

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402
from feelmri_amd import phantoms, sequences, trajectories  # noqa: E402


def _nodes(n, v, seed):
    rng = np.random.default_rng(seed)
    return dict(r0=rng.uniform(-0.06, 0.06, (n, 3)), phi=rng.normal(0, 2 * np.pi * 20 / 1000, n),
                T2=rng.uniform(40, 100, n), mxy=(rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))) * 0.1,
                w=rng.uniform(0.5, 1.5, n) * 1e-6)


def signal_cases():
    import scipy.sparse as sp

    cases = {}
    specs = {
        "signal_cartesian": (trajectories.cartesian((0.24, 0.24, 0.24), (12, 10, 3), t_first=1.2), 700, 2, 1),
        "signal_radial": (trajectories.radial(0.24, 48, 20, t_first=0.3), 600, 1, 2),
        "signal_spiral": (trajectories.spiral(0.24, 32, 40, 32, t_first=0.2), 500, 1, 3),
    }
    for name, ((kloc, t), n, v, seed) in specs.items():
        nd = _nodes(n, v, seed)
        out = oracle.Signal(0, sp.diags(nd["w"]), kloc, t, nd["r0"], nd["phi"], nd["T2"], nd["mxy"], None)
        cases[name] = dict(kx=kloc[0], ky=kloc[1], kz=kloc[2], t=t, **nd, out=out)
    # linear motion x(t) = x0 + vel * t, evaluated at each sample's t
    (kloc, t), n, v = trajectories.cartesian((0.24, 0.24, 0.24), (10, 8, 2), t_first=0.5), 400, 1
    nd = _nodes(n, v, 4)
    vel = np.array([1e-3, -2e-3, 0.5e-3])

    def motion(tt):
        return np.broadcast_to(vel * tt, (n, 3))

    out = oracle.Signal(0, sp.diags(nd["w"]), kloc, t, nd["r0"], nd["phi"], nd["T2"], nd["mxy"], motion)
    cases["signal_linear_motion"] = dict(kx=kloc[0], ky=kloc[1], kz=kloc[2], t=t, **nd, vel=vel, out=out)
    return cases


def bloch_case():
    rng = np.random.default_rng(7)
    n = 300
    blk = sequences.make_block(0.0, 1.5, pulses=[sequences.HardPulse(0.1, 0.2, np.deg2rad(30), 0.4)],
                               gradients=[sequences.Trapezoid(0, 0.4, 8.0, 0.1, 0.3),
                                          sequences.Trapezoid(2, 0.5, -5.0, 0.05, 0.2)],
                               dt_rf=0.02, dt_gr=0.05, dt=0.1)
    r0 = rng.uniform(-0.05, 0.05, (n, 3))
    T1, T2 = rng.uniform(500, 1500, n), rng.uniform(40, 100, n)
    dB = rng.normal(0, 1e-4, n)
    mxy0 = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.05
    mz0 = rng.uniform(0.5, 1.0, n)
    mxy, mz = oracle.solve_mri(r0, T1, T2, dB, 1.0, sequences.GAMMA, blk.rf, blk.G, blk.dt, None, mxy0, mz0, None)
    return dict(r0=r0, T1=T1, T2=T2, dB=dB, mxy0=mxy0, mz0=mz0, rf=blk.rf, G=blk.G, dt=blk.dt,
                gamma=np.float64(sequences.GAMMA), mxy=mxy, mz=mz)


def mass_case():
    m = phantoms.ellipsoid_mesh(n=9)
    return dict(elems=m.elems, nodes=m.nodes, lumped=oracle.mass_p1_lumped(m.elems, m.nodes),
                element_matrices=oracle.mass_p1_element_matrices(m.elems, m.nodes))


def sinc_block():
    """Slice-selective excitation: an apodized sinc on a slice gradient, then
    the refocusing lobe (half the slice gradient's area, negative)."""
    rf = sequences.RFPulse(0.1, 2.0, np.deg2rad(60), n_lobes=(2, 2), phase=0.2)
    gs = sequences.Trapezoid(2, 0.0, 6.0, 0.1, 2.0)
    gr = sequences.Trapezoid(2, 2.2, -6.0, 0.1, 0.95)
    return sequences.make_block(0.0, 3.5, pulses=[rf], gradients=[gs, gr], dt_rf=0.02, dt_gr=0.05, dt=0.5)


def bloch_sinc_case():
    rng = np.random.default_rng(8)
    n = 200
    blk = sinc_block()
    r0 = np.zeros((n, 3))
    r0[:, 2] = np.linspace(-0.01, 0.01, n)
    r0[:, :2] = rng.uniform(-0.02, 0.02, (n, 2))
    T1, T2 = rng.uniform(500, 1500, n), rng.uniform(40, 100, n)
    dB = rng.normal(0, 5e-5, n)
    mxy0 = np.zeros(n, complex)
    mz0 = np.ones(n)
    mxy, mz = oracle.solve_mri(r0, T1, T2, dB, 1.0, sequences.GAMMA, blk.rf, blk.G, blk.dt, None, mxy0, mz0, None)
    return dict(r0=r0, T1=T1, T2=T2, dB=dB, mxy0=mxy0, mz0=mz0, rf=blk.rf, G=blk.G, dt=blk.dt,
                gamma=np.float64(sequences.GAMMA), mxy=mxy, mz=mz)


def pod_case():
    rng = np.random.default_rng(11)
    modes = rng.normal(size=(500, 3, 6))
    w = rng.normal(size=6)
    return dict(modes=modes, weights=w, out=oracle.tensordot_modes_weights(modes, w))


def main():
    cases = dict(signal_cases())
    cases["bloch_hard_pulse_gradients"] = bloch_case()
    cases["bloch_sinc_slice"] = bloch_sinc_case()
    cases["mass_p1_ellipsoid"] = mass_case()
    cases["pod_tensordot"] = pod_case()
    for name, arrays in cases.items():
        np.savez_compressed(HERE / f"{name}.npz", **{k: np.asarray(v) for k, v in arrays.items()})
        print(name, {k: np.asarray(v).shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
