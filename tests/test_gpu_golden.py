"""The HIP path against the committed golden vectors (tests/golden/*.npz):
no oracle at run time.  Tolerances (relative L2): FP64 forms 1e-10 (1e-7
with a generic motion callable, tabulated in float32), mixed 1e-4, the
float32 Bloch history 1e-6, masses and POD 1e-12."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("name", ["signal_cartesian", "signal_radial", "signal_spiral", "signal_linear_motion"])
def test_signal_matches_golden(gpu, name):
    from feelmri_amd import MRI

    g = load(name)
    motion = None
    if "vel" in g:
        vel, n = g["vel"], g["r0"].shape[0]
        motion = lambda tt: np.broadcast_to(vel * tt, (n, 3))  # noqa: E731
    errs = {}
    for prec in ("f64", "mixed"):
        got = MRI.Signal(0, sp.diags(g["w"]), (g["kx"], g["ky"], g["kz"]), g["t"], g["r0"], g["phi"], g["T2"],
                         g["mxy"], motion, precision=prec)
        assert got.shape == g["out"].shape
        errs[prec] = rel(got, g["out"])
    print(name, errs)
    # a generic motion callable reaches the device as float32 displacement
    # tables (the precision of the reference's float32 motion objects): the
    # FP64 form then carries their rounding, 4e-9 on this fixture
    tol64 = 1e-7 if motion is not None else 1e-10
    assert errs["f64"] <= tol64 and errs["mixed"] <= 1e-4


def test_bloch_matches_golden(gpu):
    from feelmri_amd import BlochSimulator

    g = load("bloch_hard_pulse_gradients")
    mxy, mz = BlochSimulator.solve_mri(g["r0"], g["T1"], g["T2"], g["dB"], 1.0, float(g["gamma"]), g["rf"], g["G"],
                                       g["dt"], None, g["mxy0"], g["mz0"], None, precision="f64")
    # the drop-in returns the reference's complex64 / float32 history
    assert rel(mxy, g["mxy"]) <= 1e-6 and rel(mz, g["mz"]) <= 1e-6


def test_bloch_sinc_slice_matches_golden(gpu):
    """The slice-selective sinc excitation on the GPU: FP64 form (complex64 /
    float32 history) and the FP32 form against the fixture."""
    from feelmri_amd import BlochSimulator

    g = load("bloch_sinc_slice")
    args = (g["r0"], g["T1"], g["T2"], g["dB"], 1.0, float(g["gamma"]), g["rf"], g["G"], g["dt"], None, g["mxy0"],
            g["mz0"], None)
    mxy, mz = BlochSimulator.solve_mri(*args, precision="f64")
    assert rel(mxy, g["mxy"]) <= 1e-6 and rel(mz, g["mz"]) <= 1e-6
    mxy, mz = BlochSimulator.solve_mri(*args, precision="f32")
    assert rel(mxy, g["mxy"]) <= 1e-4 and rel(mz, g["mz"]) <= 1e-4


def test_mass_and_pod_match_golden(gpu):
    import torch

    from feelmri_amd import device

    g = load("mass_p1_ellipsoid")
    col = device.P1Colouring(g["elems"], g["nodes"].shape[0], device=gpu)
    w = device.lumped_mass(col, torch.from_numpy(g["nodes"]).to(gpu)).cpu().numpy()
    assert rel(w, g["lumped"]) <= 1e-12
    em = device.element_matrices(torch.from_numpy(g["elems"]).to(gpu), torch.from_numpy(g["nodes"]).to(gpu))
    assert rel(em.cpu().numpy(), g["element_matrices"]) <= 1e-12
    p = load("pod_tensordot")
    out = device.pod_tensordot(torch.from_numpy(p["modes"]).to(gpu), torch.from_numpy(p["weights"]).to(gpu))
    assert rel(out.cpu().numpy(), p["out"]) <= 1e-12
