"""Cartesian factorised GEMM path vs the FP64 oracle and vs the direct sum.

FP64 (v_mfma_f64_16x16x4) relL2 <= 1e-10; mixed (bf16 hi+lo MFMA, FP32
accumulation) relL2 <= 1e-4.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from feelmri_amd import trajectories

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _nodes(n, v, seed):
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(-0.1, 0.1, (n, 3))
    phi = rng.normal(0, 2 * np.pi * 20 / 1000, n)
    T2 = 1.0 / (1.0 / rng.uniform(40, 100, n) + rng.uniform(10, 50, n) / 1000)
    mxy = (rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))) * 0.1
    w = rng.uniform(0.5, 1.5, n) * 1e-6
    return r0, phi, T2, mxy, w


def _run(kloc, t, n, v, seed, method="cartesian"):
    from feelmri_amd import MRI, device
    import oracle

    r0, phi, T2, mxy, w = _nodes(n, v, seed)
    plan = device.SignalPlan(kloc, t, method=method)
    assert plan.method == method
    ref = oracle.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy)
    got = {p: MRI.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy, precision=p, method=method)
           for p in ("f64", "mixed")}
    errs = {p: rel_l2(g, ref) for p, g in got.items()}
    print(kloc[0].shape, n, v, errs)
    assert errs["f64"] <= 1e-10, errs
    assert errs["mixed"] <= 1e-4, errs
    return got, ref


# Column tiles of 128 (lines x kz x encodings): the tiles whose distinct Ey
# and Ezb rows fit the LDS staging take the staged kernel, the others (many
# lines per tile with one kz, or more than 128 kz x encodings) the per-lane
# kernel; both kinds are covered below.  Shapes with 97 or more readout rows
# use 4 row blocks per wave: (128, 9, 1), (100, 20, 6) (ragged rows and
# columns), (256, 10, 4) (two row tiles) on the staged kernel, (128, 200, 1)
# on the per-lane kernel.
@pytest.mark.parametrize("res,n,v", [((16, 8, 2), 3000, 1), ((40, 12, 3), 2049, 2), ((128, 9, 1), 700, 3),
                                     ((8, 6, 1), 1, 1), ((33, 17, 5), 5000, 1), ((32, 16, 32), 4000, 1),
                                     ((16, 200, 1), 3000, 1), ((12, 3, 50), 2000, 3), ((100, 20, 6), 5000, 2),
                                     ((256, 10, 4), 3000, 1), ((128, 200, 1), 2000, 1)])
def test_gemm_shapes(gpu, res, n, v):
    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), res, t_first=1.0, t_readout=3.0)
    _run(kloc, t, n, v, seed=sum(res) + n)


def test_gemm_many_splits(gpu):
    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), (32, 16, 4), t_first=1.0, t_readout=3.0)
    _run(kloc, t, 60000, 1, seed=3)


def test_rotated_grid_uses_general_split(gpu):
    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), (24, 20, 3), t_first=1.0, t_readout=3.0)
    a, b = 0.3, -0.2
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    R = Rz @ Rx
    k = np.stack(kloc, -1) @ R.T
    kr = tuple(np.ascontiguousarray(k[..., c]) for c in range(3))
    _run(kr, t, 4000, 1, seed=4)


def test_gemm_matches_direct(gpu):
    from feelmri_amd import MRI

    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), (64, 32, 4), t_first=1.0, t_readout=3.0)
    r0, phi, T2, mxy, w = _nodes(20000, 2, 5)
    M = sp.diags(w)
    d = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, precision="f64", method="direct")
    c = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, precision="f64", method="cartesian")
    assert rel_l2(c, d) <= 1e-12
    cm = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, precision="mixed", method="cartesian")
    assert rel_l2(cm, d) <= 1e-4


def test_non_cartesian_falls_back_to_direct(gpu):
    from feelmri_amd import device

    kloc, t = trajectories.radial(0.24, 64, 8)
    assert device.SignalPlan(kloc, t, method="auto").method == "direct"
    with pytest.raises(Exception, match="factorise"):
        device.SignalPlan(kloc, t, method="cartesian")
    # t varying along lines breaks the factorisation too
    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), (8, 8, 2))
    t = t + np.arange(8)[None, :, None] * 0.1
    assert device.SignalPlan(kloc, t, method="auto").method == "direct"


def test_large_mesh_mixed_vs_f64(gpu):
    """3M nodes on a small Cartesian grid: the mixed GEMM keeps its FP32
    accumulation to bounded node splits (summed in FP64), so its error
    against the FP64 GEMM stays at the small-mesh level."""
    from feelmri_amd import device

    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), (32, 8, 4), t_first=1.0, t_readout=3.0)
    r0, phi, T2, mxy, w = _nodes(3_000_000, 1, seed=77)
    plan = device.SignalPlan(kloc, t, method="cartesian", device=gpu)
    nodes = device.SignalNodes.from_arrays(r0, phi, T2, mxy, weights=w, device=gpu)
    a = plan.execute(nodes, precision="mixed").cpu().numpy()
    b = plan.execute(nodes, precision="f64").cpu().numpy()
    err = rel_l2(a, b)
    print("3M nodes mixed vs f64", err)
    assert err <= 1e-4, err

