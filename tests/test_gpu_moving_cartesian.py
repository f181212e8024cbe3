"""Moving nodes on Cartesian stacks: the per-readout GEMMs (mixed form on bf16
MFMA, FP64 form on FP64 MFMA) against the FP64 oracle and the direct kernels.

With POD motion the positions change with t (MRI.cpp:109-134); on a Cartesian
stack t depends on the readout index only, so each readout's samples share
one set of positions and the sum over nodes is a (lines x nodes) by
(nodes x kz*V) product that both forms run on the matrix cores with
in-kernel operands.

Shapes cover every workgroup configuration of the kernel (1-4 row blocks per
workgroup, 1-4 column blocks, several row and column tiles, partial blocks),
node counts around the 64 W-node record chunk and the 16-node step, one to
four encodings, translation-only motion, and readouts that share a t.
Tolerances: mixed <= 1e-4 (bf16 hi+lo products, FP32 accumulation), FP64
<= 1e-10.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import motion_fixtures
from feelmri_amd import Motion, trajectories

from test_gpu_signal import _nodes, _pod, descriptor_callable, rel_l2

pytestmark = pytest.mark.gpu

TOL_F64 = 1e-10
TOL_MIXED = 1e-4


def _plan_method(kloc, t):
    from feelmri_amd import device

    return device.analyse_trajectory(kloc, t, "auto")["method"]


def _run(kloc, t, n, v, motion, seed=3, oracle_motion=None, gpu=None):
    from feelmri_amd import MRI
    import oracle

    r0, phi, T2, mxy, w = _nodes(n, v, seed)
    M = sp.diags(w)
    ref = oracle.Signal(0, M, kloc, t, r0, phi, T2, mxy, oracle_motion or descriptor_callable(motion))
    gemm = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, motion, precision="mixed")
    direct = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, motion, precision="mixed", method="direct")
    f64 = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, motion, precision="f64")
    f64d = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, motion, precision="f64", method="direct")
    assert gemm.shape == ref.shape == (*np.asarray(t).shape, v)
    errs = {"gemm": rel_l2(gemm, ref), "direct": rel_l2(direct, ref), "f64_gemm": rel_l2(f64, ref),
            "f64_direct": rel_l2(f64d, ref)}
    print(f"n={n} v={v} shape={np.asarray(t).shape}:", {k: f"{e:.2e}" for k, e in errs.items()})
    assert errs["gemm"] <= TOL_MIXED, errs
    assert errs["direct"] <= TOL_MIXED, errs
    assert errs["f64_gemm"] <= TOL_F64, errs
    assert errs["f64_direct"] <= TOL_F64, errs
    return errs


@pytest.mark.parametrize("shape,v", [
    ((12, 6, 2), 1),      # 1 row block, 1 column block
    ((10, 40, 3), 4),     # 2 row blocks, KV = 12
    ((6, 96, 24), 4),     # the 4-D flow tile: 3 x 3 blocks
    ((5, 130, 5), 3),     # 5 row blocks: two row tiles of 4 waves
    ((4, 20, 70), 2),     # KV = 140: two column tiles of 4 blocks
    ((3, 33, 129), 1),    # V = 1: one kz row per column
])
def test_shapes(gpu, shape, v):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), shape, t_first=0.4)
    assert _plan_method(kloc, t) == "cartesian"
    n = 1500
    m = _pod(n)
    m.update_timeshift(2.0)
    _run(kloc, t, n, v, m)


@pytest.mark.parametrize("n", [1, 17, 191, 192, 193, 4097])
def test_ragged_node_counts(gpu, n):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (8, 96, 4), t_first=0.4)
    _run(kloc, t, n, 2, _pod(n, velocity=True, seed=n, family="ref" if n % 2 else "own"))


def test_translation_only(gpu):
    """A motion without modes (uniform translation): the R = 0 path."""
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (10, 32, 4), t_first=0.4)
    for family in ("own", "ref"):
        _run(kloc, t, 2000, 1, motion_fixtures.resp(direction=(0, 1, 0), family=family))


def test_readouts_sharing_t(gpu):
    """Readouts with equal t share a t-group (one motion coefficient row)."""
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (8, 24, 3), t_first=0.4)
    t = np.asarray(t).copy()
    t[1::2] = t[0::2]
    assert _plan_method(kloc, t) == "cartesian"
    n = 1200
    _run(kloc, t, n, 3, _pod(n, seed=5))


def test_large_mesh_against_f64(gpu):
    """300 000 nodes (many node splits) on the 4-D flow tile: mixed GEMM and
    FP64 GEMM against the FP64 direct kernels, which the tests above pin to
    the oracle."""
    from feelmri_amd import MRI

    n = 300_000
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (4, 96, 24), t_first=0.4)
    r0, phi, T2, mxy, w = _nodes(n, 4, 9)
    m = _pod(n, seed=9)
    m.update_timeshift(1.0)
    M = sp.diags(w)
    got = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, m, precision="mixed")
    g64 = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, m, precision="f64")
    ref = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, m, precision="f64", method="direct")
    err, err64 = rel_l2(got, ref), rel_l2(g64, ref)
    print(f"300k nodes: mixed GEMM vs FP64 relL2 {err:.2e}, FP64 GEMM vs FP64 direct {err64:.2e}")
    assert err <= TOL_MIXED
    assert err64 <= TOL_F64


def test_lines_out_of_order(gpu):
    """Phase-encode lines in shuffled order (still an additive Cartesian
    split, with the line coordinates in arbitrary order)."""
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (6, 40, 3), t_first=0.4)
    perm = np.random.default_rng(4).permutation(40)
    kloc = tuple(np.ascontiguousarray(np.asarray(k)[:, perm, :]) for k in kloc)
    assert _plan_method(kloc, t) == "cartesian"
    n = 1500
    _run(kloc, t, n, 2, _pod(n, seed=8))


def test_workspace_per_path(gpu):
    """The workspace query follows the path execute takes: static nodes on a
    Cartesian plan do not pay for the moving GEMM's partial tiles, and each
    path runs with exactly the bytes reported."""
    import torch

    from feelmri_amd import Motion, device

    n = 40_000
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (6, 64, 8), t_first=0.4)
    plan = device.SignalPlan(kloc, t, method="auto")
    assert plan.method == "cartesian"
    r0, phi, T2, mxy, w = _nodes(n, 2, 21)
    nodes = device.SignalNodes.from_arrays(r0, phi, T2, mxy, weights=w)
    m = _pod(n, seed=21)
    modes = device.modes_soa(torch.from_numpy(Motion.motion_modes(m)).cuda())
    coef = torch.from_numpy(Motion.motion_coefficients(m, plan.tgroup_t)).cuda()
    ws_static = plan.workspace_bytes("mixed", n, 2)
    ws_moving = plan.workspace_bytes("mixed", n, 2, moving=True)
    assert ws_static != ws_moving
    for prec in ("mixed", "f64"):
        for motion in (None, (modes, coef)):
            need = plan.workspace_bytes(prec, n, 2, moving=motion is not None)
            ws = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
            out = plan.execute(nodes, precision=prec, motion=motion, workspace=ws)
            assert torch.isfinite(torch.view_as_real(out)).all()
