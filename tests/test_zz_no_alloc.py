"""fm_signal_execute_noalloc: the signal evaluation with its workspace in the
library's static device array, against fm_signal_execute with a caller
workspace -- bitwise the same output (same kernels, same launch shapes), on
two calls in a row, the second smaller, in the FP64 and mixed forms; and a
clear FM_ERR_WORKSPACE for an evaluation whose workspace exceeds the array.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(gpu, kloc, t, n, v, prec, seed):
    import torch

    from feelmri_amd import _lib, device

    rng = np.random.default_rng(seed)
    r0 = rng.uniform(-0.06, 0.06, (n, 3))
    phi = rng.normal(0, 0.1, n)
    T2 = rng.uniform(40, 100, n)
    mxy = (rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))) * 0.1
    w = rng.uniform(0.5, 1.5, n) * 1e-6
    plan = device.SignalPlan(kloc, t, device=gpu)
    nodes = device.SignalNodes.from_arrays(r0, phi, T2, mxy, weights=w, device=gpu)
    want = plan.execute(nodes, precision=prec)
    got = torch.full_like(want, complex(float("nan"), float("nan")))
    lib = _lib.load()
    code = _lib.PREC_F64 if prec == "f64" else _lib.PREC_MIXED
    rc = lib.fm_signal_execute_noalloc(plan._h, code, n, v, ctypes.c_void_p(nodes.packed.data_ptr()), None, 0,
                                       None, ctypes.c_void_p(got.data_ptr()),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.FM_OK, _lib.last_error()
    torch.cuda.synchronize()
    return plan, got, want


def test_noalloc_equals_execute_twice(gpu):
    from feelmri_amd import trajectories

    print("no-alloc call 1: 61206 nodes, Cartesian 64 x 64, mixed", flush=True)
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (64, 64, 1), t_first=1.25)
    _, got, want = _run(gpu, kloc, t, 61_206, 1, "mixed", 1)
    assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())
    print("no-alloc call 2: 20000 nodes, radial 128 x 32, FP64", flush=True)
    kloc, t = trajectories.radial(0.24, 128, 32, t_first=0.5)
    _, got, want = _run(gpu, kloc, t, 20_000, 2, "f64", 2)
    assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())
    print("no-alloc call 3: 30000 nodes, radial 512 x 16, mixed (per-line GEMM)", flush=True)
    kloc, t = trajectories.radial(0.24, 512, 16, t_first=0.3)
    plan, got, want = _run(gpu, kloc, t, 30_000, 1, "mixed", 4)
    assert plan.line_gemm
    assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())
    print("no-alloc call 4: 3000 nodes, Cartesian 16 x 8 x 4, FP64", flush=True)
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (16, 8, 4), t_first=1.0)
    _, got, want = _run(gpu, kloc, t, 3_000, 1, "f64", 3)
    assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())


def test_noalloc_rejects_oversize_workspace(gpu):
    import torch

    from feelmri_amd import _lib, device, trajectories

    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), (128, 128, 32), t_first=1.0)
    plan = device.SignalPlan(kloc, t, device=gpu)
    n = 512_000
    assert plan.workspace_bytes("mixed", n, 1) > 128 << 20
    out = torch.empty((plan.n_samples, 1), dtype=torch.complex128, device=gpu)
    packed = torch.zeros((7, n), dtype=torch.float64, device=gpu)
    rc = _lib.load().fm_signal_execute_noalloc(plan._h, _lib.PREC_MIXED, n, 1, ctypes.c_void_p(packed.data_ptr()),
                                               None, 0, None, ctypes.c_void_p(out.data_ptr()), None)
    assert rc == _lib.FM_ERR_WORKSPACE
    assert "scratch" in _lib.last_error()
