"""Drop-in for the reference's ``feelmri.MRI`` extension module.

``Signal(MPI_rank, M, kloc, t, r0, phi_dB0, T2, Mxy, pod_trajectory)`` keeps
the reference's signature, argument meaning and return layout
(cpp/feelmri/MRI.h:43-82, MRI.cpp:38-297): it returns the k-space signal of
shape (n_meas, n_lines, n_kz, n_enc) as a new numpy array.

Overload resolution follows the reference's pybind11 registration order
(MRI.h:46-82: float/callable, float/None, double/callable, double/None):
all-float32 inputs pick the float overload, all-float64 inputs the double
one, and mixed dtypes fall through to the conversion pass, whose first
viable overload is float.  Here the float overload runs the mixed form (FP64
range reduction, FP32 sincos; complex64 output) and the double overload the
FP64 form (complex128 output); ``precision=`` overrides the choice.

``M`` is the reference's sparse mass matrix: a diagonal (lumped) matrix is
applied as node weights inside the pack kernel, any other matrix through the
device CSR SpMV (MRI.cpp:243).  ``pod_trajectory`` may be None, one of the
motion classes of ``feelmri_amd.Motion`` (or the reference's, duck-typed),
or any callable ``t -> (N, 3)``: the latter is sampled at the distinct
sample times, which is exact because the reference re-evaluates positions
only when t changes (MRI.cpp:109-134).
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np
import scipy.sparse as sp
import torch
import xxhash

from . import Motion, device, resident

_F32_KINDS = (np.float32, np.complex64)


def _is_f32(a) -> bool:
    return np.asarray(a).dtype in (np.dtype(np.float32), np.dtype(np.complex64))


def _is_f64(a) -> bool:
    return np.asarray(a).dtype in (np.dtype(np.float64), np.dtype(np.complex128))


def resolve_precision(kloc, t, r0, phi_dB0, T2, Mxy) -> str:
    dense = [*kloc, t, r0, phi_dB0, T2]
    f32 = all(_is_f32(a) for a in dense) and np.asarray(Mxy).dtype == np.complex64
    f64 = all(_is_f64(a) for a in dense) and np.asarray(Mxy).dtype == np.complex128
    if f64 and not f32:
        return "f64"
    return "mixed"


def _lumped_diagonal(M):
    """Return the diagonal if M is diagonal, else None."""
    if M is None:
        return None
    if isinstance(M, np.ndarray) and M.ndim == 1:
        return M.astype(np.float64)
    if not sp.issparse(M):
        return None
    if M.format == "dia":
        off = np.asarray(M.offsets)
        if np.all((off == 0) | ~np.any(M.data != 0, axis=1)):
            return np.asarray(M.diagonal(), dtype=np.float64)
        return None
    if M.format in ("csr", "csc"):
        # every stored entry of row (column) r sits in column (row) r
        counts = np.diff(M.indptr)
        if np.array_equal(M.indices, np.repeat(np.arange(counts.size, dtype=M.indices.dtype), counts)):
            return np.asarray(M.diagonal(), dtype=np.float64)
        return None
    Mc = sp.coo_matrix(M)
    if Mc.nnz == 0 or np.all(Mc.row == Mc.col):
        return np.asarray(M.diagonal(), dtype=np.float64)
    return None


def weighted_mxy(M, Mxy, dev, stream=None):
    """(M Mxy) on the device, or (Mxy, diag weights) when M is diagonal.  The
    matrix and the diagonal stay resident across calls (by content)."""
    Mxy = np.asarray(Mxy)
    if Mxy.ndim == 1:
        Mxy = Mxy[:, None]
    x = torch.from_numpy(np.ascontiguousarray(Mxy, dtype=np.complex128)).to(dev)
    if M is None:
        return x, None
    diag = _lumped_diagonal(M)
    if diag is not None:
        return x, resident.to_device(diag, torch.float64, dev)
    Mc = sp.csr_matrix(M, dtype=np.float64)
    Mc.sort_indices()
    y = device.csr_spmv(resident.to_device(Mc.indptr.astype(np.int64), torch.int64, dev),
                        resident.to_device(Mc.indices.astype(np.int64), torch.int64, dev),
                        resident.to_device(Mc.data, torch.float64, dev), x, stream=stream)
    return y, None


def motion_tables(pod_trajectory, n_nodes: int, times: np.ndarray, dev):
    """(modes_soa, coef) device tensors for a motion object with a modal form
    (``Motion.is_supported``) at ``times``."""
    if pod_trajectory is None:
        return None
    if not Motion.is_supported(pod_trajectory):
        raise TypeError("motion without a modal form: Signal samples it per distinct t instead")
    modes = Motion.motion_modes(pod_trajectory)
    R = modes.shape[2]
    coef = Motion.motion_coefficients(pod_trajectory, times)
    if R == 0:
        modes_soa = torch.zeros((0, n_nodes), dtype=torch.float32, device=dev)
    else:
        if modes.shape[0] != n_nodes:
            raise ValueError(f"motion modes have {modes.shape[0]} nodes, expected {n_nodes}")
        modes_soa = device.modes_soa(torch.from_numpy(modes).to(dev))
    return modes_soa, torch.from_numpy(coef).to(dev)


# Plans of recently used trajectories, keyed by device, method, shape and a
# digest of the sampling arrays: callers that evaluate the same trajectory
# repeatedly (per line, per cardiac phase, per encoding) analyse and upload it
# once.
_PLAN_CACHE: "OrderedDict[tuple, device.SignalPlan]" = OrderedDict()
_PLAN_CACHE_SIZE = 256


def _plan_for(kloc, t_arr, method: str, dev) -> "device.SignalPlan":
    arrs = [np.ascontiguousarray(np.asarray(k, dtype=np.float64)) for k in kloc]
    tt = np.ascontiguousarray(np.asarray(t_arr, dtype=np.float64))
    h = xxhash.xxh3_128()
    for a in (*arrs, tt):
        if a.shape != tt.shape:
            raise ValueError(f"kloc shape {a.shape} differs from t shape {tt.shape}")
        h.update(memoryview(a).cast("B"))
    key = (str(dev), method, tt.shape, h.digest())
    plan = _PLAN_CACHE.get(key)
    if plan is not None:
        _PLAN_CACHE.move_to_end(key)
        return plan
    plan = device.SignalPlan(arrs, tt, method=method, device=dev)
    _PLAN_CACHE[key] = plan
    while len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
        _PLAN_CACHE.popitem(last=False)
    return plan


def Signal(MPI_rank, M, kloc, t, r0, phi_dB0, T2, Mxy, pod_trajectory=None, *, precision=None,
           method: str = "auto", dev=None):
    """k-space signal (n_meas, n_lines, n_kz, n_enc); see the module docstring."""
    if len(kloc) != 3:
        raise TypeError("kloc must hold the three arrays (kx, ky, kz)")
    if pod_trajectory is not None and not callable(pod_trajectory):
        raise TypeError("pod_trajectory must be callable or None")
    prec = precision or resolve_precision(kloc, t, r0, phi_dB0, T2, Mxy)
    dev = torch.device(dev) if dev is not None else device.default_device()
    t_arr = np.asarray(t)
    if t_arr.ndim != 3:
        raise TypeError(f"t must be a 3-D array (n_meas, n_lines, n_kz), got {t_arr.shape}")
    r0 = np.asarray(r0).reshape(-1, 3)
    N = r0.shape[0]
    Mxy_arr = np.asarray(Mxy)
    V = 1 if Mxy_arr.ndim == 1 else Mxy_arr.shape[1]

    moving = pod_trajectory is not None
    generic = moving and not Motion.is_supported(pod_trajectory)
    # Supported motion classes keep the requested method: on a Cartesian plan
    # the mixed form runs the per-readout GEMM and the FP64 form the direct
    # kernels.  Generic callables are evaluated in passes over the distinct
    # times, which the direct kernels serve.
    plan = _plan_for(kloc, t_arr, "direct" if generic else method, dev)
    with torch.cuda.device(dev):
        mmxy, diag = weighted_mxy(M, Mxy_arr, dev)
        # positions, off-resonance and T2 are usually the same on every call
        # (one call per line or per phase): resident by content
        nodes = device.SignalNodes.from_arrays(resident.to_device(r0, torch.float64, dev),
                                               resident.to_device(np.asarray(phi_dB0).ravel(), torch.float64, dev),
                                               resident.to_device(np.asarray(T2).ravel(), torch.float64, dev),
                                               mmxy, weights=diag, dtype=torch.float64, device=dev)
        if generic:
            out = _signal_generic_motion(plan, nodes, prec, pod_trajectory, r0, dev)
        else:
            motion = motion_tables(pod_trajectory, N, plan.tgroup_t, dev) if moving else None
            if motion is not None and motion[0].shape[0] == 0:
                # translation only: keep the R = 0 path but pass a valid pointer
                motion = (torch.zeros((3, N), dtype=torch.float32, device=dev),
                          torch.cat([torch.zeros((plan.n_tgroups, 1), dtype=torch.float64, device=dev),
                                     motion[1]], dim=1))
            out = plan.execute(nodes, precision=prec, motion=motion)
        res = out.cpu().numpy()
    res = res.reshape(*plan.shape, V)
    return res.astype(np.complex64) if prec != "f64" else res


def _signal_generic_motion(plan, nodes, prec, fn, r0, dev):
    """Arbitrary callable: its displacement at each distinct t becomes one
    mode whose coefficient is 1 for that t only.  The kernel fuses up to 32
    modes, so the distinct times are processed 32 at a time, each pass on a
    plan of only the samples at its own times: every sample is evaluated
    once, as in the reference's loop."""
    G, N, V = plan.n_tgroups, nodes.n_nodes, nodes.n_enc
    total = torch.empty((plan.n_samples, V), dtype=torch.complex128, device=dev)
    groups = plan.sample_groups()
    k_flat = [np.ascontiguousarray(np.asarray(k, dtype=np.float64)).ravel() for k in plan.kloc]
    t_flat = np.ascontiguousarray(plan.t, dtype=np.float64).ravel()
    chunk = 32
    for c0 in range(0, G, chunk):
        c1 = min(G, c0 + chunk)
        times = plan.tgroup_t[c0:c1]
        sel = np.flatnonzero((groups >= c0) & (groups < c1))
        sub = device.SignalPlan(tuple(k[sel].reshape(-1, 1, 1) for k in k_flat), t_flat[sel].reshape(-1, 1, 1),
                                method="direct", device=dev)
        R = c1 - c0
        modes = np.stack([np.asarray(fn(float(tg)), dtype=np.float32).reshape(N, 3) for tg in times], axis=2)
        # the sub-plan orders its distinct times itself: one-hot rows by time
        slot = {float(tg): r for r, tg in enumerate(times)}
        coef = np.zeros((sub.n_tgroups, R + 3), dtype=np.float64)
        coef[np.arange(sub.n_tgroups), [slot[float(tg)] for tg in sub.tgroup_t]] = 1.0
        modes_soa = device.modes_soa(torch.from_numpy(np.ascontiguousarray(modes)).to(dev))
        part = sub.execute(nodes, precision=prec, motion=(modes_soa, torch.from_numpy(coef).to(dev)))
        total[torch.from_numpy(sel).to(dev)] = part
    return total
