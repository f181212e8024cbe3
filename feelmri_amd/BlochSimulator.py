"""Drop-in for the reference's ``feelmri.BlochSimulator`` extension module.

``solve_mri(r0, T1, T2, delta_B, M0, gamma, rf_all, G_all, dt, regime_idx,
Mxy_initial, Mz_initial, pod_trajectory)`` keeps the reference's signature
and return value (cpp/feelmri/BlochSimulator.h:16-80,
BlochSimulator.cpp:12-276): the full magnetisation history
``(Mxy (N, T) complex64, Mz (N, T) float32)`` with column 0 the initial
state.  The reference binds only the float32 instantiation; here the
integration runs in FP64 by default (``precision="f64"``) or FP32
(``precision="f32"``) on the GPU and the result keeps the reference's
dtypes.  ``regime_idx`` is accepted and unused, as in the reference
(BlochSimulator.cpp:253-267).  Motion: ``pod_trajectory`` is evaluated at
the end of every step, ti = dt[1] + ... + dt[i+1] (BlochSimulator.cpp:
81-89), through the mode/coefficient form of ``feelmri_amd.Motion``;
generic callables are sampled per step on the host.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Motion, device


def step_times(dt) -> np.ndarray:
    """ti after each step i (the reference's running ``ti += dt[i+1]``)."""
    dt = np.asarray(dt, dtype=np.float64).ravel()
    return np.cumsum(dt[1:]) if dt.size > 1 else np.zeros(0)


def bloch_motion_tables(pod_trajectory, n_pos: int, times, dev):
    """(modes_soa, coef (len(times), R+3)) for a supported motion object."""
    modes = Motion.motion_modes(pod_trajectory)
    coef = Motion.motion_coefficients(pod_trajectory, times)
    if modes.shape[2] == 0:  # translation only
        modes = np.zeros((n_pos, 3, 1), dtype=np.float32)
        coef = np.concatenate([np.zeros((coef.shape[0], 1)), coef], axis=1)
    if modes.shape[0] != n_pos:
        raise ValueError(f"motion modes have {modes.shape[0]} nodes, expected {n_pos}")
    modes_soa = device.modes_soa(torch.from_numpy(np.ascontiguousarray(modes)).to(dev))
    return modes_soa, torch.from_numpy(np.ascontiguousarray(coef, dtype=np.float64)).to(dev)


# Steps per launch when a generic callable's displacements are passed as
# one-hot modes (the kernel fuses at most 16 modes).
_GENERIC_STEPS = 16


def solve_mri(r0, T1, T2, delta_B, M0, gamma, rf_all, G_all, dt, regime_idx, Mxy_initial,
              Mz_initial, pod_trajectory=None, *, precision: str = "f64", history: bool = True,
              dev=None):
    if pod_trajectory is not None and not callable(pod_trajectory):
        raise TypeError("pod_trajectory must be callable or None")
    dev = torch.device(dev) if dev is not None else device.default_device()
    r0 = np.asarray(r0, dtype=np.float64).reshape(-1, 3)
    N = r0.shape[0]
    rf = np.asarray(rf_all).reshape(-1)
    G = np.asarray(G_all, dtype=np.float64).reshape(-1, 3)
    dtv = np.asarray(dt, dtype=np.float64).reshape(-1)
    T = rf.size
    if T == 0:
        raise ValueError("rf_all is empty")
    if G.shape[0] != T or dtv.size != T:
        raise ValueError(f"rf_all ({T}), G_all ({G.shape[0]}) and dt ({dtv.size}) lengths differ")
    real = torch.float64 if precision == "f64" else torch.float32
    cplx = torch.complex128 if precision == "f64" else torch.complex64
    with torch.cuda.device(dev):
        nodes = device.BlochNodes.from_arrays(r0, np.asarray(T1).reshape(-1), np.asarray(T2).reshape(-1),
                                              np.asarray(delta_B).reshape(-1), device=dev)
        mxy0 = device.as_device(np.asarray(Mxy_initial).reshape(-1), torch.complex128, dev)
        mz0 = device.as_device(np.asarray(Mz_initial).reshape(-1), torch.float64, dev)
        if mxy0.numel() != N or mz0.numel() != N:
            raise ValueError("initial magnetisation length differs from the node count")
        if pod_trajectory is not None:
            pod_trajectory(0.0)  # BlochSimulator.cpp:44 evaluates once at t = 0
        generic = pod_trajectory is not None and not Motion.is_supported(pod_trajectory)
        if not generic:
            ev = device.EventTable(rf, G, dtv, device=dev)
            motion = None
            if pod_trajectory is not None:
                motion = bloch_motion_tables(pod_trajectory, N, step_times(dtv), dev)
            mxy, mz = device.bloch(nodes, ev, float(M0), float(gamma), mxy0, mz0, precision=precision,
                                   motion=motion, history=history)
        else:
            # generic callable: launches of <= 16 steps, each step's
            # displacement a one-hot mode; state carried between launches
            times = step_times(dtv)
            hist_c = torch.empty((T, N), dtype=cplx, device=dev) if history else None
            hist_r = torch.empty((T, N), dtype=real, device=dev) if history else None
            cur_xy, cur_z = mxy0, mz0
            for c0 in range(0, T - 1, _GENERIC_STEPS):
                c1 = min(T - 1, c0 + _GENERIC_STEPS)
                n = c1 - c0
                disp = np.stack([np.asarray(pod_trajectory(float(t)), dtype=np.float32).reshape(N, 3)
                                 for t in times[c0:c1]], axis=2)
                coef = np.zeros((n, n + 3))
                coef[np.arange(n), np.arange(n)] = 1.0
                motion = (device.modes_soa(torch.from_numpy(np.ascontiguousarray(disp)).to(dev)),
                          torch.from_numpy(coef).to(dev))
                ev = device.EventTable(rf[c0:c1 + 1], G[c0:c1 + 1], dtv[c0:c1 + 1], device=dev)
                mo = hist_c[c0:c1 + 1] if history else None
                zo = hist_r[c0:c1 + 1] if history else None
                a, b = device.bloch(nodes, ev, float(M0), float(gamma), cur_xy, cur_z,
                                    precision=precision, motion=motion, history=history,
                                    mxy_out=mo, mz_out=zo)
                last_xy = a[:, -1] if history else a
                last_z = b[:, -1] if history else b
                cur_xy = last_xy.to(torch.complex128).contiguous()
                cur_z = last_z.to(torch.float64).contiguous()
            if T == 1:
                a, b = device.bloch(nodes, device.EventTable(rf, G, dtv, device=dev), float(M0),
                                    float(gamma), mxy0, mz0, precision=precision, history=history)
                mxy, mz = a, b
            elif history:
                mxy, mz = hist_c.T, hist_r.T
            else:
                mxy, mz = cur_xy, cur_z
        mxy_h = mxy.to(torch.complex64).cpu().numpy()
        mz_h = mz.to(torch.float32).cpu().numpy()
    return mxy_h, mz_h
