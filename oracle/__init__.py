"""FP64 CPU oracle for FEelMRI's signal-generation hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP path in
``feelmri_amd``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product never does.

It wraps ``oracle/_build/liboracle.so`` (plain C, strict IEEE FP64, built by
``oracle/Makefile``) and adds the host-side pieces the reference keeps in
Python: the ``M * Mxy`` product (cpp/feelmri/MRI.cpp:243), the per-``t``
position update through a motion callable (MRI.cpp:109-134), and the block
chaining / artificial spoiling of ``BlochSolver.solve``
(python/feelmri/Bloch.py:298-380).

Parity pins (see tests/test_oracle_*.py): closed forms (single node,
Dirichlet-kernel lattice, hard pulse, free precession, relaxation, Ernst
steady state, tet volumes), brute force in extended precision, scipy/numpy
library routines (PPoly, tensordot, sparse products).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_BUILD = _HERE / "_build"

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int
_p = ctypes.c_void_p

_libs: dict[str, ctypes.CDLL] = {}


def build() -> None:
    """Compile the oracle libraries with ``make`` (gcc only)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def _lib(fast: bool = False) -> ctypes.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = _BUILD / name
    if not path.exists():
        build()
    lib = ctypes.CDLL(str(path))
    lib.orc_signal.argtypes = [_i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _int]
    lib.orc_signal.restype = None
    lib.orc_bloch.argtypes = [_i64, _i64, _p, _p, _p, _p, _dbl, _dbl, _p, _p, _p, _p, _p, _p,
                              _p, _p, _int, _int]
    lib.orc_bloch.restype = None
    lib.orc_mass_p1_element_matrices.argtypes = [_i64, _p, _p, _p]
    lib.orc_mass_p1_lumped.argtypes = [_i64, _i64, _p, _p, _p]
    lib.orc_pod_tensordot.argtypes = [_i64, _i64, _i64, _p, _p, _p]
    lib.orc_csr_spmv_c128.argtypes = [_i64, _i64, _p, _p, _p, _p, _p]
    lib.orc_max_threads.restype = _int
    _libs[name] = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def max_threads(fast: bool = False) -> int:
    return int(_lib(fast).orc_max_threads())


# ---------------------------------------------------------------------------
# M * Mxy  (cpp/feelmri/MRI.cpp:243)
# ---------------------------------------------------------------------------
def weighted_mxy(M, Mxy) -> np.ndarray:
    """Return ``M @ Mxy`` in complex128.

    ``M`` may be a scipy sparse matrix (the reference's ``SparseMatrix``), a
    1-D array of lumped diagonal weights, or None (identity).
    """
    Mxy = _c128(Mxy)
    if Mxy.ndim == 1:
        Mxy = Mxy[:, None]
    if M is None:
        return Mxy.copy()
    if isinstance(M, np.ndarray) and M.ndim == 1:
        return _f64(M)[:, None] * Mxy
    import scipy.sparse as sp

    if sp.issparse(M):
        Mc = sp.csr_matrix(M, dtype=np.float64)
        Mc.sort_indices()
        out = np.empty_like(Mxy)
        indptr = np.ascontiguousarray(Mc.indptr, dtype=np.int64)
        indices = np.ascontiguousarray(Mc.indices, dtype=np.int64)
        data = _f64(Mc.data)
        _lib().orc_csr_spmv_c128(Mc.shape[0], Mxy.shape[1], _ptr(indptr), _ptr(indices),
                                 _ptr(data), _ptr(Mxy), _ptr(out))
        return out
    return _f64(M) @ Mxy


# ---------------------------------------------------------------------------
# Signal  (cpp/feelmri/MRI.cpp:38-159 moving, :196-297 static)
# ---------------------------------------------------------------------------
def signal_samples(kx, ky, kz, t, r, phi_dB0, T2, mmxy, n_threads: int = 0,
                   fast: bool = False) -> np.ndarray:
    """Flat-sample form: samples (S,), nodes r (N,3), mmxy (N,V) -> (S,V)."""
    kx, ky, kz, t = (_f64(a).ravel() for a in (kx, ky, kz, t))
    r = _f64(r).reshape(-1, 3)
    phi = _f64(phi_dB0).ravel()
    T2 = _f64(T2).ravel()
    mmxy = _c128(mmxy)
    if mmxy.ndim == 1:
        mmxy = mmxy[:, None]
    S, (N, V) = kx.size, mmxy.shape
    out = np.empty((S, V), dtype=np.complex128)
    _lib(fast).orc_signal(N, S, V, _ptr(kx), _ptr(ky), _ptr(kz), _ptr(t), _ptr(r), _ptr(phi),
                          _ptr(T2), _ptr(mmxy), _ptr(out), int(n_threads))
    return out


def Signal(MPI_rank, M, kloc, t, r0, phi_dB0, T2, Mxy, pod_trajectory=None,
           n_threads: int = 0) -> np.ndarray:
    """FP64 restatement of ``MRI.Signal`` (reference signature).

    Returns (n_meas, n_lines, n_kz, n_enc) complex128.  With a motion
    callable, positions are ``r0 + pod_trajectory(t)`` re-evaluated whenever
    ``t`` changes between consecutive samples in (i, j, k) order
    (MRI.cpp:109-134).
    """
    t = _f64(t)
    shape = t.shape
    kx, ky, kz = (_f64(k).ravel() for k in kloc)
    tf = t.ravel()
    mmxy = weighted_mxy(M, Mxy)
    r0 = _f64(r0).reshape(-1, 3)
    out = np.empty((tf.size, mmxy.shape[1]), dtype=np.complex128)
    if pod_trajectory is None:
        out[:] = signal_samples(kx, ky, kz, tf, r0, phi_dB0, T2, mmxy, n_threads)
    else:
        # runs of equal t in sample order: the reference's t_old cache
        starts = np.flatnonzero(np.r_[True, tf[1:] != tf[:-1]])
        ends = np.r_[starts[1:], tf.size]
        for s0, s1 in zip(starts, ends):
            r = r0 + _f64(pod_trajectory(tf[s0])).reshape(-1, 3)
            out[s0:s1] = signal_samples(kx[s0:s1], ky[s0:s1], kz[s0:s1], tf[s0:s1], r,
                                        phi_dB0, T2, mmxy, n_threads)
    return out.reshape(*shape, mmxy.shape[1])


# ---------------------------------------------------------------------------
# Bloch  (cpp/feelmri/BlochSimulator.cpp:12-148 moving, :151-276 static)
# ---------------------------------------------------------------------------
def solve_mri(r0, T1, T2, delta_B, M0, gamma, rf_all, G_all, dt, regime_idx,
              Mxy_initial, Mz_initial, pod_trajectory=None, history: bool = True,
              n_threads: int = 0):
    """FP64 restatement of ``BlochSimulator.solve_mri``.

    history=True returns (Mxy (N,T) complex128, Mz (N,T) float64) like the
    reference; history=False returns only the last column as (N,) arrays.
    """
    r0 = _f64(r0).reshape(-1, 3)
    N = r0.shape[0]
    T1, T2, dB = (_f64(a).ravel() for a in (T1, T2, delta_B))
    rf = _c128(rf_all).ravel()
    G = _f64(G_all).reshape(-1, 3)
    dt = _f64(dt).ravel()
    T = rf.size
    mxy0 = _c128(Mxy_initial).ravel()
    mz0 = _f64(Mz_initial).ravel()
    disp = None
    if pod_trajectory is not None:
        pod_trajectory(0.0)  # BlochSimulator.cpp:44 evaluates once at t=0
        disp = np.empty((max(T - 1, 0), N, 3), dtype=np.float64)
        ti = 0.0
        for i in range(T - 1):
            ti += dt[i + 1]
            disp[i] = _f64(pod_trajectory(ti)).reshape(N, 3)
    if history:
        mxy = np.empty((N, T), dtype=np.complex128, order="F")
        mz = np.empty((N, T), dtype=np.float64, order="F")
    else:
        mxy = np.empty(N, dtype=np.complex128)
        mz = np.empty(N, dtype=np.float64)
    _lib().orc_bloch(N, T, _ptr(r0), _ptr(T1), _ptr(T2), _ptr(dB), float(M0), float(gamma),
                     _ptr(rf), _ptr(G), _ptr(dt), _ptr(mxy0), _ptr(mz0), _ptr(disp),
                     _ptr(mxy), _ptr(mz), int(history), int(n_threads))
    return mxy, mz


def bloch_solver_solve(blocks, r0, T1, T2, delta_B, M0, gamma, initial_Mxy, initial_Mz,
                       pod_trajectory=None):
    """Block chaining of ``BlochSolver.solve`` (python/feelmri/Bloch.py:298-380).

    ``blocks``: sequence of dicts with keys ``rf`` (T,), ``G`` (T,3), ``dt``
    (T,), ``empty`` (bool), ``store`` (bool) and ``t0`` (block start time,
    used as the motion time shift, Bloch.py:335-336).  Returns (Mxy, Mz, Mxy_state,
    Mz_state): the stored last columns (N, n_store) and the carried state.
    """
    N = _f64(r0).reshape(-1, 3).shape[0]
    mxy_state = _c128(initial_Mxy).ravel().copy()
    mz_state = _f64(initial_Mz).ravel().copy()
    stored_xy, stored_z = [], []
    for blk in blocks:
        if pod_trajectory is not None:
            pod_trajectory.update_timeshift(blk["t0"])
        mxy, mz = solve_mri(r0, T1, T2, delta_B, M0, gamma, blk["rf"], blk["G"], blk["dt"],
                            None, mxy_state, mz_state, pod_trajectory, history=False)
        if blk.get("store", False):
            stored_xy.append(mxy.copy())
            stored_z.append(mz.copy())
        # Bloch.py:363-368: carry Mxy only across empty blocks, always carry Mz
        mxy_state = mxy.copy() if blk.get("empty", False) else np.zeros(N, np.complex128)
        mz_state = mz.copy()
    Mxy = np.stack(stored_xy, axis=1) if stored_xy else np.zeros((N, 0), np.complex128)
    Mz = np.stack(stored_z, axis=1) if stored_z else np.zeros((N, 0), np.float64)
    return Mxy, Mz, mxy_state, mz_state


# ---------------------------------------------------------------------------
# P1 mass  (cpp/feelmri/Assemble.cpp:15-116, python/feelmri/Phantom.py:379-399)
# ---------------------------------------------------------------------------
def mass_p1_element_matrices(elems, nodes) -> np.ndarray:
    elems = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 4)
    nodes = _f64(nodes).reshape(-1, 3)
    vals = np.empty((elems.shape[0], 4, 4), dtype=np.float64)
    _lib().orc_mass_p1_element_matrices(elems.shape[0], _ptr(elems), _ptr(nodes), _ptr(vals))
    return vals


def MassAssemble(elems, nodes):
    """Consistent P1 mass matrix as scipy CSC (duplicates summed)."""
    import scipy.sparse as sp

    elems = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 4)
    N = _f64(nodes).reshape(-1, 3).shape[0]
    vals = mass_p1_element_matrices(elems, nodes)
    rows = np.repeat(elems, 4, axis=1).ravel()
    cols = np.tile(elems, (1, 4)).ravel()
    return sp.csc_matrix((vals.ravel(), (rows, cols)), shape=(N, N))


def mass_p1_lumped(elems, nodes, fast: bool = False) -> np.ndarray:
    elems = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 4)
    nodes = _f64(nodes).reshape(-1, 3)
    w = np.empty(nodes.shape[0], dtype=np.float64)
    _lib(fast).orc_mass_p1_lumped(elems.shape[0], nodes.shape[0], _ptr(elems), _ptr(nodes), _ptr(w))
    return w


# ---------------------------------------------------------------------------
# POD  (cpp/feelmri/POD.cpp:5-38)
# ---------------------------------------------------------------------------
def tensordot_modes_weights(modes, weights) -> np.ndarray:
    modes = _f64(modes)
    P, C, M = modes.shape
    w = _f64(weights).ravel()
    out = np.empty((P, C), dtype=np.float64)
    _lib().orc_pod_tensordot(P, C, M, _ptr(modes), _ptr(w), _ptr(out))
    return out
