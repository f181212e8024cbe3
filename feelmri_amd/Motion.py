"""POD motion: the reference's motion classes and their device descriptors.

The classes mirror python/feelmri/Motion.py (same constructors, same
``__call__(t) -> (N, 3)`` host semantics, same ``update_timeshift``):

* ``RespiratoryMotion`` -- Motion.py:15-109: a spline of one scalar times a
  direction, evaluated at fold(t + timeshift).
* ``POD`` -- Motion.py:112-261: snapshot POD (``calculate_pod`` :174-217),
  per-mode PCHIP/Akima/cubic splines batched into one ``PPoly`` with
  ``extrapolate=False`` (:129-137), trajectory = modes . weights(fold(t+shift))
  (:237-254).
* ``PODVelocity`` -- Motion.py:264-290: weights scaled by the unshifted t.
* ``PODSum`` -- Motion.py:293-317: sum of two trajectories.

On the device the trajectory is never formed per node on the host: a
motion object becomes ``modes`` -- (3R, N) float32 rows -- plus, per
evaluation time, R mode weights and one uniform translation
(``motion_coefficients``).  The fused Signal and Bloch loaders then form
r_n(t) = r0_n + sum_r modes_r[n] w_r(t) + trans(t).  ``motion_modes`` and
``motion_coefficients`` also accept the reference's own objects (duck-typed
on ``_modes``, ``_pp_batch``, ``_fold_time``, ``timeshift``, ``pod1/pod2``,
``interpolator``/``direction``).
"""
from __future__ import annotations

import time
from collections.abc import Callable
from typing import Literal

import numpy as np
from scipy.interpolate import Akima1DInterpolator, CubicSpline, PchipInterpolator, PPoly

_INTERP = {"AkimaSpline": Akima1DInterpolator, "CubicSpline": CubicSpline, "Pchip": PchipInterpolator}


def _folder(period):
    if period is None:
        return lambda x: x

    def _fold(x, T=period):
        return x - T * np.floor(x / T)

    return _fold


class RespiratoryMotion:
    def __init__(self, times: np.ndarray, data: np.ndarray, timeshift: float = 0.0,
                 is_periodic: bool = False, remove_mean: bool = False,
                 direction: np.ndarray = np.array([1, 0, 0], dtype=np.float32),
                 interpolation_method: Literal["AkimaSpline", "CubicSpline", "Pchip"] = "Pchip"):
        self.times = np.asarray(times).astype(np.float32)
        self.data = np.asarray(data).astype(np.float32)
        self.timeshift = timeshift
        self.is_periodic = is_periodic
        self.remove_mean = remove_mean
        d = np.asarray(direction, dtype=np.float64).reshape((1, 3))
        self.direction = (d / np.linalg.norm(d)).astype(np.float32)
        self.interpolation_method = interpolation_method
        self.interpolator = self.calculate_interpolator()
        self._period = self.times[-1] if self.is_periodic else None
        self._fold_time = _folder(self._period)

    def __add__(self, other):
        return PODSum(self, other)

    def calculate_interpolator(self):
        data = self.data
        if self.remove_mean:
            data = data - np.mean(self.data, axis=0, dtype=np.float32)
        if self.interpolation_method not in _INTERP:
            raise ValueError(f"Interpolation method '{self.interpolation_method}' not recognized. "
                             "Choose from 'AkimaSpline', 'CubicSpline', or 'Pchip'.")
        if self.interpolation_method == "CubicSpline":
            return CubicSpline(self.times, data, bc_type="natural")
        return _INTERP[self.interpolation_method](self.times, data)

    def _evaluate_motion(self, t):
        return self.interpolator(self._fold_time(t + self.timeshift)).astype(np.float32)

    def __call__(self, t):
        return self._evaluate_motion(t) * self.direction

    def update_timeshift(self, timeshift):
        self.timeshift = timeshift


class POD:
    def __init__(self, times: np.ndarray, data: np.ndarray, global_to_local: np.ndarray = None,
                 n_modes: int = 5, is_periodic: bool = False,
                 interpolation_method: Literal["AkimaSpline", "CubicSpline", "Pchip"] = "Pchip",
                 timeshift: float = 0.0):
        self.times = np.asarray(times)
        self.data = data
        self.local_to_global_map = global_to_local
        self.n_modes = n_modes
        self.timeshift = timeshift
        self.is_periodic = is_periodic
        self.interpolation_method = interpolation_method
        self.modes, self.weights = self.calculate_pod(remove_mean=False)
        self.spline_coeffs = self.spline_fit()
        x0 = self.spline_coeffs[0].x
        if not all(np.array_equal(pp.x, x0) for pp in self.spline_coeffs):
            raise ValueError("knots differ")
        C = np.stack([pp.c for pp in self.spline_coeffs], axis=-1)
        self._pp_batch = PPoly(C, x0, extrapolate=False)
        self._modes = np.asarray(self.modes, dtype=np.float32, order="C")
        self._period = self.times[-1] if self.is_periodic else None
        self._weights = np.zeros([self.n_modes], dtype=np.float32, order="C")
        self._fold_time = _folder(self._period)

    def __repr__(self):
        return (f"POD(n_modes={self.n_modes}, interpolation_method='{self.interpolation_method}', "
                f"is_periodic={self.is_periodic})")

    def __add__(self, other):
        return PODSum(self, other)

    def __call__(self, t):
        return self._evaluate_trajectory(t)

    def calculate_pod(self, remove_mean: bool = False):
        """Snapshot POD (Motion.py:174-217): eigh of the t x t covariance."""
        n_t = self.times.shape[0]
        flat = np.asarray(self.data).reshape(-1, n_t)
        if remove_mean:
            flat = flat - np.mean(flat, axis=1, keepdims=True)
        cov = np.dot(flat.T, flat)
        evals, evecs = np.linalg.eigh(cov)
        idx = np.argsort(evals)[::-1][0:self.n_modes]
        evals, evecs = evals[idx], evecs[:, idx]
        phi = np.dot(flat, evecs / np.sqrt(evals).reshape(1, -1))
        weights = np.einsum("pn, pt -> nt", flat, phi)
        phi = phi.reshape((np.asarray(self.data).shape[0], -1, self.n_modes))
        if self.local_to_global_map is not None:
            phi = phi[self.local_to_global_map, :, :]
        return phi, weights

    def spline_fit(self):
        if self.interpolation_method not in _INTERP:
            raise ValueError(f"Interpolation method '{self.interpolation_method}' not recognized. "
                             "Choose from 'AkimaSpline', 'CubicSpline', or 'Pchip'.")
        interp = _INTERP[self.interpolation_method]
        return [interp(self.times, self.weights[:, i]) for i in range(self.n_modes)]

    def _evaluate_weights(self, t):
        self._weights[:] = self._pp_batch(t).astype(self._weights.dtype, copy=False)

    def _evaluate_trajectory(self, t):
        self._evaluate_weights(self._fold_time(t + self.timeshift))
        return np.tensordot(self._modes, self._weights, axes=([2], [0]))

    def update_timeshift(self, timeshift):
        self.timeshift = timeshift


class PODVelocity(POD):
    def _evaluate_trajectory(self, t):
        self._evaluate_weights(self._fold_time(t + self.timeshift))
        self._weights *= t
        return np.tensordot(self._modes, self._weights, axes=([2], [0]))


class PODSum:
    def __init__(self, pod1, pod2: Callable):
        self.pod1 = pod1
        self.pod2 = pod2
        self.timeshift = 0.0

    def __call__(self, t):
        return self.pod1(t) + self.pod2(t)

    def update_timeshift(self, timeshift):
        self.pod1.update_timeshift(timeshift)
        self.pod2.update_timeshift(timeshift)
        self.timeshift = timeshift


# ---------------------------------------------------------------------------
# Device descriptors
# ---------------------------------------------------------------------------
def _kind(obj) -> str:
    if hasattr(obj, "pod1") and hasattr(obj, "pod2"):
        return "sum"
    if hasattr(obj, "_modes") and hasattr(obj, "_pp_batch"):
        return "velocity" if type(obj).__name__ == "PODVelocity" or isinstance(obj, PODVelocity) else "pod"
    if hasattr(obj, "interpolator") and hasattr(obj, "direction"):
        return "resp"
    return "other"


def is_supported(obj) -> bool:
    """True when ``obj`` maps to modes + coefficients (no per-node host work)."""
    k = _kind(obj)
    if k == "sum":
        return is_supported(obj.pod1) and is_supported(obj.pod2)
    return k in ("pod", "velocity", "resp")


def motion_modes(obj) -> np.ndarray:
    """(P, 3, R) float32 modes of a supported motion object (R may be 0)."""
    k = _kind(obj)
    if k == "sum":
        a, b = motion_modes(obj.pod1), motion_modes(obj.pod2)
        if a.shape[2] == 0:
            return b
        if b.shape[2] == 0:
            return a
        return np.concatenate([a, b], axis=2)
    if k in ("pod", "velocity"):
        return np.ascontiguousarray(obj._modes, dtype=np.float32)
    if k == "resp":
        return np.zeros((0, 3, 0), dtype=np.float32)
    raise TypeError(f"motion object of type {type(obj).__name__} has no mode representation")


def n_modes(obj) -> int:
    k = _kind(obj)
    if k == "sum":
        return n_modes(obj.pod1) + n_modes(obj.pod2)
    if k in ("pod", "velocity"):
        return int(np.asarray(obj._modes).shape[2])
    return 0


def motion_coefficients(obj, times) -> np.ndarray:
    """(len(times), R + 3) float64: per time the mode weights (rounded to
    float32 as the reference's ``_weights``) and a uniform translation."""
    times = np.asarray(times, dtype=np.float64).ravel()
    k = _kind(obj)
    R = n_modes(obj)
    out = np.zeros((times.size, R + 3), dtype=np.float64)
    if k == "sum":
        a = motion_coefficients(obj.pod1, times)
        b = motion_coefficients(obj.pod2, times)
        ra = n_modes(obj.pod1)
        out[:, :ra] = a[:, :ra]
        out[:, ra:R] = b[:, ra:-3]
        out[:, R:] = a[:, -3:] + b[:, -3:]
        return out
    if k in ("pod", "velocity"):
        for i, t in enumerate(times.tolist()):
            w = np.asarray(obj._pp_batch(obj._fold_time(t + obj.timeshift))).astype(np.float32)
            if k == "velocity":
                w *= t  # Python float: float32 arithmetic, as the reference's in-place scale
            out[i, :R] = w
        return out
    if k == "resp":
        d = np.asarray(obj.direction, dtype=np.float32).reshape(3)
        for i, t in enumerate(times.tolist()):
            s = np.asarray(obj.interpolator(obj._fold_time(t + obj.timeshift))).astype(np.float32)
            out[i, R:] = (s.reshape(-1)[0] * d).astype(np.float32)
        return out
    raise TypeError(f"motion object of type {type(obj).__name__} has no mode representation")
