"""Node motion: modal displacement / velocity fields and their device form.

The signal and Bloch kernels never receive per-node positions from the host.
A motion becomes ``modes`` -- (P, 3, R) float32 spatial fields -- plus, per
evaluation time, one coefficient row of R mode weights and a uniform
translation (``motion_coefficients``); the fused loaders form
r_n(t) = r0_n + sum_r modes[n, :, r] w_r(t) + trans(t) on the device.

Two families of objects map to that form:

* the reference's own motion objects (python/feelmri/Motion.py: ``POD``
  :112-261, ``PODVelocity`` :264-290, ``PODSum`` :293-317,
  ``RespiratoryMotion`` :15-109), read by duck typing on the attributes
  their evaluation uses (``_modes``, ``_pp_batch``, ``_fold_time``,
  ``timeshift``, ``pod1``/``pod2``, ``interpolator``/``direction``), so a
  user keeps the reference's classes and passes them as they are;
* this module's classes, which implement ``modal_form()`` /
  ``coefficient_rows(times)`` directly:

  - ``ModalMotion``: a field sampled at snapshot times (displacement, or
    velocity scaled by t), reduced to its R leading modes by a truncated SVD
    of the (3P x T) snapshot matrix, with one vector-valued interpolant of
    the R coefficient series in time (PCHIP, Akima or cubic), folded into
    the period for periodic motion and NaN outside the snapshot span
    otherwise;
  - ``TranslationMotion``: a rigid translation, a scalar waveform times a
    unit direction;
  - ``MotionSum``: the sum of two motions.

The modes and coefficients are rounded to float32 as the reference's
``_modes`` / ``_weights`` arrays are, so both families give the device the
same numbers for the same field.
"""
from __future__ import annotations

from collections.abc import Callable

import numpy as np
from scipy.interpolate import Akima1DInterpolator, CubicSpline, PchipInterpolator

_INTERPOLANTS = ("pchip", "akima", "cubic")


def _interpolant(kind: str, times: np.ndarray, values: np.ndarray, natural: bool = False):
    """Interpolant of ``values`` (T, ...) over ``times`` along axis 0; NaN
    outside [times[0], times[-1]]."""
    if kind == "pchip":
        return PchipInterpolator(times, values, axis=0, extrapolate=False)
    if kind == "akima":
        return Akima1DInterpolator(times, values, axis=0, extrapolate=False)
    if kind == "cubic":
        return CubicSpline(times, values, axis=0, bc_type="natural" if natural else "not-a-knot",
                           extrapolate=False)
    raise ValueError(f"interpolation must be one of {_INTERPOLANTS}, got {kind!r}")


class _Clock:
    """Time shift and optional fold into [0, period) shared by the motions."""

    def __init__(self, period: float | None, timeshift: float):
        self.period = None if period is None else float(period)
        self.timeshift = float(timeshift)

    def local(self, t: float) -> float:
        u = float(t) + self.timeshift
        if self.period is not None:
            u -= self.period * np.floor(u / self.period)
        return u

    def update_timeshift(self, timeshift: float) -> None:
        self.timeshift = float(timeshift)


class ModalMotion:
    """Snapshot field reduced to ``rank`` spatial modes.

    ``snapshots`` (P, 3, T) sampled at ``times`` (T,) in ms: a displacement
    in m, or with ``velocity=True`` a velocity in m/ms whose evaluation is
    scaled by the (unshifted) time, as the reference's ``PODVelocity`` does.
    ``periodic``: fold times into [0, times[-1]).

    Decomposition: X = snapshots as (3P, T) = U S V^T (thin SVD); the modes
    are the first ``rank`` columns of U and the coefficient series at the
    snapshot times are the first ``rank`` rows of S V^T, so the field at a
    snapshot time is the rank-R projection of its snapshot.
    """

    def __init__(self, times, snapshots, rank: int = 5, periodic: bool = False, velocity: bool = False,
                 interpolation: str = "pchip", timeshift: float = 0.0):
        times = np.asarray(times, dtype=np.float64).ravel()
        X = np.asarray(snapshots, dtype=np.float64)
        if X.ndim != 3 or X.shape[1] != 3 or X.shape[2] != times.size:
            raise ValueError(f"snapshots must be (P, 3, T) with T = {times.size}, got {X.shape}")
        if not 1 <= rank <= times.size:
            raise ValueError(f"rank must be in [1, {times.size}]")
        P, T = X.shape[0], times.size
        U, S, Vt = np.linalg.svd(X.reshape(3 * P, T), full_matrices=False)
        self.rank = int(rank)
        self.times = times
        self.singular_values = S
        self.modes = np.ascontiguousarray(U[:, :rank].reshape(P, 3, rank), dtype=np.float32)
        self.series = (S[:rank, None] * Vt[:rank]).T            # (T, R) coefficients at the snapshots
        self.velocity = bool(velocity)
        self.interpolation = interpolation
        self._interp = _interpolant(interpolation, times, self.series)
        self.clock = _Clock(times[-1] if periodic else None, timeshift)

    @property
    def timeshift(self) -> float:
        return self.clock.timeshift

    def update_timeshift(self, timeshift: float) -> None:
        self.clock.update_timeshift(timeshift)

    def coefficients(self, t: float) -> np.ndarray:
        """(R,) float32 mode weights at time ``t``."""
        w = np.asarray(self._interp(self.clock.local(t)), dtype=np.float64).astype(np.float32)
        if self.velocity:
            w = (w * np.float32(t)).astype(np.float32)
        return w

    def __call__(self, t: float) -> np.ndarray:
        """(P, 3) float32 field at time ``t``."""
        return np.tensordot(self.modes, self.coefficients(t), axes=([2], [0]))

    def __add__(self, other):
        return MotionSum(self, other)

    # device form
    def modal_form(self) -> np.ndarray:
        return self.modes

    def coefficient_rows(self, times) -> np.ndarray:
        times = np.asarray(times, dtype=np.float64).ravel()
        out = np.zeros((times.size, self.rank + 3), dtype=np.float64)
        for i, t in enumerate(times.tolist()):
            out[i, :self.rank] = self.coefficients(t)
        return out


class TranslationMotion:
    """Rigid translation: ``waveform`` (T,) in m sampled at ``times`` times a
    unit ``direction``; ``remove_mean`` subtracts the waveform's mean."""

    def __init__(self, times, waveform, direction=(1.0, 0.0, 0.0), periodic: bool = False,
                 remove_mean: bool = False, interpolation: str = "pchip", timeshift: float = 0.0):
        times = np.asarray(times, dtype=np.float32).astype(np.float64).ravel()
        wave = np.asarray(waveform, dtype=np.float32).ravel()
        if wave.size != times.size:
            raise ValueError("waveform and times differ in length")
        if remove_mean:
            wave = wave - np.mean(wave, dtype=np.float32)
        d = np.asarray(direction, dtype=np.float64).reshape(3)
        self.direction = (d / np.linalg.norm(d)).astype(np.float32)
        self.times = times
        self._interp = _interpolant(interpolation, times, wave.astype(np.float64), natural=True)
        self.clock = _Clock(times[-1] if periodic else None, timeshift)

    @property
    def timeshift(self) -> float:
        return self.clock.timeshift

    def update_timeshift(self, timeshift: float) -> None:
        self.clock.update_timeshift(timeshift)

    def offset(self, t: float) -> np.ndarray:
        """(3,) float32 translation at time ``t``."""
        s = np.float32(np.asarray(self._interp(self.clock.local(t))).reshape(-1)[0])
        return (s * self.direction).astype(np.float32)

    def __call__(self, t: float) -> np.ndarray:
        return self.offset(t).reshape(1, 3)

    def __add__(self, other):
        return MotionSum(self, other)

    def modal_form(self) -> np.ndarray:
        return np.zeros((0, 3, 0), dtype=np.float32)

    def coefficient_rows(self, times) -> np.ndarray:
        times = np.asarray(times, dtype=np.float64).ravel()
        out = np.zeros((times.size, 3), dtype=np.float64)
        for i, t in enumerate(times.tolist()):
            out[i] = self.offset(t)
        return out


class MotionSum:
    """a(t) + b(t); time shifts are set on both."""

    def __init__(self, a, b: Callable):
        self.a, self.b = a, b

    def __call__(self, t: float) -> np.ndarray:
        return self.a(t) + self.b(t)

    def __add__(self, other):
        return MotionSum(self, other)

    def update_timeshift(self, timeshift: float) -> None:
        self.a.update_timeshift(timeshift)
        self.b.update_timeshift(timeshift)


# ---------------------------------------------------------------------------
# Device form of any supported motion (this module's or the reference's)
# ---------------------------------------------------------------------------
def _kind(obj) -> str:
    if isinstance(obj, MotionSum):
        return "own-sum"
    if hasattr(obj, "coefficient_rows") and hasattr(obj, "modal_form"):
        return "own"
    # the reference's objects, by the attributes their evaluation uses
    if hasattr(obj, "pod1") and hasattr(obj, "pod2"):
        return "sum"
    if hasattr(obj, "_modes") and hasattr(obj, "_pp_batch"):
        return "velocity" if type(obj).__name__ == "PODVelocity" else "pod"
    if hasattr(obj, "interpolator") and hasattr(obj, "direction"):
        return "resp"
    return "other"


def _parts(obj):
    k = _kind(obj)
    if k == "own-sum":
        return obj.a, obj.b
    return obj.pod1, obj.pod2


def is_supported(obj) -> bool:
    """True when ``obj`` maps to modes + coefficients (no per-node host work)."""
    k = _kind(obj)
    if k in ("sum", "own-sum"):
        return all(is_supported(p) for p in _parts(obj))
    return k in ("own", "pod", "velocity", "resp")


def motion_modes(obj) -> np.ndarray:
    """(P, 3, R) float32 modes of a supported motion (R may be 0)."""
    k = _kind(obj)
    if k in ("sum", "own-sum"):
        a, b = (motion_modes(p) for p in _parts(obj))
        if a.shape[2] == 0:
            return b
        if b.shape[2] == 0:
            return a
        return np.concatenate([a, b], axis=2)
    if k == "own":
        return np.ascontiguousarray(obj.modal_form(), dtype=np.float32)
    if k in ("pod", "velocity"):
        return np.ascontiguousarray(obj._modes, dtype=np.float32)
    if k == "resp":
        return np.zeros((0, 3, 0), dtype=np.float32)
    raise TypeError(f"motion object of type {type(obj).__name__} has no mode representation")


def n_modes(obj) -> int:
    k = _kind(obj)
    if k in ("sum", "own-sum"):
        return sum(n_modes(p) for p in _parts(obj))
    if k == "own":
        return int(np.asarray(obj.modal_form()).shape[2])
    if k in ("pod", "velocity"):
        return int(np.asarray(obj._modes).shape[2])
    return 0


def motion_coefficients(obj, times) -> np.ndarray:
    """(len(times), R + 3) float64: per time the mode weights (float32
    values) followed by a uniform translation."""
    times = np.asarray(times, dtype=np.float64).ravel()
    k = _kind(obj)
    R = n_modes(obj)
    out = np.zeros((times.size, R + 3), dtype=np.float64)
    if k in ("sum", "own-sum"):
        pa, pb = _parts(obj)
        a, b = motion_coefficients(pa, times), motion_coefficients(pb, times)
        ra = n_modes(pa)
        out[:, :ra] = a[:, :ra]
        out[:, ra:R] = b[:, :R - ra]
        out[:, R:] = a[:, -3:] + b[:, -3:]
        return out
    if k == "own":
        return np.asarray(obj.coefficient_rows(times), dtype=np.float64)
    if k in ("pod", "velocity"):
        # the reference evaluates its batched PPoly at fold(t + timeshift)
        # into a float32 array, scaled in place by t for velocities
        for i, t in enumerate(times.tolist()):
            w = np.asarray(obj._pp_batch(obj._fold_time(t + obj.timeshift))).astype(np.float32)
            if k == "velocity":
                w *= t
            out[i, :R] = w
        return out
    if k == "resp":
        d = np.asarray(obj.direction, dtype=np.float32).reshape(3)
        for i, t in enumerate(times.tolist()):
            s = np.asarray(obj.interpolator(obj._fold_time(t + obj.timeshift))).astype(np.float32)
            out[i, R:] = (s.reshape(-1)[0] * d).astype(np.float32)
        return out
    raise TypeError(f"motion object of type {type(obj).__name__} has no mode representation")
