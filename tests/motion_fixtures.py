"""Motion fixtures for the tests: this package's classes, and objects shaped
like the reference's motion classes for the duck-typed readers of
``feelmri_amd.Motion``.

The reference-shaped objects carry the attributes the reference's
evaluation uses (python/feelmri/Motion.py: ``POD._modes`` / ``_pp_batch`` /
``_fold_time`` / ``timeshift`` :112-261, the in-place ``*= t`` of
``PODVelocity`` :264-290, ``PODSum.pod1`` / ``pod2`` :293-317,
``RespiratoryMotion.interpolator`` / ``direction`` :15-109).  They are
built here from given modes and weight series, with scipy's PCHIP batched
into one ``PPoly`` with ``extrapolate=False``; they do not restate how the
reference derives its modes.
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import PchipInterpolator, PPoly

from feelmri_amd import Motion


def _fold(period):
    if period is None:
        return lambda x: x
    return lambda x: x - period * np.floor(x / period)


class RefPOD:
    def __init__(self, modes, times, series, periodic: bool = True):
        self._modes = np.ascontiguousarray(modes, dtype=np.float32)
        pp = PchipInterpolator(times, series, axis=0)
        self._pp_batch = PPoly(pp.c, pp.x, extrapolate=False)
        self._fold_time = _fold(float(times[-1]) if periodic else None)
        self.timeshift = 0.0

    def update_timeshift(self, timeshift):
        self.timeshift = timeshift

    def _w(self, t):
        return self._pp_batch(self._fold_time(t + self.timeshift)).astype(np.float32)

    def __call__(self, t):
        return np.tensordot(self._modes, self._w(t), axes=([2], [0]))

    def __add__(self, other):
        return RefSum(self, other)


class PODVelocity(RefPOD):
    """Named as the reference's class: the readers recognise velocities by it."""

    def __call__(self, t):
        w = self._w(t)
        w *= t
        return np.tensordot(self._modes, w, axes=([2], [0]))


class RefSum:
    def __init__(self, pod1, pod2):
        self.pod1, self.pod2 = pod1, pod2
        self.timeshift = 0.0

    def __call__(self, t):
        return self.pod1(t) + self.pod2(t)

    def update_timeshift(self, timeshift):
        self.pod1.update_timeshift(timeshift)
        self.pod2.update_timeshift(timeshift)
        self.timeshift = timeshift


class RefRespiratory:
    def __init__(self, times, data, direction, periodic: bool = True):
        times = np.asarray(times, dtype=np.float32)
        self.interpolator = PchipInterpolator(times, np.asarray(data, dtype=np.float32))
        d = np.asarray(direction, dtype=np.float64).reshape(1, 3)
        self.direction = (d / np.linalg.norm(d)).astype(np.float32)
        self._fold_time = _fold(times[-1] if periodic else None)
        self.timeshift = 0.0

    def update_timeshift(self, timeshift):
        self.timeshift = timeshift

    def __call__(self, t):
        return self.interpolator(self._fold_time(t + self.timeshift)).astype(np.float32) * self.direction

    def __add__(self, other):
        return RefSum(self, other)


def snapshots(n: int, seed: int, n_t: int = 10, period: float = 40.0, amp: float = 1e-3, rank: int = 3):
    """(times, data (n, 3, n_t), base modes, series): a rank-``rank`` periodic field."""
    rng = np.random.default_rng(seed)
    times = np.linspace(0, period, n_t)
    base = rng.normal(size=(n, 3, rank)) * amp
    series = np.stack([np.sin(2 * np.pi * times / period * (k + 1)) for k in range(rank)], 1)   # (T, R)
    return times, np.tensordot(base, series.T, axes=([2], [0])), base, series


def pod(n: int, seed: int = 11, velocity: bool = False, family: str = "own", periodic: bool = True):
    """A 3-mode field on n nodes: ``family`` "own" (ModalMotion) or "ref"
    (reference-shaped)."""
    times, data, base, series = snapshots(n, seed)
    if family == "own":
        return Motion.ModalMotion(times, data, rank=3, periodic=periodic, velocity=velocity)
    cls = PODVelocity if velocity else RefPOD
    return cls(base, times, series, periodic=periodic)


def resp(direction=(0, 1, 0), family: str = "own"):
    times = np.linspace(0, 100, 25)
    data = np.sin(np.linspace(0, 6, 25)) * 0.004
    if family == "own":
        return Motion.TranslationMotion(times, data, direction=direction, periodic=True)
    return RefRespiratory(times, data, direction)
