"""k-space trajectory producers: k locations (cycles/m) and sample times (ms).

Restated from the reference's trajectory classes for ``lines_per_shot = 1``
and no planning rotation (python/feelmri/KSpaceTraj.py):

* Cartesian stack  -- KSpaceTraj.py:40-44 (extents), :183-185 (kx, ky, kz),
  :209-220 (filling, times).
* Golden-angle radial, half spoke -- KSpaceTraj.py:335-368.
* Variable-density spiral stack -- ``spiral_stack``: KSpaceTraj.py:386-562,
  the base spiral in its parameter u (:423-434), the time law from the
  gradient amplitude and slew limits (:436-473), resampling onto the uniform
  ADC grid (:493-505), interleaf rotation, kz planes and times (:507-558).
  ``spiral`` is a simpler form kept for small test shapes: the spiral sampled
  uniformly in u at uniform times.

Every producer returns ``(kx, ky, kz), t`` as arrays of shape
(n_meas, n_lines, n_kz), the layout ``MRI.Signal`` takes.
"""
from __future__ import annotations

import numpy as np

GOLDEN_RATIO = 1.61803398875


def cartesian(fov, res, oversampling: int = 1, t_first: float = 0.0, t_readout: float = 2.0,
              dtype=np.float64):
    """Cartesian stack: kx along readout (dim 0), ky per line (dim 1), kz per slice (dim 2).

    ``t`` is ``t_first + linspace(0, t_readout, n_meas)`` ms for every line
    and slice, so it depends on the readout index only.
    """
    fov = np.asarray(fov, dtype=np.float64)
    res = np.asarray(res, dtype=np.int64)
    ro = int(oversampling * res[0])
    ph = int(res[1])
    sl = int(res[2])
    k_spa = 1.0 / (np.array([oversampling, 1.0, 1.0]) * fov)
    kx_ext = (np.array([0, ro - 1]) - ro // 2) * k_spa[0] + float(res[0] % 2 != 0) * k_spa[0]
    ky_ext = (np.array([0, ph - 1]) - ph // 2) * k_spa[1]
    kz_ext = (np.array([0, sl - 1]) - sl // 2) * k_spa[2]
    kx1 = np.linspace(kx_ext[0], kx_ext[1], ro)
    ky1 = ky_ext[0] + k_spa[1] * np.arange(ph)
    kz1 = np.linspace(kz_ext[0], kz_ext[1], sl)
    kx = np.broadcast_to(kx1[:, None, None], (ro, ph, sl))
    ky = np.broadcast_to(ky1[None, :, None], (ro, ph, sl))
    kz = np.broadcast_to(kz1[None, None, :], (ro, ph, sl))
    t1 = t_first + np.linspace(0.0, t_readout, ro)
    t = np.broadcast_to(t1[:, None, None], (ro, ph, sl))
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (kx, ky, kz)), \
        np.ascontiguousarray(t, dtype=dtype)


def radial(fov: float, n_samples: int, n_spokes: int, t_first: float = 0.0,
           t_readout: float = 2.56, oversampling: int = 2, dtype=np.float64):
    """2-D golden-angle half-spoke radial (KSpaceTraj.py:335-368), one slice.

    ``n_samples`` readout samples = ``oversampling`` x the image resolution
    (the reference's default oversampling is 2, KSpaceTraj.py:16); the
    spoke runs from k = 0 to the readout extent of KSpaceTraj.py:42,
    (n_samples - 1 - n_samples // 2) / (oversampling FOV), plus one step for
    an odd resolution."""
    res0 = n_samples // oversampling
    k_spa = 1.0 / (oversampling * fov)
    kmax = (n_samples - 1 - n_samples // 2) * k_spa + float(res0 % 2 != 0) * k_spa
    kr = np.linspace(0.0, kmax, n_samples)
    n = np.arange(n_spokes)
    theta = np.mod((2.0 * np.pi - 2.0 * np.pi / GOLDEN_RATIO) * n, 2.0 * np.pi)
    kx = kr[:, None] * np.cos(theta)[None, :]
    ky = -kr[:, None] * np.sin(theta)[None, :]
    kz = np.zeros_like(kx)
    t1 = t_first + np.linspace(0.0, t_readout, n_samples)
    t = np.broadcast_to(t1[:, None], kx.shape)
    return tuple(np.ascontiguousarray(a[..., None], dtype=dtype) for a in (kx, ky, kz)), \
        np.ascontiguousarray(t[..., None], dtype=dtype)


def spiral(fov: float, res: int, n_samples: int, n_interleaves: int, density_exponent: float = 2.0,
           turns: float | None = None, t_first: float = 0.0, t_readout: float = 8.0,
           dtype=np.float64):
    """Variable-density spiral stack: kr = kmax u^p, phi = 2 pi turns u^(1/p),
    interleaf m rotated by 2 pi m / M (KSpaceTraj.py:423-434, :530-553)."""
    kmax = 0.5 * res / fov
    if turns is None:
        turns = max(1.0, kmax * fov / n_interleaves)
    p = float(density_exponent)
    u = np.linspace(0.0, 1.0, n_samples)
    kr = kmax * u ** p
    phi = 2.0 * np.pi * turns * u ** (1.0 / p)
    K = kr * np.exp(1j * phi)
    theta = np.linspace(0.0, 2.0 * np.pi, n_interleaves, endpoint=False)
    Kr = K[:, None] * np.exp(1j * theta)[None, :]
    kx, ky = np.real(Kr), np.imag(Kr)
    kz = np.zeros_like(kx)
    t1 = t_first + np.linspace(0.0, t_readout, n_samples)
    t = np.broadcast_to(t1[:, None], kx.shape)
    return tuple(np.ascontiguousarray(a[..., None], dtype=dtype) for a in (kx, ky, kz)), \
        np.ascontiguousarray(t[..., None], dtype=dtype)


# Scanner defaults of the reference (MRObjects.py:33-39): 33 mT/m, 180 mT/m/ms,
# gammabar 42.58 MHz/T.
SCANNER_G_MAX = 33.0          # mT/m
SCANNER_SLEW = 180.0          # mT/m/ms (= T/m/s)
GAMMABAR = 42.58e6            # Hz/T


def spiral_time_law(u: np.ndarray, K: np.ndarray, g_max: float = SCANNER_G_MAX, slew: float = SCANNER_SLEW,
                    gammabar: float = GAMMABAR, safety_margin: float = 0.95):
    """Readout time t(u) (s) of the base spiral K(u) (1/m) under the gradient
    amplitude and slew limits (KSpaceTraj.py:436-473), and its total T.

    Per unit of u the time is the larger of |dK/du| / (gamma G) and
    sqrt(|d2K/du2| / (gamma S)), gamma = 2 pi gammabar, G and S the limits
    times the safety margin; derivatives are second-order finite differences
    in u.  t is the cumulative trapezoid rule with the previous node taken
    cyclically (so the first interval pairs the first and the last rate), and
    t[0] = 0 -- the reference's integration, kept as it is.
    """
    gamma = 2.0 * np.pi * gammabar
    G = g_max * 1e-3 * safety_margin              # T/m
    S = slew * safety_margin                      # T/m/s
    du = np.gradient(u)
    d1 = np.gradient(K, u, edge_order=2)
    d2 = np.gradient(d1, u, edge_order=2)
    rate = np.maximum(np.abs(d1) / (gamma * G), np.sqrt(np.maximum(np.abs(d2), 0.0) / (gamma * S)))
    t = np.cumsum(0.5 * (rate + np.roll(rate, 1)) * du)
    t[0] = 0.0
    return t, float(t[-1])


def spiral_stack(fov=(0.3, 0.3, 0.08), res=(100, 100, 1), oversampling: int = 2, t_start: float = 0.0,
                 receiver_bw: float = 128.0e3, density_exponent: float = 1.0, safety_margin: float = 0.95,
                 g_max: float = SCANNER_G_MAX, slew: float = SCANNER_SLEW, gammabar: float = GAMMABAR,
                 dtype=np.float32):
    """Stack of variable-density spirals under gradient limits
    (KSpaceTraj.py:386-562, ``lines_per_shot = 1``).

    * base spiral of res[0] points: kr = kmax u^p, phi = 2 pi turns u^(1/p),
      u = linspace(0, 1) in ``dtype``, kmax = res[0] / (2 FOV[0]) and turns =
      min(max(1, kmax FOV[0]), max(1, res[0] / 32));
    * time law t(u) from ``spiral_time_law``, total T;
    * ADC grid: round(T receiver_bw) samples of the receiver, times
      oversampling, uniform on [0, T] (both ends); k by linear interpolation
      of Re K and Im K in t;
    * res[1] interleaves rotated by 2 pi m / res[1]; res[2] kz planes at
      (s - res[2] // 2) / FOV[2]; t = t_start + ADC time (ms) for every
      interleave and plane.

    Returns ``(kx, ky, kz), t`` of shape (samples, res[1], res[2]) in
    ``dtype`` (cycles/m, ms) and the readout duration T (ms).
    """
    fov = np.asarray(fov, dtype=np.float64)
    n0, n_il, n_sl = (int(r) for r in res)
    p = float(density_exponent)
    kmax = 0.5 * n0 / fov[0]
    turns = min(max(1.0, kmax * fov[0]), max(1.0, n0 / 32.0))
    u = np.linspace(0.0, 1.0, n0, dtype=dtype)
    K = (kmax * u ** p) * np.exp(1j * (2.0 * np.pi * turns * u ** (1.0 / p)))
    t_cont, T = spiral_time_law(u, K, g_max, slew, gammabar, safety_margin)
    n_adc = int(np.round(T * receiver_bw))
    t_adc = np.linspace(0.0, T, n_adc * int(oversampling), endpoint=True)
    Ka = np.interp(t_adc, t_cont, K.real) + 1j * np.interp(t_adc, t_cont, K.imag)
    theta = np.linspace(0.0, 2.0 * np.pi, n_il, endpoint=False, dtype=dtype)
    Kr = Ka[:, None] * np.exp(1j * theta)[None, :]
    kz1 = np.linspace((0 - n_sl // 2) / fov[2], (n_sl - 1 - n_sl // 2) / fov[2], n_sl)
    shape = (t_adc.size, n_il, n_sl)
    kx = np.broadcast_to(Kr.real[:, :, None], shape)
    ky = np.broadcast_to(Kr.imag[:, :, None], shape)
    kz = np.broadcast_to(kz1[None, None, :], shape)
    t = np.broadcast_to((t_start + 1e3 * t_adc)[:, None, None], shape)
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (kx, ky, kz)), \
        np.ascontiguousarray(t, dtype=dtype), 1e3 * T

