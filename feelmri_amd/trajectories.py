"""k-space trajectory producers: k locations (cycles/m) and sample times (ms).

Restated from the reference's trajectory classes for ``lines_per_shot = 1``
and no planning rotation (python/feelmri/KSpaceTraj.py):

* Cartesian stack  -- KSpaceTraj.py:40-44 (extents), :183-185 (kx, ky, kz),
  :209-220 (filling, times).
* Golden-angle radial, half spoke -- KSpaceTraj.py:335-368.
* Variable-density spiral stack -- KSpaceTraj.py:423-434 (base spiral),
  :530-553 (interleaf rotation, times).  The reference derives the time law
  from gradient limits and resamples onto the ADC grid; here the spiral is
  sampled uniformly in its parameter and the ADC times are uniform.

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
