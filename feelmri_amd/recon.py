"""Image-domain error of a Cartesian k-space result.

``ktoi`` is the centred inverse FFT of python/feelmri/Math.py:10-14
(fftshift(ifftn(ifftshift(x)))) over the first three axes.  The image
error of a k-space result against a reference is the relative L2 of the
two reconstructions -- without the reference's Tukey filter and padding
(python/feelmri/Recon.py:13-90), which are linear and change both images
alike.  Host numpy: this is a measurement helper, not part of the path.
"""
from __future__ import annotations

import numpy as np
from numpy.fft import fftshift, ifftn, ifftshift


def ktoi(x: np.ndarray, axes=None) -> np.ndarray:
    if axes is None:
        axes = [i for i in range(x.ndim) if i < 3]
    return fftshift(ifftn(ifftshift(x, axes=axes), axes=axes), axes=axes)


def image_rel_l2(k_test: np.ndarray, k_ref: np.ndarray) -> float:
    a = ktoi(np.asarray(k_test, dtype=np.complex128))
    b = ktoi(np.asarray(k_ref, dtype=np.complex128))
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def adjoint_image(kx, ky, s, fov: float, n: int = 128, dcf: bool = True, device=None,
                  chunk: int = 4096) -> np.ndarray:
    """Image of 2-D non-Cartesian k-space samples on an n x n grid over the FOV
    by the direct adjoint DFT, I(x) = sum_s w_s S_s exp(+2 pi i k_s . x), with
    the radial density compensation w_s = |k_s| (``dcf``); the reference
    reconstructs radial and spiral data with a NUFFT adjoint
    (python/feelmri/Recon.py:151-290), which approximates this sum.  Runs on
    the GPU when one is visible (a measurement helper, not part of the path).
    k in cycles/m, fov in m; returns (n, n) complex128."""
    import torch

    dev = torch.device(device) if device is not None else (
        torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu"))
    kx = torch.as_tensor(np.ascontiguousarray(np.ravel(kx), dtype=np.float64), device=dev)
    ky = torch.as_tensor(np.ascontiguousarray(np.ravel(ky), dtype=np.float64), device=dev)
    sv = torch.as_tensor(np.ascontiguousarray(np.ravel(s), dtype=np.complex128), device=dev)
    if dcf:
        sv = sv * torch.sqrt(kx * kx + ky * ky)
    g = (torch.arange(n, dtype=torch.float64, device=dev) - n // 2) * (fov / n)
    X, Y = torch.meshgrid(g, g, indexing="ij")
    X, Y = X.reshape(-1), Y.reshape(-1)
    out = torch.zeros(n * n, dtype=torch.complex128, device=dev)
    for c0 in range(0, kx.numel(), chunk):
        c1 = min(kx.numel(), c0 + chunk)
        ph = 2.0 * np.pi * (kx[c0:c1, None] * X[None, :] + ky[c0:c1, None] * Y[None, :])
        out += sv[c0:c1] @ torch.polar(torch.ones_like(ph), ph)
    return out.reshape(n, n).cpu().numpy()


def adjoint_image_rel_l2(kx, ky, s_test, s_ref, fov: float, n: int = 128) -> float:
    a = adjoint_image(kx, ky, s_test, fov, n)
    b = adjoint_image(kx, ky, s_ref, fov, n)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
