"""Image-domain reconstruction and error of k-space results (measurement
helpers beside the signal path; they run on the GPU through torch.fft when
one is visible, else on the CPU).

* ``ktoi``: the centred inverse FFT of python/feelmri/Math.py:10-14,
  fftshift(ifftn(ifftshift(x))) over the first three axes.
* ``tukey_filter`` / ``riesz_filter``: the k-space apodisation windows of
  python/feelmri/Filters.py:5-51.
* ``cartesian_recon``: CartesianRecon of python/feelmri/Recon.py:13-90 for
  stacks with one line per shot: zero-pad each even dimension by one sample,
  weight by the outer product of the readout and line windows, ``ktoi``.
  The reference's final skimage ``resize`` to the oversampled resolution is
  not reproduced (skimage is absent); it is linear and acts alike on both
  images of an error comparison.
* ``image_rel_l2``: relative L2 between the reconstructions of two k-space
  results.
* ``adjoint_image``: the direct adjoint DFT for radial / spiral samples.
"""
from __future__ import annotations

import numpy as np


def _device(device=None):
    import torch

    if device is not None:
        return torch.device(device)
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def ktoi(x, axes=None, device=None) -> np.ndarray:
    """Centred inverse FFT over ``axes`` (default: the first three)."""
    import torch

    x = np.asarray(x, dtype=np.complex128)
    if axes is None:
        axes = [i for i in range(x.ndim) if i < 3]
    t = torch.as_tensor(x, device=_device(device))
    t = torch.fft.ifftshift(t, dim=axes)
    t = torch.fft.ifftn(t, dim=axes)
    t = torch.fft.fftshift(t, dim=axes)
    return t.cpu().numpy()


def tukey_filter(size: int, width: float = 0.6, lift: float = 0.7) -> np.ndarray:
    """Filters.py:34-51: a Tukey window with taper fraction 1 - width,
    lifted to a floor of ``lift``."""
    from scipy.signal.windows import tukey

    return tukey(size, alpha=1.0 - width) * (1.0 - lift) + lift


def riesz_filter(size: int, width: float = 0.6, lift: float = 0.7) -> np.ndarray:
    """Filters.py:5-31: quadratic roll-off to ``lift`` over the outer
    round((1 - width) / 2 * size) samples at each end."""
    s20 = int(np.round((1.0 - width) / 2 * size))
    h = np.ones(size)
    if s20 > 0:
        s1 = np.linspace(0, s20 / 2, s20)
        w1 = 1.0 - np.power(np.abs(s1 / (s20 / 2)), 2) * (1.0 - lift)
        h[0:s20] *= np.flip(w1)
        h[size - s20:size] *= w1
    return h


def cartesian_recon(K, filter_type: str | None = "Tukey", device=None) -> np.ndarray:
    """Image of Cartesian k-space ``K`` (n_meas, n_lines, n_kz[, ...]) as
    CartesianRecon forms it (Recon.py:13-90), before the final resize.  The
    windows use width 0.9 and lift 0.3 whatever the filter dict says, as the
    reference does (Recon.py:64-69)."""
    K = np.asarray(K, dtype=np.complex128)
    while K.ndim < 5:
        K = K[..., None]
    pad = [(0, 1 if K.shape[a] % 2 == 0 else 0) if a < 3 else (0, 0) for a in range(K.ndim)]
    K = np.pad(K, pad, mode="constant")
    if filter_type == "Tukey":
        h0, h1 = tukey_filter(K.shape[0], 0.9, 0.3), tukey_filter(K.shape[1], 0.9, 0.3)
    elif filter_type == "Riesz":
        h0, h1 = riesz_filter(K.shape[0], 0.9, 0.3), riesz_filter(K.shape[1], 0.9, 0.3)
    else:
        h0, h1 = np.ones(K.shape[0]), np.ones(K.shape[1])
    H = np.outer(h0, h1)[:, :, None, None, None]
    return ktoi(H * K, [0, 1, 2], device=device)


def image_rel_l2(k_test, k_ref, filter_type: str | None = None, device=None) -> float:
    """Relative L2 between the images of two Cartesian k-space arrays:
    plain ``ktoi`` (filter_type None) or ``cartesian_recon`` with a window."""
    if filter_type is None:
        a, b = ktoi(k_test, device=device), ktoi(k_ref, device=device)
    else:
        a, b = cartesian_recon(k_test, filter_type, device), cartesian_recon(k_ref, filter_type, device)
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

    dev = _device(device)
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
