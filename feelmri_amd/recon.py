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
