"""Drop-in for the reference's ``feelmri.POD`` extension module.

``tensordot_modes_weights(modes, weights)`` computes (P, C, M) . (M,) ->
(P, C) like cpp/feelmri/POD.cpp:5-38.  As in the reference binding
(POD.h:17-23, ``.noconvert()``), both arrays must already be C-contiguous
and of the same dtype, float32 or float64; anything else raises TypeError.
"""
from __future__ import annotations

import numpy as np
import torch

from . import device, resident


def tensordot_modes_weights(modes, weights, *, dev=None):
    if not isinstance(modes, np.ndarray) or not isinstance(weights, np.ndarray):
        raise TypeError("tensordot_modes_weights(): incompatible function arguments (numpy arrays required)")
    if modes.dtype != weights.dtype or modes.dtype not in (np.float32, np.float64):
        raise TypeError("tensordot_modes_weights(): incompatible function arguments "
                        f"(modes {modes.dtype}, weights {weights.dtype}; both float32 or both float64)")
    if not (modes.flags.c_contiguous and weights.flags.c_contiguous):
        raise TypeError("tensordot_modes_weights(): arrays must be C-contiguous")
    if modes.ndim != 3 or weights.ndim != 1 or weights.shape[0] != modes.shape[2]:
        raise ValueError(f"shapes {modes.shape} and {weights.shape} do not contract")
    dev = torch.device(dev) if dev is not None else device.default_device()
    tdt = torch.float32 if modes.dtype == np.float32 else torch.float64
    with torch.cuda.device(dev):
        # the caller passes the same modes at every time step
        # (python/feelmri/Motion.py:254, :290): resident by content
        out = device.pod_tensordot(resident.to_device(modes, tdt, dev), torch.from_numpy(weights).to(dev))
        return out.cpu().numpy()
