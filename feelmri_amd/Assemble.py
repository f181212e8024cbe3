"""Drop-in for the reference's ``feelmri.Assemble`` extension module.

``MassAssemble(elems, nodes, finite_element, quadrature_rule)`` returns the
consistent mass matrix as scipy CSC, like the reference
(cpp/feelmri/Assemble.h:21-39, Assemble.cpp:78-116).  Only P1 tetrahedra
are supported: their element matrices |T|/20 (1 + delta_ij) are exact for
the order-2 rule the reference uses, so ``finite_element`` and
``quadrature_rule`` are checked (4 dofs) and otherwise unused.  The element
matrices come from the GPU; the sparse structure is assembled by scipy.

``lumped_mass(elems, nodes)`` is ``FEMPhantom.mass_matrix(lumped=True)``
(python/feelmri/Phantom.py:379-399): the row sums, computed directly on the
GPU by the coloured scatter, as a diagonal sparse matrix (or the weight
vector with ``as_vector=True``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import device


def _check_p1(elems, finite_element):
    el = np.asarray(elems)
    if el.ndim != 2 or el.shape[1] != 4:
        raise TypeError(f"only P1 tetrahedra (elems of shape (E, 4)) are supported, got {el.shape}")
    dim = getattr(finite_element, "dimension", 4) if finite_element is not None else 4
    if int(dim) != 4:
        raise TypeError(f"only the P1 tetrahedron (4 dofs) is supported, element has {dim}")


def MassAssemble(elems, nodes, finite_element=None, quadrature_rule=None, *, dev=None):
    _check_p1(elems, finite_element)
    nodes = np.asarray(nodes)
    if nodes.dtype not in (np.float32, np.float64):
        raise TypeError("nodes must be float32 or float64")
    dev = torch.device(dev) if dev is not None else device.default_device()
    el = np.ascontiguousarray(elems, dtype=np.int32)
    N = nodes.shape[0]
    if el.size and (el.min() < 0 or el.max() >= N):
        raise ValueError(f"element connectivity references nodes outside [0, {N})")
    with torch.cuda.device(dev):
        vals = device.element_matrices(torch.from_numpy(el).to(dev),
                                       torch.from_numpy(np.ascontiguousarray(nodes)).to(dev))
        v = vals.cpu().numpy().astype(nodes.dtype, copy=False)
    rows = np.repeat(el, 4, axis=1).ravel()
    cols = np.tile(el, (1, 4)).ravel()
    return sp.csc_matrix((v.ravel(), (rows, cols)), shape=(N, N))


def lumped_mass(elems, nodes, *, as_vector: bool = False, dev=None):
    _check_p1(elems, None)
    nodes = np.asarray(nodes)
    dev = torch.device(dev) if dev is not None else device.default_device()
    with torch.cuda.device(dev):
        col = device.P1Colouring(elems, nodes.shape[0], device=dev)
        nd = torch.from_numpy(np.ascontiguousarray(nodes, dtype=np.float64)).to(dev)
        w = device.lumped_mass(col, nd).cpu().numpy()
    if as_vector:
        return w
    return sp.diags(w, 0, format="csr")
