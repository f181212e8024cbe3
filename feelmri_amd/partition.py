"""Element partitioning of a P1 mesh into node-sharded parts.

The reference distributes its phantom over MPI ranks by partitioning the
element dual graph with METIS and giving each rank the nodes of its elements
(python/feelmri/Phantom.py:194-239).  Nodes on a part boundary are
duplicated: every part assembles its own lumped weights from its own
elements only, so a boundary node carries a partial weight in each part.
Because the k-space sum is linear in the weights, the per-part k-spaces add
up to the k-space of the whole mesh; the reference adds them with one
``Reduce(SUM)`` (python/feelmri/MPIUtilities.py:13-26).

METIS is not available here, so two partitioners of this module's own
design take its place:

* ``"blocks"`` -- contiguous element ranges of equal size.  The synthetic
  meshes number their elements along the grid, so a range is a slab.
* ``"rcb"`` -- recursive coordinate bisection of the element centroids:
  split the longest extent at the weighted median, recurse until there
  are ``n_parts`` parts.  Parts are balanced to one element and compact,
  which keeps the duplicated boundary small.

Both return a membership array like METIS's (part index per element), and
``local_mesh`` extracts a part exactly as ``distribute_mesh`` does: the
part's nodes in ascending global order, and its elements renumbered to them.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _blocks(n_elems: int, n_parts: int) -> np.ndarray:
    bounds = (np.arange(n_parts + 1, dtype=np.int64) * n_elems) // n_parts
    member = np.empty(n_elems, dtype=np.int32)
    for p in range(n_parts):
        member[bounds[p]:bounds[p + 1]] = p
    return member


def _rcb(centroids: np.ndarray, n_parts: int) -> np.ndarray:
    member = np.empty(centroids.shape[0], dtype=np.int32)
    # work list of (element indices, first part id, number of parts)
    todo = [(np.arange(centroids.shape[0]), 0, n_parts)]
    while todo:
        idx, first, k = todo.pop()
        if k == 1 or idx.size == 0:
            member[idx] = first
            continue
        pts = centroids[idx]
        axis = int(np.argmax(pts.max(axis=0) - pts.min(axis=0))) if idx.size else 0
        k_lo = k // 2
        n_lo = (idx.size * k_lo) // k
        # stable order along the axis so ties split deterministically
        order = np.argsort(pts[:, axis], kind="stable")
        todo.append((idx[order[n_lo:]], first + k_lo, k - k_lo))
        todo.append((idx[order[:n_lo]], first, k_lo))
    return member


def partition_elements(elems: np.ndarray, nodes: np.ndarray, n_parts: int,
                       method: str = "rcb") -> np.ndarray:
    """Part index (int32, 0..n_parts-1) of every element."""
    elems = np.asarray(elems)
    if elems.ndim != 2 or elems.shape[1] != 4:
        raise ValueError(f"elems must be (E, 4) P1 tetrahedra, got {elems.shape}")
    if n_parts < 1:
        raise ValueError("n_parts must be >= 1")
    E = elems.shape[0]
    if method == "blocks":
        return _blocks(E, n_parts)
    if method == "rcb":
        cen = np.asarray(nodes, dtype=np.float64)[elems].mean(axis=1)
        return _rcb(cen, n_parts)
    raise ValueError(f"unknown partition method {method!r} (blocks, rcb)")


@dataclass
class LocalMesh:
    """One part: ``nodes_l2g`` maps local to global node indices (ascending),
    ``elems_l2g`` local to global element indices, and ``elems`` is the
    part's connectivity in local node numbers."""
    part: int
    nodes_l2g: np.ndarray
    elems_l2g: np.ndarray
    elems: np.ndarray

    @property
    def n_nodes(self) -> int:
        return int(self.nodes_l2g.size)

    @property
    def n_elems(self) -> int:
        return int(self.elems_l2g.size)


def local_mesh(elems: np.ndarray, membership: np.ndarray, part: int) -> LocalMesh:
    """The elements of ``part`` and the nodes they touch, renumbered."""
    elems = np.asarray(elems)
    e_idx = np.flatnonzero(np.asarray(membership) == part)
    le = elems[e_idx]
    n_idx = np.unique(le.ravel())
    g2l = np.full(int(elems.max()) + 1 if elems.size else 0, -1, dtype=np.int64)
    g2l[n_idx] = np.arange(n_idx.size)
    return LocalMesh(part=part, nodes_l2g=n_idx.astype(np.int64), elems_l2g=e_idx.astype(np.int64),
                     elems=g2l[le].astype(np.int32) if le.size else np.zeros((0, 4), np.int32))


def split_mesh(elems: np.ndarray, nodes: np.ndarray, n_parts: int, method: str = "rcb") -> list[LocalMesh]:
    """All parts of a partition (see ``partition_elements``)."""
    member = partition_elements(elems, nodes, n_parts, method)
    return [local_mesh(elems, member, p) for p in range(n_parts)]


def duplication(parts: list[LocalMesh], n_nodes: int) -> float:
    """Stored nodes over distinct nodes: 1.0 for one part, above 1 by the
    share of boundary copies."""
    return float(sum(p.n_nodes for p in parts)) / max(1, n_nodes)
