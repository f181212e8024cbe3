"""Seeded synthetic P1 tetrahedral phantoms and tissue maps.

The reference reads FE phantoms from XDMF files that are not shipped
(python/feelmri/Phantom.py:76-139, examples/phantoms/README.md); these
generators build structured meshes of the same kind instead: a box grid
whose hexahedra are split into 6 Kuhn tetrahedra (all sharing the hex
diagonal, so the split is conforming), optionally clipped to a region.

Node order is C order over the grid indices (ix, iy, iz), iz fastest.
Tissue is drawn with ``numpy.random.default_rng(seed)`` in node-index order.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Kuhn split of the unit cube: one tetrahedron per axis permutation, each
# walking v000 -> v000+e_a -> v000+e_a+e_b -> v111.
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def _kuhn_corner_offsets() -> np.ndarray:
    offs = np.zeros((6, 4, 3), dtype=np.int64)
    for t, perm in enumerate(_PERMS):
        cur = np.zeros(3, dtype=np.int64)
        offs[t, 0] = cur
        for k, ax in enumerate(perm):
            cur = cur.copy()
            cur[ax] = 1
            offs[t, k + 1] = cur
    return offs


@dataclass
class Mesh:
    nodes: np.ndarray      # (N, 3) float64, metres
    elems: np.ndarray      # (E, 4) int32
    grid_shape: tuple = ()
    spacing: tuple = ()
    info: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.elems.shape[0])


def box_mesh(lo, hi, shape) -> Mesh:
    """Structured Kuhn 6-tet mesh of the box [lo, hi] with ``shape`` nodes per axis."""
    nx, ny, nz = (int(s) for s in shape)
    axes = [np.linspace(lo[d], hi[d], (nx, ny, nz)[d]) for d in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    base = np.stack([ci.ravel(), cj.ravel(), ck.ravel()], axis=1)  # (H,3)
    offs = _kuhn_corner_offsets()                                     # (6,4,3)
    corner = base[:, None, None, :] + offs[None]                      # (H,6,4,3)
    idx = (corner[..., 0] * ny + corner[..., 1]) * nz + corner[..., 2]
    elems = idx.reshape(-1, 4).astype(np.int32)
    spacing = tuple(float((hi[d] - lo[d]) / ((nx, ny, nz)[d] - 1)) for d in range(3))
    return Mesh(nodes=nodes, elems=elems, grid_shape=(nx, ny, nz), spacing=spacing)


def clip_mesh(mesh: Mesh, keep_node: np.ndarray) -> Mesh:
    """Keep elements whose 4 nodes all satisfy ``keep_node``; drop unused nodes."""
    keep_e = keep_node[mesh.elems].all(axis=1)
    elems = mesh.elems[keep_e]
    used = np.zeros(mesh.n_nodes, dtype=bool)
    used[elems.ravel()] = True
    new_id = np.full(mesh.n_nodes, -1, dtype=np.int64)
    new_id[used] = np.arange(int(used.sum()))
    return Mesh(nodes=mesh.nodes[used], elems=new_id[elems].astype(np.int32),
                grid_shape=mesh.grid_shape, spacing=mesh.spacing, info=dict(mesh.info))


def slab_mesh(n_xy: int = 101, n_z: int = 6, size_xy: float = 0.10, size_z: float = 0.01) -> Mesh:
    """0.10 x 0.10 x 0.01 m slab, 101 x 101 x 6 nodes (61 206), 300 000 tets."""
    h = size_xy / 2.0
    hz = size_z / 2.0
    return box_mesh((-h, -h, -hz), (h, h, hz), (n_xy, n_xy, n_z))


def ellipsoid_mesh(n: int = 124, semi_axes=(0.10, 0.08, 0.06)) -> Mesh:
    a = np.asarray(semi_axes, dtype=np.float64)
    m = box_mesh(-a, a, (n, n, n))
    q = ((m.nodes / a) ** 2).sum(axis=1)
    return clip_mesh(m, q <= 1.0 + 1e-12)


def curved_tube_mesh(radius: float = 0.02, length: float = 0.20, n_len: int = 400,
                     n_cross: int = 80, bend_radius: float = 0.15) -> Mesh:
    """Tube of circular cross-section whose centreline is a circular arc.

    A structured grid in (s, u, v) -- arc length and cross-section
    coordinates -- is clipped to the disk u^2+v^2 <= R^2 and mapped onto the
    arc of radius ``bend_radius`` in the x-y plane.
    """
    box = box_mesh((0.0, -radius, -radius), (length, radius, radius), (n_len, n_cross, n_cross))
    s, u, v = box.nodes.T
    keep = u * u + v * v <= radius * radius * (1.0 + 1e-9)
    m = clip_mesh(box, keep)
    s, u, v = m.nodes.T
    ang = s / bend_radius
    rr = bend_radius + u
    x = rr * np.sin(ang) - 0.5 * length
    y = bend_radius - rr * np.cos(ang)
    z = v
    m.info.update(dict(s=s.copy(), u=u.copy(), v=v.copy(), radius=radius, length=length,
                       bend_radius=bend_radius))
    m.nodes = np.stack([x, y, z], axis=1)
    return m


# ---------------------------------------------------------------------------
# Tissue
# ---------------------------------------------------------------------------
@dataclass
class Tissue:
    M0: np.ndarray      # (N,)
    T1: np.ndarray      # (N,) ms
    T2: np.ndarray      # (N,) ms
    df: np.ndarray      # (N,) Hz, off-resonance
    R2s: np.ndarray     # (N,) 1/s  (T2* = 1000 / R2s ms); zeros if unused

    @property
    def delta_B_mT(self) -> np.ndarray:
        """Off-resonance as a field offset in mT (gammabar 42.58 MHz/T)."""
        return self.df / 42.58e6 * 1e3

    @property
    def delta_omega(self) -> np.ndarray:
        """Off-resonance in rad/ms, as ``MRI.Signal``'s ``phi_dB0``."""
        return 2.0 * np.pi * self.df / 1000.0

    @property
    def T2star(self) -> np.ndarray:
        """T2* in ms: 1/(1/T2 + R2s/1000) when R2s is set, else T2."""
        if not np.any(self.R2s):
            return self.T2.copy()
        return 1.0 / (1.0 / self.T2 + self.R2s / 1000.0)


def draw_tissue(n: int, seed: int, df_sigma_hz: float = 20.0, r2s_range=None) -> Tissue:
    """Per-node draws in node order: M0~U[0.8,1], T1~U[500,1500] ms,
    T2~U[40,100] ms, dB0~N(0, df_sigma) Hz, then R2*~U[r2s_range] 1/s."""
    rng = np.random.default_rng(seed)
    M0 = rng.uniform(0.8, 1.0, n)
    T1 = rng.uniform(500.0, 1500.0, n)
    T2 = rng.uniform(40.0, 100.0, n)
    df = rng.normal(0.0, df_sigma_hz, n)
    R2s = rng.uniform(*r2s_range, n) if r2s_range is not None else np.zeros(n)
    return Tissue(M0=M0, T1=T1, T2=T2, df=df, R2s=R2s)


def draw_class_tissue(labels: np.ndarray, n_classes: int, seed: int, df_sigma_hz: float = 20.0,
                      r2s_range=(10.0, 50.0)) -> Tissue:
    """Class-wise M0/T1/T2 (one draw per class, same ranges as draw_tissue);
    per-node dB0 and R2*."""
    rng = np.random.default_rng(seed)
    M0c = rng.uniform(0.8, 1.0, n_classes)
    T1c = rng.uniform(500.0, 1500.0, n_classes)
    T2c = rng.uniform(40.0, 100.0, n_classes)
    n = labels.size
    df = rng.normal(0.0, df_sigma_hz, n)
    R2s = rng.uniform(*r2s_range, n)
    return Tissue(M0=M0c[labels], T1=T1c[labels], T2=T2c[labels], df=df, R2s=R2s)


def sphere_inclusion_labels(nodes: np.ndarray, centres, radii) -> np.ndarray:
    lab = np.zeros(nodes.shape[0], dtype=np.int64)
    for c, (cen, rad) in enumerate(zip(centres, radii), start=1):
        d2 = ((nodes - np.asarray(cen)) ** 2).sum(axis=1)
        lab[d2 <= rad * rad] = c
    return lab
