"""Seeded synthetic scan workloads (meshes, tissue, trajectories, sequences).

c0  0.10 x 0.10 x 0.01 m slab, 101 x 101 x 6 nodes (61 206), 300 000 tets;
    per-node tissue; 2-D spoiled GRE 15 deg, TR 5 ms, 16 dummy TRs, 64 x 64
    Cartesian over FOV 0.24 m, 2 ms readout.
c1  80^3 = 512 000-node box of side 0.2 m with two spherical inclusions
    (3 tissue classes), R2* ~ U[10, 50] 1/s; 3-D Cartesian 128 x 128 x 32,
    3 ms readout.
c2  ~1 M-node ellipsoid, dB0 ~ N(0, 50 Hz); golden-angle radial 256 x 512
    or variable-density spiral 32 x 2048.
c3  curved tube R = 2 cm, L = 20 cm (~2 M nodes), POD motion, 4 encodings,
    Cartesian 96 x 96 x 24 per phase.

Every size can be scaled down (``scale``) for tests; seeds are 1000 + c.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import phantoms, sequences, trajectories


@dataclass
class Workload:
    name: str
    mesh: phantoms.Mesh
    tissue: phantoms.Tissue
    kloc: tuple
    t: np.ndarray
    blocks: list = field(default_factory=list)
    n_enc: int = 1
    info: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return self.mesh.n_nodes

    @property
    def n_samples(self) -> int:
        return int(self.t.size)

    @property
    def node_ksamples(self) -> int:
        return self.n_nodes * self.n_samples * self.n_enc


def c0(scale: float = 1.0) -> Workload:
    n_xy = max(3, int(round(101 * scale)))
    m = phantoms.slab_mesh(n_xy=n_xy, n_z=6)
    tis = phantoms.draw_tissue(m.n_nodes, seed=1000)
    res = max(2, int(round(64 * scale)))
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (res, res, 1), t_first=1.25, t_readout=2.0)
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=16, n_lines=res, rf_dur=0.2)
    return Workload("c0", m, tis, kloc, t, blocks,
                    info=dict(desc=f"2-D spoiled GRE slab, {m.n_nodes} nodes, {res}x{res} Cartesian"))


def c1(scale: float = 1.0, n_kspace=(128, 128, 32)) -> Workload:
    n = max(4, int(round(80 * scale)))
    m = phantoms.box_mesh((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), (n, n, n))
    labels = phantoms.sphere_inclusion_labels(m.nodes, [(-0.04, 0.0, 0.0), (0.045, 0.02, 0.0)], [0.035, 0.025])
    tis = phantoms.draw_class_tissue(labels, 3, seed=1001, df_sigma_hz=20.0, r2s_range=(10.0, 50.0))
    kloc, t = trajectories.cartesian((0.22, 0.22, 0.22), n_kspace, t_first=1.0, t_readout=3.0)
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=16, n_lines=1, rf_dur=0.2)
    return Workload("c1", m, tis, kloc, t, blocks,
                    info=dict(desc=f"3-D Cartesian spoiled GRE, {m.n_nodes}-node P1 box with inclusions, "
                                   f"{'x'.join(map(str, n_kspace))} k-space, 3 ms readout"))


def c2(scale: float = 1.0, kind: str = "radial") -> Workload:
    n = max(6, int(round(124 * scale)))
    m = phantoms.ellipsoid_mesh(n=n)
    tis = phantoms.draw_tissue(m.n_nodes, seed=1002, df_sigma_hz=50.0, r2s_range=(10.0, 50.0))
    if kind == "radial":
        ns, nl = max(2, int(round(512 * scale))), max(2, int(round(256 * scale)))
        kloc, t = trajectories.radial(0.24, ns, nl, t_first=0.5, t_readout=2.56)
    else:
        ns, nl = max(2, int(round(2048 * scale))), max(2, int(round(32 * scale)))
        kloc, t = trajectories.spiral(0.24, 256, ns, nl, density_exponent=2.0, t_first=0.5, t_readout=8.0)
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=16, n_lines=1, rf_dur=0.2)
    return Workload(f"c2-{kind}", m, tis, kloc, t, blocks,
                    info=dict(desc=f"{kind} direct sum, {m.n_nodes}-node ellipsoid, {t.shape[0]}x{t.shape[1]}"))
