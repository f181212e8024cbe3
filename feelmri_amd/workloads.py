"""Seeded synthetic scan workloads (meshes, tissue, trajectories, sequences).

c0  0.10 x 0.10 x 0.01 m slab, 101 x 101 x 6 nodes (61 206), 300 000 tets;
    per-node tissue; 2-D spoiled GRE 15 deg, TR 5 ms, 16 dummy TRs, 64 x 64
    Cartesian over FOV 0.24 m, 2 ms readout.
c1  80^3 = 512 000-node box of side 0.2 m with two spherical inclusions
    (3 tissue classes), R2* ~ U[10, 50] 1/s; 3-D Cartesian 128 x 128 x 32,
    3 ms readout.
c2  ~1 M-node ellipsoid, dB0 ~ N(0, 50 Hz); golden-angle radial 256 x 512
    or variable-density (p = 2) spiral stack 32 x 2048 under the default
    scanner's gradient limits (res 256 over FOV 0.24 m: T_ro 1.82 ms; the
    receiver bandwidth is set to 1024 / T_ro so that the readout holds 2048
    samples at oversampling 2).
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


@dataclass
class FlowWorkload(Workload):
    """c3: per cardiac phase, one Bloch block per velocity encoding (bipolar
    gradients, nodes moving with the Poiseuille velocity), then one Signal
    call with all encodings and the POD displacement as motion."""
    n_phases: int = 20
    period: float = 1000.0                 # ms
    venc: float = 1.5                      # m/s
    motion: object = None                  # ModalMotion displacement, R modes
    velocity: object = None                # ModalMotion velocity (Poiseuille field)
    enc_blocks: list = field(default_factory=list)   # one Block per encoding

    @property
    def node_ksamples(self) -> int:
        return self.n_nodes * self.n_samples * self.n_enc * self.n_phases


def _smooth_modes(nodes: np.ndarray, n_modes: int, rng, amp: float = 5e-4, scale: float = 0.05):
    """Random smooth displacement fields: one low-frequency sinusoid of
    position per component and mode, amplitude ``amp`` metres."""
    P = nodes.shape[0]
    phi = np.empty((P, 3, n_modes))
    for r in range(n_modes):
        for c in range(3):
            k = rng.normal(0, 1.0 / scale, 3)
            phi[:, c, r] = amp * np.sin(2 * np.pi * (nodes @ k) * 0.1 + rng.uniform(0, 2 * np.pi))
    return phi


def encoding_block(enc: dict, d: int, t0: float, store: bool = True) -> sequences.Block:
    """Phase-contrast imaging block starting at t0 (ms): a hard pulse, then for
    encoding d > 0 a bipolar gradient pair along axis d-1 (VENC set by the
    lobe amplitude ``gmax``); d = 0 is the reference encoding."""
    pulse = sequences.HardPulse(start=t0, dur=enc["rf_dur"], flip=np.deg2rad(enc["flip_deg"]))
    grads = []
    if d > 0:
        grads = [sequences.Trapezoid(d - 1, t0 + 0.25, enc["gmax"], enc["ramp"], enc["flat"]),
                 sequences.Trapezoid(d - 1, t0 + 0.25 + enc["sep"], -enc["gmax"], enc["ramp"], enc["flat"])]
    return sequences.make_block(t0, enc["rf_dur"], pulses=[pulse], gradients=grads, dt_rf=0.01, dt_gr=0.01,
                                dt=1.0, store=store)


def flow_sequence(wl: "FlowWorkload", n_dummy: int = 8, TR: float = 5.0, n_phases: int | None = None) -> list:
    """The 4-D flow acquisition as Bloch blocks (like examples/4dflow.py:
    dummy TRs to steady state, then per cardiac phase one stored imaging block
    per encoding, TR apart, and a delay to the next phase).  Stored columns
    come out in (phase, encoding) order."""
    enc = wl.info["encoding"]
    n_phases = wl.n_phases if n_phases is None else n_phases
    spacing = wl.period / wl.n_phases
    blocks, t = [], 0.0
    for k in range(n_dummy):
        b = encoding_block(enc, k % 4, t, store=False)
        blocks += [b, sequences.delay_block(b.info["extent"][1], TR - (b.info["extent"][1] - t), dt=1.0)]
        t += TR
    for p in range(n_phases):
        start = n_dummy * TR + p * spacing
        if start > t:
            blocks.append(sequences.delay_block(t, start - t, dt=1.0))
            t = start
        for d in range(4):
            b = encoding_block(enc, d, t, store=True)
            blocks += [b, sequences.delay_block(b.info["extent"][1], TR - (b.info["extent"][1] - t), dt=1.0)]
            t += TR
    return blocks


def c3(scale: float = 1.0, n_phases: int = 20, n_modes: int = 10, kspace=(96, 96, 24)) -> FlowWorkload:
    from . import Motion

    n_len = max(8, int(round(400 * scale)))
    n_cross = max(6, int(round(80 * scale)))
    m = phantoms.curved_tube_mesh(radius=0.02, length=0.20, n_len=n_len, n_cross=n_cross)
    rng = np.random.default_rng(1003)
    tis = phantoms.draw_tissue(m.n_nodes, seed=1003, r2s_range=(10.0, 50.0))
    period = 1000.0
    times = np.linspace(0.0, period, n_phases + 1)[:-1]
    # Poiseuille velocity along the arc tangent, peak 1 m/s, cardiac waveform
    s, u, v = m.info["s"], m.info["u"], m.info["v"]
    rho2 = (u * u + v * v) / m.info["radius"] ** 2
    ang = s / m.info["bend_radius"]
    tangent = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], axis=1)
    wave = np.maximum(0.0, np.sin(2 * np.pi * times / period)) ** 2 + 0.1
    speed = (1.0 - np.clip(rho2, 0, 1))[:, None] * wave[None, :]          # m/s
    vel = tangent[:, :, None] * speed[:, None, :] * 1e-3                    # m/ms
    # the field is one spatial profile times a waveform: one mode holds it
    velocity = Motion.ModalMotion(times, vel, rank=1, periodic=True, velocity=True)
    # node displacement from random smooth modes with periodic weights
    base = _smooth_modes(m.nodes, n_modes, rng)
    n_modes = min(n_modes, max(1, n_phases // 2))  # harmonics below Nyquist of the phase grid
    base = base[:, :, :n_modes]
    h = np.arange(1, (n_modes + 1) // 2 + 1)  # enough harmonics for rank n_modes in time
    coef = rng.normal(size=(n_modes, 2, h.size))
    a = (coef[:, 0, :, None] * np.cos(2 * np.pi * h[None, :, None] * times / period) +
         coef[:, 1, :, None] * np.sin(2 * np.pi * h[None, :, None] * times / period)).sum(1)
    disp = np.tensordot(base, a, axes=([2], [0]))
    motion = Motion.ModalMotion(times, disp, rank=n_modes, periodic=True)
    # encodings: reference + bipolar along x, y, z; VENC 1.5 m/s
    venc = 1.5
    ramp, flat = 0.1, 0.4
    area_lobe = flat + ramp                       # ms, per unit amplitude
    sep = flat + 2 * ramp                         # lobe centre separation
    gmax = np.pi / (sequences.GAMMA * area_lobe * sep * venc * 1e-3)   # mT/m
    enc = dict(gmax=gmax, ramp=ramp, flat=flat, sep=sep, flip_deg=15.0, rf_dur=0.2)
    enc_blocks = [encoding_block(enc, d, 0.0) for d in range(4)]
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.06), kspace, t_first=1.0, t_readout=2.0)
    wl = FlowWorkload("c3", m, tis, kloc, t, n_enc=4, n_phases=n_phases, period=period, venc=venc,
                      motion=motion, velocity=velocity, enc_blocks=enc_blocks,
                      info=dict(encoding=enc, desc=f"4-D flow, {m.n_nodes}-node curved tube, {n_modes} POD modes, 4 encodings, "
                                     f"{'x'.join(map(str, kspace))} x {n_phases} phases, per-readout GEMM with moving nodes"))
    return wl


def c2(scale: float = 1.0, kind: str = "radial") -> Workload:
    n = max(6, int(round(124 * scale)))
    m = phantoms.ellipsoid_mesh(n=n)
    tis = phantoms.draw_tissue(m.n_nodes, seed=1002, df_sigma_hz=50.0, r2s_range=(10.0, 50.0))
    if kind == "radial":
        ns, nl = max(2, int(round(512 * scale))), max(2, int(round(256 * scale)))
        kloc, t = trajectories.radial(0.24, ns, nl, t_first=0.5, t_readout=2.56)
    else:
        ns, nl = max(2, int(round(2048 * scale))), max(2, int(round(32 * scale)))
        fov, res = (0.24, 0.24, 0.24), (256, nl, 1)
        # readout duration of the base spiral, then the bandwidth that gives ns samples
        _, _, T_ms = trajectories.spiral_stack(fov, res, density_exponent=2.0, dtype=np.float64)
        kloc, t, _ = trajectories.spiral_stack(fov, res, oversampling=2, t_start=0.5,
                                               receiver_bw=(ns // 2) / (T_ms * 1e-3), density_exponent=2.0,
                                               dtype=np.float64)
        assert t.shape[0] == 2 * (ns // 2), t.shape
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=16, n_lines=1, rf_dur=0.2)
    return Workload(f"c2-{kind}", m, tis, kloc, t, blocks,
                    info=dict(desc=f"{kind} signal sum, {m.n_nodes}-node ellipsoid, {t.shape[0]}x{t.shape[1]}"))
