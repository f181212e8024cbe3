"""Sequence event tables for the Bloch integrator.

A block is an event table ``(rf (T,) complex mT, G (T,3) mT/m, dt (T,) ms)``
plus the flags ``BlochSolver.solve`` reads (``empty``, ``store``) and its
start time ``t0`` (the motion time shift, python/feelmri/Bloch.py:335-336).

The event grid follows ``SequenceBlock._discretization``
(python/feelmri/Bloch.py:105-129): the union of the RF start/end times and a
``dt_rf`` grid over each pulse, the gradient knots and a ``dt_gr`` grid over
each gradient, and a ``dt`` grid over the block extent (``arange``, so the
extent's end is not on the grid), sorted and de-duplicated; then
``dt = diff(times, prepend=0)`` (Bloch.py:340).  Step i -> i+1 of the
integrator uses the RF/gradient values at ``times[i+1]``.

RF pulses follow the reference's ``RF`` (python/feelmri/MRObjects.py:466-620):
the unit pulse (hard, sinc or apodized sinc, optional phase and frequency
offsets) is tabulated on ``nb_samples`` points over [start, start + dur],
scaled so that gamma times its trapezoid-rule area is the flip angle, and
read by linear interpolation, 0 outside the closed interval (so a pulse is
on at its start and end knots).  ``HardPulse`` is the hard shape with one
constant value, which the device event builder also forms; ``RFPulse``
takes any shape (host tables).  Gradients are trapezoids.  The slice-profile
machinery (MRImaging.py) is host-side sequence design and is not restated.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GAMMA = 42.58e6 * 2.0 * np.pi / 1e6   # rad/(ms*mT), python/feelmri/MRObjects.py:39-40


def _flip_scale(tt: np.ndarray, unit: np.ndarray, flip: float) -> float:
    """|flip / (gamma * trapezoid area of the unit pulse)| on the table grid
    (MRObjects.py:591-600: the step is the grid's first spacing)."""
    area = np.sum((unit[1:] + unit[:-1]) / 2) * (tt[1] - tt[0]) * GAMMA
    return float(np.abs(flip / area))


@dataclass
class HardPulse:
    """The reference's RF with ``shape='hard'`` and no frequency offset:
    B1 = scale * exp(i phase) on [start, start + dur], 0 outside."""
    start: float        # ms
    dur: float          # ms
    flip: float         # rad
    phase: float = 0.0  # rad
    nb_samples: int = 200

    @property
    def b1(self) -> float:
        tt = np.linspace(self.start, self.start + self.dur, self.nb_samples)
        return _flip_scale(tt, np.ones(tt.size), self.flip)

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        on = (t >= self.start) & (t <= self.start + self.dur)
        return np.where(on, self.b1 * np.exp(1j * self.phase), 0.0 + 0.0j)


@dataclass
class RFPulse:
    """The reference's ``RF`` (MRObjects.py:466-620) for ``time - ref =
    start``: shape ``"apodized_sinc"`` (default, alpha 0.46), ``"sinc"``
    (alpha 0) or ``"hard"``; ``n_lobes`` = (left, right) lobes; phase (rad)
    and frequency offset (Hz) applied as exp(i (phase + 2 pi f t')), t' the
    time from the main lobe's centre.

    Unit pulse: with bw = (n_left + n_right + 2) / dur and t' = t - start -
    (n_left + 1) / bw, (1 / bw) ((1 - alpha) + alpha cos(pi bw t' / N))
    sinc(bw t'), N = max lobes, on [start, start + dur].  Tabulated on
    ``nb_samples`` points, scaled to the flip angle and interpolated
    linearly, 0 outside."""
    start: float
    dur: float
    flip: float
    shape: str = "apodized_sinc"
    n_lobes: tuple = (2, 2)
    alpha: float = 0.46
    nb_samples: int = 200
    phase: float = 0.0
    freq_offset: float = 0.0   # Hz

    def __post_init__(self):
        if self.shape not in ("apodized_sinc", "sinc", "hard"):
            raise ValueError(f"unknown RF shape {self.shape!r}")
        if self.shape == "sinc":
            self.alpha = 0.0
        self._t = np.linspace(self.start, self.start + self.dur, self.nb_samples)
        unit = self.unit(self._t)
        scale = _flip_scale(self._t, unit, self.flip)
        self._re, self._im = scale * unit.real, scale * unit.imag

    def unit(self, t):
        """The unscaled pulse (MRObjects.py:522-589)."""
        t = np.asarray(t, dtype=np.float64)
        nl, nr = self.n_lobes
        half1 = (nl + 1) / (nl + nr + 2) * self.dur
        ts = t - self.start - half1
        window = (t >= self.start) & (t <= self.start + self.dur)
        if self.shape == "hard":
            b = np.where(window, 1.0, 0.0)
        else:
            bw = (nl + nr + 2) / self.dur
            n = max(nl, nr)
            b = (1.0 / bw) * ((1 - self.alpha) + self.alpha * np.cos(np.pi * bw * ts / n)) * np.sinc(bw * ts)
            b = b * window
        b = b + 0j
        if self.phase != 0.0 or self.freq_offset != 0.0:
            b = b * np.exp(1j * (self.phase + 2 * np.pi * (self.freq_offset * 1e-3) * ts))
        return b

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        return np.interp(t, self._t, self._re, left=0.0, right=0.0) + \
            1j * np.interp(t, self._t, self._im, left=0.0, right=0.0)


@dataclass
class Trapezoid:
    axis: int           # 0, 1, 2
    start: float        # ms
    amp: float          # mT/m
    ramp: float         # ms
    flat: float         # ms

    @property
    def timings(self) -> np.ndarray:
        s = self.start
        return np.array([s, s + self.ramp, s + self.ramp + self.flat, s + 2 * self.ramp + self.flat])

    @property
    def amplitudes(self) -> np.ndarray:
        return np.array([0.0, self.amp, self.amp, 0.0])

    def __call__(self, t):
        return np.interp(np.asarray(t, dtype=np.float64), self.timings, self.amplitudes,
                         left=0.0, right=0.0)


@dataclass
class Block:
    times: np.ndarray
    rf: np.ndarray
    G: np.ndarray
    dt: np.ndarray
    empty: bool = False
    store: bool = False
    t0: float = 0.0
    info: dict = field(default_factory=dict)

    @property
    def n_events(self) -> int:
        return int(self.times.size)

    def as_dict(self) -> dict:
        return dict(rf=self.rf, G=self.G, dt=self.dt, empty=self.empty, store=self.store,
                    t0=self.t0)


def make_block(t0: float, dur: float, pulses=(), gradients=(), dt_rf: float = 0.01,
               dt_gr: float = -1.0, dt: float = 10.0, empty: bool = False,
               store: bool = False) -> Block:
    """Discretise one block starting at ``t0`` (ms) with extent ``dur``."""
    parts = []
    if gradients:
        parts.append(np.concatenate([g.timings for g in gradients]))
        if dt_gr > 0:
            parts += [np.arange(g.timings[0], g.timings[-1], dt_gr) for g in gradients]
    if pulses:
        parts.append(np.concatenate([[p.start, p.start + p.dur] for p in pulses]))
        if dt_rf > 0:
            parts += [np.arange(p.start, p.start + p.dur, dt_rf) for p in pulses]
    ext0 = min([t0] + [p.start for p in pulses] + [g.timings[0] for g in gradients])
    ext1 = max([t0 + dur] + [p.start + p.dur for p in pulses] + [g.timings[-1] for g in gradients])
    parts.append(np.arange(ext0, ext1, dt))
    times = np.unique(np.sort(np.concatenate(parts)))
    dts = np.diff(times, prepend=0.0)
    rf = np.zeros(times.size, dtype=np.complex128)
    for p in pulses:
        rf += p(times)
    G = np.zeros((times.size, 3), dtype=np.float64)
    for g in gradients:
        G[:, g.axis] += g(times)
    spec = dict(t0=float(t0), dur=float(dur), dt=float(dt), dt_rf=float(dt_rf), dt_gr=float(dt_gr),
                pulses=list(pulses), gradients=list(gradients))
    return Block(times=times, rf=rf, G=G, dt=dts, empty=empty, store=store, t0=ext0,
                 info=dict(extent=(ext0, ext1), spec=spec))


def block_descriptions(blocks):
    """The blocks' specifications (``make_block``'s arguments, kept in
    ``Block.info["spec"]``) as the records of the device event builder:
    (desc, pulses, trapezoids, tables) -- numpy structured arrays in the C
    layouts of fm_block_desc, fm_rf_pulse and fm_trapezoid, and the
    (points, 2) float64 table points of the tabulated pulses.  A
    ``HardPulse`` is one constant value; an ``RFPulse`` is its interpolation
    table.  Store columns are numbered over the storing blocks in order; a
    non-empty block spoils."""
    from .device import BLOCK_DESC, RF_PULSE, TRAPEZOID

    desc = np.zeros(len(blocks), dtype=BLOCK_DESC)
    pl, tl, tabs = [], [], []
    n_tab = 0
    col = 0
    for b, blk in enumerate(blocks):
        spec = blk.info.get("spec")
        if spec is None:
            raise ValueError(f"block {b} has no specification (build blocks with make_block)")
        d = desc[b]
        d["t0"], d["dur"], d["dt"], d["dt_rf"], d["dt_gr"] = (spec["t0"], spec["dur"], spec["dt"], spec["dt_rf"],
                                                             spec["dt_gr"])
        d["pulse0"], d["n_pulses"] = len(pl), len(spec["pulses"])
        d["trap0"], d["n_traps"] = len(tl), len(spec["gradients"])
        d["spoil"] = 0 if blk.empty else 1
        d["store_col"] = col if blk.store else -1
        col += int(blk.store)
        for p in spec["pulses"]:
            if isinstance(p, HardPulse):
                amp = complex(p.b1 * np.exp(1j * p.phase))   # the value HardPulse.__call__ forms
                pl.append((p.start, p.dur, amp.real, amp.imag, 0, 0, 0))
            elif isinstance(p, RFPulse):
                tabs.append(np.stack([p._re, p._im], axis=1))
                pl.append((p.start, p.dur, 0.0, 0.0, n_tab, p._re.size, 0))
                n_tab += p._re.size
            else:
                raise NotImplementedError(f"block {b}: the device event builder takes HardPulse and RFPulse, "
                                          f"not {type(p).__name__}")
        for g in spec["gradients"]:
            tl.append((g.start, g.ramp, g.flat, g.amp, g.axis, 0))
    return (desc, np.array(pl, dtype=RF_PULSE), np.array(tl, dtype=TRAPEZOID),
            np.concatenate(tabs) if tabs else np.zeros((0, 2)))


def delay_block(t0: float, dur: float, dt: float = 1.0) -> Block:
    """Empty (relaxation-only) block, as ``Sequence.add_block(Quantity)``
    builds it (Bloch.py:195-204)."""
    return make_block(t0, dur, dt=dt, empty=True)


def spoiled_gre(flip_deg: float = 15.0, TR: float = 5.0, n_dummy: int = 16, n_lines: int = 64,
                rf_dur: float = 0.2, dt_rf: float = 0.01, dt_delay: float = 1.0,
                phase_cycle: bool = False) -> list[Block]:
    """Spoiled gradient echo: per TR one imaging block (hard pulse) and one
    delay block.  Dummy TRs are not stored; each of the ``n_lines`` imaging
    blocks stores its post-pulse Mxy (``store_magnetization``).  Spoiling is
    the reference's artificial one: Mxy is zeroed after every non-empty block
    (Bloch.py:363-368)."""
    blocks: list[Block] = []
    t = 0.0
    flip = np.deg2rad(flip_deg)
    for k in range(n_dummy + n_lines):
        ph = (np.pi * k * k / 2.0) if phase_cycle else 0.0
        pulse = HardPulse(start=t, dur=rf_dur, flip=flip, phase=ph)
        blk = make_block(t, rf_dur, pulses=[pulse], dt_rf=dt_rf, dt=1.0,
                         store=(k >= n_dummy))
        blocks.append(blk)
        t_end = blk.info["extent"][1]
        blocks.append(delay_block(t_end, TR - rf_dur, dt=dt_delay))
        t = t + TR
    return blocks


def ernst_steady_state(flip_deg: float, TR_eff: float, T1):
    """Ideal-spoiling FLASH steady state just before a pulse:
    Mz- = M0 (1 - E1) / (1 - E1 cos a), E1 = exp(-TR_eff / T1)."""
    a = np.deg2rad(flip_deg)
    E1 = np.exp(-TR_eff / np.asarray(T1, dtype=np.float64))
    return (1.0 - E1) / (1.0 - E1 * np.cos(a))
