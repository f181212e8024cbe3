"""Serving loop: continuous batching of synthetic scan requests on one GPU.

A request is one synthetic scan: a phantom of N nodes, a spoiled-GRE Bloch
solve to steady state, lumped-mass weighting and the k-space signal of its
trajectory in FP64 or mixed precision.  The loop

* admits requests in arrival order while their HBM footprint fits the
  budget (288 GB by default) and a stream slot is free; a request larger
  than the whole budget is rejected;
* keeps phantoms (mesh, tissue, lumped weights; 9 size classes) and
  trajectory plans resident across requests, and each admitted request's
  packed nodes and output resident until it completes;
* runs each request's spoiled-GRE steady state as the replay of a HIP graph
  captured once per phantom class and stream slot (``SteadyStateGraph``);
* cuts each request's signal into chunks of readout rows and interleaves the
  chunks of all active requests over their own HIP streams, so a request
  that arrives while others run starts at the next chunk boundary and its
  work overlaps theirs on the device;
* records latency from arrival to the result on the host, and aggregate
  node*k-sample/s.

Chunk launch shapes depend only on the chunk's plan and the request's
sizes, and every reduction runs in a fixed order, so a request's result is
bitwise the same alone and in any mix.

The loop is written against a small backend interface (``start``,
``step``, ``done``, ``finish``, ``footprint``); ``GpuBackend`` is the one
that runs on the device.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class ScanRequest:
    rid: int
    arrival: float          # s from the start of the stream
    n_nodes: int
    traj: str               # "cart2d" | "cart3d" | "radial"
    shape: tuple            # (n_meas, n_lines, n_kz)
    precision: str          # "f64" | "mixed"
    seed: int = 0

    @property
    def n_samples(self) -> int:
        return int(np.prod(self.shape))

    @property
    def work(self) -> int:
        return self.n_nodes * self.n_samples


def poisson_stream(n_requests: int, rate: float = 20.0, seed: int = 1004, n_min: int = 10_000,
                   n_max: int = 2_000_000, precisions=("f64", "mixed")) -> list[ScanRequest]:
    """Mixed-size synthetic scan requests with Poisson arrivals (rate per s).

    N log-uniform in [n_min, n_max]; trajectory Cartesian 2-D (64^2..256^2),
    Cartesian 3-D (64^2..256^2 x 32) or radial (64..512 spokes x 256..512
    samples); FP64 or mixed precision.
    """
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = []
    for i in range(n_requests):
        N = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        kind = rng.choice(["cart2d", "cart3d", "radial"])
        if kind == "cart2d":
            r = int(rng.choice([64, 128, 192, 256]))
            shape = (r, r, 1)
        elif kind == "cart3d":
            r = int(rng.choice([64, 128, 192, 256]))
            shape = (r, r, 32)
        else:
            shape = (int(rng.choice([256, 384, 512])), int(rng.choice([64, 128, 256, 512])), 1)
        out.append(ScanRequest(i, float(t[i]), N, str(kind), shape, str(rng.choice(list(precisions))),
                               seed=int(rng.integers(1 << 30))))
    return out


@dataclass
class RequestResult:
    rid: int
    latency: float
    start: float
    end: float
    kspace: np.ndarray | None
    work: int
    error: str | None = None


@dataclass
class ServeReport:
    results: dict = field(default_factory=dict)
    rejected: list = field(default_factory=list)
    wall: float = 0.0
    peak_bytes: int = 0
    max_active: int = 0

    def latencies(self) -> np.ndarray:
        return np.array([r.latency for r in self.results.values() if r.error is None])

    def summary(self) -> dict:
        lat = self.latencies()
        work = sum(r.work for r in self.results.values() if r.error is None)
        return {
            "requests": len(self.results),
            "rejected": len(self.rejected),
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "node_ksamples_per_s": work / self.wall if self.wall > 0 else None,
            "wall_s": self.wall,
            "peak_admitted_bytes": self.peak_bytes,
            "max_concurrent": self.max_active,
        }


class Scheduler:
    """Arrival-ordered admission against a byte budget, round-robin chunk issue."""

    def __init__(self, backend, budget_bytes: int = 288 * 1024**3, max_active: int = 8,
                 ahead: int = 2, realtime: bool = True):
        self.backend = backend
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.ahead = int(ahead)
        self.realtime = realtime

    def run(self, requests: list[ScanRequest], keep_kspace: bool = True) -> ServeReport:
        rep = ServeReport()
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        active: list = []
        used = 0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if self.realtime else None
        virtual = 0.0
        while pending or active:
            now = clock() if clock else virtual
            # admission: strictly in arrival order (no overtaking)
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                req = pending[0]
                fp = self.backend.footprint(req)
                if fp > self.budget:
                    pending.pop(0)
                    rep.rejected.append(req.rid)
                    rep.results[req.rid] = RequestResult(req.rid, 0.0, now, now, None, 0,
                                                         error=f"footprint {fp} B exceeds budget {self.budget} B")
                    continue
                if used + fp > self.budget:
                    break  # wait for memory to free up
                pending.pop(0)
                busy = {getattr(a, "slot", -1) for a in active}
                slot = next(i for i in range(self.max_active + 1) if i not in busy)
                st = self.backend.start(req, slot=slot)
                st.slot = slot
                st.t_admit = now
                st.fp = fp
                active.append(st)
                used += fp
                rep.peak_bytes = max(rep.peak_bytes, used)
                rep.max_active = max(rep.max_active, len(active))
            progressed = False
            for st in active:
                progressed |= self.backend.step(st, self.ahead)
            for st in list(active):
                if self.backend.done(st):
                    kspace = self.backend.finish(st)
                    end = clock() if clock else virtual
                    req = st.req
                    rep.results[req.rid] = RequestResult(req.rid, end - req.arrival, st.t_admit, end,
                                                         kspace if keep_kspace else None,
                                                         getattr(st, "work", req.work))
                    active.remove(st)
                    used -= st.fp
                    progressed = True
            if not progressed:
                if clock:
                    if not active and pending:
                        time.sleep(max(0.0, min(0.01, pending[0].arrival - clock())))
                    else:
                        time.sleep(0.0002)
                else:
                    virtual = max(virtual, pending[0].arrival) if (pending and not active) else virtual + 1e-3
        rep.wall = (clock() if clock else virtual)
        return rep


# ---------------------------------------------------------------------------
# GPU backend
# ---------------------------------------------------------------------------
@dataclass
class _State:
    req: ScanRequest
    stream: object
    phantom: object
    plans: list
    rows: list                    # (i0, i1) readout rows per chunk
    out: object = None
    ws: object = None
    nodes: object = None
    next_chunk: int = 0
    events: list = field(default_factory=list)
    host: object = None
    copy_event: object = None
    t_admit: float = 0.0
    fp: int = 0
    work: int = 0                 # node*k-samples actually computed (phantom class size)
    slot: int = -1


# Cube sides of the resident phantom classes: k^3 nodes from 10.6 k to 2.25 M
# in steps of 1.25 in k (about 2x in nodes).
PHANTOM_CLASSES = (22, 27, 34, 43, 54, 67, 84, 105, 131)


class PhantomStore:
    """Resident phantoms by size class: mesh, tissue, lumped weights, Bloch nodes.

    A request's node count is served by the class whose k^3 is nearest in
    log scale (``classes=None``: a class per cube side, i.e. the nearest
    k^3)."""

    def __init__(self, dev, classes=PHANTOM_CLASSES):
        self.dev = dev
        self.cache = {}
        self.classes = tuple(classes) if classes else None

    def key(self, n_nodes: int) -> int:
        # cube side of a box mesh with about n_nodes nodes
        k = max(4, int(round(n_nodes ** (1.0 / 3.0))))
        if not self.classes:
            return k
        return min(self.classes, key=lambda c: abs(math.log(c) - math.log(k)))

    def n_nodes(self, n_nodes: int) -> int:
        return self.key(n_nodes) ** 3

    def get(self, n_nodes: int):
        import torch

        from . import device, phantoms

        k = self.key(n_nodes)
        if k not in self.cache:
            m = phantoms.box_mesh((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), (k, k, k))
            tis = phantoms.draw_tissue(m.n_nodes, seed=2000 + k, r2s_range=(10.0, 50.0))
            col = device.P1Colouring(m.elems, m.n_nodes, device=self.dev)
            nodes_d = torch.from_numpy(m.nodes).to(self.dev)
            w = device.lumped_mass(col, nodes_d)
            bn = device.BlochNodes.from_arrays(m.nodes, tis.T1, tis.T2, tis.delta_B_mT, device=self.dev)
            self.cache[k] = dict(mesh=m, tissue=tis, nodes=nodes_d, w=w, bloch=bn,
                                 phi=torch.from_numpy(tis.delta_omega).to(self.dev),
                                 T2s=torch.from_numpy(tis.T2star).to(self.dev),
                                 M0=torch.from_numpy(tis.M0).to(self.dev))
        return self.cache[k]


class SteadyStateGraph:
    """Spoiled-GRE steady state of one resident phantom, captured as a HIP graph.

    The TR loop is recorded once per phantom class and stream slot, then
    replayed per request as one graph launch:
    - per TR one imaging block and one delay block;
    - all blocks run as one chained integrator launch, with the state in
      registers and the store and spoiling in the kernel.

    The graph owns its state and store buffers. The initial state
    (Mxy = 0, Mz = M0 = 1) is written by fill kernels inside the graph, so
    every replay starts from the same state. The stored Mxy of the last TR
    is ``store_xy[0]``, valid once the replay has run on the slot's stream.
    """

    def __init__(self, ph: dict, blocks: list, dev, stream):
        import torch

        from . import device, sequences

        N = ph["mesh"].n_nodes
        self.chain = device.BlockChain([(b.rf, b.G, b.dt) for b in blocks], [0 if b.store else -1 for b in blocks],
                                       [0 if b.empty else 1 for b in blocks], device=dev)
        self.blocks = blocks
        self.nodes = ph["bloch"]
        self.mxy = torch.empty(N, dtype=torch.complex128, device=dev)
        self.mz = torch.empty(N, dtype=torch.float64, device=dev)
        self.store_xy = torch.empty((1, N), dtype=torch.complex64, device=dev)
        self.store_z = torch.empty((1, N), dtype=torch.float32, device=dev)
        self.gamma = sequences.GAMMA
        with torch.cuda.stream(stream):
            self._body()           # eager run: loads the kernels before capture
        stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            self._body()

    def _body(self):
        from . import device

        self.mxy.zero_()
        self.mz.fill_(1.0)
        self.store_xy.zero_()
        self.store_z.zero_()
        # the TR loop as one chained integrator launch (state in registers)
        device.bloch_chain(self.nodes, self.chain, 1.0, self.gamma, self.mxy, self.mz, self.store_xy,
                           self.store_z)

    def replay(self):
        self.graph.replay()


def request_trajectory(req: ScanRequest):
    from . import trajectories

    r0, r1, r2 = req.shape
    if req.traj == "radial":
        return trajectories.radial(0.24, r0, r1, t_first=0.5, t_readout=2.56)
    return trajectories.cartesian((0.22, 0.22, 0.22), (r0, r1, r2), t_first=1.0, t_readout=3.0)


class GpuBackend:
    """Runs requests on the device; phantoms and plans stay resident."""

    def __init__(self, dev=None, chunk_rows: int = 32, n_dummy: int = 8):
        import torch

        from . import device

        self.dev = torch.device(dev) if dev is not None else device.default_device()
        self.store = PhantomStore(self.dev)
        self.plans = {}
        self.chunk_rows = chunk_rows
        self.n_dummy = n_dummy
        self.streams = []
        self.graphs = {}          # (phantom class, slot) -> SteadyStateGraph
        self._blocks = None

    def _stream(self, slot: int):
        import torch

        while len(self.streams) <= slot:
            self.streams.append(torch.cuda.Stream(device=self.dev))  # default priority
        return self.streams[slot]

    def _chunks(self, req: ScanRequest):
        from . import device

        key = (req.traj, req.shape)
        if key not in self.plans:
            kloc, t = request_trajectory(req)
            n = req.shape[0]
            rows, plans = [], []
            for i0 in range(0, n, self.chunk_rows):
                i1 = min(n, i0 + self.chunk_rows)
                sub = tuple(np.ascontiguousarray(k[i0:i1]) for k in kloc)
                plans.append(device.SignalPlan(sub, np.ascontiguousarray(t[i0:i1]), device=self.dev))
                rows.append((i0, i1))
            self.plans[key] = (plans, rows)
        return self.plans[key]

    def prepare(self, requests, slots: int = 0) -> None:
        """Make the phantoms and plans of ``requests`` resident (service
        start-up), and capture the steady-state graphs of their phantom
        classes for stream slots 0 .. slots-1."""
        for req in requests:
            self.store.get(req.n_nodes)
            self._chunks(req)
            for slot in range(slots):
                self._graph(req.n_nodes, slot)

    def _steady_blocks(self):
        from . import sequences

        if self._blocks is None:
            self._blocks = sequences.spoiled_gre(flip_deg=15.0, TR=5.0, n_dummy=self.n_dummy, n_lines=1,
                                                 rf_dur=0.2)
        return self._blocks

    def _graph(self, n_nodes: int, slot: int) -> SteadyStateGraph:
        key = (self.store.key(n_nodes), slot)
        g = self.graphs.get(key)
        if g is None:
            g = SteadyStateGraph(self.store.get(n_nodes), self._steady_blocks(), self.dev, self._stream(slot))
            self.graphs[key] = g
        return g

    def footprint(self, req: ScanRequest) -> int:
        N = self.store.n_nodes(req.n_nodes)
        S = req.n_samples
        plans, _ = self._chunks(req)
        ws = max(p.workspace_bytes(req.precision, N, 1) for p in plans)
        return N * (24 + 7 * 8) + ws + S * 16

    def start(self, req: ScanRequest, slot: int):
        import torch

        from . import device, sequences

        ph = self.store.get(req.n_nodes)
        stream = self._stream(slot)
        plans, rows = self._chunks(req)
        st = _State(req=req, stream=stream, phantom=ph, plans=plans, rows=rows)
        N = ph["mesh"].n_nodes
        st.work = N * req.n_samples
        graph = self._graph(req.n_nodes, slot)
        with torch.cuda.stream(stream):
            # spoiled-GRE steady state: the phantom class's captured TR loop,
            # replayed on this slot's stream, then copied out of the graph's
            # buffers before the slot's next replay (same stream, in order)
            graph.replay()
            m = (graph.store_xy[0].to(torch.complex128) * ph["M0"]).reshape(-1, 1)
            st.nodes = device.SignalNodes.from_arrays(ph["nodes"], ph["phi"], ph["T2s"], m, weights=ph["w"],
                                                      device=self.dev, stream=stream)
            st.out = torch.empty((req.n_samples, 1), dtype=torch.complex128, device=self.dev)
            need = max(p.workspace_bytes(req.precision, N, 1) for p in plans)
            st.ws = torch.empty(max(need, 1), dtype=torch.uint8, device=self.dev)
        return st

    def step(self, st: _State, ahead: int) -> bool:
        import torch

        pending = [e for e in st.events if not e.query()]
        if st.next_chunk >= len(st.plans) or len(pending) >= ahead:
            return False
        i0, i1 = st.rows[st.next_chunk]
        per_row = st.req.n_samples // st.req.shape[0]
        plan = st.plans[st.next_chunk]
        with torch.cuda.stream(st.stream):
            view = st.out[i0 * per_row:i1 * per_row]
            plan.execute(st.nodes, precision=st.req.precision, out=view, workspace=st.ws, stream=st.stream)
            ev = torch.cuda.Event()
            ev.record(st.stream)
        st.events.append(ev)
        st.next_chunk += 1
        if st.next_chunk == len(st.plans):
            with torch.cuda.stream(st.stream):
                st.host = torch.empty(st.out.shape, dtype=st.out.dtype, pin_memory=True)
                st.host.copy_(st.out, non_blocking=True)
                st.copy_event = torch.cuda.Event()
                st.copy_event.record(st.stream)
        return True

    def done(self, st: _State) -> bool:
        return st.copy_event is not None and st.copy_event.query()

    def finish(self, st: _State):
        res = st.host.numpy().reshape(*st.req.shape, 1).copy()
        st.nodes = st.out = st.ws = st.host = None
        return res
