"""Serving loop: continuous batching of synthetic scan requests on one GPU.

A request is one synthetic scan: a phantom of exactly N nodes, a spoiled-GRE
Bloch solve to steady state, lumped-mass weighting and the k-space signal of
its trajectory in FP64 or mixed precision.  The loop

* admits requests in arrival order while their HBM footprint fits the
  budget (288 GB by default) and a stream slot is free; a request larger
  than the whole budget is rejected;
* serves the requested node count exactly: every phantom is the first N
  nodes of one resident box mesh (``NodePool``, grid order) with the
  elements among them, weighted on the device from the pool's incidence
  table, and its tissue is drawn on the device from the request's seed;
* batches the steady-state Bloch solves of all requests admitted in one
  scheduling pass into one launch: their node tables are concatenated and
  the whole TR loop runs as one chained integrator launch (state in
  registers, stores and spoiling in the kernel), replayed from a HIP graph
  captured once per capacity class (``SteadyStateGraphs``: powers of two of
  nodes, the batch padded up to the class);
* keeps trajectory plans resident across requests and each admitted
  request's packed nodes and output resident until it completes;
* cuts each request's signal into chunks -- readout rows of a Cartesian
  trajectory, whole spokes of a radial one (written to a staging buffer and
  copied into place), so that the spokes keep the linear structure the
  per-line GEMM uses -- and interleaves
  the chunks of all active requests over their own HIP streams, so a
  request that arrives while others run starts at the next chunk boundary
  and its work overlaps theirs on the device;
* records latency from arrival to the result on the host, and aggregate
  node*k-sample/s.

Every per-node computation is independent of the other nodes in its launch,
chunk launch shapes depend only on the chunk's plan and the request's sizes,
and every reduction runs in a fixed order, so a request's result is bitwise
the same alone and in any mix.

The loop is written against a small backend interface (``start`` or
``start_batch``, ``step``, ``done``, ``finish``, ``footprint``);
``GpuBackend`` is the one that runs on the device.
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
    batches: list = field(default_factory=list)   # requests per admission batch

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
            "admission_batches": len(self.batches),
            "max_batch": max(self.batches) if self.batches else 0,
        }


class Scheduler:
    """Arrival-ordered admission against a byte budget, round-robin chunk issue.

    All requests admitted in one pass start together (``backend.start_batch``
    when the backend has it), so their set-up can share launches."""

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
            admitted = []
            while pending and pending[0].arrival <= now and len(active) + len(admitted) < self.max_active:
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
                busy = {getattr(a, "slot", -1) for a in active} | {s for _, s, _ in admitted}
                slot = next(i for i in range(self.max_active + 1) if i not in busy)
                admitted.append((req, slot, fp))
                used += fp
            if admitted:
                if hasattr(self.backend, "start_batch"):
                    states = self.backend.start_batch([(r, s) for r, s, _ in admitted])
                else:
                    states = [self.backend.start(r, slot=s) for r, s, _ in admitted]
                for st, (req, slot, fp) in zip(states, admitted):
                    st.slot = slot
                    st.t_admit = now
                    st.fp = fp
                    active.append(st)
                rep.batches.append(len(admitted))
                rep.peak_bytes = max(rep.peak_bytes, used)
                rep.max_active = max(rep.max_active, len(active))
            progressed = bool(admitted)
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
    plans: list
    rows: list                    # per chunk: (0, i0, i1) readout rows or (1, j0, j1) whole lines
    out: object = None
    ws: object = None
    nodes: object = None
    tissue: dict = None
    bloch: object = None
    w: object = None
    mxy: object = None            # the request's steady-state Mxy (M0 applied), (N, 1) complex128
    stage: object = None          # line-chunk output staging (radial requests)
    next_chunk: int = 0
    events: list = field(default_factory=list)
    host: object = None
    copy_event: object = None
    t_admit: float = 0.0
    fp: int = 0
    work: int = 0                 # node*k-samples computed
    slot: int = -1


class NodePool:
    """The resident box phantom every request takes its nodes from.

    A box grid of at least ``n_max`` nodes (C order, the last axis fastest)
    whose elements are ordered by their largest node id: the first N nodes
    and the leading elements -- those whose nodes all lie below N -- form a
    mesh of their own (a slab of the box), weighted by the lumped-mass gather
    from the pool's incidence table (``lumped_mass(n_nodes=N)``)."""

    def __init__(self, dev, n_max: int = 2_000_000):
        import torch

        from . import device, phantoms

        k = max(4, int(math.ceil(round(n_max ** (1.0 / 3.0), 9))))
        while k ** 3 < n_max:
            k += 1
        m = phantoms.box_mesh((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), (k, k, k))
        order = np.argsort(m.elems.max(axis=1), kind="stable")
        self.elems = np.ascontiguousarray(m.elems[order])
        self.n_max = m.n_nodes
        self.colouring = device.P1Colouring(self.elems, m.n_nodes, device=dev, order="element")
        assert self.colouring.prefix_ordered
        self.nodes = torch.from_numpy(m.nodes).to(dev)          # (n_max, 3) float64
        self.dev = dev

    def weights(self, n: int, stream=None):
        from . import device

        return device.lumped_mass(self.colouring, self.nodes, n_nodes=n, stream=stream)


def draw_tissue_device(n: int, seed: int, dev) -> dict:
    """Per-node tissue on the device from a seed (the serving stream's draw:
    M0~U[0.8,1], T1~U[500,1500] ms, T2~U[40,100] ms, dB0~N(0,20 Hz),
    R2*~U[10,50] 1/s), with the derived off-resonance in mT and rad/ms and
    T2* in ms.  The same (n, seed) gives the same values in any mix."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    u = torch.rand((4, n), generator=g, device=dev, dtype=torch.float64)
    df = 20.0 * torch.randn(n, generator=g, device=dev, dtype=torch.float64)
    T2 = 40.0 + 60.0 * u[2]
    R2s = 10.0 + 40.0 * u[3]
    return {"M0": 0.8 + 0.2 * u[0], "T1": 500.0 + 1000.0 * u[1], "T2": T2,
            "dB": df / 42.58e6 * 1e3, "phi": 2.0 * math.pi * df / 1000.0,
            "T2s": 1.0 / (1.0 / T2 + R2s / 1000.0)}


class SteadyStateGraphs:
    """The spoiled-GRE TR loop as HIP graphs, one per capacity class.

    A class holds node tables for 2^k nodes.  A batch of requests is copied
    into the class's node table (the padding beyond the batch keeps earlier,
    finite node data and its results are ignored), and one replay runs the
    initial-state fills and the whole block chain as one chained integrator
    launch.  The stored Mxy of the last TR is ``store_xy[0]``."""

    MIN_LOG2 = 14

    def __init__(self, blocks: list, dev, stream):
        from . import device, sequences

        self.blocks = blocks
        self.chain = device.BlockChain([(b.rf, b.G, b.dt) for b in blocks], [0 if b.store else -1 for b in blocks],
                                       [0 if b.empty else 1 for b in blocks], device=dev)
        self.gamma = sequences.GAMMA
        self.dev = dev
        self.stream = stream
        self.classes = {}

    def capacity(self, n: int) -> int:
        return 1 << max(self.MIN_LOG2, int(math.ceil(math.log2(max(n, 1)))))

    def get(self, n: int) -> dict:
        import torch

        from . import device

        cap = self.capacity(n)
        c = self.classes.get(cap)
        if c is not None:
            return c
        c = {"cap": cap,
             "packed": torch.zeros((6, cap), dtype=torch.float64, device=self.dev),
             "mxy": torch.empty(cap, dtype=torch.complex128, device=self.dev),
             "mz": torch.empty(cap, dtype=torch.float64, device=self.dev),
             "store_xy": torch.empty((1, cap), dtype=torch.complex64, device=self.dev),
             "store_z": torch.empty((1, cap), dtype=torch.float32, device=self.dev)}
        nodes = device.BlochNodes(c["packed"], cap)

        def body():
            c["mxy"].zero_()
            c["mz"].fill_(1.0)
            device.bloch_chain(nodes, self.chain, 1.0, self.gamma, c["mxy"], c["mz"], c["store_xy"], c["store_z"])

        with torch.cuda.stream(self.stream):
            body()                      # eager run: loads the kernels before capture
        self.stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            body()
        c["graph"] = g
        self.classes[cap] = c
        return c


def request_trajectory(req: ScanRequest):
    from . import trajectories

    r0, r1, r2 = req.shape
    if req.traj == "radial":
        return trajectories.radial(0.24, r0, r1, t_first=0.5, t_readout=2.56)
    return trajectories.cartesian((0.22, 0.22, 0.22), (r0, r1, r2), t_first=1.0, t_readout=3.0)


class GpuBackend:
    """Runs requests on the device; the node pool, trajectory plans and the
    steady-state graphs stay resident."""

    def __init__(self, dev=None, chunk_rows: int = 32, n_dummy: int = 8, n_max: int = 2_000_000,
                 chunk_lines: int = 64):
        import torch

        from . import device

        self.dev = torch.device(dev) if dev is not None else device.default_device()
        self.n_max = n_max
        self._pool = None
        self.plans = {}
        self.chunk_rows = chunk_rows
        self.chunk_lines = chunk_lines
        self.n_dummy = n_dummy
        self.streams = []
        self._graphs = None
        self._blocks = None
        self.batch_stream = None

    @property
    def pool(self) -> NodePool:
        if self._pool is None:
            self._pool = NodePool(self.dev, self.n_max)
        return self._pool

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
            n, nl = req.shape[0], req.shape[1]
            rows, plans = [], []
            if req.traj == "radial":
                # whole spokes per chunk: the lines stay linear over all their
                # samples, which the per-line GEMM of the mixed form needs
                for j0 in range(0, nl, self.chunk_lines):
                    j1 = min(nl, j0 + self.chunk_lines)
                    sub = tuple(np.ascontiguousarray(k[:, j0:j1]) for k in kloc)
                    plans.append(device.SignalPlan(sub, np.ascontiguousarray(t[:, j0:j1]), device=self.dev))
                    rows.append((1, j0, j1))
            else:
                for i0 in range(0, n, self.chunk_rows):
                    i1 = min(n, i0 + self.chunk_rows)
                    sub = tuple(np.ascontiguousarray(k[i0:i1]) for k in kloc)
                    plans.append(device.SignalPlan(sub, np.ascontiguousarray(t[i0:i1]), device=self.dev))
                    rows.append((0, i0, i1))
            self.plans[key] = (plans, rows)
        return self.plans[key]

    def _line_buffer_samples(self, req: ScanRequest) -> int:
        """Samples of the staging buffer a line chunk is written to (0 for
        readout-row chunks, which write their rows of the output in place)."""
        return req.shape[0] * min(req.shape[1], self.chunk_lines) * req.shape[2] if req.traj == "radial" else 0

    def steady_blocks(self):
        from . import sequences

        if self._blocks is None:
            self._blocks = sequences.spoiled_gre(flip_deg=15.0, TR=5.0, n_dummy=self.n_dummy, n_lines=1,
                                                 rf_dur=0.2)
        return self._blocks

    @property
    def graphs(self) -> SteadyStateGraphs:
        import torch

        if self._graphs is None:
            self.batch_stream = torch.cuda.Stream(device=self.dev)      # default priority
            self._graphs = SteadyStateGraphs(self.steady_blocks(), self.dev, self.batch_stream)
        return self._graphs

    def prepare(self, requests, slots: int = 0, max_batch: int = 8) -> None:
        """Service start-up: the node pool, the plans of ``requests``' trajectories
        and the steady-state graphs of every capacity class a batch of up to
        ``max_batch`` of the largest request can need."""
        if any(r.n_nodes > self.pool.n_max for r in requests):
            raise ValueError("a request exceeds the node pool")
        for req in requests:
            self._chunks(req)
        top = max((r.n_nodes for r in requests), default=1)
        cap = self.graphs.capacity(1)
        while cap <= self.graphs.capacity(max_batch * top):
            self.graphs.get(cap)
            cap *= 2
        for slot in range(slots):
            self._stream(slot)

    def footprint(self, req: ScanRequest) -> int:
        N, S = req.n_nodes, req.n_samples
        plans, _ = self._chunks(req)
        ws = max(p.workspace_bytes(req.precision, N, 1) for p in plans)
        # tissue (6 rows), Bloch table (6), signal table (7), weights, output,
        # line-chunk staging
        return N * 8 * (6 + 6 + 7 + 1) + ws + S * 16 + self._line_buffer_samples(req) * 16

    def start(self, req: ScanRequest, slot: int):
        return self.start_batch([(req, slot)])[0]

    def start_batch(self, admitted):
        """Set up the admitted requests together: tissue, weights and Bloch
        node tables on the batch stream, one steady-state replay for all of
        them, then each request's signal nodes on its own slot stream."""
        import torch

        from . import device

        pool = self.pool
        graphs = self.graphs
        bs = self.batch_stream
        bs.wait_stream(torch.cuda.current_stream(self.dev))
        states, offs, off = [], [], 0
        with torch.cuda.stream(bs):
            for req, slot in admitted:
                if req.n_nodes > pool.n_max:
                    raise ValueError(f"request {req.rid}: {req.n_nodes} nodes exceed the pool ({pool.n_max})")
                plans, rows = self._chunks(req)
                st = _State(req=req, stream=self._stream(slot), plans=plans, rows=rows)
                N = req.n_nodes
                st.work = N * req.n_samples
                st.tissue = draw_tissue_device(N, req.seed, self.dev)
                st.w = pool.weights(N, stream=bs)
                st.bloch = device.BlochNodes.from_arrays(pool.nodes[:N], st.tissue["T1"], st.tissue["T2"],
                                                         st.tissue["dB"], device=self.dev, stream=bs)
                states.append(st)
                offs.append(off)
                off += N
            cls = graphs.get(off)
            for st, o in zip(states, offs):
                cls["packed"][:, o:o + st.req.n_nodes].copy_(st.bloch.packed)
            cls["graph"].replay()
            store = cls["store_xy"][0]
            for st, o in zip(states, offs):
                N = st.req.n_nodes
                st.mxy = (store[o:o + N].to(torch.complex128) * st.tissue["M0"]).reshape(-1, 1)
            done = torch.cuda.Event()
            done.record(bs)
        for st in states:
            req, N = st.req, st.req.n_nodes
            s = st.stream
            s.wait_event(done)
            with torch.cuda.stream(s):
                st.nodes = device.SignalNodes.from_arrays(pool.nodes[:N], st.tissue["phi"], st.tissue["T2s"],
                                                          st.mxy, weights=st.w, device=self.dev, stream=s)
                st.out = torch.empty((req.n_samples, 1), dtype=torch.complex128, device=self.dev)
                nb = self._line_buffer_samples(req)
                st.stage = torch.empty((nb, 1), dtype=torch.complex128, device=self.dev) if nb else None
                need = max(p.workspace_bytes(req.precision, N, 1) for p in st.plans)
                st.ws = torch.empty(max(need, 1), dtype=torch.uint8, device=self.dev)
            # tensors made on the batch stream and used on this one
            for t in (st.w, st.mxy, st.tissue["phi"], st.tissue["T2s"]):
                t.record_stream(s)
            st.tissue = st.bloch = None
        return states

    def step(self, st: _State, ahead: int) -> bool:
        import torch

        pending = [e for e in st.events if not e.query()]
        if st.next_chunk >= len(st.plans) or len(pending) >= ahead:
            return False
        axis, a0, a1 = st.rows[st.next_chunk]
        plan = st.plans[st.next_chunk]
        with torch.cuda.stream(st.stream):
            if axis == 0:
                per_row = st.req.n_samples // st.req.shape[0]
                view = st.out[a0 * per_row:a1 * per_row]
                plan.execute(st.nodes, precision=st.req.precision, out=view, workspace=st.ws, stream=st.stream)
            else:
                n, nl, nk = st.req.shape
                stage = st.stage[:n * (a1 - a0) * nk]
                plan.execute(st.nodes, precision=st.req.precision, out=stage, workspace=st.ws, stream=st.stream)
                st.out.view(n, nl, nk)[:, a0:a1].copy_(stage.view(n, a1 - a0, nk))
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
        st.nodes = st.out = st.ws = st.host = st.w = st.mxy = st.stage = None
        return res
