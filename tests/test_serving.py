"""Scheduler logic of the serving loop, on the CPU with a stand-in backend.

The stand-in finishes a chunk per step call; it checks admission order,
the byte budget, rejection of oversize requests and overlap of requests.
"""
from dataclasses import dataclass, field

import numpy as np

from feelmri_amd.serving import ScanRequest, Scheduler, poisson_stream


@dataclass
class _St:
    req: ScanRequest
    chunks: int
    issued: int = 0
    t_admit: float = 0.0
    fp: int = 0


@dataclass
class FakeBackend:
    chunks_per_req: int = 3
    starts: list = field(default_factory=list)
    live: int = 0
    peak_live: int = 0
    used: int = 0
    peak_used: int = 0

    def footprint(self, req):
        return req.n_nodes

    def start(self, req, slot):
        self.starts.append(req.rid)
        self.live += 1
        self.used += req.n_nodes
        self.peak_live = max(self.peak_live, self.live)
        self.peak_used = max(self.peak_used, self.used)
        return _St(req, self.chunks_per_req)

    def step(self, st, ahead):
        if st.issued < st.chunks:
            st.issued += 1
            return True
        return False

    def done(self, st):
        return st.issued == st.chunks

    def finish(self, st):
        self.live -= 1
        self.used -= st.req.n_nodes
        return np.full((1,), st.req.rid)


def _reqs(sizes, arrivals):
    return [ScanRequest(i, a, n, "cart2d", (4, 4, 1), "f64") for i, (n, a) in enumerate(zip(sizes, arrivals))]


def test_all_requests_complete_in_arrival_order():
    be = FakeBackend()
    reqs = _reqs([100, 200, 300, 400], [0.0, 0.0, 0.001, 0.002])
    rep = Scheduler(be, budget_bytes=10_000, realtime=False).run(reqs)
    assert be.starts == [0, 1, 2, 3]
    assert sorted(rep.results) == [0, 1, 2, 3]
    assert all(r.error is None for r in rep.results.values())
    assert rep.max_active >= 2  # overlapping requests are served together


def test_budget_is_never_exceeded_and_blocks_in_order():
    be = FakeBackend(chunks_per_req=5)
    reqs = _reqs([600, 600, 100, 300], [0.0, 0.0, 0.0, 0.0])
    rep = Scheduler(be, budget_bytes=1000, realtime=False).run(reqs)
    assert be.peak_used <= 1000
    # request 1 (600) must wait for 0; request 2 may not overtake it
    assert be.starts == [0, 1, 2, 3]
    assert rep.peak_bytes <= 1000


def test_oversize_request_is_rejected_others_served():
    be = FakeBackend()
    reqs = _reqs([100, 5000, 100], [0.0, 0.0, 0.0])
    rep = Scheduler(be, budget_bytes=1000, realtime=False).run(reqs)
    assert rep.rejected == [1]
    assert rep.results[1].error and rep.results[0].error is None and rep.results[2].error is None


def test_max_active_slots():
    be = FakeBackend(chunks_per_req=50)
    reqs = _reqs([1] * 10, [0.0] * 10)
    Scheduler(be, budget_bytes=1 << 30, max_active=3, realtime=False).run(reqs)
    assert be.peak_live <= 3


def test_poisson_stream_distribution():
    reqs = poisson_stream(400, rate=20.0, seed=3)
    gaps = np.diff([r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 20.0) < 0.01
    n = np.array([r.n_nodes for r in reqs])
    assert n.min() >= 10_000 and n.max() <= 2_000_000
    # log-uniform: about half below the geometric mean
    assert 0.35 < np.mean(n < np.sqrt(10_000 * 2_000_000)) < 0.65
    assert {r.traj for r in reqs} == {"cart2d", "cart3d", "radial"}
    assert {r.precision for r in reqs} == {"f64", "mixed"}
