"""Block-chained Bloch solves with device-resident state.

Mirrors ``BlochSolver`` (python/feelmri/Bloch.py:275-380): for each block
the motion time shift is set to the block start (:335-336), the block's
event table is integrated from the carried state (:355), the last column is
kept when the block stores magnetisation (:358-359, :380), and the carried
Mxy is zeroed after every non-empty block -- the reference's artificial
spoiling -- while Mz is always carried (:363-368).

The state (Mxy complex128, Mz float64) stays in HBM between ``solve``
calls (the reference keeps it in ``initial_Mxy`` / ``initial_Mz``).  Static
nodes run a whole block range as one chained launch: the state stays in
registers across blocks and the store/spoil steps happen in the kernel.
Moving nodes take one integrator launch plus one store/spoil launch per
block, with their motion tables built per block on the host.  Event tables
are uploaded once and cached.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Motion, device
from .BlochSimulator import bloch_motion_tables, step_times
from .sequences import GAMMA, Block


class BlochSolver:
    def __init__(self, blocks: list[Block], r0, T1, T2, delta_B=0.0, M0: float = 1.0,
                 gamma: float = GAMMA, pod_trajectory=None, initial_Mxy=0.0, initial_Mz=None,
                 precision: str = "f64", dev=None):
        self.dev = torch.device(dev) if dev is not None else device.default_device()
        self.blocks = list(blocks)
        r0 = np.asarray(r0, dtype=np.float64).reshape(-1, 3)
        self.n_nodes = N = r0.shape[0]
        ones = np.ones(N)
        self.M0 = float(M0)
        self.gamma = float(gamma)
        self.precision = precision
        self.pod_trajectory = pod_trajectory
        if pod_trajectory is not None and not Motion.is_supported(pod_trajectory):
            raise TypeError("BlochSolver motion must have a modal form: the reference's POD / PODVelocity / "
                            "PODSum / RespiratoryMotion, or feelmri_amd.Motion's classes")
        with torch.cuda.device(self.dev):
            self.nodes = device.BlochNodes.from_arrays(r0, np.asarray(T1) * ones, np.asarray(T2) * ones,
                                                       np.asarray(delta_B).reshape(-1) * ones,
                                                       device=self.dev)
            mxy0 = np.asarray(initial_Mxy).reshape(-1) * ones
            mz0 = (np.asarray(initial_Mz).reshape(-1) * ones) if initial_Mz is not None else self.M0 * ones
            self.mxy = device.as_device(mxy0.astype(np.complex128), torch.complex128, self.dev)
            self.mz = device.as_device(mz0, torch.float64, self.dev)
        self._events: dict[int, device.EventTable] = {}
        self._chains: dict[tuple[int, int], device.BlockChain] = {}

    def add_block(self, block: Block) -> None:
        self.blocks.append(block)

    def _event_table(self, i: int) -> device.EventTable:
        ev = self._events.get(i)
        if ev is None:
            b = self.blocks[i]
            ev = device.EventTable(b.rf, b.G, b.dt, device=self.dev)
            self._events[i] = ev
        return ev

    def _chain(self, start: int, end: int) -> device.BlockChain:
        key = (start, end)
        ch = self._chains.get(key)
        if ch is None:
            idx = range(start, end)
            cols, col = [], 0
            for i in idx:
                cols.append(col if self.blocks[i].store else -1)
                col += int(self.blocks[i].store)
            ch = device.BlockChain([(self.blocks[i].rf, self.blocks[i].G, self.blocks[i].dt) for i in idx], cols,
                                   [0 if self.blocks[i].empty else 1 for i in idx], device=self.dev)
            self._chains[key] = ch
        return ch

    def solve_device(self, start: int = 0, end: int | None = None, stream=None, chain: bool = True):
        """Run blocks [start, end); returns device (Mxy (N, n_store) complex64,
        Mz (N, n_store) float32) of the storing blocks.

        Static nodes run the whole range as one chained launch (the state
        stays in registers across blocks; equal, bit for bit, to one
        integrator launch and one store/spoil launch per block, which moving
        nodes and ``chain=False`` take)."""
        nb = len(self.blocks)
        if start < 0:
            start += nb
        if end is None:
            end = nb
        n_store = sum(1 for i in range(start, end) if self.blocks[i].store)
        N = self.n_nodes
        # column-major (N, n_store) stores, as (n_store, N) row-major buffers
        mxy_store = torch.zeros((max(n_store, 1), N), dtype=torch.complex64, device=self.dev)
        mz_store = torch.zeros((max(n_store, 1), N), dtype=torch.float32, device=self.dev)
        self._run(start, end, mxy_store, mz_store, stream=stream, chain=chain)
        return mxy_store[:n_store].T, mz_store[:n_store].T

    def _run(self, start: int, end: int, mxy_store, mz_store, stream=None, chain: bool = True):
        if end <= start:
            return
        if chain and self.pod_trajectory is None:
            with torch.cuda.device(self.dev):
                device.bloch_chain(self.nodes, self._chain(start, end), self.M0, self.gamma, self.mxy, self.mz,
                                   mxy_store, mz_store, precision=self.precision, stream=stream)
            return
        idx = list(range(start, end))
        N = self.n_nodes
        col = 0
        with torch.cuda.device(self.dev):
            for i in idx:
                blk = self.blocks[i]
                ev = self._event_table(i)
                motion = None
                if self.pod_trajectory is not None:
                    self.pod_trajectory.update_timeshift(blk.t0)
                    motion = bloch_motion_tables(self.pod_trajectory, N, step_times(blk.dt), self.dev)
                device.bloch(self.nodes, ev, self.M0, self.gamma, self.mxy, self.mz,
                             precision=self.precision, motion=motion, history=False,
                             mxy_out=self.mxy, mz_out=self.mz, stream=stream)
                c = -1
                if blk.store:
                    c = col
                    col += 1
                device.bloch_store_spoil(self.mxy, self.mz, mxy_store, mz_store, col=c,
                                         spoil=not blk.empty, stream=stream)

    def solve(self, start: int = 0, end: int | None = None):
        mxy, mz = self.solve_device(start, end)
        return mxy.cpu().numpy(), mz.cpu().numpy()

    def capture(self, stream=None) -> "BlochSolveGraph":
        """Record the whole block chain as one HIP graph (static nodes).

        Each replay integrates every block from the state then held in
        ``self.mxy`` / ``self.mz`` and leaves the carried state there, as
        ``solve_device()`` does, with one graph launch.
        """
        return BlochSolveGraph(self, stream)


class BlochSolveGraph:
    """A ``BlochSolver``'s block chain captured once and replayed.

    Event tables are uploaded and the kernels loaded (one eager run, after
    which the carried state is restored) before the capture. The store
    buffers belong to the graph; every replay rewrites all stored columns.
    Moving nodes are not covered: their per-block tables come from the host.
    """

    def __init__(self, solver: BlochSolver, stream=None):
        if solver.pod_trajectory is not None:
            raise ValueError("graph capture covers static nodes; moving nodes build per-block motion tables "
                             "on the host")
        self.solver = solver
        N = solver.n_nodes
        nb = len(solver.blocks)
        self.n_store = sum(1 for b in solver.blocks if b.store)
        solver._chain(0, nb)   # uploaded before the capture
        with torch.cuda.device(solver.dev):
            self.stream = stream if stream is not None else torch.cuda.Stream(device=solver.dev)
            self.mxy_store = torch.zeros((max(self.n_store, 1), N), dtype=torch.complex64, device=solver.dev)
            self.mz_store = torch.zeros((max(self.n_store, 1), N), dtype=torch.float32, device=solver.dev)
            mxy0, mz0 = solver.mxy.clone(), solver.mz.clone()
            self.stream.wait_stream(torch.cuda.current_stream(solver.dev))
            with torch.cuda.stream(self.stream):
                self._body()
                solver.mxy.copy_(mxy0)
                solver.mz.copy_(mz0)
            self.stream.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph, stream=self.stream):
                self._body()

    def _body(self):
        s = self.solver
        s._run(0, len(s.blocks), self.mxy_store, self.mz_store)

    def replay(self):
        """Replay on the current stream's device; returns device (Mxy (N, n_store)
        complex64, Mz (N, n_store) float32), valid once the replay has run."""
        self.graph.replay()
        return self.mxy_store[: self.n_store].T, self.mz_store[: self.n_store].T
