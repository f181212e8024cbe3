"""Block-chained Bloch solves with device-resident state.

Mirrors ``BlochSolver`` (python/feelmri/Bloch.py:275-380): for each block
the motion time shift is set to the block start (:335-336), the block's
event table is integrated from the carried state (:355), the last column is
kept when the block stores magnetisation (:358-359, :380), and the carried
Mxy is zeroed after every non-empty block -- the reference's artificial
spoiling -- while Mz is always carried (:363-368).

The state (Mxy complex128, Mz float64) stays in HBM between ``solve``
calls (the reference keeps it in ``initial_Mxy`` / ``initial_Mz``).  A whole
block range runs as one chained launch: the state stays in registers across
blocks and the store/spoil steps happen in the kernel.  Moving nodes chain
too: the motion modes are uploaded once and the per-step coefficients of
every block (each at its block's time shift) form one table.  The event
tables come from the host (uploaded once and cached) or, with
``events="device"``, are built on the device from the blocks'
specifications (``sequences.make_block``), bitwise the same tables, so that
a captured graph rebuilds and integrates the whole chain with no host work.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Motion, device
from .BlochSimulator import step_times
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
        self._dev_events: dict[tuple[int, int], device.DeviceEvents] = {}
        self._coefs: dict[tuple[int, int], torch.Tensor] = {}
        self._modes = None

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

    def _device_events(self, start: int, end: int) -> device.DeviceEvents:
        key = (start, end)
        de = self._dev_events.get(key)
        if de is None:
            from .sequences import block_descriptions

            de = device.DeviceEvents(*block_descriptions(self.blocks[start:end]), device=self.dev)
            de.build()
            de.check()   # once: every block fits (synchronises)
            self._dev_events[key] = de
        return de

    def _motion_modes(self):
        if self._modes is None:
            modes = Motion.motion_modes(self.pod_trajectory)
            if modes.shape[2] == 0:  # translation only: one zero mode
                modes = np.zeros((self.n_nodes, 3, 1), dtype=np.float32)
            if modes.shape[0] != self.n_nodes:
                raise ValueError(f"motion modes have {modes.shape[0]} nodes, expected {self.n_nodes}")
            self._modes = device.modes_soa(torch.from_numpy(np.ascontiguousarray(modes)).to(self.dev))
        return self._modes

    def _block_coefs(self, i: int) -> np.ndarray:
        """Motion coefficients at the steps of block i, at its time shift."""
        blk = self.blocks[i]
        self.pod_trajectory.update_timeshift(blk.t0)
        coef = Motion.motion_coefficients(self.pod_trajectory, step_times(blk.dt))
        if Motion.n_modes(self.pod_trajectory) == 0:
            coef = np.concatenate([np.zeros((coef.shape[0], 1)), coef], axis=1)
        return coef

    def _chain_coefs(self, start: int, end: int, min_rows: int = 0) -> torch.Tensor:
        """One coefficient row per step of the blocks' concatenated tables:
        block b's steps take rows off_b .. off_b + n_b - 2; the row of the
        step that starts block b > 0 is not read (zeros); at least
        ``min_rows`` rows (zero padded)."""
        key = (start, end, min_rows)
        c = self._coefs.get(key)
        if c is None:
            lens = [self.blocks[i].n_events for i in range(start, end)]
            rows = []
            for k, i in enumerate(range(start, end)):
                if k:
                    rows.append(None)
                rows.append(self._block_coefs(i))
            width = next(r.shape[1] for r in rows if r is not None)
            table = np.zeros((max(sum(lens) - 1, min_rows, 1), width), dtype=np.float64)
            r0 = 0
            for r in rows:
                if r is None:
                    r0 += 1
                    continue
                table[r0:r0 + r.shape[0]] = r
                r0 += r.shape[0]
            c = torch.from_numpy(table).to(self.dev)
            self._coefs[key] = c
        return c

    def solve_device(self, start: int = 0, end: int | None = None, stream=None, chain: bool = True,
                     events: str = "host"):
        """Run blocks [start, end); returns device (Mxy (N, n_store) complex64,
        Mz (N, n_store) float32) of the storing blocks.

        The range runs as one chained launch (the state stays in registers
        across blocks; equal, bit for bit, to one integrator launch and one
        store/spoil launch per block, which ``chain=False`` takes).
        ``events="device"`` builds the event tables on the device from the
        blocks' specifications first."""
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
        self._run(start, end, mxy_store, mz_store, stream=stream, chain=chain, events=events)
        return mxy_store[:n_store].T, mz_store[:n_store].T

    def _run(self, start: int, end: int, mxy_store, mz_store, stream=None, chain: bool = True,
             events: str = "host"):
        if end <= start:
            return
        if events not in ("host", "device"):
            raise ValueError("events must be 'host' or 'device'")
        if events == "device" and not chain:
            raise ValueError("device-built event tables run as a chain")
        moving = self.pod_trajectory is not None
        n_store = sum(1 for i in range(start, end) if self.blocks[i].store)
        if chain and (moving or events == "device"):
            with torch.cuda.device(self.dev):
                if events == "device":
                    de = self._device_events(start, end)
                    motion = (self._motion_modes(), self._chain_coefs(start, end, de.capacity - 1)) if moving \
                        else None
                    de.build(stream=stream)
                    tabs = (de.rf, de.G, de.dt, de.marks, de.n_events, de.capacity)
                else:
                    motion = (self._motion_modes(), self._chain_coefs(start, end)) if moving else None
                    ch = self._chain(start, end)
                    ev = ch.events
                    tabs = (ev.rf, ev.G, ev.dt, ch.marks, ch.n_events, ev.n_time)
                device.bloch_chain_marked(self.nodes, *tabs, self.M0, self.gamma, self.mxy, self.mz,
                                          mxy_store, mz_store, n_store, precision=self.precision, motion=motion,
                                          stream=stream)
            return
        if chain:
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
                    coef = torch.from_numpy(np.ascontiguousarray(self._block_coefs(i))).to(self.dev)
                    motion = (self._motion_modes(), coef)
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

    def capture(self, stream=None, events: str = "host") -> "BlochSolveGraph":
        """Record the whole block chain as one HIP graph.

        Each replay integrates every block from the state then held in
        ``self.mxy`` / ``self.mz`` and leaves the carried state there, as
        ``solve_device()`` does, with one graph launch; with
        ``events="device"`` the replay also rebuilds the event tables from
        the device block descriptions.
        """
        return BlochSolveGraph(self, stream, events=events)


class BlochSolveGraph:
    """A ``BlochSolver``'s block chain captured once and replayed.

    Event tables (or, with ``events="device"``, their descriptions), motion
    modes and coefficients are uploaded and the kernels loaded (one eager
    run, after which the carried state is restored) before the capture.  The
    store buffers belong to the graph; every replay rewrites all stored
    columns.
    """

    def __init__(self, solver: BlochSolver, stream=None, events: str = "host"):
        self.solver = solver
        self.events = events
        N = solver.n_nodes
        self.n_store = sum(1 for b in solver.blocks if b.store)
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
        s._run(0, len(s.blocks), self.mxy_store, self.mz_store, stream=torch.cuda.current_stream(s.dev),
               events=self.events)

    def replay(self):
        """Replay on the current stream's device; returns device (Mxy (N, n_store)
        complex64, Mz (N, n_store) float32), valid once the replay has run."""
        self.graph.replay()
        return self.mxy_store[: self.n_store].T, self.mz_store[: self.n_store].T
