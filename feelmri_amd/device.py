"""Device-resident API over the C ABI: torch tensors in HBM in, tensors out.

The reference-shaped, numpy-in/numpy-out calls (``feelmri_amd.MRI.Signal``
and friends) are thin shims over these.  torch provides device memory and
streams only; every computation runs in libfeelmri_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

PRECISIONS = {"f64": _lib.PREC_F64, "mixed": _lib.PREC_MIXED, "f32": _lib.PREC_F32}
METHODS = {"auto": _lib.METHOD_AUTO, "direct": _lib.METHOD_DIRECT, "cartesian": _lib.METHOD_CARTESIAN}
METHOD_NAMES = {v: k for k, v in METHODS.items()}


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("feelmri_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def as_device(a, dtype: torch.dtype, device) -> torch.Tensor:
    """Contiguous device tensor of ``dtype`` from numpy / torch / scalars."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    arr = np.ascontiguousarray(np.asarray(a))
    return torch.from_numpy(arr).to(device=device, dtype=dtype).contiguous()


def _prec(precision) -> int:
    if isinstance(precision, int):
        return precision
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"unknown precision {precision!r}; expected one of {list(PRECISIONS)}") from None


# ---------------------------------------------------------------------------
# Signal
# ---------------------------------------------------------------------------
def analyse_trajectory(kloc, t, method: str = "auto") -> dict:
    """Host-only analysis of a trajectory (no device needed), as
    ``SignalPlan`` performs it: the method it would choose, the number of
    distinct t, of readout runs and of t-grouped tiles, and the t-groups per
    packed slot and per conjugate-pair tile (1: none)."""
    kx, ky, kz = (np.ascontiguousarray(np.asarray(k, dtype=np.float64)) for k in kloc)
    tt = np.ascontiguousarray(np.asarray(t, dtype=np.float64))
    if tt.ndim != 3 or any(k.shape != tt.shape for k in (kx, ky, kz)):
        raise ValueError("kloc arrays and t must share one 3-D shape (n_meas, n_lines, n_kz)")
    meth, ng, nr, nt = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ex, gp, cp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.call("fm_signal_plan_analyse", *tt.shape, kx.ctypes.data_as(ctypes.c_void_p),
              ky.ctypes.data_as(ctypes.c_void_p), kz.ctypes.data_as(ctypes.c_void_p),
              tt.ctypes.data_as(ctypes.c_void_p), METHODS[method], ctypes.byref(meth), ctypes.byref(ng),
              ctypes.byref(nr), ctypes.byref(nt), ctypes.byref(ex), ctypes.byref(gp), ctypes.byref(cp))
    return {"method": METHOD_NAMES[meth.value], "n_tgroups": ng.value, "n_runs": nr.value,
            "n_tiles_grouped": nt.value, "runs_exact": bool(ex.value), "group_pack": gp.value,
            "conj_pack": cp.value}


class SignalPlan:
    """Resident sampling tables for one trajectory (``fm_signal_plan``).

    ``kloc`` is the reference's ``(kx, ky, kz)`` of shape (n_meas, n_lines,
    n_kz) in cycles/m and ``t`` the sample times in ms (MRI.h:19-20).
    """

    def __init__(self, kloc, t, method: str = "auto", device=None):
        self.device = torch.device(device) if device is not None else default_device()
        kx, ky, kz = (np.ascontiguousarray(np.asarray(k, dtype=np.float64)) for k in kloc)
        tt = np.ascontiguousarray(np.asarray(t, dtype=np.float64))
        if tt.ndim != 3:
            raise ValueError(f"t must be 3-D (n_meas, n_lines, n_kz), got shape {tt.shape}")
        for name, k in (("kx", kx), ("ky", ky), ("kz", kz)):
            if k.shape != tt.shape:
                raise ValueError(f"{name} shape {k.shape} differs from t shape {tt.shape}")
        self.shape = tuple(int(s) for s in tt.shape)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("fm_signal_plan_create", ctypes.byref(handle), *self.shape,
                      kx.ctypes.data_as(ctypes.c_void_p), ky.ctypes.data_as(ctypes.c_void_p),
                      kz.ctypes.data_as(ctypes.c_void_p), tt.ctypes.data_as(ctypes.c_void_p),
                      METHODS[method])
        self._h = handle
        ns, ng, meth = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        _lib.call("fm_signal_plan_info", self._h, ctypes.byref(ns), ctypes.byref(ng),
                  ctypes.byref(meth), None)
        self.n_samples, self.n_tgroups, self.method = ns.value, ng.value, METHOD_NAMES[meth.value]
        tg = np.empty(self.n_tgroups, dtype=np.float64)
        _lib.call("fm_signal_plan_info", self._h, None, None, None,
                  tg.ctypes.data_as(ctypes.c_void_p))
        self.tgroup_t = tg
        self._t = tt
        self.kloc, self.t = (kx, ky, kz), tt   # the host sampling, (n_meas, n_lines, n_kz) float64
        nr, rl, ex = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _lib.call("fm_signal_plan_runs", self._h, ctypes.byref(nr), ctypes.byref(rl), ctypes.byref(ex))
        # readout runs (k and t linear along every readout line): the mixed
        # form of static nodes takes the run kernel when n_runs > 0, the FP64
        # form its FP64 run kernel when the runs are also exact to 1e-14
        self.n_runs, self.run_len, self.runs_exact = nr.value, rl.value, bool(ex.value)
        nl, lg = ctypes.c_int64(), ctypes.c_int()
        _lib.call("fm_signal_plan_lines", self._h, ctypes.byref(nl), ctypes.byref(lg))
        # whole linear lines: the mixed form of static nodes takes the
        # per-line GEMM when line_gemm (it then precedes the run kernel)
        self.n_line_tabs, self.line_gemm = nl.value, bool(lg.value)

    def sample_groups(self) -> np.ndarray:
        """Index into ``tgroup_t`` of every sample, in sample order."""
        order = np.argsort(self.tgroup_t, kind="stable")
        idx = np.searchsorted(self.tgroup_t[order], self._t.ravel())
        return order[np.clip(idx, 0, max(self.n_tgroups - 1, 0))]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().fm_signal_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_timing(self, max_launches: int) -> None:
        """Record HIP events around the dominant kernel of the next
        ``max_launches`` executes (0 disables)."""
        _lib.call("fm_signal_plan_set_timing", self._h, int(max_launches))

    def kernel_ms(self) -> tuple[float, int]:
        """(summed ms, number of timed executes) since the last query."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        _lib.call("fm_signal_plan_kernel_ms", self._h, ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def workspace_bytes(self, precision, n_nodes: int, n_enc: int, moving: bool = False) -> int:
        """Workspace bytes ``execute`` needs, for static or moving nodes."""
        out = ctypes.c_size_t()
        _lib.call("fm_signal_workspace_size", self._h, _prec(precision), int(n_nodes), int(n_enc),
                  int(bool(moving)), ctypes.byref(out))
        return int(out.value)

    def execute(self, nodes: "SignalNodes", precision="f64", motion=None, out=None,
                workspace=None, stream=None) -> torch.Tensor:
        """Evaluate; returns (n_samples, n_enc) complex128 on the device.

        ``motion``: (modes_soa (3R, N) float32 tensor, coef (n_tgroups, R+3)
        float64 tensor) or None.
        """
        prec = _prec(precision)
        S, V, N = self.n_samples, nodes.n_enc, nodes.n_nodes
        if out is None:
            out = torch.empty((S, V), dtype=torch.complex128, device=self.device)
        elif out.dtype != torch.complex128 or tuple(out.shape) != (S, V) or not out.is_contiguous():
            raise ValueError("out must be a contiguous complex128 tensor of shape (n_samples, n_enc)")
        modes = coef = None
        R = 0
        if motion is not None:
            modes, coef = motion
            R = int(modes.shape[0]) // 3
            if modes.dtype != torch.float32 or modes.shape[1] != N:
                raise ValueError("motion modes must be float32 of shape (3R, n_nodes)")
            if coef.dtype != torch.float64 or tuple(coef.shape) != (self.n_tgroups, R + 3):
                raise ValueError("motion coefficients must be float64 of shape (n_tgroups, R+3)")
        need = self.workspace_bytes(prec, N, V, moving=R > 0)
        own_ws = workspace is None or workspace.numel() < need
        if own_ws:
            workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("fm_signal_execute", self._h, prec, N, V, ptr(nodes.packed), ptr(modes), R,
                      ptr(coef), ptr(workspace), workspace.numel(), ptr(out), stream_handle(stream))
            if own_ws and isinstance(stream, torch.cuda.Stream) and stream != torch.cuda.current_stream():
                # the workspace is freed on return: keep it from being reused
                # before the kernels queued on `stream` have run
                workspace.record_stream(stream)
        return out


class SignalNodes:
    """Packed SoA node table of ``fm_signal_nodes_pack``."""

    def __init__(self, packed: torch.Tensor, n_nodes: int, n_enc: int):
        self.packed, self.n_nodes, self.n_enc = packed, n_nodes, n_enc

    @classmethod
    def from_arrays(cls, r0, phi_dB0, T2, mxy, weights=None, dtype=torch.float64, device=None,
                    stream=None) -> "SignalNodes":
        """r0 (N,3), phi_dB0 (N,) rad/ms, T2 (N,) ms, mxy (N,V) complex (= M Mxy,
        or Mxy with ``weights`` the lumped diagonal of M)."""
        device = torch.device(device) if device is not None else default_device()
        cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
        r0 = as_device(r0, dtype, device).reshape(-1, 3)
        N = r0.shape[0]
        phi = as_device(phi_dB0, dtype, device).reshape(-1)
        T2 = as_device(T2, dtype, device).reshape(-1)
        mx = as_device(mxy, cdtype, device)
        if mx.ndim == 1:
            mx = mx.reshape(-1, 1)
        if phi.numel() != N or T2.numel() != N or mx.shape[0] != N:
            raise ValueError(f"node arrays disagree on the node count ({N}, {phi.numel()}, "
                             f"{T2.numel()}, {mx.shape[0]})")
        V = int(mx.shape[1])
        w = None if weights is None else as_device(weights, torch.float64, device).reshape(-1)
        if w is not None and w.numel() != N:
            raise ValueError("weights length differs from the node count")
        packed = torch.empty((5 + 2 * V, N), dtype=torch.float64, device=device)
        dt_code = _lib.DTYPE_F64 if dtype == torch.float64 else _lib.DTYPE_F32
        with torch.cuda.device(device):
            _lib.call("fm_signal_nodes_pack", N, V, dt_code, ptr(r0), ptr(phi), ptr(T2),
                      ptr(torch.view_as_real(mx)), ptr(w), ptr(packed), stream_handle(stream))
        return cls(packed, N, V)


def csr_spmv(indptr, indices, data, x: torch.Tensor, stream=None) -> torch.Tensor:
    """y = M x on the device (M in CSR, x (n, V) complex128)."""
    device = x.device
    indptr = as_device(indptr, torch.int64, device)
    indices = as_device(indices, torch.int64, device)
    data = as_device(data, torch.float64, device)
    x = x.to(torch.complex128).contiguous()
    if x.ndim == 1:
        x = x.reshape(-1, 1)
    n_rows = indptr.numel() - 1
    y = torch.empty((n_rows, x.shape[1]), dtype=torch.complex128, device=device)
    with torch.cuda.device(device):
        _lib.call("fm_csr_spmv_c128", n_rows, int(x.shape[1]), ptr(indptr), ptr(indices), ptr(data),
                  ptr(x), ptr(y), stream_handle(stream))
    return y


# ---------------------------------------------------------------------------
# Bloch
# ---------------------------------------------------------------------------
class BlochNodes:
    """Packed SoA node table of ``fm_bloch_nodes_pack`` (x, y, z, 1/T1, 1/T2, dB)."""

    def __init__(self, packed: torch.Tensor, n_pos: int):
        self.packed, self.n_pos = packed, n_pos

    @classmethod
    def from_arrays(cls, r0, T1, T2, delta_B, dtype=torch.float64, device=None, stream=None):
        device = torch.device(device) if device is not None else default_device()
        r0 = as_device(r0, dtype, device).reshape(-1, 3)
        N = r0.shape[0]
        T1 = as_device(T1, dtype, device).reshape(-1)
        T2 = as_device(T2, dtype, device).reshape(-1)
        dB = as_device(delta_B, dtype, device).reshape(-1)
        for name, a in (("T1", T1), ("T2", T2), ("delta_B", dB)):
            if a.numel() != N:
                raise ValueError(f"{name} has {a.numel()} entries, expected {N}")
        packed = torch.empty((6, N), dtype=torch.float64, device=device)
        dt_code = _lib.DTYPE_F64 if dtype == torch.float64 else _lib.DTYPE_F32
        with torch.cuda.device(device):
            _lib.call("fm_bloch_nodes_pack", N, dt_code, ptr(r0), ptr(T1), ptr(T2), ptr(dB),
                      ptr(packed), stream_handle(stream))
        return cls(packed, N)


class EventTable:
    """Device copy of one block's (rf, G, dt) event arrays."""

    def __init__(self, rf, G, dt, device=None):
        device = torch.device(device) if device is not None else default_device()
        self.rf = as_device(np.asarray(rf).reshape(-1), torch.complex128, device)
        self.G = as_device(np.asarray(G).reshape(-1, 3), torch.float64, device)
        self.dt = as_device(np.asarray(dt).reshape(-1), torch.float64, device)
        self.n_time = int(self.rf.numel())
        if self.G.shape[0] != self.n_time or self.dt.numel() != self.n_time:
            raise ValueError("rf, G and dt must have the same number of events")


def bloch(nodes: BlochNodes, events: EventTable, M0: float, gamma: float, mxy0: torch.Tensor,
          mz0: torch.Tensor, precision="f64", motion=None, history: bool = False, mxy_out=None,
          mz_out=None, stream=None):
    """Integrate one event table.  ``mxy0`` complex128 (N,), ``mz0`` float64 (N,).

    Returns (mxy, mz): (N,) final state, or (N, T) column-major history
    (returned as a (T, N) tensor transposed to (N, T)).
    """
    prec = _prec(precision)
    N, T = nodes.n_pos, events.n_time
    device = nodes.packed.device
    real = torch.float64 if prec == _lib.PREC_F64 else torch.float32
    cplx = torch.complex128 if prec == _lib.PREC_F64 else torch.complex64
    if history:
        shape_c, shape_r = (T, N), (T, N)
    else:
        # final state is always complex128 / float64 (chainable)
        shape_c, shape_r = (N,), (N,)
        real, cplx = torch.float64, torch.complex128
    if mxy_out is None:
        mxy_out = torch.empty(shape_c, dtype=cplx, device=device)
    if mz_out is None:
        mz_out = torch.empty(shape_r, dtype=real, device=device)
    modes = coef = None
    R = 0
    if motion is not None:
        modes, coef = motion
        R = int(modes.shape[0]) // 3
        if coef.dtype != torch.float64 or tuple(coef.shape) != (max(T - 1, 0), R + 3):
            raise ValueError("Bloch motion coefficients must be float64 of shape (n_time-1, R+3)")
    with torch.cuda.device(device):
        _lib.call("fm_bloch_execute", prec, N, T, ptr(nodes.packed), float(M0), float(gamma),
                  ptr(torch.view_as_real(events.rf)), ptr(events.G), ptr(events.dt),
                  ptr(torch.view_as_real(mxy0)), ptr(mz0), ptr(modes), R, ptr(coef),
                  ptr(mxy_out), ptr(mz_out), int(bool(history)), stream_handle(stream))
    if history:
        return mxy_out.T, mz_out.T
    return mxy_out, mz_out


# C layouts of fm_block_desc, fm_rf_pulse, fm_trapezoid
BLOCK_DESC = np.dtype([("t0", "f8"), ("dur", "f8"), ("dt", "f8"), ("dt_rf", "f8"), ("dt_gr", "f8"),
                       ("pulse0", "i4"), ("n_pulses", "i4"), ("trap0", "i4"), ("n_traps", "i4"),
                       ("spoil", "i4"), ("store_col", "i4")], align=True)
RF_PULSE = np.dtype([("start", "f8"), ("dur", "f8"), ("amp_re", "f8"), ("amp_im", "f8"), ("table0", "i8"),
                     ("n_table", "i4"), ("reserved", "i4")], align=True)
TRAPEZOID = np.dtype([("start", "f8"), ("ramp", "f8"), ("flat", "f8"), ("amp", "f8"), ("axis", "i4"),
                      ("reserved", "i4")], align=True)


def marks_from_blocks(block_len, codes) -> np.ndarray:
    """Per-event block-boundary marks of concatenated tables: 1 + code of the
    block ending before event e when e starts block b > 0, 0 elsewhere, and
    1 + code of the last block after the last event (code = spoil |
    (store_col + 1) << 1)."""
    block_len = np.asarray(block_len, dtype=np.int64)
    codes = np.asarray(codes, dtype=np.int32)
    off = np.concatenate([[0], np.cumsum(block_len)])
    marks = np.zeros(int(off[-1]) + 1, dtype=np.int32)
    marks[off[1:-1]] = 1 + codes[:-1]
    marks[off[-1]] = 1 + codes[-1]
    return marks


class DeviceEvents:
    """A block chain's event tables built on the device (``fm_events_build``)
    from block descriptions: ``desc`` (BLOCK_DESC records), ``pulses``
    (RF_PULSE), ``traps`` (TRAPEZOID) and ``rf_tables`` ((n, 2) float64 or
    complex table points of the tabulated pulses).  Holds the device
    descriptions and output tables; ``build()`` (re)writes the tables on a
    stream with no host round trip, so it can be recorded into a HIP graph,
    and reads nothing back: ``n_events`` stays on the device."""

    def __init__(self, desc, pulses, traps, rf_tables=None, device=None):
        device = torch.device(device) if device is not None else default_device()
        desc = np.ascontiguousarray(desc, dtype=BLOCK_DESC)
        pulses = np.ascontiguousarray(pulses, dtype=RF_PULSE)
        traps = np.ascontiguousarray(traps, dtype=TRAPEZOID)
        tab = np.zeros((0, 2)) if rf_tables is None else np.asarray(rf_tables)
        if np.iscomplexobj(tab):
            tab = np.stack([tab.real, tab.imag], axis=-1)
        tab = np.ascontiguousarray(tab, dtype=np.float64).reshape(-1, 2)
        self.n_blocks = int(desc.size)
        if self.n_blocks == 0:
            raise ValueError("no block")
        mx, tot = ctypes.c_int64(), ctypes.c_int64()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a.size else None  # noqa: E731
        _lib.call("fm_events_bound", self.n_blocks, vp(desc), vp(pulses), int(tab.shape[0]), vp(traps),
                  ctypes.byref(mx), ctypes.byref(tot))
        self.max_candidates, self.capacity = int(mx.value), max(int(tot.value), 1)
        if self.max_candidates > _lib.EVENTS_MAX_CANDIDATES:
            raise ValueError(f"a block has {self.max_candidates} grid candidates; the builder sorts at most "
                             f"{_lib.EVENTS_MAX_CANDIDATES} per block")
        self.n_store = int(desc["store_col"].max()) + 1 if self.n_blocks else 0
        u8 = lambda a: torch.from_numpy(a.view(np.uint8).copy()).to(device)  # noqa: E731
        self.desc, self.pulses, self.traps = u8(desc), u8(pulses), u8(traps)
        self.rf_tables = torch.from_numpy(tab.copy()).to(device)
        C = self.capacity
        self.offsets = torch.zeros(self.n_blocks + 1, dtype=torch.int64, device=device)
        self.times = torch.zeros(C, dtype=torch.float64, device=device)
        self.rf = torch.zeros(C, dtype=torch.complex128, device=device)
        self.G = torch.zeros((C, 3), dtype=torch.float64, device=device)
        self.dt = torch.zeros(C, dtype=torch.float64, device=device)
        self.marks = torch.zeros(C + 1, dtype=torch.int32, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)
        self.device = device

    @property
    def n_events(self) -> torch.Tensor:
        """Total event count (device, (1,) int64 view of offsets)."""
        return self.offsets[self.n_blocks:]

    def build(self, stream=None) -> "DeviceEvents":
        with torch.cuda.device(self.device):
            _lib.call("fm_events_build", self.n_blocks, ptr(self.desc), ptr(self.pulses) if self.pulses.numel() else None,
                      ptr(self.rf_tables) if self.rf_tables.numel() else None,
                      ptr(self.traps) if self.traps.numel() else None, self.max_candidates, self.capacity,
                      ptr(self.offsets), ptr(self.times), ptr(torch.view_as_real(self.rf)), ptr(self.G), ptr(self.dt),
                      ptr(self.marks), ptr(self.status), stream_handle(stream))
        return self

    def check(self) -> None:
        """Read the builder's status (synchronises): raise when a block did not
        fit or the events exceeded the capacity."""
        st = int(self.status.item())
        if st:
            raise RuntimeError(f"fm_events_build status {st}: " + ", ".join(
                m for b, m in ((1, "a block exceeded max_candidates"), (2, "events exceed the capacity")) if st & b))

    def tables(self):
        """Host copies (times, rf, G, dt, offsets, marks) of the built tables
        (synchronises)."""
        n = int(self.offsets[-1].item())
        return (self.times[:n].cpu().numpy(), self.rf[:n].cpu().numpy(), self.G[:n].cpu().numpy(),
                self.dt[:n].cpu().numpy(), self.offsets.cpu().numpy(), self.marks[:n + 1].cpu().numpy())


def bloch_chain_marked(nodes: BlochNodes, rf, G, dt, marks, n_events, capacity: int, M0: float, gamma: float,
                       mxy, mz, mxy_store=None, mz_store=None, n_store: int = 0, precision="f64", motion=None,
                       stream=None):
    """A block chain over concatenated tables with per-event marks and a
    device event count (``fm_bloch_chain_marked``): static nodes, or moving
    nodes with ``motion`` = (modes_soa, coef (capacity - 1, R+3) float64,
    one row per step).  The state is updated in place."""
    prec = _prec(precision)
    N = nodes.n_pos
    if mxy.dtype != torch.complex128 or mz.dtype != torch.float64 or mxy.numel() != N or mz.numel() != N:
        raise ValueError("chain state must be complex128 / float64 of n_pos entries")
    if n_store and (mxy_store is None or mz_store is None or mxy_store.dtype != torch.complex64
                    or mz_store.dtype != torch.float32 or mxy_store.shape[0] < n_store
                    or mxy_store.shape[1] != N or tuple(mz_store.shape) != tuple(mxy_store.shape)):
        raise ValueError("stores must be (n_store, N) complex64 / float32")
    if rf.numel() < capacity or G.shape[0] < capacity or dt.numel() < capacity or marks.numel() < capacity + 1:
        raise ValueError("event tables shorter than the capacity")
    modes = coef = None
    R = 0
    if motion is not None:
        modes, coef = motion
        R = int(modes.shape[0]) // 3
        if coef.dtype != torch.float64 or coef.shape[1] != R + 3 or coef.shape[0] < capacity - 1:
            raise ValueError("chain motion coefficients must be float64 of shape (>= capacity - 1, R+3)")
    with torch.cuda.device(nodes.packed.device):
        _lib.call("fm_bloch_chain_marked", prec, N, ptr(nodes.packed), float(M0), float(gamma), ptr(n_events),
                  int(capacity), ptr(torch.view_as_real(rf)), ptr(G), ptr(dt), ptr(marks), ptr(modes), R, ptr(coef),
                  ptr(torch.view_as_real(mxy)), ptr(mz), ptr(mxy_store), ptr(mz_store), int(n_store),
                  ptr(torch.view_as_real(mxy)), ptr(mz), stream_handle(stream))


class BlockChain:
    """Consecutive blocks for ``bloch_chain``: their event tables concatenated
    on the device, and per block (host) its event count, store column (-1:
    none) and spoil flag; also the per-event marks and event count
    (device) that ``bloch_chain_marked`` reads."""

    def __init__(self, events: list, store_col, spoil, device=None):
        device = torch.device(device) if device is not None else default_device()
        rf = np.concatenate([np.asarray(e[0]).reshape(-1) for e in events])
        G = np.concatenate([np.asarray(e[1]).reshape(-1, 3) for e in events])
        dt = np.concatenate([np.asarray(e[2]).reshape(-1) for e in events])
        self.events = EventTable(rf, G, dt, device=device)
        self.block_len = np.ascontiguousarray([np.asarray(e[2]).size for e in events], dtype=np.int64)
        self.store_col = np.ascontiguousarray(store_col, dtype=np.int32)
        self.spoil = np.ascontiguousarray(spoil, dtype=np.int32)
        self.n_blocks = int(self.block_len.size)
        if self.store_col.size != self.n_blocks or self.spoil.size != self.n_blocks:
            raise ValueError("one store column and one spoil flag per block")
        self.n_store = int(self.store_col.max()) + 1 if self.n_blocks else 0
        if self.n_blocks:
            codes = (self.spoil != 0).astype(np.int32) | ((self.store_col + 1) << 1)
            self.marks = torch.from_numpy(marks_from_blocks(self.block_len, codes)).to(device)
            self.n_events = torch.tensor([int(self.block_len.sum())], dtype=torch.int64, device=device)


def bloch_chain(nodes: BlochNodes, chain: BlockChain, M0: float, gamma: float, mxy, mz, mxy_store, mz_store,
                precision="f64", stream=None):
    """Integrate a block chain of static nodes in one launch (per 128 blocks):
    the state (``mxy`` complex128, ``mz`` float64, (N,)) is updated in place
    to the state after the last block; stored columns go to ``mxy_store`` /
    ``mz_store`` ((n_store, N) complex64 / float32, row-major)."""
    prec = _prec(precision)
    N = nodes.n_pos
    if mxy.dtype != torch.complex128 or mz.dtype != torch.float64 or mxy.numel() != N or mz.numel() != N:
        raise ValueError("chain state must be complex128 / float64 of n_pos entries")
    if chain.n_store and (mxy_store is None or mz_store is None or mxy_store.dtype != torch.complex64
                          or mz_store.dtype != torch.float32 or mxy_store.shape[0] < chain.n_store
                          or mxy_store.shape[1] != N or tuple(mz_store.shape) != tuple(mxy_store.shape)):
        raise ValueError("stores must be (n_store, N) complex64 / float32")
    ev = chain.events
    with torch.cuda.device(nodes.packed.device):
        _lib.call("fm_bloch_chain", prec, N, ptr(nodes.packed), float(M0), float(gamma), chain.n_blocks,
                  chain.block_len.ctypes.data_as(ctypes.c_void_p), chain.store_col.ctypes.data_as(ctypes.c_void_p),
                  chain.spoil.ctypes.data_as(ctypes.c_void_p), ptr(torch.view_as_real(ev.rf)), ptr(ev.G),
                  ptr(ev.dt), ptr(torch.view_as_real(mxy)), ptr(mz), ptr(mxy_store), ptr(mz_store), chain.n_store,
                  ptr(torch.view_as_real(mxy)), ptr(mz), stream_handle(stream))


def bloch_store_spoil(mxy_state, mz_state, mxy_store=None, mz_store=None, col: int = -1,
                      spoil: bool = False, stream=None):
    N = int(mz_state.numel())
    with torch.cuda.device(mz_state.device):
        _lib.call("fm_bloch_store_spoil", N, ptr(torch.view_as_real(mxy_state)), ptr(mz_state),
                  ptr(mxy_store), ptr(mz_store), int(col), int(bool(spoil)), stream_handle(stream))


# ---------------------------------------------------------------------------
# P1 mass
# ---------------------------------------------------------------------------
class P1Colouring:
    """Host greedy colouring of a tet mesh and each node's incident elements
    in colour order, resident on the device.

    The gather writes each element's determinant at its position in the
    colour order (``pos``), so that the incidence lists, which hold
    positions, gather neighbouring entries for neighbouring nodes.  ``order="element"`` keeps the elements in their own order
    (identity permutation): a node then adds its contributions in ascending
    element order, and on a mesh whose elements are ordered by their largest
    node id (``prefix_ordered``) a node prefix is weighted from the same
    tables.  A mesh whose greedy colouring needs more than 64 colours (a node
    with many incident elements, e.g. an unstructured Delaunay mesh) falls
    back to element order (``method == "element-order"``,
    ``n_colours == 0``): deterministic, but not bitwise equal to a coloured
    scatter, which is unavailable for such a mesh."""

    def __init__(self, elems, n_nodes: int, device=None, order: str = "colour"):
        if order not in ("colour", "element"):
            raise ValueError(f"order must be 'colour' or 'element', got {order!r}")
        device = torch.device(device) if device is not None else default_device()
        el = np.ascontiguousarray(np.asarray(elems), dtype=np.int32)
        if el.ndim != 2 or el.shape[1] != 4:
            raise ValueError(f"P1 tetrahedra expected: elems of shape (E, 4), got {el.shape}")
        E = el.shape[0]
        perm = np.empty(E, dtype=np.int32)
        offs = np.zeros(65, dtype=np.int64)
        nc = ctypes.c_int()
        rc = _lib.load().fm_mass_p1_color(E, int(n_nodes), el.ctypes.data_as(ctypes.c_void_p),
                                          perm.ctypes.data_as(ctypes.c_void_p),
                                          offs.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nc))
        if rc == _lib.FM_ERR_UNSUPPORTED:
            self.n_colours = 0
            self.offsets = np.zeros(1, dtype=np.int64)
        else:
            _lib.check(rc, "fm_mass_p1_color")
            self.n_colours = nc.value
            self.offsets = np.ascontiguousarray(offs[: self.n_colours + 1])
        if self.n_colours == 0 or order == "element":
            gather_perm = np.arange(E, dtype=np.int32)
            self.method = "element-order"
        else:
            gather_perm = perm
            self.method = "colour-order"
        self.n_elems, self.n_nodes = E, int(n_nodes)
        chunk_offsets = np.empty((self.n_nodes + 63) // 64 + 1, dtype=np.int64)
        degree = np.empty(self.n_nodes, dtype=np.int32)
        n_slots = ctypes.c_int64()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        # size query, then the table
        _lib.call("fm_mass_p1_incidence", E, self.n_nodes, vp(el), vp(gather_perm), vp(chunk_offsets), vp(degree),
                  None, 0, ctypes.byref(n_slots))
        incident = np.empty(max(n_slots.value, 1), dtype=np.int32)
        _lib.call("fm_mass_p1_incidence", E, self.n_nodes, vp(el), vp(gather_perm), vp(chunk_offsets), vp(degree),
                  vp(incident), incident.size, ctypes.byref(n_slots))
        self.n_slots = int(n_slots.value)
        # elements ordered by their largest node id, kept in that order: a
        # node prefix's sub-mesh is an element prefix (see lumped_mass)
        self._emax = el.max(axis=1) if E else np.zeros(0, dtype=np.int32)
        self.prefix_ordered = self.method == "element-order" and bool(np.all(np.diff(self._emax) >= 0))
        self.elems = torch.from_numpy(el).to(device)
        if self.method == "element-order":
            self.pos = None
        else:  # each element's position in the colour order
            pos = np.empty(E, dtype=np.int32)
            pos[perm] = np.arange(E, dtype=np.int32)
            self.pos = torch.from_numpy(pos).to(device)
        # the colouring itself serves the coloured scatter whatever the order
        self.perm = torch.from_numpy(perm if self.n_colours else gather_perm).to(device)
        self.chunk_offsets = torch.from_numpy(chunk_offsets).to(device)
        self.degree = torch.from_numpy(degree).to(device)
        self.incident = torch.from_numpy(incident).to(device)
        self.device = device

    def n_elems_below(self, n: int) -> int:
        """Elements whose nodes all lie below node ``n`` (the leading ones
        when ``prefix_ordered``; all elements for n = n_nodes)."""
        if n >= self.n_nodes:
            return self.n_elems
        return int(np.searchsorted(self._emax, n, side="left"))


def _mass_args(colouring: P1Colouring, nodes: torch.Tensor, out):
    nodes = nodes.contiguous()
    if nodes.dtype not in (torch.float32, torch.float64) or nodes.shape != (colouring.n_nodes, 3):
        raise ValueError("nodes must be float32/float64 of shape (n_nodes, 3)")
    if out is None:
        out = torch.empty(colouring.n_nodes, dtype=torch.float64, device=nodes.device)
    dt_code = _lib.DTYPE_F64 if nodes.dtype == torch.float64 else _lib.DTYPE_F32
    return nodes, out, dt_code


def lumped_mass(colouring: P1Colouring, nodes: torch.Tensor, out=None, stream=None,
                n_nodes: int | None = None) -> torch.Tensor:
    """w_i = sum_{e ni i} |T_e| / 4 (float64, (N,)): each element writes
    |det J_e| once, then each node adds its elements' determinants in colour
    order (element order for ``order="element"`` or a mesh without a
    colouring).

    ``n_nodes`` below the mesh's node count weights the sub-mesh of the first
    ``n_nodes`` nodes: the elements whose nodes all lie below it, which are
    the leading elements when the mesh's elements are ordered by their
    largest node id and kept in element order (``colouring.prefix_ordered``).
    ``nodes`` then needs only those rows."""
    n = colouring.n_nodes if n_nodes is None else int(n_nodes)
    if n != colouring.n_nodes:
        if not colouring.prefix_ordered:
            raise ValueError("a node prefix needs elements ordered by their largest node id, kept in element "
                             "order (P1Colouring(..., order='element'))")
        if not 0 <= n <= colouring.n_nodes:
            raise ValueError(f"n_nodes {n} outside [0, {colouring.n_nodes}]")
    E = colouring.n_elems_below(n)
    nodes = nodes.contiguous()
    if nodes.dtype not in (torch.float32, torch.float64) or nodes.ndim != 2 or nodes.shape[1] != 3 \
            or nodes.shape[0] < n:
        raise ValueError(f"nodes must be float32/float64 of shape (>= {n}, 3)")
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=nodes.device)
    dt_code = _lib.DTYPE_F64 if nodes.dtype == torch.float64 else _lib.DTYPE_F32
    q = torch.empty(max(E, 1), dtype=torch.float64, device=nodes.device)
    with torch.cuda.device(nodes.device):
        _lib.call("fm_mass_p1_lumped_gather", E, n, ptr(colouring.elems),
                  ptr(colouring.pos) if colouring.pos is not None else None, dt_code, ptr(nodes),
                  ptr(colouring.chunk_offsets), ptr(colouring.degree), ptr(colouring.incident), colouring.n_slots,
                  ptr(q), ptr(out), stream_handle(stream))
    if isinstance(stream, torch.cuda.Stream):
        q.record_stream(stream)  # the scratch is freed here but used on that stream
    return out


def lumped_mass_coloured(colouring: P1Colouring, nodes: torch.Tensor, out=None, stream=None) -> torch.Tensor:
    """The same weights by the coloured scatter: one launch per colour."""
    if colouring.n_colours == 0:
        raise ValueError("this mesh has no colouring of at most 64 colours: use lumped_mass (the gather)")
    nodes, out, dt_code = _mass_args(colouring, nodes, out)
    with torch.cuda.device(nodes.device):
        _lib.call("fm_mass_p1_lumped", colouring.n_elems, colouring.n_nodes, ptr(colouring.elems),
                  dt_code, ptr(nodes), ptr(colouring.perm),
                  colouring.offsets.ctypes.data_as(ctypes.c_void_p), colouring.n_colours, ptr(out),
                  stream_handle(stream))
    return out


def element_matrices(elems: torch.Tensor, nodes: torch.Tensor, stream=None) -> torch.Tensor:
    E = int(elems.shape[0])
    vals = torch.empty((E, 4, 4), dtype=torch.float64, device=nodes.device)
    dt_code = _lib.DTYPE_F64 if nodes.dtype == torch.float64 else _lib.DTYPE_F32
    # converted inputs are bound to names until the launch: a temporary freed
    # while the argument list is built could be handed to the next one
    elems_c, nodes_c = elems.contiguous(), nodes.contiguous()
    with torch.cuda.device(nodes.device):
        _lib.call("fm_mass_p1_element_matrices", E, ptr(elems_c), dt_code, ptr(nodes_c), ptr(vals),
                  stream_handle(stream))
    return vals


# ---------------------------------------------------------------------------
# POD
# ---------------------------------------------------------------------------
def pod_tensordot(modes: torch.Tensor, weights: torch.Tensor, stream=None) -> torch.Tensor:
    P, C, M = (int(s) for s in modes.shape)
    out = torch.empty((P, C), dtype=modes.dtype, device=modes.device)
    dt_code = _lib.DTYPE_F64 if modes.dtype == torch.float64 else _lib.DTYPE_F32
    modes_c, weights_c = modes.contiguous(), weights.contiguous()
    with torch.cuda.device(modes.device):
        _lib.call("fm_pod_tensordot", dt_code, P, C, M, ptr(modes_c), ptr(weights_c), ptr(out),
                  stream_handle(stream))
    return out


def modes_soa(modes_pcr: torch.Tensor, stream=None) -> torch.Tensor:
    """(P, 3, R) float32 modes -> (3R, P) rows for the fused loaders."""
    P, C, R = (int(s) for s in modes_pcr.shape)
    if C != 3:
        raise ValueError("modes must have 3 components")
    out = torch.empty((3 * R, P), dtype=torch.float32, device=modes_pcr.device)
    src = modes_pcr.to(torch.float32).contiguous()
    with torch.cuda.device(modes_pcr.device):
        _lib.call("fm_modes_pack", P, R, ptr(src), ptr(out), stream_handle(stream))
    return out


def pod_positions(xyz_soa: torch.Tensor, modes: torch.Tensor | None, coef: torch.Tensor,
                  stream=None) -> torch.Tensor:
    """r0 + sum_r modes_r w_r + trans -> (N, 3) float64."""
    N = int(xyz_soa.shape[1])
    R = 0 if modes is None else int(modes.shape[0]) // 3
    out = torch.empty((N, 3), dtype=torch.float64, device=xyz_soa.device)
    xyz_c, coef_c = xyz_soa.contiguous(), coef.contiguous()
    with torch.cuda.device(xyz_soa.device):
        _lib.call("fm_pod_positions", N, ptr(xyz_c), ptr(modes), R, ptr(coef_c), ptr(out),
                  stream_handle(stream))
    return out
