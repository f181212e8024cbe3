"""Forked-rank CPU baseline: the reference's MPI layout, run on the oracle.

MEASUREMENT INFRASTRUCTURE ONLY (part of the oracle package): bench.py's
``cpu_baseline`` leg starts this module as a child process; nothing in the
product loads it.

The reference runs its signal path under ``mpirun``: each rank owns the
elements of one mesh part and the nodes they touch (interface nodes
duplicated, python/feelmri/Phantom.py:194-239), assembles lumped weights from
its own elements only (Phantom.py:379-399), evaluates ``MRI.Signal`` on its
nodes, and the partial k-spaces are summed with ``Reduce(SUM)``
(python/feelmri/MPIUtilities.py:13-26).  This image has no MPI (no mpirun,
no mpi4py), so the ranks are forked processes, one per core, started
together behind a barrier.  Each rank runs single-threaded on the oracle
built with the reference's release flags (``liboracle_fast.so``, -O3
-ffast-math).  The k-space partials are summed in rank order.

Time = the slowest rank's compute (weights + M*Mxy + positions + signal)
plus the reduction, measured with ``time.perf_counter``.

The signal cost is exactly proportional to nodes x samples, so the run
covers a deterministic subset of readout indices (evenly spaced, all lines
and kz rows of each), sized from a probe run so that the timed part takes
about ``--target-s`` seconds.

    python -m oracle.cpu_ranks IN.npz OUT.json [--ranks 16] [--method blocks]
                               [--target-s 15]

IN.npz holds ``nodes`` (N,3), ``elems`` (E,4), ``phi`` (N,), ``T2`` (N,),
``mxy`` (N,V) complex (Mxy before weighting), the trajectory ``kx, ky, kz,
t`` of shape (ro, lines, kz), and for moving nodes ``modes`` (N,3,R) float32
with ``coef_t`` (G,) and ``coef`` (G, R+3): per distinct t the mode weights
and a uniform translation.  OUT.json receives the timings and the readout
indices used; OUT.json's stem + ``.npy`` receives the reduced k-space of
those readouts, (n_readouts * lines * kz, V) complex in sample order.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def _rank_main(rank: int, data: dict, part, sel, barrier, conn) -> None:
    import oracle

    oracle._lib(fast=True)  # load before the barrier
    nodes = data["nodes"][part.nodes_l2g]
    phi = data["phi"][part.nodes_l2g]
    T2 = data["T2"][part.nodes_l2g]
    mxy = data["mxy"][part.nodes_l2g]
    modes = data.get("modes")
    if modes is not None:
        modes = np.ascontiguousarray(modes[part.nodes_l2g]).astype(np.float64)
    kx, ky, kz, t = (np.ascontiguousarray(data[k][sel]).ravel() for k in ("kx", "ky", "kz", "t"))
    barrier.wait()
    t0 = time.perf_counter()
    w = oracle.mass_p1_lumped(part.elems, nodes, fast=True)
    mm = w[:, None] * mxy
    out = np.empty((t.size, mm.shape[1]), dtype=np.complex128)
    if modes is None:
        out[:] = oracle.signal_samples(kx, ky, kz, t, nodes, phi, T2, mm, n_threads=1, fast=True)
    else:
        # positions re-evaluated whenever t changes between consecutive samples
        R = modes.shape[2]
        starts = np.flatnonzero(np.r_[True, t[1:] != t[:-1]])
        ends = np.r_[starts[1:], t.size]
        ct = data["coef_t"]
        for s0, s1 in zip(starts, ends):
            c = data["coef"][int(np.flatnonzero(ct == t[s0])[0])]
            r = nodes + np.tensordot(modes, c[:R], axes=([2], [0])) + c[R:R + 3]
            out[s0:s1] = oracle.signal_samples(kx[s0:s1], ky[s0:s1], kz[s0:s1], t[s0:s1], r, phi, T2, mm,
                                               n_threads=1, fast=True)
    dt = time.perf_counter() - t0
    conn.send((rank, dt, out))
    conn.close()


def run(data: dict, sel, n_ranks: int, method: str = "blocks", parts=None) -> tuple[np.ndarray, dict]:
    """Run the ranks on the samples ``data[k][sel]``; returns the reduced
    (S, V) k-space and the timings."""
    from feelmri_amd import partition

    if parts is None:
        parts = partition.split_mesh(data["elems"], data["nodes"], n_ranks, method)
    ctx = mp.get_context("fork")
    barrier = ctx.Barrier(n_ranks)
    pipes, procs = [], []
    for r, part in enumerate(parts):
        a, b = ctx.Pipe(duplex=False)
        p = ctx.Process(target=_rank_main, args=(r, data, part, sel, barrier, b))
        p.start()
        pipes.append(a)
        procs.append(p)
    results = {}
    for a in pipes:
        rank, dt, out = a.recv()
        results[rank] = (dt, out)
    for p in procs:
        p.join()
        if p.exitcode != 0:
            raise RuntimeError(f"rank process exited with {p.exitcode}")
    t0 = time.perf_counter()
    total = np.zeros_like(results[0][1])
    for r in range(n_ranks):            # Reduce(SUM) in rank order
        total += results[r][1]
    t_red = time.perf_counter() - t0
    per = [results[r][0] for r in range(n_ranks)]
    info = {"ranks": n_ranks, "partition": method, "per_rank_s": per, "max_rank_s": max(per),
            "reduce_s": t_red, "seconds": max(per) + t_red,
            "nodes_stored": int(sum(p.n_nodes for p in parts)), "nodes": int(data["nodes"].shape[0]),
            "elements": int(data["elems"].shape[0])}
    return total, info


def timed_subset(data: dict, n_ranks: int, method: str, target_s: float, probe_lines: int = 8):
    """Probe one readout index on a few lines, then run evenly spaced
    readout indices sized to ``target_s``."""
    from feelmri_amd import partition

    parts = partition.split_mesh(data["elems"], data["nodes"], n_ranks, method)
    n_ro, n_lines, n_kz = data["t"].shape
    pl = min(probe_lines, n_lines)
    _, probe = run(data, (slice(0, 1), slice(0, pl), slice(None)), n_ranks, method, parts)
    per_ro = probe["max_rank_s"] * n_lines / pl
    n_pick = int(max(1, min(n_ro, round(target_s / max(per_ro, 1e-9)))))
    ro = np.unique(np.linspace(0, n_ro - 1, n_pick).round().astype(int))
    total, info = run(data, (ro,), n_ranks, method, parts)
    info.update(readouts=ro.tolist(), n_readouts_total=int(n_ro), samples=int(ro.size * n_lines * n_kz),
                probe_s=probe["seconds"])
    return total, info


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("inp")
    ap.add_argument("out")
    ap.add_argument("--ranks", type=int, default=16)
    ap.add_argument("--method", default="blocks", choices=["blocks", "rcb"])
    ap.add_argument("--target-s", type=float, default=15.0)
    a = ap.parse_args()
    os.environ["OMP_NUM_THREADS"] = "1"
    with np.load(a.inp, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    total, info = timed_subset(data, a.ranks, a.method, a.target_s)
    out = Path(a.out)
    np.save(out.with_suffix(".npy"), total)
    out.write_text(json.dumps(info))
    print(f"cpu_ranks: {a.ranks} ranks, {len(info['readouts'])} readouts, {info['seconds']:.2f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
