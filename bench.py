#!/usr/bin/env python
"""Signal-generation throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form mixed|f64]
                    [--config c3|c1|c0|c2-radial|c2-spiral|c4] [--method auto|direct|cartesian]

A step is one evaluation of the k-space signal of the workload (all
samples, all encodings, all cardiac phases) from node data already resident
in HBM.  The default workload is c3, the largest single-GPU configuration:
4-D phase-contrast flow on a 1 948 800-node P1 curved tube with 10 POD
displacement modes, 4 velocity encodings and 20 cardiac phases of a
96 x 96 x 24 Cartesian k-space, the nodes moving with every readout.  Set-up
on the GPU: lumped-mass weights and the whole phase-contrast acquisition
through the Bloch integrator with the nodes moving along a Poiseuille
velocity field.

Prints one JSON line: node*k-sample/s of the chosen form, its relative L2
error against the FP64 GPU form over the full k-space, against the FP64
oracle on sampled k-points and against the CPU baseline's readouts, the
roofline of the dominant kernel (timed live with HIP events on the launch
stream, traffic from committed hardware counters), and the FP64 CPU oracle
run the reference's way (forked ranks over an element partition, partial
lumped weights, a final sum) on the host on a bounded sample of readouts.
The c1 3-D Cartesian workload follows as a secondary block, then the other
rows of the path (Bloch, lumped mass, POD) and a serving stream.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAKS = {  # MI355X published dense peaks
    "bf16_mfma_tflops": 2500.0,
    "fp64_mfma_tflops": 78.6,
    "fp64_valu_tflops": 78.6,
    "fp32_valu_tflops": 157.3,
    "hbm_gbs": 8000.0,
}
CPU_RANKS = 16   # the GPU box's host share for one GPU
METRIC = "node*k-sample/s at relL2<=1e-4 vs FP64 oracle"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_l2(a, b) -> float:
    import numpy as np

    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


def build_workload(name: str, scale: float):
    from feelmri_amd import workloads

    if name == "c0":
        return workloads.c0(scale)
    if name == "c1":
        return workloads.c1(scale)
    if name.startswith("c2"):
        return workloads.c2(scale, kind=name.split("-")[1] if "-" in name else "radial")
    raise SystemExit(f"unknown config {name}")


def prepare_nodes(wl, dev):
    """Lumped weights, spoiled-GRE steady state, packed signal nodes (GPU)."""
    import torch

    from feelmri_amd import device
    from feelmri_amd.Bloch import BlochSolver

    m, tis = wl.mesh, wl.tissue
    col = device.P1Colouring(m.elems, m.n_nodes, device=dev)
    nodes_d = torch.from_numpy(m.nodes).to(dev)
    w = device.lumped_mass(col, nodes_d)
    solver = BlochSolver(wl.blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT, dev=dev)
    mxy_store, _ = solver.solve_device()
    mxy = mxy_store[:, -1].to(torch.complex128) * torch.from_numpy(tis.M0).to(dev)
    mxy = mxy.reshape(-1, 1).repeat(1, wl.n_enc)
    nodes = device.SignalNodes.from_arrays(nodes_d, torch.from_numpy(tis.delta_omega).to(dev),
                                           torch.from_numpy(tis.T2star).to(dev), mxy, weights=w, device=dev)
    torch.cuda.synchronize()
    return nodes, w, mxy


def time_form(plan, nodes, form, steps, warmup):
    import torch

    out = torch.empty((plan.n_samples, nodes.n_enc), dtype=torch.complex128, device=plan.device)
    ws = torch.empty(max(plan.workspace_bytes(form, nodes.n_nodes, nodes.n_enc), 1), dtype=torch.uint8,
                     device=plan.device)
    for _ in range(warmup):
        plan.execute(nodes, precision=form, out=out, workspace=ws)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    plan.set_timing(steps)  # HIP events around the dominant kernel, on its stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        plan.execute(nodes, precision=form, out=out, workspace=ws, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_event = ev0.elapsed_time(ev1) / steps
    k_ms, k_n = plan.kernel_ms()
    plan.set_timing(0)
    return out, ms_event, wall / steps * 1e3, (k_ms / k_n if k_n else None)


def time_host_io(plan, wl, w_h, mxy_h, form, steps):
    """Whole call from host arrays: node arrays (numpy) to the device, pack,
    execute, k-space back to host memory; the plan is reused."""
    import numpy as np
    import torch

    from feelmri_amd import device

    host = (np.ascontiguousarray(wl.mesh.nodes), np.ascontiguousarray(wl.tissue.delta_omega),
            np.ascontiguousarray(wl.tissue.T2star), np.ascontiguousarray(mxy_h), np.ascontiguousarray(w_h))
    ws = torch.empty(max(plan.workspace_bytes(form, wl.n_nodes, mxy_h.shape[1]), 1), dtype=torch.uint8,
                     device=plan.device)

    def one():
        nodes = device.SignalNodes.from_arrays(*host[:4], weights=host[4], device=plan.device)
        out = torch.empty((plan.n_samples, nodes.n_enc), dtype=torch.complex128, device=plan.device)
        plan.execute(nodes, precision=form, out=out, workspace=ws)
        return out.cpu().numpy()

    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    return (time.perf_counter() - t0) / steps * 1e3


def time_shim_per_line(wl, w_h, mxy_h, form, reps: int = 3):
    """The reference examples' call pattern: one ``MRI.Signal`` call per
    phase-encoding line from host arrays (free_running.py:203-209), through
    the drop-in shim (plans and packed nodes cached across calls).  Returns
    ms per full k-space."""
    import numpy as np
    import scipy.sparse as sp

    from feelmri_amd import MRI

    M = sp.diags(w_h)
    lines = wl.t.shape[1]
    args = (wl.mesh.nodes, wl.tissue.delta_omega, wl.tissue.T2star, mxy_h)

    def once():
        for j in range(lines):
            MRI.Signal(0, M, [np.ascontiguousarray(k[:, j:j + 1, :]) for k in wl.kloc],
                       np.ascontiguousarray(wl.t[:, j:j + 1, :]), *args, precision=form)

    once()
    t0 = time.perf_counter()
    for _ in range(reps):
        once()
    return (time.perf_counter() - t0) / reps * 1e3, lines


def time_scan(wl, dev, plan, form, reps: int = 3) -> dict:
    """End-to-end scan on the device: lumped weights, the Bloch solve of the
    workload's sequence to its stored magnetisation, the packed signal
    nodes and the k-space.  Mesh colouring, event tables, the plan and the
    HIP graph of the Bloch block chain are set up once beforehand."""
    import torch

    from feelmri_amd import device
    from feelmri_amd.Bloch import BlochSolver

    m, tis = wl.mesh, wl.tissue
    col = device.P1Colouring(m.elems, m.n_nodes, device=dev)
    nodes_d = torch.from_numpy(m.nodes).to(dev)
    M0 = torch.from_numpy(tis.M0).to(dev)
    phi = torch.from_numpy(tis.delta_omega).to(dev)
    T2s = torch.from_numpy(tis.T2star).to(dev)
    # the mixed form integrates the Bloch equations in FP32, the FP64 form in FP64
    solver = BlochSolver(wl.blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT,
                         precision="f32" if form == "mixed" else "f64", dev=dev)
    mxy0, mz0 = solver.mxy.clone(), solver.mz.clone()
    graph = solver.capture()   # the block chain as one graph launch
    out = torch.empty((plan.n_samples, wl.n_enc), dtype=torch.complex128, device=dev)
    ws = torch.empty(max(plan.workspace_bytes(form, m.n_nodes, wl.n_enc), 1), dtype=torch.uint8, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def once(record):
        solver.mxy.copy_(mxy0)
        solver.mz.copy_(mz0)
        if record:
            ev[0].record()
        w = device.lumped_mass(col, nodes_d)
        if record:
            ev[1].record()
        store, _ = graph.replay()
        if record:
            ev[2].record()
        mxy = (store[:, -1].to(torch.complex128) * M0).reshape(-1, 1).repeat(1, wl.n_enc)
        nodes = device.SignalNodes.from_arrays(nodes_d, phi, T2s, mxy, weights=w, device=dev)
        plan.execute(nodes, precision=form, out=out, workspace=ws)
        if record:
            ev[3].record()

    once(False)
    torch.cuda.synchronize()
    parts = [0.0, 0.0, 0.0]
    for _ in range(reps):
        once(True)
        torch.cuda.synchronize()
        for i in range(3):
            parts[i] += ev[i].elapsed_time(ev[i + 1]) / reps
    total = sum(parts)
    bloch_err = None
    if form == "mixed":
        # FP32 steady state against the FP64 integration of the same sequence
        ref = BlochSolver(wl.blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT, precision="f64", dev=dev)
        solver.mxy.copy_(mxy0)
        solver.mz.copy_(mz0)
        a = solver.solve_device()[0][:, -1].to(torch.complex128)
        b = ref.solve_device()[0][:, -1].to(torch.complex128)
        bloch_err = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    return {"ms": total, "node_ksamples_per_s": wl.node_ksamples / (total * 1e-3),
            "bloch_relL2_vs_f64": bloch_err,
            "parts_ms": {"lumped_mass": parts[0], "bloch": parts[1], "pack_and_signal": parts[2]},
            "bloch_blocks": len(wl.blocks), "bloch_graph": True,
            "bloch_precision": "f32" if form == "mixed" else "f64"}


def load_traffic(kernel: str, config: str, form: str):
    """HBM bytes per launch of ``kernel`` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a copy kernel of
    known size), or None when no summary matches."""
    path = ROOT / "profiles" / "traffic.json"
    try:
        table = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    rec = table.get(f"{kernel}/{config}/{form}")
    return None if rec is None else rec.get("bytes_per_launch")


def cpu_baseline(wl, w_h, mxy_h, target_s: float, motion=None):
    """The FP64 oracle run the reference's way (``oracle.cpu_ranks``): forked
    ranks, one per host core, each owning one block of elements and the
    nodes they touch, with partial lumped weights, summed at the end.  Runs
    as a child process on evenly spaced readout indices (all lines and kz
    rows of each), sized to about ``target_s`` seconds.

    ``motion``: (modes (N,3,R) float32, coef_t (G,), coef (G,R+3)) for
    moving nodes.  Returns (baseline dict, readout indices, reduced k-space
    (n_readouts * lines * kz, V))."""
    import numpy as np

    tmp = Path(tempfile.mkdtemp(prefix="fm_cpu_"))
    try:
        arrays = dict(nodes=wl.mesh.nodes, elems=wl.mesh.elems, phi=wl.tissue.delta_omega, T2=wl.tissue.T2star,
                      mxy=mxy_h, kx=np.asarray(wl.kloc[0]), ky=np.asarray(wl.kloc[1]), kz=np.asarray(wl.kloc[2]),
                      t=np.asarray(wl.t))
        if motion is not None:
            arrays.update(modes=motion[0], coef_t=motion[1], coef=motion[2])
        np.savez(tmp / "in.npz", **arrays)
        ranks = min(CPU_RANKS, os.cpu_count() or 1)
        cmd = [sys.executable, "-m", "oracle.cpu_ranks", str(tmp / "in.npz"), str(tmp / "out.json"),
               "--ranks", str(ranks), "--method", "blocks", "--target-s", str(target_s)]
        res = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True,
                             env={**os.environ, "OMP_NUM_THREADS": "1"})
        if res.returncode != 0:
            raise RuntimeError(f"cpu_ranks failed ({res.returncode}): {res.stderr[-2000:]}")
        info = json.loads((tmp / "out.json").read_text())
        ks = np.load(tmp / "out.npy")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    ro = info["readouts"]
    work = wl.n_nodes * info["samples"] * mxy_h.shape[1]
    cpu = {"value": work / info["seconds"], "unit": "node*k-sample/s", "cores": info["ranks"],
           "kind": "port",
           "sample": (f"{len(ro)} of {info['n_readouts_total']} readout indices ({info['samples']} k-samples x "
                      f"{mxy_h.shape[1]} encodings x {wl.n_nodes} nodes), {info['ranks']} forked ranks over "
                      f"contiguous element blocks ({info['nodes_stored']} stored nodes incl. interface copies), "
                      f"partial lumped weights per rank, rank-order sum; max rank {info['max_rank_s']:.2f} s + "
                      f"reduce {info['reduce_s'] * 1e3:.1f} ms; FP64 oracle at -O3 -ffast-math, 1 thread per rank; "
                      "no MPI in this image"),
           "seconds": info["seconds"], "ranks": info["ranks"], "readouts": ro}
    return cpu, ro, ks


def serving_run(dev, n_requests: int, rate: float, max_active: int = 8):
    """Poisson stream of mixed synthetic scans through the continuous-batching
    loop; phantoms and plans are made resident first (service start-up)."""
    import numpy as np

    from feelmri_amd.serving import GpuBackend, ScanRequest, Scheduler, poisson_stream

    reqs = poisson_stream(n_requests, rate=rate, seed=1004)
    warm = [ScanRequest(10**6 + i, 0.0, 10_000, kind, shape, prec, 1)
            for i, (kind, shape, prec) in enumerate([("cart2d", (64, 64, 1), "f64"), ("cart3d", (64, 64, 32), "mixed"),
                                                     ("radial", (256, 64, 1), "mixed")])]
    be = GpuBackend(dev=dev)
    t0 = time.perf_counter()
    be.prepare(reqs + warm, slots=max_active)
    prep = time.perf_counter() - t0
    Scheduler(be, max_active=max_active).run(warm, keep_kspace=False)
    log(f"serving: {len(reqs)} requests at {rate}/s (resident set-up {prep:.1f} s)")
    rep = Scheduler(be, max_active=max_active).run(reqs, keep_kspace=False)
    out = rep.summary()
    lat = rep.latencies()
    by = {}
    for r in reqs:
        res = rep.results[r.rid]
        if res.error is None:
            by.setdefault(r.precision, []).append(res.latency)
    out.update({"rate_per_s": rate, "max_active": max_active, "resident_setup_s": prep,
                "mean_latency_s": float(lat.mean()) if lat.size else None,
                "p99_latency_by_form_s": {k: float(np.percentile(v, 99)) for k, v in by.items()},
                "stream": "Poisson arrivals; N log-uniform 10k-2M, served exactly (the first N nodes of a resident "
                          "box phantom, tissue drawn per request); steady states of one admission pass batched "
                          "into one graph replay; "
                          "Cartesian 2-D 64^2-256^2, 3-D 64^2-256^2 x 32, radial 64-512 spokes x 256-512; "
                          "FP64 or mixed"})
    return out


def _dtype(method: str, form: str, line_gemm: bool = False) -> str:
    """The type the dominant kernel computes in: bf16 MFMA inputs on the
    Cartesian and per-line GEMMs, FP32 on the direct kernels, FP64 in the
    FP64 form."""
    if form == "f64":
        return "f64"
    return "bf16" if method == "cartesian" or line_gemm else "f32"


def _arithmetic(method: str, form: str, line_gemm: bool = False) -> str:
    if form == "f64":
        return "FP64 throughout (FP64 MFMA on Cartesian plans)"
    if method == "cartesian":
        return "bf16 hi+lo operands, 3 products per real product, FP32 accumulation, FP64 split reduction"
    if line_gemm:
        return ("per-line GEMM S[Qp+q] = sum_n (c e Z^p)(z^q): FP64 phases range-reduced for FP32 sine/cosine, "
                "FP32 operand chains, bf16 hi+lo operands with 3 products per real product, FP32 accumulation, "
                "FP64 split reduction")
    return "FP64 phase and range reduction, FP32 sine/cosine and accumulation, FP64 chunk sums"


def run_flow(args, dev) -> dict:
    """c3: 4-D flow.  Set-up on the GPU: lumped weights, the whole
    phase-contrast acquisition through the Bloch integrator with nodes moving
    along the Poiseuille velocity field, then per cardiac phase the
    packed nodes (4 encodings) and the POD displacement tables.  A step is
    the k-space of every phase (all encodings) with moving nodes."""
    import numpy as np
    import torch

    import oracle
    from feelmri_amd import Motion, device, workloads
    from feelmri_amd.Bloch import BlochSolver

    wl = workloads.c3(args.scale)
    n_ph = wl.n_phases if args.phases <= 0 else min(args.phases, wl.n_phases)
    m, tis = wl.mesh, wl.tissue
    log(f"c3: {m.n_nodes} nodes, {m.n_elems} tets, {wl.n_samples} k-samples x 4 encodings x {n_ph} phases; "
        "Bloch set-up on GPU")
    col = device.P1Colouring(m.elems, m.n_nodes, device=dev)
    nodes_d = torch.from_numpy(m.nodes).to(dev)
    blocks = workloads.flow_sequence(wl, n_dummy=8, n_phases=n_ph)
    solver = BlochSolver(blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT, pod_trajectory=wl.velocity,
                         dev=dev)
    mxy0, mz0 = solver.mxy.clone(), solver.mz.clone()
    # Cartesian plan: with moving nodes both forms run the per-readout GEMM
    plan = device.SignalPlan(wl.kloc, wl.t, method="auto", device=dev)
    modes_pcr = Motion.motion_modes(wl.motion)
    modes = device.modes_soa(torch.from_numpy(modes_pcr).to(dev))
    M0 = torch.from_numpy(tis.M0).to(dev)
    phi = torch.from_numpy(tis.delta_omega).to(dev)
    T2s = torch.from_numpy(tis.T2star).to(dev)
    coefs = []
    for p in range(n_ph):
        wl.motion.update_timeshift(p * wl.period / wl.n_phases)
        coefs.append(torch.from_numpy(Motion.motion_coefficients(wl.motion, plan.tgroup_t)).to(dev))
    wl.motion.update_timeshift(0.0)

    # the whole scan once untimed (kernels loaded), then timed: weights, the
    # moving-node Bloch acquisition, and the packed nodes of every phase
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    scan_ms = None
    for rep in range(2):
        solver.mxy.copy_(mxy0)
        solver.mz.copy_(mz0)
        torch.cuda.synchronize()
        ev[0].record()
        w = device.lumped_mass(col, nodes_d)
        ev[1].record()
        mxy_store, _ = solver.solve_device()                       # (N, 4 n_ph) complex64
        ev[2].record()
        torch.cuda.synchronize()
        scan_ms = (ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]))
    # the mixed form's acquisition: the same blocks with FP32 Bloch rotations
    # (timed like the FP64 one; its Mxy feeds the mixed end-to-end scan)
    solver32 = BlochSolver(blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT, pod_trajectory=wl.velocity,
                           precision="f32", dev=dev)
    m32xy0, m32z0 = solver32.mxy.clone(), solver32.mz.clone()
    bloch32_ms = None
    for rep in range(2):
        solver32.mxy.copy_(m32xy0)
        solver32.mz.copy_(m32z0)
        torch.cuda.synchronize()
        ev[1].record()
        mxy32_store, _ = solver32.solve_device()
        ev[2].record()
        torch.cuda.synchronize()
        bloch32_ms = ev[1].elapsed_time(ev[2])
    del solver32
    bloch32_err = float(torch.linalg.vector_norm(mxy32_store - mxy_store) / torch.linalg.vector_norm(mxy_store))
    log(f"FP32 moving-node acquisition {bloch32_ms:.1f} ms, Mxy relL2 {bloch32_err:.2e} vs FP64")
    phases = []
    for p in range(n_ph):
        mxy = mxy_store[:, 4 * p:4 * p + 4].to(torch.complex128) * M0[:, None]
        nodes = device.SignalNodes.from_arrays(nodes_d, phi, T2s, mxy, weights=w, device=dev)
        phases.append((nodes, coefs[p], mxy))
    torch.cuda.synchronize()
    # packing every phase as a scan does it, one phase at a time into reused
    # buffers (the second pass is timed; the first one fills the allocator)
    pack_ms = None
    for rep in range(2):
        torch.cuda.synchronize()
        t_pack = time.perf_counter()
        for p in range(n_ph):
            mxy_p = mxy_store[:, 4 * p:4 * p + 4].to(torch.complex128) * M0[:, None]
            nodes_p = device.SignalNodes.from_arrays(nodes_d, phi, T2s, mxy_p, weights=w, device=dev)
            del nodes_p, mxy_p
        torch.cuda.synchronize()
        pack_ms = (time.perf_counter() - t_pack) * 1e3
    work = m.n_nodes * wl.n_samples * 4 * n_ph
    log(f"set-up: lumped mass {scan_ms[0]:.3f} ms, moving-node Bloch acquisition ({len(blocks)} blocks) "
        f"{scan_ms[1]:.1f} ms")

    def time_it(form, steps, warmup):
        outs = [torch.empty((plan.n_samples, 4), dtype=torch.complex128, device=dev) for _ in phases]
        ws = torch.empty(max(plan.workspace_bytes(form, m.n_nodes, 4, moving=True), 1), dtype=torch.uint8,
                         device=dev)

        def step():
            for (nodes, coef, _), o in zip(phases, outs):
                plan.execute(nodes, precision=form, motion=(modes, coef), out=o, workspace=ws)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        plan.set_timing(steps * len(phases))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        k_ms, k_n = plan.kernel_ms()
        plan.set_timing(0)
        return outs, e0.elapsed_time(e1) / steps, (k_ms / k_n if k_n else None)

    outs, ms, k_ms = time_it(args.form, args.steps, args.warmup)
    value = work / (ms * 1e-3)
    log(f"{args.form}: {ms:.1f} ms/step ({n_ph} phases), {value:.3e} node*k/s, dominant kernel {k_ms:.2f} ms/phase")
    forms = {args.form: {"ms_per_step": ms, "node_ksamples_per_s": value, "kernel_ms_per_phase": k_ms}}
    err64 = err64_img = scan_err = None
    out0 = outs[0].cpu().numpy()
    o64_0 = None
    if not args.no_f64 and args.form != "f64":
        o64, ms64, k64 = time_it("f64", 1, 0)
        forms["f64"] = {"ms_per_step": ms64, "node_ksamples_per_s": work / (ms64 * 1e-3), "kernel_ms_per_phase": k64}
        num = den = 0.0
        inum = iden = 0.0
        from feelmri_amd import recon
        kshape = tuple(np.asarray(wl.kloc[0]).shape)
        geo = recon.CartesianGeometry(kshape)          # one readout sample per voxel (no oversampling)
        for a, b in zip(outs, o64):
            num += float(torch.linalg.vector_norm(a - b)) ** 2
            den += float(torch.linalg.vector_norm(b)) ** 2
            # CartesianRecon image of all 4 encodings, mixed vs FP64
            ia = recon.CartesianRecon(a.cpu().numpy().reshape(*kshape, 4), geo, device=dev)
            ib = recon.CartesianRecon(b.cpu().numpy().reshape(*kshape, 4), geo, device=dev)
            inum += float(np.linalg.norm(ia - ib)) ** 2
            iden += float(np.linalg.norm(ib)) ** 2
        err64 = (num / den) ** 0.5
        err64_img = (inum / iden) ** 0.5
        o64_0 = o64[0].cpu().numpy()
        # the mixed scan end to end (FP32 Bloch Mxy, mixed signal) against
        # the FP64 scan (FP64 Bloch Mxy, FP64 signal), all phases
        num = 0.0
        ws32 = torch.empty(max(plan.workspace_bytes(args.form, m.n_nodes, 4, moving=True), 1), dtype=torch.uint8,
                           device=dev)
        o32 = torch.empty((plan.n_samples, 4), dtype=torch.complex128, device=dev)
        for p, b in enumerate(o64):
            mxy = mxy32_store[:, 4 * p:4 * p + 4].to(torch.complex128) * M0[:, None]
            nd32 = device.SignalNodes.from_arrays(nodes_d, phi, T2s, mxy, weights=w, device=dev)
            plan.execute(nd32, precision=args.form, motion=(modes, coefs[p]), out=o32, workspace=ws32)
            num += float(torch.linalg.vector_norm(o32 - b)) ** 2
        scan_err = (num / den) ** 0.5
        del ws32, o32, nd32
        log(f"mixed scan (FP32 Bloch + {args.form} signal) vs FP64 scan: relL2 {scan_err:.2e}")
        del o64
        log(f"f64: {ms64:.1f} ms/step; {args.form} vs f64 relL2 {err64:.2e} over all phases (image "
            f"{err64_img:.2e}); speed-up {ms64 / ms:.2f}x")
    # FP64 oracle on sampled k-points of phase 0: positions at each sample's t
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(wl.n_samples, size=12, replace=False))
    kx, ky, kz = (np.asarray(k).ravel()[idx] for k in wl.kloc)
    tt = np.asarray(wl.t).ravel()[idx]
    w_h = w.cpu().numpy()
    mxy0_h = phases[0][2].cpu().numpy()
    mm = w_h[:, None] * mxy0_h
    ref = np.stack([oracle.signal_samples(kx[i:i + 1], ky[i:i + 1], kz[i:i + 1], tt[i:i + 1],
                                          m.nodes + np.asarray(wl.motion(tt[i])), tis.delta_omega, tis.T2star,
                                          mm, n_threads=min(16, os.cpu_count() or 1))[0]
                    for i in range(idx.size)])
    err_or = rel_l2(out0[idx], ref)
    log(f"relL2 vs FP64 oracle on {idx.size} sampled k-points of phase 0: {err_or:.2e}")
    k_s = (k_ms or ms / n_ph) * 1e-3   # dominant kernel, per phase
    if plan.method == "cartesian" and args.form == "mixed":
        # 8 flops per complex MAC per node*k-sample*encoding; the kernel
        # executes 3 bf16 products per real product (24)
        achieved = 8.0 * m.n_nodes * wl.n_samples * 4 / k_s / 1e12
        roof = {"bound": "mfma", "kernel": "k_cart_moving_bf16", "achieved": achieved,
                "peak": PEAKS["bf16_mfma_tflops"], "unit": "TFLOP/s",
                "frac": achieved / PEAKS["bf16_mfma_tflops"],
                "traffic": load_traffic("k_cart_moving_bf16", "c3", "mixed"), "kernel_ms": k_ms,
                "executed": 3.0 * achieved,
                "per_unit": "8 flops per node*k-sample*encoding (complex MAC); one launch is one cardiac phase "
                            "(1948800 nodes x 221184 k-samples x 4 encodings); operands generated in the kernel "
                            "(FP64 phase, FP32 sin/cos per node and line / kz, bf16 hi+lo split)"}
    elif plan.method == "cartesian":
        achieved = 8.0 * m.n_nodes * wl.n_samples * 4 / k_s / 1e12
        roof = {"bound": "mfma", "kernel": "k_cart_moving_f64", "achieved": achieved,
                "peak": PEAKS["fp64_mfma_tflops"], "unit": "TFLOP/s",
                "frac": achieved / PEAKS["fp64_mfma_tflops"],
                "traffic": load_traffic("k_cart_moving_f64", "c3", "f64"), "kernel_ms": k_ms,
                "per_unit": "8 flops per node*k-sample*encoding (complex MAC) on FP64 MFMA; operands "
                            "generated in the kernel (FP64 sincospi per node and line / kz)"}
    else:
        peak = PEAKS["fp32_valu_tflops"] if args.form == "mixed" else PEAKS["fp64_valu_tflops"]
        achieved = (8.0 + 8.0 * 4) * m.n_nodes * wl.n_samples / k_s / 1e12
        roof = {"bound": "valu", "kernel": "k_signal_direct (moving nodes)", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "per_unit": "8 flops + 8 per encoding and 2 transcendentals per node*k-sample"}

    cpu = None
    err_cpu = {}
    if not args.no_cpu:
        log(f"CPU baseline: {CPU_RANKS} forked ranks (~{args.cpu_seconds:.0f} s)")
        mot = (modes_pcr, plan.tgroup_t, coefs[0].cpu().numpy())
        cpu, ro, ks = cpu_baseline(wl, w_h, mxy0_h, args.cpu_seconds, motion=mot)
        shape = (wl.t.shape[0], -1, 4)
        sel = lambda o: o.reshape(shape)[ro].reshape(-1, 4)  # noqa: E731
        err_cpu[args.form] = rel_l2(sel(out0), ks)
        if o64_0 is not None:
            err_cpu["f64"] = rel_l2(sel(o64_0), ks)
        log(f"CPU: {cpu['value']:.3e} node*k/s on {cpu['cores']} ranks; GPU vs CPU on its {ks.shape[0]} k-samples "
            f"x 4 encodings: {err_cpu}")
    return {
        "metric": METRIC, "value": value, "unit": "node*k-sample/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": _dtype(plan.method, args.form), "data": "synthetic",
        "config": {"workload": wl.info["desc"], "config": "c3", "nodes": m.n_nodes, "elements": m.n_elems,
                   "k_samples": wl.n_samples, "encodings": 4, "phases": n_ph, "method": plan.method,
                   "form": args.form, "arithmetic": _arithmetic(plan.method, args.form)},
        "relL2_vs_f64_gpu": err64, "image_relL2_vs_f64_gpu": err64_img, "relL2_vs_oracle_sampled": err_or,
        "relL2_vs_cpu_baseline_readouts": err_cpu,
        "forms": forms, "roofline": roof, "cpu_baseline": cpu,
        "setup_ms": {"lumped_mass": scan_ms[0], "bloch_moving_acquisition": scan_ms[1],
                     "bloch_moving_acquisition_f32": bloch32_ms, "bloch_blocks": len(blocks),
                     "pack_all_phases": pack_ms},
        # end to end: weights, the moving-node Bloch acquisition, the packed
        # nodes of every phase and the k-space of every phase (FP32 Bloch in
        # the mixed form, FP64 in the FP64 form)
        "scan": scan_entry(args.form, scan_ms[0], bloch32_ms if args.form != "f64" else scan_ms[1], pack_ms, ms,
                           work, scan_err, bloch32_err),
    }


def scan_entry(form, mass_ms, bloch_ms, pack_ms, signal_ms, work, err_vs_f64_scan, bloch_err) -> dict:
    total = mass_ms + bloch_ms + pack_ms + signal_ms
    return {"ms": total, "node_ksamples_per_s": work / (total * 1e-3),
            "parts_ms": {"lumped_mass": mass_ms, "bloch": bloch_ms, "pack": pack_ms, "signal": signal_ms},
            "bloch_precision": "f64" if form == "f64" else "f32",
            "relL2_vs_f64_scan": err_vs_f64_scan if form != "f64" else None,
            "bloch_mxy_relL2_f32_vs_f64": bloch_err}


def measure_static(args, dev, wl, form: str, with_cpu: bool) -> dict:
    """c0 / c1 / c2: the static-node signal of one Bloch steady state."""
    from feelmri_amd import device

    log(f"{wl.name}: {wl.n_nodes} nodes, {wl.mesh.n_elems} tets, {wl.n_samples} k-samples; preparing on GPU")
    nodes, w, mxy = prepare_nodes(wl, dev)
    plan = device.SignalPlan(wl.kloc, wl.t, method=args.method, device=dev)
    log(f"plan: method {plan.method}, {plan.n_tgroups} distinct t")
    work = wl.node_ksamples
    steps = max(1, args.steps)
    out, ms_ev, ms_wall, ms_kernel = time_form(plan, nodes, form, steps, args.warmup)
    value = work / (ms_ev * 1e-3)
    log(f"{form}: {ms_ev:.3f} ms/step (events), {ms_wall:.3f} ms/step (wall), {value:.3e} node*k/s")
    out_h = out.cpu().numpy()

    forms = {form: {"ms_per_step": ms_ev, "node_ksamples_per_s": value, "kernel_ms": ms_kernel}}
    err_f64 = err_img = err_img_tukey = None
    out64_h = None
    if not args.no_f64 and form != "f64":
        ref_steps = 1 if plan.method == "direct" else steps
        out64, ms64, _, ms64_kernel = time_form(plan, nodes, "f64", ref_steps, 1)
        forms["f64"] = {"ms_per_step": ms64, "node_ksamples_per_s": work / (ms64 * 1e-3),
                        "kernel_ms": ms64_kernel}
        out64_h = out64.cpu().numpy()
        err_f64 = rel_l2(out_h, out64_h)
        if wl.name in ("c0", "c1"):
            from feelmri_amd.recon import image_rel_l2

            shape = (*wl.t.shape, wl.n_enc)
            err_img = image_rel_l2(out_h.reshape(shape), out64_h.reshape(shape))
            # CartesianRecon's Tukey-windowed image (Recon.py:13-90, before its resize)
            err_img_tukey = image_rel_l2(out_h.reshape(shape), out64_h.reshape(shape), "Tukey")
        elif wl.name.startswith("c2"):
            import numpy as np

            from feelmri_amd.recon import nufft_image_rel_l2

            # 2-D radial / spiral: gridding-NUFFT adjoint with ramp density compensation on the
            # smallest even grid over the 0.24 m FOV that holds the trajectory's band
            kmax = float(np.max(np.hypot(wl.kloc[0], wl.kloc[1])))
            n_img = 2 * int(np.ceil(kmax * 0.24)) + 2
            err_img = nufft_image_rel_l2(wl.kloc[0], wl.kloc[1], out_h[:, 0], out64_h[:, 0], 0.24, n=n_img,
                                         device=dev)
        log(f"f64: {ms64:.3f} ms/step; {form} vs f64 relL2 {err_f64:.2e} (image {err_img}); "
            f"speed-up {ms64 / ms_ev:.2f}x")
    w_h = w.cpu().numpy()
    mxy_h = mxy.cpu().numpy()
    ms_host = time_host_io(plan, wl, w_h, mxy_h, form, max(3, steps // 4))
    log(f"{form} from host arrays (H2D nodes + pack + execute + D2H k-space): {ms_host:.3f} ms/step")
    forms[form]["host_io_ms_per_step"] = ms_host
    forms[form]["host_io_node_ksamples_per_s"] = work / (ms_host * 1e-3)
    if wl.name == "c0":
        ms_lines, n_lines = time_shim_per_line(wl, w_h, mxy_h, form)
        log(f"{form} per-line drop-in calls ({n_lines} x MRI.Signal from host arrays): {ms_lines:.2f} ms "
            f"per k-space, {ms_lines / n_lines:.3f} ms per call")
        forms[form]["per_line_shim_ms_per_kspace"] = ms_lines
        forms[form]["per_line_shim_calls"] = n_lines
    scan = time_scan(wl, dev, plan, form)
    log(f"end-to-end scan (weights + Bloch + pack + signal): {scan['ms']:.3f} ms, "
        f"{scan['node_ksamples_per_s']:.3e} node*k/s, parts {scan['parts_ms']}")
    err_oracle = oracle_subset_error(wl, w_h, mxy_h, out_h)
    log(f"relL2 vs FP64 oracle on sampled k-points: {err_oracle:.2e}")

    # roofline of the dominant kernel
    if plan.method == "cartesian":
        bound, unit = "mfma", "TFLOP/s"
        flops = 8.0 * work
        peak = PEAKS["bf16_mfma_tflops"] if form == "mixed" else PEAKS["fp64_mfma_tflops"]
        dom = "k_cart_gemm"
    elif form == "mixed" and plan.line_gemm:
        bound, unit = "mfma", "TFLOP/s"
        flops = 8.0 * work
        peak = PEAKS["bf16_mfma_tflops"]
        dom = "k_signal_lines_bf16"
    elif form == "mixed" and plan.n_runs > 0:
        bound, unit = "valu", "TFLOP/s"
        flops = (8.0 + 8.0 * wl.n_enc) * wl.n_nodes * wl.n_samples
        peak = PEAKS["fp32_valu_tflops"]
        dom = "k_signal_runs"
    else:
        bound, unit = "valu", "TFLOP/s"
        flops = (8.0 + 8.0 * wl.n_enc) * wl.n_nodes * wl.n_samples
        peak = PEAKS["fp64_valu_tflops"] if form == "f64" else PEAKS["fp32_valu_tflops"]
        dom = "k_signal_direct"
    ms_dom = ms_kernel if ms_kernel else ms_ev
    achieved = flops / (ms_dom * 1e-3) / 1e12
    roof = {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": load_traffic(dom, wl.name, form), "kernel_ms": ms_dom,
            "per_unit": ("8 flops per node*k-sample*encoding (complex MAC); the mixed GEMM executes 3 bf16 "
                         "products (hi*hi + hi*lo + lo*hi) per real product, i.e. 24" if bound == "mfma"
                         else "8 flops (k.x 3 FMA + phi t 1 FMA) + 8 per encoding (complex MAC) and 2 "
                              "transcendentals per node*k-sample")}
    if bound == "mfma" and form == "mixed":
        roof["executed"] = 3.0 * achieved
    cpu = None
    err_cpu = {}
    if with_cpu:
        log(f"CPU baseline: {CPU_RANKS} forked ranks (~{args.cpu_seconds:.0f} s)")
        cpu, ro, ks = cpu_baseline(wl, w_h, mxy_h, args.cpu_seconds)
        V = wl.n_enc
        sel = lambda o: o.reshape(wl.t.shape[0], -1, V)[ro].reshape(-1, V)  # noqa: E731
        err_cpu[form] = rel_l2(sel(out_h), ks)
        if out64_h is not None:
            err_cpu["f64"] = rel_l2(sel(out64_h), ks)
        log(f"CPU: {cpu['value']:.3e} node*k/s on {cpu['cores']} ranks; GPU vs CPU: {err_cpu}")
    return {
        "value": value, "ms_per_step": ms_ev, "dtype": _dtype(plan.method, form, plan.line_gemm),
        "config": {"workload": wl.info["desc"], "config": wl.name, "nodes": wl.n_nodes,
                   "k_samples": wl.n_samples, "encodings": wl.n_enc, "method": plan.method, "form": form,
                   "arithmetic": _arithmetic(plan.method, form, plan.line_gemm)},
        "relL2_vs_f64_gpu": err_f64, "image_relL2_vs_f64_gpu": err_img,
        "image_relL2_tukey_vs_f64_gpu": err_img_tukey, "relL2_vs_oracle_sampled": err_oracle,
        "relL2_vs_cpu_baseline_readouts": err_cpu, "forms": forms, "roofline": roof, "cpu_baseline": cpu,
        "scan": scan,
    }


def _time_cuda(fn, reps: int) -> float:
    import torch

    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


BLOCH_FLOPS = 80.0   # per node*event: the rotation, relaxation and |B| of BlochSimulator.cpp:231-271 (static)


def row_benchmarks(wl, dev) -> dict:
    """Throughput and roofline of the other rows of the path on the workload's
    mesh: the Bloch integrator (RF-sparse event rows, the RF-heavy scan chain
    and full history), P1 lumped mass, and the POD contraction at the 4-D
    flow size (2 M nodes, 10 modes).  Rooflines count the algorithmic work
    per unit: 80 flops per node*event for Bloch whatever the event, 16 B per
    element + 32 B per node for the lumped mass."""
    import numpy as np
    import torch

    from feelmri_amd import device
    from feelmri_amd.Bloch import BlochSolver
    from feelmri_amd.sequences import GAMMA

    m, tis = wl.mesh, wl.tissue
    N = m.n_nodes
    rows = {}
    # Bloch: 1000 events of 10 us (hard pulse over 20 events, then random gradients)
    T = 1000
    rng = np.random.default_rng(0)
    rf = np.zeros(T, dtype=np.complex128)
    rf[1:21] = 0.01 * np.exp(0.3j)
    G = np.zeros((T, 3))
    G[21:] = rng.normal(0.0, 5.0, (T - 21, 3))
    ev = device.EventTable(rf, G, np.full(T, 0.01), device=dev)
    bn = device.BlochNodes.from_arrays(m.nodes, tis.T1, tis.T2, tis.delta_B_mT, device=dev)
    mxy0 = torch.zeros(N, dtype=torch.complex128, device=dev)
    mz0 = torch.ones(N, dtype=torch.float64, device=dev)
    for prec, peak, key in (("f64", PEAKS["fp64_valu_tflops"], "bloch_f64"), ("f32", PEAKS["fp32_valu_tflops"], "bloch_f32")):
        ms = _time_cuda(lambda: device.bloch(bn, ev, 1.0, GAMMA, mxy0, mz0, precision=prec), 5)
        ach = BLOCH_FLOPS * N * (T - 1) / (ms * 1e-3) / 1e12
        rows[key] = {"node_events_per_s": N * (T - 1) / (ms * 1e-3), "ms": ms, "nodes": N, "events": T,
                     "rf_events": 20,
                     "roofline": {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                                  "frac": ach / peak,
                                  "per_unit": "80 flops + sqrt, rcp, sin, cos per node*event (final state kept); "
                                              "979 of 999 steps carry no RF and take the precession the rotation "
                                              "reduces to (~18 flops + sin, cos), so this row flatters the kernel; "
                                              "see bloch_scan_chain for an RF-heavy chain"}}
    # the RF-heavy scan chain: the workload's spoiled-GRE blocks in one chained launch
    blocks = wl.blocks
    n_ev = sum(len(b.dt) - 1 for b in blocks)
    n_rf = sum(int(np.count_nonzero(np.asarray(b.rf)[1:])) for b in blocks)
    for prec, peak in (("f32", PEAKS["fp32_valu_tflops"]), ("f64", PEAKS["fp64_valu_tflops"])):
        solver = BlochSolver(blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT, precision=prec, dev=dev)
        x0, z0 = solver.mxy.clone(), solver.mz.clone()

        def run():
            solver.mxy.copy_(x0)
            solver.mz.copy_(z0)
            solver.solve_device()
        ms_copy = _time_cuda(lambda: (solver.mxy.copy_(x0), solver.mz.copy_(z0)), 5)
        ms = _time_cuda(run, 5) - ms_copy
        ach = BLOCH_FLOPS * N * n_ev / (ms * 1e-3) / 1e12
        rows[f"bloch_scan_chain_{prec}"] = {
            "node_events_per_s": N * n_ev / (ms * 1e-3), "ms": ms, "nodes": N, "events": n_ev, "rf_events": n_rf,
            "blocks": len(blocks),
            "roofline": {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                         "per_unit": f"80 flops + sqrt, rcp, sin, cos per node*event; {n_rf} of {n_ev} steps carry "
                                     "RF; the spoiled-GRE scan chain (16 dummy TRs + 1) as one chained launch"}}
        del solver
    # event tables on the device (fm_events_build) against the host grid rule
    # (make_block, as the reference's SequenceBlock discretises) for the
    # same block descriptions: the scan's spoiled-GRE blocks
    from feelmri_amd import sequences
    desc, pulses, traps, tabs = sequences.block_descriptions(blocks)
    de = device.DeviceEvents(desc, pulses, traps, rf_tables=tabs, device=dev)
    ms_dev = _time_cuda(lambda: de.build(), 20)
    de.check()
    n_built = int(de.offsets[-1].item())
    specs = [b.info["spec"] for b in blocks]
    t0 = time.perf_counter()
    for _ in range(5):
        for sp in specs:
            sequences.make_block(**sp)
    ms_host = (time.perf_counter() - t0) * 1e3 / 5
    rows["events_build"] = {"blocks": len(blocks), "events": n_built, "ms": ms_dev, "host_rule_ms": ms_host,
                            "max_candidates": de.max_candidates,
                            "note": "fm_events_build (three launches: candidates sorted per block in LDS, "
                                    "offsets scanned, tables written) for the scan's blocks, against the "
                                    "host grid rule on one core for the same specifications"}
    Th = 200
    evh = device.EventTable(rf[:Th], G[:Th], np.full(Th, 0.01), device=dev)
    ms = _time_cuda(lambda: device.bloch(bn, evh, 1.0, GAMMA, mxy0, mz0, precision="f32", history=True), 3)
    gbs = 12.0 * N * Th / (ms * 1e-3) / 1e9
    rows["bloch_f32_history"] = {"node_events_per_s": N * Th / (ms * 1e-3), "ms": ms, "nodes": N, "events": Th,
                                 "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAKS["hbm_gbs"],
                                              "unit": "GB/s", "frac": gbs / PEAKS["hbm_gbs"],
                                              "per_unit": "12 B written per node*event (complex64 Mxy + float32 Mz "
                                                          "column-major history)"}}
    # P1 lumped mass
    col = device.P1Colouring(m.elems, N, device=dev)
    nodes_d = torch.from_numpy(m.nodes).to(dev)
    out = torch.empty(N, dtype=torch.float64, device=dev)
    ms = _time_cuda(lambda: device.lumped_mass(col, nodes_d, out=out), 5)
    ms_coloured = _time_cuda(lambda: device.lumped_mass_coloured(col, nodes_d, out=out), 5)
    E = m.n_elems
    by = E * 16 + N * 32   # compulsory: connectivity read, coordinates read, weights written
    rows["lumped_mass_p1"] = {"elements_per_s": E / (ms * 1e-3), "ms": ms, "elements": E, "colours": col.n_colours,
                              "coloured_scatter_ms": ms_coloured, "method": col.method,
                              "roofline": {"bound": "hbm", "achieved": by / (ms * 1e-3) / 1e9,
                                           "peak": PEAKS["hbm_gbs"], "unit": "GB/s",
                                           "frac": by / (ms * 1e-3) / 1e9 / PEAKS["hbm_gbs"],
                                           "traffic": load_traffic("lumped_mass", "c1", "f64"),
                                           "per_unit": "compulsory bytes: 16 B per element (4 x int32 connectivity) "
                                                       "+ 24 B per node coordinates read + 8 B per node weight "
                                                       "written"}}
    # consistent P1 mass: element matrices on the device, the CSR assembled
    # once on the host (scipy), then M . Mxy as a device SpMV
    import scipy.sparse as sp
    elems_d = torch.from_numpy(np.ascontiguousarray(m.elems, dtype=np.int32)).to(dev)
    ms_em = _time_cuda(lambda: device.element_matrices(elems_d, nodes_d), 5)
    vals = device.element_matrices(elems_d, nodes_d).cpu().numpy()
    el = np.ascontiguousarray(m.elems, dtype=np.int64)
    Mc = sp.csr_matrix((vals.ravel(), (np.repeat(el, 4, axis=1).ravel(), np.tile(el, (1, 4)).ravel())), shape=(N, N))
    ip = torch.from_numpy(Mc.indptr.astype(np.int64)).to(dev)
    ix = torch.from_numpy(Mc.indices.astype(np.int64)).to(dev)
    dv = torch.from_numpy(Mc.data).to(dev)
    xv = torch.randn(N, 1, dtype=torch.complex128, device=dev)
    ms_spmv = _time_cuda(lambda: device.csr_spmv(ip, ix, dv, xv), 10)
    by_em = E * (16 + 128) + N * 24
    by_spmv = Mc.nnz * 16 + (N + 1) * 8 + N * 32
    rows["consistent_mass_p1"] = {
        "elements": E, "nnz": int(Mc.nnz), "ms": ms_em + ms_spmv, "element_matrices_ms": ms_em, "spmv_ms": ms_spmv,
        "roofline": {"bound": "hbm", "achieved": by_spmv / (ms_spmv * 1e-3) / 1e9, "peak": PEAKS["hbm_gbs"],
                     "unit": "GB/s", "frac": by_spmv / (ms_spmv * 1e-3) / 1e9 / PEAKS["hbm_gbs"],
                     "element_matrices_gbs": by_em / (ms_em * 1e-3) / 1e9,
                     "per_unit": "SpMV: 16 B per stored entry (value + column) + 8 B per row pointer + 32 B per "
                                 "node (complex x read, y written); element matrices: 16 B connectivity + 128 B "
                                 "written per element + 24 B per node coordinates"}}
    del Mc, ip, ix, dv, xv, vals
    # POD contraction at 2 M nodes, 10 modes (float32)
    P, R = 2_000_000, 10
    modes = torch.randn((P, 3, R), dtype=torch.float32, device=dev)
    wts = torch.randn(R, dtype=torch.float32, device=dev)
    ms = _time_cuda(lambda: device.pod_tensordot(modes, wts), 10)
    by = P * 3 * R * 4 + P * 3 * 4
    rows["pod_tensordot"] = {"nodes": P, "modes": R, "ms": ms,
                             "roofline": {"bound": "hbm", "achieved": by / (ms * 1e-3) / 1e9, "peak": PEAKS["hbm_gbs"],
                                          "unit": "GB/s", "frac": by / (ms * 1e-3) / 1e9 / PEAKS["hbm_gbs"],
                                          "per_unit": "12R B read + 12 B written per node"}}
    del modes
    return rows


def oracle_subset_error(wl, w, mxy_host, out, n_pick=96):
    import numpy as np

    import oracle

    rng = np.random.default_rng(7)
    S = wl.n_samples
    idx = np.sort(rng.choice(S, size=min(n_pick, S), replace=False))
    kx, ky, kz = (np.asarray(k).ravel()[idx] for k in wl.kloc)
    t = np.asarray(wl.t).ravel()[idx]
    ref = oracle.signal_samples(kx, ky, kz, t, wl.mesh.nodes, wl.tissue.delta_omega, wl.tissue.T2star,
                                w[:, None] * mxy_host, n_threads=min(16, os.cpu_count() or 1))
    return rel_l2(out[idx], ref)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps first; the clock ramps up over the first few launches")
    ap.add_argument("--form", default="mixed", choices=["mixed", "f64"])
    ap.add_argument("--method", default="auto", choices=["auto", "direct", "cartesian"])
    ap.add_argument("--config", default="c3")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--phases", type=int, default=0, help="c3: cardiac phases per step (0: all)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 comparison form")
    ap.add_argument("--no-rows", action="store_true", help="skip the Bloch / mass / POD row measurements")
    ap.add_argument("--no-secondary", action="store_true", help="c3: skip the c1 secondary block")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="requests in the serving stream reported beside the throughput (0: none)")
    ap.add_argument("--rate", type=float, default=20.0, help="serving stream arrival rate (requests/s)")
    ap.add_argument("--serve-slots", type=int, default=8, help="concurrent requests (stream slots) in serving")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size kernel for "
                         "scaling hardware traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "node-sharded k-space partial sums would be combined with one reduce over the ranks",
              file=sys.stderr)
        return 2

    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + written")
        del src, dst
    if args.config == "c4":
        srv = serving_run(dev, max(1, args.serve_requests), args.rate, args.serve_slots)
        log(f"serving: p50 {srv['p50_latency_s']:.3f} s, p99 {srv['p99_latency_s']:.3f} s, "
            f"{srv['node_ksamples_per_s']:.3e} node*k/s")
        print(json.dumps({
            "metric": "p99 request latency, Poisson stream of mixed synthetic scans",
            "value": srv["p99_latency_s"], "unit": "s", "n_gpus": 1, "steps": srv["requests"], "warmup": 3,
            "ms_per_step": 1e3 * srv["wall_s"] / max(1, srv["requests"]), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 / mixed per request", "data": "synthetic",
            "config": {"workload": f"serving stream, {srv['requests']} requests at {args.rate}/s", "config": "c4"},
            "serving": srv}), flush=True)
        return 0

    if args.config == "c3":
        line = run_flow(args, dev)
        torch.cuda.empty_cache()
        aux_wl = None
        if not args.no_secondary or not args.no_rows:
            aux_wl = build_workload("c1", args.scale)
        if not args.no_secondary:
            line["secondary"] = {"c1": measure_static(args, dev, aux_wl, args.form, with_cpu=False)}
    else:
        aux_wl = build_workload(args.config, args.scale)
        sec = measure_static(args, dev, aux_wl, args.form, with_cpu=not args.no_cpu)
        line = {"metric": METRIC, "value": sec.pop("value"), "unit": "node*k-sample/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec.pop("ms_per_step"),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": sec.pop("dtype"),
                "data": "synthetic", **sec}
    if not args.no_rows and aux_wl is not None:
        rows = row_benchmarks(aux_wl, dev)
        log("rows: " + ", ".join(f"{k} {v['ms']:.3f} ms" + (f" ({v['roofline']['frac']:.2f} of {v['roofline']['bound']})"
                                                             if "roofline" in v else "")
                                  for k, v in rows.items()))
        line["rows"] = rows
    if args.serve_requests > 0:
        srv = serving_run(dev, args.serve_requests, args.rate, args.serve_slots)
        log(f"serving: p50 {srv['p50_latency_s']:.3f} s, p99 {srv['p99_latency_s']:.3f} s, "
            f"{srv['node_ksamples_per_s']:.3e} node*k/s")
        line["serving"] = srv
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
