"""Parity at the benchmark workloads' full sizes.

The whole FP64 oracle is too slow at these sizes, so each case checks what
does not need it:
  * the mixed form against the FP64 GPU form over every k-sample (<= 1e-4);
  * the FP64 GPU form against the FP64 oracle at sampled k-points (<= 1e-10;
    1e-9 for moving nodes, whose positions the oracle forms in a different
    order);
  * additivity over node partitions (the reference's MPI sum of per-rank
    partial k-spaces, MPIUtilities.py:13-26): the FP64 sum over all nodes
    equals the sum of the two halves' (<= 1e-12).
Node coefficients are random (the set-up that produces them -- lumped mass,
Bloch -- has its own tests).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _coefficients(n, v, seed):
    rng = np.random.default_rng(seed)
    mxy = (rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))) * 0.1
    w = rng.uniform(0.5, 1.5, n) * 1e-7
    return mxy, w


def _sampled(wl, n_pick, seed):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(wl.n_samples, size=n_pick, replace=False))
    kx, ky, kz = (np.asarray(k).ravel()[idx] for k in wl.kloc)
    return idx, kx, ky, kz, np.asarray(wl.t).ravel()[idx]


def _static_case(wl, method, seed, dev):
    import torch

    import oracle
    from feelmri_amd import device

    m, tis = wl.mesh, wl.tissue
    mxy, w = _coefficients(m.n_nodes, 1, seed)
    plan = device.SignalPlan(wl.kloc, wl.t, method=method, device=dev)
    nodes = device.SignalNodes.from_arrays(m.nodes, tis.delta_omega, tis.T2star, mxy, weights=w, device=dev)
    mixed = plan.execute(nodes, precision="mixed").cpu().numpy()
    f64 = plan.execute(nodes, precision="f64").cpu().numpy()
    e_mixed = rel_l2(mixed, f64)
    idx, kx, ky, kz, t = _sampled(wl, 48, seed)
    ref = oracle.signal_samples(kx, ky, kz, t, m.nodes, tis.delta_omega, tis.T2star, w[:, None] * mxy,
                                n_threads=16)
    e_f64 = rel_l2(f64[idx], ref)
    # additivity over node partitions, FP64
    h = m.n_nodes // 2
    parts = []
    for sl in (slice(0, h), slice(h, None)):
        nd = device.SignalNodes.from_arrays(m.nodes[sl], tis.delta_omega[sl], tis.T2star[sl], mxy[sl],
                                            weights=w[sl], device=dev)
        parts.append(plan.execute(nd, precision="f64").cpu().numpy())
    e_add = rel_l2(parts[0] + parts[1], f64)
    torch.cuda.synchronize()
    print(wl.name, plan.method, f"mixed vs f64 {e_mixed:.2e}, f64 vs oracle {e_f64:.2e}, halves {e_add:.2e}")
    assert e_mixed <= 1e-4
    assert e_f64 <= 1e-10
    assert e_add <= 1e-12


def test_c1_full_size(gpu):
    """512 000 nodes, 128 x 128 x 32 Cartesian: the bf16 and FP64 GEMMs."""
    from feelmri_amd import workloads

    _static_case(workloads.c1(), "cartesian", 11, gpu)


def test_c1_full_size_fp32_direct_form(gpu):
    """c1's FP32 form: the direct sum (FP32 readout runs) at full size against
    the FP64 Cartesian GEMM over every sample; the FP64 direct form (FP64
    readout runs) against the FP64 GEMM too."""
    import torch

    from feelmri_amd import device, workloads

    wl = workloads.c1()
    m, tis = wl.mesh, wl.tissue
    mxy, w = _coefficients(m.n_nodes, 1, 14)
    nodes = device.SignalNodes.from_arrays(m.nodes, tis.delta_omega, tis.T2star, mxy, weights=w, device=gpu)
    gemm = device.SignalPlan(wl.kloc, wl.t, method="cartesian", device=gpu)
    direct = device.SignalPlan(wl.kloc, wl.t, method="direct", device=gpu)
    assert direct.method == "direct" and direct.n_runs > 0
    ref = gemm.execute(nodes, precision="f64").cpu().numpy()
    f32 = direct.execute(nodes, precision="mixed").cpu().numpy()
    f64 = direct.execute(nodes, precision="f64").cpu().numpy()
    torch.cuda.synchronize()
    e32, e64 = rel_l2(f32, ref), rel_l2(f64, ref)
    print(f"c1 direct FP32 vs FP64 GEMM {e32:.2e}, direct FP64 vs FP64 GEMM {e64:.2e}")
    assert e32 <= 1e-4 and e64 <= 1e-10


@pytest.mark.parametrize("kind", ["radial", "spiral"])
def test_c2_full_size(gpu, kind):
    """973 272 nodes; radial 256 x 512 (readout runs), spiral 32 x 2048
    (packed t-groups)."""
    from feelmri_amd import workloads

    _static_case(workloads.c2(kind=kind), "direct", 12, gpu)


def test_c3_full_size_one_phase(gpu):
    """1 948 800 moving nodes, 96 x 96 x 24, 4 encodings, one cardiac phase:
    the per-readout GEMMs (mixed, FP64) against each other, and the FP64
    form against the oracle with the positions the device tables give at
    each sample's t."""
    import torch

    import oracle
    from feelmri_amd import Motion, device, workloads
    from test_gpu_signal import descriptor_callable

    dev = gpu
    wl = workloads.c3()
    m, tis = wl.mesh, wl.tissue
    mxy, w = _coefficients(m.n_nodes, 4, 13)
    plan = device.SignalPlan(wl.kloc, wl.t, method="auto", device=dev)
    assert plan.method == "cartesian"
    modes = device.modes_soa(torch.from_numpy(Motion.motion_modes(wl.motion)).to(dev))
    wl.motion.update_timeshift(0.0)
    coef = torch.from_numpy(Motion.motion_coefficients(wl.motion, plan.tgroup_t)).to(dev)
    nodes = device.SignalNodes.from_arrays(m.nodes, tis.delta_omega, tis.T2star, mxy, weights=w, device=dev)
    mixed = plan.execute(nodes, precision="mixed", motion=(modes, coef)).cpu().numpy()
    f64 = plan.execute(nodes, precision="f64", motion=(modes, coef)).cpu().numpy()
    e_mixed = rel_l2(mixed, f64)
    disp = descriptor_callable(wl.motion)
    idx, kx, ky, kz, t = _sampled(wl, 12, 14)
    ref = np.stack([oracle.signal_samples(kx[i:i + 1], ky[i:i + 1], kz[i:i + 1], t[i:i + 1], m.nodes + disp(t[i]),
                                          tis.delta_omega, tis.T2star, w[:, None] * mxy, n_threads=16)[0]
                    for i in range(idx.size)])
    e_f64 = rel_l2(f64[idx], ref)
    print("c3 mixed vs f64", f"{e_mixed:.2e}", "f64 vs oracle", f"{e_f64:.2e}")
    assert e_mixed <= 1e-4
    assert e_f64 <= 1e-9
