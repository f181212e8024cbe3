"""Bloch, P1 mass and POD parity on the GPU against the FP64 oracle.

Bloch FP64 relL2 <= 1e-10, FP32 <= 1e-4; mass and POD FP64 <= 1e-12.
"""
import numpy as np
import pytest

import motion_fixtures
from feelmri_amd import Motion, phantoms, sequences

pytestmark = pytest.mark.gpu
GAMMA = sequences.GAMMA


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _bloch_inputs(n, T, seed=0):
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(-0.05, 0.05, (n, 3))
    T1 = rng.uniform(500, 1500, n)
    T2 = rng.uniform(40, 100, n)
    dB = rng.normal(0, 20 / 42.58e6 * 1e3, n)
    rf = (rng.normal(size=T) + 1j * rng.normal(size=T)) * 2e-3
    rf[rng.uniform(size=T) < 0.4] = 0
    G = rng.normal(size=(T, 3)) * 5
    dt = rng.choice([0.01, 0.01, 0.01, 0.02, 1.0], T)
    mxy0 = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.1
    mz0 = rng.uniform(0.5, 1.0, n)
    return r0, T1, T2, dB, rf, G, dt, mxy0, mz0


@pytest.mark.parametrize("n,T", [(1000, 300), (257, 513), (1, 2), (300, 1)])
def test_bloch_history(gpu, n, T):
    from feelmri_amd import BlochSimulator
    import oracle

    r0, T1, T2, dB, rf, G, dt, mxy0, mz0 = _bloch_inputs(n, T, seed=n + T)
    ref = oracle.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, history=True)
    for prec, tol in (("f64", 1e-10), ("f32", 1e-4)):
        mxy, mz = BlochSimulator.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0,
                                           precision=prec)
        assert mxy.shape == (n, T) and mxy.dtype == np.complex64 and mz.dtype == np.float32
        # outputs keep the reference's float32 dtypes: compare at float32 resolution
        e1, e2 = rel_l2(mxy, ref[0]), rel_l2(mz, ref[1])
        print(prec, n, T, f"{e1:.2e} {e2:.2e}")
        assert e1 <= max(tol, 1e-6) and e2 <= max(tol, 1e-6)


def test_bloch_free_precession_zero_field(gpu):
    """rf-free events take the precession step Mxy exp(-i gamma dt Bz); with
    Bz = 0 (node at the origin, no off-resonance, no gradient) the
    reference floors |B| at 1e-12 (BlochSimulator.cpp:239) and turns
    by ~1e-12 rad: both leave Mxy unrotated."""
    from feelmri_amd import BlochSimulator
    import oracle

    n, T = 300, 60
    r0, T1, T2, dB, rf, G, dt, mxy0, mz0 = _bloch_inputs(n, T, seed=33)
    r0[:100] = 0.0
    dB[:100] = 0.0
    G[::3] = 0.0
    rf[10:] = 0.0
    ref = oracle.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, history=True)
    for prec, tol in (("f64", 1e-12), ("f32", 1e-5)):
        mxy, mz = BlochSimulator.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0,
                                           precision=prec)
        e1, e2 = rel_l2(mxy, ref[0]), rel_l2(mz, ref[1])
        print(prec, f"{e1:.2e} {e2:.2e}")
        assert e1 <= max(tol, 1e-6) and e2 <= max(tol, 1e-6)


def test_bloch_repeated_events(gpu):
    """Runs of equal events (rf, G and dt repeated, as in a hard pulse or on a
    gradient plateau, at 10-20 us spacing): the oracle's result, history and
    final state."""
    import torch

    from feelmri_amd import BlochSimulator, device
    import oracle

    n, T = 700, 400
    r0, T1, T2, dB, rf, G, dt, mxy0, mz0 = _bloch_inputs(n, T, seed=41)
    rng = np.random.default_rng(41)
    dt = rng.choice([0.01, 0.02], T)   # event spacing of pulses and plateaus
    i = 0
    while i < T:
        L = int(rng.integers(1, 30))
        rf[i:i + L], G[i:i + L], dt[i:i + L] = rf[i], G[i], dt[i]
        i += L
    ref = oracle.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, history=True)
    for prec, tol in (("f64", 1e-10), ("f32", 1e-4)):
        mxy, mz = BlochSimulator.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0,
                                           precision=prec)
        e1, e2 = rel_l2(mxy, ref[0]), rel_l2(mz, ref[1])
        print(prec, f"{e1:.2e} {e2:.2e}")
        assert e1 <= max(tol, 1e-6) and e2 <= max(tol, 1e-6)
    fin = oracle.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, history=False)
    nodes = device.BlochNodes.from_arrays(r0, T1, T2, dB, device=gpu)
    ev = device.EventTable(rf, G, dt, device=gpu)
    gx, gz = device.bloch(nodes, ev, 1.0, GAMMA, torch.from_numpy(mxy0).to(gpu), torch.from_numpy(mz0).to(gpu))
    assert rel_l2(gx.cpu().numpy(), fin[0]) <= 1e-10 and rel_l2(gz.cpu().numpy(), fin[1]) <= 1e-10


def test_bloch_final_state_f64_exact(gpu):
    import torch

    from feelmri_amd import device
    import oracle

    n, T = 3000, 700
    r0, T1, T2, dB, rf, G, dt, mxy0, mz0 = _bloch_inputs(n, T, seed=7)
    ref = oracle.solve_mri(r0, T1, T2, dB, 0.9, GAMMA, rf, G, dt, None, mxy0, mz0, history=False)
    nodes = device.BlochNodes.from_arrays(r0, T1, T2, dB, device=gpu)
    ev = device.EventTable(rf, G, dt, device=gpu)
    mxy, mz = device.bloch(nodes, ev, 0.9, GAMMA, torch.from_numpy(mxy0).to(gpu), torch.from_numpy(mz0).to(gpu))
    assert rel_l2(mxy.cpu().numpy(), ref[0]) <= 1e-10
    assert rel_l2(mz.cpu().numpy(), ref[1]) <= 1e-10


def _pod(n, velocity=False, seed=3, family="own"):
    return motion_fixtures.pod(n, seed=seed, velocity=velocity, family=family)


@pytest.mark.parametrize("kind", ["pod", "velocity", "sum", "callable", "ref-velocity", "ref-sum"])
def test_bloch_moving(gpu, kind):
    from feelmri_amd import BlochSimulator
    import oracle

    n, T = 800, 90
    r0, T1, T2, dB, rf, G, dt, mxy0, mz0 = _bloch_inputs(n, T, seed=21)
    family = "ref" if kind.startswith("ref-") else "own"
    base = kind.split("-")[-1]
    resp = motion_fixtures.resp(direction=(1, 1, 0), family=family)
    pod = _pod(n, velocity=base == "velocity", family=family)
    m = {"pod": pod, "velocity": pod, "sum": resp + pod, "callable": (lambda tt: pod(tt) * 2.0)}[base]
    # kernel parity: the oracle moves the nodes with the device's float32 tables
    from test_gpu_signal import descriptor_callable

    om = m if kind == "callable" else descriptor_callable(m)
    ref = oracle.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, om, history=True)
    mxy, mz = BlochSimulator.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, m)
    assert rel_l2(mxy, ref[0]) <= 1e-6 and rel_l2(mz, ref[1]) <= 1e-6
    # semantics: against the object's own float32 __call__
    ref2 = oracle.solve_mri(r0, T1, T2, dB, 1.0, GAMMA, rf, G, dt, None, mxy0, mz0, m, history=True)
    assert rel_l2(mxy, ref2[0]) <= 1e-5 and rel_l2(mz, ref2[1]) <= 1e-5


def test_bloch_solver_spoiled_gre_chain(gpu):
    from feelmri_amd.Bloch import BlochSolver
    import oracle

    n = 2000
    rng = np.random.default_rng(5)
    r0 = rng.uniform(-0.05, 0.05, (n, 3))
    T1 = rng.uniform(500, 1500, n)
    T2 = rng.uniform(40, 100, n)
    dB = rng.normal(0, 1e-4, n)
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=6, n_lines=4, rf_dur=0.2)
    want_xy, want_z, _, _ = oracle.bloch_solver_solve([b.as_dict() for b in blocks], r0, T1, T2, dB, 1.0, GAMMA,
                                                      np.zeros(n), np.ones(n))
    solver = BlochSolver(blocks, r0, T1, T2, delta_B=dB)
    got_xy, got_z = solver.solve()
    assert got_xy.shape == want_xy.shape == (n, 4)
    assert rel_l2(got_xy, want_xy) <= 1e-6 and rel_l2(got_z, want_z) <= 1e-6


@pytest.mark.parametrize("n_dummy", [3, 70])
def test_bloch_chain_equals_block_launches(gpu, n_dummy):
    """Static nodes run a block range as one chained launch (one per 128
    blocks), with the store and the artificial spoiling in the kernel.  It
    equals one integrator launch plus one store/spoil launch per block bit
    for bit (stores and carried state, both precisions) over spoiling and
    non-spoiling blocks, stored and unstored ones, a one-event block, random
    gradients, and with 70 dummy TRs (149 blocks) across a launch boundary;
    a second solve continues from the carried state."""
    import torch

    from feelmri_amd.Bloch import BlochSolver

    n = 2000
    rng = np.random.default_rng(13)
    r0 = rng.uniform(-0.05, 0.05, (n, 3))
    T1 = rng.uniform(500, 1500, n)
    T2 = rng.uniform(40, 100, n)
    dB = rng.normal(0, 1e-4, n)
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=n_dummy, n_lines=3, rf_dur=0.2)
    one = sequences.Block(times=np.array([0.0]), rf=np.zeros(1, complex), G=np.zeros((1, 3)), dt=np.array([0.5]),
                          store=True)
    grad = sequences.Block(times=np.arange(7) * 0.1, rf=np.zeros(7, complex), G=rng.normal(0, 5, (7, 3)),
                           dt=np.full(7, 0.1), store=True)
    blocks = blocks[:4] + [one, grad] + blocks[4:] + [grad]
    n_store = sum(b.store for b in blocks)
    for prec in ("f64", "f32"):
        a = BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, dev=gpu)
        b = BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, dev=gpu)
        for _ in range(2):
            xa, za = a.solve_device(chain=True)
            xb, zb = b.solve_device(chain=False)
            torch.cuda.synchronize()
            assert xa.shape == (n, n_store)
            assert torch.equal(xa, xb) and torch.equal(za, zb)
            assert torch.equal(a.mxy, b.mxy) and torch.equal(a.mz, b.mz)
        # a sub-range, as solve(start, end) runs it
        xa, za = a.solve_device(2, 9)
        xb, zb = b.solve_device(2, 9, chain=False)
        torch.cuda.synchronize()
        assert torch.equal(xa, xb) and torch.equal(za, zb)


def test_bloch_chain_without_blocks_passes_the_state(gpu):
    import ctypes

    import torch

    from feelmri_amd import _lib, device

    n = 1000
    rng = np.random.default_rng(2)
    nodes = device.BlochNodes.from_arrays(rng.uniform(-0.05, 0.05, (n, 3)), np.full(n, 900.0),
                                          np.full(n, 60.0), np.zeros(n), device=gpu)
    mxy = torch.randn(n, dtype=torch.complex128, device=gpu)
    mz = torch.rand(n, dtype=torch.float64, device=gpu)
    mxy_o = torch.zeros_like(mxy)
    mz_o = torch.zeros_like(mz)
    ptr = device.ptr
    _lib.call("fm_bloch_chain", _lib.PREC_F64, n, ptr(nodes.packed), 1.0, 1.0, 0, None, None, None, None, None,
              None, ptr(torch.view_as_real(mxy)), ptr(mz), None, None, 0, ptr(torch.view_as_real(mxy_o)),
              ptr(mz_o), None)
    torch.cuda.synchronize()
    assert torch.equal(mxy_o, mxy) and torch.equal(mz_o, mz)


def test_bloch_solver_graph_replays_match_eager(gpu):
    """BlochSolver.capture() records the block chain as one HIP graph.  Every
    replay, with eager work and synchronisation between replays, equals the
    eager solve_device() bitwise: stored columns and carried state, in both
    precisions."""
    import torch

    from feelmri_amd.Bloch import BlochSolver

    n = 3000
    rng = np.random.default_rng(9)
    r0 = rng.uniform(-0.05, 0.05, (n, 3))
    T1 = rng.uniform(500, 1500, n)
    T2 = rng.uniform(40, 100, n)
    dB = rng.normal(0, 1e-4, n)
    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=4, n_lines=3, rf_dur=0.2)
    for prec in ("f64", "f32"):
        solver = BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, dev=gpu)
        mxy0, mz0 = solver.mxy.clone(), solver.mz.clone()
        want_xy, want_z = (t.clone() for t in solver.solve_device())
        carried = (solver.mxy.clone(), solver.mz.clone())
        graph = solver.capture()
        for _ in range(3):
            solver.mxy.copy_(mxy0)
            solver.mz.copy_(mz0)
            got_xy, got_z = graph.replay()
            torch.cuda.synchronize()
            assert got_xy.shape == want_xy.shape == (n, 3)
            assert torch.equal(got_xy, want_xy) and torch.equal(got_z, want_z)
            assert torch.equal(solver.mxy, carried[0]) and torch.equal(solver.mz, carried[1])
            # eager work between replays
            assert float((torch.arange(4096, device=gpu, dtype=torch.float64) * 2.0).sum()) == 4095 * 4096


def test_lumped_mass_c0(gpu):
    from feelmri_amd import Assemble
    import oracle

    m = phantoms.slab_mesh()
    w = Assemble.lumped_mass(m.elems, m.nodes, as_vector=True)
    ref = oracle.mass_p1_lumped(m.elems, m.nodes)
    assert np.max(np.abs(w - ref) / ref) <= 1e-12
    np.testing.assert_allclose(w.sum(), 1e-4, rtol=1e-12)


@pytest.mark.parametrize("mesh,dtype", [("slab", np.float64), ("slab", np.float32), ("ellipsoid", np.float64),
                                        ("ellipsoid_shuffled", np.float64), ("box_spare_node", np.float64)])
def test_lumped_gather_equals_coloured_scatter(gpu, mesh, dtype):
    """The gather does the coloured scatter's additions in the same order
    and rounding: bitwise the same weights, and the oracle's to 1e-12.  The
    box mesh gets one node no element references (weight 0); the shuffled
    ellipsoid lists its elements in random order (another colouring, slots
    scattered over the whole table)."""
    import torch

    import oracle
    from feelmri_amd import device

    if mesh == "slab":
        m = phantoms.slab_mesh()
        nodes = m.nodes
    elif mesh.startswith("ellipsoid"):
        m = phantoms.ellipsoid_mesh(n=14)
        nodes = m.nodes
        if mesh.endswith("shuffled"):
            m.elems = m.elems[np.random.default_rng(5).permutation(m.n_elems)]
    else:
        m = phantoms.box_mesh((0, 0, 0), (0.02, 0.02, 0.01), (9, 8, 5))
        nodes = np.vstack([m.nodes, [[0.5, 0.5, 0.5]]])
    nodes = nodes.astype(dtype)
    col = device.P1Colouring(m.elems, nodes.shape[0], device=gpu)
    nd = torch.from_numpy(nodes).to(gpu)
    g = device.lumped_mass(col, nd)
    c = device.lumped_mass_coloured(col, nd)
    torch.cuda.synchronize()
    assert torch.equal(g, c)
    ref = oracle.mass_p1_lumped(m.elems, nodes.astype(np.float64))
    w = g.cpu().numpy()
    assert ref.shape == w.shape or mesh == "box_spare_node"
    if mesh == "box_spare_node":
        assert w[-1] == 0.0
        w = w[:-1]
        ref = ref[: w.size]
    assert np.max(np.abs(w - ref) / ref) <= 1e-12


def test_lumped_mass_unstructured_without_colouring(gpu):
    """A Delaunay mesh of random points (node valence above 64, so no
    colouring of at most 64 colours): the gather adds in element order and
    matches the oracle, which sums in element order too, to 1e-13; the
    weights sum to the convex hull's volume."""
    import torch
    from scipy.spatial import ConvexHull, Delaunay

    import oracle
    from feelmri_amd import device

    pts = np.random.default_rng(0).uniform(-0.05, 0.05, size=(6000, 3))
    el = np.ascontiguousarray(Delaunay(pts).simplices, dtype=np.int32)
    col = device.P1Colouring(el, pts.shape[0], device=gpu)
    assert col.method == "element-order"
    w = device.lumped_mass(col, torch.from_numpy(pts).to(gpu)).cpu().numpy()
    ref = oracle.mass_p1_lumped(el, pts)
    assert np.max(np.abs(w - ref) / ref) <= 1e-13
    np.testing.assert_allclose(w.sum(), ConvexHull(pts).volume, rtol=1e-12)
    with pytest.raises(ValueError, match="colouring"):
        device.lumped_mass_coloured(col, torch.from_numpy(pts).to(gpu))


def test_lumped_gather_element_order_and_node_prefix(gpu):
    """order="element" on a colourable mesh: each node adds its elements in
    ascending element order -- the oracle's weights to 1e-13; with the
    elements sorted by their largest node id, every node prefix weighs the
    sub-mesh of the elements below it, as the oracle does on that sub-mesh."""
    import torch

    import oracle
    from feelmri_amd import device

    m = phantoms.ellipsoid_mesh(n=14)
    el = np.ascontiguousarray(m.elems[np.argsort(m.elems.max(axis=1), kind="stable")])
    col = device.P1Colouring(el, m.n_nodes, device=gpu, order="element")
    assert col.method == "element-order" and col.n_colours > 0 and col.prefix_ordered
    nd = torch.from_numpy(m.nodes).to(gpu)
    w = device.lumped_mass(col, nd).cpu().numpy()
    ref = oracle.mass_p1_lumped(el, m.nodes)
    assert np.max(np.abs(w - ref) / ref) <= 1e-13
    # the coloured scatter stays available and agrees to rounding
    c = device.lumped_mass_coloured(col, nd).cpu().numpy()
    assert np.max(np.abs(c - ref) / ref) <= 1e-13
    for n in (1, 97, m.n_nodes // 3, m.n_nodes - 1):
        E = col.n_elems_below(n)
        assert el[:E].max(initial=-1) < n and (E == el.shape[0] or el[E].max() >= n)
        wp = device.lumped_mass(col, nd[:n].contiguous(), n_nodes=n).cpu().numpy()
        rp = oracle.mass_p1_lumped(el[:E], m.nodes[:n]) if E else np.zeros(n)
        rp = np.pad(rp, (0, n - rp.size))
        assert np.allclose(wp, rp, rtol=1e-13, atol=0.0), n
    with pytest.raises(ValueError, match="element order"):
        device.lumped_mass(device.P1Colouring(el, m.n_nodes, device=gpu), nd, n_nodes=5)


def test_lumped_gather_without_elements(gpu):
    import torch

    from feelmri_amd import device

    col = device.P1Colouring(np.zeros((0, 4), np.int32), 7, device=gpu)
    w = device.lumped_mass(col, torch.ones((7, 3), dtype=torch.float64, device=gpu))
    torch.cuda.synchronize()
    assert torch.equal(w.cpu(), torch.zeros(7, dtype=torch.float64))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_consistent_mass_assembly(gpu, dtype):
    from feelmri_amd import Assemble
    import oracle

    m = phantoms.ellipsoid_mesh(n=14)
    nodes = m.nodes.astype(dtype)
    M = Assemble.MassAssemble(m.elems, nodes, None, None)
    ref = oracle.MassAssemble(m.elems, nodes.astype(np.float64))
    diff = abs(M.astype(np.float64) - ref).max()
    assert diff <= (1e-12 if dtype == np.float64 else 1e-6) * abs(ref).max()


# CSR SpMV (16 lanes per row): empty rows, rows longer than the lane group,
# more than 4 encodings (two passes), a row count that ends in a partial
# workgroup; FP64 against scipy, and the same bits on a second call
@pytest.mark.parametrize("n_rows,V", [(1003, 1), (257, 3), (64, 5), (1, 2)])
def test_csr_spmv_against_scipy(gpu, n_rows, V):
    import scipy.sparse as sp
    import torch
    from feelmri_amd import device

    rng = np.random.default_rng(n_rows + V)
    lengths = rng.integers(0, 40, n_rows)
    lengths[:: 7] = 0
    lengths[3 % n_rows] = 150
    rows = np.repeat(np.arange(n_rows), lengths)
    cols = rng.integers(0, n_rows, rows.size)
    M = sp.csr_matrix((rng.normal(size=rows.size), (rows, cols)), shape=(n_rows, n_rows))
    M.sort_indices()
    x = rng.normal(size=(n_rows, V)) + 1j * rng.normal(size=(n_rows, V))
    xd = torch.from_numpy(x).to(gpu)
    y = device.csr_spmv(M.indptr, M.indices, M.data, xd)
    y2 = device.csr_spmv(M.indptr, M.indices, M.data, xd)
    want = M @ x
    got = y.cpu().numpy()
    assert got.shape == want.shape
    scale = np.abs(M).sum(axis=1).A.ravel()[:, None] * np.abs(x).max() + 1e-300
    assert np.max(np.abs(got - want) / scale) <= 1e-14
    assert torch.equal(y, y2)


# M <= 32 modes take the LDS-tiled kernel, more modes the per-row one; row
# counts that are not a multiple of the 256-row tile end in a partial tile
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("P,C,M", [(5000, 3, 10), (777, 3, 1), (301, 2, 32), (129, 3, 33), (1, 1, 5)])
def test_pod_tensordot(gpu, dtype, P, C, M):
    from feelmri_amd import POD

    rng = np.random.default_rng(9 + P + M)
    modes = rng.normal(size=(P, C, M)).astype(dtype)
    w = rng.normal(size=M).astype(dtype)
    got = POD.tensordot_modes_weights(modes, w)
    want = np.tensordot(modes.astype(np.float64), w.astype(np.float64), axes=([2], [0]))
    assert got.dtype == dtype
    np.testing.assert_allclose(got, want, rtol=1e-5 if dtype == np.float32 else 1e-12, atol=1e-5)


def test_pod_positions_match_call(gpu):
    import torch

    from feelmri_amd import device

    n = 3000
    pod = _pod(n)
    pod.update_timeshift(1.5)
    r0 = np.random.default_rng(1).uniform(-0.05, 0.05, (n, 3))
    modes = device.modes_soa(torch.from_numpy(Motion.motion_modes(pod)).to(gpu))
    coef = torch.from_numpy(Motion.motion_coefficients(pod, [7.0])[0]).to(gpu)
    xyz = torch.from_numpy(np.ascontiguousarray(r0.T)).to(gpu)
    got = device.pod_positions(xyz, modes, coef).cpu().numpy()
    np.testing.assert_allclose(got, r0 + pod(7.0), rtol=1e-6, atol=1e-9)
