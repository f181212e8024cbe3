"""4-D flow chain (the c3 workload at reduced size) on the GPU against the FP64 oracle.

Bloch: the whole phase-contrast acquisition (dummy TRs, then per cardiac phase one stored
imaging block per velocity encoding), with nodes moving along the Poiseuille field of a
``PODVelocity`` (BlochSimulator.cpp:12-148 with a motion callable, chained as in
Bloch.py:298-380). Signal: per phase, the four encodings' magnetisation with POD displacement
motion (MRI.cpp:38-159), FP64 and mixed forms.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_flow_chain(gpu):
    from feelmri_amd import MRI, workloads
    from feelmri_amd.Bloch import BlochSolver
    from feelmri_amd.sequences import GAMMA
    import oracle

    wl = workloads.c3(scale=0.05, n_phases=4, kspace=(12, 8, 2))
    blocks = workloads.flow_sequence(wl, n_dummy=4)
    m, tis = wl.mesh, wl.tissue
    N = m.n_nodes
    assert sum(b.store for b in blocks) == 4 * wl.n_phases

    solver = BlochSolver(blocks, m.nodes, tis.T1, tis.T2, delta_B=tis.delta_B_mT, pod_trajectory=wl.velocity,
                         dev=gpu)
    got_xy, got_z = solver.solve()
    ref_xy, ref_z, _, _ = oracle.bloch_solver_solve([b.as_dict() for b in blocks], m.nodes, tis.T1, tis.T2,
                                                    tis.delta_B_mT, 1.0, GAMMA, np.zeros(N), np.ones(N),
                                                    pod_trajectory=wl.velocity)
    assert got_xy.shape == ref_xy.shape == (N, 4 * wl.n_phases)
    # complex64 / float32 outputs of an FP64 integration
    assert rel_l2(got_xy, ref_xy) <= 1e-6, rel_l2(got_xy, ref_xy)
    assert rel_l2(got_z, ref_z) <= 1e-6
    # the velocity encodings differ from the reference encoding (flow is visible)
    assert rel_l2(ref_xy[:, 1], ref_xy[:, 0]) > 1e-3

    M = sp.diags(oracle.mass_p1_lumped(m.elems, m.nodes))
    for p in range(wl.n_phases):
        wl.motion.update_timeshift(p * wl.period / wl.n_phases)
        mxy = ref_xy[:, 4 * p:4 * p + 4] * tis.M0[:, None]
        ref = oracle.Signal(0, M, wl.kloc, wl.t, m.nodes, tis.delta_omega, tis.T2star, mxy, wl.motion)
        assert ref.shape == (*wl.t.shape, 4)
        # FP64 bar 1e-6: the motion object's own __call__ (which the oracle
        # evaluates) forms float32 modes x float32 weights in float32, the
        # device accumulates the same float32 products in float64
        for prec, tol in (("f64", 1e-6), ("mixed", 1e-4)):
            got = MRI.Signal(0, M, wl.kloc, wl.t, m.nodes, tis.delta_omega, tis.T2star, mxy, wl.motion,
                             precision=prec)
            err = rel_l2(got, ref)
            print(f"phase {p} {prec}: relL2 {err:.2e}")
            assert err <= tol, (p, prec, err)
