"""Node-sharded k-space on the GPU: each part of an element partition runs
the lumped weights and the signal on its own nodes (interface nodes carry
partial weights, python/feelmri/Phantom.py:194-239), and the per-part
k-spaces summed in part order equal the whole mesh's k-space -- the
Reduce(SUM) of python/feelmri/MPIUtilities.py:13-26 -- to FP64 rounding
(FP64 form) and to the mixed bar (mixed form), on a Cartesian GEMM plan and
a radial direct-sum plan.
"""
import numpy as np
import pytest

from feelmri_amd import partition, phantoms, trajectories

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("traj", ["cartesian", "radial"])
def test_part_kspaces_sum_to_global(gpu, traj):
    import torch

    from feelmri_amd import device

    m = phantoms.ellipsoid_mesh(n=40)
    tis = phantoms.draw_tissue(m.n_nodes, seed=11, r2s_range=(10.0, 50.0))
    rng = np.random.default_rng(2)
    mxy = (rng.normal(size=(m.n_nodes, 2)) + 1j * rng.normal(size=(m.n_nodes, 2))) * 0.1
    if traj == "cartesian":
        kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (32, 24, 4), t_first=1.0)
    else:
        kloc, t = trajectories.radial(0.24, 128, 32, t_first=0.5)
    plan = device.SignalPlan(kloc, t, device=gpu)

    def kspace(idx, elems, prec):
        nodes = torch.from_numpy(np.ascontiguousarray(m.nodes[idx])).to(gpu)
        col = device.P1Colouring(elems, idx.size, device=gpu)
        w = device.lumped_mass(col, nodes)
        sn = device.SignalNodes.from_arrays(nodes, tis.delta_omega[idx], tis.T2star[idx], mxy[idx], weights=w,
                                            device=gpu)
        return plan.execute(sn, precision=prec)

    full = np.arange(m.n_nodes)
    for prec, tol in (("f64", 1e-12), ("mixed", 1e-4)):
        ref = kspace(full, m.elems, "f64")
        acc = torch.zeros_like(ref)
        for p in partition.split_mesh(m.elems, m.nodes, 6, "rcb"):
            acc += kspace(p.nodes_l2g, p.elems, prec)
        err = float(torch.linalg.vector_norm(acc - ref) / torch.linalg.vector_norm(ref))
        print(f"{traj} {prec}: 6-part sum vs whole mesh relL2 {err:.2e}")
        assert err <= tol
