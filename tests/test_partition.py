"""Element partitioning into node-sharded parts (python/feelmri/Phantom.py:
194-239 distributes the phantom over MPI ranks this way, with METIS).

CPU: every element lands in exactly one part; a part's nodes are the sorted
unique nodes of its elements and its connectivity maps back to the global
one; RCB parts are balanced to one element; boundary nodes are duplicated;
partial lumped weights (oracle) scattered back sum to the global weights;
per-part k-spaces (oracle) sum to the global k-space -- the Reduce(SUM) of
python/feelmri/MPIUtilities.py:13-26.
"""
import numpy as np
import pytest

from feelmri_amd import partition, phantoms, trajectories


@pytest.mark.parametrize("method", ["blocks", "rcb"])
@pytest.mark.parametrize("n_parts", [1, 3, 8])
def test_partition_invariants(method, n_parts):
    m = phantoms.ellipsoid_mesh(n=12)
    parts = partition.split_mesh(m.elems, m.nodes, n_parts, method)
    assert len(parts) == n_parts
    seen = np.zeros(m.n_elems, dtype=int)
    for p in parts:
        seen[p.elems_l2g] += 1
        assert np.all(np.diff(p.nodes_l2g) > 0)
        np.testing.assert_array_equal(p.nodes_l2g[p.elems], m.elems[p.elems_l2g])
        assert set(np.unique(p.elems).tolist()) == set(range(p.n_nodes))
    assert np.all(seen == 1)
    sizes = [p.n_elems for p in parts]
    if method == "rcb":
        assert max(sizes) - min(sizes) <= 1
    dup = partition.duplication(parts, m.n_nodes)
    assert dup == 1.0 if n_parts == 1 else dup > 1.0


def test_rcb_parts_are_compact():
    """RCB keeps the duplicated boundary below that of interleaved parts."""
    m = phantoms.box_mesh((0, 0, 0), (1, 1, 1), (12, 12, 12))
    rcb = partition.split_mesh(m.elems, m.nodes, 8, "rcb")
    member = np.arange(m.n_elems) % 8
    inter = [partition.local_mesh(m.elems, member, p) for p in range(8)]
    assert partition.duplication(rcb, m.n_nodes) < 0.5 * partition.duplication(inter, m.n_nodes)


def test_partial_weights_and_kspace_sum_to_global(oracle):
    m = phantoms.ellipsoid_mesh(n=12)
    tis = phantoms.draw_tissue(m.n_nodes, seed=7)
    rng = np.random.default_rng(3)
    mxy = rng.normal(size=(m.n_nodes, 2)) + 1j * rng.normal(size=(m.n_nodes, 2))
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (6, 5, 3), t_first=1.0)
    k = [np.asarray(a).ravel() for a in kloc]
    w = oracle.mass_p1_lumped(m.elems, m.nodes)
    ref = oracle.signal_samples(*k, t.ravel(), m.nodes, tis.delta_omega, tis.T2, w[:, None] * mxy)
    acc_w = np.zeros(m.n_nodes)
    acc_s = np.zeros_like(ref)
    for p in partition.split_mesh(m.elems, m.nodes, 5, "rcb"):
        idx = p.nodes_l2g
        wl = oracle.mass_p1_lumped(p.elems, m.nodes[idx])
        np.add.at(acc_w, idx, wl)
        acc_s += oracle.signal_samples(*k, t.ravel(), m.nodes[idx], tis.delta_omega[idx], tis.T2[idx],
                                       wl[:, None] * mxy[idx])
    np.testing.assert_allclose(acc_w, w, rtol=1e-13)
    assert np.linalg.norm(acc_s - ref) / np.linalg.norm(ref) < 1e-13


def test_bad_arguments():
    m = phantoms.box_mesh((0, 0, 0), (1, 1, 1), (3, 3, 3))
    with pytest.raises(ValueError):
        partition.partition_elements(m.elems[:, :3], m.nodes, 2)
    with pytest.raises(ValueError):
        partition.partition_elements(m.elems, m.nodes, 0)
    with pytest.raises(ValueError):
        partition.partition_elements(m.elems, m.nodes, 2, "metis")
