"""An XDMF phantom (geometry, topology and per-node tissue in HDF5) ingested
by feelmri_amd.mesh_io drives the GPU path: lumped weights on the device
against the FP64 oracle of the mesh it was written from (<= 1e-12), then
the k-space of the ingested phantom in the FP64 form against the oracle
(<= 1e-10) and in the mixed form (<= 1e-4)."""
import numpy as np
import pytest
import scipy.sparse as sp

from feelmri_amd import mesh_io, phantoms, trajectories

from h5write import H5Writer

pytestmark = pytest.mark.gpu


def test_xdmf_phantom_weights_and_signal(gpu, tmp_path):
    import oracle
    from feelmri_amd import MRI, Assemble

    m = phantoms.box_mesh((-0.03, -0.03, -0.01), (0.03, 0.03, 0.01), (25, 25, 7))
    tis = phantoms.draw_tissue(m.n_nodes, seed=5)
    w = H5Writer()
    w.create_dataset("/Mesh/mesh/geometry", m.nodes)
    w.create_dataset("/Mesh/mesh/topology", m.elems.astype(np.int64), chunks=(1024, 4), shuffle=True, deflate=4)
    w.create_dataset("/T2star", tis.T2star)
    w.create_dataset("/dw", tis.delta_omega)
    w.save(tmp_path / "ph.h5")
    n, e = m.n_nodes, len(m.elems)
    (tmp_path / "ph.xdmf").write_text(f"""<?xml version="1.0"?><Xdmf Version="3.0"><Domain><Grid Name="mesh">
<Geometry GeometryType="XYZ"><DataItem Dimensions="{n} 3" Format="HDF" Precision="8">ph.h5:/Mesh/mesh/geometry</DataItem></Geometry>
<Topology TopologyType="Tetrahedron" NumberOfElements="{e}"><DataItem DataType="Int" Dimensions="{e} 4" Format="HDF" Precision="8">ph.h5:/Mesh/mesh/topology</DataItem></Topology>
<Attribute Name="T2star" Center="Node"><DataItem Dimensions="{n}" Format="HDF" Precision="8">ph.h5:/T2star</DataItem></Attribute>
<Attribute Name="dw" Center="Node"><DataItem Dimensions="{n}" Format="HDF" Precision="8">ph.h5:/dw</DataItem></Attribute>
</Grid></Domain></Xdmf>""")
    print(f"ingesting a {n}-node, {e}-tet XDMF phantom", flush=True)
    ph = mesh_io.load_phantom(tmp_path / "ph.xdmf")
    elems = ph.elements.astype(np.int32)
    wts = Assemble.lumped_mass(elems, ph.nodes, as_vector=True)
    ref_w = oracle.mass_p1_lumped(m.elems, m.nodes)
    assert np.max(np.abs(wts - ref_w) / ref_w) <= 1e-12
    T2, dw = ph.point_data["T2star"], ph.point_data["dw"]
    rng = np.random.default_rng(3)
    mxy = (rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))) * 0.1
    kloc, t = trajectories.radial(0.24, 128, 12, t_first=0.4)
    ref = oracle.Signal(0, sp.diags(ref_w), kloc, t, m.nodes, tis.delta_omega, tis.T2star, mxy)
    for prec, tol in (("f64", 1e-10), ("mixed", 1e-4)):
        got = MRI.Signal(0, sp.diags(wts), kloc, t, ph.nodes, dw, T2, mxy, precision=prec)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(prec, f"{err:.2e}")
        assert err <= tol, (prec, err)


def test_fem_phantom_mass_matrices(gpu, tmp_path):
    """FEMPhantom.mass_matrix on the GPU: the consistent matrix's row sums
    and the lumped matrix's diagonal equal the oracle's lumped weights, and
    the lumped weights of a 3-way partition sum to the whole mesh's (partial
    weights on shared nodes)."""
    import oracle
    from feelmri_amd.Phantom import FEMPhantom, lumped_weights

    m = phantoms.box_mesh((-0.02, -0.02, -0.01), (0.02, 0.02, 0.01), (12, 11, 6))
    w = H5Writer()
    w.create_dataset("/g", m.nodes)
    w.create_dataset("/t", m.elems.astype(np.int64))
    w.save(tmp_path / "m.h5")
    n, e = m.n_nodes, len(m.elems)
    (tmp_path / "m.xdmf").write_text(f"""<Xdmf Version="3.0"><Domain><Grid Name="mesh">
<Geometry GeometryType="XYZ"><DataItem Dimensions="{n} 3" Format="HDF" Precision="8">m.h5:/g</DataItem></Geometry>
<Topology TopologyType="Tetrahedron" NumberOfElements="{e}"><DataItem DataType="Int" Dimensions="{e} 4" Format="HDF" Precision="8">m.h5:/t</DataItem></Topology>
</Grid></Domain></Xdmf>""")
    want = oracle.mass_p1_lumped(m.elems, m.nodes)
    ph = FEMPhantom(str(tmp_path / "m.xdmf"), dtype=np.float64)
    M = ph.mass_matrix()
    np.testing.assert_allclose(np.asarray(M.sum(axis=1)).ravel(), want, rtol=1e-12)
    np.testing.assert_allclose(lumped_weights(ph.mass_matrix(lumped=True)), want, rtol=1e-12)
    acc = np.zeros(n)
    for part in range(3):
        p = FEMPhantom(str(tmp_path / "m.xdmf"), dtype=np.float64, n_parts=3, part=part)
        acc[p.local_to_global_nodes] += lumped_weights(p.mass_matrix(lumped=True))
    np.testing.assert_allclose(acc, want, rtol=1e-12)
