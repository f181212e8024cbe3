"""Mesh ingestion (feelmri_amd.mesh_io, feelmri_amd.hdf5_lite) on the CPU.

* The HDF5 reader against files the HDF5 library itself wrote: the sample
  files shipped with PyTables' test suite in this image's conda tree
  (/TestArray = i + j in six numeric types and both byte orders; a chunked
  extendible int32 array with unwritten chunks), when present.
* The reader against this suite's own writer (tests/h5write.py): nested
  groups, byte orders, contiguous and chunked storage with shuffle and
  deflate, ragged edge chunks -- exact round trips.
* XDMF in the layouts meshio and dolfin write (single grid; time series
  with xi:include of the mesh grid), with HDF, inline XML and raw binary
  heavy data: points, cells, node and cell data, steps and times.
* A phantom read back from XDMF gives the same lumped weights as the mesh
  it was written from (FP64 oracle), and its element partition keeps the
  weight sum (the reference's MPI invariant).
"""
from pathlib import Path

import numpy as np
import pytest

from feelmri_amd import hdf5_lite, mesh_io, partition, phantoms

from h5write import H5Writer

PYTABLES = Path("/opt/conda/lib/python3.9/site-packages/tables/tests")


@pytest.mark.skipif(not (PYTABLES / "smpl_f64le.h5").exists(), reason="PyTables sample files not in this image")
@pytest.mark.parametrize("name,dtype", [("f64le", "<f8"), ("f64be", ">f8"), ("i32le", "<i4"), ("i32be", ">i4"),
                                        ("i64le", "<i8"), ("i64be", ">i8")])
def test_reader_on_library_written_files(name, dtype):
    f = hdf5_lite.File(PYTABLES / f"smpl_{name}.h5")
    assert f.keys() == ["TestArray"]
    a = f["/TestArray"]
    assert a.shape == (6, 5) and a.dtype == np.dtype(dtype).newbyteorder("=")
    np.testing.assert_array_equal(a, np.add.outer(np.arange(6), np.arange(5)))


@pytest.mark.skipif(not (PYTABLES / "smpl_SDSextendible.h5").exists(), reason="PyTables sample files not in this image")
def test_reader_on_library_written_chunked_file():
    a = hdf5_lite.File(PYTABLES / "smpl_SDSextendible.h5")["/ExtendibleArray"]
    want = np.zeros((10, 5), dtype=np.int32)
    want[:3, :3] = 1
    want[:2, 3:] = 3
    want[3:, 0] = 2
    assert a.dtype == np.int32
    np.testing.assert_array_equal(a, want)


def test_writer_reader_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    arrays = {
        "/a": rng.normal(size=(7, 3)),
        "/g/b": rng.integers(-1000, 1000, size=(11,)).astype(np.int32),
        "/g/h/c": rng.normal(size=(5, 4, 3)).astype(np.float32),
        "/g/h/d": rng.integers(0, 2 ** 40, size=(9, 2)).astype(np.uint64),
        "/chunked": rng.normal(size=(37, 3)),
        "/chunked_filtered": rng.integers(-5, 5, size=(23, 6)).astype(np.int64),
        "/big_endian": rng.normal(size=(13, 3)),
    }
    w = H5Writer()
    w.create_dataset("/a", arrays["/a"])
    w.create_dataset("/g/b", arrays["/g/b"], byteorder=">")
    w.create_dataset("/g/h/c", arrays["/g/h/c"])
    w.create_dataset("/g/h/d", arrays["/g/h/d"], byteorder=">")
    w.create_dataset("/chunked", arrays["/chunked"], chunks=(8, 2))
    w.create_dataset("/chunked_filtered", arrays["/chunked_filtered"], chunks=(5, 4), shuffle=True, deflate=4)
    w.create_dataset("/big_endian", arrays["/big_endian"], chunks=(4, 3), deflate=1, byteorder=">")
    path = tmp_path / "t.h5"
    w.save(path)
    f = hdf5_lite.File(path)
    assert f.keys() == ["a", "big_endian", "chunked", "chunked_filtered", "g"]
    assert f.keys("/g") == ["b", "h"] and f.keys("/g/h") == ["c", "d"]
    for k, v in arrays.items():
        got = f[k]
        assert got.shape == v.shape and got.dtype == v.dtype, k
        np.testing.assert_array_equal(got, v, err_msg=k)
    with pytest.raises(KeyError):
        f["/g/missing"]


def _mesh():
    m = phantoms.box_mesh((-0.02, -0.01, -0.005), (0.02, 0.01, 0.005), (9, 6, 4))
    rng = np.random.default_rng(1)
    return m, rng.uniform(500, 1500, m.n_nodes), np.arange(len(m.elems)) % 3, rng.normal(size=(m.n_nodes, 3))


def _item(arr, fmt, h5name, key, tmp, writer):
    kind = "Float" if arr.dtype.kind == "f" else "Int"
    dims = " ".join(map(str, arr.shape))
    prec = arr.dtype.itemsize
    if fmt == "HDF":
        writer.create_dataset(key, arr)
        body = f"{h5name}:{key}"
        extra = ""
    elif fmt == "XML":
        body = " ".join(repr(float(x)) if kind == "Float" else str(int(x)) for x in arr.ravel())
        extra = ""
    else:
        name = key.strip("/").replace("/", "_") + ".bin"
        arr.astype(arr.dtype.newbyteorder(">")).tofile(tmp / name)
        body, extra = name, ' Endian="Big"'
    return (f'<DataItem DataType="{kind}" Dimensions="{dims}" Format="{fmt}" Precision="{prec}"{extra}>'
            f"{body}</DataItem>")


@pytest.mark.parametrize("fmt", ["HDF", "XML", "Binary"])
def test_xdmf_single_grid(tmp_path, fmt):
    m, T1, region, _ = _mesh()
    w = H5Writer()
    items = [_item(a, fmt, "mesh.h5", k, tmp_path, w) for a, k in
             [(m.nodes, "/data0"), (m.elems.astype(np.int64), "/data1"), (T1, "/data2"), (region.astype(np.int64), "/data3")]]
    if fmt == "HDF":
        w.save(tmp_path / "mesh.h5")
    (tmp_path / "mesh.xdmf").write_text(f"""<?xml version="1.0"?>
<Xdmf Version="3.0"><Domain><Grid Name="Grid">
<Geometry GeometryType="XYZ">{items[0]}</Geometry>
<Topology TopologyType="Tetrahedron" NumberOfElements="{len(m.elems)}">{items[1]}</Topology>
<Attribute Name="T1" AttributeType="Scalar" Center="Node">{items[2]}</Attribute>
<Attribute Name="region" AttributeType="Scalar" Center="Cell">{items[3]}</Attribute>
</Grid></Domain></Xdmf>""")
    r = mesh_io.read(tmp_path / "mesh.xdmf")
    np.testing.assert_array_equal(r.points, m.nodes)
    assert r.cells[0].type == "tetra"
    np.testing.assert_array_equal(r.cells[0].data, m.elems)
    np.testing.assert_array_equal(r.point_data["T1"], T1)
    np.testing.assert_array_equal(r.cell_data["region"][0], region)
    ph = mesh_io.load_phantom(tmp_path / "mesh.xdmf", scale_factor=10.0, dtype=np.float32)
    assert ph.n_frames == 1 and ph.reader is None and ph.cell_type == "tetra"
    np.testing.assert_array_equal(ph.nodes, (m.nodes * 10.0).astype(np.float32))


def _series_xdmf(m, vel, times, layout, h5):
    w = H5Writer()
    w.create_dataset("/Mesh/mesh/geometry", m.nodes)
    w.create_dataset("/Mesh/mesh/topology", m.elems.astype(np.int64), chunks=(64, 4), deflate=3)
    geo = f'<Geometry GeometryType="XYZ"><DataItem Dimensions="{m.n_nodes} 3" Format="HDF" Precision="8">{h5}:/Mesh/mesh/geometry</DataItem></Geometry>'
    top = (f'<Topology TopologyType="Tetrahedron" NumberOfElements="{len(m.elems)}"><DataItem DataType="Int" '
           f'Dimensions="{len(m.elems)} 4" Format="HDF" Precision="8">{h5}:/Mesh/mesh/topology</DataItem></Topology>')
    steps = []
    for k, t in enumerate(times):
        w.create_dataset(f"/VisualisationVector/{k}", vel[k], chunks=(50, 3), shuffle=True, deflate=2)
        att = (f'<Attribute Name="velocity" AttributeType="Vector" Center="Node"><DataItem Dimensions="{m.n_nodes} 3" '
               f'Format="HDF" Precision="8">{h5}:/VisualisationVector/{k}</DataItem></Attribute>')
        if layout == "meshio":
            inc = '<xi:include xpointer="xpointer(//Grid[@Name=&quot;mesh&quot;]/*[self::Topology or self::Geometry])"/>'
            steps.append(f'<Grid>{inc}<Time Value="{t!r}"/>{att}</Grid>')
        else:  # dolfin: the first step carries the mesh, the others include it
            if k == 0:
                steps.append(f'<Grid Name="mesh" GridType="Uniform">{top}{geo}<Time Value="{t!r}"/>{att}</Grid>')
            else:
                inc = ('<xi:include xpointer="xpointer(//Grid[@Name=&quot;TimeSeries_velocity&quot;]/Grid[1]/*'
                       '[self::Topology or self::Geometry])"/>')
                steps.append(f'<Grid Name="mesh" GridType="Uniform">{inc}<Time Value="{t!r}"/>{att}</Grid>')
    mesh_grid = f'<Grid Name="mesh" GridType="Uniform">{geo}{top}</Grid>' if layout == "meshio" else ""
    xml = (f'<?xml version="1.0"?><Xdmf Version="3.0" xmlns:xi="http://www.w3.org/2001/XInclude"><Domain>{mesh_grid}'
           f'<Grid Name="TimeSeries_velocity" GridType="Collection" CollectionType="Temporal">{"".join(steps)}</Grid>'
           f'</Domain></Xdmf>')
    return w, xml


@pytest.mark.parametrize("layout", ["meshio", "dolfin"])
def test_xdmf_time_series(tmp_path, layout):
    m, _, _, _ = _mesh()
    rng = np.random.default_rng(2)
    times = [0.0, 0.05, 0.1, 0.15]
    vel = [rng.normal(size=(m.n_nodes, 3)) for _ in times]
    w, xml = _series_xdmf(m, vel, times, layout, "flow.h5")
    w.save(tmp_path / "flow.h5")
    (tmp_path / "flow.xdmf").write_text(xml)
    r = mesh_io.TimeSeriesReader(tmp_path / "flow.xdmf")
    assert r.num_steps == len(times)
    pts, cells = r.read_points_cells()
    np.testing.assert_array_equal(pts, m.nodes)
    np.testing.assert_array_equal(cells[0].data, m.elems)
    for k, t in enumerate(times):
        tt, pd, cd = r.read_data(k)
        assert tt == t and cd == {}
        np.testing.assert_array_equal(pd["velocity"], vel[k])
    ph = mesh_io.load_phantom(tmp_path / "flow.xdmf", scale_factor=0.01)
    assert ph.n_frames == len(times) and ph.reader is not None
    np.testing.assert_allclose(ph.nodes, m.nodes * 0.01, rtol=0, atol=0)


def test_phantom_from_xdmf_weights_and_partition(tmp_path):
    """The ingested mesh gives the written mesh's lumped weights, and its
    element partition keeps their sum (partial weights on shared nodes)."""
    import oracle

    m, _, _, _ = _mesh()
    w, xml = _series_xdmf(m, [np.zeros((m.n_nodes, 3))], [0.0], "meshio", "p.h5")
    w.save(tmp_path / "p.h5")
    (tmp_path / "p.xdmf").write_text(xml)
    ph = mesh_io.load_phantom(tmp_path / "p.xdmf")
    want = oracle.mass_p1_lumped(m.elems, m.nodes)
    got = oracle.mass_p1_lumped(ph.elements.astype(np.int32), ph.nodes)
    np.testing.assert_array_equal(got, want)
    parts = partition.split_mesh(ph.elements.astype(np.int32), ph.nodes, 3)
    acc = np.zeros(m.n_nodes)
    for p in parts:
        acc[p.nodes_l2g] += oracle.mass_p1_lumped(p.elems, ph.nodes[p.nodes_l2g])
    np.testing.assert_allclose(acc, want, rtol=1e-13, atol=0)


def test_fem_phantom_reads_scales_and_partitions(tmp_path):
    """FEMPhantom (Phantom.py:49-422) on an XDMF time series: nodes scaled
    and cast, frames read with the displacement/velocity scaled by the scale
    factor, parts of a 3-way partition that cover every element once, a
    submesh and the data carried to it, orient/reorient round trip."""
    from feelmri_amd.Phantom import FEMPhantom

    m, _, _, _ = _mesh()
    rng = np.random.default_rng(4)
    times = [0.0, 0.1]
    vel = [rng.normal(size=(m.n_nodes, 3)) for _ in times]
    w, xml = _series_xdmf(m, vel, times, "meshio", "ph.h5")
    w.save(tmp_path / "ph.h5")
    (tmp_path / "ph.xdmf").write_text(xml)
    ph = FEMPhantom(str(tmp_path / "ph.xdmf"), scale_factor=0.01, dtype=np.float64)
    assert ph.Nfr == 2 and ph.cell_type == "tetra"
    np.testing.assert_array_equal(ph.global_nodes, m.nodes * 0.01)
    np.testing.assert_array_equal(ph.local_nodes, ph.global_nodes)       # one part: the whole mesh
    ph.read_data(1)
    np.testing.assert_allclose(ph.point_data["velocity"], vel[1] * 0.01, rtol=1e-15)
    seen = np.zeros(len(m.elems), dtype=int)
    for part in range(3):
        p = FEMPhantom(str(tmp_path / "ph.xdmf"), dtype=np.float64, n_parts=3, part=part)
        seen[p.local_to_global_elems] += 1
        np.testing.assert_array_equal(p.global_nodes[p.local_to_global_nodes][p.local_elements],
                                      p.global_nodes[p.global_elements[p.local_to_global_elems]])
    assert np.all(seen == 1)
    # submesh of the elements left of x = 0, node data carried to it
    cx = ph.global_nodes[ph.global_elements].mean(axis=1)[:, 0]
    full = ph.global_nodes.copy()
    ph.create_submesh(cx < 0)
    assert ph.global_elements.max() == ph.global_nodes.shape[0] - 1
    np.testing.assert_array_equal(ph.to_submesh(full, global_mesh=True), ph.global_nodes)
    R = np.array([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    before = ph.global_nodes.copy()
    ph.orient(R, np.array([0.001, -0.002, 0.0]))
    ph.reorient(R, np.array([0.001, -0.002, 0.0]))
    np.testing.assert_allclose(ph.global_nodes, before, atol=1e-15)
