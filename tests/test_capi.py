"""The C-ABI library loads and exports every entry point include/*.h declares.

No compute calls: this runs on machines without a GPU.
"""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "fm_signal_execute" in names and "fm_bloch_execute" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from feelmri_amd import _lib

    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    # the ctypes table covers the same set
    assert sorted(_lib.exported_symbols()) == _declared()


def test_error_string_is_available():
    from feelmri_amd import _lib

    lib = _lib.load()
    assert isinstance(lib.fm_last_error(), bytes)
    assert lib.fm_version() >= 1


def test_argument_errors_do_not_need_a_gpu():
    from feelmri_amd import _lib

    with pytest.raises(_lib.FeelmriHipError, match="plan is NULL"):
        _lib.call("fm_signal_execute", None, 0, 1, 1, None, None, 99, None, None, 0, None, None)


def test_bloch_chain_argument_checks_do_not_need_a_gpu():
    """fm_bloch_chain rejects bad block descriptions before any launch."""
    import numpy as np

    from feelmri_amd import _lib

    d = ctypes.c_void_p(1)   # never dereferenced: the checks fail first
    lens = np.array([3, 0], np.int64)
    cols = np.array([-1, 0], np.int32)
    spoil = np.array([1, 0], np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    with pytest.raises(_lib.FeelmriHipError, match="has no event"):
        _lib.call("fm_bloch_chain", _lib.PREC_F32, 10, d, 1.0, 1.0, 2, p(lens), p(cols), p(spoil), d, d, d, d, d, d, d, 1,
                  d, d, None)
    lens[1] = 2
    cols[1] = 1
    with pytest.raises(_lib.FeelmriHipError, match="store column"):
        _lib.call("fm_bloch_chain", _lib.PREC_F32, 10, d, 1.0, 1.0, 2, p(lens), p(cols), p(spoil), d, d, d, d, d, d, d, 1,
                  d, d, None)
    with pytest.raises(_lib.FeelmriHipError, match="precision"):
        _lib.call("fm_bloch_chain", 7, 10, d, 1.0, 1.0, 2, p(lens), p(cols), p(spoil), d, d, d, d, d, d, d, 1,
                  d, d, None)


def test_block_chain_encoding():
    """BlochSolver's chain description: per block its event count, store
    column (stored blocks numbered in order, -1 otherwise) and spoil flag
    (set after every non-empty block, Bloch.py:363-367)."""
    import numpy as np

    from feelmri_amd import device, sequences

    blocks = sequences.spoiled_gre(flip_deg=15, TR=5.0, n_dummy=2, n_lines=3, rf_dur=0.2)
    ch = device.BlockChain([(b.rf, b.G, b.dt) for b in blocks],
                           [-1] * 4 + [0, -1, 1, -1, 2, -1], [0 if b.empty else 1 for b in blocks], device="cpu")
    assert ch.n_blocks == len(blocks) == 10 and ch.n_store == 3
    assert ch.block_len.tolist() == [b.dt.size for b in blocks]
    assert ch.events.n_time == sum(b.dt.size for b in blocks)
    assert ch.spoil.tolist() == [1, 0] * 5            # imaging blocks spoil, delay blocks are empty
    np.testing.assert_array_equal(ch.events.dt.numpy(), np.concatenate([b.dt for b in blocks]))


def test_colouring_is_a_valid_partition():
    import numpy as np

    from feelmri_amd import _lib, phantoms

    m = phantoms.box_mesh((0, 0, 0), (1, 1, 1), (6, 5, 4))
    E = m.n_elems
    perm = np.empty(E, np.int32)
    offs = np.zeros(65, np.int64)
    nc = ctypes.c_int()
    el = np.ascontiguousarray(m.elems)
    _lib.call("fm_mass_p1_color", E, m.n_nodes, el.ctypes.data_as(ctypes.c_void_p),
              perm.ctypes.data_as(ctypes.c_void_p), offs.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nc))
    assert sorted(perm.tolist()) == list(range(E))
    for c in range(nc.value):
        ids = el[perm[offs[c]:offs[c + 1]]].ravel()
        assert len(ids) == len(set(ids.tolist())), "two elements of one colour share a node"


def test_incidence_lists_each_nodes_elements_in_colour_order():
    """The host incidence table behind the gather lumped mass: node n's slots
    hold the colour-order positions k (perm[k] = e) of exactly the elements
    containing n, ascending (brute force on a small mesh); padding slots hold
    -1; the size query returns the same slot count."""
    import numpy as np

    from feelmri_amd import _lib, phantoms

    m = phantoms.box_mesh((0, 0, 0), (1, 1, 1), (5, 4, 3))
    E, N = m.n_elems, m.n_nodes + 1  # one node no element references
    el = np.ascontiguousarray(m.elems)
    perm = np.empty(E, np.int32)
    offs = np.zeros(65, np.int64)
    nc = ctypes.c_int()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    _lib.call("fm_mass_p1_color", E, N, vp(el), vp(perm), vp(offs), 64, ctypes.byref(nc))
    chunk_offsets = np.empty((N + 63) // 64 + 1, np.int64)
    degree = np.empty(N, np.int32)
    n_slots = ctypes.c_int64()
    _lib.call("fm_mass_p1_incidence", E, N, vp(el), vp(perm), vp(chunk_offsets), vp(degree), None, 0,
              ctypes.byref(n_slots))
    assert n_slots.value == chunk_offsets[-1] >= 4 * E
    inc = np.empty(n_slots.value, np.int32)
    args = lambda p: (E, N, vp(el), vp(p), vp(chunk_offsets), vp(degree), vp(inc), inc.size, ctypes.byref(n_slots))
    _lib.call("fm_mass_p1_incidence", *args(perm))
    for n in range(N):
        want = [k for k, e in enumerate(perm) if n in el[e]]
        assert degree[n] == len(want)
        c = n // 64
        width = (chunk_offsets[c + 1] - chunk_offsets[c]) // 64
        got = [int(inc[chunk_offsets[c] + 64 * j + n % 64]) for j in range(width)]
        assert got[:degree[n]] == want, n
        assert all(g == -1 for g in got[degree[n]:]), n
    assert degree[N - 1] == 0  # the unreferenced node
    bad = perm.copy()
    bad[0] = bad[1]
    with pytest.raises(RuntimeError, match="permutation"):
        _lib.call("fm_mass_p1_incidence", *args(bad))
    small = np.empty(10, np.int32)
    with pytest.raises(RuntimeError, match="slots"):
        _lib.call("fm_mass_p1_incidence", E, N, vp(el), vp(perm), vp(chunk_offsets), vp(degree), vp(small),
                  small.size, ctypes.byref(n_slots))


def test_colouring_fallback_for_high_valence_meshes():
    """A Delaunay tet mesh of random points has nodes with more incident
    elements than 64 colours can separate: the colouring reports
    FM_ERR_UNSUPPORTED and P1Colouring falls back to element order (the
    host tables only; the GPU weights are tested against the oracle)."""
    import numpy as np
    from scipy.spatial import Delaunay

    from feelmri_amd import _lib

    pts = np.random.default_rng(0).uniform(size=(6000, 3))
    el = np.ascontiguousarray(Delaunay(pts).simplices, dtype=np.int32)
    E, N = el.shape[0], pts.shape[0]
    val = np.bincount(el.ravel(), minlength=N)
    perm = np.empty(E, np.int32)
    offs = np.zeros(65, np.int64)
    nc = ctypes.c_int()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = _lib.load().fm_mass_p1_color(E, N, vp(el), vp(perm), vp(offs), 64, ctypes.byref(nc))
    assert val.max() > 64
    assert rc == _lib.FM_ERR_UNSUPPORTED and "64 colours" in _lib.last_error()
    from feelmri_amd import device

    col = device.P1Colouring(el, N, device="cpu")
    assert col.method == "element-order" and col.n_colours == 0
    inc, co, deg = col.incident.numpy(), col.chunk_offsets.numpy(), col.degree.numpy()
    for n in range(0, N, 97):
        ids = [int(inc[co[n // 64] + 64 * j + n % 64]) for j in range(deg[n])]
        assert ids == sorted(np.flatnonzero((el == n).any(axis=1)).tolist())


def _declared_arity():
    """name -> number of parameters, from the prototypes in include/*.h."""
    out = {}
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(fm_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
            params = m.group(2).strip()
            out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_table_matches_the_prototypes():
    """Every ctypes signature has as many arguments as its C prototype, so a
    changed declaration cannot pass the CPU suite with a stale binding."""
    from feelmri_amd import _lib

    arity = _declared_arity()
    assert set(arity) == set(_lib.exported_symbols())
    bad = {n: (arity[n], len(a)) for n, a in _lib._SIGNATURES.items() if arity[n] != len(a)}
    assert not bad, f"(prototype, ctypes) argument counts differ: {bad}"
