"""Host logic of the drop-in shims (no GPU): recognising a lumped (diagonal)
mass matrix in every sparse format the reference's callers produce
(``lil`` with ``setdiag`` in python/feelmri/Phantom.py:395-398,
``scipy.sparse.diags``, CSR / CSC / COO), and the content-keyed residency
table behind the repeated node arrays of per-line ``Signal`` calls."""
import numpy as np
import pytest
import scipy.sparse as sp

from feelmri_amd.MRI import _lumped_diagonal


def test_lumped_diagonal_formats():
    w = np.random.default_rng(0).uniform(0.5, 1.5, 50)
    for M in (sp.diags(w), sp.csr_matrix(sp.diags(w)), sp.csc_matrix(sp.diags(w)), sp.coo_matrix(sp.diags(w))):
        np.testing.assert_array_equal(_lumped_diagonal(M), w)
    lil = sp.lil_matrix((50, 50))
    lil.setdiag(w)
    np.testing.assert_array_equal(_lumped_diagonal(lil), w)
    np.testing.assert_array_equal(_lumped_diagonal(w), w)
    assert _lumped_diagonal(None) is None
    off = sp.diags([w, w[:-1]], [0, 1])
    for M in (off, sp.csr_matrix(off), sp.csc_matrix(off), sp.coo_matrix(off)):
        assert _lumped_diagonal(M) is None
    # explicitly stored zeros off the diagonal of a DIA matrix are still diagonal
    z = sp.dia_matrix((np.vstack([w, np.zeros(50)]), [0, 2]), shape=(50, 50))
    np.testing.assert_array_equal(_lumped_diagonal(z), w)


def test_resident_table_is_keyed_by_content():
    torch = pytest.importorskip("torch")
    from feelmri_amd import resident

    resident.clear()
    a = np.arange(40_000, dtype=np.float64)
    t1 = resident.to_device(a, torch.float64, "cpu")
    t2 = resident.to_device(a.copy(), torch.float64, "cpu")
    assert t1 is t2 and resident.size()[0] == 1
    a[123] += 1.0                                  # changed in place: a new upload
    t3 = resident.to_device(a, torch.float64, "cpu")
    assert t3 is not t1 and float(t3[123]) == 124.0 and float(t1[123]) == 123.0
    small = np.ones(10)
    assert resident.to_device(small, torch.float64, "cpu") is not resident.to_device(small, torch.float64, "cpu")
    old = resident.budget_bytes
    try:
        resident.budget_bytes = 400_000
        for k in range(4):
            resident.to_device(np.full(40_000, float(k)), torch.float64, "cpu")
        assert resident.size()[1] <= 400_000
    finally:
        resident.budget_bytes = old
        resident.clear()
