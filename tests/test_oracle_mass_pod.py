"""Pins for the P1 mass and POD parts of the oracle, and for the motion
classes of feelmri_amd.Motion (host semantics).

Mass: element matrices equal the closed form |T|/20 (1 + delta_ij); lumped
weights sum to the mesh volume and equal h^3 at interior nodes of a Kuhn
lattice; consistent row sums equal the lumped weights.
POD: the contraction equals numpy.tensordot; the spline weights equal
scipy's interpolators; fold, NaN outside the knots and the velocity scaling
follow python/feelmri/Motion.py.
"""
import numpy as np
import pytest
from scipy.interpolate import PchipInterpolator

from feelmri_amd import Motion, phantoms


def test_element_matrix_closed_form(oracle):
    rng = np.random.default_rng(0)
    nodes = rng.normal(size=(4, 3))
    vals = oracle.mass_p1_element_matrices(np.array([[0, 1, 2, 3]]), nodes)[0]
    vol = abs(np.linalg.det(nodes[1:] - nodes[0])) / 6
    np.testing.assert_allclose(vals, vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-13)


def test_lumped_volume_and_interior_weight(oracle):
    m = phantoms.box_mesh((0, 0, 0), (0.03, 0.02, 0.01), (7, 5, 4))
    w = oracle.mass_p1_lumped(m.elems, m.nodes)
    np.testing.assert_allclose(w.sum(), 0.03 * 0.02 * 0.01, rtol=1e-13)
    h = np.prod(m.spacing)
    g = np.stack(np.meshgrid(*[np.arange(s) for s in m.grid_shape], indexing="ij"), -1).reshape(-1, 3)
    interior = np.all((g > 0) & (g < np.array(m.grid_shape) - 1), axis=1)
    np.testing.assert_allclose(w[interior], h, rtol=1e-12)
    # consistent mass row sums == lumped weights
    M = oracle.MassAssemble(m.elems, m.nodes)
    np.testing.assert_allclose(np.asarray(M.sum(axis=1)).ravel(), w, rtol=1e-12)
    assert abs(M - M.T).max() < 1e-18


def test_c0_slab_mesh_sizes(oracle):
    m = phantoms.slab_mesh()
    assert m.n_nodes == 61206 and m.n_elems == 300000
    w = oracle.mass_p1_lumped(m.elems, m.nodes)
    np.testing.assert_allclose(w.sum(), 1e-4, rtol=1e-12)


def test_tensordot_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    modes = rng.normal(size=(50, 3, 7))
    w = rng.normal(size=7)
    np.testing.assert_allclose(oracle.tensordot_modes_weights(modes, w), np.tensordot(modes, w, axes=([2], [0])),
                               rtol=1e-13)


def _pod(periodic=True, cls=Motion.POD):
    rng = np.random.default_rng(2)
    P, nt = 40, 12
    times = np.linspace(0, 55.0, nt)
    base = rng.normal(size=(P, 3, 3))
    coeff = np.stack([np.sin(2 * np.pi * times / 55.0 * (k + 1)) for k in range(3)], 0)
    data = np.einsum("pcr,rt->pct", base, coeff) + 0.01 * rng.normal(size=(P, 3, nt))
    return cls(times=times, data=data, n_modes=3, is_periodic=periodic), times


def test_pod_weights_are_scipy_pchip():
    pod, times = _pod()
    for i in range(pod.n_modes):
        pp = PchipInterpolator(times, pod.weights[:, i])
        t = np.linspace(0, 55, 17)
        np.testing.assert_allclose(pod._pp_batch(t)[:, i], pp(t), rtol=1e-12, atol=1e-12)


def test_pod_call_fold_and_nan():
    pod, times = _pod(periodic=True)
    pod.update_timeshift(3.0)
    a = pod(10.0)
    b = pod(10.0 + 55.0)  # folded back by the period
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-7)
    w = pod._pp_batch(13.0).astype(np.float32)
    np.testing.assert_allclose(a, np.tensordot(pod._modes, w, axes=([2], [0])), rtol=1e-6)
    nonper, _ = _pod(periodic=False)
    assert np.all(np.isnan(nonper(80.0)))  # extrapolate=False


def test_pod_velocity_scales_by_t():
    vel, _ = _pod(cls=Motion.PODVelocity)
    pos, _ = _pod(cls=Motion.POD)
    np.testing.assert_allclose(vel(7.5), pos(7.5) * 7.5, rtol=1e-5, atol=1e-5)


def test_descriptors_reproduce_call():
    pod, _ = _pod()
    resp = Motion.RespiratoryMotion(times=np.linspace(0, 100, 30), data=np.sin(np.linspace(0, 3, 30)) * 0.004,
                                    is_periodic=True, direction=np.array([0, 1, 0]))
    for obj in (pod, resp + pod, Motion.PODSum(pod, resp)):
        obj.update_timeshift(2.0)
        ts = np.array([0.3, 4.0, 17.2])
        modes = Motion.motion_modes(obj)
        coef = Motion.motion_coefficients(obj, ts)
        R = Motion.n_modes(obj)
        for i, t in enumerate(ts):
            want = obj(float(t))
            got = (np.tensordot(modes, coef[i, :R], axes=([2], [0])) if R else 0) + coef[i, R:]
            np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-8)
