"""Image-domain helpers (CPU): the Cartesian centred inverse FFT and the
direct adjoint DFT used for radial / spiral k-space."""
import numpy as np

from feelmri_amd import recon, trajectories


def test_adjoint_image_places_a_point_source():
    kloc, t = trajectories.radial(0.24, 64, 96, t_first=0.3)
    kx, ky = kloc[0][..., 0], kloc[1][..., 0]
    x0 = (0.03, -0.045)
    s = np.exp(-2j * np.pi * (kx * x0[0] + ky * x0[1]))
    img = recon.adjoint_image(kx, ky, s, 0.24, n=48, device="cpu")
    i, j = np.unravel_index(np.abs(img).argmax(), img.shape)
    assert ((i - 24) * 0.24 / 48, (j - 24) * 0.24 / 48) == (0.03, -0.045) or \
        np.allclose(((i - 24) * 0.24 / 48, (j - 24) * 0.24 / 48), x0)


def test_adjoint_image_is_linear_and_error_is_scale_free():
    rng = np.random.default_rng(0)
    kloc, t = trajectories.spiral(0.24, 32, 128, 4, t_first=0.2)
    kx, ky = kloc[0][..., 0], kloc[1][..., 0]
    a = rng.normal(size=kx.shape) + 1j * rng.normal(size=kx.shape)
    b = rng.normal(size=kx.shape) + 1j * rng.normal(size=kx.shape)
    ia = recon.adjoint_image(kx, ky, a, 0.24, n=16, device="cpu")
    ib = recon.adjoint_image(kx, ky, b, 0.24, n=16, device="cpu")
    iab = recon.adjoint_image(kx, ky, 2.0 * a - 1j * b, 0.24, n=16, device="cpu")
    np.testing.assert_allclose(iab, 2.0 * ia - 1j * ib, rtol=1e-10, atol=1e-10)
    e = recon.adjoint_image_rel_l2(kx, ky, a * (1 + 1e-3), a, 0.24, n=16)
    assert abs(e - 1e-3) < 1e-9


def test_ktoi_of_a_cartesian_point_source():
    n = 16
    k = (np.arange(n) - n // 2)[:, None, None] * np.ones((1, n, 1))
    s = np.exp(-2j * np.pi * k * 3 / n)          # shift of 3 pixels along the first axis
    img = recon.ktoi(s)
    i = np.unravel_index(np.abs(img).argmax(), img.shape)
    assert i[0] == n // 2 + 3


def test_ktoi_matches_numpy():
    from numpy.fft import fftshift, ifftn, ifftshift

    rng = np.random.default_rng(3)
    x = rng.normal(size=(9, 8, 3, 2)) + 1j * rng.normal(size=(9, 8, 3, 2))
    ref = fftshift(ifftn(ifftshift(x, axes=[0, 1, 2]), axes=[0, 1, 2]), axes=[0, 1, 2])
    np.testing.assert_allclose(recon.ktoi(x, device="cpu"), ref, rtol=1e-12, atol=1e-14)


def test_windows():
    # Tukey: scipy's window lifted to the floor `lift` (Filters.py:49-50)
    h = recon.tukey_filter(64, width=0.9, lift=0.3)
    assert h.shape == (64,) and abs(h.min() - 0.3) < 1e-12 and abs(h.max() - 1.0) < 1e-12
    assert np.allclose(h, h[::-1])
    assert np.all(h[8:56] == 1.0)                    # flat over the central 90 %
    # Riesz: quadratic roll-off to `lift` over round(0.05 * 64) = 3 samples per end
    r = recon.riesz_filter(64, width=0.9, lift=0.3)
    assert np.all(r[3:61] == 1.0) and abs(r[0] - 0.3) < 1e-12 and abs(r[-1] - 0.3) < 1e-12
    np.testing.assert_allclose(r[:3], r[-3:][::-1])


def test_cartesian_recon_pads_even_dimensions_and_filters():
    rng = np.random.default_rng(4)
    K = rng.normal(size=(8, 6, 1)) + 1j * rng.normal(size=(8, 6, 1))
    img = recon.cartesian_recon(K, None, device="cpu")
    assert img.shape == (9, 7, 1, 1, 1)               # even readout / line counts padded by one
    Kp = np.pad(K, ((0, 1), (0, 1), (0, 0)))[..., None, None]
    np.testing.assert_allclose(img, recon.ktoi(Kp, [0, 1, 2], device="cpu"), atol=1e-14)
    # the window is applied in k-space before the transform
    imgt = recon.cartesian_recon(K, "Tukey", device="cpu")
    H = np.outer(recon.tukey_filter(9, 0.9, 0.3), recon.tukey_filter(7, 0.9, 0.3))[:, :, None, None, None]
    np.testing.assert_allclose(imgt, recon.ktoi(H * Kp, [0, 1, 2], device="cpu"), atol=1e-14)
    assert recon.image_rel_l2(K, K, "Tukey", device="cpu") == 0.0
