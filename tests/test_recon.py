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
