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


# ---------------------------------------------------------------------------
# gridding NUFFT, density weights, resize, CartesianRecon (torch on the CPU)
# ---------------------------------------------------------------------------
import pytest  # noqa: E402


@pytest.mark.parametrize("Nd,sigma,tol", [((32,), 2.0, 2e-5), ((24, 20), 2.0, 2e-5), ((25, 21), 2.0, 2e-5),
                                          ((24, 20), 1.25, 1e-3), ((12, 10, 8), 1.25, 1e-3),
                                          ((12, 10, 8), 2.0, 2e-5)])
def test_nufft_matches_the_direct_sums(Nd, sigma, tol):
    """Forward and adjoint against the sums they approximate, evaluated
    directly; the Kaiser-Bessel kernel of width 6 reaches ~1e-5 at 2x
    oversampling and ~4e-4 at 1.25x.  The adjoint is the exact conjugate
    transpose: <A x, y> = <x, A^H y> to rounding."""
    rng = np.random.default_rng(len(Nd) * 7 + int(10 * sigma))
    D, M = len(Nd), 300
    om = rng.uniform(-np.pi, np.pi, (M, D))
    plan = recon.NUFFT(om, Nd, [int(np.ceil(n * sigma)) for n in Nd], [6] * D, device="cpu")
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    x = rng.normal(size=Nd) + 1j * rng.normal(size=Nd)
    a, f = plan.adjoint(y).numpy(), plan.forward(x).numpy()
    ad, fd = recon.nufft_direct(om, Nd, y=y, device="cpu"), recon.nufft_direct(om, Nd, x=x, device="cpu")
    assert np.linalg.norm(a - ad) / np.linalg.norm(ad) < tol
    assert np.linalg.norm(f - fd) / np.linalg.norm(fd) < tol
    lhs, rhs = np.vdot(f, y), np.vdot(x.ravel(), a.ravel())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_nufft_on_the_cartesian_grid_is_the_dft():
    """Frequencies on the image's own DFT grid: the forward transform is the
    centred DFT (an independent FFT), to the kernel's accuracy."""
    N = 16
    g = 2 * np.pi * (np.arange(N) - N // 2) / N
    om = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).reshape(-1, 2)
    x = np.random.default_rng(1).normal(size=(N, N)) + 0j
    y = recon.NUFFT(om, (N, N), (2 * N, 2 * N), (6, 6), device="cpu").forward(x).numpy().reshape(N, N)
    # sum_n x[n] exp(-i 2 pi (k - N/2)(n - N/2) / N) is the centred FFT
    ref = recon.itok(x, [0, 1], device="cpu")
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-5


def test_density_weights():
    kloc, _ = trajectories.radial(0.24, 64, 48)
    kx, ky = kloc[0], kloc[1]
    w = recon.dcf_radial_stack(kx, ky)
    r = np.hypot(kx, ky).ravel()
    np.testing.assert_allclose(w, r / (r.mean() + 1e-6), rtol=1e-6)
    # a readout of uniform speed has uniform weights; a readout that doubles
    # its step halfway gets the steps, normalised to mean 1
    kxl = np.linspace(0.0, 1.0, 9)[:, None, None] * np.ones((1, 3, 2))
    np.testing.assert_allclose(recon.dcf_local_speed_readout(kxl, 0 * kxl), 1.0, rtol=1e-6)
    kxs = np.concatenate([np.arange(5.0), 4.0 + 2.0 * np.arange(1, 5)])[:, None, None]
    w = recon.dcf_local_speed_readout(kxs, 0 * kxs)
    steps = np.array([1, 1, 1, 1, 1.5, 2, 2, 2, 2])
    np.testing.assert_allclose(w, steps / steps.mean(), rtol=1e-6)
    # Pipe-Menon on a uniform grid converges to uniform weights
    N = 12
    g = 2 * np.pi * (np.arange(N) - N // 2) / N
    om = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).reshape(-1, 2)
    wpm = recon.dcf_pipe_menon(recon.NUFFT(om, (N, N), (2 * N, 2 * N), (6, 6), device="cpu"))
    assert np.max(np.abs(wpm - 1.0)) < 1e-3
    # on half-spoke radial data the weights stay positive, grow with the
    # radius away from the band edge, and make C C^H w nearly uniform
    import torch

    kloc, _ = trajectories.radial(0.24, 96, 64)
    sc = 0.24 / 48
    plan = recon.NUFFT(2 * np.pi * np.stack([(kloc[0] * sc).ravel(), (kloc[1] * sc).ravel()], 1), (48, 48),
                       (60, 60), (6, 6), device="cpu")
    wr = recon.dcf_pipe_menon(plan)
    assert wr.min() > 0 and abs(wr.mean() - 1.0) < 1e-5
    prof = wr.reshape(96, 64).mean(axis=1)
    assert np.all(np.diff(prof[:70]) > 0)
    g = plan.grid_interp(plan.grid_spread(torch.from_numpy(wr.astype(np.float64)))).numpy()
    assert g.max() / g.min() < 1.1


def test_reconstruct_nufft_places_a_point_source():
    n_img, fov = 32, 0.24
    kloc, _ = trajectories.radial(fov, 64, 96)
    kx, ky, kz = kloc
    x0 = (0.03, -0.045)
    s = np.exp(-2j * np.pi * (kx * x0[0] + ky * x0[1]))[..., None]        # (R, L, 1, 1)
    scale = fov / n_img                                                    # cycles/m -> cycles/sample
    img = recon.reconstruct_nufft(s, (kx * scale, ky * scale, kz), (n_img, n_img, 1), auto_dcw="radial-2d",
                                  oversamp=2.0, device="cpu")
    assert img.shape == (n_img, n_img) and img.dtype == np.complex64
    i, j = np.unravel_index(np.abs(img).argmax(), img.shape)
    np.testing.assert_allclose(((i - n_img // 2) * scale, (j - n_img // 2) * scale), x0, atol=1e-12)
    # two channels, root-sum-of-squares; least squares by CG fits the data better than the adjoint
    s2 = np.concatenate([s, 2 * s], axis=3)
    rss = recon.reconstruct_nufft(s2, (kx * scale, ky * scale, kz), (n_img, n_img, 1), auto_dcw=None,
                                  combine="rss", device="cpu")
    single = recon.reconstruct_nufft(s, (kx * scale, ky * scale, kz), (n_img, n_img, 1), auto_dcw=None,
                                     device="cpu")
    np.testing.assert_allclose(np.abs(rss), np.sqrt(5.0) * np.abs(single), rtol=1e-5, atol=1e-5 * np.abs(single).max())
    plan = recon.NUFFT(2 * np.pi * np.stack([(kx * scale).ravel(), (ky * scale).ravel()], 1), (n_img, n_img),
                       (40, 40), (6, 6), device="cpu")
    y = s.ravel()
    for mode in ("adjoint", "cg"):
        im = recon.reconstruct_nufft(s, (kx * scale, ky * scale, kz), (n_img, n_img, 1), auto_dcw=None, mode=mode,
                                     maxiter=10, device="cpu").astype(np.complex128)
        scale_fit = np.vdot(plan.forward(im).numpy(), y) / np.vdot(plan.forward(im).numpy(), plan.forward(im).numpy())
        res = np.linalg.norm(scale_fit * plan.forward(im).numpy() - y) / np.linalg.norm(y)
        if mode == "adjoint":
            res_adj = res
    assert res < res_adj


def test_resize_identity_constants_and_ramps():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(6, 5, 3))
    np.testing.assert_array_equal(recon.resize(a, a.shape), a)
    np.testing.assert_allclose(recon.resize(np.full((7, 6), 2.5), (13, 4)), 2.5, rtol=0, atol=1e-14)
    # linear interpolation reproduces a ramp inside the grid: output sample o
    # reads input coordinate (o + 1/2) in/out - 1/2
    ramp = np.arange(8.0)[:, None] * np.ones((1, 3))
    up = recon.resize(ramp, (16, 3))
    o = np.arange(16)
    c = (o + 0.5) * 8 / 16 - 0.5
    inside = (c >= 0) & (c <= 7)
    np.testing.assert_allclose(up[inside, 0], c[inside], atol=1e-12)
    assert up.min() >= 0.0 and up.max() <= 7.0          # clipped to the input range
    # trailing axes not named keep their length
    assert recon.resize(rng.normal(size=(9, 8, 2, 3)), (5, 4)).shape == (5, 4, 2, 3)


def test_cartesian_recon_round_trip_and_chop():
    rng = np.random.default_rng(5)
    # odd resolution, no oversampling, no window: CartesianRecon inverts itok
    img = rng.normal(size=(9, 7, 5)) + 1j * rng.normal(size=(9, 7, 5))
    K = recon.itok(img, device="cpu")
    geo = recon.CartesianGeometry((9, 7, 5))
    with pytest.warns(UserWarning, match="No filtering"):
        out = recon.CartesianRecon(K, geo, {"type": "none"}, device="cpu")
    np.testing.assert_allclose(out[..., 0, 0], img, atol=1e-12)
    # oversampled readout (2 x 8 samples, padded to 17): resized to 16, chopped to the central 8
    K2 = rng.normal(size=(16, 6, 1)) + 1j * rng.normal(size=(16, 6, 1))
    geo2 = recon.CartesianGeometry((8, 6, 1), oversampling=2)
    out2 = recon.CartesianRecon(K2, geo2, device="cpu")
    full = recon.cartesian_recon(K2, "Tukey", device="cpu")           # (17, 7, 1, 1, 1)
    rs = recon.resize(full.real, (16, 6, 1)) + 1j * recon.resize(full.imag, (16, 6, 1))
    assert out2.shape == (8, 6, 1, 1, 1)
    np.testing.assert_allclose(out2, rs[4:12], atol=1e-14)


def test_cartesian_recon_turns_reversed_lines_round():
    class TwoLineShots:
        res = np.array([9, 4, 1])
        oversampling_arr = np.array([1.0, 1.0, 1.0])
        lines_per_shot = 2
        shots = [[0, 1], [2, 3]]

    rng = np.random.default_rng(6)
    K = rng.normal(size=(9, 4, 1)) + 1j * rng.normal(size=(9, 4, 1))
    Kr = K.copy()
    Kr[:, 1] = K[::-1, 1]
    Kr[:, 3] = K[::-1, 3]
    a = recon.CartesianRecon(Kr, TwoLineShots(), device="cpu")
    b = recon.CartesianRecon(K, recon.CartesianGeometry((9, 4, 1)), device="cpu")
    np.testing.assert_allclose(a, b, atol=1e-14)
