"""Image-domain error path on the GPU: centred FFTs, CartesianRecon, the
gridding NUFFT and reconstruct_nufft against CPU references (numpy FFT, the
directly evaluated sums, the same algorithm on the CPU), and the image error
of GPU signal results against the FP64 oracle.

Tolerances:
  FFTs, CartesianRecon                    <= 1e-12 relative (same transform)
  NUFFT vs direct sums, 2x oversampling   <= 2e-5 (the kernel's accuracy)
  NUFFT on the GPU vs on the CPU          <= 1e-12 (atomic-add order only)
  reconstruct_nufft GPU vs CPU            <= 1e-6 (complex64 output)
  image error of the mixed signal form    <= 1e-4 (the k-space bar)
"""
import numpy as np
import pytest
import scipy.sparse as sp

from feelmri_amd import recon, trajectories

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_centred_ffts_match_numpy(gpu):
    from numpy.fft import fftn, fftshift, ifftn, ifftshift

    rng = np.random.default_rng(0)
    x = rng.normal(size=(33, 24, 8, 2)) + 1j * rng.normal(size=(33, 24, 8, 2))
    ax = [0, 1, 2]
    assert rel(recon.ktoi(x, device=gpu), fftshift(ifftn(ifftshift(x, axes=ax), axes=ax), axes=ax)) <= 1e-12
    assert rel(recon.itok(x, device=gpu), fftshift(fftn(ifftshift(x, axes=ax), axes=ax), axes=ax)) <= 1e-12


def test_cartesian_recon_gpu_equals_cpu(gpu):
    rng = np.random.default_rng(1)
    K = rng.normal(size=(32, 24, 6, 2)) + 1j * rng.normal(size=(32, 24, 6, 2))
    geo = recon.CartesianGeometry((16, 24, 6), oversampling=2)
    a = recon.CartesianRecon(K, geo, device=gpu)
    b = recon.CartesianRecon(K, geo, device="cpu")
    assert a.shape == (16, 24, 6, 2, 1)
    assert rel(a, b) <= 1e-12


@pytest.mark.parametrize("Nd", [(40, 36), (20, 18, 12)])
def test_nufft_gpu(gpu, Nd):
    rng = np.random.default_rng(len(Nd))
    D, M = len(Nd), 2000
    om = rng.uniform(-np.pi, np.pi, (M, D))
    Kd = [2 * n for n in Nd]
    g = recon.NUFFT(om, Nd, Kd, [6] * D, device=gpu)
    c = recon.NUFFT(om, Nd, Kd, [6] * D, device="cpu")
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    x = rng.normal(size=Nd) + 1j * rng.normal(size=Nd)
    ag, fg = g.adjoint(y).cpu().numpy(), g.forward(x).cpu().numpy()
    assert rel(ag, recon.nufft_direct(om, Nd, y=y, device="cpu")) <= 2e-5
    assert rel(fg, recon.nufft_direct(om, Nd, x=x, device="cpu")) <= 2e-5
    assert rel(ag, c.adjoint(y).numpy()) <= 1e-12
    assert rel(fg, c.forward(x).numpy()) <= 1e-12


@pytest.mark.parametrize("dcw", ["radial-2d", "pipe-menon", "speed"])
def test_reconstruct_nufft_gpu_equals_cpu(gpu, dcw):
    n_img, fov = 48, 0.24
    kloc, _ = trajectories.radial(fov, 96, 64)
    kx, ky, kz = kloc
    rng = np.random.default_rng(2)
    s = (rng.normal(size=kx.shape) + 1j * rng.normal(size=kx.shape))[..., None]
    sc = fov / n_img
    args = (s, (kx * sc, ky * sc, kz), (n_img, n_img, 1))
    a = recon.reconstruct_nufft(*args, auto_dcw=dcw, device=gpu)
    b = recon.reconstruct_nufft(*args, auto_dcw=dcw, device="cpu")
    assert a.shape == (n_img, n_img)
    assert rel(a, b) <= 1e-6


def _signal_pair(kloc, t, n=4000, seed=0):
    """GPU mixed-form k-space and the FP64 oracle's, same seeded nodes."""
    import oracle
    from feelmri_amd import MRI

    rng = np.random.default_rng(seed)
    r0 = rng.uniform(-0.06, 0.06, (n, 3))
    phi = rng.normal(0, 2 * np.pi * 20 / 1000, n)
    T2 = rng.uniform(40, 100, n)
    mxy = (rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))) * 0.1
    M = sp.diags(rng.uniform(0.5, 1.5, n) * 1e-6)
    ref = oracle.Signal(0, M, kloc, t, r0, phi, T2, mxy, None)
    got = MRI.Signal(0, M, kloc, t, r0, phi, T2, mxy, None, precision="mixed")
    return got, ref


def test_image_error_of_mixed_cartesian_signal(gpu):
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (24, 20, 6), oversampling=2, t_first=1.2)
    got, ref = _signal_pair(kloc, t)
    geo = recon.CartesianGeometry((24, 20, 6), oversampling=2)
    e_k = rel(got, ref)
    e_img = rel(recon.CartesianRecon(got, geo, device=gpu), recon.CartesianRecon(ref, geo, device=gpu))
    print(f"k-space {e_k:.2e}, image {e_img:.2e}")
    assert e_k <= 1e-4 and e_img <= 1e-4


def test_image_error_of_mixed_radial_signal(gpu):
    fov, n_img = 0.24, 32
    kloc, t = trajectories.radial(fov, 64, 48, t_first=0.3)
    got, ref = _signal_pair(kloc, t)
    sc = fov / n_img
    kk = (kloc[0] * sc, kloc[1] * sc, kloc[2])
    a = recon.reconstruct_nufft(got, kk, (n_img, n_img, 1), auto_dcw="radial-2d", device=gpu)
    b = recon.reconstruct_nufft(ref, kk, (n_img, n_img, 1), auto_dcw="radial-2d", device=gpu)
    e_k, e_img = rel(got, ref), rel(a, b)
    e_img2 = recon.nufft_image_rel_l2(kloc[0], kloc[1], got, ref, fov, n_img, device=gpu)
    print(f"k-space {e_k:.2e}, image {e_img:.2e} / {e_img2:.2e}")
    assert e_k <= 1e-4 and e_img <= 1e-4 and e_img2 <= 1e-4
