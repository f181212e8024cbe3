"""Image-domain reconstruction and error of k-space results (measurement
helpers beside the signal path).  They run on the GPU through torch (FFTs,
gathers, atomic adds) when one is visible, else on the CPU.

* ``ktoi`` / ``itok``: the centred inverse / forward FFTs of
  python/feelmri/Math.py:4-14, over the first three axes by default.
* ``tukey_filter`` / ``riesz_filter``: the k-space apodisation windows of
  python/feelmri/Filters.py:5-51.
* ``CartesianRecon``: python/feelmri/Recon.py:13-90 -- lines read in the
  reverse direction turned round, even dimensions zero-padded by one
  sample, the readout x line window, ``ktoi``, the resize to the
  oversampled resolution and the readout chop.  ``cartesian_recon`` is the
  same without a trajectory object and without the resize.
* ``resize``: the linear resize the reference takes from scikit-image
  (absent here), composed as that library composes it from scipy.ndimage:
  a Gaussian anti-aliasing filter when any axis shrinks, then a
  linear-spline zoom with grid-mode coordinates and mirror edges, clipped to
  the input's range.
* ``NUFFT``: a Kaiser-Bessel gridding non-uniform FFT (forward and exact
  adjoint) on an oversampled grid, in place of the pynufft plan the
  reference builds (absent here); ``reconstruct_nufft`` follows
  python/feelmri/Recon.py:151-290 on it, with the density compensation
  weights of Recon.py:95-148 (Pipe-Menon on the gridding convolution, see
  ``dcf_pipe_menon``).
* ``image_rel_l2`` / ``adjoint_image`` / ``adjoint_image_rel_l2``: the
  image-domain relative L2 between two k-space results.
"""
from __future__ import annotations

import math
import warnings

import numpy as np


def _device(device=None):
    import torch

    if device is not None:
        return torch.device(device)
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def _axes(ndim, axes):
    return [i for i in range(ndim) if i < 3] if axes is None else list(axes)


def ktoi(x, axes=None, device=None) -> np.ndarray:
    """Centred inverse FFT over ``axes`` (default: the first three):
    fftshift(ifftn(ifftshift(x)))."""
    import torch

    x = np.asarray(x, dtype=np.complex128)
    axes = _axes(x.ndim, axes)
    t = torch.as_tensor(x, device=_device(device))
    t = torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(t, dim=axes), dim=axes), dim=axes)
    return t.cpu().numpy()


def itok(x, axes=None, device=None) -> np.ndarray:
    """Centred forward FFT over ``axes`` (default: the first three):
    fftshift(fftn(ifftshift(x)))."""
    import torch

    x = np.asarray(x, dtype=np.complex128)
    axes = _axes(x.ndim, axes)
    t = torch.as_tensor(x, device=_device(device))
    t = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(t, dim=axes), dim=axes), dim=axes)
    return t.cpu().numpy()


def tukey_filter(size: int, width: float = 0.6, lift: float = 0.7) -> np.ndarray:
    """Filters.py:34-51: a Tukey window with taper fraction 1 - width,
    lifted to a floor of ``lift``."""
    from scipy.signal.windows import tukey

    return tukey(size, alpha=1.0 - width) * (1.0 - lift) + lift


def riesz_filter(size: int, width: float = 0.6, lift: float = 0.7) -> np.ndarray:
    """Filters.py:5-31: quadratic roll-off to ``lift`` over the outer
    round((1 - width) / 2 * size) samples at each end."""
    n_edge = int(np.round((1.0 - width) / 2 * size))
    h = np.ones(size)
    ramp = 1.0 - (np.linspace(0, n_edge / 2, n_edge) / (n_edge / 2)) ** 2 * (1.0 - lift) if n_edge else \
        np.zeros(0)
    h[:n_edge] *= ramp[::-1]
    h[size - n_edge:] *= ramp
    return h


def _window(kind, n0, n1):
    if kind == "Tukey":
        return tukey_filter(n0, 0.9, 0.3), tukey_filter(n1, 0.9, 0.3)
    if kind == "Riesz":
        return riesz_filter(n0, 0.9, 0.3), riesz_filter(n1, 0.9, 0.3)
    return None


def cartesian_recon(K, filter_type: str | None = "Tukey", device=None) -> np.ndarray:
    """Image of Cartesian k-space ``K`` (n_meas, n_lines, n_kz[, ...]) as
    CartesianRecon forms it (Recon.py:13-90), for one line per shot and
    before the final resize.  The windows use width 0.9 and lift 0.3 whatever
    the filter dict says, as the reference does (Recon.py:64-69)."""
    K = np.asarray(K, dtype=np.complex128)
    while K.ndim < 5:
        K = K[..., None]
    pad = [(0, 1 if K.shape[a] % 2 == 0 else 0) if a < 3 else (0, 0) for a in range(K.ndim)]
    K = np.pad(K, pad, mode="constant")
    w = _window(filter_type, K.shape[0], K.shape[1])
    if w is not None:
        K = np.outer(*w)[:, :, None, None, None] * K
    return ktoi(K, [0, 1, 2], device=device)


# ---------------------------------------------------------------------------
# resize (scikit-image's linear resize, composed from scipy.ndimage)
# ---------------------------------------------------------------------------
def resize(image, output_shape) -> np.ndarray:
    """Linear resize of a real array to ``output_shape`` (trailing axes not
    named keep their length), composed as scikit-image's ``resize`` with its
    defaults (order 1, mode 'reflect', anti-aliasing when any axis shrinks,
    clipping to the input range): Gaussian filter of sigma
    max(0, (in/out - 1)/2) per axis with mirror edges, then a linear-spline
    zoom by out/in with grid-mode coordinates -- output sample o reads input
    coordinate (o + 1/2) in/out - 1/2 -- and mirror edges.

    scikit-image is absent here, so this composition is not checked against
    it; the tests pin identity, constants, linear ramps and the grid-mode
    coordinate map."""
    from scipy import ndimage as ndi

    image = np.asarray(image, dtype=np.float64)
    out_shape = tuple(int(s) for s in output_shape) + image.shape[len(output_shape):]
    if out_shape == image.shape:
        return image.copy()
    factors = np.divide(image.shape, out_shape)
    if np.any(np.asarray(out_shape) < np.asarray(image.shape)):
        sigma = np.maximum(0.0, (factors - 1.0) / 2.0)
        filtered = ndi.gaussian_filter(image, sigma, cval=0.0, mode="mirror")
    else:
        filtered = image
    out = ndi.zoom(filtered, [1.0 / f for f in factors], order=1, mode="mirror", cval=0.0, grid_mode=True)
    return np.clip(out, image.min(), image.max())


def CartesianRecon(K, trajectory, filter=None, device=None) -> np.ndarray:  # noqa: A002 (reference name)
    """python/feelmri/Recon.py:13-90.  ``trajectory`` needs ``res`` and
    ``oversampling_arr`` (the reference's CartesianStack has both; so does
    ``CartesianGeometry``); lines of shots with several lines per shot
    (``lines_per_shot``, ``shots``) read in the reverse direction are turned
    round first.  Returns the complex image at the oversampled resolution,
    chopped to ``res[0]`` along the readout when the readout was
    oversampled."""
    if filter is None:
        filter = {"type": "Tukey", "width": 0.9, "lift": 0.3}  # noqa: A001
    K = np.array(K, dtype=np.complex128)
    while K.ndim < 5:
        K = K[..., None]
    if int(getattr(trajectory, "lines_per_shot", 1)) > 1:
        for shot in trajectory.shots:
            for i, line in enumerate(shot):
                if i % 2 == 1:   # read out in the reverse direction
                    K[:, line, ...] = K[::-1, line, ...].copy()
    res = np.asarray(trajectory.res, dtype=np.int64)
    pad = [(0, 1 if (a < 3 and res[a] % 2 == 0) else 0) for a in range(5)]
    K = np.pad(K, pad, mode="constant")
    w = _window(filter.get("type"), K.shape[0], K.shape[1])
    if w is None:
        warnings.warn("Unknown filter type. No filtering applied.")
    else:
        K = np.outer(*w)[:, :, None, None, None] * K
    img = ktoi(K, [0, 1, 2], device=device)
    shape = tuple(np.asarray(trajectory.oversampling_arr * res).astype(np.int64)) + img.shape[3:]
    img = resize(img.real, shape) + 1j * resize(img.imag, shape)
    enc, rec = K.shape[0], int(res[0])
    if enc != rec:
        lo = (enc - rec) // 2
        img = img[lo:lo + rec]
    return img


class CartesianGeometry:
    """The attributes ``CartesianRecon`` reads from a Cartesian trajectory:
    resolution per axis and the readout oversampling (one line per shot)."""

    def __init__(self, res, oversampling: int = 1):
        self.res = np.asarray(res, dtype=np.int64)
        self.oversampling = int(oversampling)
        self.oversampling_arr = np.array([oversampling, 1.0, 1.0])
        self.lines_per_shot = 1


# ---------------------------------------------------------------------------
# NUFFT: Kaiser-Bessel gridding
# ---------------------------------------------------------------------------
def kaiser_bessel_beta(width: int, sigma: float) -> float:
    """Shape parameter of the Kaiser-Bessel gridding kernel for kernel width
    ``width`` (grid samples) and oversampling ``sigma`` (Beatty, Nishimura and
    Pauly, IEEE TMI 24(6), 2005, eq. 5)."""
    return math.pi * math.sqrt((width / sigma) ** 2 * (sigma - 0.5) ** 2 - 0.8)


class NUFFT:
    """Non-uniform FFT of an image on ``Nd`` samples at frequencies ``om``
    (M, D) in radians per image sample:

        forward: y_m = sum_n x[n] exp(-i om_m . (n - Nd // 2))
        adjoint: x[n] = sum_m y_m exp(+i om_m . (n - Nd // 2))

    by gridding onto the ``Kd`` oversampled grid with a separable
    Kaiser-Bessel kernel of ``Jd`` samples per axis, an FFT of that grid, and
    division by the kernel's Fourier transform over the image.  The adjoint
    is the exact conjugate transpose of the forward operator.  Device tensors
    throughout; the spreading of the adjoint uses atomic adds, so its
    rounding order varies between runs."""

    def __init__(self, om, Nd, Kd=None, Jd=None, device=None, chunk: int = 1 << 15):
        import torch

        om = np.asarray(om, dtype=np.float64)
        if om.ndim == 1:
            om = om[:, None]
        self.M, self.D = om.shape
        self.Nd = tuple(int(n) for n in Nd)
        if len(self.Nd) != self.D:
            raise ValueError(f"om has {self.D} dimensions, Nd {len(self.Nd)}")
        self.Kd = tuple(int(k) for k in (Kd if Kd is not None else [2 * n for n in self.Nd]))
        self.Jd = tuple(int(j) for j in (Jd if Jd is not None else [6] * self.D))
        if any(k < n for k, n in zip(self.Kd, self.Nd)) or any(j > k for j, k in zip(self.Jd, self.Kd)):
            raise ValueError("need Nd <= Kd and Jd <= Kd")
        self.dev = _device(device)
        self.chunk = int(chunk)
        self.beta = [kaiser_bessel_beta(j, k / n) if k > n else kaiser_bessel_beta(j, 1.0 + 1e-9)
                     for j, k, n in zip(self.Jd, self.Kd, self.Nd)]
        om_t = torch.as_tensor(om, device=self.dev)
        self._idx, self._w = [], []
        for d in range(self.D):
            K, J = self.Kd[d], self.Jd[d]
            u = om_t[:, d] * (K / (2.0 * math.pi))                    # grid units
            g = (torch.floor(u - J / 2.0) + 1.0)[:, None] + torch.arange(J, device=self.dev, dtype=torch.float64)[None, :]
            s = u[:, None] - g                                         # in [-J/2, J/2)
            q = torch.clamp(1.0 - (2.0 * s / J) ** 2, min=0.0)
            # kernel scaled to a peak of 1 (its transform below likewise)
            self._w.append(torch.special.i0(self.beta[d] * torch.sqrt(q)) / float(np.i0(self.beta[d])))
            self._idx.append(torch.remainder(g.to(torch.int64), K))
        # deapodisation: the kernel's Fourier transform at the image offsets
        deap = torch.ones(self.Nd, dtype=torch.float64, device=self.dev)
        self._place = []
        for d in range(self.D):
            N, K, J, b = self.Nd[d], self.Kd[d], self.Jd[d], self.beta[d]
            p = torch.arange(N, dtype=torch.float64, device=self.dev) - (N // 2)
            a2 = b * b - (math.pi * J * p / K) ** 2
            r = torch.sqrt(torch.abs(a2))
            ft = torch.where(a2 > 0, torch.sinh(r) / torch.where(r > 0, r, 1.0),
                             torch.where(a2 < 0, torch.sin(r) / torch.where(r > 0, r, 1.0),
                                         torch.ones_like(r))) * (J / float(np.i0(b)))
            shape = [1] * self.D
            shape[d] = N
            deap = deap * ft.reshape(shape)
            self._place.append(torch.remainder(p.to(torch.int64), K))
        self.deap = deap

    def _chunk_tables(self, m0, m1):
        import torch

        idx = torch.zeros((m1 - m0, 1), dtype=torch.int64, device=self.dev)
        w = torch.ones((m1 - m0, 1), dtype=torch.float64, device=self.dev)
        for d in range(self.D):
            idx = (idx[:, :, None] * self.Kd[d] + self._idx[d][m0:m1, None, :]).reshape(m1 - m0, -1)
            w = (w[:, :, None] * self._w[d][m0:m1, None, :]).reshape(m1 - m0, -1)
        return idx, w

    def _image_to_grid(self, x):
        import torch

        grid = torch.zeros(self.Kd, dtype=torch.complex128, device=self.dev)
        grid[torch.meshgrid(*self._place, indexing="ij")] = x / self.deap
        return grid

    def forward(self, x) -> "torch.Tensor":  # noqa: F821
        import torch

        x = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x, device=self.dev)
        x = x.to(torch.complex128).reshape(self.Nd)
        X = torch.fft.fftn(self._image_to_grid(x)).reshape(-1)
        y = torch.empty(self.M, dtype=torch.complex128, device=self.dev)
        for m0 in range(0, self.M, self.chunk):
            m1 = min(self.M, m0 + self.chunk)
            idx, w = self._chunk_tables(m0, m1)
            y[m0:m1] = (X[idx] * w).sum(dim=1)
        return y

    def adjoint(self, y) -> "torch.Tensor":  # noqa: F821
        import torch

        y = torch.as_tensor(np.asarray(y) if not isinstance(y, torch.Tensor) else y, device=self.dev)
        y = y.to(torch.complex128).reshape(-1)
        if y.numel() != self.M:
            raise ValueError(f"expected {self.M} samples, got {y.numel()}")
        G = torch.zeros((int(np.prod(self.Kd)), 2), dtype=torch.float64, device=self.dev)
        yr = torch.view_as_real(y)
        for m0 in range(0, self.M, self.chunk):
            m1 = min(self.M, m0 + self.chunk)
            idx, w = self._chunk_tables(m0, m1)
            G.index_add_(0, idx.reshape(-1), (yr[m0:m1, None, :] * w[:, :, None]).reshape(-1, 2))
        Z = torch.fft.ifftn(torch.view_as_complex(G).reshape(self.Kd)) * float(np.prod(self.Kd))
        return Z[torch.meshgrid(*self._place, indexing="ij")] / self.deap

    def grid_spread(self, v):
        """C^H v: real sample values spread onto the oversampled grid with
        the kernel (no FFT, no deapodisation); flat (prod Kd,) float64."""
        import torch

        G = torch.zeros(int(np.prod(self.Kd)), dtype=torch.float64, device=self.dev)
        for m0 in range(0, self.M, self.chunk):
            m1 = min(self.M, m0 + self.chunk)
            idx, w = self._chunk_tables(m0, m1)
            G.index_add_(0, idx.reshape(-1), (v[m0:m1, None] * w).reshape(-1))
        return G

    def grid_interp(self, G):
        """C G: kernel-weighted sums of the grid at every sample, (M,)."""
        import torch

        out = torch.empty(self.M, dtype=torch.float64, device=self.dev)
        for m0 in range(0, self.M, self.chunk):
            m1 = min(self.M, m0 + self.chunk)
            idx, w = self._chunk_tables(m0, m1)
            out[m0:m1] = (G[idx] * w).sum(dim=1)
        return out

    def solve(self, y, maxiter: int = 30, tol: float = 1e-6):
        """Least-squares image: conjugate gradients on A^H A x = A^H y."""
        import torch

        b = self.adjoint(y)
        x = torch.zeros_like(b)
        r = b.clone()
        p = r.clone()
        rs = torch.vdot(r.reshape(-1), r.reshape(-1)).real
        b_norm = math.sqrt(float(rs)) or 1.0
        for _ in range(int(maxiter)):
            Ap = self.adjoint(self.forward(p))
            alpha = rs / torch.vdot(p.reshape(-1), Ap.reshape(-1)).real
            x = x + alpha * p
            r = r - alpha * Ap
            rs_new = torch.vdot(r.reshape(-1), r.reshape(-1)).real
            if math.sqrt(float(rs_new)) <= tol * b_norm:
                break
            p = r + (rs_new / rs) * p
            rs = rs_new
        return x


def nufft_direct(om, Nd, y=None, x=None, device=None):
    """The sums ``NUFFT`` approximates, evaluated directly (tests and small
    cases): the adjoint of ``y`` or the forward transform of ``x``."""
    import torch

    dev = _device(device)
    om = torch.as_tensor(np.asarray(om, dtype=np.float64).reshape(len(om), -1), device=dev)
    axes = [torch.arange(n, dtype=torch.float64, device=dev) - (n // 2) for n in Nd]
    P = torch.stack([a.reshape(-1) for a in torch.meshgrid(*axes, indexing="ij")], dim=1)   # (prod Nd, D)
    E = torch.polar(torch.ones((om.shape[0], P.shape[0]), dtype=torch.float64, device=dev), -(om @ P.T))
    if y is not None:
        yv = torch.as_tensor(np.asarray(y, dtype=np.complex128).reshape(-1), device=dev)
        return (E.conj().T @ yv).reshape(tuple(Nd)).cpu().numpy()
    xv = torch.as_tensor(np.asarray(x, dtype=np.complex128).reshape(-1), device=dev)
    return (E @ xv).cpu().numpy()


# ---------------------------------------------------------------------------
# density compensation (Recon.py:95-148)
# ---------------------------------------------------------------------------
def dcf_pipe_menon(nufft: NUFFT, n_iter: int = 20, eps: float = 1e-8) -> np.ndarray:
    """Iterative weights w <- w / (C C^H w) from w = 1, normalised to mean 1,
    with C the plan's Kaiser-Bessel interpolation from the oversampled grid
    (Pipe and Menon, MRM 41(1), 1999).  Recon.py:95-107 iterates with the
    full forward(adjoint(w)) of its NUFFT plan instead; that operator samples
    the point-spread function, whose negative side lobes make the iteration
    diverge on half-spoke radial data (weights turn negative within a few
    iterations), so this form keeps the published one, whose kernel is
    positive."""
    import torch

    w = torch.ones(nufft.M, dtype=torch.float64, device=nufft.dev)
    for _ in range(int(n_iter)):
        w = w / (nufft.grid_interp(nufft.grid_spread(w)) + eps)
    return (w / (w.mean() + eps)).cpu().numpy().astype(np.float32)


def dcf_radial_stack(kx, ky, eps: float = 1e-6, per_slice_normalize: bool = True) -> np.ndarray:
    """In-plane radius sqrt(kx^2 + ky^2) normalised per slice (or overall)
    to mean 1 (Recon.py:110-122); kx, ky (R, L, S); returns (R*L*S,)."""
    kx, ky = np.asarray(kx, dtype=np.float64), np.asarray(ky, dtype=np.float64)
    r = np.hypot(kx, ky)
    mean = r.mean(axis=(0, 1), keepdims=True) if per_slice_normalize else r.mean()
    return (r / (mean + eps)).astype(np.float32).ravel()


def dcf_local_speed_readout(kx, ky, kz=None, eps: float = 1e-8) -> np.ndarray:
    """Sample spacing along each readout, per (line, slice): the step to the
    next sample at the first, the step from the previous at the last, the
    mean of both inside, normalised per (line, slice) to mean 1
    (Recon.py:124-148)."""
    d = [np.asarray(kx, dtype=np.float64), np.asarray(ky, dtype=np.float64)]
    if kz is not None:
        d.append(np.asarray(kz, dtype=np.float64))
    step = np.sqrt(sum(np.diff(a, axis=0) ** 2 for a in d))
    w = np.empty(d[0].shape)
    w[0], w[-1] = step[0], step[-1]
    w[1:-1] = 0.5 * (step[:-1] + step[1:])
    return (w / (w.mean(axis=0, keepdims=True) + eps)).astype(np.float32).ravel()


def reconstruct_nufft(kdata, ktraj, img_shape, *, dcw=None, auto_dcw: str | None = "pipe-menon",
                      oversamp: float = 1.25, kernel_size: int = 6, mode: str = "adjoint", maxiter: int = 30,
                      tol: float = 1e-6, combine: str | None = None, device=None) -> np.ndarray:
    """python/feelmri/Recon.py:151-290 on the gridding ``NUFFT``: kdata
    (R, L, S, C), ktraj = (kx, ky, kz) each (R, L, S) in cycles per image
    sample (|k| <= 1/2), img_shape (Nx, Ny[, Nz]).  The samples are taken in
    C order, a trajectory with kz = 0 throughout is reconstructed in 2-D,
    density weights multiply the data before the adjoint (``mode="adjoint"``)
    or the conjugate-gradient solve (``mode="cg"``).  Returns complex64
    (*img_shape,) for one channel or ``combine="rss"``, else (C, *img_shape)."""
    kx, ky, kz = ktraj
    kx = np.asarray(kx, dtype=np.float64)
    ky = np.asarray(ky, dtype=np.float64)
    kz = np.zeros_like(kx) if kz is None else np.asarray(kz, dtype=np.float64)
    if kx.shape != ky.shape or kx.shape != kz.shape:
        raise ValueError("kx, ky, kz must share the same shape (R, L, S).")
    kdata = np.asarray(kdata)
    if kdata.shape[:3] != kx.shape:
        raise ValueError("kdata (R,L,S, C) must match ktraj shapes on first 3 dims.")
    M = kx.size
    C = kdata.shape[3]
    om = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    is_2d = np.allclose(om[:, 2], 0.0)
    Nd = tuple(int(n) for n in img_shape)
    if is_2d:
        om = om[:, :2]
        if len(Nd) == 3 and Nd[2] == 1:
            Nd = Nd[:2]
    if len(Nd) not in (2, 3) or len(Nd) != om.shape[1]:
        raise ValueError("img_shape must be 2D or 3D, matching the trajectory")
    Kd = tuple(int(np.ceil(n * oversamp)) for n in Nd)
    plan = NUFFT(2.0 * np.pi * om, Nd, Kd, [kernel_size] * len(Nd), device=device)
    if dcw is None and auto_dcw is not None:
        method = auto_dcw.lower()
        if method == "pipe-menon":
            dcw = dcf_pipe_menon(plan, n_iter=20)
        elif method == "radial-2d":
            dcw = dcf_radial_stack(kx, ky)
        elif method == "speed":
            dcw = dcf_local_speed_readout(kx, ky, None if is_2d else kz)
        else:
            raise ValueError("auto_dcw must be None, 'pipe-menon', 'radial-2d' or 'speed'")
    if dcw is not None:
        dcw = np.asarray(dcw, dtype=np.float32)
        if dcw.shape != (M,):
            raise ValueError(f"dcw must be shape ({M},), got {dcw.shape}")
    samples = kdata.reshape(M, C)
    imgs = []
    for c in range(C):
        y = samples[:, c].astype(np.complex64)
        if dcw is not None:
            y = y * dcw
        if mode.lower() == "adjoint":
            x = plan.adjoint(y)
        elif mode.lower() == "cg":
            x = plan.solve(y, maxiter=maxiter, tol=tol)
        else:
            raise ValueError("mode must be 'adjoint' or 'cg'.")
        imgs.append(x.cpu().numpy().astype(np.complex64))
    img = np.stack(imgs, axis=0)
    if C == 1:
        return img[0]
    if combine is None:
        return img
    if combine.lower() == "rss":
        return np.sqrt(np.sum(np.abs(img) ** 2, axis=0)).astype(np.complex64)
    raise ValueError("combine must be None or 'rss'")


# ---------------------------------------------------------------------------
# image-domain error
# ---------------------------------------------------------------------------
def _rel(a, b) -> float:
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def image_rel_l2(k_test, k_ref, filter_type: str | None = None, device=None) -> float:
    """Relative L2 between the images of two Cartesian k-space arrays:
    plain ``ktoi`` (filter_type None) or ``cartesian_recon`` with a window."""
    if filter_type is None:
        return _rel(ktoi(k_test, device=device), ktoi(k_ref, device=device))
    return _rel(cartesian_recon(k_test, filter_type, device), cartesian_recon(k_ref, filter_type, device))


def adjoint_image(kx, ky, s, fov: float, n: int = 128, dcf: bool = True, device=None,
                  chunk: int = 4096) -> np.ndarray:
    """Image of 2-D non-Cartesian k-space samples on an n x n grid over the FOV
    by the direct adjoint DFT, I(x) = sum_s w_s S_s exp(+2 pi i k_s . x), with
    the radial density compensation w_s = |k_s| (``dcf``).  k in cycles/m,
    fov in m; returns (n, n) complex128."""
    import torch

    dev = _device(device)
    kx = torch.as_tensor(np.ascontiguousarray(np.ravel(kx), dtype=np.float64), device=dev)
    ky = torch.as_tensor(np.ascontiguousarray(np.ravel(ky), dtype=np.float64), device=dev)
    sv = torch.as_tensor(np.ascontiguousarray(np.ravel(s), dtype=np.complex128), device=dev)
    if dcf:
        sv = sv * torch.sqrt(kx * kx + ky * ky)
    g = (torch.arange(n, dtype=torch.float64, device=dev) - n // 2) * (fov / n)
    X, Y = torch.meshgrid(g, g, indexing="ij")
    X, Y = X.reshape(-1), Y.reshape(-1)
    out = torch.zeros(n * n, dtype=torch.complex128, device=dev)
    for c0 in range(0, kx.numel(), chunk):
        c1 = min(kx.numel(), c0 + chunk)
        ph = 2.0 * np.pi * (kx[c0:c1, None] * X[None, :] + ky[c0:c1, None] * Y[None, :])
        out += sv[c0:c1] @ torch.polar(torch.ones_like(ph), ph)
    return out.reshape(n, n).cpu().numpy()


def adjoint_image_rel_l2(kx, ky, s_test, s_ref, fov: float, n: int = 128) -> float:
    return _rel(adjoint_image(kx, ky, s_test, fov, n), adjoint_image(kx, ky, s_ref, fov, n))


def nufft_image_rel_l2(kx, ky, s_test, s_ref, fov: float, n: int = 128, device=None) -> float:
    """Relative L2 between the gridding-NUFFT adjoint images (radial ramp
    weights) of two 2-D non-Cartesian sample sets; k in cycles/m, on an
    n x n grid over ``fov``."""
    kx = np.asarray(kx, dtype=np.float64).ravel()
    ky = np.asarray(ky, dtype=np.float64).ravel()
    om = 2.0 * np.pi * np.stack([kx, ky], axis=1) * (fov / n)
    if np.abs(om).max() > np.pi:
        raise ValueError("k beyond the grid's Nyquist band: enlarge n")
    plan = NUFFT(om, (n, n), (2 * n, 2 * n), (6, 6), device=device)
    w = np.hypot(kx, ky)
    a = plan.adjoint(np.asarray(s_test).ravel() * w).cpu().numpy()
    b = plan.adjoint(np.asarray(s_ref).ravel() * w).cpu().numpy()
    return _rel(a, b)
