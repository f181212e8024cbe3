"""Per-line GEMM (mixed form, static nodes): lines whose k and t are linear
over their whole length (radial spokes, straight readouts) take
S[Q p + q] = sum_n (c e Z^p)[n] (z^q)[n] on the matrix cores.

Checked against the FP64 oracle at the mixed-form bar (relative L2 <= 1e-4)
for every (P, Q) tiling the kernel instantiates (128 .. 2048 samples per
line), ragged node counts (not a multiple of the 32-node build, several
splits), several encodings, lines through the centre (k0 != 0), kz stacks
with per-line start times, and a plan whose lines are not linear (which
takes the other kernels).
"""
import numpy as np
import pytest

from feelmri_amd import trajectories

from test_gpu_signal import _check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_meas,n_spokes,n,v", [
    (128, 12, 3001, 1),     # Q 16, P 8 (one row tile, half used)
    (256, 9, 2500, 2),      # Q 16, P 16
    (384, 8, 4099, 1),      # Q 32, P 12
    (512, 16, 20000, 1),    # Q 32, P 16: the c2 radial tiling, many splits
    (1024, 5, 3333, 3),     # Q 32, P 32
    (1536, 4, 2000, 1),     # Q 32, P 48
    (2048, 3, 2600, 1),     # Q 32, P 64
    (256, 5, 1, 1),         # one node: one split, one build of mostly empty lanes
    (300, 4, 33, 2),        # two builds, the second with one node
])
def test_line_gemm_radial(gpu, n_meas, n_spokes, n, v):
    from feelmri_amd import device

    kloc, t = trajectories.radial(0.24, n_meas, n_spokes, t_first=0.3)
    p = device.SignalPlan(kloc, t)
    assert p.method == "direct" and p.line_gemm and p.n_line_tabs == n_spokes
    _check(kloc, t, n=n, v=v, seed=n_meas, precisions=("mixed",))


def _through_centre(n_meas, n_lines, n_kz, seed):
    """Full-diameter lines at random angles on kz planes, each line with its
    own start time: k0 = -kmax u, dk = 2 kmax u / (n_meas - 1)."""
    rng = np.random.default_rng(seed)
    kmax = 0.5 * 128 / 0.24
    th = rng.uniform(0, 2 * np.pi, n_lines)
    u = np.stack([np.cos(th), np.sin(th)], axis=1)
    s = np.linspace(-1.0, 1.0, n_meas)
    kx = np.empty((n_meas, n_lines, n_kz))
    ky = np.empty_like(kx)
    kz = np.empty_like(kx)
    t = np.empty_like(kx)
    t0 = rng.uniform(0.2, 3.0, (n_lines, n_kz))
    for k in range(n_kz):
        kx[:, :, k] = kmax * s[:, None] * u[None, :, 0]
        ky[:, :, k] = kmax * s[:, None] * u[None, :, 1]
        kz[:, :, k] = (k - n_kz // 2) / 0.12
        t[:, :, k] = t0[None, :, k] + np.linspace(0.0, 2.5, n_meas)[:, None]
    return (kx, ky, kz), t


def test_line_gemm_through_centre_kz_stack(gpu):
    from feelmri_amd import device

    kloc, t = _through_centre(300, 7, 3, seed=4)
    p = device.SignalPlan(kloc, t)
    assert p.line_gemm and p.n_line_tabs == 21
    _check(kloc, t, n=5000, v=2, seed=9, precisions=("mixed",))


def test_non_linear_line_takes_other_kernels(gpu):
    from feelmri_amd import device

    (kx, ky, kz), t = _through_centre(256, 5, 1, seed=6)
    kx = kx.copy()
    kx[100, 2, 0] += 1e-3          # one sample off its line
    p = device.SignalPlan((kx, ky, kz), t)
    assert not p.line_gemm and p.n_line_tabs == 0
    _check((kx, ky, kz), t, n=3000, seed=2, precisions=("mixed",))
