"""Host analysis of trajectories (runs on the CPU, no device): the Cartesian
factorisation test, distinct-t groups, readout runs and t-grouped slots."""
import numpy as np
import pytest

from feelmri_amd import trajectories
from feelmri_amd._lib import FeelmriHipError
from feelmri_amd.device import analyse_trajectory


def test_cartesian_stack_factorises():
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (16, 12, 4), t_first=1.2, t_readout=2.0)
    a = analyse_trajectory(kloc, t)
    assert a["method"] == "cartesian"
    assert a["n_tgroups"] == 16                      # t depends on the readout index only
    assert a["n_runs"] == 12 * 4 * 1                 # one 16-sample run per line
    # forcing the direct method keeps the same groups and runs
    d = analyse_trajectory(kloc, t, method="direct")
    assert d["method"] == "direct" and d["n_runs"] == a["n_runs"]


def test_rotated_grid_factorises_by_general_split():
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (10, 6, 3), t_first=1.0)
    c, s = np.cos(0.3), np.sin(0.3)
    kx, ky, kz = kloc
    rot = (c * kx - s * ky, s * kx + c * ky, kz)
    assert analyse_trajectory(rot, t)["method"] == "cartesian"


def test_time_along_lines_breaks_factorisation():
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (8, 8, 2), t_first=1.0)
    t2 = t + np.arange(8)[None, :, None] * 5.0       # each line a shot later
    assert analyse_trajectory(kloc, t2)["method"] == "direct"
    with pytest.raises(FeelmriHipError, match="does not factorise"):
        analyse_trajectory(kloc, t2, method="cartesian")


def test_non_separable_k_is_direct():
    rng = np.random.default_rng(1)
    kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (8, 6, 2), t_first=1.0)
    kx = kloc[0] + rng.normal(0, 1e-3, kloc[0].shape)    # jitter breaks the additive split
    assert analyse_trajectory((kx, kloc[1], kloc[2]), t)["method"] == "direct"


def test_radial_runs_and_slots():
    kloc, t = trajectories.radial(0.24, 100, 7, t_first=0.3)
    a = analyse_trajectory(kloc, t)
    assert a["method"] == "direct"
    assert a["n_tgroups"] == 100                     # one group per readout index
    assert a["n_runs"] == 7 * 7                      # ceil(100 / 16) runs per spoke
    assert a["n_tiles_grouped"] == 50                # 100 slots of <= 256 samples, two per tile


def test_spiral_and_random_have_no_runs():
    kloc, t = trajectories.spiral(0.24, 64, 200, 8, t_first=0.2)
    assert analyse_trajectory(kloc, t)["n_runs"] == 0
    rng = np.random.default_rng(5)
    kl = tuple(rng.uniform(-200, 200, (30, 7, 2)) for _ in range(3))
    a = analyse_trajectory(kl, rng.uniform(0.0, 4.0, (30, 7, 2)))
    assert a["n_runs"] == 0 and a["method"] == "direct" and a["n_tgroups"] == 30 * 7 * 2


def test_groups_merge_signed_zero_and_bad_method():
    kloc = tuple(np.zeros((2, 1, 1)) for _ in range(3))
    t = np.array([[[0.0]], [[-0.0]]])
    assert analyse_trajectory(kloc, t, method="direct")["n_tgroups"] == 1
    with pytest.raises(KeyError):
        analyse_trajectory(kloc, t, method="nufft")


def test_runs_exact_for_fp64():
    """The FP64 form takes the readout runs only when they are linear to
    1e-14 of the largest |k'| and |t|; the mixed form accepts 1e-12."""
    kloc, t = trajectories.radial(0.24, 100, 7, t_first=0.3)
    a = analyse_trajectory(kloc, t)
    assert a["n_runs"] > 0 and a["runs_exact"]
    kc, tc = trajectories.cartesian((0.24, 0.24, 0.24), (40, 6, 2), t_first=1.0)
    assert analyse_trajectory(kc, tc)["runs_exact"]
    # a departure of ~1e-13 of the largest |k| keeps the runs but not for FP64
    kx = np.asarray(kloc[0]).copy()
    kmax = max(np.abs(np.asarray(k)).max() for k in kloc)
    kx[5, 2, 0] += 1e-13 * kmax
    b = analyse_trajectory((kx, kloc[1], kloc[2]), t)
    assert b["n_runs"] == a["n_runs"] and not b["runs_exact"]
    # ~1e-11 breaks the runs altogether
    kx[5, 2, 0] += 1e-11 * kmax
    assert analyse_trajectory((kx, kloc[1], kloc[2]), t)["n_runs"] == 0


def test_packed_t_groups():
    """Equal-size t-groups of 32, 64 or 128 samples pack 8, 4 or 2 to a slot
    (spiral interleaves share t per readout index); other sizes do not."""
    for n_il, gp in ((32, 8), (64, 4), (128, 2)):
        kloc, t = trajectories.spiral(0.24, 64, 40, n_il, t_first=0.2)
        a = analyse_trajectory(kloc, t)
        assert a["n_runs"] == 0 and a["n_tgroups"] == 40 and a["group_pack"] == gp, (n_il, a)
    for n_il in (16, 48, 256):
        kloc, t = trajectories.spiral(0.24, 64, 40, n_il, t_first=0.2)
        assert analyse_trajectory(kloc, t)["group_pack"] == 1
    # one group of a different size breaks the packing
    kloc, t = trajectories.spiral(0.24, 64, 40, 32, t_first=0.2)
    t = np.asarray(t).copy()
    t[3, 0, 0] = 99.0
    assert analyse_trajectory(kloc, t)["group_pack"] == 1


def _rotated_interleaves(angles, n_samples=90, res=64, fov=0.24, t_first=0.2):
    """One spiral interleave rotated by each of ``angles``."""
    kloc, t = trajectories.spiral(fov, res, n_samples, 1, t_first=t_first)
    base = kloc[0][:, 0, 0] + 1j * kloc[1][:, 0, 0]
    K = base[:, None] * np.exp(1j * np.asarray(angles))[None, :]
    kx, ky = np.ascontiguousarray(K.real[..., None]), np.ascontiguousarray(K.imag[..., None])
    tt = np.ascontiguousarray(np.broadcast_to(t[:, :1, :], kx.shape))
    return (kx, ky, np.zeros_like(kx)), tt


@pytest.mark.parametrize("n_il", [32, 64, 128])
def test_conjugate_pairs_of_spiral_interleaves(n_il):
    """Interleaves half a turn apart have k and -k at equal t: every packed
    group pairs up, 512 / n_il groups per conjugate-pair tile."""
    kloc, t = trajectories.spiral(0.24, 64, 90, n_il, t_first=0.2)
    a = analyse_trajectory(kloc, t)
    assert a["group_pack"] == 256 // n_il and a["conj_pack"] == 512 // n_il


def test_no_conjugate_pairs_without_antipodal_interleaves():
    """32 interleaves at irregular angles pack (equal groups) but do not pair."""
    ang = 2 * np.pi * np.sort(np.random.default_rng(3).uniform(size=32))
    kloc, t = _rotated_interleaves(ang)
    a = analyse_trajectory(kloc, t)
    assert a["group_pack"] == 8 and a["conj_pack"] == 1
    # pairs present but one interleave nudged off by more than the 1e-12 tolerance
    ang = 2 * np.pi * np.arange(32) / 32
    kloc, t = _rotated_interleaves(ang)
    assert analyse_trajectory(kloc, t)["conj_pack"] == 16
    kx = kloc[0].copy()
    kx[40, 5, 0] *= 1.0 + 1e-9
    assert analyse_trajectory((kx, kloc[1], kloc[2]), t)["conj_pack"] == 1
