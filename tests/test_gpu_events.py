"""Device event tables and marked block chains on the GPU.

* fm_events_build's tables equal, bit for bit, the host rule's
  (sequences.make_block, restating python/feelmri/Bloch.py:105-129, :44-49,
  :340): times, RF (hard pulses and tabulated sinc pulses read as numpy's
  interp), gradients, dt, block offsets and chain marks.
* A chain over device-built tables equals the chain over host tables and the
  per-block launches, bit for bit (static nodes, both precisions).
* Moving nodes chain in one launch, equal bit for bit to one launch per
  block with that block's motion coefficients, from host or device tables.
* A captured graph that rebuilds the tables and integrates the chain equals
  the eager solve on every replay.
"""
import numpy as np
import pytest

import motion_fixtures
from feelmri_amd import sequences
from feelmri_amd.sequences import HardPulse, RFPulse, Trapezoid, make_block

pytestmark = pytest.mark.gpu


def _mixed_blocks(seed=0):
    rng = np.random.default_rng(seed)
    out = sequences.spoiled_gre(flip_deg=20, TR=4.3, n_dummy=3, n_lines=2, rf_dur=0.25, phase_cycle=True)
    t = out[-1].info["extent"][1]
    g = [Trapezoid(0, t + 0.1, 12.5, 0.13, 0.77), Trapezoid(1, t + 0.3, -7.0, 0.0, 0.4),
         Trapezoid(2, t + 0.9, 3.3, 0.21, 0.1), Trapezoid(0, t + 2.05, -4.0, 0.11, 0.3)]
    out.append(make_block(t, 2.9, pulses=[HardPulse(t + 0.3, 0.3, 0.2, 0.4), HardPulse(t + 0.5, 0.35, 0.1)],
                          gradients=g, dt_rf=0.01, dt_gr=0.05, dt=0.1, store=True))
    t = out[-1].info["extent"][1]
    # tabulated pulses: an apodized sinc beside a slice gradient, a plain sinc
    # with phase and frequency offsets, read between and on their knots
    out.append(make_block(t, 1.6, pulses=[RFPulse(t + 0.2, 1.2, 0.4, n_lobes=(2, 2)),
                                          RFPulse(t + 0.5, 0.6, 0.1, shape="sinc", n_lobes=(1, 3), phase=0.3,
                                                  freq_offset=250.0, nb_samples=73)],
                          gradients=[Trapezoid(2, t + 0.1, 6.0, 0.1, 1.2)], dt_rf=0.01, dt_gr=0.04, dt=0.2,
                          store=True))
    t = out[-1].info["extent"][1]
    out.append(make_block(t, 0.0, gradients=[Trapezoid(1, t, float(rng.normal(0, 5)), 0.07, 0.2)], dt=0.03))
    out.append(sequences.delay_block(out[-1].info["extent"][1], 1.7, dt=0.45))
    return out


def _nodes(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-0.05, 0.05, (n, 3)), rng.uniform(500, 1500, n), rng.uniform(40, 100, n),
            rng.normal(0, 1e-4, n))


def test_device_tables_equal_host_tables(gpu):
    from feelmri_amd import device

    blocks = _mixed_blocks()
    de = device.DeviceEvents(*sequences.block_descriptions(blocks), device=gpu).build()
    de.check()
    times, rf, G, dt, offsets, marks = de.tables()
    lens = [b.n_events for b in blocks]
    assert offsets.tolist() == np.concatenate([[0], np.cumsum(lens)]).tolist()
    assert np.array_equal(times, np.concatenate([b.times for b in blocks]))
    assert np.array_equal(dt, np.concatenate([b.dt for b in blocks]))
    assert np.array_equal(rf, np.concatenate([b.rf for b in blocks]))
    assert np.array_equal(G, np.concatenate([b.G for b in blocks]))
    desc = sequences.block_descriptions(blocks)[0]
    assert (sequences.block_descriptions(blocks)[1]["n_table"] > 0).sum() == 2
    codes = desc["spoil"] | ((desc["store_col"] + 1) << 1)
    assert np.array_equal(marks, device.marks_from_blocks(lens, codes))
    # rebuilt in place: the same tables
    de.build()
    assert np.array_equal(de.tables()[0], times)


def test_builder_flags_overflow(gpu):
    """A capacity below the event count is reported (status 2) and nothing is
    written past it."""
    import torch

    from feelmri_amd import device

    blocks = _mixed_blocks()
    de = device.DeviceEvents(*sequences.block_descriptions(blocks), device=gpu)
    n = sum(b.n_events for b in blocks)
    de.capacity = n - 5
    sentinel = de.times[n - 5:].clone().fill_(123.0)
    de.times[n - 5:] = sentinel
    de.build()
    torch.cuda.synchronize()
    assert int(de.status.item()) == 2
    with pytest.raises(RuntimeError, match="capacity"):
        de.check()
    assert torch.equal(de.times[n - 5:], sentinel)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_chain_on_device_tables_equals_host_chain(gpu, prec):
    import torch

    from feelmri_amd.Bloch import BlochSolver

    n = 2500
    r0, T1, T2, dB = _nodes(n, 3)
    blocks = _mixed_blocks(1)
    a = BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, dev=gpu)
    b = BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, dev=gpu)
    c = BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, dev=gpu)
    for _ in range(2):   # the second solve continues from the carried state
        xa, za = a.solve_device(events="device")
        xb, zb = b.solve_device()
        xc, zc = c.solve_device(chain=False)
        torch.cuda.synchronize()
        assert xa.shape == (n, sum(bl.store for bl in blocks))
        assert torch.equal(xa, xb) and torch.equal(za, zb)
        assert torch.equal(xa, xc) and torch.equal(za, zc)
        assert torch.equal(a.mxy, c.mxy) and torch.equal(a.mz, c.mz)


@pytest.mark.parametrize("kind", ["pod", "translation", "velocity"])
def test_moving_chain_equals_block_launches(gpu, kind):
    """Moving nodes: the marked chain (one launch, coefficients of every
    block in one table) equals the per-block launches bit for bit, from host
    and from device tables."""
    import torch

    from feelmri_amd.Bloch import BlochSolver

    n = 2000
    r0, T1, T2, dB = _nodes(n, 5)
    blocks = _mixed_blocks(2)
    if kind == "pod":
        motion = motion_fixtures.pod(n, seed=1)
    elif kind == "translation":
        motion = motion_fixtures.resp()
    else:
        motion = motion_fixtures.pod(n, seed=3, velocity=True)
    for prec in ("f64", "f32"):
        s = [BlochSolver(blocks, r0, T1, T2, delta_B=dB, precision=prec, pod_trajectory=motion, dev=gpu)
             for _ in range(3)]
        for _ in range(2):
            xa, za = s[0].solve_device(chain=True)
            xb, zb = s[1].solve_device(chain=False)
            xc, zc = s[2].solve_device(events="device")
            torch.cuda.synchronize()
            assert torch.equal(xa, xb) and torch.equal(za, zb), (kind, prec)
            assert torch.equal(xc, xb) and torch.equal(zc, zb), (kind, prec)
            assert torch.equal(s[0].mxy, s[1].mxy) and torch.equal(s[2].mz, s[1].mz)


@pytest.mark.parametrize("moving", [False, True])
def test_graph_rebuilds_tables_and_chain(gpu, moving):
    import torch

    from feelmri_amd.Bloch import BlochSolver

    n = 3000
    r0, T1, T2, dB = _nodes(n, 7)
    blocks = _mixed_blocks(3)
    motion = motion_fixtures.pod(n, seed=4) if moving else None
    solver = BlochSolver(blocks, r0, T1, T2, delta_B=dB, pod_trajectory=motion, dev=gpu)
    mxy0, mz0 = solver.mxy.clone(), solver.mz.clone()
    want_xy, want_z = (t.clone() for t in solver.solve_device())
    carried = (solver.mxy.clone(), solver.mz.clone())
    graph = solver.capture(events="device")
    de = next(iter(solver._dev_events.values()))
    for _ in range(3):
        solver.mxy.copy_(mxy0)
        solver.mz.copy_(mz0)
        de.times.fill_(-1.0)      # the replay rebuilds the tables
        got_xy, got_z = graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(got_xy, want_xy) and torch.equal(got_z, want_z)
        assert torch.equal(solver.mxy, carried[0]) and torch.equal(solver.mz, carried[1])
        assert float(de.times[0]) == blocks[0].times[0]
        assert float((torch.arange(4096, device=gpu, dtype=torch.float64) * 2.0).sum()) == 4095 * 4096
