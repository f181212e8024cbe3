"""Host side of the device event-table builder: the candidate bound of
fm_events_bound against the grid parts the host rule concatenates
(python/feelmri/Bloch.py:105-129), the per-event chain marks, and the block
records handed to the device."""
import ctypes

import numpy as np
import pytest

from feelmri_amd import device, sequences
from feelmri_amd.sequences import HardPulse, Trapezoid, make_block


def _blocks():
    rng = np.random.default_rng(0)
    out = sequences.spoiled_gre(flip_deg=20, TR=4.3, n_dummy=2, n_lines=2, rf_dur=0.25, phase_cycle=True)
    g = [Trapezoid(0, 1.1, 12.5, 0.13, 0.77), Trapezoid(1, 1.3, -7.0, 0.0, 0.4), Trapezoid(2, 0.9, 3.3, 0.21, 0.1),
         Trapezoid(0, 2.05, -4.0, 0.11, 0.3)]
    out.append(make_block(0.7, 2.9, pulses=[HardPulse(1.0, 0.3, 0.2, 0.4), HardPulse(1.2, 0.35, 0.1)],
                          gradients=g, dt_rf=0.01, dt_gr=0.05, dt=0.1, store=True))
    out.append(make_block(3.1, 0.0, gradients=[Trapezoid(1, 3.1, float(rng.normal()), 0.07, 0.2)], dt=0.3))
    return out


def _candidates(spec):
    """Lengths of the parts make_block concatenates, before de-duplication."""
    n = 0
    for g in spec["gradients"]:
        n += 4 + (len(np.arange(g.timings[0], g.timings[-1], spec["dt_gr"])) if spec["dt_gr"] > 0 else 0)
    for p in spec["pulses"]:
        n += 2 + (len(np.arange(p.start, p.start + p.dur, spec["dt_rf"])) if spec["dt_rf"] > 0 else 0)
    ext0 = min([spec["t0"]] + [p.start for p in spec["pulses"]] + [g.timings[0] for g in spec["gradients"]])
    ext1 = max([spec["t0"] + spec["dur"]] + [p.start + p.dur for p in spec["pulses"]] +
               [g.timings[-1] for g in spec["gradients"]])
    return n + len(np.arange(ext0, ext1, spec["dt"]))


def test_events_bound_counts_the_host_rules_candidates():
    from feelmri_amd import _lib

    blocks = _blocks()
    desc, pulses, traps, _ = sequences.block_descriptions(blocks)
    mx, tot = ctypes.c_int64(), ctypes.c_int64()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    _lib.call("fm_events_bound", desc.size, vp(desc), vp(pulses), 0, vp(traps), ctypes.byref(mx), ctypes.byref(tot))
    want = [_candidates(b.info["spec"]) for b in blocks]
    assert mx.value == max(want) and tot.value == sum(want)
    assert all(b.n_events <= w for b, w in zip(blocks, want))
    # a block with too many parts, and one with no event, are refused
    bad = desc[:1].copy()
    bad["n_pulses"] = _lib.EVENTS_MAX_PARTS + 1
    with pytest.raises(RuntimeError, match="at most"):
        _lib.call("fm_events_bound", 1, vp(bad), vp(pulses), 0, vp(traps), ctypes.byref(mx), ctypes.byref(tot))
    empty = np.zeros(1, dtype=device.BLOCK_DESC)
    empty["dt"], empty["store_col"] = 1.0, -1
    with pytest.raises(RuntimeError, match="no event"):
        _lib.call("fm_events_bound", 1, vp(empty), None, 0, None, ctypes.byref(mx), ctypes.byref(tot))
    # a pulse table outside the table points is refused
    tp = pulses[:1].copy()
    tp["n_table"], tp["table0"] = 200, 10
    one = desc[2:3].copy()
    one["pulse0"], one["n_pulses"], one["n_traps"] = 0, 1, 0
    with pytest.raises(RuntimeError, match="table"):
        _lib.call("fm_events_bound", 1, vp(one), vp(tp), 205, None, ctypes.byref(mx), ctypes.byref(tot))
    _lib.call("fm_events_bound", 1, vp(one), vp(tp), 210, None, ctypes.byref(mx), ctypes.byref(tot))


def test_block_records():
    blocks = _blocks()
    desc, pulses, traps, _ = sequences.block_descriptions(blocks)
    assert desc.dtype.itemsize == 64 and pulses.dtype.itemsize == 48 and traps.dtype.itemsize == 40
    assert list(desc["store_col"]) == [-1, -1, -1, -1, 0, -1, 1, -1, 2, -1]
    assert list(desc["spoil"]) == [0 if b.empty else 1 for b in blocks]
    assert desc["n_pulses"].sum() == pulses.size and desc["n_traps"].sum() == traps.size
    # the pulse amplitude is the value HardPulse.__call__ forms
    p = blocks[2].info["spec"]["pulses"][0]
    on = p.start + 0.5 * p.dur
    q = pulses[desc[2]["pulse0"]]
    assert complex(q["amp_re"], q["amp_im"]) == complex(p(np.array([on]))[0]) and q["n_table"] == 0


def test_block_records_with_tabulated_pulses():
    """RFPulse tables go into the table array in block and pulse order;
    their records point at them."""
    s1 = sequences.RFPulse(start=0.2, dur=1.0, flip=0.3, n_lobes=(1, 1), nb_samples=50)
    s2 = sequences.RFPulse(start=2.1, dur=0.8, flip=0.5, shape="sinc", phase=0.4, nb_samples=80)
    h = HardPulse(1.5, 0.2, 0.2)
    blocks = [make_block(0.0, 1.5, pulses=[s1, h], dt_rf=0.05, dt=0.5), make_block(2.0, 1.0, pulses=[s2], dt=0.5)]
    desc, pulses, traps, tab = sequences.block_descriptions(blocks)
    assert pulses["n_table"].tolist() == [50, 0, 80] and pulses["table0"].tolist() == [0, 0, 50]
    assert tab.shape == (130, 2)
    np.testing.assert_array_equal(tab[:50, 0] + 1j * tab[:50, 1], s1(np.linspace(s1.start, s1.start + s1.dur, 50)))
    np.testing.assert_array_equal(tab[50:, 0] + 1j * tab[50:, 1], s2(np.linspace(s2.start, s2.start + s2.dur, 80)))


def test_marks_from_blocks():
    m = device.marks_from_blocks([3, 1, 2], [1, 2 | 1, (2 + 1) << 1])
    assert m.tolist() == [0, 0, 0, 2, 4, 0, 7]
