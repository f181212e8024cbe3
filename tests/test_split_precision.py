"""Why the mixed Cartesian GEMMs run three split products per real product:
a numpy emulation of their operand rounding on a Cartesian k-space sum
(bf16 hi + lo = bf16(x) + bf16(x - bf16(x)), exact accumulation).  Three
products (hi*hi + hi*lo + lo*hi) meet the 1e-4 bar with room; dropping
either cross product leaves one operand rounded once, and the k-space error
inherits that rounding -- above the bar for bf16 and also for fp16."""
import numpy as np
import pytest


def _bf16(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)


def _split(x, kind):
    x = np.asarray(x, dtype=np.float32)
    rnd = _bf16 if kind == "bf16" else (lambda v: v.astype(np.float16).astype(np.float32))
    h = rnd(x)
    return h.astype(np.float64), rnd(x - h).astype(np.float64)


def _problem(n=20000, L=32, C=32, seed=0):
    rng = np.random.default_rng(seed)
    y, z, x = rng.uniform(-0.1, 0.1, n), rng.uniform(-0.03, 0.03, n), rng.uniform(-0.1, 0.1, n)
    ky = (np.arange(L) - L // 2) / 0.24
    kz = (np.arange(C) - C // 2) / 0.06
    m = rng.uniform(0.4, 1.5, n) * np.exp(1j * rng.normal(0, 0.3, n))
    A = np.exp(-2j * np.pi * ky[:, None] * y[None, :])
    B = np.exp(-2j * np.pi * kz[None, :] * z[:, None]) * (np.exp(-2j * np.pi * 7 / 0.24 * x) * m)[:, None]
    return A, B


@pytest.mark.parametrize("kind", ["bf16", "fp16"])
def test_three_products_meet_the_bar_two_do_not(kind):
    A, B = _problem()
    ref = A @ B

    def parts(M):
        (rh, rl), (ih, il) = _split(M.real, kind), _split(M.imag, kind)
        return rh + 1j * ih, rl + 1j * il

    (Ah, Al), (Bh, Bl) = parts(A), parts(B)
    rel = lambda S: np.linalg.norm(S - ref) / np.linalg.norm(ref)  # noqa: E731
    three = rel(Ah @ Bh + Ah @ Bl + Al @ Bh)
    two_a = rel(Ah @ Bh + Ah @ Bl)          # A rounded once
    two_b = rel(Ah @ Bh + Al @ Bh)          # B rounded once
    assert three < 1e-5
    assert two_a > 1e-4 and two_b > 1e-4
