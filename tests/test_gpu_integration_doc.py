"""The ctypes stub that INTEGRATION.md shows a maintainer runs as written:
its code block is executed against the built library and its result is
checked against the FP64 oracle, in both precision forms."""
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


def _stub_namespace():
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```python\n(.*?)```", text, flags=re.S)
    code = next(b for b in blocks if "def signal_lumped" in b)
    code = code.replace('"feelmri_amd/_lib/libfeelmri_hip.so"',
                        repr(str(ROOT / "feelmri_amd" / "_lib" / "libfeelmri_hip.so")))
    ns = {}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    return ns


def test_documented_ctypes_stub(gpu):
    import oracle
    from feelmri_amd import trajectories

    ns = _stub_namespace()
    rng = np.random.default_rng(17)
    n, v = 2000, 2
    r0 = rng.uniform(-0.06, 0.06, (n, 3))
    phi = rng.normal(0, 0.1, n)
    T2 = rng.uniform(40, 100, n)
    mxy = (rng.normal(size=(n, v)) + 1j * rng.normal(size=(n, v))) * 0.1
    w = rng.uniform(0.5, 1.5, n) * 1e-6
    for traj in ("cartesian", "radial"):
        if traj == "cartesian":
            kloc, t = trajectories.cartesian((0.24, 0.24, 0.24), (16, 8, 2), t_first=0.4)
        else:
            kloc, t = trajectories.radial(0.24, 48, 12, t_first=0.3)
        ref = oracle.Signal(0, sp.diags(w), kloc, t, r0, phi, T2, mxy)
        for prec, tol in ((0, 1e-10), (1, 1e-4)):
            got = ns["signal_lumped"](*(np.asarray(k) for k in kloc), np.asarray(t), r0, phi, T2, mxy, w,
                                      precision=prec)
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(traj, prec, f"{err:.2e}")
            assert got.shape == ref.shape and err <= tol
