"""The HIP kernels against the committed golden fixtures (tests/golden/):
FP64 mode to rounding, the FP32-pair modes to the benchmark's gate (median
relative error <= 1e-6), 20 KDK steps, and a planetary system with its
star's terms in FP64 (needs an MI355X)."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1) if a.ndim == 2 else np.abs(a - b) / np.abs(b)


@pytest.mark.parametrize("mode,med,mx", [("fp64", 1e-13, 1e-12), ("split", 1e-6, 1e-4), ("fp32", 2e-6, 1e-4)])
def test_plummer256_force_potential_and_steps(pkg, mode, med, mx):
    nat = pkg._native
    g = load("plummer256.npz")
    with nat.Context(256, nat.MODES[mode]) as c:
        c.set_params(float(g["G"]), [float(g["h"])] * 6)
        c.upload(g["pos"], g["vel"], g["mass"])
        c.compute_force()
        c.compute_potential(self_term=False)
        d = c.download(pos=False, vel=False, acc=True, pot=True)
        c.half_kick(+1, float(g["dt"]))
        c.kdk_steps(int(g["steps"]), float(g["dt"]), 0.0, use_graph=True)
        c.half_kick(-1, float(g["dt"]))
        s = c.download(pos=True, vel=True, acc=False)
    ea, ep = rel(d["acc"], g["acc"]), rel(d["pot"], g["pot"])
    # the FP32 state rounds positions to FP32 at every drift: relative to |x|
    dx = np.max(np.abs(s["pos"] - g["pos20"]) / np.maximum(1.0, np.abs(g["pos20"])))
    print(f"{mode}: acc median {np.median(ea):.3e} max {ea.max():.3e}; pot median {np.median(ep):.3e}; "
          f"20 steps max |dx| {dx:.3e}")
    assert np.median(ea) <= med and ea.max() <= mx
    assert np.median(ep) <= med and ep.max() <= mx
    assert dx <= {"fp64": 1e-13, "split": 1e-8, "fp32": 3e-6}[mode]


def test_planetary73_central_star(pkg):
    nat = pkg._native
    p = load("planetary73.npz")
    with nat.Context(73, nat.NB_MODE_SPLIT_CENTRAL) as c:
        c.set_params(1.0, [0.0] * 6)
        c.upload(p["pos"], p["vel"], p["mass"])
        c.set_central(1)
        c.compute_force()
        c.compute_potential(self_term=False)
        d = c.download(pos=False, vel=False, acc=True, pot=True)
    ea, ep = rel(d["acc"], p["acc"]), rel(d["pot"], p["pot"])
    print(f"planetary: acc median {np.median(ea):.3e} max {ea.max():.3e}; pot max {ep.max():.3e}")
    assert np.median(ea) <= 1e-9 and ea.max() <= 1e-6 and ep.max() <= 1e-6
