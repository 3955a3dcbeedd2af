"""The committed golden fixtures (tests/golden/, written by make_golden.py)
against the seeded generators and the FP64 oracle, on the CPU: a change to
either shows up here (allow_pickle=False: the fixtures are plain arrays)."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_generators_reproduce_the_fixture_inputs(pkg):
    g = load("plummer256.npz")
    pos, vel, mass = pkg.ics.plummer(256, seed=1)
    assert np.array_equal(pos, g["pos"]) and np.array_equal(vel, g["vel"])
    assert np.array_equal(mass, g["mass"])
    p = load("planetary73.npz")
    ens = pkg.ics.planetary_ensemble(nsys=1, seed=6)
    assert np.array_equal(ens["pos"], p["pos"]) and np.array_equal(ens["mass"], p["mass"])


def test_oracle_reproduces_the_fixtures(orc):
    g = load("plummer256.npz")
    soft = [float(g["h"])] * 6
    acc = orc.compute_force(g["pos"], g["mass"], soft, float(g["G"]))
    pot = orc.compute_potential(g["pos"], g["mass"], soft, float(g["G"]))
    assert np.allclose(acc, g["acc"], rtol=1e-13, atol=0)
    assert np.allclose(pot, g["pot"], rtol=1e-13, atol=0)
    K, W, _, _ = orc.energy(g["pos"], g["vel"], g["mass"], pot)
    assert K == pytest.approx(float(g["K"]), rel=1e-14) and W == pytest.approx(float(g["W"]), rel=1e-13)
    x, v, a, info = orc.run_const_kdk(g["pos"], g["vel"], g["mass"], soft, 1.0, t_begin=0.0,
                                      t_end=0.02, dt=float(g["dt"]))
    assert info["steps"] == int(g["steps"])
    assert np.allclose(x, g["pos20"], rtol=0, atol=1e-14)
    assert np.allclose(v, g["vel20"], rtol=0, atol=1e-13)
    p = load("planetary73.npz")
    a = orc.compute_force(p["pos"], p["mass"], [0.0] * 6, 1.0, skip_self=True)
    assert np.allclose(a, p["acc"], rtol=1e-13, atol=0)
