"""Write the golden fixtures of this directory: seeded inputs and the FP64
oracle's outputs for them (small: a few hundred particles).

    python tests/golden/make_golden.py

The reference (Julia) ships no golden vectors for this path, so these hold
the oracle's own results.  They freeze the seeded input generators and the
oracle (a change to either shows up as a fixture mismatch in the CPU tests),
and the GPU tests compare the HIP kernels with them directly.

plummer256.npz   Plummer sphere N = 256 (seed 1), G = 1, h = 0.01:
                 pos, vel, mass, acc, pot (compute_force / compute_potential,
                 src/directsum/gravity.jl:80-169, potential.jl:142-148),
                 K, W (src/base/Energy.jl), and pos20, vel20, acc20 after
                 run(sim) with 20 KDK steps at dt = 1e-3 (src/run.jl:46-184).
planetary73.npz  one planetary system (seed 6): a star, 8 planets, 64 test
                 particles, eps = 0 (self pair skipped): pos, vel, mass, acc,
                 pot.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))


def main():
    import __graft_entry__ as ge
    from oracle import oracle as orc
    pkg = ge.load_package()
    pos, vel, mass = pkg.ics.plummer(256, seed=1)
    soft = [0.01] * 6
    acc = orc.compute_force(pos, mass, soft, 1.0)
    pot = orc.compute_potential(pos, mass, soft, 1.0)
    K, W, _, _ = orc.energy(pos, vel, mass, pot)
    x20, v20, a20, info = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.02,
                                            dt=1e-3)
    assert info["steps"] == 20
    np.savez(HERE / "plummer256.npz", pos=pos, vel=vel, mass=mass, h=0.01, G=1.0, acc=acc,
             pot=pot, K=K, W=W, pos20=x20, vel20=v20, acc20=a20, dt=1e-3, steps=20)
    ens = pkg.ics.planetary_ensemble(nsys=1, seed=6)
    p, v, m = ens["pos"], ens["vel"], ens["mass"]
    a = orc.compute_force(p, m, [0.0] * 6, 1.0, skip_self=True)
    q = orc.compute_potential(p, m, [0.0] * 6, 1.0, skip_self=True)
    np.savez(HERE / "planetary73.npz", pos=p, vel=v, mass=m, acc=a, pot=q)
    print("wrote", HERE / "plummer256.npz", HERE / "planetary73.npz")


if __name__ == "__main__":
    main()
