"""Many planetary systems in one resident launch (the Ensemble API).

2048 systems, each a central mass 1 with 8 planets (a log-uniform in
[0.4, 30], e < 0.1, i <= 2 deg, m/M log-uniform in [1e-7, 1e-3]) and 64
massless test particles on circular orbits (r in [1, 40]), eps = 0, G = 1:
the star's pair terms in FP64, the planets' in FP32 split precision, KDK
steps at dt = 1e-3 in the resident kernel (the whole run in one launch;
four systems per workgroup, their 36 massive bodies in one wavefront and
each system's 64 test particles in another), with every system's total
energy tracked.  The reference runs one Simulation per system.

    python astronbodysim.jl_amd/examples/planetary_ensemble.py [--nsys 2048] [--steps 10000]
        [--every 100] [--dt 1e-3]

It prints the time per KDK step (all systems), the pair interactions per
second (every body against the 9 massive ones, and all pairs), and the
distribution of the relative energy errors.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np


def _package():
    root = Path(__file__).resolve().parents[2]
    if str(root) not in sys.path:
        sys.path.insert(0, str(root))
    import __graft_entry__ as ge
    return ge.load_package()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--nsys", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=6)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--dt", type=float, default=1e-3)
    args = ap.parse_args(argv)
    nb = _package()
    ens = nb.ics.planetary_ensemble(nsys=args.nsys, seed=args.seed)
    off, per = ens["offsets"], ens["per_system"]
    with nb.Ensemble(off, "split_central") as e:
        e.set_params(1.0, 0.0, args.dt)
        e.upload(ens["pos"], ens["vel"], ens["mass"])
        e.set_central(1)
        e.compute_force()
        en = e.energy()
        E0 = en["K"] + en["W"]
        e.half_kick(+1)
        e.sync()
        t0 = time.perf_counter()
        E = e.kdk_steps_tracked(args.steps, args.every)
        e.sync()
        t = time.perf_counter() - t0
        tun = e.get_tuning()
    rel = np.abs(E - E0[:, None]) / np.abs(E0[:, None])
    nmass = int((ens["mass"][:per] > 0).sum())
    print(f"{args.nsys} systems x {per} bodies ({nmass} massive), {args.steps} steps; tuning {tun}")
    print(f"KDK step: {t / args.steps * 1e6:.3f} us; {args.nsys * per * nmass * args.steps / t:.4g} "
          f"evaluated pairs/s, {args.nsys * per * per * args.steps / t:.4g} all pairs/s")
    print(f"relative energy error: median {np.median(rel):.3e}, max per system: median "
          f"{np.median(rel.max(1)):.3e}, {int((rel.max(1) > 1e-3).sum())} systems above 1e-3")
    return 0


if __name__ == "__main__":
    sys.exit(main())
