"""Many independent star clusters in one ragged launch (the Ensemble API).

512 Plummer clusters, N log-uniform in [256, 16384] (about 2e6 particles in
total), scale radius a ~ U[0.5, 2], softening h = 0.01 a, M = 1 each, G = 1:
the force of every cluster in one persistent launch, then KDK steps
captured in HIP graphs.  The reference runs one Simulation per system.

    python astronbodysim.jl_amd/examples/cluster_ensemble.py [--steps 200] [--dt 1e-3]
        [--forces 5] [--autotune]

It prints the time per force evaluation, the time per KDK step and the pair
interactions per second of both (sum over clusters of N^2 per force).
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
    ap.add_argument("--nsys", type=int, default=512)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--forces", type=int, default=5, help="force evaluations to time")
    ap.add_argument("--autotune", action="store_true", help="autotune the launch first")
    args = ap.parse_args(argv)
    nb = _package()
    ens = nb.ics.cluster_ensemble(nsys=args.nsys, seed=args.seed)
    off = ens["offsets"]
    ns = np.diff(off)
    pairs = float((ns.astype(np.float64) ** 2).sum())
    with nb.Ensemble(off, "split") as e:
        e.set_params(1.0, ens["h"], args.dt)
        e.upload(ens["pos"], ens["vel"], ens["mass"])
        if args.autotune:
            e.autotune()
        e.compute_force()
        e.sync()
        t0 = time.perf_counter()
        for _ in range(args.forces):
            e.compute_force()
        e.sync()
        t_force = (time.perf_counter() - t0) / max(args.forces, 1)
        e.half_kick(+1)
        t0 = time.perf_counter()
        e.kdk_steps(args.steps)
        e.sync()
        t_steps = (time.perf_counter() - t0) / max(args.steps, 1)
        e.half_kick(-1)
        tun = e.get_tuning()
    print(f"{args.nsys} clusters, {int(off[-1])} particles, {pairs:.4g} pairs per force; tuning {tun}")
    print(f"force: {t_force * 1e3:.3f} ms, {pairs / t_force:.4g} pairs/s")
    print(f"KDK step: {t_steps * 1e3:.3f} ms, {pairs / t_steps:.4g} pairs/s over {args.steps} steps")
    return 0


if __name__ == "__main__":
    sys.exit(main())
