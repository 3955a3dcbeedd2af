"""One Plummer sphere through the one-system drop-in path.

N particles (G = M = a = 1, isotropic velocities from the Plummer
distribution function), softening h, in a chosen precision mode: one force
evaluation timed on its own, then KDK steps (graph-captured) timed as a
whole, with the energy drift.  The defaults are the FP32 case the reference's
floattype = Float32 GPU runs use (N = 65536, h = 0.005, 50 steps at
dt = 5e-4); --precision fp64 --n 4096 --h 0.01 --steps 100 --dt 1e-3 is its
FP64 CPU-runnable case.

    python astronbodysim.jl_amd/examples/plummer.py [--n 65536] [--precision fp32]
        [--h 0.005] [--steps 50] [--dt 5e-4] [--forces 5]
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
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "split"])
    ap.add_argument("--h", type=float, default=0.005)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--forces", type=int, default=5, help="force evaluations to time")
    args = ap.parse_args(argv)
    nb = _package()
    nat = nb._native
    pos, vel, mass = nb.ics.plummer(args.n, seed=args.seed)
    pairs = float(args.n) ** 2
    with nat.Context(args.n, nat.MODES[args.precision]) as ctx:
        ctx.set_params(1.0, [args.h] * 6)
        ctx.upload(pos, vel, mass)
        e0 = ctx.energy()
        ctx.compute_force()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.forces):
            ctx.compute_force()
        ctx.sync()
        t_force = (time.perf_counter() - t0) / max(args.forces, 1)
        ctx.half_kick(+1, args.dt)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.kdk_steps(args.steps, args.dt, 0.0, use_graph=True)
        ctx.sync()
        t_steps = (time.perf_counter() - t0) / max(args.steps, 1)
        ctx.half_kick(-1, args.dt)
        e1 = ctx.energy()
        tun = ctx.get_tuning()
    E0, E1 = e0["K"] + e0["W"], e1["K"] + e1["W"]
    print(f"Plummer N={args.n}, {args.precision}, h={args.h}; tuning {tun}")
    print(f"force: {t_force * 1e3:.3f} ms, {pairs / t_force:.4g} pairs/s")
    print(f"KDK step: {t_steps * 1e3:.3f} ms, {pairs / t_steps:.4g} pairs/s over {args.steps} steps; "
          f"|dE/E| {abs(E1 - E0) / abs(E0):.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
