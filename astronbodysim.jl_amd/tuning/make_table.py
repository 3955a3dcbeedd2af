"""Regenerate the launch-configuration table that the package loads at import
(tuning/mi355x.json): autotune one seeded system per size and precision mode
and save what the autotuner measured.

    python astronbodysim.jl_amd/tuning/make_table.py [out.json] [--modes=split,fp32,fp64]

(the package directory's name has a dot, so it does not run with -m).  Sizes 2^12 .. 2^20 in the split mode on
two-Plummer collisions (the benchmark's geometry), 2^12 .. 2^18 in the fp32
mode and 2^12 .. 2^15 in the fp64 mode on Plummer spheres.  With --modes= only those modes are
measured again and the table's other entries are kept.  Needs the GPU; a
few minutes.
"""
from __future__ import annotations

import sys
import time
from pathlib import Path


def _package():
    here = Path(__file__).resolve().parents[2]
    if str(here) not in sys.path:
        sys.path.insert(0, str(here))
    import __graft_entry__ as ge
    return ge.load_package()


def main(out=None, modes=None):
    pkg = _package()
    nat = pkg._native
    out = out or str(nat.DEFAULT_TUNING)
    nat.tuner_clear()
    if modes:    # keep the entries of the modes not measured again
        nat.tuner_load(nat.DEFAULT_TUNING)
    plan = [("split", k) for k in range(12, 21)] + [("fp32", k) for k in range(12, 19)] + \
           [("fp64", k) for k in range(12, 16)]
    plan = [(m, k) for m, k in plan if not modes or m in modes]
    for mode, k in plan:
        n = 1 << k
        if mode == "split":
            pos, vel, mass = pkg.ics.collision(n // 2)
            h = 0.002
        else:
            pos, vel, mass = pkg.ics.plummer(n, seed=2)
            h = 0.005
        ctx = nat.Context(n, nat.MODES[mode])
        ctx.set_params(1.0, [h] * 6)
        ctx.upload(pos, vel, mass)
        t0 = time.time()
        ctx.autotune()
        print(f"{mode} N=2^{k}: {ctx.get_tuning()} ({time.time() - t0:.1f} s)", flush=True)
        ctx.close()
    nat.tuner_save(out)
    print("saved", out, flush=True)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--modes")]
    sel = [a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--modes=")]
    main(args[0] if args else None, sel[0].split(",") if sel else None)
