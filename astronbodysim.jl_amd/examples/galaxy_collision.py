"""The reference's galaxy-collision example on this library.

examples/04-collisions/04-collisions.jl:22-79 reads a Gadget-1 file of two
disk galaxies (40000 halo + 20000 disk particles), converts it to astro units
(kpc, Msun, Gyr) and runs it on the GPU in FP32 with adaptive (block)
timesteps, a 0.4 kpc softening, snapshots every 0.03 Gyr up to 3 Gyr:

    python astronbodysim.jl_amd/examples/galaxy_collision.py galaxy_littleendian.dat \\
        [--time-end 3.0] [--between-snapshots 0.03] [--precision fp32]

It prints the number of steps, the active-particle pairs evaluated, the wall
time and the relative energy change.

Unit conversion: the file is in Gadget-2's internal units (kpc, 1e10 Msun,
km/s); velocities become kpc/Gyr with 1 kpc = 3.0856775814913673e16 km and
1 Gyr = 3.15576e16 s (Julian years), masses Msun, and G = 6.6743e-11
m^3 kg^-1 s^-2 with Msun = 1.98847e30 kg.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

KPC_KM = 3.0856775814913673e16
GYR_S = 3.15576e16
KMS_TO_KPC_PER_GYR = GYR_S / KPC_KM
MSUN_KG = 1.98847e30
G_ASTRO = 6.6743e-11 * MSUN_KG * GYR_S ** 2 / (KPC_KM * 1e3) ** 3   # kpc^3 Msun^-1 Gyr^-2


def _package():
    root = Path(__file__).resolve().parents[2]
    if str(root) not in sys.path:
        sys.path.insert(0, str(root))
    import __graft_entry__ as ge
    return ge.load_package()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("ic", help="Gadget-1/2 initial conditions (galaxy_littleendian.dat)")
    ap.add_argument("--time-end", type=float, default=3.0, help="Gyr")
    ap.add_argument("--between-snapshots", type=float, default=0.03, help="Gyr")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "split", "fp64"])
    ap.add_argument("--output", default="output/Collision-DirectSumAdaptiveGPU")
    args = ap.parse_args(argv)
    nb = _package()
    header, blk = nb.read_gadget2(args.ic)
    data = nb.StarData(blk["POS"], Vel=blk["VEL"] * KMS_TO_KPC_PER_GYR, Mass=blk["Mass"] * 1e10,
                       Collection=blk["Collection"], ID=blk["ID"])
    sim = nb.Simulation(data, precision=args.precision, G=G_ASTRO, TimeEnd=args.time_end,
                        TimeBetweenSnapshots=args.between_snapshots, TimeStep=0.0,
                        ForceSofteningTable=[0.4] * 6, OutputDir=args.output)
    n = len(data)
    e0 = nb.total_energy(sim)
    t0 = time.perf_counter()
    nb.run(sim)
    wall = time.perf_counter() - t0
    e1 = nb.total_energy(sim)
    hist = sim._adaptive.nactive_history
    pairs = float(np.sum(np.asarray(hist, dtype=np.float64))) * n
    print(f"N = {n}, {header.npart[1]} halo + {header.npart[2]} disk; t = {sim.timeinfo.system_time_float:.4f} Gyr")
    print(f"{sim.timeinfo.stepcount} steps ({len(hist)} block steps), {pairs:.4g} active pairs, "
          f"{wall:.2f} s wall, {pairs / wall:.4g} pairs/s")
    print(f"snapshots: {sim.outputinfo['snapshotcount']}, relative energy change {abs(e1 - e0) / abs(e0):.3e}")
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
