"""Benchmark: pair interactions/s of the direct-summation KDK step on one MI355X.

Workload (default): head-on collision of two Plummer spheres, N = 2^20 in
total (524288 each), log-normal masses, h = 0.002, split precision (FP32
pair arithmetic, FP64 state), leapfrog KDK at dt = 1e-3.  One "step" is one
KDK step (drift, all-pairs force, kick): N^2 pair interactions.

Printed: one JSON line with the throughput, the force kernel's roofline
figures (HIP-event timed on the library's stream), the force error of the
initial state against the FP64 oracle on a sample of targets, and the CPU
oracle's throughput on that same sample (host cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pair interactions/s at force rel err ≤1e-6 vs FP64; batched vs plain speedup"
FLOPS_PER_PAIR = 20
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector, published (packed)
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # one FP32 op per lane per cycle


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=1 << 20)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target duration of the CPU-oracle sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-batch", action="store_true")
    return p.parse_args()


def traffic_from_profile():
    f = ROOT / "profiles" / "force_c3_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding of "
              "this path is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as ge
    pkg = ge.load_package()
    nat = pkg._native
    n = args.n
    h, dt = 0.002, 1e-3
    pos, vel, mass = pkg.ics.collision(n // 2, seeds=(3, 4))
    n = pos.shape[0]
    dev = nat.device_info()

    ctx = nat.Context(n, nat.NB_MODE_SPLIT)
    ctx.set_params(1.0, [h] * 6)
    ctx.upload(pos, vel, mass)
    # initial force (also the accuracy sample) and the first half kick
    ctx.compute_force()
    acc0 = ctx.download(pos=False, vel=False, acc=True)["acc"]
    ctx.half_kick(+1, dt)
    t = 0.0
    for _ in range(args.warmup):
        ctx.kdk_steps(1, dt, t)
        t += dt
    ctx.sync()

    ctx.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.kdk_steps(args.steps, dt, t)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    nforce, force_ms = ctx.get_timing()
    ctx.set_timing(False)
    tuning = ctx.get_tuning()
    ctx.close()

    pairs_per_step = float(n) * float(n)
    value = pairs_per_step * args.steps / elapsed
    avg_force_s = force_ms / max(nforce, 1) * 1e-3
    achieved = FLOPS_PER_PAIR * pairs_per_step / avg_force_s / 1e12

    # accuracy + CPU baseline on a seeded sample of targets against all sources
    cpu = None
    err = None
    if not args.no_cpu:
        from oracle import oracle as orc
        threads = orc.num_threads()
        rng = np.random.default_rng(7)
        probe = rng.choice(n, 64, replace=False)
        t1 = time.perf_counter()
        _ = orc.force_at_points(pos[probe], h, pos, mass, variant="simd")
        per_target = (time.perf_counter() - t1) / 64
        k = int(min(n, max(256, args.cpu_seconds / max(per_target, 1e-9))))
        sample = np.sort(rng.choice(n, k, replace=False))
        t1 = time.perf_counter()
        ref = orc.force_at_points(pos[sample], h, pos, mass, variant="simd")
        cpu_s = time.perf_counter() - t1
        # the probe path evaluates the self pair at r = 0 (contributes 0 for h > 0)
        e = orc.rel_err(acc0[sample], ref)
        err = {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
               "max": float(np.max(e)), "targets": int(k)}
        cpu = {"value": float(k) * n / cpu_s, "unit": "pair interactions/s", "cores": threads,
               "kind": "port",
               "sample": f"{k} seeded targets x all {n} sources of the initial state, FP64 "
                         f"C restatement of src/directsum (SIMD variant, OpenMP {threads} threads)"}

    batch = None
    if not args.no_batch:
        try:
            from bench_batch import run_batch_comparison  # noqa: WPS433
            batch = run_batch_comparison(pkg)
        except ImportError:
            batch = None

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Plummer spheres)",
        "config": {"workload": "two-Plummer head-on collision, N=2^20 (2x524288), log-normal "
                               "masses, h=0.002, split precision, KDK dt=1e-3",
                   "n": n, "softening": h, "dt": dt, "precision": "split",
                   "tuning": tuning},
        "force_rel_err": err,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "frac_unpacked": achieved / PEAK_FP32_UNPACKED_TFLOPS,
                     "traffic": traffic_from_profile(),
                     "kernel": "k_force_f32<double,TPT>", "flops_per_pair": FLOPS_PER_PAIR,
                     "avg_kernel_ms": avg_force_s * 1e3, "launches": nforce},
        "cpu_baseline": cpu,
        "batched": batch,
        "device": {"name": dev["name"], "arch": dev["arch"], "cu": dev["compute_units"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
