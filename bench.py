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
PEAK_FP64_TFLOPS = 78.6           # MI355X FP64 vector, published


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
    p.add_argument("--tune", action="store_true",
                   help="autotune on this device instead of using the package's tuning table")
    p.add_argument("--no-tune", action="store_true",
                   help="use the package's tuning table (the default)")
    p.add_argument("--save-tuning", default=None, help="write the autotuner's table here")
    p.add_argument("--tuning-file", default=None, help="load a tuning table before creating contexts")
    return p.parse_args()


def traffic_from_profile(tuning):
    """HBM bytes per force launch from the committed counter summary, when it
    was taken with the launch configuration this run uses."""
    f = ROOT / "profiles" / "force_c3_traffic.json"
    if not f.exists():
        return None
    try:
        prof = json.loads(f.read_text())
    except ValueError:
        return None
    keys = ("targets_per_thread", "j_segments", "far_path")
    if any(prof.get("tuning", {}).get(k) != tuning.get(k) for k in keys):
        return None
    return prof.get("bytes_per_launch")


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _force_timing(ctx, reps=20):
    """Wall-clock time (ms) per force evaluation over `reps` evaluations
    enqueued back to back with one host wait at the end: the force kernel with
    its tile boxes and segment reduction.  For kernels of a millisecond or
    less, HIP events around each launch add their own overhead (at N = 65536
    they read 1.43 ms where the kernel trace gives 1.27 ms)."""
    ctx.compute_force()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.compute_force()
    ctx.sync()
    return (time.perf_counter() - t0) / reps * 1e3


def bench_c1(pkg, nat, args):
    """Plummer N = 4096, eps = 0.01, FP64, 100 KDK steps at dt = 1e-3: the
    reference's own CPU-runnable case.  Both the FP64 mode and split precision
    run the 100 steps (graph-captured), and the oracle's run loop
    (src/run.jl:46-184 restated) runs them on the host for the comparison."""
    from oracle import oracle as orc
    n, h, dt, steps = 4096, 0.01, 1e-3, 100
    pos, vel, mass = pkg.ics.plummer(n, seed=1)
    t0 = time.perf_counter()
    xo, vo, _, info = orc.run_const_kdk(pos, vel, mass, [h] * 6, 1.0, t_begin=0.0,
                                        t_end=steps * dt, dt=dt)
    t_cpu = time.perf_counter() - t0
    E0o = sum(orc.energy(pos, vel, mass, orc.compute_potential(pos, mass, [h] * 6))[:2])
    E1o = sum(orc.energy(xo, vo, mass, orc.compute_potential(xo, mass, [h] * 6))[:2])
    out = {"workload": "Plummer N=4096, eps=0.01, 100 KDK steps at dt=1e-3",
           "steps": info["steps"], "oracle": {"cpu_s": t_cpu, "cores": orc.num_threads(),
                                              "pairs_per_s": float(n) * n * steps / t_cpu,
                                              "rel_energy_drift": abs(E1o - E0o) / abs(E0o)}}
    for mode in ("fp64", "split"):
        ctx = nat.Context(n, nat.MODES[mode])
        ctx.set_params(1.0, [h] * 6)
        ctx.upload(pos, vel, mass)
        e0 = ctx.energy()
        force_ms = _force_timing(ctx)
        # the 100 steps from the initial state (checked against the oracle) ...
        ctx.compute_force()
        ctx.half_kick(+1, dt)
        ctx.kdk_steps(steps, dt, 0.0, use_graph=True)
        ctx.half_kick(-1, dt)
        e1 = ctx.energy()
        d = ctx.download(pos=True, vel=True, acc=False)
        # ... then 100 more, timed, replaying the graphs the first run built
        ctx.half_kick(+1, dt)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.kdk_steps(steps, dt, steps * dt, use_graph=True)
        ctx.sync()
        t_gpu = time.perf_counter() - t0
        ctx.close()
        E0, E1 = e0["K"] + e0["W"], e1["K"] + e1["W"]
        out[mode] = {"s": t_gpu, "ms_per_step": t_gpu / steps * 1e3,
                     "pairs_per_s": float(n) * n * steps / t_gpu,
                     "force_ms": force_ms,
                     "force_frac_of_peak": FLOPS_PER_PAIR * float(n) * n / (force_ms * 1e-3) / 1e12
                                           / (PEAK_FP64_TFLOPS if mode == "fp64" else PEAK_FP32_TFLOPS),
                     "peak_tflops": PEAK_FP64_TFLOPS if mode == "fp64" else PEAK_FP32_TFLOPS,
                     "max_abs_dx_vs_oracle": float(np.max(np.abs(d["pos"] - xo))),
                     "rel_energy_drift": abs(E1 - E0) / abs(E0),
                     "drift_minus_oracle": abs((E1 - E0) - (E1o - E0o)) / abs(E0o)}
    return out


def bench_c2(pkg, nat, args):
    """Plummer N = 65536, eps = 0.005, FP32 state (the reference's floattype =
    Float32 GPU mode): one force evaluation (mean of 20 back to back) and 50
    KDK steps at dt = 5e-4; forces of a seeded target sample against the FP64
    oracle on the FP32-rounded inputs; energy drift (nb_energy)."""
    n, h, dt, steps = 65536, 0.005, 5e-4, 50
    pos, vel, mass = pkg.ics.plummer(n, seed=2)
    ctx = nat.Context(n, nat.NB_MODE_FP32)
    ctx.set_params(1.0, [h] * 6)
    ctx.upload(pos, vel, mass)
    tun = ctx.get_tuning()
    e0 = ctx.energy()
    force_ms = _force_timing(ctx)
    ctx.compute_force()
    acc = ctx.download(pos=False, vel=False)["acc"]
    ctx.half_kick(+1, dt)
    ctx.kdk_steps(steps, dt, 0.0, use_graph=True)
    ctx.half_kick(-1, dt)
    e1 = ctx.energy()
    # 50 more steps, timed, replaying the graphs the first 50 built
    ctx.half_kick(+1, dt)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.kdk_steps(steps, dt, steps * dt, use_graph=True)
    ctx.sync()
    t_gpu = time.perf_counter() - t0
    ctx.close()
    pairs = float(n) * n
    out = {"workload": "Plummer N=65536, eps=0.005, FP32 state, 1 force + 50 KDK steps at dt=5e-4",
           "tuning": tun, "force_ms": force_ms, "force_pairs_per_s": pairs / (force_ms * 1e-3),
           "force_frac_of_peak": FLOPS_PER_PAIR * pairs / (force_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
           "steps": steps, "s": t_gpu, "ms_per_step": t_gpu / steps * 1e3,
           "pairs_per_s": pairs * steps / t_gpu,
           "rel_energy_drift": abs((e1["K"] + e1["W"]) - (e0["K"] + e0["W"])) / abs(e0["K"] + e0["W"])}
    if not args.no_cpu:
        from oracle import oracle as orc
        p32 = pos.astype(np.float32).astype(np.float64)
        m32 = mass.astype(np.float32).astype(np.float64)
        rng = np.random.default_rng(22)
        sample = np.sort(rng.choice(n, 2048, replace=False))
        ref = orc.force_at_points(p32[sample], h, p32, m32, variant="simd")
        e = orc.rel_err(acc[sample], ref)
        out["force_rel_err"] = {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
                                "max": float(e.max()), "targets": int(sample.size)}
    return out


def bench_clusters(pkg, nat, args):
    """512 Plummer clusters, N log-uniform in [256, 16384], a ~ U[0.5, 2],
    h = 0.01 a, 200 KDK steps at dt = 1e-3: one ragged persistent launch per
    force, graph-captured steps, versus sequential one-system launches."""
    steps, dt = 200, 1e-3
    progress("cluster ensemble: generating 512 clusters")
    ens = pkg.ics.cluster_ensemble(nsys=512, seed=5)
    off = ens["offsets"]
    ns = np.diff(off)
    pairs = float((ns.astype(np.float64) ** 2).sum())
    e = pkg.Ensemble(off, "split")
    e.set_params(1.0, ens["h"], dt)
    e.upload(ens["pos"], ens["vel"], ens["mass"])
    e.autotune()
    e.compute_force()
    acc = e.download(pos=False, vel=False)["acc"]
    e.half_kick(+1)
    e.kdk_steps(16)
    e.sync()
    t0 = time.perf_counter()
    e.kdk_steps(steps)
    e.sync()
    t_batch = time.perf_counter() - t0
    tun = e.get_tuning()
    e.close()
    # accuracy on a few clusters (oracle, FP64)
    err = None
    try:
        from oracle import oracle as orc
        es = []
        for s in (0, 100, 200, 300, 400, 511):
            sl = slice(off[s], off[s + 1])
            ref = orc.compute_force(ens["pos"][sl], ens["mass"][sl], [ens["h"][s]] * 6)
            es.append(orc.rel_err(acc[sl], ref))
        es = np.concatenate(es)
        err = {"median": float(np.median(es)), "max": float(es.max()), "clusters": 6}
    except Exception as exc:  # pragma: no cover - reported, not fatal
        err = {"error": str(exc)}
    # sequential plain launches: one context per cluster, every cluster, the
    # same 200 graph-captured steps
    subset = np.arange(512)
    sub_steps = steps
    ctxs = []
    for s in subset:
        sl = slice(off[s], off[s + 1])
        c = nat.Context(int(ns[s]), nat.NB_MODE_SPLIT)
        c.set_params(1.0, [ens["h"][s]] * 6)
        c.upload(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl])
        c.compute_force()
        c.half_kick(+1, dt)
        c.kdk_steps(2, dt, 0.0, use_graph=True)
        c.sync()
        ctxs.append(c)
    t0 = time.perf_counter()
    for c in ctxs:
        c.kdk_steps(sub_steps, dt, 0.0, use_graph=True)
        c.sync()
    t_sub = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    t_plain = t_sub
    return {"workload": "512 Plummer clusters, N logU[256,16384], 200 KDK steps",
            "n_total": int(off[-1]), "pairs_per_step": pairs, "steps": steps,
            "batched_s": t_batch, "batched_pairs_per_s": pairs * steps / t_batch,
            "batched_frac_of_peak": FLOPS_PER_PAIR * pairs * steps / t_batch / 1e12 / PEAK_FP32_TFLOPS,
            "plain_s": t_plain, "plain_pairs_per_s": pairs * steps / t_plain,
            "speedup": t_plain / t_batch, "force_rel_err": err, "tuning": tun,
            "plain": "all 512 clusters, one context each, 200 graph-captured steps each, "
                     "one after the other"}


def bench_planetary(pkg, nat, args):
    """2048 planetary systems (central mass + 8 planets + 64 test particles),
    eps = 0, 10^4 KDK steps at dt = 1e-3, energy tracked every 100 steps: the
    resident multi-step kernel versus sequential one-system launches."""
    steps, every, dt = 10000, 100, 1e-3
    progress("planetary ensemble: generating 2048 systems")
    ens = pkg.ics.planetary_ensemble(nsys=2048, seed=6)
    off = ens["offsets"]
    per = ens["per_system"]
    nmass = 9
    e = pkg.Ensemble(off, "split_central")
    e.set_params(1.0, 0.0, dt)
    e.upload(ens["pos"], ens["vel"], ens["mass"])
    e.set_central(1)
    e.compute_force()
    e0 = e.energy()
    E0 = e0["K"] + e0["W"]
    e.half_kick(+1)
    e.sync()
    t0 = time.perf_counter()
    E = e.kdk_steps_tracked(steps, every)
    e.sync()
    t_batch = time.perf_counter() - t0
    e.close()
    rel = np.abs(E - E0[:, None]) / np.abs(E0[:, None])
    # the same 2048 systems through the general batched path (one ragged force
    # launch per step, graph-captured), without the resident kernel
    progress("planetary ensemble: batched path without the resident kernel")
    gen_steps = 1000
    e = pkg.Ensemble(off, "split_central")
    e.set_params(1.0, 0.0, dt)
    e.upload(ens["pos"], ens["vel"], ens["mass"])
    e.set_central(1)
    e.set_tuning(resident=0)
    e.compute_force()
    e.half_kick(+1)
    e.kdk_steps(32)
    e.sync()
    t0 = time.perf_counter()
    e.kdk_steps(gen_steps)
    e.sync()
    t_gen = (time.perf_counter() - t0) / gen_steps * steps
    e.close()
    # sequential plain launches on a seeded subset of systems
    rng = np.random.default_rng(12)
    subset = rng.choice(2048, 8, replace=False)
    sub_steps = 200
    ctxs = []
    for s in subset:
        sl = slice(off[s], off[s + 1])
        c = nat.Context(per, nat.NB_MODE_SPLIT_CENTRAL)
        c.set_params(1.0, [0.0] * 6)
        c.upload(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl])
        c.set_central(1)
        c.compute_force()
        c.half_kick(+1, dt)
        c.kdk_steps(16, dt, 0.0, use_graph=True)
        c.sync()
        ctxs.append(c)
    t0 = time.perf_counter()
    for c in ctxs:
        c.kdk_steps(sub_steps, dt, 0.0, use_graph=True)
        c.sync()
    t_sub = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    t_plain = t_sub / (len(subset) * sub_steps) * 2048 * steps
    allpairs = 2048.0 * per * per
    evaluated = 2048.0 * per * nmass
    out = {"workload": "2048 planetary systems x 73 bodies, eps=0, 1e4 KDK steps, "
                       "energy every 100 steps",
           "steps": steps, "batched_s": t_batch,
           "us_per_step": t_batch / steps * 1e6,
           "batched_pairs_per_s_allpairs": allpairs * steps / t_batch,
           "batched_pairs_per_s_evaluated": evaluated * steps / t_batch,
           # 20 flops per pair against the FP32 peak; the evaluated pairs (every
           # body against the 9 massive ones) are the work the kernel does
           "evaluated_frac_of_peak": FLOPS_PER_PAIR * evaluated * steps / t_batch / 1e12 / PEAK_FP32_TFLOPS,
           "allpairs_frac_of_peak": FLOPS_PER_PAIR * allpairs * steps / t_batch / 1e12 / PEAK_FP32_TFLOPS,
           "plain_s_estimate": t_plain, "speedup": t_plain / t_batch,
           "general_batched_s_estimate": t_gen, "speedup_vs_general_batched": t_gen / t_batch,
           "general_batched_sample": f"all 2048 systems x {gen_steps} graph-captured steps, "
                                     "resident kernel off, scaled",
           "rel_energy_err": {"median": float(np.median(rel)),
                              "p99": float(np.percentile(rel.max(1), 99)),
                              "max": float(rel.max()),
                              "systems_above_1e-3": int((rel.max(1) > 1e-3).sum())},
           "plain_sample": f"{len(subset)} seeded systems x {sub_steps} steps, scaled"}
    if not args.no_cpu:
        # the FP64 oracle integrates the same ensemble: CPU baseline and a
        # system-by-system check of the tracked energy errors
        progress("planetary ensemble: FP64 oracle run")
        from oracle import oracle as orc
        t0 = time.perf_counter()
        _, _, E0o, Eo = orc.run_ensemble_kdk(off, ens["pos"], ens["vel"], ens["mass"], 1.0, 0.0,
                                             dt, nsteps=steps, every=every)
        t_cpu = time.perf_counter() - t0
        relo = np.abs(Eo - E0o[:, None]) / np.abs(E0o[:, None])
        calm = relo.max(1) < 1e-3
        out["oracle"] = {"cpu_s": t_cpu, "cores": orc.num_threads(),
                         "rel_energy_err_median": float(np.median(relo)),
                         "rel_energy_err_max": float(relo.max()),
                         "systems_above_1e-3": int((~calm).sum()),
                         "gpu_minus_oracle_energy_err_calm_systems_max":
                             float(np.max(np.abs(rel[calm] - relo[calm]))),
                         "evaluated_pairs_per_s": evaluated * steps / t_cpu}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); sharding this path over "
              "several GPUs (targets split, sources all-gathered per step) is not built",
              file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as ge
    pkg = ge.load_package()
    nat = pkg._native
    if args.tuning_file:
        nat.tuner_load(args.tuning_file)
    n = args.n
    h, dt = 0.002, 1e-3
    pos, vel, mass = pkg.ics.collision(n // 2, seeds=(3, 4))
    n = pos.shape[0]
    dev = nat.device_info()

    ctx = nat.Context(n, nat.NB_MODE_SPLIT)
    ctx.set_params(1.0, [h] * 6)
    ctx.upload(pos, vel, mass)
    if args.tune and not args.no_tune:
        progress("autotuning the all-pairs launch configuration")
        ctx.autotune()
        if args.save_tuning:
            nat.tuner_save(args.save_tuning)
    progress(f"tuning {ctx.get_tuning()}")
    # initial force (also the accuracy sample) and the first half kick
    ctx.compute_force()
    acc0 = ctx.download(pos=False, vel=False, acc=True)["acc"]
    ctx.half_kick(+1, dt)
    t = 0.0
    for _ in range(args.warmup):
        ctx.kdk_steps(1, dt, t)
        t += dt
    ctx.sync()

    progress(f"N={n}: {args.warmup} warm-up steps done, timing {args.steps} steps")
    ctx.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.kdk_steps(args.steps, dt, t)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    nforce, force_ms = ctx.get_timing()
    # five separate force evaluations, each timed with HIP events: median
    singles = []
    for _ in range(5):
        c0, m0 = ctx.get_timing()
        ctx.compute_force()
        ctx.sync()
        c1, m1 = ctx.get_timing()
        singles.append((m1 - m0) / max(c1 - c0, 1))
    ctx.set_timing(False)
    tuning = ctx.get_tuning()
    # the same force evaluation with the boundary's host copies: upload of
    # positions, velocities and masses, the force, download of the accelerations
    ctx.sync()
    t1 = time.perf_counter()
    ctx.upload(pos, vel, mass)
    ctx.compute_force()
    _ = ctx.download(pos=False, vel=False, acc=True)["acc"]
    pcie_s = time.perf_counter() - t1
    ctx.close()

    pairs_per_step = float(n) * float(n)
    value = pairs_per_step * args.steps / elapsed
    avg_force_s = force_ms / max(nforce, 1) * 1e-3
    achieved = FLOPS_PER_PAIR * pairs_per_step / avg_force_s / 1e12

    # accuracy + CPU baseline on a seeded sample of targets against all sources
    cpu = None
    err = None
    progress(f"timed steps done ({elapsed:.2f} s)")
    if not args.no_cpu:
        progress("CPU oracle sample")
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
        # the faithful variant (per-target scratch array + pairwise sum, as
        # src/directsum/gravity.jl:40-45) on a smaller sample of the same targets
        progress("CPU oracle sample (faithful variant)")
        kf = int(min(k, max(32, k // 8)))
        t1 = time.perf_counter()
        ref_f = orc.force_at_points(pos[sample[:kf]], h, pos, mass, variant="faithful")
        faithful_s = time.perf_counter() - t1
        cpu = {"value": float(k) * n / cpu_s, "unit": "pair interactions/s", "cores": threads,
               "kind": "port",
               "sample": f"{k} seeded targets x all {n} sources of the initial state, FP64 "
                         f"C restatement of src/directsum (SIMD variant, OpenMP {threads} threads)",
               "variants": {"simd": float(k) * n / cpu_s,
                            "faithful": float(kf) * n / faithful_s,
                            "faithful_sample": f"{kf} of the same targets",
                            "faithful_vs_simd_max_rel": float(np.max(orc.rel_err(ref_f, ref[:kf])))}}

    batch = None
    if not args.no_batch:
        progress("C1: Plummer N=4096, 100 steps, FP64 and split precision, with the oracle run")
        c1 = bench_c1(pkg, nat, args) if not args.no_cpu else None
        progress("C2: Plummer N=65536, FP32 state")
        c2 = bench_c2(pkg, nat, args)
        batch = {"clusters": bench_clusters(pkg, nat, args),
                 "planetary": bench_planetary(pkg, nat, args)}

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
        # one force evaluation including the host-buffer boundary (PCIe): not `value`
        "host_inclusive": {"pairs_per_s": float(n) * float(n) / pcie_s, "s": pcie_s,
                           "what": "nb_upload (pos, vel, mass) + nb_compute_force + nb_download (acc)"},
        # compute-bound on the vector ALUs: priced against MI355X's FP32 peak
        # (157.3 TFLOP/s); the pair loop issues unpacked FP32 VALU + v_rsq_f32,
        # and with far path 4 the far tiles' squared distances come from bf16
        # MFMAs beside it (a quarter MFMA per 16 pairs), so VALU issue sets the rate
        "roofline": {"bound": "valu",
                     "compute_unit": "FP32 VALU (unpacked) + v_rsq_f32; bf16 MFMA r^2 on far tiles"
                                     if tuning["far_path"] == 4 else "FP32 VALU (unpacked) + v_rsq_f32",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profile(tuning),
                     # far path 4 with 1 target per lane runs as a far pass and a
                     # near pass (one force evaluation = both, timed together)
                     "kernel": ("k_force_f32<double,1,false,4> far pass + k_force_near<double>"
                                if tuning["far_path"] == 4 and tuning["targets_per_thread"] == 1 else
                                "k_force_f32<double,%d,false,%s>" % (
                                    tuning["targets_per_thread"], tuning["far_path"])),
                     "flops_per_pair": FLOPS_PER_PAIR,
                     "avg_kernel_ms": avg_force_s * 1e3, "launches": nforce,
                     "median_single_force_ms": float(np.median(singles)), "single_forces": len(singles)},
        "cpu_baseline": cpu,
        "configs": None if args.no_batch else {"plummer_4096_fp64": c1, "plummer_65536_fp32": c2},
        "batched": batch,
        "device": {"name": dev["name"], "arch": dev["arch"], "cu": dev["compute_units"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
