"""The benchmark configurations' full integration legs (needs an MI355X).

Energy and momentum drift over the integration are part of the oracle for
this path (the constant-dt KDK loop of src/base/Timestep.jl:84-125,181-215
and src/run.jl:100-184).  Each test runs a configuration's whole leg on the
GPU and checks it against the FP64 oracle where the oracle can follow it
(sampled targets, sampled systems, every planetary system), and against the
conservation laws where it cannot:

* Plummer N = 65536, FP32 state, 50 KDK steps at dt = 5e-4: step count,
  sampled forces at the final state against the oracle on the FP32 state,
  total energy at both ends against the oracle, energy drift;
* two-sphere collision N = 2^20, split precision, 20 KDK steps at dt = 1e-3:
  energy, momentum and angular momentum drift (nb_energy), sampled forces at
  the final state;
* 512 Plummer clusters, 200 graph-captured KDK steps: six seeded clusters
  against the oracle's run loop;
* 2048 planetary systems, 10^4 KDK steps with the energy tracked every 100:
  every system's energy error against the oracle's own integration.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _energy_oracle(orc, pos, vel, mass, h):
    pot = orc.compute_potential(pos, mass, [h] * 6)
    K, W, P, L = orc.energy(pos, vel, mass, pot)
    return K + W


def test_c2_plummer_fp32_fifty_steps(pkg, orc, tmp_path):
    n, h, dt, steps = 65536, 0.005, 5e-4, 50
    pos, vel, mass = pkg.ics.plummer(n, seed=2)
    t_end = steps * dt
    assert orc.count_steps(0.0, t_end, dt) == steps
    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), precision="fp32", TimeStep=dt,
                         TimeEnd=t_end, TimeBetweenSnapshots=t_end, ForceSofteningTable=[h] * 6,
                         output_func=lambda s: None, OutputDir=str(tmp_path))
    try:
        E0 = pkg.total_energy(sim)
        d0 = pkg.get_all_data(sim)
        pkg.run(sim)
        E1 = pkg.total_energy(sim)
        d = pkg.get_all_data(sim)
        assert sim.timeinfo.stepcount == steps
    finally:
        sim.close()
    # the state is FP32: the oracle works on the values the GPU holds
    m32 = mass.astype(np.float32).astype(np.float64)
    E0o = _energy_oracle(orc, d0.Pos, d0.Vel, m32, h)
    E1o = _energy_oracle(orc, d.Pos, d.Vel, m32, h)
    rng = np.random.default_rng(31)
    sample = np.sort(rng.choice(n, 2048, replace=False))
    ref = orc.force_at_points(d.Pos[sample], h, d.Pos, m32, variant="simd")
    e = orc.rel_err(d.Acc[sample], ref)
    drift = abs(E1 - E0) / abs(E0)
    print(f"C2 after {steps} steps: force median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e}; "
          f"E0 {E0:.12f} (oracle {E0o:.12f}), E1 {E1:.12f} (oracle {E1o:.12f}), |dE/E| {drift:.3e}")
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-4
    assert abs(E0 - E0o) / abs(E0o) <= 1e-6 and abs(E1 - E1o) / abs(E1o) <= 1e-6
    assert drift <= 1e-4


def test_c3_collision_twenty_steps_conservation(pkg, orc):
    pos, vel, mass = pkg.ics.collision(1 << 19, seeds=(3, 4))
    n, h, dt, steps = len(mass), 0.002, 1e-3, 20
    nat = pkg._native
    with nat.Context(n, nat.NB_MODE_SPLIT) as ctx:
        ctx.set_params(1.0, [h] * 6)
        ctx.upload(pos, vel, mass)
        ctx.compute_force()
        e0 = ctx.energy()
        ctx.half_kick(+1, dt)
        ctx.kdk_steps(steps, dt, 0.0, use_graph=True)
        ctx.half_kick(-1, dt)
        e1 = ctx.energy()
        d = ctx.download(pos=True, vel=True, acc=True)
    E0, E1 = e0["K"] + e0["W"], e1["K"] + e1["W"]
    dE = abs(E1 - E0) / abs(E0)
    dP = np.max(np.abs(e1["P"] - e0["P"]))
    dL = np.max(np.abs(e1["L"] - e0["L"]))
    rng = np.random.default_rng(32)
    sample = np.sort(rng.choice(n, 256, replace=False))
    ref = orc.force_at_points(d["pos"][sample], h, d["pos"], mass, variant="simd")
    err = orc.rel_err(d["acc"][sample], ref)
    moved = np.max(np.abs(d["pos"] - pos))
    print(f"C3 {steps} steps: |dE/E| {dE:.3e}, |dP| {dP:.3e}, |dL| {dL:.3e} (total mass 2); "
          f"moved {moved:.3e}; final force median {np.median(err):.3e} max {err.max():.3e}")
    assert moved > 1e-3
    assert dE <= 1e-6 and dP <= 1e-10 and dL <= 1e-9
    assert np.median(err) <= 2e-7 and err.max() <= 1e-5


def test_c4_clusters_two_hundred_graph_steps(pkg, orc):
    ens = pkg.ics.cluster_ensemble(nsys=512, seed=5)
    off = ens["offsets"]
    ns = np.diff(off)
    dt, steps = 1e-3, 200
    assert orc.count_steps(0.0, steps * dt, dt) == steps
    e = pkg.Ensemble(off, "split")
    try:
        e.set_params(1.0, ens["h"], dt)
        e.upload(ens["pos"], ens["vel"], ens["mass"])
        tun = e.get_tuning()
        e.compute_force()
        e.half_kick(+1)
        e.kdk_steps(steps, use_graph=True)
        e.half_kick(-1)
        d = e.download(pos=True, vel=True, acc=False)
    finally:
        e.close()
    rng = np.random.default_rng(33)
    small = np.nonzero(ns <= 3000)[0]
    worst_x = worst_v = 0.0
    for s in np.sort(rng.choice(small, 6, replace=False)):
        sl = slice(off[s], off[s + 1])
        xo, vo, _, info = orc.run_const_kdk(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl],
                                            [ens["h"][s]] * 6, 1.0, t_begin=0.0, t_end=steps * dt,
                                            dt=dt)
        assert info["steps"] == steps
        dx = np.max(np.abs(d["pos"][sl] - xo))
        dv = np.max(np.abs(d["vel"][sl] - vo)) / np.max(np.abs(vo))
        print(f"cluster {s} (N={ns[s]}): max |dx| {dx:.3e}, max |dv|/max|v| {dv:.3e}")
        worst_x, worst_v = max(worst_x, dx), max(worst_v, dv)
    print(f"C4 {steps} graph-captured steps, tuning {tun}: worst |dx| {worst_x:.3e}, |dv| {worst_v:.3e}")
    assert worst_x <= 1e-6 and worst_v <= 1e-5


def test_c5_planetary_ten_thousand_steps_energy(pkg, orc):
    steps, every, dt = 10000, 100, 1e-3
    ens = pkg.ics.planetary_ensemble(nsys=2048, seed=6)
    off = ens["offsets"]
    e = pkg.Ensemble(off, "split_central")
    try:
        e.set_params(1.0, 0.0, dt)
        e.upload(ens["pos"], ens["vel"], ens["mass"])
        e.set_central(1)
        e.compute_force()
        en = e.energy()
        E0 = en["K"] + en["W"]
        e.half_kick(+1)
        E = e.kdk_steps_tracked(steps, every)
    finally:
        e.close()
    _, _, E0o, Eo = orc.run_ensemble_kdk(off, ens["pos"], ens["vel"], ens["mass"], 1.0, 0.0, dt,
                                         nsteps=steps, every=every)
    rel = np.abs(E - E0[:, None]) / np.abs(E0[:, None])
    relo = np.abs(Eo - E0o[:, None]) / np.abs(E0o[:, None])
    calm = relo.max(1) < 1e-3
    diff = np.max(np.abs(rel[calm] - relo[calm]))
    print(f"C5 {steps} steps: {int(calm.sum())} calm systems; energy error median {np.median(rel):.3e} "
          f"(oracle {np.median(relo):.3e}); max |gpu - oracle| over calm systems {diff:.3e}")
    # nb_batch_energy takes every pair term of the star in FP64 (its own
    # potential included); the planet-planet terms are FP32, ~1e-10 of E
    assert np.max(np.abs(E0 - E0o) / np.abs(E0o)) <= 1e-8
    assert calm.sum() >= 2000
    assert diff <= 5e-5
    assert np.median(np.abs(rel - relo)) <= 1e-6
