"""End-to-end runs through the Simulation / run mirror on the GPU, compared
with the oracle's restatement of run(sim) (src/run.jl:46-184) on identical
seeded inputs (needs an MI355X)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _in_tmp(tmp_path, monkeypatch):
    """Default snapshots go to ./output, as in the reference: run in a scratch dir."""
    monkeypatch.chdir(tmp_path)


@pytest.fixture(scope="module")
def c1_inputs(pkg):
    """Config 1: Plummer N=4096, G=M=a=1, eps=0.01, dt=1e-3, 100 KDK steps."""
    return pkg.ics.plummer(4096, seed=1)


@pytest.fixture(scope="module")
def c1_oracle(orc, c1_inputs):
    pos, vel, mass = c1_inputs
    soft = [0.01] * 6
    x, v, a, info = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.1,
                                      dt=1e-3, between_snapshots=0.1)
    pot0 = orc.compute_potential(pos, mass, soft)
    pot1 = orc.compute_potential(x, mass, soft)
    E0 = sum(orc.energy(pos, vel, mass, pot0)[:2])
    E1 = sum(orc.energy(x, v, mass, pot1)[:2])
    return x, v, info, E0, E1


# FP32 state: positions are rounded to FP32 at every one of the 100 drifts
@pytest.mark.parametrize("precision,tol_x", [("fp64", 1e-11), ("split", 2e-8), ("fp32", 1e-5)])
def test_config1_run_matches_oracle(pkg, orc, c1_inputs, c1_oracle, precision, tol_x):
    pos, vel, mass = c1_inputs
    xo, vo, info, E0o, E1o = c1_oracle
    data = pkg.StarData(pos, Vel=vel, Mass=mass)
    sim = pkg.Simulation(data, precision=precision, TimeStep=1e-3, TimeEnd=0.1,
                         TimeBetweenSnapshots=0.1, ForceSofteningTable=[0.01] * 6)
    E0 = pkg.total_energy(sim)
    pkg.run(sim)
    d = pkg.get_all_data(sim)
    E1 = pkg.total_energy(sim)
    assert sim.timeinfo.stepcount == info["steps"] == 100
    # FP32 state rounds positions to FP32 at every drift: compare relative to |x|
    scale = np.maximum(1.0, np.abs(xo)) if precision == "fp32" else 1.0
    err = np.max(np.abs(d.Pos - xo) / scale)
    print(f"{precision}: max |dx| {err:.3e}, dE/E gpu {abs(E1 - E0) / abs(E0):.3e} "
          f"oracle {abs(E1o - E0o) / abs(E0o):.3e}")
    assert err < tol_x
    assert abs(E0 - E0o) / abs(E0o) < (1e-12 if precision == "fp64" else 1e-6)
    assert abs((E1 - E0) - (E1o - E0o)) / abs(E0o) < (1e-12 if precision == "fp64" else 1e-6)
    sim.close()


def test_step_count_by_float_accumulation(pkg, orc):
    """run(sim) counts steps as the reference does, by adding dt to the time
    until it reaches TimeEnd (src/base/Timestep.jl:198, src/run.jl:158-160):
    10^4 steps of 1e-3 up to 10.0 take 10001 steps, also when stretches of
    steps are handed to nb_kdk_steps in one call."""
    pos, vel, mass = pkg.ics.plummer(512, seed=9)
    data = pkg.StarData(pos, Vel=vel, Mass=mass)
    sim = pkg.Simulation(data, precision="split", TimeStep=1e-3, TimeEnd=10.0,
                         TimeBetweenSnapshots=10.0, ForceSofteningTable=[0.05] * 6)
    pkg.run(sim)
    t, k = 0.0, 0
    while t < 10.0:
        t += 1e-3
        k += 1
    assert sim.timeinfo.stepcount == orc.count_steps(0.0, 10.0, 1e-3) == k == 10001
    assert sim.timeinfo.system_time_float == t
    sim.close()


def test_snapshots_analysers_and_stop_file(pkg, orc, tmp_path):
    pos, vel, mass = pkg.ics.plummer(512, seed=3)
    soft = [0.02] * 6
    seen = []

    def out(sim):
        seen.append(sim.timeinfo.system_time_float)

    oracle_times = []
    orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.02, dt=3e-3,
                      between_snapshots=0.0045, on_output=lambda t, p, v: oracle_times.append(t))
    data = pkg.StarData(pos, Vel=vel, Mass=mass)
    sim = pkg.Simulation(data, TimeStep=3e-3, TimeEnd=0.02, TimeBetweenSnapshots=0.0045,
                         ForceSofteningTable=soft, output_func=out, OutputDir=str(tmp_path),
                         analysers={"K": pkg.total_kinetic})
    res = pkg.run(sim)
    assert np.allclose(seen, oracle_times, rtol=0, atol=1e-15)
    nsteps = orc.count_steps(0.0, 0.02, 3e-3)
    assert res.shape == (nsteps, 2)                  # time, then each analyser (Analyse.jl:7-22)
    lines = open(os.path.join(str(tmp_path), "analysis.csv")).read().splitlines()
    assert lines[0] == "time,K" and len(lines) == nsteps + 1
    assert np.allclose([float(l.split(",")[0]) for l in lines[1:]], res[:, 0], rtol=1e-15)
    # stop file: run ends after the current step (src/run.jl:153-157)
    sim2 = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), TimeStep=1e-3, TimeEnd=1.0,
                          ForceSofteningTable=soft, OutputDir=str(tmp_path),
                          poststep=[lambda s: open(os.path.join(str(tmp_path), "stop"), "w").close()])
    with pytest.warns(UserWarning, match="Interrupted by user"):   # src/run.jl:153-157
        pkg.run(sim2)
    assert sim2.timeinfo.stepcount == 1
    sim.close()
    sim2.close()


def test_stop_file_without_host_events(pkg, orc, tmp_path):
    """A run with no analyser, poststep or restart and one snapshot interval
    hands stretches of steps to nb_kdk_steps.  A stop file present when the
    loop starts (here written by the output function at snapshot 0) ends it
    after one step, as src/run.jl:153-157 does; one that appears during a
    stretch ends it at the end of that stretch (at most 256 steps)."""
    pos, vel, mass = pkg.ics.plummer(256, seed=5)
    soft = [0.02] * 6
    stop = os.path.join(str(tmp_path), "stop")
    calls = []

    def out(sim):
        calls.append(sim.timeinfo.system_time_float)
        open(stop, "w").close()

    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), TimeStep=1e-3, TimeEnd=10.0,
                         TimeBetweenSnapshots=10.0, ForceSofteningTable=soft, output_func=out,
                         OutputDir=str(tmp_path))
    with pytest.warns(UserWarning, match="Interrupted by user"):
        pkg.run(sim)
    assert calls == [0.0] and sim.timeinfo.stepcount == 1
    sim.close()

    # the file appears while a stretch runs: a thread writes it after the
    # loop has started; the run ends within one stretch of 256 steps
    import threading
    os.remove(stop)
    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), TimeStep=1e-3, TimeEnd=100.0,
                         TimeBetweenSnapshots=100.0, ForceSofteningTable=soft,
                         output_func=lambda s: threading.Timer(0.2, lambda: open(stop, "w").close()).start(),
                         OutputDir=str(tmp_path))
    with pytest.warns(UserWarning, match="Interrupted by user"):
        pkg.run(sim)
    n = sim.timeinfo.stepcount
    assert 1 <= n < orc.count_steps(0.0, 100.0, 1e-3)
    print(f"stop file during a stretch: run ended after {n} steps")
    sim.close()


def test_public_api_pieces(pkg, orc):
    pos, vel, mass = pkg.ics.plummer(800, seed=4)
    soft = [0.01] * 6
    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), precision="fp64", TimeStep=1e-3,
                         ForceSofteningTable=soft)
    pkg.compute_force(sim)
    d = pkg.get_all_data(sim)
    assert np.max(orc.rel_err(d.Acc, orc.compute_force(pos, mass, soft))) < 1e-12
    x, y, z = pkg.pack_pos(sim)      # src/base/Plot.jl:1-5, caller order
    assert np.array_equal(np.stack([x, y, z], axis=1), pos) and pkg.numlocal(sim) == 800
    pkg.compute_potential(sim)      # with the reference's self term
    d = pkg.get_all_data(sim)
    ref = orc.compute_potential(pos, mass, soft, self_term=True)
    assert np.max(np.abs(d.Potential - ref) / np.abs(ref)) < 1e-12
    pts = np.array([[0.1, 0.2, 0.3], [5.0, 0.0, 0.0]])
    a = pkg.compute_force(sim, pts, 0.05)
    assert np.max(orc.rel_err(a, orc.force_at_points(pts, 0.05, pos, mass))) < 1e-12
    p = pkg.compute_potential(sim, pts[0], 0.05)
    assert p == pytest.approx(orc.potential_at_points(pts[:1], 0.05, pos, mass)[0], rel=1e-12)
    pkg.set_softlen(sim, 0.03)
    assert pkg.softlen(pkg.STAR, sim) == 0.03
    pkg.compute_force(sim)
    d = pkg.get_all_data(sim)
    assert np.max(orc.rel_err(d.Acc, orc.compute_force(pos, mass, [0.03] * 6))) < 1e-12
    K, W, P, L = orc.energy(pos, vel, mass, orc.compute_potential(pos, mass, [0.03] * 6))
    assert pkg.total_kinetic(sim) == pytest.approx(K, rel=1e-13)
    assert pkg.total_potential(sim) == pytest.approx(W, rel=1e-12)
    assert np.allclose(pkg.total_momentum(sim), P, atol=1e-15)
    assert np.allclose(pkg.total_angular_momentum(sim), L, rtol=1e-12, atol=1e-15)
    sim.close()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_adaptive_run_matches_oracle(pkg, orc, precision):
    pos, vel, mass = pkg.ics.plummer(256, seed=9)
    soft = [0.05] * 6
    kw = dict(t_begin=0.0, t_end=0.05, time_base=1 << 16, err_tol=1e-4)
    xo, vo, ao, info = orc.run_adaptive(pos, vel, mass, soft, 1.0, **kw)
    data = pkg.StarData(pos, Vel=vel, Mass=mass)
    sim = pkg.Simulation(data, precision=precision, TimeStep=0.0, TimeEnd=0.05, TimeBase=1 << 16,
                         ErrTolTimestep=1e-4, TimeBetweenSnapshots=0.05,
                         ForceSofteningTable=soft)
    pkg.run(sim)
    d = pkg.get_all_data(sim)
    assert sim._adaptive.nactive_history == info["nactive"][: len(sim._adaptive.nactive_history)]
    assert np.array_equal(d.Ti_endstep, info["ti_end"]) and np.array_equal(d.Ti_begstep, info["ti_beg"])
    tol = 1e-12 if precision == "fp64" else 1e-7
    assert np.max(np.abs(d.Pos - xo)) < tol
    assert np.max(np.abs(d.Vel - vo)) < 100 * tol
    sim.close()


def test_adaptive_timestep_error_surfaces(pkg):
    pos, vel, mass = pkg.ics.binary(1.0, 1e-6, 1.0)   # huge acceleration at tiny h
    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), TimeStep=0.0, TimeEnd=1.0,
                         MinStep=1e-3, ForceSofteningTable=[1e-9] * 6)
    with pytest.raises(pkg.NBError) as ei:
        pkg.run(sim)
    assert "MinStep" in str(ei.value)
    sim.close()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_mond_postprocessing(pkg, orc, precision):
    """MOND1983Milgrom: Newtonian force, background field, then nu(|a|/ACC0)
    (src/base/PostProcessing/Force.jl:33-46)."""
    pos, vel, mass = pkg.ics.plummer(1024, seed=21)
    soft = [0.01] * 6
    acc0, nidx = 0.3, 1.5
    bg = (1, [0.5, 2.0, 0.0, 0.0])
    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), precision=precision,
                         TimeStep=1e-3, ForceSofteningTable=soft, model=pkg.MOND1983Milgrom(),
                         constants=pkg.Constants(1.0, acc0), MOND_nuIndex=nidx, bgforce=[bg])
    pkg.compute_force(sim)
    a = pkg.get_all_data(sim).Acc
    ref = orc.compute_force(pos, mass, soft)
    orc.background_point_mass(pos, 0.5, np.array([2.0, 0.0, 0.0]), acc=ref)
    ref = orc.mond(ref, acc0, nidx)
    e = orc.rel_err(a, ref)
    # split: FP32 pair terms, gated like the force itself (median 1e-6, max 1e-5)
    if precision == "fp64":
        assert e.max() < 1e-12
    else:
        assert np.median(e) <= 1e-6 and e.max() < 1e-5
    sim.close()


def test_mond_adaptive_active_subset(pkg, orc):
    """MOND on the active targets of adaptive steps (src/mond/milgrom1983.jl:40-44)."""
    pos, vel, mass = pkg.ics.plummer(256, seed=23)
    soft = [0.05] * 6
    acc0, nidx = 0.5, 2.0
    kw = dict(t_begin=0.0, t_end=0.03, time_base=1 << 16, err_tol=1e-4)
    xo, vo, ao, info = orc.run_adaptive(pos, vel, mass, soft, 1.0,
                                        post=lambda p, a: orc.mond(a, acc0, nidx), **kw)
    assert len(set(info["nactive"])) > 1          # some steps move a strict subset
    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), precision="fp64", TimeStep=0.0,
                         TimeEnd=0.03, TimeBase=1 << 16, ErrTolTimestep=1e-4,
                         TimeBetweenSnapshots=0.03, ForceSofteningTable=soft,
                         model=pkg.MOND1983Milgrom(), constants=pkg.Constants(1.0, acc0),
                         MOND_nuIndex=nidx)
    pkg.run(sim)
    d = pkg.get_all_data(sim)
    assert sim._adaptive.nactive_history == info["nactive"][: len(sim._adaptive.nactive_history)]
    assert np.array_equal(d.Ti_endstep, info["ti_end"])
    assert np.max(np.abs(d.Pos - xo)) < 1e-12
    sim.close()


@pytest.mark.parametrize("format2", [True, False])
def test_gadget2_snapshots_match_oracle(pkg, orc, format2):
    """Default output (src/base/Output.jl:36-63): asynchronous Gadget-2 snapshots
    hold the state at each output time while the run goes on."""
    from astronbodysim_jl_amd.snapshot import read_gadget2
    pos, vel, mass = pkg.ics.plummer(2000, seed=31)
    soft = [0.02] * 6
    coll = np.where(np.arange(2000) % 3 == 0, pkg.HALO, pkg.DISK).astype(np.uint8)
    snaps = []
    orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.05, dt=2e-3,
                      between_snapshots=0.01, collection=coll,
                      on_output=lambda t, p, v: snaps.append((t, p.copy(), v.copy())))
    data = pkg.StarData(pos, Vel=vel, Mass=mass, Collection=coll)
    sim = pkg.Simulation(data, precision="fp64", TimeStep=2e-3, TimeEnd=0.05,
                         TimeBetweenSnapshots=0.01, ForceSofteningTable=soft, format2=format2,
                         acc=format2, pot=format2)
    pkg.run(sim)
    assert sim.outputinfo["snapshotcount"] == len(snaps)
    order = np.argsort(coll, kind="stable")                 # files are grouped by type
    for k, (t, p, v) in enumerate(snaps):
        h, b = read_gadget2(os.path.join("output", "snapshot_%04d.gadget2" % k))
        assert h.time == t and h.npart == [0, 667, 1333, 0, 0, 0] and h.Omega0 == 0.3
        assert np.array_equal(b["ID"], np.arange(1, 2001)[order])
        assert np.array_equal(b["POS"], p[order].astype(np.float32).astype(np.float64))
        assert np.array_equal(b["VEL"], v[order].astype(np.float32).astype(np.float64))
        assert np.array_equal(b["Mass"], mass[order].astype(np.float32).astype(np.float64))
        if format2:
            assert "POT" in b and "ACCE" in b
            ref = orc.compute_potential(p, mass, soft, collection=coll, self_term=True)
            assert np.allclose(b["POT"], ref[order], rtol=1e-6, atol=0)
    sim.close()


@pytest.mark.parametrize("adaptive", [False, True])
def test_restart_resumes_bit_identical(pkg, orc, tmp_path, adaptive):
    """saverestart inside run (src/run.jl:138-141), loadrestart
    (src/restart.jl:127-136) and restart(sim) (src/run.jl:186-189): in FP64 the
    resumed run ends in exactly the state of the uninterrupted one, and both
    end where the oracle's uninterrupted run loop (constant or block steps)
    ends on the same inputs."""
    import shutil
    pos, vel, mass = pkg.ics.plummer(700, seed=41)
    soft = [0.03] * 6
    kw = dict(TimeStep=0.0, TimeBase=1 << 14, ErrTolTimestep=1e-3) if adaptive else dict(TimeStep=2e-3)
    rfile = os.path.join(str(tmp_path), "out", "restart.1.npz")
    mid = os.path.join(str(tmp_path), "mid.npz")

    def grab(sim):
        if os.path.isfile(rfile) and not os.path.isfile(mid):
            shutil.copy(rfile, mid)

    sim = pkg.Simulation(pkg.StarData(pos, Vel=vel, Mass=mass), precision="fp64", TimeEnd=0.04,
                         TimeBetweenSnapshots=0.01, ForceSofteningTable=soft,
                         OutputDir=os.path.join(str(tmp_path), "out"), SaveRestart=True,
                         SaveRestartFreq=0.5, poststep=[grab], **kw)
    pkg.run(sim)
    ref = pkg.get_all_data(sim)
    ref_ti = sim.timeinfo
    nsnap = sim.outputinfo["snapshotcount"]
    sim.close()
    assert os.path.isfile(mid)
    sim2 = pkg.loadrestart(mid)
    assert 0.02 <= sim2.timeinfo.system_time_float < 0.04
    pkg.restart(sim2)
    d = pkg.get_all_data(sim2)
    assert np.array_equal(d.Pos, ref.Pos) and np.array_equal(d.Vel, ref.Vel)
    assert np.array_equal(d.Acc, ref.Acc)
    assert sim2.timeinfo.stepcount == ref_ti.stepcount
    assert sim2.timeinfo.system_time_float == ref_ti.system_time_float
    assert sim2.outputinfo["snapshotcount"] == nsnap
    if adaptive:
        assert np.array_equal(d.Ti_endstep, ref.Ti_endstep)
        assert np.array_equal(d.Ti_begstep, ref.Ti_begstep)
    sim2.close()
    # the resumed run against the FP64 oracle, not only against itself
    if adaptive:
        xo, vo, _, info = orc.run_adaptive(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.04,
                                           time_base=1 << 14, err_tol=1e-3, between_snapshots=0.01)
        assert np.array_equal(d.Ti_endstep, info["ti_end"])
        assert np.array_equal(d.Ti_begstep, info["ti_beg"])
    else:
        xo, vo, _, info = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.04,
                                            dt=2e-3, between_snapshots=0.01)
    assert sim2.timeinfo.stepcount == info["steps"]
    print(f"restart ({'adaptive' if adaptive else 'constant'}): max |dx| vs oracle "
          f"{np.max(np.abs(d.Pos - xo)):.3e}, max |dv| {np.max(np.abs(d.Vel - vo)):.3e}")
    assert np.max(np.abs(d.Pos - xo)) < 1e-11
    assert np.max(np.abs(d.Vel - vo)) < 1e-9
