"""GPU parity of the batched ensemble engine against the FP64 oracle, system
by system (needs an MI355X).  Gate: median per-particle force relative error
<= 1e-6 in the FP32-pair modes, <= 1e-12 in FP64 mode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GATE = 1e-6


@pytest.fixture(scope="module")
def clusters(pkg):
    return pkg.ics.cluster_ensemble(nsys=8, seed=5, nmin=200, nmax=3000)


def oracle_forces(orc, ens, G=1.0, skip_self=False, pot=False):
    off = ens["offsets"]
    acc = np.zeros_like(ens["pos"])
    phi = np.zeros(len(ens["mass"]))
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        h = ens["h"][s] if "h" in ens else 0.0
        acc[sl] = orc.compute_force(ens["pos"][sl], ens["mass"][sl], [h] * 6, G, skip_self=skip_self)
        if pot:
            phi[sl] = orc.compute_potential(ens["pos"][sl], ens["mass"][sl], [h] * 6, G,
                                            skip_self=skip_self)
    return acc, phi


def make(pkg, ens, mode="split", dt=1e-3, **tun):
    e = pkg.Ensemble(ens["offsets"], mode)
    e.set_params(1.0, ens.get("h", 0.0), dt)
    e.upload(ens["pos"], ens["vel"], ens["mass"])
    if tun:
        e.set_tuning(**tun)
    return e


@pytest.mark.parametrize("mode,tol,tpt,far", [("split", GATE, 1, 0), ("split", GATE, 2, 0),
                                               ("split", GATE, 4, 0), ("split", GATE, 1, 1),
                                               ("split", GATE, 1, 2), ("split", GATE, 1, 3),
                                               ("split", GATE, 2, 3), ("split", GATE, 4, 3),
                                               ("split", GATE, 1, 4),
                                               ("fp64", 1e-12, 1, 0)])
def test_ragged_force(pkg, orc, clusters, mode, tol, tpt, far):
    ref, _ = oracle_forces(orc, clusters)
    with make(pkg, clusters, mode, targets_per_thread=tpt, blocks=37, resident=0, far_path=far) as e:
        e.compute_force()
        acc = e.download(pos=False, vel=False)["acc"]
    err = orc.rel_err(acc, ref)
    off = clusters["offsets"]
    for s in range(len(off) - 1):
        assert np.median(err[off[s]:off[s + 1]]) <= tol, s


def test_ragged_queue_reused_across_launches(pkg, orc, clusters):
    ref, _ = oracle_forces(orc, clusters)
    with make(pkg, clusters, "split", blocks=5, resident=0) as e:
        for _ in range(4):     # the queue is reset by the last block of each launch
            e.compute_force()
        acc = e.download(pos=False, vel=False)["acc"]
    assert np.median(orc.rel_err(acc, ref)) <= GATE


def test_batch_matches_plain_contexts(pkg, orc, clusters):
    nat = pkg._native
    with make(pkg, clusters, "split", resident=0) as e:
        e.compute_force()
        e.half_kick(+1)
        e.kdk_steps(8, use_graph=False)
        d = e.download()
    off = clusters["offsets"]
    for s in (0, 3, 7):
        sl = slice(off[s], off[s + 1])
        with nat.Context(off[s + 1] - off[s], nat.NB_MODE_SPLIT) as c:
            c.set_params(1.0, [clusters["h"][s]] * 6)
            c.upload(clusters["pos"][sl], clusters["vel"][sl], clusters["mass"][sl])
            c.compute_force()
            c.half_kick(+1, 1e-3)
            c.kdk_steps(8, 1e-3, 0.0)
            p = c.download()
        assert np.max(np.abs(p["pos"] - d["pos"][sl])) < 1e-9
        assert np.median(orc.rel_err(p["acc"], d["acc"][sl])) < 1e-6


def test_graph_and_eager_steps_identical(pkg, clusters):
    outs = []
    for use_graph in (False, True):
        with make(pkg, clusters, "split", steps_per_graph=4, resident=0, sort_interval=0) as e:
            e.compute_force()
            e.half_kick(+1)
            for _ in range(3):
                e.kdk_steps(8, use_graph=use_graph)
                e.sync()
            outs.append(e.download())
    assert np.array_equal(outs[0]["pos"], outs[1]["pos"])
    assert np.array_equal(outs[0]["vel"], outs[1]["vel"])


def test_batch_energy(pkg, orc, clusters):
    ref, phi = oracle_forces(orc, clusters, pot=True)
    with make(pkg, clusters, "split", resident=0) as e:
        en = e.energy()
    off = clusters["offsets"]
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        K, W, P, L = orc.energy(clusters["pos"][sl], clusters["vel"][sl], clusters["mass"][sl],
                                phi[sl])
        assert en["K"][s] == pytest.approx(K, rel=1e-12)
        assert en["W"][s] == pytest.approx(W, rel=1e-6)


def _degenerate(pkg, sizes, seed=31):
    """An ensemble with empty, one- and two-particle systems among larger ones."""
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = np.zeros((off[-1], 3))
    vel = np.zeros((off[-1], 3))
    mass = np.zeros(off[-1])
    for s, n in enumerate(sizes):
        if n:
            p, v, m = pkg.ics.plummer(int(n), seed=seed + s)
            pos[off[s]:off[s + 1]], vel[off[s]:off[s + 1]], mass[off[s]:off[s + 1]] = p, v, m
    return {"offsets": off, "pos": pos, "vel": vel, "mass": mass, "h": np.full(len(sizes), 0.01)}


@pytest.mark.parametrize("sizes,tun", [
    ([0, 1, 2, 257, 0, 600, 1, 255], dict(resident=0)),
    ([0, 1, 2, 257, 0, 600, 1, 255], dict(resident=0, targets_per_thread=2)),
    ([0, 1, 2, 257, 0, 600, 1, 255], dict(resident=0, far_path=4)),
    ([0, 1, 2, 255, 0, 73, 1], dict(resident=1))])
def test_empty_and_tiny_systems(pkg, orc, sizes, tun):
    """Systems of 0, 1 and 2 particles in the ragged and the resident launch:
    forces match the oracle system by system, a lone particle feels nothing,
    and KDK steps and energies stay finite."""
    ens = _degenerate(pkg, sizes)
    ref, _ = oracle_forces(orc, ens)
    with make(pkg, ens, "split", **tun) as e:
        e.compute_force()
        acc = e.download(pos=False, vel=False)["acc"]
        e.half_kick(+1)
        e.kdk_steps(8)
        e.half_kick(-1)
        d = e.download(pos=True, vel=True, acc=True)
        en = e.energy()
    off = ens["offsets"]
    for s, n in enumerate(sizes):
        sl = slice(off[s], off[s + 1])
        if n == 1:
            assert np.all(acc[sl] == 0.0)
        elif n > 1:
            assert np.median(orc.rel_err(acc[sl], ref[sl])) <= GATE, s
        if n == 0:
            assert en["K"][s] == 0.0 and en["W"][s] == 0.0
    assert all(np.isfinite(v).all() for v in d.values())
    assert np.isfinite(en["K"]).all() and np.isfinite(en["W"]).all()


@pytest.mark.parametrize("tpt,far", [(1, 0), (2, 0), (1, 4)])
def test_potential_alone(pkg, orc, clusters, tpt, far):
    """compute_potential writes the potential alone (with far path 4 on the
    matrix cores: m/r = rsq(w r^2) m^(2/3)); the accelerations stay as the
    last force left them."""
    ref_acc, phi = oracle_forces(orc, clusters, pot=True)
    with make(pkg, clusters, "split", resident=0, targets_per_thread=tpt, far_path=far) as e:
        e.compute_force()
        acc0 = e.download(pos=False, vel=False)["acc"]
        e.compute_potential()
        d = e.download(pos=False, vel=False, acc=True, pot=True)
    assert np.array_equal(d["acc"], acc0)
    off = clusters["offsets"]
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        ep = np.abs(d["pot"][sl] - phi[sl]) / np.abs(phi[sl])
        assert np.median(ep) <= GATE and ep.max() <= 1e-5, s


def test_potential_alone_with_central_bodies(pkg, orc):
    """split_central ensembles: the FP64 central pass adds the star's
    potential and leaves the accelerations alone."""
    ens = pkg.ics.planetary_ensemble(nsys=6, seed=6)
    _, phi = oracle_forces(orc, ens, skip_self=True, pot=True)
    with make(pkg, ens, "split_central", resident=0) as e:
        e.set_central(1)
        e.compute_force()
        acc0 = e.download(pos=False, vel=False)["acc"]
        e.compute_potential()
        d = e.download(pos=False, vel=False, acc=True, pot=True)
    assert np.array_equal(d["acc"], acc0)
    massive = np.abs(phi) > 0
    ep = np.abs(d["pot"][massive] - phi[massive]) / np.abs(phi[massive])
    print(f"planetary potential: median {np.median(ep):.3e} max {ep.max():.3e}")
    assert np.median(ep) <= 1e-9 and ep.max() <= 1e-5


def test_energy_reports_nonfinite_system(pkg, clusters):
    """A NaN velocity in system 3: nb_batch_energy writes every system's values,
    returns NB_ERR_NONFINITE naming system 3, and the Python host warns."""
    vel = clusters["vel"].copy()
    off = clusters["offsets"]
    vel[off[3] + 5, 0] = np.nan
    ens = dict(clusters, vel=vel)
    with make(pkg, ens, "split", resident=0) as e:
        with pytest.warns(RuntimeWarning, match="system 3"):
            en = e.energy()
    assert np.isnan(en["K"][3]) and np.isfinite(np.delete(en["K"], 3)).all()
    assert np.isfinite(en["W"]).all()


def test_more_systems_than_the_sort_key_once_held(pkg, orc):
    """9000 systems on the ragged (Morton-sorted) path: the sort key keeps
    the system index above the Morton code for any number of systems, so
    every particle stays in its own system (13 key bits held 8192)."""
    rng = np.random.default_rng(41)
    sizes = rng.integers(2, 41, 9000)
    ens = _degenerate(pkg, list(sizes), seed=1000)
    ref, _ = oracle_forces(orc, ens)
    with make(pkg, ens, "split", resident=0, sort_interval=1) as e:
        e.compute_force()
        d = e.download(pos=True, vel=False, acc=True)
    assert np.array_equal(d["pos"], ens["pos"])
    off = ens["offsets"]
    worst = max(np.median(orc.rel_err(d["acc"][off[s]:off[s + 1]], ref[off[s]:off[s + 1]]))
                for s in range(len(sizes)))
    print(f"9000 systems: worst per-system median {worst:.3e}")
    assert worst <= GATE


@pytest.fixture(scope="module")
def planets(pkg):
    return pkg.ics.planetary_ensemble(nsys=6, seed=6)


def test_planetary_force_central_fp64(pkg, orc, planets):
    ref, _ = oracle_forces(orc, planets, skip_self=True)
    for resident in (0, 1):
        with make(pkg, planets, "split_central", resident=resident) as e:
            e.set_central(1)
            e.compute_force()
            acc = e.download(pos=False, vel=False)["acc"]
        err = orc.rel_err(acc, ref)
        assert np.max(err) < 1e-6 and np.median(err) < 1e-9


def test_resident_matches_oracle_run(pkg, orc, planets):
    """Resident multi-step kernel vs the oracle's KDK run loop, 300 steps."""
    steps, dt = 300, 1e-3
    with make(pkg, planets, "split_central", dt=dt, resident=1) as e:
        e.set_central(1)
        e.compute_force()
        e.half_kick(+1)
        e.kdk_steps(steps, use_graph=True)
        e.half_kick(-1)
        d = e.download()
    off = planets["offsets"]
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        xo, vo, _, info = orc.run_const_kdk(planets["pos"][sl], planets["vel"][sl],
                                            planets["mass"][sl], [0.0] * 6, 1.0, t_begin=0.0,
                                            t_end=steps * dt, dt=dt, between_snapshots=1.0,
                                            skip_self=True)
        assert info["steps"] == steps
        assert np.max(np.abs(d["pos"][sl] - xo)) < 1e-9
        assert np.max(np.abs(d["vel"][sl] - vo)) < 1e-8


def test_resident_after_sorted_ragged_force(pkg, orc, planets):
    """An ensemble Morton-sorted by the ragged path (resident = 0) and then
    switched to the resident kernel goes back to the caller's order first:
    the resident kernel takes the central body by its position in the system,
    so the FP64 central terms (and the oracle agreement) survive the switch."""
    steps, dt = 100, 1e-3
    with make(pkg, planets, "split_central", dt=dt, resident=0, sort_interval=16) as e:
        e.set_central(1)
        e.compute_force()                    # sorts every system (ragged path)
        e.set_tuning(resident=1, sort_interval=16)
        e.half_kick(+1)
        e.kdk_steps(steps, use_graph=True)   # resident kernel
        e.half_kick(-1)
        d = e.download()
    off = planets["offsets"]
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        xo, vo, _, _ = orc.run_const_kdk(planets["pos"][sl], planets["vel"][sl],
                                         planets["mass"][sl], [0.0] * 6, 1.0, t_begin=0.0,
                                         t_end=steps * dt, dt=dt, between_snapshots=1.0,
                                         skip_self=True)
        assert np.max(np.abs(d["pos"][sl] - xo)) < 1e-9
        assert np.max(np.abs(d["vel"][sl] - vo)) < 1e-8


def test_resident_and_ragged_paths_agree(pkg, planets):
    outs = []
    for resident in (1, 0):
        with make(pkg, planets, "split_central", resident=resident) as e:
            e.set_central(1)
            e.compute_force()
            e.half_kick(+1)
            e.kdk_steps(40, use_graph=True)
            outs.append(e.download())
    assert np.max(np.abs(outs[0]["pos"] - outs[1]["pos"])) < 1e-10


def _mixed_planetary(pkg):
    """Planetary systems of different shapes in one ensemble: the C5 system
    (star, 8 planets, 64 test particles), one of 201 bodies whose massive and
    massless bodies are interleaved (indices 1..200 shuffled), one of 609
    bodies (more than 256: its test particles go to three or four waves), a
    star with three planets and no test particles, a star with test particles
    only, a system of massless particles only, and a lone star."""
    shapes = [(8, 64, False), (20, 180, True), (8, 600, True), (3, 0, False), (0, 40, False),
              (0, 10, False), (0, 0, False)]
    rng = np.random.default_rng(31)
    parts = []
    for k, (npl, nt, shuffle) in enumerate(shapes):
        e = pkg.ics.planetary_ensemble(nsys=1, seed=40 + k, nplanets=npl, ntest=nt)
        pos, vel, mass = e["pos"], e["vel"], e["mass"]
        if shuffle:
            idx = np.concatenate([[0], 1 + rng.permutation(len(mass) - 1)])
            pos, vel, mass = pos[idx], vel[idx], mass[idx]
        if (npl, nt) == (0, 10):
            mass = np.zeros_like(mass)
        parts.append((pos, vel, mass))
    off = np.concatenate([[0], np.cumsum([len(p[2]) for p in parts])]).astype(np.int64)
    return {"offsets": off, "pos": np.concatenate([p[0] for p in parts]),
            "vel": np.concatenate([p[1] for p in parts]), "mass": np.concatenate([p[2] for p in parts])}


@pytest.mark.parametrize("waves", [0, 1, 2, 3, 4])
def test_resident_wave_layouts(pkg, orc, waves):
    """The resident kernels with every layout (0: several systems per
    workgroup, massive and massless particles in separate waves; 1-4: waves
    per system, each holding the massive bodies and a share of the massless
    ones, a count whose lanes do not fit falling back to the choice): 200
    steps against the oracle's KDK run loop, system by system, and
    bit-identical to one wave per system, since each particle's arithmetic
    does not depend on the layout (zero-mass pad records add exact zeros)."""
    ens = _mixed_planetary(pkg)
    steps, dt = 200, 1e-3

    def run(w):
        with make(pkg, ens, "split_central", dt=dt, resident=1, resident_waves=w) as e:
            e.set_central(1)
            e.compute_force()
            e.half_kick(+1)
            e.kdk_steps(steps, use_graph=True)
            e.half_kick(-1)
            return e.download()

    d = run(waves)
    off = ens["offsets"]
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        xo, vo, _, _ = orc.run_const_kdk(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl], [0.0] * 6,
                                         1.0, t_begin=0.0, t_end=steps * dt, dt=dt,
                                         between_snapshots=1.0, skip_self=True)
        assert np.max(np.abs(d["pos"][sl] - xo)) < 1e-9, s
        assert np.max(np.abs(d["vel"][sl] - vo)) < 1e-8, s
    if waves != 1:
        d1 = run(1)
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(d[k], d1[k]), k


@pytest.mark.parametrize("mode", ["split", "split_central"])
def test_resident_heavy_binary(pkg, orc, mode):
    """The StaticTest binary (masses 1e10, test/StaticTest.jl:18-123) in the
    resident kernel without softening, beside a planetary system: the r^2
    offset that keeps the self pair finite scales with the largest mass, and
    both systems follow the oracle's run loop for 100 steps."""
    bp, bv, bm = pkg.ics.binary()
    pl = pkg.ics.planetary_ensemble(nsys=1, seed=9)
    ens = {"offsets": np.array([0, 2, 2 + len(pl["mass"])], dtype=np.int64),
           "pos": np.concatenate([bp, pl["pos"]]), "vel": np.concatenate([bv, pl["vel"]]),
           "mass": np.concatenate([bm, pl["mass"]])}
    steps, dt = 100, 2.0 ** -20   # exact in binary: the oracle's t += dt loop runs 100 steps
    with make(pkg, ens, mode, dt=dt, resident=1) as e:
        if mode == "split_central":
            e.set_central(1)
        e.compute_force()
        e.half_kick(+1)
        e.kdk_steps(steps, use_graph=True)
        e.half_kick(-1)
        d = e.download()
    assert all(np.isfinite(v).all() for v in d.values())
    off = ens["offsets"]
    # the binary's pair is FP32 (hi/lo) at least one way: a relative force
    # error of a few 1e-8 over 0.8 orbits; the planetary system as before
    for s, tol in ((0, 1e-5), (1, 1e-9)):
        sl = slice(off[s], off[s + 1])
        xo, vo, _, info = orc.run_const_kdk(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl], [0.0] * 6,
                                            1.0, t_begin=0.0, t_end=steps * dt, dt=dt,
                                            between_snapshots=1.0, skip_self=True)
        assert info["steps"] == steps
        assert np.max(np.abs(d["pos"][sl] - xo)) <= tol * np.max(np.abs(xo)), s
        assert np.max(np.abs(d["vel"][sl] - vo)) <= tol * np.max(np.abs(vo)), s


def test_tracked_energy_wave_layouts(pkg):
    """Tracked energies of the mixed planetary shapes agree between the
    packed layout (0) and one or two waves per system; the system of
    massless particles has zero energy in every layout."""
    ens = _mixed_planetary(pkg)
    steps, every, dt = 300, 100, 1e-3
    Es = []
    for w in (0, 1, 2):
        with make(pkg, ens, "split_central", dt=dt, resident=1, resident_waves=w) as e:
            e.set_central(1)
            e.compute_force()
            e.half_kick(+1)
            Es.append(e.kdk_steps_tracked(steps, every))
    nsys = len(ens["offsets"]) - 1
    zero = [s for s in range(nsys) if not ens["mass"][ens["offsets"][s]:ens["offsets"][s + 1]].any()]
    assert zero and Es[0].shape == (nsys, 3)
    for E in Es:
        assert np.all(E[zero] == 0.0)
    # The trajectories are identical; the energies are summed in different
    # orders.  (A lone star at rest has exactly zero energy: zero stays zero.)
    for E in Es[1:]:
        assert np.all(np.abs(E - Es[0]) <= 1e-13 * np.abs(Es[0]))


def test_packed_resident_mixed_shapes(pkg, orc):
    """1201 planetary systems of four shapes (3 or 8 planets, 0 to 100 test
    particles) in the packed layout: several systems per workgroup (5 on 256
    CUs), sorted by massive count, so the workgroup that holds the last
    4-body system (1200) and the first 9-body one (1) pads the former's
    records to the latter's.  100 steps of 20 sampled systems (those two
    among them) against the oracle's run loop, and tracked energies against
    one wave per system."""
    shapes = [(3, 0), (8, 64), (3, 100), (8, 20)]
    parts = []
    for k in range(1201):
        npl, nt = shapes[k % 4]
        e = pkg.ics.planetary_ensemble(nsys=1, seed=1000 + k, nplanets=npl, ntest=nt)
        parts.append((e["pos"], e["vel"], e["mass"]))
    off = np.concatenate([[0], np.cumsum([len(p[2]) for p in parts])]).astype(np.int64)
    ens = {"offsets": off, "pos": np.concatenate([p[0] for p in parts]),
           "vel": np.concatenate([p[1] for p in parts]), "mass": np.concatenate([p[2] for p in parts])}
    steps, every, dt = 100, 50, 1e-3
    out = {}
    for w in (0, 1):
        with make(pkg, ens, "split_central", dt=dt, resident=1, resident_waves=w) as e:
            e.set_central(1)
            e.compute_force()
            e.half_kick(+1)
            E = e.kdk_steps_tracked(steps, every)
            e.half_kick(-1)
            out[w] = (e.download(), E)
    d, E0 = out[0]
    assert np.max(np.abs(out[1][1] - E0) / np.abs(E0)) < 1e-13   # (summation order)
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(d[k], out[1][0][k]), k
    rng = np.random.default_rng(3)
    sample = sorted(set(rng.choice(1201, 18, replace=False).tolist()) | {1200, 1})
    for s in sample:
        sl = slice(off[s], off[s + 1])
        xo, vo, _, _ = orc.run_const_kdk(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl], [0.0] * 6,
                                         1.0, t_begin=0.0, t_end=steps * dt, dt=dt,
                                         between_snapshots=1.0, skip_self=True)
        assert np.max(np.abs(d["pos"][sl] - xo)) < 1e-9, s
        assert np.max(np.abs(d["vel"][sl] - vo)) < 1e-8, s


def test_packed_resident_mixed_central_counts(pkg, orc):
    """Two central bodies per system (set_central(2): the star and particle
    1), where a quarter of the systems have a massless test particle at index
    1 and so one central body.  Sorted by central count, the workgroup that
    holds the last one-central system (1203) and the first two-central ones
    (0, 2, ...) pads the former's FP64 slots with a zero record.  Every system
    is bit-identical to one wave per system (which has no padding), and a
    sample follows the oracle's run loop."""
    shapes = [(3, 0), (8, 64), (3, 100), (8, 20)]
    parts = []
    for k in range(1204):
        npl, nt = shapes[k % 4]
        e = pkg.ics.planetary_ensemble(nsys=1, seed=3000 + k, nplanets=npl, ntest=nt)
        pos, vel, mass = e["pos"], e["vel"], e["mass"]
        if k % 4 == 3:   # a test particle at index 1
            idx = np.arange(len(mass))
            idx[[1, 10]] = idx[[10, 1]]
            pos, vel, mass = pos[idx], vel[idx], mass[idx]
        parts.append((pos, vel, mass))
    off = np.concatenate([[0], np.cumsum([len(p[2]) for p in parts])]).astype(np.int64)
    ens = {"offsets": off, "pos": np.concatenate([p[0] for p in parts]),
           "vel": np.concatenate([p[1] for p in parts]), "mass": np.concatenate([p[2] for p in parts])}
    steps, dt = 100, 1e-3
    out = {}
    for w in (0, 1):
        with make(pkg, ens, "split_central", dt=dt, resident=1, resident_waves=w) as e:
            e.set_central(2)
            e.compute_force()
            e.half_kick(+1)
            e.kdk_steps(steps, use_graph=True)
            e.half_kick(-1)
            out[w] = e.download()
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(out[0][k], out[1][k]), k
    for s in (0, 1, 2, 3, 1199, 1203):
        sl = slice(off[s], off[s + 1])
        xo, vo, _, _ = orc.run_const_kdk(ens["pos"][sl], ens["vel"][sl], ens["mass"][sl], [0.0] * 6,
                                         1.0, t_begin=0.0, t_end=steps * dt, dt=dt,
                                         between_snapshots=1.0, skip_self=True)
        assert np.max(np.abs(out[0]["pos"][sl] - xo)) < 1e-9, s
        assert np.max(np.abs(out[0]["vel"][sl] - vo)) < 1e-8, s


@pytest.mark.parametrize("mode", ["split", "fp64"])
def test_packed_resident_softened_clusters(pkg, orc, mode):
    """1100 softened Plummer clusters of 8 to 64 bodies (h = 0.01 a, every
    body massive) in the packed layout (several clusters per massive wave),
    with one cluster in four given 30 massless test particles: the softened
    branch of the resident kernel (the spline inside h, in FP32 pairs for
    split precision and in FP64 for the FP64 mode).  Bit-identical to one
    wave per system; a sample follows the oracle's softened run loop for
    100 steps, and the tracked energies agree."""
    ens = pkg.ics.cluster_ensemble(nsys=1100, seed=17, nmin=8, nmax=64)
    off, pos, vel, mass = ens["offsets"], ens["pos"], ens["vel"], ens["mass"]
    rng = np.random.default_rng(4)
    parts, hs = [], []
    for s in range(len(off) - 1):
        sl = slice(off[s], off[s + 1])
        p, v, m = pos[sl], vel[sl], mass[sl]
        if s % 4 == 1:
            tp = rng.normal(0.0, ens["a"][s], (30, 3))
            p, v, m = np.vstack([p, tp]), np.vstack([v, np.zeros((30, 3))]), np.concatenate([m, np.zeros(30)])
        parts.append((p, v, m))
    off2 = np.concatenate([[0], np.cumsum([len(q[2]) for q in parts])]).astype(np.int64)
    e2 = {"offsets": off2, "pos": np.concatenate([q[0] for q in parts]),
          "vel": np.concatenate([q[1] for q in parts]), "mass": np.concatenate([q[2] for q in parts]),
          "h": ens["h"]}
    steps, every, dt = 100, 50, 1e-3
    out = {}
    for w in (0, 1):
        with make(pkg, e2, mode, dt=dt, resident=1, resident_waves=w) as e:
            e.compute_force()
            e.half_kick(+1)
            E = e.kdk_steps_tracked(steps, every)
            e.half_kick(-1)
            out[w] = (e.download(), E)
    d = out[0][0]
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(d[k], out[1][0][k]), k
    assert np.all(np.abs(out[1][1] - out[0][1]) <= 1e-13 * np.abs(out[0][1]))
    # FP32 pairs: a relative force error of about 1e-7 moves a body by about
    # 0.5e-7 a t^2 ~ 5e-10 in 0.1 time units; close encounters amplify it
    tol_x, tol_v = (1e-8, 1e-7) if mode == "split" else (1e-12, 1e-11)
    for s in (0, 1, 2, 5, 500, 1099):
        sl = slice(off2[s], off2[s + 1])
        xo, vo, _, _ = orc.run_const_kdk(e2["pos"][sl], e2["vel"][sl], e2["mass"][sl], [ens["h"][s]] * 6,
                                         1.0, t_begin=0.0, t_end=steps * dt, dt=dt, between_snapshots=1.0)
        assert np.max(np.abs(d["pos"][sl] - xo)) < tol_x, s
        assert np.max(np.abs(d["vel"][sl] - vo)) < tol_v, s


def test_tracked_energy(pkg, orc, planets):
    steps, every, dt = 400, 100, 1e-3
    with make(pkg, planets, "split_central", dt=dt, resident=1) as e:
        e.set_central(1)
        e.compute_force()
        e0 = e.energy()
        E0 = e0["K"] + e0["W"]
        e.half_kick(+1)
        E = e.kdk_steps_tracked(steps, every)
    assert E.shape == (6, 4)
    rel = np.abs(E - E0[:, None]) / np.abs(E0[:, None])
    assert np.all(rel < 1e-4)
