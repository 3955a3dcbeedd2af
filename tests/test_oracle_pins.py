"""Pins of the FP64 oracle (oracle/ds_oracle.c) -- CPU only.

The reference ships no golden vectors for this path (its tests are closed
forms and cross-layout self-consistency, test/StaticTest.jl), so each oracle
function is pinned by something independent of its own text: closed forms,
the published Gadget-2 kernel with exact fractions, continuity and symmetry
invariants, an mpmath brute force, the reference's hand-written leapfrog
integrator (test/Local-Timestep.jl:54-91), reversibility and dt^2 scaling.
"""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 40


# -- independent formulas -------------------------------------------------------
def spline_fac_exact(r, m, h):
    """Gadget-2 cubic-spline force factor with exact fractions (Springel 2005,
    MNRAS 364, 1105, eq. A2 derivative), mpmath."""
    r, m, h = mp.mpf(r), mp.mpf(m), mp.mpf(h)
    if r >= h:
        return m / r ** 3
    u = r / h
    if u < mp.mpf(1) / 2:
        return m / h ** 3 * (mp.mpf(32) / 3 + u ** 2 * (32 * u - mp.mpf(192) / 5))
    return m / h ** 3 * (mp.mpf(64) / 3 - 48 * u + mp.mpf(192) / 5 * u ** 2
                         - mp.mpf(32) / 3 * u ** 3 - mp.mpf(1) / 15 / u ** 3)


def spline_pot_exact(r, m, h):
    r, m, h = mp.mpf(r), mp.mpf(m), mp.mpf(h)
    if r == 0:
        return mp.mpf(0)
    if r >= h:
        return -m / r
    u = r / h
    if u < mp.mpf(1) / 2:
        wp = -mp.mpf(14) / 5 + u ** 2 * (mp.mpf(16) / 3 + u ** 2 * (mp.mpf(32) / 5 * u - mp.mpf(48) / 5))
    else:
        wp = (-mp.mpf(16) / 5 + mp.mpf(1) / 15 / u
              + u ** 2 * (mp.mpf(32) / 3 + u * (-16 + u * (mp.mpf(48) / 5 - mp.mpf(32) / 15 * u))))
    return m / h * wp


def brute_force_acc(pos, mass, h, G=1.0, skip_self=False):
    n = len(mass)
    out = np.zeros((n, 3))
    for i in range(n):
        acc = [mp.mpf(0)] * 3
        for j in range(n):
            if skip_self and i == j:
                continue
            d = [mp.mpf(pos[j][k]) - mp.mpf(pos[i][k]) for k in range(3)]
            r = mp.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
            if r == 0:
                continue
            f = spline_fac_exact(r, mass[j], h) * G
            acc = [acc[k] + d[k] * f for k in range(3)]
        out[i] = [float(a) for a in acc]
    return out


def brute_force_pot(pos, mass, h, G=1.0):
    n = len(mass)
    out = np.zeros(n)
    for i in range(n):
        s = mp.mpf(0)
        for j in range(n):
            d = [mp.mpf(pos[j][k]) - mp.mpf(pos[i][k]) for k in range(3)]
            r = mp.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
            s += spline_pot_exact(r, mass[j], h) * G
        out[i] = float(s)
    return out


# -- a1 / a2 pair functions ------------------------------------------------------
@pytest.mark.parametrize("r", [0.011, 0.5, 3.0, 1e3])
def test_pair_force_newtonian_outside_h(orc, r):
    p = np.zeros(3)
    s = np.array([r / math.sqrt(3)] * 3)
    a = orc.pair_force(p, s, 2.5, 1.7, 0.01)
    expect = 1.7 * 2.5 / r ** 2
    assert np.linalg.norm(a) == pytest.approx(expect, rel=1e-14)
    assert np.allclose(a / np.linalg.norm(a), s / np.linalg.norm(s), rtol=1e-14)


@pytest.mark.parametrize("u", [1e-3, 0.1, 0.3, 0.4999, 0.5, 0.5001, 0.7, 0.9, 0.9999])
def test_pair_force_spline_matches_exact_fractions(orc, u):
    h = 0.02
    r = u * h
    s = np.array([r, 0.0, 0.0])
    a = orc.pair_force(np.zeros(3), s, 1.3, 1.0, h)
    exact = float(spline_fac_exact(r, 1.3, h)) * r
    # the reference's 12-digit decimals differ from the exact fractions at ~1e-12
    assert a[0] == pytest.approx(exact, rel=5e-11)
    pot = orc.pair_potential(np.zeros(3), s, 1.3, 1.0, h)
    assert pot == pytest.approx(float(spline_pot_exact(r, 1.3, h)), rel=5e-11)


@pytest.mark.parametrize("edge", [0.5, 1.0])
def test_spline_continuity(orc, edge):
    h = 0.5
    lo = orc.pair_force(np.zeros(3), np.array([edge * h * (1 - 1e-13), 0, 0]), 1.0, 1.0, h)[0]
    hi = orc.pair_force(np.zeros(3), np.array([edge * h * (1 + 1e-13), 0, 0]), 1.0, 1.0, h)[0]
    assert lo == pytest.approx(hi, rel=1e-10)
    plo = orc.pair_potential(np.zeros(3), np.array([edge * h * (1 - 1e-13), 0, 0]), 1.0, 1.0, h)
    phi = orc.pair_potential(np.zeros(3), np.array([edge * h * (1 + 1e-13), 0, 0]), 1.0, 1.0, h)
    assert plo == pytest.approx(phi, rel=1e-10)


def test_self_pair(orc):
    p = np.array([0.3, -0.2, 0.1])
    assert np.all(orc.pair_force(p, p, 1.0, 1.0, 0.01) == 0.0)
    assert orc.pair_potential(p, p, 1.0, 1.0, 0.01) == 0.0
    # h == 0: the reference computes 0*Inf (src/directsum/gravity.jl:16-27)
    assert np.all(np.isnan(orc.pair_force(p, p, 1.0, 1.0, 0.0)))
    assert np.isnan(orc.pair_potential(p, p, 1.0, 1.0, 0.0))


def test_pair_antisymmetry(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        p, s = rng.normal(size=3), rng.normal(size=3)
        mi, mj = rng.uniform(0.1, 2, 2)
        h = rng.uniform(0.01, 3.0)
        aij = orc.pair_force(p, s, mj, 1.0, h)
        aji = orc.pair_force(s, p, mi, 1.0, h)
        assert np.allclose(mi * aij, -mj * aji, rtol=1e-13, atol=0)


# -- a3-a6 sums --------------------------------------------------------------------
def test_sums_match_mpmath_bruteforce(orc):
    rng = np.random.default_rng(1)
    n = 24
    pos = rng.normal(scale=0.05, size=(n, 3))
    pos[5] = pos[4] + [0.003, 0.0, 0.0]     # a pair inside h
    pos[7] = pos[6] + [0.0, 0.012, 0.0]     # a pair in the u > 0.5 band
    mass = rng.uniform(0.5, 1.5, n)
    h = 0.02
    ref = brute_force_acc(pos, mass, h, G=1.3)
    for variant in ("faithful", "tuned", "longdouble", "simd"):
        acc = orc.compute_force(pos, mass, [h] * 6, 1.3, variant=variant)
        assert np.allclose(acc, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max()), variant
    pref = brute_force_pot(pos, mass, h, G=1.3)
    for variant in ("faithful", "tuned", "longdouble"):
        pot = orc.compute_potential(pos, mass, [h] * 6, 1.3, variant=variant)
        assert np.allclose(pot, pref, rtol=1e-11), variant


def test_target_side_softening(orc):
    """h comes from the target's collection (src/directsum/gravity.jl:97)."""
    pos = np.array([[0.0, 0.0, 0.0], [0.004, 0.0, 0.0]])
    mass = np.array([1.0, 1.0])
    coll = np.array([1, 5], dtype=np.uint8)
    table = [0.01, 1e-4, 1e-4, 1e-4, 1e-5, 1e-4]
    acc = orc.compute_force(pos, mass, table, 1.0, coll)
    a0 = orc.pair_force(pos[0], pos[1], 1.0, 1.0, 0.01)
    a1 = orc.pair_force(pos[1], pos[0], 1.0, 1.0, 1e-5)
    assert np.array_equal(acc[0], a0) and np.array_equal(acc[1], a1)
    assert abs(acc[1, 0]) == pytest.approx(1.0 / 0.004 ** 2, rel=1e-14)   # Newtonian for particle 2


def test_variants_agree_plummer(pkg, orc):
    pos, vel, mass = pkg.ics.plummer(2000, seed=3)
    ref = orc.compute_force(pos, mass, [0.01] * 6, variant="longdouble")
    for v in ("faithful", "tuned", "simd"):
        a = orc.compute_force(pos, mass, [0.01] * 6, variant=v)
        assert np.max(orc.rel_err(a, ref)) < 1e-12, v


def test_staticcase_binary_closed_forms(pkg, orc):
    """test/StaticTest.jl:18-52,163-207 (unitless: G = 1 here, m = 1e10, D = 2, h = 1e-4)."""
    G, m, D, h = 1.0, 1.0e10, 2.0, 1e-4
    pos, vel, mass = pkg.ics.binary(m, D, G)
    acc = orc.compute_force(pos, mass, [h] * 6, G)
    binary_force = G * m / (D ** 2 + h ** 2) / math.sqrt(D ** 2 + h ** 2) * D
    assert acc[0, 0] + acc[1, 0] == pytest.approx(0.0, abs=1e-6)
    assert abs(acc[0, 0]) == pytest.approx(binary_force, rel=1.5e-8)   # Julia's default rtol sqrt(eps)
    pot = orc.compute_potential(pos, mass, [h] * 6, G)
    K, W, P, L = orc.energy(pos, vel, mass, pot)
    v = math.sqrt(G * m / D / 2.0)
    assert W == pytest.approx(-G * m ** 2 / math.sqrt(D ** 2 + h ** 2), rel=1.5e-8)
    assert K == pytest.approx(m * v ** 2, rel=1e-14)
    assert np.allclose(P, 0.0, atol=1e-6 * m * v)
    assert np.allclose(L, [0.0, 0.0, m * v * D], rtol=1e-14)


def test_staticcase_box_probe(pkg, orc):
    """test/StaticTest.jl:93-123,333-349: force/potential at the origin."""
    pos, _, mass = pkg.ics.box10(1e9)
    h = 1e-4
    pt = np.zeros((1, 3))
    a = orc.force_at_points(pt, h, pos, mass, 1.0)[0]
    ref = np.zeros(3)
    for j in range(10):
        r = np.linalg.norm(pos[j])
        ref += mass[j] * pos[j] / r ** 3
    assert np.allclose(a, ref, rtol=1e-13)
    p = orc.potential_at_points(pt, h, pos, mass, 1.0)[0]
    assert p == pytest.approx(-sum(mass[j] / np.linalg.norm(pos[j]) for j in range(10)), rel=1e-13)


def test_momentum_conservation(pkg, orc):
    pos, vel, mass = pkg.ics.plummer(1500, seed=9, masses=pkg.ics.lognormal_masses(1500, 2))
    acc = orc.compute_force(pos, mass, [0.02] * 6)
    net = (mass[:, None] * acc).sum(0)
    scale = (mass[:, None] * np.abs(acc)).sum(0)
    assert np.all(np.abs(net) < 1e-13 * scale)


def test_plummer_virial_and_potential(pkg, orc):
    """Continuum Plummer: W = -3 pi G M^2 / (32 a); virial 2K/|W| = 1."""
    pos, vel, mass = pkg.ics.plummer(4096, seed=1)
    pot = orc.compute_potential(pos, mass, [0.01] * 6)
    K, W, _, _ = orc.energy(pos, vel, mass, pot)
    assert W == pytest.approx(-3 * math.pi / 32, rel=0.05)
    assert 2 * K / abs(W) == pytest.approx(1.0, rel=0.06)


def test_self_potential_term(pkg, orc):
    """public compute_potential(sim) adds G m_i / (h_i/2.8) (src/base/Potential.jl:5-24):
    the constant is minus the spline's centre value, -W_p(0) = 14/5, taken here
    from the exact-fraction kernel (Springel 2005) at r -> 0, not from the
    oracle's literal."""
    h, G = 0.05, 2.0
    wp0 = float(spline_pot_exact(mp.mpf(h) * mp.mpf("1e-30"), 1, h) * h)   # W_p(0+)
    assert wp0 == pytest.approx(-2.8, rel=1e-15)
    pos, _, mass = pkg.ics.plummer(300, seed=4)
    p0 = orc.compute_potential(pos, mass, [h] * 6, G)
    p1 = orc.compute_potential(pos, mass, [h] * 6, G, self_term=True)
    assert np.allclose(p1 - p0, G * mass * (-wp0) / h, rtol=1e-12)
    # the other branch's constant, 16/5, would fail
    assert not np.allclose(p1 - p0, G * mass * 3.2 / h, rtol=1e-3)


def test_active_filter_and_ranges(pkg, orc):
    pos, _, mass = pkg.ics.plummer(400, seed=5)
    full = orc.compute_force(pos, mass, [0.01] * 6)
    part = orc.compute_force(pos, mass, [0.01] * 6, targets=(100, 250))
    assert np.array_equal(part[100:250], full[100:250]) and np.all(part[:100] == 0)
    active = np.zeros(400, dtype=np.uint8)
    active[::3] = 1
    sel = orc.compute_force(pos, mass, [0.01] * 6, active=active)
    assert np.array_equal(sel[::3], full[::3]) and np.all(sel[1::3] == 0)


def test_zero_softening_skip_self(pkg, orc):
    pos, _, mass = pkg.ics.plummer(64, seed=8)
    acc = orc.compute_force(pos, mass, [0.0] * 6, skip_self=True)
    assert np.all(np.isfinite(acc))
    ref = brute_force_acc(pos, mass, 0.0, skip_self=True)
    assert np.allclose(acc, ref, rtol=1e-12)
    assert np.all(np.isnan(orc.compute_force(pos, mass, [0.0] * 6, skip_self=False)))


# -- a11-a13 integration -------------------------------------------------------------
def test_step_count_float_accumulation(orc):
    """run() stops when t += dt reaches End (src/base/Timestep.jl:198, src/run.jl:158)."""
    assert orc.count_steps(0.0, 0.1, 1e-3) == 100
    assert orc.count_steps(0.0, 50 * 5e-4, 5e-4) == 50
    assert orc.count_steps(0.0, 20 * 1e-3, 1e-3) == 20
    assert orc.count_steps(0.0, 0.2, 1e-3) == 200
    assert orc.count_steps(0.0, 10.0, 1e-3) == 10001


def integrate_const_leapfrog(N, x1, v1, dt, compute_acc):
    """test/Local-Timestep.jl:54-91, restated (state vectors instead of a DataFrame)."""
    x = [x1]
    v = [v1]
    a = [compute_acc(x1)]
    v[0] = v[0] + 0.5 * a[0] * dt
    for i in range(N - 1):
        x.append(x[i] + v[i] * dt)
        a.append(compute_acc(x[i + 1]))
        v.append(v[i] + a[i + 1] * dt)
    v[-1] = v[-1] - 0.5 * a[-1] * dt
    return x[-1], v[-1]


def test_run_loop_matches_hand_integrator(pkg, orc):
    """test/Local-Timestep.jl:402-455: run(sim) vs the hand integrator on a
    circular binary (two-body acceleration from the oracle's own pair
    function for both)."""
    G, m, D, h = 1.0, 1.0, 1.0, 1e-4
    pos, vel, mass = pkg.ics.binary(m, D, G)
    dt, steps = 1e-3, 500
    t_end = steps * dt

    def acc2(x):
        a = np.zeros((2, 3))
        a[0] = orc.pair_force(x[0], x[1], m, G, h)
        a[1] = orc.pair_force(x[1], x[0], m, G, h)
        return a

    nsteps = orc.count_steps(0.0, t_end, dt)
    xh, vh = integrate_const_leapfrog(nsteps + 1, pos.copy(), vel.copy(), dt, acc2)
    xr, vr, _, info = orc.run_const_kdk(pos, vel, mass, [h] * 6, G, t_begin=0.0, t_end=t_end,
                                        dt=dt, between_snapshots=t_end)
    assert info["steps"] == nsteps
    assert np.allclose(xr, xh, rtol=0, atol=1e-12)
    assert np.allclose(vr, vh, rtol=0, atol=1e-12)


def test_time_reversibility(pkg, orc):
    pos, vel, mass = pkg.ics.plummer(200, seed=12)
    soft = [0.05] * 6
    dt = 1e-3
    acc0 = orc.compute_force(pos, mass, soft)
    v = orc.half_kick(vel, acc0, +1, dt)
    p1, v1, a1 = orc.kdk_steps(pos, v, mass, soft, 1.0, nsteps=40, dt=dt)
    v1 = orc.half_kick(v1, a1, -1, dt)
    v1 = -v1
    v = orc.half_kick(v1, a1, +1, dt)
    p2, v2, a2 = orc.kdk_steps(p1, v, mass, soft, 1.0, nsteps=40, dt=dt)
    v2 = orc.half_kick(v2, a2, -1, dt)
    assert np.max(np.abs(p2 - pos)) < 1e-11
    assert np.max(np.abs(-v2 - vel)) < 1e-10


def test_energy_error_scales_dt2(pkg, orc):
    pos, vel, mass = pkg.ics.binary(1.0, 1.0, 1.0)
    vel = vel * 0.8   # eccentric orbit
    soft = [1e-4] * 6

    def energy_err(dt, t_end):
        e0 = orc.energy(pos, vel, mass, orc.compute_potential(pos, mass, soft))
        xr, vr, _, _ = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=t_end,
                                         dt=dt, between_snapshots=t_end)
        e1 = orc.energy(xr, vr, mass, orc.compute_potential(xr, mass, soft))
        E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
        return abs(E1 - E0) / abs(E0)

    e1 = energy_err(4e-3, 0.6)
    e2 = energy_err(2e-3, 0.6)
    assert 2.5 < e1 / e2 < 6.0


def test_snapshot_split_drift(pkg, orc):
    """Snapshots inside a step split the drift; the total drift equals the
    un-split one (CPU semantics, src/base/Timestep.jl:181-215)."""
    pos, vel, mass = pkg.ics.plummer(100, seed=2)
    soft = [0.05] * 6
    times = []
    a, b, _, info = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.02,
                                      dt=3e-3, between_snapshots=0.0045,
                                      on_output=lambda t, p, v: times.append(t))
    c, d, _, info2 = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=0.02,
                                       dt=3e-3, between_snapshots=1.0)
    assert info["steps"] == info2["steps"] == orc.count_steps(0.0, 0.02, 3e-3)
    assert times[0] == 0.0 and len(times) == 1 + 4
    assert np.allclose(a, c, atol=1e-14) and np.allclose(b, d, atol=1e-13)


def test_ensemble_runner_matches_run_loop(pkg, orc):
    """ds_run_ensemble_kdk vs the run-loop restatement, system by system."""
    ens = pkg.ics.planetary_ensemble(nsys=3, seed=6, ntest=8)
    off = ens["offsets"]
    steps, dt = 150, 1e-3
    p, v, E0, E = orc.run_ensemble_kdk(off, ens["pos"], ens["vel"], ens["mass"], 1.0, 0.0, dt,
                                       nsteps=steps, every=50)
    assert E.shape == (3, 3)
    for s in range(3):
        sl = slice(off[s], off[s + 1])
        acc0 = orc.compute_force(ens["pos"][sl], ens["mass"][sl], [0.0] * 6, skip_self=True)
        v0 = orc.half_kick(ens["vel"][sl], acc0, +1, dt)
        xo, vo, ao = orc.kdk_steps(ens["pos"][sl], v0, ens["mass"][sl], [0.0] * 6, 1.0,
                                   nsteps=steps, dt=dt, skip_self=True)
        assert np.allclose(p[sl], xo, rtol=0, atol=1e-13)
        assert np.allclose(v[sl], vo, rtol=0, atol=1e-12)
        pot = orc.compute_potential(xo, ens["mass"][sl], [0.0] * 6, skip_self=True)
        K, W, _, _ = orc.energy(xo, vo - 0.5 * dt * ao, ens["mass"][sl], pot)
        assert E[s, -1] == pytest.approx(K + W, rel=1e-12)


def test_adaptive_timestep_rule(orc):
    """get_timestep: largest TimeBase/2^k not above sqrt(2 eta h/|a|)/TimeInterval
    (src/base/Timestep.jl:1-23), and the StaticTest timestep check
    dt_sim < sqrt(2 eta h/|a|) (test/StaticTest.jl:190-198)."""
    interval = 0.1 / (1 << 30)
    kw = dict(err_tol=0.025, min_step=1e-8, max_step=math.inf, time_interval=interval,
              time_base=1 << 30)
    acc = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 1e-3]])
    h = np.array([1e-4, 1e-2])
    steps = orc.get_timestep(acc, h, **kw)
    for k in range(2):
        dt_theory = math.sqrt(2 * 0.025 * h[k] / np.linalg.norm(acc[k]))
        dt_sim = steps[k] * interval
        assert dt_sim <= dt_theory
        assert dt_theory < 2 * dt_sim or steps[k] == 1 << 30   # capped at TimeBase
        assert steps[k] & (steps[k] - 1) == 0   # a power of two
    with pytest.raises(RuntimeError):
        orc.get_timestep(np.array([[1e12, 0, 0]]), np.array([1e-4]), **kw)


def test_adaptive_run_binary_energy(pkg, orc):
    """Adaptive run of an eccentric binary: block steps shrink near pericentre
    and the energy error stays small (the integrator is symplectic per step)."""
    pos, vel, mass = pkg.ics.binary(1.0, 1.0, 1.0)
    vel = vel * 0.7
    soft = [1e-3] * 6
    x, v, a, info = orc.run_adaptive(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=2.0,
                                     time_base=1 << 20, err_tol=0.002)
    pot0 = orc.compute_potential(pos, mass, soft)
    pot1 = orc.compute_potential(x, mass, soft)
    E0 = sum(orc.energy(pos, vel, mass, pot0)[:2])
    E1 = sum(orc.energy(x, v, mass, pot1)[:2])
    assert abs(E1 - E0) / abs(E0) < 1e-3
    assert info["steps"] > 50 and info["outputs"][-1] == pytest.approx(2.0)
    assert np.all(info["ti_end"] == 1 << 20)


def test_adaptive_reduces_to_constant_kdk(pkg, orc):
    """With MaxStep forcing every particle onto one power-of-two step (the
    timestep criterion never binds), the block-step loop is the constant-dt
    KDK loop: every particle is active at every block step, each kick spans
    (tstart, tend) = one step and the last one half a step
    (src/base/Timestep.jl:47-82,105-114 against :84-133).  On this
    power-of-two timeline both loops drift by exactly dt, so positions agree
    bit for bit; velocities differ only by the rounding of v + a dt - a dt/2
    against v + a dt/2 at the end.  A kick interval off by half a step in
    either loop would move the positions by about a dt^2."""
    pos, vel, mass = pkg.ics.plummer(64, seed=31)
    soft = [0.05] * 6
    dt, t_end = 2.0 ** -10, 0.125
    xc, vc, _, ic = orc.run_const_kdk(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=t_end, dt=dt)
    xa, va, _, ia = orc.run_adaptive(pos, vel, mass, soft, 1.0, t_begin=0.0, t_end=t_end,
                                     time_base=1 << 20, err_tol=1e6, max_step=dt)
    assert ic["steps"] == ia["steps"] == 128
    assert ia["nactive"] == [64] * 128
    assert np.array_equal(xa, xc)
    assert np.max(np.abs(va - vc)) <= 4e-16 * np.max(np.abs(vc))


def compute_dt_pow2(acc, h, err_tol, T, time_base):
    """The timestep criterion sqrt(2 eta h/|a|) of test/Local-Timestep.jl:9-19,
    rounded down onto the power-of-two timeline T/2^k (k <= log2 time_base)
    of the block scheme."""
    dt = math.sqrt(2.0 * err_tol * h / math.sqrt(float(np.dot(acc, acc))))
    k = 0
    while T * 2.0 ** -k > dt and 2 ** k < time_base:
        k += 1
    return T * 2.0 ** -k


def integrate_adapt_leapfrog(x1, v1, compute_acc, compute_dt, t_end):
    """test/Local-Timestep.jl:125-164, restated: kick v by a_{i+1} (dt_i +
    dt_{i+1})/2 between drifts of dt_i, with a forward half kick first and a
    backward half kick last; dt_{i+1} is clipped to the time left, so the last
    one is 0 (the run loop's final step, src/base/Timestep.jl:55-60)."""
    x, v, t = x1.copy(), v1.copy(), 0.0
    a = compute_acc(x)
    dt = min(compute_dt(a), t_end)
    v = v + 0.5 * a * dt
    steps = []
    while t < t_end:
        x = x + v * dt
        t = t + dt
        a = compute_acc(x)
        dt_new = min(compute_dt(a), t_end - t)
        v = v + a * 0.5 * (dt + dt_new)
        steps.append(dt)
        dt = dt_new
    v = v - 0.5 * a * dt
    return x, v, steps


def test_adaptive_matches_variable_step_hand_integrator(orc):
    """One test particle on an eccentric Kepler orbit in a point-mass
    background field: the block-step run loop against the reference's own
    adaptive leapfrog hand integrator (test/Local-Timestep.jl:125-164,
    compared with run(sim) at :618-680), both taking the acceleration from
    the same closed form.  The step changes between 2^-11 and 2^-8 of T along
    the orbit (pericentre near t = 1.2), so a wrong kick interval (tstart/tend of
    src/base/Timestep.jl:64-72) cannot pass."""
    GM, h, eta, T, tb = 1.0, 0.01, 0.01, 2.0, 1 << 20

    def field(p):
        p = np.atleast_2d(p)
        r = np.linalg.norm(p, axis=1)
        return -GM * p / r[:, None] ** 3

    pos = np.array([[1.0, 0.0, 0.0]])
    vel = np.array([[0.0, 0.4, 0.03]])
    mass = np.array([1e-12])
    xa, va, _, info = orc.run_adaptive(pos, vel, mass, [h] * 6, 1.0, t_begin=0.0, t_end=T,
                                       time_base=tb, err_tol=eta,
                                       post=lambda p, a: a + field(p))
    xh, vh, steps = integrate_adapt_leapfrog(pos[0], vel[0], lambda x: field(x)[0],
                                             lambda a: compute_dt_pow2(a, h, eta, T, tb), T)
    assert info["steps"] == len(steps)
    assert len(set(steps)) >= 3                   # the step does change along the orbit
    assert np.max(np.abs(xa[0] - xh)) < 1e-12
    assert np.max(np.abs(va[0] - vh)) < 1e-11


def test_mond_nu_inverts_the_mu_functions(orc):
    """g = nu(gN/a0) gN must satisfy Milgrom's law gN = mu(g/a0) g with the
    'simple' mu(x) = x/(1+x) for index 1 and the 'standard' mu(x) = x/sqrt(1+x^2)
    for index 2 (the closed forms src/mond/milgrom1983.jl:13-26 name)."""
    gN = np.logspace(-6, 6, 200)
    for n, mu in ((1.0, lambda x: x / (1.0 + x)), (2.0, lambda x: x / np.sqrt(1.0 + x * x))):
        g = orc.mond_nu(gN, n) * gN
        assert np.allclose(mu(g) * g, gN, rtol=1e-12, atol=0)
    # Newtonian limit nu -> 1 and deep-MOND limit nu -> y^-1/2
    assert orc.mond_nu(1e12, 2.0) == pytest.approx(1.0, abs=1e-12)
    assert orc.mond_nu(1e-12, 3.0) * 1e-6 == pytest.approx(1.0, rel=1e-5)


def test_mond_point_mass_flat_rotation(orc):
    """Deep-MOND point mass: v_c^4 = G M a0 independently of r."""
    G, M, a0 = 1.0, 1.0, 1e-2
    r = np.array([1e3, 1e4, 1e5])
    pos = np.zeros((4, 3))
    pos[1:, 0] = r
    mass = np.array([M, 0.0, 0.0, 0.0])
    acc = orc.compute_force(pos, mass, [1e-6] * 6, G)
    am = orc.mond(acc, a0, 2.0)
    assert np.all(am[0] == 0.0)                  # massless sources: the centre feels nothing
    v4 = (np.linalg.norm(am[1:], axis=1) * r) ** 2
    assert np.allclose(v4, G * M * a0, rtol=1e-3)


@pytest.mark.parametrize("kind,params", [
    (1, [2.0, 0.5, -1.0, 0.2]), (2, [2.0, 0.5, -1.0, 0.2, 0.3]),
    (3, [0.1, 0.5, -1.0, 0.2, 0.7, -0.2, 0.05]), (4, [0.3, -0.2, 0.9])])
def test_background_fields_are_gradients(orc, kind, params):
    """a = -grad(phi) by central differences, for every field of the menu."""
    rng = np.random.default_rng(kind)
    pos = rng.normal(size=(50, 3)) * 2.0 + np.array([3.0, 0.0, 0.0])
    acc, _ = orc.background_field(pos, kind, params)
    hstep = 1e-5
    num = np.zeros_like(pos)
    for d in range(3):
        e = np.zeros(3)
        e[d] = hstep
        num[:, d] = -(orc.background_field(pos + e, kind, params)[1]
                      - orc.background_field(pos - e, kind, params)[1]) / (2 * hstep)
    assert np.allclose(acc, num, rtol=1e-6, atol=1e-9)


def test_background_fields_closed_forms(orc):
    rng = np.random.default_rng(4)
    pos = rng.normal(size=(40, 3)) * 3.0
    # Paczynski-Wiita with r_g = 0 is the point mass of the C oracle
    acc_pm, pot_pm = np.zeros((40, 3)), np.zeros(40)
    orc.background_point_mass(pos, 1.5, np.array([0.2, 0.1, -0.3]), acc=acc_pm, pot=pot_pm)
    a2, p2 = orc.background_field(pos, 2, [1.5, 0.2, 0.1, -0.3, 0.0])
    assert np.allclose(a2, acc_pm, rtol=1e-13) and np.allclose(p2, pot_pm, rtol=1e-13)
    # radial polynomial with only c1 = k is the harmonic field a = -k (x - c)
    a3, p3 = orc.background_field(pos, 3, [0.0, 1.0, 2.0, 3.0, 0.7, 0.0, 0.0])
    d = pos - np.array([1.0, 2.0, 3.0])
    assert np.allclose(a3, -0.7 * d, rtol=1e-13)
    assert np.allclose(p3, 0.35 * np.sum(d * d, axis=1), rtol=1e-13)
