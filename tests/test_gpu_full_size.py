"""The benchmark configurations at their full sizes (needs an MI355X).

The FP64 oracle cannot sum all N^2 pairs of these in a test's time, so each
check is either a seeded sample of targets against every source, or a
property that holds at any size:

* two-sphere collision, N = 2^20, split precision (the benchmark): a sample of
  1024 targets against the oracle (median <= 2e-7, p99 <= 2e-6 relative per
  particle), momentum balance sum_i m_i a_i ~ 0 (equal softening, so the
  pair forces are antisymmetric; relative to sum_i m_i |a_i|, <= 1e-6), and
  time reversal of the KDK step (2 steps forward, velocities flipped, 2
  steps back: positions within 1e-9, velocities within 1e-6, absolute), and
  the potential of 512 sampled targets (median <= 1e-6);
* the same collision at N = 2^22 (four times the benchmark): 256 sampled
  targets and the momentum balance;
* Plummer sphere, N = 65536, FP32 state (the FP32 configuration): a sample of
  2048 targets against the oracle on the FP32-rounded inputs (median <= 1e-6);
* the two ensembles at their full sizes: eight sampled clusters of the
  512-cluster ensemble, and all 2048 planetary systems (the oracle sums each
  73-body system whole).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(pkg):
    pos, vel, mass = pkg.ics.collision(1 << 19, seeds=(3, 4))
    return pos, vel, mass


def _ctx(pkg, pos, vel, mass, h, mode):
    nat = pkg._native
    ctx = nat.Context(len(mass), nat.MODES[mode])
    ctx.set_params(1.0, [h] * 6)
    ctx.upload(pos, vel, mass)
    return ctx


def test_collision_full_size_sampled_parity_and_momentum(pkg, orc, c3):
    pos, vel, mass = c3
    n, h = len(mass), 0.002
    assert n == 1 << 20
    ctx = _ctx(pkg, pos, vel, mass, h, "split")
    try:
        ctx.compute_force()
        acc = ctx.download(pos=False, vel=False)["acc"]
        tuning = ctx.get_tuning()
    finally:
        ctx.close()
    rng = np.random.default_rng(21)
    sample = np.sort(rng.choice(n, 1024, replace=False))
    ref = orc.force_at_points(pos[sample], h, pos, mass, variant="simd")
    e = orc.rel_err(acc[sample], ref)
    print(f"N=2^20 {tuning}: median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e} max {e.max():.3e}")
    assert np.median(e) <= 2e-7 and np.percentile(e, 99) <= 2e-6
    ma = mass[:, None] * acc
    bal = np.linalg.norm(ma.sum(axis=0)) / np.linalg.norm(ma, axis=1).sum()
    print(f"momentum balance {bal:.3e}")
    assert bal <= 1e-6


def test_collision_4x_size_sampled_parity(pkg, orc):
    """Four times the benchmark, N = 2^22 (no table entry: the default launch,
    16 source segments of 1024 tiles, far path 4): 256 sampled targets
    against every source, and the momentum balance."""
    pos, vel, mass = pkg.ics.collision(1 << 21, seeds=(3, 4))
    n, h = len(mass), 0.002
    assert n == 1 << 22
    ctx = _ctx(pkg, pos, vel, mass, h, "split")
    try:
        tuning = ctx.get_tuning()
        ctx.compute_force()
        acc = ctx.download(pos=False, vel=False)["acc"]
    finally:
        ctx.close()
    assert np.isfinite(acc).all()
    rng = np.random.default_rng(25)
    sample = np.sort(rng.choice(n, 256, replace=False))
    ref = orc.force_at_points(pos[sample], h, pos, mass, variant="simd")
    e = orc.rel_err(acc[sample], ref)
    ma = mass[:, None] * acc
    bal = np.linalg.norm(ma.sum(axis=0)) / np.linalg.norm(ma, axis=1).sum()
    print(f"N=2^22 {tuning}: median {np.median(e):.3e} max {e.max():.3e}; balance {bal:.3e}")
    assert np.median(e) <= 2e-7 and e.max() <= 1e-5
    assert bal <= 1e-6


def test_collision_full_size_time_reversal(pkg, c3):
    pos, vel, mass = c3
    h, dt, k = 0.002, 1e-3, 2
    ctx = _ctx(pkg, pos, vel, mass, h, "split")
    try:
        ctx.compute_force()
        ctx.half_kick(+1, dt)
        ctx.kdk_steps(k, dt, 0.0)
        ctx.half_kick(-1, dt)
        mid = ctx.download(pos=True, vel=True, acc=False)
        ctx.upload(mid["pos"], -mid["vel"], mass)
        ctx.compute_force()
        ctx.half_kick(+1, dt)
        ctx.kdk_steps(k, dt, k * dt)
        ctx.half_kick(-1, dt)
        back = ctx.download(pos=True, vel=True, acc=False)
    finally:
        ctx.close()
    dx = np.max(np.abs(back["pos"] - pos))
    dv = np.max(np.abs(back["vel"] + vel))
    moved = np.max(np.abs(mid["pos"] - pos))
    print(f"moved {moved:.3e}; after reversal max |dx| {dx:.3e}, max |dv| {dv:.3e}")
    assert moved > 1e-4
    assert dx <= 1e-9 and dv <= 1e-6


def test_plummer_full_size_fp32_sampled_parity(pkg, orc):
    n, h = 65536, 0.005
    pos, vel, mass = pkg.ics.plummer(n, seed=2)
    ctx = _ctx(pkg, pos, vel, mass, h, "fp32")
    try:
        ctx.compute_force()
        acc = ctx.download(pos=False, vel=False)["acc"]
    finally:
        ctx.close()
    # the FP32 state holds the inputs rounded to FP32: compare on those
    pos32 = pos.astype(np.float32).astype(np.float64)
    m32 = mass.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(22)
    sample = np.sort(rng.choice(n, 2048, replace=False))
    ref = orc.force_at_points(pos32[sample], h, pos32, m32, variant="simd")
    e = orc.rel_err(acc[sample], ref)
    print(f"N=65536 fp32: median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e} max {e.max():.3e}")
    assert np.median(e) <= 1e-6


def test_cluster_ensemble_full_size_sampled_parity(pkg, orc):
    """512 clusters (N log-uniform in [256, 16384], about 2e6 particles) in
    one ragged launch: eight seeded clusters against the oracle, each with
    its own softening (median <= 1e-6, max <= 1e-4)."""
    ens = pkg.ics.cluster_ensemble(nsys=512, seed=5)
    off = ens["offsets"]
    e = pkg.Ensemble(off, "split")
    try:
        e.set_params(1.0, ens["h"], 1e-3)
        e.upload(ens["pos"], ens["vel"], ens["mass"])
        e.compute_force()
        acc = e.download(pos=False, vel=False)["acc"]
    finally:
        e.close()
    rng = np.random.default_rng(23)
    errs = []
    for s in np.sort(rng.choice(512, 8, replace=False)):
        sl = slice(off[s], off[s + 1])
        ref = orc.compute_force(ens["pos"][sl], ens["mass"][sl], [ens["h"][s]] * 6)
        errs.append(orc.rel_err(acc[sl], ref))
    err = np.concatenate(errs)
    print(f"clusters: median {np.median(err):.3e} max {err.max():.3e} over {len(err)} particles")
    assert np.median(err) <= 1e-6 and err.max() <= 1e-4


def test_planetary_ensemble_full_size_parity(pkg, orc):
    """2048 planetary systems x 73 bodies, eps = 0, central star in FP64:
    every system against the oracle (self pair skipped)."""
    ens = pkg.ics.planetary_ensemble(nsys=2048, seed=6)
    off = ens["offsets"]
    e = pkg.Ensemble(off, "split_central")
    try:
        e.set_params(1.0, 0.0, 1e-3)
        e.upload(ens["pos"], ens["vel"], ens["mass"])
        e.set_central(1)
        e.compute_force()
        acc = e.download(pos=False, vel=False)["acc"]
    finally:
        e.close()
    ref = np.empty_like(acc)
    for s in range(2048):
        sl = slice(off[s], off[s + 1])
        ref[sl] = orc.compute_force(ens["pos"][sl], ens["mass"][sl], [0.0] * 6, skip_self=True)
    err = orc.rel_err(acc, ref)
    print(f"planetary: median {np.median(err):.3e} p99 {np.percentile(err, 99):.3e} max {err.max():.3e}")
    assert np.median(err) <= 1e-9 and err.max() <= 1e-5


def test_collision_full_size_sampled_potential(pkg, orc, c3):
    """The potential of the N = 2^20 collision (FP32 pairs, FP64 sums): 512
    sampled targets against the oracle (median <= 1e-6, max <= 1e-4)."""
    pos, vel, mass = c3
    n, h = len(mass), 0.002
    ctx = _ctx(pkg, pos, vel, mass, h, "split")
    try:
        ctx.compute_potential(False)
        pot = ctx.download(pos=False, vel=False, acc=False, pot=True)["pot"]
    finally:
        ctx.close()
    rng = np.random.default_rng(24)
    sample = np.sort(rng.choice(n, 512, replace=False))
    ref = orc.potential_at_points(pos[sample], h, pos, mass)
    e = np.abs(pot[sample] - ref) / np.abs(ref)
    print(f"N=2^20 potential: median {np.median(e):.3e} max {e.max():.3e}")
    assert np.median(e) <= 1e-6 and e.max() <= 1e-4
