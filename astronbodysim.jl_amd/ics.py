"""Seeded synthetic initial conditions for the benchmark workloads.

The reference draws Plummer spheres from the external AstroIC package
(test/Project.toml:2); its sample stream is not reproducible here, so these
generators produce their own seeded inputs (numpy PCG64) and every comparison
feeds identical inputs to the GPU path and to the FP64 oracle.

Plummer sampling follows Aarseth, Henon & Wielen (1974, A&A 37, 183):
radius from the inverted cumulative mass M(<r) = r^3/(r^2+a^2)^{3/2}
(truncated at 99.9% of the mass, as NEMO's mkplummer does), speed from the
isotropic distribution function by rejection on g(q) = q^2 (1-q^2)^{7/2},
isotropic directions; units G = M = a = 1 scaled to the requested G, M, a.
"""
from __future__ import annotations

import math

import numpy as np

MASS_FRACTION_CUT = 0.999


def _isotropic(rng, n):
    u = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(1.0 - u * u)
    return np.stack([s * np.cos(phi), s * np.sin(phi), u], axis=1)


def plummer(n: int, seed: int, *, G: float = 1.0, M: float = 1.0, a: float = 1.0,
            masses=None, recenter: bool = True):
    """Plummer sphere: returns (pos (n,3), vel (n,3), mass (n,)) in float64."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x1 = rng.uniform(0.0, MASS_FRACTION_CUT, n)
    x1 = np.maximum(x1, 1e-300)
    r = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    pos = _isotropic(rng, n) * r[:, None]
    # speeds: q in [0,1) with density ~ q^2 (1-q^2)^3.5 (max 0.092 at q=sqrt(2/9))
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.uniform(0.0, 1.0, todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    vesc = math.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    vel = _isotropic(rng, n) * (q * vesc)[:, None]
    if masses is None:
        mass = np.full(n, 1.0 / n)
    else:
        mass = np.asarray(masses, dtype=np.float64)
        mass = mass / mass.sum()
    if recenter:
        pos -= (mass[:, None] * pos).sum(0) / mass.sum()
        vel -= (mass[:, None] * vel).sum(0) / mass.sum()
    vscale = math.sqrt(G * M / a)
    return pos * a, vel * vscale, mass * M


def lognormal_masses(n: int, seed: int, sigma: float = 0.5):
    rng = np.random.Generator(np.random.PCG64(seed))
    m = np.exp(sigma * rng.standard_normal(n))
    return m / m.sum()


def collision(n_each: int, seeds=(3, 4), *, a: float = 1.0, separation: float = 20.0,
              rel_speed: float = 0.5, sigma: float = 0.5, G: float = 1.0):
    """Two Plummer spheres (mass 1 each, log-normal masses) at x = -/+ separation/2
    approaching along x with relative speed `rel_speed` (each moves at half of it)."""
    out_p, out_v, out_m = [], [], []
    for k, seed in enumerate(seeds):
        m = lognormal_masses(n_each, seed + 1000, sigma)
        p, v, m = plummer(n_each, seed, G=G, M=1.0, a=a, masses=m)
        sgn = -1.0 if k == 0 else 1.0
        p[:, 0] += sgn * 0.5 * separation
        v[:, 0] += -sgn * 0.5 * rel_speed
        out_p.append(p)
        out_v.append(v)
        out_m.append(m)
    return np.concatenate(out_p), np.concatenate(out_v), np.concatenate(out_m)


def cluster_ensemble(nsys: int = 512, seed: int = 5, nmin: int = 256, nmax: int = 16384):
    """Independent Plummer clusters: N log-uniform in [nmin, nmax], a ~ U[0.5, 2],
    h = 0.01 a, M = 1.  Returns dict with offsets (nsys+1), pos, vel, mass, a, h."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nsys))).astype(np.int64)
    ns = np.clip(ns, nmin, nmax)
    a = rng.uniform(0.5, 2.0, nsys)
    seeds = rng.integers(0, 2**31 - 1, nsys)
    offs = np.zeros(nsys + 1, dtype=np.int64)
    offs[1:] = np.cumsum(ns)
    pos = np.empty((offs[-1], 3))
    vel = np.empty((offs[-1], 3))
    mass = np.empty(offs[-1])
    for s in range(nsys):
        p, v, m = plummer(int(ns[s]), int(seeds[s]), a=float(a[s]))
        sl = slice(offs[s], offs[s + 1])
        pos[sl], vel[sl], mass[sl] = p, v, m
    return {"offsets": offs, "pos": pos, "vel": vel, "mass": mass, "a": a, "h": 0.01 * a}


def _kepler_E(M, e):
    E = M + e * np.sin(M)
    for _ in range(50):
        dE = (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
        E -= dE
        if dE.size == 0 or np.max(np.abs(dE)) < 1e-15:
            break
    return E


def _elements_to_cartesian(a, e, inc, Omega, omega, M, mu):
    E = _kepler_E(M, e)
    cE, sE = np.cos(E), np.sin(E)
    b = a * np.sqrt(1.0 - e * e)
    xp = a * (cE - e)
    yp = b * sE
    n = np.sqrt(mu / a ** 3)
    vxp = -a * n * sE / (1.0 - e * cE)
    vyp = b * n * cE / (1.0 - e * cE)
    cO, sO = np.cos(Omega), np.sin(Omega)
    co, so = np.cos(omega), np.sin(omega)
    ci, si = np.cos(inc), np.sin(inc)
    r11 = cO * co - sO * so * ci
    r12 = -cO * so - sO * co * ci
    r21 = sO * co + cO * so * ci
    r22 = -sO * so + cO * co * ci
    r31 = so * si
    r32 = co * si
    pos = np.stack([r11 * xp + r12 * yp, r21 * xp + r22 * yp, r31 * xp + r32 * yp], axis=-1)
    vel = np.stack([r11 * vxp + r12 * vyp, r21 * vxp + r22 * vyp, r31 * vxp + r32 * vyp], axis=-1)
    return pos, vel


def planetary_ensemble(nsys: int = 2048, seed: int = 6, nplanets: int = 8, ntest: int = 64):
    """Planetary systems: central mass 1 (index 0 of each system), `nplanets`
    planets (a ~ logU[0.4, 30], e ~ U[0, 0.1], i <= 2 deg, m/M ~ logU[1e-7, 1e-3]),
    `ntest` massless particles on circular orbits at r ~ U[1, 40] (i <= 2 deg).
    Positions/velocities are barycentric per system.  G = 1, eps = 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per = 1 + nplanets + ntest
    N = nsys * per
    pos = np.zeros((N, 3))
    vel = np.zeros((N, 3))
    mass = np.zeros(N)
    deg = math.pi / 180.0
    for s in range(nsys):
        base = s * per
        mass[base] = 1.0
        ap = np.exp(rng.uniform(math.log(0.4), math.log(30.0), nplanets))
        ep = rng.uniform(0.0, 0.1, nplanets)
        ip = rng.uniform(0.0, 2.0 * deg, nplanets)
        mp = np.exp(rng.uniform(math.log(1e-7), math.log(1e-3), nplanets))
        ang = rng.uniform(0.0, 2.0 * math.pi, (3, nplanets))
        p, v = _elements_to_cartesian(ap, ep, ip, ang[0], ang[1], ang[2], 1.0 + mp)
        pos[base + 1:base + 1 + nplanets] = p
        vel[base + 1:base + 1 + nplanets] = v
        mass[base + 1:base + 1 + nplanets] = mp
        rt = rng.uniform(1.0, 40.0, ntest)
        it = rng.uniform(0.0, 2.0 * deg, ntest)
        angt = rng.uniform(0.0, 2.0 * math.pi, (3, ntest))
        pt, vt = _elements_to_cartesian(rt, np.zeros(ntest), it, angt[0], angt[1], angt[2], 1.0)
        t0 = base + 1 + nplanets
        pos[t0:t0 + ntest] = pt
        vel[t0:t0 + ntest] = vt
        sl = slice(base, base + per)
        msum = mass[sl].sum()
        pos[sl] -= (mass[sl, None] * pos[sl]).sum(0) / msum
        vel[sl] -= (mass[sl, None] * vel[sl]).sum(0) / msum
    offs = np.arange(nsys + 1, dtype=np.int64) * per
    return {"offsets": offs, "pos": pos, "vel": vel, "mass": mass, "per_system": per}


def binary(mass: float = 1.0e10, D: float = 2.0, G: float = 1.0):
    """The StaticTest binary (test/StaticTest.jl:18-123): equal masses at +-D/2
    on x with circular velocities +-sqrt(G m / (2 D)) along y."""
    v = math.sqrt(G * mass / D / 2.0)
    pos = np.array([[+0.5 * D, 0.0, 0.0], [-0.5 * D, 0.0, 0.0]])
    vel = np.array([[0.0, +v, 0.0], [0.0, -v, 0.0]])
    return pos, vel, np.array([mass, mass])


def box10(mass: float = 1.0e9):
    """The StaticTest Box particles (test/StaticTest.jl:93-123)."""
    pts = [
        (+0.1 + 1.9, +0.1 + 1.9, +0.1 + 1.9),
        (-0.1 + 1.9, +0.1 + 1.9, +0.1 + 1.9),
        (+0.1 + 1.9, -0.1 + 1.9, +0.1 + 1.9),
        (+0.1 + 1.9, +0.1 + 1.9, -0.1 + 1.9),
        (-0.1 + 1.9, -0.1 + 1.9, +0.1 + 1.9),
        (-0.1 + 1.9, +0.1 + 1.9, -0.1 + 1.9),
        (+0.1 + 1.9, -0.1 + 1.9, -0.1 + 1.9),
        (-0.1 + 1.9, -0.1 + 1.9, -0.1 + 1.9),
        (+1.0, +1.0, +1.0),
        (-1.0, -1.0, -1.0),
    ]
    pos = np.array(pts, dtype=np.float64)
    return pos, np.zeros_like(pos), np.full(10, mass)
