"""FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of liboracle_ds.so (ds_oracle.c, a C restatement of the
reference's src/directsum CPU path) plus a Python restatement of the
constant-timestep run loop (src/run.jl:46-184).  Only the tests, the smoke
check and the CPU-baseline leg of bench.py may import this module; the
product path never does.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_ds.so"

FAITHFUL, TUNED, LONGDOUBLE, SIMD = 0, 1, 2, 3
VARIANTS = {"faithful": FAITHFUL, "tuned": TUNED, "longdouble": LONGDOUBLE, "simd": SIMD}

_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)


def build(force: bool = False):
    """Compile liboracle_ds.so with the committed Makefile."""
    cmd = ["make", "-C", str(HERE)] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, capture_output=True)


def _lib():
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    I64, I = C.c_int64, C.c_int
    lib.ds_pair_force.argtypes = [_D, _D, C.c_double, C.c_double, C.c_double, _D]
    lib.ds_pair_force.restype = None
    lib.ds_pair_potential.argtypes = [_D, _D, C.c_double, C.c_double, C.c_double]
    lib.ds_pair_potential.restype = C.c_double
    lib.ds_compute_force.argtypes = [I64, _D, _D, _D, _D, _U8, _D, C.c_double, _U8, I64, I64,
                                     I, I, I, _D]
    lib.ds_compute_force.restype = I
    lib.ds_compute_potential.argtypes = [I64, _D, _D, _D, _D, _U8, _D, C.c_double, I64, I64,
                                         I, I, I, _D]
    lib.ds_compute_potential.restype = I
    lib.ds_remove_self_potential.argtypes = [I64, _D, _U8, _D, C.c_double, _D]
    lib.ds_remove_self_potential.restype = None
    lib.ds_force_at_points.argtypes = [I64, _D, C.c_double, I64, _D, _D, _D, _D, C.c_double,
                                       I, I, _D]
    lib.ds_force_at_points.restype = I
    lib.ds_potential_at_points.argtypes = [I64, _D, C.c_double, I64, _D, _D, _D, _D,
                                           C.c_double, I, I, _D]
    lib.ds_potential_at_points.restype = I
    lib.ds_background_point_mass.argtypes = [I64, _D, _D, _D, C.c_double, C.c_double, _D, _D, _D]
    lib.ds_background_point_mass.restype = None
    lib.ds_half_kick.argtypes = [I64, _D, _D, I, C.c_double]
    lib.ds_half_kick.restype = None
    lib.ds_kick.argtypes = [I64, _D, _D, C.c_double]
    lib.ds_kick.restype = None
    lib.ds_drift.argtypes = [I64, _D, _D, _D, _D, C.c_double]
    lib.ds_drift.restype = None
    lib.ds_energy.argtypes = [I64, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
    lib.ds_energy.restype = I
    lib.ds_run_ensemble_kdk.argtypes = [C.c_int32, C.POINTER(C.c_int64), _D, _D, _D, _D, _D, _D,
                                        _D, _D, C.c_int64, C.c_int32, _D, _D, I]
    lib.ds_run_ensemble_kdk.restype = I
    lib.ds_num_threads.argtypes = []
    lib.ds_num_threads.restype = I
    return lib


lib = _lib()


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _soa(pos):
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    return (np.ascontiguousarray(pos[:, 0]), np.ascontiguousarray(pos[:, 1]),
            np.ascontiguousarray(pos[:, 2]))


def _table(soft):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(soft, dtype=np.float64), (6,)))


def num_threads() -> int:
    return lib.ds_num_threads()


def pair_force(p, s, mass, G, h):
    """a1: src/directsum/gravity.jl:8-28 for one pair."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(3)
    lib.ds_pair_force(_p(p), _p(s), float(mass), float(G), float(h), _p(out))
    return out


def pair_potential(p, s, mass, G, h):
    """a2: src/directsum/potential.jl:2-27 for one pair."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib.ds_pair_potential(_p(p), _p(s), float(mass), float(G), float(h))


def compute_force(pos, mass, soft, G=1.0, collection=None, *, variant="tuned", skip_self=False,
                  targets=None, active=None, nthreads=0):
    """a5: Acc[i] for targets [t0, t1) (all by default); returns (n, 3) with
    zeros for targets outside the range."""
    x, y, z = _soa(pos)
    n = x.size
    m = np.ascontiguousarray(mass, dtype=np.float64)
    coll = np.ones(n, dtype=np.uint8) if collection is None else np.ascontiguousarray(
        collection, dtype=np.uint8)
    t0, t1 = (0, n) if targets is None else targets
    acc = np.zeros((n, 3))
    act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    rc = lib.ds_compute_force(n, _p(x), _p(y), _p(z), _p(m), coll.ctypes.data_as(_U8),
                              _p(_table(soft)), float(G),
                              None if act is None else act.ctypes.data_as(_U8), int(t0), int(t1),
                              1 if skip_self else 0, VARIANTS[variant], int(nthreads), _p(acc))
    if rc:
        raise RuntimeError(f"ds_compute_force rc={rc}")
    return acc


def compute_potential(pos, mass, soft, G=1.0, collection=None, *, variant="tuned",
                      skip_self=False, targets=None, self_term=False, nthreads=0):
    """a6 (and a7 with self_term=True): Potential[i]."""
    x, y, z = _soa(pos)
    n = x.size
    m = np.ascontiguousarray(mass, dtype=np.float64)
    coll = np.ones(n, dtype=np.uint8) if collection is None else np.ascontiguousarray(
        collection, dtype=np.uint8)
    t0, t1 = (0, n) if targets is None else targets
    pot = np.zeros(n)
    tab = _table(soft)
    rc = lib.ds_compute_potential(n, _p(x), _p(y), _p(z), _p(m), coll.ctypes.data_as(_U8),
                                  _p(tab), float(G), int(t0), int(t1), 1 if skip_self else 0,
                                  VARIANTS[variant], int(nthreads), _p(pot))
    if rc:
        raise RuntimeError(f"ds_compute_potential rc={rc}")
    if self_term:
        lib.ds_remove_self_potential(n, _p(m), coll.ctypes.data_as(_U8), _p(tab), float(G),
                                     _p(pot))
    return pot


def force_at_points(pts, h, pos, mass, G=1.0, *, variant="tuned", nthreads=0):
    """a8: probe-point accelerations."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    x, y, z = _soa(pos)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    acc = np.zeros_like(pts)
    rc = lib.ds_force_at_points(pts.shape[0], _p(pts), float(h), x.size, _p(x), _p(y), _p(z),
                                _p(m), float(G), VARIANTS[variant], int(nthreads), _p(acc))
    if rc:
        raise RuntimeError(f"ds_force_at_points rc={rc}")
    return acc


def potential_at_points(pts, h, pos, mass, G=1.0, *, variant="tuned", nthreads=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    x, y, z = _soa(pos)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    pot = np.zeros(pts.shape[0])
    rc = lib.ds_potential_at_points(pts.shape[0], _p(pts), float(h), x.size, _p(x), _p(y),
                                    _p(z), _p(m), float(G), VARIANTS[variant], int(nthreads),
                                    _p(pot))
    if rc:
        raise RuntimeError(f"ds_potential_at_points rc={rc}")
    return pot


def background_point_mass(pos, GM, center, acc=None, pot=None, G=1.0):
    """a10: analytic point-mass field added to acc/pot in place (G*M = G*GM/G)."""
    x, y, z = _soa(pos)
    c = np.ascontiguousarray(center, dtype=np.float64)
    lib.ds_background_point_mass(x.size, _p(x), _p(y), _p(z), float(G), float(GM / G), _p(c),
                                 _p(acc), _p(pot))


def background_field(pos, kind, params):
    """Analytic background fields (acc, pot) at `pos` in FP64.
    kind 1: point mass {GM, cx, cy, cz}; kind 2: Paczynski-Wiita
    pseudo-Newtonian mass {GM, cx, cy, cz, r_g}, a = -GM/(R - r_g)^2 n,
    phi = -GM/(R - r_g) (src/tools/BlackHole.jl:18-60); kind 3: radial
    polynomial {c0, cx, cy, cz, c1, c2, c3}, |a| = c0 + c1 R + c2 R^2 + c3 R^3
    toward the centre; kind 4: uniform {gx, gy, gz}, phi = -g.x."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    p = list(params) + [0.0] * (8 - len(params))
    if kind == 4:
        g = np.array(p[:3])
        return np.broadcast_to(g, pos.shape).copy(), -(pos @ g)
    d = pos - np.array(p[1:4])
    R = np.sqrt(np.sum(d * d, axis=1))
    with np.errstate(divide="ignore", invalid="ignore"):
        n = np.where(R[:, None] > 0, d / R[:, None], 0.0)
    if kind == 1:
        mag, phi = p[0] / R ** 2, -p[0] / R
    elif kind == 2:
        mag, phi = p[0] / (R - p[4]) ** 2, -p[0] / (R - p[4])
    elif kind == 3:
        c0, c1, c2, c3 = p[0], p[4], p[5], p[6]
        mag = c0 + c1 * R + c2 * R ** 2 + c3 * R ** 3
        phi = c0 * R + c1 * R ** 2 / 2 + c2 * R ** 3 / 3 + c3 * R ** 4 / 4
    else:
        raise ValueError(f"unknown background kind {kind}")
    return -mag[:, None] * n, phi


def mond_nu(y, n=2.0):
    """Generalised MOND interpolation nu(y, n) (src/mond/milgrom1983.jl:6-8)."""
    y = np.asarray(y, dtype=np.float64)
    return (0.5 * (1.0 + np.sqrt(1.0 + 4.0 * y ** (-n)))) ** (1.0 / n)


def mond(acc, acc0, n=2.0):
    """Milgrom-1983 rescaling Acc * nu(|Acc|/ACC0, n) of every row of `acc`
    (src/mond/milgrom1983.jl:33-47, all particles active).  A zero row stays
    zero."""
    acc = np.asarray(acc, dtype=np.float64)
    g = np.sqrt(np.sum(acc * acc, axis=1))
    out = acc.copy()
    nz = g > 0
    out[nz] *= mond_nu(g[nz] / acc0, n)[:, None]
    return out


def energy(pos, vel, mass, pot=None):
    """a14: (K, W, P, L) with W = 0.5 * sum(m * phi) as total_potential(sim)."""
    x, y, z = _soa(pos)
    v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    p = None if pot is None else np.ascontiguousarray(pot, dtype=np.float64)
    K, Ws = C.c_double(), C.c_double()
    P, L = np.zeros(3), np.zeros(3)
    rc = lib.ds_energy(x.size, _p(x), _p(y), _p(z), _p(v), _p(m), _p(p), C.byref(K), C.byref(Ws),
                       _p(P), _p(L))
    if rc:
        raise RuntimeError(f"ds_energy rc={rc}")
    return K.value, 0.5 * Ws.value, P, L


# ---------------------------------------------------------------------------
# Constant-timestep run loop: src/run.jl:46-184 with
#   find_next_sync_point_and_drift (const)  src/base/Timestep.jl:197-215
#   drift_particles_kernel (const)          src/base/Timestep.jl:181-191
#   advance_and_find_timestep_local (const) src/base/Timestep.jl:84-93
#   leapfrog_half_kick_local (const)        src/base/Timestep.jl:116-125
#   TimeInfo initial values                 src/base/Config.jl:746-766
# ---------------------------------------------------------------------------
def count_steps(t_begin: float, t_end: float, dt: float) -> int:
    """Number of loop iterations of run(): t += dt (IEEE FP64) until t >= End."""
    t = float(t_begin)
    k = 0
    while True:
        t = t + dt
        k += 1
        if t >= t_end:
            return k


def run_const_kdk(pos, vel, mass, soft, G, *, t_begin, t_end, dt, between_snapshots=None,
                  collection=None, on_output=None, analysers=None, skip_self=False,
                  variant="tuned", nthreads=0):
    """Restatement of run(sim) for DirectSum + ConstantTimestep + Leapfrog.

    Returns (pos, vel, acc, info) with info = {"steps", "outputs", "analysis"}.
    on_output(t, pos, vel) is called at each snapshot event (snapshot 0 after
    the first half kick, then whenever next_output_time <= t + dt)."""
    pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.array(vel, dtype=np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    n = pos.shape[0]
    x, y, z = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
    v = np.ascontiguousarray(vel)
    if between_snapshots is None:
        between_snapshots = t_end - t_begin

    def force():
        return compute_force(np.stack([x, y, z], axis=1), m, soft, G, collection,
                             variant=variant, skip_self=skip_self, nthreads=nthreads)

    outputs = []
    analysis = []

    def output(t):
        outputs.append(t)
        if on_output is not None:
            on_output(t, np.stack([x, y, z], axis=1), v.copy())

    system_time = float(t_begin)
    next_output = float(t_begin) + between_snapshots
    acc = force()                                   # run.jl:76
    lib.ds_half_kick(n, _p(v), _p(acc), +1, float(dt))   # run.jl:85-87
    output(system_time)                             # run.jl:90
    steps = 0
    state = True
    while state:                                    # run.jl:101
        next_time = system_time + dt                # Timestep.jl:198
        while next_output <= next_time:             # Timestep.jl:200-210
            d = next_output - system_time
            lib.ds_drift(n, _p(x), _p(y), _p(z), _p(v), float(d))
            system_time = next_output
            output(system_time)
            next_output += between_snapshots
        d = next_time - system_time                 # Timestep.jl:183
        lib.ds_drift(n, _p(x), _p(y), _p(z), _p(v), float(d))
        system_time = next_time
        acc = force()                               # compute_force
        lib.ds_kick(n, _p(v), _p(acc), float(dt))   # Timestep.jl:84-93
        if analysers:
            analysis.append([f(np.stack([x, y, z], axis=1), v, m) for f in analysers])
        if system_time >= t_end:                    # run.jl:158-160
            state = False
        steps += 1
    lib.ds_half_kick(n, _p(v), _p(acc), -1, float(dt))   # run.jl:171-173
    return np.stack([x, y, z], axis=1), v.reshape(n, 3), acc, {
        "steps": steps, "outputs": outputs, "analysis": analysis}


def kdk_steps(pos, vel, mass, soft, G, *, nsteps, dt, t_begin=0.0, acc=None, collection=None,
              skip_self=False, variant="tuned", nthreads=0):
    """nsteps x {drift by (t+dt)-t, force, kick by dt} (no half kicks): the body
    of step(sim, DirectSum, CPU)  src/directsum/timestep.jl:43-55."""
    pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(np.array(vel, dtype=np.float64).reshape(-1, 3))
    m = np.ascontiguousarray(mass, dtype=np.float64)
    n = pos.shape[0]
    x, y, z = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
    t = float(t_begin)
    for _ in range(nsteps):
        t_new = t + dt
        lib.ds_drift(n, _p(x), _p(y), _p(z), _p(v), float(t_new - t))
        t = t_new
        acc = compute_force(np.stack([x, y, z], axis=1), m, soft, G, collection,
                            variant=variant, skip_self=skip_self, nthreads=nthreads)
        lib.ds_kick(n, _p(v), _p(acc), float(dt))
    return np.stack([x, y, z], axis=1), v, acc


def run_ensemble_kdk(offsets, pos, vel, mass, G, soft, dt, *, nsteps, every=0, nthreads=0):
    """Independent systems: force, +1/2 kick, nsteps x (drift, force, kick),
    per-system energies every `every` steps (velocities synchronised to the
    step).  Returns (pos, vel, E0 (nsys,), energy (nsys, nsteps // every))."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nsys = off.size - 1
    x, y, z = [np.ascontiguousarray(np.asarray(pos, dtype=np.float64)[:, k]) for k in range(3)]
    v = np.ascontiguousarray(np.array(vel, dtype=np.float64).reshape(-1, 3))
    m = np.ascontiguousarray(mass, dtype=np.float64)
    vec = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (nsys,)))
    Gs, hs, dts = vec(G), vec(soft), vec(dt)
    slots = nsteps // every if every > 0 else 0
    energy = np.zeros(max(slots, 1) * nsys)
    E0 = np.zeros(nsys)
    rc = lib.ds_run_ensemble_kdk(nsys, off.ctypes.data_as(C.POINTER(C.c_int64)), _p(x), _p(y),
                                 _p(z), _p(v), _p(m), _p(Gs), _p(hs), _p(dts), int(nsteps),
                                 int(every), _p(energy), _p(E0), int(nthreads))
    if rc:
        raise RuntimeError(f"ds_run_ensemble_kdk rc={rc}")
    return (np.stack([x, y, z], axis=1), v, E0,
            energy[: slots * nsys].reshape(nsys, slots) if slots else np.zeros((nsys, 0)))


def half_kick(vel, acc, forward, dt):
    v = np.ascontiguousarray(np.array(vel, dtype=np.float64))
    a = np.ascontiguousarray(acc, dtype=np.float64)
    lib.ds_half_kick(v.shape[0], _p(v), _p(a), int(forward), float(dt))
    return v


# ---------------------------------------------------------------------------
# Adaptive (block) timestep run loop: src/run.jl:46-184 with AdaptiveTimestep
#   get_timestep                      src/base/Timestep.jl:1-23
#   init_timesteps_local              :29-39
#   advance_and_find_timestep_local   :47-82
#   leapfrog_half_kick_local          :105-114
#   find_min_endstep_local            :135-144
#   drift_particles_kernel            :169-179
#   find_next_sync_point_and_drift    :224-245
# ---------------------------------------------------------------------------
def get_timestep(acc, h, *, err_tol, min_step, max_step, time_interval, time_base):
    """Integer steps for accelerations acc (n,3) and softening h (n,)."""
    acc = np.atleast_2d(acc)
    out = np.empty(acc.shape[0], dtype=np.int64)
    for i in range(acc.shape[0]):
        ac = math.sqrt(acc[i, 0] ** 2 + acc[i, 1] ** 2 + acc[i, 2] ** 2)
        with np.errstate(divide="ignore"):
            dt = math.sqrt(2 * err_tol * h[i] / ac) if ac > 0 else math.inf
        if dt > max_step:
            dt = max_step
        if dt < min_step:
            raise RuntimeError(f"Timestep < MinStep = {min_step}, dt = {dt}")
        ti_step = dt / time_interval
        if ti_step <= 0:
            raise RuntimeError(f"Wrong integer timestep: {ti_step}")
        ti_min = time_base
        while ti_min > ti_step:
            ti_min >>= 1
        out[i] = ti_min
    return out


def run_adaptive(pos, vel, mass, soft, G, *, t_begin, t_end, time_base=1 << 30, err_tol=0.025,
                 min_step=1e-8, max_step=math.inf, between_snapshots=None, collection=None,
                 on_output=None, nthreads=0, max_steps=10 ** 7, post=None):
    """Restatement of run(sim) for DirectSum + AdaptiveTimestep + Leapfrog.
    `post(pos, acc)` (optional) returns the post-processed accelerations
    (src/base/PostProcessing/Force.jl:33-46) of the rows it is given.
    Returns (pos, vel, acc, info) with info = {"steps", "nactive", "outputs",
    "ti_end", "ti_beg"}."""
    pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(np.array(vel, dtype=np.float64).reshape(-1, 3))
    m = np.ascontiguousarray(mass, dtype=np.float64)
    n = pos.shape[0]
    coll = np.ones(n, dtype=np.uint8) if collection is None else np.asarray(collection, np.uint8)
    table = _table(soft)
    h = table[coll - 1]
    interval = (t_end - t_begin) / time_base
    if between_snapshots is None:
        between_snapshots = t_end - t_begin
    bsi = int(math.floor(between_snapshots / interval))
    x, y, z = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
    kw = dict(err_tol=err_tol, min_step=min_step, max_step=max_step, time_interval=interval,
              time_base=time_base)
    te = np.zeros(n, dtype=np.int64)
    tb = np.zeros(n, dtype=np.int64)
    sys_int, next_out = 0, bsi
    sys_float = float(t_begin)
    outputs = []

    def force(active):
        p = np.stack([x, y, z], axis=1)
        a = compute_force(p, m, soft, G, coll, active=active.astype(np.uint8), nthreads=nthreads)
        if post is not None:
            a[active] = post(p[active], a[active])
        return a

    def drift(dt):
        lib.ds_drift(n, _p(x), _p(y), _p(z), _p(v), float(dt))

    acc = force(te == sys_int)                               # run.jl:76
    te = get_timestep(acc, h, **kw)                          # run.jl:78
    tb[:] = 0
    v += acc * 1.0 * 0.5 * ((te - tb) * interval)[:, None]   # run.jl:85-87
    outputs.append(sys_float)
    steps = 0
    nactive = []
    while steps < max_steps:
        min_end = min(int(te.min()), time_base)
        while next_out <= min_end:
            drift((next_out - sys_int) * interval)
            sys_int = next_out
            sys_float = t_begin + sys_int * interval
            outputs.append(sys_float)
            if on_output is not None:
                on_output(sys_float, np.stack([x, y, z], axis=1), v.copy())
            next_out += bsi
        drift((min_end - sys_int) * interval)
        sys_int = min_end
        sys_float = t_begin + sys_int * interval
        active = te == sys_int
        nactive.append(int(active.sum()))
        new = force(active)
        acc[active] = new[active]
        for i in np.nonzero(active)[0]:
            ti_step = int(get_timestep(acc[i:i + 1], h[i:i + 1], **kw)[0])
            if sys_int == time_base:
                ti_step = 0
            if time_base - sys_int < ti_step:
                ti_step = time_base - sys_int
            tstart = (tb[i] + te[i]) / 2
            tend = te[i] + 0.5 * ti_step
            dt = (tend - tstart) * interval
            v[i] += acc[i] * dt
            tb[i] = te[i]
            te[i] = tb[i] + ti_step
        steps += 1
        if sys_float >= t_end:
            break
    v += acc * -1.0 * 0.5 * ((te - tb) * interval)[:, None]  # run.jl:171-173
    return np.stack([x, y, z], axis=1), v, acc, {
        "steps": steps, "nactive": nactive, "outputs": outputs, "ti_end": te, "ti_beg": tb}


def rel_err(a, b):
    """Per-particle |a-b|/|b| for (n,3) vectors or (n,) scalars."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 2:
        return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    return np.abs(a - b) / np.abs(b)


def err_stats(e):
    e = np.asarray(e)
    return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
            "max": float(np.max(e))}
