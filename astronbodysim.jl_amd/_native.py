"""ctypes binding of libnbody_mi355x.so (the C-ABI in include/nbody_mi355x.h).

The shared library is built in-tree by ``__graft_entry__.build()``.  Nothing
here falls back to CPU code: if the library is missing, importing this module
raises, and every failed call raises :class:`NBError` with the library's
message.
"""
from __future__ import annotations

import ctypes as C
import os
import warnings
from pathlib import Path

import numpy as np

LIB_NAME = "libnbody_mi355x.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

NB_OK = 0
NB_ERR_INVALID = -1
NB_ERR_HIP = -2
NB_ERR_NOMEM = -3
NB_ERR_STATE = -4
NB_ERR_NONFINITE = -5
NB_ERR_TIMESTEP = -6

NB_MODE_FP64 = 0
NB_MODE_FP32 = 1
NB_MODE_SPLIT = 2
NB_MODE_SPLIT_CENTRAL = 3

MODES = {
    "fp64": NB_MODE_FP64,
    "fp32": NB_MODE_FP32,
    "split": NB_MODE_SPLIT,
    "split_central": NB_MODE_SPLIT_CENTRAL,
}


class NBError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code
        self.func = func


class DevInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 128),
        ("arch", C.c_char * 64),
        ("device", C.c_int32),
        ("compute_units", C.c_int32),
        ("clock_khz", C.c_int32),
        ("lds_per_block", C.c_int32),
        ("warp_size", C.c_int32),
        ("max_threads_per_block", C.c_int32),
        ("global_mem_bytes", C.c_int64),
        ("l2_bytes", C.c_int64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("targets_per_thread", C.c_int32),
        ("j_segments", C.c_int32),
        ("sort_interval", C.c_int32),
        ("far_path", C.c_int32),
    ]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_I64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
_SIGS = {
    "nb_version": (C.c_int, []),
    "nb_last_error": (C.c_char_p, []),
    "nb_device_info": (C.c_int, [C.POINTER(DevInfo)]),
    "nb_create": (C.c_int, [C.POINTER(_P), C.c_int64, C.c_int32]),
    "nb_destroy": (C.c_int, [_P]),
    "nb_set_params": (C.c_int, [_P, C.c_double, _D]),
    "nb_upload": (C.c_int, [_P, _D, _D, _D, _U8]),
    "nb_download": (C.c_int, [_P, _D, _D, _D, _D]),
    "nb_upload_acc": (C.c_int, [_P, _D, _D]),
    "nb_snapshot_begin": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "nb_snapshot_fetch": (C.c_int, [_P, C.c_int32, _D, _D, _D, _D]),
    "nb_compute_force": (C.c_int, [_P]),
    "nb_compute_potential": (C.c_int, [_P, C.c_int32]),
    "nb_force_at_points": (C.c_int, [_P, C.c_int64, _D, C.c_double, _D]),
    "nb_potential_at_points": (C.c_int, [_P, C.c_int64, _D, C.c_double, _D]),
    "nb_half_kick": (C.c_int, [_P, C.c_int32, C.c_double]),
    "nb_drift": (C.c_int, [_P, C.c_double]),
    "nb_kick": (C.c_int, [_P, C.c_double]),
    "nb_kdk_steps": (C.c_int, [_P, C.c_int64, C.c_double, C.c_double, C.c_int32]),
    "nb_energy": (C.c_int, [_P, _D, _D, _D, _D]),
    "nb_set_central": (C.c_int, [_P, C.c_int64]),
    "nb_add_background": (C.c_int, [_P, C.c_int32, _D]),
    "nb_set_mond": (C.c_int, [_P, C.c_double, C.c_double]),
    "nb_sync": (C.c_int, [_P]),
    "nb_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "nb_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "nb_autotune": (C.c_int, [_P]),
    "nb_set_timing": (C.c_int, [_P, C.c_int32]),
    "nb_get_timing": (C.c_int, [_P, _I64, _D]),
    "nb_tuner_save": (C.c_int, [C.c_char_p]),
    "nb_tuner_load": (C.c_int, [C.c_char_p]),
    "nb_tuner_clear": (C.c_int, []),
    # device pointers (and the stream) travel as plain addresses
    "nb_compute_force_noalloc": (C.c_int, [C.c_int32, C.c_int64, _P, _P, _P, _P, _P, C.c_double,
                                           C.c_double, _P, _P, _P, _P, _P]),
}

_OPTIONAL = {}


def _load():
    global LIB_PATH
    override = os.environ.get("NBODY_MI355X_LIB")
    if override:
        LIB_PATH = Path(override)
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: the HIP extension has not been built "
            "(run __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, (res, args) in {**_SIGS, **_OPTIONAL}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Launch configurations measured by the autotuner on MI355X, keyed by
# (precision mode, ceil(log2 N)); nb_create starts from the matching entry.
DEFAULT_TUNING = Path(__file__).resolve().parent / "tuning" / "mi355x.json"
if DEFAULT_TUNING.exists():
    lib.nb_tuner_load(os.fsencode(str(DEFAULT_TUNING)))


def exported_symbols():
    """Names of the C-ABI entry points this binding declares."""
    return sorted(_SIGS)


def register(sigs: dict):
    """Declare further entry points (used by the batched-ensemble module)."""
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
        _SIGS[name] = (res, args)


def last_error() -> str:
    return lib.nb_last_error().decode(errors="replace")


def check(rc: int, func: str):
    if rc != NB_OK:
        raise NBError(func, rc, last_error())


def call(func: str, *args):
    rc = getattr(lib, func)(*args)
    check(rc, func)
    return rc


def dptr(a):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_D)


def u8ptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_U8)


def i64ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I64)


def device_info() -> dict:
    info = DevInfo()
    call("nb_device_info", C.byref(info))
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "device": info.device,
        "compute_units": info.compute_units,
        "clock_khz": info.clock_khz,
        "lds_per_block": info.lds_per_block,
        "warp_size": info.warp_size,
        "max_threads_per_block": info.max_threads_per_block,
        "global_mem_bytes": info.global_mem_bytes,
        "l2_bytes": info.l2_bytes,
    }


def version() -> tuple:
    v = lib.nb_version()
    return (v // 10000, (v // 100) % 100, v % 100)


class Context:
    """Owning wrapper of one nb_ctx (one system on the GPU)."""

    def __init__(self, n: int, mode: int = NB_MODE_SPLIT):
        h = _P()
        call("nb_create", C.byref(h), int(n), int(mode))
        self._h = h
        self.n = int(n)
        self.mode = int(mode)

    @property
    def handle(self):
        if self._h is None:
            raise NBError("context", NB_ERR_STATE, "context destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.nb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ---------------------------------------------------------------
    def set_params(self, G: float, soft):
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(soft, dtype=np.float64), (6,)))
        call("nb_set_params", self.handle, float(G), dptr(s))

    def upload(self, pos, vel, mass, collection=None):
        n = self.n
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(n)
        if collection is None:
            collection = np.ones(n, dtype=np.uint8)
        collection = np.ascontiguousarray(collection, dtype=np.uint8).reshape(n)
        call("nb_upload", self.handle, dptr(pos), dptr(vel), dptr(mass), u8ptr(collection))

    def download(self, pos=True, vel=True, acc=True, pot=False):
        n = self.n
        out = {}
        bufs = {}
        for key, want, shape in (("pos", pos, (n, 3)), ("vel", vel, (n, 3)),
                                 ("acc", acc, (n, 3)), ("pot", pot, (n,))):
            bufs[key] = np.empty(shape, dtype=np.float64) if want else None
        call("nb_download", self.handle, dptr(bufs["pos"]), dptr(bufs["vel"]),
             dptr(bufs["acc"]), dptr(bufs["pot"]))
        for k, v in bufs.items():
            if v is not None:
                out[k] = v
        return out

    def upload_acc(self, acc=None, pot=None):
        """Set Acc (n,3) and/or Potential (n,) in the caller's order."""
        n = self.n
        a = None if acc is None else np.ascontiguousarray(acc, dtype=np.float64).reshape(n, 3)
        p = None if pot is None else np.ascontiguousarray(pot, dtype=np.float64).reshape(n)
        call("nb_upload_acc", self.handle, dptr(a), dptr(p))

    SNAP_FIELDS = {"pos": 1, "vel": 2, "acc": 4, "pot": 8}

    def snapshot_begin(self, slot: int, fields=("pos", "vel")):
        """Enqueue an asynchronous copy of `fields` into pinned slot `slot`."""
        mask = 0
        for f in fields:
            mask |= self.SNAP_FIELDS[f]
        call("nb_snapshot_begin", self.handle, int(slot), mask)

    def snapshot_fetch(self, slot: int, fields=("pos", "vel")):
        """Wait for slot `slot` and return its fields (caller order).  Safe to
        call from another thread while this context keeps working."""
        n = self.n
        bufs = {k: (np.empty((n, 3) if k != "pot" else (n,), dtype=np.float64) if k in fields else None)
                for k in self.SNAP_FIELDS}
        call("nb_snapshot_fetch", self.handle, int(slot), dptr(bufs["pos"]), dptr(bufs["vel"]),
             dptr(bufs["acc"]), dptr(bufs["pot"]))
        return {k: v for k, v in bufs.items() if v is not None}

    # -- forces --------------------------------------------------------------
    def compute_force(self):
        call("nb_compute_force", self.handle)

    def compute_potential(self, self_term: bool = False):
        call("nb_compute_potential", self.handle, 1 if self_term else 0)

    def force_at_points(self, pts, h: float):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        acc = np.empty_like(pts)
        call("nb_force_at_points", self.handle, pts.shape[0], dptr(pts), float(h), dptr(acc))
        return acc

    def potential_at_points(self, pts, h: float):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        pot = np.empty(pts.shape[0], dtype=np.float64)
        call("nb_potential_at_points", self.handle, pts.shape[0], dptr(pts), float(h), dptr(pot))
        return pot

    # -- integration ---------------------------------------------------------
    def half_kick(self, forward: int, dt: float):
        call("nb_half_kick", self.handle, int(forward), float(dt))

    def drift(self, dt: float):
        call("nb_drift", self.handle, float(dt))

    def kick(self, dt: float):
        call("nb_kick", self.handle, float(dt))

    def kdk_steps(self, nsteps: int, dt: float, t_begin: float = 0.0, use_graph: bool = False):
        call("nb_kdk_steps", self.handle, int(nsteps), float(dt), float(t_begin),
             1 if use_graph else 0)

    def energy(self, potential: bool = True):
        K = C.c_double()
        W = C.c_double()
        P = np.zeros(3)
        L = np.zeros(3)
        rc = lib.nb_energy(self.handle, C.byref(K), C.byref(W) if potential else None,
                           dptr(P), dptr(L))
        if rc == NB_ERR_NONFINITE:
            # the reference returns NaN / Inf energies as they are: so does
            # this host, with a warning
            warnings.warn(last_error(), RuntimeWarning, stacklevel=2)
        else:
            check(rc, "nb_energy")
        return {"K": K.value, "W": W.value if potential else None, "P": P, "L": L}

    def set_central(self, count: int):
        call("nb_set_central", self.handle, int(count))

    def add_background(self, kind: int, params=()):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        if p.size < 8:
            p = np.concatenate([p, np.zeros(8 - p.size)])
        call("nb_add_background", self.handle, int(kind), dptr(p))

    def set_mond(self, acc0: float, nu_index: float = 2.0):
        call("nb_set_mond", self.handle, float(acc0), float(nu_index))

    def sync(self):
        call("nb_sync", self.handle)

    # -- tuning / timing -----------------------------------------------------
    def set_tuning(self, targets_per_thread: int, j_segments: int, sort_interval: int = 0,
                   far_path: int = 0):
        t = Tuning(int(targets_per_thread), int(j_segments), int(sort_interval), int(far_path))
        call("nb_set_tuning", self.handle, C.byref(t))

    def get_tuning(self) -> dict:
        t = Tuning()
        call("nb_get_tuning", self.handle, C.byref(t))
        return {"targets_per_thread": t.targets_per_thread, "j_segments": t.j_segments,
                "sort_interval": t.sort_interval, "far_path": t.far_path}

    def autotune(self):
        call("nb_autotune", self.handle)

    def set_timing(self, enable: bool):
        call("nb_set_timing", self.handle, 1 if enable else 0)

    def get_timing(self):
        cnt = C.c_int64()
        ms = C.c_double()
        call("nb_get_timing", self.handle, C.byref(cnt), C.byref(ms))
        return cnt.value, ms.value


def compute_force_noalloc(mode: int, n: int, x: int, y: int, z: int, m: int, ax: int, ay: int,
                          az: int, pot: int = 0, h: int = 0, h_const: float = 0.0, G: float = 1.0,
                          stream: int = 0):
    """nb_compute_force_noalloc on caller-owned device arrays, given as device
    addresses (0 = NULL for pot, h and stream).  Enqueued on `stream`; the
    caller synchronises."""
    call("nb_compute_force_noalloc", int(mode), int(n), x, y, z, m, h or None, float(h_const),
         float(G), ax, ay, az, pot or None, stream or None)


def tuner_save(path):
    call("nb_tuner_save", os.fsencode(str(path)))


def tuner_load(path):
    call("nb_tuner_load", os.fsencode(str(path)))


def tuner_clear():
    call("nb_tuner_clear")
