"""C-ABI library checks that need no GPU: it loads, exports every symbol the
public headers declare, and rejects bad arguments before touching a device."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for hdr in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"^NB_API\s+[^;(]*?\b(nb_\w+)\s*\(", hdr.read_text(), re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert {"nb_create", "nb_compute_force", "nb_kdk_steps", "nb_energy"} <= names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg._native.lib
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header(pkg):
    bound = set(pkg._native.exported_symbols())
    assert declared_symbols() <= bound, sorted(declared_symbols() - bound)


def test_version_and_last_error(pkg):
    assert pkg.version() == (1, 0, 0)
    assert pkg._native.lib.nb_version() == 10000


def test_invalid_arguments_rejected_without_device(pkg):
    nat = pkg._native
    h = C.c_void_p()
    assert nat.lib.nb_create(C.byref(h), 10, 7) == nat.NB_ERR_INVALID
    assert "precision mode" in nat.last_error()
    assert nat.lib.nb_create(C.byref(h), -1, 2) == nat.NB_ERR_INVALID
    assert nat.lib.nb_create(None, 10, 2) == nat.NB_ERR_INVALID
    assert nat.lib.nb_destroy(None) == nat.NB_OK
    assert nat.lib.nb_compute_force(None) == nat.NB_ERR_INVALID
    assert nat.lib.nb_energy(None, None, None, None, None) == nat.NB_ERR_INVALID
    with pytest.raises(nat.NBError):
        nat.call("nb_kdk_steps", None, 1, 1e-3, 0.0, 0)


def test_last_error_is_per_thread(pkg):
    """nb_last_error belongs to the calling thread: an error raised on another
    thread in between does not replace it."""
    import threading
    nat = pkg._native
    step = threading.Barrier(2)
    seen = {}

    def first():
        h = C.c_void_p()
        assert nat.lib.nb_create(C.byref(h), 10, 7) == nat.NB_ERR_INVALID
        step.wait()          # the other thread fails now
        step.wait()
        seen["first"] = nat.last_error()

    def second():
        step.wait()
        h = C.c_void_p()
        assert nat.lib.nb_create(C.byref(h), -1, 2) == nat.NB_ERR_INVALID
        seen["second"] = nat.last_error()
        step.wait()

    ts = [threading.Thread(target=first), threading.Thread(target=second)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=30)
    assert "precision mode" in seen["first"]
    assert seen["second"] and "precision mode" not in seen["second"]


def test_tuner_table_roundtrip(pkg, tmp_path):
    nat = pkg._native
    p = tmp_path / "tune.json"
    p.write_text('{"version": 1, "entries": [\n {"mode": 2, "log2n": 16, "targets_per_thread": 2, '
                 '"j_segments": 8, "sort_interval": 32, "ms": 1.5}\n]}\n')
    nat.tuner_clear()
    nat.tuner_load(p)
    out = tmp_path / "out.json"
    nat.tuner_save(out)
    text = out.read_text()
    assert '"log2n": 16' in text and '"j_segments": 8' in text and '"targets_per_thread": 2' in text
    bad = tmp_path / "bad.json"
    bad.write_text('{"entries": [{"mode": 2, "log2n": 16, "targets_per_thread": 3, "j_segments": 1}]}')
    with pytest.raises(nat.NBError):
        nat.tuner_load(bad)
    nat.tuner_clear()


def test_cpu_device_rejected(pkg):
    data = pkg.StarData([[0.0, 0.0, 0.0]], Mass=[1.0])
    with pytest.raises(NotImplementedError):
        pkg.Simulation(data, device=pkg.CPU(), TimeStep=1e-3)
