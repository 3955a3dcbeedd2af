"""Gadget-1/2 snapshot format (src/base/Output.jl:9-50 via AstroIO's
write_gadget2 / read_gadget2), checked on the CPU: the reference's own initial
conditions file is rewritten byte for byte, and format-2 / big-endian files
round-trip."""
import filecmp
import os
import struct

import numpy as np
import pytest

IC = "/root/reference/examples/04-collisions/galaxy_littleendian.dat"


@pytest.fixture(scope="module")
def snap(pkg):
    from astronbodysim_jl_amd import snapshot
    return snapshot


def _sample(n=300, seed=0):
    rng = np.random.default_rng(seed)
    coll = rng.integers(1, 7, n).astype(np.uint8)
    return (rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), np.arange(1, n + 1),
            rng.uniform(0.5, 1.5, n), coll)


def test_header_layout(snap):
    h = snap.HeaderGadget2(npart=[1, 2, 3, 4, 5, 6], time=0.5, num_files=1, Omega0=0.3)
    b = h.pack()
    assert len(b) == 256
    assert struct.unpack("<6i", b[:24]) == (1, 2, 3, 4, 5, 6)
    assert struct.unpack("<d", b[72:80])[0] == 0.5          # time after npart and mass[6]
    assert struct.unpack("<i", b[124:128])[0] == 1          # num_files
    assert struct.unpack("<d", b[136:144])[0] == 0.3        # Omega0 after BoxSize
    assert snap.HeaderGadget2.unpack(b) == h


@pytest.mark.skipif(not os.path.exists(IC), reason="reference initial conditions not present")
def test_reference_ic_rewrites_byte_identical(snap, tmp_path):
    """examples/04-collisions/galaxy_littleendian.dat: 40000 halo + 20000 disk
    particles, masses in the header, format 1, uint32 IDs."""
    h, b = snap.read_gadget2(IC)
    assert h.npart == [0, 40000, 20000, 0, 0, 0] and h.time == 0.0
    assert b["POS"].shape == (60000, 3)
    assert np.array_equal(b["ID"], np.arange(1, 60001))
    assert np.all(b["Mass"][:40000] == h.mass[1]) and np.all(b["Mass"][40000:] == h.mass[2])
    assert np.array_equal(b["Collection"], np.repeat([2, 3], [40000, 20000]))
    out = tmp_path / "ic.dat"
    snap.write_gadget2(str(out), h, b["POS"], b["VEL"], b["ID"], b["Mass"], b["Collection"],
                       format2=False)
    assert filecmp.cmp(str(out), IC, shallow=False)


@pytest.mark.parametrize("format2,endian", [(True, "<"), (False, "<"), (True, ">"), (False, ">")])
def test_round_trip(snap, tmp_path, format2, endian):
    pos, vel, ids, mass, coll = _sample()
    h = snap.header_for(coll, 1.25)
    f = str(tmp_path / "s.g2")
    pot = np.linspace(-1, 0, coll.size)
    acc = pos * 2.0
    snap.write_gadget2(f, h, pos, vel, ids, mass, coll, Acc=acc, Potential=pot, format2=format2,
                       endian=endian)
    h2, b = snap.read_gadget2(f)
    assert h2 == h and h2.npart == snap.count_gadget_types(coll)
    order = np.argsort(coll, kind="stable")
    f32 = lambda a: np.asarray(a)[order].astype(np.float32).astype(np.float64)
    assert np.array_equal(b["POS"], f32(pos)) and np.array_equal(b["VEL"], f32(vel))
    assert np.array_equal(b["ID"], ids[order]) and np.array_equal(b["Mass"], f32(mass))
    assert np.array_equal(b["Collection"], coll[order])
    if format2:
        assert np.array_equal(b["POT"], f32(pot)) and np.array_equal(b["ACCE"], f32(acc))
    # gas particles carry U, RHO, HSML blocks: file size is fixed by the layout
    n, ng = coll.size, int(np.count_nonzero(coll == 1))
    rec = lambda nb: nb + 8 + (16 if format2 else 0)
    size = (rec(256) + 2 * rec(12 * n) + rec(4 * n) + rec(4 * n) + 3 * rec(4 * ng)
            + rec(4 * n) + rec(12 * n))
    assert os.path.getsize(f) == size


def test_large_ids_and_header_masses(snap, tmp_path):
    pos, vel, ids, mass, coll = _sample(50, seed=3)
    coll[:] = 2
    ids = ids + (1 << 33)
    h = snap.header_for(coll, 0.0)
    h.mass[1] = 0.75                                       # halo mass in the header: no MASS block
    f = str(tmp_path / "s.g2")
    snap.write_gadget2(f, h, pos, vel, ids, mass, coll, format2=True)
    h2, b = snap.read_gadget2(f)
    assert np.array_equal(b["ID"], ids) and np.all(b["Mass"] == 0.75)
    with open(f, "rb") as fh:
        assert b"MASS" not in fh.read()


def test_corrupt_and_mismatched_input(snap, tmp_path):
    pos, vel, ids, mass, coll = _sample(20)
    h = snap.header_for(coll, 0.0)
    h.npart[1] += 1
    with pytest.raises(ValueError):
        snap.write_gadget2(str(tmp_path / "x"), h, pos, vel, ids, mass, coll)
    bad = tmp_path / "bad"
    bad.write_bytes(struct.pack("<i", 12) + bytes(12) + struct.pack("<i", 12))
    with pytest.raises(ValueError):
        snap.read_gadget2(str(bad))


def test_gadget_internal_G(snap):
    """Gadget-2's internal units (kpc, 1e10 Msun, km/s) give G = 43007.1
    (Gadget-2 user guide, 'System of units')."""
    assert snap.gadget_internal_G() == pytest.approx(43007.1, rel=1e-6)
