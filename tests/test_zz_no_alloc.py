"""nb_compute_force_noalloc (no device allocation, fixed scratch array) against
nb_compute_force on the same particles, through the C-ABI, on two calls in a
row with the second smaller (needs an MI355X).  Device arrays are torch
tensors: torch only provides the device memory and the stream here."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _existing(pkg, mode, pos, mass, h, G, pot):
    nat = pkg._native
    with nat.Context(len(mass), nat.MODES[mode]) as c:
        c.set_params(G, [h] * 6)
        c.upload(pos, None, mass)
        c.compute_force()
        if pot:
            c.compute_potential(self_term=False)
        d = c.download(pos=False, vel=False, acc=True, pot=pot)
    return d["acc"], d.get("pot")


def _noalloc(pkg, torch, mode, pos, mass, h, G, pot, per_particle_h):
    nat = pkg._native
    n = len(mass)
    x, y, z, m = (_dev(torch, pos[:, 0]), _dev(torch, pos[:, 1]), _dev(torch, pos[:, 2]),
                  _dev(torch, mass))
    hh = _dev(torch, np.full(n, h)) if per_particle_h else None
    out = [torch.full((n,), float("nan"), dtype=torch.float64, device="cuda") for _ in range(4)]
    stream = torch.cuda.current_stream().cuda_stream
    nat.compute_force_noalloc(nat.MODES[mode], n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                              m.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                              pot=out[3].data_ptr() if pot else 0,
                              h=hh.data_ptr() if hh is not None else 0, h_const=h, G=G,
                              stream=stream)
    torch.cuda.synchronize()
    acc = np.stack([o.cpu().numpy() for o in out[:3]], axis=1)
    return acc, (out[3].cpu().numpy() if pot else None)


@pytest.mark.parametrize("mode,tol", [("split", 1e-6), ("fp64", 1e-12)])
def test_noalloc_matches_compute_force(pkg, orc, mode, tol):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    G, h = 1.3, 0.01
    # two calls in a row, the second with a smaller input (the scratch keeps
    # the first call's contents; nothing in it has to start at zero)
    for k, (n, seed, pot, per_h) in enumerate([(4096, 1, True, False), (1000, 2, True, True),
                                               (777, 3, False, False)]):
        pos, _, mass = pkg.ics.plummer(n, seed=seed)
        print(f"noalloc {mode} call {k}: n={n} pot={pot} per-particle h={per_h}", flush=True)
        acc_ref, pot_ref = _existing(pkg, mode, pos, mass, h, G, pot)
        acc, phi = _noalloc(pkg, torch, mode, pos, mass, h, G, pot, per_h)
        assert np.all(np.isfinite(acc))
        e = orc.rel_err(acc, acc_ref)
        print(f"  vs nb_compute_force: median {np.median(e):.3e} max {e.max():.3e}", flush=True)
        assert np.median(e) <= tol and e.max() <= 100 * tol
        eo = orc.rel_err(acc, orc.compute_force(pos, mass, [h] * 6, G))
        assert np.median(eo) <= tol
        if pot:
            ep = np.abs(phi - pot_ref) / np.abs(pot_ref)
            assert np.median(ep) <= tol and ep.max() <= 100 * tol
