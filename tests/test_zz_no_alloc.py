"""The allocation-free energy_and_forces (agpu_potential_energy_and_forces_noalloc,
no counterpart in the reference) through the C-ABI against
agpu_potential_energy_and_forces on the same bound potential and list: the
same kernels, so the results agree bit for bit, on two calls in a row with
the second system smaller than the first (the shared scratch array keeps
the first call's contents).  Also its refusals: a stale list, a screened
class, and a system beyond the scratch size."""
import numpy as np
import pytest
import torch

from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError, call, ERR_ARG, ERR_STATE, ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _system(n, seed, kind="tersoff"):
    rng = np.random.default_rng(seed)
    if kind == "brenner":
        pos, cell, pbc, _ = W.diamond_si(n, a=4.6, sigma=0.1, seed=seed)
        prm = P.Albe_PRB_65_195124_PtC
        Z = np.array([P.ATOMIC_NUMBERS[prm["el"][e]] for e in rng.integers(0, 2, len(pos))], np.int32)
        pot = N.Brenner(**{k: v for k, v in prm.items() if k not in ("__ref__", "family")})
    else:
        pos, cell, pbc, _ = W.diamond_si(n, sigma=0.1, seed=seed)
        Z = np.where(rng.random(len(pos)) < 0.3, 6, 14).astype(np.int32)
        pot = N.Tersoff()
    p = N.from_atoms(pos, cell, pbc, Z)
    r = torch.tensor(pos, dtype=torch.float64, device=DEV)
    p.set_positions_device(r)
    nl = N.Neighbors(100)
    pot.bind_to(p, nl)
    return p, nl, pot, r


def _outputs(nat):
    return (torch.zeros(1, dtype=torch.float64, device=DEV),
            torch.zeros((nat, 3), dtype=torch.float64, device=DEV),
            torch.zeros(9, dtype=torch.float64, device=DEV))


def _noalloc(p, nl, pot, mask=None):
    e, f, w = _outputs(p.nat)
    call("agpu_potential_energy_and_forces_noalloc", pot._h, p._h, nl._h, N._dev(e), N._dev(f), N._dev(w),
         N._dev(mask))
    torch.cuda.synchronize()
    return e.cpu().numpy(), f.cpu().numpy(), w.cpu().numpy()


def _regular(p, nl, pot, mask=None):
    e, f, w = _outputs(p.nat)
    pot.energy_and_forces_device(p, nl, f, e, w, mask=mask)
    torch.cuda.synchronize()
    return e.cpu().numpy(), f.cpu().numpy(), w.cpu().numpy()


@pytest.mark.parametrize("kind", ["tersoff", "brenner"])
def test_noalloc_equals_energy_and_forces(kind):
    N.use_torch_stream()
    for n, seed in ((5, 1), (3, 2)):            # 1000 atoms, then 216
        p, nl, pot, r = _system(n, seed, kind)
        print("noalloc", kind, p.nat, "atoms", flush=True)
        nl.update(p)
        mask = torch.tensor((np.arange(p.nat) % 4 != 0).astype(np.int32), device=DEV)
        for m in (None, mask):
            a = _noalloc(p, nl, pot, m)
            b = _regular(p, nl, pot, m)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        assert np.abs(a[1]).max() > 0


def test_noalloc_refusals():
    N.use_torch_stream()
    p, nl, pot, r = _system(3, 3)
    nl.update(p)
    _noalloc(p, nl, pot)
    # positions moved: the list is stale and nothing is rebuilt
    r += 0.01
    p.changed_positions()
    with pytest.raises(AgpuError) as ei:
        _noalloc(p, nl, pot)
    assert ei.value.code == ERR_STATE
    nl.update(p)
    _noalloc(p, nl, pot)
    # a screened class is not covered
    ps = N.from_atoms(*W.diamond_si(2, sigma=0.05, seed=4))
    nls = N.Neighbors(100)
    scr = N.TersoffScr()
    scr.bind_to(ps, nls)
    nls.update(ps)
    with pytest.raises(AgpuError) as ei:
        _noalloc(ps, nls, scr)
    assert ei.value.code == ERR_UNSUPPORTED


def test_noalloc_size_limit():
    """729 000 Si atoms (2.9 M list entries, about 88 MB of scratch) exceed
    the 64 MiB scratch."""
    N.use_torch_stream()
    pos, cell, pbc, Z = W.diamond_si(45, sigma=0.02, seed=5)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(16)
    pot = N.Tersoff()
    pot.bind_to(p, nl)
    nl.update(p)
    e, f, w = _outputs(p.nat)
    with pytest.raises(AgpuError) as ei:
        call("agpu_potential_energy_and_forces_noalloc", pot._h, p._h, nl._h, N._dev(e), N._dev(f), N._dev(w),
             None)
    assert ei.value.code == ERR_ARG
    assert "maximum" in str(ei.value)
