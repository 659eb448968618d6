"""Per-bond properties (epot_per_bond, f_per_bond, wpot_per_bond) of the
bond-order classes and REBO2 on the GPU against the oracle through the
C-ABI, entry by entry in list order, within 1e-10 relative (scaled by the
largest entry): bop_kernel.f90:1370-1373,1406-1408,1511-1512,1593-1594 and
bop_kernel_rebo2.f90:2546-2547,2571-2572,2762-2763.  The oracle runs on the
list the GPU built, after that list is checked identical to the oracle's
own.  Also: the device-pointer path equals the host path, per-bond arrays
accumulate (+=), the mixed form stays within its force bar, and the classes
the reference gives no per-bond output (EAM, Coulomb) refuse them."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError
from helpers import bop_cutoff, read_cfg

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, rel=REL):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max()) <= rel * max(1.0, float(np.abs(b).max()))


def _cls(prm):
    fam = prm.get("family")
    scr = "or1" in prm
    if fam == "kumagai":
        return N.KumagaiScr if scr else N.Kumagai
    if fam == "brenner":
        return N.BrennerScr if scr else N.Brenner
    if fam == "juslin":
        return N.JuslinScr if scr else N.Juslin
    return N.TersoffScr if scr else N.Tersoff


def _build(prm, pos, cell, pbc, el, precision="fp64"):
    names = prm["el"] if isinstance(prm["el"], (list, tuple)) else [prm["el"]]
    Z = np.array([P.ATOMIC_NUMBERS[names[e]] for e in el], np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(300)
    pot = _cls(prm)(**{k: v for k, v in prm.items() if k not in ("__ref__", "family")})
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    return p, nl, pot


def _si(n=3, sigma=0.1, seed=1):
    pos, cell, pbc, _ = W.diamond_si(n, sigma=sigma, seed=seed)
    return pos, cell, pbc, np.ones(len(pos), int)


def _case(name):
    rng = np.random.default_rng(11)
    if name == "tersoff_si":
        return (P.Tersoff_PRB_39_5566_Si_C,) + _si() + (None,)
    if name == "tersoff_sic_mask":
        pos, cell, pbc, _ = _si(seed=2)
        el = rng.integers(0, 2, len(pos))
        return P.Tersoff_PRB_39_5566_Si_C, pos, cell, pbc, el, (rng.random(len(pos)) < 0.6).astype(np.int32)
    if name == "kumagai":
        pos, cell, pbc, _ = _si(seed=3)
        return P.Kumagai_CompMaterSci_39_457_Si, pos, cell, pbc, np.zeros(len(pos), int), None
    if name == "brenner_ptc":
        pos, cell, pbc, _ = W.diamond_si(3, a=4.6, sigma=0.1, seed=2)
        return P.Albe_PRB_65_195124_PtC, pos, cell, pbc, rng.integers(0, 2, len(pos)), None
    if name == "juslin_wch":
        pos, cell, pbc, _ = W.diamond_si(3, a=4.4, sigma=0.1, seed=4)
        return P.Juslin_JAP_98_123520_WCH, pos, cell, pbc, rng.integers(0, 3, len(pos)), None
    if name == "tersoff_scr_aC":
        pos, cell, pbc, _ = read_cfg("aC_small.cfg")
        pos = pos + rng.normal(0, 0.03, pos.shape)
        return P.Tersoff_PRB_39_5566_Si_C__Scr, pos, cell, pbc, np.zeros(len(pos), int), None
    if name == "brenner_scr_mask":
        pos, cell, pbc, _ = read_cfg("aC_small.cfg")
        pos = pos + rng.normal(0, 0.03, pos.shape)
        return (P.Erhart_PRB_71_035211_SiC__Scr, pos, cell, pbc, np.zeros(len(pos), int),
                (rng.random(len(pos)) < 0.7).astype(np.int32))
    if name == "juslin_scr":
        pos, cell, pbc, _ = W.diamond_si(3, a=4.4, sigma=0.12, seed=5)
        return P.Juslin_JAP_98_123520_WCH__Scr, pos, cell, pbc, rng.integers(0, 3, len(pos)), None
    raise KeyError(name)


BOP_CASES = ["tersoff_si", "tersoff_sic_mask", "kumagai", "brenner_ptc", "juslin_wch", "tersoff_scr_aC",
             "brenner_scr_mask", "juslin_scr"]


@pytest.mark.parametrize("name", BOP_CASES)
def test_bop_per_bond_parity(name):
    prm, pos, cell, pbc, el, mask = _case(name)
    print("per-bond case", name, len(pos), flush=True)
    p, nl, pot = _build(prm, pos, cell, pbc, el)
    e, f, w, _, epb, fpb, _, wpb = pot.energy_and_forces(p, nl, epot_per_bond=True, f_per_bond=True,
                                                         wpot_per_bond=True, mask=mask)
    glist = nl.get_list(p)
    olist = O.nl_build(pos, cell, pbc, bop_cutoff(prm), avgn=300)
    for a, b in zip(glist, olist):
        assert np.array_equal(a, b)
    re_, rf, rw, repb, rfpb, rwpb = O.tersoff(prm, pos, np.asarray(el, np.int32), cell, olist, mask=mask,
                                              per_bond=True)
    assert abs(e - re_) <= REL * max(1.0, abs(re_))
    assert close(f, rf) and close(w, rw)
    assert len(epb) == len(repb)
    assert close(epb, repb) and close(fpb, rfpb) and close(wpb, rwpb)
    assert abs(epb.sum() - e) <= 1e-12 * max(1.0, abs(e))


@pytest.mark.parametrize("dihedral", [False, True], ids=["plain", "torsion"])
@pytest.mark.parametrize("name", ["aC_small", "random_CH"])
def test_rebo2_per_bond_parity(name, dihedral):
    if name == "aC_small":
        pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    else:
        pos, cell, pbc, Z = W.random_ch(300, 100, 2.3, seed=8, strain=0.15)
    print("rebo2 per-bond case", name, len(pos), dihedral, flush=True)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.Rebo2(dihedral=dihedral)
    pot.bind_to(p, nl)
    e, f, w, _, epb, fpb, _, wpb = pot.energy_and_forces(p, nl, epot_per_bond=True, f_per_bond=True,
                                                         wpot_per_bond=True)
    glist = nl.get_list(p)
    olist = O.nl_build(pos, cell, pbc, 2.0)
    for a, b in zip(glist, olist):
        assert np.array_equal(a, b)
    re_, rf, rw, repb, rfpb, rwpb = O.Rebo2(dihedral=dihedral)(pos, Z, cell, olist, per_bond=True)
    assert abs(e - re_) <= REL * max(1.0, abs(re_))
    assert close(f, rf) and close(w, rw)
    assert close(epb, repb) and close(fpb, rfpb) and close(wpb, rwpb)
    assert np.array_equal(epb != 0, repb != 0)


def test_device_path_accumulates():
    """Device pointers, per-bond arrays accumulated: a second call into the
    same arrays doubles them; equal to the host path."""
    prm, pos, cell, pbc, el, mask = _case("tersoff_scr_aC")
    p, nl, pot = _build(prm, pos, cell, pbc, el)
    e, f, w, _, epb, fpb, _, wpb = pot.energy_and_forces(p, nl, epot_per_bond=True, f_per_bond=True,
                                                         wpot_per_bond=True)
    nb = len(epb)
    dev = torch.device("cuda:0")
    N.use_torch_stream()
    r = torch.tensor(pos, dtype=torch.float64, device=dev)
    p.set_positions_device(r)
    fd = torch.zeros((len(pos), 3), dtype=torch.float64, device=dev)
    ed = torch.zeros(1, dtype=torch.float64, device=dev)
    wd = torch.zeros(9, dtype=torch.float64, device=dev)
    eb = torch.zeros(nb, dtype=torch.float64, device=dev)
    fb = torch.zeros((nb, 3), dtype=torch.float64, device=dev)
    wb = torch.zeros(9 * nb, dtype=torch.float64, device=dev)
    for _ in range(2):
        pot.energy_and_forces_device(p, nl, fd, ed, wd, epot_per_bond=eb, f_per_bond=fb, wpot_per_bond=wb)
    torch.cuda.synchronize()
    assert close(eb.cpu().numpy(), 2 * epb, 1e-13)
    assert close(fb.cpu().numpy(), 2 * fpb, 1e-13)
    assert close(wb.cpu().numpy().reshape(nb, 3, 3).transpose(0, 2, 1), 2 * wpb, 1e-13)


def test_mixed_per_bond_error():
    prm, pos, cell, pbc, el, _ = _case("tersoff_si")
    out = {}
    for prec in ("fp64", "mixed"):
        p, nl, pot = _build(prm, pos, cell, pbc, el, precision=prec)
        out[prec] = pot.energy_and_forces(p, nl, epot_per_bond=True, f_per_bond=True, wpot_per_bond=True)
    e64, e32 = out["fp64"][4], out["mixed"][4]
    f64, f32 = out["fp64"][5], out["mixed"][5]
    assert np.abs(e32 - e64).max() < 1e-5
    assert np.sqrt(np.mean((f32 - f64) ** 2)) < 1e-5


def test_per_bond_refused_by_classes_without_it():
    pos, cell, pbc, Z = W.fcc_cu(3, sigma=0.05, seed=2)
    import tempfile, os
    d = tempfile.mkdtemp()
    fn = os.path.join(d, "cu.eam.alloy")
    with open(fn, "w") as fh:
        fh.write(W.fs_cu_setfl_text())
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.TabulatedAlloyEAM(fn=fn)
    pot.bind_to(p, nl)
    with pytest.raises(AgpuError):
        pot.energy_and_forces(p, nl, epot_per_bond=True)


def test_calculator_set_per_bond():
    """The ASE calculator's set_per_bond (aseinterface.py:290-297): per-bond
    energies in get_neighbors() order that add up to the energy and match
    the oracle."""
    from atomistica_amd import aseinterface as A
    pos, cell, pbc, _ = W.diamond_si(3, sigma=0.08, seed=9)
    a = A.Structure(pos, cell, True, np.full(len(pos), 14))
    c = A.Tersoff()
    a.calc = c
    c.set_per_bond(epot=True, f=True, wpot=True)
    e = a.get_potential_energy()
    olist = O.nl_build(pos, cell, pbc, 3.0)
    re_, rf, rw, repb, rfpb, rwpb = O.tersoff(P.Tersoff_PRB_39_5566_Si_C, pos, np.ones(len(pos), np.int32),
                                              cell, olist, per_bond=True)
    assert close(c.epot_per_bond, repb) and close(c.f_per_bond, rfpb) and close(c.wpot_per_bond, rwpb)
    assert abs(c.epot_per_bond.sum() - e) < 1e-10
    i, j, absdr = c.get_neighbors()
    assert len(i) == len(c.epot_per_bond)
