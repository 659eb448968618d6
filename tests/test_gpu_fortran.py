"""The Fortran host path (iso_c_binding module atomistica_gpu) drives the
HIP kernels through init / bind_to / energy_and_forces; a perfect diamond Si
crystal must give the closed-form Tersoff energy (-4.629593127667 eV/atom,
checked independently in test_oracle_potentials) and zero forces."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "atomistica_amd", "_build", "tersoff_check")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_host_tersoff_diamond():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    vals = {}
    for line in out.stdout.splitlines():
        k, v = line.rsplit(None, 1)
        vals[k.strip()] = float(v)
    assert abs(vals["E/atom"] - (-4.629593127667)) < 1e-10
    assert vals["max|f|"] < 1e-9
    assert abs(vals["sum(epot_per_at)-epot"]) < 1e-9
    assert abs(vals["E/atom TersoffScr"] - (-4.629593127667)) < 1e-10
    assert vals["max|f| TersoffScr"] < 1e-9


DROPIN = os.path.join(ROOT, "atomistica_amd", "_build", "dropin_check")


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="Fortran drop-in driver not built")
def test_fortran_dropin_types_tersoff(tmp_path):
    """The reference-shaped Fortran types (particles_t / neighbors_t of
    python_particles.f90:84-170 and python_neighbors.f90:51-139, filled by
    the host, changes signalled with I_changed_positions) with Tersoff
    through bind_to / energy_and_forces (default_compute_func.f90:25-45):
    energy, virial, forces and per-atom energies against the oracle at 1e-10,
    the per-atom and per-bond sums against the totals, and a second call
    after one atom moved in place."""
    import numpy as np
    from oracle import oracle as O
    from atomistica_amd import parameters as P
    from atomistica_amd import workloads as W

    pos, cell, pbc, _ = W.diamond_si(3, sigma=0.08, seed=4)
    nat = len(pos)
    fn = tmp_path / "structure.txt"
    with open(fn, "w") as fh:
        fh.write("%d\n" % nat)
        fh.write(" ".join("%.17g" % v for v in np.asarray(cell).reshape(-1)) + "\n")
        fh.write(" ".join("1" if b else "0" for b in pbc) + "\n")
        for x in pos:
            fh.write("14 %.17g %.17g %.17g\n" % tuple(x))
    out = subprocess.run([DROPIN, str(fn)], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr)
    assert out.returncode == 0
    res = {"A": {"F": {}}, "B": {"F": {}}}
    for line in out.stdout.splitlines():
        t = line.split()
        if len(t) < 3 or t[0] not in res:
            continue
        if t[1] == "F":
            res[t[0]]["F"][int(t[2]) - 1] = [float(v) for v in t[3:7]]
        else:
            res[t[0]][t[1]] = [float(v) for v in t[2:]]
    prm = P.Tersoff_PRB_39_5566_Si_C
    el = np.ones(nat, np.int32)                  # Si in the (C, Si) set
    moved = pos.copy()
    moved[0] += [0.013, -0.021, 0.017]
    for tag, r in (("A", pos), ("B", moved)):
        nl = O.nl_build(r, cell, pbc, 3.0)
        e, f, w, epa, wpa, epb, fpb, wpb = O.tersoff(prm, r, el, cell, nl, per_at=True, per_bond=True)
        d = res[tag]
        fg = np.array([d["F"][i][:3] for i in range(nat)])
        eg = np.array([d["F"][i][3] for i in range(nat)])
        scale = max(1.0, float(np.abs(f).max()))
        assert abs(d["E"][0] - e) <= 1e-10 * abs(e)
        assert np.abs(np.array(d["W"]).reshape(3, 3, order="F") - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
        assert np.abs(fg - f).max() <= 1e-10 * scale
        assert np.abs(eg - epa).max() <= 1e-10 * max(1.0, np.abs(epa).max())
        assert abs(d["sum_epb"][0] - e) <= 1e-10 * abs(e)
        assert abs(d["sum_epa"][0] - e) <= 1e-10 * abs(e)
        assert np.abs(np.array(d["sum_wpa"]).reshape(3, 3, order="F") - wpa.sum(0)).max() <= 1e-9
        assert np.abs(np.array(d["sum_wpb"]).reshape(3, 3, order="F") - wpb.sum(0)).max() <= 1e-9
    assert res["A"]["E"][0] != res["B"]["E"][0]
