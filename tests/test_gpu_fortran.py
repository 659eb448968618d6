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
