"""The drop-in from the Fortran side: executables built with amdflang from
starsmasher_amd/fortran call the library exactly as StarSmasher's call sites
do (implicit interfaces, by reference, grav.f / balAV3.f conventions), and
through the iso_c_binding module.  Results are checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from starsmasher_amd import ics

from conftest import rel_errors

pytestmark = pytest.mark.gpu

FDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "starsmasher_amd", "fortran")


def _write_input(path, d, nlo, nhi, nkernel):
    n = d["x"].size
    with open(path, "wb") as f:
        np.array([n, nlo, nhi, nkernel], dtype="<i4").tofile(f)
        for k in ("x", "y", "z", "m", "hp", "u"):
            np.ascontiguousarray(d[k], dtype="<f8").tofile(f)


def _read_output(path, n, narrays=4):
    a = np.fromfile(path, dtype="<f8")
    assert a.size == narrays * n
    return a.reshape(narrays, n)


@pytest.mark.parametrize("exe", ["gravity_driver", "sphgrav_check"])
@pytest.mark.parametrize("nlo,nhi,nk", [(1, 3000, 0), (701, 2200, 0), (1, 3000, 2)])
def test_fortran_callers(tmp_path, exe, nlo, nhi, nk):
    path = os.path.join(FDIR, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (make -C starsmasher_amd/fortran)")
    d = ics.plummer(3000, seed=21)
    d["u"][-2:] = 0.0                           # two point particles
    d["range2"] = ics.range2_of(d["hp"], d["u"])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, d, nlo, nhi, nk)
    env = dict(os.environ, SPHGRAV_QUIET="1")
    r = subprocess.run([path, str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    if exe == "sphgrav_check":
        assert "status-check ok" in r.stdout
    out = _read_output(fout, 3000, 5 if exe == "sphgrav_check" else 4)
    lo, hi = nlo - 1, nhi
    assert np.all(out[:, :lo] == -7777.0) and np.all(out[:, hi:] == -7777.0)
    ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=nk,
                        rows=np.arange(lo, hi))
    ea, ep = rel_errors(out[:3, lo:hi].T, out[3, lo:hi], np.stack(ref[:3], -1), ref[3])
    print(f"[fortran {exe}] acc max {ea.max():.3e} pot max {ep.max():.3e}")
    assert ea.max() < 1e-5 and ep.max() < 1e-6
    if exe == "sphgrav_check":
        psi = out[4]
        assert np.all(psi[:lo] == -7777.0) and np.all(psi[hi:] == -7777.0)
        want = oracle.psi(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=nk,
                          rows=np.arange(lo, hi))
        gas = want != 0.0
        assert np.all(psi[lo:hi][~gas] == 0.0)
        assert np.max(np.abs(psi[lo:hi][gas] - want[gas]) / np.abs(want[gas])) < 1e-11
