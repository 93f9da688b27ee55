"""The drop-in from the Fortran side.

* gravity_driver / sphgrav_check: executables built with amdflang from
  starsmasher_amd/fortran that call the library as StarSmasher's call sites
  do (implicit interfaces, by reference, grav.f / balAV3.f conventions), and
  through the iso_c_binding module.
* refsite_gravity: the reference's own grav.f (gravforce) and gpu_grav.f,
  compiled unmodified, called through their common blocks on a real
  StarSmasher star (example_input/collision/sph.start1u, N = 19,938).
* starsmasher_gpu: the whole StarSmasher program, compiled unmodified and
  linked against libsphgrav_hip.so, runs a short collision and writes the
  dumps its own CPU-gravity build writes (tests/golden/collision_run.npz).

Results are checked against the oracle, the potentials the reference stored
in its dumps, and the reference program's own output.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from starsmasher_amd import ics

from conftest import ROOT, rel_errors

pytestmark = pytest.mark.gpu

FDIR = os.path.join(ROOT, "starsmasher_amd", "fortran")
GOLDEN = os.path.join(ROOT, "tests", "golden")
MPICH_LIB = "/opt/conda/lib/libmpifort.so"


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


def _ref_exe(name):
    """A reference-built executable, or a skip with the reason it cannot run."""
    path = os.path.join(FDIR, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built: it compiles the reference's own sources "
                    "(make -C starsmasher_amd/fortran ref, where the source tree is present)")
    if not os.path.exists(MPICH_LIB):
        pytest.skip(f"MPICH ({MPICH_LIB}) is absent; the reference's sources include mpif.h")
    return path


@pytest.mark.parametrize("nlo,nhi", [(1, 19938), (4001, 13000)])
def test_reference_gravforce_call_site(tmp_path, golden_dumps, nlo, nhi):
    """The reference's own gravforce (grav.f:1-31: range = +-4 hp^2, then
    firsthalf_grav_forces over the rank's chunk) and the lasthalf call of
    uvdots (balAV3.f:418), compiled unmodified, on the collision star the
    reference ships: potentials against the grpot the reference stored in
    the dump, accelerations and potentials against the FP64 oracle."""
    exe = _ref_exe("refsite_gravity")
    d = golden_dumps["collision_start1u"]
    n = d["x"].size
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, {"x": d["x"], "y": d["y"], "z": d["z"], "m": d["am"], "hp": d["hp"],
                       "u": d["u"]}, nlo, nhi, 0)
    env = dict(os.environ, SPHGRAV_QUIET="1")
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "1sthalf:" in r.stdout            # grav.f's own timing line
    out = _read_output(fout, n)
    lo, hi = nlo - 1, nhi
    assert np.all(out[:, :lo] == -7777.0) and np.all(out[:, hi:] == -7777.0)
    rel = np.abs(out[3, lo:hi] - d["grpot"][lo:hi]) / np.abs(d["grpot"][lo:hi])
    ref = oracle.direct(d["x"], d["y"], d["z"], d["am"], d["range2"], nkernel=0,
                        rows=np.arange(lo, hi))
    ea, ep = rel_errors(out[:3, lo:hi].T, out[3, lo:hi], np.stack(ref[:3], -1), ref[3])
    print(f"[grav.f call site {nlo}..{nhi}] grpot vs dump max {rel.max():.3e}; vs oracle acc max "
          f"{ea.max():.3e} pot max {ep.max():.3e}")
    assert rel.max() < 4e-7                  # the dump's own FP32 floor
    assert ea.max() < 1e-5 and ep.max() < 1e-6


def test_reference_gravforce_call_site_at_2p20(tmp_path):
    """The reference's gravforce and the lasthalf call of uvdots (grav.f,
    gpu_grav.f, compiled unmodified against starsmasher.h with nmax raised to
    1,100,000, the one change a run at N = 2^20 needs) on the 2^20-particle
    giant star with its central point mass, for the second half of the
    particles as one rank's chunk.  The result is bit-identical to the
    library's own evaluation of that i-range (grav.f forms range2 as
    4 hp^2, negated for u = 0, which is what the set holds), and agrees with
    the oracle on sampled rows and at the point mass.  grav.f's automatic
    range(ntot) takes 8 MB of stack at this size: the run gets 64 MB."""
    import resource
    exe = _ref_exe("refsite_gravity_1m")
    need = 64 << 20
    _, hard = resource.getrlimit(resource.RLIMIT_STACK)
    if hard != resource.RLIM_INFINITY and hard < need:
        pytest.skip(f"the stack hard limit ({hard} B) is below the {need} B grav.f needs at 2^20")
    d = ics.giant_with_core(1 << 20)
    n = d["x"].size
    nlo, nhi = n // 2 + 1, n
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, d, nlo, nhi, 0)
    env = dict(os.environ, SPHGRAV_QUIET="1")
    env.pop("SPHGRAV_PRECISION", None)
    print(f"[grav.f at 2^20] running rows {nlo}..{nhi}", flush=True)
    r = subprocess.run(["bash", "-c", f'ulimit -s {need >> 10} && "$0" "$1" "$2"', exe,
                        str(fin), str(fout)], capture_output=True, text=True, timeout=200,
                       env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "1sthalf:" in r.stdout
    out = _read_output(fout, n)
    lo, hi = nlo - 1, nhi
    assert np.all(out[:, :lo] == -7777.0)
    from starsmasher_amd import Gravity
    with Gravity(0) as g:
        want = g.compute(d["x"], d["y"], d["z"], d["m"], d["range2"], ibeg=lo, ni=hi - lo)
    for a, b in zip(out[:, lo:hi], want):
        np.testing.assert_array_equal(a, b)
    rng = np.random.default_rng(7)
    rows = np.unique(np.concatenate([rng.choice(np.arange(lo, hi), 48, replace=False), [n - 1]]))
    ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=0, rows=rows)
    ea, ep = rel_errors(out[:3, rows].T, out[3, rows], np.stack(ref[:3], -1), ref[3])
    print(f"[grav.f at 2^20] vs oracle on {rows.size} rows: acc max {ea.max():.3e} "
          f"pot max {ep.max():.3e}")
    assert ea.max() < 1e-5 and ep.max() < 1e-6


def _collision_fixture():
    g = np.load(os.path.join(GOLDEN, "collision_run.npz"))
    return {k: g[k] for k in g.files}


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_starsmasher_program_collision_run(tmp_path, precision):
    """The whole StarSmasher program (the GPUOBJS of pbe/src/Makefile:24-38,
    compiled unmodified) with libsphgrav_hip.so in place of the CUDA library:
    a short collision of two copies of the reference's small star, about
    five steps of gravity, hydrodynamics and integration.  Its dumps match
    those of the same program built with its own CPU gravity
    (get_gravity_using_cpus, FP64).  The run is not chaotic over these
    steps: two CPU runs that differ only in the order of the gravity sums
    (2 and 3 ranks) agree to 3e-15 in position and 1.2e-14 in potential, so
    the differences seen are the library's own arithmetic."""
    exe = _ref_exe("starsmasher_gpu")
    sys.path.insert(0, GOLDEN)
    from collision_case import write_case
    from make_golden import read_dump
    want = _collision_fixture()
    write_case(str(tmp_path), want["start_bytes"], ngravprocs=1)
    env = dict(os.environ, SPHGRAV_QUIET="1")
    if precision == "fp64":
        env["SPHGRAV_PRECISION"] = "fp64"
    else:
        env.pop("SPHGRAV_PRECISION", None)
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=100,
                       env=env)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "with gpu" in r.stdout                 # main.f:270-275 took the GPU branch
    d0 = read_dump(str(tmp_path / "out0000.sph"))
    d1 = read_dump(str(tmp_path / "out0001.sph"))
    assert d1["x"].size == want["out0001/x"].size
    e0 = np.abs(d0["grpot"] - want["out0000/grpot"]) / np.abs(want["out0000/grpot"])
    dt = abs(float(d1["t"]) - float(want["out0001/t"])) / float(want["out0001/t"])
    pos = np.stack([d1[k] for k in ("x", "y", "z")], 1)
    pos_w = np.stack([want[f"out0001/{k}"] for k in ("x", "y", "z")], 1)
    vel = np.stack([d1[k] for k in ("vx", "vy", "vz")], 1)
    vel_w = np.stack([want[f"out0001/{k}"] for k in ("vx", "vy", "vz")], 1)
    # position and velocity errors relative to the set's extent and speed scale
    ep = np.max(np.linalg.norm(pos - pos_w, axis=1)) / np.max(np.linalg.norm(pos_w, axis=1))
    ev = np.max(np.linalg.norm(vel - vel_w, axis=1)) / np.max(np.linalg.norm(vel_w, axis=1))
    e1 = np.abs(d1["grpot"] - want["out0001/grpot"]) / np.abs(want["out0001/grpot"])
    print(f"[starsmasher {precision}] out0000 grpot max {e0.max():.3e}; out0001 t {dt:.3e} "
          f"pos {ep:.3e} vel {ev:.3e} grpot max {e1.max():.3e}")
    tol = {"fp64": (1e-12, 1e-12, 1e-11), "fp32c": (1e-6, 1e-6, 1e-5)}[precision]
    assert e0.max() <= tol[0] and e1.max() <= tol[0] * 10
    assert dt <= tol[1] and ep <= tol[1] and ev <= tol[2]
