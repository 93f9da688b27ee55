"""Write the golden fixtures under tests/golden/ from the reference checkout.

Run in the container that holds the reference (``python tests/golden/make_golden.py``).
It writes data only -- inputs and expected outputs:

* ``dumps.npz``: for each of the four distinct start dumps shipped in the
  reference's ``example_input/`` (``collision/sph.start{1,2}u`` and
  ``3_stars/sph.start{1,2}u``; the ``collision_elliptical`` copies are
  md5-identical), the per-particle x, y, z, am, hp, u and the ``grpot`` the
  reference's own gravity path stored.  The file format is the binary dump of
  pbe/src/output.f:53-92 (Fortran sequential records: header, one record per
  particle of 16 x f64 + i32 + f64 [+ 2 x f64 when ncooling != 0], trailer).

* ``collision_run.npz``: the short collision run of the whole StarSmasher
  program (collision_case.py) as the reference program with its own CPU
  gravity writes it: its start dump and the expected potentials, positions
  and velocities of its two output dumps.

* ``twin_vectors.npz``: outputs of the reference's own CPU twin
  ``get_gravity_using_cpus`` (pbe/src/cpu_grav.f) compiled from its source
  into oracle/_ref (``make -C oracle ref``), on small seeded inputs for
  nkernel 0/1/2, including point particles, a partial i-range and the
  nselfgravity=0 mode.  These keep the twin's answers available where
  amdflang or the reference checkout is absent.
"""
from __future__ import annotations

import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
DUMPS = [
    ("collision_start2u", "example_input/collision/sph.start2u"),
    ("3stars_start1u", "example_input/3_stars/sph.start1u"),
    ("3stars_start2u", "example_input/3_stars/sph.start2u"),
    ("collision_start1u", "example_input/collision/sph.start1u"),
]


def read_dump(path: str) -> dict:
    """Parse an output.f:53-92 dump (little-endian, 4-byte record markers)."""
    with open(path, "rb") as f:
        buf = f.read()
    recs = []
    off = 0
    while off < len(buf):
        (n1,) = struct.unpack_from("<i", buf, off)
        off += 4
        recs.append(buf[off:off + n1])
        off += n1
        (n2,) = struct.unpack_from("<i", buf, off)
        off += 4
        if n1 != n2:
            raise ValueError(f"{path}: corrupt record markers at byte {off}")
    ntot = struct.unpack_from("<i", recs[0], 0)[0]
    parts = recs[1:1 + ntot]
    if len(parts) != ntot or struct.unpack_from("<i", recs[1 + ntot], 0)[0] != ntot:
        raise ValueError(f"{path}: particle count / trailer mismatch")
    rec_len = len(parts[0])
    if rec_len not in (140, 156):
        raise ValueError(f"{path}: unexpected particle record of {rec_len} bytes")
    f64 = np.frombuffer(b"".join(p[:120] for p in parts), dtype="<f8").reshape(ntot, 15)
    # header: ntot, nnopt (i4); hco, hfloor, sep0, tf, dtout (f8); nout, nit (i4); t (f8)
    (t,) = struct.unpack_from("<d", recs[0], 56)
    return {
        "x": f64[:, 0].copy(), "y": f64[:, 1].copy(), "z": f64[:, 2].copy(),
        "am": f64[:, 3].copy(), "hp": f64[:, 4].copy(), "u": f64[:, 12].copy(),
        "grpot": f64[:, 14].copy(), "header_bytes": np.int64(len(recs[0])),
        "vx": f64[:, 6].copy(), "vy": f64[:, 7].copy(), "vz": f64[:, 8].copy(),
        "rho": f64[:, 5].copy(), "t": np.float64(t),
    }


DUMP_KEYS = ("x", "y", "z", "am", "hp", "u", "grpot", "header_bytes")


def make_dumps() -> None:
    out = {}
    for name, rel in DUMPS:
        d = read_dump(os.path.join(REF, rel))
        for k in DUMP_KEYS:
            out[f"{name}/{k}"] = d[k]
        print(f"{rel}: N={d['x'].size} header={int(d['header_bytes'])} B")
    np.savez_compressed(os.path.join(HERE, "dumps.npz"), **out)


def make_twin_vectors() -> None:
    sys.path.insert(0, REPO)
    import oracle

    oracle.build(ref=True)
    rng = np.random.default_rng(20261202)
    out = {}
    cases = []
    for nk in (0, 1, 2):
        cases.append((f"nk{nk}", 96, nk, 1, None, 1, 0))
    cases.append(("nk0_points", 80, 0, 1, None, 1, 3))
    cases.append(("nk2_points", 80, 2, 1, None, 1, 3))
    cases.append(("nk0_range", 90, 0, 21, 57, 1, 0))
    cases.append(("nk0_noself", 70, 0, 1, None, 0, 1))
    for name, n, nk, lo, hi, nsg, npoint in cases:
        pos = rng.normal(size=(n, 3))
        am = rng.uniform(0.5, 2.0, n) / n
        hp = rng.uniform(0.15, 0.6, n)
        u = np.ones(n)
        if npoint:
            u[-npoint:] = 0.0
        if name == "nk0_noself":
            u[:] = 1.0
            u[-1] = 0.0
        # one exactly coincident pair and one self-like duplicate exercise r = 0
        pos[5] = pos[4]
        res = oracle.ref_cpu_gravity(pos[:, 0], pos[:, 1], pos[:, 2], am, hp, u, nkernel=nk,
                                     ngrav_lower=lo, ngrav_upper=hi, nselfgravity=nsg)
        for k, v in zip(("x", "y", "z", "am", "hp", "u"), (pos[:, 0], pos[:, 1], pos[:, 2], am, hp, u)):
            out[f"{name}/{k}"] = v
        out[f"{name}/args"] = np.array([nk, lo, hi if hi is not None else n, nsg], dtype=np.int64)
        for k, v in zip(("gx", "gy", "gz", "grpot"), res):
            out[f"{name}/{k}"] = v
        print(f"twin {name}: n={n} nkernel={nk} rows {lo}..{hi or n} nselfgravity={nsg}")
    np.savez_compressed(os.path.join(HERE, "twin_vectors.npz"), **out)


def make_collision_run() -> None:
    """collision_run.npz: the start dump of the short collision run
    (collision_case.py) and what the reference program, built with its own
    CPU gravity (``make -C oracle ref-app``), writes for it on 2 MPI ranks:
    the potentials of out0000.sph (the first gravity call) and positions,
    velocities and potentials of out0001.sph (about five steps later)."""
    import subprocess
    import tempfile

    from collision_case import write_case

    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref-app"], check=True)
    exe = os.path.join(REPO, "oracle", "_ref", "starsmasher_cpu")
    start = np.fromfile(os.path.join(REF, "example_input/collision/sph.start2u"), dtype=np.uint8)
    out = {"start_bytes": start}
    with tempfile.TemporaryDirectory() as tmp:
        write_case(tmp, start, ngravprocs=2)
        subprocess.run(["/opt/conda/bin/mpiexec", "-np", "2", exe], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, timeout=600)
        d0 = read_dump(os.path.join(tmp, "out0000.sph"))
        d1 = read_dump(os.path.join(tmp, "out0001.sph"))
    out["out0000/grpot"] = d0["grpot"]
    out["out0000/t"] = d0["t"]
    for k in ("x", "y", "z", "vx", "vy", "vz", "grpot", "t"):
        out[f"out0001/{k}"] = d1[k]
    np.savez_compressed(os.path.join(HERE, "collision_run.npz"), **out)
    print(f"collision run: N={d1['x'].size}, out0001 at t={float(d1['t']):.6f}")


if __name__ == "__main__":
    sys.path.insert(0, HERE)
    which = sys.argv[1:] or ["dumps", "twin", "collision"]
    if "dumps" in which:
        make_dumps()
    if "twin" in which:
        make_twin_vectors()
    if "collision" in which:
        make_collision_run()
