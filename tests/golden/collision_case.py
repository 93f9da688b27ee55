"""The short collision run of the whole StarSmasher program, shared by the
fixture writer (make_golden.py, the program built with its own CPU gravity)
and the GPU test (tests/test_gpu_fortran.py, the same program linked against
libsphgrav_hip.so).

Two copies of the reference's small collision star
(example_input/collision/sph.start2u, 1,974 particles) collide on the
hyperbolic orbit that the example's own sph.input sets up (init 'hyp',
sep0 = 15, rp = 3.9, semimajoraxis = 20), cut to about five time steps:
tf = 0.25 with one dump interval, so the program writes out0000.sph (t = 0)
and out0001.sph (the first step past t = 0.25).
"""
from __future__ import annotations

import os

import numpy as np

SPH_INIT = " &INITT\n INAME='hyp'\n &END\n"


def sph_input(ngravprocs: int) -> str:
    return (" &input\n tf=0.25,\n dtout=0.25,\n nrelax=0,\n sep0=15,\n rp=3.9d0,\n"
            " semimajoraxis=20d0,\n"
            f" ngravprocs={int(ngravprocs)},\n computeexclusivemode=0,\n ppn=1,\n &end\n")


def write_case(path: str, start_bytes: np.ndarray, ngravprocs: int) -> None:
    """Input files of the run in directory `path`: both stars are the same
    start dump (the bytes of collision/sph.start2u)."""
    raw = np.asarray(start_bytes, dtype=np.uint8).tobytes()
    for name in ("sph.start1u", "sph.start2u"):
        with open(os.path.join(path, name), "wb") as f:
            f.write(raw)
    with open(os.path.join(path, "sph.init"), "w") as f:
        f.write(SPH_INIT)
    with open(os.path.join(path, "sph.input"), "w") as f:
        f.write(sph_input(ngravprocs))
