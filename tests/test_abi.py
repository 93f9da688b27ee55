"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/sphgrav.h declares (and the reference's three Fortran symbols),
argument errors come back as status codes before any device work, and the
legacy symbols exit on error like the reference unless SPHGRAV_NOABORT=1."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from starsmasher_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sphgrav.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)) - {"if", "sizeof"}


def test_header_declares_reference_abi():
    syms = _declared_symbols()
    for s in ("firsthalf_grav_forces_", "lasthalf_grav_forces_", "gpu_init_dev_"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = _declared_symbols()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= syms


def test_exported_symbols_in_dynamic_table():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert _declared_symbols() <= names


def test_status_strings():
    for st in range(0, 11):
        assert isinstance(_lib.status_string(st), str) and _lib.status_string(st)


def _arrays(n):
    return [np.zeros(n) for _ in range(5)]


def test_argument_errors_are_status_codes():
    L = _lib.load()
    dp = C.POINTER(C.c_double)
    x, y, z, m, r2 = _arrays(4)
    m[:] = 1.0
    r2[:] = 1.0
    p = [a.ctypes.data_as(dp) for a in (x, y, z, m, r2)]
    assert L.sg_firsthalf(4, 1, 4, *p, 3) == _lib.SG_ERR_NKERNEL
    assert L.sg_firsthalf(4, 0, 4, *p, 0) == _lib.SG_ERR_ARG        # n_lower is 1-based
    assert L.sg_firsthalf(4, 2, 4, *p, 0) == _lib.SG_ERR_ARG        # range past n
    assert L.sg_firsthalf(0, 1, 0, *p, 0) == _lib.SG_ERR_ARG
    m[1] = 0.0
    assert L.sg_firsthalf(4, 1, 4, *p, 0) == _lib.SG_ERR_MASS       # grav_force_direct.cu:602
    m[1] = 1.0
    r2[2] = 0.0
    assert L.sg_firsthalf(4, 1, 4, *p, 0) == _lib.SG_ERR_RANGE2     # :601
    r2[2] = 1.0
    x[3] = np.nan
    assert L.sg_firsthalf(4, 1, 4, *p, 0) == _lib.SG_ERR_NONFINITE
    assert L.sg_lasthalf(*[a.ctypes.data_as(dp) for a in _arrays(4)[:4]]) == _lib.SG_ERR_STATE
    assert L.sg_set_config(None, 0, 64, 2, 0) == _lib.SG_ERR_ARG
    assert L.sg_set_config(None, 0, 0, 1, 0) == _lib.SG_ERR_ARG     # automatic size, no context
    assert L.sg_last_config(None, None, None, None) == _lib.SG_ERR_ARG
    assert L.sg_compute_device(None, 4, 0, 4, *([None] * 5), 0, 0, *([None] * 4)) == _lib.SG_ERR_ARG
    e = C.c_double()
    assert L.sg_potential_energy(None, 4, None, None, C.byref(e)) == _lib.SG_ERR_ARG


_CHILD = r"""
import ctypes as C, numpy as np, sys
sys.path.insert(0, {root!r})
from starsmasher_amd import _lib
L = _lib.load()
ip = lambda v: C.byref(C.c_int(v))
a = [np.ones(3) for _ in range(5)]
dp = C.POINTER(C.c_double)
L.firsthalf_grav_forces_(ip(3), ip(1), ip(3), *[v.ctypes.data_as(dp) for v in a], ip(1), ip(7))
print("returned", L.sg_last_status())
"""


def test_legacy_symbol_exits_on_error_like_reference():
    code = _CHILD.format(root=ROOT)
    env = dict(os.environ)
    env.pop("SPHGRAV_NOABORT", None)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode != 0 and "nkernel" in r.stderr
    env["SPHGRAV_NOABORT"] = "1"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0 and "returned 2" in r.stdout
