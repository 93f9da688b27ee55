"""Oracle for the direct-sum softened gravity path.

TEST INFRASTRUCTURE ONLY: this package is the checker.  It may be imported
by ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` and by nothing else; the product library in
``starsmasher_amd`` never routes through it.

Contents
--------
* ctypes bindings over ``oracle/liboracle.so`` (built from ``sg_oracle.c``),
  an FP64 C restatement of

  - ``grav_force_direct.cu:55-502`` (pbe/src/SPHgrav_lib2): full i x j sum
    with per-pair softening on H_i/H_j, nkernel 0/1/2 -> :func:`direct`;
  - ``cpu_grav.f:1-381``: the symmetric j >= i CPU twin -> :func:`cpu_twin`;
  - ``main.f:297-349``: the interaction-balanced split -> :func:`partition`;
  - an FP64 adjoint of L = sum_i w_i.a_i + u_i phi_i -> :func:`vjp`.

* ctypes binding over ``oracle/_ref/libref_cpugrav.so``: the reference's own
  ``cpu_grav.f`` compiled from its source by ``oracle/Makefile``
  -> :func:`ref_cpu_gravity` (None when not built).

* :func:`direct_numpy`: a pure-numpy restatement for tiny cases, written
  independently of the C file, used to cross-check it.

Pinning: see the header of ``sg_oracle.c`` and ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libref_cpugrav.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

SELF_LIMIT = 0   # i == j self term; distinct coincident pairs -> finite r->0 limit
SELF_GPU = 1     # the reference kernel's rule (NaN for distinct coincident pairs)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and the reference twin when its source exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/parallel_bleeding_edge/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build(ref=False)
        L = C.CDLL(_LIB)
        L.sgo_direct_rows.argtypes = [C.c_int, _i64p, C.c_int64, _dp, _dp, _dp, _dp, _dp,
                                      C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int]
        L.sgo_direct_rows.restype = C.c_int
        L.sgo_cpu_twin.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.sgo_cpu_twin.restype = C.c_int
        L.sgo_cpu_twin_parallel.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                            _dp, _dp, _dp, _dp, C.c_int]
        L.sgo_cpu_twin_parallel.restype = C.c_int
        L.sgo_cpu_twin_rows.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                        _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.sgo_cpu_twin_rows.restype = C.c_int64
        L.sgo_cpu_twin_rows_parallel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                                 _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                                 C.c_int]
        L.sgo_cpu_twin_rows_parallel.restype = C.c_int64
        L.sgo_partition.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
        L.sgo_partition.restype = C.c_int
        L.sgo_vjp.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp,
                              _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.sgo_vjp.restype = C.c_int
        L.sgo_vjp_rows.argtypes = [C.c_int, _i64p, C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.sgo_vjp_rows.restype = C.c_int
        L.sgo_loss_particle.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                        _dp, _dp, _dp, _dp]
        L.sgo_loss_particle.restype = C.c_double
        L.sgo_fd_grad.argtypes = [C.c_int, _i64p, C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                  _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_int]
        L.sgo_fd_grad.restype = C.c_int
        L.sgo_psi_rows.argtypes = [C.c_int, _i64p, C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                   C.c_int, _dp, C.c_int]
        L.sgo_psi_rows.restype = C.c_int
        L.sgo_kernel_AP.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.sgo_kernel_dAP.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.sgo_epot.argtypes = [C.c_int64, _dp, _dp]
        L.sgo_epot.restype = C.c_double
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def kernel_AP(nkernel: int, q: float):
    A, P = C.c_double(), C.c_double()
    lib().sgo_kernel_AP(nkernel, float(q), C.byref(A), C.byref(P))
    return A.value, P.value


def kernel_dAP(nkernel: int, q: float):
    A, P = C.c_double(), C.c_double()
    lib().sgo_kernel_dAP(nkernel, float(q), C.byref(A), C.byref(P))
    return A.value, P.value


def direct(x, y, z, m, range2, nkernel=0, rows=None, self_rule=SELF_LIMIT, nthreads=0):
    """FP64 full i x j sum.  rows: 0-based i indices (default all).
    Returns (ax, ay, az, pot) for those rows."""
    x, y, z, m, range2 = map(_f64, (x, y, z, m, range2))
    n = x.size
    rows = np.arange(n, dtype=np.int64) if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    out = [np.empty(rows.size) for _ in range(4)]
    rc = lib().sgo_direct_rows(n, rows, rows.size, x, y, z, m, range2, int(nkernel), int(self_rule),
                               *out, int(nthreads or os.cpu_count() or 1))
    if rc != 0:
        raise ValueError(f"sgo_direct_rows rejected its arguments (rc={rc})")
    return tuple(out)


def cpu_twin(x, y, z, am, hp, u, nkernel=0, ngrav_lower=1, ngrav_upper=None, nselfgravity=1):
    """Serial restatement of get_gravity_using_cpus (cpu_grav.f:1-381)."""
    x, y, z, am, hp, u = map(_f64, (x, y, z, am, hp, u))
    n = x.size
    ngrav_upper = n if ngrav_upper is None else ngrav_upper
    out = [np.empty(n) for _ in range(4)]
    rc = lib().sgo_cpu_twin(n, ngrav_lower, ngrav_upper, nkernel, nselfgravity, x, y, z, am, hp, u, *out)
    if rc != 0:
        raise ValueError(f"sgo_cpu_twin rc={rc}")
    return tuple(out)


def cpu_twin_parallel(x, y, z, am, hp, u, nkernel=0, nthreads=None):
    x, y, z, am, hp, u = map(_f64, (x, y, z, am, hp, u))
    n = x.size
    out = [np.empty(n) for _ in range(4)]
    rc = lib().sgo_cpu_twin_parallel(n, nkernel, x, y, z, am, hp, u, *out,
                                     int(nthreads or os.cpu_count() or 1))
    if rc != 0:
        raise ValueError(f"sgo_cpu_twin_parallel rc={rc}")
    return tuple(out)


def cpu_twin_rows(x, y, z, am, hp, u, row_lo, row_hi, nkernel=0):
    """Rows [row_lo, row_hi] (0-based) of the symmetric loop; returns
    (evaluations, gx, gy, gz, grpot) -- partial sums, for timing samples."""
    x, y, z, am, hp, u = map(_f64, (x, y, z, am, hp, u))
    n = x.size
    out = [np.zeros(n) for _ in range(4)]
    ev = lib().sgo_cpu_twin_rows(n, row_lo, row_hi, nkernel, x, y, z, am, hp, u, *out)
    if ev < 0:
        raise ValueError("sgo_cpu_twin_rows rejected its arguments")
    return (int(ev), *out)


def cpu_twin_rows_parallel(x, y, z, am, hp, u, row_lo, row_hi, nkernel=0, nthreads=1):
    """Rows [row_lo, row_hi] of the symmetric loop on nthreads threads;
    returns (evaluations, gx, gy, gz, grpot) partial sums."""
    x, y, z, am, hp, u = map(_f64, (x, y, z, am, hp, u))
    n = x.size
    out = [np.zeros(n) for _ in range(4)]
    ev = lib().sgo_cpu_twin_rows_parallel(n, row_lo, row_hi, nkernel, x, y, z, am, hp, u, *out,
                                          int(nthreads))
    if ev < 0:
        raise ValueError("sgo_cpu_twin_rows_parallel rejected its arguments")
    return (int(ev), *out)


def partition(n: int, nranks: int):
    lo = np.zeros(nranks, dtype=np.int32)
    hi = np.zeros(nranks, dtype=np.int32)
    rc = lib().sgo_partition(n, nranks, lo, hi)
    if rc != 0:
        raise ValueError("partition failed (main.f:341-344)")
    return lo, hi


def vjp(x, y, z, m, range2, w, u, nkernel=0, nthreads=0, rows=None):
    """FP64 gradient of L = sum w_i.a_i + u_i phi_i w.r.t. (x, y, z, m, H),
    for all particles or the 0-based ``rows`` (outputs indexed like rows)."""
    x, y, z, m, range2 = map(_f64, (x, y, z, m, range2))
    w = np.asarray(w, dtype=np.float64)
    wx, wy, wz = (np.ascontiguousarray(w[:, k]) for k in range(3))
    u = _f64(u)
    n = x.size
    nt = int(nthreads or os.cpu_count() or 1)
    if rows is not None:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = [np.empty(rows.size) for _ in range(5)]
        rc = lib().sgo_vjp_rows(n, rows, rows.size, x, y, z, m, range2, nkernel, wx, wy, wz, u,
                                *out, nt)
        if rc != 0:
            raise ValueError("sgo_vjp_rows rejected its arguments")
        return tuple(out)
    out = [np.empty(n) for _ in range(5)]
    rc = lib().sgo_vjp(n, x, y, z, m, range2, nkernel, wx, wy, wz, u, *out, nt)
    if rc != 0:
        raise ValueError("sgo_vjp rejected its arguments")
    return tuple(out)


def loss_particle(k, x, y, z, m, range2, w, u, nkernel=0):
    x, y, z, m, range2 = map(_f64, (x, y, z, m, range2))
    w = np.asarray(w, dtype=np.float64)
    wx, wy, wz = (np.ascontiguousarray(w[:, c]) for c in range(3))
    return lib().sgo_loss_particle(x.size, int(k), x, y, z, m, range2, nkernel, wx, wy, wz, _f64(u))


def fd_grad(ks, x, y, z, m, range2, w, u, nkernel=0, step=1e-6, nthreads=0):
    """Central finite differences of L w.r.t. (x, y, z, m, H) of the
    particles ks; returns an array of shape (len(ks), 5)."""
    x, y, z, m, range2 = map(_f64, (x, y, z, m, range2))
    w = np.asarray(w, dtype=np.float64)
    wx, wy, wz = (np.ascontiguousarray(w[:, c]) for c in range(3))
    ks = np.ascontiguousarray(ks, dtype=np.int64)
    out = np.empty((ks.size, 5))
    rc = lib().sgo_fd_grad(x.size, ks, ks.size, x, y, z, m, range2, nkernel, wx, wy, wz, _f64(u),
                           float(step), out.reshape(-1), int(nthreads or os.cpu_count() or 1))
    if rc != 0:
        raise ValueError("sgo_fd_grad rejected its arguments")
    return out


def psi(x, y, z, m, range2, nkernel=0, rows=None, tabulated=False, nthreads=0):
    """Grad-h softening term psi_i = hp_i^-2 sum_j m_j dphidh(r_ij/hp_i) over
    r_ij < 2 hp_i (advance.f90:716-737), for gas particles; 0 for point
    particles.  tabulated=True reproduces the reference's 10^7-entry table."""
    x, y, z, m, range2 = map(_f64, (x, y, z, m, range2))
    n = x.size
    rows = np.arange(n, dtype=np.int64) if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    out = np.empty(rows.size)
    rc = lib().sgo_psi_rows(n, rows, rows.size, x, y, z, m, range2, int(nkernel), int(bool(tabulated)),
                            out, int(nthreads or os.cpu_count() or 1))
    if rc != 0:
        raise ValueError("sgo_psi_rows rejected its arguments")
    return out


def epot(m, pot):
    m, pot = _f64(m), _f64(pot)
    return lib().sgo_epot(m.size, m, pot)


_ref = None


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref_cpu_gravity(x, y, z, am, hp, u, nkernel=0, ngrav_lower=1, ngrav_upper=None, nselfgravity=1):
    """The reference's own get_gravity_using_cpus (cpu_grav.f), compiled from
    its source into oracle/_ref.  Raises FileNotFoundError when not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(_REF)
        R = C.CDLL(_REF)
        R.ref_cpu_gravity.argtypes = [C.c_int] * 5 + [_dp] * 10 + [C.POINTER(C.c_int)]
        R.ref_cpu_gravity.restype = None
        _ref = R
    x, y, z, am, hp, u = map(_f64, (x, y, z, am, hp, u))
    n = x.size
    ngrav_upper = n if ngrav_upper is None else ngrav_upper
    out = [np.empty(n) for _ in range(4)]
    st = C.c_int(0)
    _ref.ref_cpu_gravity(n, ngrav_lower, ngrav_upper, nkernel, nselfgravity, x, y, z, am, hp, u, *out,
                         C.byref(st))
    if st.value != 0:
        raise ValueError("ref_cpu_gravity: n outside 1..nmax (starsmasher.h:4)")
    return tuple(out)


# --------------------------------------------------------------------------
# Independent pure-numpy restatement for tiny inputs (the cubic-spline
# closed forms of grav_force_direct.cu:55-182; cubic only).
# --------------------------------------------------------------------------
def _cubic_A(q):
    q = np.asarray(q, dtype=np.float64)
    inner = 32.0 / 3.0 + q * q * (32.0 * q - 38.4)
    with np.errstate(divide="ignore", invalid="ignore"):
        outer = 64.0 / 3.0 - 48.0 * q + 38.4 * q * q - 32.0 / 3.0 * q ** 3 - 1.0 / 15.0 / q ** 3
    return np.where(q < 0.5, inner, outer)


def _cubic_P(q):
    q = np.asarray(q, dtype=np.float64)
    inner = -2.8 + q * q * (16.0 / 3.0 + q * q * (6.4 * q - 9.6))
    with np.errstate(divide="ignore", invalid="ignore"):
        outer = -3.2 + 1.0 / 15.0 / q + q * q * (32.0 / 3.0 + q * (-16.0 + q * (9.6 - 32.0 / 15.0 * q)))
    return np.where(q < 0.5, inner, outer)


def direct_numpy(pos, m, H, flag):
    """O(N^2) numpy cubic-spline direct sum (exact fractions, not the
    reference's 12-digit literals).  pos: (N,3); H: support radius."""
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    acc = np.zeros((n, 3))
    pot = np.zeros(n)
    for i in range(n):
        dr = pos - pos[i]
        r = np.sqrt((dr * dr).sum(1))
        same = np.arange(n) == i
        with np.errstate(divide="ignore", invalid="ignore"):
            newt_f = 1.0 / r ** 3
            newt_p = -1.0 / r
            F = np.zeros(n)
            P = np.zeros(n)
            for Hk in (np.full(n, H[i]), H):
                q = r / Hk
                soft = r < Hk
                F += np.where(soft, _cubic_A(q) / Hk ** 3, newt_f)
                P += np.where(soft, _cubic_P(q) / Hk, newt_p)
        F *= 0.5
        P *= 0.5
        zero = r == 0
        F[zero] = 0.0
        P[zero] = np.where(same[zero], flag[zero], 1.0) * (-1.4) * (1.0 / H[i] + 1.0 / H[zero])
        acc[i] = (m[:, None] * F[:, None] * dr).sum(0)
        pot[i] = (m * P).sum()
    return acc, pot
