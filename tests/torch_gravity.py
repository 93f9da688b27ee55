"""Test-side torch binding of the gravity library over ctypes.

``GravityFunction`` is a ``torch.autograd.Function`` whose forward calls the
C-ABI's ``sg_compute_host`` and whose backward calls ``sg_vjp_host``, the
hand-written HIP adjoint.  Tensors are CPU float64 (the library copies them
to the device): torch ships its own HIP runtime, and device pointers of one
runtime are not valid in the other, so the binding hands over host memory.

Inputs: positions (N, 3), masses (N,), smoothing lengths h = hp (N,) and a
boolean mask of point particles; range2 = +-4 h^2 as grav.f:15-18 builds it.
Outputs: accelerations (N, 3) and potentials (N,).  Gradients flow to
positions, masses and h (dL/dh = 2 dL/dH, H = 2 h).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "starsmasher_amd", "libsphgrav_hip.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_L = None


def lib():
    global _L
    if _L is None:
        L = C.CDLL(LIB_PATH)
        L.sg_create.argtypes = [C.c_int, C.POINTER(_vp)]
        L.sg_create.restype = C.c_int
        L.sg_destroy.argtypes = [_vp]
        L.sg_destroy.restype = C.c_int
        L.sg_compute_host.argtypes = ([_vp, C.c_int, C.c_int, C.c_int] + [_dp] * 5
                                      + [C.c_int, C.c_int] + [_dp] * 4)
        L.sg_compute_host.restype = C.c_int
        L.sg_vjp_host.argtypes = [_vp, C.c_int] + [_dp] * 5 + [C.c_int, C.c_int] + [_dp] * 9
        L.sg_vjp_host.restype = C.c_int
        L.sg_status_string.argtypes = [C.c_int]
        L.sg_status_string.restype = C.c_char_p
        _L = L
    return _L


class Context:
    """One sg_context on a device."""

    def __init__(self, device: int = 0):
        h = _vp()
        st = lib().sg_create(device, C.byref(h))
        if st != 0:
            raise RuntimeError(f"sg_create: {lib().sg_status_string(st).decode()}")
        self.h = h

    def close(self):
        if self.h:
            lib().sg_destroy(self.h)
            self.h = None


def _check(st, where):
    if st != 0:
        raise RuntimeError(f"{where}: {lib().sg_status_string(st).decode()} (status {st})")


def _np(t):
    return np.ascontiguousarray(t.detach().cpu().numpy(), dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_dp)


def _range2(h, point):
    r2 = 4.0 * h * h
    return np.where(point, -r2, r2)


class GravityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, m, h, point, nkernel, precision, context):
        P, M, Hh = _np(pos), _np(m), _np(h)
        pt = point.detach().cpu().numpy().astype(bool)
        r2 = _range2(Hh, pt)
        n = M.size
        x, y, z = (np.ascontiguousarray(P[:, c]) for c in range(3))
        out = [np.empty(n) for _ in range(4)]
        _check(lib().sg_compute_host(context.h, n, 0, n, _p(x), _p(y), _p(z), _p(M), _p(r2),
                                     int(nkernel), int(precision), *[_p(o) for o in out]),
               "sg_compute_host")
        ctx.save_for_backward(pos, m, h, point)
        ctx.nkernel, ctx.precision, ctx.context = int(nkernel), int(precision), context
        acc = torch.from_numpy(np.stack(out[:3], 1))
        return acc, torch.from_numpy(out[3])

    @staticmethod
    def backward(ctx, g_acc, g_pot):
        pos, m, h, point = ctx.saved_tensors
        P, M, Hh = _np(pos), _np(m), _np(h)
        r2 = _range2(Hh, point.detach().cpu().numpy().astype(bool))
        n = M.size
        x, y, z = (np.ascontiguousarray(P[:, c]) for c in range(3))
        W = _np(g_acc)
        wx, wy, wz = (np.ascontiguousarray(W[:, c]) for c in range(3))
        u = _np(g_pot)
        g = [np.empty(n) for _ in range(5)]
        _check(lib().sg_vjp_host(ctx.context.h, n, _p(x), _p(y), _p(z), _p(M), _p(r2),
                                 ctx.nkernel, ctx.precision, _p(wx), _p(wy), _p(wz), _p(u),
                                 *[_p(o) for o in g]),
               "sg_vjp_host")
        gpos = torch.from_numpy(np.stack(g[:3], 1))
        return gpos, torch.from_numpy(g[3]), torch.from_numpy(2.0 * g[4]), None, None, None, None


def gravity(pos, m, h, point, context, nkernel=0, precision=0):
    """(acc, pot) with autograd to pos, m and h."""
    return GravityFunction.apply(pos, m, h, point, nkernel, precision, context)
