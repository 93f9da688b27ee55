"""Host-side mirror of the reference's gravity-step interface.

The Fortran host calls three C symbols (pbe/src/SPHgrav_lib2/grav_force_direct.cu:648-677)
through ``gravforce`` (pbe/src/grav.f:1-31) and ``uvdots`` (pbe/src/balAV3.f:415-421).
This module offers the same calls to Python, with the same names, argument
meaning and index conventions (1-based ``ngrav_lower``), backed by
``libsphgrav_hip.so``:

* :func:`gpu_init_dev`          -- gpu_init_dev_(device, theta)
* :func:`firsthalf_grav_forces` -- firsthalf_grav_forces_(...): copy in, launch, return
* :func:`lasthalf_grav_forces`  -- lasthalf_grav_forces_(...): wait, write the i-range
* :func:`gravforce`             -- grav.f: range = +-4 hp^2 then firsthalf

The reference asserts/exits on bad input; here the status-code twins
(``sg_firsthalf`` / ``sg_lasthalf``) are called and a :class:`SphgravError`
is raised instead.  :class:`Gravity` wraps the device-resident sg_* API
(inputs already in HBM, e.g. torch tensors on ``cuda:0``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import SG_PREC_FP32C, SG_PREC_FP64, SphgravError, check

__all__ = [
    "gpu_init_dev", "firsthalf_grav_forces", "lasthalf_grav_forces", "gravforce",
    "Gravity", "SphgravError", "SG_PREC_FP32C", "SG_PREC_FP64", "device_count",
]

_dp = C.POINTER(C.c_double)


def _host_f64(a, name, n=None, writable=False):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        if writable:
            raise TypeError(f"{name} must be a C-contiguous float64 numpy array")
        a = np.ascontiguousarray(a, dtype=np.float64)
    if writable and not a.flags.writeable:
        raise TypeError(f"{name} must be writable")
    if n is not None and a.size < n:
        raise ValueError(f"{name} has {a.size} entries, needs {n}")
    return a


def _ptr(a):
    return a.ctypes.data_as(_dp)


def device_count() -> int:
    n = C.c_int(0)
    _lib.load().sg_device_count(C.byref(n))
    return n.value


def gpu_init_dev(device: int, theta: float = 0.3) -> None:
    """Select the HIP device (grav_force_direct.cu:672-675; theta is ignored)."""
    check(_lib.load().sg_init_device(int(device)), "gpu_init_dev")


_pending_keepalive = None


def firsthalf_grav_forces(ntot, ngrav_lower, mygravlength, x, y, z, am, range_, q=1, nkernel=0):
    """Copy positions, masses and range2 = +-4 hp^2 of particles 1..ntot and
    start the gravity of particles ngrav_lower..ngrav_lower+mygravlength-1
    (1-based).  Returns at once; :func:`lasthalf_grav_forces` collects."""
    global _pending_keepalive
    ntot = int(ntot)
    arrs = [_host_f64(a, nm, ntot) for a, nm in
            ((x, "x"), (y, "y"), (z, "z"), (am, "am"), (range_, "range"))]
    st = _lib.load().sg_firsthalf(ntot, int(ngrav_lower), int(mygravlength),
                                  *[_ptr(a) for a in arrs], int(nkernel))
    check(st, "firsthalf_grav_forces")
    _pending_keepalive = arrs   # inputs are copied synchronously; kept only for symmetry


def lasthalf_grav_forces(ntot, gx, gy, gz, grpot, myrank=0):
    """Wait for the pending evaluation and write gx, gy, gz, grpot over the
    i-range of the preceding firsthalf (other entries untouched)."""
    outs = [_host_f64(a, nm, None, writable=True) for a, nm in
            ((gx, "gx"), (gy, "gy"), (gz, "gz"), (grpot, "grpot"))]
    check(_lib.load().sg_lasthalf(*[_ptr(a) for a in outs]), "lasthalf_grav_forces")


def gravforce(x, y, z, am, hp, u, ngrav_lower=1, ngrav_upper=None, q=1, nkernel=0):
    """grav.f:10-31: build range = 4 hp^2 (negated where u == 0, the point
    particles) and call firsthalf for this rank's chunk."""
    ntot = len(x)
    ngrav_upper = ntot if ngrav_upper is None else int(ngrav_upper)
    hp = np.asarray(hp, dtype=np.float64)
    rng = 4.0 * hp * hp
    rng = np.where(np.asarray(u) == 0.0, -rng, rng)
    firsthalf_grav_forces(ntot, ngrav_lower, ngrav_upper - ngrav_lower + 1, x, y, z, am, rng, q,
                          nkernel)


def bonet_psi(x, y, z, am, hp, u, nkernel=0, ngrav_lower=1, ngrav_upper=None):
    """The grad-h softening term psi_i of particles ngrav_lower..ngrav_upper
    (1-based; advance.f90:716-737's bonet1_psi, with nselfgravity = 1), from
    the GPU through sg_grav_psi.  Returns an array of length ntot with the
    i-range filled and zeros elsewhere."""
    ntot = len(x)
    ngrav_upper = ntot if ngrav_upper is None else int(ngrav_upper)
    hp = np.asarray(hp, dtype=np.float64)
    rng = 4.0 * hp * hp
    rng = np.where(np.asarray(u) == 0.0, -rng, rng)
    ins = [_host_f64(a, nm, ntot) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (am, "am"),
                                                 (rng, "range"))]
    out = np.zeros(ntot)
    check(_lib.load().sg_grav_psi(ntot, int(ngrav_lower), ngrav_upper - ngrav_lower + 1,
                                  *[_ptr(a) for a in ins], int(nkernel), _ptr(out)),
          "sg_grav_psi")
    return out


def hip_runtimes() -> list:
    """The distinct HIP runtime libraries (libamdhip64) mapped into this
    process, by real path."""
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 6 and "libamdhip64.so" in parts[-1]:
                paths.add(os.path.realpath(parts[-1]))
    return sorted(paths)


def _check_shared_runtime(name):
    """Device pointers and streams of torch are only meaningful to the library
    when both use one HIP runtime.  torch wheels bundle their own
    libamdhip64, and a pointer or stream of one runtime is not valid in the
    other, so a second runtime in the process makes torch tensors unusable
    here (allocate through starsmasher_amd.hipmem instead)."""
    _lib.load()
    rts = hip_runtimes()
    if len(rts) != 1:
        raise TypeError(f"{name}: torch tensors cannot be passed: this process maps "
                        f"{len(rts)} HIP runtimes ({', '.join(rts)}), and device pointers "
                        "and streams of one are invalid in another; allocate with "
                        "starsmasher_amd.hipmem.DeviceArray")


def _is_torch(t) -> bool:
    return type(t).__module__.startswith("torch")


def _dev_ptr(t, name, n=None):
    """Device pointer of a :class:`starsmasher_amd.hipmem.DeviceArray`, a raw
    integer address, or a torch tensor (float64, contiguous, on a GPU; only
    accepted when torch and the library share one HIP runtime)."""
    if isinstance(t, int):
        return t
    if hasattr(t, "ptr") and hasattr(t, "size"):
        if n is not None and t.size < n:
            raise ValueError(f"{name} has {t.size} entries, needs {n}")
        return int(t.ptr)
    try:
        import torch
    except ImportError as e:   # pragma: no cover
        raise TypeError(f"{name}: pass an integer device address or a torch tensor") from e
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU or a device address")
    _check_shared_runtime(name)
    if t.device.type != "cuda" or t.dtype != torch.float64 or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float64 tensor on a GPU")
    if n is not None and t.numel() < n:
        raise ValueError(f"{name} has {t.numel()} entries, needs {n}")
    return t.data_ptr()


class Gravity:
    """A library context bound to one HIP device (device buffers are reused
    across calls).  ``compute`` takes host numpy arrays; ``compute_device``
    takes device tensors and only enqueues work (call ``synchronize``)."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        check(self._L.sg_create(int(device), C.byref(h)), "sg_create")
        self._h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.sg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ config
    _stream = 0          # 0: the context's own stream

    def set_stream(self, stream):
        """Enqueue on `stream` (a hipStream_t address; 0/None = own stream)."""
        stream = int(stream or 0)
        if stream != self._stream:
            check(self._L.sg_set_stream(self._h, C.c_void_p(stream or None)), "sg_set_stream")
            self._stream = stream

    def set_config(self, precision, wg, ipl, jsplit=0):
        check(self._L.sg_set_config(self._h, int(precision), int(wg), int(ipl), int(jsplit)),
              "sg_set_config")

    def get_config(self, precision):
        """(wg, ipl, jsplit) as configured; 0 = chosen per evaluation."""
        wg, ipl, js = C.c_int(), C.c_int(), C.c_int()
        check(self._L.sg_get_config(self._h, int(precision), C.byref(wg), C.byref(ipl), C.byref(js)),
              "sg_get_config")
        return wg.value, ipl.value, js.value

    def last_config(self):
        """(wg, ipl, jsplit) the last single-set evaluation ran with; raises
        SphgravError (SG_ERR_STATE) before one or after a call of another
        kind (batch, psi-only, vector-Jacobian product)."""
        wg, ipl, js = C.c_int(), C.c_int(), C.c_int()
        check(self._L.sg_last_config(self._h, C.byref(wg), C.byref(ipl), C.byref(js)),
              "sg_last_config")
        return wg.value, ipl.value, js.value

    # ----------------------------------------------------------- compute
    def compute(self, x, y, z, m, range2, nkernel=0, precision=SG_PREC_FP32C, ibeg=0, ni=None):
        """Blocking host-array evaluation.  Returns (ax, ay, az, pot) of the
        0-based i-range [ibeg, ibeg+ni)."""
        n = len(x)
        ni = n - ibeg if ni is None else int(ni)
        ins = [_host_f64(a, nm, n) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                  (range2, "range2"))]
        outs = [np.empty(max(ni, 0)) for _ in range(4)]
        st = self._L.sg_compute_host(self._h, n, int(ibeg), ni, *[_ptr(a) for a in ins],
                                     int(nkernel), int(precision), *[_ptr(o) for o in outs])
        check(st, "sg_compute_host")
        return tuple(outs)

    def compute_psi(self, x, y, z, m, range2, nkernel=0, precision=SG_PREC_FP32C, ibeg=0,
                    ni=None):
        """Blocking host-array evaluation that also forms the grad-h softening
        term psi (advance.f90:716-737).  Returns (ax, ay, az, pot, psi) of the
        0-based i-range [ibeg, ibeg+ni)."""
        n = len(x)
        ni = n - ibeg if ni is None else int(ni)
        ins = [_host_f64(a, nm, n) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                  (range2, "range2"))]
        outs = [np.empty(max(ni, 0)) for _ in range(5)]
        st = self._L.sg_compute_psi_host(self._h, n, int(ibeg), ni, *[_ptr(a) for a in ins],
                                         int(nkernel), int(precision), *[_ptr(o) for o in outs])
        check(st, "sg_compute_psi_host")
        return tuple(outs)

    def psi(self, x, y, z, m, range2, nkernel=0, ibeg=0, ni=None):
        """Blocking grad-h term alone (FP64) of the 0-based i-range."""
        n = len(x)
        ni = n - ibeg if ni is None else int(ni)
        ins = [_host_f64(a, nm, n) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                  (range2, "range2"))]
        out = np.empty(max(ni, 0))
        st = self._L.sg_psi_host(self._h, n, int(ibeg), ni, *[_ptr(a) for a in ins],
                                 int(nkernel), _ptr(out))
        check(st, "sg_psi_host")
        return out

    def vjp(self, x, y, z, m, range2, wx, wy, wz, u, nkernel=0, precision=SG_PREC_FP32C):
        """Blocking host-array adjoint: the gradient of
        L = sum_i w_i . a_i + u_i phi_i over all particles with respect to
        (x, y, z, m, H), H = sqrt|range2|.  Returns (gx, gy, gz, gm, gH)."""
        n = len(x)
        ins = [_host_f64(a, nm, n) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                  (range2, "range2"), (wx, "wx"), (wy, "wy"),
                                                  (wz, "wz"), (u, "u"))]
        outs = [np.empty(n) for _ in range(5)]
        st = self._L.sg_vjp_host(self._h, n, *[_ptr(a) for a in ins[:5]], int(nkernel),
                                 int(precision), *[_ptr(a) for a in ins[5:]],
                                 *[_ptr(o) for o in outs])
        check(st, "sg_vjp_host")
        return tuple(outs)

    def vjp_device(self, n, x, y, z, m, range2, wx, wy, wz, u, gx, gy, gz, gm, gH, nkernel=0,
                   precision=SG_PREC_FP32C):
        """Enqueue a device-resident adjoint (addresses or device arrays of
        length n); call ``synchronize``."""
        n = int(n)
        ptrs = [_dev_ptr(a, nm, n) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                  (range2, "range2"), (wx, "wx"), (wy, "wy"),
                                                  (wz, "wz"), (u, "u"), (gx, "gx"), (gy, "gy"),
                                                  (gz, "gz"), (gm, "gm"), (gH, "gH"))]
        vp = [C.c_void_p(p) for p in ptrs]
        self._stream_for(x)
        st = self._L.sg_vjp_device(self._h, n, *vp[:5], int(nkernel), int(precision), *vp[5:])
        check(st, "sg_vjp_device")

    def _stream_for(self, t):
        """Enqueue on the caller's current torch stream when `t` is a torch
        tensor (ordered after the work that produced it), else on the
        context's own stream, so a torch call's stream does not linger into
        later calls on device arrays."""
        if _is_torch(t):
            _check_shared_runtime("stream")
            import torch
            self.set_stream(torch.cuda.current_stream(t.device).cuda_stream)
        else:
            self.set_stream(None)

    def compute_device(self, n, x, y, z, m, range2, ax, ay, az, pot, nkernel=0,
                       precision=SG_PREC_FP32C, ibeg=0, ni=None):
        """Enqueue a device-resident evaluation (device arrays, addresses, or
        torch tensors when torch shares the library's HIP runtime)."""
        n = int(n)
        ni = n - ibeg if ni is None else int(ni)
        ins = [_dev_ptr(a, nm, n) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                 (range2, "range2"))]
        outs = [_dev_ptr(a, nm, ni) for a, nm in ((ax, "ax"), (ay, "ay"), (az, "az"), (pot, "pot"))]
        self._stream_for(x)
        st = self._L.sg_compute_device(self._h, n, int(ibeg), ni, *[C.c_void_p(p) for p in ins],
                                       int(nkernel), int(precision),
                                       *[C.c_void_p(p) for p in outs])
        check(st, "sg_compute_device")

    # ------------------------------------------------------------ batch
    def set_batch_config(self, precision, wg, ipl):
        check(self._L.sg_set_batch_config(self._h, int(precision), int(wg), int(ipl)),
              "sg_set_batch_config")

    def get_batch_config(self, precision):
        wg, ipl = C.c_int(), C.c_int()
        check(self._L.sg_get_batch_config(self._h, int(precision), C.byref(wg), C.byref(ipl)),
              "sg_get_batch_config")
        return wg.value, ipl.value

    def compute_batch_device(self, offsets, x, y, z, m, range2, ax, ay, az, pot, nkernel=0,
                             precision=SG_PREC_FP32C, use_graph=True):
        """Enqueue a batch of independent sets: set s is particles
        offsets[s]..offsets[s+1]-1 of the device arrays."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        ntot = int(off[-1])
        ins = [_dev_ptr(a, nm, ntot) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (m, "m"),
                                                    (range2, "range2"))]
        outs = [_dev_ptr(a, nm, ntot) for a, nm in ((ax, "ax"), (ay, "ay"), (az, "az"),
                                                     (pot, "pot"))]
        self._stream_for(x)
        st = self._L.sg_compute_batch_device(
            self._h, off.size - 1, off.ctypes.data_as(C.POINTER(C.c_longlong)),
            *[C.c_void_p(p) for p in ins], int(nkernel), int(precision), int(bool(use_graph)),
            *[C.c_void_p(p) for p in outs])
        check(st, "sg_compute_batch_device")

    def compute_batch(self, sets, nkernel=0, precision=SG_PREC_FP32C, use_graph=False):
        """Blocking host-array batch: `sets` is a list of dicts with x, y, z,
        m, range2.  Returns a list of (ax, ay, az, pot) per set."""
        from .hipmem import DeviceArray
        sizes = [len(s["x"]) for s in sets]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        dev = [DeviceArray.from_host(np.concatenate([s[k] for s in sets]))
               for k in ("x", "y", "z", "m", "range2")]
        out = [DeviceArray(int(off[-1])) for _ in range(4)]
        self.compute_batch_device(off, *dev, *out, nkernel=nkernel, precision=precision,
                                  use_graph=use_graph)
        self.synchronize()
        res = [o.download() for o in out]
        return [tuple(r[off[s]:off[s + 1]] for r in res) for s in range(len(sets))]

    def synchronize(self):
        check(self._L.sg_synchronize(self._h), "sg_synchronize")

    def potential_energy(self, n, m, pot) -> float:
        """epot = 1/2 sum m pot of device arrays (torch tensors or addresses),
        summed in FP64 in a fixed order on the device: the energy diagnostic
        of output.f:232-252 without copying the potentials to the host."""
        n = int(n)
        mp, pp = _dev_ptr(m, "m", n), _dev_ptr(pot, "pot", n)
        self._stream_for(m)
        e = C.c_double()
        check(self._L.sg_potential_energy(self._h, n, mp, pp, C.byref(e)), "sg_potential_energy")
        return e.value

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(self._L.sg_last_kernel_ms(self._h, C.byref(ms)), "sg_last_kernel_ms")
        return ms.value

    def last_timings(self):
        """(prep_ms, far_ms, near_ms) of the last evaluation (HIP events)."""
        t = [C.c_float() for _ in range(3)]
        check(self._L.sg_last_timings(self._h, *[C.byref(v) for v in t]), "sg_last_timings")
        return tuple(v.value for v in t)

    def last_tile_stats(self):
        v, nv = C.c_longlong(), C.c_longlong()
        check(self._L.sg_last_tile_stats(self._h, C.byref(v), C.byref(nv)), "sg_last_tile_stats")
        return v.value, nv.value

    def last_precision_stats(self):
        """(chunk visits, near chunk visits, wide chunk visits, FP64 pairs) of
        the last evaluation (sg_last_precision_stats)."""
        v = [C.c_longlong() for _ in range(4)]
        check(self._L.sg_last_precision_stats(self._h, *[C.byref(x) for x in v]),
              "sg_last_precision_stats")
        return tuple(x.value for x in v)
