"""sg_compute_device_static: the evaluation of sg_compute_device with every
scratch buffer in one file-scope device array (no device allocation).

Its results are compared with sg_compute_device on a context at the default
configuration (automatic workgroup size, one i per lane, automatic j-split),
which runs the same
kernels with the same j-partitions and reduction order: the outputs must be
bit-identical, in both precisions, on two calls in a row (the second with a
smaller set and an i-subrange, so stale contents of the shared array from the
first call would show).  The oracle checks the first call's FP32C results at
the usual tolerance (acceleration max 1e-5, potential 1e-6).  The entry point
is also captured in a HIP graph and replayed (the Morton sort is rocPRIM's
merge sort there, which issues kernels only; it gives the radix sort's order,
so the results stay bit-identical).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity, _lib, ics
from starsmasher_amd.hipmem import DeviceArray, runtime

from conftest import rel_errors


def _set(n, seed):
    d = ics.plummer(n, seed=seed)
    d["range2"][-3:] *= -1.0            # point particles
    for k in ("x", "y", "z"):
        d[k][11] = d[k][10]             # a distinct coincident pair
    return d


def _static(L, d, ibeg, ni, nk, prec, err=None):
    n = d["x"].size
    ins = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    outs = [DeviceArray(ni) for _ in range(4)]
    rc = L.sg_compute_device_static(n, ibeg, ni, *[C.c_void_p(a.ptr) for a in ins], nk, prec,
                                    *[C.c_void_p(o.ptr) for o in outs],
                                    C.c_void_p(err.ptr) if err is not None else None, None)
    assert rc == 0, _lib.status_string(rc)
    assert runtime().hipDeviceSynchronize() == 0
    res = [o.download() for o in outs]
    for a in ins + outs:
        a.free()
    return res


def test_static_rejects_sizes_beyond_its_maximum():
    """n above SG_STATIC_MAX_N is refused before any device work."""
    L = _lib.load()
    big = _lib.SG_STATIC_MAX_N + 1
    dummy = C.c_void_p(16)   # never dereferenced: the size check comes first
    rc = L.sg_compute_device_static(big, 0, big, *([dummy] * 5), 0, SG_PREC_FP32C,
                                    *([dummy] * 4), None, None)
    assert rc == _lib.SG_ERR_SIZE
    assert "SG_STATIC_MAX_N" in _lib.status_string(rc)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_static_matches_context_entry_point(prec):
    L = _lib.load()
    cases = [(_set(20000, 61), 0, 20000, 0), (_set(5000, 62), 1200, 3000, 2)]
    with Gravity(0) as g:
        assert g.get_config(prec) == (0, 1, 0)   # the static entry point's own shape rule
        for k, (d, ibeg, ni, nk) in enumerate(cases):
            print(f"[static call {k}] n={d['x'].size} ibeg={ibeg} ni={ni} nk={nk} prec={prec}",
                  flush=True)
            got = _static(L, d, ibeg, ni, nk, prec)
            want = g.compute(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=nk,
                             precision=prec, ibeg=ibeg, ni=ni)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), k
            if k == 0:
                ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=nk)
                ea, ep = rel_errors(np.stack(got[:3], -1), got[3], np.stack(ref[:3], -1), ref[3])
                lim = (1e-5, 1e-6) if prec == SG_PREC_FP32C else (1e-11, 1e-12)
                print(f"[static call {k}] vs oracle: acc {ea.max():.3e} pot {ep.max():.3e}")
                assert ea.max() <= lim[0] and ep.max() <= lim[1]


@pytest.mark.gpu
def test_static_at_its_maximum_size():
    """n = SG_STATIC_MAX_N (the largest partial-sum slab set and sort storage
    the scratch holds) against the context entry point, FP32C, Wendland C4."""
    L = _lib.load()
    n = _lib.SG_STATIC_MAX_N
    d = _set(n, 64)
    got = _static(L, d, 0, n, 2, SG_PREC_FP32C)
    with Gravity(0) as g:
        want = g.compute(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=2,
                         precision=SG_PREC_FP32C)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_static_reports_validation_flags():
    """The device int given as err receives the validation bits: 0 for a
    valid set, SG_VALID_MASS after a set with a non-positive mass."""
    L = _lib.load()
    err = DeviceArray(1, np.int32)
    err.upload(np.array([-1], dtype=np.int32))
    d = _set(3000, 63)
    _static(L, d, 0, 3000, 0, SG_PREC_FP32C, err)
    assert err.download()[0] == 0
    d["m"][7] = -1.0
    _static(L, d, 0, 3000, 0, SG_PREC_FP32C, err)
    assert err.download()[0] == _lib.SG_VALID_MASS
    err.free()


def _hip():
    h = runtime()
    vp = C.c_void_p
    for name, args in (("hipStreamCreate", [C.POINTER(vp)]), ("hipStreamDestroy", [vp]),
                       ("hipStreamSynchronize", [vp]),
                       ("hipStreamBeginCapture", [vp, C.c_int]),
                       ("hipStreamEndCapture", [vp, C.POINTER(vp)]),
                       ("hipGraphInstantiate", [C.POINTER(vp), vp, vp, vp, C.c_size_t]),
                       ("hipGraphLaunch", [vp, vp]), ("hipGraphExecDestroy", [vp]),
                       ("hipGraphDestroy", [vp])):
        f = getattr(h, name)
        f.argtypes = args
        f.restype = C.c_int
    return h


def _gset(n, seed):
    d = ics.plummer(n, seed=seed)
    d["range2"][-2:] *= -1.0
    return d


def _call(L, ins, outs, n, nk, prec, stream):
    rc = L.sg_compute_device_static(n, 0, n, *[C.c_void_p(a.ptr) for a in ins], nk, prec,
                                    *[C.c_void_p(o.ptr) for o in outs], None, stream)
    assert rc == 0, _lib.status_string(rc)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_static_entry_point_in_a_graph(prec):
    """Captured in a HIP graph (global capture mode, which refuses device
    allocation while the capture is open) and replayed several times, with
    eager calls on another set between the replays (they overwrite the
    shared scratch): every replay reproduces the eager result bit for bit."""
    L = _lib.load()
    h = _hip()
    s = C.c_void_p()
    assert h.hipStreamCreate(C.byref(s)) == 0
    n, n2 = 12000, 7000
    d, d2 = _gset(n, 71), _gset(n2, 72)
    ins = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    ins2 = [DeviceArray.from_host(d2[k]) for k in ("x", "y", "z", "m", "range2")]
    eager = [DeviceArray(n) for _ in range(4)]
    outs = [DeviceArray(n) for _ in range(4)]
    outs2 = [DeviceArray(n2) for _ in range(4)]
    # eager reference (also looks up the scratch address before any capture)
    _call(L, ins, eager, n, 0, prec, s)
    assert h.hipStreamSynchronize(s) == 0
    want = [o.download() for o in eager]
    graph, exe = C.c_void_p(), C.c_void_p()
    print(f"[graph prec={prec}] capture", flush=True)
    assert h.hipStreamBeginCapture(s, 0) == 0          # hipStreamCaptureModeGlobal
    _call(L, ins, outs, n, 0, prec, s)
    assert h.hipStreamEndCapture(s, C.byref(graph)) == 0
    assert h.hipGraphInstantiate(C.byref(exe), graph, None, None, 0) == 0
    try:
        for rep in range(4):
            for o in outs:
                o.upload(np.full(n, np.nan))
            assert h.hipGraphLaunch(exe, s) == 0
            assert h.hipStreamSynchronize(s) == 0
            got = [o.download() for o in outs]
            print(f"[graph prec={prec}] replay {rep}", flush=True)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), rep
            # eager work between replays: another set through the same scratch
            _call(L, ins2, outs2, n2, 2, prec, s)
            assert h.hipStreamSynchronize(s) == 0
    finally:
        h.hipGraphExecDestroy(exe)
        h.hipGraphDestroy(graph)
        h.hipStreamDestroy(s)
        for a in ins + ins2 + eager + outs + outs2:
            a.free()
