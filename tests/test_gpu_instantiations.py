"""Every kernel instantiation a caller can select, against the FP64 oracle.

The tile shape (workgroup size wg, i-particles per lane ipl, j-partitions)
is chosen per precision with sg_set_config / sg_set_batch_config, the
softening kernel per call, and the legacy entry points' precision by the
SPHGRAV_PRECISION environment variable.  Each combination instantiates its
own kernels (k_far_fp32c<wg, ipl, nk>, k_far_fp64<wg, ipl> + k_near<ipl, nk>,
k_far_batch<wg, ipl, fp64, nk>, k_adj_far<wg, fp64, nk>,
k_far_fp32c_split<wg> + k_near_psi<fast, nk>), so each is run here on a set
with point particles and a distinct coincident pair.

Reference semantics: grav_force_direct.cu:55-396 (the three softening
kernels), :110-112 (self term), cpu_grav.f:139-145 (point particles).

Tolerances (per particle, relative): as tests/test_gpu_parity.py --
FP64 acceleration <= 1e-11, potential <= 1e-12; FP32C acceleration max
<= 1e-5, median <= 1e-6, potential <= 1e-6.  Adjoint and grad-h term: as
tests/test_gpu_adjoint.py and test_grad_h_psi.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity, ics, _lib

from conftest import rel_errors

pytestmark = pytest.mark.gpu

TOL = {SG_PREC_FP64: (1e-11, 1e-11, 1e-12), SG_PREC_FP32C: (1e-5, 1e-6, 1e-6)}
SHAPES = [(w, i) for w in (64, 128, 256) for i in (1, 2)]
PREC = [SG_PREC_FP32C, SG_PREC_FP64]


def _check(tag, prec, res, ref):
    ea, ep = rel_errors(np.stack(res[:3], -1), res[3], np.stack(ref[:3], -1), ref[3])
    amax, amed, pmax = TOL[prec]
    print(f"[{tag}] acc max {ea.max():.3e} med {np.median(ea):.3e} pot max {ep.max():.3e}",
          flush=True)
    assert np.all(np.isfinite(ea)) and np.all(np.isfinite(ep)), tag
    assert ea.max() <= amax and np.median(ea) <= amed and ep.max() <= pmax, tag


def _args(d):
    return d["x"], d["y"], d["z"], d["m"], d["range2"]


@pytest.fixture(scope="module")
def grav():
    g = Gravity(0)
    yield g
    g.close()


@pytest.fixture(scope="module")
def mixed_set():
    """6,000-particle Plummer sphere with five point particles and a
    distinct coincident pair, and its oracle results for each kernel."""
    d = ics.plummer(6000, seed=31)
    d["range2"][-5:] *= -1.0
    for k in ("x", "y", "z"):
        d[k][41] = d[k][40]
    refs = {nk: oracle.direct(*_args(d), nkernel=nk) for nk in (0, 1, 2)}
    return d, refs


@pytest.mark.parametrize("jsplit", [0, 1])
@pytest.mark.parametrize("wg,ipl", SHAPES)
@pytest.mark.parametrize("nk", [1, 2])
@pytest.mark.parametrize("prec", PREC)
def test_forward_shapes_wendland(grav, mixed_set, prec, nk, wg, ipl, jsplit):
    """Wendland W3,3 (nk=1) and C4 (nk=2) at every tile shape, with the
    automatic j-split and with one j-partition."""
    d, refs = mixed_set
    old = grav.get_config(prec)
    grav.set_config(prec, wg, ipl, jsplit)
    try:
        res = grav.compute(*_args(d), nkernel=nk, precision=prec)
    finally:
        grav.set_config(prec, *old)
    _check(f"fwd nk={nk} {wg}x{ipl}/{jsplit} prec={prec}", prec, res, refs[nk])


def _batch_sets():
    out = []
    for k, n in enumerate((2600, 700, 1)):
        d = ics.plummer(n, seed=40 + k) if n > 1 else None
        if d is None:
            d = {"x": np.array([0.3]), "y": np.array([-0.1]), "z": np.array([0.2]),
                 "m": np.array([0.5]), "range2": np.array([0.04])}
        out.append(d)
    out[0]["range2"][-2:] *= -1.0      # point particles
    for k in ("x", "y", "z"):
        out[1][k][9] = out[1][k][8]    # a distinct coincident pair
    return out


@pytest.fixture(scope="module")
def batch_sets():
    sets = _batch_sets()
    refs = {nk: [oracle.direct(*_args(d), nkernel=nk) for d in sets] for nk in (1, 2)}
    return sets, refs


@pytest.mark.parametrize("wg,ipl", SHAPES)
@pytest.mark.parametrize("nk", [1, 2])
@pytest.mark.parametrize("prec", PREC)
def test_batch_shapes_wendland(grav, batch_sets, prec, nk, wg, ipl):
    """The persistent batch kernel at every tile shape with both Wendland
    kernels, graph-captured."""
    sets, refs = batch_sets
    old = grav.get_batch_config(prec)
    grav.set_batch_config(prec, wg, ipl)
    try:
        res = grav.compute_batch(sets, nkernel=nk, precision=prec, use_graph=True)
    finally:
        grav.set_batch_config(prec, *old)
    for k, (r, ref) in enumerate(zip(res, refs[nk])):
        if sets[k]["x"].size == 1:
            assert r[0][0] == 0.0 and abs(r[3][0] - ref[3][0]) <= 1e-12 * abs(ref[3][0])
            continue
        _check(f"batch nk={nk} {wg}x{ipl} set{k} prec={prec}", prec, r, ref)


@pytest.fixture(scope="module")
def adj_set():
    d = ics.adjoint_set(131_072, seed=5)
    n = 4096
    x = np.stack([d["x"][:n], d["y"][:n], d["z"][:n]], 1)
    hp = ics.knn_hp(x)
    point = np.zeros(n, dtype=bool)
    point[-3:] = True
    x[7] = x[6]
    u = (~point).astype(float)
    r2 = ics.range2_of(hp, u)
    w = d["w"][:n].copy()
    cu = d["cot_u"][:n].copy()
    refs = {nk: oracle.vjp(x[:, 0], x[:, 1], x[:, 2], d["m"][:n], r2, w, cu, nkernel=nk)
            for nk in (0, 1, 2)}
    return x, d["m"][:n].copy(), r2, w, cu, refs


@pytest.mark.parametrize("wg", [64, 128])
@pytest.mark.parametrize("nk", [0, 1, 2])
@pytest.mark.parametrize("prec", PREC)
def test_adjoint_workgroup_sizes(grav, adj_set, prec, nk, wg):
    """The adjoint's far kernel takes the forward workgroup size of its
    precision: k_adj_far<64|128, fp64, nk> against the oracle's full FP64
    gradient (median and rms-normalised maximum as in test_gpu_adjoint)."""
    x, m, r2, w, cu, refs = adj_set
    old = grav.get_config(prec)
    grav.set_config(prec, wg, 1, 0)
    try:
        g = grav.vjp(x[:, 0], x[:, 1], x[:, 2], m, r2, np.ascontiguousarray(w[:, 0]),
                     np.ascontiguousarray(w[:, 1]), np.ascontiguousarray(w[:, 2]), cu,
                     nkernel=nk, precision=prec)
    finally:
        grav.set_config(prec, *old)
    lim = 1e-5 if prec == SG_PREC_FP32C else 1e-10
    ox, oy, oz, om, oH = refs[nk]
    for key, got, ref in (("x", np.stack(g[:3], 1), np.stack([ox, oy, oz], 1)), ("m", g[3], om),
                          ("H", g[4], oH)):
        dd = np.abs(got - ref) if got.ndim == 1 else np.linalg.norm(got - ref, axis=1)
        mag = np.abs(ref) if ref.ndim == 1 else np.linalg.norm(ref, axis=1)
        med = float(np.median(dd / np.maximum(mag, 1e-300)))
        rms = float(dd.max() / np.sqrt(np.mean(mag ** 2)))
        print(f"[adj wg={wg} nk={nk} prec={prec} {key}] median {med:.3e} rms-max {rms:.3e}")
        assert med <= lim and rms <= lim, (key, med, rms)


@pytest.mark.parametrize("wg", [64, 128])
@pytest.mark.parametrize("nk", [0, 1, 2])
@pytest.mark.parametrize("prec", PREC)
def test_grad_h_pass_workgroup_sizes(grav, mixed_set, prec, nk, wg):
    """The pass that forms the grad-h term runs the far kernel at the
    forward workgroup size with one i per lane (k_far_fp32c_split<64|128> or
    k_far_fp64<64|128, 1>) and the psi near kernel: a, phi against the
    oracle and psi against the oracle's exact sum."""
    d, refs = mixed_set
    old = grav.get_config(prec)
    grav.set_config(prec, wg, 1, 0)
    try:
        ax, ay, az, pot, psi = grav.compute_psi(*_args(d), nkernel=nk, precision=prec)
    finally:
        grav.set_config(prec, *old)
    _check(f"psi-pass nk={nk} wg={wg} prec={prec}", prec, (ax, ay, az, pot), refs[nk])
    rows = np.arange(0, d["x"].size, 7)
    want = oracle.psi(*_args(d), nkernel=nk, rows=rows)
    gas = want != 0.0
    rel = np.abs(psi[rows][gas] - want[gas]) / np.abs(want[gas])
    print(f"[psi wg={wg} nk={nk} prec={prec}] max {rel.max():.3e}")
    assert rel.max() <= 1e-11 and np.all(psi[rows][~gas] == 0.0)


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_legacy_entry_points_fp64(mixed_set, nk, monkeypatch):
    """firsthalf_grav_forces_ / lasthalf_grav_forces_ with
    SPHGRAV_PRECISION=fp64 run the plain FP64 kernels (read at each
    firsthalf): results at the FP64 tolerance, written over the i-range only."""
    d, refs = mixed_set
    L = _lib.load()
    monkeypatch.setenv("SPHGRAV_PRECISION", "fp64")
    monkeypatch.setenv("SPHGRAV_QUIET", "1")
    n, n_lower, my_len = d["x"].size, 1201, 3500
    ip = lambda v: C.byref(C.c_int(v))
    dp = C.POINTER(C.c_double)
    arrs = [np.ascontiguousarray(a) for a in _args(d)]
    L.firsthalf_grav_forces_(ip(n), ip(n_lower), ip(my_len), *[a.ctypes.data_as(dp) for a in arrs],
                             ip(1), ip(nk))
    assert L.sg_last_status() == 0
    outs = [np.full(n, 12345.0) for _ in range(4)]
    L.lasthalf_grav_forces_(ip(n), *[o.ctypes.data_as(dp) for o in outs], ip(0))
    assert L.sg_last_status() == 0
    lo, hi = n_lower - 1, n_lower - 1 + my_len
    for o in outs:
        assert np.all(o[:lo] == 12345.0) and np.all(o[hi:] == 12345.0)
    ref = [r[lo:hi] for r in refs[nk]]
    _check(f"legacy fp64 nk={nk}", SG_PREC_FP64, [o[lo:hi] for o in outs], ref)
    # and the default (FP32C) again once the variable is gone
    monkeypatch.delenv("SPHGRAV_PRECISION")
    L.firsthalf_grav_forces_(ip(n), ip(n_lower), ip(my_len), *[a.ctypes.data_as(dp) for a in arrs],
                             ip(1), ip(nk))
    L.lasthalf_grav_forces_(ip(n), *[o.ctypes.data_as(dp) for o in outs], ip(0))
    assert L.sg_last_status() == 0
    _check(f"legacy fp32c nk={nk}", SG_PREC_FP32C, [o[lo:hi] for o in outs], ref)


@pytest.mark.parametrize("wg,ipl", [(256, 1), (128, 2)])
def test_precision_stats_count_softened_pairs(grav, mixed_set, wg, ipl):
    """sg_last_precision_stats of the FP32C pass: the pairs it evaluates in
    FP64 are those with r^2 < max(H_i^2, H_j^2) (self pairs included),
    counted exactly on the CPU here (the kernel decides in FP32, so pairs
    within rounding of the boundary may move: at most 0.1%); wide and near
    chunk visits are bounded by all visits."""
    d, _ = mixed_set
    old = grav.get_config(SG_PREC_FP32C)
    grav.set_config(SG_PREC_FP32C, wg, ipl, 0)
    try:
        grav.compute(*_args(d), nkernel=0, precision=SG_PREC_FP32C)
        visits, nearv, widev, fp64 = grav.last_precision_stats()
    finally:
        grav.set_config(SG_PREC_FP32C, *old)
    pos = np.stack([d["x"], d["y"], d["z"]], 1)
    h2 = np.abs(d["range2"])
    want = 0
    for a in range(0, pos.shape[0], 1000):
        r2 = ((pos[a:a + 1000, None, :] - pos[None, :, :]) ** 2).sum(-1)
        want += int(np.count_nonzero(r2 < np.maximum(h2[a:a + 1000, None], h2[None, :])))
    print(f"[stats {wg}x{ipl}] visits {visits} near {nearv} wide {widev} fp64 pairs {fp64} "
          f"(exact count {want})")
    assert 0 < nearv <= visits and 0 <= widev <= visits - nearv
    assert abs(fp64 - want) <= 1e-3 * want
