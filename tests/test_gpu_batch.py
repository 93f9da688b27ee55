"""Batched evaluation of independent particle sets in one persistent launch
(sg_compute_batch_device), eager and captured in a HIP graph.

Checks: each set against the FP64 oracle; against a separate single-set
launch; graph replays (several, with other device work in between) bitwise
equal to the eager batch; new inputs in the same buffers; argument errors.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity, ics, _lib
from starsmasher_amd.hipmem import DeviceArray

from conftest import rel_errors

pytestmark = pytest.mark.gpu

TOL = {SG_PREC_FP32C: (1e-5, 1e-6), SG_PREC_FP64: (1e-11, 1e-12)}


@pytest.fixture(scope="module")
def grav():
    g = Gravity(0)
    yield g
    g.close()


def _sets(sizes, seed=0):
    out = []
    for k, n in enumerate(sizes):
        if k % 2 == 0:
            d = ics.plummer(int(n), seed=seed + k)
        else:
            pos, m = ics.polytrope(int(n), 1.5, 1.0, 1.0, seed=seed + k)
            d = {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy(), "m": m,
                 "hp": ics.knn_hp(pos), "u": np.ones(int(n))}
            d["range2"] = ics.range2_of(d["hp"], d["u"])
        if k == 1:
            d["range2"][-1] = -abs(d["range2"][-1])     # a point particle
        out.append(d)
    return out


def _check(tag, prec, res, ref):
    ea, ep = rel_errors(np.stack(res[:3], -1), res[3], np.stack(ref[:3], -1), ref[3])
    amax, pmax = TOL[prec]
    print(f"[{tag}] acc max {ea.max():.3e} pot max {ep.max():.3e}")
    assert ea.max() <= amax and ep.max() <= pmax, tag


SIZES = [300, 5000, 1, 2049, 777, 4096]


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle(grav, prec, graph):
    sets = _sets(SIZES, seed=3)
    res = grav.compute_batch(sets, nkernel=0, precision=prec, use_graph=graph)
    for k, (d, r) in enumerate(zip(sets, res)):
        ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=0)
        if d["x"].size == 1:
            assert r[0][0] == 0.0 and abs(r[3][0] - ref[3][0]) <= 1e-12 * abs(ref[3][0])
            continue
        _check(f"batch set{k} n={d['x'].size} prec={prec} graph={graph}", prec, r, ref)


@pytest.mark.parametrize("nk", [1, 2])
def test_batch_wendland_kernels(grav, nk):
    """The batch kernels' Wendland W3,3 and C4 instantiations, with a point
    particle, against the FP64 oracle in both precisions."""
    sets = _sets([1500, 2049, 700], seed=11)
    for prec in (SG_PREC_FP32C, SG_PREC_FP64):
        res = grav.compute_batch(sets, nkernel=nk, precision=prec)
        for k, (d, r) in enumerate(zip(sets, res)):
            ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=nk)
            _check(f"batch nk={nk} set{k} n={d['x'].size} prec={prec}", prec, r, ref)


def test_batch_matches_single_launches(grav):
    sets = _sets([3000, 8000, 1200], seed=5)
    for prec in (SG_PREC_FP32C, SG_PREC_FP64):
        res = grav.compute_batch(sets, precision=prec)
        for d, r in zip(sets, res):
            one = grav.compute(d["x"], d["y"], d["z"], d["m"], d["range2"], precision=prec)
            _check(f"batch-vs-single prec={prec}", prec, r, one)


def test_graph_replays_equal_eager(grav):
    sets = _sets([4000, 2500, 6100], seed=7)
    sizes = [len(s["x"]) for s in sets]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    dev = [DeviceArray.from_host(np.concatenate([s[k] for s in sets]))
           for k in ("x", "y", "z", "m", "range2")]
    out = [DeviceArray(int(off[-1])) for _ in range(4)]
    grav.compute_batch_device(off, *dev, *out, use_graph=False)
    grav.synchronize()
    eager = [o.download() for o in out]
    other = sets[0]
    for rep in range(4):
        grav.compute_batch_device(off, *dev, *out, use_graph=True)
        grav.synchronize()
        got = [o.download() for o in out]
        for a, b in zip(got, eager):
            np.testing.assert_array_equal(a, b, err_msg=f"replay {rep}")
        # unrelated eager work between replays
        grav.compute(other["x"], other["y"], other["z"], other["m"], other["range2"])
    # new positions in the same buffers: the graph is replayed on new data
    moved = {k: np.concatenate([s[k] for s in sets]) for k in ("x", "y", "z", "m", "range2")}
    moved["x"] = moved["x"] * 1.1 + 0.01
    dev[0].upload(moved["x"])
    grav.compute_batch_device(off, *dev, *out, use_graph=True)
    grav.synchronize()
    got = [o.download() for o in out]
    grav.compute_batch_device(off, *dev, *out, use_graph=False)
    grav.synchronize()
    for a, b in zip(got, [o.download() for o in out]):
        np.testing.assert_array_equal(a, b)
    r = oracle.direct(moved["x"][:4000], moved["y"][:4000], moved["z"][:4000], moved["m"][:4000],
                      moved["range2"][:4000], nkernel=0, rows=np.arange(0, 4000, 37))
    _check("moved set0", SG_PREC_FP32C, [g[:4000][::37] for g in got], r)


@pytest.mark.parametrize("prec,wg,ipl", [(p, w, i) for p in (SG_PREC_FP32C, SG_PREC_FP64)
                                          for w in (64, 128, 256) for i in (1, 2)])
def test_batch_configurations(grav, prec, wg, ipl):
    sets = _sets([2600, 700], seed=9)
    old = grav.get_batch_config(prec)
    grav.set_batch_config(prec, wg, ipl)
    try:
        res = grav.compute_batch(sets, precision=prec, use_graph=True)
    finally:
        grav.set_batch_config(prec, *old)
    for k, (d, r) in enumerate(zip(sets, res)):
        _check(f"cfg {wg}x{ipl} set{k}", prec, r,
               oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=0))


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_batch_partitioned_sets(grav, nk):
    """Sets of more than 64 tiles (16,384 particles) have their j range cut
    into strided partitions with partial sums of their own: 40,000 particles
    (three partitions), 20,000 and 16,385 (65 tiles: two) beside a
    one-partition set, both precisions and the three softening kernels
    against oracle rows; the graph replay equals the eager batch."""
    sets = _sets([40000, 3000, 20000, 16385], seed=11)
    rng = np.random.default_rng(2)
    for prec in (SG_PREC_FP32C, SG_PREC_FP64):
        res = grav.compute_batch(sets, nkernel=nk, precision=prec)
        again = grav.compute_batch(sets, nkernel=nk, precision=prec, use_graph=True)
        for k, (d, r, r2) in enumerate(zip(sets, res, again)):
            for a, b in zip(r, r2):
                np.testing.assert_array_equal(a, b)
            n = d["x"].size
            rows = np.unique(np.concatenate([rng.choice(n, min(n, 300), replace=False),
                                             [0, n - 1]]))
            ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=nk,
                                rows=rows)
            _check(f"partitioned nk={nk} set{k} n={n} prec={prec}", prec, [x[rows] for x in r],
                   ref)


def test_batch_argument_errors(grav):
    L = _lib.load()
    a = [DeviceArray(10) for _ in range(9)]
    ptr = [C.c_void_p(x.ptr) for x in a]

    def call(offs):
        o = np.asarray(offs, dtype=np.int64)
        return L.sg_compute_batch_device(grav._h, o.size - 1, o.ctypes.data_as(C.POINTER(C.c_longlong)),
                                         *ptr[:5], 0, 0, 0, *ptr[5:])
    assert call([0, 4, 4, 10]) == _lib.SG_ERR_ARG        # empty set
    assert call([1, 4, 10]) == _lib.SG_ERR_ARG           # offsets[0] != 0
    assert call([0, 6, 3]) == _lib.SG_ERR_ARG            # decreasing


def test_batch64_full_config(grav):
    """The 64-star relaxation batch at full size (2.6M particles, N per star
    log-uniform in [2048, 131072]) in one graph-captured launch: every star
    against the FP64 batch on all particles and against the oracle on 16
    random rows plus its 4 worst FP32C rows."""
    stars = ics.batch_set(64)
    res32 = grav.compute_batch(stars, precision=SG_PREC_FP32C, use_graph=True)
    res64 = grav.compute_batch(stars, precision=SG_PREC_FP64, use_graph=True)
    rng = np.random.default_rng(0)
    worst = 0.0
    for k, (d, r32, r64) in enumerate(zip(stars, res32, res64)):
        ea, ep = rel_errors(np.stack(r32[:3], -1), r32[3], np.stack(r64[:3], -1), r64[3])
        worst = max(worst, ea.max())
        assert ea.max() <= 1e-5 and ep.max() <= 1e-6, (k, ea.max(), ep.max())
        n = d["x"].size
        rows = np.unique(np.concatenate([rng.choice(n, 16, replace=False), np.argsort(ea)[-4:]]))
        ref = oracle.direct(d["x"], d["y"], d["z"], d["m"], d["range2"], nkernel=0, rows=rows)
        _check(f"batch64 star{k} n={n}", SG_PREC_FP32C, [x[rows] for x in r32], ref)
        _check(f"batch64 star{k} n={n}", SG_PREC_FP64, [x[rows] for x in r64], ref)
    print(f"[batch64] worst star acc max vs FP64 batch {worst:.3e}")
