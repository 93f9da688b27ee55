"""GPU parity of the HIP gravity library against the FP64 oracle.

Tolerances (relative, per particle; |a - a_ref| / |a_ref| and |phi - phi_ref| / |phi_ref|):
  FP64 kernel : acceleration <= 1e-11, potential <= 1e-12 (summation order only)
  FP32C kernel: acceleration max <= 1e-5 (the path's accuracy target) and
                median <= 1e-6; potential max <= 1e-6
Golden potentials stored by the reference in its dumps: <= 4e-7 (their own
FP32 floor), both precisions.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle
from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity, bonet_psi, ics, _lib
from starsmasher_amd.gravity import firsthalf_grav_forces, lasthalf_grav_forces
from starsmasher_amd.hipmem import DeviceArray

from conftest import rel_errors

pytestmark = pytest.mark.gpu

TOL = {SG_PREC_FP64: (1e-11, 1e-11, 1e-12), SG_PREC_FP32C: (1e-5, 1e-6, 1e-6)}
PNAME = {SG_PREC_FP32C: "fp32c", SG_PREC_FP64: "fp64"}


@pytest.fixture(scope="module")
def grav():
    g = Gravity(0)
    yield g
    g.close()


def _check(tag, prec, res, ref):
    ax, ay, az, pot = res
    rx, ry, rz, rp = ref
    ea, ep = rel_errors(np.stack([ax, ay, az], -1), pot, np.stack([rx, ry, rz], -1), rp)
    amax, amed, pmax = TOL[prec]
    print(f"[{tag} {PNAME[prec]}] acc max {ea.max():.3e} med {np.median(ea):.3e} "
          f"pot max {ep.max():.3e} med {np.median(ep):.3e}", flush=True)
    assert np.all(np.isfinite(ea)) and np.all(np.isfinite(ep))
    assert ea.max() <= amax, (tag, ea.max())
    assert np.median(ea) <= amed, (tag, np.median(ea))
    assert ep.max() <= pmax, (tag, ep.max())


def _args(d):
    return d["x"], d["y"], d["z"], d["m"], d["range2"]


@pytest.fixture(scope="module")
def plummer16k():
    d = ics.plummer(16384, seed=1)
    ref = oracle.direct(*_args(d), nkernel=0)
    return d, ref


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_plummer_16k(grav, plummer16k, prec):
    d, ref = plummer16k
    _check("plummer16k", prec, grav.compute(*_args(d), nkernel=0, precision=prec), ref)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_golden_dump_potentials(grav, golden_dumps, prec):
    for name, d in golden_dumps.items():
        args = (d["x"], d["y"], d["z"], d["am"], d["range2"])
        ax, ay, az, pot = grav.compute(*args, nkernel=0, precision=prec)
        rel = np.abs(pot - d["grpot"]) / np.abs(d["grpot"])
        print(f"[dump {name} {PNAME[prec]}] grpot max {rel.max():.3e} med {np.median(rel):.3e}")
        assert rel.max() < 4e-7
        _check(f"dump {name}", prec, (ax, ay, az, pot), oracle.direct(*args, nkernel=0))


@pytest.mark.parametrize("nk", [1, 2])
@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_wendland_kernels_with_point_particles(grav, nk, prec):
    d = ics.plummer(6000, seed=7)
    d["range2"][-5:] *= -1.0          # five point particles
    _check(f"nkernel{nk}", prec, grav.compute(*_args(d), nkernel=nk, precision=prec),
           oracle.direct(*_args(d), nkernel=nk))


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("ibeg,ni", [(0, 1), (1234, 1700), (0, 4999), (4999, 1), (2500, 2500),
                                     (17, 0)])
def test_i_ranges(grav, prec, ibeg, ni):
    d = ics.plummer(5000, seed=8)
    res = grav.compute(*_args(d), nkernel=0, precision=prec, ibeg=ibeg, ni=ni)
    assert all(r.size == ni for r in res)
    if ni == 0:
        return
    ref = oracle.direct(*_args(d), nkernel=0, rows=np.arange(ibeg, ibeg + ni))
    _check(f"irange{ibeg}+{ni}", prec, res, ref)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("n", [1, 2, 3, 63, 255, 256, 257, 1000])
def test_small_and_ragged_sizes(grav, prec, n):
    r = np.random.default_rng(n)
    pos = r.normal(size=(n, 3))
    m = r.uniform(0.5, 1.5, n) / n
    hp = r.uniform(0.2, 0.8, n)
    u = np.ones(n)
    u[0] = 0.0 if n > 1 else 1.0
    r2 = ics.range2_of(hp, u)
    res = grav.compute(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, nkernel=0, precision=prec)
    ref = oracle.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, nkernel=0)
    if n == 1:
        assert res[0][0] == 0.0 and abs(res[3][0] - ref[3][0]) <= 1e-12 * abs(ref[3][0])
        return
    _check(f"n{n}", prec, res, ref)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_coincident_particles_are_finite(grav, prec):
    """A distinct pair at r = 0 returns the finite r->0 limit (the reference
    kernel returns NaN there, grav_force_direct.cu:110-112)."""
    d = ics.plummer(3000, seed=9)
    for k in (10, 11, 12):
        d["x"][k], d["y"][k], d["z"][k] = d["x"][5], d["y"][5], d["z"][5]
    res = grav.compute(*_args(d), nkernel=0, precision=prec)
    assert all(np.all(np.isfinite(r)) for r in res)
    _check("coincident", prec, res, oracle.direct(*_args(d), nkernel=0))


def test_legacy_fortran_abi(grav):
    """The reference's three symbols, by reference, through ctypes: results
    land in [n_lower-1, n_lower-1+my_length) and nothing else is touched."""
    L = _lib.load()
    d = ics.plummer(4000, seed=11)
    n, n_lower, my_len = 4000, 801, 1500
    ip = lambda v: C.byref(C.c_int(v))
    dp = C.POINTER(C.c_double)
    L.gpu_init_dev_(ip(0), C.byref(C.c_double(0.3)))
    arrs = [np.ascontiguousarray(d[k]) for k in ("x", "y", "z", "m", "range2")]
    os.environ["SPHGRAV_QUIET"] = "1"
    L.firsthalf_grav_forces_(ip(n), ip(n_lower), ip(my_len), *[a.ctypes.data_as(dp) for a in arrs],
                             ip(1), ip(0))
    assert L.sg_last_status() == 0
    outs = [np.full(n, 12345.0) for _ in range(4)]
    L.lasthalf_grav_forces_(ip(n), *[o.ctypes.data_as(dp) for o in outs], ip(0))
    assert L.sg_last_status() == 0
    lo, hi = n_lower - 1, n_lower - 1 + my_len
    for o in outs:
        assert np.all(o[:lo] == 12345.0) and np.all(o[hi:] == 12345.0)
    ref = oracle.direct(*arrs, nkernel=0, rows=np.arange(lo, hi))
    _check("legacy-abi", SG_PREC_FP32C, [o[lo:hi] for o in outs], ref)


def test_python_mirror_of_reference_interface():
    d = ics.plummer(2000, seed=12)
    firsthalf_grav_forces(2000, 1, 2000, d["x"], d["y"], d["z"], d["m"], d["range2"], 1, 0)
    outs = [np.zeros(2000) for _ in range(4)]
    lasthalf_grav_forces(2000, *outs)
    _check("mirror", SG_PREC_FP32C, outs, oracle.direct(*_args(d), nkernel=0))


def test_device_api_matches_host_api(grav):
    d = ics.plummer(7000, seed=13)
    dev = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    out = [DeviceArray(7000) for _ in range(4)]
    grav.compute_device(7000, *dev, *out, nkernel=0, precision=SG_PREC_FP32C)
    grav.synchronize()
    host = grav.compute(*_args(d), nkernel=0, precision=SG_PREC_FP32C)
    for a, b in zip(out, host):
        np.testing.assert_array_equal(a.download(), b)   # same kernels, same order


def test_device_validation_errors(grav):
    d = ics.plummer(1000, seed=14)
    for key, val, status in (("m", 0.0, _lib.SG_ERR_MASS), ("range2", 0.0, _lib.SG_ERR_RANGE2),
                             ("x", np.inf, _lib.SG_ERR_NONFINITE)):
        arrs = {k: d[k].copy() for k in ("x", "y", "z", "m", "range2")}
        arrs[key][37] = val
        dev = [DeviceArray.from_host(arrs[k]) for k in ("x", "y", "z", "m", "range2")]
        out = [DeviceArray(1000) for _ in range(4)]
        grav.compute_device(1000, *dev, *out)
        with pytest.raises(_lib.SphgravError) as e:
            grav.synchronize()
        assert e.value.status == status


@pytest.mark.parametrize("n", [1, 1000, 262_145, 2_000_003])
def test_potential_energy_diagnostic(grav, n):
    """epot = 1/2 sum m phi (output.f:232-252) summed on the device in FP64:
    against the correctly rounded sum of the products (math.fsum; the
    products' own rounding is below 1.2e-16 relative with terms of one sign)
    and the oracle's loop, for sizes from one particle to beyond the 1024
    block sums, and bit-for-bit the same on a second call."""
    rng = np.random.default_rng(n)
    m = rng.uniform(1e-6, 1.0, n)
    pot = -rng.uniform(0.1, 10.0, n)
    dm, dp = DeviceArray.from_host(m), DeviceArray.from_host(pot)
    e1 = grav.potential_energy(n, dm, dp)
    e2 = grav.potential_energy(n, dm, dp)
    assert e1 == e2
    exact = 0.5 * math.fsum((m * pot).tolist())
    assert abs(e1 - exact) <= 1e-13 * abs(exact)
    assert abs(e1 - oracle.epot(m, pot)) <= 1e-12 * abs(exact)


def test_potential_energy_of_gravity_outputs(grav, plummer16k):
    """The diagnostic on the potentials the pair kernel leaves in HBM."""
    d, _ = plummer16k
    n = d["x"].size
    dev = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    out = [DeviceArray(n) for _ in range(4)]
    grav.compute_device(n, *dev, *out, nkernel=0, precision=SG_PREC_FP64)
    grav.synchronize()
    e = grav.potential_energy(n, dev[3], out[3])
    want = oracle.epot(d["m"], out[3].download())
    assert e < 0.0 and abs(e - want) <= 1e-12 * abs(want)
    with pytest.raises(_lib.SphgravError) as err:
        grav.potential_energy(0, dev[3], out[3])
    assert err.value.status == _lib.SG_ERR_ARG


def test_deterministic(grav, plummer16k):
    d, _ = plummer16k
    a = grav.compute(*_args(d), nkernel=0)
    b = grav.compute(*_args(d), nkernel=0)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("prec,wg,ipl", [(SG_PREC_FP32C, w, i) for w in (64, 128, 256) for i in (1, 2)]
                         + [(SG_PREC_FP64, w, i) for w in (64, 128, 256) for i in (1, 2)])
@pytest.mark.parametrize("jsplit", [0, 1, 3])
def test_tile_configurations(grav, plummer16k, prec, wg, ipl, jsplit):
    d, ref = plummer16k
    old = grav.get_config(prec)
    grav.set_config(prec, wg, ipl, jsplit)
    try:
        _check(f"cfg{wg}x{ipl}/{jsplit}", prec, grav.compute(*_args(d), nkernel=0, precision=prec), ref)
    finally:
        grav.set_config(prec, *old)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_automatic_workgroup_size(grav, prec):
    """The default configuration (wg 0) takes 128-lane workgroups for an
    i-range of up to 65,536 particles and 256 above, and gives the same bits
    as asking for that shape explicitly."""
    old = grav.get_config(prec)
    assert old == (0, 1, 0)
    for n, ibeg, ni, want_wg in ((20000, 0, 20000, 128), (70000, 0, 70000, 256),
                                 (70000, 2000, 65536, 128)):
        d = ics.plummer(n, seed=30 + want_wg)
        auto = grav.compute(*_args(d), nkernel=1, precision=prec, ibeg=ibeg, ni=ni)
        wg, ipl, js = grav.last_config()
        print(f"[auto n={n} ni={ni}] resolved wg={wg} ipl={ipl} jsplit={js}")
        assert (wg, ipl) == (want_wg, 1) and js >= 1
        grav.set_config(prec, want_wg, 1, 0)
        try:
            expl = grav.compute(*_args(d), nkernel=1, precision=prec, ibeg=ibeg, ni=ni)
            assert grav.last_config() == (wg, ipl, js)
        finally:
            grav.set_config(prec, *old)
        for a, b in zip(auto, expl):
            np.testing.assert_array_equal(a, b)


def test_last_config_follows_single_set_calls(grav):
    """sg_last_config reports the shape of the last single-set pass; after a
    batch or a psi-only call it reports SG_ERR_STATE instead of a stale one."""
    d = ics.plummer(3000, seed=4)
    grav.compute(*_args(d), nkernel=0)
    assert grav.last_config()[:2] == (128, 1)
    grav.compute_batch([d])
    with pytest.raises(_lib.SphgravError) as e:
        grav.last_config()
    assert e.value.status == _lib.SG_ERR_STATE
    grav.compute(*_args(d), nkernel=0)
    assert grav.last_config()[:2] == (128, 1)
    grav.psi(*_args(d), nkernel=0)
    with pytest.raises(_lib.SphgravError):
        grav.last_config()


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_far_path_dominates_with_small_h(grav, prec):
    """Small softening lengths make almost every tile Newtonian, so the far
    kernel's FP32 body (not the FP64 near path) produces the result."""
    d = ics.plummer(20000, seed=15)
    d["range2"] = d["range2"] * 1e-4          # H / 100
    res = grav.compute(*_args(d), nkernel=0, precision=prec)
    visits, nearv = grav.last_tile_stats()
    print(f"[smallH] near tiles {nearv}/{visits}")
    # boxes of neighbouring curve segments overlap, so some tiles stay near
    assert nearv < 0.5 * visits
    _check("smallH", prec, res, oracle.direct(*_args(d), nkernel=0))


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("wg,ipl,jsplit", [(64, 1, 1), (256, 1, 1), (128, 2, 1), (256, 1, 0)])
def test_waves_near_every_tile(grav, prec, wg, ipl, jsplit):
    """40 gas particles with H = 100 (ten times the set's radius): every pair
    they take part in is softened, so their waves find every one of the 256
    tiles near and their chunks are near for every wave.  With one
    j-partition a wave meets more near tiles than its deferred near list
    holds (128), which sends the rest through the list-overflow path."""
    d = ics.plummer(65536, seed=21)
    rng = np.random.default_rng(5)
    big = rng.choice(65536, 40, replace=False)
    d["range2"][big] = 1.0e4
    old = grav.get_config(prec)
    grav.set_config(prec, wg, ipl, jsplit)
    try:
        res = grav.compute(*_args(d), nkernel=0, precision=prec)
        visits, nearv = grav.last_tile_stats()
    finally:
        grav.set_config(prec, *old)
    print(f"[bigH {wg}x{ipl}/{jsplit}] near chunk visits {nearv}/{visits}")
    assert nearv > 0
    rows = np.unique(np.concatenate([big, rng.choice(65536, 400, replace=False)]))
    ref = oracle.direct(*_args(d), nkernel=0, rows=rows)
    _check(f"bigH {wg}x{ipl}/{jsplit}", prec, [r[rows] for r in res], ref)


@pytest.fixture(scope="module")
def giant64k():
    return ics.giant_with_core(65536, seed=16)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_giant_with_point_mass_rows(grav, giant64k, prec, nk):
    """Masses over 4 decades, H over 3 decades and a 30% point mass (no self
    potential, softening H = 0.2), for the three softening kernels, against
    the oracle on sampled rows."""
    d = giant64k
    res = grav.compute(*_args(d), nkernel=nk, precision=prec)
    rng = np.random.default_rng(1)
    rows = np.unique(np.concatenate([rng.choice(65536, 1500, replace=False), [65535]]))
    ref = oracle.direct(*_args(d), nkernel=nk, rows=rows)
    _check(f"giant64k rows nk={nk}", prec, [r[rows] for r in res], ref)


@pytest.fixture(scope="module")
def two_poly():
    return ics.two_polytropes()


def test_two_polytropes_262k(grav, two_poly):
    """Full size: FP32C against the FP64 kernel on every particle, against the
    oracle on sampled rows (including the worst FP32C rows), and momentum
    conservation of both."""
    d = two_poly
    r64 = grav.compute(*_args(d), nkernel=0, precision=SG_PREC_FP64)
    r32 = grav.compute(*_args(d), nkernel=0, precision=SG_PREC_FP32C)
    ea, ep = rel_errors(np.stack(r32[:3], -1), r32[3], np.stack(r64[:3], -1), r64[3])
    print(f"[twopoly fp32c vs fp64 kernel] acc max {ea.max():.3e} med {np.median(ea):.3e} "
          f"pot max {ep.max():.3e}")
    rng = np.random.default_rng(0)
    rows = np.unique(np.concatenate([rng.choice(d["x"].size, 384, replace=False),
                                     np.argsort(ea)[-64:], np.argsort(ep)[-64:]]))
    ref = oracle.direct(*_args(d), nkernel=0, rows=rows)
    _check("twopoly rows", SG_PREC_FP64, [r[rows] for r in r64], ref)
    _check("twopoly rows", SG_PREC_FP32C, [r[rows] for r in r32], ref)
    assert ea.max() <= 1e-5 and ep.max() <= 1e-6
    m = d["m"]
    for r in (r64, r32):
        a = np.stack(r[:3], -1)
        tot = np.linalg.norm((m[:, None] * a).sum(0))
        assert tot <= 1e-7 * (m * np.linalg.norm(a, axis=1)).sum()


def test_giant_1m_full_size(grav):
    """The 1,048,576-particle giant star with a 30% point mass at full size:
    FP32C against the FP64 kernel on every particle, both against the oracle
    on 256 random rows plus the 32 worst FP32C rows (the FP32C bound is the
    max relative acceleration error 1e-5 against FP64)."""
    d = ics.giant_with_core(1 << 20)
    n = d["x"].size
    r64 = grav.compute(*_args(d), nkernel=0, precision=SG_PREC_FP64)
    r32 = grav.compute(*_args(d), nkernel=0, precision=SG_PREC_FP32C)
    ea, ep = rel_errors(np.stack(r32[:3], -1), r32[3], np.stack(r64[:3], -1), r64[3])
    print(f"[giant1m fp32c vs fp64 kernel] acc max {ea.max():.3e} med {np.median(ea):.3e} "
          f"pot max {ep.max():.3e} med {np.median(ep):.3e}")
    rng = np.random.default_rng(0)
    rows = np.unique(np.concatenate([rng.choice(n, 256, replace=False), np.argsort(ea)[-32:],
                                     np.argsort(ep)[-16:]]))
    ref = oracle.direct(*_args(d), nkernel=0, rows=rows)
    _check("giant1m rows", SG_PREC_FP64, [r[rows] for r in r64], ref)
    _check("giant1m rows", SG_PREC_FP32C, [r[rows] for r in r32], ref)
    assert ea.max() <= 1e-5 and ep.max() <= 1e-6


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_grad_h_psi(grav, plummer16k, nk, prec):
    """psi_i = hp_i^-2 sum_j m_j dphidh(r_ij/hp_i) (advance.f90:716-737) from
    the near-field pass, with point particles and a coincident pair, against
    the oracle's exact sum and the reference's 10^7-entry table lookup; the
    accelerations and potentials of the same call match the plain call's."""
    d = {k: np.array(v, copy=True) for k, v in plummer16k[0].items()}
    d["range2"][-5:] = -np.abs(d["range2"][-5:])      # point particles: psi = 0
    for k in ("x", "y", "z"):
        d[k][11] = d[k][10]                            # a distinct coincident pair
    n = d["x"].size
    ax, ay, az, pot, psi = grav.compute_psi(*_args(d), nkernel=nk, precision=prec)
    plain = grav.compute(*_args(d), nkernel=nk, precision=prec)
    if prec == SG_PREC_FP64:
        # the same FP64 far kernel and the same FP64 near pair in both passes
        for a, b in zip((ax, ay, az, pot), plain):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
    else:
        # the plain FP32C pass evaluates near chunks inside its far kernel, the
        # psi pass in the near kernel: the same pairs, summed in another order
        ea, ep = rel_errors(np.stack([ax, ay, az], -1), pot, np.stack(plain[:3], -1), plain[3])
        assert ea.max() <= 1e-6 and ep.max() <= 1e-7, (ea.max(), ep.max())
    rows = np.unique(np.concatenate([np.arange(0, n, 61), [10, 11, n - 1]]))
    ref = oracle.psi(*_args(d), nkernel=nk, rows=rows)
    tab = oracle.psi(*_args(d), nkernel=nk, rows=rows, tabulated=True)
    gas = ref != 0.0
    rel = np.abs(psi[rows][gas] - ref[gas]) / np.abs(ref[gas])
    print(f"[psi nk={nk} prec={prec}] vs exact max {rel.max():.3e}; "
          f"exact vs table max {np.max(np.abs(tab[gas] - ref[gas]) / np.abs(ref[gas])):.3e}")
    # the Wendland polynomials' coefficients (up to 9009) cancel near q = 1,
    # so summation order alone moves psi by up to ~1e-12 in FP64
    assert rel.max() <= 1e-11
    assert np.all(psi[rows][~gas] == 0.0) and np.all(psi[-5:] == 0.0)
    assert np.max(np.abs(psi[rows][gas] - tab[gas]) / np.abs(tab[gas])) <= 1e-6
    # an i-range of the same set
    sub = grav.compute_psi(*_args(d), nkernel=nk, precision=prec, ibeg=3000, ni=5000)
    np.testing.assert_allclose(sub[4], psi[3000:8000], rtol=1e-11, atol=0)
    # the psi-only pass (sg_psi_host), whole set and an i-range, and the
    # Fortran-callable form through the Python mirror (range from hp and u)
    alone = grav.psi(*_args(d), nkernel=nk)
    np.testing.assert_allclose(alone, psi, rtol=1e-11, atol=0)
    np.testing.assert_allclose(grav.psi(*_args(d), nkernel=nk, ibeg=3000, ni=5000),
                               psi[3000:8000], rtol=1e-11, atol=0)
    u = np.where(d["range2"] > 0, 1.0, 0.0)
    hp = 0.5 * np.sqrt(np.abs(d["range2"]))
    mirror = bonet_psi(d["x"], d["y"], d["z"], d["m"], hp, u, nkernel=nk, ngrav_lower=101,
                       ngrav_upper=9000)
    np.testing.assert_allclose(mirror[100:9000], psi[100:9000], rtol=1e-11, atol=0)
    assert np.all(mirror[:100] == 0.0) and np.all(mirror[9000:] == 0.0)
