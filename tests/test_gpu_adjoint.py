"""The HIP adjoint (sg_vjp_host / sg_vjp_device) behind a torch.autograd.Function.

Loss L = sum_i w_i . a_i + u_i phi_i with random FP32 cotangents (unit
Gaussian).  Checks:
  * autograd gradients w.r.t. positions, masses and h against the FP64
    oracle's full gradient (oracle.vjp) on a 4,096-particle subset of the
    131,072-particle Plummer sphere, for every kernel, point particles and a
    coincident pair;
  * on 256 sampled particles of the full 131,072-particle sphere, against
    the oracle's FP64 adjoint rows and FP64 central finite differences
    (step 1e-6) of the oracle's loss;
  * torch.autograd.gradcheck of the FP64 path on a small set;
  * the device-pointer entry point against the host one, argument errors.

Error measures per particle k: |g_k - g_ref,k| / |g_ref,k| (3-vector norm
for positions), and the same difference over the root-mean-square of the
reference gradient over all particles (errors of sums whose terms cancel
are judged against the size of the gradient field, not of one entry).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity, ics, _lib
from starsmasher_amd.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

N_FULL = 131_072
N_SUB = 4_096

# Per-particle relative median and rms-normalised max, per precision.  The
# per-particle maximum is not bounded: a gradient entry whose pair terms
# cancel (random-sign cotangents) magnifies any rounding difference, e.g. the
# oracle differentiates each decimal coefficient term by term while the
# kernels use P'(q) = q A(q), coefficients that differ in the 13th digit.
TOL = {
    SG_PREC_FP32C: dict(rel_median=1e-5, rms_max=1e-5),
    SG_PREC_FP64: dict(rel_median=1e-10, rms_max=1e-10),
}


@pytest.fixture(scope="module")
def tg():
    _lib.load()     # bind the library (and its HIP runtime) before torch loads its own
    import torch_gravity
    ctx = torch_gravity.Context(0)
    yield torch_gravity, ctx
    ctx.close()


@pytest.fixture(scope="module")
def full_set():
    return ics.adjoint_set(N_FULL, seed=5)


def _subset(d, n=N_SUB, nkernel_point=True):
    x = np.stack([d["x"][:n], d["y"][:n], d["z"][:n]], 1)
    hp = ics.knn_hp(x)                          # 64 neighbours within the subset
    point = np.zeros(n, dtype=bool)
    if nkernel_point:
        point[-3:] = True                       # point particles (range2 < 0)
    x[7] = x[6]                                 # a distinct coincident pair
    return x, d["m"][:n].copy(), hp, point, d["w"][:n].copy(), d["cot_u"][:n].copy()


def _autograd(tg, x, m, hp, point, w, u, nkernel, prec):
    import torch
    torch_gravity, ctx = tg
    X = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    M = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    Hh = torch.tensor(hp, dtype=torch.float64, requires_grad=True)
    P = torch.tensor(point)
    acc, pot = torch_gravity.gravity(X, M, Hh, P, ctx, nkernel=nkernel, precision=prec)
    loss = (acc * torch.tensor(w)).sum() + (pot * torch.tensor(u)).sum()
    loss.backward()
    return X.grad.numpy(), M.grad.numpy(), Hh.grad.numpy()


def _errors(tag, got, ref):
    """got/ref: (n, 3) position gradients or (n,) scalars."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    d = np.abs(got - ref) if got.ndim == 1 else np.linalg.norm(got - ref, axis=1)
    mag = np.abs(ref) if ref.ndim == 1 else np.linalg.norm(ref, axis=1)
    rel = d / np.maximum(mag, 1e-300)
    rms = np.sqrt(np.mean(mag ** 2))
    out = dict(rel_max=float(rel.max()), rel_median=float(np.median(rel)),
               rms_max=float(d.max() / rms))
    print(f"[{tag}] rel max {out['rel_max']:.3e} median {out['rel_median']:.3e} "
          f"rms-normalised max {out['rms_max']:.3e}", flush=True)
    return out


def _assert_tol(tag, e, prec):
    for key, lim in TOL[prec].items():
        assert e[key] <= lim, (tag, key, e[key], lim)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_autograd_matches_oracle_gradient(tg, full_set, nk, prec):
    x, m, hp, point, w, u = _subset(full_set)
    gx, gm, gh = _autograd(tg, x, m, hp, point, w, u, nk, prec)
    r2 = ics.range2_of(hp, (~point).astype(float))
    ox, oy, oz, om, oH = oracle.vjp(x[:, 0], x[:, 1], x[:, 2], m, r2, w, u, nkernel=nk)
    tag = f"subset n={N_SUB} nk={nk} prec={prec}"
    _assert_tol(tag + " x", _errors(tag + " x", gx, np.stack([ox, oy, oz], 1)), prec)
    _assert_tol(tag + " m", _errors(tag + " m", gm, om), prec)
    _assert_tol(tag + " h", _errors(tag + " h", gh, 2.0 * oH), prec)


@pytest.mark.parametrize("prec", [SG_PREC_FP32C, SG_PREC_FP64])
def test_full_sphere_oracle_rows_and_finite_differences(tg, full_set, prec):
    """All 131,072 particles (the far kernel's FP32 path runs here): 256
    sampled gradients against the oracle's FP64 adjoint rows and against
    FP64 central differences of the oracle's loss."""
    d = full_set
    n = N_FULL
    x = np.stack([d["x"], d["y"], d["z"]], 1)
    point = np.zeros(n, dtype=bool)
    gx, gm, gh = _autograd(tg, x, d["m"], d["hp"], point, d["w"], d["cot_u"], 0, prec)
    ks = np.random.default_rng(11).choice(n, 256, replace=False)
    ox, oy, oz, om, oH = oracle.vjp(d["x"], d["y"], d["z"], d["m"], d["range2"], d["w"],
                                    d["cot_u"], nkernel=0, rows=ks)
    tag = f"oracle rows n={n} prec={prec}"
    _assert_tol(tag + " x", _errors(tag + " x", gx[ks], np.stack([ox, oy, oz], 1)), prec)
    _assert_tol(tag + " m", _errors(tag + " m", gm[ks], om), prec)
    _assert_tol(tag + " h", _errors(tag + " h", gh[ks], 2.0 * oH), prec)
    fd = oracle.fd_grad(ks, d["x"], d["y"], d["z"], d["m"], d["range2"], d["w"], d["cot_u"],
                        nkernel=0, step=1e-6)
    tag = f"FD n={n} 256 sampled prec={prec}"
    # central differences carry ~1e-8 relative truncation and rounding error
    lim = 2e-5 if prec == SG_PREC_FP32C else 1e-6
    for key, got, ref in (("x", gx[ks], fd[:, :3]), ("m", gm[ks], fd[:, 3]),
                          ("h", gh[ks], 2.0 * fd[:, 4])):
        e = _errors(f"{tag} {key}", got, ref)
        assert e["rel_median"] <= lim and e["rms_max"] <= lim, (tag, key, e)


def test_gradcheck_fp64(tg):
    import torch
    torch_gravity, ctx = tg
    rng = np.random.default_rng(3)
    n = 24
    x = rng.normal(size=(n, 3))
    x[5] = x[4]
    X = torch.tensor(x, requires_grad=True)
    M = torch.tensor(rng.uniform(0.5, 2.0, n) / n, requires_grad=True)
    Hh = torch.tensor(rng.uniform(0.3, 0.9, n), requires_grad=True)
    P = torch.tensor(np.arange(n) >= n - 2)
    for nk in (0, 1, 2):
        fn = lambda a, b, c: torch_gravity.gravity(a, b, c, P, ctx, nkernel=nk,
                                                   precision=SG_PREC_FP64)
        assert torch.autograd.gradcheck(fn, (X, M, Hh), eps=1e-6, atol=1e-6, rtol=1e-5)


def test_device_entry_point_matches_host():
    d = ics.adjoint_set(20_000, seed=9)
    g = Gravity(0)
    try:
        ins = [d[k] for k in ("x", "y", "z", "m", "range2")]
        cot = [np.ascontiguousarray(d["w"][:, c]) for c in range(3)] + [d["cot_u"]]
        for prec in (SG_PREC_FP32C, SG_PREC_FP64):
            host = g.vjp(*ins, *cot, precision=prec)
            dev_in = [DeviceArray.from_host(a) for a in ins + cot]
            dev_out = [DeviceArray(len(d["x"])) for _ in range(5)]
            g.vjp_device(len(d["x"]), *dev_in, *dev_out, precision=prec)
            g.synchronize()
            for a, b in zip(host, (o.download() for o in dev_out)):
                np.testing.assert_array_equal(a, b)
            # deterministic: a second call gives the same bits
            again = g.vjp(*ins, *cot, precision=prec)
            for a, b in zip(host, again):
                np.testing.assert_array_equal(a, b)
    finally:
        g.close()


def test_argument_errors():
    g = Gravity(0)
    try:
        d = ics.adjoint_set(1000, seed=2)
        ins = [d[k] for k in ("x", "y", "z", "m", "range2")]
        cot = [np.ascontiguousarray(d["w"][:, c]) for c in range(3)] + [d["cot_u"].copy()]
        with pytest.raises(_lib.SphgravError) as e:
            g.vjp(*ins, *cot, nkernel=3)
        assert e.value.status == _lib.SG_ERR_NKERNEL
        cot[3][17] = np.nan
        with pytest.raises(_lib.SphgravError) as e:
            g.vjp(*ins, *cot)
        assert e.value.status == _lib.SG_ERR_NONFINITE
        L = _lib.load()
        z = C.c_void_p(0)
        assert L.sg_vjp_device(g._h, 0, *[z] * 5, 0, 0, *[z] * 9) == _lib.SG_ERR_ARG
    finally:
        g.close()
