"""Pinning the oracle (CPU).  Each check is chosen so that a plausible slip in
the restatement -- a dropped term, a wrong sign or index, a swapped H_i/H_j,
a wrong coefficient -- fails one of them:

* golden potentials stored by the reference's own gravity path in its example
  dumps (cubic kernel) -- and the Wendland kernels must *miss* them;
* the reference's own Fortran CPU twin (outputs committed as fixtures, and the
  live build in oracle/_ref when present);
* the kernel polynomials against the potential of the smoothing density,
  integrated by quadrature (independent of the decimal coefficients);
* a = -grad(phi) by central differences, momentum conservation, the self
  term, and an independent numpy restatement on tiny inputs;
* the adjoint against central differences and kernels.f's dphi/dh forms;
* the rank split against a hand-worked case of main.f:297-349.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle


def _rng(seed=0):
    return np.random.default_rng(seed)


def _cloud(n, seed=0, npoint=0, hscale=(0.2, 0.6)):
    r = _rng(seed)
    pos = r.normal(size=(n, 3))
    m = r.uniform(0.5, 2.0, n) / n
    hp = r.uniform(*hscale, n)
    u = np.ones(n)
    if npoint:
        u[-npoint:] = 0.0
    r2 = 4 * hp ** 2
    r2 = np.where(u == 0, -r2, r2)
    return pos, m, hp, u, r2


# ------------------------------------------------------------ golden dumps
@pytest.mark.parametrize("name", ["collision_start2u", "3stars_start1u", "3stars_start2u",
                                  "collision_start1u"])
def test_dump_potentials_cubic(golden_dumps, name):
    d = golden_dumps[name]
    _, _, _, pot = oracle.direct(d["x"], d["y"], d["z"], d["am"], d["range2"], nkernel=0)
    rel = np.abs(pot - d["grpot"]) / np.abs(d["grpot"])
    # the dumps hold the reference's FP32-kernel potentials (some rounded to FP32)
    assert rel.max() < 4e-7, rel.max()
    assert np.median(rel) < 1e-7


def test_dump_potentials_reject_wrong_kernel(golden_dumps):
    d = golden_dumps["collision_start2u"]
    for nk in (1, 2):
        _, _, _, pot = oracle.direct(d["x"], d["y"], d["z"], d["am"], d["range2"], nkernel=nk)
        rel = np.abs(pot - d["grpot"]) / np.abs(d["grpot"])
        assert rel.max() > 1e-4


# ------------------------------------------------------ reference CPU twin
def test_twin_matches_reference_fortran_fixtures(twin_vectors):
    for name, v in twin_vectors.items():
        nk, lo, hi, nsg = (int(a) for a in v["args"])
        res = oracle.cpu_twin(v["x"], v["y"], v["z"], v["am"], v["hp"], v["u"], nkernel=nk,
                              ngrav_lower=lo, ngrav_upper=hi, nselfgravity=nsg)
        for got, key in zip(res, ("gx", "gy", "gz", "grpot")):
            np.testing.assert_array_equal(got, v[key], err_msg=f"{name}/{key}")


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs amdflang + reference)")
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_twin_matches_reference_fortran_live(nk):
    pos, m, hp, u, _ = _cloud(150, seed=10 + nk, npoint=2)
    pos[7] = pos[3]
    ref = oracle.ref_cpu_gravity(pos[:, 0], pos[:, 1], pos[:, 2], m, hp, u, nkernel=nk)
    got = oracle.cpu_twin(pos[:, 0], pos[:, 1], pos[:, 2], m, hp, u, nkernel=nk)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_direct_equals_symmetric_twin(nk):
    """The library's full i x j semantics and the CPU twin's symmetric loop
    describe the same sums (no coincident distinct pairs here)."""
    pos, m, hp, u, r2 = _cloud(300, seed=20 + nk, npoint=3)
    ax, ay, az, pot = oracle.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, nkernel=nk)
    gx, gy, gz, gp = oracle.cpu_twin(pos[:, 0], pos[:, 1], pos[:, 2], m, hp, u, nkernel=nk)
    a = np.stack([ax, ay, az], 1)
    g = np.stack([gx, gy, gz], 1)
    assert np.max(np.linalg.norm(a - g, axis=1) / np.linalg.norm(g, axis=1)) < 1e-12
    assert np.max(np.abs(pot - gp) / np.abs(gp)) < 1e-13


def test_parallel_twin_equals_serial():
    pos, m, hp, u, _ = _cloud(400, seed=3, npoint=1)
    ser = oracle.cpu_twin(pos[:, 0], pos[:, 1], pos[:, 2], m, hp, u)
    par = oracle.cpu_twin_parallel(pos[:, 0], pos[:, 1], pos[:, 2], m, hp, u, nthreads=5)
    for a, b in zip(par, ser):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15 * np.abs(b).max())


# ------------------------------------------------------ kernel polynomials
def _density_shape(nk):
    """Smoothing density of each kernel on q = r/H in [0, 1] (unnormalised):
    cubic M4 spline (Monaghan & Lattanzio 1985) and the Wendland functions
    W3,3 (kernels.f:283, commented w33) and C4 (Wendland 1995, 3-D psi_{3,2})."""
    if nk == 0:
        return lambda q: np.where(q < 0.5, 1 - 6 * q ** 2 + 6 * q ** 3, 2 * (1 - q) ** 3)
    if nk == 1:
        return lambda q: (1 - q) ** 8 * (32 * q ** 3 + 25 * q ** 2 + 8 * q + 1)
    return lambda q: (1 - q) ** 6 * (35 / 3 * q ** 2 + 6 * q + 1)


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_polynomials_are_the_potential_of_the_kernel_density(nk):
    """A(q) = M(<q)/q^3 and P(q) = -M(<q)/q - 4 pi int_q^1 rho q' dq' for the
    unit-mass density of the kernel, integrated by quadrature."""
    f = _density_shape(nk)
    norm = integrate.quad(lambda q: 4 * np.pi * q * q * f(q), 0, 1, points=[0.5], epsabs=0,
                          epsrel=1e-13)[0]
    for q in (0.05, 0.2, 0.37, 0.5, 0.63, 0.8, 0.95, 1.0):
        M = integrate.quad(lambda s: 4 * np.pi * s * s * f(s), 0, q, epsabs=0, epsrel=1e-13)[0] / norm
        tail = integrate.quad(lambda s: 4 * np.pi * s * f(s), q, 1, epsabs=0, epsrel=1e-13)[0] / norm
        A, P = oracle.kernel_AP(nk, q)
        assert abs(A - M / q ** 3) < 2e-10 * max(1, abs(A)), (nk, q, A, M / q ** 3)
        assert abs(P - (-M / q - tail)) < 2e-10 * max(1, abs(P)), (nk, q, P, -M / q - tail)


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_polynomial_derivatives(nk):
    """dP/dq = q A(q) (so that a = -grad phi) and the coded dA/dq, dP/dq agree
    with central differences of A and P."""
    for q in np.linspace(0.03, 0.99, 23):
        A, _ = oracle.kernel_AP(nk, q)
        dA, dP = oracle.kernel_dAP(nk, q)
        e = 1e-6
        Ap, Pp = oracle.kernel_AP(nk, q + e)
        Am, Pm = oracle.kernel_AP(nk, q - e)
        assert abs(dP - q * A) < 1e-9 * max(1.0, abs(q * A))
        assert abs(dA - (Ap - Am) / (2 * e)) < 1e-5 * max(1.0, abs(dA))
        assert abs(dP - (Pp - Pm) / (2 * e)) < 1e-5 * max(1.0, abs(dP))


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_continuity_at_support_radius(nk):
    A, P = oracle.kernel_AP(nk, 1.0)
    assert abs(A - 1.0) < 1e-10 and abs(P + 1.0) < 1e-10


# --------------------------------------------------------------- invariants
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_acceleration_is_minus_gradient_of_potential(nk):
    pos, m, hp, u, r2 = _cloud(120, seed=5, npoint=2, hscale=(0.4, 1.0))
    x, y, z = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
    rows = np.array([0, 17, 60, 119])
    ax, ay, az, _ = oracle.direct(x, y, z, m, r2, nkernel=nk, rows=rows)
    acc = np.stack([ax, ay, az], 1)
    eps = 1e-6
    for r, i in enumerate(rows):
        for c, arr in enumerate((x, y, z)):
            keep = arr[i]
            arr[i] = keep + eps
            pp = oracle.direct(x, y, z, m, r2, nkernel=nk, rows=[i])[3][0]
            arr[i] = keep - eps
            pm = oracle.direct(x, y, z, m, r2, nkernel=nk, rows=[i])[3][0]
            arr[i] = keep
            g = -(pp - pm) / (2 * eps)
            assert abs(g - acc[r, c]) < 1e-6 * np.linalg.norm(acc[r]), (i, c, g, acc[r, c])


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_momentum_conservation(nk):
    pos, m, hp, u, r2 = _cloud(500, seed=6, npoint=3)
    ax, ay, az, _ = oracle.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, nkernel=nk)
    a = np.stack([ax, ay, az], 1)
    tot = (m[:, None] * a).sum(0)
    assert np.linalg.norm(tot) < 1e-13 * (m * np.linalg.norm(a, axis=1)).sum()


@pytest.mark.parametrize("nk,c", [(0, 1.4), (1, 1.9140625), (2, 1.71875)])
def test_self_term(nk, c):
    x = np.array([0.3]); m = np.array([2.5]); H = 0.7
    for r2, want in ((H * H, -2 * c * 2.5 / H), (-H * H, 0.0)):
        ax, ay, az, pot = oracle.direct(x, x, x, m, np.array([r2]), nkernel=nk)
        assert ax[0] == 0.0 and abs(pot[0] - want) < 1e-14


def test_coincident_pair_rules():
    """Distinct coincident pair: finite r->0 limit (library rule) vs the
    reference kernel's NaN (grav_force_direct.cu:110-112)."""
    x = np.array([0.0, 0.0]); m = np.array([1.0, 2.0]); r2 = np.array([1.0, 1.0])
    _, _, _, pot = oracle.direct(x, x, x, m, r2, nkernel=0, self_rule=oracle.SELF_LIMIT)
    assert abs(pot[0] - (-1.4 * (1.0 * 2 + 2.0 * 2))) < 1e-14   # self + limit of the pair
    _, _, _, potg = oracle.direct(x, x, x, m, r2, nkernel=0, self_rule=oracle.SELF_GPU)
    assert np.isnan(potg).all()


def test_numpy_restatement_agrees():
    pos, m, hp, u, r2 = _cloud(60, seed=9, npoint=2, hscale=(0.5, 1.5))
    H = np.sqrt(np.abs(r2))
    acc, pot = oracle.direct_numpy(pos, m, H, (r2 > 0).astype(float))
    ax, ay, az, p = oracle.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, nkernel=0)
    a = np.stack([ax, ay, az], 1)
    # 12-digit decimal literals vs exact fractions: ~1e-12 relative
    assert np.max(np.linalg.norm(a - acc, axis=1) / np.linalg.norm(acc, axis=1)) < 1e-11
    assert np.max(np.abs(p - pot) / np.abs(pot)) < 1e-11


def test_newtonian_limit():
    r = _rng(11)
    pos = r.uniform(-10, 10, size=(40, 3))
    m = r.uniform(1, 2, 40)
    r2 = np.full(40, -1e-8)    # point particles: no self potential
    ax, ay, az, pot = oracle.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, r2)
    for i in (0, 13, 39):
        d = pos - pos[i]
        rr = np.linalg.norm(d, axis=1)
        rr[i] = np.inf
        want = (m[:, None] * d / rr[:, None] ** 3).sum(0)
        np.testing.assert_allclose([ax[i], ay[i], az[i]], want, rtol=1e-13)
        np.testing.assert_allclose(pot[i], -(m / rr).sum(), rtol=1e-13)


# ------------------------------------------------------------------ adjoint
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_vjp_matches_finite_differences(nk):
    n = 48
    pos, m, hp, u, r2 = _cloud(n, seed=30 + nk, npoint=2, hscale=(0.4, 1.2))
    r = _rng(40 + nk)
    w = r.normal(size=(n, 3))
    cu = r.normal(size=n)
    x, y, z = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
    gx, gy, gz, gm, gH = oracle.vjp(x, y, z, m, r2, w, cu, nkernel=nk)
    eps = 1e-6
    for k in (0, 11, 30, n - 1):
        def L():
            return oracle.loss_particle(k, x, y, z, m, r2, w, cu, nkernel=nk)
        for arr, g in ((x, gx), (y, gy), (z, gz), (m, gm)):
            keep = arr[k]
            h = eps * max(1.0, abs(keep))
            arr[k] = keep + h; lp = L()
            arr[k] = keep - h; lm = L()
            arr[k] = keep
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])
        # H derivative: perturb |range2| = H^2 keeping the sign
        H = np.sqrt(abs(r2[k])); sgn = np.sign(r2[k])
        keep = r2[k]
        r2[k] = sgn * (H + eps) ** 2; lp = L()
        r2[k] = sgn * (H - eps) ** 2; lm = L()
        r2[k] = keep
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - gH[k]) < 1e-6 * max(1.0, abs(gH[k])), (k, fd, gH[k])


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_vjp_coincident_pair_finite_differences(nk):
    """Two distinct particles at the same position: the acceleration of each
    is m F(r) dr with F finite at r = 0, so the x-gradient keeps F(0) terms."""
    n = 24
    pos, m, hp, u, r2 = _cloud(n, seed=50 + nk, npoint=1, hscale=(0.4, 1.2))
    pos[5] = pos[4]
    r = _rng(60 + nk)
    w = r.normal(size=(n, 3))
    cu = r.normal(size=n)
    x, y, z = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
    gx, gy, gz, gm, gH = oracle.vjp(x, y, z, m, r2, w, cu, nkernel=nk)
    eps = 1e-6
    for k in (4, 5):
        for arr, g in ((x, gx), (y, gy), (z, gz), (m, gm)):
            keep = arr[k]
            h = eps * max(1.0, abs(keep))
            arr[k] = keep + h; lp = oracle.loss_particle(k, x, y, z, m, r2, w, cu, nkernel=nk)
            arr[k] = keep - h; lm = oracle.loss_particle(k, x, y, z, m, r2, w, cu, nkernel=nk)
            arr[k] = keep
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_fd_grad_agrees_with_vjp():
    """The threaded central-difference helper used by the GPU adjoint tests
    reproduces the adjoint on a small set (and so sums the same loss)."""
    n = 200
    pos, m, hp, u, r2 = _cloud(n, seed=70, npoint=3, hscale=(0.3, 0.9))
    r = _rng(71)
    w = r.normal(size=(n, 3))
    cu = r.normal(size=n)
    g = np.stack(oracle.vjp(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, w, cu), 1)
    ks = np.array([0, 3, 50, 199])
    fd = oracle.fd_grad(ks, pos[:, 0], pos[:, 1], pos[:, 2], m, r2, w, cu, step=1e-6)
    assert np.abs(fd - g[ks]).max() < 1e-7 * np.abs(g).max()


def _dphidh_reference(nk, u):
    """kernels.f:215-279: dphidh (cubic), dphi332hdh, dphic42hdh, u = r/hp."""
    if nk == 0:
        if u <= 1.0:
            return -2 * u ** 2 + 1.5 * u ** 4 - 0.6 * u ** 5 + 1.4
        return (0.1 + 0.2 * u) * (2.0 - u) ** 4 if u < 2.0 else 0.0
    q = u / 2.0
    if nk == 1:
        return 35 * (1 - q) ** 9 * (7 + 63 * q + 237 * q ** 2 + 453 * q ** 3 + 384 * q ** 4) / 128
    return 55 * (1 - q) ** 7 * (1 + 7 * q + 19 * q ** 2 + 21 * q ** 3) / 32


@pytest.mark.parametrize("nk", [0, 1, 2])
def test_h_derivative_matches_kernels_f(nk):
    """kernels.f's dphidh(u) = -(hp^2/m) d(phi_one_sided)/d(hp): with
    phi = m P(q)/H, H = 2 hp, q = u/2 that is -(P(q) + q P'(q))/2."""
    for u in np.linspace(0.0, 1.98, 34):
        q = u / 2
        _, P = oracle.kernel_AP(nk, q)
        _, dP = oracle.kernel_dAP(nk, q)
        ours = -(P + q * dP) / 2
        assert abs(ours - _dphidh_reference(nk, u)) < 2e-9, (nk, u, ours, _dphidh_reference(nk, u))


# ---------------------------------------------------------------- partition
def test_partition_hand_worked():
    lo, hi = oracle.partition(10, 3)
    assert lo.tolist() == [1, 3, 5] and hi.tolist() == [2, 4, 10]


def test_partition_reference_failure_modes():
    with pytest.raises(ValueError):
        oracle.partition(10, 1)       # main.f:341-344 stops for a single rank


@pytest.mark.parametrize("n,nr", [(1000, 2), (1000, 8), (4097, 16)])
def test_partition_invariants(n, nr):
    lo, hi = oracle.partition(n, nr)
    assert lo[0] == 1 and hi[-1] == n
    assert np.all(lo[1:] == hi[:-1] + 1) and np.all(hi >= lo)
    work = np.array([sum(n - i + 1 for i in range(a, b + 1)) for a, b in zip(lo, hi)])
    assert work.sum() == n * (n + 1) // 2
    assert work[:-1].max() / work[:-1].min() < 1.2


# ------------------------------------------------------- grad-h softening psi
@pytest.mark.parametrize("nk", [0, 1, 2])
def test_psi_matches_softened_potential(nk):
    """psi_i = hp^-2 sum_j m_j dphidh(r/hp) (advance.f90:716-737) against the
    same sum formed from the potential polynomial: dphidh(u) =
    -(P(q) + q^2 A(q))/2 with q = u/2, P' = q A (kernels.f:215-279)."""
    n = 150
    pos, m, hp, u, r2 = _cloud(n, seed=80 + nk, npoint=4, hscale=(0.4, 1.2))
    got = oracle.psi(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, nkernel=nk)
    for i in (0, 17, 90, n - 5):
        d = np.linalg.norm(pos - pos[i], axis=1)
        H = 2 * hp[i]
        want = 0.0
        for j in np.nonzero(d < H)[0]:
            q = d[j] / H
            A, P = oracle.kernel_AP(nk, q)
            want += m[j] * (-(P + q * q * A) / 2)
        want /= hp[i] ** 2
        assert abs(got[i] - want) <= 1e-10 * abs(want), (i, got[i], want)
    assert np.all(got[-4:] == 0.0)          # point particles carry no psi


def test_psi_self_term_and_table():
    # an isolated gas particle: psi = m dphidh(0) / hp^2 = m c / hp^2
    x = np.array([0.0, 50.0]); m = np.array([2.0, 1.0]); r2 = np.array([4 * 0.3 ** 2, 4.0])
    for nk, c in ((0, 1.4), (1, 1.9140625), (2, 1.71875)):
        p = oracle.psi(x, x, x, m, r2, nkernel=nk)
        assert abs(p[0] - 2.0 * c / 0.09) < 1e-12 * p[0]
    # the 10^7-entry table of the reference differs from the closed form by
    # less than its spacing in u^2 (4e-7) times the slope
    pos, m, hp, u, r2 = _cloud(200, seed=91, hscale=(0.4, 1.2))
    a = oracle.psi(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, tabulated=False)
    b = oracle.psi(pos[:, 0], pos[:, 1], pos[:, 2], m, r2, tabulated=True)
    assert np.max(np.abs(a - b) / np.abs(a)) < 1e-6


def test_epot_matches_correctly_rounded_sum():
    """sgo_epot (output.f:232-252: epot = 1/2 sum m phi) against math.fsum of
    the same products; a dropped factor 1/2 or a wrong array fails."""
    import math
    rng = np.random.default_rng(21)
    m = rng.uniform(1e-6, 1.0, 50_001)
    pot = -rng.uniform(0.1, 10.0, 50_001)
    want = 0.5 * math.fsum((m * pot).tolist())
    got = oracle.epot(m, pot)
    assert abs(got - want) <= 1e-12 * abs(want)
    assert oracle.epot(m[:1], pot[:1]) == 0.5 * m[0] * pot[0]
