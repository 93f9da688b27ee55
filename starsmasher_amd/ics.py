"""Seeded synthetic particle sets for the gravity workloads.

The reference builds stars on an hcp lattice (pbe/src/initialize_polyes.f:66-147)
and sizes hp from a neighbour count (:281-282).  These generators produce the
five workloads the library is measured and tested on:

* :func:`plummer`          -- Plummer sphere, a = 1, truncated at r = 10a, m = 1/N
* :func:`two_polytropes`   -- two n = 1.5 polytropes on jittered hcp lattices
* :func:`giant_with_core`  -- n = 3 polytrope core + sparse envelope + point mass
* :func:`batch_set`        -- many independent stars of uneven N
* :func:`adjoint_set`      -- Plummer sphere plus random cotangents

Every set is a dict of float64 arrays ``x, y, z, m, hp, u`` (``u == 0`` marks a
point particle, as in grav.f:17) and ``range2`` = +-4 hp^2 (grav.f:15-18).
``hp`` is half the distance to the 64th nearest neighbour, i.e. each particle
has 64 neighbours within 2 hp (scipy cKDTree).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

NNEIGH = 64


# ----------------------------------------------------------------- helpers
def range2_of(hp, u):
    """grav.f:15-18: range = 4 hp^2, negated for point particles (u == 0)."""
    r2 = 4.0 * np.asarray(hp, dtype=np.float64) ** 2
    return np.where(np.asarray(u) == 0.0, -r2, r2)


def knn_hp(pos, k=NNEIGH):
    """hp such that each particle has k neighbours within 2 hp."""
    pos = np.asarray(pos, dtype=np.float64)
    kk = min(k, pos.shape[0] - 1)
    if kk < 1:
        return np.ones(pos.shape[0])
    d, _ = cKDTree(pos).query(pos, k=kk + 1, workers=-1)
    return 0.5 * d[:, kk]


def _pack(pos, m, hp, u, **extra):
    pos = np.asarray(pos, dtype=np.float64)
    out = {
        "x": np.ascontiguousarray(pos[:, 0]), "y": np.ascontiguousarray(pos[:, 1]),
        "z": np.ascontiguousarray(pos[:, 2]), "m": np.asarray(m, dtype=np.float64),
        "hp": np.asarray(hp, dtype=np.float64), "u": np.asarray(u, dtype=np.float64),
    }
    out["range2"] = range2_of(out["hp"], out["u"])
    out.update(extra)
    return out


def _isotropic(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def lane_emden(n, dxi=1e-4):
    """Solve theta'' + 2 theta'/xi + theta^n = 0 by RK4.  Returns (xi, theta,
    mass) up to the first zero xi1, mass(xi) = -xi^2 theta'(xi)."""
    xi0 = 1e-6
    y = np.array([1.0 - xi0 ** 2 / 6.0, -xi0 / 3.0])

    def f(xi, y):
        th = max(y[0], 0.0)
        return np.array([y[1], -th ** n - 2.0 * y[1] / xi])

    xs, ths, ms = [xi0], [y[0]], [-xi0 ** 2 * y[1]]
    xi = xi0
    while y[0] > 0.0:
        k1 = f(xi, y)
        k2 = f(xi + dxi / 2, y + dxi / 2 * k1)
        k3 = f(xi + dxi / 2, y + dxi / 2 * k2)
        k4 = f(xi + dxi, y + dxi * k3)
        ynew = y + dxi / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        if ynew[0] <= 0.0:     # linear step to the surface
            frac = y[0] / (y[0] - ynew[0])
            xi += frac * dxi
            y = y + frac * (ynew - y)
            y[0] = 0.0
        else:
            xi += dxi
            y = ynew
        xs.append(xi)
        ths.append(max(y[0], 0.0))
        ms.append(-xi ** 2 * y[1])
    return np.array(xs), np.array(ths), np.array(ms)


_LE_CACHE: dict = {}


def _lane_emden_cached(n):
    if n not in _LE_CACHE:
        _LE_CACHE[n] = lane_emden(n, dxi=1e-3)
    return _LE_CACHE[n]


def hcp_lattice(n_target, rng, jitter=0.05):
    """~n_target points of a hexagonal close-packed lattice filling the unit
    sphere, each coordinate jittered uniformly by +-jitter * spacing.  Returns
    exactly n_target points (the innermost ones)."""
    # hcp packing fraction: one point per a^3/sqrt(2) volume
    vol = 4.0 / 3.0 * np.pi
    a = (vol * np.sqrt(0.5) / n_target) ** (1.0 / 3.0) * 0.98
    dz = a * np.sqrt(2.0 / 3.0)
    dy = a * np.sqrt(3.0) / 2.0
    nx = int(np.ceil(1.0 / a)) + 2
    ny = int(np.ceil(1.0 / dy)) + 2
    nz = int(np.ceil(1.0 / dz)) + 2
    ii, jj, kk = np.meshgrid(np.arange(-nx, nx + 1), np.arange(-ny, ny + 1),
                             np.arange(-nz, nz + 1), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    x = a * (ii + 0.5 * (jj % 2) + 0.5 * (kk % 2))
    y = dy * (jj + (kk % 2) / 3.0)
    z = dz * kk
    pos = np.stack([x, y, z], axis=1)
    r = np.linalg.norm(pos, axis=1)
    keep = np.argsort(r, kind="stable")[:n_target]
    if keep.size < n_target:
        raise ValueError("lattice too small")
    pos = pos[keep]
    pos += rng.uniform(-jitter * a, jitter * a, size=pos.shape)
    return pos, a


def _stretch_to_profile(pos, n_poly, radius):
    """Map a uniform-density ball onto an equal-mass polytrope of index
    n_poly: the particle with radial rank k moves to the radius enclosing a
    fraction (k + 0.5)/N of the mass, keeping its direction."""
    xi, _, mass = _lane_emden_cached(n_poly)
    r = np.linalg.norm(pos, axis=1)
    order = np.argsort(r, kind="stable")
    frac = (np.arange(pos.shape[0]) + 0.5) / pos.shape[0]
    target = np.interp(frac, mass / mass[-1], xi) / xi[-1] * radius
    out = np.empty_like(pos)
    dirs = pos[order] / np.maximum(r[order], 1e-300)[:, None]
    out[order] = dirs * target[:, None]
    return out


# -------------------------------------------------------------- workloads
def plummer(n, seed=1, a=1.0, rmax=10.0, mass=1.0):
    """Plummer sphere of scale a truncated at rmax; equal masses mass/n."""
    rng = np.random.default_rng(seed)
    fmax = (rmax / a) ** 3 / (1.0 + (rmax / a) ** 2) ** 1.5   # M(<rmax) / M
    f = rng.uniform(0.0, fmax, n)
    r = a / np.sqrt(f ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _isotropic(rng, n)
    m = np.full(n, mass / n)
    hp = knn_hp(pos)
    return _pack(pos, m, hp, np.ones(n))


def polytrope(n, n_poly=1.5, radius=1.0, mass=1.0, seed=1, jitter=0.05, center=(0, 0, 0)):
    """Equal-mass polytrope on a jittered hcp lattice radially stretched to
    the Lane-Emden mass profile."""
    rng = np.random.default_rng(seed)
    pos, _ = hcp_lattice(n, rng, jitter)
    pos = _stretch_to_profile(pos, n_poly, radius) + np.asarray(center, dtype=np.float64)
    return pos, np.full(n, mass / n)


def two_polytropes(n=262_144, seed=2, mass_ratio=0.6, radii=(1.0, 0.8), separation=3.0,
                   vrel=0.5):
    """Two n = 1.5 polytropes, M2/M1 = mass_ratio, equal particle masses, the
    centres `separation` apart along x, star 2 approaching star 1 at vrel."""
    n1 = int(round(n / (1.0 + mass_ratio)))
    n2 = n - n1
    m_tot = 1.0
    m1 = m_tot / (1.0 + mass_ratio)
    m2 = m_tot - m1
    p1, mm1 = polytrope(n1, 1.5, radii[0], m1, seed=seed, center=(-separation * m2 / m_tot, 0, 0))
    p2, mm2 = polytrope(n2, 1.5, radii[1], m2, seed=seed + 1000,
                        center=(separation * m1 / m_tot, 0, 0))
    pos = np.concatenate([p1, p2])
    m = np.concatenate([mm1, mm2])
    hp = knn_hp(pos)
    vel = np.zeros_like(pos)
    vel[:n1, 0] = vrel * m2 / m_tot
    vel[n1:, 0] = -vrel * m1 / m_tot
    return _pack(pos, m, hp, np.ones(n), vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2])


def giant_with_core(n=1_048_576, seed=3, core_frac=0.6, r_env=300.0, point_mass_frac=0.3,
                    point_hp=0.1):
    """Giant-star model: an n = 3 polytrope (radius 1) sampled at equal
    number density with masses proportional to the local density, an
    extended envelope out to r_env whose number density falls as r^-2.5
    (masses ~ r^-0.5), and one point particle (u = 0) of point_mass_frac of
    the total mass at the centre with hp = point_hp (range2 = -4 hp^2).
    Masses span ~4 decades and hp ~3 decades."""
    rng = np.random.default_rng(seed)
    ng = n - 1
    nc = int(core_frac * ng)
    ne = ng - nc
    # core: uniform positions in the unit ball, m ~ rho_n3(r)
    pos_c, _ = hcp_lattice(nc, rng, 0.05)
    xi, th, mass = _lane_emden_cached(3.0)
    rc = np.linalg.norm(pos_c, axis=1)
    rho = np.interp(rc * xi[-1], xi, th) ** 3
    rho = np.maximum(rho, 1e-3)
    m_core = rho
    # envelope: n(r) ~ r^-2.5 between 1 and r_env -> p(r) ~ r^-0.5
    u = rng.uniform(0, 1, ne)
    re = (1.0 + u * (r_env ** 0.5 - 1.0)) ** 2
    pos_e = re[:, None] * _isotropic(rng, ne)
    m_env = 1e-3 * re ** -0.5
    m_gas = np.concatenate([m_core, m_env])
    m_gas *= (1.0 - point_mass_frac) / m_gas.sum()
    pos = np.concatenate([pos_c, pos_e, np.zeros((1, 3))])
    m = np.concatenate([m_gas, [point_mass_frac]])
    hp = knn_hp(pos[:-1])
    hp = np.concatenate([hp, [point_hp]])
    uu = np.ones(n)
    uu[-1] = 0.0
    return _pack(pos, m, hp, uu)


def batch_sizes(nstars=64, nmin=2048, nmax=131072, seed=4):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(nmin), np.log(nmax), nstars)).astype(np.int64)


def batch_set(nstars=64, nmin=2048, nmax=131072, seed=4):
    """Independent stars of log-uniform N; even entries Plummer spheres, odd
    entries n = 1.5 polytropes.  Returns a list of particle dicts."""
    sizes = batch_sizes(nstars, nmin, nmax, seed)
    stars = []
    for s, ns in enumerate(sizes):
        if s % 2 == 0:
            stars.append(plummer(int(ns), seed=100 + s))
        else:
            pos, m = polytrope(int(ns), 1.5, 1.0, 1.0, seed=100 + s)
            stars.append(_pack(pos, m, knn_hp(pos), np.ones(int(ns))))
    return stars


def adjoint_set(n=131_072, seed=5):
    """Plummer sphere plus FP32 cotangents w (unit Gaussian 3-vectors) and u
    (unit Gaussian scalars) for the loss sum_i w_i.a_i + u_i phi_i."""
    d = plummer(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    d["w"] = rng.normal(size=(n, 3)).astype(np.float32).astype(np.float64)
    d["cot_u"] = rng.normal(size=n).astype(np.float32).astype(np.float64)
    return d
