"""The oracle's UAngle (Src/energy.F90:5409-5481) pinned by what the
mathematics fixes: the closed form u = k/p (theta - eq)^p with theta and eq in
rad and k the namelist's value per degree^2 (IOPotChain's scaling,
Src/potential.F90:3558-3562, divides by (pi/180)^2 whatever p is), forces
equal to minus central differences of the energy, zero net force,
invariance under translation across the periodic boundary, and its
place in UTotal."""
import numpy as np
import pytest

import oracle as O

L = 40.0


def _three_beads(theta_deg, r1=4.0, r2=5.5):
    t = np.deg2rad(theta_deg)
    c = np.array([1.0, -2.0, 0.5])
    return np.array([c + [r1, 0.0, 0.0], c, c + [r2 * np.cos(t), r2 * np.sin(t), 0.0]])


@pytest.mark.parametrize("k,p,eq,theta", [(0.05, 2, 120.0, 100.0), (0.002, 4, 109.5, 150.0),
                                          (0.3, 1, 90.0, 135.0), (0.01, 3, 180.0, 60.0)])
def test_closed_form(k, p, eq, theta):
    ro = _three_beads(theta)
    f, u = O.uangle(L, ro, [(1, 3)], k, p, eq)
    # IOPotChain divides k by (pi/180)^2 whatever p is
    d = theta - eq
    expect = k / p * d ** 2 * np.deg2rad(d) ** (p - 2)
    assert abs(u - expect) <= 1e-11 * max(abs(expect), 1.0)
    # the three forces cancel and the centre's balances the ends
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).max()
    # forces lie in the plane of the angle
    assert np.abs(f[:, 2]).max() <= 1e-12 * np.abs(f).max()


def _chains_system(seed=7):
    rng = np.random.default_rng(seed)
    chains, ro = [], []
    start = 1
    for n in (1, 2, 3, 7, 12):
        x = rng.uniform(-L / 2, L / 2, 3)
        pts = [x]
        for _ in range(n - 1):
            v = rng.normal(size=3)
            pts.append(pts[-1] + 4.5 * v / np.linalg.norm(v))
        chains.append((start, n))
        start += n
        ro.extend(pts)
    ro = np.array(ro)
    ro = ro - L * np.round(ro / L)  # inside [-L/2, L/2]
    return ro, chains


@pytest.mark.parametrize("p", [2, 3])
def test_forces_are_minus_gradient(p):
    ro, chains = _chains_system()
    k, eq = 0.04, 112.0
    f, u0 = O.uangle(L, ro, chains, k, p, eq)
    h = 1e-6
    g = np.zeros_like(ro)
    for i in range(ro.shape[0]):
        for m in range(3):
            a = ro.copy()
            b = ro.copy()
            a[i, m] += h
            b[i, m] -= h
            g[i, m] = (O.uangle(L, a, chains, k, p, eq)[1] - O.uangle(L, b, chains, k, p, eq)[1]) / (2 * h)
    scale = np.abs(f).max()
    assert scale > 0
    assert np.abs(f + g).max() <= 1e-6 * scale
    # the angle forces of each chain sum to zero
    assert np.abs(f.sum(axis=0)).max() <= 1e-11 * scale


def test_translation_across_the_boundary():
    ro, chains = _chains_system(seed=3)
    k, p, eq = 0.05, 2, 100.0
    f0, u0 = O.uangle(L, ro, chains, k, p, eq)
    shifted = ro + np.array([17.3, -9.1, 25.0])
    shifted = shifted - L * np.round(shifted / L)
    f1, u1 = O.uangle(L, shifted, chains, k, p, eq)
    assert abs(u1 - u0) <= 1e-11 * abs(u0)
    assert np.abs(f1 - f0).max() <= 1e-10 * np.abs(f0).max()


def test_zero_k_and_short_chains_give_nothing():
    ro, chains = _chains_system()
    f, u = O.uangle(L, ro, chains, 0.0, 2, 100.0)
    assert u == 0.0 and not f.any()
    f, u = O.uangle(L, ro[:3], [(1, 1), (2, 2)], 0.05, 2, 100.0)
    assert u == 0.0 and not f.any()


def test_utotal_adds_the_angle_term():
    ro, chains = _chains_system()
    n = ro.shape[0]
    iptpn = np.ones(n, np.int32)
    az = np.zeros(n)
    from molsim_amd import engine as E
    tab = E.build_tables([0.0], [1.0], [0.0], rcut=10.0, epsi_four_pi=1.0, alpha=0.0)
    t = O.Tables(tab.ubuf, tab.iubuflow, tab.nugrid, tab.r2umin, None)
    kw = dict(ewald="none", nlist="vlist", chains=chains, bond=(0.4, 2, 5.0),
              maxnneigh=n - 1)
    base = O.Oracle(L, iptpn, az, t, 10.0, **kw)
    base.set_list(ro)
    fb, eb = base.utotal(ro)
    wa = O.Oracle(L, iptpn, az, t, 10.0, angle=(0.05, 2, 100.0), **kw)
    wa.set_list(ro)
    fa, ea = wa.utotal(ro)
    fang, uang = O.uangle(L, ro, chains, 0.05, 2, 100.0)
    assert eb["angle"] == 0.0
    assert ea["angle"] == uang
    assert abs(ea["tot"] - (eb["tot"] + uang)) <= 1e-12 * abs(ea["tot"])
    assert ea["virial"] == eb["virial"]
    fac = np.abs(fa - fb).max() / np.abs(fang).max()   # the force unit factor
    assert np.allclose(fa - fb, fang * fac, rtol=1e-9, atol=1e-12 * np.abs(fa).max())
