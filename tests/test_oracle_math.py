"""SPME parts of the oracle, pinned by mathematics instead of the reference.

The reference's SPME needs FFTW3, which is absent here, so its outputs
cannot be recorded.  These checks fix what the oracle computes through
independent means: numpy.fft for the FFT, closed forms of the B-splines and
of Essmann et al. (JCP 103, 8577 (1995)) eqs. 3.9/4.4/4.8 for the influence
function, finite differences of the SPME energy for the gather, the
direct (standard) Ewald sum, and the NaCl Madelung constant.
"""
import math

import numpy as np
import pytest

import oracle as O

EFP = None


def efp():
    return O.epsi0_four_pi() / 78.4


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 10, 12), (6, 18, 14), (32, 32, 32)])
def test_fft_r2c_matches_numpy(shape):
    rng = np.random.default_rng(1)
    q = rng.standard_normal(shape)
    got = O.fft_r2c_3d(q)
    ref = np.fft.rfftn(q)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 10, 12), (6, 18, 14)])
def test_fft_c2r_is_unnormalised_inverse(shape):
    rng = np.random.default_rng(2)
    q = rng.standard_normal(shape)
    f = np.fft.rfftn(q)
    back = O.fft_c2r_3d(f, shape[2])
    assert np.abs(back - q * q.size).max() <= 1e-11 * q.size


# ------------------------------------------------------------- B-splines
def test_bspline_knot_values():
    """M_4 at knots 1,2,3 = 1/6, 2/3, 1/6; M_6 = (1, 26, 66, 26, 1)/120."""
    _, v4, _ = O.cardinal_bspline(4, 0.0)
    assert np.allclose(v4, [0, 1 / 6, 2 / 3, 1 / 6], rtol=0, atol=1e-15)
    _, v6, _ = O.cardinal_bspline(6, 0.0)
    assert np.allclose(v6, np.array([0, 1, 26, 66, 26, 1]) / 120, rtol=0, atol=1e-15)


@pytest.mark.parametrize("order", [3, 4, 5, 6, 7, 8])
def test_bspline_partition_and_derivative(order):
    rng = np.random.default_rng(order)
    for pos in rng.uniform(-40, 40, 50):
        k, v, d = O.cardinal_bspline(order, pos)
        assert k == math.floor(pos)
        assert abs(v.sum() - 1.0) < 1e-14
        assert abs(d.sum()) < 1e-14
        h = 1e-6
        _, vp, _ = O.cardinal_bspline(order, pos + h)
        _, vm, _ = O.cardinal_bspline(order, pos - h)
        if math.floor(pos + h) == k == math.floor(pos - h):
            assert np.allclose((vp - vm) / (2 * h), d, atol=1e-8)


def _mn(n, x):
    """Cardinal B-spline M_n(x) by the Cox-de Boor recursion."""
    if n == 1:
        return 1.0 if 0 <= x < 1 else 0.0
    return (x * _mn(n - 1, x) + (n - x) * _mn(n - 1, x - 1)) / (n - 1)


@pytest.mark.parametrize("order,s", [(4, 32), (6, 64), (5, 20), (8, 16)])
def test_influence_function_closed_form(order, s):
    """b^2 (eq. 4.4) and meshfac/energyfac/virfac (eqs. 3.9, 4.8)."""
    L = (37.0, 41.0, 45.0)
    alpha = 0.27
    sp = O.SpmeSetup(L, alpha, order, (s, s, s))
    b2, mesh, ener, vir = sp.tables()
    for i in range(3):
        for m in range(s):
            if order % 2 == 1 and 2 * m == s:
                continue
            a = sum(_mn(order, k + 1) * np.exp(2j * np.pi * m * k / s)
                    for k in range(order - 1))
            # the reference evaluates exp(i*2*pi*m*k/s) in single precision
            assert b2[i][m] == pytest.approx(1 / abs(a) ** 2, rel=2e-6)
    vol = L[0] * L[1] * L[2]
    h = s // 2 + 1
    rng = np.random.default_rng(0)
    for _ in range(200):
        nx, ny, nz = rng.integers(0, h, 3)
        if nx == ny == nz == 0:
            assert mesh[0, 0, 0] == 0
            continue
        k2 = sum((2 * np.pi * n / l) ** 2 for n, l in zip((nx, ny, nz), L))
        mf = 4 * np.pi / vol * np.exp(-k2 / (4 * alpha ** 2)) / k2 * \
            b2[0][nx] * b2[1][ny] * b2[2][nz]
        assert mesh[nz, ny, nx] == pytest.approx(mf, rel=1e-13)
        halves = sum(1 for n in (nx, ny, nz) if n in (0, s // 2))
        assert ener[nz, ny, nx] == pytest.approx(mf / 2 ** halves, rel=1e-13)
        assert vir[nz, ny, nx] == pytest.approx(
            (1 - k2 / (2 * alpha ** 2)) * mf / 2 ** halves, rel=1e-12)


# ------------------------------------------------------------------ SPME
def _ions(n, L, seed, charges=(1.0, -1.0)):
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < n:
        p = rng.uniform(-L / 2, L / 2, 3)
        if all(np.linalg.norm((p - q) - L * np.round((p - q) / L)) > 2.5 for q in pos):
            pos.append(p)
    az = np.array([charges[i % len(charges)] for i in range(n)])
    return np.array(pos), az


def test_spme_gather_is_energy_gradient():
    """F = -EpsiFourPi * dU_rec/dr from central differences of U_rec."""
    L = 24.0
    pos, az = _ions(40, L, 3)
    sp = O.SpmeSetup(L, 0.3, 6, (24, 24, 24))
    f, u, _ = O.spme_rec(sp, pos, az, 1.0)
    h = 1e-5
    for i in (0, 7, 19):
        for m in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, m] += h
            pm[i, m] -= h
            up = O.spme_rec(sp, pp, az, 1.0)[1]
            um = O.spme_rec(sp, pm, az, 1.0)[1]
            assert f[i, m] == pytest.approx(-(up - um) / (2 * h), rel=2e-6, abs=1e-8)


def test_spme_converges_to_standard_ewald():
    """TestEwaldSPM-style check (Src/potential.F90:3211-3378): SPME at high
    order on a fine mesh agrees with the direct k-space sum."""
    L = 20.0
    pos, az = _ions(30, L, 5, charges=(2.0, -1.0, -1.0))
    alpha = 0.35
    sp = O.SpmeSetup(L, alpha, 10, (64, 64, 64))
    fs, us, vs = O.spme_rec(sp, pos, az, 1.0)
    fe, ue, ve = O.ewald_rec_std(L, alpha, 14, pos, az, 1.0, cube=True)
    assert us == pytest.approx(ue, rel=1e-6)
    assert vs == pytest.approx(ve, rel=1e-6)
    # std Ewald seeds exp(ikx) in single precision; allow for that
    assert np.abs(fs - fe).max() <= 2e-5 * np.abs(fe).max()


def test_madelung_nacl():
    """Rock salt: E = -(N/2) * 1.747565 * q^2 / d with real space by
    direct summation, reciprocal space by the SPME oracle."""
    d = 2.8
    a = 2 * d
    base = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)])
    pos = base * d - a / 2 + 0.01
    az = np.array([1.0 if (i + j + k) % 2 == 0 else -1.0 for i, j, k in base])
    alpha = 1.2
    sp = O.SpmeSetup(a, alpha, 10, (32, 32, 32))
    _, urec, _ = O.spme_rec(sp, pos, az, 1.0)
    uself = O.ewald_self(alpha, az)
    ureal = 0.0
    n = 3
    for i in range(8):
        for j in range(8):
            for sx in range(-n, n + 1):
                for sy in range(-n, n + 1):
                    for sz in range(-n, n + 1):
                        if i == j and sx == sy == sz == 0:
                            continue
                        r = np.linalg.norm(pos[i] - pos[j] + a * np.array([sx, sy, sz]))
                        ureal += 0.5 * az[i] * az[j] * math.erfc(alpha * r) / r
    e = urec + uself + ureal
    assert e / 4 == pytest.approx(-1.747565 / d, rel=1e-6)


def test_spme_translation_invariance():
    L = 30.0
    pos, az = _ions(24, L, 9)
    sp = O.SpmeSetup(L, 0.3, 8, (48, 48, 48))
    _, u0, _ = O.spme_rec(sp, pos, az, 1.0)
    shifted = pos + np.array([0.37, -1.21, 2.05])
    shifted -= L * np.round(shifted / L)
    _, u1, _ = O.spme_rec(sp, shifted, az, 1.0)
    assert u1 == pytest.approx(u0, rel=1e-6)


# ------------------------------------------------------------- skin test
def test_skin_test_second_max_rule():
    """NList's r2max2 is only updated when r2 does not exceed r2max1
    (Src/nlist.F90:602-608), so an earlier smaller displacement is lost
    when a larger one follows."""
    drostep = np.zeros(6)
    drostep[0] = math.sqrt(3.0)   # r2 = 3 first
    drostep[3] = math.sqrt(5.0)   # then r2 = 5 becomes r2max1; r2max2 stays 0
    drosum = np.zeros(6)
    assert O.skin_test(drosum, drostep, 2.5) is False
    drosum = np.zeros(6)
    drostep2 = drostep[[3, 4, 5, 0, 1, 2]]  # larger first: r2max2 = 3
    assert O.skin_test(drosum, drostep2, 2.5) is True
    assert not drosum.any()


# ---- engine extensions without a reference counterpart -------------------
def test_philox_known_answers():
    """Philox4x32-10 against the Random123 known-answer vectors (counter,
    key -> output) for the all-zero, all-ones and pi-digit inputs."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert O.philox4x32_10(ctr, key) == want


def test_bd_noise_is_standard_normal_and_counter_based():
    n0 = O.bd_noise(12345, 100000, 0)
    # moments of N(0, 1) within 5 standard errors
    se = 1.0 / np.sqrt(n0.shape[0])
    assert np.all(np.abs(n0.mean(0)) < 5 * se)
    assert np.all(np.abs(n0.var(0) - 1.0) < 5 * np.sqrt(2.0) * se)
    assert abs(np.corrcoef(n0[:, 0], n0[:, 1])[0, 1]) < 5 * se
    # Kolmogorov-Smirnov distance to the normal CDF
    from scipy.stats import kstest
    assert kstest(n0[:, 2], "norm").pvalue > 1e-4
    # counter-based: a particle's draw depends only on (seed, i, step)
    sub = O.bd_noise(12345, 1000, 0)
    assert np.array_equal(sub, n0[:1000])
    assert not np.allclose(O.bd_noise(12345, 1000, 1), sub)
    assert not np.allclose(O.bd_noise(12346, 1000, 0), sub)


def test_knot_weights_closed_forms():
    """Uniform cubic B-spline segments: partition of unity, linear
    reproduction (coefficients c_k = a + b k give a + b (x + 1) on segment
    coordinates x = (r - r0)/h), derivatives by central differences, zero
    outside [r0, r0 + (nknot - 3) h)."""
    import ctypes as C
    lib = O.lib()
    lib.orc_knot_weights.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
    r0, r1, nk = 3.0, 12.0, 16
    dinv = (nk - 3) / (r1 - r0)

    def wts(r):
        i = C.c_int()
        w = (C.c_double * 4)()
        dw = (C.c_double * 4)()
        ok = lib.orc_knot_weights(r, r0, dinv, nk, C.byref(i), w, dw)
        return ok, i.value, np.array(w[:]), np.array(dw[:])

    rng = np.random.default_rng(3)
    for r in rng.uniform(r0, r1, 200):
        ok, i, w, dw = wts(r)
        assert ok and 0 <= i <= nk - 4
        assert abs(w.sum() - 1.0) < 1e-15 and abs(dw.sum()) < 1e-15
        x = (r - r0) * dinv
        a, b = 0.7, -1.3
        c = a + b * (i + np.arange(4))
        assert abs((c * w).sum() - (a + b * (x + 1.0))) < 1e-13
        h = 1e-6
        _, i2, wp, _ = wts(r + h)
        _, i3, wm, _ = wts(r - h)
        if i2 == i and i3 == i:
            assert np.allclose((wp - wm) / (2 * h), dw * dinv, atol=1e-8)
    assert wts(r0 - 1e-9)[0] == 0 and wts(r1)[0] == 0


def test_model_pair_energy_two_particles():
    """Two ions: E = q1 q2 efp erfc(alpha r)/r + 4 eps (s^12 - s^6) + eps
    inside the WCA cut, 0 LJ beyond; Lorentz-Berthelot mixing; minimum
    image; pairs beyond rc excluded."""
    L = 40.0
    efp = 138.935485 / 78.4
    zat, sig, eps = [1.0, -2.0], [3.0, 5.0], [1.5, 0.6]
    alpha, rc = 0.3, 12.0
    s, e = 0.5 * (sig[0] + sig[1]), math.sqrt(eps[0] * eps[1])
    ljrcut = 2.0 ** float(np.float32(1.0) / np.float32(6.0))
    for r in [3.7, 4.4, 4.6, 9.0, 11.9, 12.5]:
        # across the periodic boundary: x = -L/2 + 0.1 and x = L/2 + 0.1 - r
        ro = np.array([[-L / 2 + 0.1, 1.0, -2.0], [L / 2 + 0.1 - r, 1.0, -2.0]])
        got, _ = O.model_pair_energy(L, [1, 2], ro, zat, sig, eps, efp, alpha, rc)
        want = 0.0
        if r <= rc:
            want = zat[0] * zat[1] * efp * math.erfc(alpha * r) / r
            if r < ljrcut * s:
                x6 = (s / r) ** 6
                want += 4 * e * (x6 * x6 - x6) + e
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), r


def test_model_spline_forces_are_minus_gradient():
    rng = np.random.default_rng(11)
    L, n = 30.0, 40
    ro = rng.uniform(-L / 2, L / 2, (n, 3))
    iptpn = rng.integers(1, 3, n)
    c = rng.normal(size=(3, 12))
    knots = (2.0, 10.0, c)
    z = [0.0, 0.0]
    e0, fk = O.model_pair_energy(L, iptpn, ro, z, [1.0, 1.0], [0.0, 0.0], 1.0, 0.3, 10.0,
                                 lljcut=False, knots=knots)
    h = 1e-6
    for i, m in [(0, 0), (5, 1), (17, 2), (31, 0)]:
        rp, rm = ro.copy(), ro.copy()
        rp[i, m] += h
        rm[i, m] -= h
        ep, _ = O.model_pair_energy(L, iptpn, rp, z, [1.0, 1.0], [0.0, 0.0], 1.0, 0.3, 10.0,
                                    lljcut=False, knots=knots)
        em, _ = O.model_pair_energy(L, iptpn, rm, z, [1.0, 1.0], [0.0, 0.0], 1.0, 0.3, 10.0,
                                    lljcut=False, knots=knots)
        assert abs(-(ep - em) / (2 * h) - fk[i, m]) < 1e-6 * max(1.0, abs(fk[i, m]))
