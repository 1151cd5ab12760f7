"""The parameter adjoint (no reference counterpart) and the spline pair term.

The adjoint differentiates the model the tables interpolate -- Pot_1_6_12
with exact erfc, the spline term, SPME with its self term, the bonds.  The
oracle evaluates the same model energy independently (brute-force pair sum
with Erf2, its own SPME); every gradient component is checked against
central differences of it with h = 1e-4 (relative), to 1e-4 relative, on
both precision paths (the mixed path evaluates the pair terms in FP32).
The position adjoint is -F, checked against the oracle's UTotal forces.
"""
import numpy as np
import pytest

import oracle as O
from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp, oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def small_c2():
    """A 1500-particle slice of the polyelectrolyte case with every type:
    the first ten chains (500 beads, type 1), 750 counterions (type 2) and
    250 coions (type 3), a neutral set, so that every per-type gradient is
    exercised."""
    s = S.c2()
    keep = np.concatenate([np.arange(0, 500), np.arange(2000, 2750), np.arange(5000, 5250)])
    assert np.bincount(s.iptpn[keep]).tolist() == [0, 500, 750, 250]
    s.ro, s.rod, s.iptpn = s.ro[keep].copy(), s.rod[keep].copy(), s.iptpn[keep].copy()
    s.chains = [c for c in s.chains if c[0] - 1 + c[1] <= 500]
    return s


def model_energy(s, zat, sig, eps, knots=None):
    """Oracle model energy: exact pair sum + efp (u_SPME + u_self)."""
    e_pair, _ = O.model_pair_energy(s.boxlen, s.iptpn, s.ro, zat, sig, eps, efp(), s.alpha,
                                    s.rcut, knots=knots)
    az = np.asarray(zat)[s.iptpn - 1]
    setup = O.SpmeSetup(s.boxlen, s.alpha, s.order, (s.mesh,) * 3)
    _, u_rec, _ = O.spme_rec(setup, s.ro, az, efp())
    u_self = O.ewald_self(s.alpha, az)
    return e_pair + efp() * (u_rec + u_self)


# per precision: spline-term forces (rms error over the rms of the total
# force), energies (relative to the total), knot gradients, position adjoint
TOL = {"fp64": dict(f=1e-10, e=1e-9, knot=1e-6, dro=1e-10),
       "mixed": dict(f=1e-5, e=2e-7, knot=1e-5, dro=1e-5)}


@pytest.fixture(scope="module")
def slice_sys():
    return small_c2()


@pytest.fixture(scope="module", params=["fp64", "mixed"])
def setup(request):
    prec = request.param
    s = small_c2()
    s.chains = []          # bonds do not depend on the parameters
    rng = np.random.default_rng(21)
    nptpt = 6
    knots = (4.0, s.rcut, 0.05 * rng.normal(size=(nptpt, 12)))
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab)
    eng.set_pair_params(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                        alpha=s.alpha)
    yield s, tab, eng, knots, prec
    eng.close()


def test_spline_term_forces_match_oracle(setup):
    s, tab, eng, knots, prec = setup
    tol = TOL[prec]
    say(f"case: spline pair term, N={s.np}, 12 knots on [4, {s.rcut}], {prec}")
    eng.set_knots(knots[0], knots[1], None)
    f0, e0 = eng.utotal()
    eng.set_knots(*knots)
    f1, e1 = eng.utotal()
    _, fk = O.model_pair_energy(s.boxlen, s.iptpn, s.ro, [0.0] * 3, s.sigat, [0.0] * 3, efp(),
                                s.alpha, s.rcut, lljcut=False, knots=knots)
    fk_o = fk * E.force_unit_fac()
    ferr = np.sqrt(np.mean(((f1 - f0) - fk_o) ** 2))
    if prec == "fp64":
        err = rel_rms(f1 - f0, fk_o)
    else:
        # FP32 rounds each total force, of which the spline part is a
        # fraction: its error is measured on the scale of the total force
        err = ferr / np.sqrt(np.mean(f0 ** 2))
    say(f"  spline-term force error {err:.2e} (rel-RMS {rel_rms(f1 - f0, fk_o):.2e})")
    assert err <= tol["f"]
    ek, _ = O.model_pair_energy(s.boxlen, s.iptpn, s.ro, [0.0] * 3, s.sigat, [0.0] * 3, efp(),
                                s.alpha, s.rcut, lljcut=False, knots=knots)
    de = abs((e1.tot - e0.tot) - ek)
    say(f"  spline-term energy {ek:.10e}, error {de:.2e} (total {e0.tot:.6e})")
    scale = max(1.0, abs(ek)) if prec == "fp64" else abs(e0.tot)
    assert de <= tol["e"] * scale


def test_adjoint_energy_and_gradients_match_finite_differences(setup):
    s, tab, eng, knots, prec = setup
    tol = TOL[prec]
    eng.set_knots(*knots)
    say(f"case: adjoint, N={s.np} SPME {s.mesh}^3 order {s.order}, {prec}")
    a = eng.adjoint()
    zat, sig, eps = list(s.zat), list(s.sigat), list(s.epsat)
    e_ref = model_energy(s, zat, sig, eps, knots)
    say(f"  model energy engine {a['E']:.12e} oracle {e_ref:.12e}")
    assert abs(a["E"] - e_ref) <= tol["e"] * abs(e_ref)

    def fd(which, t, h):
        p = {"zat": list(zat), "sig": list(sig), "eps": list(eps)}
        p[which][t] += h
        ep = model_energy(s, p["zat"], p["sig"], p["eps"], knots)
        p[which][t] -= 2 * h
        em = model_energy(s, p["zat"], p["sig"], p["eps"], knots)
        return (ep - em) / (2 * h)

    for t in range(3):
        for which, key, base in (("zat", "dzat", zat), ("sig", "dsigat", sig),
                                 ("eps", "depsat", eps)):
            h = 1e-4 * abs(base[t])
            g_fd = fd(which, t, h)
            g = a[key][t]
            say(f"  d/d{which}[{t}]: adjoint {g:.10e} FD {g_fd:.10e}")
            assert g_fd != 0.0  # every type takes part
            assert abs(g - g_fd) <= 1e-4 * max(abs(g_fd), 1e-3 * abs(e_ref))
    # knots: the energy is linear in c, so differences are exact up to rounding
    c = knots[2]
    rng = np.random.default_rng(2)
    for p, k in zip(rng.integers(0, 6, 5), rng.integers(0, 12, 5)):
        cp, cm = c.copy(), c.copy()
        h = 1e-3
        cp[p, k] += h
        cm[p, k] -= h
        g_fd = (model_energy(s, zat, sig, eps, (knots[0], knots[1], cp)) -
                model_energy(s, zat, sig, eps, (knots[0], knots[1], cm))) / (2 * h)
        say(f"  d/dc[{p}, {k}]: adjoint {a['dknot'][p, k]:.10e} FD {g_fd:.10e}")
        assert abs(a["dknot"][p, k] - g_fd) <= tol["knot"] * max(1.0, abs(g_fd))
    # dE/dzat is the per-type sum of dE/dq
    for t in range(3):
        assert abs(a["dq"][s.iptpn == t + 1].sum() - a["dzat"][t]) <= 1e-9 * abs(a["dzat"][t])


def test_position_adjoint_is_minus_force(setup):
    s, tab, eng, knots, prec = setup
    eng.set_knots(knots[0], knots[1], None)
    a = eng.adjoint()
    orc = oracle_for(s, tab)
    orc.set_list(s.ro)
    fo, _ = orc.utotal(s.ro)
    grad = -fo / E.force_unit_fac()
    say(f"case: position adjoint vs -F oracle ({prec}), rel-RMS {rel_rms(a['dro'], grad):.2e}")
    assert rel_rms(a["dro"], grad) <= TOL[prec]["dro"]


def test_torch_autograd_bridge(slice_sys):
    import torch
    from molsim_amd import torch_adjoint as T
    s = slice_sys
    say("case: torch.autograd ModelEnergy, 16 knots")
    model = T.PairModel(s, nknot=16)
    ro = torch.tensor(s.ro, dtype=torch.float64, requires_grad=True)
    zat = torch.tensor(s.zat, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(s.sigat, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(s.epsat, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(0.02 * np.random.default_rng(4).normal(size=(6, 16)), requires_grad=True)
    e = T.ModelEnergy.apply(ro, zat, sig, eps, c, model)
    e.backward()
    # the energy is consistent with its gradient: central differences of
    # the engine's own model energy
    h = 1e-4
    for t in range(3):
        zp, zm = s.zat.copy(), s.zat.copy()
        zp[t] += h * abs(s.zat[t])
        zm[t] -= h * abs(s.zat[t])
        ep = model.evaluate(s.ro, zp, s.sigat, s.epsat, c.detach().numpy())["E"]
        em = model.evaluate(s.ro, zm, s.sigat, s.epsat, c.detach().numpy())["E"]
        g_fd = (ep - em) / (2 * h * abs(s.zat[t]))
        say(f"  dE/dzat[{t}] autograd {zat.grad[t].item():.10e} FD {g_fd:.10e}")
        assert abs(zat.grad[t].item() - g_fd) <= 1e-4 * abs(g_fd)
    assert ro.grad.shape == ro.shape and torch.isfinite(ro.grad).all()
    assert torch.isfinite(sig.grad).all() and torch.isfinite(eps.grad).all()
    assert c.grad.shape == c.shape
    model.close()
