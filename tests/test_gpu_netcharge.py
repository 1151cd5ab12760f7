"""The opt-in neutralising-background term of a net charge Q (MOLSIM only
warns about a net charge, Src/potential.F90:621-625).

Pinned by its closed form -- u%rec gains EpsiFourPi (-pi Q^2 / (2 V alpha^2)),
the virial -3 times that, forces none, dE/dq_i EpsiFourPi (-pi Q/(V alpha^2))
-- and by the property that fixes it: with the term, the Ewald energy of a
charged system does not depend on the splitting parameter alpha; without
it, it changes by exactly the term's change.
"""
import numpy as np
import pytest

from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def charged_c1(n_drop=40):
    """The 1:1 electrolyte of 1000 ions with n_drop anions removed: Q = +n_drop."""
    s = S.c1()
    keep = np.ones(s.np, bool)
    keep[np.nonzero(s.iptpn == 2)[0][:n_drop]] = False
    s.ro, s.rod, s.iptpn = s.ro[keep].copy(), s.rod[keep].copy(), s.iptpn[keep].copy()
    s.ewald = "spm"
    s.mesh, s.order = 64, 8
    return s


def engine(s, prec, alpha):
    s.alpha = alpha
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=alpha)
    eng = S.configure(s, prec, tables=tab)
    eng.set_pair_params(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                        alpha=alpha)
    return eng


def net_term(s, alpha):
    q = s.az.sum()
    return efp() * (-np.pi * q * q / (2.0 * s.boxlen ** 3 * alpha * alpha))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_net_charge_term_closed_form(prec):
    s = charged_c1()
    say(f"case: net charge Q={s.az.sum():+.0f}, closed form, {prec}")
    eng = engine(s, prec, 0.25)
    f0, e0 = eng.utotal()
    eng.set_net_charge(True)
    f1, e1 = eng.utotal()
    t = net_term(s, 0.25)
    say(f"  term {t:.12e}: d rec {e1.rec - e0.rec:.12e}, d virial {e1.virial - e0.virial:.12e}")
    assert abs((e1.rec - e0.rec) - t) <= 1e-9 * abs(t) + 1e-12 * abs(e0.rec)
    assert abs((e1.tot - e0.tot) - t) <= 1e-9 * abs(t) + 1e-12 * abs(e0.tot)
    assert abs((e1.virial - e0.virial) + 3.0 * t) <= 1e-9 * abs(t) + 1e-12 * abs(e0.virial)
    # no force; the FP64 spreading adds with LDS atomics in a run-dependent
    # order, so FP64 forces repeat to rounding only
    if prec == "mixed":
        assert np.array_equal(f0, f1)
    else:
        assert float(np.sqrt(((f1 - f0) ** 2).sum() / (f0 ** 2).sum())) <= 1e-14
    # off again: back to the reference's energy
    eng.set_net_charge(False)
    _, e2 = eng.utotal()
    assert abs(e2.tot - e0.tot) <= (0.0 if prec == "mixed" else 1e-13) * abs(e0.tot)
    eng.close()


def test_net_charge_term_makes_the_energy_alpha_independent():
    """Model energies (exact real-space pairs, SPME 64^3 order 8, self term)
    of a system with Q = +40 at alpha = 0.30 and 0.36 (rc 15 A): equal with
    the term to far below the term's change; without it they differ by that
    change.  dE/dq gains the term's derivative on every particle."""
    s = charged_c1()
    res = {}
    for alpha in (0.30, 0.36):
        for on in (False, True):
            eng = engine(s, "fp64", alpha)
            eng.set_net_charge(on)
            res[(alpha, on)] = eng.adjoint()
            eng.close()
    d_term = net_term(s, 0.30) - net_term(s, 0.36)
    d_on = res[(0.30, True)]["E"] - res[(0.36, True)]["E"]
    d_off = res[(0.30, False)]["E"] - res[(0.36, False)]["E"]
    say(f"case: Q=+40 energy change alpha 0.30 -> 0.36: with term {d_on:.3e}, without "
        f"{d_off:.3e}, term change {d_term:.3e}")
    assert abs(d_on) <= 0.02 * abs(d_term)
    assert abs(d_off + d_term) <= 0.02 * abs(d_term)
    q = s.az.sum()
    for alpha in (0.30, 0.36):
        dd = res[(alpha, True)]["dq"] - res[(alpha, False)]["dq"]
        want = efp() * (-np.pi * q / (s.boxlen ** 3 * alpha * alpha))
        assert np.allclose(dd, want, rtol=1e-9, atol=1e-12 * abs(want))
