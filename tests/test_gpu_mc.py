"""Monte Carlo trial moves: the SPME reciprocal-energy change of moving
some particles (DUTwoBodyEwaldRecSPM, Src/denergy.F90:1413-1495) against
the oracle's reciprocal energy (UEwaldRecSPM restated) of both
configurations.  The engine forms the change from the kept spectrum of the
current mesh and the transform of the moved stencils (no difference of two
energies), so its error follows the change, not |E_rec|.

Tolerances: FP64 |du - du_oracle| <= 1e-9 |E_rec|; mixed <= 1e-5 |E_rec|
and, what a Metropolis decision sees, <= 1e-4 kT at 298 K (measured
<= 6e-6 kT).
"""
import numpy as np
import pytest

import oracle as O
from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp, rel_rms

pytestmark = pytest.mark.gpu


KT_298 = S.KT298   # kJ/mol


def oracle_rec(s, ro):
    setup = O.SpmeSetup(s.boxlen, s.alpha, s.order, (s.mesh,) * 3)
    _, u, _ = O.spme_rec(setup, ro, s.az, efp())
    return efp() * u


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("mixed", 1e-5)])
def test_mc_delta_rec_vs_oracle(prec, tol):
    s = S.c2()
    print("case: MC trial moves c2", prec, flush=True)
    eng = S.configure(s, prec)
    f_before, en_before = eng.utotal()
    e0 = oracle_rec(s, s.ro)
    rng = np.random.default_rng(11)
    for natm in (1, 3, 8):
        ia = rng.choice(s.np, natm, replace=False)
        rtm = s.ro[ia] + rng.uniform(-2.5, 2.5, (natm, 3))
        du = eng.mc_delta_rec(ia + 1, rtm)
        ro1 = s.ro.copy()
        ro1[ia] = rtm
        du0 = oracle_rec(s, ro1) - e0
        print(f"  natm {natm}: du {du:.10e} oracle {du0:.10e} "
              f"({abs(du - du0) / KT_298:.1e} kT)", flush=True)
        assert abs(du - du0) <= tol * abs(e0)
        if prec == "mixed":
            assert abs(du - du0) <= 1e-4 * KT_298
    # a particle moved across the box edge (the mesh is periodic)
    rtm = s.ro[[9]].copy()
    rtm[0, 0] += 0.5 * s.boxlen
    du = eng.mc_delta_rec([10], rtm)
    ro1 = s.ro.copy()
    ro1[9] = rtm[0]
    assert abs(du - (oracle_rec(s, ro1) - e0)) <= tol * abs(e0)
    # no move, no change
    assert abs(eng.mc_delta_rec([5], s.ro[4:5])) <= tol * abs(e0)
    # the engine's own state is untouched (bitwise on the mixed path; the
    # FP64 spreading adds in the order of its LDS atomics, which varies
    # from run to run at the 1e-16 level)
    f_after, en_after = eng.utotal()
    if prec == "mixed":
        assert np.array_equal(f_after, f_before)
        assert en_after.tot == en_before.tot
    else:
        assert rel_rms(f_after, f_before) <= 1e-13
        assert abs(en_after.tot - en_before.tot) <= 1e-13 * abs(en_before.tot)
    eng.close()


def test_mc_delta_rec_rejects_bad_moves():
    s = S.c2()
    eng = S.configure(s, "fp64")
    for ia, rtm in (([3, 3], s.ro[[2, 2]]), ([0], s.ro[:1]), ([s.np + 1], s.ro[:1]),
                    ([4], np.full((1, 3), np.nan))):
        with pytest.raises(E.MfeError) as ei:
            eng.mc_delta_rec(ia, rtm)
        assert E.ERRORS[ei.value.code] == "MFE_E_BAD_PARAM"
    eng.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 0.0)])
def test_mc_reuse_current_mesh(prec, tol):
    """Trials that start from the kept spectrum (reuse_current) give the
    energy changes of trials that transform the current mesh anew, and the
    engine transforms anew once its positions change (bitwise on the mixed
    path, whose spreading does not depend on the order of its adds)."""
    s = S.c2()
    print("case: MC trials from the kept mesh c2", prec, flush=True)
    eng = S.configure(s, prec)
    e0 = abs(oracle_rec(s, s.ro))
    rng = np.random.default_rng(3)
    moves = [(rng.choice(s.np, n, replace=False), None) for n in (1, 2, 5)]
    moves = [(ia, s.ro[ia] + rng.uniform(-2.0, 2.0, (ia.size, 3))) for ia, _ in moves]
    full = [eng.mc_delta_rec(ia + 1, rtm) for ia, rtm in moves]
    eng.mc_delta_rec(moves[0][0] + 1, moves[0][1])
    kept = [eng.mc_delta_rec(ia + 1, rtm, reuse_current=True) for ia, rtm in moves]
    for a, b in zip(full, kept):
        assert abs(a - b) <= tol * e0, (a, b)
    # an accepted move: new positions uploaded; a reuse request spreads anew
    ia, rtm = moves[1]
    ro1 = s.ro.copy()
    ro1[ia] = rtm
    ro1 -= s.boxlen * np.round(ro1 / s.boxlen)
    eng.upload_positions(ro1)
    ia2 = np.array([17])
    rtm2 = ro1[ia2] + 1.0
    du = eng.mc_delta_rec(ia2 + 1, rtm2, reuse_current=True)
    ro2 = ro1.copy()
    ro2[ia2] = rtm2
    du0 = oracle_rec(s, ro2) - oracle_rec(s, ro1)
    assert abs(du - du0) <= (1e-9 if prec == "fp64" else 1e-5) * e0
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_mc_delta_pair_vs_oracle(prec):
    """UTwoBodyANew - UTwoBodyAOld (Src/denergy.F90:289-495) against the
    oracle's two-body energy of both configurations over the same list
    (FP64 on both engine paths: <= 1e-10 of |u_twob|)."""
    from helpers import oracle_for
    s = S.c2()
    print("case: MC pair deltas c2", prec, flush=True)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab)
    eng.build_list()
    orc = oracle_for(s, tab)
    orc.set_list(s.ro)
    _, eo = orc.utotal(s.ro)
    u0 = eo["twob0"]
    rng = np.random.default_rng(5)
    for natm in (1, 2, 6):
        ia = rng.choice(s.np, natm, replace=False)
        rtm = s.ro[ia] + rng.uniform(-0.8, 0.8, (natm, 3))
        rtm -= s.boxlen * np.round(rtm / s.boxlen)
        du, ov = eng.mc_delta_pair(ia + 1, rtm)
        ro1 = s.ro.copy()
        ro1[ia] = rtm
        _, e1 = orc.utotal(ro1)
        print(f"  natm {natm}: du {du:.10e} oracle {e1['twob0'] - u0:.10e} overlap {ov}",
              flush=True)
        assert not ov and du != 0.0
        assert abs(du - (e1["twob0"] - u0)) <= 1e-10 * abs(u0)
    # a particle moved onto another one overlaps
    du, ov = eng.mc_delta_pair([1], s.ro[[1]] + 0.1)
    assert ov and du == 0.0
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_mc_delta_pair_cell_list_vs_oracle(prec):
    """UTwoBodyANewCellList - UTwoBodyAOldCellList (Src/denergy.F90:728-928,
    chosen for lCellList at :145-147): every particle within rcut (r2 <
    rcut2) of the old and the new position, including moves far beyond the
    Verlet skin, against the oracle's UTwoBodyACellList energy of both
    configurations (<= 1e-10 of |u_twob|)."""
    from helpers import oracle_for
    s = S.c2()
    print("case: MC pair deltas, cell list, c2", prec, flush=True)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab, nlist="cell")
    eng.build_list()
    orc = oracle_for(s, tab, nlist="cell")
    orc.set_list(s.ro)
    _, eo = orc.utotal(s.ro)
    u0 = eo["twob0"]
    rng = np.random.default_rng(9)
    done = 0
    for natm, step in ((1, 0.8), (3, 0.8), (1, 12.0), (4, 25.0)):
        while True:
            ia = rng.choice(s.np, natm, replace=False)
            rtm = _wrap(s.ro[ia] + rng.uniform(-step, step, (natm, 3)), s.boxlen)
            du, ov = eng.mc_delta_pair(ia + 1, rtm)
            if not ov:
                break
        ro1 = s.ro.copy()
        ro1[ia] = rtm
        orc.set_list(ro1)
        _, e1 = orc.utotal(ro1)
        print(f"  natm {natm} step {step}: du {du:.10e} oracle {e1['twob0'] - u0:.10e}",
              flush=True)
        assert du != 0.0
        assert abs(du - (e1["twob0"] - u0)) <= 1e-10 * abs(u0)
        done += 1
    assert done == 4
    du, ov = eng.mc_delta_pair([1], s.ro[[1]] + 0.1)
    assert ov and du == 0.0
    eng.close()


def _wrap(r, L):
    return r - L * np.round(r / L)


def _free_move(eng, s, ro, rng, natm, step):
    """A trial move that the reference's MC would not reject for overlap
    (lhsoverlap of UTwoBodyANew), drawn again until it does not overlap."""
    while True:
        ia = rng.choice(s.np, natm, replace=False)
        rtm = _wrap(ro[ia] + rng.uniform(-step, step, (natm, 3)), s.boxlen)
        if not eng.mc_delta_pair(ia + 1, rtm)[1]:
            return ia, rtm


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_mc_accept_keeps_the_trial_mesh(prec):
    """EwaldUpdateArray (Src/denergy.F90:1776-1819) with the accepted
    positions: after mc_accept of the move just tried, the next trial with
    reuse_current, from the kept FP64 spectrum with the accepted change
    added, equals a fresh spread of the accepted configuration uploaded
    whole (FP64: <= 1e-12 |E_rec|; mixed, whose kept spectrum carries the
    first FP32 transform's rounding where a fresh one carries its own:
    <= 1e-4 kT) and the oracle's energy change; positions, forces and
    energies after the accepts equal those of the uploaded configuration."""
    s = S.c2()
    print("case: MC accept c2", prec, flush=True)
    eng = S.configure(s, prec)
    ref = S.configure(s, prec)
    eng.utotal()
    ref.utotal()
    rng = np.random.default_rng(23)
    ro = s.ro.copy()
    tol = 1e-9 if prec == "fp64" else 1e-5
    for it, natm in enumerate((1, 3, 2)):
        ia, rtm = _free_move(eng, s, ro, rng, natm, 2.0)
        eng.mc_delta_rec(ia + 1, rtm, reuse_current=it > 0)
        eng.mc_accept(ia + 1, rtm)
        ro[ia] = rtm
        ref.upload_positions(ro)
        ib = rng.choice(s.np, 2, replace=False)
        rtb = _wrap(ro[ib] + rng.uniform(-2.0, 2.0, (2, 3)), s.boxlen)
        d_keep = eng.mc_delta_rec(ib + 1, rtb, reuse_current=True)
        d_ref = ref.mc_delta_rec(ib + 1, rtb)
        e0 = oracle_rec(s, ro)
        ro_b = ro.copy()
        ro_b[ib] = rtb
        d_orc = oracle_rec(s, ro_b) - e0
        print(f"  accept {it} natm {natm}: next trial kept {d_keep:.12e} "
              f"fresh {d_ref:.12e} oracle {d_orc:.12e}", flush=True)
        if prec == "mixed":
            assert abs(d_keep - d_ref) <= 1e-4 * KT_298
            assert abs(d_keep - d_orc) <= 1e-4 * KT_298
        else:
            assert abs(d_keep - d_ref) <= 1e-12 * abs(e0)
        assert abs(d_keep - d_orc) <= tol * abs(e0)
    # an accept of a move that was not the last one tried: the next trial
    # spreads anew and still agrees
    ia, rtm = _free_move(eng, s, ro, rng, 2, 1.5)
    eng.mc_accept(ia + 1, rtm)
    ro[ia] = rtm
    ref.upload_positions(ro)
    r0 = _wrap(ro[[0]] + 0.7, s.boxlen)
    d_keep = eng.mc_delta_rec([1], r0, reuse_current=True)
    d_ref = ref.mc_delta_rec([1], r0)
    assert abs(d_keep - d_ref) <= 1e-12 * abs(oracle_rec(s, ro))
    # the engine's state is that of the uploaded configuration
    r1 = eng.download_state()[0].reshape(-1, 3)
    assert np.array_equal(r1, ro)
    f1, en1 = eng.utotal()
    f2, en2 = ref.utotal()
    if prec == "mixed":
        assert np.array_equal(f1, f2)
        assert en1.tot == en2.tot
    else:
        assert rel_rms(f1, f2) <= 1e-12
        assert abs(en1.tot - en2.tot) <= 1e-12 * abs(en2.tot)
    # a move outside the box is refused
    with pytest.raises(RuntimeError):
        eng.mc_accept([1], np.array([[0.6 * s.boxlen, 0.0, 0.0]]))
    eng.close()
    ref.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_mc_large_moves(prec):
    """Moves of more than 64 particles, which the pair delta and the accept
    take through device copies of the move (smaller ones are staged in LDS
    straight from host memory): the pair delta against the oracle's
    two-body energies (<= 1e-10 of |u_twob|), and the accepted
    configuration equal to the same positions uploaded whole (bitwise on
    the mixed path, to 1e-12 on the FP64 path)."""
    from helpers import oracle_for
    s = S.c2()
    natm = 70
    print("case: MC moves of", natm, "particles c2", prec, flush=True)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab)
    ref = S.configure(s, prec, tables=tab)
    eng.utotal()
    ref.utotal()
    orc = oracle_for(s, tab)
    orc.set_list(s.ro)
    _, eo = orc.utotal(s.ro)
    u0 = eo["twob0"]
    rng = np.random.default_rng(31)
    ia = rng.choice(s.np, natm, replace=False)
    rtm = _wrap(s.ro[ia] + rng.uniform(-0.3, 0.3, (natm, 3)), s.boxlen)
    du, ov = eng.mc_delta_pair(ia + 1, rtm)
    ro1 = s.ro.copy()
    ro1[ia] = rtm
    _, e1 = orc.utotal(ro1)
    print(f"  du {du:.10e} oracle {e1['twob0'] - u0:.10e} overlap {ov}", flush=True)
    assert not ov
    assert abs(du - (e1["twob0"] - u0)) <= 1e-10 * abs(u0)
    eng.mc_delta_rec(ia + 1, rtm)
    eng.mc_accept(ia + 1, rtm)
    ref.upload_positions(ro1)
    assert np.array_equal(eng.download_state()[0].reshape(-1, 3), ro1)
    f1, en1 = eng.utotal()
    f2, en2 = ref.utotal()
    if prec == "mixed":
        assert np.array_equal(f1, f2) and en1.tot == en2.tot
    else:
        assert rel_rms(f1, f2) <= 1e-12
        assert abs(en1.tot - en2.tot) <= 1e-12 * abs(en2.tot)
    eng.close()
    ref.close()
