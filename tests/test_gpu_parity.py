"""HIP engine against the oracle and the reference's golden vectors.

Tolerances (stated per result):
  FP64 path:  forces rel-RMS <= 1e-10 and energies rel <= 1e-11 against the
              oracle for table pairs and SPME (summation order differs);
              standard Ewald <= 2e-6, because the reference seeds
              exp(i k x) in single precision (CMPLX without KIND,
              Src/energy.F90:1404-1409) while the engine uses FP64 sincos.
  mixed path: forces rel-RMS <= 1e-5, energies rel <= 1e-7.
  pair sets:  bit-exact.
"""
import os

import numpy as np
import pytest

import oracle as O
from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp, oracle_for, owned_pairs, rel_rms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def golden():
    here = os.path.join(os.path.dirname(__file__), "golden")
    return {k: np.load(os.path.join(here, k + ".npz"))
            for k in ("c1_std", "c1_cell_surf", "chains", "chain_angles")}


def c1_from_golden(z):
    s = S.c1()
    s.ro = z["probe_pos"].copy()
    return s


# ------------------------------------------------------------- pair sets
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pair_set_matches_reference(golden, prec):
    z = golden["c1_std"]
    s = c1_from_golden(z)
    say("case: pair set c1 golden", prec)
    eng = S.configure(s, prec)
    eng.build_list()
    nn, jl = eng.get_pairs(1024)
    assert np.array_equal(nn, z["ref_nneighpn"])
    import hashlib
    flat = np.concatenate([jl[i, :nn[i]] for i in range(s.np)]).astype("<i4")
    assert hashlib.sha256(flat.tobytes()).hexdigest() == str(z["ref_jpnlist_sha256"])
    eng.close()


def test_pair_set_bruteforce_c2():
    s = S.c2()
    say("case: pair set c2 vs brute force")
    eng = S.configure(s, "mixed")
    eng.build_list()
    nn, jl = eng.get_pairs(2048)
    nn0, jl0 = O.set_vlist_md(s.boxlen, s.ro, s.rcut + s.drnlist, 2048)
    assert np.array_equal(nn, nn0)
    for i in range(s.np):
        assert np.array_equal(jl[i, :nn[i]], jl0[i, :nn0[i]]), i
    eng.close()


# --------------------------------------------------------- table forces
@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 1e-12, 1e-12), ("mixed", 1e-5, 1e-7)])
def test_pair_forces_vs_oracle(prec, ftol, etol):
    s = S.c1()
    s.ewald = "none"
    say("case: table pair forces c1", prec)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab)
    f, en = eng.utotal()
    orc = oracle_for(s, tab, ewald="none")
    orc.set_list(s.ro)
    fo, eo = orc.utotal(s.ro)
    assert rel_rms(f, fo) <= ftol
    assert abs(en.twob0 - eo["twob0"]) <= etol * abs(eo["twob0"])
    assert np.allclose(en.twob, eo["twob"], rtol=etol * 10, atol=0)
    assert abs(en.virial - eo["virial"]) <= etol * abs(eo["virial"])
    eng.close()


@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 2e-6, 1e-7), ("mixed", 1e-5, 1e-7)])
def test_utotal_std_ewald_vs_reference(golden, prec, ftol, etol):
    z = golden["c1_std"]
    s = c1_from_golden(z)
    say("case: UTotal std Ewald c1 golden", prec)
    eng = S.configure(s, prec)
    f, en = eng.utotal()
    assert rel_rms(f, z["ref_forceo"]) <= ftol
    assert abs(en.tot - z["ref_u_tot"]) <= etol * abs(z["ref_u_tot"])
    assert abs(en.rec - z["ref_u_rec"]) <= etol * 10 * abs(z["ref_u_rec"])
    eng.close()


@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 2e-6, 1e-7), ("mixed", 1e-5, 1e-7)])
def test_reference_cnf_drives_the_engine(prec, ftol, etol):
    """IOCnf (Src/molsim.F90:840-963): the configuration file the
    reference's C1 run wrote, read with molsim_amd.cnf, drives the engine
    to the forces and energy the reference's UTotal gives at the same
    positions (c1_final_utotal.npz, drawn by the reference probe)."""
    from molsim_amd import cnf
    here = os.path.join(os.path.dirname(__file__), "golden")
    c = cnf.read_cnf(os.path.join(here, "c1_final.cnf"))
    z = np.load(os.path.join(here, "c1_final_utotal.npz"))
    s = S.c1()
    assert c.np == s.np and np.array_equal(c.boxlen, [s.boxlen] * 3)
    s.ro = c.ro.copy()
    s.rod = c.rod.copy()
    say("case: UTotal at the reference's c1_final.cnf", prec)
    eng = S.configure(s, prec)
    f, en = eng.utotal()
    r = rel_rms(f, z["ref_forceo"])
    say(f"  force rel-RMS {r:.2e}, u_tot {en.tot:.12e} vs {float(z['ref_u_tot']):.12e}")
    assert r <= ftol
    assert abs(en.tot - z["ref_u_tot"]) <= etol * abs(z["ref_u_tot"])
    eng.close()


@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 2e-6, 1e-7), ("mixed", 1e-5, 1e-7)])
def test_utotal_cell_list_surface_vs_reference(golden, prec, ftol, etol):
    zc = golden["c1_cell_surf"]
    s = c1_from_golden(golden["c1_std"])
    s.ro = zc["probe_pos"].copy()
    s.lsurf = True
    say("case: UTotal cell list + surface c1 golden", prec)
    eng = S.configure(s, prec, nlist="cell")
    f, en = eng.utotal()
    say(f"  force rel-RMS {rel_rms(f, zc['ref_forceo']):.2e}, energy "
        f"{abs(en.tot - zc['ref_u_tot']) / abs(zc['ref_u_tot']):.2e}")
    assert rel_rms(f, zc["ref_forceo"]) <= ftol
    assert abs(en.tot - zc["ref_u_tot"]) <= etol * abs(zc["ref_u_tot"])
    eng.close()


def test_utotal_bonds_vs_reference(golden):
    z = golden["chains"]
    s = S.System("chains", 60.0, z["probe_pos"].copy(), np.zeros_like(z["probe_pos"]),
                 np.repeat([1, 2], z["nppt"]).astype(np.int32), z["zat"], z["sigat"],
                 z["epsat"], np.array([30.0, 23.0]), rcut=20.0, alpha=float(z["alpha"]),
                 ewald="std", ncut=5, chains=[(1 + c * 10, 10) for c in range(4)],
                 bond=(float(z["bond"][0]), 2, float(z["bond"][2])))
    say("case: UTotal with bonds, golden chains")
    eng = S.configure(s, "fp64")
    f, en = eng.utotal()
    assert abs(en.bond - float(z["ref_u_bond"])) <= 1e-12 * max(abs(float(z["ref_u_bond"])), 1e-3)
    assert rel_rms(f, z["ref_forceo"]) <= 2e-6
    assert abs(en.tot - float(z["ref_u_tot"])) <= 1e-7 * abs(float(z["ref_u_tot"]))
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reference_ubuf_drives_the_engine(prec):
    """MOLSIM's own ubuf/iubuflow/r2umin (the reference run's tables for the
    polyelectrolyte potential, tests/golden/tables_c2.npz) handed to
    mfe_set_tables as a MOLSIM host would hand its arrays: the engine gives
    the same forces and energies as with the tables it builds itself
    (bitwise), and matches the oracle on the reference's tables."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "tables_c2.npz"))
    s = S.c2()
    assert float(z["rcut"]) == s.rcut and float(z["alpha"]) == s.alpha
    say("case: reference ubuf through mfe_set_tables", prec)
    ref_tab = E.Tables(z["ubuf"].copy(), z["iubuflow"].astype(np.int32), z["nugrid"].copy(),
                       z["r2umin"].copy())
    own_tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                             alpha=s.alpha)
    eng = S.configure(s, prec, tables=ref_tab)
    f, en = eng.utotal()
    eng.close()
    eng = S.configure(s, prec, tables=own_tab)
    f2, en2 = eng.utotal()
    eng.close()
    if prec == "mixed":
        assert np.array_equal(f, f2) and en.tot == en2.tot
    else:
        assert rel_rms(f, f2) <= 1e-13 and abs(en.tot - en2.tot) <= 1e-13 * abs(en2.tot)
    orc = oracle_for(s, ref_tab)
    orc.set_list(s.ro)
    fo, eo = orc.utotal(s.ro)
    assert rel_rms(f, fo) <= (1e-10 if prec == "fp64" else 1e-5)
    etol = 1e-11 if prec == "fp64" else 1e-7
    assert abs(en.twob0 - eo["twob0"]) <= etol * abs(eo["twob0"])


@pytest.mark.parametrize("a", [0, 1])
def test_utotal_angles_vs_reference(golden, a):
    """UAngle on the golden chains against the reference run with bond
    angles on (p = 2 and p = 3): u%angle (the reference's u%tot less its
    u%tot without angles) and u%tot, forces within the FP64 path's bound."""
    z, za = golden["chains"], golden["chain_angles"]
    ka, pa, eqa = za["angles"][a]
    s = S.System("chains", 60.0, z["probe_pos"].copy(), np.zeros_like(z["probe_pos"]),
                 np.repeat([1, 2], z["nppt"]).astype(np.int32), z["zat"], z["sigat"],
                 z["epsat"], np.array([30.0, 23.0]), rcut=20.0, alpha=float(z["alpha"]),
                 ewald="std", ncut=5, chains=[(1 + c * 10, 10) for c in range(4)],
                 bond=(float(z["bond"][0]), 2, float(z["bond"][2])))
    say("case: UTotal with angles, golden chains", (ka, pa, eqa))
    eng = S.configure(s, "fp64")
    eng.set_angles(float(ka), int(pa), float(eqa))
    f, en = eng.utotal()
    ref_tot = float(za[f"u_tot{a}"])
    ref_angle = ref_tot - float(z["ref_u_tot"])
    ua = eng.angle_energy()
    say(f"  u%angle {ua:.15e} vs reference {ref_angle:.15e}")
    assert abs(ua - ref_angle) <= 1e-12 * abs(ref_tot)
    assert abs(en.tot - ref_tot) <= 1e-7 * abs(ref_tot)
    assert rel_rms(f, za[f"forceo{a}"]) <= 2e-6
    eng.close()


# ----------------------------------------------------------------- SPME
@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 1e-10, 1e-11), ("mixed", 1e-5, 1e-7)])
def test_spme_c2_vs_oracle(prec, ftol, etol):
    s = S.c2()
    say("case: SPME c2 (N=6000, 64^3, order 4)", prec)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab)
    f, en = eng.utotal()
    orc = oracle_for(s, tab)
    orc.set_list(s.ro)
    fo, eo = orc.utotal(s.ro)
    r = rel_rms(f, fo)
    say(f"  force rel-RMS {r:.3e}  tot {en.tot:.12e} vs {eo['tot']:.12e}  "
        f"rec {en.rec:.12e} vs {eo['rec']:.12e}")
    assert r <= ftol
    assert abs(en.rec - eo["rec"]) <= etol * abs(eo["rec"])
    assert abs(en.tot - eo["tot"]) <= etol * abs(eo["tot"])
    assert abs(en.bond - eo["bond"]) <= 1e-12 * abs(eo["bond"])
    assert abs(en.virial - eo["virial"]) <= etol * 10 * abs(eo["virial"])
    eng.close()


@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 1e-10, 1e-11), ("mixed", 1e-5, 1e-7)])
@pytest.mark.parametrize("angle", [(0.02, 2, 109.5), (0.0005, 3, 150.0)])
def test_utotal_with_angles_vs_oracle(prec, ftol, etol, angle):
    """UAngle (Src/energy.F90:5409-5481) on the chains of the polyelectrolyte
    system: u%angle within 1e-12 of the oracle on both precision paths (the
    angle term is FP64 on both), u%tot and forces within the path's bound,
    the virial unchanged by the angles, and the angle term's own forces
    (engine with minus engine without) against the oracle's UAngle."""
    s = S.c2()
    say("case: UTotal with angles, c2 chains", prec, angle)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, prec, tables=tab)
    f0, en0 = eng.utotal()
    eng.set_angles(*angle)
    f, en = eng.utotal()
    ua = eng.angle_energy()
    orc = oracle_for(s, tab)
    orc.angle = angle
    orc.set_list(s.ro)
    fo, eo = orc.utotal(s.ro)
    say(f"  u%angle {ua:.15e} vs {eo['angle']:.15e}  force rel-RMS {rel_rms(f, fo):.3e}")
    assert eo["angle"] != 0.0   # negative where an odd power meets theta < eq
    assert abs(ua - eo["angle"]) <= 1e-12 * abs(eo["angle"])
    assert abs(en.tot - eo["tot"]) <= etol * abs(eo["tot"])
    assert abs((en.tot - en0.tot) - ua) <= 1e-9 * abs(en.tot)
    # FP64 spreading's atomics reorder sums between calls (~1e-16)
    assert abs(en.virial - en0.virial) <= 1e-12 * abs(en0.virial)
    assert rel_rms(f, fo) <= ftol
    fa, _ = O.uangle(s.boxlen, s.ro, s.chains, *angle)
    assert rel_rms(f - f0, fa * E.force_unit_fac()) <= 1e-8
    # switching the angles off restores the plain result
    eng.set_angles(0.0, 2, 180.0)
    f2, en2 = eng.utotal()
    assert eng.angle_energy() == 0.0
    assert abs(en2.tot - en0.tot) <= 1e-12 * abs(en0.tot)
    eng.close()


@pytest.mark.parametrize("mesh,order", [(32, 4), (48, 5), (40, 6), (24, 8), (18, 3), (36, 7)])
def test_spme_meshes_and_orders(mesh, order):
    """Mixed-radix meshes (2, 3, 5) and every supported spline order."""
    s = S.c2()
    s.mesh, s.order = mesh, order
    say("case: SPME mesh", mesh, "order", order)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, "fp64", tables=tab)
    f, en = eng.utotal()
    orc = oracle_for(s, tab)
    orc.set_list(s.ro)
    fo, eo = orc.utotal(s.ro)
    assert rel_rms(f, fo) <= 1e-10
    assert abs(en.rec - eo["rec"]) <= 1e-11 * abs(eo["rec"])
    eng.close()


# ------------------------------------------------------------------- MD
def test_velver_vs_reference_trajectory(golden):
    """Five MD steps from the reference's initial state (rodd = 0)."""
    z = golden["c1_std"]
    s = S.c1()
    s.ro = z["traj_pos"][0].copy()
    s.rod = z["traj_vel"][0].copy()
    say("case: VelVer 5 steps vs reference trajectory")
    eng = S.configure(s, "fp64")
    eng.build_list()
    for k in range(1, z["traj_pos"].shape[0]):
        eng.step_md(1, 0.002)
        ro, rod, _ = eng.download_state()
        d = ro - z["traj_pos"][k]
        d -= 50.0 * np.round(d / 50.0)
        assert np.abs(d).max() < 1e-9, k
        assert np.abs(rod - z["traj_vel"][k]).max() < 1e-6 * np.abs(z["traj_vel"][k]).max(), k
    eng.close()


def test_md_rebuilds_match_oracle():
    """Skin-test decisions (NList, Src/nlist.F90:597-616) over 40 steps."""
    s = S.c1()
    s.ewald = "none"
    say("case: MD rebuild schedule c1, 40 steps")
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, "fp64", tables=tab)
    eng.build_list()
    orc = oracle_for(s, tab, ewald="none")
    ro = s.ro.reshape(-1).copy()
    rod = s.rod.reshape(-1).copy()
    rodd = np.zeros_like(ro)
    drosum = np.zeros_like(ro)
    massip = s.massipt[s.iptpn - 1]
    orc.set_list(ro)
    nreb = 0
    for k in range(40):
        dro = O.velver_move(s.boxlen, 0.002, ro, rod, rodd)
        f, _ = orc.utotal(ro)
        O.velver_kick(0.002, massip, f.reshape(-1), rod, rodd)
        if O.skin_test(drosum, dro, s.drnlist):
            orc.set_list(ro)
            nreb += 1
    eng.step_md(40, 0.002)
    en = eng.energy()
    ro_g, _, _ = eng.download_state()
    d = ro_g.reshape(-1) - ro
    d -= s.boxlen * np.round(d / s.boxlen)
    assert en.rebuilds == nreb
    assert np.abs(d).max() < 1e-8
    eng.close()


# --------------------------------------------------------- error paths
def test_host_positions_follow_the_skin_rule():
    """Host-driven UTotal (upload, compute) keeps the list while the NList
    skin test holds and rebuilds when it fires; forces equal a freshly
    built engine's at the same positions (the pair set is the same; the
    partner order of a list built earlier may differ, so FP64 to 1e-13)."""
    s = S.c2()
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    say("case: c2 host uploads with the skin rule, fp64")
    eng = S.configure(s, "fp64", tables=tab)
    eng.utotal()
    r0 = eng.energy().rebuilds
    rng = np.random.default_rng(8)
    ro = s.ro.copy()
    # rigid shifts (no new close contacts) plus 0.02 A jitter; the 1.1 A
    # shift moves every particle by more than skin/2 = 1 A
    for step, shift in enumerate([0.05, 0.05, 1.1, 0.05]):
        ro = ro + shift * np.array([0.6, 0.64, 0.48]) + 0.02 * rng.uniform(-1, 1, ro.shape)
        ro -= s.boxlen * np.round(ro / s.boxlen)
        eng.upload_positions(ro)
        f, en = eng.utotal()
        fresh = S.configure(s, "fp64", tables=tab)
        fresh.upload_positions(ro)
        f2, en2 = fresh.utotal()
        fresh.close()
        say(f"  step {step}: rebuilds {en.rebuilds - r0}")
        assert rel_rms(f, f2) <= 1e-13 and abs(en.tot - en2.tot) <= 1e-13 * abs(en2.tot)
    # the small moves keep the list; the 1.1 A shift rebuilds it
    assert en.rebuilds - r0 == 1
    eng.close()


def test_overlap_is_reported():
    s = S.c1()
    s.ewald = "none"
    s.ro[1] = s.ro[0] + np.array([0.2, 0.0, 0.0])  # r2 = 0.04 < r2umin = 0.1
    s.ro = s.ro - 50.0 * np.round(s.ro / 50.0)
    say("case: overlap error")
    eng = S.configure(s, "fp64")
    with pytest.raises(E.MfeError) as ei:
        eng.utotal()
    assert ei.value.code == 2
    eng.close()


def test_capacity_is_reported():
    s = S.c1()
    say("case: neighbour capacity error")
    eng = S.configure(s, "mixed", maxnneigh=4)
    with pytest.raises(E.MfeError) as ei:
        eng.build_list()
    assert ei.value.code == 4
    eng.close()


def test_odd_mesh_is_rejected():
    s = S.c2()
    eng = E.ForceEngine(s.np, "mixed")
    eng.set_box(s.boxlen)
    eng.set_particles(s.iptpn, s.az, s.massipt)
    with pytest.raises(E.MfeError) as ei:
        eng.set_ewald("spm", s.alpha, 4, (33, 32, 32))
    assert ei.value.code == 5
    eng.close()


def test_charge_out_of_range_is_rejected():
    # the mixed spreading accumulates in 2^-32 fixed point; charges are
    # bounded (|q| < 2^16) and must be finite
    s = S.c2()
    eng = E.ForceEngine(s.np, "mixed")
    eng.set_box(s.boxlen)
    for bad in (70000.0, float("nan")):
        az = np.array(s.az, dtype=np.float64)
        az[3] = bad
        with pytest.raises(E.MfeError) as ei:
            eng.set_particles(s.iptpn, az, s.massipt)
        assert E.ERRORS[ei.value.code] == "MFE_E_BAD_PARAM"
    eng.set_particles(s.iptpn, s.az, s.massipt)
    eng.close()


def test_rebuild_barrier_give_up_recovers():
    """The single-engine list rebuild waits at grid barriers with a bounded
    number of polls.  With the bound set to 0 from the host, a rebuild that
    fires gives up at its first barrier: every block leaves, the call
    reports MFE_E_HIP, and the next call builds the list from scratch (cell
    counts and barrier cleared) -- forces then equal a fresh engine's at the
    same positions, and a later gated rebuild with the default bound works."""
    s = S.c2()
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    say("case: c2 rebuild barrier gives up (poll bound 0), fp64")
    eng = S.configure(s, "fp64", tables=tab)
    eng.utotal()
    shift = 1.1 * np.array([0.6, 0.64, 0.48])

    def moved(ro):
        ro = ro + shift
        return ro - s.boxlen * np.round(ro / s.boxlen)

    def check_fresh(ro, f, en):
        fresh = S.configure(s, "fp64", tables=tab)
        fresh.upload_positions(ro)
        f2, en2 = fresh.utotal()
        fresh.close()
        assert rel_rms(f, f2) <= 1e-13 and abs(en.tot - en2.tot) <= 1e-13 * abs(en2.tot)

    # host upload: the skin test fires, the rebuild gives up
    eng.set_sync_limit(0)
    ro1 = moved(s.ro)
    with pytest.raises(E.MfeError) as ei:
        eng.upload_positions(ro1)
    assert ei.value.code == 6
    eng.set_sync_limit(-1)
    f, en = eng.utotal()
    check_fresh(ro1, f, en)
    # a gated rebuild with the default bound after the recovery
    ro2 = moved(ro1)
    r0 = eng.energy().rebuilds
    eng.upload_positions(ro2)
    f, en = eng.utotal()
    assert en.rebuilds == r0 + 1
    check_fresh(ro2, f, en)
    # MD steps: a shift of 0.99 A keeps the list (1.98 A < skin 2 A); the
    # thermal motion of the first steps then fires the skin test inside the
    # step graph, whose rebuild gives up; the rest of the call leaves the
    # positions where they are and the call reports the error
    ro3 = ro2 + 0.99 * shift / 1.1
    ro3 -= s.boxlen * np.round(ro3 / s.boxlen)
    eng.upload_positions(ro3)
    assert eng.energy().rebuilds == r0 + 1
    eng.set_sync_limit(0)
    with pytest.raises(E.MfeError) as ei:
        eng.step_md(10, s.tstep)
    assert ei.value.code == 6
    ro4, _, _ = eng.download_state()
    d = ro4 - ro3
    d -= s.boxlen * np.round(d / s.boxlen)
    assert np.isfinite(ro4).all() and np.abs(d).max() < 0.5
    eng.set_sync_limit(-1)
    eng.upload_positions(ro4)
    f, en = eng.utotal()
    check_fresh(ro4, f, en)
    eng.step_md(5, s.tstep)
    eng.close()
