"""The FP64 oracle against golden vectors produced by the MOLSIM reference.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py
from the amdflang build of the unmodified reference (tables, UTotal
forces/energies/virial, Verlet lists, MD trajectory, Erf2,
CardinalBSpline).  Every comparison here is bitwise.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import oracle as O


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


# ------------------------------------------------------------------ math
def test_erf2_bitwise(golden_dir):
    z = load(golden_dir, "math.npz")
    got = O.erf2(z["args"])
    assert np.array_equal(got, z["erf2"])


@pytest.mark.parametrize("order", range(3, 9))
def test_cardinal_bspline_bitwise(golden_dir, order):
    z = load(golden_dir, "math.npz")
    for i, x in enumerate(z["args"]):
        k, v, d = O.cardinal_bspline(order, x)
        assert k == z[f"kmax{order}"][i]
        assert np.array_equal(v, z[f"val{order}"][i])
        assert np.array_equal(d, z[f"der{order}"][i])


def test_pot_1_6_12_bitwise(golden_dir):
    """Pot_1_6_12('calc') and CalcEwaldUfac on 600 radii, 3 type pairs."""
    z = load(golden_dir, "math.npz")
    efp = O.epsi0_four_pi() / 78.4
    pairs = [(0, 0), (0, 1), (1, 1)]
    for i, r in enumerate(z["pot_r"]):
        for k, (it, jt) in enumerate(pairs):
            u = O.pot_1_6_12([1.0, -1.0], [4.0, 4.0], [2.4777, 2.4777], it, jt, r,
                             epsi_four_pi=efp, alpha=float(z["pot_alpha"]))
            assert np.array_equal(u, z["pot_1_6_12"][i, k]), (r, it, jt)


# ---------------------------------------------------------------- tables
@pytest.mark.parametrize("name,rcut", [("c1_std.npz", 15.0), ("tables_c2.npz", 18.0),
                                       ("tables_c3.npz", 12.0)])
def test_tables_bitwise(golden_dir, name, rcut):
    z = load(golden_dir, name)
    efp = float(z["epsi_four_pi"])
    assert efp == O.epsi0_four_pi() / 78.4
    t = O.build_tables(z["zat"], z["sigat"], z["epsat"], rcut=rcut,
                       epsi_four_pi=efp, alpha=float(z["alpha"]))
    assert np.array_equal(t.iubuflow, z["iubuflow"])
    assert np.array_equal(t.nugrid, z["nugrid"])
    assert np.array_equal(t.r2umin, z["r2umin"])
    assert t.ubuf.size == z["ubuf"].size
    assert np.array_equal(t.ubuf, z["ubuf"])


# ---------------------------------------------------------------- UTotal
def _c1_oracle(z, nlist="vlist", lsurf=False):
    efp = float(z["epsi_four_pi"])
    t = O.build_tables(z["zat"], z["sigat"], z["epsat"], rcut=15.0,
                       epsi_four_pi=efp, alpha=0.2)
    iptpn = np.repeat([1, 2], z["nppt"])
    az = np.repeat(z["zat"], z["nppt"])
    return O.Oracle(50.0, iptpn, az, t, 15.0, ewald="std", alpha=0.2,
                    epsi_four_pi=efp, ncut=8, lsurf=lsurf, nlist=nlist,
                    drnlist=2.0)


def test_utotal_vlist_std_ewald_bitwise(golden_dir):
    z = load(golden_dir, "c1_std.npz")
    orc = _c1_oracle(z)
    pos = z["probe_pos"]
    orc.set_list(pos)
    assert np.array_equal(orc.nn, z["ref_nneighpn"])
    flat = np.concatenate([orc.jl[i, :orc.nn[i]] for i in range(orc.np)]).astype("<i4")
    assert hashlib.sha256(flat.tobytes()).hexdigest() == str(z["ref_jpnlist_sha256"])
    f, e = orc.utotal(pos)
    assert np.array_equal(f, z["ref_forceo"])
    assert e["tot"] == z["ref_u_tot"]
    assert e["twob0"] == z["ref_u_twob"][0]
    assert np.array_equal(e["twob"], z["ref_u_twob"][1:])
    assert e["rec"] == z["ref_u_rec"]
    assert e["virial"] == z["ref_virial"]


def test_vlist_cells_equals_bruteforce(golden_dir):
    z = load(golden_dir, "c1_std.npz")
    pos = z["probe_pos"]
    nn1, jl1 = O.set_vlist_md(50.0, pos, 17.0, 1000)
    nn2, jl2 = O.set_vlist_md(50.0, pos, 17.0, 1000, cells=True)
    assert np.array_equal(nn1, nn2)
    assert np.array_equal(jl1, jl2)


def test_utotal_celllist_surface_bitwise(golden_dir):
    z = load(golden_dir, "c1_std.npz")
    zc = load(golden_dir, "c1_cell_surf.npz")
    orc = _c1_oracle(z, nlist="cell", lsurf=True)
    pos = zc["probe_pos"]
    orc.set_list(pos)
    f, e = orc.utotal(pos)
    assert np.array_equal(f, zc["ref_forceo"])
    assert e["tot"] == zc["ref_u_tot"]
    assert e["rec"] == zc["ref_u_rec"]
    assert e["virial"] == zc["ref_virial"]


def test_utotal_chains_bonds_bitwise(golden_dir):
    z = load(golden_dir, "chains.npz")
    efp = float(z["epsi_four_pi"])
    t = O.build_tables(z["zat"], z["sigat"], z["epsat"], rcut=20.0,
                       epsi_four_pi=efp, alpha=float(z["alpha"]))
    assert np.array_equal(t.ubuf, z["ubuf"])
    iptpn = np.repeat([1, 2], z["nppt"])
    az = np.repeat(z["zat"], z["nppt"])
    nper = int(z["nper"])
    chains = [(1 + c * nper, nper) for c in range(int(z["nchain"]))]
    orc = O.Oracle(60.0, iptpn, az, t, 20.0, ewald="std", alpha=float(z["alpha"]),
                   epsi_four_pi=efp, ncut=5, chains=chains, bond=tuple(z["bond"]))
    pos = z["probe_pos"]
    orc.set_list(pos)
    f, e = orc.utotal(pos)
    assert e["bond"] == z["ref_u_bond"]
    assert np.array_equal(f, z["ref_forceo"])
    assert e["tot"] == z["ref_u_tot"]
    assert e["virial"] == z["ref_virial"]


@pytest.mark.parametrize("a", [0, 1])
def test_utotal_chains_angles_bitwise(golden_dir, a):
    """The chain case with bond angles on (UAngle, Src/energy.F90:5409-5481),
    p = 2 and p = 3, against the reference run at the same positions
    (tests/golden/chain_angles.npz, drawn by the reference probe)."""
    z = load(golden_dir, "chains.npz")
    za = load(golden_dir, "chain_angles.npz")
    ka, pa, eqa = za["angles"][a]
    efp = float(z["epsi_four_pi"])
    t = O.build_tables(z["zat"], z["sigat"], z["epsat"], rcut=20.0,
                       epsi_four_pi=efp, alpha=float(z["alpha"]))
    iptpn = np.repeat([1, 2], z["nppt"])
    az = np.repeat(z["zat"], z["nppt"])
    nper = int(z["nper"])
    chains = [(1 + c * nper, nper) for c in range(int(z["nchain"]))]
    orc = O.Oracle(60.0, iptpn, az, t, 20.0, ewald="std", alpha=float(z["alpha"]),
                   epsi_four_pi=efp, ncut=5, chains=chains, bond=tuple(z["bond"]),
                   angle=(float(ka), int(pa), float(eqa)))
    pos = z["probe_pos"]
    orc.set_list(pos)
    f, e = orc.utotal(pos)
    # the reference's u%angle is its u%tot less the terms without angles
    ref_angle = za[f"u_tot{a}"] - z["ref_u_tot"]
    assert e["angle"] != 0.0
    assert abs(e["angle"] - ref_angle) <= 1e-12 * abs(za[f"u_tot{a}"])
    assert e["tot"] == za[f"u_tot{a}"]
    assert e["bond"] == za[f"u_bond{a}"]
    assert e["virial"] == za[f"virial{a}"]
    assert np.array_equal(f, za[f"forceo{a}"])


def test_velver_trajectory_bitwise(golden_dir):
    """Five VelVer steps (Src/md.F90:252-410) with the skin-test rebuilds.

    rodd starts at zero: MDDriver(iReadInput) zeroes it and nothing evaluates
    forces before the first step (Src/md.F90:130-149), so the first kick
    uses only the new acceleration."""
    z = load(golden_dir, "c1_std.npz")
    orc = _c1_oracle(z)
    massip = 1.0 / np.repeat(z["masses"], z["nppt"])
    ro = z["traj_pos"][0].copy().reshape(-1)
    rod = z["traj_vel"][0].copy().reshape(-1)
    orc.set_list(ro)
    rodd = np.zeros_like(ro)
    drosum = np.zeros_like(ro)
    dt = float(z["tstep"])
    for k in range(1, z["traj_pos"].shape[0]):
        drostep = O.velver_move(50.0, dt, ro, rod, rodd)
        f, _ = orc.utotal(ro)
        O.velver_kick(dt, massip, f.reshape(-1), rod, rodd)
        assert np.array_equal(ro.reshape(-1, 3), z["traj_pos"][k]), k
        assert np.array_equal(f, z["traj_forceo"][k]), k
        assert np.array_equal(rod.reshape(-1, 3), z["traj_vel"][k]), k
        if O.skin_test(drosum, drostep, 2.0):
            orc.set_list(ro)


def test_gaurandom_and_random_streams_bitwise(golden_dir):
    """GauRandom and Random (Src/mollib.F90:2689-2766) from fixed seeds, as
    the reference probe drew them after the C1 setup.  GauRandom keeps one
    deviate between calls (a SAVEd flag), so the setup may leave one
    pending: the stream is compared from the first deviate of the seed,
    which is at probe index 0 or 1."""
    z = load(golden_dir, "rng.npz")
    L = O.lib()
    L.orc_random.restype = C.c_double

    class RandomState(C.Structure):
        _fields_ = [("ix", C.c_int32), ("iy", C.c_int32), ("am", C.c_double)]

    L.orc_random.argtypes = [C.POINTER(C.c_int32), C.POINTER(RandomState)]
    g_ref = z["gaurandom"]
    seed = C.c_int32(int(z["iseed"]))
    st = O.GauState(0, 0.0)
    g = np.array([L.orc_gaurandom(C.byref(seed), C.byref(st)) for _ in range(g_ref.size)])
    off = 0 if g[0] == g_ref[0] else 1
    assert np.array_equal(g[:g_ref.size - off], g_ref[off:])
    r_ref = z["random"]
    idum = C.c_int32(int(z["idum"]))
    rs = RandomState(-1, -1, 0.0)
    r = np.array([L.orc_random(C.byref(idum), C.byref(rs)) for _ in range(r_ref.size)])
    assert np.array_equal(r, r_ref)


def test_bd_step_bitwise(golden_dir):
    """BDStep (Src/bd.F90:120-155) against the reference run of four steps
    (bd.npz, drawn by the reference probe): fac1/fac2 from the unit formula
    (:136-139), GauRandom's deviates drawn from the recorded iseed in the
    reference's order (x, y, z per particle, 297 per step, so the spare
    deviate carries across steps), the displacement and PBC -- every
    displacement and position bitwise, given the reference's forceo."""
    z = load(golden_dir, "bd.npz")
    fac1, fac2 = O.bd_factors(float(z["tstep"]), float(z["temp"]), z["dcoeff"])
    L = O.lib()
    seed = C.c_int32(int(z["iseed"]))
    st = O.GauState(0, 0.0)
    iptpn = z["iptpn"]
    nstep, np_ = z["forceo"].shape[:2]
    ro = z["ro_before"][0].copy()
    for k in range(nstep):
        assert np.array_equal(ro, z["ro_before"][k]), k
        noise = np.array([L.orc_gaurandom(C.byref(seed), C.byref(st))
                          for _ in range(3 * np_)])
        dro = O.bd_step(float(z["boxlen"]), iptpn, fac1, fac2,
                        z["forceo"][k].reshape(-1), noise, ro.reshape(-1))
        assert np.array_equal(dro.reshape(-1, 3), z["drostep"][k]), k
        assert np.array_equal(ro, z["ro_after"][k]), k
        ro = z["ro_before"][k + 1].copy() if k + 1 < nstep else ro
    # the displacement is not dominated by either term alone
    f_part = np.abs(fac1[iptpn - 1][:, None] * z["forceo"][0]).mean()
    n_part = np.abs(z["drostep"][0]).mean()
    assert 1e-4 < f_part / n_part < 1e2


def test_bd_factors_closed_form():
    """fac1 = dt D / (R T) and fac2 = sqrt(2 D dt), in A and kJ/(mol A):
    D = 1e-9 m^2/s = 0.1 A^2/ps; R T in kJ/mol."""
    f1, f2 = O.bd_factors(0.005, 298.0, [1.0, 2.0])
    rt = 8.314472e-3 * 298.0          # kJ/mol
    d = np.array([1.0, 2.0]) * 0.1    # A^2/ps
    # forceo is in g A/(mol ps^2) = 1e-2 kJ/(mol A)
    assert np.allclose(f1, 0.005 * d / rt * 1e-2, rtol=1e-13)
    assert np.allclose(f2, np.sqrt(2 * d * 0.005), rtol=1e-13)
