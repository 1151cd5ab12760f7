"""CPU-side checks of the product: the C ABI library loads and exports
everything include/mfe.h declares, the host table builder reproduces the
reference's tables bitwise, the Fortran binding compiles, and the seeded
systems are deterministic.  No GPU work is started here."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from molsim_amd import engine as E
from molsim_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "mfe.h")).read()
    return sorted(set(re.findall(r"\b(mfe_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(E.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(E.SYMBOLS) == syms


def test_abi_version():
    assert E.lib().mfe_abi_version() == 1


def test_pot_defaults_match_reference_literals():
    s = E.pot_defaults()
    assert s.utoltab == float(np.float32(1e-5))      # default-real 1e-5
    assert s.ljrcut == 2.0 ** float(np.float32(1.0) / np.float32(6.0))
    assert (s.umaxtab, s.fmaxtab, s.r2uminin, s.rumaxfac) == (20000.0, 20000.0, 0.1, 1.1)


@pytest.mark.parametrize("name,rcut", [("c1_std.npz", 15.0), ("tables_c2.npz", 18.0),
                                       ("tables_c3.npz", 12.0)])
def test_host_tables_bitwise_vs_reference(golden_dir, name, rcut):
    z = np.load(os.path.join(golden_dir, name))
    t = E.build_tables(z["zat"], z["sigat"], z["epsat"], rcut=rcut,
                       epsi_four_pi=float(z["epsi_four_pi"]), alpha=float(z["alpha"]))
    assert np.array_equal(t.ubuf, z["ubuf"])
    assert np.array_equal(t.iubuflow, z["iubuflow"])
    assert np.array_equal(t.nugrid, z["nugrid"])
    assert np.array_equal(t.r2umin, z["r2umin"])


def test_units():
    import oracle as O
    assert E.epsi0_four_pi() == O.epsi0_four_pi()
    assert E.force_unit_fac() == O.force_unit_fac()


def test_systems_are_seeded():
    a, b = S.c1(), S.c1()
    assert np.array_equal(a.ro, b.ro) and np.array_equal(a.rod, b.rod)
    assert np.abs(a.ro).max() <= a.boxlen / 2
    c = S.c2()
    assert c.np == 6000 and c.boxlen == 120.0
    assert abs(c.az.sum()) < 1e-12
    # bonded beads are 5 A apart
    ro = c.ro
    for f, n in c.chains[:3]:
        d = ro[f:f + n - 1] - ro[f - 1:f + n - 2]
        d -= c.boxlen * np.round(d / c.boxlen)
        assert np.allclose(np.linalg.norm(d, axis=1), 5.0)


def test_c3_composition():
    s = S.c3(n_units=400)
    assert s.np == 2000
    assert abs(s.az.sum()) < 1e-12
    big = 642.8697667596732
    assert abs(S.c3.__defaults__[1] - 40000) == 0
    assert abs((2 * 40000 / (0.5 * S.AVNO_PER_A3_MOLAR)) ** (1 / 3) - big) < 1e-9


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="no Fortran compiler")
def test_fortran_binding_compiles(tmp_path):
    mod = os.path.join(ROOT, "molsim_amd", "fortran", "mfe_mod.F90")
    r = subprocess.run(["amdflang", "-c", mod, "-o", str(tmp_path / "m.o")],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_autotune_mesh_rule():
    from molsim_amd import autotune as A
    assert A.even_ceil(15.2) == 16 and A.even_ceil(16.0) == 16 and A.even_ceil(17.0) == 18
    # alpha*h <= target: s >= alpha L / ah, even, at least 16
    assert [n for n in range(16, 50, 2) if A.smooth(n)] == [16, 18, 20, 24, 28, 30, 32, 36,
                                                             40, 42, 48]
    for L, a, ah in [(116.1, 0.25, 1.0), (298.7, 0.25, 1.3), (74.6, 0.3, 2.0),
                     (300.0, 0.3, 0.45)]:
        s = A.mesh_for(L, a, ah)
        assert s % 2 == 0 and s >= 16 and a * L / s <= ah + 1e-12 and A.smooth(s)
        # no smaller even 7-smooth mesh reaches the spacing
        assert all(a * L / m > ah for m in range(16, s, 2) if A.smooth(m))
    sysm = S.c4_replica(0)
    t = A.with_params(sysm, 0.2, 5, 30)
    assert (t.alpha, t.order, t.mesh) == (0.2, 5, 30) and abs(t.rcut - 15.0) < 1e-12


def test_ewald_parameters_match_reference(golden_dir):
    """iewaldopt 0-4 (Src/energy.F90:6644-6690, 7052-7124): the parameters
    the reference settles on for the C1 system (1000 ions, L = 50 A,
    rc = 15 A), bit for bit (tests/golden/ewald_opt.npz)."""
    z = np.load(os.path.join(golden_dir, "ewald_opt.npz"))
    for (opt, lq, ncut), (tol, ared, alpha, rcut, q2, lshort, efp) in zip(z["ints"], z["reals"]):
        inp = dict(q2sum=q2, boxlenshort=lshort, epsi_four_pi=efp, lq2sum=int(lq))
        if opt in (0, 1):
            inp.update(ualphared=3.0, rcut=15.0, ncut=8)
        elif opt == 2:
            inp.update(uewaldtol=tol, ualpha=0.21)
        elif opt == 3:
            inp.update(uewaldtol=tol, rcut=15.0)
        else:
            inp.update(uewaldtol=tol, ncut=int(ncut))
        r = E.ewald_parameters(int(opt), **inp)
        assert r["ualpha"] == alpha, (opt, r["ualpha"], alpha)
        assert r["rcut"] == rcut and r["ncut"] == ncut, (opt, r)
        assert r["uewaldtol"] == tol and (opt == 0 or r["ualphared"] == ared), (opt, r)
    # SPME takes options 0 and 3 only, and 3 gives the same alpha as std
    s = E.ewald_parameters(3, q2sum=1.0, boxlenshort=50.0, epsi_four_pi=z["reals"][0][6],
                           spme=True, uewaldtol=1e-3, rcut=15.0)
    assert s["ualpha"] == z["reals"][3][2]
    with pytest.raises(E.MfeError):
        E.ewald_parameters(2, q2sum=1.0, boxlenshort=50.0, epsi_four_pi=17.7, spme=True,
                           uewaldtol=1e-3, ualpha=0.2)
    with pytest.raises(E.MfeError):
        E.ewald_parameters(3, q2sum=1.0, boxlenshort=50.0, epsi_four_pi=17.7, rcut=15.0)


def test_host_bd_factors_equal_the_pinned_oracle():
    """molsim_amd.engine.bd_factors (what a host passes to mfe_set_bd) is the
    oracle's orc_bd_factors bitwise; the oracle's is pinned against the
    reference's BDStep displacements (test_oracle_golden.py)."""
    import oracle as O
    for ts, temp, d in [(0.005, 298.0, [1.0, 2.0, 1.33]), (0.002, 300.0, [0.7]),
                        (0.01, 350.0, [1.33, 2.03])]:
        a = E.bd_factors(ts, temp, d)
        b = O.bd_factors(ts, temp, d)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
