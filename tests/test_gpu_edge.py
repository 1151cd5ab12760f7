"""Edge cases of the SPME path and the lists against the oracle.

A strongly clustered system (all particles in one corner: most spreading
bricks empty, a few holding hundreds of particles, stencils that wrap the
periodic mesh) and a two-particle system.  Tolerances as elsewhere: FP64
forces rel-RMS <= 1e-10 and energies <= 1e-11 relative against the
oracle; mixed <= 1e-5 and <= 1e-7.
"""
import numpy as np
import pytest

from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp, oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def clustered(seed=9):
    rng = np.random.default_rng(seed)
    L, n = 120.0, 800
    # a jittered lattice (spacing 4.4 A) filling a 40 A cube around a box
    # corner, so the cluster wraps the periodic mesh
    g = np.arange(10) * 4.4 - 19.8
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[rng.permutation(pts.shape[0])[:n]] + rng.uniform(-0.4, 0.4, (n, 3))
    ro = pts + 0.5 * L
    ro -= L * np.round(ro / L)
    iptpn = np.repeat([1, 2], n // 2).astype(np.int32)
    return S.System("clustered", L, ro, np.zeros_like(ro), iptpn, np.array([1.0, -1.0]),
                    np.array([4.0, 4.0]), np.array([S.KT298] * 2), np.array([23.0, 35.4]),
                    rcut=18.0, alpha=3.0 / 18.0, ewald="spm", order=5, mesh=64,
                    drnlist=2.0)


def pair_system():
    ro = np.array([[1.0, 2.0, 3.0], [8.0, -1.0, 2.5]])
    return S.System("pair", 60.0, ro, np.zeros_like(ro), np.array([1, 2], np.int32),
                    np.array([1.0, -1.0]), np.array([4.0, 4.0]), np.array([S.KT298] * 2),
                    np.array([23.0, 35.4]), rcut=20.0, alpha=3.0 / 20.0, ewald="spm",
                    order=4, mesh=32, drnlist=2.0)


@pytest.mark.parametrize("build", [clustered, pair_system])
@pytest.mark.parametrize("prec,ftol,etol", [("fp64", 1e-10, 1e-11), ("mixed", 1e-5, 1e-7)])
def test_edge_systems_vs_oracle(build, prec, ftol, etol):
    s = build()
    print("case:", s.name, "N", s.np, prec, flush=True)
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    # the cluster's density is far above the box average, which the default
    # list capacity is sized for
    eng = S.configure(s, prec, tables=tab, maxnneigh=s.np - 1)
    f, en = eng.utotal()
    orc = oracle_for(s, tab)
    orc.set_list(s.ro)
    fo, eo = orc.utotal(s.ro)
    r = rel_rms(f, fo)
    print(f"  force rel-RMS {r:.2e}, rec {en.rec:.12e} vs {eo['rec']:.12e}", flush=True)
    assert r <= ftol
    assert abs(en.rec - eo["rec"]) <= etol * abs(eo["rec"])
    assert abs(en.tot - eo["tot"]) <= etol * abs(eo["tot"])
    eng.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_binning_alone_then_utotal(prec):
    """Binnings that no spreading follows (the stage timings run the count
    and scatter kernels alone, which then clear the brick counts
    themselves), and spreadings repeated on one binning, leave the next
    UTotal as a fresh engine computes it: bitwise on the mixed path, to
    rounding on the FP64 path (its spreading adds FP64 in LDS in a
    run-dependent order).  Brick counts left over would make the scatter
    report MFE_E_HIP."""
    s = S.c4_replica(7)
    print("case: bin-only stages, then UTotal, electrolyte N", s.np, prec, flush=True)
    a, b = S.configure(s, prec), S.configure(s, prec)
    for st, n in (("bin", 3), ("spread", 2), ("bin", 2), ("fft_x", 1), ("bin", 1)):
        a.time_stage(st, nrep=n)
    fa, ea = a.utotal()
    fb, eb = b.utotal()
    d = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fb)))
    print(f"  max |dF|/max|F| {d:.2e}, E {ea.tot:.15e} vs {eb.tot:.15e}", flush=True)
    if prec == "mixed":
        assert np.array_equal(fa, fb) and ea.tot == eb.tot
    else:
        assert d <= 1e-12 and abs(ea.tot - eb.tot) <= 1e-13 * abs(eb.tot)
    a.close()
    b.close()
