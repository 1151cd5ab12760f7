"""Seeded synthetic systems for the benchmark and parity configurations.

Every generator is deterministic for its seed.  Units follow MOLSIM: A, ps,
g/mol, kJ/mol.  Assumptions the configurations leave open are recorded in
the returned System's `notes`.

  c1  1000 monovalent ions, L = 50 A, WCA + standard Ewald (rc 15, kmax 8)
  c2  40 chains x 50 beads (q = -1) + 2000 counterions + 1000 salt pairs,
      L = 120 A, SPME 64^3, order 4, rc 18
  c3  200k ions (40k +1, 40k +2, 120k -1), 0.5 M, SPME 128^3, order 6,
      rc 12, skin 2
  c4  one replica of the batched BD set (N 500-4000, 0.1-1.0 M)
  c5  C2 composition scaled to 64k particles, SPME 96^3
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import engine as E

KT298 = 2.4777          # kJ/mol at 298 K (1 kT)
AVNO_PER_A3_MOLAR = 6.02214179e23 / 1e27  # particles per A^3 per mol/L
GAS_CONSTANT = 8.314472


@dataclass
class System:
    name: str
    boxlen: float
    ro: np.ndarray            # [np, 3] in [-L/2, L/2]
    rod: np.ndarray           # [np, 3] A/ps
    iptpn: np.ndarray         # [np] 1-based types
    zat: np.ndarray           # [npt]
    sigat: np.ndarray
    epsat: np.ndarray
    massat: np.ndarray
    rcut: float
    alpha: float
    ewald: str = "spm"
    order: int = 4
    mesh: int = 32
    ncut: int = 8
    lsurf: bool = False
    drnlist: float = 2.0
    tstep: float = 0.002
    chains: list = field(default_factory=list)   # [(first 1-based, length)]
    bond: tuple = (0.0, 2, 0.0)                   # namelist (k, p, eq)
    notes: str = ""

    @property
    def np(self):
        return self.ro.shape[0]

    @property
    def az(self):
        return self.zat[self.iptpn - 1]

    @property
    def massipt(self):
        return 1.0 / self.massat


def _uniform_no_overlap(rng, n, L, dmin, fixed=None):
    """n uniform points in the periodic box with all distances >= dmin
    (also to the `fixed` points), by redrawing offenders."""
    from scipy.spatial import cKDTree
    pts = rng.uniform(-L / 2, L / 2, (n, 3))
    base = np.zeros((0, 3)) if fixed is None else fixed
    for _ in range(200):
        allp = np.concatenate([base, pts]) + L / 2
        allp = np.mod(allp, L)
        tree = cKDTree(allp, boxsize=L)
        pairs = tree.query_pairs(dmin, output_type="ndarray")
        nb = base.shape[0]
        bad = set()
        for a, b in pairs:
            if b >= nb:
                bad.add(b - nb)
            elif a >= nb:
                bad.add(a - nb)
        if not bad:
            return pts
        idx = np.fromiter(bad, dtype=np.int64)
        pts[idx] = rng.uniform(-L / 2, L / 2, (idx.size, 3))
    raise RuntimeError("could not place particles without overlap")


def _maxwell(rng, masses):
    sig = np.sqrt(GAS_CONSTANT * 298.0 / (masses / 1000.0)) * 0.01  # A/ps
    return rng.standard_normal((masses.size, 3)) * sig[:, None]


def _wrap(x, L):
    return x - L * np.round(x / L)


def c1(seed=1) -> System:
    rng = np.random.default_rng(seed)
    L = 50.0
    iptpn = np.repeat([1, 2], [500, 500]).astype(np.int32)
    massat = np.array([23.0, 35.4])
    ro = _uniform_no_overlap(rng, 1000, L, 0.8 * 4.0)
    rod = _maxwell(rng, massat[iptpn - 1])
    return System("c1", L, ro, rod, iptpn, np.array([1.0, -1.0]), np.array([4.0, 4.0]),
                  np.array([KT298, KT298]), massat, rcut=15.0, alpha=3.0 / 15.0,
                  ewald="std", ncut=8, lsurf=False, drnlist=2.0, tstep=0.002,
                  notes="masses 23.0/35.4 and skin 2 A assumed")


def _chains(rng, nchain, nper, bondlen, L, dmin):
    """Seeded random walks with bond length `bondlen`, rejecting steps that
    come closer than dmin to any bead already placed."""
    from scipy.spatial import cKDTree
    placed = np.zeros((0, 3))
    out = []
    def clear(p, others):
        if others.shape[0] == 0:
            return True
        d = _wrap(others - p, L)
        return bool((np.einsum("ij,ij->i", d, d) >= dmin * dmin).all())

    for _ in range(nchain):
        for _try in range(100):
            first = rng.uniform(-L / 2, L / 2, 3)
            while not clear(first, placed):
                first = rng.uniform(-L / 2, L / 2, 3)
            chain = [first]
            ok = True
            for _s in range(1, nper):
                for _t in range(200):
                    v = rng.standard_normal(3)
                    p = chain[-1] + bondlen * v / np.linalg.norm(v)
                    p = _wrap(p, L)
                    others = np.concatenate([placed, np.array(chain[:-1])]) if chain[:-1] else placed
                    if others.shape[0] == 0:
                        break
                    d = _wrap(others - p, L)
                    if (np.einsum("ij,ij->i", d, d) >= dmin * dmin).all():
                        break
                else:
                    ok = False
                    break
                chain.append(p)
            if ok:
                break
        chain = np.array(chain)
        out.append(chain)
        placed = np.concatenate([placed, chain])
    del cKDTree
    return np.concatenate(out)


def c2(seed=2, scale=1) -> System:
    """Polyelectrolyte solution; scale multiplies every species count and
    the volume (scale 1: N = 6000, L = 120 A)."""
    rng = np.random.default_rng(seed)
    nchain, nper = 40 * scale, 50
    ncounter, nsalt = 2000 * scale, 1000 * scale
    L = 120.0 * scale ** (1.0 / 3.0)
    beads = _chains(rng, nchain, nper, 5.0, L, 0.8 * 4.0)
    ions = _uniform_no_overlap(rng, ncounter + 2 * nsalt, L, 0.8 * 4.0, fixed=beads)
    ro = np.concatenate([beads, ions])
    iptpn = np.concatenate([np.full(beads.shape[0], 1), np.full(ncounter + nsalt, 2),
                            np.full(nsalt, 3)]).astype(np.int32)
    massat = np.array([30.0, 23.0, 35.4])
    rod = _maxwell(rng, massat[iptpn - 1])
    chains = [(1 + c * nper, nper) for c in range(nchain)]
    k = 0.4 * KT298  # namelist k; u = (k/2)(r - 5)^2 (Src/potential.F90:3561)
    mesh = 64 if scale == 1 else 96
    return System("c2" if scale == 1 else "c5", L, _wrap(ro, L), rod, iptpn,
                  np.array([-1.0, 1.0, -1.0]), np.array([4.0] * 3),
                  np.array([KT298] * 3), massat, rcut=18.0, alpha=3.0 / 18.0,
                  ewald="spm", order=4, mesh=mesh, drnlist=2.0, tstep=0.002,
                  chains=chains, bond=(k, 2, 5.0),
                  notes="WCA sigma 4 for all pairs; bead mass 30 assumed; "
                        "bond k = 0.4 kT/A^2 read as MOLSIM's namelist k")


def c3(seed=3, n_units=40000, mesh=128) -> System:
    """Large-scale SPME throughput case: n_units each of the 1:1 and 2:1
    salts at 0.25 M each (0.5 M in all) -> L = 642.8 A at 40000 units."""
    rng = np.random.default_rng(seed)
    n1, n2, nneg = n_units, n_units, 3 * n_units
    conc = 0.5
    L = (2 * n_units / (conc * AVNO_PER_A3_MOLAR)) ** (1.0 / 3.0)
    n = n1 + n2 + nneg
    ro = _uniform_no_overlap(rng, n, L, 0.8 * 4.0)
    iptpn = np.concatenate([np.full(n1, 1), np.full(n2, 2), np.full(nneg, 3)]).astype(np.int32)
    massat = np.array([22.99, 40.08, 35.45])
    rod = _maxwell(rng, massat[iptpn - 1])
    return System("c3", L, ro, rod, iptpn, np.array([1.0, 2.0, -1.0]),
                  np.array([4.0] * 3), np.array([KT298] * 3), massat, rcut=12.0,
                  alpha=3.0 / 12.0, ewald="spm", order=6, mesh=mesh, drnlist=2.0,
                  tstep=0.002,
                  notes="0.5 M read as 0.25 M NaCl-like + 0.25 M CaCl2-like "
                        f"salt (L = {L:.1f} A); masses 22.99/40.08/35.45")


def c4_replica(r, base_seed=1000, mesh=None, order=4) -> System:
    """Replica r of the batched BD set: N ~ U{500..4000} (even), c ~ U[0.1,1]."""
    rng = np.random.default_rng(base_seed + r)
    n = 2 * int(rng.integers(250, 2001))
    c = float(rng.uniform(0.1, 1.0))
    L = ((n // 2) / (c * AVNO_PER_A3_MOLAR)) ** (1.0 / 3.0)
    ro = _uniform_no_overlap(rng, n, L, 0.8 * 4.0)
    iptpn = np.repeat([1, 2], [n // 2, n // 2]).astype(np.int32)
    massat = np.array([23.0, 35.4])
    rc = min(12.0, 0.45 * L)
    m = mesh or int(max(16, min(48, 2 * round(L / 2.0 / 2))))
    return System(f"c4_{r}", L, ro, np.zeros_like(ro), iptpn, np.array([1.0, -1.0]),
                  np.array([4.0, 4.0]), np.array([KT298, KT298]), massat, rcut=rc,
                  alpha=3.0 / rc, ewald="spm", order=order, mesh=m, drnlist=2.0,
                  tstep=0.005, notes=f"replica {r}: N={n}, c={c:.3f} M")


def configure(sys: System, precision="mixed", device=0, tables: E.Tables | None = None,
              nlist="vlist", maxnneigh=0) -> E.ForceEngine:
    """A ForceEngine holding `sys`, set up like MOLSIM's setup stages."""
    efp = E.epsi0_four_pi() / 78.4
    if tables is None:
        tables = E.build_tables(sys.zat, sys.sigat, sys.epsat, rcut=sys.rcut,
                                epsi_four_pi=efp, alpha=sys.alpha)
    eng = E.ForceEngine(sys.np, precision, device)
    eng.set_box(sys.boxlen)
    eng.set_particles(sys.iptpn, sys.az, sys.massipt)
    eng.set_tables(tables, sys.rcut)
    eng.set_ewald(sys.ewald, sys.alpha, sys.order, (sys.mesh,) * 3, sys.ncut, False,
                  sys.lsurf, efp)
    eng.set_neighbor(nlist, sys.drnlist, maxnneigh)
    if sys.chains:
        eng.set_chains(sys.chains, sys.bond[0], sys.bond[1], sys.bond[2])
    eng.upload_positions(sys.ro)
    eng.upload_velocities(sys.rod)
    return eng
