"""Shared test helpers: the oracle counterpart of a molsim_amd System."""
import numpy as np

import oracle as O
from molsim_amd import engine as E


def efp():
    return E.epsi0_four_pi() / 78.4


def oracle_for(sysm, tables: E.Tables, nlist="vlist", lsurf=None, ewald=None):
    t = O.Tables(tables.ubuf, tables.iubuflow, tables.nugrid, tables.r2umin, None)
    return O.Oracle(sysm.boxlen, sysm.iptpn, sysm.az, t, sysm.rcut,
                    ewald=ewald or sysm.ewald, alpha=sysm.alpha, epsi_four_pi=efp(),
                    ncut=sysm.ncut, lsurf=sysm.lsurf if lsurf is None else lsurf,
                    order=sysm.order, mesh=(sysm.mesh,) * 3, nlist=nlist,
                    drnlist=sysm.drnlist, chains=sysm.chains or None, bond=sysm.bond,
                    maxnneigh=min(sysm.np - 1, 2048))


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def owned_pairs(nn, jl):
    """Set of (ip, jp) of a SetVListMD-style list."""
    out = set()
    for i in range(nn.size):
        for j in jl[i, :nn[i]]:
            out.add((i + 1, int(j)))
    return out
