"""torch.autograd bridge of the engine's parameter adjoint (no reference
counterpart).

`ModelEnergy.apply(ro, zat, sigat, epsat, knots, model)` returns the model
energy of one configuration -- the pair model the tables interpolate
(Pot_1_6_12 with exact erfc, Lorentz-Berthelot mixing, WCA cut) plus the
additive spline term, SPME with its self term, and bonds -- as a
differentiable torch scalar; backward returns the engine's adjoint:
dE/dro = -F (kJ/mol/A), dE/dzat, dE/dsigat, dE/depsat and dE/dc.

The engine computes in FP64 on the device; torch only carries the numbers
(parameters move to the host for the table build, which MOLSIM also does
on the host).
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine as E
from . import systems as S


class PairModel:
    """One system (box, types, Ewald setup, list) held by a device engine;
    `evaluate` re-tabulates the pair potential for the given parameters and
    runs the adjoint at the given positions."""

    def __init__(self, sysm: S.System, nknot=64, knot_lo=None, precision="fp64", device=0):
        self.s = sysm
        self.nknot = nknot
        self.knot_lo = (0.8 * float(min(sysm.sigat))) if knot_lo is None else knot_lo
        self.knot_hi = sysm.rcut
        self.efp = E.epsi0_four_pi() / 78.4
        self.eng = S.configure(sysm, precision)

    @property
    def nptpt(self):
        n = self.s.massat.size
        return n * (n + 1) // 2

    def evaluate(self, ro, zat, sigat, epsat, knots):
        s, eng = self.s, self.eng
        zat, sigat, epsat = (np.asarray(v, float) for v in (zat, sigat, epsat))
        tab = E.build_tables(zat, sigat, epsat, rcut=s.rcut, epsi_four_pi=self.efp,
                             alpha=s.alpha)
        eng.set_particles(s.iptpn, zat[s.iptpn - 1], 1.0 / s.massat)
        eng.set_tables(tab, s.rcut)
        eng.set_neighbor("vlist", s.drnlist, 0)
        eng.set_ewald(s.ewald, s.alpha, s.order, (s.mesh,) * 3, s.ncut, False, s.lsurf,
                      self.efp)
        eng.set_pair_params(zat, sigat, epsat, rcut=s.rcut, epsi_four_pi=self.efp,
                            alpha=s.alpha)
        c = np.asarray(knots, float).reshape(self.nptpt, self.nknot)
        eng.set_knots(self.knot_lo, self.knot_hi, c)
        eng.upload_positions(np.asarray(ro, float).reshape(-1, 3))
        return eng.adjoint()

    def close(self):
        self.eng.close()


class ModelEnergy(torch.autograd.Function):
    """E(ro, zat, sigat, epsat, knots) with the adjoint as its backward."""

    @staticmethod
    def forward(ctx, ro, zat, sigat, epsat, knots, model: PairModel):
        g = model.evaluate(*(t.detach().cpu().numpy() for t in (ro, zat, sigat, epsat, knots)))
        dev, dt = ro.device, torch.float64
        ctx.save_for_backward(
            torch.as_tensor(g["dro"], dtype=dt, device=dev).reshape(ro.shape),
            torch.as_tensor(g["dzat"], dtype=dt, device=dev),
            torch.as_tensor(g["dsigat"], dtype=dt, device=dev),
            torch.as_tensor(g["depsat"], dtype=dt, device=dev),
            torch.as_tensor(g["dknot"], dtype=dt, device=dev).reshape(knots.shape))
        return torch.tensor(g["E"], dtype=dt, device=dev)

    @staticmethod
    def backward(ctx, grad_e):
        dro, dzat, dsig, deps, dknot = ctx.saved_tensors
        return (grad_e * dro, grad_e * dzat, grad_e * dsig, grad_e * deps, grad_e * dknot,
                None)
