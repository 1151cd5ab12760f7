"""SPME autotuner for batched replicas (no reference counterpart).

MOLSIM takes its Ewald parameters from the input (`ualphared`, `order`,
`nmesh` in nmlEwald; Src/energy.F90:6875-6892).  For a batch of replicas
with different sizes and concentrations this module chooses them:

  * alpha (and rc = 3/alpha, the reference's ualphared = 3 rule) and the
    B-spline order p are shared by the batch -- p is a compile-time
    parameter of the stencil kernels, so one order keeps every stage one
    launch;
  * the mesh is per replica: the smallest even s in [16, 48] without prime
    factors above 7 (fast FFT radices) with alpha * L / s <= (alpha h)*,
    where (alpha h)* is the coarsest grid spacing whose measured force
    error stays within the tolerance.

Both steps are measured on the device rather than modelled:

  1. calibration -- for each (alpha, p) on a sample of replicas, the force
     rel-RMS of the batched engine against an FP64 order-8 reference on a
     fine mesh, over a ladder of alpha*h values; (alpha h)* is the largest
     value at which every sampled replica passes;
  2. timing -- the batched BD step of a sample of replicas with the meshes
     each feasible (alpha, p) implies; the fastest wins;
  3. the launch schedule -- the chosen batch timed whole in both of the
     engine's schedules for BD steps: one persistent launch (a workgroup
     per replica at a time, all stages inside) or one graph launch per
     stage over all replicas; the faster is kept (Plan.persistent).
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import engine as E
from . import systems as S

# alpha*rc = 3: alpha 0.1 -> rc 30 A.  Boxes up to ~320 A on meshes <= 48
# (h >= 6.7 A) need small alpha for alpha*h <~ 1
ALPHAS = (0.10, 0.125, 0.15, 0.20, 0.25)
ORDERS = (4, 5, 6)
AH_LADDER = (2.0, 1.8, 1.6, 1.45, 1.3, 1.15, 1.0, 0.9, 0.8, 0.7, 0.6, 0.5)
MESH_MIN, MESH_MAX = 16, 48


def even_ceil(x):
    n = int(math.ceil(x - 1e-9))
    return n + (n % 2)


def smooth(n, primes=(2, 3, 5, 7)):
    """True when n has no prime factor above 7 (fast FFT radices)."""
    for p in primes:
        while n % p == 0:
            n //= p
    return n == 1


def mesh_for(L, alpha, ah):
    """Smallest even, 7-smooth mesh >= max(16, alpha L / ah)."""
    n = max(MESH_MIN, even_ceil(alpha * L / ah))
    while not smooth(n):
        n += 2
    return n


def with_params(sysm: S.System, alpha, order, mesh) -> S.System:
    rc = min(3.0 / alpha, 0.45 * sysm.boxlen)
    return dataclasses.replace(sysm, alpha=alpha, rcut=rc, order=order, mesh=mesh)


def _tables(sysm, cache):
    key = (round(sysm.rcut, 9), round(sysm.alpha, 9), tuple(sysm.zat), tuple(sysm.sigat),
           tuple(sysm.epsat))
    if key not in cache:
        efp = E.epsi0_four_pi() / 78.4
        cache[key] = E.build_tables(sysm.zat, sysm.sigat, sysm.epsat, rcut=sysm.rcut,
                                    epsi_four_pi=efp, alpha=sysm.alpha)
    return cache[key]


def batch_forces(systems, precision, cache):
    """forceo of every replica from one batched UTotal."""
    engs = [S.configure(s, precision, tables=_tables(s, cache)) for s in systems]
    b = E.Batch(engs)
    try:
        b.compute()
        return [e.download_forces() for e in engs]
    finally:
        b.close()
        for e in engs:
            e.close()


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def calibrate(sample, alpha, order, tol, cache, precision="mixed"):
    """Largest alpha*h on the ladder at which every sampled replica's force
    rel-RMS against the FP64 order-8 reference is <= tol (None if none),
    and the measured error of every rung tried (largest over the sample)."""
    refs = []
    for s in sample:
        s_ref = with_params(s, alpha, 8, min(256, mesh_for(s.boxlen, alpha, 0.45)))
        refs.append(s_ref)
    f_ref = batch_forces(refs, "fp64", cache)
    errs = {}
    for ah in AH_LADDER:
        meshes = [mesh_for(s.boxlen, alpha, ah) for s in sample]
        if max(meshes) > MESH_MAX:
            break
        cand = [with_params(s, alpha, order, m) for s, m in zip(sample, meshes)]
        f = batch_forces(cand, precision, cache)
        e = max(rel_rms(a, b) for a, b in zip(f, f_ref))
        errs[ah] = e
        if e <= tol:
            return ah, errs
    return None, errs


def time_bd(systems, steps, cache, precision="mixed", fac1=2.0e-7, fac2=0.0316,
            persistent=True):
    """Seconds per batched BD step of `systems` (persistent: whether the
    batch may take the one-launch schedule)."""
    engs = [S.configure(s, precision, tables=_tables(s, cache)) for s in systems]
    for r, e in enumerate(engs):
        e.set_bd(np.full(s_npt(systems[r]), fac1), np.full(s_npt(systems[r]), fac2), 77 + r)
    b = E.Batch(engs)
    try:
        b.set_persistent(persistent)
        b.compute()
        b.step_bd(max(2, steps // 5))
        b.energies()
        t0 = time.perf_counter()
        b.step_bd(steps)
        b.energies()
        return (time.perf_counter() - t0) / steps
    finally:
        b.close()
        for e in engs:
            e.close()


def s_npt(sysm):
    return sysm.massat.size


@dataclasses.dataclass
class Plan:
    alpha: float
    order: int
    ah: float
    meshes: list
    rcut: float
    sec_per_step: float
    table: dict          # (alpha, order) -> {"ah", "errors", "sec_per_step"}
    error: float = 0.0   # measured force rel-RMS at the chosen alpha*h (sample max)
    met: bool = True     # False: no candidate reached the tolerance; the
                         # most accurate feasible one was taken
    persistent: bool = True   # launch schedule: one persistent launch for all
                              # replicas, else one graph launch per stage
    schedule_sec: dict = dataclasses.field(default_factory=dict)

    def apply(self, systems):
        return [with_params(s, self.alpha, self.order, m) for s, m in zip(systems, self.meshes)]


def autotune(systems, tol=1e-4, alphas=ALPHAS, orders=ORDERS, n_sample=4, n_time=32,
             time_steps=20, precision="mixed") -> Plan:
    """Choose (alpha, order) for the batch and a mesh per replica."""
    cache = {}
    idx = np.linspace(0, len(systems) - 1, min(n_sample, len(systems))).round().astype(int)
    # the sample spans the largest boxes (finest meshes) and the densest
    order_by_L = sorted(range(len(systems)), key=lambda r: systems[r].boxlen)
    sample = [systems[order_by_L[i]] for i in sorted(set(idx))]
    tidx = np.linspace(0, len(systems) - 1, min(n_time, len(systems))).round().astype(int)
    tsample = [systems[i] for i in sorted(set(tidx))]
    table = {}
    best = None
    fallback = None     # (error, alpha, order, ah): most accurate feasible rung
    for a in alphas:
        for p in orders:
            ah, errs = calibrate(sample, a, p, tol, cache, precision)
            entry = {"ah": ah, "errors": errs, "sec_per_step": None}
            table[(a, p)] = entry
            for ah_k, e_k in errs.items():
                if fallback is None or e_k < fallback[0]:
                    fallback = (e_k, a, p, ah_k)
            if ah is None:
                continue
            meshes = [mesh_for(s.boxlen, a, ah) for s in systems]
            if max(meshes) > MESH_MAX:
                continue
            tm = [with_params(s, a, p, mesh_for(s.boxlen, a, ah)) for s in tsample]
            sec = time_bd(tm, time_steps, cache, precision)
            entry["sec_per_step"] = sec
            if best is None or sec < best.sec_per_step:
                best = Plan(a, p, ah, meshes, min(3.0 / a, 0.45 * min(s.boxlen for s in systems)),
                            sec, table, error=errs[ah])
    if best is None:
        if fallback is None:
            raise RuntimeError(f"no (alpha, order) fits meshes <= {MESH_MAX}")
        e, a, p, ah = fallback
        meshes = [mesh_for(s.boxlen, a, ah) for s in systems]
        tm = [with_params(s, a, p, mesh_for(s.boxlen, a, ah)) for s in tsample]
        best = Plan(a, p, ah, meshes, min(3.0 / a, 0.45 * min(s.boxlen for s in systems)),
                    time_bd(tm, time_steps, cache, precision), table, error=e, met=False)
    best.table = table
    # the launch schedule, timed on the whole batch: the persistent launch
    # (a workgroup per replica at a time) against the per-stage graph
    if precision == "mixed" and len(systems) > 1:
        full = best.apply(systems)
        steps = max(3, time_steps // 4)
        tp = time_bd(full, steps, cache, precision, persistent=True)
        ts = time_bd(full, steps, cache, precision, persistent=False)
        best.persistent = tp <= ts
        best.schedule_sec = {"persistent": tp, "per_stage": ts}
    return best
