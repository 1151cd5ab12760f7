"""Brownian dynamics and batched replicas on the device.

BDStep (Src/bd.F90:120-155) with host-supplied deviates is checked against
the oracle's BDStep step for step.  The device noise stream (no reference
counterpart: GauRandom is one serial LCG) is checked against the oracle's
Philox4x32-10 + Box-Muller, itself pinned by the Random123 known-answer
vectors.  Batches (no reference counterpart) are pinned by equality with
the same replicas run one engine at a time: bitwise on the mixed path,
whose charge spreading accumulates in exact fixed point; within 1e-12 on
the FP64 path, whose LDS atomics add in a run-dependent order.
"""
import numpy as np
import pytest

import oracle as O
from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp, oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


# BD factors of the right scale for 1:1 ions at dt = 5 fs, D = 1e-9 m^2/s:
# sqrt(2 D dt) = 0.0316 A; fac1 * |forceo| ~ 1e-4 A
FAC1 = np.array([2.0e-7, 2.0e-7])
FAC2 = np.array([0.0316, 0.0316])


# meshes of mixed radix, including 46 = 2 x 23 (generic prime radix, a
# brick of 23^3 points above 64 KB of LDS)
MESHES = [24, 32, 40, 46, 36, 30]


def replicas(n, order=4):
    return [S.c4_replica(r, mesh=MESHES[r % len(MESHES)], order=order) for r in range(n)]


def test_bd_host_noise_matches_oracle():
    s = S.c4_replica(0)
    say(f"case: BD host noise, replica N={s.np} L={s.boxlen:.1f} fp64")
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    eng = S.configure(s, "fp64", tables=tab)
    f0, _ = eng.utotal()
    rng = np.random.default_rng(5)
    nsteps = 5
    noise = rng.standard_normal((nsteps, s.np, 3))
    eng.step_bd(FAC1, FAC2, noise)
    ro_dev, _, _ = eng.download_state()
    eng.close()

    orc = oracle_for(s, tab)
    ro = s.ro.reshape(-1).copy()
    orc.set_list(ro)
    f, _ = orc.utotal(ro)
    drosum = np.zeros_like(ro)
    for k in range(nsteps):
        dro = O.bd_step(s.boxlen, s.iptpn, FAC1, FAC2, f.reshape(-1), noise[k].reshape(-1), ro)
        f, _ = orc.utotal(ro)
        if O.skin_test(drosum, dro, s.drnlist):
            orc.set_list(ro)
    d = np.abs(ro_dev.reshape(-1) - ro)
    d = np.minimum(d, s.boxlen - d)
    say(f"  max |dro| engine vs oracle {d.max():.3e} A")
    assert d.max() <= 1e-9


def test_device_noise_matches_oracle():
    s = S.c4_replica(1)
    say(f"case: BD device noise, replica N={s.np} fp64")
    eng = S.configure(s, "fp64")
    eng.utotal()
    seed = 0x1234_5678_9abc
    fac1 = np.zeros(2)
    eng.set_bd(fac1, FAC2, seed)
    ro = s.ro.copy()
    for step in range(2):
        eng.step_bd(nsteps=1)
        dev, _, _ = eng.download_state()
        ro = ro + FAC2[0] * O.bd_noise(seed, s.np, step)
        ro -= s.boxlen * np.round(ro / s.boxlen)
        d = np.abs(dev - ro)
        d = np.minimum(d, s.boxlen - d)
        say(f"  step {step}: max |dro| {d.max():.3e} A")
        assert d.max() <= 1e-12
    eng.close()


def _pair(sysl, prec, tab):
    single = [S.configure(s, prec, tables=tab) for s in sysl]
    inbatch = [S.configure(s, prec, tables=tab) for s in sysl]
    return single, inbatch


@pytest.fixture(scope="module")
def c4_tab():
    s = S.c4_replica(0)
    return E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                          alpha=s.alpha)


def test_batch_equals_single_mixed_bitwise(c4_tab):
    sysl = replicas(6)
    say("case: batch of 6 replicas, mixed, N =", [s.np for s in sysl],
        "mesh =", [s.mesh for s in sysl])
    single, inbatch = _pair(sysl, "mixed", c4_tab)
    b = E.Batch(inbatch)
    # UTotal
    b.compute()
    ens = b.energies()
    for r, (e1, e2) in enumerate(zip(single, inbatch)):
        f1, en1 = e1.utotal()
        assert np.array_equal(f1, e2.download_forces()), f"replica {r} forces"
        assert en1.tot == ens[r].tot and en1.rec == ens[r].rec, f"replica {r} energy"
    # VelVer with rebuilds
    for e in single:
        e.step_md(15, 0.005)
    b.step_md(15, 0.005)
    for r, (e1, e2) in enumerate(zip(single, inbatch)):
        assert np.array_equal(e1.download_state()[0], e2.download_state()[0]), f"replica {r} md"
    # BD with each replica's device noise stream
    for r, (e1, e2) in enumerate(zip(single, inbatch)):
        e1.set_bd(FAC1, FAC2, 1000 + r)
        e2.set_bd(FAC1, FAC2, 1000 + r)
    for e in single:
        e.step_bd(nsteps=15)
    # the per-stage launches (the persistent launch has its own pins)
    b.set_persistent(False)
    b.step_bd(15)
    ens = b.energies()
    for r, (e1, e2) in enumerate(zip(single, inbatch)):
        assert np.array_equal(e1.download_state()[0], e2.download_state()[0]), f"replica {r} bd"
        en1 = e1.energy()
        assert en1.tot == ens[r].tot and en1.rebuilds == ens[r].rebuilds
    say("  rebuilds per replica:", [en.rebuilds for en in ens])
    b.close()
    for e in single + inbatch:
        e.close()


def test_batch_fp64_close_to_single_and_oracle(c4_tab):
    sysl = replicas(3)
    say("case: batch of 3 replicas, fp64, N =", [s.np for s in sysl])
    single, inbatch = _pair(sysl, "fp64", c4_tab)
    b = E.Batch(inbatch)
    b.compute()
    for r, (s, e1, e2) in enumerate(zip(sysl, single, inbatch)):
        f1, _ = e1.utotal()
        f2 = e2.download_forces()
        assert rel_rms(f2, f1) <= 1e-12, f"replica {r}"
        orc = oracle_for(s, c4_tab)
        orc.set_list(s.ro)
        fo, _ = orc.utotal(s.ro)
        assert rel_rms(f2, fo) <= 1e-10, f"replica {r} vs oracle"
    b.close()
    for e in single + inbatch:
        e.close()


def test_batch_reports_failing_replica(c4_tab):
    sysl = replicas(3)
    bad = sysl[2]
    bad.ro[1] = bad.ro[0] + np.array([0.5, 0.0, 0.0])   # r2 < r2umin
    bad.ro -= bad.boxlen * np.round(bad.ro / bad.boxlen)
    say("case: batch with an overlapping pair in replica 2")
    engs = [S.configure(s, "mixed", tables=c4_tab) for s in sysl]
    b = E.Batch(engs)
    with pytest.raises(E.MfeError) as ei:
        b.compute()
    assert ei.value.code == 2 and ei.value.replica == 2
    b.close()
    for e in engs:
        e.close()


def test_autotuner_plan_meets_tolerance():
    from molsim_amd import autotune as A
    sysl = [S.c4_replica(r) for r in range(6)]
    tol = 1e-3
    say("case: autotune 6 replicas, alpha 0.1/0.125/0.2, orders 4/6, tol", tol)
    plan = A.autotune(sysl, tol=tol, alphas=(0.1, 0.125, 0.2), orders=(4, 6), n_sample=3,
                      n_time=6, time_steps=10)
    say(f"  plan: alpha {plan.alpha} order {plan.order} alpha*h {plan.ah} meshes {plan.meshes}"
        f" {plan.sec_per_step * 1e3:.3f} ms/step, sample error {plan.error:.2e}, met {plan.met}")
    for k, v in plan.table.items():
        say("   ", k, v["ah"], v["sec_per_step"], {a: "%.2e" % e for a, e in v["errors"].items()})
    assert plan.met
    # every replica -- not only the calibration sample -- meets the tolerance
    cache = {}
    tuned = plan.apply(sysl)
    refs = [A.with_params(s, plan.alpha, 8, min(256, A.mesh_for(s.boxlen, plan.alpha, 0.45)))
            for s in sysl]
    f = A.batch_forces(tuned, "mixed", cache)
    fr = A.batch_forces(refs, "fp64", cache)
    errs = [A.rel_rms(a, b) for a, b in zip(f, fr)]
    say("  errors:", ["%.2e" % e for e in errs])
    assert max(errs) <= 2 * tol
    assert all(A.MESH_MIN <= m <= A.MESH_MAX and m % 2 == 0 for m in plan.meshes)


def test_engines_step_alone_after_their_batch_is_closed():
    """A batch holds views of its engines, never their memory: after
    mfe_batch_destroy each engine steps alone -- host-noise BD steps equal a
    fresh engine's from the same positions bitwise, and device-noise steps
    (a captured graph of the engine's own) run and stay in the box."""
    reps = [S.c4_replica(r, mesh=MESHES[r], order=4) for r in range(3)]
    engs = []
    for r, s in enumerate(reps):
        tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                             alpha=s.alpha)
        e = S.configure(s, "mixed", tables=tab)
        e.set_bd(FAC1, FAC2, 100 + r)
        engs.append((e, tab))
    say("case: engines after their batch is closed, 3 replicas, mixed")
    b = E.Batch([e for e, _ in engs])
    b.compute()
    b.step_bd(5)
    b.energies()
    b.close()
    rng = np.random.default_rng(31)
    for (e, tab), s in zip(engs, reps):
        ro, _, _ = e.download_state()
        # the list built anew at these positions, as the fresh engine builds it
        e.build_list()
        e.utotal()
        noise = rng.standard_normal((3, s.np, 3))
        e.step_bd(FAC1, FAC2, noise)
        ro_a, _, _ = e.download_state()
        s2 = S.c4_replica(reps.index(s), mesh=s.mesh, order=4)
        s2.ro = ro
        fresh = S.configure(s2, "mixed", tables=tab)
        fresh.utotal()
        fresh.step_bd(FAC1, FAC2, noise)
        ro_b, _, _ = fresh.download_state()
        fresh.close()
        assert np.array_equal(ro_a, ro_b)
        e.step_bd(nsteps=5)
        ro_c, _, _ = e.download_state()
        assert np.isfinite(ro_c).all() and np.abs(ro_c).max() <= s.boxlen / 2
        e.close()
