"""Step graphs (no reference counterpart): a run of n MD or BD steps replays
a graph of eight captured steps n // 8 times and a one-step graph for the
rest.  The steps are static (the list rebuild is gated on the device), so a
run must give what n single-step calls give: bitwise on the mixed path,
whose spreading sums in exact fixed point; to rounding on the FP64 path,
whose spreading adds FP64 in LDS in a run-dependent order.  MD on two
systems: the polyelectrolyte (bonds: the pair forces on a second stream
beside the reciprocal stages) and an electrolyte replica (no bonds: the
mixed path's pair forces inside the gather's launch, the FP64 path's on
the second stream); BD on the electrolyte."""
import numpy as np
import pytest

from molsim_amd import systems as S

pytestmark = pytest.mark.gpu

NSTEPS = 19  # two replays of the eight-step graph and three single steps


def say(*a):
    print(*a, flush=True)


def systems():
    return {"polyelectrolyte": S.c2(), "electrolyte": S.c4_replica(7)}


def compare(a, b, prec, what):
    ra, va, _ = a.download_state()
    rb, vb, _ = b.download_state()
    fa, fb = a.download_forces(), b.download_forces()
    ea, eb = a.energy(), b.energy()
    dr = float(np.max(np.abs(ra - rb)))
    df = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fb)))
    say(f"  {what}: max |dr| {dr:.2e} A, max |dF|/max|F| {df:.2e}, "
        f"rebuilds {ea.rebuilds} / {eb.rebuilds}")
    assert ea.rebuilds == eb.rebuilds
    if prec == "mixed":
        assert np.array_equal(ra, rb) and np.array_equal(va, vb) and np.array_equal(fa, fb)
        assert ea.tot == eb.tot
    else:
        assert dr <= 1e-10 and df <= 1e-10
        assert abs(ea.tot - eb.tot) <= 1e-12 * abs(eb.tot)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("name", ["polyelectrolyte", "electrolyte"])
def test_md_run_equals_single_steps(name, prec):
    s = systems()[name]
    say(f"case: {NSTEPS} MD steps at once vs one by one, {name} N={s.np}, {prec}")
    a, b = S.configure(s, prec), S.configure(s, prec)
    a.step_md(nsteps=NSTEPS, tstep=s.tstep)
    for _ in range(NSTEPS):
        b.step_md(nsteps=1, tstep=s.tstep)
    compare(a, b, prec, "MD")
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("name", ["electrolyte"])
def test_bd_run_equals_single_steps(name, prec):
    s = systems()[name]
    say(f"case: {NSTEPS} BD steps at once vs one by one, {name} N={s.np}, {prec}")
    npt = int(np.max(s.iptpn))
    # factors of the scale the batch tests use (noise 0.03 A per step, a
    # drift of ~1e-4 A), gentle on the start configuration's close pairs
    fac1, fac2 = np.full(npt, 2.0e-7), np.full(npt, 0.0316)
    a, b = S.configure(s, prec), S.configure(s, prec)
    a.set_bd(fac1, fac2, 11)
    b.set_bd(fac1, fac2, 11)
    a.step_bd(nsteps=NSTEPS)
    for _ in range(NSTEPS):
        b.step_bd(nsteps=1)
    compare(a, b, prec, "BD")
    a.close()
    b.close()


def test_bd_first_step_takes_the_current_forces():
    """A BD run without a UTotal before it evaluates the forces first (the
    reference's UTotal before its BD loop): the same trajectory, bitwise,
    as with an explicit UTotal."""
    s = S.c4_replica(7)
    say(f"case: BD without and with a UTotal before it, electrolyte N={s.np}, mixed")
    npt = int(np.max(s.iptpn))
    fac1, fac2 = np.full(npt, 2.0e-7), np.full(npt, 0.0316)
    a, b = S.configure(s, "mixed"), S.configure(s, "mixed")
    a.set_bd(fac1, fac2, 5)
    b.set_bd(fac1, fac2, 5)
    b.utotal()
    a.step_bd(nsteps=3)
    b.step_bd(nsteps=3)
    compare(a, b, "mixed", "BD")
    a.close()
    b.close()
