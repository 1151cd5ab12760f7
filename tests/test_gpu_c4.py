"""The batched-replica configuration at its full size (512 BD replicas,
N ~ U{500..4000}, 0.1-1 M salt; no reference counterpart): one batch of
all of them, on the autotuner's grid rule for alpha 0.15 and order 4,
advanced in the persistent launch; a sample of replicas is then checked
against the oracle's UTotal at their final positions."""
import numpy as np
import pytest

from molsim_amd import autotune as A
from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def capacity(s):
    mean = 4.0 / 3.0 * np.pi * (s.rcut + s.drnlist) ** 3 * s.np / s.boxlen ** 3
    return int(min(s.np - 1, 4 * mean + 64))


def test_512_replicas_in_one_persistent_launch():
    alpha, order = 0.15, 4
    reps = [S.c4_replica(r) for r in range(512)]
    tuned = [A.with_params(s, alpha, order, A.mesh_for(s.boxlen, alpha, 1.0)) for s in reps]
    say(f"case: 512 replicas, N total {sum(s.np for s in tuned)}, meshes "
        f"{min(s.mesh for s in tuned)}-{max(s.mesh for s in tuned)}")
    cache = {}
    fac1, fac2 = (float(v[0]) for v in E.bd_factors(0.005, 298.0, [1.0]))
    engs = [S.configure(s, "mixed", tables=A._tables(s, cache), maxnneigh=capacity(s))
            for s in tuned]
    for r, e in enumerate(engs):
        e.set_bd(np.full(2, fac1), np.full(2, fac2), 4000 + r)
    b = E.Batch(engs)
    b.compute()
    assert b.uses_persistent()
    b.step_bd(10)
    b.energies()
    b.step_bd(10)
    ens = b.energies()
    b.close()
    assert all(np.isfinite(en.tot) for en in ens)
    res = []
    for r in (0, 63, 127, 255, 383, 511):
        s = tuned[r]
        ro, _, _ = engs[r].download_state()
        assert np.isfinite(ro).all() and np.abs(ro).max() <= s.boxlen / 2
        orc = oracle_for(s, A._tables(s, cache))
        orc.set_list(ro)
        fo, eo = orc.utotal(ro)
        fr = rel_rms(engs[r].download_forces(), fo)
        # energies against the scale of the terms: the total (and the pair sum
        # of +/- ion pairs) is a small difference of larger pair and
        # reciprocal contributions, and FP32 rounding follows those
        scale = abs(eo["twob0"]) + abs(eo["rec"])
        dp = abs(ens[r].twob0 - eo["twob0"]) / scale
        dr = abs(ens[r].rec - eo["rec"]) / scale
        dt = abs(ens[r].tot - eo["tot"]) / scale
        say(f"  replica {r}: N {s.np} mesh {s.mesh}: force rel-RMS {fr:.2e}, pair {dp:.2e},"
            f" rec {dr:.2e}, tot {dt:.2e} of |pair|+|rec| {scale:.4e}")
        res.append((fr, dp, dr, dt))
    for fr, dp, dr, dt in res:
        assert fr <= 1e-5
        assert max(dp, dr, dt) <= 1e-6
    for e in engs:
        e.close()
