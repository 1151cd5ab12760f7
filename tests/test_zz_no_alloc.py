"""The no-allocation variant of the Monte Carlo pair delta,
mfe_mc_delta_pair_noalloc, called through the C ABI, against
mfe_mc_delta_pair on the same engine: two calls in a row (the second with
fewer moved particles), on both list semantics and both precision paths.
The two entry points run the same kernels on the same inputs, so their
results are bitwise equal."""
import ctypes as C

import numpy as np
import pytest

from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def say(*a):
    print(*a, flush=True)


def noalloc(eng, ia, rtm):
    L = E.lib()
    f = L.mfe_mc_delta_pair_noalloc
    f.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp, _ip]
    ia = np.ascontiguousarray(ia, np.int32)
    rtm = np.ascontiguousarray(rtm, np.float64).reshape(-1)
    du = C.c_double()
    ov = C.c_int32()
    rc = f(eng.h, ia.size, ia.ctypes.data_as(_ip), rtm.ctypes.data_as(_dp), C.byref(du),
           C.byref(ov))
    return rc, du.value, ov.value


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("nlist", ["vlist", "cell"])
def test_noalloc_pair_delta_equals_allocating_entry_point(prec, nlist):
    s = S.c2()
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    say(f"case: no-allocation MC pair delta, c2 {prec} {nlist}")
    eng = S.configure(s, prec, tables=tab, nlist=nlist)
    eng.build_list()
    rng = np.random.default_rng(77)
    for natm in (8, 3):
        ia = rng.choice(s.np, natm, replace=False) + 1
        rtm = s.ro[ia - 1] + rng.uniform(-0.6, 0.6, (natm, 3))
        rtm -= s.boxlen * np.round(rtm / s.boxlen)
        du_ref, ov_ref = eng.mc_delta_pair(ia, rtm)
        rc, du, ov = noalloc(eng, ia, rtm)
        say(f"  natm {natm}: du {du!r} vs {du_ref!r}, overlap {ov} vs {ov_ref}")
        assert rc == 0
        assert bool(ov) == ov_ref and du == du_ref
    # beyond the scratch limit: rejected before any device work
    ia = np.arange(1, 1026, dtype=np.int32)
    rc, _, _ = noalloc(eng, ia, s.ro[:1025])
    assert E.ERRORS[rc] == "MFE_E_BAD_PARAM"
    eng.close()
