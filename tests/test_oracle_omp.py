"""The oracle's OpenMP twin (the same C source with its particle loops, FFT
rows and spreading split over threads), used only for the all-cores CPU
baseline, against the serial oracle: the same UTotal up to the summation
order of its threaded reductions."""
import oracle as O
from molsim_amd import engine as E
from molsim_amd import systems as S

from helpers import efp, oracle_for, rel_rms


def test_openmp_twin_equals_serial_oracle():
    s = S.c2()
    tab = E.build_tables(s.zat, s.sigat, s.epsat, rcut=s.rcut, epsi_four_pi=efp(),
                         alpha=s.alpha)
    ro = s.ro.reshape(-1).copy()
    res = {}
    for omp in (False, True):
        O.use_all_cores(omp)
        try:
            orc = oracle_for(s, tab)
            orc.set_list(ro)
            res[omp] = orc.utotal(ro)
        finally:
            O.use_all_cores(False)
    (f0, e0), (f1, e1) = res[False], res[True]
    assert rel_rms(f1, f0) <= 1e-12
    for k in ("tot", "twob0", "rec"):
        assert abs(e1[k] - e0[k]) <= 1e-12 * abs(e0[k]), k
