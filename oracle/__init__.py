"""ctypes bindings of the FP64 oracle (oracle/mfe_oracle.c).

Test infrastructure: imported only by tests/, the smoke check and the CPU
baseline leg of bench.py, to check results against.  The product never
imports it.

Every wrapper mirrors one reference routine (cited in mfe_oracle.c) and uses
MOLSIM's conventions: xyz-interleaved positions (ro(3,np)), 1-based particle
types, 1-based ubuf offsets.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
# the all-cores twin: the same source built with -fopenmp (particle loops,
# FFT rows and spreading split over threads; summation order not the
# reference's, so a timing baseline only)
LIB_OMP_PATH = os.path.join(_HERE, "liboracle_omp.so")
SRC_PATH = os.path.join(_HERE, "mfe_oracle.c")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction, like the reference),
    and its all-cores twin with -fopenmp."""
    for path, extra in ((LIB_PATH, []), (LIB_OMP_PATH, ["-fopenmp"])):
        if force or not os.path.exists(path) or (
                os.path.getmtime(path) < os.path.getmtime(SRC_PATH)):
            subprocess.check_call([
                "gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC"] + extra +
                ["-shared", "-o", path, SRC_PATH, "-lm"])
    return LIB_PATH


_lib = None
_lib_serial = None
_lib_omp = None


def use_all_cores(on: bool) -> None:
    """Route every binding through the serial oracle (default) or its
    all-cores OpenMP twin."""
    global _lib, _lib_serial, _lib_omp
    if on:
        if _lib_omp is None:
            if not os.path.exists(LIB_OMP_PATH):
                build()
            _lib_omp = C.CDLL(LIB_OMP_PATH)
            _declare(_lib_omp)
        _lib = _lib_omp
    else:
        _lib = None
        lib()


def lib():
    global _lib, _lib_serial
    if _lib is None:
        if _lib_serial is None:
            if not os.path.exists(LIB_PATH):
                build()
            _lib_serial = C.CDLL(LIB_PATH)
            _declare(_lib_serial)
        _lib = _lib_serial
    return _lib


class Box(C.Structure):
    _fields_ = [("boxlen", C.c_double * 3), ("boxlen2", C.c_double * 3),
                ("boxleni", C.c_double * 3), ("dpbc", C.c_double * 3)]


class PotSpec(C.Structure):
    _fields_ = [("npt", C.c_int), ("zat", _dp), ("sigat", _dp), ("epsat", _dp),
                ("epsi_four_pi", C.c_double), ("rcut", C.c_double),
                ("alpha", C.c_double), ("lewald", C.c_int), ("lljcut", C.c_int),
                ("ljrcut", C.c_double), ("ljushift", C.c_double),
                ("r2uminin", C.c_double), ("utoltab", C.c_double),
                ("ftoltab", C.c_double), ("umaxtab", C.c_double),
                ("fmaxtab", C.c_double), ("rumaxfac", C.c_double)]


class TableSet(C.Structure):
    _fields_ = [("ubuf", _dp), ("nbuf", C.c_int), ("iubuflow", _ip),
                ("r2atat", _dp), ("r2umin", _dp), ("rcut2", C.c_double),
                ("npt", C.c_int)]


class CellList(C.Structure):
    _fields_ = [("ncell", C.c_int * 3), ("cellsize", C.c_double * 3),
                ("cellsizei", C.c_double * 3), ("ndir", C.c_int),
                ("dirs", _ip), ("iphead", _ip), ("npart", _ip),
                ("ipnext", _ip), ("cellip", _ip), ("np", C.c_int)]


class Spme(C.Structure):
    _fields_ = [("order", C.c_int), ("s", C.c_int * 3), ("alpha", C.c_double),
                ("dkdr", C.c_double * 3), ("vol", C.c_double),
                ("boxlen", C.c_double * 3), ("b2", _dp * 3), ("meshfac", _dp),
                ("energyfac", _dp), ("virfac", _dp)]


class System(C.Structure):
    _fields_ = [("box", Box), ("np", C.c_int), ("npt", C.c_int),
                ("iptpn", _ip), ("az", _dp), ("tab", TableSet),
                ("nlkind", C.c_int), ("maxnneigh", C.c_int),
                ("nneighpn", _ip), ("jpnlist", _ip),
                ("cells", C.POINTER(CellList)), ("ewald", C.c_int),
                ("alpha", C.c_double), ("epsi_four_pi", C.c_double),
                ("ncut", C.c_int), ("ncut_cube", C.c_int), ("lsurf", C.c_int),
                ("spme", C.POINTER(Spme)), ("nchain", C.c_int),
                ("chain_first", _ip), ("chain_len", _ip),
                ("bond_k", C.c_double), ("bond_p", C.c_double),
                ("bond_eq", C.c_double), ("angle_k", C.c_double),
                ("angle_p", C.c_double), ("angle_eq", C.c_double)]


class Energy(C.Structure):
    _fields_ = [("tot", C.c_double), ("twob0", C.c_double), ("rec", C.c_double),
                ("bond", C.c_double), ("virial", C.c_double),
                ("twob", C.c_double * 64), ("angle", C.c_double)]


class GauState(C.Structure):
    _fields_ = [("flag", C.c_int), ("save", C.c_double)]


def _declare(L):
    L.orc_epsi0_four_pi.restype = C.c_double
    L.orc_force_unit_fac.restype = C.c_double
    L.orc_gas_constant.restype = C.c_double
    L.orc_erf2.restype = C.c_double
    L.orc_erf2.argtypes = [C.c_double]
    L.orc_cardinal_bspline.argtypes = [C.c_int, C.c_double, _ip, _dp, _dp]
    L.orc_box_init.argtypes = [C.POINTER(Box), _dp]
    L.orc_pot_spec_defaults.argtypes = [C.POINTER(PotSpec)]
    L.orc_pot_1_6_12.argtypes = [C.POINTER(PotSpec), C.c_int, C.c_int,
                                 C.c_double, _dp, _dp, _dp]
    L.orc_build_tables.argtypes = [C.POINTER(PotSpec), _dp, C.c_int, _ip, _ip,
                                   _dp, _dp]
    L.orc_table_eval.argtypes = [_dp, C.c_int, C.c_int, C.c_double, _dp, _dp]
    L.orc_set_vlist_md.argtypes = [C.POINTER(Box), C.c_int, _dp, C.c_double,
                                   C.c_int, _ip, _ip]
    L.orc_set_vlist_md_cells.argtypes = L.orc_set_vlist_md.argtypes
    L.orc_nlist_skin_test.argtypes = [C.c_int, _dp, _dp, C.c_double]
    L.orc_celllist_init.argtypes = [C.POINTER(CellList), C.POINTER(Box),
                                    C.c_double, C.c_double, C.c_int]
    L.orc_celllist_set.argtypes = [C.POINTER(CellList), C.POINTER(Box), _dp]
    L.orc_celllist_free.argtypes = [C.POINTER(CellList)]
    L.orc_heapsort_index.argtypes = [C.c_int, _dp, _ip]
    L.orc_fft_r2c_3d.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.orc_fft_c2r_3d.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.orc_spme_setup.argtypes = [C.POINTER(Spme), _dp, C.c_double, C.c_int, _ip]
    L.orc_spme_free.argtypes = [C.POINTER(Spme)]
    L.orc_spme_rec.argtypes = [C.POINTER(Spme), C.c_int, _dp, _dp, C.c_double,
                               _dp, _dp, _dp, _dp, _dp]
    L.orc_ewald_rec_std.argtypes = [_dp, C.c_double, C.c_int, C.c_int, C.c_int,
                                    _dp, _dp, C.c_double, _dp, _dp, _dp]
    L.orc_ewald_self.argtypes = [C.c_double, C.c_int, _dp, _dp]
    L.orc_ewald_surf.argtypes = [_dp, C.c_int, _dp, _dp, C.c_double, _dp, _dp,
                                 _dp]
    L.orc_ubond.argtypes = [C.POINTER(Box), _dp, C.c_int, _ip, _ip, C.c_double,
                            C.c_double, C.c_double, _dp, _dp, _dp]
    L.orc_uangle.argtypes = [C.POINTER(Box), _dp, C.c_int, _ip, _ip, C.c_double,
                             C.c_double, C.c_double, _dp, _dp]
    L.orc_utotal.argtypes = [C.POINTER(System), _dp, _dp, C.POINTER(Energy)]
    L.orc_velver_move.argtypes = [C.POINTER(Box), C.c_int, C.c_double, _dp,
                                  _dp, _dp, _dp]
    L.orc_velver_kick.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp]
    L.orc_bd_step.argtypes = [C.POINTER(Box), C.c_int, _ip, _dp, _dp, _dp, _dp,
                              _dp, _dp]
    L.orc_bd_factors.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]
    L.orc_bd_noise.argtypes = [C.c_uint64, C.c_int, C.c_uint32, _dp]
    L.orc_model_pair_energy.restype = C.c_double
    L.orc_model_pair_energy.argtypes = [C.POINTER(Box), C.c_int, _ip, _dp, C.c_int, _dp, _dp,
                                        _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_double, _dp, _dp]
    L.orc_model_pair_forces.restype = C.c_double
    L.orc_model_pair_forces.argtypes = [C.POINTER(Box), C.c_int, _ip, _dp, C.c_int, _dp, _dp,
                                        _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_double, _dp, _dp]
    L.orc_gaurandom.restype = C.c_double
    L.orc_gaurandom.argtypes = [C.POINTER(C.c_int32), C.POINTER(GauState)]


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ units
def epsi0_four_pi() -> float:
    return lib().orc_epsi0_four_pi()


def force_unit_fac() -> float:
    return lib().orc_force_unit_fac()


# ------------------------------------------------------------------- math
def erf2(x) -> np.ndarray:
    L = lib()
    return np.array([L.orc_erf2(float(v)) for v in np.atleast_1d(x)])


def cardinal_bspline(order: int, pos: float):
    v = np.zeros(order)
    d = np.zeros(order)
    k = C.c_int()
    lib().orc_cardinal_bspline(order, float(pos), C.byref(k), _d(v), _d(d))
    return k.value, v, d


# -------------------------------------------------------------------- box
def make_box(boxlen) -> Box:
    b = Box()
    bl = f64(np.broadcast_to(np.asarray(boxlen, float), (3,)))
    lib().orc_box_init(C.byref(b), _d(bl))
    return b


# ----------------------------------------------------------------- tables
class Tables:
    """Output of PotTwoBodyTab1 for the (1,6,12) potential."""

    def __init__(self, ubuf, iubuflow, nugrid, r2umin, r2umax):
        self.ubuf, self.iubuflow, self.nugrid = ubuf, iubuflow, nugrid
        self.r2umin, self.r2umax = r2umin, r2umax


def build_tables(zat, sigat, epsat, *, rcut, epsi_four_pi, alpha=0.0,
                 lewald=True, lljcut=True, **overrides) -> Tables:
    L = lib()
    spec = PotSpec()
    L.orc_pot_spec_defaults(C.byref(spec))
    zat, sigat, epsat = f64(zat), f64(sigat), f64(epsat)
    spec.npt = len(zat)
    spec.zat, spec.sigat, spec.epsat = _d(zat), _d(sigat), _d(epsat)
    spec.epsi_four_pi = epsi_four_pi
    spec.rcut = rcut
    spec.alpha = alpha
    spec.lewald = int(lewald)
    spec.lljcut = int(lljcut)
    for k, v in overrides.items():
        setattr(spec, k, v)
    nptpt = spec.npt * (spec.npt + 1) // 2
    cap = 100000
    ubuf = np.zeros(cap)
    iub = np.zeros(nptpt, np.int32)
    nug = np.zeros(nptpt, np.int32)
    r2min = np.zeros(nptpt)
    r2max = np.zeros(nptpt)
    n = L.orc_build_tables(C.byref(spec), _d(ubuf), cap, _i(iub), _i(nug),
                           _d(r2min), _d(r2max))
    if n < 0:
        raise RuntimeError(f"orc_build_tables failed ({n})")
    return Tables(ubuf[:n].copy(), iub, nug, r2min, r2max)


def pot_1_6_12(zat, sigat, epsat, it, jt, r1, *, epsi_four_pi, alpha=0.0,
               lewald=True, lljcut=True):
    L = lib()
    spec = PotSpec()
    L.orc_pot_spec_defaults(C.byref(spec))
    zat, sigat, epsat = f64(zat), f64(sigat), f64(epsat)
    spec.npt = len(zat)
    spec.zat, spec.sigat, spec.epsat = _d(zat), _d(sigat), _d(epsat)
    spec.epsi_four_pi, spec.alpha = epsi_four_pi, alpha
    spec.lewald, spec.lljcut = int(lewald), int(lljcut)
    u = np.zeros(3)
    L.orc_pot_1_6_12(C.byref(spec), it, jt, float(r1), _d(u[0:]), _d(u[1:]),
                     _d(u[2:]))
    return u


# ------------------------------------------------------- neighbour lists
def set_vlist_md(boxlen, ro, rs, maxnneigh=None, cells=False):
    """SetVListMD: returns (nneighpn[np], jpnlist[np, maxnneigh]) 1-based."""
    ro = f64(ro).reshape(-1)
    np_ = ro.size // 3
    if maxnneigh is None:
        maxnneigh = max(16, np_ - 1) if not cells else 1024
    b = make_box(boxlen)
    nn = np.zeros(np_, np.int32)
    jl = np.zeros((np_, maxnneigh), np.int32)
    fn = lib().orc_set_vlist_md_cells if cells else lib().orc_set_vlist_md
    rc = fn(C.byref(b), np_, _d(ro), float(rs), maxnneigh, _i(nn), _i(jl))
    if rc:
        raise RuntimeError(f"SetVListMD error {rc}")
    return nn, jl


def skin_test(drosum, drostep, drnlist) -> bool:
    drosum = np.asarray(drosum)
    assert drosum.dtype == np.float64 and drosum.flags.c_contiguous
    ds = f64(drostep)
    return bool(lib().orc_nlist_skin_test(drosum.size // 3, _d(drosum), _d(ds),
                                          float(drnlist)))


def heapsort_index(vec):
    v = f64(vec)
    idx = np.zeros(v.size, np.int32)
    lib().orc_heapsort_index(v.size, _d(v), _i(idx))
    return idx


class Cells:
    def __init__(self, boxlen, rcell, rcut, ro):
        self.box = make_box(boxlen)
        self.c = CellList()
        ro = f64(ro).reshape(-1)
        L = lib()
        L.orc_celllist_init(C.byref(self.c), C.byref(self.box), float(rcell),
                            float(rcut), ro.size // 3)
        rc = L.orc_celllist_set(C.byref(self.c), C.byref(self.box), _d(ro))
        if rc:
            raise RuntimeError("particle outside the box")

    def dirs(self):
        return np.ctypeslib.as_array(self.c.dirs, (self.c.ndir, 3)).copy()

    def __del__(self):
        try:
            lib().orc_celllist_free(C.byref(self.c))
        except Exception:
            pass


# -------------------------------------------------------------------- FFT
def fft_r2c_3d(q):
    """q: real array [s3, s2, s1] -> complex [s3, s2, s1//2+1] (FFTW r2c)."""
    q = f64(q)
    s3, s2, s1 = q.shape
    out = np.zeros((s3, s2, s1 // 2 + 1), np.complex128)
    lib().orc_fft_r2c_3d(s1, s2, s3, _d(q), out.ctypes.data_as(_dp))
    return out


def fft_c2r_3d(f, s1):
    f = np.ascontiguousarray(f, np.complex128)
    s3, s2, _ = f.shape
    out = np.zeros((s3, s2, s1))
    lib().orc_fft_c2r_3d(s1, s2, s3, f.ctypes.data_as(_dp), _d(out))
    return out


# -------------------------------------------------------------------- SPME
class SpmeSetup:
    def __init__(self, boxlen, alpha, order, mesh):
        self.p = Spme()
        bl = f64(np.broadcast_to(np.asarray(boxlen, float), (3,)))
        s = i32(np.broadcast_to(np.asarray(mesh), (3,)))
        rc = lib().orc_spme_setup(C.byref(self.p), _d(bl), float(alpha),
                                  int(order), _i(s))
        if rc:
            raise ValueError("mesh must be even")
        self.s = tuple(int(v) for v in s)

    def tables(self):
        h = [v // 2 + 1 for v in self.s]
        n = h[0] * h[1] * h[2]
        shp = (h[2], h[1], h[0])
        get = lambda p: np.ctypeslib.as_array(p, (n,)).reshape(shp).copy()
        b2 = [np.ctypeslib.as_array(self.p.b2[i], (self.s[i],)).copy()
              for i in range(3)]
        return b2, get(self.p.meshfac), get(self.p.energyfac), get(self.p.virfac)

    def __del__(self):
        try:
            lib().orc_spme_free(C.byref(self.p))
        except Exception:
            pass


def spme_rec(setup: SpmeSetup, r, az, epsi_four_pi, want_meshes=False):
    """UEwaldRecSPM: returns (force[np,3] contribution, u_rec, virrec[, Q, conv])."""
    r = f64(r).reshape(-1)
    az = f64(az)
    na = az.size
    force = np.zeros(3 * na)
    u = np.zeros(1)
    vir = np.zeros(1)
    M = setup.s[0] * setup.s[1] * setup.s[2]
    qm = np.zeros(M) if want_meshes else None
    cv = np.zeros(M) if want_meshes else None
    lib().orc_spme_rec(C.byref(setup.p), na, _d(r), _d(az), float(epsi_four_pi),
                       _d(force), _d(u), _d(vir),
                       _d(qm) if want_meshes else None,
                       _d(cv) if want_meshes else None)
    out = (force.reshape(na, 3), float(u[0]), float(vir[0]))
    if want_meshes:
        shp = (setup.s[2], setup.s[1], setup.s[0])
        out = out + (qm.reshape(shp), cv.reshape(shp))
    return out


def ewald_rec_std(boxlen, alpha, ncut, r, az, epsi_four_pi, cube=False):
    r = f64(r).reshape(-1)
    az = f64(az)
    bl = f64(np.broadcast_to(np.asarray(boxlen, float), (3,)))
    na = az.size
    force = np.zeros(3 * na)
    u = np.zeros(1)
    vir = np.zeros(1)
    lib().orc_ewald_rec_std(_d(bl), float(alpha), int(ncut), int(cube), na,
                            _d(r), _d(az), float(epsi_four_pi), _d(force),
                            _d(u), _d(vir))
    return force.reshape(na, 3), float(u[0]), float(vir[0])


def ewald_self(alpha, az):
    u = np.zeros(1)
    az = f64(az)
    lib().orc_ewald_self(float(alpha), az.size, _d(az), _d(u))
    return float(u[0])


def ewald_surf(boxlen, r, az, epsi_four_pi):
    r = f64(r).reshape(-1)
    az = f64(az)
    bl = f64(np.broadcast_to(np.asarray(boxlen, float), (3,)))
    force = np.zeros(r.size)
    u = np.zeros(1)
    vir = np.zeros(1)
    lib().orc_ewald_surf(_d(bl), az.size, _d(r), _d(az), float(epsi_four_pi),
                         _d(force), _d(u), _d(vir))
    return force.reshape(-1, 3), float(u[0]), float(vir[0])


# ----------------------------------------------------------------- UTotal
class Oracle:
    """UTotal for one monoatomic charged system (Src/energy.F90:280-434).

    Holds the tables, neighbour list and Ewald setup like MolModule does, and
    exposes utotal(ro) -> (forceo[np,3], energies dict).
    """

    def __init__(self, boxlen, iptpn, az, tables: Tables, rcut, *, r2atat=None,
                 ewald="spm", alpha=0.0, epsi_four_pi=None, ncut=8,
                 ncut_cube=False, lsurf=False, order=4, mesh=(32, 32, 32),
                 nlist="vlist", drnlist=2.0, chains=None, bond=(0.0, 2, 0.0),
                 maxnneigh=None, angle=(0.0, 2, 180.0)):
        self.boxlen = f64(np.broadcast_to(np.asarray(boxlen, float), (3,)))
        self.iptpn = i32(iptpn)
        self.az = f64(az)
        self.np = self.az.size
        self.npt = int(self.iptpn.max())
        self.tables = tables
        nptpt = self.npt * (self.npt + 1) // 2
        self.r2atat = f64(np.zeros(nptpt) if r2atat is None else r2atat)
        self.ubuf = f64(tables.ubuf)
        self.iub = i32(tables.iubuflow)
        self.r2umin = f64(tables.r2umin)
        self.rcut = float(rcut)
        self.drnlist = float(drnlist)
        self.nlist = nlist
        self.maxnneigh = maxnneigh
        self.epsi_four_pi = epsi0_four_pi() / 78.4 if epsi_four_pi is None else epsi_four_pi
        self.alpha = alpha
        self.ewald = {"none": 0, "std": 1, "spm": 2}[ewald]
        self.spme = SpmeSetup(self.boxlen, alpha, order, mesh) if ewald == "spm" else None
        self.ncut, self.ncut_cube, self.lsurf = ncut, ncut_cube, lsurf
        if chains is not None:
            self.chain_first = i32([c[0] for c in chains])
            self.chain_len = i32([c[1] for c in chains])
        else:
            self.chain_first = self.chain_len = None
        self.bond = bond
        self.angle = angle          # namelist (k per degree^2, p, eq in degrees)
        self.nn = self.jl = None
        self.cells = None

    def set_list(self, ro):
        """SetVListMD (or InitCellList+SetCellList) at positions ro."""
        ro = f64(ro)
        if self.nlist == "vlist":
            big = self.np > 6000
            self.nn, self.jl = set_vlist_md(self.boxlen, ro,
                                            self.rcut + self.drnlist,
                                            self.maxnneigh, cells=big)
        else:
            self.cells = Cells(self.boxlen, self.rcut + self.drnlist, self.rcut, ro)

    def utotal(self, ro):
        ro = f64(ro).reshape(-1)
        s = System()
        s.box = make_box(self.boxlen)
        s.np, s.npt = self.np, self.npt
        s.iptpn, s.az = _i(self.iptpn), _d(self.az)
        t = TableSet()
        t.ubuf, t.nbuf = _d(self.ubuf), self.ubuf.size
        t.iubuflow, t.r2atat, t.r2umin = _i(self.iub), _d(self.r2atat), _d(self.r2umin)
        t.rcut2 = self.rcut ** 2
        t.npt = self.npt
        s.tab = t
        if self.nlist == "vlist":
            s.nlkind = 0
            s.maxnneigh = self.jl.shape[1]
            s.nneighpn, s.jpnlist = _i(self.nn), _i(self.jl)
        else:
            s.nlkind = 1
            s.cells = C.pointer(self.cells.c)
        s.ewald = self.ewald
        s.alpha, s.epsi_four_pi = self.alpha, self.epsi_four_pi
        s.ncut, s.ncut_cube, s.lsurf = self.ncut, int(self.ncut_cube), int(self.lsurf)
        if self.spme is not None:
            s.spme = C.pointer(self.spme.p)
        if self.chain_first is not None:
            s.nchain = self.chain_first.size
            s.chain_first, s.chain_len = _i(self.chain_first), _i(self.chain_len)
            # namelist k is divided by p at setup (Src/potential.F90:3561)
            kin, pw, eq = float(self.bond[0]), int(self.bond[1]), float(self.bond[2])
            s.bond_k, s.bond_p, s.bond_eq = kin / pw, float(pw), eq
            # angle k per rad^p divided by p, eq in rad (Src/potential.F90:3558-3562)
            sclang = 3.1415926535897932 / 180.0
            ka, pa, eqa = float(self.angle[0]), int(self.angle[1]), float(self.angle[2])
            s.angle_k, s.angle_p, s.angle_eq = ka / sclang ** 2 / pa, float(pa), eqa * sclang
        forceo = np.zeros(3 * self.np)
        e = Energy()
        rc = lib().orc_utotal(C.byref(s), _d(ro), _d(forceo), C.byref(e))
        if rc:
            raise RuntimeError(f"UTotal error {rc}")
        nptpt = self.npt * (self.npt + 1) // 2
        en = dict(tot=e.tot, twob0=e.twob0, twob=np.array(e.twob[:nptpt]),
                  rec=e.rec, bond=e.bond, angle=e.angle, virial=e.virial)
        return forceo.reshape(self.np, 3), en


def uangle(boxlen, ro, chains, k, p, eq):
    """UAngle (Src/energy.F90:5409-5481) of chains [(first, len), ...]
    (1-based) with the namelist k (per degree^2 whatever p is), p and eq (degrees).
    Returns (force[np,3] in kJ/mol/A, u_angle)."""
    ro = f64(ro).reshape(-1)
    b = make_box(boxlen)
    first = i32([c[0] for c in chains])
    ln = i32([c[1] for c in chains])
    sclang = 3.1415926535897932 / 180.0
    force = np.zeros_like(ro)
    u = C.c_double(0.0)
    lib().orc_uangle(C.byref(b), _d(ro), first.size, _i(first), _i(ln),
                     float(k) / sclang ** 2 / int(p), float(int(p)), float(eq) * sclang,
                     _d(force), C.byref(u))
    return force.reshape(-1, 3), u.value


def ubond(boxlen, ro, chains, k, p, eq):
    """UBond (Src/energy.F90:5335-5400) of chains [(first, len), ...]
    (1-based), namelist k (divided by p here, Src/potential.F90:3561), p and
    eq: (force[np, 3] in kJ/(mol A), u_bond)."""
    ro = f64(ro).reshape(-1)
    b = make_box(boxlen)
    first = i32([c[0] for c in chains])
    ln = i32([c[1] for c in chains])
    force = np.zeros_like(ro)
    u = np.zeros(1)
    vir = np.zeros(1)
    lib().orc_ubond(C.byref(b), _d(ro), first.size, _i(first), _i(ln), float(k) / int(p),
                    float(int(p)), float(eq), _d(force), _d(u), _d(vir))
    return force.reshape(-1, 3), float(u[0])


def model_forces(boxlen, iptpn, ro, zat, sigat, epsat, efp, alpha, rcut, *, spme=None,
                 lsurf=False, chains=None, bond=(0.0, 2, 0.0), angle=(0.0, 2, 180.0),
                 knots=None, lewald=True, lljcut=True):
    """Forces of the model the engine's force matching uses: the pair model
    (model_pair_forces), SPME (spme_rec with the SpmeSetup `spme`), the
    surface term, bonds and angles -- -dE/dro [np, 3] in kJ/(mol A)."""
    _, f = model_pair_forces(boxlen, iptpn, ro, zat, sigat, epsat, efp, alpha, rcut,
                             lewald=lewald, lljcut=lljcut, knots=knots)
    az = f64(zat)[i32(iptpn) - 1]
    if spme is not None:
        f = f + spme_rec(spme, ro, az, efp)[0]
    if lsurf:
        f = f + ewald_surf(boxlen, ro, az, efp)[0]
    if chains:
        f = f + ubond(boxlen, ro, chains, *bond)[0]
        if angle[0] != 0.0:
            f = f + uangle(boxlen, ro, chains, *angle)[0]
    return f


def velver_move(boxlen, tstep, ro, rod, rodd):
    b = make_box(boxlen)
    drostep = np.zeros_like(ro)
    lib().orc_velver_move(C.byref(b), ro.size // 3, float(tstep), _d(ro),
                          _d(f64(rod)), _d(f64(rodd)), _d(drostep))
    return drostep


def velver_kick(tstep, massip, forceo, rod, rodd):
    roddo = np.zeros_like(rodd)
    lib().orc_velver_kick(rod.size // 3, float(tstep), _d(f64(massip)),
                          _d(f64(forceo)), _d(rod), _d(rodd), _d(roddo))


def philox4x32_10(ctr, key):
    """Philox4x32-10 of one counter (4 words) under one key (2 words)."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def bd_noise(seed, np_, step):
    """The engine's device BD noise of one step: [np, 3]."""
    out = np.zeros(3 * np_)
    lib().orc_bd_noise(int(seed), int(np_), int(step), _d(out))
    return out.reshape(-1, 3)


def model_pair_energy(boxlen, iptpn, ro, zat, sigat, epsat, efp, alpha, rcut, *,
                      lewald=True, lljcut=True, ljrcut=None, ljushift=1.0, knots=None):
    """The pair model of the engine's adjoint (exact erfc, brute force):
    (energy, spline-term forces [np, 3]).  knots = (r0, r1, c[nptpt, nknot])."""
    b = make_box(boxlen)
    ro = f64(ro).reshape(-1)
    np_ = ro.size // 3
    if ljrcut is None:
        # ljrcut = Two**(1.0/6.0) with a default-real exponent (Src/potential.F90)
        ljrcut = 2.0 ** float(np.float32(1.0) / np.float32(6.0))
    fk = np.zeros(3 * np_)
    if knots is None:
        nk, r0, r1, c = 0, 0.0, 1.0, np.zeros(1)
    else:
        r0, r1, c = knots
        c = f64(c)
        nk = c.shape[-1]
    e = lib().orc_model_pair_energy(C.byref(b), np_, _i(i32(iptpn)), _d(ro), len(zat),
                                    _d(f64(zat)), _d(f64(sigat)), _d(f64(epsat)), float(efp),
                                    float(alpha), int(lewald), int(lljcut), float(ljrcut),
                                    float(ljushift), float(rcut), int(nk), float(r0),
                                    float(r1), _d(c.reshape(-1)), _d(fk))
    return e, fk.reshape(-1, 3)


def model_pair_forces(boxlen, iptpn, ro, zat, sigat, epsat, efp, alpha, rcut, *,
                      lewald=True, lljcut=True, ljrcut=None, ljushift=1.0, knots=None):
    """The same pair model through a cell grid: (energy, force[np, 3]) with
    force = -dE/dro in kJ/(mol A).  knots = (r0, r1, c[nptpt, nknot])."""
    b = make_box(boxlen)
    ro = f64(ro).reshape(-1)
    np_ = ro.size // 3
    if ljrcut is None:
        ljrcut = 2.0 ** float(np.float32(1.0) / np.float32(6.0))
    f = np.zeros(3 * np_)
    if knots is None:
        nk, r0, r1, c = 0, 0.0, 1.0, np.zeros(1)
    else:
        r0, r1, c = knots
        c = f64(c)
        nk = c.shape[-1]
    e = lib().orc_model_pair_forces(C.byref(b), np_, _i(i32(iptpn)), _d(ro), len(zat),
                                    _d(f64(zat)), _d(f64(sigat)), _d(f64(epsat)), float(efp),
                                    float(alpha), int(lewald), int(lljcut), float(ljrcut),
                                    float(ljushift), float(rcut), int(nk), float(r0),
                                    float(r1), _d(c.reshape(-1)), _d(f))
    return e, f.reshape(-1, 3)


def bd_factors(tstep, temp, dcoeff):
    """BDStep's fac1, fac2 per type (Src/bd.F90:136-139)."""
    d = f64(dcoeff)
    f1, f2 = np.zeros(d.size), np.zeros(d.size)
    lib().orc_bd_factors(d.size, float(tstep), float(temp), _d(d), _d(f1), _d(f2))
    return f1, f2


def bd_step(boxlen, iptpn, fac1, fac2, forceo, noise, ro):
    b = make_box(boxlen)
    drostep = np.zeros_like(ro)
    lib().orc_bd_step(C.byref(b), ro.size // 3, _i(i32(iptpn)), _d(f64(fac1)),
                      _d(f64(fac2)), _d(f64(forceo)), _d(f64(noise)), _d(ro),
                      _d(drostep))
    return drostep
