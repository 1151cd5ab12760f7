"""Python host of the C ABI in include/mfe.h (libmfe.so, gfx950).

Mirrors MOLSIM's own call structure for this path: PotTwoBodyTab1 builds
the tables (`build_tables`), EwaldSetup/IONList/Particle set the engine up,
SetVList builds the list (`build_list`), UTotal evaluates forces and
energies (`utotal`), and MDDriver/BDDriver advance the state (`step_md`,
`step_bd`).  Errors raise MfeError carrying the reference's abort condition.

There is no fallback: if libmfe.so cannot be loaded the import of the
engine fails.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmfe.so")

MFE_OK = 0
ERRORS = {1: "MFE_E_BAD_PARAM", 2: "MFE_E_OVERLAP_R2UMIN", 3: "MFE_E_TABLE_RANGE",
          4: "MFE_E_NEIGH_CAPACITY", 5: "MFE_E_MESH_ODD", 6: "MFE_E_HIP",
          7: "MFE_E_STATE", 8: "MFE_E_TABLE_BUILD", 9: "MFE_E_NOMEM"}
PREC = {"fp64": 0, "mixed": 1}
EWALD = {"none": 0, "std": 1, "spm": 2}
NLIST = {"vlist": 0, "cell": 1}
FORCES, ENERGY, VIRIAL = 1, 2, 4
STAGES = {"pair": 0, "spread": 1, "fft_x": 2, "fft_y": 3, "conv": 4,
          "gather": 5, "utotal": 6, "nlist": 7, "ifft": 8, "bin": 9}
MAX_TYPES = 8
MAX_PAIRS = MAX_TYPES * (MAX_TYPES + 1) // 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MfeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32),
                ("max_np", C.c_int32), ("reserved", C.c_int32)]


class EnergyStruct(C.Structure):
    _fields_ = [("tot", C.c_double), ("twob0", C.c_double), ("rec", C.c_double),
                ("bond", C.c_double), ("virial", C.c_double),
                ("nptpt", C.c_int32), ("rebuilds", C.c_int32),
                ("twob", C.c_double * MAX_PAIRS)]


class PotSpec(C.Structure):
    _fields_ = [("npt", C.c_int32), ("zat", _dp), ("sigat", _dp), ("epsat", _dp),
                ("epsi_four_pi", C.c_double), ("rcut", C.c_double),
                ("alpha", C.c_double), ("lewald", C.c_int32), ("lljcut", C.c_int32),
                ("ljrcut", C.c_double), ("ljushift", C.c_double),
                ("r2uminin", C.c_double), ("utoltab", C.c_double),
                ("ftoltab", C.c_double), ("umaxtab", C.c_double),
                ("fmaxtab", C.c_double), ("rumaxfac", C.c_double)]


class EwaldOpt(C.Structure):
    _fields_ = [("iewaldopt", C.c_int32), ("spme", C.c_int32), ("lq2sum", C.c_int32),
                ("ncut", C.c_int32), ("q2sum", C.c_double), ("boxlenshort", C.c_double),
                ("epsi_four_pi", C.c_double), ("uewaldtol", C.c_double),
                ("ualphared", C.c_double), ("ualpha", C.c_double), ("rcut", C.c_double)]


# every symbol include/mfe.h declares
SYMBOLS = ["mfe_abi_version", "mfe_pot_defaults", "mfe_build_tables_1_6_12",
           "mfe_create", "mfe_destroy", "mfe_last_error", "mfe_set_box",
           "mfe_set_particles", "mfe_set_tables", "mfe_set_ewald",
           "mfe_set_neighbor", "mfe_set_chains", "mfe_set_angles",
           "mfe_get_angle_energy", "mfe_upload_positions",
           "mfe_upload_velocities", "mfe_download_state", "mfe_build_list",
           "mfe_compute", "mfe_step_md", "mfe_step_bd", "mfe_get_energy",
           "mfe_get_pairs", "mfe_time_stage", "mfe_set_bd", "mfe_download_forces",
           "mfe_batch_create", "mfe_batch_destroy", "mfe_batch_last_error",
           "mfe_batch_failed_replica", "mfe_batch_build_list", "mfe_batch_compute",
           "mfe_batch_step_md", "mfe_batch_step_bd", "mfe_batch_get_energies",
           "mfe_batch_set_persistent", "mfe_batch_uses_persistent",
           "mfe_set_pair_params", "mfe_set_knots", "mfe_adjoint", "mfe_mc_delta_rec",
           "mfe_mc_delta_pair", "mfe_mc_accept", "mfe_ewald_parameters",
           "mfe_set_sync_limit", "mfe_set_net_charge", "mfe_model_forces", "mfe_fm_set_frames", "mfe_fm_loss_grad",
           "mfe_mc_delta_pair_noalloc"]

_lib = None


def lib():
    """Load libmfe.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built: run python -m molsim_amd.build")
        L = C.CDLL(LIB_PATH)
        L.mfe_abi_version.restype = C.c_int
        L.mfe_pot_defaults.argtypes = [C.POINTER(PotSpec)]
        L.mfe_build_tables_1_6_12.argtypes = [C.POINTER(PotSpec), _dp, C.c_int32,
                                              _ip, _ip, _ip, _dp]
        L.mfe_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.mfe_destroy.argtypes = [C.c_void_p]
        L.mfe_last_error.argtypes = [C.c_void_p]
        L.mfe_last_error.restype = C.c_char_p
        L.mfe_set_sync_limit.argtypes = [C.c_void_p, C.c_int32]
        L.mfe_set_net_charge.argtypes = [C.c_void_p, C.c_int32]
        L.mfe_model_forces.argtypes = [C.c_void_p, _dp]
        L.mfe_fm_set_frames.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
        L.mfe_fm_loss_grad.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.mfe_set_box.argtypes = [C.c_void_p, _dp]
        L.mfe_set_particles.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, _dp, _dp]
        L.mfe_set_tables.argtypes = [C.c_void_p, _dp, C.c_int32, _ip, _dp, _dp, C.c_double]
        L.mfe_set_ewald.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32, _ip,
                                    C.c_int32, C.c_int32, C.c_int32, C.c_double]
        L.mfe_set_neighbor.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32]
        L.mfe_set_angles.argtypes = [C.c_void_p, C.c_double, C.c_int32, C.c_double]
        L.mfe_get_angle_energy.argtypes = [C.c_void_p, _dp]
        L.mfe_set_chains.argtypes = [C.c_void_p, C.c_int32, _ip, _ip, C.c_double,
                                     C.c_int32, C.c_double]
        L.mfe_upload_positions.argtypes = [C.c_void_p, _dp]
        L.mfe_upload_velocities.argtypes = [C.c_void_p, _dp, _dp]
        L.mfe_download_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.mfe_build_list.argtypes = [C.c_void_p]
        L.mfe_compute.argtypes = [C.c_void_p, C.c_uint32, _dp, C.POINTER(EnergyStruct)]
        L.mfe_mc_delta_rec.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, C.c_int32, _dp]
        L.mfe_mc_delta_pair.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp, _ip]
        L.mfe_mc_accept.argtypes = [C.c_void_p, C.c_int32, _ip, _dp]
        L.mfe_step_md.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_uint32]
        L.mfe_step_bd.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp, C.c_uint32]
        L.mfe_get_energy.argtypes = [C.c_void_p, C.POINTER(EnergyStruct)]
        L.mfe_get_pairs.argtypes = [C.c_void_p, _ip, _ip, C.c_int32]
        L.mfe_time_stage.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp]
        L.mfe_set_bd.argtypes = [C.c_void_p, _dp, _dp, C.c_uint64]
        L.mfe_download_forces.argtypes = [C.c_void_p, _dp]
        L.mfe_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
        L.mfe_batch_destroy.argtypes = [C.c_void_p]
        L.mfe_batch_last_error.argtypes = [C.c_void_p]
        L.mfe_batch_last_error.restype = C.c_char_p
        L.mfe_batch_failed_replica.argtypes = [C.c_void_p]
        L.mfe_batch_failed_replica.restype = C.c_int32
        L.mfe_batch_build_list.argtypes = [C.c_void_p]
        L.mfe_batch_compute.argtypes = [C.c_void_p, C.c_uint32]
        L.mfe_batch_step_md.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_uint32]
        L.mfe_batch_step_bd.argtypes = [C.c_void_p, C.c_int32, C.c_uint32]
        L.mfe_batch_set_persistent.argtypes = [C.c_void_p, C.c_int32]
        L.mfe_batch_uses_persistent.argtypes = [C.c_void_p]
        L.mfe_batch_uses_persistent.restype = C.c_int32
        L.mfe_batch_get_energies.argtypes = [C.c_void_p, C.POINTER(EnergyStruct)]
        L.mfe_set_pair_params.argtypes = [C.c_void_p, C.POINTER(PotSpec)]
        L.mfe_set_knots.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, _dp]
        L.mfe_adjoint.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ tables
@dataclass
class Tables:
    """PotTwoBodyTab1 output: ubuf and, per type pair, iubuflow (1-based),
    nugrid and r2umin (Src/potential.F90:868-1131)."""
    ubuf: np.ndarray
    iubuflow: np.ndarray
    nugrid: np.ndarray
    r2umin: np.ndarray


def ewald_parameters(iewaldopt, *, q2sum, boxlenshort, epsi_four_pi, lq2sum=0, spme=False,
                     uewaldtol=0.0, ualphared=0.0, ualpha=0.0, rcut=0.0, ncut=0) -> dict:
    """EwaldSetup's parameter choice for iewaldopt 0-4 (mfe_ewald_parameters):
    the inputs the option uses, the rest determined; returns every field."""
    p = EwaldOpt(int(iewaldopt), int(bool(spme)), int(lq2sum), int(ncut), float(q2sum),
                 float(boxlenshort), float(epsi_four_pi), float(uewaldtol), float(ualphared),
                 float(ualpha), float(rcut))
    L = lib()
    L.mfe_ewald_parameters.argtypes = [C.POINTER(EwaldOpt)]
    rc = L.mfe_ewald_parameters(C.byref(p))
    if rc:
        raise MfeError(rc, "mfe_ewald_parameters: input out of range or no convergence")
    return {k: getattr(p, k) for k, _ in EwaldOpt._fields_}


def pot_defaults() -> PotSpec:
    s = PotSpec()
    lib().mfe_pot_defaults(C.byref(s))
    return s


def pot_spec(zat, sigat, epsat, *, rcut, epsi_four_pi, alpha=0.0, lewald=True,
             lljcut=True, **overrides):
    """(spec, kept arrays): IOPotTwoBody's (1,6,12) specification."""
    zat, sigat, epsat = _f64(zat), _f64(sigat), _f64(epsat)
    s = pot_defaults()
    s.npt = zat.size
    s.zat, s.sigat, s.epsat = _d(zat), _d(sigat), _d(epsat)
    s.epsi_four_pi, s.rcut, s.alpha = epsi_four_pi, rcut, alpha
    s.lewald, s.lljcut = int(lewald), int(lljcut)
    for k, v in overrides.items():
        setattr(s, k, v)
    return s, (zat, sigat, epsat)


def build_tables(zat, sigat, epsat, *, rcut, epsi_four_pi, alpha=0.0,
                 lewald=True, lljcut=True, **overrides) -> Tables:
    """Tables of the (1,6,12) potential for monoatomic types."""
    s, _keep = pot_spec(zat, sigat, epsat, rcut=rcut, epsi_four_pi=epsi_four_pi, alpha=alpha,
                        lewald=lewald, lljcut=lljcut, **overrides)
    nptpt = s.npt * (s.npt + 1) // 2
    cap = 200000
    ubuf = np.zeros(cap)
    n = C.c_int32()
    iub = np.zeros(nptpt, np.int32)
    nug = np.zeros(nptpt, np.int32)
    r2 = np.zeros(nptpt)
    rc = lib().mfe_build_tables_1_6_12(C.byref(s), _d(ubuf), cap, C.byref(n),
                                       _i(iub), _i(nug), _d(r2))
    if rc:
        raise MfeError(rc, "table build failed")
    return Tables(ubuf[:n.value].copy(), iub, nug, r2)


# ------------------------------------------------------------------ units
def bd_factors(tstep, temp, dcoeff):
    """BDStep's per-type factors (Src/bd.F90:136-139) with MOLSIM's default
    units: fac1 = dt D/(R T) (A per unit of forceo), fac2 = sqrt(2 D dt)
    (A); tstep in ps, temp in K, dcoeff in 1e-9 m^2/s.  Evaluated left to
    right as the reference writes them."""
    scllen, sclmas, scltem, scltim, scldif = 1.0e-10, 1.0e-03, 1.0, 1.0e-12, 1.0e-09
    sclfor = scllen / scltim / scltim * sclmas
    gas = 8.314472
    d = np.asarray(dcoeff, dtype=np.float64)
    f1 = np.array([tstep * scltim * v * scldif * sclfor / (gas * temp * scltem) / scllen
                   for v in d])
    f2 = np.array([np.sqrt(2.0 * v * scldif * tstep * scltim) / scllen for v in d])
    return f1, f2


def epsi0_four_pi() -> float:
    """Epsi0FourPi in kJ A/mol (Src/molsim.F90:805, constants Src/mol.F90:70-76)."""
    pi = 3.1415926535897932
    ech, eps0, avno = 1.602176487e-19, 8.854187817e-12, 6.02214179e23
    return ((ech * ech) / ((4.0 * pi * eps0) * 1.0e-10)) * avno / 1.0e+03


def force_unit_fac() -> float:
    """UTotal's fac = (sclene/scllen)/sclfor (Src/energy.F90:408)."""
    sclvel = 1.0e-10 / 1.0e-12
    return (1.0e+03 / 1.0e-10) / ((sclvel / 1.0e-12) * 1.0e-03)


# ------------------------------------------------------------------ engine
@dataclass
class Energies:
    tot: float
    twob0: float
    rec: float
    bond: float
    virial: float
    twob: np.ndarray
    rebuilds: int


class ForceEngine:
    """One engine = one device stream and its device-resident system."""

    def __init__(self, max_np: int, precision: str = "mixed", device: int = 0):
        self.L = lib()
        cfg = Config(PREC[precision], device, int(max_np), 0)
        h = C.c_void_p()
        self._check(self.L.mfe_create(C.byref(cfg), C.byref(h)), "mfe_create", handle=None)
        self.h = h
        self.precision = precision
        self.np = 0

    def _check(self, rc, what, handle="self"):
        if rc != MFE_OK:
            msg = ""
            if handle == "self" and getattr(self, "h", None):
                msg = self.L.mfe_last_error(self.h).decode()
            raise MfeError(rc, f"{what}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mfe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- setup
    def set_box(self, boxlen):
        b = _f64(np.broadcast_to(np.asarray(boxlen, float), (3,)))
        self._check(self.L.mfe_set_box(self.h, _d(b)), "mfe_set_box")

    def set_particles(self, iptpn, az, massipt):
        iptpn, az, massipt = _i32(iptpn), _f64(az), _f64(massipt)
        self.np = iptpn.size
        self.npt = massipt.size
        self._check(self.L.mfe_set_particles(self.h, self.np, self.npt, _i(iptpn), _d(az),
                                             _d(massipt)), "mfe_set_particles")

    def set_tables(self, tables: Tables, rcut, r2atat=None):
        nptpt = tables.iubuflow.size
        r2atat = _f64(np.zeros(nptpt) if r2atat is None else r2atat)
        ub = _f64(tables.ubuf)
        self._check(self.L.mfe_set_tables(self.h, _d(ub), ub.size, _i(_i32(tables.iubuflow)),
                                          _d(r2atat), _d(_f64(tables.r2umin)), float(rcut)),
                    "mfe_set_tables")

    def set_ewald(self, kind="spm", alpha=0.0, order=4, mesh=(32, 32, 32), ncut=8,
                  ncut_cube=False, lsurf=False, epsi_four_pi=None):
        mesh = _i32(np.broadcast_to(np.asarray(mesh), (3,)))
        efp = epsi0_four_pi() / 78.4 if epsi_four_pi is None else epsi_four_pi
        self._check(self.L.mfe_set_ewald(self.h, EWALD[kind], float(alpha), int(order),
                                         _i(mesh), int(ncut), int(ncut_cube), int(lsurf),
                                         float(efp)), "mfe_set_ewald")

    def set_neighbor(self, kind="vlist", drnlist=2.0, maxnneigh=0):
        self._check(self.L.mfe_set_neighbor(self.h, NLIST[kind], float(drnlist),
                                            int(maxnneigh)), "mfe_set_neighbor")

    def set_chains(self, chains, k, p, eq):
        first = _i32([c[0] for c in chains])
        ln = _i32([c[1] for c in chains])
        self._check(self.L.mfe_set_chains(self.h, first.size, _i(first), _i(ln), float(k),
                                          int(p), float(eq)), "mfe_set_chains")

    def set_angles(self, k, p, eq):
        """UAngle of the chains: namelist k (per degree^2), p, eq (degrees)."""
        self._check(self.L.mfe_set_angles(self.h, float(k), int(p), float(eq)),
                    "mfe_set_angles")

    def angle_energy(self):
        """u%angle of the last call that reduced energies."""
        u = np.zeros(1)
        self._check(self.L.mfe_get_angle_energy(self.h, _d(u)), "mfe_get_angle_energy")
        return float(u[0])

    # --- state
    def upload_positions(self, ro):
        ro = _f64(ro).reshape(-1)
        self._check(self.L.mfe_upload_positions(self.h, _d(ro)), "mfe_upload_positions")

    def upload_velocities(self, rod=None, rodd=None):
        a = None if rod is None else _f64(rod).reshape(-1)
        b = None if rodd is None else _f64(rodd).reshape(-1)
        self._check(self.L.mfe_upload_velocities(self.h, _d(a) if a is not None else None,
                                                 _d(b) if b is not None else None),
                    "mfe_upload_velocities")

    def download_state(self):
        ro = np.zeros(3 * self.np)
        rod = np.zeros(3 * self.np)
        rodd = np.zeros(3 * self.np)
        self._check(self.L.mfe_download_state(self.h, _d(ro), _d(rod), _d(rodd)),
                    "mfe_download_state")
        return ro.reshape(-1, 3), rod.reshape(-1, 3), rodd.reshape(-1, 3)

    # --- hot path
    def build_list(self):
        self._check(self.L.mfe_build_list(self.h), "mfe_build_list")

    def utotal(self, flags=FORCES | ENERGY | VIRIAL):
        """UTotal: (forceo[np,3], Energies)."""
        f = np.zeros(3 * self.np)
        e = EnergyStruct()
        self._check(self.L.mfe_compute(self.h, flags, _d(f), C.byref(e)), "mfe_compute")
        return f.reshape(-1, 3), self._energies(e)

    def mc_delta_rec(self, ianatm, rtm, reuse_current=False):
        """DUTwoBodyEwaldRecSPM: change of the SPME reciprocal energy (units of
        Energies.rec) when particles ianatm (1-based) move to rtm[natm, 3].
        reuse_current: start from the mesh and energy of the previous call
        (valid only if nothing changed the configuration since)."""
        ia = _i32(np.atleast_1d(ianatm))
        r = _f64(np.asarray(rtm, dtype=np.float64).reshape(-1))
        if r.size != 3 * ia.size:
            raise ValueError("rtm must hold three coordinates per moved particle")
        du = np.zeros(1)
        self._check(self.L.mfe_mc_delta_rec(self.h, ia.size, _i(ia), _d(r),
                                            1 if reuse_current else 0, _d(du)),
                    "mfe_mc_delta_rec")
        return float(du[0])

    def mc_delta_pair(self, ianatm, rtm):
        """UTwoBodyANew - UTwoBodyAOld: (change of the two-body energy, overlap)
        when particles ianatm (1-based) move to rtm[natm, 3]."""
        ia = _i32(np.atleast_1d(ianatm))
        r = _f64(np.asarray(rtm, dtype=np.float64).reshape(-1))
        if r.size != 3 * ia.size:
            raise ValueError("rtm must hold three coordinates per moved particle")
        du = np.zeros(1)
        ov = np.zeros(1, np.int32)
        self._check(self.L.mfe_mc_delta_pair(self.h, ia.size, _i(ia), _d(r), _d(du), _i(ov)),
                    "mfe_mc_delta_pair")
        return float(du[0]), bool(ov[0])

    def mc_accept(self, ianatm, rtm):
        """Accepted move (EwaldUpdateArray + ro(:, ianatm) = rtm + NList):
        after mc_delta_rec of the same move, the next mc_delta_rec may pass
        reuse_current=True."""
        ia = _i32(np.atleast_1d(ianatm))
        r = _f64(np.asarray(rtm, dtype=np.float64).reshape(-1))
        if r.size != 3 * ia.size:
            raise ValueError("rtm must hold three coordinates per moved particle")
        self._check(self.L.mfe_mc_accept(self.h, ia.size, _i(ia), _d(r)), "mfe_mc_accept")

    def step_md(self, nsteps, tstep, flags=FORCES | ENERGY | VIRIAL):
        self._check(self.L.mfe_step_md(self.h, int(nsteps), float(tstep), flags), "mfe_step_md")

    def set_net_charge(self, on=True):
        """Neutralising-background term of a net charge (opt-in)."""
        self._check(self.L.mfe_set_net_charge(self.h, 1 if on else 0), "mfe_set_net_charge")

    def set_sync_limit(self, polls):
        """Poll bound of the list rebuild's grid-barrier waits (< 0: default)."""
        self._check(self.L.mfe_set_sync_limit(self.h, int(polls)), "mfe_set_sync_limit")

    def set_bd(self, fac1, fac2, seed=0):
        """BDStep factors per type and the seed of the device noise stream."""
        self._check(self.L.mfe_set_bd(self.h, _d(_f64(fac1)), _d(_f64(fac2)), int(seed)),
                    "mfe_set_bd")

    def step_bd(self, fac1=None, fac2=None, noise=None, nsteps=None,
                flags=FORCES | ENERGY | VIRIAL):
        """BDStep x nsteps: with host noise[nsteps, np, 3], or (noise None)
        the device's counter-based stream for nsteps steps."""
        f1 = None if fac1 is None else _f64(fac1)
        f2 = None if fac2 is None else _f64(fac2)
        if noise is not None:
            noise = _f64(noise)
            nsteps = noise.size // (3 * self.np)
        self._check(self.L.mfe_step_bd(self.h, int(nsteps),
                                       _d(f1) if f1 is not None else None,
                                       _d(f2) if f2 is not None else None,
                                       _d(noise) if noise is not None else None, flags),
                    "mfe_step_bd")

    def set_pair_params(self, zat, sigat, epsat, *, rcut, epsi_four_pi, alpha=0.0,
                        lewald=True, lljcut=True, **overrides):
        """Parameters of the pair model the adjoint differentiates."""
        s, keep = pot_spec(zat, sigat, epsat, rcut=rcut, epsi_four_pi=epsi_four_pi,
                           alpha=alpha, lewald=lewald, lljcut=lljcut, **overrides)
        self._check(self.L.mfe_set_pair_params(self.h, C.byref(s)), "mfe_set_pair_params")
        del keep
        self.nptpt = self.npt * (self.npt + 1) // 2

    def set_knots(self, r_lo, r_hi, c):
        """Additive cubic B-spline pair term c[nptpt, nknot] on [r_lo, r_hi)
        (c None removes it)."""
        if c is None:
            self._check(self.L.mfe_set_knots(self.h, 0, 0.0, 0.0, None), "mfe_set_knots")
            self.nknot = 0
            return
        c = _f64(c)
        self.nknot = c.shape[-1]
        self._check(self.L.mfe_set_knots(self.h, self.nknot, float(r_lo), float(r_hi),
                                         _d(c.reshape(-1))), "mfe_set_knots")

    def adjoint(self):
        """Model energy and its gradient: dict(E, dq[np], dzat[npt],
        dsigat[npt], depsat[npt], dknot[nptpt, nknot], dro[np, 3])."""
        nk = getattr(self, "nknot", 0)
        nptpt = self.npt * (self.npt + 1) // 2
        out = dict(dq=np.zeros(self.np), dzat=np.zeros(self.npt), dsigat=np.zeros(self.npt),
                   depsat=np.zeros(self.npt), dknot=np.zeros(max(nptpt * nk, 1)),
                   dro=np.zeros(3 * self.np))
        e = C.c_double()
        self._check(self.L.mfe_adjoint(self.h, _d(out["dq"]), _d(out["dzat"]),
                                       _d(out["dsigat"]), _d(out["depsat"]),
                                       _d(out["dknot"]), _d(out["dro"]), C.byref(e)),
                    "mfe_adjoint")
        out["dknot"] = out["dknot"][:nptpt * nk].reshape(nptpt, nk)
        out["dro"] = out["dro"].reshape(-1, 3)
        out["E"] = e.value
        return out

    def model_forces(self):
        """Forces of the adjoint's model at the current positions:
        -dE/dro [np, 3] in kJ/(mol A)."""
        f = np.zeros(3 * self.np)
        self._check(self.L.mfe_model_forces(self.h, _d(f)), "mfe_model_forces")
        return f.reshape(-1, 3)

    def fm_set_frames(self, ro, f_target):
        """Frames ro[nf, np, 3] and their target forces (kJ/(mol A)), kept
        on the device for fm_loss_grad."""
        ro = _f64(ro).reshape(-1)
        ft = _f64(f_target).reshape(-1)
        nf = ro.size // (3 * self.np)
        if nf < 1 or ro.size != 3 * self.np * nf or ft.size != ro.size:
            raise ValueError("frames and targets must be [nf, np, 3]")
        self._check(self.L.mfe_fm_set_frames(self.h, nf, _d(ro), _d(ft)), "mfe_fm_set_frames")
        self.fm_nframes = nf

    def fm_loss_grad(self):
        """Force-matching loss over the stored frames and its gradient:
        dict(loss, dzat[npt], dsigat[npt], depsat[npt], dknot[nptpt, nknot])."""
        nk = getattr(self, "nknot", 0)
        nptpt = self.npt * (self.npt + 1) // 2
        out = dict(dzat=np.zeros(self.npt), dsigat=np.zeros(self.npt),
                   depsat=np.zeros(self.npt), dknot=np.zeros(max(nptpt * nk, 1)))
        loss = C.c_double()
        self._check(self.L.mfe_fm_loss_grad(self.h, C.byref(loss), _d(out["dzat"]),
                                            _d(out["dsigat"]), _d(out["depsat"]),
                                            _d(out["dknot"])), "mfe_fm_loss_grad")
        out["dknot"] = out["dknot"][:nptpt * nk].reshape(nptpt, nk)
        out["loss"] = loss.value
        return out

    def download_forces(self):
        f = np.zeros(3 * self.np)
        self._check(self.L.mfe_download_forces(self.h, _d(f)), "mfe_download_forces")
        return f.reshape(-1, 3)

    def energy(self) -> Energies:
        e = EnergyStruct()
        self._check(self.L.mfe_get_energy(self.h, C.byref(e)), "mfe_get_energy")
        return self._energies(e)

    def get_pairs(self, maxnneigh=None):
        """SetVListMD's (nneighpn, jpnlist) from the device list (1-based)."""
        m = maxnneigh or 1024
        nn = np.zeros(self.np, np.int32)
        jl = np.zeros((self.np, m), np.int32)
        self._check(self.L.mfe_get_pairs(self.h, _i(nn), _i(jl), m), "mfe_get_pairs")
        return nn, jl

    def time_stage(self, stage, nrep=20) -> float:
        ms = C.c_double()
        self._check(self.L.mfe_time_stage(self.h, STAGES[stage], int(nrep), C.byref(ms)),
                    "mfe_time_stage")
        return ms.value

    @staticmethod
    def _energies(e: EnergyStruct) -> Energies:
        return Energies(e.tot, e.twob0, e.rec, e.bond, e.virial,
                        np.array(e.twob[:e.nptpt]), e.rebuilds)


class Batch:
    """Replicas driven together (mfe_batch_*): every stage is one launch
    over all replicas.  The engines stay owned by the caller and must
    outlive the batch."""

    def __init__(self, engines):
        self.L = lib()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        rc = self.L.mfe_batch_create(arr, len(self.engines), C.byref(h))
        if rc != 0:
            raise MfeError(rc, "mfe_batch_create failed (see stderr)")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.mfe_batch_last_error(self.h).decode()
            err = MfeError(rc, f"{what}: {msg}")
            err.replica = self.L.mfe_batch_failed_replica(self.h)
            raise err

    def close(self):
        if getattr(self, "h", None):
            self.L.mfe_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def build_list(self):
        self._check(self.L.mfe_batch_build_list(self.h), "mfe_batch_build_list")

    def compute(self, flags=FORCES | ENERGY | VIRIAL):
        self._check(self.L.mfe_batch_compute(self.h, flags), "mfe_batch_compute")

    def step_md(self, nsteps, tstep, flags=FORCES | ENERGY | VIRIAL):
        self._check(self.L.mfe_batch_step_md(self.h, int(nsteps), float(tstep), flags),
                    "mfe_batch_step_md")

    def step_bd(self, nsteps, flags=FORCES | ENERGY | VIRIAL):
        self._check(self.L.mfe_batch_step_bd(self.h, int(nsteps), flags), "mfe_batch_step_bd")

    def set_persistent(self, on=True):
        """Let step_bd take the persistent launch where eligible (default)."""
        self._check(self.L.mfe_batch_set_persistent(self.h, 1 if on else 0),
                    "mfe_batch_set_persistent")

    def uses_persistent(self):
        """Whether step_bd takes the persistent launch for this batch."""
        return bool(self.L.mfe_batch_uses_persistent(self.h))

    def energies(self):
        arr = (EnergyStruct * len(self.engines))()
        self._check(self.L.mfe_batch_get_energies(self.h, arr), "mfe_batch_get_energies")
        return [ForceEngine._energies(e) for e in arr]
