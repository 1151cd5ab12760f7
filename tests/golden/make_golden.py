"""Generate the golden fixtures in tests/golden/ from the MOLSIM reference.

Runs the amdflang build of the unmodified reference (oracle/ref.mk builds it
from /root/reference into oracle/_ref/) on small inputs written here, and
stores what it computes -- potential tables, UTotal forces/energies/virial,
Verlet-list digests, MD trajectories, Erf2 and CardinalBSpline values -- as
compressed .npz fixtures.  Only the data is committed; this script needs the
reference tree and is run by hand in the build container:

    make -f oracle/ref.mk REF=/root/reference -j8
    python tests/golden/make_golden.py

The reference is run in its serial, FFTW-free configuration, so SPME cannot
be exercised; the SPME oracle is pinned by closed forms, numpy.fft and the
direct Ewald sum instead (tests/test_oracle_math.py).
"""
from __future__ import annotations

import hashlib
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(REPO, "oracle", "_ref")
MOLSIM = os.path.join(REF_DIR, "molsim")
PROBE = os.path.join(REF_DIR, "ref_probe")

KT298 = 2.4777  # kJ/mol at 298 K


def records(path):
    """Fortran unformatted sequential records (4-byte markers)."""
    b = open(path, "rb").read()
    i, out = 0, []
    while i < len(b):
        n = struct.unpack("<i", b[i:i + 4])[0]
        out.append(b[i + 4:i + 4 + n])
        i += 8 + n
    return out


def system_nml(title, np_list, boxlen, nstep2=1, dump=True, method="md"):
    ens = "nvt" if method == "bd" else "nve"
    return (f" &nmlSystem txtitle='{title}', txmethod='{method}', txensemb='{ens}',"
            f" txbc='xyz', txstart='setconf', nstep1=1, nstep2={nstep2},"
            f" boxlen=3*{boxlen}, temp=298.0, prsr=0.1013, iseed=1,"
            f" ldump={'.true.' if dump else '.false.'}, itest=0, ipart=0,"
            f" iatom=0, iaver=0, ishow=0, iplot=0, ilist=0, ltime=.false. /\n"
            " &nmlScale /\n")


def particle_nml(types, chain=None):
    """types: list of (name, count, mass, charge, sigma, eps)."""
    n = len(types)
    s = " &nmlParticle"
    if chain is not None:
        nct, ncct, nper = chain
        s += f" nct=1, txct='chain', ncct={ncct}, npptct(1,1)={nper},"
    s += f" npt={n}, txpt=" + ",".join(f"'{t[0]}'" for t in types)
    s += ", nppt=" + ",".join(str(t[1]) for t in types)
    s += ", natpt=" + ",".join("1" for _ in types)
    s += ", txat=" + ",".join(f"'{t[0][:2]}'" for t in types)
    s += ", massat=" + ",".join(str(t[2]) for t in types)
    s += ", radat=" + ",".join("0.0" for _ in types)
    s += ", zat=" + ",".join(str(t[3]) for t in types)
    s += ", sigat=" + ",".join(str(t[4]) for t in types)
    s += ", epsat=" + ",".join(str(t[5]) for t in types)
    for i, t in enumerate(types, 1):
        s += f", naatpt(1,{i})=1, txaat(1,{i})='{t[0][:2]}'"
    return s + " /\n"


def run(exe, proj, workdir):
    r = subprocess.run([exe, proj], cwd=workdir, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{exe} {proj} failed:\n{r.stdout}\n{r.stderr}")


def read_tab(path):
    b = open(path, "rb").read()
    nused, nptpt = struct.unpack("<ii", b[:8])
    o = 8
    ubuf = np.frombuffer(b[o:o + 8 * nused], "<f8").copy()
    o += 8 * nused
    iub = np.frombuffer(b[o:o + 4 * nptpt], "<i4").copy()
    o += 4 * nptpt
    nug = np.frombuffer(b[o:o + 4 * nptpt], "<i4").copy()
    o += 4 * nptpt
    r2umin = np.frombuffer(b[o:o + 8 * nptpt], "<f8").copy()
    o += 8 * nptpt
    rcut, alpha, efp = struct.unpack("<ddd", b[o:o + 24])
    return dict(ubuf=ubuf, iubuflow=iub, nugrid=nug, r2umin=r2umin,
                rcut=rcut, alpha=alpha, epsi_four_pi=efp)


def read_force(path, nptpt):
    b = open(path, "rb").read()
    np_, maxn = struct.unpack("<ii", b[:8])
    o = 8
    force = np.frombuffer(b[o:o + 24 * np_], "<f8").reshape(np_, 3).copy()
    o += 24 * np_
    forceo = np.frombuffer(b[o:o + 24 * np_], "<f8").reshape(np_, 3).copy()
    o += 24 * np_
    ne = nptpt + 5
    ev = np.frombuffer(b[o:o + 8 * ne], "<f8").copy()
    o += 8 * ne
    nn = np.frombuffer(b[o:o + 4 * np_], "<i4").copy()
    o += 4 * np_
    jl = np.frombuffer(b[o:o + 4 * np_ * maxn], "<i4").reshape(np_, maxn)
    flat = np.concatenate([jl[i, :nn[i]] for i in range(np_)]).astype("<i4")
    return dict(force=force, forceo=forceo, u_tot=ev[0], u_twob=ev[1:2 + nptpt],
                u_rec=ev[2 + nptpt], u_bond=ev[3 + nptpt], virial=ev[4 + nptpt],
                nneighpn=nn, jpnlist_sha256=hashlib.sha256(flat.tobytes()).hexdigest())


def dumps(workdir, proj):
    pos = [np.frombuffer(r, "<f8").reshape(-1, 3) for r in records(
        os.path.join(workdir, proj + ".pos"))[1::2]]
    liv = [np.frombuffer(r, "<f8").reshape(-1, 3) for r in records(
        os.path.join(workdir, proj + ".liv"))]
    frc = [np.frombuffer(r, "<f8").reshape(-1, 3) for r in records(
        os.path.join(workdir, proj + ".for"))]
    return np.array(pos), np.array(liv), np.array(frc)


def potential_nml(npt, rcut, ewald="std", ncut=8, lsurf=False, ualphared=3.0,
                  extra=""):
    nptpt = npt * (npt + 1) // 2
    s = (f" &nmlPotential relpermitt=78.4, rcut={rcut}, lljcut=.true.,"
         f" txpot={nptpt}*'(1,6,12)',"
         f" lewald=.true., txewaldrec='{ewald}', iewaldopt=0,"
         f" ualphared={ualphared}, ncut={ncut}, ncutregion='sphere',"
         f" lsurf={'.true.' if lsurf else '.false.'}{extra} /\n")
    return s


def case(name, text, workdir, probe_pos=None, md=True):
    d = os.path.join(workdir, name)
    os.makedirs(d, exist_ok=True)
    open(os.path.join(d, name + ".in"), "w").write(text)
    if md:
        run(MOLSIM, name, d)
    if probe_pos is not None:
        np.asarray(probe_pos, "<f8").tofile(os.path.join(d, name + ".probe_pos"))
    run(PROBE, name, d)
    return d



def c1_text():
    ions = [("cat", 500, 23.0, 1.0, 4.0, KT298), ("an", 500, 35.4, -1.0, 4.0, KT298)]
    body = (particle_nml(ions) + potential_nml(2, 15.0) +
            " &nmlSetConfiguration txsetconf='random','random', radatset=1.6,1.6 /\n"
            " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
            " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n"
            " &nmlDump idump=1, ldpos=.true., ldliv=.true., ldfor=.true. /\n")
    return system_nml("golden c1", [500, 500], 50.0, nstep2=5) + body


def rng_streams():
    """GauRandom and Random (Src/mollib.F90:2689-2766) from fixed seeds, run
    inside the reference probe after the C1 setup: tests/golden/rng.npz."""
    nrng, iseed, idum = 1000, 12345, -4711
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        d = os.path.join(work, "c1")
        os.makedirs(d, exist_ok=True)
        np.array([nrng, iseed, idum], "<i4").tofile(os.path.join(d, "c1.probe_rng"))
        case("c1", c1_text(), work, md=False)
        v = np.fromfile(os.path.join(d, "c1.probe_rngout"), "<f8")
        np.savez_compressed(os.path.join(HERE, "rng.npz"), iseed=iseed, idum=idum,
                            gaurandom=v[:2 * nrng], random=v[2 * nrng:])
    finally:
        shutil.rmtree(work, ignore_errors=True)


def c1_cnf():
    """The configuration file the reference writes at the end of the C1 run
    (IOCnf('write'), Src/molsim.F90:945-955): tests/golden/c1_final.cnf, with
    the run's last dumped positions and velocities for comparison."""
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        d = case("c1", c1_text(), work)
        pos, liv, _ = dumps(d, "c1")
        shutil.copyfile(os.path.join(d, "c1.cnf"), os.path.join(HERE, "c1_final.cnf"))
        np.savez_compressed(os.path.join(HERE, "c1_final.npz"), ro=pos[-1], rod=liv[-1])
    finally:
        shutil.rmtree(work, ignore_errors=True)


def cnf_forces():
    """UTotal of the reference at the positions its C1 run wrote to
    tests/golden/c1_final.cnf (read with molsim_amd.cnf): force, forceo and
    energies -> tests/golden/c1_final_utotal.npz."""
    sys.path.insert(0, REPO)
    from molsim_amd import cnf
    c = cnf.read_cnf(os.path.join(HERE, "c1_final.cnf"))
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        d = case("c1", c1_text(), work, probe_pos=c.ro, md=False)
        f = read_force(os.path.join(d, "c1.probe_f"), 3)
        np.savez_compressed(os.path.join(HERE, "c1_final_utotal.npz"),
                            **{"ref_" + k: v for k, v in f.items()})
        print("cnf forces: u_tot", f["u_tot"])
    finally:
        shutil.rmtree(work, ignore_errors=True)


def ewald_options():
    """EwaldSetup's parameter choice for iewaldopt 0-4 (std,
    Src/energy.F90:6644-6690, 7052-7124) on the C1 system: what the
    reference settles on for each set of inputs (tests/golden/ewald_opt.npz)."""
    ions = [("cat", 500, 23.0, 1.0, 4.0, KT298), ("an", 500, 35.4, -1.0, 4.0, KT298)]
    cases = [(0, ""), (1, ""), (2, ", uewaldtol=1.0d-3, ualpha=0.21"),
             (3, ", uewaldtol=1.0d-3"), (3, ", uewaldtol=1.0d-5"), (4, ", uewaldtol=1.0d-3"),
             (4, ", uewaldtol=1.0d-4, ncut=12")]
    rows_i, rows_d = [], []
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        for k, (opt, extra) in enumerate(cases):
            body = (particle_nml(ions) +
                    potential_nml(2, 15.0, extra=f", iewaldopt={opt}{extra}") +
                    " &nmlSetConfiguration txsetconf='random','random', radatset=1.6,1.6 /\n"
                    " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
                    " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n")
            text = system_nml("golden ewald", [500, 500], 50.0, nstep2=1, dump=False) + body
            d = case(f"e{k}", text, work, md=False)
            b = open(os.path.join(d, f"e{k}.probe_ewald"), "rb").read()
            rows_i.append(struct.unpack("<3i", b[:12]))
            rows_d.append(struct.unpack("<7d", b[12:68]))
        np.savez_compressed(os.path.join(HERE, "ewald_opt.npz"), ints=np.array(rows_i),
                            reals=np.array(rows_d))
    finally:
        shutil.rmtree(work, ignore_errors=True)


def chain_angles():
    """UTotal of the chain case's probe positions (chains.npz) with bond
    angles on (UAngle, Src/energy.F90:5409-5481): for each (k, p, eq) of the
    namelist, the reference's force, forceo and energies -> chain_angles.npz."""
    z = np.load(os.path.join(HERE, "chains.npz"))
    ch = [("pe", 40, 30.0, -1.0, 4.0, KT298), ("cat", 40, 23.0, 1.0, 4.0, KT298)]
    k_bond = float(z["bond"][0])
    angles = [(0.02, 2, 109.5), (0.0005, 3, 150.0)]
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        res = {}
        for a, (ka, pa, eqa) in enumerate(angles):
            text = (system_nml("golden chain angles", None, 60.0, nstep2=1) +
                    particle_nml(ch, chain=(1, 4, 10)) + potential_nml(2, 20.0, ncut=5) +
                    f" &nmlPotentialChain bond = {k_bond!r}, 2, 5.0,"
                    f" angle = {ka!r}, {pa}, {eqa!r} /\n"
                    " &nmlSetConfiguration txsetconf='chainrandom','random', radatset=1.6,1.6 /\n"
                    " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
                    " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n")
            d = case(f"ang{a}", text, work, probe_pos=z["probe_pos"], md=False)
            f = read_force(os.path.join(d, f"ang{a}.probe_f"), 3)
            for k_ in ("force", "forceo", "u_tot", "u_twob", "u_rec", "u_bond", "virial"):
                res[f"{k_}{a}"] = f[k_]
            print("angles", (ka, pa, eqa), "u_tot", f["u_tot"], "u_bond", f["u_bond"])
        np.savez_compressed(os.path.join(HERE, "chain_angles.npz"),
                            angles=np.array(angles, float), **res)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def bd_steps():
    """BDStep (Src/bd.F90:120-155) as BDDriver runs it: a 2:1 electrolyte of
    99 ions (297 deviates per step, an odd count, so GauRandom's spare
    deviate carries across steps) with per-type diffusion coefficients; the
    reference's forceo, positions and displacements over four steps from a
    recorded iseed -> tests/golden/bd.npz."""
    ions = [("c2", 33, 40.0, 2.0, 4.0, KT298), ("an", 66, 35.4, -1.0, 4.0, KT298)]
    dcoeff, tstep, nbd = (1.33, 2.03), 0.005, 4
    body = (particle_nml(ions) + potential_nml(2, 12.0, ncut=5) +
            " &nmlSetConfiguration txsetconf='random','random', radatset=1.6,1.6 /\n"
            f" &nmlBD tstep={tstep!r}, dcoeff={dcoeff[0]!r},{dcoeff[1]!r} /\n"
            " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n")
    text = system_nml("golden bd", None, 40.0, nstep2=1, dump=False, method="bd") + body
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        # positions: the setup's own random configuration, read back from
        # a first probe run with the force dump
        d = os.path.join(work, "bd")
        os.makedirs(d, exist_ok=True)
        rng = np.random.default_rng(11)
        pos = []
        while len(pos) < 99:
            p = rng.uniform(-20.0, 20.0, 3)
            if all(np.sum((np.asarray(q) - p - 40.0 * np.round((np.asarray(q) - p) / 40.0)) ** 2)
                   > 3.2 ** 2 for q in pos):
                pos.append(p)
        pos = np.array(pos)
        np.array([nbd], "<i4").tofile(os.path.join(d, "bd.probe_bd"))
        case("bd", text, work, probe_pos=pos, md=False)
        b = open(os.path.join(d, "bd.probe_bdout"), "rb").read()
        np_, nb, iseed = struct.unpack("<3i", b[:12])
        v = np.frombuffer(b[12:], "<f8").reshape(nb, 4, np_, 3)
        np.savez_compressed(os.path.join(HERE, "bd.npz"), boxlen=40.0, tstep=tstep,
                            temp=298.0, dcoeff=np.array(dcoeff),
                            iptpn=np.array([1] * 33 + [2] * 66, np.int32),
                            iseed=iseed, forceo=v[:, 0], ro_before=v[:, 1],
                            drostep=v[:, 2], ro_after=v[:, 3])
        print("bd: iseed", iseed, "steps", nb, "max |drostep|", np.abs(v[:, 2]).max())
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main():
    if not os.path.exists(PROBE):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    if sys.argv[1:] == ["rng"]:
        rng_streams()
        return
    if sys.argv[1:] == ["bd"]:
        bd_steps()
        return
    if sys.argv[1:] == ["cnf_forces"]:
        cnf_forces()
        return
    if sys.argv[1:] == ["cnf"]:
        c1_cnf()
        return
    if sys.argv[1:] == ["ewald"]:
        ewald_options()
        return
    if sys.argv[1:] == ["angles"]:
        chain_angles()
        return
    work = tempfile.mkdtemp(prefix="mfe_golden_")
    try:
        out = {}
        # ---------------------------------------------------------- C1 (std)
        ions = [("cat", 500, 23.0, 1.0, 4.0, KT298), ("an", 500, 35.4, -1.0, 4.0, KT298)]
        body = (particle_nml(ions) + potential_nml(2, 15.0) +
                " &nmlSetConfiguration txsetconf='random','random', radatset=1.6,1.6 /\n"
                " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
                " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n"
                " &nmlDump idump=1, ldpos=.true., ldliv=.true., ldfor=.true. /\n")
        text = system_nml("golden c1", [500, 500], 50.0, nstep2=5) + body
        d = case("c1", text, work)
        pos, liv, frc = dumps(d, "c1")
        d = case("c1", text, work, probe_pos=pos[1], md=False)
        tab = read_tab(os.path.join(d, "c1.probe_tab"))
        f = read_force(os.path.join(d, "c1.probe_f"), 3)
        np.savez_compressed(
            os.path.join(HERE, "c1_std.npz"), boxlen=50.0, drnlist=2.0,
            ncut=8, tstep=0.002, masses=np.array([23.0, 35.4]),
            zat=np.array([1.0, -1.0]), sigat=np.array([4.0, 4.0]),
            epsat=np.array([KT298, KT298]), nppt=np.array([500, 500]),
            traj_pos=pos, traj_vel=liv, traj_forceo=frc, probe_pos=pos[1],
            **{k: v for k, v in tab.items()}, **{"ref_" + k: v for k, v in f.items()})
        out["c1_std"] = (tab["nugrid"].tolist(), f["u_tot"])

        # ------------------------------------------- C1 system, cell list + surf
        body_c = body.replace("txintlist='vlist'", "txintlist='celllist'").replace(
            "lsurf=.false.", "lsurf=.true.")
        text = system_nml("golden c1 cell", [500, 500], 50.0, nstep2=1, dump=False) + body_c
        d = case("c1cell", text, work, probe_pos=pos[1], md=False)
        f = read_force(os.path.join(d, "c1cell.probe_f"), 3)
        np.savez_compressed(os.path.join(HERE, "c1_cell_surf.npz"), probe_pos=pos[1],
                            **{"ref_" + k: v for k, v in f.items()
                               if k not in ("nneighpn", "jpnlist_sha256")})
        out["c1_cell"] = f["u_tot"]

        # ---------------------------------------------- tables for C2 and C3
        c3 = [("c1", 10, 23.0, 1.0, 4.0, KT298), ("c2", 10, 40.0, 2.0, 4.0, KT298),
              ("an", 30, 35.4, -1.0, 4.0, KT298)]
        text = (system_nml("golden c3 tables", None, 200.0, dump=False) + particle_nml(c3) +
                potential_nml(3, 12.0, ncut=4) +
                " &nmlSetConfiguration txsetconf='random','random','random' /\n"
                " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
                " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n")
        d = case("c3tab", text, work, md=False)
        tab = read_tab(os.path.join(d, "c3tab.probe_tab"))
        np.savez_compressed(os.path.join(HERE, "tables_c3.npz"),
                            zat=[1.0, 2.0, -1.0], sigat=[4.0] * 3, epsat=[KT298] * 3, **tab)
        out["tables_c3"] = tab["nugrid"].tolist()

        c2 = [("pe", 40, 30.0, -1.0, 4.0, KT298), ("cat", 60, 23.0, 1.0, 4.0, KT298),
              ("an", 20, 35.4, -1.0, 4.0, KT298)]
        text = (system_nml("golden c2 tables", None, 200.0, dump=False) + particle_nml(c2) +
                potential_nml(3, 18.0, ncut=4) +
                " &nmlSetConfiguration txsetconf='random','random','random' /\n"
                " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
                " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n")
        d = case("c2tab", text, work, md=False)
        tab = read_tab(os.path.join(d, "c2tab.probe_tab"))
        np.savez_compressed(os.path.join(HERE, "tables_c2.npz"),
                            zat=[-1.0, 1.0, -1.0], sigat=[4.0] * 3, epsat=[KT298] * 3, **tab)
        out["tables_c2"] = tab["nugrid"].tolist()

        # --------------------------------------------- chains with bonds (UBond)
        ch = [("pe", 40, 30.0, -1.0, 4.0, KT298), ("cat", 40, 23.0, 1.0, 4.0, KT298)]
        k_bond = 0.4 * KT298
        text = (system_nml("golden chains", None, 60.0, nstep2=1) +
                particle_nml(ch, chain=(1, 4, 10)) + potential_nml(2, 20.0, ncut=5) +
                f" &nmlPotentialChain bond = {k_bond!r}, 2, 5.0, angle = 0.0 /\n"
                " &nmlSetConfiguration txsetconf='chainrandom','random', radatset=1.6,1.6 /\n"
                " &nmlMD integ='velver', tstep=0.002, nvvite=2, lsetvel=.true. /\n"
                " &nmlIntList txintlist='vlist', inlist=0, drnlist=2.0 /\n"
                " &nmlDump idump=1, ldpos=.true., ldliv=.true., ldfor=.true. /\n")
        d = case("chain", text, work)
        cpos, _, _ = dumps(d, "chain")
        d = case("chain", text, work, probe_pos=cpos[1], md=False)
        tab = read_tab(os.path.join(d, "chain.probe_tab"))
        f = read_force(os.path.join(d, "chain.probe_f"), 3)
        np.savez_compressed(
            os.path.join(HERE, "chains.npz"), boxlen=60.0, ncut=5,
            bond=np.array([k_bond, 2.0, 5.0]), nchain=4, nper=10,
            zat=[-1.0, 1.0], sigat=[4.0] * 2, epsat=[KT298] * 2, nppt=[40, 40],
            probe_pos=cpos[1], **tab, **{"ref_" + k: v for k, v in f.items()})
        out["chains"] = (f["u_bond"], f["u_tot"])

        # ------------------------------------------- Erf2 and CardinalBSpline
        rng = np.random.default_rng(7)
        args = np.concatenate([np.linspace(-7, 7, 701), rng.uniform(-3, 3, 300),
                               rng.uniform(-70, 70, 100), [0.0, 2.0, -2.0, 1e-8]])
        d = os.path.join(work, "c1")
        args.astype("<f8").tofile(os.path.join(d, "c1.probe_args"))
        rr = np.sort(rng.uniform(0.32, 16.5, 600))
        rr.astype("<f8").tofile(os.path.join(d, "c1.probe_r"))
        run(PROBE, "c1", d)
        m = open(os.path.join(d, "c1.probe_math"), "rb").read()
        n = args.size
        erf = np.frombuffer(m[:8 * n], "<f8").copy()
        o = 8 * n
        bs = {}
        for order in range(3, 9):
            rec = 4 + 16 * order
            k = np.zeros(n, np.int32)
            v = np.zeros((n, order))
            dv = np.zeros((n, order))
            for i in range(n):
                chunk = m[o:o + rec]
                o += rec
                k[i] = struct.unpack("<i", chunk[:4])[0]
                v[i] = np.frombuffer(chunk[4:4 + 8 * order], "<f8")
                dv[i] = np.frombuffer(chunk[4 + 8 * order:], "<f8")
            bs[f"kmax{order}"], bs[f"val{order}"], bs[f"der{order}"] = k, v, dv
        pot = np.fromfile(os.path.join(d, "c1.probe_pot"), "<f8").reshape(rr.size, 4, 3)
        np.savez_compressed(os.path.join(HERE, "math.npz"), args=args, erf2=erf,
                            pot_r=rr, ewald_ufac=pot[:, 0], pot_1_6_12=pot[:, 1:],
                            pot_alpha=0.2, **bs)
        for k_, v_ in out.items():
            print(k_, v_)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
