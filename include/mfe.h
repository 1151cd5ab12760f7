/*
 * mfe.h -- C ABI of the MI355X force engine for MOLSIM's nonbonded path.
 *
 * MOLSIM has no plugin interface for this path: its seam is a set of
 * Fortran subroutines acting on module state.  Each entry point below
 * replaces one of them and takes MOLSIM's own arrays by pointer, in its own
 * layout (column-major ro(3,np) == xyz-interleaved, 1-based type and
 * table indices), so a Fortran host can bind them with ISO_C_BINDING.
 *
 * Conventions
 *  - Every function returns an int status: MFE_OK (0) or an MFE_E_* code;
 *    mfe_last_error() returns the message of the last failure.
 *  - Inputs are copied to device memory; no host pointer is kept after a
 *    call returns.  Calls are synchronous with respect to the host.
 *  - One engine owns one HIP stream and is not thread-safe: call it from a
 *    single host thread, as MOLSIM (single-threaded) does.
 *  - Energies are in kJ/mol and forces follow UTotal: mfe_compute returns
 *    forceo = force * (sclene/scllen)/sclfor (Src/energy.F90:405-418).
 */
#ifndef MFE_H
#define MFE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFE_ABI_VERSION 1

/* status codes: the abort conditions of the reference path */
enum {
  MFE_OK = 0,
  MFE_E_BAD_PARAM = 1,       /* argument out of range / inconsistent     */
  MFE_E_OVERLAP_R2UMIN = 2,  /* r2 < r2umin: StopUTwoBodyA, Src/energy.F90:478 */
  MFE_E_TABLE_RANGE = 3,     /* ibuf > nbuf, Src/energy.F90:485            */
  MFE_E_NEIGH_CAPACITY = 4,  /* nneighpn > maxnneigh, Src/nlist.F90:752    */
  MFE_E_MESH_ODD = 5,        /* nmesh odd, Src/energy.F90:6891             */
  MFE_E_HIP = 6,             /* HIP runtime failure                        */
  MFE_E_STATE = 7,           /* call order: a required set_* is missing    */
  MFE_E_TABLE_BUILD = 8,     /* PotTwoBodyTab3 did not converge            */
  MFE_E_NOMEM = 9            /* device allocation failed                   */
};

enum { MFE_PREC_FP64 = 0, MFE_PREC_MIXED = 1 };
enum { MFE_EWALD_NONE = 0, MFE_EWALD_STD = 1, MFE_EWALD_SPME = 2 };
/* pair-set semantics: MFE_NL_VLIST = Verlet list of SetVListMD + UTwoBodyA
   (pairs with r2 <= rcut2); MFE_NL_CELL = UTwoBodyACellList (r2 < rcut2) */
enum { MFE_NL_VLIST = 0, MFE_NL_CELL = 1 };
/* mfe_compute flags */
enum { MFE_FORCES = 1u, MFE_ENERGY = 2u, MFE_VIRIAL = 4u };

#define MFE_MAX_TYPES 8
#define MFE_MAX_PAIRS (MFE_MAX_TYPES * (MFE_MAX_TYPES + 1) / 2)

typedef struct mfe_engine mfe_engine;

typedef struct {
  int32_t precision; /* MFE_PREC_FP64 or MFE_PREC_MIXED                      */
  int32_t device;    /* HIP device ordinal                                    */
  int32_t max_np;    /* capacity in particles                                 */
  int32_t reserved;
} mfe_config;

/* u%tot, u%twob(0:nptpt), u%rec, u%bond and virial of UTotal */
typedef struct {
  double tot;
  double twob0;
  double rec;
  double bond;
  double virial;
  int32_t nptpt;
  int32_t rebuilds; /* neighbour-list rebuilds since the last reset */
  double twob[MFE_MAX_PAIRS];
} mfe_energy;

/* Pair-potential specification for the (1,6,12) potential, the arguments
   of IOPotTwoBody that Pot_1_6_12 reads (Src/potential.F90:307-392). */
typedef struct {
  int32_t npt;
  const double *zat;   /* [npt] */
  const double *sigat; /* [npt] */
  const double *epsat; /* [npt] */
  double epsi_four_pi; /* Epsi0FourPi/relpermitt */
  double rcut;
  double alpha;        /* ualpha (used when lewald) */
  int32_t lewald;
  int32_t lljcut;
  double ljrcut, ljushift;
  double r2uminin, utoltab, ftoltab, umaxtab, fmaxtab, rumaxfac;
} mfe_pot_1_6_12;

/* EwaldSetup's parameter choice (iewaldopt; Src/energy.F90:6644-6690 for
   the standard sum, 6863-6886 for SPME). Which fields are inputs and
   which are outputs depends on iewaldopt, as in the reference's table:
     0: ualphared, rcut (, ncut)      -> ualpha = ualphared/rcut
     1: ualphared, rcut               -> ualpha, uewaldtol, ncut
     2: uewaldtol, ualpha             -> rcut, ncut, ualphared
     3: uewaldtol, rcut               -> ualpha, ualphared (, ncut: std)
     4: uewaldtol, ncut               -> ualpha, rcut, ualphared
   (SPME: 0 and 3 only.) The error estimates are Kolafa-Perram for charges
   (lq2sum 0) and Wang-Holm for dipoles (lq2sum 1): AlphaToRcut,
   AlphaToNcut, RcutToAlpha and NcutToAlpha (Src/energy.F90:7052-7124),
   EwaldErrorUReal (Src/potential.F90:2995-3018). q2sum = sum(q^2)/np,
   boxlenshort = minval(boxlen). Inputs out of range and iterations that do
   not converge in 40 steps give MFE_E_BAD_PARAM, as they stop the
   reference. */
typedef struct {
  int32_t iewaldopt;
  int32_t spme;         /* 1: txewaldrec = 'spm' */
  int32_t lq2sum;       /* 0: charges, 1: dipoles and no charges */
  int32_t ncut;
  double q2sum, boxlenshort, epsi_four_pi;
  double uewaldtol, ualphared, ualpha, rcut;
} mfe_ewald_opt;

/* ---- library ---------------------------------------------------------- */
int mfe_abi_version(void);
int mfe_ewald_parameters(mfe_ewald_opt *p);
/* Fills spec with IOPotTwoBody's defaults (Src/potential.F90:340-392,411). */
int mfe_pot_defaults(mfe_pot_1_6_12 *spec);
/* PotTwoBodyTab1/2/3 + SetUBuffer + CheckUBuffer for Pot_1_6_12 with
   monoatomic particle types (Src/potential.F90:868-1258, 1349-1425): fills
   ubuf (capacity cap doubles) and, per type pair iptjpt (Set_iptpt order),
   iubuflow (1-based), nugrid, r2umin.  *nbuf_used receives the length. */
int mfe_build_tables_1_6_12(const mfe_pot_1_6_12 *spec, double *ubuf,
                            int32_t cap, int32_t *nbuf_used,
                            int32_t *iubuflow, int32_t *nugrid,
                            double *r2umin);

/* ---- engine lifetime (replaces MolModule allocation) ------------------- */
int mfe_create(const mfe_config *cfg, mfe_engine **out);
int mfe_destroy(mfe_engine *e);
const char *mfe_last_error(const mfe_engine *e);
/* No reference counterpart (MOLSIM only warns about a net charge,
   Src/potential.F90:621-625).  on = 1 adds the neutralising-background
   term of a system with net charge Q = sum az to the Ewald energy (both
   kinds): u%rec += EpsiFourPi * (-pi Q^2 / (2 V ualpha^2)), virial
   -3 times that term (it scales as 1/V), no force; mfe_adjoint's dE/dq
   gains EpsiFourPi * (-pi Q / (V ualpha^2)).  Off by default, as the
   reference computes it.  Zero for neutral systems. */
int mfe_set_net_charge(mfe_engine *e, int32_t on);
/* No reference counterpart.  Bound on the polls of each grid-barrier wait
   of the single-engine list rebuild (polls < 0: the default, 2^22).  A
   wait that runs out ends the rebuild in every block; the call then
   returns MFE_E_HIP, the positions stay where they were, and the next
   call builds the list from scratch.  Small bounds exist to exercise
   that path from the host. */
int mfe_set_sync_limit(mfe_engine *e, int32_t polls);

/* ---- setup (replaces IOSystem box setup, Particle, IOPotTwoBody,
        EwaldSetup, IONList) ------------------------------------------- */
/* Orthorhombic periodic box, txbc='xyz' (Src/molaux.F90:326-335). */
int mfe_set_box(mfe_engine *e, const double boxlen[3]);
/* np particles (monoatomic), npt types: iptpn[np] 1-based, az[np]
   charges, massipt[npt] inverse masses (Src/particle.F90:2189-2278). */
int mfe_set_particles(mfe_engine *e, int32_t np, int32_t npt,
                      const int32_t *iptpn, const double *az,
                      const double *massipt);
/* Potential table in the reference's 12-double interval format
   (Src/potential.F90:1181-1210): ubuf[nbuf], iubuflow[nptpt] (1-based),
   r2atat[nptpt], r2umin[nptpt]; rcut is the real-space cut-off. */
int mfe_set_tables(mfe_engine *e, const double *ubuf, int32_t nbuf,
                   const int32_t *iubuflow, const double *r2atat,
                   const double *r2umin, double rcut);
/* Ewald sum (EwaldSetup, Src/energy.F90:6601-7046): kind MFE_EWALD_*,
   alpha = ualpha, order and mesh[3] for SPME, ncut (+ ncut_cube) for the
   standard sum, lsurf the surface term, epsi_four_pi = EpsiFourPi. */
int mfe_set_ewald(mfe_engine *e, int32_t kind, double alpha, int32_t order,
                  const int32_t mesh[3], int32_t ncut, int32_t ncut_cube,
                  int32_t lsurf, double epsi_four_pi);
/* Neighbour list (IONList, Src/nlist.F90:184-333): kind MFE_NL_*,
   drnlist the skin, maxnneigh the per-particle capacity (0 = automatic). */
int mfe_set_neighbor(mfe_engine *e, int32_t kind, double drnlist,
                     int32_t maxnneigh);
/* Chains for UBond (Src/energy.F90:5335-5400): chain c holds particles
   chain_first[c] .. chain_first[c]+chain_len[c]-1 (1-based); k is the
   namelist force constant (divided by p here, as IOPotChain does at
   Src/potential.F90:3561), p the integer power, eq the bond length. */
int mfe_set_chains(mfe_engine *e, int32_t nchain, const int32_t *chain_first,
                   const int32_t *chain_len, double k, int32_t p, double eq);
/* Bond angles of the chains for UAngle (Src/energy.F90:5409-5481, called
   after UBond from UTotal, Src/energy.F90:387): angle(ict)%k, %p and %eq
   of nmlPotentialChain, k per degree^2 and eq in degrees, scaled as
   IOPotChain does (Src/potential.F90:3558-3562): k / (pi/180)^2 / p
   whatever p is, eq * pi/180.  k = 0 (the reference's
   default) leaves the term out.  u%angle enters u%tot and carries no
   virial; read it with mfe_get_angle_energy. */
int mfe_set_angles(mfe_engine *e, double k, int32_t p, double eq);
/* u%angle of the last call that reduced energies (kJ/mol, like u%bond). */
int mfe_get_angle_energy(mfe_engine *e, double *u_angle);

/* ---- state ------------------------------------------------------------ */
/* ro(3,np), positions inside the box (Src/mol.F90:1087). */
int mfe_upload_positions(mfe_engine *e, const double *ro);
/* rod(3,np) velocities and rodd(3,np) accelerations (NULL = zero). */
int mfe_upload_velocities(mfe_engine *e, const double *rod,
                          const double *rodd);
int mfe_download_state(mfe_engine *e, double *ro, double *rod, double *rodd);
/* forceo(3,np) of the last UTotal (mfe_compute or a step). */
int mfe_download_forces(mfe_engine *e, double *forceo);

/* ---- the hot path ----------------------------------------------------- */
/* SetVList / InitCellList+SetCellList at the current positions. */
int mfe_build_list(mfe_engine *e);
/* UTotal (Src/energy.F90:280-434) at the current positions with the
   current list: forceo(3,np) (may be NULL) and energies (may be NULL). */
int mfe_compute(mfe_engine *e, uint32_t flags, double *forceo,
                mfe_energy *energy);
/* Monte Carlo trial move, DUTwoBodyEwaldRecSPM (Src/denergy.F90:1413-1495,
   reached from DUTwoBodyEwald, Src/denergy.F90:1258-1260): the change of
   the SPME reciprocal energy when the natm distinct particles ianatm[]
   (1-based) move from their current positions to rtm(3,natm), in the
   units of mfe_energy.rec (EpsiFourPi included; the self term does not
   change).  Nothing of the engine's state changes (the next force call
   spreads the mesh again).  The reference forms the trial mesh and takes
   its energy minus urecold; the transform is linear, so the engine keeps
   the spectrum F of the current configuration's mesh (as the reference
   keeps QMesh) and adds the transform D of the moved stencils, a product
   of one-dimensional sums per particle: the change is the sum over the
   spectrum of the convolution's weights times 2 Re(conj(F) D) + |D|^2 --
   one pass, no transform per trial, no difference of two large energies.
   reuse_current = 1 starts from the kept spectrum and skips spreading and
   transforming the current configuration (the caller states that nothing
   was accepted since, other than through mfe_mc_accept); the engine still
   spreads anew after any upload, step or setup call, and on its first
   call.  The change agrees with the difference of the two energies to
   ~1e-11 of |E_rec| on FP64 engines and ~1e-6 of itself (a few 1e-6 kT
   for one-particle moves) on mixed ones.  SPME only (MFE_E_STATE
   otherwise). */
int mfe_mc_delta_rec(mfe_engine *e, int32_t natm, const int32_t *ianatm,
                     const double *rtm, int32_t reuse_current, double *du_rec);
/* Monte Carlo trial move, pair part: UTwoBodyANew - UTwoBodyAOld
   (Src/denergy.F90:289-495, from DUTwoBody :246-279).  The natm distinct
   particles ianatm[] (1-based) against their partners in the engine's
   current Verlet list that do not move, plus the moved pairs among
   themselves; table walk and checks as the reference's, in FP64 on both
   precision paths.  du_twob = change of the two-body energy summed over
   pair types (the units of mfe_energy.twob0); overlap = 1 when a trial
   pair is inside the hard core or below r2umin (the reference's
   lhsoverlap: the move is rejected; du_twob is then 0).  The list is the
   caller's to keep current (the reference's NList policy).  On an engine
   set up with MFE_NL_CELL the partners are, as for UTwoBodyANewCellList /
   UTwoBodyAOldCellList (Src/denergy.F90:728-928, chosen for lCellList at
   :145-147), every particle with r2 < rcut2 of the new or old position,
   however far the move. */
int mfe_mc_delta_pair(mfe_engine *e, int32_t natm, const int32_t *ianatm,
                      const double *rtm, double *du_twob, int32_t *overlap);
/* No reference counterpart: mfe_mc_delta_pair for callers that cannot
   allocate device memory (code under HIP graph capture, tight MC serving
   loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
   results; it calls no hipMalloc or hipFree: the moved particles, the
   results and the kernels' block partials live in a fixed device array
   (about 38 KB per device, held until the process ends).  natm <= 1024
   (MFE_E_BAD_PARAM beyond).  The engine's list must exist already
   (mfe_build_list or any force call first; MFE_E_STATE otherwise).  All
   calls on one device share that array: calls that overlap on one device
   (from several host threads or streams) race. */
int mfe_mc_delta_pair_noalloc(mfe_engine *e, int32_t natm, const int32_t *ianatm,
                              const double *rtm, double *du_twob, int32_t *overlap);
/* Monte Carlo accepted move: EwaldUpdateArray for 'spm'
   (Src/denergy.F90:1776-1819: QMesh = QMeshTM, urecold = urecnew) with the
   MC driver's position update ro(:,ianatm) = rtm and NList's skin
   bookkeeping (the list is rebuilt on the device only when the
   reference's skin test fires, as after mfe_upload_positions).  rtm must
   lie inside the box.  When the last mfe_mc_delta_rec was this same move,
   its change D is added to the kept spectrum (FP64 on both precision
   paths), which then is the current one, so the next trial spreads
   nothing; after any other move the next trial spreads anew.
   reuse_current = 1 is valid after this call in both cases.  On a mixed
   engine the kept spectrum then differs from a fresh FP32 transform of the
   accepted configuration by that transform's rounding (the next trial's
   change agrees to ~1e-6 kT for one-particle moves), and each accepted
   change adds the rounding of its FP32 stencil factors: a long run of
   accepts without an upload or step can refresh the spectrum by passing
   reuse_current = 0 now and then. */
int mfe_mc_accept(mfe_engine *e, int32_t natm, const int32_t *ianatm,
                  const double *rtm);
/* nsteps of MDDriver(iSimulationStep) + NList(iSimulationStep): VelVer
   (Src/md.F90:252-410) with the skin-test rebuild (Src/nlist.F90:597-616),
   all on the device.  flags select which UTotal terms are reduced each
   step (forces are always computed); energies of the last step are kept
   for mfe_get_energy. */
int mfe_step_md(mfe_engine *e, int32_t nsteps, double tstep, uint32_t flags);
/* nsteps of BDStep (Src/bd.F90:120-155) + UTotal + NList: fac1/fac2 per
   type (NULL: keep those of mfe_set_bd), noise[nsteps][np][3] unit
   Gaussian deviates in the reference's draw order (ip outer, xyz inner).
   With noise == NULL the deviates come from the engine's counter-based
   stream (see mfe_set_bd) instead of GauRandom's serial LCG
   (Src/mollib.F90:2739-2766), so replicas draw independent, reproducible
   noise without a host round trip.  The reference's first BD step uses
   the forces of the preceding UTotal: when the engine's forces are not
   those of its current positions (no mfe_compute or step since the
   positions or the setup changed), the call evaluates them first. */
int mfe_step_bd(mfe_engine *e, int32_t nsteps, const double *fac1,
                const double *fac2, const double *noise, uint32_t flags);
/* BDStep's per-type factors (fac1 = dt D/kT in force units, fac2 =
   sqrt(2 D dt)/scllen, Src/bd.F90:120-140) and the seed of the device
   noise stream: particle i at BD step n draws Philox4x32-10 with counter
   (i, n, 0, 0) and key seed, and takes three Box-Muller deviates.  Resets
   the step counter n to 0.  No reference counterpart (replaces the
   GauRandom seed iseed of nmlSystem). */
int mfe_set_bd(mfe_engine *e, const double *fac1, const double *fac2,
               uint64_t seed);
int mfe_get_energy(mfe_engine *e, mfe_energy *energy);
/* The Verlet list in SetVListMD's form (Src/nlist.F90:724-759): for each
   ip the partners jp it owns under the checkerboard rule, ascending.
   nneighpn[np]; jpnlist[np][maxnneigh] row-major (row ip-1).  */
int mfe_get_pairs(mfe_engine *e, int32_t *nneighpn, int32_t *jpnlist,
                  int32_t maxnneigh);

/* ---- measurement ------------------------------------------------------ */
enum {
  MFE_STAGE_PAIR = 0,   /* table pair forces (the pair kernel)             */
  MFE_STAGE_SPREAD = 1, /* charge spreading (the spreading kernel)        */
  MFE_STAGE_FFT_X = 2,  /* r2c along x (forward)                          */
  MFE_STAGE_FFT_Y = 3,  /* complex along y (forward)                      */
  MFE_STAGE_CONV = 4,   /* z forward + influence/energy/virial + z inverse */
  MFE_STAGE_GATHER = 5, /* force back-interpolation                       */
  MFE_STAGE_STEP = 6,   /* one whole MD step                              */
  MFE_STAGE_NLIST = 7,  /* neighbour-list rebuild                         */
  MFE_STAGE_IFFT = 8,   /* y inverse + c2r along x                        */
  MFE_STAGE_BIN = 9,    /* particles into mesh bricks (count, scan, records) */
  MFE_NSTAGES = 10
};
/* Runs one stage nrep times on the engine's stream, bracketed by HIP
   events on that stream, and returns the mean time per run in ms.  One
   full UTotal runs first, so each stage finds its inputs in place. */
int mfe_time_stage(mfe_engine *e, int32_t stage, int32_t nrep, double *ms);

/* ---- parameter adjoint ------------------------------------------------ */
/* No reference counterpart.  The model the adjoint differentiates is the
   pair potential the tables interpolate -- Pot_1_6_12 with exact erfc
   (Src/potential.F90:1349-1425, Lorentz-Berthelot mixing sig = (sig_i +
   sig_j)/2, eps = sqrt(eps_i eps_j), optional WCA cut at ljrcut*sig with
   shift ljushift*eps) -- plus the optional spline term below, SPME (or no
   Ewald sum) with its self and surface terms, and the bonds. */
/* The parameters of that model: spec as for mfe_build_tables_1_6_12; its
   zat must equal the charges az of mfe_set_particles. */
int mfe_set_pair_params(mfe_engine *e, const mfe_pot_1_6_12 *spec);
/* An additive pair term per type pair p (Set_iptpt order), a uniform cubic
   B-spline in r with nknot coefficients c[p][k] on [r_lo, r_hi) and zero
   outside: u_p(r) = sum_k c[p][k] B_k(r), B_k the cubic B-spline on knots
   r_lo + (j-3) h, h = (r_hi - r_lo)/(nknot-3).  Applied in every UTotal
   (forces and energies); nknot = 0 removes it. */
int mfe_set_knots(mfe_engine *e, int32_t nknot, double r_lo, double r_hi,
                  const double *c);
/* One UTotal at the current positions and the gradient of the model
   energy: dE/dq[np] (per particle), dE/dzat[npt], dE/dsigat[npt],
   dE/depsat[npt], dE/dc[nptpt][nknot] and the position adjoint
   dE/dro(3,np) = -force (kJ/mol and A; any output may be NULL).  energy
   receives the model energy: exact pair sum + u%rec + u%bond. */
int mfe_adjoint(mfe_engine *e, double *dE_dq, double *dE_dzat,
                double *dE_dsigat, double *dE_depsat, double *dE_dknot,
                double *dE_dro, double *energy);

/* ---- force matching --------------------------------------------------- */
/* No reference counterpart.  The forces of the adjoint's model (the pair
   model of mfe_set_pair_params and mfe_set_knots evaluated exactly, in
   place of the tables, plus SPME, bonds, angles and the surface term) at
   the current positions: force(3,np) = -dE/dro in kJ/(mol A), UTotal's
   force before the unit scaling.  The arithmetic is the engine's: FP32
   model terms with FP64 sums on the mixed path, FP64 on the FP64 path. */
int mfe_model_forces(mfe_engine *e, double *force);
/* nframes configurations ro[nframes][np][3] (inside the box) and their
   target forces f_target[nframes][np][3] (kJ/(mol A)), kept in device
   memory for mfe_fm_loss_grad. */
int mfe_fm_set_frames(mfe_engine *e, int32_t nframes, const double *ro,
                      const double *f_target);
/* Loss L = sum over frames and particles of |F_model - F_target|^2 and its
   gradient with respect to the per-type charges zat, sigat and epsat
   [npt] and the spline coefficients c[nptpt][nknot] (any output may be
   NULL).  Per frame, on the device: list build, model forces with the
   reciprocal field kept, residual, the pair part of the gradient, the
   reciprocal part through a second spreading (of the residual-weighted
   B-spline derivatives), convolution and potential gather, and the
   surface term.  The engine's positions are those of the last frame
   afterwards.  SPME or no Ewald sum. */
int mfe_fm_loss_grad(mfe_engine *e, double *loss, double *dL_dzat,
                     double *dL_dsigat, double *dL_depsat, double *dL_dknot);

/* ---- batched replicas ------------------------------------------------- */
/* No reference counterpart: MOLSIM runs one system per process.  A batch
   drives nrep engines, each set up as above (own box, particles, tables,
   list, mesh), as one system: every stage of UTotal and VelVer is one
   launch over all replicas; BD steps take one persistent launch where the
   replicas allow it (mfe_batch_step_bd).  The replicas must share device,
   precision, Ewald kind and SPME order; box, N, mesh, tables and cutoffs
   are per replica.  Each replica computes exactly what its engine computes
   alone.  Results stay in the engines (mfe_download_forces,
   mfe_download_state, mfe_get_energy) or come back together through
   mfe_batch_get_energies.  The engines must outlive the batch; a set-up
   call on an engine is picked up at the batch's next call.  On an error
   mfe_batch_failed_replica names the first failing replica. */
typedef struct mfe_batch mfe_batch;
int mfe_batch_create(mfe_engine *const *engines, int32_t nrep, mfe_batch **out);
int mfe_batch_destroy(mfe_batch *b);
const char *mfe_batch_last_error(const mfe_batch *b);
int32_t mfe_batch_failed_replica(const mfe_batch *b);
/* SetVList for every replica */
int mfe_batch_build_list(mfe_batch *b);
/* UTotal for every replica (energies and virials always reduced) */
int mfe_batch_compute(mfe_batch *b, uint32_t flags);
/* nsteps of VelVer + NList for every replica, one graph launch per step */
int mfe_batch_step_md(mfe_batch *b, int32_t nsteps, double tstep, uint32_t flags);
/* nsteps of BDStep + NList with each replica's device noise stream and
   factors (mfe_set_bd).  Mixed-precision SPME replicas without chains,
   surface term or spline knots are advanced in one persistent launch: workgroups take replicas from a
   queue (costliest first) and run all nsteps of a replica with workgroup
   barriers only -- every stage of BDStep, UTotal and the gated SetVListMD.
   A replica's result there depends on that replica alone.  Other batches
   replay one captured graph per step.  As in mfe_step_bd, replicas whose
   forces are not those of their current positions are evaluated first. */
int mfe_batch_step_bd(mfe_batch *b, int32_t nsteps, uint32_t flags);
/* No reference counterpart: whether mfe_batch_step_bd may take the
   persistent launch (on = 1, the default) or always replays the per-stage
   graph (on = 0); and whether it would take it for this batch now. */
int mfe_batch_set_persistent(mfe_batch *b, int32_t on);
int32_t mfe_batch_uses_persistent(mfe_batch *b);
/* energies of the last UTotal of every replica: out[nrep] */
int mfe_batch_get_energies(mfe_batch *b, mfe_energy *out);

#ifdef __cplusplus
}
#endif
#endif /* MFE_H */
