/*
 * mfe_oracle.h -- FP64 CPU restatement of MOLSIM's nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (molsim_amd/, include/)
 * links, loads or calls this code; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py use it, to check results against.
 *
 * Every function follows one reference routine loop for loop, in the same
 * arithmetic order, and cites it as Src/<file>:<lines>.  Conventions:
 *   - positions/forces are xyz-interleaved [3*np], i.e. MOLSIM's ro(3,np);
 *   - type indices passed in are 1-based like iptpn(:), and pair indices
 *     iptjpt follow Set_iptpt (Src/particle.F90:1535-1541);
 *   - ubuf offsets (iubuflow) are 1-based like the reference.
 */
#ifndef MFE_ORACLE_H
#define MFE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- units and constants (Src/mol.F90:70-76, Src/molsim.F90:786-805) ---- */
double orc_epsi0_four_pi(void);
double orc_force_unit_fac(void);
double orc_gas_constant(void);

/* ---- math library (Src/mollib.F90) ---- */
double orc_erf2(double x);
void orc_cardinal_bspline(int order, double pos, int *kmax, double *values,
                          double *derivs);

/* ---- periodic box (Src/molaux.F90:104-184) ---- */
typedef struct {
  double boxlen[3], boxlen2[3], boxleni[3], dpbc[3];
} orc_box;
void orc_box_init(orc_box *b, const double boxlen[3]);
void orc_pbcr2(const orc_box *b, double *dx, double *dy, double *dz,
               double *r2);
void orc_pbc(const orc_box *b, double *x, double *y, double *z);

/* ---- (1,6,12) potential and its tables (Src/potential.F90) ---- */
typedef struct {
  int npt;                 /* number of (monoatomic) particle types */
  const double *zat;       /* [npt] charges */
  const double *sigat;     /* [npt] LJ sigma (A) */
  const double *epsat;     /* [npt] LJ epsilon (kJ/mol) */
  double epsi_four_pi;     /* Epsi0FourPi/relpermitt */
  double rcut;
  double alpha;            /* ualpha; used only if lewald */
  int lewald;
  int lljcut;
  double ljrcut, ljushift;
  double r2uminin, utoltab, ftoltab, umaxtab, fmaxtab, rumaxfac;
} orc_pot_spec;
void orc_pot_spec_defaults(orc_pot_spec *s);
void orc_calc_ewald_ufac(double r1, double alpha, double *c0, double *c1,
                         double *c2);
void orc_pot_1_6_12(const orc_pot_spec *s, int it, int jt, double r1,
                    double *u0, double *u1, double *u2);
void orc_set_ubuffer(double zlow, double u0low, double u1low, double u2low,
                     double zupp, double u0upp, double u1upp, double u2upp,
                     double ubuft[12]);
/* Returns the number of doubles used in ubuf (>0), or -1 (capacity), -2
 * (no convergence in dr), -3 (too many grids).  iubuflow/nugrid/r2umin/
 * r2umax are indexed by iptjpt-1 (nptpt entries). */
int orc_build_tables(const orc_pot_spec *s, double *ubuf, int cap,
                     int *iubuflow, int *nugrid, double *r2umin,
                     double *r2umax);
/* Table evaluation of one pair (Src/energy.F90:476-491).  Returns 0, or 1
 * for r2 < r2umin (StopUTwoBodyA), 2 for ibuf > nbuf. */
int orc_table_eval(const double *ubuf, int nbuf, int iubuflow, double r2,
                   double *usum, double *fsum);

/* ---- neighbour lists (Src/nlist.F90, Src/celllist.F90) ---- */
/* SetVListMD (Src/nlist.F90:724-759) with ipmyid = [1, np].  jpnlist is
 * [np][maxnneigh] (row ip-1 holds the 1-based partners of ip).  Returns 0
 * or 4 (nneighpn > maxnneigh). */
int orc_set_vlist_md(const orc_box *b, int np, const double *ro, double rs,
                     int maxnneigh, int *nneighpn, int *jpnlist);
/* Same pair set and list order as orc_set_vlist_md, found through a cell
 * grid of edge >= rs (for sizes where the O(N^2) sweep is too slow). */
int orc_set_vlist_md_cells(const orc_box *b, int np, const double *ro,
                           double rs, int maxnneigh, int *nneighpn,
                           int *jpnlist);
/* NList(iSimulationStep) skin test, inlist == 0 (Src/nlist.F90:597-616).
 * Accumulates drostep into drosum; returns 1 (and zeroes drosum) when the
 * list must be rebuilt. */
int orc_nlist_skin_test(int np, double *drosum, const double *drostep,
                        double drnlist);

/* InitCellList + SetCellList (Src/celllist.F90:76-265, 384-410). */
typedef struct {
  int ncell[3];
  double cellsize[3], cellsizei[3];
  int ndir;                /* number of neighbour directions kept */
  int *dirs;               /* [ndir][3], sorted by HeapSortIndex */
  int *iphead;             /* [ncells], 1-based particle or 0 */
  int *npart;              /* [ncells] */
  int *ipnext;             /* [np], 1-based or 0 */
  int *cellip;             /* [np], 0-based cell id (idCellPos-1) */
  int np;
} orc_celllist;
int orc_celllist_init(orc_celllist *c, const orc_box *b, double rcell,
                      double rcut, int np);
int orc_celllist_set(orc_celllist *c, const orc_box *b, const double *ro);
void orc_celllist_free(orc_celllist *c);
void orc_heapsort_index(int n, const double *vec, int *index);

/* ---- pair energy (Src/energy.F90:445-544, 669-777) ---- */
typedef struct {
  const double *ubuf;
  int nbuf;
  const int *iubuflow;     /* [nptpt], 1-based */
  const double *r2atat;    /* [nptpt] */
  const double *r2umin;    /* [nptpt] */
  double rcut2;
  int npt;
} orc_table_set;
/* UTwoBodyA over a Verlet list (lmc = .false.).  force is accumulated,
 * u_twob[0..nptpt] and *virial are accumulated into. */
int orc_utwobody_a(const orc_table_set *t, const orc_box *b, int np,
                   const double *ro, const int *iptpn, int maxnneigh,
                   const int *nneighpn, const int *jpnlist, double *force,
                   double *u_twob, double *virial);
int orc_utwobody_a_celllist(const orc_table_set *t, const orc_box *b, int np,
                            const double *ro, const int *iptpn,
                            const orc_celllist *c, double *force,
                            double *u_twob, double *virial);

/* ---- FFT in FFTW's r2c/c2r convention (unnormalised) ----
 * in: real [s3][s2][s1] (x fastest, Fortran Qmesh(s1,s2,s3));
 * out: complex interleaved [s3][s2][s1/2+1]. */
void orc_fft_r2c_3d(int s1, int s2, int s3, const double *in, double *out);
void orc_fft_c2r_3d(int s1, int s2, int s3, const double *in, double *out);

/* ---- Ewald (Src/energy.F90:905-1371, 6601-7046) ---- */
typedef struct {
  int order, s[3];
  double alpha, dkdr[3], vol, boxlen[3];
  double *b2[3];                   /* [s(i)] */
  double *meshfac, *energyfac, *virfac; /* [(s3/2+1)][(s2/2+1)][(s1/2+1)] */
} orc_spme;
int orc_spme_setup(orc_spme *p, const double boxlen[3], double alpha,
                   int order, const int s[3]);
void orc_spme_free(orc_spme *p);
/* UEwaldRecSPM + SPMFFTRec.  r: atom positions [3*na] (already PBC'd).
 * force accumulates -az*esum*dkdr*EpsiFourPi; *u_rec and *virrec receive
 * the unscaled sums (u%rec, virrec before the EpsiFourPi scaling).  If
 * qmesh_out is non-NULL the charge mesh (before the FFT) is copied there;
 * if conv_out is non-NULL the convolved mesh (after c2r) is copied there. */
void orc_spme_rec(const orc_spme *p, int na, const double *r,
                  const double *az, double epsi_four_pi, double *force,
                  double *u_rec, double *virrec, double *qmesh_out,
                  double *conv_out);
/* UEwaldRecStd with its kfac table (serial branch). */
void orc_ewald_rec_std(const double boxlen[3], double alpha, int ncut,
                       int cube, int na, const double *r, const double *az,
                       double epsi_four_pi, double *force, double *u_rec,
                       double *virrec);
/* UEwaldSelf (monoatomic: atomic term only) and UEwaldSurf (3D). */
void orc_ewald_self(double alpha, int na, const double *az, double *u_rec);
void orc_ewald_surf(const double boxlen[3], int na, const double *r,
                    const double *az, double epsi_four_pi, double *force,
                    double *u_rec, double *virrec);

/* ---- angles (Src/energy.F90:5409-5481) ---- */
/* k and eq in the stored units of IOPotChain (Src/potential.F90:3558-3562):
 * k = namelist k / (pi/180)^2 / p, eq = namelist eq * pi/180. */
void orc_uangle(const orc_box *b, const double *ro, int nchain,
                const int *chain_first, const int *chain_len, double k,
                double p, double eq, double *force, double *u_angle);

/* ---- bonds (Src/energy.F90:5335-5400) ---- */
/* Chains are runs of consecutive particles: chain c spans particles
 * chain_first[c] .. chain_first[c]+chain_len[c]-1 (1-based).  k is the
 * stored bond%k, i.e. the namelist value divided by p (IOPotChain,
 * Src/potential.F90:3561), so u = (k_in/p)(r-eq)**p. */
void orc_ubond(const orc_box *b, const double *ro, int nchain,
               const int *chain_first, const int *chain_len, double k,
               double p, double eq, double *force, double *u_bond,
               double *virial);

/* ---- whole force evaluation (UTotal, Src/energy.F90:280-434) ---- */
enum { ORC_EW_NONE = 0, ORC_EW_STD = 1, ORC_EW_SPM = 2 };
enum { ORC_NL_VLIST = 0, ORC_NL_CELL = 1 };
typedef struct {
  orc_box box;
  int np, npt;
  const int *iptpn;          /* [np] 1-based types */
  const double *az;          /* [np] charges */
  orc_table_set tab;
  int nlkind;                /* ORC_NL_VLIST / ORC_NL_CELL */
  int maxnneigh;
  const int *nneighpn;       /* vlist */
  const int *jpnlist;
  const orc_celllist *cells; /* cell list */
  int ewald;                 /* ORC_EW_* */
  double alpha, epsi_four_pi;
  int ncut, ncut_cube, lsurf;
  const orc_spme *spme;
  int nchain;                /* bonds */
  const int *chain_first, *chain_len;
  double bond_k, bond_p, bond_eq;
  double angle_k, angle_p, angle_eq; /* angles (k = 0: none), stored units */
} orc_system;
typedef struct {
  double tot, twob0, rec, bond, virial;
  double twob[64];           /* u%twob(1:nptpt) */
  double angle;              /* u%angle */
} orc_energy;
/* Evaluates UTotal.  forceo receives force*fac (kJ/mol/A -> internal
 * units of Src/energy.F90:408). Returns 0 or an error code. */
int orc_utotal(const orc_system *s, const double *ro, double *forceo,
               orc_energy *e);

/* ---- integrators ---- */
/* VelVer position half (Src/md.F90:287-298). */
void orc_velver_move(const orc_box *b, int np, double tstep, double *ro,
                     const double *rod, const double *rodd, double *drostep);
/* GetLinAcc + velocity update (Src/md.F90:348-358, 601-617). */
void orc_velver_kick(int np, double tstep, const double *massip,
                     const double *forceo, double *rod, double *rodd,
                     double *roddo);
/* BDStep with caller-supplied unit Gaussian noise [3*np] in the order the
 * reference draws it (Src/bd.F90:120-155). */
void orc_bd_step(const orc_box *b, int np, const int *iptpn,
                 const double *fac1, const double *fac2, const double *forceo,
                 const double *noise, double *ro, double *drostep);
/* BDStep's per-type factors (Src/bd.F90:136-139) with IOScale's default
 * units (Src/molsim.F90:786-801): fac1 = dt D/(R T) in A per unit forceo,
 * fac2 = sqrt(2 D dt) in A; tstep in ps, dcoeff in 1e-9 m^2/s, temp in K. */
void orc_bd_factors(int npt, double tstep, double temp, const double *dcoeff,
                    double *fac1, double *fac2);

/* ---- random numbers (Src/mollib.F90:2689-2766) ---- */
typedef struct { int flag; double save; } orc_gau_state;
double orc_gaurandom(int32_t *iseed, orc_gau_state *st);

/* ---- engine extensions without a reference counterpart ---- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the generator of the
   engine's per-replica BD noise; pinned by the Random123 known-answer
   vectors. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* noise[3*np] of BD step n: particle i draws Philox4x32-10 with counter
   (i, n, 0, 0) and key (seed lo, seed hi); uniforms u_k = (x_k + 1/2) 2^-32;
   g0, g1 = sqrt(-2 ln u0) (cos, sin)(2 pi u1), g2 = sqrt(-2 ln u2) cos(2 pi u3) */
void orc_bd_noise(uint64_t seed, int np, uint32_t step, double *noise);
/* Uniform cubic B-spline of the engine's additive pair term: weights and
   f-derivatives of coefficients i..i+3 at r (0 outside [r0, r0+(nknot-3)/dinv)). */
int orc_knot_weights(double r, double r0, double dinv, int nknot, int *i, double w[4],
                     double dw[4]);
/* The pair model the engine's adjoint differentiates, by brute force over
   all pairs with r2 <= rcut^2 (minimum image): Pot_1_6_12 with Erf2 for the
   Ewald damping (Src/potential.F90:1349-1425) plus the spline term with
   coefficients c[nptpt][nknot] (Set_iptpt pair order).  Returns the energy;
   fknot (3*np, may be NULL) receives the spline term's forces. */
double orc_model_pair_energy(const orc_box *b, int np, const int *iptpn, const double *ro,
                             int npt, const double *zat, const double *sigat,
                             const double *epsat, double efp, double alpha, int lewald,
                             int lljcut, double ljrcut, double ljushift, double rcut,
                             int nknot, double r0, double r1, const double *c,
                             double *fknot);
/* The same pair model, its energy (return value) and forces force[3*np]
   (kJ/(mol A), -dE/dro), over the pairs with r2 <= rcut^2 found through a
   cell grid of edge >= rcut (all pairs when an axis has fewer than three
   cells); the force-matching loss of the engine is built on these model
   forces. */
double orc_model_pair_forces(const orc_box *b, int np, const int *iptpn, const double *ro,
                             int npt, const double *zat, const double *sigat,
                             const double *epsat, double efp, double alpha, int lewald,
                             int lljcut, double ljrcut, double ljushift, double rcut,
                             int nknot, double r0, double r1, const double *c,
                             double *force);
typedef struct { int32_t ix, iy; double am; } orc_random_state;
double orc_random(int32_t *idum, orc_random_state *st);

#ifdef __cplusplus
}
#endif
#endif
