/*
 * mfe_oracle.c -- FP64 CPU restatement of MOLSIM's nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY (results are checked against it): used by tests/,
 * the smoke check and bench.py's cpu_baseline leg; never by the product.  Build with
 * gcc -O2 -ffp-contract=off so every product of the reference's FP64
 * expressions is rounded separately, as the reference's x86-64 build does.
 *
 * Each routine follows the cited reference routine loop for loop and keeps
 * its operation order (Fortran evaluates a+b+c as (a+b)+c, x**6 as
 * a left-to-right product chain, and CMPLX(...) without KIND in default (single)
 * precision -- all three are reproduced).
 */
#include "mfe_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* constants: Src/mol.F90:60-76                                            */
/* ---------------------------------------------------------------------- */
static const double Pi = 3.1415926535897932;
static const double AvNo = 6.02214179e23;
static const double ECh = 1.602176487e-19;
static const double Epsi0 = 8.854187817e-12;
static const double GasConstant = 8.314472;

/* default scale factors, Src/molsim.F90:786-803 */
static const double scllen = 1.0e-10, sclmas = 1.0e-03, scltim = 1.0e-12,
                    sclene = 1.0e+03;

double orc_epsi0_four_pi(void) {
  /* Epsi0FourPi = ECh**2/(FourPi*Epsi0*scllen)*AvNo/sclene (:805) */
  const double FourPi = 4.0 * Pi;
  return ((ECh * ECh) / ((FourPi * Epsi0) * scllen)) * AvNo / sclene;
}

double orc_force_unit_fac(void) {
  /* fac = (sclene/scllen)/sclfor; sclfor = sclacc*sclmas,
     sclacc = sclvel/scltim, sclvel = scllen/scltim (Src/molsim.F90:797-801,
     Src/energy.F90:408) */
  double sclvel = scllen / scltim;
  double sclacc = sclvel / scltim;
  double sclfor = sclacc * sclmas;
  return (sclene / scllen) / sclfor;
}

double orc_gas_constant(void) { return GasConstant; }

static double sq(double x) { return x * x; }
/* x**n for a constant integer n as the reference's compiler evaluates it:
   a left-to-right chain of products x*x*...*x (checked bitwise against
   the amdflang build of Src/potential.F90:1399-1404) */
static double ipow(double x, int n) {
  double res = 1.0;
  if (n <= 0) return 1.0;
  res = x;
  for (int k = 1; k < n; ++k) res = res * x;
  return res;
}

/* ---------------------------------------------------------------------- */
/* Erf2: Src/mollib.F90:215-274 (Chebyshev series, NAG mark 4)            */
/* ---------------------------------------------------------------------- */
static const double erf2_c[18] = {
    1.9449071068178803e0, 4.20186582324414e-2, -1.86866103976769e-2,
    5.1281061839107e-3,   -1.0683107461726e-3, 1.744737872522e-4,
    -2.15642065714e-5,    1.7282657974e-6,     -2.00479241e-8,
    -1.64782105e-8,       2.0008475e-9,        2.57716e-11,
    -3.06343e-11,         1.9158e-12,          3.703e-13,
    -5.43e-14,            -4.0e-15,            1.2e-15};
static const double erf2_d[17] = {
    1.4831105640848036e0, -3.010710733865950e-1, 6.89948306898316e-2,
    -1.39162712647222e-2, 2.4207995224335e-3,    -3.658639685849e-4,
    4.86209844323e-5,     -5.7492565580e-6,      6.113243578e-7,
    -5.89910153e-8,       5.2070091e-9,          -4.232976e-10,
    3.18811e-11,          -2.2361e-12,           1.467e-13,
    -9.0e-15,             5.0e-16};

double orc_erf2(double x) {
  const double sqrtpi = 1.7724538509055160, xup = 6.25;
  const int ncfc = 18, ncfd = 17;
  double xv = fabs(x), bj, bjp1, bjp2, x2;
  int j;
  if (xv <= 2.0) {
    x2 = x * x - 2.0;
    bjp2 = 0.0;
    bjp1 = erf2_d[ncfd - 1];
    for (j = ncfd - 1; j >= 2; --j) {
      bj = x2 * bjp1 - bjp2 + erf2_d[j - 1];
      bjp2 = bjp1;
      bjp1 = bj;
    }
    bj = x2 * bjp1 - bjp2 + erf2_d[0];
    return 0.5 * (bj - bjp2) * x;
  } else if (xv < xup) {
    x2 = 2.0 - 20.0 / (xv + 3.0);
    bjp2 = 0.0;
    bjp1 = erf2_c[ncfc - 1];
    /* the reference's loop runs over ncfd-1..2 (not ncfc-1..2), so c(17)
       never enters the series; after the loop j == 1 */
    for (j = ncfd - 1; j >= 2; --j) {
      bj = x2 * bjp1 - bjp2 + erf2_c[j - 1];
      bjp2 = bjp1;
      bjp1 = bj;
    }
    bj = x2 * bjp1 - bjp2 + erf2_c[0];
    x2 = 0.5 * (bj - bjp2) / xv * exp(-x * x) / sqrtpi;
    return (1.0 - x2) * (x >= 0 ? 1.0 : -1.0);
  } else {
    /* xv > xup; at xv == xup exactly the reference leaves Erf2 undefined */
    return x >= 0 ? 1.0 : -1.0;
  }
}

/* ---------------------------------------------------------------------- */
/* CardinalBSpline: Src/mollib.F90:1629-1683                              */
/* ---------------------------------------------------------------------- */
void orc_cardinal_bspline(int order, double pos, int *kmax, double *values,
                          double *derivs) {
  double last[32];
  int v = (int)floor(pos), n, i;
  double x = pos - v, a;
  *kmax = v;
  for (i = 0; i < order; ++i) last[i] = 0.0;
  values[0] = x;
  if (order > 1) {
    values[1] = 1.0 - x;
    for (n = 3; n <= order; ++n) {
      for (i = 0; i < order; ++i) last[i] = values[i];
      values[0] = x * last[0] / (n - 1);
      for (i = 1; i <= n - 2; ++i) {
        a = i + x;
        values[i] = (a * last[i] + (n - a) * last[i - 1]) / (n - 1);
      }
      values[n - 1] = (1 - x) * last[n - 2] / (n - 1);
    }
    n = order;
    if (derivs) {
      derivs[0] = last[0];
      for (i = 1; i <= n - 2; ++i) derivs[i] = last[i] - last[i - 1];
      derivs[n - 1] = -last[n - 2];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* PBC / PBCr2 (box-like cell): Src/molaux.F90:104-184; box setup         */
/* Src/molaux.F90:326-335                                                 */
/* ---------------------------------------------------------------------- */
void orc_box_init(orc_box *b, const double boxlen[3]) {
  for (int m = 0; m < 3; ++m) {
    b->boxlen[m] = boxlen[m];
    b->boxlen2[m] = boxlen[m] / 2.0;
    b->boxleni[m] = 1.0 / boxlen[m];
    b->dpbc[m] = boxlen[m];
  }
}

static double fsign(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

void orc_pbcr2(const orc_box *b, double *dx, double *dy, double *dz,
               double *r2) {
  if (fabs(*dx) > b->boxlen2[0]) *dx = *dx - fsign(b->dpbc[0], *dx);
  if (fabs(*dy) > b->boxlen2[1]) *dy = *dy - fsign(b->dpbc[1], *dy);
  if (fabs(*dz) > b->boxlen2[2]) *dz = *dz - fsign(b->dpbc[2], *dz);
  *r2 = (*dx) * (*dx) + (*dy) * (*dy) + (*dz) * (*dz);
}

void orc_pbc(const orc_box *b, double *x, double *y, double *z) {
  /* ANINT rounds half away from zero, like C round() */
  if (fabs(*x) > b->boxlen2[0]) *x = *x - b->dpbc[0] * round(*x * b->boxleni[0]);
  if (fabs(*y) > b->boxlen2[1]) *y = *y - b->dpbc[1] * round(*y * b->boxleni[1]);
  if (fabs(*z) > b->boxlen2[2]) *z = *z - b->dpbc[2] * round(*z * b->boxleni[2]);
}

/* ---------------------------------------------------------------------- */
/* (1,6,12) potential: Src/potential.F90:1349-1425; CalcEwaldUfac         */
/* Src/potential.F90:2938-2960; defaults Src/potential.F90:340-392, 411   */
/* ---------------------------------------------------------------------- */
void orc_pot_spec_defaults(orc_pot_spec *s) {
  memset(s, 0, sizeof(*s));
  s->r2uminin = 0.1;
  s->utoltab = (double)1e-5f; /* default-real literal 1e-5 */
  s->ftoltab = (double)1e-5f;
  s->umaxtab = 20000.0;
  s->fmaxtab = 20000.0;
  s->ljrcut = pow(2.0, (double)(1.0f / 6.0f)); /* Two**(1.0/6.0) */
  s->ljushift = 1.0;
  s->rumaxfac = 1.1;
}

void orc_calc_ewald_ufac(double r1, double alpha, double *c0, double *c1,
                         double *c2) {
  const double fac = 1.12837916709;
  double ra = alpha * r1;
  double er = orc_erf2(ra);
  double ex = fac * exp(-(ra * ra));
  *c0 = 1.0 - (er);
  *c1 = 1.0 - (er - ra * ex);
  *c2 = 1.0 - (er - ra * (1.0 + ra * ra) * ex);
}

void orc_pot_1_6_12(const orc_pot_spec *s, int it, int jt, double r1,
                    double *u0, double *u1, double *u2) {
  /* it, jt are 0-based type indices */
  double qq = s->zat[it] * s->zat[jt] * s->epsi_four_pi;
  double sig = 0.5 * (s->sigat[it] + s->sigat[jt]);
  double eps = sqrt(s->epsat[it] * s->epsat[jt]);
  double r1i = 1.0 / r1;
  double fac = ipow(sig * r1i, 6);
  double c0fac, c1fac, c2fac;
  double us0, us1, us2;
  if (s->lewald) {
    orc_calc_ewald_ufac(r1, s->alpha, &c0fac, &c1fac, &c2fac);
  } else {
    c0fac = 1.0;
    c1fac = 1.0;
    c2fac = 1.0;
  }
  us0 = +4.0 * eps * (sq(fac) - fac);
  us1 = -(24.0 * eps * (2.0 * sq(fac) - fac) * r1i);
  us2 = +24.0 * eps * (26.0 * sq(fac) - 7.0 * fac) * sq(r1i);
  if (s->lljcut) {
    if (r1 < s->ljrcut * sig) {
      us0 = us0 + s->ljushift * eps;
    } else {
      us0 = 0.0;
      us1 = 0.0;
      us2 = 0.0;
    }
  }
  *u0 = qq * c0fac * r1i + us0;
  *u1 = -(qq * c1fac * sq(r1i)) + us1;
  *u2 = 2.0 * qq * c2fac * ipow(r1i, 3) + us2;
}

/* SetUBuffer: Src/potential.F90:1139-1216 */
void orc_set_ubuffer(double zlow, double u0low, double u1low, double u2low,
                     double zupp, double u0upp, double u1upp, double u2upp,
                     double t[12]) {
  double rlow, rupp, dz1, dz2, dz3, w0low, w1low, w2low, w0upp, w1upp, w2upp;
  double a, b, c;
  if (u0low == 0.0 && u0upp == 0.0) {
    t[0] = zlow;
    for (int i = 1; i < 12; ++i) t[i] = 0.0;
    return;
  }
  rlow = sqrt(zlow);
  rupp = sqrt(zupp);
  dz1 = zupp - zlow;
  dz2 = sq(dz1);
  dz3 = dz2 * dz1;
  w0low = u0low;
  w0upp = u0upp;
  w1low = 0.50 * u1low / rlow;
  w1upp = 0.50 * u1upp / rupp;
  w2low = 0.25 * (u2low - u1low / rlow) / zlow;
  w2upp = 0.25 * (u2upp - u1upp / rupp) / zupp;
  a = 6.0 * (w0upp - w0low - w1low * dz1 - 0.5 * w2low * dz2) / dz3;
  b = 2.0 * (w1upp - w1low - w2low * dz1) / dz2;
  c = (w2upp - w2low) / dz1;
  t[0] = zlow;
  t[1] = w0low;
  t[2] = w1low;
  t[3] = 0.5 * w2low;
  t[4] = (10.0 * a - 12.0 * b + 3.0 * c) / (6.0);
  t[5] = (-(15.0 * a) + 21.0 * b - 6.0 * c) / (6.0 * dz1);
  t[6] = (2.0 * a - 3.0 * b + c) / (2.0 * dz2);
  t[7] = -2.0 * t[2];
  t[8] = -4.0 * t[3];
  t[9] = -6.0 * t[4];
  t[10] = -8.0 * t[5];
  t[11] = -10.0 * t[6];
}

/* CheckUBuffer: Src/potential.F90:1224-1258 */
static void check_ubuffer(const orc_pot_spec *s, int it, int jt, double rlow,
                          double rupp, const double t[12], int *ok,
                          int *repul) {
  const int ncheck = 11;
  double dr, dz, r1, u0, u1, u2, usum, fsum;
  *ok = 0;
  *repul = 0;
  dr = (rupp - rlow) / (ncheck - 1);
  for (int icheck = 1; icheck <= ncheck; ++icheck) {
    r1 = rlow + dr * (icheck - 1);
    orc_pot_1_6_12(s, it, jt, r1, &u0, &u1, &u2);
    dz = sq(r1) - sq(rlow);
    usum = t[1] + dz * (t[2] + dz * (t[3] + dz * (t[4] + dz * (t[5] + dz * t[6]))));
    fsum = t[7] + dz * (t[8] + dz * (t[9] + dz * (t[10] + dz * t[11])));
    fsum = r1 * fsum;
    if (fabs(usum - u0) > s->utoltab) return;
    if (fabs(fsum + u1) > s->ftoltab) return;
    if (fabs(usum) > s->umaxtab) *repul = 1;
    if (fabs(fsum) > s->fmaxtab) *repul = 1;
  }
  *ok = 1;
}

/* PotTwoBodyTab3: Src/potential.F90:1036-1131 */
static int pot_two_body_tab3(const orc_pot_spec *s, int it, int jt,
                             double *ubuf, int cap, int *ibuf, double *rumin,
                             double rumax, int *iubuflow, int *nugrid) {
  const int mnugrid = 1200, ndr = 100;
  const double drfac = 0.9;
  double rupp, rlow = 0, zupp, zlow, dr, u0l, u1l, u2l, u0u, u1u, u2u;
  double t[12];
  int igrid, idr, ok = 0, repul = 0;
  *iubuflow = *ibuf;
  rupp = rumax;
  dr = 0.5 * (rumax - *rumin);
  for (igrid = 1; igrid <= mnugrid; ++igrid) {
    zupp = sq(rupp);
    for (idr = 1; idr <= ndr; ++idr) {
      rlow = rupp - dr;
      rlow = fmax(*rumin, rlow);
      zlow = sq(rlow);
      orc_pot_1_6_12(s, it, jt, rlow, &u0l, &u1l, &u2l);
      orc_pot_1_6_12(s, it, jt, rupp, &u0u, &u1u, &u2u);
      orc_set_ubuffer(zlow, u0l, u1l, u2l, zupp, u0u, u1u, u2u, t);
      check_ubuffer(s, it, jt, rlow, rupp, t, &ok, &repul);
      if (ok) {
        rupp = rlow;
        if (s->lljcut) dr = 2.0 * dr;
        break;
      }
      dr = dr * drfac;
    }
    if (idr >= ndr + 1) return -2;
    if (*ibuf + 11 > cap) return -1;
    memcpy(&ubuf[*ibuf - 1], t, sizeof(t));
    *ibuf += 12;
    if (repul) *rumin = rlow;
    if (rlow <= *rumin) break;
    if (repul) break;
  }
  if (igrid >= mnugrid + 1) return -3;
  *nugrid = igrid;
  return 0;
}

/* PotTwoBodyTab1/2 for the (1,6,12) potential with monoatomic particle
   types (iat == ipt): Src/potential.F90:868-1014 and 'init' :1366-1373. */
int orc_build_tables(const orc_pot_spec *s, double *ubuf, int cap,
                     int *iubuflow, int *nugrid, double *r2umin,
                     double *r2umax) {
  int ibuf = 1, iptjpt = 0, rc;
  memset(ubuf, 0, sizeof(double) * (size_t)cap);
  for (int ipt = 0; ipt < s->npt; ++ipt) {
    for (int jpt = ipt; jpt < s->npt; ++jpt) {
      double rumin, rumax;
      r2umin[iptjpt] = s->r2uminin;
      r2umax[iptjpt] = sq(s->rumaxfac * (s->rcut + 0.0 + 0.0));
      rumax = sqrt(r2umax[iptjpt]);
      rumin = sqrt(r2umin[iptjpt]);
      rc = pot_two_body_tab3(s, ipt, jpt, ubuf, cap, &ibuf, &rumin, rumax,
                             &iubuflow[iptjpt], &nugrid[iptjpt]);
      if (rc) return rc;
      r2umin[iptjpt] = sq(rumin);
      ++iptjpt;
    }
  }
  return ibuf - 1;
}

/* table evaluation, Src/energy.F90:476-491 */
int orc_table_eval(const double *ubuf, int nbuf, int iubuflow, double r2,
                   double *usum, double *fsum) {
  int ibuf = iubuflow;
  const double *u;
  double d;
  for (;;) {
    if (r2 >= ubuf[ibuf - 1]) break;
    ibuf += 12;
    if (ibuf > nbuf) return 2;
  }
  u = &ubuf[ibuf - 1];
  d = r2 - u[0];
  *usum = u[1] + d * (u[2] + d * (u[3] + d * (u[4] + d * (u[5] + d * u[6]))));
  *fsum = u[7] + d * (u[8] + d * (u[9] + d * (u[10] + d * u[11])));
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Verlet lists: SetVListMD Src/nlist.F90:724-759; NList Src/nlist.F90:597 */
/* ---------------------------------------------------------------------- */
static int owns(int ip, int jp) {
  /* ip, jp 1-based; Src/nlist.F90:742-745 */
  int itemp = (ip + jp) % 2;
  if (ip < jp && itemp == 0) return 0;
  if (ip > jp && itemp != 0) return 0;
  if (ip == jp) return 0;
  return 1;
}

int orc_set_vlist_md(const orc_box *b, int np, const double *ro, double rs,
                     int maxnneigh, int *nneighpn, int *jpnlist) {
  double rs2cut = sq(rs), dx, dy, dz, r2;
  for (int ip = 1; ip <= np; ++ip) {
    int n = 0;
    for (int jp = 1; jp <= np; ++jp) {
      if (!owns(ip, jp)) continue;
      dx = ro[3 * (ip - 1)] - ro[3 * (jp - 1)];
      dy = ro[3 * (ip - 1) + 1] - ro[3 * (jp - 1) + 1];
      dz = ro[3 * (ip - 1) + 2] - ro[3 * (jp - 1) + 2];
      orc_pbcr2(b, &dx, &dy, &dz, &r2);
      if (r2 < rs2cut) {
        ++n;
        if (n > maxnneigh) return 4;
        jpnlist[(size_t)(ip - 1) * maxnneigh + n - 1] = jp;
      }
    }
    nneighpn[ip - 1] = n;
  }
  return 0;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int orc_set_vlist_md_cells(const orc_box *b, int np, const double *ro,
                           double rs, int maxnneigh, int *nneighpn,
                           int *jpnlist) {
  int nc[3], ncells, *head, *next, *cand, ncand_cap = 4096;
  double cs[3], rs2cut = sq(rs);
  for (int m = 0; m < 3; ++m) {
    nc[m] = (int)floor(b->boxlen[m] / rs);
    if (nc[m] < 3) nc[m] = 1;
    cs[m] = b->boxlen[m] / nc[m];
  }
  ncells = nc[0] * nc[1] * nc[2];
  head = malloc(sizeof(int) * (size_t)ncells);
  next = malloc(sizeof(int) * (size_t)np);
  cand = malloc(sizeof(int) * (size_t)ncand_cap);
  for (int c = 0; c < ncells; ++c) head[c] = -1;
  int *cellof = malloc(sizeof(int) * (size_t)np);
  for (int i = 0; i < np; ++i) {
    int ci[3];
    for (int m = 0; m < 3; ++m) {
      int k = (int)floor((ro[3 * i + m] + b->boxlen2[m]) / cs[m]);
      if (k < 0) k = 0;
      if (k >= nc[m]) k = nc[m] - 1;
      ci[m] = k;
    }
    int c = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
    cellof[i] = c;
    next[i] = head[c];
    head[c] = i;
  }
  int rc = 0;
#ifdef _OPENMP
  free(cand);
  cand = NULL;
#pragma omp parallel reduction(max : rc)
  {
    int cap = 4096;
    int *cd = malloc(sizeof(int) * (size_t)cap);
#pragma omp for schedule(dynamic, 256)
    for (int ip = 1; ip <= np; ++ip) {
      int c = cellof[ip - 1];
      int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
      int ncand = 0;
      int rx = nc[0] > 1 ? 1 : 0, ry = nc[1] > 1 ? 1 : 0, rz = nc[2] > 1 ? 1 : 0;
      for (int dzc = -rz; dzc <= rz; ++dzc)
        for (int dyc = -ry; dyc <= ry; ++dyc)
          for (int dxc = -rx; dxc <= rx; ++dxc) {
            int x = (cx + dxc + nc[0]) % nc[0], y = (cy + dyc + nc[1]) % nc[1],
                z = (cz + dzc + nc[2]) % nc[2];
            for (int j = head[x + nc[0] * (y + nc[1] * z)]; j >= 0; j = next[j]) {
              int jp = j + 1;
              double dx, dy, dz, r2;
              if (!owns(ip, jp)) continue;
              dx = ro[3 * (ip - 1)] - ro[3 * j];
              dy = ro[3 * (ip - 1) + 1] - ro[3 * j + 1];
              dz = ro[3 * (ip - 1) + 2] - ro[3 * j + 2];
              orc_pbcr2(b, &dx, &dy, &dz, &r2);
              if (r2 < rs2cut) {
                if (ncand == cap) {
                  cap *= 2;
                  cd = realloc(cd, sizeof(int) * (size_t)cap);
                }
                cd[ncand++] = jp;
              }
            }
          }
      qsort(cd, (size_t)ncand, sizeof(int), cmp_int);
      if (ncand > maxnneigh) {
        rc = 4;
        continue;
      }
      memcpy(&jpnlist[(size_t)(ip - 1) * maxnneigh], cd, sizeof(int) * (size_t)ncand);
      nneighpn[ip - 1] = ncand;
    }
    free(cd);
  }
  np = 0; /* the serial loop below is skipped */
#endif
  for (int ip = 1; ip <= np && !rc; ++ip) {
    int c = cellof[ip - 1];
    int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
    int ncand = 0;
    int rx = nc[0] > 1 ? 1 : 0, ry = nc[1] > 1 ? 1 : 0, rz = nc[2] > 1 ? 1 : 0;
    for (int dzc = -rz; dzc <= rz; ++dzc)
      for (int dyc = -ry; dyc <= ry; ++dyc)
        for (int dxc = -rx; dxc <= rx; ++dxc) {
          int x = (cx + dxc + nc[0]) % nc[0], y = (cy + dyc + nc[1]) % nc[1],
              z = (cz + dzc + nc[2]) % nc[2];
          for (int j = head[x + nc[0] * (y + nc[1] * z)]; j >= 0; j = next[j]) {
            int jp = j + 1;
            double dx, dy, dz, r2;
            if (!owns(ip, jp)) continue;
            dx = ro[3 * (ip - 1)] - ro[3 * j];
            dy = ro[3 * (ip - 1) + 1] - ro[3 * j + 1];
            dz = ro[3 * (ip - 1) + 2] - ro[3 * j + 2];
            orc_pbcr2(b, &dx, &dy, &dz, &r2);
            if (r2 < rs2cut) {
              if (ncand == ncand_cap) {
                ncand_cap *= 2;
                cand = realloc(cand, sizeof(int) * (size_t)ncand_cap);
              }
              cand[ncand++] = jp;
            }
          }
        }
    qsort(cand, (size_t)ncand, sizeof(int), cmp_int);
    if (ncand > maxnneigh) {
      rc = 4;
      break;
    }
    memcpy(&jpnlist[(size_t)(ip - 1) * maxnneigh], cand, sizeof(int) * (size_t)ncand);
    nneighpn[ip - 1] = ncand;
  }
  free(head);
  free(next);
  free(cand);
  free(cellof);
  return rc;
}

int orc_nlist_skin_test(int np, double *drosum, const double *drostep,
                        double drnlist) {
  double r2max1 = 0.0, r2max2 = 0.0, r2;
  for (int ip = 0; ip < np; ++ip) {
    for (int m = 0; m < 3; ++m) drosum[3 * ip + m] = drosum[3 * ip + m] + drostep[3 * ip + m];
    r2 = sq(drosum[3 * ip]) + sq(drosum[3 * ip + 1]) + sq(drosum[3 * ip + 2]);
    if (r2 > r2max1) {
      r2max1 = r2;
    } else if (r2 > r2max2) {
      r2max2 = r2;
    }
  }
  if (sqrt(r2max1) + sqrt(r2max2) > drnlist) {
    memset(drosum, 0, sizeof(double) * 3 * (size_t)np);
    return 1;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* HeapSortIndex: Src/mollib.F90:1749-1796 (Numerical Recipes)            */
/* ---------------------------------------------------------------------- */
void orc_heapsort_index(int n, const double *vec, int *index) {
  /* vec and index are used with 1-based semantics internally */
  int l, i, j, ir, indext;
  double vect;
  for (i = 1; i <= n; ++i) index[i - 1] = i;
  if (n == 1) return;
  l = n / 2 + 1;
  ir = n;
  for (;;) {
    if (l > 1) {
      l = l - 1;
      indext = index[l - 1];
      vect = vec[indext - 1];
    } else {
      indext = index[ir - 1];
      vect = vec[indext - 1];
      index[ir - 1] = index[0];
      ir = ir - 1;
      if (ir == 1) {
        index[0] = indext;
        return;
      }
    }
    i = l;
    j = l + l;
    while (j <= ir) {
      if (j < ir) {
        if (vec[index[j - 1] - 1] < vec[index[j] - 1]) j = j + 1;
      }
      if (vect < vec[index[j - 1] - 1]) {
        index[i - 1] = index[j - 1];
        i = j;
        j = j + j;
      } else {
        j = ir + 1;
      }
    }
    index[i - 1] = indext;
  }
}

/* ---------------------------------------------------------------------- */
/* Cell lists: InitCellList Src/celllist.F90:76-265, pcellro :295-305,    */
/* AddIpToCell :307-327, SetCellList :384-410                             */
/* ---------------------------------------------------------------------- */
int orc_celllist_init(orc_celllist *c, const orc_box *b, double rcell,
                      double rcut, int np) {
  int dirRange[3], maxdir, idir = 0, *dirs, *idx;
  double ircell = 1.0 / rcell, dr[3], *man;
  memset(c, 0, sizeof(*c));
  for (int m = 0; m < 3; ++m) {
    int n = (int)floor(b->boxlen[m] * ircell);
    c->ncell[m] = n > 1 ? n : 1;
    c->cellsize[m] = b->boxlen[m] / c->ncell[m];
    c->cellsizei[m] = 1.0 / c->cellsize[m];
  }
  for (int m = 0; m < 3; ++m) dirRange[m] = (int)ceil(rcut * c->cellsizei[m]);
  for (int m = 0; m < 3; ++m) {
    if (2 * dirRange[m] + 1 > c->ncell[m]) {
      c->ncell[m] = 1;
      dirRange[m] = 0;
      c->cellsize[m] = b->boxlen[m];
      c->cellsizei[m] = b->boxleni[m];
    }
  }
  maxdir = (2 * dirRange[0] + 1) * (2 * dirRange[1] + 1) * (2 * dirRange[2] + 1);
  dirs = malloc(sizeof(int) * 3 * (size_t)maxdir);
  for (int ix = -dirRange[0]; ix <= dirRange[0]; ++ix)
    for (int iy = -dirRange[1]; iy <= dirRange[1]; ++iy)
      for (int iz = -dirRange[2]; iz <= dirRange[2]; ++iz) {
        int d[3] = {ix, iy, iz};
        for (int m = 0; m < 3; ++m) {
          int a = abs(d[m]) - 1;
          dr[m] = (a > 0 ? a : 0) * c->cellsize[m];
        }
        if (sq(dr[0]) + sq(dr[1]) + sq(dr[2]) > rcut * rcut) continue;
        dirs[3 * idir] = ix;
        dirs[3 * idir + 1] = iy;
        dirs[3 * idir + 2] = iz;
        ++idir;
      }
  maxdir = idir;
  man = malloc(sizeof(double) * (size_t)maxdir);
  idx = malloc(sizeof(int) * (size_t)maxdir);
  for (int k = 0; k < maxdir; ++k)
    man[k] = (double)(abs(dirs[3 * k]) + abs(dirs[3 * k + 1]) + abs(dirs[3 * k + 2]));
  orc_heapsort_index(maxdir, man, idx);
  c->dirs = malloc(sizeof(int) * 3 * (size_t)maxdir);
  for (int k = 0; k < maxdir; ++k)
    for (int m = 0; m < 3; ++m) c->dirs[3 * k + m] = dirs[3 * (idx[k] - 1) + m];
  c->ndir = maxdir;
  free(dirs);
  free(man);
  free(idx);
  int ncells = c->ncell[0] * c->ncell[1] * c->ncell[2];
  c->iphead = calloc((size_t)ncells, sizeof(int));
  c->npart = calloc((size_t)ncells, sizeof(int));
  c->ipnext = calloc((size_t)np, sizeof(int));
  c->cellip = calloc((size_t)np, sizeof(int));
  c->np = np;
  return 0;
}

int orc_celllist_set(orc_celllist *c, const orc_box *b, const double *ro) {
  int ncells = c->ncell[0] * c->ncell[1] * c->ncell[2];
  memset(c->iphead, 0, sizeof(int) * (size_t)ncells);
  memset(c->npart, 0, sizeof(int) * (size_t)ncells);
  memset(c->ipnext, 0, sizeof(int) * (size_t)c->np);
  for (int ip = 1; ip <= c->np; ++ip) {
    int i[3];
    for (int m = 0; m < 3; ++m) {
      i[m] = (int)floor((ro[3 * (ip - 1) + m] + b->boxlen2[m]) * c->cellsizei[m]);
      if (i[m] < 0 || i[m] >= c->ncell[m]) return 8; /* outside the box */
    }
    int id = i[0] + i[1] * c->ncell[0] + i[2] * c->ncell[0] * c->ncell[1];
    c->npart[id] += 1;
    c->ipnext[ip - 1] = c->iphead[id];
    c->iphead[id] = ip;
    c->cellip[ip - 1] = id;
  }
  return 0;
}

void orc_celllist_free(orc_celllist *c) {
  free(c->dirs);
  free(c->iphead);
  free(c->npart);
  free(c->ipnext);
  free(c->cellip);
  memset(c, 0, sizeof(*c));
}

/* neighbour cells of cell id in InitCellList's order (:224-251) */
static int neighbour_cells(const orc_celllist *c, const orc_box *b, int id,
                           int *out) {
  int ix = id % c->ncell[0], iy = (id / c->ncell[0]) % c->ncell[1],
      iz = id / (c->ncell[0] * c->ncell[1]);
  int n = 0;
  for (int k = 0; k < c->ndir; ++k) {
    int nb[3] = {ix + c->dirs[3 * k], iy + c->dirs[3 * k + 1], iz + c->dirs[3 * k + 2]};
    int skip = 0;
    for (int m = 0; m < 3; ++m) {
      if ((nb[m] >= c->ncell[m] || nb[m] < 0) && b->dpbc[m] == b->boxlen[m]) {
        nb[m] = ((nb[m] % c->ncell[m]) + c->ncell[m]) % c->ncell[m];
      }
      if (nb[m] >= c->ncell[m] || nb[m] < 0) skip = 1;
    }
    if (skip) continue;
    out[n++] = nb[0] + nb[1] * c->ncell[0] + nb[2] * c->ncell[0] * c->ncell[1];
  }
  return n;
}

/* ---------------------------------------------------------------------- */
/* UTwoBodyA (Src/energy.F90:445-544) and UTwoBodyACellList (:669-777)   */
/* ---------------------------------------------------------------------- */
static int pair_type(int npt, int ipt, int jpt) {
  /* iptpt(ipt,jpt), Src/particle.F90:1535-1541; 1-based in, 1-based out */
  int a = ipt < jpt ? ipt : jpt, b = ipt < jpt ? jpt : ipt;
  return (a - 1) * npt - (a - 1) * (a - 2) / 2 + (b - a) + 1;
}

#ifdef _OPENMP
/* All-cores twin (built with -fopenmp only; the serial build keeps the
   reference's summation order): particles split over threads, each with
   its own force array, pair-type energies and virial, summed afterwards. */
static int utwobody_a_omp(const orc_table_set *t, const orc_box *b, int np,
                          const double *ro, const int *iptpn, int maxnneigh,
                          const int *nneighpn, const int *jpnlist, double *force,
                          double *u_twob, double *virial) {
  int nptpt = t->npt * (t->npt + 1) / 2, err = 0;
  double vir = 0.0;
  for (int k = 0; k <= nptpt; ++k) u_twob[k] = 0.0;
#pragma omp parallel reduction(+ : vir) reduction(max : err)
  {
    double *f = calloc(3 * (size_t)np, sizeof(double));
    double ut[64] = {0};
    double usum = 0.0, fsum = 0.0;
#pragma omp for schedule(dynamic, 256)
    for (int ip = 1; ip <= np; ++ip) {
      int ipt = iptpn[ip - 1];
      for (int jl = 0; jl < nneighpn[ip - 1]; ++jl) {
        int jp = jpnlist[(size_t)(ip - 1) * maxnneigh + jl];
        int iptjpt = pair_type(t->npt, ipt, iptpn[jp - 1]);
        double dx = ro[3 * (ip - 1)] - ro[3 * (jp - 1)];
        double dy = ro[3 * (ip - 1) + 1] - ro[3 * (jp - 1) + 1];
        double dz = ro[3 * (ip - 1) + 2] - ro[3 * (jp - 1) + 2];
        double r2;
        orc_pbcr2(b, &dx, &dy, &dz, &r2);
        if (r2 > t->rcut2) continue;
        if (r2 < t->r2atat[iptjpt - 1]) {
          usum = 1e10;
        } else if (r2 < t->r2umin[iptjpt - 1]) {
          err = 2;
          continue;
        } else if (orc_table_eval(t->ubuf, t->nbuf, t->iubuflow[iptjpt - 1], r2, &usum, &fsum)) {
          err = 3;
          continue;
        }
        ut[iptjpt] += usum;
        f[3 * (ip - 1)] += fsum * dx;
        f[3 * (ip - 1) + 1] += fsum * dy;
        f[3 * (ip - 1) + 2] += fsum * dz;
        f[3 * (jp - 1)] -= fsum * dx;
        f[3 * (jp - 1) + 1] -= fsum * dy;
        f[3 * (jp - 1) + 2] -= fsum * dz;
        vir -= fsum * r2;
      }
    }
#pragma omp critical
    {
      for (size_t k = 0; k < 3 * (size_t)np; ++k) force[k] += f[k];
      for (int k = 1; k <= nptpt; ++k) u_twob[k] += ut[k];
    }
    free(f);
  }
  if (err) return err;
  u_twob[0] = 0.0;
  for (int k = 1; k <= nptpt; ++k) u_twob[0] += u_twob[k];
  *virial += vir;
  return 0;
}
#endif

int orc_utwobody_a(const orc_table_set *t, const orc_box *b, int np,
                   const double *ro, const int *iptpn, int maxnneigh,
                   const int *nneighpn, const int *jpnlist, double *force,
                   double *u_twob, double *virial) {
#ifdef _OPENMP
  return utwobody_a_omp(t, b, np, ro, iptpn, maxnneigh, nneighpn, jpnlist, force, u_twob,
                        virial);
#endif
  int nptpt = t->npt * (t->npt + 1) / 2;
  double virtwob = 0.0, usum = 0.0, fsum = 0.0;
  for (int k = 0; k <= nptpt; ++k) u_twob[k] = 0.0;
  for (int ip = 1; ip <= np; ++ip) {
    int ipt = iptpn[ip - 1];
    for (int jl = 0; jl < nneighpn[ip - 1]; ++jl) {
      int jp = jpnlist[(size_t)(ip - 1) * maxnneigh + jl];
      int jpt = iptpn[jp - 1];
      int iptjpt = pair_type(t->npt, ipt, jpt);
      double dx = ro[3 * (ip - 1)] - ro[3 * (jp - 1)];
      double dy = ro[3 * (ip - 1) + 1] - ro[3 * (jp - 1) + 1];
      double dz = ro[3 * (ip - 1) + 2] - ro[3 * (jp - 1) + 2];
      double r2;
      orc_pbcr2(b, &dx, &dy, &dz, &r2);
      if (r2 > t->rcut2) continue;
      if (r2 < t->r2atat[iptjpt - 1]) {
        usum = 1e10; /* fsum keeps the previous pair's value (:476-477) */
      } else if (r2 < t->r2umin[iptjpt - 1]) {
        return 2;
      } else {
        if (orc_table_eval(t->ubuf, t->nbuf, t->iubuflow[iptjpt - 1], r2, &usum, &fsum))
          return 3;
      }
      u_twob[iptjpt] = u_twob[iptjpt] + usum;
      force[3 * (ip - 1)] = force[3 * (ip - 1)] + (fsum * dx);
      force[3 * (ip - 1) + 1] = force[3 * (ip - 1) + 1] + (fsum * dy);
      force[3 * (ip - 1) + 2] = force[3 * (ip - 1) + 2] + (fsum * dz);
      force[3 * (jp - 1)] = force[3 * (jp - 1)] - (fsum * dx);
      force[3 * (jp - 1) + 1] = force[3 * (jp - 1) + 1] - (fsum * dy);
      force[3 * (jp - 1) + 2] = force[3 * (jp - 1) + 2] - (fsum * dz);
      virtwob = virtwob - (fsum * r2);
    }
  }
  u_twob[0] = 0.0;
  for (int k = 1; k <= nptpt; ++k) u_twob[0] += u_twob[k];
  *virial += virtwob;
  return 0;
}

int orc_utwobody_a_celllist(const orc_table_set *t, const orc_box *b, int np,
                            const double *ro, const int *iptpn,
                            const orc_celllist *c, double *force,
                            double *u_twob, double *virial) {
  int nptpt = t->npt * (t->npt + 1) / 2;
  int *ncl = malloc(sizeof(int) * (size_t)(c->ndir > 0 ? c->ndir : 1));
  double virtwob = 0.0, usum, fsum;
  for (int k = 0; k <= nptpt; ++k) u_twob[k] = 0.0;
  for (int ip = 1; ip <= np; ++ip) {
    int ipt = iptpn[ip - 1];
    int nn = neighbour_cells(c, b, c->cellip[ip - 1], ncl);
    for (int inc = 0; inc < nn; ++inc) {
      int cid = ncl[inc];
      int jp = c->iphead[cid];
      for (int jl = 1; jl <= c->npart[cid]; ++jl) {
        if (jp > ip) {
          int jpt = iptpn[jp - 1];
          int iptjpt = pair_type(t->npt, ipt, jpt);
          double dr[3], r2;
          for (int m = 0; m < 3; ++m) dr[m] = ro[3 * (ip - 1) + m] - ro[3 * (jp - 1) + m];
          orc_pbcr2(b, &dr[0], &dr[1], &dr[2], &r2);
          if (r2 < t->rcut2) {
            if (r2 < t->r2atat[iptjpt - 1]) {
              usum = 1e10;
              fsum = 1e10;
            } else if (r2 < t->r2umin[iptjpt - 1]) {
              free(ncl);
              return 2;
            } else {
              if (orc_table_eval(t->ubuf, t->nbuf, t->iubuflow[iptjpt - 1], r2, &usum, &fsum)) {
                free(ncl);
                return 3;
              }
            }
            u_twob[iptjpt] = u_twob[iptjpt] + usum;
            for (int m = 0; m < 3; ++m) {
              force[3 * (ip - 1) + m] = force[3 * (ip - 1) + m] + (fsum * dr[m]);
              force[3 * (jp - 1) + m] = force[3 * (jp - 1) + m] - (fsum * dr[m]);
            }
            virtwob = virtwob - (fsum * r2);
          }
        }
        jp = c->ipnext[jp - 1];
      }
    }
  }
  free(ncl);
  u_twob[0] = 0.0;
  for (int k = 1; k <= nptpt; ++k) u_twob[0] += u_twob[k];
  *virial += virtwob;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* FFT (plain recursive mixed radix, FP64), FFTW r2c/c2r conventions       */
/* ---------------------------------------------------------------------- */
/* X[k] = sum_r w^(r k) Sub_r[k mod m] for the smallest prime factor p of n;
   tw holds exp(sign*2*pi*i*e/N) for the top-level length N, and twm = N/n
   maps this level's exponents onto it. */
static void dft_rec(int n, const double complex *in, int stride,
                    double complex *out, const double complex *tw, int N,
                    int twm) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = 2;
  while (n % p) ++p;
  int m = n / p;
  double complex *sub = malloc(sizeof(double complex) * (size_t)n);
  for (int r = 0; r < p; ++r)
    dft_rec(m, in + (size_t)r * stride, stride * p, sub + (size_t)r * m, tw, N,
            twm * p);
  for (int k = 0; k < n; ++k) {
    double complex acc = sub[k % m];
    for (int r = 1; r < p; ++r) {
      long e = ((long)r * k) % n;
      acc += sub[(size_t)r * m + (k % m)] * tw[(e * twm) % N];
    }
    out[k] = acc;
  }
  free(sub);
}

static void fft1d(int n, double complex *data, int sign) {
  double complex *tw = malloc(sizeof(double complex) * (size_t)n);
  double complex *res = malloc(sizeof(double complex) * (size_t)n);
  for (int e = 0; e < n; ++e) {
    double ang = sign * 2.0 * Pi * (double)e / (double)n;
    tw[e] = cos(ang) + I * sin(ang);
  }
  dft_rec(n, data, 1, res, tw, n, 1);
  memcpy(data, res, sizeof(double complex) * (size_t)n);
  free(tw);
  free(res);
}

void orc_fft_r2c_3d(int s1, int s2, int s3, const double *in, double *out) {
  int h = s1 / 2 + 1, nmax = s1 > s2 ? (s1 > s3 ? s1 : s3) : (s2 > s3 ? s2 : s3);
  double complex *c = (double complex *)out;
  /* every row transform is independent (the all-cores twin splits them) */
#pragma omp parallel
  {
    double complex *row = malloc(sizeof(double complex) * (size_t)nmax);
#pragma omp for collapse(2)
    for (int z = 0; z < s3; ++z)
      for (int y = 0; y < s2; ++y) {
        for (int x = 0; x < s1; ++x) row[x] = in[((size_t)z * s2 + y) * s1 + x];
        fft1d(s1, row, -1);
        for (int x = 0; x < h; ++x) c[((size_t)z * s2 + y) * h + x] = row[x];
      }
#pragma omp for collapse(2)
    for (int z = 0; z < s3; ++z)
      for (int x = 0; x < h; ++x) {
        for (int y = 0; y < s2; ++y) row[y] = c[((size_t)z * s2 + y) * h + x];
        fft1d(s2, row, -1);
        for (int y = 0; y < s2; ++y) c[((size_t)z * s2 + y) * h + x] = row[y];
      }
#pragma omp for collapse(2)
    for (int y = 0; y < s2; ++y)
      for (int x = 0; x < h; ++x) {
        for (int z = 0; z < s3; ++z) row[z] = c[((size_t)z * s2 + y) * h + x];
        fft1d(s3, row, -1);
        for (int z = 0; z < s3; ++z) c[((size_t)z * s2 + y) * h + x] = row[z];
      }
    free(row);
  }
}

void orc_fft_c2r_3d(int s1, int s2, int s3, const double *in, double *out) {
  int h = s1 / 2 + 1, nmax = s1 > s2 ? (s1 > s3 ? s1 : s3) : (s2 > s3 ? s2 : s3);
  size_t nc = (size_t)h * s2 * s3;
  double complex *c = malloc(sizeof(double complex) * nc);
  memcpy(c, in, sizeof(double complex) * nc);
#pragma omp parallel
  {
    double complex *row = malloc(sizeof(double complex) * (size_t)nmax);
#pragma omp for collapse(2)
    for (int y = 0; y < s2; ++y)
      for (int x = 0; x < h; ++x) {
        for (int z = 0; z < s3; ++z) row[z] = c[((size_t)z * s2 + y) * h + x];
        fft1d(s3, row, +1);
        for (int z = 0; z < s3; ++z) c[((size_t)z * s2 + y) * h + x] = row[z];
      }
#pragma omp for collapse(2)
    for (int z = 0; z < s3; ++z)
      for (int x = 0; x < h; ++x) {
        for (int y = 0; y < s2; ++y) row[y] = c[((size_t)z * s2 + y) * h + x];
        fft1d(s2, row, +1);
        for (int y = 0; y < s2; ++y) c[((size_t)z * s2 + y) * h + x] = row[y];
      }
#pragma omp for collapse(2)
    for (int z = 0; z < s3; ++z)
      for (int y = 0; y < s2; ++y) {
        const double complex *src = &c[((size_t)z * s2 + y) * h];
        for (int x = 0; x < h; ++x) row[x] = src[x];
        for (int x = h; x < s1; ++x) row[x] = conj(src[s1 - x]);
        fft1d(s1, row, +1);
        for (int x = 0; x < s1; ++x) out[((size_t)z * s2 + y) * s1 + x] = creal(row[x]);
      }
    free(row);
  }
  free(c);
}

/* ---------------------------------------------------------------------- */
/* SPME: EwaldSetup Src/energy.F90:6855-7019; UEwaldRecSPM :1207-1297;   */
/* SPMFFTRec :6525-6587                                                   */
/* ---------------------------------------------------------------------- */
int orc_spme_setup(orc_spme *p, const double boxlen[3], double alpha,
                   int order, const int s[3]) {
  double spl[32], TwoPiBoxi[3];
  int z;
  const double TwoPi = 2.0 * Pi, FourPi = 4.0 * Pi;
  memset(p, 0, sizeof(*p));
  if (s[0] % 2 || s[1] % 2 || s[2] % 2) return 5;
  p->order = order;
  p->alpha = alpha;
  for (int m = 0; m < 3; ++m) {
    p->s[m] = s[m];
    p->boxlen[m] = boxlen[m];
    p->dkdr[m] = s[m] / boxlen[m];
    TwoPiBoxi[m] = TwoPi * (1.0 / boxlen[m]);
  }
  p->vol = boxlen[0] * boxlen[1] * boxlen[2];
  orc_cardinal_bspline(order, 0.0, &z, spl, NULL);
  for (int i = 0; i < 3; ++i) {
    p->b2[i] = calloc((size_t)s[i], sizeof(double));
    for (int m = 0; m < s[i]; ++m) {
      double complex a = 0;
      if ((order % 2 == 1) && (2 * m == s[i])) continue;
      for (int k = 0; k <= order - 2; ++k) {
        /* exp(cmplx(0, x)): default-kind (single precision) complex */
        float th = (float)(TwoPi * m * k / s[i]);
        double complex e = (double)cosf(th) + I * (double)sinf(th);
        a = a + spl[k + 1] * e;
      }
      p->b2[i][m] = 1.0 / (sq(creal(a)) + sq(cimag(a)));
    }
  }
  int h1 = s[0] / 2 + 1, h2 = s[1] / 2 + 1, h3 = s[2] / 2 + 1;
  size_t nf = (size_t)h1 * h2 * h3;
  p->meshfac = calloc(nf, sizeof(double));
  p->energyfac = calloc(nf, sizeof(double));
  p->virfac = calloc(nf, sizeof(double));
  double fac0 = FourPi / p->vol, expf_ = -1.0 / (4.0 * sq(alpha));
  for (int nz = 0; nz < h3; ++nz)
    for (int ny = 0; ny < h2; ++ny)
      for (int nx = 0; nx < h1; ++nx) {
        size_t o = ((size_t)nz * h2 + ny) * h1 + nx;
        if (nx == 0 && ny == 0 && nz == 0) continue;
        double m2 = sq(nx * TwoPiBoxi[0]) + sq(ny * TwoPiBoxi[1]) + sq(nz * TwoPiBoxi[2]);
        double c = exp(m2 * expf_) / m2;
        double mf = fac0 * c * p->b2[0][nx] * p->b2[1][ny] * p->b2[2][nz];
        p->meshfac[o] = mf;
        if (nx == (s[0] - nx) % s[0]) mf = mf / 2;
        if (ny == (s[1] - ny) % s[1]) mf = mf / 2;
        if (nz == (s[2] - nz) % s[2]) mf = mf / 2;
        p->energyfac[o] = mf;
        p->virfac[o] = (1.0 + 2 * m2 * expf_) * mf;
      }
  return 0;
}

void orc_spme_free(orc_spme *p) {
  for (int i = 0; i < 3; ++i) free(p->b2[i]);
  free(p->meshfac);
  free(p->energyfac);
  free(p->virfac);
  memset(p, 0, sizeof(*p));
}

void orc_spme_rec(const orc_spme *p, int na, const double *r,
                  const double *az, double epsi_four_pi, double *force,
                  double *u_rec, double *virrec, double *qmesh_out,
                  double *conv_out) {
  const int order = p->order, s1 = p->s[0], s2 = p->s[1], s3 = p->s[2];
  const int h1 = s1 / 2 + 1, h2 = s2 / 2 + 1;
  size_t M = (size_t)s1 * s2 * s3, Mc = (size_t)h1 * s2 * s3;
  double *Q = calloc(M, sizeof(double));
  double *FQ = malloc(sizeof(double) * 2 * Mc);
  double *spline = malloc(sizeof(double) * (size_t)order * 3 * (size_t)na);
  double *deriv = malloc(sizeof(double) * (size_t)order * 3 * (size_t)na);
  int *meshmax = malloc(sizeof(int) * 3 * (size_t)na);
  /* MakeGQMesh (:1226-1245); the all-cores twin spreads into a mesh per
     thread and adds them up afterwards */
#ifdef _OPENMP
#pragma omp parallel
  {
    double *Qt = calloc(M, sizeof(double));
#pragma omp for
    for (int ia = 0; ia < na; ++ia) {
      for (int m = 0; m < 3; ++m) {
        int km;
        orc_cardinal_bspline(order, p->dkdr[m] * r[3 * ia + m], &km,
                             &spline[((size_t)ia * 3 + m) * order],
                             &deriv[((size_t)ia * 3 + m) * order]);
        meshmax[3 * ia + m] = (km + p->s[m]) % p->s[m];
      }
      const double *sx = &spline[((size_t)ia * 3 + 0) * order];
      const double *sy = &spline[((size_t)ia * 3 + 1) * order];
      const double *sz = &spline[((size_t)ia * 3 + 2) * order];
      for (int iz = 0; iz < order; ++iz) {
        int nz = (s3 + meshmax[3 * ia + 2] - iz) % s3;
        for (int iy = 0; iy < order; ++iy) {
          int ny = (s2 + meshmax[3 * ia + 1] - iy) % s2;
          double qyz = az[ia] * sz[iz] * sy[iy];
          for (int ix = 0; ix < order; ++ix) {
            int nx = (s1 + meshmax[3 * ia] - ix) % s1;
            Qt[((size_t)nz * s2 + ny) * s1 + nx] += qyz * sx[ix];
          }
        }
      }
    }
#pragma omp critical
    for (size_t o = 0; o < M; ++o) Q[o] += Qt[o];
    free(Qt);
  }
  if (0)
#endif
  for (int ia = 0; ia < na; ++ia) {
    for (int m = 0; m < 3; ++m) {
      int km;
      orc_cardinal_bspline(order, p->dkdr[m] * r[3 * ia + m], &km,
                           &spline[((size_t)ia * 3 + m) * order],
                           &deriv[((size_t)ia * 3 + m) * order]);
      meshmax[3 * ia + m] = (km + p->s[m]) % p->s[m];
    }
    const double *sx = &spline[((size_t)ia * 3 + 0) * order];
    const double *sy = &spline[((size_t)ia * 3 + 1) * order];
    const double *sz = &spline[((size_t)ia * 3 + 2) * order];
    for (int iz = 0; iz < order; ++iz) {
      int nz = (s3 + meshmax[3 * ia + 2] - iz) % s3;
      double qz = az[ia] * sz[iz];
      for (int iy = 0; iy < order; ++iy) {
        int ny = (s2 + meshmax[3 * ia + 1] - iy) % s2;
        double qyz = qz * sy[iy];
        for (int ix = 0; ix < order; ++ix) {
          int nx = (s1 + meshmax[3 * ia] - ix) % s1;
          double q = qyz * sx[ix];
          size_t o = ((size_t)nz * s2 + ny) * s1 + nx;
          Q[o] = Q[o] + q;
        }
      }
    }
  }
  if (qmesh_out) memcpy(qmesh_out, Q, sizeof(double) * M);
  /* SPMFFTRec (:6543-6585) */
  orc_fft_r2c_3d(s1, s2, s3, Q, FQ);
  double complex *F = (double complex *)FQ;
  double uloc = 0.0, virloc = *virrec;
  /* distinct nz touch distinct z planes (nz and s3 - nz) */
#pragma omp parallel for reduction(+ : uloc) reduction(- : virloc)
  for (int nz = 0; nz <= s3 / 2; ++nz) {
    int nza = (s3 - nz) % s3;
    for (int ny = 0; ny <= s2 / 2; ++ny) {
      int nya = (s3 - ny) % s3; /* the reference uses s(3) on the y axis */
      for (int nx = 0; nx <= s1 / 2; ++nx) {
        size_t o1 = ((size_t)nz * s2 + ny) * h1 + nx;
        size_t o2 = ((size_t)nza * s2 + ny) * h1 + nx;
        size_t o3 = ((size_t)nz * s2 + nya) * h1 + nx;
        size_t o4 = ((size_t)nza * s2 + nya) * h1 + nx;
        double complex v1 = F[o1], v2 = F[o2], v3 = F[o3], v4 = F[o4];
        size_t of = ((size_t)nz * h2 + ny) * h1 + nx;
        double Qsum = sq(creal(v1)) + sq(cimag(v1)) + sq(creal(v2)) + sq(cimag(v2)) +
                      sq(creal(v3)) + sq(cimag(v3)) + sq(creal(v4)) + sq(cimag(v4));
        uloc = uloc + Qsum * p->energyfac[of];
        virloc = virloc - Qsum * p->virfac[of];
        double mf = p->meshfac[of];
        F[o1] = v1 * mf;
        F[o2] = v2 * mf;
        F[o3] = v3 * mf;
        F[o4] = v4 * mf;
      }
    }
  }
  orc_fft_c2r_3d(s1, s2, s3, FQ, Q);
  if (conv_out) memcpy(conv_out, Q, sizeof(double) * M);
  *u_rec = uloc;
  *virrec = virloc;
  /* CalcProp (:1259-1285) */
#pragma omp parallel for
  for (int ia = 0; ia < na; ++ia) {
    double esum[3] = {0, 0, 0};
    const double *sx = &spline[((size_t)ia * 3 + 0) * order];
    const double *sy = &spline[((size_t)ia * 3 + 1) * order];
    const double *sz = &spline[((size_t)ia * 3 + 2) * order];
    const double *dx_ = &deriv[((size_t)ia * 3 + 0) * order];
    const double *dy_ = &deriv[((size_t)ia * 3 + 1) * order];
    const double *dz_ = &deriv[((size_t)ia * 3 + 2) * order];
    for (int iz = 0; iz < order; ++iz) {
      int nz = (s3 + meshmax[3 * ia + 2] - iz) % s3;
      double splz = sz[iz], dsdz = dz_[iz];
      for (int iy = 0; iy < order; ++iy) {
        int ny = (s2 + meshmax[3 * ia + 1] - iy) % s2;
        double sply = sy[iy], dsdy = dy_[iy];
        for (int ix = 0; ix < order; ++ix) {
          int nx = (s1 + meshmax[3 * ia] - ix) % s1;
          double splx = sx[ix], dsdx = dx_[ix];
          double Qval = Q[((size_t)nz * s2 + ny) * s1 + nx];
          esum[0] = esum[0] + dsdx * sply * splz * Qval;
          esum[1] = esum[1] + splx * dsdy * splz * Qval;
          esum[2] = esum[2] + splx * sply * dsdz * Qval;
        }
      }
    }
    for (int m = 0; m < 3; ++m)
      force[3 * ia + m] = force[3 * ia + m] - az[ia] * esum[m] * p->dkdr[m] * epsi_four_pi;
  }
  free(Q);
  free(FQ);
  free(spline);
  free(deriv);
  free(meshmax);
}

/* UEwaldRecStd, serial branch (Src/energy.F90:1031-1096) with EwaldSetArray
   (:1404-1421) and the kfac table (:6757-6783) */
void orc_ewald_rec_std(const double boxlen[3], double alpha, int ncut,
                       int cube, int na, const double *r, const double *az,
                       double epsi_four_pi, double *force, double *u_rec,
                       double *virrec) {
  const double TwoPi = 2.0 * Pi;
  double TwoPiBoxi[3], vol = boxlen[0] * boxlen[1] * boxlen[2];
  int ncut2 = cube ? 3 * ncut * ncut : ncut * ncut;
  for (int m = 0; m < 3; ++m) TwoPiBoxi[m] = TwoPi * (1.0 / boxlen[m]);
  size_t nk1 = (size_t)ncut + 1;
  double complex *ex = malloc(sizeof(double complex) * (size_t)na * nk1);
  double complex *ey = malloc(sizeof(double complex) * (size_t)na * nk1);
  double complex *ez = malloc(sizeof(double complex) * (size_t)na * nk1);
  for (int ia = 0; ia < na; ++ia) {
    double complex *e[3] = {ex + ia * nk1, ey + ia * nk1, ez + ia * nk1};
    for (int m = 0; m < 3; ++m) {
      double arg = TwoPiBoxi[m] * r[3 * ia + m];
      e[m][0] = 1.0;
      /* cmplx(cos(.), sin(.)) without KIND: rounded to single precision */
      if (ncut >= 1) e[m][1] = (double)(float)cos(arg) + I * (double)(float)sin(arg);
      for (int ic = 2; ic <= ncut; ++ic) e[m][ic] = e[m][ic - 1] * e[m][1];
    }
  }
  /* kfac in the reference's (nz, ny, nx) order */
  int nkvec = 0;
  double *kfac = malloc(sizeof(double) * nk1 * nk1 * nk1);
  double fac0 = TwoPi / vol;
  for (int nz = 0; nz <= ncut; ++nz) {
    double facz = nz == 0 ? 1.0 : 2.0;
    for (int ny = 0; ny <= ncut; ++ny) {
      double facy = ny == 0 ? 0.5 : 1.0;
      for (int nx = 0; nx <= ncut; ++nx) {
        if (nx * nx + ny * ny + nz * nz > ncut2) continue;
        if (nx == 0 && ny == 0 && nz == 0) continue;
        double facx = nx == 0 ? 0.5 : 1.0;
        double k2 = sq(nx * TwoPiBoxi[0]) + sq(ny * TwoPiBoxi[1]) + sq(nz * TwoPiBoxi[2]);
        kfac[nkvec++] = fac0 * facz * facy * facx * exp(-k2 / (4.0 * sq(alpha))) / k2;
      }
    }
  }
  double complex *eyzm = malloc(sizeof(double complex) * (size_t)na);
  double complex *eyzp = malloc(sizeof(double complex) * (size_t)na);
  double complex *eikr = malloc(sizeof(double complex) * 4 * (size_t)na);
  double ualpha2 = sq(alpha);
  int kn = 0;
  for (int nz = 0; nz <= ncut; ++nz) {
    for (int ny = 0; ny <= ncut; ++ny) {
      if (ny * ny + nz * nz > ncut2) continue;
      for (int ia = 0; ia < na; ++ia) {
        eyzm[ia] = conj(ey[ia * nk1 + ny]) * ez[ia * nk1 + nz];
        eyzp[ia] = ey[ia * nk1 + ny] * ez[ia * nk1 + nz];
      }
      for (int nx = 0; nx <= ncut; ++nx) {
        if (nx * nx + ny * ny + nz * nz > ncut2) continue;
        if (nx == 0 && ny == 0 && nz == 0) continue;
        double complex S[4] = {0, 0, 0, 0};
        for (int ia = 0; ia < na; ++ia) {
          double complex exx = ex[ia * nk1 + nx];
          eikr[4 * ia + 0] = az[ia] * conj(exx) * eyzm[ia];
          eikr[4 * ia + 1] = az[ia] * conj(exx) * eyzp[ia];
          eikr[4 * ia + 2] = az[ia] * exx * eyzm[ia];
          eikr[4 * ia + 3] = az[ia] * exx * eyzp[ia];
          for (int q = 0; q < 4; ++q) S[q] = S[q] + eikr[4 * ia + q];
        }
        double ssum = 0.0;
        for (int q = 0; q < 4; ++q) ssum += sq(creal(S[q])) + sq(cimag(S[q]));
        double term = kfac[kn] * ssum;
        *u_rec = *u_rec + term;
        double kfac2 = epsi_four_pi * 2.0 * kfac[kn];
        double kx = kfac2 * nx * TwoPiBoxi[0];
        double ky = kfac2 * ny * TwoPiBoxi[1];
        double kz = kfac2 * nz * TwoPiBoxi[2];
        for (int ia = 0; ia < na; ++ia) {
          double f[4];
          for (int q = 0; q < 4; ++q)
            f[q] = -creal(eikr[4 * ia + q]) * cimag(S[q]) + cimag(eikr[4 * ia + q]) * creal(S[q]);
          force[3 * ia] = force[3 * ia] + kx * (-f[0] - f[1] + f[2] + f[3]);
          force[3 * ia + 1] = force[3 * ia + 1] + ky * (-f[0] + f[1] - f[2] + f[3]);
          force[3 * ia + 2] = force[3 * ia + 2] + kz * (+f[0] + f[1] + f[2] + f[3]);
        }
        double k2 = sq(nx * TwoPiBoxi[0]) + sq(ny * TwoPiBoxi[1]) + sq(nz * TwoPiBoxi[2]);
        *virrec = *virrec - (1.0 - k2 / (2.0 * ualpha2)) * term;
        ++kn;
      }
    }
  }
  free(ex);
  free(ey);
  free(ez);
  free(kfac);
  free(eyzm);
  free(eyzp);
  free(eikr);
}

/* UEwaldSelf, atomic term (Src/energy.F90:1309-1311) */
void orc_ewald_self(double alpha, int na, const double *az, double *u_rec) {
  double fac = alpha / sqrt(Pi), s = 0.0;
  for (int ia = 0; ia < na; ++ia) s += sq(az[ia]);
  *u_rec = *u_rec - fac * s;
}

/* UEwaldSurf, 3D periodic (Src/energy.F90:1352-1362) */
void orc_ewald_surf(const double boxlen[3], int na, const double *r,
                    const double *az, double epsi_four_pi, double *force,
                    double *u_rec, double *virrec) {
  double vol = boxlen[0] * boxlen[1] * boxlen[2];
  double fac = 2.0 * Pi / (3.0 * vol), sx = 0, sy = 0, sz = 0;
  for (int ia = 0; ia < na; ++ia) sx += az[ia] * r[3 * ia];
  for (int ia = 0; ia < na; ++ia) sy += az[ia] * r[3 * ia + 1];
  for (int ia = 0; ia < na; ++ia) sz += az[ia] * r[3 * ia + 2];
  double term = fac * (sq(sx) + sq(sy) + sq(sz));
  *u_rec = *u_rec + term;
  for (int ia = 0; ia < na; ++ia) {
    force[3 * ia] = force[3 * ia] - epsi_four_pi * (2.0 * fac) * az[ia] * sx;
    force[3 * ia + 1] = force[3 * ia + 1] - epsi_four_pi * (2.0 * fac) * az[ia] * sy;
    force[3 * ia + 2] = force[3 * ia + 2] - epsi_four_pi * (2.0 * fac) * az[ia] * sz;
  }
  *virrec = *virrec - term;
}

/* UBond (Src/energy.F90:5335-5372) with integer power p */
void orc_ubond(const orc_box *b, const double *ro, int nchain,
               const int *chain_first, const int *chain_len, double k,
               double p, double eq, double *force, double *u_bond,
               double *virial) {
  double ub = 0.0, vir = 0.0;
  int ip_pow = (int)p;
  for (int ic = 0; ic < nchain; ++ic) {
    int ip = chain_first[ic];
    for (int iseg = 2; iseg <= chain_len[ic]; ++iseg) {
      int jp = chain_first[ic] + iseg - 1;
      double dx = ro[3 * (ip - 1)] - ro[3 * (jp - 1)];
      double dy = ro[3 * (ip - 1) + 1] - ro[3 * (jp - 1) + 1];
      double dz = ro[3 * (ip - 1) + 2] - ro[3 * (jp - 1) + 2];
      double r2, r1, term, fac;
      orc_pbcr2(b, &dx, &dy, &dz, &r2);
      r1 = sqrt(r2);
      term = r1 - eq;
      fac = k * ipow(term, ip_pow - 1);
      ub = ub + fac * term;
      fac = -ip_pow * fac / r1;
      force[3 * (ip - 1)] += (fac * dx);
      force[3 * (ip - 1) + 1] += (fac * dy);
      force[3 * (ip - 1) + 2] += (fac * dz);
      force[3 * (jp - 1)] -= (fac * dx);
      force[3 * (jp - 1) + 1] -= (fac * dy);
      force[3 * (jp - 1) + 2] -= (fac * dz);
      vir = vir - (fac * r2);
      ip = jp;
    }
  }
  *u_bond = ub;
  *virial += vir;
}

/* UAngle (Src/energy.F90:5409-5481): k already divided by (pi/180)^2 and p,
 * eq in rad.  The loop carries the previous bond's unit vector, reversed,
 * as the reference does; no virial. */
void orc_uangle(const orc_box *b, const double *ro, int nchain,
                const int *chain_first, const int *chain_len, double k,
                double p, double eq, double *force, double *u_angle) {
  double ua = 0.0;
  int ip_pow = (int)p;
  if (k == 0.0) {
    *u_angle = 0.0;
    return;
  }
  for (int ic = 0; ic < nchain; ++ic) {
    if (chain_len[ic] < 3) continue;
    int jm = chain_first[ic] - 1, ip = chain_first[ic];
    double dxm = ro[3 * ip] - ro[3 * jm], dym = ro[3 * ip + 1] - ro[3 * jm + 1];
    double dzm = ro[3 * ip + 2] - ro[3 * jm + 2], r2;
    orc_pbcr2(b, &dxm, &dym, &dzm, &r2);
    double dmi = 1.0 / sqrt(r2);
    dxm = dxm * dmi;
    dym = dym * dmi;
    dzm = dzm * dmi;
    for (int iseg = 2; iseg <= chain_len[ic] - 1; ++iseg) {
      int jp = chain_first[ic] - 1 + iseg;
      double dxp = ro[3 * ip] - ro[3 * jp], dyp = ro[3 * ip + 1] - ro[3 * jp + 1];
      double dzp = ro[3 * ip + 2] - ro[3 * jp + 2];
      orc_pbcr2(b, &dxp, &dyp, &dzp, &r2);
      double dpi = 1.0 / sqrt(r2);
      dxp = dxp * dpi;
      dyp = dyp * dpi;
      dzp = dzp * dpi;
      double cosine = dxm * dxp + dym * dyp + dzm * dzp;
      double theta = acos(fmax(-1.0, fmin(cosine, 1.0)));
      double term = theta - eq;
      double fac = k * ipow(term, ip_pow - 1);
      ua = ua + fac * term;
      fac = -ip_pow * fac / sin(theta);
      double fxm = (fac * dmi) * (dxp - cosine * dxm);
      double fym = (fac * dmi) * (dyp - cosine * dym);
      double fzm = (fac * dmi) * (dzp - cosine * dzm);
      double fxp = (fac * dpi) * (dxm - cosine * dxp);
      double fyp = (fac * dpi) * (dym - cosine * dyp);
      double fzp = (fac * dpi) * (dzm - cosine * dzp);
      force[3 * jm] += fxm;
      force[3 * jm + 1] += fym;
      force[3 * jm + 2] += fzm;
      force[3 * jp] += fxp;
      force[3 * jp + 1] += fyp;
      force[3 * jp + 2] += fzp;
      force[3 * ip] -= (fxm + fxp);
      force[3 * ip + 1] -= (fym + fyp);
      force[3 * ip + 2] -= (fzm + fzp);
      jm = ip;
      ip = jp;
      dmi = dpi;
      dxm = -dxp;
      dym = -dyp;
      dzm = -dzp;
    }
  }
  *u_angle = ua;
}

/* ---------------------------------------------------------------------- */
/* UTotal: Src/energy.F90:280-434 (monoatomic, charged, serial)           */
/* ---------------------------------------------------------------------- */
int orc_utotal(const orc_system *s, const double *ro, double *forceo,
               orc_energy *e) {
  int np = s->np, rc = 0;
  int nptpt = s->npt * (s->npt + 1) / 2;
  double *force = calloc(3 * (size_t)np, sizeof(double));
  double u_twob[65];
  memset(e, 0, sizeof(*e));
  double virial = 0.0;
  if (s->nlkind == ORC_NL_VLIST)
    rc = orc_utwobody_a(&s->tab, &s->box, np, ro, s->iptpn, s->maxnneigh,
                        s->nneighpn, s->jpnlist, force, u_twob, &virial);
  else
    rc = orc_utwobody_a_celllist(&s->tab, &s->box, np, ro, s->iptpn, s->cells,
                                 force, u_twob, &virial);
  if (rc) {
    free(force);
    return rc;
  }
  e->twob0 = u_twob[0];
  for (int k = 1; k <= nptpt && k < 64; ++k) e->twob[k - 1] = u_twob[k];
  e->tot = e->tot + u_twob[0];
  if (s->ewald != ORC_EW_NONE) {
    double urec = 0.0, virrec = 0.0;
    if (s->ewald == ORC_EW_STD)
      orc_ewald_rec_std(s->box.boxlen, s->alpha, s->ncut, s->ncut_cube, np, ro,
                        s->az, s->epsi_four_pi, force, &urec, &virrec);
    else
      orc_spme_rec(s->spme, np, ro, s->az, s->epsi_four_pi, force, &urec,
                   &virrec, NULL, NULL);
    orc_ewald_self(s->alpha, np, s->az, &urec);
    if (s->lsurf)
      orc_ewald_surf(s->box.boxlen, np, ro, s->az, s->epsi_four_pi, force,
                     &urec, &virrec);
    e->rec = s->epsi_four_pi * urec;
    e->tot = e->tot + e->rec;
    virial = virial + s->epsi_four_pi * virrec;
  }
  if (s->nchain > 0) {
    orc_ubond(&s->box, ro, s->nchain, s->chain_first, s->chain_len, s->bond_k,
              s->bond_p, s->bond_eq, force, &e->bond, &virial);
    e->tot = e->tot + e->bond;
    orc_uangle(&s->box, ro, s->nchain, s->chain_first, s->chain_len,
               s->angle_k, s->angle_p, s->angle_eq, force, &e->angle);
    e->tot = e->tot + e->angle;
  }
  double fac = orc_force_unit_fac();
  for (int i = 0; i < 3 * np; ++i) forceo[i] = force[i] * fac;
  e->virial = virial;
  free(force);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* integrators                                                            */
/* ---------------------------------------------------------------------- */
void orc_velver_move(const orc_box *b, int np, double tstep, double *ro,
                     const double *rod, const double *rodd, double *drostep) {
  double fac = 0.5 * sq(tstep);
  for (int ip = 0; ip < np; ++ip) {
    for (int m = 0; m < 3; ++m) {
      drostep[3 * ip + m] = tstep * rod[3 * ip + m] + fac * rodd[3 * ip + m];
      ro[3 * ip + m] = ro[3 * ip + m] + drostep[3 * ip + m];
    }
    orc_pbc(b, &ro[3 * ip], &ro[3 * ip + 1], &ro[3 * ip + 2]);
  }
}

void orc_velver_kick(int np, double tstep, const double *massip,
                     const double *forceo, double *rod, double *rodd,
                     double *roddo) {
  for (int i = 0; i < 3 * np; ++i) roddo[i] = rodd[i];
  for (int ip = 0; ip < np; ++ip)
    for (int m = 0; m < 3; ++m) rodd[3 * ip + m] = massip[ip] * forceo[3 * ip + m];
  for (int i = 0; i < 3 * np; ++i) rod[i] = rod[i] + 0.5 * tstep * (rodd[i] + roddo[i]);
}

void orc_bd_step(const orc_box *b, int np, const int *iptpn,
                 const double *fac1, const double *fac2, const double *forceo,
                 const double *noise, double *ro, double *drostep) {
  for (int ip = 0; ip < np; ++ip) {
    int ipt = iptpn[ip] - 1;
    for (int m = 0; m < 3; ++m)
      drostep[3 * ip + m] = fac1[ipt] * forceo[3 * ip + m] + fac2[ipt] * noise[3 * ip + m];
    for (int m = 0; m < 3; ++m) ro[3 * ip + m] = ro[3 * ip + m] + drostep[3 * ip + m];
    orc_pbc(b, &ro[3 * ip], &ro[3 * ip + 1], &ro[3 * ip + 2]);
  }
}

void orc_bd_factors(int npt, double tstep, double temp, const double *dcoeff,
                    double *fac1, double *fac2) {
  /* IOScale defaults and derived scales, in its order of operations */
  const double scllen = 1.0e-10, sclmas = 1.0e-03, scltem = 1.0, scltim = 1.0e-12;
  const double scldif = 1.0e-09, gas = 8.314472;
  const double sclvel = scllen / scltim, sclacc = sclvel / scltim;
  const double sclfor = sclacc * sclmas;
  /* left to right, as Fortran evaluates a*b*c/d/e */
  for (int t = 0; t < npt; ++t) {
    fac1[t] = tstep * scltim * dcoeff[t] * scldif * sclfor / (gas * temp * scltem) / scllen;
    fac2[t] = sqrt(2.0 * dcoeff[t] * scldif * tstep * scltim) / scllen;
  }
}

/* GauRandom (Src/mollib.F90:2739-2766): polar Box-Muller over the inline
   Lehmer generator iseed = ibclr(iseed*16087, 31) (32-bit wrap-around) */
double orc_gaurandom(int32_t *iseed, orc_gau_state *st) {
  const double f = 1.0 / 2147483648.0;
  double rx, ry, r, fac, g;
  if (st->flag == 0) {
    for (;;) {
      *iseed = (int32_t)(((uint32_t)*iseed * 16087u) & 0x7fffffffu);
      rx = 2.0 * (f * *iseed - 0.5);
      *iseed = (int32_t)(((uint32_t)*iseed * 16087u) & 0x7fffffffu);
      ry = 2.0 * (f * *iseed - 0.5);
      r = sq(rx) + sq(ry);
      if (r < 1.0) break;
    }
    fac = sqrt(-2.0 * log(r) / r);
    g = rx * fac;
    st->save = ry * fac;
    st->flag = 1;
    return g;
  }
  st->flag = 0;
  return st->save;
}

/* Random (Src/mollib.F90:2689-2709): Marsaglia xorshift combined with a
   Park-Miller generator */
double orc_random(int32_t *idum, orc_random_state *st) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  int32_t k;
  if (*idum <= 0 || st->iy < 0) {
    st->am = (double)(nextafterf(1.0f, -1.0f) / (float)im); /* default-real division */
    int32_t a = *idum < 0 ? -*idum : *idum;
    st->iy = (888889999 ^ a) | 1;
    st->ix = 777755555 ^ a;
    *idum = a + 1;
  }
  uint32_t x = (uint32_t)st->ix;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  st->ix = (int32_t)x;
  k = st->iy / iq;
  st->iy = ia * (st->iy - k * iq) - ir * k;
  if (st->iy < 0) st->iy += im;
  return st->am * (double)((im & (st->ix ^ st->iy)) | 1);
}

/* ---- engine extensions without a reference counterpart ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_bd_noise(uint64_t seed, int np, uint32_t step, double *noise) {
  const double s = 1.0 / 4294967296.0;
  const double twopi = 6.283185307179586476925286766559;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < np; ++i) {
    const uint32_t ctr[4] = {(uint32_t)i, step, 0u, 0u};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    const double u0 = ((double)x[0] + 0.5) * s, u1 = ((double)x[1] + 0.5) * s;
    const double u2 = ((double)x[2] + 0.5) * s, u3 = ((double)x[3] + 0.5) * s;
    const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
    noise[3 * i] = r0 * cos(twopi * u1);
    noise[3 * i + 1] = r0 * sin(twopi * u1);
    noise[3 * i + 2] = r1 * cos(twopi * u3);
  }
}

int orc_knot_weights(double r, double r0, double dinv, int nknot, int *i, double w[4],
                     double dw[4]) {
  const double x = (r - r0) * dinv;
  if (!(x >= 0.0) || !(x < (double)(nknot - 3))) return 0;
  *i = (int)floor(x);
  const double f = x - *i, g = 1.0 - f;
  /* de Boor's uniform cubic B-spline segments */
  w[0] = g * g * g / 6.0;
  w[1] = (3.0 * f * f * f - 6.0 * f * f + 4.0) / 6.0;
  w[2] = (-3.0 * f * f * f + 3.0 * f * f + 3.0 * f + 1.0) / 6.0;
  w[3] = f * f * f / 6.0;
  dw[0] = -0.5 * g * g;
  dw[1] = 1.5 * f * f - 2.0 * f;
  dw[2] = -1.5 * f * f + f + 0.5;
  dw[3] = 0.5 * f * f;
  return 1;
}

double orc_model_pair_energy(const orc_box *b, int np, const int *iptpn, const double *ro,
                             int npt, const double *zat, const double *sigat,
                             const double *epsat, double efp, double alpha, int lewald,
                             int lljcut, double ljrcut, double ljushift, double rcut,
                             int nknot, double r0, double r1, const double *c,
                             double *fknot) {
  const double dinv = nknot > 3 ? (nknot - 3) / (r1 - r0) : 0.0;
  double e = 0.0;
  if (fknot)
    for (int k = 0; k < 3 * np; ++k) fknot[k] = 0.0;
  for (int i = 0; i < np; ++i) {
    const int ti = iptpn[i] - 1;
    for (int j = i + 1; j < np; ++j) {
      const int tj = iptpn[j] - 1;
      double dx = ro[3 * i] - ro[3 * j], dy = ro[3 * i + 1] - ro[3 * j + 1],
             dz = ro[3 * i + 2] - ro[3 * j + 2];
      double r2;
      orc_pbcr2(b, &dx, &dy, &dz, &r2);
      if (r2 > rcut * rcut) continue;
      const double r = sqrt(r2);
      const double c0 = lewald ? 1.0 - orc_erf2(alpha * r) : 1.0;
      e += zat[ti] * zat[tj] * efp * c0 / r;
      const double sig = 0.5 * (sigat[ti] + sigat[tj]);
      const double eps = sqrt(epsat[ti] * epsat[tj]);
      if (!lljcut || r < ljrcut * sig) {
        const double x = sig / r, s6 = x * x * x * x * x * x;
        e += 4.0 * eps * (s6 * s6 - s6) + (lljcut ? ljushift * eps : 0.0);
      }
      if (nknot > 3) {
        int ki;
        double w[4], dw[4];
        if (orc_knot_weights(r, r0, dinv, nknot, &ki, w, dw)) {
          /* Set_iptpt order: (1,1), (1,2), ..., (1,npt), (2,2), ... */
          const int a = ti < tj ? ti : tj, bb = ti < tj ? tj : ti;
          const int p = a * npt - a * (a - 1) / 2 + (bb - a);
          const double *cp = c + (size_t)p * nknot + ki;
          e += cp[0] * w[0] + cp[1] * w[1] + cp[2] * w[2] + cp[3] * w[3];
          if (fknot) {
            const double dup = (cp[0] * dw[0] + cp[1] * dw[1] + cp[2] * dw[2] + cp[3] * dw[3]) * dinv;
            const double fr = -dup / r;
            fknot[3 * i] += fr * dx; fknot[3 * i + 1] += fr * dy; fknot[3 * i + 2] += fr * dz;
            fknot[3 * j] -= fr * dx; fknot[3 * j + 1] -= fr * dy; fknot[3 * j + 2] -= fr * dz;
          }
        }
      }
    }
  }
  return e;
}

/* one pair of the model: energy and fr = -(du/dr)/r */
static double model_pair(double r2, int ti, int tj, int npt, const double *zat,
                         const double *sigat, const double *epsat, double efp, double alpha,
                         int lewald, int lljcut, double ljrcut, double ljushift, int nknot,
                         double r0, double dinv, const double *c, double *fr) {
  const double r = sqrt(r2);
  const double qq = zat[ti] * zat[tj] * efp;
  double u, dudr;
  if (lewald) {
    const double ar = alpha * r;
    const double ec = 1.0 - orc_erf2(ar);
    u = qq * ec / r;
    dudr = -qq * (ec / r2 + 1.1283791670955126 * alpha * exp(-ar * ar) / r);
  } else {
    u = qq / r;
    dudr = -qq / r2;
  }
  const double sig = 0.5 * (sigat[ti] + sigat[tj]);
  const double eps = sqrt(epsat[ti] * epsat[tj]);
  if (!lljcut || r < ljrcut * sig) {
    const double x = sig / r, s6 = x * x * x * x * x * x;
    u += 4.0 * eps * (s6 * s6 - s6) + (lljcut ? ljushift * eps : 0.0);
    dudr -= 24.0 * eps * (2.0 * s6 * s6 - s6) / r;
  }
  if (nknot > 3) {
    int ki;
    double w[4], dw[4];
    if (orc_knot_weights(r, r0, dinv, nknot, &ki, w, dw)) {
      const int a = ti < tj ? ti : tj, bb = ti < tj ? tj : ti;
      const int p = a * npt - a * (a - 1) / 2 + (bb - a);
      const double *cp = c + (size_t)p * nknot + ki;
      u += cp[0] * w[0] + cp[1] * w[1] + cp[2] * w[2] + cp[3] * w[3];
      dudr += (cp[0] * dw[0] + cp[1] * dw[1] + cp[2] * dw[2] + cp[3] * dw[3]) * dinv;
    }
  }
  *fr = -dudr / r;
  return u;
}

double orc_model_pair_forces(const orc_box *b, int np, const int *iptpn, const double *ro,
                             int npt, const double *zat, const double *sigat,
                             const double *epsat, double efp, double alpha, int lewald,
                             int lljcut, double ljrcut, double ljushift, double rcut,
                             int nknot, double r0, double r1, const double *c,
                             double *force) {
  const double dinv = nknot > 3 ? (nknot - 3) / (r1 - r0) : 0.0;
  const double rc2 = rcut * rcut;
  double e = 0.0;
  for (int k = 0; k < 3 * np; ++k) force[k] = 0.0;
  int n[3];
  int cells = 1;
  for (int m = 0; m < 3; ++m) {
    n[m] = (int)floor(b->boxlen[m] / rcut);
    if (n[m] < 3) cells = 0;
  }
#define MODEL_PAIR(i, j)                                                                   \
  do {                                                                                     \
    double dx = ro[3 * (i)] - ro[3 * (j)], dy = ro[3 * (i) + 1] - ro[3 * (j) + 1],         \
           dz = ro[3 * (i) + 2] - ro[3 * (j) + 2], r2, fr;                                 \
    orc_pbcr2(b, &dx, &dy, &dz, &r2);                                                      \
    if (r2 <= rc2) {                                                                       \
      e += model_pair(r2, iptpn[i] - 1, iptpn[j] - 1, npt, zat, sigat, epsat, efp, alpha,  \
                      lewald, lljcut, ljrcut, ljushift, nknot, r0, dinv, c, &fr);          \
      force[3 * (i)] += fr * dx; force[3 * (i) + 1] += fr * dy; force[3 * (i) + 2] += fr * dz; \
      force[3 * (j)] -= fr * dx; force[3 * (j) + 1] -= fr * dy; force[3 * (j) + 2] -= fr * dz; \
    }                                                                                      \
  } while (0)
  if (!cells) {
    for (int i = 0; i < np; ++i)
      for (int j = i + 1; j < np; ++j) MODEL_PAIR(i, j);
    return e;
  }
  const int nc = n[0] * n[1] * n[2];
  int *head = (int *)malloc(sizeof(int) * (size_t)nc);
  int *next = (int *)malloc(sizeof(int) * (size_t)np);
  int *cof = (int *)malloc(sizeof(int) * 3 * (size_t)np);
  for (int k = 0; k < nc; ++k) head[k] = -1;
  for (int i = np - 1; i >= 0; --i) {
    int cc[3];
    for (int m = 0; m < 3; ++m) {
      int v = (int)floor((ro[3 * i + m] + 0.5 * b->boxlen[m]) * n[m] / b->boxlen[m]);
      cc[m] = v < 0 ? 0 : (v >= n[m] ? n[m] - 1 : v);
      cof[3 * i + m] = cc[m];
    }
    const int id = cc[0] + n[0] * (cc[1] + n[1] * cc[2]);
    next[i] = head[id];
    head[id] = i;
  }
  for (int i = 0; i < np; ++i) {
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = (cof[3 * i] + dx + n[0]) % n[0];
          const int y = (cof[3 * i + 1] + dy + n[1]) % n[1];
          const int z = (cof[3 * i + 2] + dz + n[2]) % n[2];
          for (int j = head[x + n[0] * (y + n[1] * z)]; j >= 0; j = next[j])
            if (j > i) MODEL_PAIR(i, j);
        }
  }
#undef MODEL_PAIR
  free(head);
  free(next);
  free(cof);
  return e;
}
