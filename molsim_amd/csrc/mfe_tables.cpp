// mfe_tables.cpp -- host side of the boundary: MOLSIM's two-body potential
// tables for the (1,6,12) potential, built exactly as the reference builds
// them so that a table made here and one exported by MOLSIM are identical.
//
//   PotTwoBodyTab1/2 ........ Src/potential.F90:868-1014
//   PotTwoBodyTab3 .......... Src/potential.F90:1036-1131 (Andrea-Swope-
//                             Andersen adaptive intervals in z = r^2)
//   SetUBuffer .............. Src/potential.F90:1139-1216
//   CheckUBuffer ............ Src/potential.F90:1224-1258
//   Pot_1_6_12 .............. Src/potential.F90:1349-1425
//   CalcEwaldUfac ........... Src/potential.F90:2938-2960
//   Erf2 .................... Src/mollib.F90:215-274
//
// Compiled with -ffp-contract=off: every product is rounded on its own, as
// in the reference's x86-64 build, so the tables agree bit for bit.
#include <cmath>
#include <cstring>
#include <vector>

#include "mfe.h"

namespace {

// Chebyshev coefficients of the published NAG (mark 4) erf/erfc expansion
// used by Src/mollib.F90:230-245.
constexpr double kErfC[18] = {
    1.9449071068178803e0, 4.20186582324414e-2, -1.86866103976769e-2,
    5.1281061839107e-3,   -1.0683107461726e-3, 1.744737872522e-4,
    -2.15642065714e-5,    1.7282657974e-6,     -2.00479241e-8,
    -1.64782105e-8,       2.0008475e-9,        2.57716e-11,
    -3.06343e-11,         1.9158e-12,          3.703e-13,
    -5.43e-14,            -4.0e-15,            1.2e-15};
constexpr double kErfD[17] = {
    1.4831105640848036e0, -3.010710733865950e-1, 6.89948306898316e-2,
    -1.39162712647222e-2, 2.4207995224335e-3,    -3.658639685849e-4,
    4.86209844323e-5,     -5.7492565580e-6,      6.113243578e-7,
    -5.89910153e-8,       5.2070091e-9,          -4.232976e-10,
    3.18811e-11,          -2.2361e-12,           1.467e-13,
    -9.0e-15,             5.0e-16};

// Clenshaw recurrence over coefficients cf[last..0]; the reference runs
// the loop index from 16 down to 2 for both series, so in the |x| > 2
// branch the 18-term series skips its 17th coefficient.
double clenshaw(double t, const double *cf, int top, int loop_hi) {
  double b2 = 0.0, b1 = cf[top], b0 = 0.0;
  for (int j = loop_hi; j >= 2; --j) {
    b0 = t * b1 - b2 + cf[j - 1];
    b2 = b1;
    b1 = b0;
  }
  b0 = t * b1 - b2 + cf[0];
  return 0.5 * (b0 - b2);
}

double erf_molsim(double x) {
  const double ax = std::fabs(x);
  if (ax <= 2.0) return clenshaw(x * x - 2.0, kErfD, 16, 16) * x;
  if (ax < 6.25) {
    double t = 2.0 - 20.0 / (ax + 3.0);
    double e = clenshaw(t, kErfC, 17, 16) / ax * std::exp(-x * x) /
               1.7724538509055160;
    return (1.0 - e) * (x >= 0 ? 1.0 : -1.0);
  }
  return x >= 0 ? 1.0 : -1.0;
}

inline double p2(double x) { return x * x; }
// integer powers as the reference's compiler evaluates them: x*x*...*x
inline double p3(double x) { return x * x * x; }
inline double p6(double x) { return x * x * x * x * x * x; }

struct Pot {
  const mfe_pot_1_6_12 &s;
  // u0, u(r); u1, du/dr; u2, d2u/dr2 of Pot_1_6_12('calc') for types it, jt
  void eval(int it, int jt, double r1, double &u0, double &u1,
            double &u2) const {
    const double qq = s.zat[it] * s.zat[jt] * s.epsi_four_pi;
    const double sig = 0.5 * (s.sigat[it] + s.sigat[jt]);
    const double eps = std::sqrt(s.epsat[it] * s.epsat[jt]);
    const double r1i = 1.0 / r1;
    const double fac = p6(sig * r1i);
    double c0 = 1.0, c1 = 1.0, c2 = 1.0;
    if (s.lewald) {
      const double ra = s.alpha * r1;
      const double er = erf_molsim(ra);
      const double ex = 1.12837916709 * std::exp(-(ra * ra));
      c0 = 1.0 - (er);
      c1 = 1.0 - (er - ra * ex);
      c2 = 1.0 - (er - ra * (1.0 + ra * ra) * ex);
    }
    double s0 = 4.0 * eps * (p2(fac) - fac);
    double s1 = -(24.0 * eps * (2.0 * p2(fac) - fac) * r1i);
    double s2 = 24.0 * eps * (26.0 * p2(fac) - 7.0 * fac) * p2(r1i);
    if (s.lljcut) {
      if (r1 < s.ljrcut * sig) {
        s0 = s0 + s.ljushift * eps;
      } else {
        s0 = s1 = s2 = 0.0;
      }
    }
    u0 = qq * c0 * r1i + s0;
    u1 = -(qq * c1 * p2(r1i)) + s1;
    u2 = 2.0 * qq * c2 * p3(r1i) + s2;
  }
};

// quintic (energy) and quartic (force) coefficients of one z-interval
void interval_coeffs(double zlow, const double lo[3], double zupp,
                     const double hi[3], double t[12]) {
  if (lo[0] == 0.0 && hi[0] == 0.0) {
    t[0] = zlow;
    for (int i = 1; i < 12; ++i) t[i] = 0.0;
    return;
  }
  const double rlow = std::sqrt(zlow), rupp = std::sqrt(zupp);
  const double dz1 = zupp - zlow, dz2 = p2(dz1), dz3 = dz2 * dz1;
  // derivatives with respect to z = r^2
  const double w1l = 0.50 * lo[1] / rlow, w1u = 0.50 * hi[1] / rupp;
  const double w2l = 0.25 * (lo[2] - lo[1] / rlow) / zlow;
  const double w2u = 0.25 * (hi[2] - hi[1] / rupp) / zupp;
  const double a = 6.0 * (hi[0] - lo[0] - w1l * dz1 - 0.5 * w2l * dz2) / dz3;
  const double b = 2.0 * (w1u - w1l - w2l * dz1) / dz2;
  const double c = (w2u - w2l) / dz1;
  t[0] = zlow;
  t[1] = lo[0];
  t[2] = w1l;
  t[3] = 0.5 * w2l;
  t[4] = (10.0 * a - 12.0 * b + 3.0 * c) / (6.0);
  t[5] = (-(15.0 * a) + 21.0 * b - 6.0 * c) / (6.0 * dz1);
  t[6] = (2.0 * a - 3.0 * b + c) / (2.0 * dz2);
  for (int k = 0; k < 5; ++k) t[7 + k] = -2.0 * (k + 1) * t[2 + k];
}

// accuracy of one interval on 11 equidistant radii
void check_interval(const Pot &pot, int it, int jt, double rlow, double rupp,
                    const double t[12], bool &ok, bool &repul) {
  ok = repul = false;
  const double dr = (rupp - rlow) / 10;
  for (int ic = 1; ic <= 11; ++ic) {
    const double r1 = rlow + dr * (ic - 1);
    double u0, u1, u2;
    pot.eval(it, jt, r1, u0, u1, u2);
    const double dz = p2(r1) - p2(rlow);
    const double us = t[1] + dz * (t[2] + dz * (t[3] + dz * (t[4] + dz * (t[5] + dz * t[6]))));
    double fs = t[7] + dz * (t[8] + dz * (t[9] + dz * (t[10] + dz * t[11])));
    fs = r1 * fs;
    if (std::fabs(us - u0) > pot.s.utoltab) return;
    if (std::fabs(fs + u1) > pot.s.ftoltab) return;
    if (std::fabs(us) > pot.s.umaxtab) repul = true;
    if (std::fabs(fs) > pot.s.fmaxtab) repul = true;
  }
  ok = true;
}

}  // namespace

extern "C" int mfe_pot_defaults(mfe_pot_1_6_12 *s) {
  if (!s) return MFE_E_BAD_PARAM;
  std::memset(s, 0, sizeof(*s));
  s->r2uminin = 0.1;
  s->utoltab = static_cast<double>(1e-5f);  // default-real literal
  s->ftoltab = static_cast<double>(1e-5f);
  s->umaxtab = 20000.0;
  s->fmaxtab = 20000.0;
  s->ljrcut = std::pow(2.0, static_cast<double>(1.0f / 6.0f));  // Two**(1.0/6.0)
  s->ljushift = 1.0;
  s->rumaxfac = 1.1;
  s->lewald = 1;
  return MFE_OK;
}

extern "C" int mfe_build_tables_1_6_12(const mfe_pot_1_6_12 *s, double *ubuf,
                                       int32_t cap, int32_t *nbuf_used,
                                       int32_t *iubuflow, int32_t *nugrid,
                                       double *r2umin) {
  if (!s || !ubuf || !iubuflow || !nugrid || !r2umin || s->npt < 1 ||
      s->npt > MFE_MAX_TYPES || !s->zat || !s->sigat || !s->epsat)
    return MFE_E_BAD_PARAM;
  const Pot pot{*s};
  std::memset(ubuf, 0, sizeof(double) * static_cast<size_t>(cap));
  int ibuf = 1, pair = 0;  // 1-based like the reference
  for (int it = 0; it < s->npt; ++it) {
    for (int jt = it; jt < s->npt; ++jt, ++pair) {
      const double rumax = std::sqrt(p2(s->rumaxfac * (s->rcut + 0.0 + 0.0)));
      double rumin = std::sqrt(s->r2uminin);
      iubuflow[pair] = ibuf;
      double rupp = rumax, dr = 0.5 * (rumax - rumin), rlow = 0.0;
      int igrid = 1;
      for (; igrid <= 1200; ++igrid) {
        const double zupp = p2(rupp);
        bool ok = false, repul = false;
        double t[12];
        int idr = 1;
        for (; idr <= 100; ++idr) {
          rlow = std::fmax(rumin, rupp - dr);
          double lo[3], hi[3];
          pot.eval(it, jt, rlow, lo[0], lo[1], lo[2]);
          pot.eval(it, jt, rupp, hi[0], hi[1], hi[2]);
          interval_coeffs(p2(rlow), lo, zupp, hi, t);
          check_interval(pot, it, jt, rlow, rupp, t, ok, repul);
          if (ok) {
            rupp = rlow;
            if (s->lljcut) dr = 2.0 * dr;
            break;
          }
          dr = dr * 0.9;
        }
        if (idr > 100) return MFE_E_TABLE_BUILD;
        if (ibuf + 11 > cap) return MFE_E_BAD_PARAM;
        std::memcpy(&ubuf[ibuf - 1], t, sizeof(t));
        ibuf += 12;
        if (repul) rumin = rlow;
        if (rlow <= rumin || repul) break;
      }
      if (igrid > 1200) return MFE_E_TABLE_BUILD;
      nugrid[pair] = igrid;
      r2umin[pair] = p2(rumin);
    }
  }
  *nbuf_used = ibuf - 1;
  return MFE_OK;
}

extern "C" int mfe_abi_version(void) { return MFE_ABI_VERSION; }

namespace {

constexpr double kPiE = 3.1415926535897932;  // MolModule's Pi

// The fixed-point iterations of EwaldSetup (Src/energy.F90:7052-7124),
// each expression in the reference's order of operations; 40 iterations
// at most (the reference stops otherwise).
bool alpha_to_rcut(const mfe_ewald_opt &p, double &rcut) {
  const double L = p.boxlenshort, a = p.ualpha, t = p.uewaldtol / (p.q2sum * p.epsi_four_pi);
  const double y = p.lq2sum == 0 ? t * std::sqrt(2.0 * (L * L * L)) * (a * a)
                                 : (std::sqrt(15.0) / 4.0) * t * std::sqrt(L * L * L) / (a * a);
  double xold = 1.0;
  for (int i = 1; i <= 40; ++i) {
    const double xnew = p.lq2sum == 0 ? (1.0 / a) * std::sqrt(-std::log(y * std::pow(xold, 1.5)))
                                      : (1.0 / a) * std::sqrt(-std::log(y / std::pow(xold, 0.5)));
    if (std::fabs(xold - xnew) / xnew < 1.0e-6) {
      rcut = xnew;
      return true;
    }
    xold = xnew;
  }
  return false;
}

bool alpha_to_ncut(const mfe_ewald_opt &p, int &ncut) {
  const double L = p.boxlenshort, a = p.ualpha, t = p.uewaldtol / (p.q2sum * p.epsi_four_pi);
  const double y = p.lq2sum == 0 ? t * (kPiE * kPiE / ((a * a) * L))
                                 : (std::sqrt(15.0) / 4.0) * t * L / (a * a);
  double xold = 1.0;
  for (int i = 1; i <= 40; ++i) {
    const double g = 1.0 + 1.0 / (a * L * std::sqrt(xold));
    const double xnew = p.lq2sum == 0 ? (a * L / kPiE) * std::sqrt(-std::log(y * xold / g))
                                      : (a * L / kPiE) * std::sqrt(-std::log(y / xold / g));
    if (std::fabs(xold - xnew) / xnew < 1.0e-6) {
      ncut = (int)(xnew + 1.0);
      return true;
    }
    xold = xnew;
  }
  return false;
}

bool rcut_to_alpha(const mfe_ewald_opt &p, double &alpha) {
  const double L = p.boxlenshort, r = p.rcut, t = p.uewaldtol / (p.q2sum * p.epsi_four_pi);
  const double lr = L * r;
  const double y = p.lq2sum == 0 ? t * std::sqrt(2.0 * (lr * lr * lr))
                                 : (3.0 / 4.0) * t * std::sqrt((L * L * L) / r);
  double xold = 1.0e-1;
  for (int i = 1; i <= 40; ++i) {
    const double xnew = p.lq2sum == 0 ? (1.0 / r) * std::sqrt(-std::log(y * (xold * xold)))
                                      : (1.0 / r) * std::sqrt(-std::log(y / (xold * xold)));
    if (std::fabs(xold - xnew) / xnew < 1.0e-6) {
      alpha = xnew;
      return true;
    }
    xold = xnew;
  }
  return false;
}

bool ncut_to_alpha(const mfe_ewald_opt &p, double &alpha) {
  const double L = p.boxlenshort, t = p.uewaldtol / (p.q2sum * p.epsi_four_pi);
  const int n = p.ncut;
  const double y = p.lq2sum == 0 ? t * (kPiE * kPiE * n) / L : (3.0 / 4.0) * t * L / n;
  // sqrt(real(ncut)): default (single) precision in the reference
  const double sn = (double)std::sqrt((float)n);
  double xold = 1.0;
  for (int i = 1; i <= 40; ++i) {
    const double g = 1.0 + 1.0 / (xold * L * sn);
    const double xnew = (kPiE * n / L) / std::sqrt(-std::log(y / (xold * xold) / g));
    if (std::fabs(xold - xnew) / xnew < 1.0e-6) {
      alpha = xnew;
      return true;
    }
    xold = xnew;
  }
  return false;
}

// EwaldErrorUReal (Src/potential.F90:2995-3018), leading order for dipoles
double ewald_error_ureal(const mfe_ewald_opt &p) {
  const double L = p.boxlenshort, r = p.rcut, fac = (p.ualpha * r) * (p.ualpha * r);
  if (p.lq2sum == 0)
    return p.epsi_four_pi * p.q2sum * std::sqrt(r / (2.0 * (L * L * L))) / fac * std::exp(-fac);
  const double rl = r * L;
  return p.epsi_four_pi * p.q2sum * 4.0 / std::sqrt(15.0 * (rl * rl * rl)) * 1.0 * fac *
         std::exp(-fac);
}

}  // namespace

extern "C" int mfe_ewald_parameters(mfe_ewald_opt *p) {
  if (!p || (p->lq2sum != 0 && p->lq2sum != 1) || !(p->q2sum > 0.0) ||
      !(p->boxlenshort > 0.0) || !(p->epsi_four_pi > 0.0))
    return MFE_E_BAD_PARAM;
  const int opt = p->iewaldopt;
  if (p->spme ? (opt != 0 && opt != 3) : (opt < 0 || opt > 4)) return MFE_E_BAD_PARAM;
  bool ok = true;
  switch (opt) {
    case 0:
      if (!(p->ualphared > 0.0) || !(p->rcut > 0.0) || (!p->spme && p->ncut <= 0))
        return MFE_E_BAD_PARAM;
      p->ualpha = p->ualphared / p->rcut;
      break;
    case 1:
      if (!(p->ualphared > 0.0) || !(p->rcut > 0.0)) return MFE_E_BAD_PARAM;
      p->ualpha = p->ualphared / p->rcut;
      p->uewaldtol = ewald_error_ureal(*p);
      ok = alpha_to_ncut(*p, p->ncut);
      break;
    case 2:
      if (!(p->uewaldtol > 0.0) || !(p->ualpha > 0.0)) return MFE_E_BAD_PARAM;
      ok = alpha_to_rcut(*p, p->rcut) && alpha_to_ncut(*p, p->ncut);
      p->ualphared = p->ualpha * p->rcut;
      break;
    case 3:
      if (!(p->uewaldtol > 0.0) || !(p->rcut > 0.0)) return MFE_E_BAD_PARAM;
      ok = rcut_to_alpha(*p, p->ualpha) && (p->spme || alpha_to_ncut(*p, p->ncut));
      p->ualphared = p->ualpha * p->rcut;
      break;
    case 4:
      if (!(p->uewaldtol > 0.0) || p->ncut <= 0) return MFE_E_BAD_PARAM;
      ok = ncut_to_alpha(*p, p->ualpha) && alpha_to_rcut(*p, p->rcut);
      p->ualphared = p->ualpha * p->rcut;
      break;
  }
  return ok ? MFE_OK : MFE_E_BAD_PARAM;
}
