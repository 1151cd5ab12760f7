// mfe_engine.hip -- the C ABI of include/mfe.h: engine state, setup of the
// device tables (the parts of EwaldSetup, IONList and InitCellList the
// kernels need) and the call sequences of UTotal, VelVer and BDStep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mfe_internal.h"

using mfe::HipError;

namespace {

constexpr double kPi = 3.1415926535897932;

int fail(mfe_engine *e, int code, const std::string &msg) {
  if (e) e->err = msg;
  return code;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

template <typename T>
void dalloc(T *&p, size_t n) {
  dfree(p);
  if (n == 0) n = 1;
  MFE_HIP_TRY(hipMalloc((void **)&p, n * sizeof(T)));
}

template <typename T>
void upload(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) MFE_HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

// guarded entry: converts HIP failures and bad_alloc into status codes
template <typename F>
int guarded(mfe_engine *e, F &&f) {
  try {
    return f();
  } catch (const HipError &h) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)h.code,
                  hipGetErrorString(h.code), h.expr, h.file, h.line);
    return fail(e, MFE_E_HIP, buf);
  } catch (const std::bad_alloc &) {
    return fail(e, MFE_E_NOMEM, "host allocation failed");
  }
}

mfe::FftPlan make_plan(int n) {
  mfe::FftPlan p{};
  p.n = n;
  int m = n;
  while (m % 4 == 0 && p.nstages < 20) {
    p.radix[p.nstages++] = 4;
    m /= 4;
  }
  while (m % 2 == 0) {
    p.radix[p.nstages++] = 2;
    m /= 2;
  }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) {
      p.radix[p.nstages++] = f;
      m /= f;
    }
  }
  return p;
}

bool plan_ok(const mfe::FftPlan &p) {
  for (int i = 0; i < p.nstages; ++i)
    if (p.radix[i] > 32) return false;
  return true;
}

template <typename R>
void* make_twiddles(int n, hipStream_t s) {
  std::vector<mfe::cplx<R>> h(n);
  for (int e = 0; e < n; ++e) {
    const double a = -2.0 * kPi * (double)e / (double)n;
    h[e].x = (R)std::cos(a);
    h[e].y = (R)std::sin(a);
  }
  mfe::cplx<R> *d = nullptr;
  dalloc(d, (size_t)n);
  upload(d, h.data(), (size_t)n, s);
  MFE_HIP_TRY(hipStreamSynchronize(s));
  return d;
}

// CardinalBSpline at pos = 0 in FP64 (Src/mollib.F90:1629-1683)
void bspline0(int order, double *v) {
  double last[32];
  const double x = 0.0;
  v[0] = x;
  v[1] = 1.0 - x;
  for (int n = 3; n <= order; ++n) {
    for (int i = 0; i < order; ++i) last[i] = v[i];
    v[0] = x * last[0] / (n - 1);
    for (int i = 1; i <= n - 2; ++i) {
      const double a = i + x;
      v[i] = (a * last[i] + (n - a) * last[i - 1]) / (n - 1);
    }
    v[n - 1] = (1 - x) * last[n - 2] / (n - 1);
  }
}

int smallest_divisor_at_least(int s, int lo) {
  for (int d = std::max(lo, 1); d <= s; ++d)
    if (s % d == 0) return d;
  return s;
}

void free_all(mfe_engine *e) {
  if (e->md_graph) (void)hipGraphExecDestroy(e->md_graph);
  e->md_graph = nullptr;
  dfree(e->ro); dfree(e->rod); dfree(e->rodd); dfree(e->force); dfree(e->forceo);
  dfree(e->drosum); dfree(e->r2disp); dfree(e->type); dfree(e->q); dfree(e->massip);
  dfree(e->t_zlow); dfree(e->t_coef); dfree(e->t_coef_f); dfree(e->t_off);
  dfree(e->t_nint); dfree(e->t_pairidx); dfree(e->t_r2atat); dfree(e->t_r2umin);
  dfree(e->c_count); dfree(e->c_start); dfree(e->c_of); dfree(e->c_slot);
  dfree(e->c_sorted); dfree(e->nn); dfree(e->nbr); dfree(e->flags);
  dfree(e->ro_s); dfree(e->type_s);
  dfree(e->mesh_r); dfree(e->mesh_c); dfree(e->b_count); dfree(e->b_start);
  dfree(e->b_id); dfree(e->b_slot); dfree(e->rec_frac); dfree(e->rec_meta);
  dfree(e->ax_g); dfree(e->ax_kk);
  for (auto &t : e->tw) dfree(t);
  dfree(e->tw_half);
  dfree(e->k_vec); dfree(e->k_sum); dfree(e->ch_prev); dfree(e->ch_next);
  dfree(e->part); dfree(e->en);
  if (e->h_en) (void)hipHostFree(e->h_en);
  if (e->h_flags) (void)hipHostFree(e->h_flags);
  e->h_en = nullptr;
  e->h_flags = nullptr;
}

// partial-sum regions of the reductions (see k_reduce)
void layout_partials(mfe_engine *e) {
  const int nb = (e->np + 255) / 256;
  e->nblk_pair = (e->np * mfe::kPairLanes + 255) / 256;
  e->nblk_bond = nb;
  e->nblk_surf = nb;
  e->nblk_move = nb;
  int conv = 0;
  if (e->ewald == MFE_EWALD_SPME && e->hx > 0) conv = mfe::conv_blocks(e);
  int off = 0;
  e->off_pair = off; off += e->nblk_pair * (e->nptpt + 1);
  e->off_conv = off; off += 2 * conv;
  e->off_bond = off; off += 2 * nb;
  e->off_surf = off; off += 3 * nb;
  e->off_std = off; off += 2 * e->nkvec;
  e->off_move = off; off += nb;
  e->off_skin = off; off += nb;
  if (off > e->part_cap) {
    dalloc(e->part, (size_t)off);
    e->part_cap = off;
  }
  mfe::launch_fill_double(e->part, e->part_cap, 0.0, e->stream);
}

// Neighbour-list cell grid: cells of edge >= rcut + drnlist; an axis with
// fewer than three such cells is not divided.
void set_cell_geom(mfe_engine *e) {
  const double rs = e->rcut + e->drnlist;
  mfe::CellGeom g{};
  for (int m = 0; m < 3; ++m) {
    int n = (int)std::floor(e->box.L[m] / rs);
    if (n < 3) n = 1;
    g.n[m] = n;
    g.cinv[m] = n / e->box.L[m];
  }
  g.ncells = g.n[0] * g.n[1] * g.n[2];
  if (g.ncells > e->ncells_alloc) {
    dalloc(e->c_count, (size_t)g.ncells);
    dalloc(e->c_start, (size_t)g.ncells + 1);
    e->ncells_alloc = g.ncells;
  }
  mfe::launch_fill_int(e->c_count, g.ncells, 0, e->stream);
  e->cg = g;
}

int auto_maxnn(const mfe_engine *e) {
  const double rs = e->rcut + e->drnlist;
  const double vol = e->box.L[0] * e->box.L[1] * e->box.L[2];
  const double mean = 4.0 / 3.0 * kPi * rs * rs * rs * e->np / vol;
  int m = (int)(2.0 * mean + 64.0);
  return std::min(m, std::max(e->np - 1, 1));
}

int prepare_neighbor(mfe_engine *e) {
  if (!e->have_neighbor || !e->have_tables || !e->have_box || !e->have_particles)
    return fail(e, MFE_E_STATE, "neighbour list needs box, particles, tables and mfe_set_neighbor");
  set_cell_geom(e);
  return MFE_OK;
}

int check_ready(mfe_engine *e) {
  if (!e->have_box) return fail(e, MFE_E_STATE, "mfe_set_box not called");
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles not called");
  if (!e->have_tables) return fail(e, MFE_E_STATE, "mfe_set_tables not called");
  if (!e->have_neighbor) return fail(e, MFE_E_STATE, "mfe_set_neighbor not called");
  if (!e->have_positions) return fail(e, MFE_E_STATE, "positions not uploaded");
  return MFE_OK;
}

int read_flags(mfe_engine *e) {
  MFE_HIP_TRY(hipMemcpyAsync(e->h_flags, e->flags, 3 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  MFE_HIP_TRY(hipStreamSynchronize(e->stream));
  const int err = e->h_flags[0];
  if (err) {
    mfe::launch_fill_int(e->flags, 1, 0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    if (err & mfe::ERR_CAPACITY) {
      e->list_valid = false;
      return fail(e, MFE_E_NEIGH_CAPACITY, "nneighpn > maxnneigh: raise maxnneigh");
    }
    if (err & mfe::ERR_OVERLAP) return fail(e, MFE_E_OVERLAP_R2UMIN, "r2 < r2umin(iptjpt)");
    if (err & mfe::ERR_TABLE) return fail(e, MFE_E_TABLE_RANGE, "ibuf > nbuf");
  }
  return MFE_OK;
}

// UTotal's device sequence; mode 0 = forces only, 1 = with VelVer kick;
// skin: also take the NList skin decision (MD/BD steps)
void launch_utotal(mfe_engine *e, unsigned flags, int mode, double dt, bool skin = false) {
  const bool energy = (flags & (MFE_ENERGY | MFE_VIRIAL)) != 0;
  mfe::launch_pair(e, energy);
  if (e->ewald == MFE_EWALD_SPME) mfe::launch_spme(e, energy, SP_BIN, SP_GATHER);
  else if (e->ewald == MFE_EWALD_STD) mfe::launch_std_ewald(e);
  if (e->nchain > 0) mfe::launch_bonds(e);
  if (e->lsurf && e->ewald != MFE_EWALD_NONE) mfe::launch_surf(e);
  mfe::launch_reduce(e, flags, skin);
  mfe::launch_finalize(e, mode, dt);
}

void launch_md_step(mfe_engine *e, double dt, unsigned flags) {
  mfe::launch_move(e, dt, 0, nullptr, nullptr, nullptr);
  mfe::launch_skin(e);
  launch_utotal(e, flags, 1, dt, true);
  mfe::launch_nlist_rebuild(e, true);
}

void energy_out(mfe_engine *e, mfe_energy *en) {
  std::memset(en, 0, sizeof(*en));
  en->tot = e->h_en[0];
  en->twob0 = e->h_en[1];
  en->rec = e->h_en[2];
  en->bond = e->h_en[3];
  en->virial = e->h_en[4];
  en->nptpt = e->nptpt;
  en->rebuilds = e->h_flags[2];
  for (int k = 0; k < e->nptpt; ++k) en->twob[k] = e->h_en[8 + k];
}

void fetch_energy(mfe_engine *e) {
  MFE_HIP_TRY(hipMemcpyAsync(e->h_en, e->en, (8 + MFE_MAX_PAIRS) * sizeof(double),
                             hipMemcpyDeviceToHost, e->stream));
}

}  // namespace

void mfe_allow_big_spread_lds(int prec, int bytes);

// ======================================================================= ABI
extern "C" {

int mfe_create(const mfe_config *cfg, mfe_engine **out) {
  if (!cfg || !out) return MFE_E_BAD_PARAM;
  *out = nullptr;
  if (cfg->max_np <= 0 || (cfg->precision != MFE_PREC_FP64 && cfg->precision != MFE_PREC_MIXED))
    return MFE_E_BAD_PARAM;
  mfe_engine *e = new (std::nothrow) mfe_engine();
  if (!e) return MFE_E_NOMEM;
  int rc = guarded(e, [&]() -> int {
    e->device = cfg->device;
    e->prec = cfg->precision;
    e->max_np = cfg->max_np;
    MFE_HIP_TRY(hipSetDevice(e->device));
    MFE_HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    dalloc(e->flags, 4);
    dalloc(e->en, 8 + MFE_MAX_PAIRS);
    MFE_HIP_TRY(hipHostMalloc((void **)&e->h_en, (8 + MFE_MAX_PAIRS) * sizeof(double)));
    MFE_HIP_TRY(hipHostMalloc((void **)&e->h_flags, 4 * sizeof(int)));
    mfe::launch_fill_int(e->flags, 4, 0, e->stream);
    mfe::launch_fill_double(e->en, 8 + MFE_MAX_PAIRS, 0.0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    return MFE_OK;
  });
  if (rc != MFE_OK) {
    free_all(e);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return rc;
  }
  *out = e;
  return MFE_OK;
}

int mfe_destroy(mfe_engine *e) {
  if (!e) return MFE_OK;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  free_all(e);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return MFE_OK;
}

const char *mfe_last_error(const mfe_engine *e) { return e ? e->err.c_str() : "null engine"; }

int mfe_set_box(mfe_engine *e, const double boxlen[3]) {
  if (!e || !boxlen) return MFE_E_BAD_PARAM;
  for (int m = 0; m < 3; ++m)
    if (!(boxlen[m] > 0.0)) return fail(e, MFE_E_BAD_PARAM, "boxlen must be positive");
  for (int m = 0; m < 3; ++m) {
    e->box.L[m] = boxlen[m];
    e->box.L2[m] = boxlen[m] / 2.0;
    e->box.Li[m] = 1.0 / boxlen[m];
  }
  e->have_box = true;
  e->list_valid = false;
  e->graph_dirty = true;
  return MFE_OK;
}

int mfe_set_particles(mfe_engine *e, int32_t np, int32_t npt, const int32_t *iptpn,
                      const double *az, const double *massipt) {
  if (!e || !iptpn || !az || !massipt) return MFE_E_BAD_PARAM;
  if (np < 1 || np > e->max_np) return fail(e, MFE_E_BAD_PARAM, "np out of range (0, max_np]");
  if (npt < 1 || npt > MFE_MAX_TYPES) return fail(e, MFE_E_BAD_PARAM, "npt out of range");
  if (np > (1 << 30)) return fail(e, MFE_E_BAD_PARAM, "np too large");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    std::vector<int> t(np);
    std::vector<double> mi(np);
    for (int i = 0; i < np; ++i) {
      if (iptpn[i] < 1 || iptpn[i] > npt) return fail(e, MFE_E_BAD_PARAM, "iptpn out of range");
      t[i] = iptpn[i] - 1;
      mi[i] = massipt[t[i]];
    }
    if (np != e->np) {
      dalloc(e->ro, 3 * (size_t)np); dalloc(e->rod, 3 * (size_t)np);
      dalloc(e->rodd, 3 * (size_t)np); dalloc(e->force, 3 * (size_t)np);
      dalloc(e->forceo, 3 * (size_t)np); dalloc(e->drosum, 3 * (size_t)np);
      dalloc(e->r2disp, (size_t)np); dalloc(e->type, (size_t)np);
      dalloc(e->q, (size_t)np); dalloc(e->massip, (size_t)np);
      dalloc(e->c_of, (size_t)np); dalloc(e->c_slot, (size_t)np);
      dalloc(e->c_sorted, (size_t)np); dalloc(e->nn, (size_t)np);
      dalloc(e->ro_s, 3 * (size_t)np); dalloc(e->type_s, (size_t)np);
      dalloc(e->b_id, (size_t)np); dalloc(e->b_slot, (size_t)np);
      dalloc(e->rec_meta, (size_t)np);
      dfree(e->rec_frac);
      MFE_HIP_TRY(hipMalloc(&e->rec_frac, 4 * (size_t)np * sizeof(double)));
      dfree(e->nbr);
      e->maxnn = 0;
      e->have_positions = false;
      e->have_neighbor = false;
    }
    e->np = np;
    e->npt = npt;
    e->nptpt = npt * (npt + 1) / 2;
    e->h_type = t;
    e->h_q.assign(az, az + np);
    e->h_massipt.assign(massipt, massipt + npt);
    upload(e->type, t.data(), (size_t)np, e->stream);
    {
      std::vector<int> ident(np);
      for (int i = 0; i < np; ++i) ident[i] = i;
      upload(e->c_sorted, ident.data(), (size_t)np, e->stream);
    }
    upload(e->q, az, (size_t)np, e->stream);
    upload(e->massip, mi.data(), (size_t)np, e->stream);
    mfe::launch_fill_double(e->rod, 3 * np, 0.0, e->stream);
    mfe::launch_fill_double(e->rodd, 3 * np, 0.0, e->stream);
    mfe::launch_fill_double(e->drosum, 3 * np, 0.0, e->stream);
    double q2 = 0.0;
    for (int i = 0; i < np; ++i) q2 += az[i] * az[i];
    e->self_energy = -(e->alpha / std::sqrt(kPi)) * q2;
    e->have_particles = true;
    e->list_valid = false;
    e->graph_dirty = true;
    layout_partials(e);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    return MFE_OK;
  });
}

int mfe_set_tables(mfe_engine *e, const double *ubuf, int32_t nbuf, const int32_t *iubuflow,
                   const double *r2atat, const double *r2umin, double rcut) {
  if (!e || !ubuf || !iubuflow || !r2atat || !r2umin || nbuf < 12)
    return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  if (!(rcut > 0.0)) return fail(e, MFE_E_BAD_PARAM, "rcut must be positive");
  return guarded(e, [&]() -> int {
    const int P = e->nptpt;
    // the intervals of pair p run from iubuflow(p) up to the next table start
    std::vector<int> starts(iubuflow, iubuflow + P);
    std::vector<int> off(P), nint(P);
    std::vector<double> zl, cf;
    for (int p = 0; p < P; ++p) {
      const int s0 = starts[p];
      if (s0 < 1 || s0 + 11 > nbuf) return fail(e, MFE_E_BAD_PARAM, "iubuflow out of range");
      int end = nbuf + 1;
      for (int q = 0; q < P; ++q)
        if (starts[q] > s0) end = std::min(end, starts[q]);
      int n = (end - s0) / 12;
      // keep intervals down to the one whose zlow is r2umin (trailing
      // zero-filled slots of MOLSIM's allocation are not intervals)
      int used = 0;
      for (int k = 0; k < n; ++k) {
        ++used;
        if (ubuf[s0 - 1 + 12 * k] <= r2umin[p]) break;
      }
      off[p] = (int)zl.size();
      nint[p] = used;
      for (int k = 0; k < used; ++k) {
        const double *u = &ubuf[s0 - 1 + 12 * k];
        zl.push_back(u[0]);
        for (int c = 1; c < 12; ++c) cf.push_back(u[c]);
      }
    }
    std::vector<float> cff(cf.begin(), cf.end());
    std::vector<int> pidx(e->npt * e->npt);
    int k = 0;
    for (int i = 0; i < e->npt; ++i)
      for (int j = i; j < e->npt; ++j, ++k) pidx[i * e->npt + j] = pidx[j * e->npt + i] = k;
    dalloc(e->t_zlow, zl.size()); dalloc(e->t_coef, cf.size()); dalloc(e->t_coef_f, cf.size());
    dalloc(e->t_off, (size_t)P); dalloc(e->t_nint, (size_t)P); dalloc(e->t_pairidx, pidx.size());
    dalloc(e->t_r2atat, (size_t)P); dalloc(e->t_r2umin, (size_t)P);
    upload(e->t_zlow, zl.data(), zl.size(), e->stream);
    upload(e->t_coef, cf.data(), cf.size(), e->stream);
    upload(e->t_coef_f, cff.data(), cff.size(), e->stream);
    upload(e->t_off, off.data(), (size_t)P, e->stream);
    upload(e->t_nint, nint.data(), (size_t)P, e->stream);
    upload(e->t_pairidx, pidx.data(), pidx.size(), e->stream);
    upload(e->t_r2atat, r2atat, (size_t)P, e->stream);
    upload(e->t_r2umin, r2umin, (size_t)P, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->rcut = rcut;
    e->have_tables = true;
    e->list_valid = false;
    e->graph_dirty = true;
    return MFE_OK;
  });
}

int mfe_set_ewald(mfe_engine *e, int32_t kind, double alpha, int32_t order,
                  const int32_t mesh[3], int32_t ncut, int32_t ncut_cube, int32_t lsurf,
                  double epsi_four_pi) {
  if (!e) return MFE_E_BAD_PARAM;
  if (!e->have_box || !e->have_particles)
    return fail(e, MFE_E_STATE, "mfe_set_box and mfe_set_particles first");
  if (kind != MFE_EWALD_NONE && kind != MFE_EWALD_STD && kind != MFE_EWALD_SPME)
    return fail(e, MFE_E_BAD_PARAM, "unknown Ewald kind");
  if (kind != MFE_EWALD_NONE && !(alpha > 0.0)) return fail(e, MFE_E_BAD_PARAM, "ualpha <= 0");
  return guarded(e, [&]() -> int {
    e->ewald = kind;
    e->alpha = alpha;
    e->efp = epsi_four_pi;
    e->lsurf = lsurf ? 1 : 0;
    double q2 = 0.0;
    for (double v : e->h_q) q2 += v * v;
    e->self_energy = -(alpha / std::sqrt(kPi)) * q2;
    e->nkvec = 0;
    if (kind == MFE_EWALD_SPME) {
      if (!mesh || order < 3 || order > 8)
        return fail(e, MFE_E_BAD_PARAM, "SPME order must be 3..8");
      for (int m = 0; m < 3; ++m) {
        if (mesh[m] <= 0) return fail(e, MFE_E_BAD_PARAM, "nmesh <= 0");
        if (mesh[m] % 2) return fail(e, MFE_E_MESH_ODD, "nmesh must be even");
        if (mesh[m] > 1024) return fail(e, MFE_E_BAD_PARAM, "nmesh > 1024");
        if (mesh[m] < order) return fail(e, MFE_E_BAD_PARAM, "nmesh < order");
      }
      e->order = order;
      for (int m = 0; m < 3; ++m) e->s[m] = mesh[m];
      e->hx = e->s[0] / 2 + 1;
      mfe::FftPlan ph = make_plan(e->s[0] / 2);
      if (!plan_ok(ph)) return fail(e, MFE_E_BAD_PARAM, "mesh has a prime factor > 32");
      e->plan_half = ph;
      for (int m = 0; m < 3; ++m) {
        e->plan[m] = make_plan(e->s[m]);
        if (!plan_ok(e->plan[m])) return fail(e, MFE_E_BAD_PARAM, "mesh has a prime factor > 32");
      }
      // bricks owned by the spreading blocks: B >= order - 1 divides s
      size_t tile = 1;
      for (int m = 0; m < 3; ++m) {
        int b = smallest_divisor_at_least(e->s[m], std::max(order - 1, 8));
        if (b > 16) b = smallest_divisor_at_least(e->s[m], order - 1);
        e->B[m] = b;
        e->nb[m] = e->s[m] / b;
        tile *= (size_t)b;
      }
      e->nbricks = e->nb[0] * e->nb[1] * e->nb[2];
      const size_t rsz = e->prec == MFE_PREC_MIXED ? 4 : 8;
      e->spread_lds = (int)(tile * 8);
      size_t gext = rsz;
      for (int m = 0; m < 3; ++m) gext *= (size_t)(e->B[m] + order - 1);
      const int lds_need = (int)std::max<size_t>((size_t)e->spread_lds, gext);
      if (lds_need > 150 * 1024)
        return fail(e, MFE_E_BAD_PARAM, "mesh brick too large for LDS (mesh 2q, q prime > 16?)");
      const size_t M = (size_t)e->s[0] * e->s[1] * e->s[2];
      const size_t Mc = (size_t)e->hx * e->s[1] * e->s[2];
      dfree(e->mesh_r);
      dfree(e->mesh_c);
      MFE_HIP_TRY(hipMalloc(&e->mesh_r, M * rsz));
      MFE_HIP_TRY(hipMalloc(&e->mesh_c, Mc * 2 * rsz));
      dalloc(e->b_count, (size_t)e->nbricks);
      dalloc(e->b_start, (size_t)e->nbricks + 1);
      mfe::launch_fill_int(e->b_count, e->nbricks, 0, e->stream);
      // per-axis influence tables (Essmann eqs. 3.9, 4.4; EwaldSetup
      // Src/energy.F90:6980-7018), b^2 with the reference's single-precision
      // exp(i 2 pi m k / s)
      const int smax2 = std::max(e->s[0], std::max(e->s[1], e->s[2])) / 2 + 1;
      std::vector<double> g(3 * smax2, 0.0), kk(3 * smax2, 0.0);
      double spl[16];
      bspline0(order, spl);
      const double expf_ = -1.0 / (4.0 * alpha * alpha);
      for (int a = 0; a < 3; ++a) {
        const double tpb = 2.0 * kPi * (1.0 / e->box.L[a]);
        for (int m = 0; m <= e->s[a] / 2; ++m) {
          double b2 = 0.0;
          if (!((order % 2 == 1) && (2 * m == e->s[a]))) {
            std::complex<double> acc(0.0, 0.0);
            for (int k = 0; k <= order - 2; ++k) {
              const float th = (float)(2.0 * kPi * m * k / e->s[a]);
              acc += spl[k + 1] * std::complex<double>((double)std::cos(th), (double)std::sin(th));
            }
            b2 = 1.0 / (acc.real() * acc.real() + acc.imag() * acc.imag());
          }
          const double k2 = (m * tpb) * (m * tpb);
          kk[a * smax2 + m] = k2;
          g[a * smax2 + m] = b2 * std::exp(k2 * expf_);
        }
      }
      dalloc(e->ax_g, g.size());
      dalloc(e->ax_kk, kk.size());
      upload(e->ax_g, g.data(), g.size(), e->stream);
      upload(e->ax_kk, kk.data(), kk.size(), e->stream);
      for (auto &t : e->tw) dfree(t);
      dfree(e->tw_half);
      for (int m = 0; m < 3; ++m)
        e->tw[m] = e->prec == MFE_PREC_MIXED ? make_twiddles<float>(e->s[m], e->stream)
                                             : make_twiddles<double>(e->s[m], e->stream);
      e->tw_half = e->prec == MFE_PREC_MIXED ? make_twiddles<float>(e->s[0] / 2, e->stream)
                                             : make_twiddles<double>(e->s[0] / 2, e->stream);
      if (lds_need > 64 * 1024) {
        // large bricks (meshes 2q with q prime) need the LDS above 64 KB
        mfe_allow_big_spread_lds(e->prec, lds_need);
      }
    } else if (kind == MFE_EWALD_STD) {
      if (ncut < 1) return fail(e, MFE_E_BAD_PARAM, "ncut < 1");
      e->ncut = ncut;
      e->ncut_cube = ncut_cube ? 1 : 0;
      const int ncut2 = ncut_cube ? 3 * ncut * ncut : ncut * ncut;
      const double vol = e->box.L[0] * e->box.L[1] * e->box.L[2];
      const double fac0 = 2.0 * kPi / vol;
      double tpb[3];
      for (int m = 0; m < 3; ++m) tpb[m] = 2.0 * kPi * (1.0 / e->box.L[m]);
      std::vector<double> kv;
      // kfac in the reference's order (Src/energy.F90:6757-6783)
      for (int nz = 0; nz <= ncut; ++nz) {
        const double facz = nz == 0 ? 1.0 : 2.0;
        for (int ny = 0; ny <= ncut; ++ny) {
          const double facy = ny == 0 ? 0.5 : 1.0;
          for (int nx = 0; nx <= ncut; ++nx) {
            if (nx * nx + ny * ny + nz * nz > ncut2) continue;
            if (nx == 0 && ny == 0 && nz == 0) continue;
            const double facx = nx == 0 ? 0.5 : 1.0;
            const double k2 = (nx * tpb[0]) * (nx * tpb[0]) + (ny * tpb[1]) * (ny * tpb[1]) +
                              (nz * tpb[2]) * (nz * tpb[2]);
            kv.push_back(nx * tpb[0]);
            kv.push_back(ny * tpb[1]);
            kv.push_back(nz * tpb[2]);
            kv.push_back(fac0 * facz * facy * facx * std::exp(-k2 / (4.0 * alpha * alpha)) / k2);
          }
        }
      }
      e->nkvec = (int)(kv.size() / 4);
      dalloc(e->k_vec, kv.size());
      dalloc(e->k_sum, 8 * (size_t)e->nkvec);
      upload(e->k_vec, kv.data(), kv.size(), e->stream);
    }
    layout_partials(e);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->have_ewald = true;
    e->graph_dirty = true;
    return MFE_OK;
  });
}

int mfe_set_neighbor(mfe_engine *e, int32_t kind, double drnlist, int32_t maxnneigh) {
  if (!e) return MFE_E_BAD_PARAM;
  if (kind != MFE_NL_VLIST && kind != MFE_NL_CELL) return fail(e, MFE_E_BAD_PARAM, "unknown list kind");
  if (!(drnlist >= 0.0)) return fail(e, MFE_E_BAD_PARAM, "drnlist < 0");
  if (!e->have_box || !e->have_particles || !e->have_tables)
    return fail(e, MFE_E_STATE, "box, particles and tables first");
  return guarded(e, [&]() -> int {
    e->nlkind = kind;
    e->drnlist = drnlist;
    e->have_neighbor = true;
    const int m = maxnneigh > 0 ? maxnneigh : auto_maxnn(e);
    if (m != e->maxnn || !e->nbr) {
      dalloc(e->nbr, (size_t)m * e->np);
      e->maxnn = m;
    }
    int rc = prepare_neighbor(e);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->list_valid = false;
    e->graph_dirty = true;
    return rc;
  });
}

int mfe_set_chains(mfe_engine *e, int32_t nchain, const int32_t *chain_first,
                   const int32_t *chain_len, double k, int32_t p, double eq) {
  if (!e) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  if (nchain < 0 || (nchain > 0 && (!chain_first || !chain_len)) || p < 1)
    return fail(e, MFE_E_BAD_PARAM, "bad chain description");
  return guarded(e, [&]() -> int {
    std::vector<int> prev(e->np, -1), next(e->np, -1);
    for (int c = 0; c < nchain; ++c) {
      const int f = chain_first[c] - 1, n = chain_len[c];
      if (f < 0 || n < 1 || f + n > e->np) return fail(e, MFE_E_BAD_PARAM, "chain out of range");
      for (int s = 0; s + 1 < n; ++s) {
        next[f + s] = f + s + 1;
        prev[f + s + 1] = f + s;
      }
    }
    dalloc(e->ch_prev, (size_t)e->np);
    dalloc(e->ch_next, (size_t)e->np);
    upload(e->ch_prev, prev.data(), (size_t)e->np, e->stream);
    upload(e->ch_next, next.data(), (size_t)e->np, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->nchain = nchain;
    e->bond_k = k / p;  // IOPotChain, Src/potential.F90:3561
    e->bond_p = p;
    e->bond_eq = eq;
    e->graph_dirty = true;
    return MFE_OK;
  });
}

int mfe_upload_positions(mfe_engine *e, const double *ro) {
  if (!e || !ro) return MFE_E_BAD_PARAM;
  if (!e->have_particles || !e->have_box) return fail(e, MFE_E_STATE, "box and particles first");
  for (int i = 0; i < 3 * e->np; ++i) {
    const int m = i % 3;
    if (!(std::fabs(ro[i]) <= e->box.L2[m]))
      return fail(e, MFE_E_BAD_PARAM, "particle outside the box [-L/2, L/2]");
  }
  return guarded(e, [&]() -> int {
    upload(e->ro, ro, 3 * (size_t)e->np, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->have_positions = true;
    e->list_valid = false;
    return MFE_OK;
  });
}

int mfe_upload_velocities(mfe_engine *e, const double *rod, const double *rodd) {
  if (!e) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  return guarded(e, [&]() -> int {
    const int n = 3 * e->np;
    if (rod) upload(e->rod, rod, (size_t)n, e->stream); else mfe::launch_fill_double(e->rod, n, 0.0, e->stream);
    if (rodd) upload(e->rodd, rodd, (size_t)n, e->stream); else mfe::launch_fill_double(e->rodd, n, 0.0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    return MFE_OK;
  });
}

int mfe_download_state(mfe_engine *e, double *ro, double *rod, double *rodd) {
  if (!e) return MFE_E_BAD_PARAM;
  return guarded(e, [&]() -> int {
    const size_t n = 3 * (size_t)e->np * sizeof(double);
    if (ro) MFE_HIP_TRY(hipMemcpyAsync(ro, e->ro, n, hipMemcpyDeviceToHost, e->stream));
    if (rod) MFE_HIP_TRY(hipMemcpyAsync(rod, e->rod, n, hipMemcpyDeviceToHost, e->stream));
    if (rodd) MFE_HIP_TRY(hipMemcpyAsync(rodd, e->rodd, n, hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    return MFE_OK;
  });
}

int mfe_build_list(mfe_engine *e) {
  if (!e) return MFE_E_BAD_PARAM;
  int rc = check_ready(e);
  if (rc) return rc;
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    mfe::launch_nlist_rebuild(e, false);
    int r = read_flags(e);
    e->list_valid = (r == MFE_OK);
    return r;
  });
}

int mfe_compute(mfe_engine *e, uint32_t flags, double *forceo, mfe_energy *energy) {
  if (!e) return MFE_E_BAD_PARAM;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->have_ewald) return fail(e, MFE_E_STATE, "mfe_set_ewald not called");
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    unsigned f = flags;
    if (energy) f |= MFE_ENERGY | MFE_VIRIAL;
    launch_utotal(e, f, 0, 0.0);
    if (forceo)
      MFE_HIP_TRY(hipMemcpyAsync(forceo, e->forceo, 3 * (size_t)e->np * sizeof(double),
                                 hipMemcpyDeviceToHost, e->stream));
    fetch_energy(e);
    int r = read_flags(e);
    if (r) return r;
    if (energy) energy_out(e, energy);
    return MFE_OK;
  });
}

int mfe_step_md(mfe_engine *e, int32_t nsteps, double tstep, uint32_t flags) {
  if (!e || nsteps < 0) return MFE_E_BAD_PARAM;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->have_ewald) return fail(e, MFE_E_STATE, "mfe_set_ewald not called");
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    // one step is captured once into a graph and replayed; the rebuild
    // kernels read the device-side skin-test flag, so the graph is static
    if (e->graph_dirty || !e->md_graph || e->md_graph_dt != tstep || e->md_graph_flags != flags) {
      if (e->md_graph) (void)hipGraphExecDestroy(e->md_graph);
      e->md_graph = nullptr;
      hipGraph_t g;
      MFE_HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      launch_md_step(e, tstep, flags);
      MFE_HIP_TRY(hipStreamEndCapture(e->stream, &g));
      hipError_t ie = hipGraphInstantiate(&e->md_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      MFE_HIP_TRY(ie);
      e->md_graph_dt = tstep;
      e->md_graph_flags = flags;
      e->graph_dirty = false;
    }
    for (int s = 0; s < nsteps; ++s) MFE_HIP_TRY(hipGraphLaunch(e->md_graph, e->stream));
    fetch_energy(e);
    return read_flags(e);
  });
}

int mfe_step_bd(mfe_engine *e, int32_t nsteps, const double *fac1, const double *fac2,
                const double *noise, uint32_t flags) {
  if (!e || nsteps < 0 || !fac1 || !fac2 || !noise) return MFE_E_BAD_PARAM;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->have_ewald) return fail(e, MFE_E_STATE, "mfe_set_ewald not called");
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    double *d_f1 = nullptr, *d_f2 = nullptr, *d_noise = nullptr;
    dalloc(d_f1, (size_t)e->npt);
    dalloc(d_f2, (size_t)e->npt);
    dalloc(d_noise, 3 * (size_t)e->np);
    upload(d_f1, fac1, (size_t)e->npt, e->stream);
    upload(d_f2, fac2, (size_t)e->npt, e->stream);
    int r = MFE_OK;
    for (int s = 0; s < nsteps && r == MFE_OK; ++s) {
      upload(d_noise, noise + 3 * (size_t)e->np * s, 3 * (size_t)e->np, e->stream);
      mfe::launch_move(e, 0.0, 1, d_f1, d_f2, d_noise);
      mfe::launch_skin(e);
      launch_utotal(e, flags, 0, 0.0, true);
      mfe::launch_nlist_rebuild(e, true);
    }
    fetch_energy(e);
    r = read_flags(e);
    dfree(d_f1);
    dfree(d_f2);
    dfree(d_noise);
    return r;
  });
}

int mfe_get_energy(mfe_engine *e, mfe_energy *energy) {
  if (!e || !energy) return MFE_E_BAD_PARAM;
  return guarded(e, [&]() -> int {
    fetch_energy(e);
    MFE_HIP_TRY(hipMemcpyAsync(e->h_flags, e->flags, 3 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    energy_out(e, energy);
    return MFE_OK;
  });
}

int mfe_get_pairs(mfe_engine *e, int32_t *nneighpn, int32_t *jpnlist, int32_t maxnneigh) {
  if (!e || !nneighpn || !jpnlist || maxnneigh < 1) return MFE_E_BAD_PARAM;
  if (!e->list_valid) return fail(e, MFE_E_STATE, "no neighbour list built");
  return guarded(e, [&]() -> int {
    const int np = e->np;
    std::vector<int> nn(np), nbr((size_t)e->maxnn * np), sorted(np);
    MFE_HIP_TRY(hipMemcpyAsync(nn.data(), e->nn, np * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipMemcpyAsync(nbr.data(), e->nbr, nbr.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipMemcpyAsync(sorted.data(), e->c_sorted, np * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<std::vector<int>> rows(np);
    for (int t = 0; t < np; ++t) {
      const int ip = sorted[t] + 1;
      auto &row = rows[ip - 1];
      for (int k = 0; k < nn[t]; ++k) {
        const int jp = sorted[nbr[(size_t)k * np + t]] + 1;
        const int par = (ip + jp) % 2;
        // checkerboard ownership of SetVListMD (Src/nlist.F90:742-745)
        if (ip < jp && par == 0) continue;
        if (ip > jp && par != 0) continue;
        row.push_back(jp);
      }
    }
    for (int i = 0; i < np; ++i) {
      auto &row = rows[i];
      std::sort(row.begin(), row.end());
      if ((int)row.size() > maxnneigh) return fail(e, MFE_E_NEIGH_CAPACITY, "jpnlist capacity too small");
      nneighpn[i] = (int)row.size();
      std::copy(row.begin(), row.end(), jpnlist + (size_t)i * maxnneigh);
    }
    return MFE_OK;
  });
}

int mfe_time_stage(mfe_engine *e, int32_t stage, int32_t nrep, double *ms) {
  if (!e || !ms || nrep < 1 || stage < 0 || stage >= MFE_NSTAGES) return MFE_E_BAD_PARAM;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  const bool spme = e->ewald == MFE_EWALD_SPME;
  if (!spme && stage != MFE_STAGE_PAIR && stage != MFE_STAGE_STEP && stage != MFE_STAGE_NLIST)
    return fail(e, MFE_E_STATE, "SPME stage requested without SPME");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    // one full evaluation first so every stage has valid inputs
    launch_utotal(e, MFE_ENERGY, 0, 0.0);
    hipEvent_t a, b;
    MFE_HIP_TRY(hipEventCreate(&a));
    MFE_HIP_TRY(hipEventCreate(&b));
    MFE_HIP_TRY(hipEventRecord(a, e->stream));
    for (int r = 0; r < nrep; ++r) {
      switch (stage) {
        case MFE_STAGE_PAIR: mfe::launch_pair(e, true); break;
        case MFE_STAGE_SPREAD: mfe::launch_spme(e, true, SP_BIN, SP_SPREAD); break;
        case MFE_STAGE_FFT_X: mfe::launch_spme(e, true, SP_FFTX, SP_FFTX); break;
        case MFE_STAGE_FFT_Y: mfe::launch_spme(e, true, SP_FFTY, SP_FFTY); break;
        case MFE_STAGE_CONV: mfe::launch_spme(e, true, SP_CONV, SP_CONV); break;
        case MFE_STAGE_IFFT: mfe::launch_spme(e, true, SP_IFFTY, SP_IFFTX); break;
        case MFE_STAGE_GATHER: mfe::launch_spme(e, true, SP_GATHER, SP_GATHER); break;
        case MFE_STAGE_NLIST: mfe::launch_nlist_rebuild(e, false); break;
        case MFE_STAGE_STEP: launch_utotal(e, MFE_ENERGY, 0, 0.0); break;
      }
    }
    MFE_HIP_TRY(hipEventRecord(b, e->stream));
    MFE_HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    MFE_HIP_TRY(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)t / nrep;
    // leave a consistent state (the stage runs may have transformed meshes)
    launch_utotal(e, MFE_ENERGY, 0, 0.0);
    return read_flags(e);
  });
}

}  // extern "C"
