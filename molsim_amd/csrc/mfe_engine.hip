// mfe_engine.hip -- the C ABI of include/mfe.h: engine state, setup of the
// device tables (the parts of EwaldSetup, IONList and InitCellList the
// kernels need) and the call sequences of UTotal, VelVer and BDStep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
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
  while (m % 8 == 0 && p.nstages < 20) {
    p.radix[p.nstages++] = 8;
    m /= 8;
  }
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
  std::vector<mfe::Twiddle<R>> h(n);
  for (int e = 0; e < n; ++e) {
    // the angle reduced to the first octant's accuracy: cos/sin of
    // 2 pi e / n in long double
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)e /
                          (long double)n;
    const long double c = std::cos(a), sn = std::sin(a);
    h[e].x = (R)c;
    h[e].y = (R)sn;
    if constexpr (sizeof(R) == 4) {
      h[e].lx = (float)(c - (long double)h[e].x);
      h[e].ly = (float)(sn - (long double)h[e].y);
    }
  }
  mfe::Twiddle<R> *d = nullptr;
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
  e->md_graph.clear();
  dfree(e->ro); dfree(e->rod); dfree(e->rodd); dfree(e->force); dfree(e->forceo);
  dfree(e->force_rec);
  dfree(e->drosum); dfree(e->r2disp); dfree(e->type); dfree(e->q); dfree(e->massip);
  dfree(e->t_zlow); dfree(e->t_coef); dfree(e->t_coef_f); dfree(e->t_off);
  dfree(e->t_nint); dfree(e->t_pairidx); dfree(e->t_r2atat); dfree(e->t_r2umin);
  dfree(e->t_bin_k); dfree(e->t_rec);
  dfree(e->c_count); dfree(e->c_start); dfree(e->c_of); dfree(e->c_slot);
  dfree(e->c_sorted); dfree(e->c_pos); dfree(e->nn); dfree(e->nbr); dfree(e->flags);
  dfree(e->rebuild_bar); dfree(e->rebuild_sum);
  dfree(e->ro_s); dfree(e->type_s); dfree(e->pos_f); dfree(e->t_rec_f); dfree(e->t_zlow_f);
  dfree(e->mesh_r); dfree(e->mesh_c); dfree(e->b_count); dfree(e->b_start); dfree(e->b_done);
  dfree(e->b_id); dfree(e->b_slot); dfree(e->rec_frac); dfree(e->rec_meta);
  dfree(e->spread_tiles);
  dfree(e->mc_ia); dfree(e->mc_rtm); dfree(e->mc_out);
  if (e->mc_spec) (void)hipFree(e->mc_spec);
  e->mc_spec = nullptr;
  e->mc_spec_bytes = 0;
  dfree(e->mc_fac);
  if (e->h_mc) (void)hipHostFree(e->h_mc);
  e->h_mc = nullptr;
  e->h_mc_bytes = 0;
  dfree(e->ax_g); dfree(e->ax_kk);
  for (auto &t : e->tw) dfree(t);
  dfree(e->tw_half);
  dfree(e->k_vec); dfree(e->k_sum); dfree(e->ch_prev); dfree(e->ch_next);
  dfree(e->part); dfree(e->en);
  dfree(e->bd_fac1); dfree(e->bd_fac2); dfree(e->bd_noise); dfree(e->d_sys);
  dfree(e->ro_stage);
  dfree(e->knot_c); dfree(e->phi_rec); dfree(e->adj_phi); dfree(e->adj_dq);
  dfree(e->adj_part); dfree(e->adj_out);
  dfree(e->fm_ro); dfree(e->fm_tgt); dfree(e->fm_fld); dfree(e->fm_res); dfree(e->fm_res_s);
  dfree(e->fm_vec); dfree(e->fm_part); dfree(e->fm_acc); dfree(e->fm_qr);
  if (e->h_en) (void)hipHostFree(e->h_en);
  if (e->h_flags) (void)hipHostFree(e->h_flags);
  e->h_en = nullptr;
  e->h_flags = nullptr;
}

// partial-sum regions of the reductions (see k_reduce)
void layout_partials(mfe_engine *e) {
  const int nb = (e->np + 255) / 256;
  e->nblk_pair = std::min(mfe::kPairGridMax,
                          (e->np * mfe::kPairLanes + mfe::kPairBlock - 1) / mfe::kPairBlock);
  e->nblk_bond = nb;
  e->nblk_surf = nb;
  e->nblk_move = nb;
  int conv = 0;
  if (e->ewald == MFE_EWALD_SPME && e->hx > 0) {
    int a, b, c, d, fz_T, fz_nbx;
    size_t l0, l1, l2;
    mfe::fft_geometry(e->prec, e->s, &a, &b, &c, &d, &fz_T, &fz_nbx, &l0, &l1, &l2);
    conv = fz_nbx * e->s[1];
  }
  e->nblk_conv = conv;
  int off = 0;
  e->off_pair = off; off += e->nblk_pair * (e->nptpt + 1);
  e->off_conv = off; off += 2 * conv;
  e->off_bond = off; off += 3 * nb;
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
  for (int m = 0; m < 3; ++m) {
    const int n = g.n[m];
    const double L = e->box.L[m];
    g.hf[m] = (double)(float)(L / n);
    // {-W, -hf, 0, hf, W}: W the offset across the periodic boundary
    const double W = n > 1 ? L - (n - 1) * g.hf[m] : L;
    g.off[m][0] = (float)(-W);
    g.off[m][1] = (float)(-g.hf[m]);
    g.off[m][2] = 0.0f;
    g.off[m][3] = (float)g.hf[m];
    g.off[m][4] = (float)W;
    g.L[m] = (float)L;
    g.L2[m] = (float)(0.5 * L);
  }
  if (g.ncells > e->ncells_alloc) {
    dalloc(e->c_count, (size_t)g.ncells);
    dalloc(e->c_start, (size_t)g.ncells + 1);
    e->ncells_alloc = g.ncells;
  }
  mfe::launch_fill_int(e->c_count, g.ncells, 0, e->stream);
  e->cg = g;
}

// expected Verlet-list length (partners within rcut + drnlist)
double mean_partners(const mfe_engine *e) {
  const double rs = e->rcut + e->drnlist;
  const double vol = e->box.L[0] * e->box.L[1] * e->box.L[2];
  return 4.0 / 3.0 * kPi * rs * rs * rs * e->np / vol;
}

// lanes per particle of the mixed pair kernel: a lane evaluates batches of
// four entries, so lists of up to ~12 keep one lane busy, up to ~48 two.
// Boxes with an axis of fewer than three list cells take four (their
// kernel variant forms separations in FP64).
bool pair_pos64_for(const mfe_engine *e) {
  return e->cg.n[0] < 3 || e->cg.n[1] < 3 || e->cg.n[2] < 3;
}
int pair_lanes_for(const mfe_engine *e) {
  if (pair_pos64_for(e)) return 4;
  const double m = mean_partners(e);
  return m <= 12.0 ? 1 : (m <= 48.0 ? 2 : 4);
}

int auto_maxnn(const mfe_engine *e) {
  int m = (int)(2.0 * mean_partners(e) + 64.0);
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
  return mfe_map_flag_error(e, e->h_flags[0]);
}

inline int nblocks256(int n) { return (n + 255) / 256; }

}  // namespace

// the engine's device view (one replica of the kernels' SysDev array)
void mfe_fill_view(const mfe_engine *e, mfe::SysDev &v) {
  v = mfe::SysDev{};
  v.np = e->np;
  v.npt = e->npt;
  v.nptpt = e->nptpt;
  v.nlsem = e->nlkind == MFE_NL_CELL ? 1 : 0;
  v.box = e->box;
  MFE_SETP(v.ro, e->ro); MFE_SETP(v.rod, e->rod); MFE_SETP(v.rodd, e->rodd); MFE_SETP(v.force, e->force); MFE_SETP(v.forceo, e->forceo);
  MFE_SETP(v.force_rec, e->force_rec);
  MFE_SETP(v.drosum, e->drosum); MFE_SETP(v.r2disp, e->r2disp); MFE_SETP(v.type, e->type); MFE_SETP(v.q, e->q);
  MFE_SETP(v.massip, e->massip);
  mfe::TableDev &t = v.tab;
  MFE_SETP(t.zlow, e->t_zlow); MFE_SETP(t.coef, e->t_coef); MFE_SETP(t.coef_f, e->t_coef_f); MFE_SETP(t.off, e->t_off);
  MFE_SETP(t.nint, e->t_nint); MFE_SETP(t.r2atat, e->t_r2atat); MFE_SETP(t.r2umin, e->t_r2umin); MFE_SETP(t.pairidx, e->t_pairidx);
  MFE_SETP(t.bin_k, e->t_bin_k); MFE_SETP(t.rec, e->t_rec);
  MFE_SETP(t.rec_f, e->t_rec_f); MFE_SETP(t.zlow_f, e->t_zlow_f);
  t.nbin = e->t_nbin;
  t.lds_mode = e->t_lds_mode;
  t.npt = e->npt; t.nptpt = e->nptpt; t.rcut2 = e->rcut * e->rcut; t.nzl = e->t_nzl;
  const double rs = e->rcut + e->drnlist;
  v.rs2 = rs * rs;
  v.drnlist = e->drnlist;
  v.maxnn = e->maxnn;
  v.cg = e->cg;
  MFE_SETP(v.c_count, e->c_count); MFE_SETP(v.c_start, e->c_start); MFE_SETP(v.c_of, e->c_of); MFE_SETP(v.c_slot, e->c_slot);
  MFE_SETP(v.c_sorted, e->c_sorted); MFE_SETP(v.c_pos, e->c_pos); MFE_SETP(v.ro_s, e->ro_s); MFE_SETP(v.type_s, e->type_s); MFE_SETP(v.nn, e->nn);
  MFE_SETP(v.pos_f, e->pos_f);
  MFE_SETP(v.nbr, e->nbr); MFE_SETP(v.flags, e->flags);
  v.ewald = e->ewald;
  v.lsurf = (e->lsurf && e->ewald != MFE_EWALD_NONE) ? 1 : 0;
  v.alpha = e->alpha;
  v.efp = e->efp;
  v.self_energy = e->self_energy;
  v.vol = e->box.L[0] * e->box.L[1] * e->box.L[2];
  v.net_energy = v.net_dq = 0.0;
  if (e->net_charge && e->ewald != MFE_EWALD_NONE && e->alpha > 0.0) {
    double qsum = 0.0;
    for (double qv : e->h_q) qsum += qv;
    const double a2 = e->alpha * e->alpha;
    v.net_energy = -kPi * qsum * qsum / (2.0 * v.vol * a2);
    v.net_dq = -kPi * qsum / (v.vol * a2);
  }
  if (e->ewald == MFE_EWALD_SPME) {
    for (int m = 0; m < 3; ++m) {
      v.sg.s[m] = e->s[m];
      v.sg.B[m] = e->B[m];
      v.sg.nb[m] = e->nb[m];
      v.sg.dkdr[m] = e->s[m] / e->box.L[m];
    }
    v.sg.order = e->order;
    for (int m = 0; m < 3; ++m) {
      v.sg.fB[m] = mfe::make_fastdiv(e->B[m]);
      v.sg.fE[m] = mfe::make_fastdiv(e->B[m] + e->order - 1);
    }
    v.hx = e->hx;
    v.nbricks = e->nbricks;
    MFE_SETP(v.mesh_r, e->mesh_r); MFE_SETP(v.mesh_c, e->mesh_c);
    MFE_SETP(v.b_count, e->b_count); MFE_SETP(v.b_done, e->b_done); MFE_SETP(v.b_start, e->b_start); MFE_SETP(v.b_id, e->b_id); MFE_SETP(v.b_slot, e->b_slot);
    MFE_SETP(v.rec_frac, e->rec_frac); MFE_SETP(v.rec_meta, e->rec_meta);
    MFE_SETP(v.spread_tiles, e->spread_tiles);
    const int smax2 = std::max(e->s[0], std::max(e->s[1], e->s[2])) / 2 + 1;
    for (int a = 0; a < 3; ++a) {
      MFE_SETP(v.ax.g[a], e->ax_g + a * smax2);
      MFE_SETP(v.ax.kk[a], e->ax_kk + a * smax2);
    }
    v.ax.fac0 = 4.0 * kPi / v.vol;
    v.ax.expf = -1.0 / (4.0 * e->alpha * e->alpha);
    for (int m = 0; m < 3; ++m) {
      v.plan[m] = e->plan[m];
      MFE_SETP(v.tw[m], e->tw[m]);
    }
    v.plan_half = e->plan_half;
    MFE_SETP(v.tw_half, e->tw_half);
    size_t l0, l1, l2;
    mfe::fft_geometry(e->prec, e->s, &v.fx_rpb, &v.fx_nblk, &v.fy_T, &v.fy_nbx, &v.fz_T,
                      &v.fz_nbx, &l0, &l1, &l2);
    v.f_s1 = mfe::make_fastdiv(e->s[0]);
    v.f_s2 = mfe::make_fastdiv(e->s[1]);
    v.f_s3 = mfe::make_fastdiv(e->s[2]);
    v.f_hx = mfe::make_fastdiv(e->s[0] / 2 + 1);
    v.f_m = mfe::make_fastdiv(e->s[0] / 2);
    v.f_fyT = mfe::make_fastdiv(v.fy_T);
    v.f_fzT = mfe::make_fastdiv(v.fz_T);
  }
  MFE_SETP(v.k_vec, e->k_vec); MFE_SETP(v.k_sum, e->k_sum); v.nkvec = e->nkvec;
  v.nchain = e->nchain; v.bond_p = e->bond_p; MFE_SETP(v.ch_prev, e->ch_prev); MFE_SETP(v.ch_next, e->ch_next);
  v.bond_k = e->bond_k; v.bond_eq = e->bond_eq;
  v.ang_k = e->ang_k; v.ang_eq = e->ang_eq; v.ang_p = e->ang_p;
  MFE_SETP(v.bd_fac1, e->bd_fac1); MFE_SETP(v.bd_fac2, e->bd_fac2); MFE_SETP(v.noise, e->bd_noise); v.seed = e->bd_seed;
  v.nknot = e->nknot;
  v.knot_r0 = e->knot_r0;
  v.knot_dinv = e->knot_dinv;
  MFE_SETP(v.knot_c, e->knot_c);
  v.pp = e->pp;
  MFE_SETP(v.phi_rec, e->phi_rec);
  MFE_SETP(v.adj_phi, e->adj_phi);
  MFE_SETP(v.adj_dq, e->adj_dq);
  MFE_SETP(v.adj_part, e->adj_part);
  v.adj_ncol = 1 + 2 * e->npt + e->nptpt * e->nknot;
  MFE_SETP(v.fm_fld, e->fm_fld);
  MFE_SETP(v.fm_res, e->fm_res);
  MFE_SETP(v.fm_res_s, e->fm_res_s);
  MFE_SETP(v.fm_vec, e->fm_vec);
  v.fm_ncol = 1 + 3 * e->npt + e->nptpt * e->nknot;
  MFE_SETP(v.part, e->part); MFE_SETP(v.en, e->en);
  v.off_pair = e->off_pair; v.off_conv = e->off_conv; v.off_bond = e->off_bond;
  v.off_surf = e->off_surf; v.off_std = e->off_std; v.off_move = e->off_move;
  v.off_skin = e->off_skin;
  v.nblk_np = nblocks256(e->np);
  v.nblk_pair = e->nblk_pair;
  v.pair_lanes = pair_lanes_for(e);
  v.pair_pos64 = pair_pos64_for(e) ? 1 : 0;
  v.nblk_conv = e->nblk_conv;
  // UTotal's force scaling (sclene/scllen)/sclfor and the surface term's
  // force factor EpsiFourPi*(4 pi/3V) (Src/energy.F90:405-418, 1352-1362)
  const double sclvel = 1.0e-10 / 1.0e-12, sclacc = sclvel / 1.0e-12;
  v.ffac = (1.0e+03 / 1.0e-10) / (sclacc * 1.0e-03);
  v.surf_fac = e->efp * (2.0 * (2.0 * kPi / (3.0 * v.vol)));
}

namespace mfe {

hipError_t side_create(Side &s) {
  hipError_t rc = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
  if (rc == hipSuccess) rc = hipEventCreateWithFlags(&s.fork, hipEventDisableTiming);
  if (rc == hipSuccess) rc = hipEventCreateWithFlags(&s.join, hipEventDisableTiming);
  if (rc != hipSuccess) side_destroy(s);
  return rc;
}

void side_destroy(Side &s) {
  if (s.fork) (void)hipEventDestroy(s.fork);
  if (s.join) (void)hipEventDestroy(s.join);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = Side{};
}

}  // namespace mfe

// the launch group of a set of engines (maxima of the per-replica grids)
int mfe_group_build(mfe_engine *const *eng, int R, const mfe::SysDev *d, hipStream_t st,
                    mfe::Group &g, std::string &err) {
  g = mfe::Group{};
  g.R = R;
  g.d = d;
  g.stream = st;
  for (int r = 0; r < R; ++r) {
    const mfe_engine *e = eng[r];
    if (r == 0) {
      g.prec = e->prec;
      g.ewald = e->ewald;
      g.order = e->ewald == MFE_EWALD_SPME ? e->order : 0;
    } else {
      if (e->prec != g.prec) { err = "replicas differ in precision"; return MFE_E_BAD_PARAM; }
      if (e->ewald != g.ewald) { err = "replicas differ in Ewald kind"; return MFE_E_BAD_PARAM; }
      if (g.ewald == MFE_EWALD_SPME && e->order != g.order) {
        err = "replicas differ in SPME order";
        return MFE_E_BAD_PARAM;
      }
    }
    if (e->nchain > 0) g.chains = true;
    if (e->lsurf && e->ewald != MFE_EWALD_NONE) g.surf = true;
    g.np_blocks = std::max(g.np_blocks, nblocks256(e->np));
    g.pair_blocks = std::max(g.pair_blocks, e->nblk_pair);
    g.lds_pair = std::max(g.lds_pair, e->t_lds_bytes);
    g.pair_npt = std::max(g.pair_npt, e->npt);
    if (e->t_lds_mode < 2) g.pair_cf_lds = false;
    if (e->nknot > 0) g.knots = true;
    if (pair_pos64_for(e)) g.pair_any64 = true;
    else g.pair_any32 = true;
    g.cell_blocks = std::max(g.cell_blocks, nblocks256(e->cg.ncells));
    g.kvec = std::max(g.kvec, e->nkvec);
    if (e->ewald == MFE_EWALD_SPME) {
      int fx_rpb, fx_nblk, fy_T, fy_nbx, fz_T, fz_nbx;
      size_t lfx, lfy, lfz;
      mfe::fft_geometry(e->prec, e->s, &fx_rpb, &fx_nblk, &fy_T, &fy_nbx, &fz_T, &fz_nbx,
                        &lfx, &lfy, &lfz);
      g.bricks = std::max(g.bricks, e->nbricks);
      for (int m = 0; m < 3; ++m) g.mc_s[m] = std::max(g.mc_s[m], e->s[m]);
      g.mesh_blocks = std::max(g.mesh_blocks, (e->s[0] * e->s[1] * e->s[2] + 255) / 256);
      g.fx_blocks = std::max(g.fx_blocks, fx_nblk);
      g.fy_x = std::max(g.fy_x, fy_nbx);
      g.fy_y = std::max(g.fy_y, e->s[2]);
      g.fz_x = std::max(g.fz_x, fz_nbx);
      g.fz_y = std::max(g.fz_y, e->s[1]);
      g.lds_fx = std::max(g.lds_fx, lfx);
      g.lds_fy = std::max(g.lds_fy, lfy);
      g.lds_fz = std::max(g.lds_fz, lfz);
      const size_t rsz = e->prec == MFE_PREC_MIXED ? 4 : 8;
      size_t gext = rsz;
      for (int m = 0; m < 3; ++m) gext *= (size_t)(e->B[m] + e->order - 1);
      g.lds_spread = std::max(g.lds_spread, (size_t)e->spread_lds);
      g.lds_gather = std::max(g.lds_gather, gext);
    }
  }
  // the fused pair + gather launch: its block holds both kernels' static
  // LDS (~6.5 KB) and the larger dynamic need, within 64 KB
  g.fuse_pg = g.prec == MFE_PREC_MIXED && g.ewald == MFE_EWALD_SPME && !g.chains && !g.surf &&
              !g.knots && !g.pair_any64 && g.pair_cf_lds &&
              std::max(g.lds_pair, g.lds_gather) <= mfe::kPairLdsBudget;
  return MFE_OK;
}

namespace {

// refresh the engine's device view and its one-replica group
void sync_view(mfe_engine *e) {
  if (!e->sys_dirty && e->d_sys) return;
  if (!e->d_sys) MFE_HIP_TRY(hipMalloc((void **)&e->d_sys, sizeof(mfe::SysDev)));
  mfe_fill_view(e, e->h_sys);
  MFE_HIP_TRY(hipMemcpyAsync(e->d_sys, &e->h_sys, sizeof(mfe::SysDev), hipMemcpyHostToDevice,
                             e->stream));
  MFE_HIP_TRY(hipStreamSynchronize(e->stream));
  std::string err;
  mfe_engine *one[1] = {e};
  mfe_group_build(one, 1, e->d_sys, e->stream, e->grp, err);
  e->grp.side = e->side;
  e->grp.flags = e->flags;
  e->grp.rebuild_bar = e->rebuild_bar;
  e->grp.rebuild_sum = e->rebuild_sum;
  e->grp.ncu = e->ncu;
  e->grp.sync_spin = e->sync_spin;
  e->sys_dirty = false;
}

// any set-up change: the device view, the step graphs and batches holding
// the engine are rebuilt before the next launch
void touch(mfe_engine *e) {
  e->graph_dirty = true;
  e->sys_dirty = true;
  e->mc_cached = false;
  e->forces_current = false;
  ++e->version;
}

}  // namespace

// error bits raised by the kernels -> status (the bits are cleared)
int mfe_map_flag_error(mfe_engine *e, int err) {
  if (!err) return MFE_OK;
  mfe::launch_fill_int(e->flags, 1, 0, e->stream);
  MFE_HIP_TRY(hipStreamSynchronize(e->stream));
  if (err & mfe::ERR_SYNC) {
    // the barrier state is undefined after a timed-out wait, and the cell
    // counts may hold the failed rebuild's increments: both are cleared
    // before the list is built again from scratch
    mfe::launch_fill_int((int *)e->rebuild_bar, mfe::kBarWords, 0, e->stream);
    if (e->c_count) mfe::launch_fill_int(e->c_count, e->ncells_alloc, 0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->list_valid = false;
    return fail(e, MFE_E_HIP, "list rebuild: a grid barrier timed out (CUs not all available)");
  }
  if (err & mfe::ERR_CAPACITY) {
    e->list_valid = false;
    return fail(e, MFE_E_NEIGH_CAPACITY, "nneighpn > maxnneigh: raise maxnneigh");
  }
  if (err & mfe::ERR_BIN) {
    // the brick counts are cleared before the next binning
    if (e->b_count) mfe::launch_fill_int(e->b_count, e->nbricks, 0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    return fail(e, MFE_E_HIP, "SPME binning: brick counts do not add up to the particles");
  }
  if (err & mfe::ERR_OVERLAP) return fail(e, MFE_E_OVERLAP_R2UMIN, "r2 < r2umin(iptjpt)");
  if (err & mfe::ERR_TABLE) return fail(e, MFE_E_TABLE_RANGE, "ibuf > nbuf");
  return MFE_OK;
}

void mfe_sync_view(mfe_engine *e) { sync_view(e); }

namespace {

// the finalize launch; with rebuild (MD/BD steps) the gated list rebuild
// rides in it where the group allows: true when it did
bool finalize_launch(const mfe::Group &g, unsigned flags, bool skin, bool bd, int mode,
                     double dt, bool split, bool rebuild) {
  if (skin && rebuild && mfe::launch_finalize_rebuild(g, flags, bd, mode, dt, split))
    return true;
  mfe::launch_reduce_finalize(g, flags, skin, bd, mode, dt, split);
  return false;
}

}  // namespace

// UTotal's device sequence; mode 0 = forces only, 1 = with VelVer kick;
// skin: also take the NList skin decision (MD/BD steps)
bool mfe_launch_utotal(const mfe::Group &g, unsigned flags, int mode, double dt, bool skin,
                       bool bd, bool rebuild) {
  const bool energy = (flags & (MFE_ENERGY | MFE_VIRIAL)) != 0;
  // inside MD/BD steps (skin) k_move has already updated the cell-ordered
  // copy, and the pair launch takes the skin test's block candidates.
  // With SPME the pair forces and the reciprocal stages up to the inverse
  // x pass share no data: the pair launch runs on the side stream,
  // alongside them, and is joined before the gather adds the reciprocal
  // forces to the pair forces.  A pair launch alone fills a small part of
  // the GPU (a lane per particle at low density).
  if (g.fuse_pg) {
    // the pair forces ride in the gather's launch (no stream fork); the
    // finalize adds the two force arrays
    if (!skin) mfe::launch_sort_pos(g);
    mfe::launch_spme(g, energy, SP_BIN, SP_IFFTX);
    mfe::launch_pair_gather(g, energy, skin);
    return finalize_launch(g, flags, skin, bd, mode, dt, true, rebuild);
  }
  const bool fork = g.ewald == MFE_EWALD_SPME && g.side.stream;
  int first = SP_BIN;
  if (fork) {
    if (!skin) mfe::launch_sort_pos(g);
    MFE_HIP_TRY(hipEventRecord(g.side.fork, g.stream));
    MFE_HIP_TRY(hipStreamWaitEvent(g.side.stream, g.side.fork, 0));
    mfe::Group gp = g;
    gp.stream = g.side.stream;
    mfe::launch_pair(gp, energy, false, skin);
    MFE_HIP_TRY(hipEventRecord(g.side.join, g.side.stream));
    mfe::launch_spme(g, energy, SP_BIN, SP_IFFTX);
    MFE_HIP_TRY(hipStreamWaitEvent(g.stream, g.side.join, 0));
    first = SP_GATHER;
  } else {
    mfe::launch_pair(g, energy, !skin, skin);
  }
  // with SPME as the last force stage and no VelVer kick, the gather
  // launch also finalizes each particle (forceo only) and reduces the
  // partials.  With the kick the finalize phase reads and writes velocities
  // and accelerations too, scattered in brick order: a separate launch in
  // particle order is faster then.
  const bool gather_fin = g.ewald == MFE_EWALD_SPME && !g.chains && !g.surf && mode == 0;
  if (gather_fin) {
    const mfe::FinArgs fa{energy ? 1 : 0, skin ? 1 : 0, bd ? 1 : 0, mode, dt};
    mfe::launch_spme(g, energy, first, SP_GATHER, mfe::GM_FORCE, &fa);
    return false;
  }
  if (g.ewald == MFE_EWALD_SPME) mfe::launch_spme(g, energy, first, SP_GATHER);
  else if (g.ewald == MFE_EWALD_STD) mfe::launch_std_ewald(g);
  if (g.chains) mfe::launch_bonds(g);
  if (g.surf) mfe::launch_surf(g);
  if (g.surf) {
    // the surface field of the reduction enters the forces
    mfe::launch_reduce(g, flags, skin, bd);
    mfe::launch_finalize(g, mode, dt);
    return false;
  }
  return finalize_launch(g, flags, skin, bd, mode, dt, false, rebuild);
}

// one VelVer step with the NList skin test and gated rebuild
void mfe_launch_md_step(const mfe::Group &g, double dt, unsigned flags) {
  mfe::launch_move(g, dt, 0);
  if (!mfe_launch_utotal(g, flags, 1, dt, true, false, true)) mfe::launch_nlist_rebuild(g, true);
}

// one BDStep (noise: 1 = host-supplied buffer, 2 = counter-based stream)
void mfe_launch_bd_step(const mfe::Group &g, int noise, unsigned flags) {
  mfe::launch_move(g, 0.0, noise);
  if (!mfe_launch_utotal(g, flags, 0, 0.0, true, true, true)) mfe::launch_nlist_rebuild(g, true);
}

namespace {

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
  MFE_HIP_TRY(hipMemcpyAsync(e->h_en, e->en, mfe::kEnLen * sizeof(double),
                             hipMemcpyDeviceToHost, e->stream));
}

}  // namespace

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
    MFE_HIP_TRY(mfe::side_create(e->side));
    hipDeviceProp_t prop;
    MFE_HIP_TRY(hipGetDeviceProperties(&prop, e->device));
    e->ncu = prop.multiProcessorCount;
    dalloc(e->rebuild_bar, mfe::kBarWords);
    dalloc(e->rebuild_sum, (size_t)e->ncu);
    mfe::launch_fill_int((int *)e->rebuild_bar, mfe::kBarWords, 0, e->stream);
    dalloc(e->flags, 4);
    dalloc(e->en, mfe::kEnLen);
    MFE_HIP_TRY(hipHostMalloc((void **)&e->h_en, mfe::kEnLen * sizeof(double)));
    MFE_HIP_TRY(hipHostMalloc((void **)&e->h_flags, 4 * sizeof(int)));
    dalloc(e->bd_fac1, MFE_MAX_TYPES);
    dalloc(e->bd_fac2, MFE_MAX_TYPES);
    mfe::launch_fill_int(e->flags, 4, 0, e->stream);
    mfe::launch_fill_double(e->en, mfe::kEnLen, 0.0, e->stream);
    mfe::launch_fill_double(e->bd_fac1, MFE_MAX_TYPES, 0.0, e->stream);
    mfe::launch_fill_double(e->bd_fac2, MFE_MAX_TYPES, 0.0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    return MFE_OK;
  });
  if (rc != MFE_OK) {
    free_all(e);
    mfe::side_destroy(e->side);
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
  if (e->side.stream) (void)hipStreamSynchronize(e->side.stream);
  free_all(e);
  mfe::side_destroy(e->side);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return MFE_OK;
}

const char *mfe_last_error(const mfe_engine *e) { return e ? e->err.c_str() : "null engine"; }

int mfe_set_net_charge(mfe_engine *e, int32_t on) {
  if (!e) return MFE_E_BAD_PARAM;
  e->net_charge = on != 0;
  touch(e);
  return MFE_OK;
}

int mfe_set_sync_limit(mfe_engine *e, int32_t polls) {
  if (!e) return MFE_E_BAD_PARAM;
  e->sync_spin = polls < 0 ? -1 : polls;
  e->sys_dirty = true;
  e->graph_dirty = true;
  ++e->version;
  return MFE_OK;
}

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
  touch(e);
  return MFE_OK;
}

int mfe_set_particles(mfe_engine *e, int32_t np, int32_t npt, const int32_t *iptpn,
                      const double *az, const double *massipt) {
  if (!e || !iptpn || !az || !massipt) return MFE_E_BAD_PARAM;
  if (np < 1 || np > e->max_np) return fail(e, MFE_E_BAD_PARAM, "np out of range (0, max_np]");
  if (npt < 1 || npt > MFE_MAX_TYPES) return fail(e, MFE_E_BAD_PARAM, "npt out of range");
  if (np > mfe::kMaxParticles) return fail(e, MFE_E_BAD_PARAM, "np > 2^25");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    std::vector<int> t(np);
    std::vector<double> mi(np);
    for (int i = 0; i < np; ++i) {
      if (iptpn[i] < 1 || iptpn[i] > npt) return fail(e, MFE_E_BAD_PARAM, "iptpn out of range");
      t[i] = iptpn[i] - 1;
      mi[i] = massipt[t[i]];
      // the mixed spreading's 2^-32 fixed point takes |q| < 2^19
      if (!(std::fabs(az[i]) < 65536.0)) return fail(e, MFE_E_BAD_PARAM, "charge not finite or |q| >= 2^16");
    }
    if (np != e->np) {
      dalloc(e->ro, 3 * (size_t)np); dalloc(e->rod, 3 * (size_t)np);
      dalloc(e->rodd, 3 * (size_t)np); dalloc(e->force, 3 * (size_t)np);
      dalloc(e->force_rec, 3 * (size_t)np);
      dalloc(e->forceo, 3 * (size_t)np); dalloc(e->drosum, 3 * (size_t)np);
      dalloc(e->r2disp, (size_t)np); dalloc(e->type, (size_t)np);
      dalloc(e->q, (size_t)np); dalloc(e->massip, (size_t)np);
      dalloc(e->c_of, (size_t)np); dalloc(e->c_slot, (size_t)np);
      dalloc(e->c_sorted, (size_t)np); dalloc(e->c_pos, (size_t)np); dalloc(e->nn, (size_t)np);
      dalloc(e->ro_s, 3 * (size_t)np); dalloc(e->type_s, (size_t)np);
      if (e->prec == MFE_PREC_MIXED) dalloc(e->pos_f, (size_t)np);
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
      upload(e->c_pos, ident.data(), (size_t)np, e->stream);
      MFE_HIP_TRY(hipStreamSynchronize(e->stream));
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
    touch(e);
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
        for (int c = 11; c < mfe::kCoefStride; ++c) cf.push_back(0.0);
      }
    }
    // FP32 copy for the mixed path.  Its kernel evaluates the polynomials
    // in d' = r2 - (float)zlow, so each interval's two polynomials are
    // re-expanded about the rounded bound (in d = d' - eps, eps = zlow -
    // (float)zlow; in long double) before rounding: otherwise every pair of
    // an interval would see the same argument shift, a bias that does not
    // average out over the sum.  The free twelfth slot carries the rounding
    // of the interval constant w0, w0 - (float)w0, so that energy passes add
    // w0 to the FP32 remainder polynomial in FP64.
    std::vector<float> cff(cf.size());
    for (size_t k = 0; k < zl.size(); ++k) {
      const size_t r = k * mfe::kCoefStride;
      const long double eps = (long double)zl[k] - (long double)(float)zl[k];
      auto shift = [&](int first, int deg) {
        long double w[6];
        for (int j = 0; j <= deg; ++j) w[j] = cf[r + first + j];
        // synthetic division by (d' - eps) repeated: Taylor coefficients
        // of p(d' - eps) in d'
        for (int i = 0; i < deg; ++i)
          for (int j = deg - 1; j >= i; --j) w[j] -= eps * w[j + 1];
        for (int j = 0; j <= deg; ++j) cff[r + first + j] = (float)w[j];
        return (double)w[0];
      };
      const double w0 = shift(0, 5);
      shift(6, 4);
      cff[r + 11] = (float)(w0 - (double)cff[r]);
    }
    // bins uniform in s = 1/r2, where the tables' intervals are roughly
    // evenly spaced.  The kernel computes the bin in FP32 from r2 clamped
    // to [lo, inf) and clamps it into [0, nbin); each bin is built over
    // its s range widened by kBinMargin bins on both sides (far more than
    // that rounding), so "interval kb or kb + 1" holds for every r2 the
    // kernel can put there -- else the bin is marked slow.
    const int nbin = std::min(mfe::kTableBinsPerPair, mfe::kTableBins / P);
    constexpr double kBinMargin = 0.05;
    std::vector<unsigned short> bin_k((size_t)P * nbin);
    std::vector<mfe::PairRec> rec(P);
    for (int p = 0; p < P; ++p) {
      const double *z = zl.data() + off[p];
      const int n = nint[p];
      if (n > mfe::kSlowBin - 1)
        return fail(e, MFE_E_BAD_PARAM, "more than 32767 intervals in one table");
      const double lo = z[n - 1], hi = std::max(rcut * rcut, z[0]);
      // first interval k with r2 >= z[k] (r2 >= lo)
      auto first_at = [&](double r2) {
        int k = n - 1;
        while (k > 0 && r2 >= z[k - 1]) --k;
        return k;
      };
      mfe::PairRec &q = rec[p];
      q.r2at = r2atat[p];
      q.r2um = r2umin[p];
      q.lo = lo;
      q.pad = 0.0;
      q.off = off[p];
      q.nint = n;
      q.s_lo = (float)(1.0 / hi);
      const double span = 1.0 / lo - (double)q.s_lo;
      q.s_inv = span > 0.0 ? (float)(nbin / span) : 0.0f;
      const double w = q.s_inv > 0.0f ? 1.0 / (double)q.s_inv : 0.0;
      for (int b = 0; b < nbin; ++b) {
        const double sa = (double)q.s_lo + (b - kBinMargin) * w;
        const double sb = (double)q.s_lo + (b + 1 + kBinMargin) * w;
        const double r2a = (b == 0 || !(sa > 0.0)) ? HUGE_VAL : 1.0 / sa;
        const double r2b = (b == nbin - 1 || !(sb > 0.0)) ? lo : std::max(lo, 1.0 / sb);
        const int kmin = first_at(r2a), kmax = first_at(r2b);
        bin_k[(size_t)p * nbin + b] =
            (unsigned short)(kmax - kmin <= 1 ? kmin : (kmin | mfe::kSlowBin));
      }
    }
    std::vector<int> pidx(e->npt * e->npt);
    int k = 0;
    for (int i = 0; i < e->npt; ++i)
      for (int j = i; j < e->npt; ++j, ++k) pidx[i * e->npt + j] = pidx[j * e->npt + i] = k;
    e->t_nzl = (int)zl.size();
    dalloc(e->t_zlow, zl.size()); dalloc(e->t_coef, cf.size()); dalloc(e->t_coef_f, cf.size());
    dalloc(e->t_off, (size_t)P); dalloc(e->t_nint, (size_t)P); dalloc(e->t_pairidx, pidx.size());
    dalloc(e->t_r2atat, (size_t)P); dalloc(e->t_r2umin, (size_t)P);
    dalloc(e->t_bin_k, bin_k.size()); dalloc(e->t_rec, (size_t)P);
    upload(e->t_bin_k, bin_k.data(), bin_k.size(), e->stream);
    upload(e->t_rec, rec.data(), (size_t)P, e->stream);
    std::vector<mfe::PairRecF> rec_f(P);
    for (int p = 0; p < P; ++p) {
      rec_f[p] = mfe::PairRecF{(float)rec[p].r2at, (float)rec[p].r2um, (float)rec[p].lo,
                               rec[p].s_lo, rec[p].s_inv, rec[p].off, rec[p].nint, 0};
    }
    std::vector<float> zl_f(zl.begin(), zl.end());
    dalloc(e->t_rec_f, (size_t)P);
    dalloc(e->t_zlow_f, zl_f.size());
    upload(e->t_rec_f, rec_f.data(), (size_t)P, e->stream);
    upload(e->t_zlow_f, zl_f.data(), zl_f.size(), e->stream);
    e->t_nbin = nbin;
    // LDS staging of the pair kernel within its budget
    const size_t csz = e->prec == MFE_PREC_MIXED ? 4 : 8;
    const size_t bb = mfe::pair_bin_bytes(P, nbin), zb = mfe::pair_zl_bytes((int)zl.size());
    const size_t cb = zl.size() * mfe::kCoefStride * csz;
    e->t_lds_mode = 0;
    e->t_lds_bytes = bb;
    if (bb + zb <= mfe::kPairLdsBudget) {
      e->t_lds_mode = 1;
      e->t_lds_bytes = bb + zb;
      if (bb + zb + cb <= mfe::kPairLdsBudget) {
        e->t_lds_mode = 2;
        e->t_lds_bytes = bb + zb + cb;
      }
    }
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
    touch(e);
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
      // bricks of the spreading and gather blocks: B divides s; 16 when s
      // allows (fewer halo points per brick), else the smallest divisor >= 8
      size_t ext = 1;
      for (int m = 0; m < 3; ++m) {
        int b = (e->s[m] % 16 == 0) ? 16 : smallest_divisor_at_least(e->s[m], 8);
        if (b > 16) b = smallest_divisor_at_least(e->s[m], order - 1);
        e->B[m] = b;
        e->nb[m] = e->s[m] / b;
        ext *= (size_t)(b + order - 1);
      }
      e->nbricks = e->nb[0] * e->nb[1] * e->nb[2];
      const size_t rsz = e->prec == MFE_PREC_MIXED ? 4 : 8;
      // spreading tile: 64-bit fixed point (mixed) or FP64 per extended point
      e->spread_lds = (int)(ext * 8);
      const size_t gext = ext * rsz;
      const int lds_need = (int)std::max<size_t>((size_t)e->spread_lds, gext);
      if ((size_t)lds_need > 160 * 1024)
        return fail(e, MFE_E_BAD_PARAM, "mesh brick too large for LDS (mesh 2q, q prime > 16?)");
      const size_t M = (size_t)e->s[0] * e->s[1] * e->s[2];
      const size_t Mc = (size_t)e->hx * e->s[1] * e->s[2];
      dfree(e->mesh_r);
      dfree(e->mesh_c);
      MFE_HIP_TRY(hipMalloc(&e->mesh_r, M * rsz));
      MFE_HIP_TRY(hipMalloc(&e->mesh_c, Mc * 2 * rsz));
      dfree(e->spread_tiles);
      e->spread_tiles_bytes = (size_t)e->nbricks * ext * rsz;
      MFE_HIP_TRY(hipMalloc(&e->spread_tiles, e->spread_tiles_bytes));
      dalloc(e->b_count, (size_t)e->nbricks);
      dalloc(e->b_start, (size_t)e->nbricks + 1);
      mfe::launch_fill_int(e->b_count, e->nbricks, 0, e->stream);
      dalloc(e->b_done, 1);
      mfe::launch_fill_int(e->b_done, 1, 0, e->stream);
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
      if ((size_t)lds_need > 64 * 1024) {
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
    touch(e);
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
    touch(e);
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
    touch(e);
    return MFE_OK;
  });
}

// UAngle's parameters (Src/energy.F90:5409-5481) for the chains of
// mfe_set_chains: the namelist force constant k (per degree^2) and
// equilibrium angle eq (degrees) scaled as IOPotChain does
// (Src/potential.F90:3558-3562): k / (pi/180)^2 / p and eq * pi/180.
int mfe_set_angles(mfe_engine *e, double k, int32_t p, double eq) {
  if (!e) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  if (p < 1 || !(k >= 0.0) || !std::isfinite(k) || !std::isfinite(eq))
    return fail(e, MFE_E_BAD_PARAM, "angle: k < 0, p < 1 or not finite");
  const double sclang = 3.14159265358979323846 / 180.0;  // DegToRad
  e->ang_k = k / (sclang * sclang) / p;
  e->ang_eq = eq * sclang;
  e->ang_p = p;
  touch(e);
  return MFE_OK;
}

// u%angle of the last force call that reduced energies
int mfe_get_angle_energy(mfe_engine *e, double *u_angle) {
  if (!e || !u_angle) return MFE_E_BAD_PARAM;
  *u_angle = e->h_en ? e->h_en[mfe::kEnAngle] : 0.0;
  return MFE_OK;
}

int mfe_upload_positions(mfe_engine *e, const double *ro) {
  if (!e || !ro) return MFE_E_BAD_PARAM;
  e->forces_current = false;
  e->mc_cached = false;
  if (!e->have_particles || !e->have_box) return fail(e, MFE_E_STATE, "box and particles first");
  for (int i = 0; i < 3 * e->np; ++i) {
    const int m = i % 3;
    if (!(std::fabs(ro[i]) <= e->box.L2[m]))
      return fail(e, MFE_E_BAD_PARAM, "particle outside the box [-L/2, L/2]");
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    if (e->list_valid && e->have_positions && e->have_neighbor) {
      // NList(iSimulationStep): displacements since the last upload, the
      // skin test and, only when it fires, the rebuild -- all on the device
      if (!e->ro_stage) dalloc(e->ro_stage, 3 * (size_t)e->max_np);
      upload(e->ro_stage, ro, 3 * (size_t)e->np, e->stream);
      sync_view(e);
      mfe::launch_set_positions(e->grp, e->ro_stage);
      mfe::launch_skin(e->grp);
      mfe::launch_reduce(e->grp, 0u, true, false);
      mfe::launch_nlist_rebuild(e->grp, true);
      return read_flags(e);
    }
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

int mfe_download_forces(mfe_engine *e, double *forceo) {
  if (!e || !forceo) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipMemcpyAsync(forceo, e->forceo, 3 * (size_t)e->np * sizeof(double),
                               hipMemcpyDeviceToHost, e->stream));
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
    sync_view(e);
    mfe::launch_nlist_rebuild(e->grp, false);
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
    sync_view(e);
    mfe_launch_utotal(e->grp, f, 0, 0.0, false, false);
    if (forceo)
      MFE_HIP_TRY(hipMemcpyAsync(forceo, e->forceo, 3 * (size_t)e->np * sizeof(double),
                                 hipMemcpyDeviceToHost, e->stream));
    fetch_energy(e);
    int r = read_flags(e);
    if (r) return r;
    e->forces_current = true;
    if (energy) energy_out(e, energy);
    return MFE_OK;
  });
}

namespace {

// checks of a trial move (1-based distinct particles, finite positions)
// and a valid list; ia gets the 0-based indices
int mc_prepare(mfe_engine *e, int natm, const int32_t *ianatm, const double *rtm,
               std::vector<int> &ia) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (natm > e->np) return fail(e, MFE_E_BAD_PARAM, "natm > np");
  ia.assign(natm, 0);
  std::vector<char> seen(e->np, 0);
  for (int a = 0; a < natm; ++a) {
    const int i = ianatm[a] - 1;
    if (i < 0 || i >= e->np) return fail(e, MFE_E_BAD_PARAM, "ianatm out of range");
    if (seen[i]) return fail(e, MFE_E_BAD_PARAM, "ianatm repeats a particle");
    seen[i] = 1;
    ia[a] = i;
    for (int m = 0; m < 3; ++m)
      if (!std::isfinite(rtm[3 * a + m])) return fail(e, MFE_E_BAD_PARAM, "rtm not finite");
  }
  if (!e->list_valid) return mfe_build_list(e);
  return MFE_OK;
}

// device room for natm moved particles and the four results
void mc_buffers(mfe_engine *e, int natm) {
  if (e->mc_cap < natm) {
    dalloc(e->mc_ia, (size_t)natm);
    dalloc(e->mc_rtm, 3 * (size_t)natm);
    e->mc_cap = natm;
  }
  // [reciprocal change, 0, du_twob, overlap, block partials]
  if (!e->mc_out) dalloc(e->mc_out, 4 + 2 * (size_t)mfe::kMcPairBlocks);
  // pinned staging: 4 results, then the indices, then the positions
  const size_t bytes = 32 + ((size_t)natm * sizeof(int) + 7) / 8 * 8 + 24 * (size_t)natm;
  if (e->h_mc_bytes < bytes) {
    if (e->h_mc) MFE_HIP_TRY(hipHostFree(e->h_mc));
    e->h_mc = nullptr;
    e->h_mc_bytes = 0;
    MFE_HIP_TRY(hipHostMalloc((void **)&e->h_mc, bytes));
    e->h_mc_bytes = bytes;
  }
}

// the move's indices and positions in the pinned staging (device-
// accessible host memory: a kernel that reads them a few times takes them
// from there directly)
void mc_stage(mfe_engine *e, const std::vector<int> &ia, const double *rtm, int **hi_out,
              double **hr_out) {
  const size_t natm = ia.size();
  int *hi = reinterpret_cast<int *>(e->h_mc + 32);
  double *hr = reinterpret_cast<double *>(e->h_mc + 32 + (natm * sizeof(int) + 7) / 8 * 8);
  std::memcpy(hi, ia.data(), natm * sizeof(int));
  std::memcpy(hr, rtm, 3 * natm * sizeof(double));
  *hi_out = hi;
  *hr_out = hr;
}

// the move's indices and positions to the device through the pinned
// staging (asynchronous copies; every MC call ends with a synchronization
// before the staging is written again), for kernels that read them often
void mc_upload(mfe_engine *e, const std::vector<int> &ia, const double *rtm) {
  int *hi;
  double *hr;
  mc_stage(e, ia, rtm, &hi, &hr);
  upload(e->mc_ia, hi, ia.size(), e->stream);
  upload(e->mc_rtm, hr, 3 * ia.size(), e->stream);
}

// the results slots of the pinned staging, which the last kernel of a
// trial writes directly
double *mc_result_slots(mfe_engine *e) { return reinterpret_cast<double *>(e->h_mc); }

// n results from the staging's result slots (waits for the stream)
void mc_results(mfe_engine *e, int n, double *out) {
  MFE_HIP_TRY(hipStreamSynchronize(e->stream));
  const double *h = mc_result_slots(e);
  for (int k = 0; k < n; ++k) out[k] = h[k];
}

// the kept spectrum (the complex mesh's size) and the stencil factors of
// natm moved particles; a new spectrum buffer is not a valid one
void mc_spectrum_buffers(mfe_engine *e, int natm) {
  // FP64 values on both precision paths
  const size_t bytes = (size_t)e->hx * e->s[1] * e->s[2] * 16;
  if (e->mc_spec_bytes != bytes) {
    if (e->mc_spec) MFE_HIP_TRY(hipFree(e->mc_spec));
    e->mc_spec = nullptr;
    e->mc_spec_bytes = 0;
    e->mc_cached = false;
    MFE_HIP_TRY(hipMalloc(&e->mc_spec, bytes));
    e->mc_spec_bytes = bytes;
  }
  const size_t nfac = 2 * 2 * (size_t)natm * (e->s[0] + e->s[1] + e->s[2]);
  if (e->mc_fac_cap < nfac) {
    dalloc(e->mc_fac, nfac);
    e->mc_fac_cap = nfac;
  }
}

}  // namespace

// UTwoBodyANew - UTwoBodyAOld (Src/denergy.F90:289-495) for a trial move
// of natm particles: partners from the engine's current list, as the
// reference walks jpnlist.  du = change of the two-body energy (the sum
// over pair types, kJ/mol); overlap = 1 when a trial pair lies inside the
// hard core or below r2umin (the reference's lhsoverlap), du then 0.
int mfe_mc_delta_pair(mfe_engine *e, int32_t natm, const int32_t *ianatm, const double *rtm,
                      double *du_twob, int32_t *overlap) {
  if (!e || !ianatm || !rtm || !du_twob || !overlap || natm < 1) return MFE_E_BAD_PARAM;
  std::vector<int> ia;
  int rc = mc_prepare(e, natm, ianatm, rtm, ia);
  if (rc) return rc;
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    mc_buffers(e, natm);
    // small moves straight from the pinned staging (the kernel stages them
    // in LDS), larger ones through device copies
    int *hi;
    double *hr;
    mc_stage(e, ia, rtm, &hi, &hr);
    const int *mia = hi;
    const double *mrtm = hr;
    if (natm > mfe::kMcStage) {
      upload(e->mc_ia, hi, (size_t)natm, e->stream);
      upload(e->mc_rtm, hr, 3 * (size_t)natm, e->stream);
      mia = e->mc_ia;
      mrtm = e->mc_rtm;
    }
    sync_view(e);
    mfe::launch_mc_pair(e->grp, natm, mia, mrtm, mc_result_slots(e), e->mc_out + 4,
                        e->nlkind == MFE_NL_CELL);
    double h[2];
    mc_results(e, 2, h);
    *overlap = h[1] != 0.0 ? 1 : 0;
    *du_twob = *overlap ? 0.0 : h[0];
    return MFE_OK;
  });
}

// ---------------------------------------------------------------------
// The pair part of a trial move without any allocation: the moved
// particles' indices and positions and the kernels' results and block
// partials live in one file-scope device array, `scratch`, of
// kNoAllocBytes.  Host code never passes the array's name to a kernel or a
// copy (on the host it is a stand-in symbol, not a device address): its
// address on each device is looked up once with hipGetSymbolAddress, with
// that device current, and kept per device.  Every call on a device shares
// that device's array, so calls that overlap on one device race; nothing
// in it needs clearing (every kernel overwrites what it later reads).
namespace {

constexpr int kNoAllocMaxAtoms = 1024;
// [ia: kNoAllocMaxAtoms int][rtm: 3 kNoAllocMaxAtoms double][out: 4 double]
// [partials: 2 kMcPairBlocks double], each part 256-byte aligned
constexpr size_t kNaIaOff = 0;
constexpr size_t kNaRtmOff = 4096;
constexpr size_t kNaOutOff = kNaRtmOff + 3 * 8 * kNoAllocMaxAtoms;
constexpr size_t kNaPartOff = kNaOutOff + 256;
constexpr size_t kNoAllocBytes = kNaPartOff + 2 * 8 * mfe::kMcPairBlocks;
static_assert(kNaRtmOff >= sizeof(int) * kNoAllocMaxAtoms, "scratch layout");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on device dev (current), looked up once
int scratch_on(int dev, void **ptr, std::string &err) {
  if (dev < 0 || dev >= kMaxDevices) {
    err = "device ordinal out of range for the no-allocation entry point";
    return MFE_E_BAD_PARAM;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc);
      return MFE_E_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return MFE_OK;
}

}  // namespace

int mfe_mc_delta_pair_noalloc(mfe_engine *e, int32_t natm, const int32_t *ianatm,
                              const double *rtm, double *du_twob, int32_t *overlap) {
  if (!e || !ianatm || !rtm || !du_twob || !overlap || natm < 1) return MFE_E_BAD_PARAM;
  if (natm > kNoAllocMaxAtoms)
    return fail(e, MFE_E_BAD_PARAM, "natm > 1024, the no-allocation scratch limit");
  int rc = check_ready(e);
  if (rc) return rc;
  if (natm > e->np) return fail(e, MFE_E_BAD_PARAM, "natm > np");
  // the device view must exist already (creating it would allocate)
  if (!e->d_sys || !e->list_valid)
    return fail(e, MFE_E_STATE, "build the list first (mfe_build_list or mfe_compute)");
  int ia[kNoAllocMaxAtoms];
  for (int a = 0; a < natm; ++a) {
    const int i = ianatm[a] - 1;
    if (i < 0 || i >= e->np) return fail(e, MFE_E_BAD_PARAM, "ianatm out of range");
    for (int b = 0; b < a; ++b)
      if (ia[b] == i) return fail(e, MFE_E_BAD_PARAM, "ianatm repeats a particle");
    ia[a] = i;
    for (int m = 0; m < 3; ++m)
      if (!std::isfinite(rtm[3 * a + m])) return fail(e, MFE_E_BAD_PARAM, "rtm not finite");
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    void *base = nullptr;
    std::string err;
    const int src = scratch_on(e->device, &base, err);
    if (src) return fail(e, src, err);
    unsigned char *sp = static_cast<unsigned char *>(base);
    int *d_ia = reinterpret_cast<int *>(sp + kNaIaOff);
    double *d_rtm = reinterpret_cast<double *>(sp + kNaRtmOff);
    double *d_out = reinterpret_cast<double *>(sp + kNaOutOff);
    double *d_part = reinterpret_cast<double *>(sp + kNaPartOff);
    MFE_HIP_TRY(hipMemcpyAsync(d_ia, ia, natm * sizeof(int), hipMemcpyHostToDevice, e->stream));
    MFE_HIP_TRY(hipMemcpyAsync(d_rtm, rtm, 3 * (size_t)natm * sizeof(double),
                               hipMemcpyHostToDevice, e->stream));
    sync_view(e);
    mfe::launch_mc_pair(e->grp, natm, d_ia, d_rtm, d_out, d_part, e->nlkind == MFE_NL_CELL);
    double h[2];
    MFE_HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    *overlap = h[1] != 0.0 ? 1 : 0;
    *du_twob = *overlap ? 0.0 : h[0];
    return MFE_OK;
  });
}

// DUTwoBodyEwaldRecSPM (Src/denergy.F90:1413-1495) for a trial move of
// natm particles: the current configuration is binned and spread, its
// forward transforms and convolution give E_rec; the moved particles'
// stencils are then exchanged in the spreading tiles and the same passes
// give E_rec of the trial mesh.  du = EpsiFourPi (E_trial - E_current), the
// units of mfe_energy.rec.  Positions, forces and the list are unchanged;
// the next force call spreads the mesh again.  The current tiles and
// E_rec are kept (as the reference keeps QMesh and urecold); with
// reuse_current a later call starts from them instead of spreading again.
int mfe_mc_delta_rec(mfe_engine *e, int32_t natm, const int32_t *ianatm, const double *rtm,
                     int32_t reuse_current, double *du_rec) {
  if (!e || !ianatm || !rtm || !du_rec || natm < 1) return MFE_E_BAD_PARAM;
  std::vector<int> ia;
  int rc = mc_prepare(e, natm, ianatm, rtm, ia);
  if (rc) return rc;
  if (!e->have_ewald || e->ewald != MFE_EWALD_SPME)
    return fail(e, MFE_E_STATE, "mfe_set_ewald with SPME first");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    mc_buffers(e, natm);
    mc_spectrum_buffers(e, natm);
    int *hi;
    double *hr;
    mc_stage(e, ia, rtm, &hi, &hr);
    sync_view(e);
    const mfe::Group &G = e->grp;
    if (!(reuse_current && e->mc_cached)) {
      // the current configuration's spectrum: its mesh (binning reads the
      // cell-ordered positions), forward along x, y and z, kept
      e->mc_cached = false;
      mfe::launch_sort_pos(G);
      mfe::launch_spme(G, true, SP_BIN, SP_FFTY);
      mfe::fft_z_forward(G, e->mc_spec);
    }
    // the trial's stencil factors, then the energy change over the spectrum
    mfe::launch_mc_kfac(G, natm, hi, hr, e->mc_fac);
    mfe::launch_mc_kdelta(G, natm, e->mc_fac, e->mc_spec, false, e->mc_out + 4,
                          mc_result_slots(e));
    double h = 0.0;
    mc_results(e, 1, &h);
    e->mc_cached = true;
    e->mc_last_ia = ia;
    e->mc_last_rtm.assign(rtm, rtm + 3 * (size_t)natm);
    *du_rec = e->efp * h;
    return MFE_OK;
  });
}

// EwaldUpdateArray (Src/denergy.F90:1776-1819, 'spm': QMesh = QMeshTM,
// urecold = urecnew) with the MC driver's position update of an accepted
// move (ro(:,ianatm) = rtm) and NList's skin bookkeeping.  When the last
// mfe_mc_delta_rec was this move, its change D (from the stored stencil
// factors) is added to the kept FP64 spectrum, which is then the current
// one, so the next trial with reuse_current transforms nothing; otherwise
// the next trial transforms anew.
int mfe_mc_accept(mfe_engine *e, int32_t natm, const int32_t *ianatm, const double *rtm) {
  if (!e || !ianatm || !rtm || natm < 1) return MFE_E_BAD_PARAM;
  e->forces_current = false;
  std::vector<int> ia;
  int rc = mc_prepare(e, natm, ianatm, rtm, ia);
  if (rc) return rc;
  for (int a = 0; a < natm; ++a)
    for (int m = 0; m < 3; ++m)
      if (!(std::fabs(rtm[3 * a + m]) <= e->box.L2[m]))
        return fail(e, MFE_E_BAD_PARAM, "rtm outside the box [-L/2, L/2]");
  if (!e->list_valid || !e->have_positions || !e->have_neighbor)
    return fail(e, MFE_E_STATE, "positions and neighbour list first");
  // the kept spectrum is FP64 on both paths: the accepted change adds to
  // it without FP32 rounding
  const bool keep = e->mc_cached && e->have_ewald &&
                    e->ewald == MFE_EWALD_SPME && e->mc_spec && e->mc_last_ia == ia &&
                    e->mc_last_rtm.size() == 3 * (size_t)natm &&
                    std::equal(e->mc_last_rtm.begin(), e->mc_last_rtm.end(), rtm);
  e->mc_cached = false;
  e->mc_last_ia.clear();
  e->mc_last_rtm.clear();
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    mc_buffers(e, natm);
    // small moves straight from the pinned staging, larger ones through
    // device copies
    const bool small = natm <= mfe::kMcStage;
    int *hi = nullptr;
    double *hr = nullptr;
    if (small) mc_stage(e, ia, rtm, &hi, &hr);
    else mc_upload(e, ia, rtm);
    sync_view(e);
    const mfe::Group &G = e->grp;
    if (keep) {
      // the trial's change added to the kept spectrum (its stencil factors
      // are those of the last trial, this move)
      mfe::launch_mc_kdelta(G, natm, e->mc_fac, e->mc_spec, true, nullptr, nullptr);
    }
    // the accepted positions, then NList(iSimulationStep) as after an upload
    if (small) {
      mfe::launch_set_positions_moved(G, natm, hi, hr);
    } else {
      if (!e->ro_stage) dalloc(e->ro_stage, 3 * (size_t)e->max_np);
      MFE_HIP_TRY(hipMemcpyAsync(e->ro_stage, e->ro, 3 * (size_t)e->np * sizeof(double),
                                 hipMemcpyDeviceToDevice, e->stream));
      mfe::launch_mc_place(G, natm, e->mc_ia, e->mc_rtm, e->ro_stage);
      mfe::launch_set_positions(G, e->ro_stage);
    }
    mfe::launch_skin(G);
    mfe::launch_reduce(G, 0u, true, false);
    mfe::launch_nlist_rebuild(G, true);
    int r = read_flags(e);
    if (r) return r;
    e->mc_cached = keep;
    return MFE_OK;
  });
}

int mfe_step_md(mfe_engine *e, int32_t nsteps, double tstep, uint32_t flags) {
  if (!e || nsteps < 0) return MFE_E_BAD_PARAM;
  e->mc_cached = false;
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
    sync_view(e);
    if (e->graph_dirty || e->md_graph_kind != 1 || e->md_graph_dt != tstep ||
        e->md_graph_flags != flags) {
      e->md_graph.clear();
      e->md_graph_kind = 1;
      e->md_graph_dt = tstep;
      e->md_graph_flags = flags;
      e->graph_dirty = false;
    }
    MFE_HIP_TRY(e->md_graph.run(e->stream, nsteps,
                                [&]() { mfe_launch_md_step(e->grp, tstep, flags); }));
    fetch_energy(e);
    const int r = read_flags(e);
    e->forces_current = r == MFE_OK;
    return r;
  });
}

int mfe_set_bd(mfe_engine *e, const double *fac1, const double *fac2, uint64_t seed) {
  if (!e || !fac1 || !fac2) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    upload(e->bd_fac1, fac1, (size_t)e->npt, e->stream);
    upload(e->bd_fac2, fac2, (size_t)e->npt, e->stream);
    // restart the counter-based stream: step counter flags[3] = 0
    mfe::launch_fill_int(e->flags + 3, 1, 0, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->bd_seed = seed;
    e->have_bd = true;
    touch(e);
    return MFE_OK;
  });
}

int mfe_step_bd(mfe_engine *e, int32_t nsteps, const double *fac1, const double *fac2,
                const double *noise, uint32_t flags) {
  if (!e || nsteps < 0) return MFE_E_BAD_PARAM;
  e->mc_cached = false;
  if ((fac1 == nullptr) != (fac2 == nullptr)) return fail(e, MFE_E_BAD_PARAM, "fac1 and fac2 go together");
  if (!fac1 && !e->have_bd) return fail(e, MFE_E_STATE, "BD factors not set (mfe_set_bd)");
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->have_ewald) return fail(e, MFE_E_STATE, "mfe_set_ewald not called");
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    if (fac1) {
      upload(e->bd_fac1, fac1, (size_t)e->npt, e->stream);
      upload(e->bd_fac2, fac2, (size_t)e->npt, e->stream);
      e->have_bd = true;
    }
    if (noise && !e->bd_noise) {
      dalloc(e->bd_noise, 3 * (size_t)e->max_np);
      e->sys_dirty = true;
      e->graph_dirty = true;
      ++e->version;
    }
    sync_view(e);
    if (!e->forces_current) {
      // the first move takes the forces at the current positions (the
      // reference's UTotal before its BD loop)
      mfe_launch_utotal(e->grp, 0u, 0, 0.0, false, false);
    }
    if (noise) {
      // BDStep with the host's deviates (one upload per step)
      for (int s2 = 0; s2 < nsteps; ++s2) {
        upload(e->bd_noise, noise + 3 * (size_t)e->np * s2, 3 * (size_t)e->np, e->stream);
        mfe_launch_bd_step(e->grp, 1, flags);
      }
    } else {
      // counter-based deviates on the device: the step is a static graph
      if (e->graph_dirty || e->md_graph_kind != 2 || e->md_graph_flags != flags) {
        e->md_graph.clear();
        e->md_graph_kind = 2;
        e->md_graph_flags = flags;
        e->graph_dirty = false;
      }
      MFE_HIP_TRY(e->md_graph.run(e->stream, nsteps,
                                  [&]() { mfe_launch_bd_step(e->grp, 2, flags); }));
    }
    fetch_energy(e);
    const int r = read_flags(e);
    e->forces_current = r == MFE_OK;
    return r;
  });
}

int mfe_set_pair_params(mfe_engine *e, const mfe_pot_1_6_12 *spec) {
  if (!e || !spec || !spec->zat || !spec->sigat || !spec->epsat) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  if (spec->npt != e->npt) return fail(e, MFE_E_BAD_PARAM, "spec npt differs from the particles' npt");
  for (int i = 0; i < e->np; ++i)
    if (e->h_q[i] != spec->zat[e->h_type[i]])
      return fail(e, MFE_E_BAD_PARAM, "az differs from zat(iptpn): the adjoint needs q = zat");
  mfe::PairParams &pp = e->pp;
  pp = mfe::PairParams{};
  for (int t = 0; t < e->npt; ++t) {
    pp.zat[t] = spec->zat[t];
    pp.sigat[t] = spec->sigat[t];
    pp.epsat[t] = spec->epsat[t];
  }
  pp.efp = spec->epsi_four_pi;
  pp.alpha = spec->alpha;
  pp.lewald = spec->lewald ? 1 : 0;
  pp.lljcut = spec->lljcut ? 1 : 0;
  pp.ljrcut = spec->ljrcut;
  pp.ljushift = spec->ljushift;
  e->have_pp = true;
  touch(e);
  return MFE_OK;
}

int mfe_set_knots(mfe_engine *e, int32_t nknot, double r_lo, double r_hi, const double *c) {
  if (!e) return MFE_E_BAD_PARAM;
  if (!e->have_particles) return fail(e, MFE_E_STATE, "mfe_set_particles first");
  if (nknot == 0) {
    e->nknot = 0;
    touch(e);
    return MFE_OK;
  }
  if (nknot < 4 || nknot > 1024 || !c || !(r_hi > r_lo) || !(r_lo >= 0.0))
    return fail(e, MFE_E_BAD_PARAM, "knots: need 4 <= nknot <= 1024, 0 <= r_lo < r_hi, coefficients");
  return guarded(e, [&]() -> int {
    const size_t n = (size_t)e->nptpt * nknot;
    if (e->nknot != nknot || !e->knot_c) dalloc(e->knot_c, n);
    upload(e->knot_c, c, n, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->nknot = nknot;
    e->knot_r0 = r_lo;
    e->knot_dinv = (nknot - 3) / (r_hi - r_lo);
    touch(e);
    return MFE_OK;
  });
}

int mfe_adjoint(mfe_engine *e, double *dE_dq, double *dE_dzat, double *dE_dsigat,
                double *dE_depsat, double *dE_dknot, double *dE_dro, double *energy) {
  if (!e) return MFE_E_BAD_PARAM;
  e->forces_current = false;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->have_ewald) return fail(e, MFE_E_STATE, "mfe_set_ewald not called");
  if (!e->have_pp) return fail(e, MFE_E_STATE, "mfe_set_pair_params not called");
  if (e->ewald == MFE_EWALD_STD)
    return fail(e, MFE_E_BAD_PARAM, "the adjoint covers SPME and no-Ewald systems");
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    const int ncol = 1 + 2 * e->npt + e->nptpt * e->nknot;
    const int nb = (e->np + 255) / 256;
    if (!e->adj_phi || e->adj_cap < nb * ncol) {
      dalloc(e->phi_rec, (size_t)e->max_np);
      dalloc(e->adj_phi, (size_t)e->max_np);
      dalloc(e->adj_dq, (size_t)e->max_np);
      dalloc(e->adj_part, (size_t)nb * ncol);
      dalloc(e->adj_out, (size_t)ncol);
      e->adj_cap = nb * ncol;
      touch(e);
    }
    sync_view(e);
    const mfe::Group &G = e->grp;
    const unsigned f = MFE_FORCES | MFE_ENERGY | MFE_VIRIAL;
    mfe::launch_pair(G, true);
    if (G.ewald == MFE_EWALD_SPME) mfe::launch_spme(G, true, SP_BIN, SP_GATHER, mfe::GM_POT);
    if (G.chains) mfe::launch_bonds(G);
    if (G.surf) mfe::launch_surf(G);
    mfe::launch_reduce(G, f, false, false);
    mfe::launch_finalize(G, 0, 0.0);
    mfe::launch_adjoint(G, ncol, e->adj_out);
    std::vector<double> out(ncol), dq(e->np), fo(3 * (size_t)e->np);
    MFE_HIP_TRY(hipMemcpyAsync(out.data(), e->adj_out, ncol * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipMemcpyAsync(dq.data(), e->adj_dq, e->np * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    MFE_HIP_TRY(hipMemcpyAsync(fo.data(), e->forceo, fo.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    fetch_energy(e);
    int r = read_flags(e);
    if (r) return r;
    if (dE_dq) std::copy(dq.begin(), dq.end(), dE_dq);
    if (dE_dzat) {
      for (int t = 0; t < e->npt; ++t) dE_dzat[t] = 0.0;
      for (int i = 0; i < e->np; ++i) dE_dzat[e->h_type[i]] += dq[i];
    }
    if (dE_dsigat) for (int t = 0; t < e->npt; ++t) dE_dsigat[t] = out[1 + t];
    if (dE_depsat) for (int t = 0; t < e->npt; ++t) dE_depsat[t] = out[1 + e->npt + t];
    if (dE_dknot) std::copy(out.begin() + 1 + 2 * e->npt, out.end(), dE_dknot);
    if (dE_dro) {
      // position adjoint: dE/dr = -force (UTotal's force before the unit scaling)
      const double ffac = e->h_sys.ffac;
      for (size_t k = 0; k < fo.size(); ++k) dE_dro[k] = -fo[k] / ffac;
    }
    // the model energy: exact pair sum + reciprocal (with self/surface) + bonds
    if (energy) *energy = out[0] + e->h_en[2] + e->h_en[3];
    return MFE_OK;
  });
}

namespace {

// device room of the force-matching passes for the engine's np
void fm_buffers(mfe_engine *e) {
  const int np = e->np;
  const int ncol = 1 + 3 * e->npt + e->nptpt * e->nknot;
  const int ncolp = 2 * e->npt + e->nptpt * e->nknot;
  const int nb = (np + 255) / 256, nb4 = mfe::fm_pair_blocks(np);
  const size_t part = (size_t)nb * 4 + (size_t)nb * e->npt + (size_t)nb4 * std::max(ncolp, 1);
  bool changed = false;
  if (e->fm_np != np || !e->fm_fld) {
    dalloc(e->fm_fld, 3 * (size_t)np);
    dalloc(e->fm_res, 3 * (size_t)np);
    dalloc(e->fm_res_s, 3 * (size_t)np);
    dfree(e->fm_vec);
    MFE_HIP_TRY(hipMalloc(&e->fm_vec, 4 * (size_t)np * sizeof(double)));
    dalloc(e->phi_rec, (size_t)e->max_np);
    dalloc(e->adj_phi, (size_t)e->max_np);
    dalloc(e->adj_dq, (size_t)e->max_np);
    dalloc(e->fm_qr, 3);
    e->fm_np = np;
    changed = true;
  }
  if (part > e->fm_part_cap) {
    dalloc(e->fm_part, part);
    e->fm_part_cap = part;
  }
  if (ncol > e->fm_acc_cap) {
    dalloc(e->fm_acc, (size_t)ncol);
    e->fm_acc_cap = ncol;
  }
  if (changed) touch(e);
}

// the model's forces at the engine's positions (list current): pair model,
// SPME with the reciprocal field kept per particle, bonds and angles, the
// surface term; forceo = force * ffac as UTotal leaves it
void fm_forward(mfe_engine *e) {
  const mfe::Group &G = e->grp;
  mfe::launch_fm_pair(G, false, nullptr, 0, e->np);
  if (G.ewald == MFE_EWALD_SPME) mfe::launch_spme(G, false, SP_BIN, SP_GATHER, mfe::GM_FIELD);
  if (G.chains) mfe::launch_bonds(G);
  if (G.surf) mfe::launch_surf(G);
  mfe::launch_reduce(G, 0u, false, false);
  mfe::launch_finalize(G, 0, 0.0);
}

int fm_check(mfe_engine *e) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->have_ewald) return fail(e, MFE_E_STATE, "mfe_set_ewald not called");
  if (!e->have_pp) return fail(e, MFE_E_STATE, "mfe_set_pair_params not called");
  if (e->ewald == MFE_EWALD_STD)
    return fail(e, MFE_E_BAD_PARAM, "force matching covers SPME and no-Ewald systems");
  return MFE_OK;
}

}  // namespace

int mfe_model_forces(mfe_engine *e, double *force) {
  if (!e || !force) return MFE_E_BAD_PARAM;
  e->forces_current = false;
  int rc = fm_check(e);
  if (rc) return rc;
  if (!e->list_valid) {
    rc = mfe_build_list(e);
    if (rc) return rc;
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    e->mc_cached = false;
    fm_buffers(e);
    sync_view(e);
    fm_forward(e);
    std::vector<double> fo(3 * (size_t)e->np);
    MFE_HIP_TRY(hipMemcpyAsync(fo.data(), e->forceo, fo.size() * sizeof(double),
                               hipMemcpyDeviceToHost, e->stream));
    int r = read_flags(e);
    if (r) return r;
    const double inv = 1.0 / e->h_sys.ffac;
    for (size_t k = 0; k < fo.size(); ++k) force[k] = fo[k] * inv;
    return MFE_OK;
  });
}

int mfe_fm_set_frames(mfe_engine *e, int32_t nframes, const double *ro, const double *f_target) {
  if (!e || nframes < 1 || !ro || !f_target) return MFE_E_BAD_PARAM;
  e->forces_current = false;
  if (!e->have_particles || !e->have_box) return fail(e, MFE_E_STATE, "box and particles first");
  const size_t n3 = 3 * (size_t)e->np;
  for (size_t k = 0; k < n3 * nframes; ++k) {
    const int m = (int)(k % 3);
    if (!(std::fabs(ro[k]) <= e->box.L2[m]))
      return fail(e, MFE_E_BAD_PARAM, "frame position outside the box [-L/2, L/2]");
    if (!std::isfinite(f_target[k])) return fail(e, MFE_E_BAD_PARAM, "target not finite");
  }
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    if (nframes != e->fm_nframes || e->fm_frame_np != e->np) {
      dalloc(e->fm_ro, n3 * nframes);
      dalloc(e->fm_tgt, n3 * nframes);
    }
    upload(e->fm_ro, ro, n3 * nframes, e->stream);
    upload(e->fm_tgt, f_target, n3 * nframes, e->stream);
    MFE_HIP_TRY(hipStreamSynchronize(e->stream));
    e->fm_nframes = nframes;
    e->fm_frame_np = e->np;
    return MFE_OK;
  });
}

int mfe_fm_loss_grad(mfe_engine *e, double *loss, double *dL_dzat, double *dL_dsigat,
                     double *dL_depsat, double *dL_dknot) {
  if (!e) return MFE_E_BAD_PARAM;
  e->forces_current = false;
  int rc = fm_check(e);
  if (rc) return rc;
  if (e->fm_nframes < 1 || e->fm_frame_np != e->np)
    return fail(e, MFE_E_STATE, "mfe_fm_set_frames first (for this np)");
  return guarded(e, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(e->device));
    e->mc_cached = false;
    fm_buffers(e);
    sync_view(e);
    const mfe::Group &G = e->grp;
    const int np = e->np, npt = e->npt;
    const int ncol = 1 + 3 * npt + e->nptpt * e->nknot;
    const int ncolp = 2 * npt + e->nptpt * e->nknot;
    const int nb = (np + 255) / 256, nb4 = mfe::fm_pair_blocks(np);
    double *part_r = e->fm_part;                       // [nb][4]
    double *part_q = part_r + (size_t)nb * 4;          // [nb][npt]
    double *part_p = part_q + (size_t)nb * npt;        // [nb4][ncolp]
    const size_t n3 = 3 * (size_t)np;
    mfe::launch_fill_double(e->fm_acc, ncol, 0.0, e->stream);
    for (int f = 0; f < e->fm_nframes; ++f) {
      MFE_HIP_TRY(hipMemcpyAsync(e->ro, e->fm_ro + n3 * f, n3 * sizeof(double),
                                 hipMemcpyDeviceToDevice, e->stream));
      mfe::launch_nlist_rebuild(G, false);
      fm_forward(e);
      mfe::launch_fm_resid(G, e->fm_tgt + n3 * f, part_r);
      mfe::launch_fm_acc(G, part_r, nb, 4, 1, e->fm_acc);
      mfe::launch_fm_pair(G, true, part_p, ncolp, np);
      mfe::launch_fm_acc(G, part_p, nb4, ncolp, ncolp, e->fm_acc + 1 + npt);
      if (G.ewald == MFE_EWALD_SPME) mfe::launch_fm_rec_backward(G);
      mfe::launch_fm_q(G, part_r, e->fm_qr, part_q);
      mfe::launch_fm_acc(G, part_q, nb, npt, npt, e->fm_acc + 1);
    }
    std::vector<double> out(ncol);
    MFE_HIP_TRY(hipMemcpyAsync(out.data(), e->fm_acc, ncol * sizeof(double),
                               hipMemcpyDeviceToHost, e->stream));
    int r = read_flags(e);
    e->list_valid = (r == MFE_OK);
    e->have_positions = true;
    if (r) return r;
    if (loss) *loss = out[0];
    if (dL_dzat) std::copy(out.begin() + 1, out.begin() + 1 + npt, dL_dzat);
    if (dL_dsigat) std::copy(out.begin() + 1 + npt, out.begin() + 1 + 2 * npt, dL_dsigat);
    if (dL_depsat) std::copy(out.begin() + 1 + 2 * npt, out.begin() + 1 + 3 * npt, dL_depsat);
    if (dL_dknot) std::copy(out.begin() + 1 + 3 * npt, out.end(), dL_dknot);
    return MFE_OK;
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
        const int jp = sorted[(unsigned)nbr[(size_t)k * np + t] & mfe::kNbrIdxMask] + 1;
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
  e->forces_current = false;
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
    sync_view(e);
    const mfe::Group &G = e->grp;
    mfe_launch_utotal(G, MFE_ENERGY, 0, 0.0, false, false);
    hipEvent_t a, b;
    MFE_HIP_TRY(hipEventCreate(&a));
    MFE_HIP_TRY(hipEventCreate(&b));
    MFE_HIP_TRY(hipEventRecord(a, e->stream));
    for (int r = 0; r < nrep; ++r) {
      switch (stage) {
        case MFE_STAGE_PAIR: mfe::launch_pair(G, true, false); break;
        case MFE_STAGE_BIN: mfe::launch_spme(G, true, SP_BIN, SP_BIN); break;
        case MFE_STAGE_SPREAD: mfe::launch_spme(G, true, SP_SPREAD, SP_SPREAD); break;
        case MFE_STAGE_FFT_X: mfe::launch_spme(G, true, SP_FFTX, SP_FFTX); break;
        case MFE_STAGE_FFT_Y: mfe::launch_spme(G, true, SP_FFTY, SP_FFTY); break;
        case MFE_STAGE_CONV: mfe::launch_spme(G, true, SP_CONV, SP_CONV); break;
        case MFE_STAGE_IFFT: mfe::launch_spme(G, true, SP_IFFTY, SP_IFFTX); break;
        case MFE_STAGE_GATHER: mfe::launch_spme(G, true, SP_GATHER, SP_GATHER); break;
        case MFE_STAGE_NLIST: mfe::launch_nlist_rebuild(G, false); break;
        case MFE_STAGE_STEP: mfe_launch_utotal(G, MFE_ENERGY, 0, 0.0, false, false); break;
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
    mfe_launch_utotal(G, MFE_ENERGY, 0, 0.0, false, false);
    return read_flags(e);
  });
}

}  // extern "C"
