// mfe_batch.hip -- replicas driven together.  A batch holds engines that
// were set up one by one (each keeps its own device buffers); every stage
// of UTotal, VelVer and BDStep is then one launch over all replicas, the
// replica taken from blockIdx.z, with grids sized for the largest replica.
// A replica computes exactly what its engine computes alone: the kernels
// and the per-replica block partition of every reduction are the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

#include "mfe_internal.h"

struct mfe_batch {
  std::vector<mfe_engine *> eng;
  std::vector<unsigned> ver;
  int device = 0;
  hipStream_t stream = nullptr;
  mfe::Side side;
  std::string err;
  mfe::SysDev *d_sys = nullptr;
  mfe::Group grp{};
  mfe::StepGraphs graph;
  int graph_kind = 0;  // 1 = MD, 2 = BD
  double graph_dt = -1.0;
  unsigned graph_flags = 0;
  int *d_flags = nullptr;  // collected [R][4]
  double *d_en = nullptr;  // collected [R][8 + MFE_MAX_PAIRS]
  int *h_flags = nullptr;
  double *h_en = nullptr;
  int failed = -1;
  // persistent BD: replica order (costliest first), queue counter, and
  // the captured graph of the launch (queue reset + kernel) for pgraph_key
  bool persist = true;
  int *d_order = nullptr, *d_queue = nullptr;
  hipGraphExec_t pgraph = nullptr;
  long long pgraph_key = -1;
};

namespace {

constexpr int kNen = 8 + MFE_MAX_PAIRS;

int bfail(mfe_batch *b, int code, const std::string &msg, int replica = -1) {
  if (b) {
    b->err = msg;
    b->failed = replica;
  }
  return code;
}

template <typename F>
int bguarded(mfe_batch *b, F &&f) {
  try {
    return f();
  } catch (const mfe::HipError &h) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)h.code,
                  hipGetErrorString(h.code), h.expr, h.file, h.line);
    return bfail(b, MFE_E_HIP, buf);
  } catch (const std::bad_alloc &) {
    return bfail(b, MFE_E_NOMEM, "host allocation failed");
  }
}

bool ready(const mfe_engine *e) {
  return e->have_box && e->have_particles && e->have_tables && e->have_neighbor &&
         e->have_positions && e->have_ewald;
}

// rebuild the device views and the group when any replica changed
int refresh(mfe_batch *b) {
  const int R = (int)b->eng.size();
  bool changed = !b->d_sys;
  for (int r = 0; r < R; ++r)
    if (b->eng[r]->version != b->ver[r] || b->eng[r]->sys_dirty) changed = true;
  if (!changed) return MFE_OK;
  std::vector<mfe::SysDev> h(R);
  for (int r = 0; r < R; ++r) {
    mfe_sync_view(b->eng[r]);
    h[r] = b->eng[r]->h_sys;
    b->ver[r] = b->eng[r]->version;
  }
  if (!b->d_sys) MFE_HIP_TRY(hipMalloc((void **)&b->d_sys, R * sizeof(mfe::SysDev)));
  MFE_HIP_TRY(hipMemcpyAsync(b->d_sys, h.data(), R * sizeof(mfe::SysDev),
                             hipMemcpyHostToDevice, b->stream));
  MFE_HIP_TRY(hipStreamSynchronize(b->stream));
  std::string err;
  int rc = mfe_group_build(b->eng.data(), R, b->d_sys, b->stream, b->grp, err);
  if (rc) return bfail(b, rc, err);
  b->grp.side = b->side;
  const size_t lds = std::max(b->grp.lds_spread, b->grp.lds_gather);
  if (lds > 32 * 1024) mfe_allow_big_spread_lds(b->grp.prec, (int)lds);
  b->graph.clear();
  // the persistent launch's order follows the replicas' sizes
  if (b->pgraph) (void)hipGraphExecDestroy(b->pgraph);
  b->pgraph = nullptr;
  b->pgraph_key = -1;
  if (b->d_order) (void)hipFree(b->d_order);
  if (b->d_queue) (void)hipFree(b->d_queue);
  b->d_order = b->d_queue = nullptr;
  return MFE_OK;
}

// flags and energies of every replica in one copy; the first replica with
// an error bit names the failure
int check(mfe_batch *b) {
  const int R = (int)b->eng.size();
  mfe::launch_collect(b->grp, b->d_flags, b->d_en, kNen);
  MFE_HIP_TRY(hipMemcpyAsync(b->h_flags, b->d_flags, 4 * R * sizeof(int),
                             hipMemcpyDeviceToHost, b->stream));
  MFE_HIP_TRY(hipMemcpyAsync(b->h_en, b->d_en, (size_t)kNen * R * sizeof(double),
                             hipMemcpyDeviceToHost, b->stream));
  MFE_HIP_TRY(hipStreamSynchronize(b->stream));
  int first = MFE_OK;
  for (int r = 0; r < R; ++r) {
    const int bits = b->h_flags[4 * r];
    if (!bits) continue;
    const int rc = mfe_map_flag_error(b->eng[r], bits);
    if (rc && first == MFE_OK) {
      first = rc;
      char buf[64];
      std::snprintf(buf, sizeof buf, "replica %d: ", r);
      bfail(b, rc, buf + b->eng[r]->err, r);
    }
  }
  return first;
}

int build_lists(mfe_batch *b, bool only_stale) {
  bool any = false;
  for (auto *e : b->eng)
    if (!only_stale || !e->list_valid) any = true;
  if (!any) return MFE_OK;
  mfe::launch_nlist_rebuild(b->grp, false);
  int rc = check(b);
  if (rc) return rc;
  for (auto *e : b->eng) e->list_valid = true;
  return MFE_OK;
}

int prepare(mfe_batch *b) {
  b->failed = -1;
  for (size_t r = 0; r < b->eng.size(); ++r)
    if (!ready(b->eng[r])) {
      char buf[96];
      std::snprintf(buf, sizeof buf, "replica %d is not fully set up", (int)r);
      return bfail(b, MFE_E_STATE, buf, (int)r);
    }
  MFE_HIP_TRY(hipSetDevice(b->device));
  int rc = refresh(b);
  if (rc) return rc;
  return build_lists(b, true);
}

}  // namespace

extern "C" {

int mfe_batch_create(mfe_engine *const *engines, int32_t nrep, mfe_batch **out) {
  if (!engines || !out || nrep < 1 || nrep > 65535) return MFE_E_BAD_PARAM;
  *out = nullptr;
  for (int r = 0; r < nrep; ++r)
    if (!engines[r]) return MFE_E_BAD_PARAM;
  mfe_batch *b = new (std::nothrow) mfe_batch();
  if (!b) return MFE_E_NOMEM;
  int rc = bguarded(b, [&]() -> int {
    b->eng.assign(engines, engines + nrep);
    b->ver.assign(nrep, 0u);
    b->device = engines[0]->device;
    for (int r = 1; r < nrep; ++r)
      if (engines[r]->device != b->device)
        return bfail(b, MFE_E_BAD_PARAM, "replicas on different devices", r);
    MFE_HIP_TRY(hipSetDevice(b->device));
    MFE_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    MFE_HIP_TRY(mfe::side_create(b->side));
    MFE_HIP_TRY(hipMalloc((void **)&b->d_flags, 4 * (size_t)nrep * sizeof(int)));
    MFE_HIP_TRY(hipMalloc((void **)&b->d_en, (size_t)kNen * nrep * sizeof(double)));
    MFE_HIP_TRY(hipHostMalloc((void **)&b->h_flags, 4 * (size_t)nrep * sizeof(int)));
    MFE_HIP_TRY(hipHostMalloc((void **)&b->h_en, (size_t)kNen * nrep * sizeof(double)));
    for (int r = 0; r < nrep; ++r)
      if (!ready(engines[r])) return bfail(b, MFE_E_STATE, "replica not fully set up", r);
    return refresh(b);
  });
  if (rc != MFE_OK) {
    // keep the message for the caller: the batch is returned only on success
    std::fprintf(stderr, "mfe_batch_create: %s\n", b->err.c_str());
    mfe_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MFE_OK;
}

int mfe_batch_destroy(mfe_batch *b) {
  if (!b) return MFE_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->graph.clear();
  if (b->d_sys) (void)hipFree(b->d_sys);
  if (b->d_flags) (void)hipFree(b->d_flags);
  if (b->d_en) (void)hipFree(b->d_en);
  if (b->pgraph) (void)hipGraphExecDestroy(b->pgraph);
  if (b->d_order) (void)hipFree(b->d_order);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  if (b->h_en) (void)hipHostFree(b->h_en);
  if (b->side.stream) (void)hipStreamSynchronize(b->side.stream);
  mfe::side_destroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MFE_OK;
}

const char *mfe_batch_last_error(const mfe_batch *b) { return b ? b->err.c_str() : "null batch"; }

int32_t mfe_batch_failed_replica(const mfe_batch *b) { return b ? b->failed : -1; }

int mfe_batch_build_list(mfe_batch *b) {
  if (!b) return MFE_E_BAD_PARAM;
  return bguarded(b, [&]() -> int {
    int rc = prepare(b);
    if (rc) return rc;
    return build_lists(b, false);
  });
}

int mfe_batch_compute(mfe_batch *b, uint32_t flags) {
  if (!b) return MFE_E_BAD_PARAM;
  return bguarded(b, [&]() -> int {
    int rc = prepare(b);
    if (rc) return rc;
    mfe_launch_utotal(b->grp, flags | MFE_ENERGY | MFE_VIRIAL, 0, 0.0, false, false);
    rc = check(b);
    for (mfe_engine *e : b->eng) e->forces_current = rc == MFE_OK;
    return rc;
  });
}

// The persistent BD launch takes replicas on the mixed path with SPME and
// no chains, surface term or spline knots, whose LDS needs fit a CU.
namespace {

bool persist_eligible(const mfe_batch *b, size_t *lds) {
  if (!b->persist) return false;
  const mfe::Group &g = b->grp;
  if (g.prec != MFE_PREC_MIXED || g.ewald != MFE_EWALD_SPME || g.chains || g.surf ||
      g.knots || g.order < 3 || g.order > 8)
    return false;
  // the pair tables, and one x row, y plane and z line of every replica's
  // mesh with its twiddles, fit the launch's LDS
  const size_t L = mfe::kPersistLds, c = sizeof(mfe::cplx<float>), t = sizeof(mfe::Twiddle<float>);
  if (g.lds_pair > L) return false;
  for (const mfe_engine *e : b->eng) {
    const size_t s1 = e->s[0], s2 = e->s[1], s3 = e->s[2], hx = s1 / 2 + 1;
    if (s1 * s2 * sizeof(unsigned long long) > L) return false;   // a spreading slab
    if (2 * hx * c + (s1 / 2 + s1) * t > L) return false;
    if (2 * hx * (s2 + 1) * c + s2 * t > L) return false;
    if (2 * hx * (s3 + 1) * c + s3 * t > L) return false;
  }
  *lds = L;
  return mfe::persist_allow_lds(g.order, L);
}

int run_persistent(mfe_batch *b, int32_t nsteps, uint32_t flags, size_t lds) {
  const int R = (int)b->eng.size();
  if (!b->d_order) {
    // costliest replicas first, so the queue's tail holds the small ones
    std::vector<std::pair<double, int>> cost(R);
    for (int r = 0; r < R; ++r) {
      const mfe_engine *e = b->eng[r];
      const double m = (double)e->s[0] * e->s[1] * e->s[2];
      cost[r] = {-(8.0 * e->np + m), r};
    }
    std::stable_sort(cost.begin(), cost.end());
    std::vector<int> ord(R);
    for (int r = 0; r < R; ++r) ord[r] = cost[r].second;
    MFE_HIP_TRY(hipMalloc((void **)&b->d_order, R * sizeof(int)));
    MFE_HIP_TRY(hipMalloc((void **)&b->d_queue, sizeof(int)));
    MFE_HIP_TRY(hipMemcpyAsync(b->d_order, ord.data(), R * sizeof(int), hipMemcpyHostToDevice,
                               b->stream));
  }
  const int ncu = std::max(b->eng[0]->ncu, 1);
  const int blocks = std::min(R, ncu);
  const int energy = (flags & (MFE_ENERGY | MFE_VIRIAL)) != 0 ? 1 : 0;
  // the queue reset and the launch, captured once per (nsteps, energy)
  const long long key = 2LL * nsteps + energy;
  if (!b->pgraph || b->pgraph_key != key) {
    if (b->pgraph) (void)hipGraphExecDestroy(b->pgraph);
    b->pgraph = nullptr;
    hipGraph_t gph;
    MFE_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    mfe::launch_bd_persistent(b->grp, b->d_order, R, nsteps, energy, b->d_queue, blocks, lds);
    MFE_HIP_TRY(hipStreamEndCapture(b->stream, &gph));
    hipError_t ie = hipGraphInstantiate(&b->pgraph, gph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gph);
    MFE_HIP_TRY(ie);
    b->pgraph_key = key;
  }
  MFE_HIP_TRY(hipGraphLaunch(b->pgraph, b->stream));
  return check(b);
}

}  // namespace

static int run_graph(mfe_batch *b, int kind, int32_t nsteps, double tstep, uint32_t flags) {
  if (b->graph_kind != kind || b->graph_flags != flags || (kind == 1 && b->graph_dt != tstep)) {
    b->graph.clear();
    b->graph_kind = kind;
    b->graph_dt = tstep;
    b->graph_flags = flags;
  }
  MFE_HIP_TRY(b->graph.run(b->stream, nsteps, [&]() {
    if (kind == 1) mfe_launch_md_step(b->grp, tstep, flags);
    else mfe_launch_bd_step(b->grp, 2, flags);
  }));
  return check(b);
}

int mfe_batch_step_md(mfe_batch *b, int32_t nsteps, double tstep, uint32_t flags) {
  if (!b || nsteps < 0) return MFE_E_BAD_PARAM;
  for (mfe_engine *e : b->eng) e->mc_cached = false;
  return bguarded(b, [&]() -> int {
    int rc = prepare(b);
    if (rc) return rc;
    rc = run_graph(b, 1, nsteps, tstep, flags);
    for (mfe_engine *e : b->eng) e->forces_current = rc == MFE_OK;
    return rc;
  });
}

int mfe_batch_step_bd(mfe_batch *b, int32_t nsteps, uint32_t flags) {
  if (!b || nsteps < 0) return MFE_E_BAD_PARAM;
  for (mfe_engine *e : b->eng) e->mc_cached = false;
  return bguarded(b, [&]() -> int {
    for (size_t r = 0; r < b->eng.size(); ++r)
      if (!b->eng[r]->have_bd) return bfail(b, MFE_E_STATE, "BD factors not set (mfe_set_bd)", (int)r);
    int rc = prepare(b);
    if (rc) return rc;
    bool current = true;
    for (mfe_engine *e : b->eng) current = current && e->forces_current;
    if (!current) {
      // the first move takes the forces at the current positions (the
      // reference's UTotal before its BD loop)
      mfe_launch_utotal(b->grp, 0u, 0, 0.0, false, false);
      rc = check(b);
      if (rc) return rc;
    }
    size_t lds = 0;
    if (nsteps > 0 && persist_eligible(b, &lds)) rc = run_persistent(b, nsteps, flags, lds);
    else rc = run_graph(b, 2, nsteps, 0.0, flags);
    for (mfe_engine *e : b->eng) e->forces_current = rc == MFE_OK;
    return rc;
  });
}

int mfe_batch_set_persistent(mfe_batch *b, int32_t on) {
  if (!b) return MFE_E_BAD_PARAM;
  b->persist = on != 0;
  return MFE_OK;
}

int32_t mfe_batch_uses_persistent(mfe_batch *b) {
  if (!b) return 0;
  size_t lds = 0;
  int32_t yes = 0;
  const int rc = bguarded(b, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(b->device));
    int rc2 = refresh(b);
    if (rc2) return rc2;
    yes = persist_eligible(b, &lds) ? 1 : 0;
    return MFE_OK;
  });
  return rc == MFE_OK ? yes : 0;
}

int mfe_batch_get_energies(mfe_batch *b, mfe_energy *out) {
  if (!b || !out) return MFE_E_BAD_PARAM;
  return bguarded(b, [&]() -> int {
    MFE_HIP_TRY(hipSetDevice(b->device));
    int rc = refresh(b);
    if (rc) return rc;
    mfe::launch_collect(b->grp, b->d_flags, b->d_en, kNen);
    MFE_HIP_TRY(hipMemcpyAsync(b->h_flags, b->d_flags, 4 * b->eng.size() * sizeof(int),
                               hipMemcpyDeviceToHost, b->stream));
    MFE_HIP_TRY(hipMemcpyAsync(b->h_en, b->d_en, (size_t)kNen * b->eng.size() * sizeof(double),
                               hipMemcpyDeviceToHost, b->stream));
    MFE_HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t r = 0; r < b->eng.size(); ++r) {
      const double *h = b->h_en + (size_t)kNen * r;
      mfe_energy &en = out[r];
      en = mfe_energy{};
      en.tot = h[0];
      en.twob0 = h[1];
      en.rec = h[2];
      en.bond = h[3];
      en.virial = h[4];
      en.nptpt = b->eng[r]->nptpt;
      en.rebuilds = b->h_flags[4 * r + 2];
      for (int k = 0; k < b->eng[r]->nptpt; ++k) en.twob[k] = h[8 + k];
    }
    return MFE_OK;
  });
}

}  // extern "C"
