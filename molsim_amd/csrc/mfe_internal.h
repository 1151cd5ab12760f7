// mfe_internal.h -- engine state and kernel interfaces shared by the
// translation units of libmfe (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "mfe.h"

// Pointers that kernels read out of a device view (SysDev) are global
// memory, but the compiler cannot know that for a pointer loaded from
// memory and would emit flat loads and stores (which also wait on the LDS
// counter).  On the device pass they are declared in the global address
// space; the host stores plain addresses (same 64-bit layout) through
// MFE_SETP.
#if defined(__HIP_DEVICE_COMPILE__)
#define MFE_DPTR(T) __attribute__((address_space(1))) T *
#else
#define MFE_DPTR(T) T *
#endif
#define MFE_SETP(dst, src) ((dst) = (std::remove_reference_t<decltype(dst)>)(src))

#define MFE_HIP_TRY(expr)                                                    \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) throw mfe::HipError(_e, #expr, __FILE__, __LINE__); \
  } while (0)

namespace mfe {

// an engine's energy vector: [tot, twob0, rec, bond, virial, S(3),
// twob(1..MFE_MAX_PAIRS), angle]
constexpr int kEnAngle = 8 + MFE_MAX_PAIRS;
constexpr int kEnLen = kEnAngle + 1;

// lanes of the pair kernel that share one particle's neighbour list
constexpr int kPairLanes = 4;

struct HipError {
  hipError_t code;
  const char *expr;
  const char *file;
  int line;
  HipError(hipError_t c, const char *e, const char *f, int l)
      : code(c), expr(e), file(f), line(l) {}
};

// error bits raised by kernels (read back by the host after a call)
enum : int {
  ERR_OVERLAP = 1,    // r2 < r2umin
  ERR_TABLE = 2,      // no table interval for r2
  ERR_CAPACITY = 4,   // neighbour capacity exceeded
  ERR_SYNC = 8,       // a grid barrier of the fused list rebuild timed out
  ERR_BIN = 16,       // SPME brick counts that do not add up to the particles
};

template <typename R>
struct cplx {
  R x, y;
};

// A twiddle factor.  FP64: the value.  FP32: its rounded value and the
// rounding residual (w - (float)w), which the products add back by FMA:
// rounded twiddles alone have |w| < 1 on average for power-of-two sizes
// (cos(pi/4) rounds down, for one), and every transform would shrink its
// output by ~2e-8 per axis -- a bias the reciprocal energy would keep.
template <typename R>
struct Twiddle;
template <>
struct Twiddle<double> {
  double x, y;
};
template <>
struct Twiddle<float> {
  float x, y, lx, ly;
};

// Orthorhombic periodic box (Src/molaux.F90:326-335); dpbc == boxlen.
struct Box {
  double L[3], L2[3], Li[3];
};

// Cell grid used to find Verlet-list partners: edge >= rcut + drnlist.
// The mixed path's pair kernel reads positions relative to a reference
// point of the particle's cell, o(c) = -L/2 + (c + 1/2) hf (hf the cell
// edge rounded to FP32, so that neighbouring cells differ by exactly hf),
// in FP32; each list entry carries the cell offset of its pair, one of
// {-W, -hf, 0, hf, W} per axis, W = L - (n-1) hf for pairs across the
// periodic boundary (off[axis][code + 2]).  Axes of one cell carry code 0
// and take the minimum image in FP32.
struct CellGeom {
  int n[3];          // cells per axis (1 when the axis has < 3 cells)
  double cinv[3];    // cells per unit length
  int ncells;
  double hf[3];      // cell edge, an FP32 value
  float off[3][5];   // pair offsets per axis and code
  float L[3], L2[3]; // box edge and half edge in FP32
};
// list entries: cell-order partner position in the low 25 bits, the pair's
// offset code (cx+2) + 5 (cy+2) + 25 (cz+2) in the high 7 bits
constexpr unsigned kNbrIdxMask = (1u << 25) - 1;
constexpr int kNbrCodeShift = 25;
constexpr int kMaxParticles = 1 << 25;

// Table intervals in device layout: for pair p, intervals k = 0..nint[p]-1
// start at index off[p]; zlow[off+k] (FP64, decreasing in k) and the 11
// coefficients coef[(off+k)*11 + c] (usum: c = 0..5, fsum: c = 6..10).
// Per type pair: what the pair kernel needs before the interval search.
struct alignas(16) PairRec {
  double r2at, r2um;  // hard core below r2at; overlap error below r2um
  double lo, pad;     // lowest interval bound zlow(nint - 1)
  float s_lo, s_inv;  // bin of r2: (1/r2 - s_lo) * s_inv
  int off, nint;      // the pair's intervals in zlow / coef
};
// the same in FP32 for the mixed path's pair kernel
struct alignas(16) PairRecF {
  float r2at, r2um, lo, s_lo;
  float s_inv;
  int off, nint, pad;
};
constexpr int kSlowBin = 0x8000;
// coefficients stored per interval (11 used, padded for 16-byte loads)
constexpr int kCoefStride = 12;

struct TableDev {
  MFE_DPTR(const double) zlow;
  MFE_DPTR(const double) coef;
  MFE_DPTR(const float) coef_f;
  MFE_DPTR(const int) off;
  MFE_DPTR(const int) nint;
  MFE_DPTR(const double) r2atat;
  MFE_DPTR(const double) r2umin;
  MFE_DPTR(const int) pairidx;  // [npt*npt] -> pair index 0..nptpt-1
  // bins of s = 1/r2 per pair (see PairRec): entry kb means the interval
  // of any r2 in the bin is kb or kb + 1; kb | kSlowBin means the bin
  // spans more intervals and the search walks from kb
  MFE_DPTR(const unsigned short) bin_k;  // [nptpt * nbin]
  MFE_DPTR(const PairRec) rec;           // [nptpt]
  MFE_DPTR(const PairRecF) rec_f;        // [nptpt]
  MFE_DPTR(const float) zlow_f;          // zlow in FP32
  int nbin;
  // what the pair kernel stages in LDS: 0 = the bins, 1 = also the
  // interval bounds, 2 = also the polynomial coefficients
  int lds_mode;
  int npt, nptpt;
  int nzl;             // intervals over all pairs (length of zlow)
  double rcut2;
};

// r2 bins per pair, the total over pairs bounded by kTableBins
constexpr int kTableBins = 4096;
constexpr int kTableBinsPerPair = 256;
// words of k_rebuild's grid barrier: kBarGroups group counters, then the
// top counter and the generation, each on a 128-byte line of its own
constexpr int kBarGroups = 8;
constexpr int kBarStride = 32;
constexpr int kBarTop = kBarGroups * kBarStride;
constexpr int kBarGen = kBarTop + kBarStride;
constexpr int kBarWords = kBarGen + kBarStride;
// threads per block of the pair kernel and its dynamic LDS budget
constexpr int kPairBlock = 512;
// blocks of the persistent pair kernel (two per CU)
constexpr int kPairGridMax = 512;
constexpr size_t kPairLdsBudget = 56 * 1024;
// dynamic LDS of the pair kernel: bins, then zlow, then coefficients,
// each part starting on 16 bytes
__host__ __device__ inline size_t pair_bin_bytes(int nptpt, int nbin) {
  return ((size_t)nptpt * nbin * 2 + 15) / 16 * 16;
}
__host__ __device__ inline size_t pair_zl_bytes(int nzl) {
  return ((size_t)nzl * 8 + 15) / 16 * 16;
}

// FFT plan of one transform length (mixed radix Stockham).
struct FftPlan {
  int n;
  int nstages;
  int radix[20];
};

// Per-axis SPME tables (FP64): g[n] = b2[n]*exp(-k_n^2/(4 alpha^2)) and
// kk[n] = k_n^2 for n = 0..s/2, so that
// meshfac(nx,ny,nz) = fac0 * gx*gy*gz / (kkx+kky+kkz)  (Essmann eq. 3.9).
struct SpmeAxes {
  MFE_DPTR(const double) g[3];
  MFE_DPTR(const double) kk[3];
  double fac0;   // 4*pi/V
  double expf;   // -1/(4 alpha^2)
};

// Parameters of Pot_1_6_12 (Src/potential.F90:1349-1425) that the adjoint
// differentiates: per-type charge, sigma and epsilon, Lorentz-Berthelot
// mixing, optional Ewald damping and WCA cut.
struct PairParams {
  double zat[MFE_MAX_TYPES], sigat[MFE_MAX_TYPES], epsat[MFE_MAX_TYPES];
  double efp, alpha, ljrcut, ljushift;
  int lewald, lljcut;
};

// SPME mesh geometry: mesh s, spreading bricks B (nb per axis), dkdr = s/L.
// Division by a divisor d fixed at set-up: q = (n * ceil(2^32 / d)) >> 32,
// exact for 0 <= n with n * d < 2^32 (the error term n * e / 2^32, e < d,
// stays below 1/d).  One 64-bit multiply instead of the ~25 instructions
// of a run-time integer division.
struct FastDiv {
  unsigned long long m;
  int d;
};
inline FastDiv make_fastdiv(int d) {
  return FastDiv{((1ull << 32) + (unsigned long long)d - 1) / (unsigned long long)d, d};
}
__host__ __device__ __forceinline__ int fdiv(int n, const FastDiv &f) {
  return (int)(((unsigned long long)(unsigned)n * f.m) >> 32);
}

// Blocks are dispatched to the 8 XCDs round-robin (block b on XCD b % 8),
// each XCD with its own L2.  This maps them to logical blocks so that each
// XCD takes one contiguous range, whose data (positions of a slab of the
// box, rows of neighbouring mesh lines) then stays in that XCD's L2.  A
// bijection of [0, nblk).
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  const int q = nblk >> 3, r = nblk & 7, x = b & 7, k = b >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

struct SpmeGeom {
  int s[3], B[3], nb[3], order;
  double dkdr[3];
  FastDiv fB[3];   // the brick edges B
  FastDiv fE[3];   // the extended tile edges B + order - 1
};

// Device-side view of one engine (one replica): every kernel reads the
// view of its replica, blockIdx.z, so a single engine and a batch of
// replicas run the same kernels -- a single engine is a batch of one.
// Per-replica block counts bound the work of each launch, whose grid is
// the maximum over the replicas.
struct SysDev {
  int np, npt, nptpt, nlsem;   // nlsem: 0 = vlist (r2 <= rcut2), 1 = cell list
  Box box;
  MFE_DPTR(double) ro;
  MFE_DPTR(double) rod;
  MFE_DPTR(double) rodd;
  MFE_DPTR(double) force;
  MFE_DPTR(double) forceo;
  MFE_DPTR(double) force_rec;  // reciprocal forces apart (fused pair + gather)
  MFE_DPTR(double) drosum;
  MFE_DPTR(double) r2disp;
  MFE_DPTR(int) type;
  MFE_DPTR(double) q;
  MFE_DPTR(double) massip;
  TableDev tab;
  // neighbour list
  double rs2, drnlist;
  int maxnn;
  CellGeom cg;
  MFE_DPTR(int) c_count;
  MFE_DPTR(int) c_start;
  MFE_DPTR(int) c_of;
  MFE_DPTR(int) c_slot;
  MFE_DPTR(int) c_sorted;
  MFE_DPTR(int) c_pos;          // inverse of c_sorted: particle -> cell-order position
  MFE_DPTR(double) ro_s;
  MFE_DPTR(int) type_s;
  // mixed path: (x, y, z) relative to the cell's reference point in FP32
  // and the type in the fourth word, in cell order (null on FP64 engines)
  MFE_DPTR(float4) pos_f;
  MFE_DPTR(int) nn;
  MFE_DPTR(int) nbr;
  MFE_DPTR(int) flags;        // [0] error bits, [1] rebuild flag, [2] rebuilds, [3] BD step
  // Ewald
  int ewald, lsurf;
  double alpha, efp, self_energy, vol;
  // neutralising-background term of a net charge Q (opt-in, 0 when off):
  // -pi Q^2 / (2 V alpha^2), its virial -3 times that, and dE/dq_i =
  // -pi Q / (V alpha^2) (all before the EpsiFourPi factor)
  double net_energy, net_dq;
  // SPME
  SpmeGeom sg;
  int hx, nbricks;
  MFE_DPTR(void) mesh_r;
  MFE_DPTR(void) mesh_c;
  MFE_DPTR(int) b_count;
  MFE_DPTR(int) b_done;    // count blocks finished (the last one scans)
  MFE_DPTR(int) b_start;
  MFE_DPTR(int) b_id;
  MFE_DPTR(int) b_slot;
  MFE_DPTR(void) rec_frac;
  MFE_DPTR(int2) rec_meta;
  MFE_DPTR(void) spread_tiles;   // [nbricks][(B+order-1)^3] spreading tiles
  SpmeAxes ax;
  FftPlan plan[3], plan_half;
  MFE_DPTR(const void) tw[3];
  MFE_DPTR(const void) tw_half;
  int fx_rpb, fx_nblk;   // x pass: rows per block, blocks
  int fy_T, fy_nbx;      // y pass: kx tile, tiles (grid y = s3)
  int fz_T, fz_nbx;      // conv pass: kx tile, tiles (grid y = s2)
  // fast divisors of the FFT index arithmetic: s1, s2, s3, hx = s1/2+1,
  // m = s1/2 and the y / conv tile widths
  FastDiv f_s1, f_s2, f_s3, f_hx, f_m, f_fyT, f_fzT;
  // standard Ewald
  MFE_DPTR(const double) k_vec;
  MFE_DPTR(double) k_sum;
  int nkvec;
  // chains
  int nchain, bond_p;
  MFE_DPTR(const int) ch_prev;
  MFE_DPTR(const int) ch_next;
  double bond_k, bond_eq;
  // UAngle: k already divided by (pi/180)^2 and p, eq in rad; ang_k == 0
  // leaves the angle term out, as the reference does
  double ang_k, ang_eq;
  int ang_p;
  // Brownian dynamics: fac1/fac2 per type, noise (host-supplied) or the
  // counter-based generator keyed by seed
  MFE_DPTR(const double) bd_fac1;
  MFE_DPTR(const double) bd_fac2;
  MFE_DPTR(const double) noise;
  unsigned long long seed;
  // additive cubic B-spline pair term on uniform knots in r (per pair
  // type, nknot coefficients, r in [knot_r0, knot_r0 + (nknot-3)/knot_dinv))
  int nknot;
  double knot_r0, knot_dinv;
  MFE_DPTR(const double) knot_c;
  // adjoint
  PairParams pp;
  MFE_DPTR(double) phi_rec;    // efp * reciprocal potential per particle (gather)
  MFE_DPTR(double) adj_phi;    // real-space dE/dq per particle
  MFE_DPTR(double) adj_dq;     // dE/dq per particle
  MFE_DPTR(double) adj_part;   // per-block partials [blocks][adj_ncol]
  int adj_ncol;
  // force matching: reciprocal field per particle (forward gather), the
  // residual force - target (particle and cell order), the target of the
  // frame in hand, and the residual spreading weights in brick order
  MFE_DPTR(double) fm_fld;
  MFE_DPTR(double) fm_res;
  MFE_DPTR(double) fm_res_s;
  MFE_DPTR(const double) fm_tgt;
  MFE_DPTR(void) fm_vec;
  int fm_ncol;
  // reductions and output scaling
  MFE_DPTR(double) part;
  MFE_DPTR(double) en;
  int off_pair, off_conv, off_bond, off_surf, off_std, off_move, off_skin;
  int nblk_np, nblk_pair, nblk_conv;
  // mixed pair kernel: lanes per particle (1, 2, 4) and whether the box
  // has an axis of fewer than three list cells (separations in FP64)
  int pair_lanes, pair_pos64;
  double ffac, surf_fac;
};

// Replicas launched together: same device, precision, Ewald kind and
// SPME order (a template parameter of the stencil kernels).
// A second stream and two events: a force evaluation forks the pair
// forces off the main stream and joins them before the stage that adds to
// them (under stream capture: two branches of the graph)
struct Side {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

hipError_t side_create(Side &s);
void side_destroy(Side &s);

// Step graphs: one step, and kGraphSteps steps captured back to back.  A
// run of nsteps replays the long graph nsteps / kGraphSteps times and the
// short one for the rest, so a graph launch's fixed cost (several us
// between replays) falls on every kGraphSteps-th step only.  The steps are
// static (the rebuild is gated on the device), so both graphs compute the
// same as launching the steps one by one.
constexpr int kGraphSteps = 8;
struct StepGraphs {
  hipGraphExec_t one = nullptr, many = nullptr;
  void clear() {
    if (one) (void)hipGraphExecDestroy(one);
    if (many) (void)hipGraphExecDestroy(many);
    one = many = nullptr;
  }
  // launch_step() enqueues one step on stream s (captured on first use)
  template <class F>
  hipError_t run(hipStream_t s, int nsteps, F &&launch_step) {
    const int nmany = nsteps / kGraphSteps, rest = nsteps % kGraphSteps;
    hipError_t rc = hipSuccess;
    if (nmany > 0 && !many) rc = capture(s, kGraphSteps, launch_step, &many);
    if (rc == hipSuccess && rest > 0 && !one) rc = capture(s, 1, launch_step, &one);
    for (int k = 0; rc == hipSuccess && k < nmany; ++k) rc = hipGraphLaunch(many, s);
    for (int k = 0; rc == hipSuccess && k < rest; ++k) rc = hipGraphLaunch(one, s);
    return rc;
  }

 private:
  template <class F>
  static hipError_t capture(hipStream_t s, int n, F &launch_step, hipGraphExec_t *out) {
    hipError_t rc = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (rc != hipSuccess) return rc;
    hipGraph_t g = nullptr;
    try {
      for (int k = 0; k < n; ++k) launch_step();
    } catch (...) {
      // leave the stream out of capture mode before the error goes on
      if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      throw;
    }
    rc = hipStreamEndCapture(s, &g);
    if (rc != hipSuccess) return rc;
    rc = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return rc;
  }
};

struct Group {
  int R = 0;
  const SysDev *d = nullptr;  // device array [R]
  hipStream_t stream = nullptr;
  Side side;                  // no stream: every launch on the main stream
  int prec = 0, ewald = 0, order = 0;
  bool chains = false, surf = false, cells = false;
  // maxima over the replicas of each launch's per-replica block counts
  int np_blocks = 0, pair_blocks = 0, cell_blocks = 0, bricks = 0, kvec = 0;
  int mesh_blocks = 0;   // blocks of 256 over the real mesh
  int fx_blocks = 0, fy_x = 0, fy_y = 0, fz_x = 0, fz_y = 0;
  size_t lds_spread = 0, lds_gather = 0, lds_fx = 0, lds_fy = 0, lds_fz = 0;
  size_t lds_pair = 0;
  int pair_npt = 0;         // most types of any replica
  // fused list rebuild (one engine): its flags, grid barrier state and CU
  // count
  int *flags = nullptr;
  unsigned *rebuild_bar = nullptr;
  int *rebuild_sum = nullptr;
  int ncu = 0;
  int sync_spin = -1;       // polls before a barrier wait gives up (-1: default)
  bool pair_cf_lds = true;  // every replica's coefficients fit in LDS
  bool knots = false;       // some replica carries the spline pair term
  // replicas whose mixed pair kernel forms separations from cell-relative
  // FP32 coordinates / in FP64 (boxes with an axis of < 3 list cells)
  bool pair_any32 = false, pair_any64 = false;
  int mc_s[3] = {0, 0, 0};  // mesh of a one-replica group (MC trial moves)
  // the pair forces and the gather in one launch (launch_pair_gather):
  // mixed SPME without bonds, surface term, spline knots or FP64
  // separations, coefficients in LDS, both LDS needs within the block limit
  bool fuse_pg = false;
};

}  // namespace mfe

// ------------------------------------------------------------------ engine
struct mfe_engine {
  int device = 0;
  int prec = MFE_PREC_MIXED;
  hipStream_t stream = nullptr;
  mfe::Side side;  // the pair forces' branch of a force evaluation
  std::string err;
  int max_np = 0;

  // system
  int np = 0, npt = 0, nptpt = 0;
  mfe::Box box{};
  bool have_box = false, have_particles = false, have_tables = false;
  bool have_ewald = false, have_neighbor = false, have_positions = false;
  bool list_valid = false;
  // forceo holds the forces at the current positions (a UTotal or a step
  // since the last change): a BD step's first move needs them
  bool forces_current = false;
  double self_energy = 0.0;  // -alpha/sqrt(pi) * sum q^2 (UEwaldSelf)
  bool net_charge = false;   // neutralising-background term on (opt-in)

  // particle state (device, xyz-interleaved FP64)
  double *ro = nullptr, *rod = nullptr, *rodd = nullptr;
  unsigned *rebuild_bar = nullptr;  // grid barrier of k_rebuild [kBarWords]
  int *rebuild_sum = nullptr;       // its per-block scan totals [ncu]
  int ncu = 0;                      // compute units of the device
  int sync_spin = -1;               // k_rebuild's barrier poll bound (-1: default)
  double *force = nullptr;   // UTotal's force(3,na) before unit scaling
  double *forceo = nullptr;  // force * fac
  double *force_rec = nullptr;  // the reciprocal forces of a fused pair + gather launch
  double *drosum = nullptr, *r2disp = nullptr;
  int *type = nullptr;       // 0-based
  double *q = nullptr;
  double *massip = nullptr;  // per particle inverse mass
  std::vector<double> h_massipt;
  std::vector<int> h_type;
  std::vector<double> h_q;

  // tables
  double *t_zlow = nullptr, *t_coef = nullptr;
  float *t_coef_f = nullptr;
  int *t_off = nullptr, *t_nint = nullptr, *t_pairidx = nullptr;
  double *t_r2atat = nullptr, *t_r2umin = nullptr;
  unsigned short *t_bin_k = nullptr;
  mfe::PairRec *t_rec = nullptr;
  mfe::PairRecF *t_rec_f = nullptr;
  float *t_zlow_f = nullptr;
  int t_nbin = 0;
  int t_lds_mode = 0;
  size_t t_lds_bytes = 0;
  int t_nzl = 0;
  double rcut = 0.0;

  // neighbour list
  int nlkind = MFE_NL_VLIST;
  double drnlist = 0.0;
  int maxnn = 0;
  mfe::CellGeom cg{};
  int *c_count = nullptr, *c_start = nullptr, *c_of = nullptr, *c_slot = nullptr;
  int *c_sorted = nullptr;   // cell order: position t -> particle index
  int *c_pos = nullptr;      // particle index -> cell-order position
  double *ro_s = nullptr;    // positions in cell order (refreshed per UTotal)
  int *type_s = nullptr;
  float4 *pos_f = nullptr;   // mixed path: cell-relative FP32 positions and types
  int *nn = nullptr, *nbr = nullptr;  // by cell-order position
  int *flags = nullptr;      // [0] error bits, [1] rebuild flag, [2] rebuild count
  int ncells_alloc = 0;

  // Ewald
  int ewald = MFE_EWALD_NONE;
  double alpha = 0.0, efp = 0.0;
  int lsurf = 0;
  // SPME
  int order = 0, s[3] = {0, 0, 0}, hx = 0;
  int B[3] = {0, 0, 0}, nb[3] = {0, 0, 0}, nbricks = 0;
  void *mesh_r = nullptr;    // real mesh [s3][s2][s1]
  void *mesh_c = nullptr;    // complex mesh [s3][s2][hx]
  int *b_count = nullptr, *b_start = nullptr, *b_id = nullptr, *b_slot = nullptr;
  // Monte Carlo trial moves: moved particles, their trial positions, and the
  // two reciprocal energies (current, trial)
  int *mc_ia = nullptr;
  double *mc_rtm = nullptr, *mc_out = nullptr;
  int mc_cap = 0;
  // the current configuration's spreading tiles and (in mc_out[0]) its
  // E_rec, kept for callers whose positions have not changed since
  // MC trial moves: the kept spectrum (forward transform of the current
  // configuration's mesh, valid while mc_cached) and the stencil factors
  void *mc_spec = nullptr;
  size_t mc_spec_bytes = 0;
  double *mc_fac = nullptr;
  size_t mc_fac_cap = 0;
  // pinned host staging of a move (indices, positions) and its results, so
  // the copies of an MC call run asynchronously on the stream
  char *h_mc = nullptr;
  size_t h_mc_bytes = 0;
  bool mc_cached = false;
  // the move of the last mfe_mc_delta_rec (0-based indices, positions):
  // mfe_mc_accept of the same move keeps its trial mesh and E_rec
  std::vector<int> mc_last_ia;
  std::vector<double> mc_last_rtm;
  int *b_done = nullptr;
  void *rec_frac = nullptr;  // per sorted particle: (fx, fy, fz, q)
  void *spread_tiles = nullptr;  // per brick: extended spreading tile
  size_t spread_tiles_bytes = 0;
  int2 *rec_meta = nullptr;  // per sorted particle: (packed meshmax, index)
  double *ax_g = nullptr, *ax_kk = nullptr;  // [3][smax/2+1]
  void *tw[3] = {nullptr, nullptr, nullptr};     // twiddles of length s[a]
  void *tw_half = nullptr;                       // twiddles of length s1/2
  mfe::FftPlan plan[3]{}, plan_half{};
  int spread_lds = 0;
  // standard Ewald
  int ncut = 0, ncut_cube = 0, nkvec = 0;
  double *k_vec = nullptr;   // [nkvec][4]: kx, ky, kz (2*pi*n/L), kfac
  double *k_sum = nullptr;   // [nkvec][2] structure factor (re, im)

  // chains
  int nchain = 0;
  int *ch_prev = nullptr, *ch_next = nullptr;
  double bond_k = 0.0, bond_eq = 0.0;
  int bond_p = 2;
  double ang_k = 0.0, ang_eq = 0.0;
  int ang_p = 2;

  // reductions
  double *part = nullptr;
  int part_cap = 0;
  double *en = nullptr;      // final energies (device)
  double *h_en = nullptr;    // pinned host copy
  int *h_flags = nullptr;    // pinned host copy of flags
  int off_pair = 0, off_conv = 0, off_bond = 0, off_surf = 0, off_std = 0;
  int off_move = 0, off_skin = 0;
  int nblk_pair = 0, nblk_conv = 0, nblk_bond = 0, nblk_surf = 0, nblk_move = 0;

  // additive spline pair term (knots) and adjoint state
  int nknot = 0;
  double knot_r0 = 0.0, knot_dinv = 0.0;
  double *knot_c = nullptr;  // [nptpt][nknot]
  mfe::PairParams pp{};
  bool have_pp = false;
  double *phi_rec = nullptr, *adj_phi = nullptr, *adj_dq = nullptr;
  double *adj_part = nullptr, *adj_out = nullptr;
  int adj_cap = 0;
  // force matching: device-resident frames (positions, target forces), the
  // model forces' reciprocal field, residuals, residual spreading weights,
  // block partials and the accumulated loss and gradient
  double *fm_ro = nullptr, *fm_tgt = nullptr;
  int fm_nframes = 0, fm_frame_np = 0;
  double *fm_fld = nullptr, *fm_res = nullptr, *fm_res_s = nullptr;
  void *fm_vec = nullptr;
  double *fm_part = nullptr, *fm_acc = nullptr, *fm_qr = nullptr;
  size_t fm_part_cap = 0;
  int fm_acc_cap = 0, fm_np = 0;

  // Brownian dynamics
  double *bd_fac1 = nullptr, *bd_fac2 = nullptr;  // [MFE_MAX_TYPES]
  double *bd_noise = nullptr;                     // [3 np] host-supplied noise
  double *ro_stage = nullptr;                     // [3 np] positions from the host
  unsigned long long bd_seed = 0;
  bool have_bd = false;

  // device view (one replica) and its launch group
  mfe::SysDev h_sys{};
  mfe::SysDev *d_sys = nullptr;
  mfe::Group grp{};
  bool sys_dirty = true;
  unsigned version = 1;      // bumped whenever the view changes

  // MD bookkeeping
  mfe::StepGraphs md_graph;
  double md_graph_dt = -1.0;
  unsigned md_graph_flags = 0;
  int md_graph_kind = 0;     // 1 = MD, 2 = BD
  bool graph_dirty = true;
};

namespace mfe {
// launchers (mfe_kernels.hip / mfe_fft.hip); every launch covers all
// replicas of the group
void launch_nlist_rebuild(const Group &g, bool gated);
// sort_pos: refresh the cell-ordered positions first (not needed inside MD
// and BD steps, whose k_move keeps them current)
// skin: the pair launch also computes the skin test's block candidates
// (MD/BD steps, after k_move), replacing launch_skin
void launch_pair(const Group &g, bool energy, bool sort_pos = true, bool skin = false);
// What the gather launch also does when it is the last force stage (SPME
// without bonds or surface term): each particle's finalize phase right
// after its force, and the reductions in one extra block.
struct FinArgs {
  int energy, skin, bd, mode;
  double dt;
};
// what the gather does beside the forces (k_gather's MODE)
enum { GM_FORCE = 0, GM_POT = 1, GM_FIELD = 2, GM_POTONLY = 3, GM_FORCE_REC = 4 };
void launch_spme(const Group &g, bool energy, int first_stage, int last_stage,
                 int gather_mode = GM_FORCE, const FinArgs *fin = nullptr);
// adjoint passes: real-space pair sums, per-particle dE/dq, column sums
void launch_adjoint(const Group &g, int ncol, double *out);
void launch_std_ewald(const Group &g);
void launch_bonds(const Group &g);
void launch_surf(const Group &g);
void launch_reduce(const Group &g, unsigned flags, bool skin, bool bd);
void launch_finalize(const Group &g, int mode, double dt);
// reduce and finalize in one launch (only without the surface term)
// split: the reciprocal forces are in force_rec (after launch_pair_gather)
void launch_reduce_finalize(const Group &g, unsigned flags, bool skin, bool bd, int mode,
                            double dt, bool split = false);
// launch_reduce_finalize (with the skin test) followed by the gated
// rebuild, in one launch; false (nothing launched) when the group cannot
// take it (several replicas, or no grid barrier state)
bool launch_finalize_rebuild(const Group &g, unsigned flags, bool bd, int mode, double dt,
                             bool split);
// the pair forces and the gather (into force_rec) in one launch
void launch_pair_gather(const Group &g, bool energy, bool skin);
void launch_move(const Group &g, double dt, int bd);
// the cell-ordered positions refreshed from the particle-ordered ones
void launch_sort_pos(const Group &g);
// ro_stage[ia[a]] = rtm[a] for a < natm (a trial move written into a copy
// of the positions)
void launch_mc_place(const Group &g, int natm, const int *ia, const double *rtm,
                     double *ro_stage);
// Monte Carlo trial move, reciprocal part (k_mc_kfac, k_mc_kdelta): the
// stencil factors of the natm moved particles (current and trial
// positions) into fac[natm][2][s1 + s2 + s3] (double2), then one pass over
// the kept spectrum spec: the energy change (without EpsiFourPi) into
// out[0] over part[2 * kMcPairBlocks], or (apply) spec += D.  One replica.
void launch_mc_kfac(const Group &g, int natm, const int *ia, const double *rtm, void *fac);
void launch_mc_kdelta(const Group &g, int natm, const void *fac, void *spec, bool apply,
                      double *part, double *out);
// the z forward transform of the complex mesh into out (no Green function)
void fft_z_forward(const Group &g, void *out);
// Monte Carlo trial move, pair part: out[0] the two-body energy change,
// out[1] 1 when a trial pair overlaps
// all: every particle a candidate partner (cell-list semantics), over
// kMcPairBlocks blocks whose partials part[2 * kMcPairBlocks] are summed
// in a fixed order; else the Verlet-list partners in one block
constexpr int kMcPairBlocks = 512;
// moves of up to this many particles: k_mc_pair stages them in LDS and may
// read them from host memory
constexpr int kMcStage = 64;
void launch_mc_pair(const Group &g, int natm, const int *ia, const double *rtm, double *out,
                    double *part, bool all);
void launch_skin(const Group &g);
void launch_set_positions(const Group &g, const double *ro_new);
// the same with the current positions, the natm <= kMcStage particles ia
// moved to rtm (device or pinned host memory)
void launch_set_positions_moved(const Group &g, int natm, const int *ia, const double *rtm);
void launch_collect(const Group &g, int *flags, double *en, int nen);
// persistent batched BD (mfe_kernels.hip): nsteps BDSteps of every replica
// in one launch, a workgroup per replica at a time, replicas taken from a
// queue in the given order
constexpr int kPersistThreads = 1024;
// dynamic LDS of the persistent launch: tables and FFT tiles
constexpr size_t kPersistLds = 144 * 1024;
void launch_bd_persistent(const Group &g, const int *order, int nrep, int nsteps, int energy,
                          int *queue, int blocks, size_t lds);
bool persist_allow_lds(int order, size_t bytes);
void launch_fill_int(int *p, int n, int v, hipStream_t s);
void launch_fill_double(double *p, int n, double v, hipStream_t s);

// force matching (mfe_fm.hip): model pair forces (backward false: written
// over force[]) or the pair part of the loss gradient (backward true:
// per-particle dL/dq into adj_dq, block partials [dsig, deps, dknot] of
// ncol columns); the residual and its partials [|r|^2, q r]; the
// reciprocal part (residual mesh spread, convolved, potential gathered);
// dL/dq per particle with per-type partials; column sums into acc
void launch_fm_pair(const Group &g, bool backward, double *part, int ncol, int np);
int fm_pair_blocks(int np);
void launch_fm_resid(const Group &g, const double *tgt, double *part);
void launch_fm_rec_backward(const Group &g);
void launch_fm_q(const Group &g, const double *rpart, double *qr, double *part);
void launch_fm_acc(const Group &g, const double *part, int nblk, int ncol, int cols,
                   double *acc);
template <typename R, int O>
const void *k_fm_spread_ptr();

// per-replica launch geometry of the FFT passes (mfe_fft.hip)
void fft_geometry(int prec, const int s[3], int *fx_rpb, int *fx_nblk, int *fy_T,
                  int *fy_nbx, int *fz_T, int *fz_nbx, size_t *lds_fx, size_t *lds_fy,
                  size_t *lds_fz);
template <typename R, int O>
const void *k_spread_ptr();
template <typename R, int O>
const void *k_gather_ptr(int mode, bool fin);

// FFT passes (mfe_fft.hip)
void fft_x_forward(const Group &g);
void fft_y(const Group &g, int inverse);
void fft_conv(const Group &g, bool energy);
void fft_x_inverse(const Group &g);

// Philox4x32-10 (Salmon et al., SC'11) on the host, for tests of the
// device generator
void philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
}  // namespace mfe

// groups and step sequences (mfe_engine.hip)
int mfe_group_build(mfe_engine *const *eng, int R, const mfe::SysDev *d, hipStream_t st,
                    mfe::Group &g, std::string &err);
void mfe_fill_view(const mfe_engine *e, mfe::SysDev &v);
void mfe_sync_view(mfe_engine *e);   // throws HipError
int mfe_map_flag_error(mfe_engine *e, int bits);
void mfe_allow_big_spread_lds(int prec, int bytes);
// rebuild (with skin): the gated list rebuild may ride in the last launch;
// true when it did (else the caller launches it)
bool mfe_launch_utotal(const mfe::Group &g, unsigned flags, int mode, double dt, bool skin,
                       bool bd, bool rebuild = false);
void mfe_launch_md_step(const mfe::Group &g, double dt, unsigned flags);
void mfe_launch_bd_step(const mfe::Group &g, int noise, unsigned flags);

// sizes of SPME stages (stage indices used by launch_spme)
enum { SP_BIN = 0, SP_SPREAD = 1, SP_FFTX = 2, SP_FFTY = 3, SP_CONV = 4,
       SP_IFFTY = 5, SP_IFFTX = 6, SP_GATHER = 7 };
