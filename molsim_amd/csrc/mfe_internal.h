// mfe_internal.h -- engine state and kernel interfaces shared by the
// translation units of libmfe (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mfe.h"

#define MFE_HIP_TRY(expr)                                                    \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) throw mfe::HipError(_e, #expr, __FILE__, __LINE__); \
  } while (0)

namespace mfe {

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
};

template <typename R>
struct cplx {
  R x, y;
};

// Orthorhombic periodic box (Src/molaux.F90:326-335); dpbc == boxlen.
struct Box {
  double L[3], L2[3], Li[3];
};

// Cell grid used to find Verlet-list partners: edge >= rcut + drnlist.
struct CellGeom {
  int n[3];          // cells per axis (1 when the axis has < 3 cells)
  double cinv[3];    // cells per unit length
  int ncells;
};

// Table intervals in device layout: for pair p, intervals k = 0..nint[p]-1
// start at index off[p]; zlow[off+k] (FP64, decreasing in k) and the 11
// coefficients coef[(off+k)*11 + c] (usum: c = 0..5, fsum: c = 6..10).
struct TableDev {
  const double *zlow;
  const double *coef;
  const float *coef_f;
  const int *off;
  const int *nint;
  const double *r2atat;
  const double *r2umin;
  const int *pairidx;  // [npt*npt] -> pair index 0..nptpt-1
  int npt, nptpt;
  double rcut2;
};

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
  const double *g[3];
  const double *kk[3];
  double fac0;   // 4*pi/V
  double expf;   // -1/(4 alpha^2)
};

struct Partials;  // forward

}  // namespace mfe

// ------------------------------------------------------------------ engine
struct mfe_engine {
  int device = 0;
  int prec = MFE_PREC_MIXED;
  hipStream_t stream = nullptr;
  std::string err;
  int max_np = 0;

  // system
  int np = 0, npt = 0, nptpt = 0;
  mfe::Box box{};
  bool have_box = false, have_particles = false, have_tables = false;
  bool have_ewald = false, have_neighbor = false, have_positions = false;
  bool list_valid = false;
  double self_energy = 0.0;  // -alpha/sqrt(pi) * sum q^2 (UEwaldSelf)

  // particle state (device, xyz-interleaved FP64)
  double *ro = nullptr, *rod = nullptr, *rodd = nullptr;
  double *force = nullptr;   // UTotal's force(3,na) before unit scaling
  double *forceo = nullptr;  // force * fac
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
  double rcut = 0.0;

  // neighbour list
  int nlkind = MFE_NL_VLIST;
  double drnlist = 0.0;
  int maxnn = 0;
  mfe::CellGeom cg{};
  int *c_count = nullptr, *c_start = nullptr, *c_of = nullptr, *c_slot = nullptr;
  int *c_sorted = nullptr;   // cell order: position t -> particle index
  double *ro_s = nullptr;    // positions in cell order (refreshed per UTotal)
  int *type_s = nullptr;
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
  void *rec_frac = nullptr;  // per sorted particle: (fx, fy, fz, q)
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

  // reductions
  double *part = nullptr;
  int part_cap = 0;
  double *en = nullptr;      // final energies (device)
  double *h_en = nullptr;    // pinned host copy
  int *h_flags = nullptr;    // pinned host copy of flags
  int off_pair = 0, off_conv = 0, off_bond = 0, off_surf = 0, off_std = 0;
  int off_move = 0, off_skin = 0;
  int nblk_pair = 0, nblk_conv = 0, nblk_bond = 0, nblk_surf = 0, nblk_move = 0;

  // MD bookkeeping
  hipGraphExec_t md_graph = nullptr;
  double md_graph_dt = -1.0;
  unsigned md_graph_flags = 0;
  bool graph_dirty = true;
};

namespace mfe {
// launchers (mfe_kernels.hip / mfe_fft.hip)
void launch_nlist_rebuild(mfe_engine *e, bool gated);
void launch_pair(mfe_engine *e, bool energy);
void launch_spme(mfe_engine *e, bool energy, int first_stage, int last_stage);
void launch_std_ewald(mfe_engine *e);
void launch_bonds(mfe_engine *e);
void launch_surf(mfe_engine *e);
void launch_reduce(mfe_engine *e, unsigned flags, bool skin);
void launch_finalize(mfe_engine *e, int mode, double dt);
void launch_move(mfe_engine *e, double dt, int bd, const double *fac1,
                 const double *fac2, const double *noise);
void launch_skin(mfe_engine *e);
void launch_fill_int(int *p, int n, int v, hipStream_t s);
void launch_fill_double(double *p, int n, double v, hipStream_t s);

int conv_blocks(mfe_engine *e);
template <typename R, int O>
const void *k_spread_ptr();
template <typename R, int O>
const void *k_gather_ptr();

// FFT passes (mfe_fft.hip)
void fft_x_forward(mfe_engine *e);
void fft_y(mfe_engine *e, int inverse);
void fft_conv(mfe_engine *e, bool energy);
void fft_x_inverse(mfe_engine *e);
}  // namespace mfe

// sizes of SPME stages (stage indices used by launch_spme)
enum { SP_BIN = 0, SP_SPREAD = 1, SP_FFTX = 2, SP_FFTY = 3, SP_CONV = 4,
       SP_IFFTY = 5, SP_IFFTX = 6, SP_GATHER = 7 };
