// mfe_fm.hip -- the force-matching loss and its parameter gradient.
//
// No reference counterpart.  The model is the pair potential MOLSIM's
// tables interpolate -- Pot_1_6_12 (Src/potential.F90:1349-1425) with
// exact erfc, Lorentz-Berthelot mixing and the WCA cut -- plus the
// additive cubic B-spline term per type pair, SPME (UEwaldRecSPM,
// Src/energy.F90:1207-1297) with its surface term, and the bonds and
// angles (UBond/UAngle, Src/energy.F90:5335-5481).  Its forces F = -dE/dro
// are those UTotal assembles (Src/energy.F90:280-434) with the pair
// tables replaced by the model itself.  For frames n with targets T:
//
//   L = sum_n sum_i |F_i - T_i|^2,   dL/dtheta = 2 sum_i r_i . dF_i/dtheta,
//
// with the residual r = F - T.  A pair (i, j) at d = r_i - r_j (minimum
// image) adds f(r) d to F_i and -f(r) d to F_j, f = -u'(r)/r, so its share
// of dL/dtheta is 2 (df/dtheta) w with w = (r_i - r_j) . d.  SPME's force
// F_i = q_i E(x_i), E = -efp grad(W^T G Q), gives
//
//   dL/dq_m = 2 r_m . E(x_m) - 2 efp W(x_m)^T G R,
//   R = sum_i q_i (r_i . grad W(x_i)),
//
// the second term one more spreading (of the residual-weighted stencil
// derivatives, recomputed in the backward pass instead of a stored mesh),
// the same convolution and a potential gather.  The surface term
// F_i = -efp (4 pi/3V) q_i S, S = sum q x, adds
// -2 efp (4 pi/3V) (r_m . S + x_m . sum_i q_i r_i).  Per-type charges,
// sigma and epsilon collect the per-particle and per-pair terms through
// q_m = zat(type m), sigma_ij = (sigma_i + sigma_j)/2 and
// eps_ij = sqrt(eps_i eps_j).
//
// R is the arithmetic type of the model terms: FP32 on the mixed path,
// FP64 on the FP64 path; sums over lanes, particles and blocks are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mfe_device.h"
#include "mfe_internal.h"

namespace mfe {
namespace {

// lanes sharing one particle's list in the model pair kernel
constexpr int kFmLanes = 4;
constexpr int kFmBlock = 256;

template <typename R>
__device__ __forceinline__ R erfc_r(R x);
template <>
__device__ __forceinline__ float erfc_r<float>(float x) { return erfcf(x); }
template <>
__device__ __forceinline__ double erfc_r<double>(double x) { return erfc(x); }
template <typename R>
__device__ __forceinline__ R exp_r(R x);
template <>
__device__ __forceinline__ float exp_r<float>(float x) { return expf(x); }
template <>
__device__ __forceinline__ double exp_r<double>(double x) { return exp(x); }

// quad sum of an FP64 value by DPP (every lane of the wavefront active)
__device__ __forceinline__ double quad_sum(double x) {
  x += quad_swap<kXor1>(x);
  x += quad_swap<kXor2>(x);
  return x;
}

// Model pair forces (BWD false): F of each particle's pairs, written over
// force[] (the other terms add to it), in particle order.  Gradient pass
// (BWD true): per visit (i, j), with w = (r_i - r_j) . d,
//   dL/dq_i         += 2 efp q_j g(r) w            (per particle, adj_dq)
//   dL/dsig[t_i]    += w df/dsig_ij
//   dL/deps[t_i]    += w df/deps_ij eps_j / eps_ij
//   dL/dc[p][k]     += w df/dc_k
// where g = f_coul / (q_i q_j efp) -- each pair is seen from both ends, so
// the per-type terms carry half of the pair's share each time.  Block
// partials [dsig(npt), deps(npt), dc(nptpt*nknot)] at part[blk * ncol].
template <typename R, bool BWD>
__global__ void __launch_bounds__(kFmBlock) k_fm_pair(const SysDev *__restrict__ sys,
                                                      double *__restrict__ part, int ncol) {
  const SysDev &V = sys[blockIdx.z];
  const int np = V.np;
  constexpr int kPer = kFmBlock / kFmLanes;
  if ((int)blockIdx.x * kPer >= np) return;
  extern __shared__ double sk[];
  __shared__ double sh[kFmBlock / 64];
  const int nk = V.nptpt * V.nknot;
  if (BWD) {
    for (int k = threadIdx.x; k < nk; k += blockDim.x) sk[k] = 0.0;
    __syncthreads();
  }
  const PairParams &pp = V.pp;
  const TableDev &t = V.tab;
  const Box b = V.box;
  const int lane = threadIdx.x % kFmLanes;
  const int tt = blockIdx.x * kPer + threadIdx.x / kFmLanes;
  const bool act = tt < np;
  const R efp = (R)pp.efp, alpha = (R)pp.alpha;
  const R two_a_pi = (R)(1.1283791670955126 * pp.alpha);  // 2 alpha / sqrt(pi)
  double fx = 0.0, fy = 0.0, fz = 0.0, dq = 0.0;
  double gs = 0.0, ge = 0.0;
  int ti = 0;
  if (act) {
    const double xi = V.ro_s[3 * tt], yi = V.ro_s[3 * tt + 1], zi = V.ro_s[3 * tt + 2];
    ti = V.type_s[tt];
    const R qi = (R)pp.zat[ti], si = (R)pp.sigat[ti], ei = (R)pp.epsat[ti];
    R rxi = 0, ryi = 0, rzi = 0;
    if (BWD) {
      rxi = (R)V.fm_res_s[3 * tt];
      ryi = (R)V.fm_res_s[3 * tt + 1];
      rzi = (R)V.fm_res_s[3 * tt + 2];
    }
    const int n = V.nn[tt];
    for (int k = lane; k < n; k += kFmLanes) {
      const int pj = (int)((unsigned)V.nbr[(size_t)k * np + tt] & kNbrIdxMask);
      double dx = xi - V.ro_s[3 * pj], dy = yi - V.ro_s[3 * pj + 1], dz = zi - V.ro_s[3 * pj + 2];
      const double r2d = pbcr2(dx, dy, dz, b);
      if (V.nlsem == 0 ? (r2d > t.rcut2) : !(r2d < t.rcut2)) continue;
      const int tj = V.type_s[pj];
      const R r2 = (R)r2d, r = sqrt(r2);
      const R dxr = (R)dx, dyr = (R)dy, dzr = (R)dz;
      const R qj = (R)pp.zat[tj];
      // Coulomb: f = q_i q_j efp g(r)
      R g;
      if (pp.lewald) {
        const R ar = alpha * r;
        g = (erfc_r<R>(ar) / r + two_a_pi * exp_r<R>(-ar * ar)) / r2;
      } else {
        g = R(1) / (r * r2);
      }
      R f = qi * qj * efp * g;
      // Lennard-Jones inside the WCA cut
      const R sig = R(0.5) * (si + (R)pp.sigat[tj]);
      const R ej = (R)pp.epsat[tj];
      const R eps = sqrt(ei * ej);
      R fs = 0, fe = 0;
      if (!pp.lljcut || r < (R)pp.ljrcut * sig) {
        const R x = sig / r, x2 = x * x, s6 = x2 * x2 * x2, s12 = s6 * s6;
        fe = R(24) * (R(2) * s12 - s6) / r2;      // df/deps_ij
        f += eps * fe;
        fs = R(144) * eps * (R(4) * s12 - s6) / (sig * r2);  // df/dsig_ij
      }
      // spline term (FP64: few pairs, exact weights)
      int ki = 0;
      double w4[4], dw4[4];
      bool kn = false;
      int pidx = 0;
      if (V.nknot) {
        kn = knot_weights((double)r2d > 0.0 ? sqrt(r2d) : 0.0, V.knot_r0, V.knot_dinv, V.nknot,
                          ki, w4, dw4);
        if (kn) {
          pidx = t.pairidx[ti * t.npt + tj];
          if (!BWD) {
            const double *c = V.knot_c + (size_t)pidx * V.nknot + ki;
            const double dup =
                (c[0] * dw4[0] + c[1] * dw4[1] + c[2] * dw4[2] + c[3] * dw4[3]) * V.knot_dinv;
            f += (R)(-dup / sqrt(r2d));
          }
        }
      }
      if (!BWD) {
        fx += (double)(f * dxr);
        fy += (double)(f * dyr);
        fz += (double)(f * dzr);
      } else {
        const R w = (rxi - (R)V.fm_res_s[3 * pj]) * dxr + (ryi - (R)V.fm_res_s[3 * pj + 1]) * dyr +
                    (rzi - (R)V.fm_res_s[3 * pj + 2]) * dzr;
        dq += (double)(R(2) * efp * qj * g * w);
        gs += (double)(w * fs);
        if (eps > R(0)) ge += (double)(w * fe * ej / eps);
        if (kn) {
          const double wr = (double)w * V.knot_dinv / sqrt(r2d);
          for (int m = 0; m < 4; ++m) atomicAdd(&sk[pidx * V.nknot + ki + m], -wr * dw4[m]);
        }
      }
    }
  }
  // the particle's lanes combined in a fixed order
  if (!BWD) {
    fx = quad_sum(fx);
    fy = quad_sum(fy);
    fz = quad_sum(fz);
    if (act && lane == 0) {
      const int i = V.c_sorted[tt];
      V.force[3 * i] = fx;
      V.force[3 * i + 1] = fy;
      V.force[3 * i + 2] = fz;
    }
    return;
  }
  dq = quad_sum(dq);
  if (act && lane == 0) V.adj_dq[V.c_sorted[tt]] = dq;
  double *out = part + (size_t)blockIdx.x * ncol;
  for (int u = 0; u < V.npt; ++u) {
    const double a = block_sum(act && ti == u ? gs : 0.0, sh);
    const double c = block_sum(act && ti == u ? ge : 0.0, sh);
    if (threadIdx.x == 0) {
      out[u] = a;
      out[V.npt + u] = c;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nk; k += blockDim.x) out[2 * V.npt + k] = sk[k];
}

// Residual r = forceo/ffac - T (particle order and cell order) and its
// block partials [sum |r|^2, sum q r (3)] at part[blk * 4].
__global__ void __launch_bounds__(kFmBlock) k_fm_resid(const SysDev *__restrict__ sys,
                                                       const double *__restrict__ tgt,
                                                       double *__restrict__ part) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nblk_np) return;
  __shared__ double sh[kFmBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double l2 = 0.0, s[3] = {0.0, 0.0, 0.0};
  if (i < V.np) {
    const double inv = 1.0 / V.ffac;
    const int cp = V.c_pos[i];
    const double q = V.q[i];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double r = V.forceo[3 * i + m] * inv - tgt[3 * i + m];
      V.fm_res[3 * i + m] = r;
      V.fm_res_s[3 * cp + m] = r;
      l2 += r * r;
      s[m] = q * r;
    }
  }
  double *out = part + (size_t)blockIdx.x * 4;
  const double a = block_sum(l2, sh);
  if (threadIdx.x == 0) out[0] = a;
  for (int m = 0; m < 3; ++m) {
    const double c = block_sum(s[m], sh);
    if (threadIdx.x == 0) out[1 + m] = c;
  }
}

// Residual spreading weights in brick order: (q r_x dkdr_x, q r_y dkdr_y,
// q r_z dkdr_z) of the particle of each record, so that the spreading of
// R = sum_i q_i r_i . grad W(x_i) takes the records' fractions as they are.
template <typename R>
__global__ void __launch_bounds__(kFmBlock) k_fm_vec(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= V.np) return;
  const int i = V.rec_meta[pos].y;
  const double q = V.q[i];
  R *v = (R *)V.fm_vec + 4 * pos;
  v[0] = (R)(q * V.fm_res[3 * i] * V.sg.dkdr[0]);
  v[1] = (R)(q * V.fm_res[3 * i + 1] * V.sg.dkdr[1]);
  v[2] = (R)(q * V.fm_res[3 * i + 2] * V.sg.dkdr[2]);
  v[3] = R(0);
}

// Spreading of R: as k_spread (one block per brick, LDS tile, fixed-point
// or FP64 LDS accumulation, no global atomics), with the stencil weight of
// point (jx, jy, jz) vx W'x Wy Wz + vy Wx W'y Wz + vz Wx Wy W'z (W' the
// B-spline derivative in mesh units; v from k_fm_vec).
template <typename R, int ORDER>
__global__ void __launch_bounds__(1024) k_fm_spread(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nbricks) return;
  const R *__restrict__ frac = (const R *)V.rec_frac;
  const R *__restrict__ vec = (const R *)V.fm_vec;
  const int2 *__restrict__ meta = V.rec_meta;
  const int *__restrict__ bstart = V.b_start;
  const SpmeGeom g = V.sg;
  typedef typename SpreadAcc<R>::T Acc;
  extern __shared__ char smem[];
  Acc *tile = reinterpret_cast<Acc *>(smem);
  const int bx = blockIdx.x % g.nb[0];
  const int by = (blockIdx.x / g.nb[0]) % g.nb[1];
  const int bz = blockIdx.x / (g.nb[0] * g.nb[1]);
  const int o0 = bx * g.B[0] - (ORDER - 1), o1 = by * g.B[1] - (ORDER - 1),
            o2 = bz * g.B[2] - (ORDER - 1);
  const int E0 = g.B[0] + ORDER - 1, E1 = g.B[1] + ORDER - 1, E2 = g.B[2] + ORDER - 1;
  const int ntile = E0 * E1 * E2;
  const int first = bstart[blockIdx.x], items = (bstart[blockIdx.x + 1] - first) * ORDER;
  for (int k = threadIdx.x; k < ntile; k += blockDim.x) tile[k] = Acc(0);
  __syncthreads();
  for (int r = threadIdx.x; r < items; r += blockDim.x) {
    const int rec = first + r / ORDER;
    const Rec4<R> f = load_rec(frac + 4 * rec);
    const Rec4<R> v = load_rec(vec + 4 * rec);
    const int code = meta[rec].x;
    const int iz = r % ORDER;
    R wx[ORDER], wy[ORDER], wz[ORDER], dx[ORDER], dy[ORDER], dz[ORDER];
    bspline<R, ORDER>(f.x, wx, dx);
    bspline<R, ORDER>(f.y, wy, dy);
    bspline<R, ORDER>(f.z, wz, dz);
    R wzi = wz[0], dzi = dz[0];
#pragma unroll
    for (int k = 1; k < ORDER; ++k)
      if (iz == k) {
        wzi = wz[k];
        dzi = dz[k];
      }
    const int lx = (code & 1023) - o0;
    const int ly = ((code >> 10) & 1023) - o1;
    const int lz = ((code >> 20) & 1023) - o2;
    Acc *plane = tile + (lz - iz) * E1 * E0;
#pragma unroll
    for (int jy = 0; jy < ORDER; ++jy) {
      // point (jx, jy, iz): dx[jx] * A + wx[jx] * Bc
      const R A = v.x * wy[jy] * wzi;
      const R Bc = v.y * dy[jy] * wzi + v.z * wy[jy] * dzi;
      Acc *row = plane + (ly - jy) * E0 + lx;
#pragma unroll
      for (int j = 0; j < ORDER; ++j) SpreadAcc<R>::add(&row[-j], dx[j] * A + wx[j] * Bc);
    }
  }
  __syncthreads();
  R *__restrict__ out = (R *)V.spread_tiles + (size_t)blockIdx.x * ntile;
  for (int k = threadIdx.x; k < ntile; k += blockDim.x) out[k] = SpreadAcc<R>::get(tile[k]);
}

// the sum q r of the residual partials (the surface term needs it whole)
__global__ void __launch_bounds__(64) k_fm_qr(const SysDev *__restrict__ sys,
                                              const double *__restrict__ part,
                                              double *__restrict__ qr) {
  const SysDev &V = sys[blockIdx.z];
  const int lane = threadIdx.x;
  for (int m = 0; m < 3; ++m) {
    double s = 0.0;
    for (int b2 = lane; b2 < V.nblk_np; b2 += 64) s += part[(size_t)b2 * 4 + 1 + m];
    for (int w = 32; w > 0; w >>= 1) s += __shfl_down(s, w, 64);
    if (lane == 0) qr[m] = s;
  }
}

// dL/dq per particle: the pair part (adj_dq) + SPME 2 r . E - 2 phi_R +
// the surface term; per-type block partials at part[blk * npt].
__global__ void __launch_bounds__(kFmBlock) k_fm_q(const SysDev *__restrict__ sys,
                                                   const double *__restrict__ qr,
                                                   double *__restrict__ part) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nblk_np) return;
  __shared__ double sh[kFmBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  int ti = -1;
  if (i < V.np) {
    ti = V.type[i];
    d = V.adj_dq[i];
    const double r[3] = {V.fm_res[3 * i], V.fm_res[3 * i + 1], V.fm_res[3 * i + 2]};
    if (V.ewald == MFE_EWALD_SPME) {
      d += 2.0 * (r[0] * V.fm_fld[3 * i] + r[1] * V.fm_fld[3 * i + 1] +
                  r[2] * V.fm_fld[3 * i + 2]);
      d -= 2.0 * V.phi_rec[i];
    }
    if (V.lsurf) {
      const double f = V.surf_fac;  // efp 4 pi / (3V)
      d -= 2.0 * f *
           (r[0] * V.en[5] + r[1] * V.en[6] + r[2] * V.en[7] + V.ro[3 * i] * qr[0] +
            V.ro[3 * i + 1] * qr[1] + V.ro[3 * i + 2] * qr[2]);
    }
    V.adj_dq[i] = d;
  }
  double *out = part + (size_t)blockIdx.x * V.npt;
  for (int u = 0; u < V.npt; ++u) {
    const double a = block_sum(ti == u ? d : 0.0, sh);
    if (threadIdx.x == 0) out[u] = a;
  }
}

// acc[off + c] += sum over blocks of part[b * ncol + c] (fixed order)
__global__ void __launch_bounds__(kFmBlock) k_fm_acc(const double *__restrict__ part, int nblk,
                                                     int ncol, int cols, double *__restrict__ acc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0.0;
  for (int b2 = 0; b2 < nblk; ++b2) s += part[(size_t)b2 * ncol + c];
  acc[c] += s;
}

inline dim3 grid1(int x, const Group &g) { return dim3(x > 0 ? x : 1, 1, g.R); }

}  // namespace

void launch_fm_pair(const Group &g, bool backward, double *part, int ncol, int np) {
  const int nb = (np + kFmBlock / kFmLanes - 1) / (kFmBlock / kFmLanes);
  const size_t lds = backward ? (size_t)std::max(ncol, 1) * sizeof(double) : 0;
  if (g.prec == MFE_PREC_MIXED) {
    if (backward)
      hipLaunchKernelGGL((k_fm_pair<float, true>), grid1(nb, g), dim3(kFmBlock), lds, g.stream,
                         g.d, part, ncol);
    else
      hipLaunchKernelGGL((k_fm_pair<float, false>), grid1(nb, g), dim3(kFmBlock), 0, g.stream,
                         g.d, part, ncol);
  } else {
    if (backward)
      hipLaunchKernelGGL((k_fm_pair<double, true>), grid1(nb, g), dim3(kFmBlock), lds, g.stream,
                         g.d, part, ncol);
    else
      hipLaunchKernelGGL((k_fm_pair<double, false>), grid1(nb, g), dim3(kFmBlock), 0, g.stream,
                         g.d, part, ncol);
  }
}

int fm_pair_blocks(int np) { return (np + kFmBlock / kFmLanes - 1) / (kFmBlock / kFmLanes); }

void launch_fm_resid(const Group &g, const double *tgt, double *part) {
  hipLaunchKernelGGL(k_fm_resid, grid1(g.np_blocks, g), dim3(kFmBlock), 0, g.stream, g.d, tgt,
                     part);
}

void launch_fm_rec_backward(const Group &g) {
  // the residual mesh: weights, spreading, the forward transforms and the
  // convolution (no energy), the inverse transforms, the potential gather
  if (g.prec == MFE_PREC_MIXED)
    hipLaunchKernelGGL(k_fm_vec<float>, grid1(g.np_blocks, g), dim3(kFmBlock), 0, g.stream, g.d);
  else
    hipLaunchKernelGGL(k_fm_vec<double>, grid1(g.np_blocks, g), dim3(kFmBlock), 0, g.stream,
                       g.d);
#define MFE_FMS(R, O)                                                                     \
  case O:                                                                                 \
    hipLaunchKernelGGL((k_fm_spread<R, O>), grid1(g.bricks, g), dim3(1024), g.lds_spread, \
                       g.stream, g.d);                                                    \
    break;
  if (g.prec == MFE_PREC_MIXED) {
    switch (g.order) { MFE_FMS(float, 3) MFE_FMS(float, 4) MFE_FMS(float, 5) MFE_FMS(float, 6)
                       MFE_FMS(float, 7) MFE_FMS(float, 8) }
  } else {
    switch (g.order) { MFE_FMS(double, 3) MFE_FMS(double, 4) MFE_FMS(double, 5)
                       MFE_FMS(double, 6) MFE_FMS(double, 7) MFE_FMS(double, 8) }
  }
#undef MFE_FMS
  launch_spme(g, false, SP_FFTX, SP_GATHER, GM_POTONLY);
}

void launch_fm_q(const Group &g, const double *rpart, double *qr, double *part) {
  if (g.surf)
    hipLaunchKernelGGL(k_fm_qr, grid1(1, g), dim3(64), 0, g.stream, g.d, rpart, qr);
  hipLaunchKernelGGL(k_fm_q, grid1(g.np_blocks, g), dim3(kFmBlock), 0, g.stream, g.d, qr, part);
}

void launch_fm_acc(const Group &g, const double *part, int nblk, int ncol, int cols,
                   double *acc) {
  if (cols <= 0) return;
  hipLaunchKernelGGL(k_fm_acc, dim3((cols + kFmBlock - 1) / kFmBlock), dim3(kFmBlock), 0,
                     g.stream, part, nblk, ncol, cols, acc);
}

template <typename R, int O>
const void *k_fm_spread_ptr() {
  return (const void *)&k_fm_spread<R, O>;
}
#define MFE_FMP(R, O) template const void *k_fm_spread_ptr<R, O>();
MFE_FMP(float, 3) MFE_FMP(float, 4) MFE_FMP(float, 5) MFE_FMP(float, 6) MFE_FMP(float, 7)
MFE_FMP(float, 8) MFE_FMP(double, 3) MFE_FMP(double, 4) MFE_FMP(double, 5) MFE_FMP(double, 6)
MFE_FMP(double, 7) MFE_FMP(double, 8)
#undef MFE_FMP

}  // namespace mfe
