// mfe_fft.hip -- the 3D FFT of SPMFFTRec (Src/energy.F90:6525-6587) on the
// mesh of UEwaldRecSPM, in FFTW's unnormalised r2c/c2r convention:
// real mesh [s3][s2][s1] (Fortran Qmesh(s1,s2,s3)) <-> complex mesh
// [s3][s2][s1/2+1].
//
// Five passes, each reading and writing the mesh once:
//   x forward  : r2c of every x row through a complex FFT of half length
//   y forward  : complex FFT along y for tiles of kx columns
//   conv       : z forward, Green function (meshfac) with the FP64 energy
//                and virial sums of the reference's CalcGIF loop, z inverse
//                -- the z transform pair is fused around the pointwise
//                product, so the mesh is not written between them
//   y inverse, x inverse (c2r)
// Each 1D transform is a mixed-radix Stockham FFT staged in LDS.
#include <hip/hip_runtime.h>

#include "mfe_fft_device.h"
#include "mfe_internal.h"

namespace mfe {
namespace {

// The spread mesh at point (x, y, z): the point's value in its own brick's
// spreading tile plus the low-side halo values of the next brick along each
// axis (at most 8 tiles), added in a fixed order.  Tiles are the bricks
// extended by order-1 points below (see k_spread).
template <typename R>
__device__ __forceinline__ R spread_point(const SysDev &V, const R *__restrict__ tiles, int x,
                                          int y, int z) {
  const SpmeGeom &g = V.sg;
  const int h = g.order - 1;
  const int E0 = g.B[0] + h, E1 = g.B[1] + h, E2 = g.B[2] + h;
  const int bc0 = fdiv(x, g.fB[0]), bc1 = fdiv(y, g.fB[1]), bc2 = fdiv(z, g.fB[2]);
  const int l0 = x - bc0 * g.B[0], l1 = y - bc1 * g.B[1], l2 = z - bc2 * g.B[2];
  R acc = R(0);
  for (int dz = 0; dz < 2; ++dz) {
    if (dz && l2 < g.B[2] - h) break;
    const int tz = dz ? (bc2 + 1 == g.nb[2] ? 0 : bc2 + 1) : bc2;
    const int ez = l2 + h - dz * g.B[2];
    for (int dy = 0; dy < 2; ++dy) {
      if (dy && l1 < g.B[1] - h) break;
      const int ty = dy ? (bc1 + 1 == g.nb[1] ? 0 : bc1 + 1) : bc1;
      const int ey = l1 + h - dy * g.B[1];
      for (int dx = 0; dx < 2; ++dx) {
        if (dx && l0 < g.B[0] - h) break;
        const int tx = dx ? (bc0 + 1 == g.nb[0] ? 0 : bc0 + 1) : bc0;
        const int ex = l0 + h - dx * g.B[0];
        const int t = tx + g.nb[0] * (ty + g.nb[1] * tz);
        acc += tiles[(size_t)t * (E0 * E1 * E2) + (ez * E1 + ey) * E0 + ex];
      }
    }
  }
  return acc;
}

// ------------------------------------------------------------ pass kernels
// The rows are read from the spreading tiles (the spread mesh is never
// stored).
template <typename R>
__global__ void __launch_bounds__(256) k_fft_x_fwd(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  // neighbouring rows (which read the same spreading tiles) on one XCD
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  if (blk >= V.fx_nblk) return;
  extern __shared__ char smem[];
  const R *__restrict__ tiles = (const R *)V.spread_tiles;
  const int nrows = V.sg.s[1] * V.sg.s[2], row0 = blk * V.fx_rpb;
  fftx_fwd_rows<R>(V, row0, min(V.fx_rpb, nrows - row0), V.fx_rpb, smem,
                   [&](int x, int y, int z) { return spread_point<R>(V, tiles, x, y, z); });
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_x_inv(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  // neighbouring rows (which the y inverse wrote together) on one XCD
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  if (blk >= V.fx_nblk) return;
  extern __shared__ char smem[];
  const int nrows = V.sg.s[1] * V.sg.s[2], row0 = blk * V.fx_rpb;
  fftx_inv_rows<R>(V, row0, min(V.fx_rpb, nrows - row0), V.fx_rpb, smem);
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_y(const SysDev *__restrict__ sys, int inv) {
  const SysDev &V = sys[blockIdx.z];
  // the tiles of one z plane share the cache lines of its rows: the blocks
  // of a contiguous range of planes go to one XCD
  const int L = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int bx = L % gridDim.x, z = L / gridDim.x;
  if (bx >= V.fy_nbx || z >= V.sg.s[2]) return;
  extern __shared__ char smem[];
  ffty_tile<R>(V, bx, z, inv, smem);
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_conv(const SysDev *__restrict__ sys, int energy) {
  const SysDev &V = sys[blockIdx.z];
  // the tiles of one y line share the cache lines of its rows: the blocks
  // of a contiguous range of lines go to one XCD
  const int L = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int bx = L % gridDim.x, y = L / gridDim.x;
  if (bx >= V.fz_nbx || y >= V.sg.s[1]) return;
  extern __shared__ char smem[];
  __shared__ double red[2][256];
  double esum = 0.0, vsum = 0.0;
  fftconv_tile<R>(V, bx, y, energy, smem, esum, vsum);
  if (energy) {
    // shuffle trees in each wavefront, then the wavefronts in order
    for (int w = 32; w > 0; w >>= 1) {
      esum += __shfl_down(esum, w, 64);
      vsum += __shfl_down(vsum, w, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
      red[0][wave] = esum;
      red[1][wave] = vsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double e = red[0][0], v = red[1][0];
      for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
        e += red[0][k];
        v += red[1][k];
      }
      double *__restrict__ part = V.part + V.off_conv;
      const int blk = y * V.fz_nbx + bx;
      part[blk] = e;
      part[V.nblk_conv + blk] = v;
    }
  }
}

// the z forward transform into out (the MC trial moves' kept spectrum,
// FP64 on both precision paths, so that accepted changes add to it without
// FP32 rounding)
template <typename R>
__global__ void __launch_bounds__(256) k_fft_z_fwd(const SysDev *__restrict__ sys, void *out) {
  const SysDev &V = sys[blockIdx.z];
  const int L = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int bx = L % gridDim.x, y = L / gridDim.x;
  if (bx >= V.fz_nbx || y >= V.sg.s[1]) return;
  extern __shared__ char smem[];
  fftz_fwd_tile<R, double>(V, bx, y, (cplx<double> *)out, smem);
}

inline int tile_width(int hx, int target) {
  const int ntiles = (hx + target - 1) / target;
  return (hx + ntiles - 1) / ntiles;
}

template <typename R>
int rows_per_block(int s1) {
  // two LDS buffers of rows*(s1/2+1) complex values, about 32 KB in all and
  // at most 16 rows; at FP64 about 48 KB and at most 32 rows (23 rows of 128
  // points measured faster there than 15, while FP32 is fastest at 16)
  constexpr bool f64 = sizeof(R) == 8;
  const int per_row = 2 * (s1 / 2 + 1) * (int)sizeof(cplx<R>);
  const int rows = ((f64 ? 48 : 32) * 1024) / per_row, cap = f64 ? 32 : 16;
  return rows < 1 ? 1 : (rows > cap ? cap : rows);
}

template <typename R>
int tile_target(int n) {
  const int per_col = 2 * (n + 1) * (int)sizeof(cplx<R>);
  int t = (32 * 1024) / per_col;
  return t < 1 ? 1 : (t > 16 ? 16 : t);
}

}  // namespace

void fft_geometry(int prec, const int s[3], int *fx_rpb, int *fx_nblk, int *fy_T, int *fy_nbx,
                  int *fz_T, int *fz_nbx, size_t *lds_fx, size_t *lds_fy, size_t *lds_fz) {
  const bool f = prec == MFE_PREC_MIXED;
  const size_t csz = f ? sizeof(cplx<float>) : sizeof(cplx<double>);
  const size_t tsz = f ? sizeof(Twiddle<float>) : sizeof(Twiddle<double>);
  const int hx = s[0] / 2 + 1;
  int rpb = f ? rows_per_block<float>(s[0]) : rows_per_block<double>(s[0]);
  // the rows and the twiddles within 64 KB of LDS
  while (rpb > 1 && 2 * (size_t)rpb * (s[0] / 2 + 1) * csz + (size_t)(s[0] / 2 + s[0]) * tsz >
                        64 * 1024)
    --rpb;
  *fx_rpb = rpb;
  *fx_nblk = (s[1] * s[2] + rpb - 1) / rpb;
  // two row buffers, then the half- and full-length twiddles
  *lds_fx = 2 * (size_t)rpb * (s[0] / 2 + 1) * csz + (size_t)(s[0] / 2 + s[0]) * tsz;
  const int Ty = tile_width(hx, f ? tile_target<float>(s[1]) : tile_target<double>(s[1]));
  *fy_T = Ty;
  *fy_nbx = (hx + Ty - 1) / Ty;
  *lds_fy = 2 * (size_t)Ty * (s[1] + 1) * csz + (size_t)s[1] * tsz;
  const int Tz = tile_width(hx, f ? tile_target<float>(s[2]) : tile_target<double>(s[2]));
  *fz_T = Tz;
  *fz_nbx = (hx + Tz - 1) / Tz;
  *lds_fz = 2 * (size_t)Tz * (s[2] + 1) * csz + (size_t)s[2] * tsz;
}

void fft_x_forward(const Group &g) {
  const dim3 gr(g.fx_blocks, 1, g.R);
  if (g.prec == MFE_PREC_MIXED)
    hipLaunchKernelGGL(k_fft_x_fwd<float>, gr, dim3(256), g.lds_fx, g.stream, g.d);
  else
    hipLaunchKernelGGL(k_fft_x_fwd<double>, gr, dim3(256), g.lds_fx, g.stream, g.d);
}
void fft_x_inverse(const Group &g) {
  const dim3 gr(g.fx_blocks, 1, g.R);
  if (g.prec == MFE_PREC_MIXED)
    hipLaunchKernelGGL(k_fft_x_inv<float>, gr, dim3(256), g.lds_fx, g.stream, g.d);
  else
    hipLaunchKernelGGL(k_fft_x_inv<double>, gr, dim3(256), g.lds_fx, g.stream, g.d);
}
void fft_y(const Group &g, int inverse) {
  const dim3 gr(g.fy_x, g.fy_y, g.R);
  if (g.prec == MFE_PREC_MIXED)
    hipLaunchKernelGGL(k_fft_y<float>, gr, dim3(256), g.lds_fy, g.stream, g.d, inverse);
  else
    hipLaunchKernelGGL(k_fft_y<double>, gr, dim3(256), g.lds_fy, g.stream, g.d, inverse);
}
void fft_z_forward(const Group &g, void *out) {
  const dim3 gr(g.fz_x, g.fz_y, g.R);
  if (g.prec == MFE_PREC_MIXED)
    hipLaunchKernelGGL(k_fft_z_fwd<float>, gr, dim3(256), g.lds_fz, g.stream, g.d, out);
  else
    hipLaunchKernelGGL(k_fft_z_fwd<double>, gr, dim3(256), g.lds_fz, g.stream, g.d, out);
}
void fft_conv(const Group &g, bool energy) {
  const dim3 gr(g.fz_x, g.fz_y, g.R);
  if (g.prec == MFE_PREC_MIXED)
    hipLaunchKernelGGL(k_fft_conv<float>, gr, dim3(256), g.lds_fz, g.stream, g.d, energy ? 1 : 0);
  else
    hipLaunchKernelGGL(k_fft_conv<double>, gr, dim3(256), g.lds_fz, g.stream, g.d,
                       energy ? 1 : 0);
}

}  // namespace mfe

