// mfe_device.h -- device helpers shared by the kernel translation units of
// libmfe (gfx950 only): minimum image, block reductions, lane exchanges,
// compensated sums, the spline pair term's weights, CardinalBSpline, mesh
// coordinates and the spreading accumulators.
#pragma once

#include <hip/hip_runtime.h>

#include "mfe_internal.h"

namespace mfe {
namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------- helpers
// PBCr2 for the box (Src/molaux.F90:159-182): one shift per axis, then
// r2 = dx**2 + dy**2 + dz**2 summed left to right without fusion, so that
// list membership is decided on the reference's FP64 value.
__device__ __forceinline__ double pbcr2(double &dx, double &dy, double &dz,
                                        const Box &b) {
#pragma clang fp contract(off)
  if (fabs(dx) > b.L2[0]) dx = dx - copysign(b.L[0], dx);
  if (fabs(dy) > b.L2[1]) dy = dy - copysign(b.L[1], dy);
  if (fabs(dz) > b.L2[2]) dz = dz - copysign(b.L[2], dz);
  double r2 = dx * dx;
  r2 = r2 + dy * dy;
  r2 = r2 + dz * dz;
  return r2;
}

// Block-wide sum (max) in a fixed order: a shuffle tree inside each
// wavefront, then the wavefront results in wavefront order.  Every thread
// gets the result; sh needs one slot per wavefront.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
  for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  T r = sh[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
  return r;
}

__device__ __forceinline__ double block_max(double v, double *sh) {
  for (int w = 32; w > 0; w >>= 1) v = fmax(v, __shfl_down(v, w, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, sh[k]);
  return r;
}

// The partner value of an FP64 register in a lane quad, by DPP (no LDS):
// kXor1 swaps lanes 0<->1 and 2<->3, kXor2 swaps 0<->2 and 1<->3 -- the
// exchanges of __shfl_xor(x, 1 or 2) within a quad.  Every lane of the
// wavefront must be active.
constexpr int kXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]
template <int CTRL>
__device__ __forceinline__ double quad_swap(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// ------------------------------------------------- spline pair term
// Uniform cubic B-spline on knots r0 + k/dinv: on interval i (fraction f)
// the weights of coefficients i..i+3 are (1-f)^3/6, (3f^3-6f^2+4)/6,
// (-3f^3+3f^2+3f+1)/6, f^3/6 (de Boor, A Practical Guide to Splines, ch. IX),
// and their derivatives in f.
__device__ __forceinline__ bool knot_weights(double r, double r0, double dinv, int nknot,
                                             int &i, double w[4], double dw[4]) {
  const double x = (r - r0) * dinv;
  if (!(x >= 0.0) || !(x < (double)(nknot - 3))) return false;
  i = (int)x;
  const double f = x - i, f2 = f * f, f3 = f2 * f, g = 1.0 - f;
  w[0] = g * g * g / 6.0;
  w[1] = (3.0 * f3 - 6.0 * f2 + 4.0) / 6.0;
  w[2] = (-3.0 * f3 + 3.0 * f2 + 3.0 * f + 1.0) / 6.0;
  w[3] = f3 / 6.0;
  dw[0] = -0.5 * g * g;
  dw[1] = 1.5 * f2 - 2.0 * f;
  dw[2] = -1.5 * f2 + f + 0.5;
  dw[3] = 0.5 * f2;
  return true;
}

struct KahanF {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// a / D for a small integer D, as the reference divides.  Powers of two
// multiply exactly.  Otherwise the product with the rounded reciprocal is
// refined by one FMA step (q + (a - q D) / D), which gives the correctly
// rounded quotient: multiplying by RN(1/3) or RN(1/5) alone rounds every
// quotient the same way, and on the FP32 path that bias (5e-8 per axis at
// order 6, so 1.4e-7 on the spread charge and 3e-7 on the reciprocal
// energy) does not average out over the mesh.
template <typename R, int D>
__device__ __forceinline__ R div_small(R a) {
  constexpr R inv = R(1.0 / D);
  if constexpr ((D & (D - 1)) == 0) {
    return a * inv;
  } else if constexpr (sizeof(R) == 4) {
    const float q = a * inv;
    const float r = __builtin_fmaf(-q, float(D), a);
    return __builtin_fmaf(r, inv, q);
  } else {
    const double q = a * inv;
    const double r = __builtin_fma(-q, double(D), a);
    return __builtin_fma(r, inv, q);
  }
}

// stages n = N..ORDER of the Schoenberg recursion (compile-time n, so the
// divisor is a constant of each stage)
template <typename R, int ORDER, int N>
__device__ __forceinline__ void bspline_stage(R x, R *v, R *last) {
  if constexpr (N <= ORDER) {
#pragma unroll
    for (int i = 0; i < N - 1; ++i) last[i] = v[i];
    v[0] = div_small<R, N - 1>(x * last[0]);
#pragma unroll
    for (int i = 1; i <= N - 2; ++i) {
      const R a = R(i) + x;
      v[i] = div_small<R, N - 1>(a * last[i] + (R(N) - a) * last[i - 1]);
    }
    v[N - 1] = div_small<R, N - 1>((R(1) - x) * last[N - 2]);
    bspline_stage<R, ORDER, N + 1>(x, v, last);
  }
}

// CardinalBSpline (Src/mollib.F90:1629-1683) at fraction x in [0,1):
// v[i] is the weight of grid point kmax - i, d[i] its derivative.
template <typename R, int ORDER>
__device__ __forceinline__ void bspline(R x, R *v, R *d) {
  R last[ORDER];
  v[0] = x;
  v[1] = R(1) - x;
#pragma unroll
  for (int i = 2; i < ORDER; ++i) v[i] = R(0);
#pragma unroll
  for (int i = 0; i < ORDER; ++i) last[i] = R(0);
  bspline_stage<R, ORDER, 3>(x, v, last);
  if (d) {
    d[0] = last[0];
#pragma unroll
    for (int i = 1; i <= ORDER - 2; ++i) d[i] = last[i] - last[i - 1];
    d[ORDER - 1] = -last[ORDER - 2];
  }
}

__device__ __forceinline__ void mesh_coords(const double *r, const SpmeGeom &g,
                                            int mm[3], double fr[3]) {
  for (int m = 0; m < 3; ++m) {
    const double u = g.dkdr[m] * r[m];
    const double f = floor(u);
    fr[m] = u - f;
    int k = (int)f % g.s[m];
    mm[m] = k < 0 ? k + g.s[m] : k;
  }
}

// LDS accumulator of the spreading: exact 64-bit fixed point (2^-32) for
// the mixed path, so the sum does not depend on the order of the adds;
// FP64 LDS atomics for the FP64 path.  The fixed-point value is
// round-to-nearest-even of v * 2^32: v * 2^32 + 1.5 * 2^52 in FP64 (the
// product is exact, the sum rounds once to an integer because
// |v| * 2^32 < 2^51), whose low mantissa bits minus those of 1.5 * 2^52 are
// the two's-complement integer.
template <typename R>
struct SpreadAcc;
template <>
struct SpreadAcc<float> {
  typedef unsigned long long T;
  static __device__ __forceinline__ void add(T *p, float v) {
    constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
    const double d = __fma_rn((double)v, 4294967296.0, kMagic);
    atomicAdd(p, (unsigned long long)__double_as_longlong(d) -
                     (unsigned long long)__double_as_longlong(kMagic));
  }
  static __device__ __forceinline__ float get(T v) {
    return (float)((double)(long long)v * (1.0 / 4294967296.0));
  }
};
template <>
struct SpreadAcc<double> {
  typedef double T;
  static __device__ __forceinline__ void add(T *p, double v) { atomicAdd(p, v); }
  static __device__ __forceinline__ double get(T v) { return v; }
};

constexpr int kMaxOrder = 8;

// one particle record in brick order: fraction (x, y, z) and charge
template <typename R>
struct Rec4 {
  R x, y, z, q;
};
__device__ __forceinline__ Rec4<float> load_rec(const float *__restrict__ p) {
  const float4 v = *reinterpret_cast<const float4 *>(p);
  return {v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ Rec4<double> load_rec(const double *__restrict__ p) {
  const double2 a = reinterpret_cast<const double2 *>(p)[0];
  const double2 b = reinterpret_cast<const double2 *>(p)[1];
  return {a.x, a.y, b.x, b.y};
}

}  // namespace
}  // namespace mfe
