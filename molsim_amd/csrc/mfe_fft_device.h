// mfe_fft_device.h -- device building blocks of the 3D FFT of SPMFFTRec
// (Src/energy.F90:6525-6587): complex arithmetic with compensated FP32
// twiddles, the LDS-staged mixed-radix Stockham transform, and the passes
// over one block's rows or tiles of one replica's mesh.  The pass kernels
// (mfe_fft.hip) run one pass over many blocks; the persistent batched step
// (mfe_kernels.hip) runs all passes of a replica in one workgroup.
#pragma once

#include <hip/hip_runtime.h>

#include "mfe_internal.h"

namespace mfe {
namespace {

template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename R>
__device__ __forceinline__ cplx<R> conjc(cplx<R> a) {
  return {a.x, -a.y};
}

// v * w (v * conj(w) with cj, the inverse transforms).  FP32 forward: the
// twiddle's residual enters the products through FMA, so the result is the
// product with the exact twiddle, rounded, and carries no magnitude bias
// (the reciprocal energy is a sum of squared forward outputs).  The inverse
// transforms feed only the forces, whose tolerance is far above that bias.
__device__ __forceinline__ cplx<float> twmul(cplx<float> v, const Twiddle<float> &w, bool cj) {
  if (cj) return {v.x * w.x + v.y * w.y, v.y * w.x - v.x * w.y};
  const float c1 = v.x * w.lx - v.y * w.ly, c2 = v.x * w.ly + v.y * w.lx;
  return {__builtin_fmaf(v.x, w.x, __builtin_fmaf(-v.y, w.y, c1)),
          __builtin_fmaf(v.x, w.y, __builtin_fmaf(v.y, w.x, c2))};
}
__device__ __forceinline__ cplx<double> twmul(cplx<double> v, const Twiddle<double> &w,
                                              bool cj) {
  const double wy = cj ? -w.y : w.y;
  return {v.x * w.x - v.y * wy, v.x * wy + v.y * w.x};
}

// x * sqrt(1/2): FP32 forward with the constant's rounding residual added
// by FMA (RN(sqrt(1/2)) is below sqrt(1/2))
__device__ __forceinline__ float mul_sqrt_half(float x, bool inv) {
  return inv ? x * 0.70710677f : __builtin_fmaf(x, 0.70710677f, x * 1.2101617e-08f);
}
__device__ __forceinline__ double mul_sqrt_half(double x, bool) {
  return x * 0.70710678118654752440;
}

constexpr int kMaxRadix = 32;

// e = threadIdx.x + k blockDim.x < n in batches of 8 per thread: the global
// loads of a batch are all issued before the first is stored to LDS, so a
// thread waits for one load latency per batch instead of per element
template <typename Load, typename Store>
__device__ __forceinline__ void batched_copy(int n, Load load, Store store) {
  constexpr int kU = 8;
  for (int e0 = threadIdx.x; e0 < n; e0 += kU * blockDim.x) {
    decltype(load(0)) v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int e = e0 + u * blockDim.x;
      if (e < n) v[u] = load(e);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int e = e0 + u * blockDim.x;
      if (e < n) store(e, v[u]);
    }
  }
}

// a twiddle table of n entries copied to LDS (read by every stage)
template <typename R>
__device__ __forceinline__ void stage_twiddles(Twiddle<R> *dst,
                                               const Twiddle<R> *__restrict__ src, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = src[k];
}

// R-point DFT of v[0..RAD-1] in registers: V[q] = sum_j v[j] w^(j q), with
// w = exp(-+2 pi i / RAD) taken from the length-n table tw (RAD divides n).
template <typename R, int RAD>
__device__ __forceinline__ void small_dft(cplx<R> *v, bool inv,
                                          const Twiddle<R> *tw, int n) {
  if constexpr (RAD == 2) {
    cplx<R> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (RAD == 4) {
    cplx<R> a = v[0], b = v[1], c = v[2], d = v[3];
    cplx<R> s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
    // forward: -i * s3 ; inverse: +i * s3
    cplx<R> js3 = inv ? cplx<R>{-s3.y, s3.x} : cplx<R>{s3.y, -s3.x};
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, js3);
    v[3] = csub(s1, js3);
  } else if constexpr (RAD == 8) {
    // radix 2 x 4: two 4-point DFTs of the even and odd samples, then the
    // odd half twiddled by w8^q (w8 = e^{-+i pi/4})
    cplx<R> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    small_dft<R, 4>(e, inv, tw, n);
    small_dft<R, 4>(o, inv, tw, n);
    // o[1] *= w8, o[2] *= w8^2 = -+i, o[3] *= w8^3
    const R sg = inv ? R(1) : R(-1);
    cplx<R> o1 = {mul_sqrt_half(o[1].x - sg * o[1].y, inv),
                  mul_sqrt_half(o[1].y + sg * o[1].x, inv)};
    cplx<R> o2 = {-sg * o[2].y, sg * o[2].x};
    cplx<R> o3 = {mul_sqrt_half(-o[3].x - sg * o[3].y, inv),
                  mul_sqrt_half(-o[3].y + sg * o[3].x, inv)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  } else {
    cplx<R> out[RAD];
    const int step = n / RAD;
#pragma unroll
    for (int qq = 0; qq < RAD; ++qq) {
      cplx<R> acc = v[0];
#pragma unroll
      for (int j = 1; j < RAD; ++j) acc = cadd(acc, twmul(v[j], tw[((j * qq) % RAD) * step], inv));
      out[qq] = acc;
    }
#pragma unroll
    for (int qq = 0; qq < RAD; ++qq) v[qq] = out[qq];
  }
}

// one Stockham stage of radix RAD over `batch` sequences (Govindaraju et
// al., SC'08): j-th butterfly reads in[j + q*n/RAD], twiddles by
// w^(k q) with k = j mod Ns, and writes out[(j/Ns)*Ns*RAD + k + q*Ns]
// lg_nbf / lg_ns: log2 of n/RAD and of Ns when both are powers of two
// (shifts replace the divisions), -1 otherwise.  k q tstep < n, so the
// twiddle index needs no reduction.
template <typename R, int RAD>
__device__ __forceinline__ void stockham_stage(const cplx<R> *a, cplx<R> *b,
                                               int n, int Ns, int batch,
                                               int ld, const Twiddle<R> *tw,
                                               bool inv, int lg_nbf, int lg_ns) {
  const int nbf = n / RAD;
  const int total = nbf * batch;
  const int tstep = n / (Ns * RAD);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    int sq, j, jb, k;
    if (lg_nbf >= 0) {
      sq = t >> lg_nbf;
      j = t & (nbf - 1);
      jb = j >> lg_ns;
      k = j & (Ns - 1);
    } else {
      sq = t / nbf;
      j = t - sq * nbf;
      jb = j / Ns;
      k = j - jb * Ns;
    }
    const cplx<R> *in = a + sq * ld;
    cplx<R> *out = b + sq * ld;
    cplx<R> v[RAD];
#pragma unroll
    for (int q = 0; q < RAD; ++q) v[q] = in[j + q * nbf];
#pragma unroll
    for (int q = 1; q < RAD; ++q) v[q] = twmul(v[q], tw[k * q * tstep], inv);
    small_dft<R, RAD>(v, inv, tw, n);
    const int base = jb * Ns * RAD + k;
#pragma unroll
    for (int q = 0; q < RAD; ++q) out[base + q * Ns] = v[q];
  }
}

// generic prime radix up to kMaxRadix (meshes with large prime factors)
template <typename R>
__device__ __forceinline__ void stockham_stage_generic(const cplx<R> *a, cplx<R> *b, int n,
                                       int Ns, int rad, int batch, int ld,
                                       const Twiddle<R> *tw, bool inv) {
  const int nbf = n / rad;
  const int total = nbf * batch;
  const int tstep = n / (Ns * rad);
  const int step = n / rad;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int sq = t / nbf, j = t - sq * nbf;
    const cplx<R> *in = a + sq * ld;
    cplx<R> *out = b + sq * ld;
    const int k = j % Ns;
    const int base = (j / Ns) * Ns * rad + k;
    for (int qq = 0; qq < rad; ++qq) {
      cplx<R> acc = {R(0), R(0)};
      for (int q = 0; q < rad; ++q) {
        acc = cadd(acc, twmul(in[j + q * nbf], tw[((k * q * tstep) + ((q * qq) % rad) * step) % n],
                              inv));
      }
      out[base + qq * Ns] = acc;
    }
  }
}

// Batched Stockham FFT staged in LDS: `batch` sequences of length p.n at
// a[s*ld + k]; b is scratch of the same shape.  Returns the buffer holding
// the result in natural order.  All threads of the block take part.
template <typename R>
__device__ __forceinline__ cplx<R> *lds_fft(cplx<R> *a, cplx<R> *b, const FftPlan &p,
                            int batch, int ld, const Twiddle<R> *tw, bool inv) {
  const int n = p.n;
  int Ns = 1;
  const bool pow2 = (n & (n - 1)) == 0;
  for (int st = 0; st < p.nstages; ++st) {
    const int r = p.radix[st];
    const int lg_nbf = pow2 ? __builtin_ctz(n / r) : -1;
    const int lg_ns = pow2 ? __builtin_ctz(Ns) : -1;
    switch (r) {
      case 8: stockham_stage<R, 8>(a, b, n, Ns, batch, ld, tw, inv, lg_nbf, lg_ns); break;
      case 4: stockham_stage<R, 4>(a, b, n, Ns, batch, ld, tw, inv, lg_nbf, lg_ns); break;
      case 2: stockham_stage<R, 2>(a, b, n, Ns, batch, ld, tw, inv, lg_nbf, lg_ns); break;
      case 3: stockham_stage<R, 3>(a, b, n, Ns, batch, ld, tw, inv, lg_nbf, lg_ns); break;
      case 5: stockham_stage<R, 5>(a, b, n, Ns, batch, ld, tw, inv, lg_nbf, lg_ns); break;
      case 7: stockham_stage<R, 7>(a, b, n, Ns, batch, ld, tw, inv, lg_nbf, lg_ns); break;
      default: stockham_stage_generic<R>(a, b, n, Ns, r, batch, ld, tw, inv); break;
    }
    __syncthreads();
    cplx<R> *tmp = a;
    a = b;
    b = tmp;
    Ns *= r;
  }
  return a;
}

// ----------------------------------------------------------------- x pass
// r2c of rows of length s1 = 2m: z[k] = x[2k] + i x[2k+1], Z = FFT_m(z),
// X[k] = (Z[k] + conj(Z[m-k]))/2 - i/2 (Z[k] - conj(Z[m-k])) W^k.
// Rows blk * fx_rpb, ... of the mesh; load(x, y, z) gives the real mesh
// value at a point.
// rows row0 .. row0 + nr - 1 (nr <= cap, the LDS buffers' row capacity)
template <typename R, typename Load>
__device__ __forceinline__ void fftx_fwd_rows(const SysDev &V, int row0, int nr, int cap,
                                              char *smem, Load load) {
  cplx<R> *__restrict__ cm = (cplx<R> *)V.mesh_c;
  const int s1 = V.sg.s[0], rows_per_blk = cap;
  const FftPlan &ph = V.plan_half;
  const Twiddle<R> *__restrict__ tw_half = (const Twiddle<R> *)V.tw_half;
  const Twiddle<R> *__restrict__ tw_full = (const Twiddle<R> *)V.tw[0];
  const int m = s1 / 2, hx = m + 1, ld = m + 1;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + rows_per_blk * ld;
  R *ar = reinterpret_cast<R *>(a);
  const int s2 = V.sg.s[1];
  Twiddle<R> *twh_s = reinterpret_cast<Twiddle<R> *>(b + rows_per_blk * ld), *twf_s = twh_s + m;
  stage_twiddles(twh_s, tw_half, m);
  stage_twiddles(twf_s, tw_full, s1);
  batched_copy(
      nr * s1,
      [&](int e) {
        const int r = fdiv(e, V.f_s1), x = e - r * s1;
        const int row = row0 + r, z = fdiv(row, V.f_s2), y = row - z * s2;
        return load(x, y, z);
      },
      [&](int e, R v) {
        const int r = fdiv(e, V.f_s1), x = e - r * s1;
        ar[2 * (r * ld) + x] = v;
      });
  __syncthreads();
  cplx<R> *z = lds_fft(a, b, ph, nr, ld, twh_s, false);
  for (int e = threadIdx.x; e < nr * hx; e += blockDim.x) {
    const int r = fdiv(e, V.f_hx), k = e - r * hx;
    // k in [0, m]: k mod m and (m - k) mod m without divisions
    cplx<R> zk = z[r * ld + (k == m ? 0 : k)];
    cplx<R> zm = conjc(z[r * ld + (k == 0 ? 0 : m - k)]);
    cplx<R> ev = {(zk.x + zm.x) * R(0.5), (zk.y + zm.y) * R(0.5)};
    cplx<R> d = csub(zk, zm);
    cplx<R> od = {d.y * R(0.5), -d.x * R(0.5)};
    cplx<R> X = cadd(ev, twmul(od, twf_s[k], false));
    cm[(size_t)(row0 + r) * hx + k] = X;
  }
}

// c2r of rows: Re(X[0]), Re(X[m]) only (FFTW's reading of a Hermitian row);
// Z[k] = (X[k] + conj(X[m-k])) + i (X[k] - conj(X[m-k])) W^-k, z = IFFT_m(Z),
// x[2j] = Re z[j], x[2j+1] = Im z[j] (unnormalised), into the real mesh.
template <typename R>
__device__ __forceinline__ void fftx_inv_rows(const SysDev &V, int row0, int nr, int cap,
                                              char *smem) {
  const cplx<R> *__restrict__ cm = (const cplx<R> *)V.mesh_c;
  R *__restrict__ real = (R *)V.mesh_r;
  const int s1 = V.sg.s[0], rows_per_blk = cap;
  const FftPlan &ph = V.plan_half;
  const Twiddle<R> *__restrict__ tw_half = (const Twiddle<R> *)V.tw_half;
  const Twiddle<R> *__restrict__ tw_full = (const Twiddle<R> *)V.tw[0];
  const int m = s1 / 2, hx = m + 1, ld = m + 1;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + rows_per_blk * ld;
  Twiddle<R> *twh_s = reinterpret_cast<Twiddle<R> *>(b + rows_per_blk * ld), *twf_s = twh_s + m;
  stage_twiddles(twh_s, tw_half, m);
  stage_twiddles(twf_s, tw_full, s1);
  batched_copy(
      nr * hx, [&](int e) { return cm[(size_t)row0 * hx + e]; },
      [&](int e, cplx<R> v) {
        const int r = fdiv(e, V.f_hx), k = e - r * hx;
        if (k == 0 || k == m) v.y = R(0);
        b[r * ld + k] = v;
      });
  __syncthreads();
  for (int e = threadIdx.x; e < nr * m; e += blockDim.x) {
    const int r = fdiv(e, V.f_m), k = e - r * m;
    cplx<R> xk = b[r * ld + k];
    cplx<R> xm = conjc(b[r * ld + (m - k)]);
    cplx<R> A = cadd(xk, xm), D = csub(xk, xm);
    cplx<R> Dw = twmul(D, twf_s[k], true);
    a[r * ld + k] = {A.x - Dw.y, A.y + Dw.x};
  }
  __syncthreads();
  cplx<R> *z = lds_fft(a, b, ph, nr, ld, twh_s, true);
  for (int e = threadIdx.x; e < nr * s1; e += blockDim.x) {
    const int r = fdiv(e, V.f_s1), x = e - r * s1;
    cplx<R> v = z[r * ld + (x >> 1)];
    real[(size_t)(row0 + r) * s1 + x] = (x & 1) ? v.y : v.x;
  }
}

// ----------------------------------------------------------------- y pass
// complex FFT along y of kx columns bx * fy_T, ... of plane z
template <typename R>
__device__ __forceinline__ void ffty_tile(const SysDev &V, int bx, int z, int inv, char *smem) {
  cplx<R> *__restrict__ cm = (cplx<R> *)V.mesh_c;
  const int s2 = V.sg.s[1], hx = V.hx, T = V.fy_T;
  const FftPlan &py = V.plan[1];
  const Twiddle<R> *__restrict__ tw = (const Twiddle<R> *)V.tw[1];
  const int ld = s2 + 1;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + T * ld;
  const int kx0 = bx * T;
  const int nt = min(T, hx - kx0);
  cplx<R> *base = cm + (size_t)z * s2 * hx + kx0;
  Twiddle<R> *tw_s = reinterpret_cast<Twiddle<R> *>(b + T * ld);
  stage_twiddles(tw_s, tw, s2);
  // elements e = y T + c over the full tile width T; columns c >= nt (the
  // last tile) are skipped
  batched_copy(
      T * s2,
      [&](int e) {
        const int y = fdiv(e, V.f_fyT), c = e - y * T;
        return c < nt ? base[(size_t)y * hx + c] : cplx<R>{R(0), R(0)};
      },
      [&](int e, cplx<R> v) {
        const int y = fdiv(e, V.f_fyT), c = e - y * T;
        if (c < nt) a[c * ld + y] = v;
      });
  __syncthreads();
  cplx<R> *res = lds_fft(a, b, py, nt, ld, tw_s, inv != 0);
  for (int e = threadIdx.x; e < T * s2; e += blockDim.x) {
    const int y = fdiv(e, V.f_fyT), c = e - y * T;
    if (c < nt) base[(size_t)y * hx + c] = res[c * ld + y];
  }
}

// ----------------------------------------------- z pass with the convolution
// For every stored element (kx, ky, kz) the reference's CalcGIF loop counts
// |F|^2 * meshfac(nx,ny,nz) once, halved when kx is 0 or s1/2 (its
// energyfac halvings along y and z only undo double visits), with
// nx = kx, ny = min(ky, s2-ky), nz = min(kz, s3-kz); the virial uses
// (1 + 2 k^2 expf) on top (Src/energy.F90:6553-6578, 6994-7018).
// Tile of kx columns bx * fz_T, ... of line y: z forward, Green function,
// z inverse; the thread's energy and virial terms are added to esum, vsum.
template <typename R>
__device__ __forceinline__ void fftconv_tile(const SysDev &V, int bx, int y, int energy,
                                             char *smem, double &esum, double &vsum) {
  cplx<R> *__restrict__ cm = (cplx<R> *)V.mesh_c;
  const int s1 = V.sg.s[0], s2 = V.sg.s[1], s3 = V.sg.s[2], hx = V.hx, T = V.fz_T;
  const FftPlan &pz = V.plan[2];
  const Twiddle<R> *__restrict__ tw = (const Twiddle<R> *)V.tw[2];
  const SpmeAxes &ax = V.ax;
  const int ld = s3 + 1;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + T * ld;
  const int kx0 = bx * T;
  const int nt = min(T, hx - kx0);
  const size_t zstride = (size_t)s2 * hx;
  cplx<R> *base = cm + (size_t)y * hx + kx0;
  Twiddle<R> *tw_s = reinterpret_cast<Twiddle<R> *>(b + T * ld);
  stage_twiddles(tw_s, tw, s3);
  batched_copy(
      T * s3,
      [&](int e) {
        const int z = fdiv(e, V.f_fzT), c = e - z * T;
        return c < nt ? base[z * zstride + c] : cplx<R>{R(0), R(0)};
      },
      [&](int e, cplx<R> v) {
        const int z = fdiv(e, V.f_fzT), c = e - z * T;
        if (c < nt) a[c * ld + z] = v;
      });
  __syncthreads();
  cplx<R> *f = lds_fft(a, b, pz, nt, ld, tw_s, false);
  const int ny = min(y, s2 - y);
  const double gy = ax.g[1][ny], kky = ax.kk[1][ny];
  for (int e = threadIdx.x; e < nt * s3; e += blockDim.x) {
    const int c = fdiv(e, V.f_s3), z = e - c * s3;
    const int nx = kx0 + c, nz = min(z, s3 - z);
    const double m2 = ax.kk[0][nx] + kky + ax.kk[2][nz];
    double mf = 0.0;
    if (m2 > 0.0) mf = ax.fac0 * ax.g[0][nx] * gy * ax.g[2][nz] / m2;
    cplx<R> v = f[c * ld + z];
    if (energy) {
      const double q2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
      const double w = (nx == 0 || 2 * nx == s1) ? 0.5 : 1.0;
      esum += w * q2 * mf;
      vsum -= w * q2 * (1.0 + 2.0 * m2 * ax.expf) * mf;
    }
    const R mfr = (R)mf;
    f[c * ld + z] = {v.x * mfr, v.y * mfr};
  }
  __syncthreads();
  cplx<R> *g = (f == a) ? b : a;
  cplx<R> *res = lds_fft(f, g, pz, nt, ld, tw_s, true);
  for (int e = threadIdx.x; e < T * s3; e += blockDim.x) {
    const int z = fdiv(e, V.f_fzT), c = e - z * T;
    if (c < nt) base[z * zstride + c] = res[c * ld + z];
  }
}

// The z forward transform alone (no Green function), for the MC trial
// moves' kept spectrum: tile of kx columns bx * fz_T, ... of line y, read
// from the complex mesh and written, transformed, to out (same layout) as
// values of type O.
template <typename R, typename O>
__device__ __forceinline__ void fftz_fwd_tile(const SysDev &V, int bx, int y, cplx<O> *out,
                                              char *smem) {
  const cplx<R> *__restrict__ cm = (const cplx<R> *)V.mesh_c;
  const int s2 = V.sg.s[1], s3 = V.sg.s[2], hx = V.hx, T = V.fz_T;
  const FftPlan &pz = V.plan[2];
  const Twiddle<R> *__restrict__ tw = (const Twiddle<R> *)V.tw[2];
  const int ld = s3 + 1;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + T * ld;
  const int kx0 = bx * T;
  const int nt = min(T, hx - kx0);
  const size_t zstride = (size_t)s2 * hx;
  const cplx<R> *base = cm + (size_t)y * hx + kx0;
  cplx<O> *obase = out + (size_t)y * hx + kx0;
  Twiddle<R> *tw_s = reinterpret_cast<Twiddle<R> *>(b + T * ld);
  stage_twiddles(tw_s, tw, s3);
  batched_copy(
      T * s3,
      [&](int e) {
        const int z = fdiv(e, V.f_fzT), c = e - z * T;
        return c < nt ? base[z * zstride + c] : cplx<R>{R(0), R(0)};
      },
      [&](int e, cplx<R> v) {
        const int z = fdiv(e, V.f_fzT), c = e - z * T;
        if (c < nt) a[c * ld + z] = v;
      });
  __syncthreads();
  const cplx<R> *f = lds_fft(a, b, pz, nt, ld, tw_s, false);
  for (int e = threadIdx.x; e < T * s3; e += blockDim.x) {
    const int z = fdiv(e, V.f_fzT), c = e - z * T;
    if (c < nt) {
      const cplx<R> v = f[c * ld + z];
      obase[z * zstride + c] = cplx<O>{(O)v.x, (O)v.y};
    }
  }
}

// ------------------------------------------- whole planes and lines (LDS)
// The same transforms over larger LDS tiles, for a workgroup that runs all
// passes of one replica: every kx column of planes z0 .. z0 + nz - 1 along
// y, of lines y0 .. y0 + ny - 1 along z.  In the complex mesh such a range
// is contiguous (planes) or nz runs of ny * hx elements (lines).
template <typename R>
__device__ __forceinline__ void ffty_planes(const SysDev &V, int z0, int nz, int inv,
                                            char *smem) {
  cplx<R> *__restrict__ base = (cplx<R> *)V.mesh_c + (size_t)z0 * V.sg.s[1] * V.hx;
  const int s2 = V.sg.s[1], hx = V.hx;
  const FftPlan &py = V.plan[1];
  const int ld = s2 + 1, nseq = nz * hx;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + nseq * ld;
  Twiddle<R> *tw_s = reinterpret_cast<Twiddle<R> *>(b + nseq * ld);
  stage_twiddles(tw_s, (const Twiddle<R> *)V.tw[1], s2);
  // element e = (zz s2 + y) hx + kx of the range
  batched_copy(
      nseq * s2, [&](int e) { return base[e]; },
      [&](int e, cplx<R> v) {
        const int q = fdiv(e, V.f_hx), kx = e - q * hx;
        const int zz = fdiv(q, V.f_s2), y = q - zz * s2;
        a[(zz * hx + kx) * ld + y] = v;
      });
  __syncthreads();
  cplx<R> *res = lds_fft(a, b, py, nseq, ld, tw_s, inv != 0);
  for (int e = threadIdx.x; e < nseq * s2; e += blockDim.x) {
    const int q = fdiv(e, V.f_hx), kx = e - q * hx;
    const int zz = fdiv(q, V.f_s2), y = q - zz * s2;
    base[e] = res[(zz * hx + kx) * ld + y];
  }
}

template <typename R>
__device__ __forceinline__ void fftconv_lines(const SysDev &V, int y0, int ny, int energy,
                                              char *smem, double &esum, double &vsum) {
  cplx<R> *__restrict__ cm = (cplx<R> *)V.mesh_c;
  const int s1 = V.sg.s[0], s2 = V.sg.s[1], s3 = V.sg.s[2], hx = V.hx;
  const FftPlan &pz = V.plan[2];
  const SpmeAxes &ax = V.ax;
  const int ld = s3 + 1, nseq = ny * hx;
  cplx<R> *a = reinterpret_cast<cplx<R> *>(smem);
  cplx<R> *b = a + nseq * ld;
  Twiddle<R> *tw_s = reinterpret_cast<Twiddle<R> *>(b + nseq * ld);
  stage_twiddles(tw_s, (const Twiddle<R> *)V.tw[2], s3);
  const size_t zstride = (size_t)s2 * hx;
  cplx<R> *base = cm + (size_t)y0 * hx;
  // element e = z nseq + (yy hx + kx): sequence c = yy hx + kx, position z
  batched_copy(
      nseq * s3,
      [&](int e) {
        const int z = e / nseq, c = e - z * nseq;
        return base[z * zstride + c];
      },
      [&](int e, cplx<R> v) {
        const int z = e / nseq, c = e - z * nseq;
        a[c * ld + z] = v;
      });
  __syncthreads();
  cplx<R> *f = lds_fft(a, b, pz, nseq, ld, tw_s, false);
  for (int e = threadIdx.x; e < nseq * s3; e += blockDim.x) {
    const int c = fdiv(e, V.f_s3), z = e - c * s3;
    const int yy = fdiv(c, V.f_hx), nx = c - yy * hx;
    const int y = y0 + yy, nyy = min(y, s2 - y), nz = min(z, s3 - z);
    const double m2 = ax.kk[0][nx] + ax.kk[1][nyy] + ax.kk[2][nz];
    double mf = 0.0;
    if (m2 > 0.0) mf = ax.fac0 * ax.g[0][nx] * ax.g[1][nyy] * ax.g[2][nz] / m2;
    cplx<R> v = f[c * ld + z];
    if (energy) {
      const double q2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
      const double w = (nx == 0 || 2 * nx == s1) ? 0.5 : 1.0;
      esum += w * q2 * mf;
      vsum -= w * q2 * (1.0 + 2.0 * m2 * ax.expf) * mf;
    }
    const R mfr = (R)mf;
    f[c * ld + z] = {v.x * mfr, v.y * mfr};
  }
  __syncthreads();
  cplx<R> *g = (f == a) ? b : a;
  cplx<R> *res = lds_fft(f, g, pz, nseq, ld, tw_s, true);
  for (int e = threadIdx.x; e < nseq * s3; e += blockDim.x) {
    const int z = e / nseq, c = e - z * nseq;
    base[z * zstride + c] = res[c * ld + z];
  }
}

}  // namespace
}  // namespace mfe
