// mfe_kernels.hip -- device kernels of the force engine except the FFT.
//
//   neighbour list ...... SetVListMD / NList skin test (Src/nlist.F90)
//   pair forces ......... UTwoBodyA / UTwoBodyACellList (Src/energy.F90)
//   SPME ................ UEwaldRecSPM spreading and gather
//   standard Ewald ...... UEwaldRecStd
//   bonds ............... UBond
//   integrator .......... VelVer (Src/md.F90), BDStep (Src/bd.F90)
//
// No global floating-point atomics anywhere: charge spreading owns its
// output bricks and accumulates in LDS; energies and virials are reduced
// per block into partial sums that one block adds up in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "mfe_device.h"
#include "mfe_fft_device.h"
#include "mfe_internal.h"

namespace mfe {
namespace {

__global__ void k_fill_int(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_double(double *p, int n, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Exclusive scan of count[0..n) into start[0..n], zeroing count for the
// next use; one block of 1024 threads per replica, each thread owning a
// contiguous chunk; the chunk sums are scanned inside each wavefront with
// shuffles and across the 16 wavefronts through LDS (one barrier).
// which: 0 = cells of the neighbour list, 1 = SPME bricks.
__device__ __forceinline__ void ph_scan(int *count, int *start, int n) {
  __shared__ int wsum[16];
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int lo = threadIdx.x * per, hi = min(n, lo + per);
  // the chunk in batches of 16 loads in flight
  constexpr int kU = 16;
  int s = 0;
  for (int i0 = lo; i0 < hi; i0 += kU) {
    int v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = i0 + u < hi ? count[i0 + u] : 0;
#pragma unroll
    for (int u = 0; u < kU; ++u) s += v[u];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = s;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int wpre = 0;
  for (int w = 0; w < wave; ++w) wpre += wsum[w];
  int run = wpre + x - s;
  for (int i0 = lo; i0 < hi; i0 += kU) {
    int v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = i0 + u < hi ? count[i0 + u] : 0;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (i0 + u < hi) {
        start[i0 + u] = run;
        count[i0 + u] = 0;
      }
      run += v[u];
    }
  }
  if (threadIdx.x == blockDim.x - 1) start[n] = wpre + x;
}

// The same scan over all blocks of a grid (k_rebuild): each block scans its
// contiguous range into start[] and leaves its total in bsum[block]; after
// a grid barrier each block adds the totals of the blocks before it.
__device__ __forceinline__ void ph_scan_local(const int *count, int *start, int n, int *bsum) {
  __shared__ int wsum[16];
  const int nb = gridDim.x, C = (n + nb - 1) / nb;
  const int lo = min(n, blockIdx.x * C), hi = min(n, lo + C);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int carry = 0;
  for (int base = lo; base < hi; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const int v = i < hi ? count[i] : 0;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int wpre = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
      wpre += w < wave ? wsum[w] : 0;
      tot += wsum[w];
    }
    if (i < hi) start[i] = carry + wpre + x - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = carry;
}

__device__ __forceinline__ void ph_scan_offset(int *count, int *start, int n, const int *bsum) {
  __shared__ int wsum[16];
  const int nb = gridDim.x, C = (n + nb - 1) / nb;
  const int lo = min(n, blockIdx.x * C), hi = min(n, lo + C);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // totals of the blocks before this one (nb <= blockDim), in a fixed order
  int v = 0;
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += blockDim.x) v += bsum[k];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  int off = 0;
  for (int w = 0; w < nw; ++w) off += wsum[w];
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    start[i] += off;
    count[i] = 0;
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) start[n] = off + bsum[nb - 1];
}

__global__ void __launch_bounds__(1024) k_scan(const SysDev *__restrict__ sys,
                                               int which, int gated) {
  const SysDev &V = sys[blockIdx.z];
  if (gated && V.flags[1] == 0) return;
  if (which) ph_scan(V.b_count, V.b_start, V.nbricks);
  else ph_scan(V.c_count, V.c_start, V.cg.ncells);
}

// ------------------------------------------------------- neighbour lists
// Cells of edge >= rcut + drnlist; a particle's cell is
// floor((x + L/2) * cinv), clamped to the grid.
__device__ __forceinline__ int cell_of(const double *r, const CellGeom &g,
                                       const Box &b, int c[3]) {
  for (int m = 0; m < 3; ++m) {
    int k = (int)floor((r[m] + b.L2[m]) * g.cinv[m]);
    c[m] = min(max(k, 0), g.n[m] - 1);
  }
  return c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
}

// The mixed path's FP32 position of a particle at x (FP64) in cell c,
// relative to the cell's reference point -L/2 + (c + 1/2) hf per axis and
// taken into [-L/2, L/2] (the particle may have crossed the box edge since
// the list was built), with its type in the fourth word.
__device__ __forceinline__ float4 rel_pos(const SysDev &V, const double *x, int c, int type) {
  const CellGeom &g = V.cg;
  const int cc[3] = {c % g.n[0], (c / g.n[0]) % g.n[1], c / (g.n[0] * g.n[1])};
  float r[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double o = -V.box.L2[m] + (cc[m] + 0.5) * g.hf[m];
    double d = x[m] - o;
    if (d > V.box.L2[m]) d -= V.box.L[m];
    else if (d < -V.box.L2[m]) d += V.box.L[m];
    r[m] = (float)d;
  }
  return make_float4(r[0], r[1], r[2], __int_as_float(type));
}

// The rebuild's phases, each over threads t0, t0 + nt, ...: run as five
// kernels (batches, explicit builds) or as the phases of k_rebuild.
__device__ __forceinline__ void ph_cell_count(const SysDev &V, int t0, int nt) {
  const double *__restrict__ ro = V.ro;
  const CellGeom g = V.cg;
  const Box b = V.box;
  for (int i = t0; i < V.np; i += nt) {
    int c3[3];
    const int c = cell_of(ro + 3 * i, g, b, c3);
    V.c_of[i] = c;
    V.c_slot[i] = atomicAdd(&V.c_count[c], 1);
  }
}

__device__ __forceinline__ void ph_cell_scatter(const SysDev &V, int t0, int nt) {
  for (int i = t0; i < V.np; i += nt) V.c_sorted[V.c_start[V.c_of[i]] + V.c_slot[i]] = i;
}

// members of each cell in ascending particle index (deterministic lists)
__device__ __forceinline__ void ph_cell_sort(const SysDev &V, int t0, int nt) {
  const int *start = V.c_start;
  int *sorted = V.c_sorted;
  for (int c = t0; c < V.cg.ncells; c += nt) {
    const int lo = start[c], hi = start[c + 1];
    for (int i = lo + 1; i < hi; ++i) {
      int v = sorted[i], j = i - 1;
      while (j >= lo && sorted[j] > v) {
        sorted[j + 1] = sorted[j];
        --j;
      }
      sorted[j + 1] = v;
    }
  }
}

// positions and types in cell order (ro_s, type_s)
__device__ __forceinline__ void ph_sort_pos(const SysDev &V, int t0, int nt) {
  for (int t = t0; t < V.np; t += nt) {
    const int i = V.c_sorted[t];
    // every load before the first store
    const double x = V.ro[3 * i], y = V.ro[3 * i + 1], z = V.ro[3 * i + 2];
    const int ty = V.type[i];
    V.ro_s[3 * t] = x;
    V.ro_s[3 * t + 1] = y;
    V.ro_s[3 * t + 2] = z;
    V.type_s[t] = ty;
    if (V.pos_f) {
      const double r[3] = {x, y, z};
      V.pos_f[t] = rel_pos(V, r, V.c_of[i], ty);
    }
  }
}

// Full Verlet list: every j != i with r2 < (rcut+drnlist)^2 (SetVListMD's
// test, Src/nlist.F90:746-750).  Threads take particles in cell order
// (t -> i = sorted[t]) and store partners by their cell-order position, so
// the pair kernel reads neighbouring positions from neighbouring memory;
// the list is column-major, nbr[k*np + t].  The half list the reference
// keeps is the checkerboard subset of this set.  POS_S: the cell-ordered
// positions are already in ro_s (read there, contiguously); else they are
// read through sorted[] and ro_s / type_s are written here.
template <bool POS_S>
__device__ __forceinline__ void ph_nlist_build(const SysDev &V, int t0, int nt) {
  const double *__restrict__ ro = V.ro;
  const double *__restrict__ ros = V.ro_s;
  const int np = V.np;
  const CellGeom g = V.cg;
  const Box b = V.box;
  const int *__restrict__ cof = V.c_of;
  const int *__restrict__ start = V.c_start;
  const int *__restrict__ sorted = V.c_sorted;
  const double rs2 = V.rs2;
  const int maxnn = V.maxnn;
  int *__restrict__ nn = V.nn;
  int *__restrict__ nbr = V.nbr;
  int *flags = V.flags;
  double *drosum = V.drosum;
  for (int t = t0; t < np; t += nt) {
    const int i = sorted[t];
    const double xi = POS_S ? ros[3 * t] : ro[3 * i];
    const double yi = POS_S ? ros[3 * t + 1] : ro[3 * i + 1];
    const double zi = POS_S ? ros[3 * t + 2] : ro[3 * i + 2];
    const int c = cof[i];
    const int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
    const int rx = g.n[0] > 1 ? 1 : 0, ry = g.n[1] > 1 ? 1 : 0, rz = g.n[2] > 1 ? 1 : 0;
    int n = 0;
    for (int dz = -rz; dz <= rz; ++dz) {
      const int z = (cz + dz + g.n[2]) % g.n[2];
      // the pair's offset code per axis (see CellGeom): -d inside the box,
      // +-2 across its edge, 0 on an axis of one cell
      const int kz = rz ? (cz + dz < 0 ? 2 : (cz + dz >= g.n[2] ? -2 : -dz)) : 0;
      for (int dy = -ry; dy <= ry; ++dy) {
        const int y = (cy + dy + g.n[1]) % g.n[1];
        const int ky = ry ? (cy + dy < 0 ? 2 : (cy + dy >= g.n[1] ? -2 : -dy)) : 0;
        // the row's cells x-1, x, x+1 (one on an axis of one cell): their
        // bounds loaded together, then their particles taken as one
        // sequence in the same order (cell by cell, p ascending), four
        // positions in flight at a time
        int cs[3] = {0, 0, 0}, cn[3] = {0, 0, 0};
        unsigned cd[3] = {0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int dx = k - 1;
          if (rx || dx == 0) {
            const int x = (cx + dx + g.n[0]) % g.n[0];
            const int kx = rx ? (cx + dx < 0 ? 2 : (cx + dx >= g.n[0] ? -2 : -dx)) : 0;
            cd[k] = (unsigned)((kx + 2) + 5 * (ky + 2) + 25 * (kz + 2)) << kNbrCodeShift;
            const int cc = x + g.n[0] * (y + g.n[1] * z);
            cs[k] = start[cc];
            cn[k] = start[cc + 1];
          }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) cn[k] -= cs[k];
        const int n01 = cn[0] + cn[1], tot = n01 + cn[2];
        for (int q0 = 0; q0 < tot; q0 += 4) {
          int pp[4], kk[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int q = q0 + u;
            const int k = q < cn[0] ? 0 : (q < n01 ? 1 : 2);
            kk[u] = k;
            // past the row's end: the particle itself (skipped below)
            pp[u] = q < tot ? cs[k] + q - (k == 0 ? 0 : (k == 1 ? cn[0] : n01)) : t;
          }
          double xj[4], yj[4], zj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (POS_S) {
              xj[u] = ros[3 * pp[u]];
              yj[u] = ros[3 * pp[u] + 1];
              zj[u] = ros[3 * pp[u] + 2];
            } else {
              const int j = sorted[pp[u]];
              xj[u] = ro[3 * j];
              yj[u] = ro[3 * j + 1];
              zj[u] = ro[3 * j + 2];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (pp[u] == t) continue;
            double ddx = xi - xj[u], ddy = yi - yj[u], ddz = zi - zj[u];
            const double r2 = pbcr2(ddx, ddy, ddz, b);
            if (r2 < rs2) {
              const unsigned code = kk[u] == 0 ? cd[0] : (kk[u] == 1 ? cd[1] : cd[2]);
              if (n < maxnn) nbr[(size_t)n * np + t] = (int)((unsigned)pp[u] | code);
              ++n;
            }
          }
        }
      }
    }
    // on overflow only the stored maxnn partners are visible to the pair
    // kernel; the call reports MFE_E_NEIGH_CAPACITY
    nn[t] = min(n, maxnn);
    if (n > maxnn) atomicOr(&flags[0], ERR_CAPACITY);
    // cell-ordered positions and types for the pair kernel; k_move keeps
    // them current through c_pos until the next rebuild
    V.c_pos[i] = t;
    if (!POS_S) {
      V.ro_s[3 * t] = xi;
      V.ro_s[3 * t + 1] = yi;
      V.ro_s[3 * t + 2] = zi;
      const int ty = V.type[i];
      V.type_s[t] = ty;
      if (V.pos_f) {
        const double r[3] = {xi, yi, zi};
        V.pos_f[t] = rel_pos(V, r, c, ty);
      }
    }
    if (drosum) {
      drosum[3 * i] = 0.0;
      drosum[3 * i + 1] = 0.0;
      drosum[3 * i + 2] = 0.0;
    }
  }
}

#define MFE_GATE                                \
  const SysDev &V = sys[blockIdx.z];           \
  if (gated && V.flags[1] == 0) return;        \
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;

__global__ void k_cell_count(const SysDev *__restrict__ sys, int gated) {
  MFE_GATE
  ph_cell_count(V, t0, nt);
}

__global__ void k_cell_scatter(const SysDev *__restrict__ sys, int gated) {
  MFE_GATE
  ph_cell_scatter(V, t0, nt);
}

__global__ void k_cell_sort(const SysDev *__restrict__ sys, int gated) {
  MFE_GATE
  ph_cell_sort(V, t0, nt);
}

__global__ void __launch_bounds__(kBlock) k_nlist_build(const SysDev *__restrict__ sys,
                                                        int gated) {
  MFE_GATE
  ph_nlist_build<false>(V, t0, nt);
  if (t0 == 0 && gated) V.flags[2] += 1;
}
#undef MFE_GATE

// Grid-wide barrier of k_rebuild, whose blocks are all resident (one per
// CU), in two levels: the blocks arrive on kBarGroups group counters
// (block b on counter b % kBarGroups, each counter on a line of its own,
// so arrivals on different counters do not queue behind one another), the
// last block of each group on the top counter, and the last of those
// advances the generation.  Each wave waits for its stores (in L2), the
// block meets at s_barrier; thread 0 then arrives with acquire-release
// increments (a release writes the XCD's L2 back, so the other XCDs see
// the block's data; a group's last block has acquired its group's data
// before it releases on the top counter), the last arrival at each level
// resets that level's counter, the final one advances the generation
// (release), the others poll it (relaxed) and take one acquire fence,
// which invalidates the CU's stale lines for the whole block.  The words
// are laid out as kBarWords in mfe_internal.h.  The wait is bounded: after `spin` polls a
// block raises ERR_SYNC and leaves.  A block goes on to the next phase
// only when every block has arrived (so the data of the finished phase is
// complete) and no block has given up; otherwise it returns false, the
// kernel ends in every block, and the host reports the error, clears the
// barrier and the cell counts, and rebuilds the list from scratch.
constexpr int kGridSpin = 1 << 22;

__device__ __forceinline__ bool grid_sync(unsigned *bar, int *flags, int spin) {
  __shared__ int ok_s;
  // every wave waits for its own stores to be acknowledged (in L2) before
  // the block barrier (a workgroup-scope barrier does not wait for them)
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned nblk = gridDim.x;
    const unsigned ng = nblk < (unsigned)kBarGroups ? nblk : (unsigned)kBarGroups;
    const unsigned grp = blockIdx.x % ng;
    const unsigned ingrp = (nblk - 1u - grp) / ng + 1u;  // blocks of this group
    unsigned *gc = bar + grp * kBarStride;
    unsigned *gen_p = bar + kBarGen, *top = bar + kBarTop;
    const unsigned gen = __hip_atomic_load(gen_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool final_block = false;
    if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == ingrp) {
      __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == ng) {
        __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(gen_p, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        final_block = true;
      }
    }
    if (!final_block) {
      int spins = 0;
      while (__hip_atomic_load(gen_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (++spins > spin) {
          atomicOr(&flags[0], ERR_SYNC);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (ok && (__hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
               ERR_SYNC))
      ok = 0;
    ok_s = ok;
  }
  __syncthreads();
  return ok_s != 0;
}

// The list rebuild's phases in order, separated by grid barriers, run by
// every block of a launch whose blocks are all resident (one 1024-thread
// block per CU).  A block leaves as soon as a barrier gives up.
__device__ __forceinline__ void rebuild_phases(const SysDev &V, unsigned *bar, int *bsum,
                                               int spin) {
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  ph_cell_count(V, t0, nt);
  if (!grid_sync(bar, V.flags, spin)) return;
  ph_scan_local(V.c_count, V.c_start, V.cg.ncells, bsum);
  if (!grid_sync(bar, V.flags, spin)) return;
  ph_scan_offset(V.c_count, V.c_start, V.cg.ncells, bsum);
  if (!grid_sync(bar, V.flags, spin)) return;
  ph_cell_scatter(V, t0, nt);
  if (!grid_sync(bar, V.flags, spin)) return;
  ph_cell_sort(V, t0, nt);
  if (!grid_sync(bar, V.flags, spin)) return;
  ph_sort_pos(V, t0, nt);
  if (!grid_sync(bar, V.flags, spin)) return;
  ph_nlist_build<true>(V, t0, nt);
  if (t0 == 0) V.flags[2] += 1;
}

// The gated list rebuild of one engine as one kernel: it returns at once
// when the skin test did not fire (one launch per step instead of five),
// or after a failed rebuild earlier in the same call (the list is void
// until the host rebuilds it); otherwise the phases run.  Launched with
// one 1024-thread block per CU, so all blocks are resident.
// gate: the engine's flags (V.flags), passed directly so that the skin
// decision is one load away rather than behind the view's.
__global__ void __launch_bounds__(1024) k_rebuild(const SysDev *__restrict__ sys,
                                                  const int *__restrict__ gate, unsigned *bar,
                                                  int *bsum, int spin) {
  if (gate[1] == 0 || (gate[0] & ERR_SYNC)) return;
  rebuild_phases(sys[0], bar, bsum, spin);
}

// positions and types in cell order for the pair kernel (every UTotal)
__global__ void k_sort_pos(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  ph_sort_pos(V, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// ------------------------------------------------------------ pair forces
// First interval (in storage order) with r2 >= zlow -- the interval the
// reference's linear walk stops at (Src/energy.F90:482-486).  zlow
// decreases along the table; the walk starts at the interval kb of r2's
// bin and moves to lower r2 while r2 < zlow, then to higher r2 while the
// next bound is still <= r2.  Both moves together reach the first k with
// r2 >= zl(k) from any start, so the bin only sets the walk's length
// (0-2 steps for bins about as wide as the intervals).
template <typename Z>
__device__ __forceinline__ int find_interval(const Z &zl, int n, double r2, int kb) {
  int k = kb;
  while (k < n - 1 && !(r2 >= zl(k))) ++k;
  while (k > 0 && r2 >= zl(k - 1)) --k;
  return k;
}

// list entries a lane loads together (indices, then positions)
constexpr int kPairBatch = 4;

// k_skin_blocks' computation for skin blocks sb0 and sb0 + 1 (256
// particles each, the k_move partition), one per half of a 512-thread
// block: the same maxima in the same order, so the same candidates.
__device__ __forceinline__ void skin_two_blocks(const SysDev &V, int sb0) {
  __shared__ double wred[kPairBlock / 64];
  const int sub = threadIdx.x >> 8, t = threadIdx.x & 255;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, w0 = 4 * sub;
  const int sb = sb0 + sub;
  const bool valid = sb < V.nblk_np;
  const double *part_max = V.part + V.off_move;
  // running maximum entering the block: the maxima of the earlier blocks
  double pm = 0.0;
  if (valid)
    for (int b2 = t; b2 < sb; b2 += 256) pm = fmax(pm, part_max[b2]);
  for (int w = 32; w > 0; w >>= 1) pm = fmax(pm, __shfl_down(pm, w, 64));
  if (lane == 0) wred[wv] = pm;
  __syncthreads();
  const double P = fmax(fmax(wred[w0], wred[w0 + 1]), fmax(wred[w0 + 2], wred[w0 + 3]));
  __syncthreads();
  const int i = sb * 256 + t;
  const double v = (valid && i < V.np) ? V.r2disp[i] : 0.0;
  // inclusive max-scan inside the wavefront, then the earlier wavefronts
  double inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const double y = __shfl_up(inc, off, 64);
    if (lane >= off) inc = fmax(inc, y);
  }
  double excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 0.0;
  if (lane == 63) wred[wv] = inc;
  __syncthreads();
  double pw = 0.0;
  for (int k = w0; k < wv; ++k) pw = fmax(pw, wred[k]);
  const double before = fmax(P, fmax(pw, excl));
  double cand = (valid && i < V.np && !(v > before)) ? v : 0.0;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) cand = fmax(cand, __shfl_down(cand, w, 64));
  if (lane == 0) wred[wv] = cand;
  __syncthreads();
  if (t == 0 && valid)
    V.part[V.off_skin + sb] =
        fmax(fmax(wred[w0], wred[w0 + 1]), fmax(wred[w0 + 2], wred[w0 + 3]));
}

// Pair forces over the Verlet list.  kPairLanes lanes share one particle
// (cell order i) and split its list k = lane, lane + kPairLanes, ...; their
// partial forces are combined by shuffles in FP64 in a fixed order, so
// forces need no atomics and stay deterministic.  Every pair is seen from
// both ends, so its energy and virial carry weight 1/2.  semantics: 0 =
// vlist (r2 <= rcut2 kept), 1 = cell list (r2 < rcut2 kept).
//
// The grid is persistent: nblk_pair (<= kPairGridMax) blocks each stage
// the tables in LDS once -- r2 bins, interval bounds and, with CF, the
// polynomial coefficients -- and then walk one contiguous range of chunks
// of kPairBlock / kPairLanes particles.  A lane loads the indices of
// kPairBatch of its entries together, then their positions, then
// evaluates them.  NT bounds the number of types: up to 4 the energies
// accumulate per pair type in registers over all chunks, else per partner
// type, converted per chunk into a per-wavefront LDS row; either way in a
// fixed order, and per block into the partials.
// With skin (MD/BD steps) blocks from pgrid on take the skin test's block
// candidates (two k_move blocks each) instead of a separate launch.
template <int NT, bool CF>
__global__ void __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_pair(const SysDev *__restrict__ sys, int energy, int pgrid, int skin) {
  using CT = double;
  const SysDev &V = sys[blockIdx.z];
  if (skin && (int)blockIdx.x >= pgrid) {
    skin_two_blocks(V, 2 * ((int)blockIdx.x - pgrid));
    return;
  }
  const int blk = xcd_block(blockIdx.x, pgrid);
  const int nb = V.nblk_pair;
  if (blk >= nb) return;
  const double *__restrict__ ro = V.ro_s;
  const int *__restrict__ type = V.type_s;
  const int np = V.np;
  const int *__restrict__ nn = V.nn;
  const int *__restrict__ nbr = V.nbr;
  const Box b = V.box;
  const TableDev t = V.tab;
  const int semantics = V.nlsem;
  int *flags = V.flags;
  extern __shared__ __align__(16) unsigned char pair_lds[];
  __shared__ PairRec rec_s[MFE_MAX_PAIRS];
  __shared__ int pidx_s[MFE_MAX_TYPES * MFE_MAX_TYPES];
  __shared__ double wacc[kPairBlock / 64][MFE_MAX_PAIRS + 1];
  const int P = t.nptpt, nbin = t.nbin, nzl = t.nzl, npt = t.npt;
  const bool lds_zl = CF || t.lds_mode >= 1;
  const size_t bb = pair_bin_bytes(P, nbin);
  unsigned short *bin_s = reinterpret_cast<unsigned short *>(pair_lds);
  double *zl_s = reinterpret_cast<double *>(pair_lds + bb);
  CT *cf_s = reinterpret_cast<CT *>(pair_lds + bb + pair_zl_bytes(nzl));
  for (int k = threadIdx.x; k < P * nbin; k += blockDim.x) bin_s[k] = t.bin_k[k];
  if (lds_zl)
    for (int k = threadIdx.x; k < nzl; k += blockDim.x) zl_s[k] = t.zlow[k];
  if (CF) {
    for (int k = threadIdx.x; k < nzl * kCoefStride; k += blockDim.x) cf_s[k] = t.coef[k];
  }
  for (int k = threadIdx.x; k < P; k += blockDim.x) rec_s[k] = t.rec[k];
  for (int k = threadIdx.x; k < npt * npt; k += blockDim.x) pidx_s[k] = t.pairidx[k];
  const int l64 = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = l64; k <= P; k += 64) wacc[wv][k] = 0.0;
  __syncthreads();
  // the 12 stored coefficients of interval iv (16-byte loads)
  auto coefs = [&](int iv, CT *c) {
    const double2 *q = CF ? reinterpret_cast<const double2 *>(cf_s) + 6 * iv
                          : reinterpret_cast<const double2 *>(t.coef) + 6 * iv;
#pragma unroll
    for (int h = 0; h < 6; ++h) {
      const double2 v = q[h];
      c[2 * h] = v.x;
      c[2 * h + 1] = v.y;
    }
  };
  auto zlow = [&](int idx) -> double {
    if constexpr (CF) return zl_s[idx];
    else return lds_zl ? zl_s[idx] : t.zlow[idx];
  };
  const int lane = threadIdx.x % kPairLanes;
  // energy accumulators: per pair type for up to 4 types (kept over all
  // chunks), else per partner type (converted to pair types per chunk)
  constexpr bool kRegE = NT <= 4;
  constexpr int NA = kRegE ? NT * (NT + 1) / 2 : NT;
  double ej[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) ej[k] = 0.0;
  double vir = 0.0;
  double fx, fy, fz;
  int ti = 0;
  // pair index of this particle with partner type u in bits 8u..8u+7
  unsigned long long ptp = 0;
  int err = 0;    // error bits, flagged once at the end
  // one list entry: partner of type tj at (dx, dy, dz) = r_i - r_j.  The
  // LDS reads come before any branch, so the entries of a batch overlap.
  auto eval = [&](double dx, double dy, double dz, int tj) {
    const double r2 = pbcr2(dx, dy, dz, b);
    const bool in = semantics == 0 ? !(r2 > t.rcut2) : (r2 < t.rcut2);
    const int p = (int)((ptp >> (8 * tj)) & 0xffu);
    const PairRec R = rec_s[p];
    // interval of r2 (of lo below the table: such an entry is an error
    // or a hard core, and only its flags count)
    const double r2c = fmax(r2, R.lo);
    const float sf = __builtin_amdgcn_rcpf((float)r2c);
    const float xb = fminf(fmaxf((sf - R.s_lo) * R.s_inv, 0.0f), (float)(nbin - 1));
    const int bk = bin_s[p * nbin + (int)xb];
    const int o = R.off, n = R.nint, kb = bk & (kSlowBin - 1);
    const double z0 = zlow(o + kb), z1 = zlow(o + min(kb + 1, n - 1));
    int iv;
    double zv;
    if (bk & kSlowBin) {
      auto zl = [&](int k) -> double { return zlow(o + k); };
      iv = find_interval(zl, n, r2c, kb);
      zv = zlow(o + iv);
    } else {
      const bool at = r2c >= z0;
      iv = at ? kb : min(kb + 1, n - 1);
      zv = at ? z0 : z1;
    }
    CT c[kCoefStride];
    coefs(o + iv, c);
    const bool hc = r2 < R.r2at;
    const bool ov = !hc && r2 < R.r2um;
    const bool te = !hc && !ov && !(r2 >= R.lo);
    if (in) err |= (ov ? ERR_OVERLAP : 0) | (te ? ERR_TABLE : 0);
    if (!in || ov || te) return;
    const double d = r2 - zv;
    double fsum = c[6] + d * (c[7] + d * (c[8] + d * (c[9] + d * c[10])));
    double usum = c[0] + d * (c[1] + d * (c[2] + d * (c[3] + d * (c[4] + d * c[5]))));
    if (hc) {
      usum = 1e10;
      fsum = 1e10;
    } else if (V.nknot) {
      // additive spline term: u += sum c w, fsum -= u'/r
      const double r = sqrt(r2);
      int ki;
      double w[4], dw[4];
      if (knot_weights(r, V.knot_r0, V.knot_dinv, V.nknot, ki, w, dw)) {
        const double *kc = V.knot_c + (size_t)p * V.nknot + ki;
        const double du = kc[0] * w[0] + kc[1] * w[1] + kc[2] * w[2] + kc[3] * w[3];
        const double dup = (kc[0] * dw[0] + kc[1] * dw[1] + kc[2] * dw[2] + kc[3] * dw[3]) *
                           V.knot_dinv;
        usum += du;
        fsum -= dup / r;
      }
    }
    fx += fsum * dx;
    fy += fsum * dy;
    fz += fsum * dz;
    if (energy) {
#pragma unroll
      for (int k2 = 0; k2 < NA; ++k2) ej[k2] += (k2 == (kRegE ? p : tj)) ? 0.5 * usum : 0.0;
      vir -= 0.5 * fsum * r2;
    }
  };
  // this block's chunks: an even share of the contiguous chunk range
  constexpr int kChunk = kPairBlock / kPairLanes;
  const int nch = (np + kChunk - 1) / kChunk;
  const int ch0 = (int)((long long)blk * nch / nb), ch1 = (int)((long long)(blk + 1) * nch / nb);
  for (int ch = ch0; ch < ch1; ++ch) {
    const int i = ch * kChunk + threadIdx.x / kPairLanes;
    if (!kRegE) {
#pragma unroll
      for (int k = 0; k < NA; ++k) ej[k] = 0.0;
    }
    fx = fy = fz = 0.0;
    ti = 0;
    if (i < np) {
      const double xi = ro[3 * i], yi = ro[3 * i + 1], zi = ro[3 * i + 2];
      ti = type[i];
      ptp = 0;
      for (int u = 0; u < npt; ++u)
        ptp |= (unsigned long long)pidx_s[ti * npt + u] << (8 * u);
      const int n = nn[i];
      // the loads are unconditional (rows clamped into the list; past its
      // end the particle itself) so no branch makes the compiler wait for
      // them one by one
      const int last = n - 1;
      for (int k0 = lane; k0 < n; k0 += kPairBatch * kPairLanes) {
        int js[kPairBatch];
#pragma unroll
        for (int g = 0; g < kPairBatch; ++g) {
          const int kk = k0 + g * kPairLanes;
          const int jr = (int)((unsigned)nbr[(size_t)max(0, min(kk, last)) * np + i] & kNbrIdxMask);
          js[g] = kk < n ? jr : i;
        }
        double xs[kPairBatch], ys[kPairBatch], zs[kPairBatch];
        int ts[kPairBatch];
#pragma unroll
        for (int g = 0; g < kPairBatch; ++g) {
          xs[g] = ro[3 * js[g]];
          ys[g] = ro[3 * js[g] + 1];
          zs[g] = ro[3 * js[g] + 2];
          ts[g] = type[js[g]];
        }
#pragma unroll
        for (int g = 0; g < kPairBatch; ++g)
          if (k0 + g * kPairLanes < n) eval(xi - xs[g], yi - ys[g], zi - zs[g], ts[g]);
      }
    }
    // the four lanes' partial forces: exchanges within the quad by DPP
    static_assert(kPairLanes == 4, "quad exchanges assume four lanes per particle");
    fx += quad_swap<kXor1>(fx);
    fy += quad_swap<kXor1>(fy);
    fz += quad_swap<kXor1>(fz);
    fx += quad_swap<kXor2>(fx);
    fy += quad_swap<kXor2>(fy);
    fz += quad_swap<kXor2>(fz);
    if (i < np && lane == 0) {
      const int io = V.c_sorted[i];
      V.force[3 * io] = fx;
      V.force[3 * io + 1] = fy;
      V.force[3 * io + 2] = fz;
    }
    if (energy && !kRegE) {
      // pair-type energies of this chunk: a shuffle tree in each
      // wavefront, added to the wavefront's own LDS row
      for (int k = 0; k < P; ++k) {
        double v = 0.0;
        if (i < np) {
#pragma unroll
          for (int u = 0; u < NA; ++u)
            if (u < npt && pidx_s[ti * npt + u] == k) v += ej[u];
        }
        for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
        if (l64 == 0) wacc[wv][k] += v;
      }
    }
  }
  if (err) atomicOr(&flags[0], err);
  if (energy) {
    if (kRegE) {
      // the pair-type energies of all chunks: a shuffle tree in each
      // wavefront
#pragma unroll
      for (int k = 0; k < NA; ++k) {
        double v = ej[k];
        for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
        if (l64 == 0 && k < P) wacc[wv][k] = v;
      }
    }
    // the virial, then per-block partials, component-major (component k
    // of block b at k * nblk_pair + b), the wavefronts summed in order
    double v = vir;
    for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
    if (l64 == 0) wacc[wv][P] = v;
    __syncthreads();
    double *out = V.part + V.off_pair + blk;
    for (int k = threadIdx.x; k <= P; k += blockDim.x) {
      double s2 = wacc[0][k];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s2 += wacc[w][k];
      out[(size_t)k * nb] = s2;
    }
  }
}

// The mixed path's tables as a pair evaluation sees them (in LDS, or the
// coefficients in global memory when they do not fit): bins, FP32 interval
// bounds, coefficient records, per-pair records, and the list semantics.
struct PairTabF {
  const unsigned short *bin;
  const float *zl;
  const float *cf;
  const PairRecF *rec;
  int nbin, semantics;
  float rc2;
};

// One FP32 pair term at separation (dx, dy, dz), pair type p: false when
// the pair is outside the cut-off or an error (overlap below r2umin, r2
// below the table: bits added to err); else the force factor fs (force =
// fs * d), the energy usum (FP64; with energy) and r2.  The interval is
// chosen on FP32 bounds and the polynomials are evaluated in d = r2 -
// (float)zlow; the FP32 coefficients were re-expanded about the rounded
// bound, so that rounding shifts no interval's argument.  The interval
// constant w0 enters in FP64 (its FP32 value plus the rounding kept in
// slot 11).  KN: the additive spline term (when the replica carries one).
template <bool KN>
__device__ __forceinline__ bool pair_eval_f(const SysDev &V, const PairTabF &T, float dx, float dy,
                                            float dz, int p, int energy, float &fs,
                                            double &usum, float &r2, int &err) {
  r2 = dx * dx + dy * dy + dz * dz;
  const bool in = T.semantics == 0 ? !(r2 > T.rc2) : (r2 < T.rc2);
  const PairRecF R = T.rec[p];
  const float r2c = fmaxf(r2, R.lo);
  const float sf = __builtin_amdgcn_rcpf(r2c);
  const float xb = fminf(fmaxf((sf - R.s_lo) * R.s_inv, 0.0f), (float)(T.nbin - 1));
  const int bk = T.bin[p * T.nbin + (int)xb];
  const int o = R.off, n = R.nint, kb = bk & (kSlowBin - 1);
  const float z0 = T.zl[o + kb], z1 = T.zl[o + min(kb + 1, n - 1)];
  int iv;
  float zv;
  if (bk & kSlowBin) {
    auto zl = [&](int k) -> float { return T.zl[o + k]; };
    iv = find_interval(zl, n, r2c, kb);
    zv = T.zl[o + iv];
  } else {
    const bool at = r2c >= z0;
    iv = at ? kb : min(kb + 1, n - 1);
    zv = at ? z0 : z1;
  }
  const float4 *q = reinterpret_cast<const float4 *>(T.cf) + 3 * (o + iv);
  const float4 c0 = q[0], c1 = q[1], c2 = q[2];
  const bool hc = r2 < R.r2at;
  const bool ov = !hc && r2 < R.r2um;
  const bool te = !hc && !ov && !(r2 >= R.lo);
  if (in) err |= (ov ? ERR_OVERLAP : 0) | (te ? ERR_TABLE : 0);
  if (!in || ov || te) return false;
  const float df = r2 - zv;
  // c0 = (w0, w1, w2, w3), c1 = (w4, w5, f0, f1), c2 = (f2, f3, f4, w0 rounding)
  fs = c1.z + df * (c1.w + df * (c2.x + df * (c2.y + df * c2.z)));
  usum = 0.0;
  if (energy) {
    const float up = df * (c0.y + df * (c0.z + df * (c0.w + df * (c1.x + df * c1.y))));
    usum = ((double)c0.x + (double)c2.w) + (double)up;
  }
  if (hc) {
    usum = 1e10;
    fs = 1e10f;
  } else if (KN && V.nknot) {
    const double r = sqrt((double)r2);
    int ki;
    double w[4], dw[4];
    if (knot_weights(r, V.knot_r0, V.knot_dinv, V.nknot, ki, w, dw)) {
      const double *kc = V.knot_c + (size_t)p * V.nknot + ki;
      usum += kc[0] * w[0] + kc[1] * w[1] + kc[2] * w[2] + kc[3] * w[3];
      fs -= (float)((kc[0] * dw[0] + kc[1] * dw[1] + kc[2] * dw[2] + kc[3] * dw[3]) *
                    V.knot_dinv / r);
    }
  }
  return true;
}

// The mixed path's pair forces: the list walk, partition and reductions of
// k_pair, in FP32 (pair_eval_f).  Positions come from pos_f (relative to
// the particle's cell reference point, 16 bytes with the type); a list
// entry's code names the pair's cell offset, kept in LDS per code, so
// dx = x_i - x_j + offset is formed in FP32 from values of at most a cell
// edge or so.  Forces are Kahan-compensated FP32 sums, energies FP64 per
// partner type, flushed into the wavefront's LDS row per chunk.  KN:
// some replica carries the additive spline term.  V.pair_lanes lanes share
// a particle's list: 4 for dense systems, fewer where lists are short (a
// lane then keeps its batch slots busy and the per-particle work is not
// repeated on idle lanes); a replica's choice is its own, so it computes
// the same alone and in a batch.  POS64: the replicas whose box has an
// axis of fewer than three list cells, where cell-relative coordinates
// would span half the box -- the pair's separation is formed in FP64 from
// the cell-ordered positions and rounded once; the other replicas' blocks
// return (a group launches both variants when it holds both kinds).
// The body takes its block index bxi (the launch's own, or its place in a
// launch that also carries other work).
template <int NT, bool CF, bool KN, bool POS64>
__device__ __forceinline__ void pair_f_body(const SysDev *__restrict__ sys, int energy, int pgrid,
                                            int skin, int bxi) {
  const SysDev &V = sys[blockIdx.z];
  if (skin && bxi >= pgrid) {
    skin_two_blocks(V, 2 * (bxi - pgrid));
    return;
  }
  const int blk = xcd_block(bxi, pgrid);
  const int nb = V.nblk_pair;
  if (blk >= nb || V.pair_pos64 != (POS64 ? 1 : 0)) return;
  const float4 *__restrict__ pos = V.pos_f;
  const int np = V.np;
  const int *__restrict__ nn = V.nn;
  const unsigned *__restrict__ nbr = reinterpret_cast<const unsigned *>((const int *)V.nbr);
  const TableDev t = V.tab;
  const int semantics = V.nlsem;
  extern __shared__ __align__(16) unsigned char pair_lds[];
  __shared__ PairRecF rec_s[MFE_MAX_PAIRS];
  __shared__ int pidx_s[MFE_MAX_TYPES * MFE_MAX_TYPES];
  __shared__ float4 off_s[125];
  __shared__ double wacc[kPairBlock / 64][MFE_MAX_PAIRS + 1];
  const int P = t.nptpt, nbin = t.nbin, nzl = t.nzl, npt = t.npt;
  const size_t bb = pair_bin_bytes(P, nbin);
  unsigned short *bin_s = reinterpret_cast<unsigned short *>(pair_lds);
  float *zl_s = reinterpret_cast<float *>(pair_lds + bb);
  float *cf_s = reinterpret_cast<float *>(pair_lds + bb + pair_zl_bytes(nzl));
  for (int k = threadIdx.x; k < P * nbin; k += blockDim.x) bin_s[k] = t.bin_k[k];
  for (int k = threadIdx.x; k < nzl; k += blockDim.x) zl_s[k] = t.zlow_f[k];
  if (CF)
    for (int k = threadIdx.x; k < nzl * kCoefStride; k += blockDim.x) cf_s[k] = t.coef_f[k];
  for (int k = threadIdx.x; k < P; k += blockDim.x) rec_s[k] = t.rec_f[k];
  for (int k = threadIdx.x; k < npt * npt; k += blockDim.x) pidx_s[k] = t.pairidx[k];
  for (int k = threadIdx.x; k < 125; k += blockDim.x)
    off_s[k] = make_float4(V.cg.off[0][k % 5], V.cg.off[1][(k / 5) % 5], V.cg.off[2][k / 25], 0.f);
  const int l64 = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = l64; k <= P; k += 64) wacc[wv][k] = 0.0;
  __syncthreads();
  const float rc2 = (float)t.rcut2;
  const float L0 = V.cg.L[0], L1 = V.cg.L[1], L2 = V.cg.L[2];
  const float H0 = V.cg.L2[0], H1 = V.cg.L2[1], H2 = V.cg.L2[2];
  // lanes per particle: 1, 2 or 4 (lsh its log2)
  const int LANES = V.pair_lanes, lsh = LANES == 4 ? 2 : LANES - 1;
  const int lane = threadIdx.x & (LANES - 1);
  double ej[NT];
  double vir = 0.0;
  KahanF fxf, fyf, fzf;
  int ti = 0;
  unsigned long long ptp = 0;
  int err = 0;
  // the separation of a partner at relative position pj with code e
  auto sep = [&](float xi, float yi, float zi, float4 pj, unsigned e, float &dx, float &dy,
                 float &dz) {
    const float4 o4 = off_s[e >> kNbrCodeShift];
    dx = (xi - pj.x) + o4.x;
    dy = (yi - pj.y) + o4.y;
    dz = (zi - pj.z) + o4.z;
    // axes of one cell: the minimum image
    if (fabsf(dx) > H0) dx -= copysignf(L0, dx);
    if (fabsf(dy) > H1) dy -= copysignf(L1, dy);
    if (fabsf(dz) > H2) dz -= copysignf(L2, dz);
  };
  const PairTabF T{bin_s, zl_s, CF ? cf_s : (const float *)t.coef_f, rec_s, nbin, semantics, rc2};
  // one list entry: a partner of type tj at separation (dx, dy, dz)
  auto eval = [&](float dx, float dy, float dz, int tj) {
    const int p = (int)((ptp >> (8 * tj)) & 0xffu);
    float fs, r2;
    double usum;
    if (!pair_eval_f<KN>(V, T, dx, dy, dz, p, energy, fs, usum, r2, err)) return;
    fxf.add(fs * dx);
    fyf.add(fs * dy);
    fzf.add(fs * dz);
    if (energy) {
#pragma unroll
      for (int k2 = 0; k2 < NT; ++k2) ej[k2] += (k2 == tj) ? 0.5 * usum : 0.0;
      vir -= 0.5 * (double)(fs * r2);
    }
  };
  const int kChunk = kPairBlock >> lsh;
  const int nch = (np + kChunk - 1) / kChunk;
  const int ch0 = (int)((long long)blk * nch / nb), ch1 = (int)((long long)(blk + 1) * nch / nb);
  const int maxrow = V.maxnn - 1;
  for (int ch = ch0; ch < ch1; ++ch) {
    const int i = ch * kChunk + (threadIdx.x >> lsh);
#pragma unroll
    for (int k = 0; k < NT; ++k) ej[k] = 0.0;
    fxf = KahanF();
    fyf = KahanF();
    fzf = KahanF();
    ti = 0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < np) {
      // the particle, its list length and its first entries issued together
      const float4 pi = POS64 ? make_float4(0.f, 0.f, 0.f, __int_as_float(V.type_s[i])) : pos[i];
      const double xi = POS64 ? V.ro_s[3 * i] : 0.0, yi = POS64 ? V.ro_s[3 * i + 1] : 0.0,
                   zi = POS64 ? V.ro_s[3 * i + 2] : 0.0;
      const int n = nn[i];
      unsigned es[kPairBatch];
#pragma unroll
      for (int g = 0; g < kPairBatch; ++g)
        es[g] = nbr[(size_t)min(lane + g * LANES, maxrow) * np + i];
      ti = __float_as_int(pi.w);
      ptp = 0;
      for (int u = 0; u < npt; ++u)
        ptp |= (unsigned long long)pidx_s[ti * npt + u] << (8 * u);
      for (int k0 = lane; k0 < n; k0 += kPairBatch * LANES) {
        if (k0 != lane) {
#pragma unroll
          for (int g = 0; g < kPairBatch; ++g)
            es[g] = nbr[(size_t)min(k0 + g * LANES, maxrow) * np + i];
        }
        if constexpr (POS64) {
          double xs[kPairBatch], ys[kPairBatch], zs[kPairBatch];
          int ts[kPairBatch];
#pragma unroll
          for (int g = 0; g < kPairBatch; ++g) {
            const int j = k0 + g * LANES < n ? (int)(es[g] & kNbrIdxMask) : i;
            xs[g] = V.ro_s[3 * j];
            ys[g] = V.ro_s[3 * j + 1];
            zs[g] = V.ro_s[3 * j + 2];
            ts[g] = V.type_s[j];
          }
#pragma unroll
          for (int g = 0; g < kPairBatch; ++g) {
            if (k0 + g * LANES < n) {
              double ddx = xi - xs[g], ddy = yi - ys[g], ddz = zi - zs[g];
              pbcr2(ddx, ddy, ddz, V.box);
              eval((float)ddx, (float)ddy, (float)ddz, ts[g]);
            }
          }
        } else {
          float4 ps[kPairBatch];
#pragma unroll
          for (int g = 0; g < kPairBatch; ++g) {
            // entries past the list end read the particle itself
            const int j = k0 + g * LANES < n ? (int)(es[g] & kNbrIdxMask) : i;
            ps[g] = pos[j];
          }
#pragma unroll
          for (int g = 0; g < kPairBatch; ++g) {
            if (k0 + g * LANES < n) {
              float dx, dy, dz;
              sep(pi.x, pi.y, pi.z, ps[g], es[g], dx, dy, dz);
              eval(dx, dy, dz, __float_as_int(ps[g].w));
            }
          }
        }
      }
      fx = fxf.value();
      fy = fyf.value();
      fz = fzf.value();
    }
    // the lanes' partial forces: exchanges within pairs / quads by DPP
    // (LANES is uniform over the block, so every lane takes part)
    if (LANES >= 2) {
      fx += quad_swap<kXor1>(fx);
      fy += quad_swap<kXor1>(fy);
      fz += quad_swap<kXor1>(fz);
    }
    if (LANES == 4) {
      fx += quad_swap<kXor2>(fx);
      fy += quad_swap<kXor2>(fy);
      fz += quad_swap<kXor2>(fz);
    }
    if (i < np && lane == 0) {
      const int io = V.c_sorted[i];
      V.force[3 * io] = fx;
      V.force[3 * io + 1] = fy;
      V.force[3 * io + 2] = fz;
    }
    if (energy) {
      // pair-type energies of this chunk: a shuffle tree in each
      // wavefront, added to the wavefront's own LDS row
      for (int k = 0; k < P; ++k) {
        double v = 0.0;
        if (i < np) {
#pragma unroll
          for (int u = 0; u < NT; ++u)
            if (u < npt && pidx_s[ti * npt + u] == k) v += ej[u];
        }
        for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
        if (l64 == 0) wacc[wv][k] += v;
      }
    }
  }
  if (err) atomicOr(&V.flags[0], err);
  if (energy) {
    double v = vir;
    for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
    if (l64 == 0) wacc[wv][P] = v;
    __syncthreads();
    double *out = V.part + V.off_pair + blk;
    for (int k = threadIdx.x; k <= P; k += blockDim.x) {
      double s2 = wacc[0][k];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s2 += wacc[w][k];
      out[(size_t)k * nb] = s2;
    }
  }
}

template <int NT, bool CF, bool KN, bool POS64>
__global__ void __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(KN || POS64 ? 4 : 5)))
k_pair_f(const SysDev *__restrict__ sys, int energy, int pgrid, int skin) {
  pair_f_body<NT, CF, KN, POS64>(sys, energy, pgrid, skin, (int)blockIdx.x);
}

// ------------------------------------------------------------------ SPME
// Brick of every particle and its slot within the brick, both indexed by
// the particle's place t in the cell order, from the cell-ordered positions
// (current whenever SPME runs: k_move or the pair launch refreshes them).
// A block takes 1024 consecutive particles of the cell order -- a compact
// piece of the box, spanning a few dozen bricks -- counts them in an LDS
// histogram and
// then adds each non-empty bin to the global count with one atomic, so
// global atomics scale with the bricks a block touches, not with its
// particles.  Meshes with more bricks than the histogram holds fall back to
// one global atomic per particle.
constexpr int kCountThreads = 1024;
constexpr int kCountBins = 8192;
// Up to this many bricks every k_spme_scatter block scans the counts itself
// (no done counter and no last-block scan in k_spme_count), k_spread clears
// each brick's count after the scatter has used it, and a binning that is
// not followed by spreading clears them with k_spme_clear.
constexpr int kScatterScan = 2048;
static_assert(kScatterScan % kBlock == 0, "the scatter's scan gives each thread whole runs");

__global__ void __launch_bounds__(kCountThreads) k_spme_count(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  const int np = V.np;
  const int nblk = (np + kCountThreads - 1) / kCountThreads;
  if ((int)blockIdx.x >= nblk) return;
  const double *__restrict__ ro_s = V.ro_s;
  const SpmeGeom g = V.sg;
  int *bcount = V.b_count, *bid = V.b_id, *bslot = V.b_slot;
  __shared__ int hist[kCountBins];
  const int t = blockIdx.x * kCountThreads + threadIdx.x;
  const bool act = t < np;
  int id = 0;
  if (act) {
    int mm[3];
    double fr[3];
    mesh_coords(ro_s + 3 * t, g, mm, fr);
    id = fdiv(mm[0], g.fB[0]) + g.nb[0] * (fdiv(mm[1], g.fB[1]) + g.nb[1] * fdiv(mm[2], g.fB[2]));
  }
  if (V.nbricks > kCountBins) {
    if (act) {
      bid[t] = id;
      bslot[t] = atomicAdd(&bcount[id], 1);
    }
  } else {
    for (int k = threadIdx.x; k < V.nbricks; k += kCountThreads) hist[k] = 0;
    __syncthreads();
    const int rank = act ? atomicAdd(&hist[id], 1) : 0;
    __syncthreads();
    // one global atomic per non-empty bin; its old value is the bin's base
    for (int k = threadIdx.x; k < V.nbricks; k += kCountThreads) {
      const int c = hist[k];
      if (c) hist[k] = atomicAdd(&bcount[k], c);
    }
    __syncthreads();
    if (act) {
      bid[t] = id;
      bslot[t] = hist[id] + rank;
    }
  }
  if (V.nbricks <= kScatterScan) return;
  // The last block to finish scans the counts into b_start and clears them
  // (no separate scan launch).  Every count atomic of this block has
  // returned -- its value was used -- before the block barrier, so the
  // done counter is incremented after them; the last block takes an
  // acquire fence before reading the counts.  Nothing waits.
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0)
    last = __hip_atomic_fetch_add(V.b_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
           nblk - 1;
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  ph_scan(bcount, V.b_start, V.nbricks);
  if (threadIdx.x == 0) __hip_atomic_store(V.b_done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// particle records in brick order: fraction (x, y, z) and charge, plus the
// packed meshmax (10 bits per axis) and the particle index; one thread per
// place t of the cell order
// Exclusive scan of the brick counts into LDS (st[k], st[n] = the total),
// by one block of kBlock threads, each taking a contiguous run of bricks.
__device__ __forceinline__ void block_scan_counts(const int *__restrict__ count, int *st,
                                                  int n) {
  constexpr int kPer = kScatterScan / kBlock;
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = threadIdx.x * kPer;
  int v[kPer];
  int x = 0;
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    v[u] = lo + u < n ? count[lo + u] : 0;
    x += v[u];
  }
  int inc = x;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int run = inc - x, tot = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    run += w < wave ? wsum[w] : 0;
    tot += wsum[w];
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    if (lo + u < n) st[lo + u] = run;
    run += v[u];
  }
  if (threadIdx.x == 0) st[n] = tot;
  __syncthreads();
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_spme_scatter(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  const double *__restrict__ ro_s = V.ro_s;
  const double *__restrict__ q = V.q;
  const SpmeGeom g = V.sg;
  const int *bid = V.b_id, *bslot = V.b_slot, *bstart = V.b_start;
  R *frac = (R *)V.rec_frac;
  int2 *meta = V.rec_meta;
  const int nbk = V.nbricks;
  // the particle's own loads first, so that they are in flight during the
  // scan of the counts
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = t < V.np;
  int i = 0, b = 0, slot = 0, mm[3] = {0, 0, 0};
  double fr[3] = {0.0, 0.0, 0.0}, qi = 0.0;
  if (act) {
    i = V.c_sorted[t];
    b = bid[t];
    slot = bslot[t];
    mesh_coords(ro_s + 3 * t, g, mm, fr);
    qi = q[i];
  }
  __shared__ int st[kScatterScan + 1];
  if (nbk <= kScatterScan) {
    // the brick starts from the counts k_spme_count left (every block the
    // same scan); block 0 stores them for the spreading and the gather
    block_scan_counts(V.b_count, st, nbk);
    if (st[nbk] != V.np) {
      if (threadIdx.x == 0) atomicOr(&V.flags[0], ERR_BIN);
      return;
    }
    if (blockIdx.x == 0)
      for (int k = threadIdx.x; k <= nbk; k += blockDim.x) V.b_start[k] = st[k];
  }
  if (!act) return;
  const int pos = (nbk <= kScatterScan ? st[b] : bstart[b]) + slot;
  frac[4 * pos] = (R)fr[0];
  frac[4 * pos + 1] = (R)fr[1];
  frac[4 * pos + 2] = (R)fr[2];
  frac[4 * pos + 3] = (R)qi;
  meta[pos] = make_int2(mm[0] | (mm[1] << 10) | (mm[2] << 20), i);
}

// the brick counts cleared after a binning that no spreading follows
__global__ void k_spme_clear(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (V.nbricks <= kScatterScan && k < V.nbricks) V.b_count[k] = 0;
}

// Spreading, each particle once and no global atomics: one block per brick
// takes the brick's own particles and adds their whole stencil -- points
// m, m-1, ..., m-order+1 along each axis from meshmax m
// (Src/energy.F90:1232-1244) -- into an LDS tile of the brick extended by
// order-1 points on the low side of each axis, then writes the tile to a
// scratch array.  The forward x FFT reads each mesh point as its own
// brick's tile value plus the low-side halos of the next bricks (at most 8
// tiles) in a fixed order (spread_point in mfe_fft.hip).  Work is handed
// out as (particle, stencil plane) items, order planes per particle; each
// item computes the particle's B-splines in registers (no staging, no
// barrier between particles) and loads the record of its next item before
// its own adds, so the load latency hides behind them.
// 1024 threads share one brick's tile; without other LDS two blocks fit a
// CU at 16-point bricks of order <= 7.
constexpr int kSpreadThreads = 1024;

template <typename R, int ORDER>
__global__ void __launch_bounds__(kSpreadThreads) k_spread(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nbricks) return;
  // the scatter has used this brick's count: cleared for the next binning
  if (V.nbricks <= kScatterScan && threadIdx.x == 0) V.b_count[blockIdx.x] = 0;
  const R *__restrict__ frac = (const R *)V.rec_frac;
  const int2 *__restrict__ meta = V.rec_meta;
  const int *__restrict__ bstart = V.b_start;
  const SpmeGeom g = V.sg;
  typedef typename SpreadAcc<R>::T Acc;
  extern __shared__ char smem[];
  Acc *tile = reinterpret_cast<Acc *>(smem);
  const int bx = blockIdx.x % g.nb[0];
  const int by = (blockIdx.x / g.nb[0]) % g.nb[1];
  const int bz = blockIdx.x / (g.nb[0] * g.nb[1]);
  // meshmax - brick origin + order-1: extended-tile coordinates
  const int o0 = bx * g.B[0] - (ORDER - 1), o1 = by * g.B[1] - (ORDER - 1),
            o2 = bz * g.B[2] - (ORDER - 1);
  const int E0 = g.B[0] + ORDER - 1, E1 = g.B[1] + ORDER - 1, E2 = g.B[2] + ORDER - 1;
  const int ntile = E0 * E1 * E2;
  const int first = bstart[blockIdx.x], items = (bstart[blockIdx.x + 1] - first) * ORDER;
  int r = threadIdx.x;
  Rec4<R> f{};
  int code = 0;
  if (r < items) {
    const int rec = first + r / ORDER;
    f = load_rec(frac + 4 * rec);
    code = meta[rec].x;
  }
  for (int k = threadIdx.x; k < ntile; k += kSpreadThreads) tile[k] = Acc(0);
  __syncthreads();
  for (; r < items; r += kSpreadThreads) {
    const int rn = r + kSpreadThreads;
    Rec4<R> fn = f;
    int cn = code;
    if (rn < items) {
      const int rec = first + rn / ORDER;
      fn = load_rec(frac + 4 * rec);
      cn = meta[rec].x;
    }
    const int iz = r % ORDER;
    R wx[ORDER], wy[ORDER], wz[ORDER];
    bspline<R, ORDER>(f.x, wx, nullptr);
    bspline<R, ORDER>(f.y, wy, nullptr);
    bspline<R, ORDER>(f.z, wz, nullptr);
    R wzi = wz[0];
#pragma unroll
    for (int k = 1; k < ORDER; ++k)
      if (iz == k) wzi = wz[k];
    const int lx = (code & 1023) - o0;
    const int ly = ((code >> 10) & 1023) - o1;
    const int lz = ((code >> 20) & 1023) - o2;
    const R qz = f.q * wzi;
    Acc *plane = tile + (lz - iz) * E1 * E0;
#pragma unroll
    for (int jy = 0; jy < ORDER; ++jy) {
      const R qyz = qz * wy[jy];
      Acc *row = plane + (ly - jy) * E0 + lx;
#pragma unroll
      for (int j = 0; j < ORDER; ++j) SpreadAcc<R>::add(&row[-j], qyz * wx[j]);
    }
    f = fn;
    code = cn;
  }
  __syncthreads();
  R *__restrict__ out = (R *)V.spread_tiles + (size_t)blockIdx.x * ntile;
  for (int k = threadIdx.x; k < ntile; k += kSpreadThreads) out[k] = SpreadAcc<R>::get(tile[k]);
}

// per-lane stencil sums: compensated FP32 on the mixed path, plain FP64
template <typename R>
struct Acc3;
template <>
struct Acc3<float> {
  KahanF a[3];
  __device__ __forceinline__ void add(int m, float v) { a[m].add(v); }
  __device__ __forceinline__ double get(int m) const { return a[m].value(); }
};
template <>
struct Acc3<double> {
  double a[3] = {0.0, 0.0, 0.0};
  __device__ __forceinline__ void add(int m, double v) { a[m] += v; }
  __device__ __forceinline__ double get(int m) const { return a[m]; }
};

// CalcProp (Src/energy.F90:1259-1285): esum over the order^3 stencil of the
// convolved mesh; force -= q * esum * dkdr * EpsiFourPi.  One block per
// brick stages the mesh the brick's particles touch -- points
// [o - (order-1), o + B) per axis -- in LDS; each particle is taken by a
// few lanes, each computing the particle's B-splines and derivatives in
// registers and summing some of its stencil planes (FP32 with compensated
// row sums on the mixed path); the lanes are combined by lane exchanges in
// FP64 in a fixed order.  No barrier between particles: each lane loads
// its next particle's record while it sums the current one.
constexpr int kGatherLanes = 512;

// the finalize phase split in its loads and its arithmetic and stores, so
// a kernel can issue the loads early; and the reductions (defined below)
struct FinState {
  double aold[3], v[3], q, mi;
};
__device__ __forceinline__ void fin_load(const SysDev &V, int i, int mode, FinState &st);
__device__ __forceinline__ void fin_apply(const SysDev &V, int i, const double f[3],
                                         const FinState &st, int mode, double dt,
                                         const double es[3]);
__device__ __forceinline__ void ph_reduce(const SysDev &V, int energy, int skin, int bd);

// MODE: GM_FORCE adds the force; GM_POT also writes efp * phi_i (the
// reciprocal potential, dE_rec/dq_i); GM_FIELD also writes the reciprocal
// field -efp grad phi_i (the force per unit charge); GM_POTONLY writes
// efp * phi_i of the mesh and leaves the forces alone (the backward pass
// of the force-matching loss interpolates its own convolved mesh).
// FIN: also the finalize phase of each particle (its owner lane, after the
// force) and, in block nbricks, the reductions, whose inputs come from
// earlier launches.
// GM_FORCE_REC writes the reciprocal force alone into force_rec (the
// finalize adds it to the pair force).  The body takes its block index bxi.
template <typename R, int ORDER, int MODE, bool FIN>
__device__ __forceinline__ void gather_body(const SysDev *__restrict__ sys, FinArgs fa, int bxi) {
  const SysDev &V = sys[blockIdx.z];
  if (FIN && bxi == V.nbricks) {
    ph_reduce(V, fa.energy, fa.skin, fa.bd);
    return;
  }
  if (bxi >= V.nbricks) return;
  const R *__restrict__ frac = (const R *)V.rec_frac;
  const int2 *__restrict__ meta = V.rec_meta;
  const int *__restrict__ bstart = V.b_start;
  const SpmeGeom g = V.sg;
  const R *__restrict__ mesh = (const R *)V.mesh_r;
  const double efp = V.efp;
  double *__restrict__ force = V.force;
  extern __shared__ char smem[];
  R *tile = reinterpret_cast<R *>(smem);
  const int bx = bxi % g.nb[0];
  const int by = (bxi / g.nb[0]) % g.nb[1];
  const int bz = bxi / (g.nb[0] * g.nb[1]);
  const int o0 = bx * g.B[0], o1 = by * g.B[1], o2 = bz * g.B[2];
  const int E0 = g.B[0] + ORDER - 1, E1 = g.B[1] + ORDER - 1, E2 = g.B[2] + ORDER - 1;
  const int nstage = E0 * E1 * E2;
  // GL lanes share a particle, lane l summing stencil planes l, l + GL,
  // ...: two lanes on the FP32 path (no idle lanes at even orders), eight
  // on the FP64 path (one plane each)
  constexpr int GL = sizeof(R) == 4 ? 2 : 8;
  constexpr int kGStep = kGatherLanes / GL;
  constexpr int kPlanes = (ORDER + GL - 1) / GL;
  const int first = bstart[bxi], n = bstart[bxi + 1] - first;
  const int l = threadIdx.x % GL;
  int p = threadIdx.x / GL;
  // the first particle's record is in flight during the staging
  Rec4<R> f{};
  int2 md = make_int2(0, 0);
  if (p < n) {
    f = load_rec(frac + 4 * (first + p));
    md = meta[first + p];
  }
  // staged in batches of kU loads per lane, so the loads of a batch are in
  // flight together (one batch for bricks of 16 and order <= 8)
  constexpr int kU = 24;
  for (int k0 = threadIdx.x; k0 < nstage; k0 += kU * blockDim.x) {
    R val[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = k0 + u * blockDim.x;
      val[u] = R(0);
      if (k < nstage) {
        // exact: the tile fits LDS (k < 40960) and its edges are short
        const int k01 = fdiv(k, g.fE[0]);
        const int ze = fdiv(k01, g.fE[1]);
        const int xe = k - k01 * E0, ye = k01 - ze * E1;
        // region coordinates lie in [o - (order-1), o + B) with o + B <= s
        // and s >= order: one wrap, below zero only
        int x = o0 - (ORDER - 1) + xe, y = o1 - (ORDER - 1) + ye, z = o2 - (ORDER - 1) + ze;
        x += x < 0 ? g.s[0] : 0;
        y += y < 0 ? g.s[1] : 0;
        z += z < 0 ? g.s[2] : 0;
        // meshes up to 1024 per axis: the offset fits 32 bits
        val[u] = mesh[(unsigned)((z * g.s[1] + y) * g.s[0] + x)];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = k0 + u * blockDim.x;
      if (k < nstage) tile[k] = val[u];
    }
  }
  __syncthreads();
  for (; p < n; p += kGStep) {
    // the force this particle's owner lane updates and the next particle's
    // record, loaded now so the loads overlap the stencil sums
    const bool owner = l == 0;
    const int idx = md.y;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    FinState st;
    if (owner && MODE != GM_POTONLY && MODE != GM_FORCE_REC) {
      f0 = force[3 * idx];
      f1 = force[3 * idx + 1];
      f2 = force[3 * idx + 2];
      if (FIN) fin_load(V, idx, fa.mode, st);
    }
    const int pn = p + kGStep;
    Rec4<R> fn = f;
    int2 mn = md;
    if (pn < n) {
      fn = load_rec(frac + 4 * (first + pn));
      mn = meta[first + pn];
    }
    double ex = 0.0, ey = 0.0, ez = 0.0, ep = 0.0;
    if (l < ORDER) {
      R v[3][ORDER], d[3][ORDER];
      bspline<R, ORDER>(f.x, v[0], d[0]);
      bspline<R, ORDER>(f.y, v[1], d[1]);
      bspline<R, ORDER>(f.z, v[2], d[2]);
      // brick-local coordinates inside the staged region
      const int lx = (md.x & 1023) - o0 + ORDER - 1;
      const int ly = ((md.x >> 10) & 1023) - o1 + ORDER - 1;
      const int lz = ((md.x >> 20) & 1023) - o2 + ORDER - 1;
      Acc3<R> acc;
      double pacc = 0.0;
#pragma unroll
      for (int kz = 0; kz < kPlanes; ++kz) {
        const int iz = l + GL * kz;
        if (iz >= ORDER) break;
        // weights of plane iz, selected among this step's GL planes
        R vz = v[2][GL * kz], dz = d[2][GL * kz];
#pragma unroll
        for (int k = 1; k < GL; ++k)
          if (GL * kz + k < ORDER && l == k) {
            vz = v[2][GL * kz + k];
            dz = d[2][GL * kz + k];
          }
        const R *plane = tile + (size_t)(lz - iz) * E1 * E0;
#pragma unroll
        for (int iy = 0; iy < ORDER; ++iy) {
          const R *row = plane + (ly - iy) * E0;
          R sx = R(0), sdx = R(0);
#pragma unroll
          for (int ix = 0; ix < ORDER; ++ix) {
            const R qv = row[lx - ix];
            sx += v[0][ix] * qv;
            sdx += d[0][ix] * qv;
          }
          if (MODE != GM_POTONLY) {
            acc.add(0, sdx * (v[1][iy] * vz));
            acc.add(1, sx * (d[1][iy] * vz));
            acc.add(2, sx * (v[1][iy] * dz));
          }
          if (MODE == GM_POT || MODE == GM_POTONLY) pacc += (double)sx * (double)(v[1][iy] * vz);
        }
      }
      ex = acc.get(0);
      ey = acc.get(1);
      ez = acc.get(2);
      ep = pacc;
    }
    // the GL lanes of a particle are all in or all out of the loop
    if constexpr (GL == 2) {
      // the lane pair within a quad, by DPP
      ex += quad_swap<kXor1>(ex);
      ey += quad_swap<kXor1>(ey);
      ez += quad_swap<kXor1>(ez);
      if (MODE == GM_POT || MODE == GM_POTONLY) ep += quad_swap<kXor1>(ep);
    } else {
#pragma unroll
      for (int w = 1; w < GL; w <<= 1) {
        ex += __shfl_xor(ex, w, GL);
        ey += __shfl_xor(ey, w, GL);
        ez += __shfl_xor(ez, w, GL);
        if (MODE == GM_POT || MODE == GM_POTONLY) ep += __shfl_xor(ep, w, GL);
      }
    }
    if (owner) {
      // q fld rounded before the pair force is added (no contraction): the
      // same sum as the finalize forms from force and force_rec
#pragma clang fp contract(off)
      const double qd = (double)f.q;
      const double fld[3] = {-ex * g.dkdr[0] * efp, -ey * g.dkdr[1] * efp,
                             -ez * g.dkdr[2] * efp};
      const double fr[3] = {qd * fld[0], qd * fld[1], qd * fld[2]};
      const double ft[3] = {f0 + fr[0], f1 + fr[1], f2 + fr[2]};
      if (MODE == GM_FORCE_REC) {
        V.force_rec[3 * idx] = fr[0];
        V.force_rec[3 * idx + 1] = fr[1];
        V.force_rec[3 * idx + 2] = fr[2];
      } else if (MODE != GM_POTONLY) {
        force[3 * idx] = ft[0];
        force[3 * idx + 1] = ft[1];
        force[3 * idx + 2] = ft[2];
      }
      // dE_rec/dq = efp * sum_stencil theta * (G * Q) (the convolved mesh)
      if (MODE == GM_POT || MODE == GM_POTONLY) V.phi_rec[idx] = efp * ep;
      if (MODE == GM_FIELD) {
        V.fm_fld[3 * idx] = fld[0];
        V.fm_fld[3 * idx + 1] = fld[1];
        V.fm_fld[3 * idx + 2] = fld[2];
      }
      if (FIN) {
        const double es[3] = {0.0, 0.0, 0.0};
        fin_apply(V, idx, ft, st, fa.mode, fa.dt, es);
      }
    }
    f = fn;
    md = mn;
  }
}

template <typename R, int ORDER, int MODE, bool FIN>
__global__ void __launch_bounds__(kGatherLanes) k_gather(const SysDev *__restrict__ sys,
                                                         FinArgs fa) {
  gather_body<R, ORDER, MODE, FIN>(sys, fa, (int)blockIdx.x);
}

// The mixed path's pair forces and the reciprocal forces in one launch:
// blocks [0, npb) run k_pair_f's blocks (pair partition, then the skin
// blocks; npb a multiple of eight, so the gather blocks keep their
// placement on the XCDs), block npb + b the gather of brick b, which
// writes force_rec.  The two share no data; the pair waves, a lane per
// particle at low density, fill what the gather leaves idle.
static_assert(kPairBlock == kGatherLanes, "the fused launch runs both bodies in one block size");

template <int NT, int ORDER>
__global__ void __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(5)))
k_pair_gather(const SysDev *__restrict__ sys, int energy, int pgrid, int skin, int npb) {
  const int b = (int)blockIdx.x;
  if (b < npb) {
    pair_f_body<NT, true, false, false>(sys, energy, pgrid, skin, b);
    return;
  }
  gather_body<float, ORDER, GM_FORCE_REC, false>(sys, FinArgs{0, 0, 0, 0, 0.0}, b - npb);
}

// an accepted trial move written into a copy of the positions
__global__ void __launch_bounds__(256) k_mc_place(int natm, const int *__restrict__ ia,
                                                  const double *__restrict__ rtm,
                                                  double *__restrict__ ro_stage) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natm) return;
  const int i = ia[a];
  for (int m = 0; m < 3; ++m) ro_stage[3 * i + m] = rtm[3 * a + m];
}

// ------------------------------------------------------- MC trial move
// DUTwoBodyEwaldRecSPM (Src/denergy.F90:1413-1495) takes the reciprocal
// energy of the trial mesh -- the current one less the moved particles' old
// stencils plus their new ones (DUMakeGQMesh, :1432-1460) -- minus the
// current energy.  The transform is linear, so the trial spectrum is the
// current one, F, plus D(k), the transform of the stencil difference: per
// moved particle q (prod_m c'_m(k_m) - prod_m c_m(k_m)), each factor a
// one-dimensional sum over the particle's stencil along axis m, c(k) =
// sum_j w_j exp(-2 pi i k (p - j) / s) (weight w_j at mesh point p - j from
// meshmax p), the forward transform's sign.  The energy change is then
// sum_k wt(k) G(k) (2 Re(conj(F) D) + |D|^2) with the convolution pass's
// weights wt and Green function G: one pass over the kept spectrum, no
// transform per trial, and no difference of two large energies.  Weights
// and charge are rounded as the spreading rounds them (the fraction and q
// to R, the B-spline in R), so D takes out the stencil the mesh holds up
// to the rounding of the spreading's products.
//
// The factor tables, fac[(a * 2 + n) * (s1 + s2 + s3) + off_m + k] for
// moved particle a (n = 0 its current position, 1 the trial one), axis m
// and wavenumber k, the charge folded into the x factor.
template <typename R, int ORDER>
__global__ void __launch_bounds__(256) k_mc_kfac(const SysDev *__restrict__ sys, int natm,
                                                 const int *__restrict__ ia,
                                                 const double *__restrict__ rtm,
                                                 double2 *__restrict__ fac) {
  const SysDev &V = sys[blockIdx.z];
  const SpmeGeom g = V.sg;
  const int s0 = g.s[0], s1 = g.s[1], s2 = g.s[2], ns = s0 + s1 + s2;
  const int total = natm * 2 * ns;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int row = t / ns, k0 = t - row * ns;
    const int a = row >> 1, trial = row & 1;
    const int m = k0 < s0 ? 0 : (k0 < s0 + s1 ? 1 : 2);
    const int k = k0 - (m == 0 ? 0 : (m == 1 ? s0 : s0 + s1));
    const int i = ia[a];
    int mm[3];
    double fr[3];
    mesh_coords(trial ? rtm + 3 * a : V.ro + 3 * i, g, mm, fr);
    R w[ORDER];
    bspline<R, ORDER>((R)fr[m], w, nullptr);
    const int sm = g.s[m];
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int j = 0; j < ORDER; ++j) {
      int pt = (mm[m] - j) % sm;
      pt += pt < 0 ? sm : 0;
      // the phase k pt / s reduced exactly
      const int ph = (int)(((long long)k * pt) % sm);
      double sn, cs;
      sincospi(-2.0 * (double)ph / (double)sm, &sn, &cs);
      re += (double)w[j] * cs;
      im += (double)w[j] * sn;
    }
    if (m == 0) {
      const double q = (double)(R)V.q[i];
      re *= q;
      im *= q;
    }
    fac[t] = make_double2(re, im);
  }
}

// One pass over the kept spectrum F (the complex mesh's layout, kx
// fastest, fewer than 2^31 points): apply == 1 adds D to F (an accepted
// move); apply == 0 writes this block's part of the energy change to
// part[2 * blockIdx.x] (and 0 to part[2 * blockIdx.x + 1]), summed by
// k_mc_pair_sum in a fixed order.  The grid is fixed, so each thread's
// points and the sums' order are too.  One replica (an engine's MC move).
template <typename R>
__global__ void __launch_bounds__(256) k_mc_kdelta(const SysDev *__restrict__ sys, int natm,
                                                   const double2 *__restrict__ fac, void *spec,
                                                   int apply, double *__restrict__ part) {
  const SysDev &V = sys[0];
  const SpmeGeom g = V.sg;
  const SpmeAxes &ax = V.ax;
  const int s0 = g.s[0], s1 = g.s[1], s2 = g.s[2], hx = V.hx, ns = s0 + s1 + s2;
  cplx<R> *__restrict__ F = (cplx<R> *)spec;
  const unsigned n = (unsigned)hx * s1 * s2;
  double acc = 0.0;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const unsigned r = e / (unsigned)hx;
    const int kx = (int)(e - r * (unsigned)hx);
    const unsigned kzu = r / (unsigned)s1;
    const int ky = (int)(r - kzu * (unsigned)s1), kz = (int)kzu;
    double dr = 0.0, di = 0.0;
    for (int a = 0; a < natm; ++a) {
#pragma unroll
      for (int trial = 0; trial < 2; ++trial) {
        const double2 *f = fac + (size_t)(2 * a + trial) * ns;
        const double2 cx = f[kx], cy = f[s0 + ky], cz = f[s0 + s1 + kz];
        const double pr = cx.x * cy.x - cx.y * cy.y, pi = cx.x * cy.y + cx.y * cy.x;
        const double tr = pr * cz.x - pi * cz.y, ti = pr * cz.y + pi * cz.x;
        dr += trial ? tr : -tr;
        di += trial ? ti : -ti;
      }
    }
    const cplx<R> v = F[e];
    if (apply) {
      F[e] = {(R)((double)v.x + dr), (R)((double)v.y + di)};
    } else {
      const int nx = kx, ny = min(ky, s1 - ky), nz = min(kz, s2 - kz);
      const double m2 = ax.kk[0][nx] + ax.kk[1][ny] + ax.kk[2][nz];
      if (m2 > 0.0) {
        const double mf = ax.fac0 * ax.g[0][nx] * ax.g[1][ny] * ax.g[2][nz] / m2;
        const double wt = (nx == 0 || 2 * nx == s0) ? 0.5 : 1.0;
        acc += wt * mf * (2.0 * ((double)v.x * dr + (double)v.y * di) + (dr * dr + di * di));
      }
    }
  }
  if (apply) return;
  __shared__ double sh[256 / 64];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// UTwoBodyANew/UTwoBodyAOld (Src/denergy.F90:289-495) for a trial move of
// natm particles: each moved particle against its list partners that stay
// (u at the trial position minus u at the current one), plus the pairs of
// moved particles among themselves, each once (both trial minus both
// current, as its lptmdutwob branch).  The table is walked as the
// reference walks ubuf, from the pair's first interval to the first with
// r2 >= zlow, in FP64 on both precision paths.  A trial pair inside the
// hard core or below r2umin is an overlap (its goto 400).  One block,
// lanes over list entries, a fixed-order sum.
__device__ __forceinline__ double mc_pair_u(const SysDev &V, int ti, int tj, double dx,
                                            double dy, double dz, int &ovl) {
  const TableDev &t = V.tab;
  const double r2 = pbcr2(dx, dy, dz, V.box);
  const bool in = V.nlsem == 0 ? !(r2 > t.rcut2) : (r2 < t.rcut2);
  if (!in) return 0.0;
  const int p = t.pairidx[ti * t.npt + tj];
  if (r2 < t.r2atat[p] || r2 < t.r2umin[p]) {
    ovl = 1;
    return 0.0;
  }
  const int o = t.off[p], n = t.nint[p];
  int k = 0;
  while (k < n - 1 && r2 < t.zlow[o + k]) ++k;
  const double *c = t.coef + (size_t)(o + k) * kCoefStride;
  const double d = r2 - t.zlow[o + k];
  return c[0] + d * (c[1] + d * (c[2] + d * (c[3] + d * (c[4] + d * c[5]))));
}

__global__ void __launch_bounds__(256) k_mc_pair(const SysDev *__restrict__ sys, int natm,
                                                const int *__restrict__ ia_in,
                                                const double *__restrict__ rtm_in,
                                                double *__restrict__ out,
                                                double *__restrict__ part) {
  const SysDev &V = sys[blockIdx.z];
  // up to kMcStage moved particles are copied into LDS once per block, so
  // the move may come from host memory (the pinned staging) and the loops
  // below read it from LDS
  __shared__ int ia_s[kMcStage];
  __shared__ double rtm_s[3 * kMcStage];
  const bool staged = natm <= kMcStage;
  if (staged) {
    for (int k = threadIdx.x; k < natm; k += blockDim.x) ia_s[k] = ia_in[k];
    for (int k = threadIdx.x; k < 3 * natm; k += blockDim.x) rtm_s[k] = rtm_in[k];
  }
  __syncthreads();
  const int *ia = staged ? ia_s : ia_in;
  const double *rtm = staged ? rtm_s : rtm_in;
  const int np = V.np;
  const int stride = gridDim.x * blockDim.x;
  // cell-list semantics (UTwoBodyANew/OldCellList, Src/denergy.F90:728-928):
  // every particle within rcut of the old or new position, as the cells
  // around each position hold them; Verlet list (UTwoBodyANew/Old): the
  // moved particle's list partners
  const bool all = V.nlsem == 1;
  double acc = 0.0;
  int ovl = 0, ovl_old = 0;
  for (int a = 0; a < natm; ++a) {
    const int i = ia[a], t = V.c_pos[i], ti = V.type[i];
    const double xo = V.ro[3 * i], yo = V.ro[3 * i + 1], zo = V.ro[3 * i + 2];
    const double xn = rtm[3 * a], yn = rtm[3 * a + 1], zn = rtm[3 * a + 2];
    const int n = all ? np : V.nn[t];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
      int jp, tj;
      double xj, yj, zj;
      if (all) {
        jp = k;
        tj = V.type[k];
        xj = V.ro[3 * k];
        yj = V.ro[3 * k + 1];
        zj = V.ro[3 * k + 2];
      } else {
        const int j = (int)((unsigned)V.nbr[(size_t)k * np + t] & kNbrIdxMask);  // cell-ordered slot
        jp = V.c_sorted[j];
        tj = V.type_s[j];
        xj = V.ro_s[3 * j];
        yj = V.ro_s[3 * j + 1];
        zj = V.ro_s[3 * j + 2];
      }
      bool moved = false;
      for (int b = 0; b < natm; ++b) moved |= ia[b] == jp;
      if (moved) continue;
      acc += mc_pair_u(V, ti, tj, xn - xj, yn - yj, zn - zj, ovl) -
             mc_pair_u(V, ti, tj, xo - xj, yo - yj, zo - zj, ovl_old);
    }
  }
  for (int q = blockIdx.x == 0 ? (int)threadIdx.x : natm * natm; q < natm * natm;
       q += blockDim.x) {
    const int a = q / natm, b = q - a * natm;
    if (b <= a) continue;
    const int i = ia[a], j = ia[b], ti = V.type[i], tj = V.type[j];
    acc += mc_pair_u(V, ti, tj, rtm[3 * a] - rtm[3 * b], rtm[3 * a + 1] - rtm[3 * b + 1],
                     rtm[3 * a + 2] - rtm[3 * b + 2], ovl) -
           mc_pair_u(V, ti, tj, V.ro[3 * i] - V.ro[3 * j], V.ro[3 * i + 1] - V.ro[3 * j + 1],
                     V.ro[3 * i + 2] - V.ro[3 * j + 2], ovl_old);
  }
  __shared__ double red[256 / 64];
  for (int w = 32; w > 0; w >>= 1) acc += __shfl_down(acc, w, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  const int any = __syncthreads_or(ovl);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    if (gridDim.x == 1) {
      out[2 * blockIdx.z] = s;
      out[2 * blockIdx.z + 1] = any ? 1.0 : 0.0;
    } else {
      part[2 * blockIdx.x] = s;
      part[2 * blockIdx.x + 1] = any ? 1.0 : 0.0;
    }
  }
}

// the block partials of a many-block k_mc_pair in a fixed order
__global__ void __launch_bounds__(64) k_mc_pair_sum(const double *__restrict__ part, int nblk,
                                                    double *__restrict__ out) {
  double s = 0.0, o = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) {
    s += part[2 * b];
    o = fmax(o, part[2 * b + 1]);
  }
  for (int w = 32; w > 0; w >>= 1) {
    s += __shfl_down(s, w, 64);
    o = fmax(o, __shfl_down(o, w, 64));
  }
  if (threadIdx.x == 0) {
    out[0] = s;
    out[1] = o;
  }
}

// --------------------------------------------------------- standard Ewald
// UEwaldRecStd (Src/energy.F90:1050-1096) over the k-vectors of its kfac
// table: one block per k-vector sums the four sign combinations of the
// structure factor; a second kernel gives each particle its force.
__global__ void __launch_bounds__(kBlock) k_ewald_sk(const SysDev *__restrict__ sys,
                                                     int energy) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nkvec) return;
  const double *__restrict__ ro = V.ro;
  const double *__restrict__ q = V.q;
  const int np = V.np;
  const double *__restrict__ kv = V.k_vec;
  double *__restrict__ ksum = V.k_sum;
  const double alpha2 = V.alpha * V.alpha;
  double *part = V.part + V.off_std;
  __shared__ double sh[kBlock];
  const int kn = blockIdx.x;
  const double kx = kv[4 * kn], ky = kv[4 * kn + 1], kz = kv[4 * kn + 2];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    const double x = kx * ro[3 * i], y = ky * ro[3 * i + 1], z = kz * ro[3 * i + 2];
    double sa, ca;
    // combinations (-x,-y,+z), (-x,+y,+z), (+x,-y,+z), (+x,+y,+z)
    sincos(-x - y + z, &sa, &ca);
    s[0] += q[i] * ca; s[1] += q[i] * sa;
    sincos(-x + y + z, &sa, &ca);
    s[2] += q[i] * ca; s[3] += q[i] * sa;
    sincos(x - y + z, &sa, &ca);
    s[4] += q[i] * ca; s[5] += q[i] * sa;
    sincos(x + y + z, &sa, &ca);
    s[6] += q[i] * ca; s[7] += q[i] * sa;
  }
  double tot[8];
  for (int c = 0; c < 8; ++c) tot[c] = block_sum(s[c], sh);
  if (threadIdx.x == 0) {
    for (int c = 0; c < 8; ++c) ksum[8 * kn + c] = tot[c];
    if (energy) {
      double ssum = 0.0;
      for (int c = 0; c < 8; ++c) ssum += tot[c] * tot[c];
      const double term = kv[4 * kn + 3] * ssum;
      const double k2 = kx * kx + ky * ky + kz * kz;
      part[kn] = term;
      part[V.nkvec + kn] = -(1.0 - k2 / (2.0 * alpha2)) * term;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_ewald_force(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  const double *__restrict__ ro = V.ro;
  const double *__restrict__ q = V.q;
  const double *__restrict__ kv = V.k_vec;
  const double *__restrict__ ksum = V.k_sum;
  const int nk = V.nkvec;
  const double efp = V.efp;
  double *__restrict__ force = V.force;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V.np) return;
  const double xi = ro[3 * i], yi = ro[3 * i + 1], zi = ro[3 * i + 2];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int kn = 0; kn < nk; ++kn) {
    const double kx = kv[4 * kn], ky = kv[4 * kn + 1], kz = kv[4 * kn + 2];
    const double kfac2 = efp * 2.0 * kv[4 * kn + 3];
    const double *S = ksum + 8 * kn;
    const double x = kx * xi, y = ky * yi, z = kz * zi;
    double f[4];
    const double args[4] = {-x - y + z, -x + y + z, x - y + z, x + y + z};
    for (int c = 0; c < 4; ++c) {
      double sa, ca;
      sincos(args[c], &sa, &ca);
      // -Re(q e) Im(S) + Im(q e) Re(S)
      f[c] = q[i] * (-ca * S[2 * c + 1] + sa * S[2 * c]);
    }
    fx += kfac2 * kx * (-f[0] - f[1] + f[2] + f[3]);
    fy += kfac2 * ky * (-f[0] + f[1] - f[2] + f[3]);
    fz += kfac2 * kz * (f[0] + f[1] + f[2] + f[3]);
  }
  force[3 * i] += fx;
  force[3 * i + 1] += fy;
  force[3 * i + 2] += fz;
}

// ------------------------------------------------------------------ bonds
// One bond angle of UAngle (Src/energy.F90:5409-5481), centred on bead c
// with neighbours jm (earlier) and jp (later): unit vectors from jm and jp
// to c, theta = acos(clamped cosine), u = k (theta - eq)^p with k already
// divided by p.  fm and fp are the forces on jm and jp; c receives
// -(fm + fp).  sin(theta) = 0 divides by zero, as in the reference.
__device__ __forceinline__ double angle_at(const double *__restrict__ ro, int c, int jm, int jp,
                                           const Box &b, double k, double eq, int p,
                                           double fm[3], double fp[3]) {
  double dxm = ro[3 * c] - ro[3 * jm], dym = ro[3 * c + 1] - ro[3 * jm + 1],
         dzm = ro[3 * c + 2] - ro[3 * jm + 2];
  const double dmi = 1.0 / sqrt(pbcr2(dxm, dym, dzm, b));
  dxm *= dmi;
  dym *= dmi;
  dzm *= dmi;
  double dxp = ro[3 * c] - ro[3 * jp], dyp = ro[3 * c + 1] - ro[3 * jp + 1],
         dzp = ro[3 * c + 2] - ro[3 * jp + 2];
  const double dpi = 1.0 / sqrt(pbcr2(dxp, dyp, dzp, b));
  dxp *= dpi;
  dyp *= dpi;
  dzp *= dpi;
  const double cosine = dxm * dxp + dym * dyp + dzm * dzp;
  const double theta = acos(fmax(-1.0, fmin(cosine, 1.0)));
  const double term = theta - eq;
  double pw = 1.0;
  for (int e = 1; e < p; ++e) pw *= term;
  double fac = k * pw;
  const double u = fac * term;
  fac = -p * fac / sin(theta);
  fm[0] = (fac * dmi) * (dxp - cosine * dxm);
  fm[1] = (fac * dmi) * (dyp - cosine * dym);
  fm[2] = (fac * dmi) * (dzp - cosine * dzm);
  fp[0] = (fac * dpi) * (dxm - cosine * dxp);
  fp[1] = (fac * dpi) * (dym - cosine * dyp);
  fp[2] = (fac * dpi) * (dzm - cosine * dzp);
  return u;
}

// UBond (Src/energy.F90:5335-5372): u = k (r - eq)^p with k already
// divided by p; each particle gathers the forces of its (at most two)
// bonds, and the bond to the next bead carries the energy.  With an angle
// force constant, also UAngle (Src/energy.F90:5409-5481): each particle
// gathers the forces of the (at most three) angles it takes part in, each
// angle evaluated by angle_at, and the centre bead carries the energy.
__global__ void __launch_bounds__(kBlock) k_bonds(const SysDev *__restrict__ sys,
                                                  int energy) {
  const SysDev &V = sys[blockIdx.z];
  if (V.nchain == 0 || (int)blockIdx.x >= V.nblk_np) return;
  const double *__restrict__ ro = V.ro;
  const int np = V.np;
  const int *__restrict__ prev = V.ch_prev;
  const int *__restrict__ next = V.ch_next;
  const Box b = V.box;
  const double k = V.bond_k, eq = V.bond_eq;
  const int p = V.bond_p;
  double *__restrict__ force = V.force;
  double *part = V.part + V.off_bond;
  __shared__ double sh[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ub = 0.0, vir = 0.0, ua = 0.0;
  if (i < np) {
    double f[3] = {0.0, 0.0, 0.0};
    const int nb2[2] = {next[i], prev[i]};
    for (int s = 0; s < 2; ++s) {
      const int j = nb2[s];
      if (j < 0) continue;
      // bond (a, c) with a the earlier bead: dx = ro(a) - ro(c)
      const int a = s == 0 ? i : j, c = s == 0 ? j : i;
      double dx = ro[3 * a] - ro[3 * c], dy = ro[3 * a + 1] - ro[3 * c + 1],
             dz = ro[3 * a + 2] - ro[3 * c + 2];
      const double r2 = pbcr2(dx, dy, dz, b);
      const double r1 = sqrt(r2);
      const double term = r1 - eq;
      double pw = 1.0;
      for (int e = 1; e < p; ++e) pw *= term;
      double fac = k * pw;
      if (s == 0) {
        ub += fac * term;
      }
      fac = -p * fac / r1;
      const double sg = s == 0 ? 1.0 : -1.0;
      f[0] += sg * fac * dx;
      f[1] += sg * fac * dy;
      f[2] += sg * fac * dz;
      if (s == 0) vir -= fac * r2;
    }
    if (V.ang_k != 0.0) {
      const double ka = V.ang_k, eqa = V.ang_eq;
      const int pa = V.ang_p;
      double fm[3], fp[3];
      const int jm = prev[i], jp = next[i];
      // i as the later bead of the angle centred on jm
      if (jm >= 0 && prev[jm] >= 0) {
        (void)angle_at(ro, jm, prev[jm], i, b, ka, eqa, pa, fm, fp);
        f[0] += fp[0];
        f[1] += fp[1];
        f[2] += fp[2];
      }
      // i as the centre
      if (jm >= 0 && jp >= 0) {
        ua += angle_at(ro, i, jm, jp, b, ka, eqa, pa, fm, fp);
        f[0] -= fm[0] + fp[0];
        f[1] -= fm[1] + fp[1];
        f[2] -= fm[2] + fp[2];
      }
      // i as the earlier bead of the angle centred on jp
      if (jp >= 0 && next[jp] >= 0) {
        (void)angle_at(ro, jp, i, next[jp], b, ka, eqa, pa, fm, fp);
        f[0] += fm[0];
        f[1] += fm[1];
        f[2] += fm[2];
      }
    }
    force[3 * i] += f[0];
    force[3 * i + 1] += f[1];
    force[3 * i + 2] += f[2];
  }
  if (energy) {
    double s1 = block_sum(ub, sh);
    double s2 = block_sum(vir, sh);
    double s3 = block_sum(ua, sh);
    if (threadIdx.x == 0) {
      part[blockIdx.x] = s1;
      part[V.nblk_np + blockIdx.x] = s2;
      part[2 * V.nblk_np + blockIdx.x] = s3;
    }
  }
}

// ------------------------------------------------------- surface + sums
__global__ void __launch_bounds__(kBlock) k_surf_sum(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  if (!V.lsurf || (int)blockIdx.x >= V.nblk_np) return;
  const double *__restrict__ ro = V.ro;
  const double *__restrict__ q = V.q;
  const int np = V.np;
  double *part = V.part + V.off_surf;
  __shared__ double sh[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s[3] = {0, 0, 0};
  if (i < np)
    for (int m = 0; m < 3; ++m) s[m] = q[i] * ro[3 * i + m];
  for (int m = 0; m < 3; ++m) {
    double t = block_sum(s[m], sh);
    if (threadIdx.x == 0) part[m * V.nblk_np + blockIdx.x] = t;
  }
}

constexpr int kRedComp = MFE_MAX_PAIRS + 12;

// Final sums in a fixed order (one block, one pass over the partials):
// writes en = [tot, twob0, rec, bond, virial, Sx, Sy, Sz, twob(1..nptpt)]
// and, in MD, the skin-test flag of NList (Src/nlist.F90:612-615).
// bd: also advance the replica's Brownian-dynamics step counter.
// Each wavefront sums whole components (partials are component-major),
// four independent accumulators per lane, then a shuffle tree; the fixed
// order keeps the result deterministic.
constexpr int kReduceThreads = 1024;

__device__ __forceinline__ void ph_reduce(const SysDev &V, int energy, int skin, int bd) {
  double *en = V.en;
  int *flags = V.flags;
  const int P = V.nptpt;
  const int nc = P + 11;
  const int nb = V.nblk_np;
  const int nrec = V.ewald == MFE_EWALD_SPME ? V.nblk_conv
                                             : (V.ewald == MFE_EWALD_STD ? V.nkvec : 0);
  const int offrec = V.ewald == MFE_EWALD_SPME ? V.off_conv : V.off_std;
  __shared__ double red[kRedComp];
  if (bd && threadIdx.x == 0) flags[3] += 1;
  // component layout: [0,P) pair energies, P pair virial, P+1..P+2 rec
  // (u, vir), P+3..P+4 bond (u, vir), P+5..P+7 S, P+8 max r2, P+9 cand,
  // P+10 angle u
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = wave; c < nc; c += (int)(blockDim.x >> 6)) {
    const double *src;
    int n;
    bool mx = false;
    if (c <= P) {
      src = V.part + V.off_pair + (size_t)c * V.nblk_pair;
      n = energy ? V.nblk_pair : 0;
    } else if (c <= P + 2) {
      src = V.part + offrec + (size_t)(c - P - 1) * nrec;
      n = energy ? nrec : 0;
    } else if (c <= P + 4) {
      src = V.part + V.off_bond + (size_t)(c - P - 3) * nb;
      n = (energy && V.nchain > 0) ? nb : 0;
    } else if (c <= P + 7) {
      src = V.part + V.off_surf + (size_t)(c - P - 5) * nb;
      n = V.lsurf ? nb : 0;
    } else if (c == P + 10) {
      src = V.part + V.off_bond + (size_t)2 * nb;
      n = (energy && V.nchain > 0 && V.ang_k != 0.0) ? nb : 0;
    } else {
      src = V.part + (c == P + 8 ? V.off_move : V.off_skin);
      n = skin ? nb : 0;
      mx = true;
    }
    // eight accumulators, eight loads in flight per lane
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int b2 = lane;
    for (; b2 + 7 * 64 < n; b2 += 8 * 64) {
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = src[b2 + u * 64];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = mx ? fmax(acc[u], a[u]) : acc[u] + a[u];
    }
    for (; b2 < n; b2 += 64) acc[0] = mx ? fmax(acc[0], src[b2]) : acc[0] + src[b2];
    double sv;
    if (mx) {
      sv = fmax(fmax(fmax(acc[0], acc[1]), fmax(acc[2], acc[3])),
                fmax(fmax(acc[4], acc[5]), fmax(acc[6], acc[7])));
    } else {
      sv = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    }
    for (int w = 32; w > 0; w >>= 1) {
      const double o = __shfl_down(sv, w, 64);
      sv = mx ? fmax(sv, o) : sv + o;
    }
    if (lane == 0) red[c] = sv;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double *v = red;
  const double S[3] = {v[P + 5], v[P + 6], v[P + 7]};
  for (int m = 0; m < 3; ++m) en[5 + m] = S[m];
  if (skin) flags[1] = (sqrt(v[P + 8]) + sqrt(v[P + 9]) > V.drnlist) ? 1 : 0;
  if (!energy) return;
  double twob0 = 0.0;
  for (int k = 0; k < P; ++k) {
    en[8 + k] = v[k];
    twob0 += v[k];
  }
  double vir = v[P];
  double urec = v[P + 1], virrec = v[P + 2];
  const double ubond = v[P + 3];
  vir += v[P + 4];
  double rec = 0.0;
  if (V.ewald != MFE_EWALD_NONE) {
    // self term; the neutralising background of a net charge scales as
    // 1/V, so its virial (dE/dlambda under r -> lambda r) is -3 times it
    urec += V.self_energy + V.net_energy;
    virrec -= 3.0 * V.net_energy;
    if (V.lsurf) {
      const double fac = 2.0 * 3.1415926535897932 / (3.0 * V.vol);
      const double term = fac * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
      urec += term;
      virrec -= term;
    }
    rec = V.efp * urec;
    vir += V.efp * virrec;
  }
  const double uangle = v[P + 10];
  en[0] = twob0 + rec + ubond + uangle;
  en[1] = twob0;
  en[2] = rec;
  en[3] = ubond;
  en[4] = vir;
  en[kEnAngle] = uangle;
}

__global__ void __launch_bounds__(kReduceThreads) k_reduce(const SysDev *__restrict__ sys,
                                                           int energy, int skin, int bd) {
  ph_reduce(sys[blockIdx.z], energy, skin, bd);
}

// --------------------------------------------------------------- adjoint
// Gradient of the pair model (Pot_1_6_12, Src/potential.F90:1349-1425,
// with exact erfc, plus the spline term) with respect to its parameters,
// summed over the same pair set as the pair kernel.  Each pair is seen from
// both ends: the visit (i, j) adds
//   dE/dq_i        q_j efp erfc(alpha r)/r                       (full)
//   dE/dsig[t_i]   1/2 dLJ/dsig_ij            (sig_ij = (sig_i+sig_j)/2)
//   dE/deps[t_i]   dLJ/deps_ij eps_j/(2 eps_ij)   (eps_ij = sqrt(eps_i eps_j))
//   dE/dc[p][k]    1/2 spline weight of coefficient k
// and 1/2 of the pair's model energy.  Per-block column partials:
// [E, dsig(npt), deps(npt), dc(nptpt*nknot)].
__global__ void __launch_bounds__(kBlock) k_adjoint(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nblk_np) return;
  extern __shared__ double sk[];
  __shared__ double sh[kBlock];
  const int nk = V.nptpt * V.nknot;
  for (int k = threadIdx.x; k < nk; k += blockDim.x) sk[k] = 0.0;
  __syncthreads();
  const PairParams &pp = V.pp;
  const TableDev t = V.tab;
  const Box b = V.box;
  const int np = V.np;
  const int tt = blockIdx.x * blockDim.x + threadIdx.x;
  double eu = 0.0, gs = 0.0, ge = 0.0;
  int ti = -1;
  if (tt < np) {
    const int i = V.c_sorted[tt];
    const double xi = V.ro_s[3 * tt], yi = V.ro_s[3 * tt + 1], zi = V.ro_s[3 * tt + 2];
    ti = V.type_s[tt];
    const double qi = pp.zat[ti];
    double phi = 0.0;
    const int n = V.nn[tt];
    for (int k = 0; k < n; ++k) {
      const int pj = (int)((unsigned)V.nbr[(size_t)k * np + tt] & kNbrIdxMask);
      double dx = xi - V.ro_s[3 * pj], dy = yi - V.ro_s[3 * pj + 1], dz = zi - V.ro_s[3 * pj + 2];
      const double r2 = pbcr2(dx, dy, dz, b);
      if (V.nlsem == 0 ? (r2 > t.rcut2) : !(r2 < t.rcut2)) continue;
      const int tj = V.type_s[pj];
      const double r = sqrt(r2);
      const double qj = pp.zat[tj];
      const double c0 = pp.lewald ? erfc(pp.alpha * r) : 1.0;
      phi += qj * pp.efp * c0 / r;
      eu += 0.5 * qi * qj * pp.efp * c0 / r;
      const double sig = 0.5 * (pp.sigat[ti] + pp.sigat[tj]);
      const double eps = sqrt(pp.epsat[ti] * pp.epsat[tj]);
      if (!pp.lljcut || r < pp.ljrcut * sig) {
        const double x = sig / r, x2 = x * x, s6 = x2 * x2 * x2, s12 = s6 * s6;
        const double sh_ = pp.lljcut ? pp.ljushift : 0.0;
        eu += 0.5 * (4.0 * eps * (s12 - s6) + sh_ * eps);
        gs += 0.5 * 4.0 * eps * (12.0 * s12 - 6.0 * s6) / sig;
        if (eps > 0.0) ge += (4.0 * (s12 - s6) + sh_) * pp.epsat[tj] / (2.0 * eps);
      }
      if (V.nknot) {
        int ki;
        double w[4], dw[4];
        if (knot_weights(r, V.knot_r0, V.knot_dinv, V.nknot, ki, w, dw)) {
          const int pidx = t.pairidx[ti * t.npt + tj];
          const double *c = V.knot_c + (size_t)pidx * V.nknot + ki;
          eu += 0.5 * (c[0] * w[0] + c[1] * w[1] + c[2] * w[2] + c[3] * w[3]);
          for (int m = 0; m < 4; ++m) atomicAdd(&sk[pidx * V.nknot + ki + m], 0.5 * w[m]);
        }
      }
    }
    V.adj_phi[i] = phi;
  }
  double *out = V.adj_part + (size_t)blockIdx.x * V.adj_ncol;
  const double se = block_sum(eu, sh);
  if (threadIdx.x == 0) out[0] = se;
  for (int u = 0; u < V.npt; ++u) {
    const double a = block_sum(ti == u ? gs : 0.0, sh);
    const double c = block_sum(ti == u ? ge : 0.0, sh);
    if (threadIdx.x == 0) {
      out[1 + u] = a;
      out[1 + V.npt + u] = c;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nk; k += blockDim.x) out[1 + 2 * V.npt + k] = sk[k];
}

// dE/dq_i = real-space potential + efp * reciprocal potential
//           - 2 efp alpha/sqrt(pi) q_i (UEwaldSelf) + surface term
__global__ void __launch_bounds__(kBlock) k_adjoint_q(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V.np) return;
  double d = V.adj_phi[i];
  if (V.ewald == MFE_EWALD_SPME) {
    d += V.phi_rec[i];
    d -= 2.0 * V.efp * (V.alpha / 1.7724538509055160273) * V.q[i];
    d += V.efp * V.net_dq;
  }
  if (V.lsurf) {
    // efp (2 pi/3V) |sum q r|^2 -> efp (4 pi/3V) (sum q r) . r_i
    const double f = V.efp * 4.0 * 3.1415926535897932 / (3.0 * V.vol);
    d += f * (V.en[5] * V.ro[3 * i] + V.en[6] * V.ro[3 * i + 1] + V.en[7] * V.ro[3 * i + 2]);
  }
  V.adj_dq[i] = d;
}

// column sums of the adjoint partials in block order
__global__ void __launch_bounds__(kBlock) k_adjoint_sum(const SysDev *__restrict__ sys,
                                                        double *out) {
  const SysDev &V = sys[blockIdx.z];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= V.adj_ncol) return;
  double s = 0.0;
  for (int b2 = 0; b2 < V.nblk_np; ++b2) s += V.adj_part[(size_t)b2 * V.adj_ncol + c];
  out[(size_t)blockIdx.z * V.adj_ncol + c] = s;
}

// ------------------------------------------------- counter-based noise
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with
// the round keys bumped by the Weyl constants.
__host__ __device__ __forceinline__ void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

__host__ __device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  uint32_t k[2] = {k0, k1};
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k);
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
}

// three standard normal deviates for particle i at step n: Box-Muller on
// the uniforms (x + 1/2) 2^-32 of the first two pairs of Philox words
__device__ __forceinline__ void gauss3(unsigned long long seed, unsigned i, unsigned n,
                                       double g[3]) {
  uint32_t c[4] = {i, n, 0u, 0u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = ((double)c[0] + 0.5) * s, u1 = ((double)c[1] + 0.5) * s;
  const double u2 = ((double)c[2] + 0.5) * s, u3 = ((double)c[3] + 0.5) * s;
  const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  double sn, cs;
  sincospi(2.0 * u1, &sn, &cs);
  g[0] = r0 * cs;
  g[1] = r0 * sn;
  g[2] = r1 * cospi(2.0 * u3);
}

// ------------------------------------------------------------ integrators
// VelVer positions (Src/md.F90:287-298): drostep = dt v + dt^2/2 a,
// ro += drostep, PBC with ANINT; the skin bookkeeping of NList
// (Src/nlist.F90:601-603) is folded in: drosum += drostep, r2 = |drosum|^2
// and the block maximum of r2 for the second-largest rule.
// For BD (Src/bd.F90:141-151): drostep = fac1 forceo + fac2 noise, with
// the noise supplied by the host (bd = 1) or drawn from the replica's
// counter-based stream (bd = 2): particle i at BD step n takes
// Philox4x32-10(counter (i, n, 0, 0), key seed) and three Box-Muller
// deviates of it.
__global__ void __launch_bounds__(kBlock) k_move(const SysDev *__restrict__ sys,
                                                 double dt, int bd) {
#pragma clang fp contract(off)
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nblk_np) return;
  // after a failed list rebuild (a grid barrier gave up) the state stays
  // where it was for the rest of the call, which then reports the error:
  // the flag is read alongside the particle's loads and holds back only
  // the stores (the block maximum it still writes feeds a skin decision
  // that no rebuild acts on after the failure)
  const bool halt = (V.flags[0] & ERR_SYNC) != 0;
  double *__restrict__ ro = V.ro;
  const double *__restrict__ rod = V.rod;
  const double *__restrict__ rodd = V.rodd;
  const double *__restrict__ forceo = V.forceo;
  const int *__restrict__ type = V.type;
  const double *__restrict__ fac1 = V.bd_fac1;
  const double *__restrict__ fac2 = V.bd_fac2;
  const int np = V.np;
  const Box b = V.box;
  double *__restrict__ drosum = V.drosum;
  double *r2disp = V.r2disp;
  double *part_max = V.part + V.off_move;
  __shared__ double sh[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double r2 = 0.0;
  if (i < np) {
    // every load before the first store (the arrays may alias as far as
    // the compiler knows, which would serialise the loads behind stores)
    const double fac = 0.5 * (dt * dt);
    double x[3], d[3], s[3];
    const int cp = V.c_pos[i];
    const int cell = V.pos_f ? V.c_of[i] : 0, ty = V.pos_f ? type[i] : 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      x[m] = ro[3 * i + m];
      s[m] = drosum[3 * i + m];
    }
    if (bd) {
      double gn[3] = {0.0, 0.0, 0.0};
      if (bd == 1) {
#pragma unroll
        for (int m = 0; m < 3; ++m) gn[m] = V.noise[3 * i + m];
      } else {
        gauss3(V.seed, (unsigned)i, (unsigned)V.flags[3], gn);
      }
      const int t = type[i];
      const double f1 = fac1[t], f2 = fac2[t];
#pragma unroll
      for (int m = 0; m < 3; ++m) d[m] = f1 * forceo[3 * i + m] + f2 * gn[m];
    } else {
      double v[3], a[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        v[m] = rod[3 * i + m];
        a[m] = rodd[3 * i + m];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) d[m] = dt * v[m] + fac * a[m];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double xm = x[m] + d[m];
      if (fabs(xm) > b.L2[m]) xm = xm - b.L[m] * round(xm * b.Li[m]);
      x[m] = xm;
      s[m] = s[m] + d[m];
    }
    r2 = s[0] * s[0];
    r2 = r2 + s[1] * s[1];
    r2 = r2 + s[2] * s[2];
    if (!halt) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        ro[3 * i + m] = x[m];
        V.ro_s[3 * cp + m] = x[m];
        drosum[3 * i + m] = s[m];
      }
      if (V.pos_f) V.pos_f[cp] = rel_pos(V, x, cell, ty);
      r2disp[i] = r2;
    }
  }
  const double m = block_max(r2, sh);
  if (threadIdx.x == 0) part_max[blockIdx.x] = m;
}

// New positions with a valid list: NList's displacement bookkeeping
// (Src/nlist.F90:601-603) on the step since the last upload, taken as the
// minimum image of new - old (one shift per axis, as PBCr2), so the skin
// test that follows decides whether the list is rebuilt.  The new
// positions are ro_new (a host upload), or, with ro_new null, the current
// ones with the natm <= kMcStage particles ia moved to rtm (an accepted MC
// move, staged in LDS, so it may come from host memory).
__global__ void __launch_bounds__(kBlock) k_set_positions(const SysDev *__restrict__ sys,
                                                          const double *__restrict__ ro_new,
                                                          int natm, const int *__restrict__ ia,
                                                          const double *__restrict__ rtm) {
#pragma clang fp contract(off)
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nblk_np) return;
  __shared__ double sh[kBlock];
  __shared__ int ia_s[kMcStage];
  __shared__ double rtm_s[3 * kMcStage];
  if (!ro_new) {
    for (int k = threadIdx.x; k < natm; k += blockDim.x) ia_s[k] = ia[k];
    for (int k = threadIdx.x; k < 3 * natm; k += blockDim.x) rtm_s[k] = rtm[k];
    __syncthreads();
  }
  const Box b = V.box;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double r2 = 0.0;
  if (i < V.np) {
    // every load before the first store
    double xn[3], xo[3], s[3];
    const int cp = V.c_pos[i];
    const int cell = V.pos_f ? V.c_of[i] : 0, ty = V.pos_f ? V.type[i] : 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      xo[m] = V.ro[3 * i + m];
      s[m] = V.drosum[3 * i + m];
    }
    if (ro_new) {
#pragma unroll
      for (int m = 0; m < 3; ++m) xn[m] = ro_new[3 * i + m];
    } else {
#pragma unroll
      for (int m = 0; m < 3; ++m) xn[m] = xo[m];
      for (int a = 0; a < natm; ++a)
        if (ia_s[a] == i) {
#pragma unroll
          for (int m = 0; m < 3; ++m) xn[m] = rtm_s[3 * a + m];
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double d = xn[m] - xo[m];
      if (fabs(d) > b.L2[m]) d = d - copysign(b.L[m], d);
      s[m] = s[m] + d;
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      V.ro[3 * i + m] = xn[m];
      V.ro_s[3 * cp + m] = xn[m];
      V.drosum[3 * i + m] = s[m];
    }
    if (V.pos_f) V.pos_f[cp] = rel_pos(V, xn, cell, ty);
    const double *ds = s;
    r2 = ds[0] * ds[0];
    r2 = r2 + ds[1] * ds[1];
    r2 = r2 + ds[2] * ds[2];
    V.r2disp[i] = r2;
  }
  const double mx = block_max(r2, sh);
  if (threadIdx.x == 0) V.part[V.off_move + blockIdx.x] = mx;
}

// NList's r2max2 (Src/nlist.F90:602-608) is the largest r2 that is not a
// new running maximum at the moment it is visited in particle order.  Each
// block owns the same particle range as in k_move; the running maximum
// entering it is the max of the earlier blocks' maxima.
__global__ void __launch_bounds__(kBlock) k_skin_blocks(const SysDev *__restrict__ sys) {
  const SysDev &V = sys[blockIdx.z];
  if ((int)blockIdx.x >= V.nblk_np) return;
  const double *__restrict__ r2disp = V.r2disp;
  const int np = V.np;
  const double *part_max = V.part + V.off_move;
  double *part_cand = V.part + V.off_skin;
  __shared__ double sh[kBlock];
  double pm = 0.0;
  for (int b2 = threadIdx.x; b2 < (int)blockIdx.x; b2 += blockDim.x) pm = fmax(pm, part_max[b2]);
  const double P = block_max(pm, sh);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double v = i < np ? r2disp[i] : 0.0;
  // running max in particle order: inclusive max-scan inside each
  // wavefront by shuffles, then the maxima of the earlier wavefronts
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const double y = __shfl_up(inc, off, 64);
    if (lane >= off) inc = fmax(inc, y);
  }
  double excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 0.0;
  __shared__ double wmax[kBlock / 64];
  if (lane == 63) wmax[wave] = inc;
  __syncthreads();
  double pw = 0.0;
  for (int k = 0; k < wave; ++k) pw = fmax(pw, wmax[k]);
  const double before = fmax(P, fmax(pw, excl));
  const double cand = (i < np && !(v > before)) ? v : 0.0;
  const double c = block_max(cand, sh);
  if (threadIdx.x == 0) part_cand[blockIdx.x] = c;
}

// UTotal's tail (Src/energy.F90:405-418) plus, for MD, GetLinAcc and the
// velocity half of VelVer (Src/md.F90:348-358):
//   mode 0: forceo = force*fac        mode 1: + VelVer kick
// one particle of k_finalize; es = the reduced surface field en[5..7]
__device__ __forceinline__ void fin_load(const SysDev &V, int i, int mode, FinState &st) {
  if (mode == 1) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      st.aold[m] = V.rodd[3 * i + m];
      st.v[m] = V.rod[3 * i + m];
    }
  }
  st.q = V.lsurf ? V.q[i] : 0.0;
  st.mi = mode == 1 ? V.massip[i] : 0.0;
}

__device__ __forceinline__ void fin_apply(const SysDev &V, int i, const double f[3],
                                         const FinState &st, int mode, double dt,
                                         const double es[3]) {
#pragma clang fp contract(off)
  double *__restrict__ forceo = V.forceo;
  const double ffac = V.ffac, surf_fac = V.surf_fac;
  const int surf = V.lsurf;
  double *__restrict__ rod = V.rod;
  double *__restrict__ rodd = V.rodd;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double fm = f[m];
    if (surf) fm = fm - surf_fac * st.q * es[m];
    const double fo = fm * ffac;
    forceo[3 * i + m] = fo;
    if (mode == 1) {
      const double anew = st.mi * fo;
      rodd[3 * i + m] = anew;
      rod[3 * i + m] = st.v[m] + 0.5 * dt * (anew + st.aold[m]);
    }
  }
}

// every load before the first store; split: the reciprocal forces are in
// force_rec (the fused pair + gather launch), added here
__device__ __forceinline__ void ph_finalize(const SysDev &V, int i, int mode, double dt,
                                            const double es[3], bool split = false) {
  double f[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) f[m] = V.force[3 * i + m];
  if (split) {
    double r[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) r[m] = V.force_rec[3 * i + m];
#pragma unroll
    for (int m = 0; m < 3; ++m) f[m] = f[m] + r[m];
  }
  FinState st;
  fin_load(V, i, mode, st);
  fin_apply(V, i, f, st, mode, dt, es);
}

__global__ void __launch_bounds__(kBlock) k_finalize(const SysDev *__restrict__ sys,
                                                     int mode, double dt) {
  const SysDev &V = sys[blockIdx.z];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V.np) return;
  double es[3] = {0.0, 0.0, 0.0};
  if (V.lsurf) {
#pragma unroll
    for (int m = 0; m < 3; ++m) es[m] = V.en[5 + m];
  }
  ph_finalize(V, i, mode, dt, es);
}

// Without the surface term the reduction and k_finalize are independent:
// one launch in which block 0 reduces the partials (dispatched first, so
// its chain of loads overlaps the rest) and blocks [1, F] finalize
// kReduceThreads particles each (F of the replica).
__global__ void __launch_bounds__(kReduceThreads) k_reduce_finalize(const SysDev *__restrict__ sys,
                                                                    int energy, int skin, int bd,
                                                                    int mode, double dt,
                                                                    int split) {
  const SysDev &V = sys[blockIdx.z];
  const int F = (V.np + kReduceThreads - 1) / kReduceThreads;
  if (blockIdx.x == 0) {
    ph_reduce(V, energy, skin, bd);
    return;
  }
  if ((int)blockIdx.x <= F) {
    const int i = ((int)blockIdx.x - 1) * kReduceThreads + threadIdx.x;
    const double es[3] = {0.0, 0.0, 0.0};
    if (i < V.np) ph_finalize(V, i, mode, dt, es, split != 0);
  }
}

// The last launch of an MD or BD step on one engine: k_reduce_finalize's
// work (block 0 reduces, the other blocks finalize the particles, grid-
// stride) and then, in the same launch, the gated list rebuild.  Every
// block takes the skin decision itself, ph_reduce's test on the same block
// maxima (Src/nlist.F90:601-608), so no block waits for another to learn
// it, and the launch ends there when the test did not fire.  The rebuild
// reads the positions and writes the cell, list and cell-ordered arrays;
// the finalize reads and writes forces, velocities and accelerations only,
// so the rebuild's first phase needs no barrier after it.  One
// 1024-thread block per CU, all resident, as the rebuild's grid barriers
// require.
__global__ void __launch_bounds__(kReduceThreads) k_finalize_rebuild(
    const SysDev *__restrict__ sys, int energy, int bd, int mode, double dt, int split,
    unsigned *bar, int *bsum, int spin) {
  const SysDev &V = sys[0];
  const int nb = V.nblk_np;
  const double *pmove = V.part + V.off_move, *pcand = V.part + V.off_skin;
  // this thread's share of the two maxima, loaded before the finalize so
  // that the loads overlap it
  double m1 = 0.0, m2 = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    m1 = fmax(m1, pmove[b]);
    m2 = fmax(m2, pcand[b]);
  }
  if (blockIdx.x == 0) {
    ph_reduce(V, energy, 1, bd);
  } else {
    const double es[3] = {0.0, 0.0, 0.0};
    const int stride = ((int)gridDim.x - 1) * kReduceThreads;
    for (int i = ((int)blockIdx.x - 1) * kReduceThreads + (int)threadIdx.x; i < V.np;
         i += stride)
      ph_finalize(V, i, mode, dt, es, split != 0);
  }
  __shared__ double mx_s[2][kReduceThreads / 64];
  for (int w = 32; w > 0; w >>= 1) {
    m1 = fmax(m1, __shfl_xor(m1, w, 64));
    m2 = fmax(m2, __shfl_xor(m2, w, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    mx_s[0][threadIdx.x >> 6] = m1;
    mx_s[1][threadIdx.x >> 6] = m2;
  }
  __syncthreads();
  double M1 = 0.0, M2 = 0.0;
  for (int w = 0; w < kReduceThreads / 64; ++w) {
    M1 = fmax(M1, mx_s[0][w]);
    M2 = fmax(M2, mx_s[1][w]);
  }
  if (!(sqrt(M1) + sqrt(M2) > V.drnlist)) return;
  // after a failed rebuild earlier in the call the list stays void
  if (__hip_atomic_load(&V.flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_SYNC)
    return;
  rebuild_phases(V, bar, bsum, spin);
}

// ------------------------------------------------- persistent batched BD
// Many small replicas advanced by BDStep in one launch: a workgroup takes
// a replica from a queue (the costliest first), runs all nsteps steps of
// it -- move with the counter-based noise, NList's skin test, FP32 pair
// forces, SPME (fixed-point spreading into the replica's own mesh, the five
// FFT passes tile by tile, Green function, gather), the finalize and the
// reductions, the gated list rebuild -- with workgroup barriers only, then
// takes the next replica.  The arithmetic is the mixed path's: the same
// pair evaluation (pair_eval_f), B-splines, FFT passes and reductions in
// FP64.  Each replica's result depends on that replica alone (integer
// spreading sums, fixed summation orders), so a batch computes exactly
// what the same kernel computes for each replica on its own.
// Eligible replicas: mixed precision, SPME, no chains, surface term or
// spline knots.

struct PersistArgs {
  const SysDev *sys;
  const int *order;                   // replica indices, costliest first
  int nrep, nsteps, energy;
  int *queue;                         // dequeue counter, zero at the launch
  int lds;                            // dynamic LDS bytes (tables, slabs, FFT tiles)
};

// inclusive prefix maximum over the block of one value per thread (wave
// shuffles, then the earlier waves' maxima from LDS); tot gets the block's
// maximum.  wm: one slot per wave.
__device__ __forceinline__ double block_prefix_max(double v, double *wm, double &excl,
                                                   double &tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const double y = __shfl_up(inc, off, 64);
    if (lane >= off) inc = fmax(inc, y);
  }
  double e = __shfl_up(inc, 1, 64);
  if (lane == 0) e = 0.0;
  if (lane == 63) wm[wave] = inc;
  __syncthreads();
  double pw = 0.0, t = 0.0;
  for (int k = 0; k < nw; ++k) {
    if (k < wave) pw = fmax(pw, wm[k]);
    t = fmax(t, wm[k]);
  }
  __syncthreads();
  excl = fmax(pw, e);
  tot = t;
  return inc;
}

// exclusive prefix sum of cnt[0..n) into start[0..n], start[n] = total;
// the counts are left zero, as every rebuild expects them (ph_scan_offset)
__device__ __forceinline__ void block_scan_counts(int *cnt, int *start, int n, int *wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += blockDim.x) {
    const int k = base + threadIdx.x;
    const int v = k < n ? cnt[k] : 0;
    int inc = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(inc, off, 64);
      if (lane >= off) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int pw = 0, t = 0;
    for (int w = 0; w < nw; ++w) {
      if (w < wave) pw += wsum[w];
      t += wsum[w];
    }
    if (k < n) {
      start[k] = carry + pw + inc - v;
      cnt[k] = 0;
    }
    carry += t;
    __syncthreads();
  }
  if (threadIdx.x == 0) start[n] = carry;
}

// BDStep for particle i at step n (k_move's arithmetic, device noise):
// position, its cell-ordered copies, NList's displacement sum; returns
// |sum|^2 (also stored in r2disp)
__device__ __forceinline__ double bd_move_particle(const SysDev &V, int i, unsigned n) {
#pragma clang fp contract(off)
  const Box b = V.box;
  double x[3], d[3], sd[3], gn[3];
  const int cp = V.c_pos[i];
  const int cell = V.c_of[i], ty = V.type[i];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    x[m] = V.ro[3 * i + m];
    sd[m] = V.drosum[3 * i + m];
  }
  gauss3(V.seed, (unsigned)i, n, gn);
  const double f1 = V.bd_fac1[ty], f2 = V.bd_fac2[ty];
#pragma unroll
  for (int m = 0; m < 3; ++m) d[m] = f1 * V.forceo[3 * i + m] + f2 * gn[m];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double xm = x[m] + d[m];
    if (fabs(xm) > b.L2[m]) xm = xm - b.L[m] * round(xm * b.Li[m]);
    x[m] = xm;
    sd[m] = sd[m] + d[m];
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    V.ro[3 * i + m] = x[m];
    V.ro_s[3 * cp + m] = x[m];
    V.drosum[3 * i + m] = sd[m];
  }
  V.pos_f[cp] = rel_pos(V, x, cell, ty);
  double r2 = sd[0] * sd[0];
  r2 = r2 + sd[1] * sd[1];
  r2 = r2 + sd[2] * sd[2];
  V.r2disp[i] = r2;
  return r2;
}

// The stages of the persistent step, each a function of its own (not
// inlined) so that register allocation works stage by stage.
#define MFE_STAGE __device__ __attribute__((noinline))

// BDStep for the replica's particles, then NList's skin test
// (Src/nlist.F90:597-621): the largest displacement and the largest one
// that was not a running maximum, in particle order.  Returns the rebuild
// decision (every thread).
MFE_STAGE int p_move_skin(const SysDev &V, unsigned stepctr, double *wm, int *flag_s) {
  const int tid = threadIdx.x, NT = blockDim.x, np = V.np;
  double run = 0.0, cand = 0.0;
  for (int base = 0; base < np; base += NT) {
    const int i = base + tid;
    const double v = i < np ? bd_move_particle(V, i, stepctr) : 0.0;
    double excl, tot;
    block_prefix_max(v, wm, excl, tot);
    const double before = fmax(run, excl);
    if (i < np && !(v > before)) cand = fmax(cand, v);
    run = fmax(run, tot);
  }
  const double c = block_max(cand, wm);
  if (tid == 0) *flag_s = (sqrt(run) + sqrt(c) > V.drnlist) ? 1 : 0;
  __syncthreads();
  return *flag_s;
}

// Where the persistent pair stage keeps the replica's cell-relative
// positions: after the tables in LDS (offset po) when they fit the launch's
// dynamic LDS, and the separations are formed in FP32; else (false) they
// are read from global memory.
__device__ __forceinline__ bool pair_pos_in_lds(const SysDev &V, int lds, size_t &po) {
  const TableDev &t = V.tab;
  size_t tb = pair_bin_bytes(t.nptpt, t.nbin) + pair_zl_bytes(t.nzl);
  if (t.lds_mode >= 2) tb += (size_t)t.nzl * kCoefStride * sizeof(float);
  po = (tb + 15) / 16 * 16;
  return V.pair_pos64 == 0 && po + (size_t)V.np * sizeof(float4) <= (size_t)lds;
}

// FP32 pair forces in cell order, one lane per particle (pair_eval_f);
// per-wavefront pair-type energies in wred[wave][0..P), virial in [P].
// NTY bounds the number of particle types (energies per partner type).
// POSL: the partners' positions are read from an LDS copy (pair_pos_in_lds),
// a few hundred cycles instead of an L2 round trip per batch of partners.
template <int NTY, bool POSL>
MFE_STAGE void p_pair(const SysDev &V, int energy, char *smem, int lds, double (*wred)[40],
                      PairRecF *rec_s, int *pidx_s, float4 *off_s) {
  const int tid = threadIdx.x, NT = blockDim.x, np = V.np;
  const int lane = tid & 63, wave = tid >> 6;
  const TableDev t = V.tab;
  const int P = t.nptpt, npt = t.npt, nbin = t.nbin, nzl = t.nzl;
  const size_t bb = pair_bin_bytes(P, nbin);
  unsigned short *bin_s = reinterpret_cast<unsigned short *>(smem);
  float *zl_s = reinterpret_cast<float *>(smem + bb);
  float *cf_s = reinterpret_cast<float *>(smem + bb + pair_zl_bytes(nzl));
  const bool cf_lds = t.lds_mode >= 2;
  for (int k = tid; k < P * nbin; k += NT) bin_s[k] = t.bin_k[k];
  for (int k = tid; k < nzl; k += NT) zl_s[k] = t.zlow_f[k];
  if (cf_lds)
    for (int k = tid; k < nzl * kCoefStride; k += NT) cf_s[k] = t.coef_f[k];
  for (int k = tid; k < P; k += NT) rec_s[k] = t.rec_f[k];
  for (int k = tid; k < npt * npt; k += NT) pidx_s[k] = t.pairidx[k];
  for (int k = tid; k < 125; k += NT)
    off_s[k] = make_float4(V.cg.off[0][k % 5], V.cg.off[1][(k / 5) % 5], V.cg.off[2][k / 25], 0.f);
  for (int k = lane; k <= P; k += 64) wred[wave][k] = 0.0;
  const float4 *__restrict__ pos = V.pos_f;
  if constexpr (POSL) {
    size_t po = 0;
    (void)pair_pos_in_lds(V, lds, po);
    float4 *pos_s = reinterpret_cast<float4 *>(smem + po);
    for (int k = tid; k < np; k += NT) pos_s[k] = V.pos_f[k];
    pos = pos_s;
  }
  __syncthreads();
  const PairTabF T{bin_s, zl_s, cf_lds ? cf_s : (const float *)t.coef_f, rec_s, nbin, V.nlsem,
                   (float)t.rcut2};
  const float L0 = V.cg.L[0], L1 = V.cg.L[1], L2 = V.cg.L[2];
  const float H0 = V.cg.L2[0], H1 = V.cg.L2[1], H2 = V.cg.L2[2];
  const unsigned *__restrict__ nbr = reinterpret_cast<const unsigned *>((const int *)V.nbr);
  const int maxrow = V.maxnn - 1;
  const bool pos64 = V.pair_pos64 != 0;
  double vir = 0.0;
  int err = 0;
  for (int base = 0; base < np; base += NT) {
    const int tt = base + tid;
    double ej[NTY];
#pragma unroll
    for (int k = 0; k < NTY; ++k) ej[k] = 0.0;
    int ti = 0;
    if (tt < np) {
      // boxes with an axis of fewer than three list cells: separations in
      // FP64 from the cell-ordered positions (as k_pair_f's POS64 variant)
      const float4 pi = pos64 ? make_float4(0.f, 0.f, 0.f, __int_as_float(V.type_s[tt]))
                              : pos[tt];
      const double xi = pos64 ? V.ro_s[3 * tt] : 0.0, yi = pos64 ? V.ro_s[3 * tt + 1] : 0.0,
                   zi = pos64 ? V.ro_s[3 * tt + 2] : 0.0;
      const int n = V.nn[tt];
      ti = __float_as_int(pi.w);
      unsigned long long ptp = 0;
      for (int u = 0; u < npt; ++u) ptp |= (unsigned long long)pidx_s[ti * npt + u] << (8 * u);
      KahanF fxf, fyf, fzf;
      // the indices of the next two batches are in flight while this one
      // is evaluated (lists of tens of partners: the list loads, not the
      // arithmetic, would otherwise set the pace)
      unsigned en[kPairBatch], en2[kPairBatch];
#pragma unroll
      for (int q = 0; q < kPairBatch; ++q) en[q] = nbr[(size_t)min(q, maxrow) * np + tt];
#pragma unroll
      for (int q = 0; q < kPairBatch; ++q)
        en2[q] = nbr[(size_t)min(kPairBatch + q, maxrow) * np + tt];
      for (int k0 = 0; k0 < n; k0 += kPairBatch) {
        unsigned es[kPairBatch];
        float4 ps[kPairBatch];
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          es[q] = en[q];
          en[q] = en2[q];
        }
        if (pos64) {
#pragma unroll
          for (int q = 0; q < kPairBatch; ++q) {
            const int j = k0 + q < n ? (int)(es[q] & kNbrIdxMask) : tt;
            double ddx = xi - V.ro_s[3 * j], ddy = yi - V.ro_s[3 * j + 1],
                   ddz = zi - V.ro_s[3 * j + 2];
            pbcr2(ddx, ddy, ddz, V.box);
            ps[q] = make_float4((float)ddx, (float)ddy, (float)ddz, __int_as_float(V.type_s[j]));
          }
        } else {
#pragma unroll
          for (int q = 0; q < kPairBatch; ++q)
            ps[q] = pos[k0 + q < n ? (int)(es[q] & kNbrIdxMask) : tt];
        }
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q)
          en2[q] = nbr[(size_t)min(k0 + 2 * kPairBatch + q, maxrow) * np + tt];
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          if (k0 + q >= n) break;
          float dx, dy, dz;
          if (pos64) {
            dx = ps[q].x;
            dy = ps[q].y;
            dz = ps[q].z;
          } else {
            const float4 o4 = off_s[es[q] >> kNbrCodeShift];
            dx = (pi.x - ps[q].x) + o4.x;
            dy = (pi.y - ps[q].y) + o4.y;
            dz = (pi.z - ps[q].z) + o4.z;
            if (fabsf(dx) > H0) dx -= copysignf(L0, dx);
            if (fabsf(dy) > H1) dy -= copysignf(L1, dy);
            if (fabsf(dz) > H2) dz -= copysignf(L2, dz);
          }
          const int tj = __float_as_int(ps[q].w);
          const int p = (int)((ptp >> (8 * tj)) & 0xffu);
          float fs, r2;
          double usum;
          if (!pair_eval_f<false>(V, T, dx, dy, dz, p, energy, fs, usum, r2, err)) continue;
          fxf.add(fs * dx);
          fyf.add(fs * dy);
          fzf.add(fs * dz);
          if (energy) {
#pragma unroll
            for (int k2 = 0; k2 < NTY; ++k2) ej[k2] += (k2 == tj) ? 0.5 * usum : 0.0;
            vir -= 0.5 * (double)(fs * r2);
          }
        }
      }
      const int io = V.c_sorted[tt];
      V.force[3 * io] = fxf.value();
      V.force[3 * io + 1] = fyf.value();
      V.force[3 * io + 2] = fzf.value();
    }
    if (energy) {
      for (int k = 0; k < P; ++k) {
        double v = 0.0;
        if (tt < np) {
#pragma unroll
          for (int u = 0; u < NTY; ++u)
            if (u < npt && pidx_s[ti * npt + u] == k) v += ej[u];
        }
        for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
        if (lane == 0) wred[wave][k] += v;
      }
    }
  }
  if (err) atomicOr(&V.flags[0], err);
  for (int w = 32; w > 0; w >>= 1) vir += __shfl_down(vir, w, 64);
  if (lane == 0) wred[wave][P] = vir;
  __syncthreads();
}

// spreading, slab by slab of z planes in LDS: each slab's points are
// accumulated in exact fixed point (LDS 64-bit adds) from every particle
// whose stencil reaches it -- the points m, m-1, ..., m-order+1 from
// meshmax m per axis (Src/energy.F90:1223-1246), each product q wz wy wx
// in FP32 as k_spread forms it -- and then stored, rounded to FP32, into
// the real mesh (the x pass reads it there)
template <int ORDER>
MFE_STAGE void p_spread(const SysDev &V, char *smem, int lds) {
  typedef SpreadAcc<float>::T Acc;
  const int tid = threadIdx.x, NT = blockDim.x, np = V.np;
  const SpmeGeom g = V.sg;
  const int s1 = g.s[0], s2 = g.s[1], s3 = g.s[2];
  const int plane = s1 * s2;
  const int nzs = max(1, min(s3, lds / (plane * (int)sizeof(Acc))));
  Acc *slab = reinterpret_cast<Acc *>(smem);
  float *__restrict__ mesh = (float *)V.mesh_r;
  for (int z0 = 0; z0 < s3; z0 += nzs) {
    const int nz = min(nzs, s3 - z0);
    for (int k = tid; k < nz * plane; k += NT) slab[k] = Acc(0);
    __syncthreads();
    for (int i = tid; i < np; i += NT) {
      int mm[3];
      double fr[3];
      mesh_coords(V.ro + 3 * i, g, mm, fr);
      // stencil planes mm2, mm2-1, ... (mod s3) inside the slab
      bool any = false;
#pragma unroll
      for (int iz = 0; iz < ORDER; ++iz) {
        const int z = mm[2] - iz < 0 ? mm[2] - iz + s3 : mm[2] - iz;
        any |= (unsigned)(z - z0) < (unsigned)nz;
      }
      if (!any) continue;
      float wx[ORDER], wy[ORDER], wz[ORDER];
      bspline<float, ORDER>((float)fr[0], wx, nullptr);
      bspline<float, ORDER>((float)fr[1], wy, nullptr);
      bspline<float, ORDER>((float)fr[2], wz, nullptr);
      const float q = (float)V.q[i];
#pragma unroll
      for (int iz = 0; iz < ORDER; ++iz) {
        const int z = mm[2] - iz < 0 ? mm[2] - iz + s3 : mm[2] - iz;
        if ((unsigned)(z - z0) >= (unsigned)nz) continue;
        const float qz = q * wz[iz];
#pragma unroll
        for (int iy = 0; iy < ORDER; ++iy) {
          const int y = mm[1] - iy < 0 ? mm[1] - iy + s2 : mm[1] - iy;
          const float qyz = qz * wy[iy];
          Acc *row = slab + ((z - z0) * s2 + y) * s1;
#pragma unroll
          for (int ix = 0; ix < ORDER; ++ix) {
            const int x = mm[0] - ix < 0 ? mm[0] - ix + s1 : mm[0] - ix;
            SpreadAcc<float>::add(&row[x], qyz * wx[ix]);
          }
        }
      }
    }
    __syncthreads();
    float *out = mesh + (size_t)z0 * plane;
    for (int k = tid; k < nz * plane; k += NT) out[k] = SpreadAcc<float>::get(slab[k]);
    __syncthreads();
  }
}

// the five FFT passes of the replica's mesh (one function per pass), each
// over as many rows, planes or lines as the lds bytes of LDS hold; the
// convolution's energy and virial terms of this thread into esum, vsum
constexpr int kCplxF = (int)sizeof(cplx<float>), kTwF = (int)sizeof(Twiddle<float>);
__device__ __forceinline__ int persist_rows(const SysDev &V, int lds) {
  const int m = V.sg.s[0] / 2;
  return max(1, (lds - (m + V.sg.s[0]) * kTwF) / (2 * (m + 1) * kCplxF));
}
__device__ __forceinline__ int persist_planes(const SysDev &V, int lds) {
  return max(1, (lds - V.sg.s[1] * kTwF) / (2 * V.hx * (V.sg.s[1] + 1) * kCplxF));
}
__device__ __forceinline__ int persist_lines(const SysDev &V, int lds) {
  return max(1, (lds - V.sg.s[2] * kTwF) / (2 * V.hx * (V.sg.s[2] + 1) * kCplxF));
}
MFE_STAGE void p_fftx_fwd(const SysDev &V, char *smem, int lds) {
  const int s1 = V.sg.s[0], s2 = V.sg.s[1], nrows = V.sg.s[1] * V.sg.s[2];
  const int cap = persist_rows(V, lds);
  const float *__restrict__ mesh = (const float *)V.mesh_r;
  for (int row0 = 0; row0 < nrows; row0 += cap) {
    fftx_fwd_rows<float>(V, row0, min(cap, nrows - row0), cap, smem, [&](int x, int y, int z) {
      return mesh[((size_t)z * s2 + y) * s1 + x];
    });
    __syncthreads();
  }
}
MFE_STAGE void p_ffty(const SysDev &V, int inv, char *smem, int lds) {
  const int s3 = V.sg.s[2], cap = persist_planes(V, lds);
  for (int z0 = 0; z0 < s3; z0 += cap) {
    ffty_planes<float>(V, z0, min(cap, s3 - z0), inv, smem);
    __syncthreads();
  }
}
MFE_STAGE void p_fftconv(const SysDev &V, int energy, char *smem, int lds, double *esum,
                         double *vsum) {
  const int s2 = V.sg.s[1], cap = persist_lines(V, lds);
  double e = 0.0, v = 0.0;
  for (int y0 = 0; y0 < s2; y0 += cap) {
    fftconv_lines<float>(V, y0, min(cap, s2 - y0), energy, smem, e, v);
    __syncthreads();
  }
  *esum = e;
  *vsum = v;
}
MFE_STAGE void p_fftx_inv(const SysDev &V, char *smem, int lds) {
  const int nrows = V.sg.s[1] * V.sg.s[2], cap = persist_rows(V, lds);
  for (int row0 = 0; row0 < nrows; row0 += cap) {
    fftx_inv_rows<float>(V, row0, min(cap, nrows - row0), cap, smem);
    __syncthreads();
  }
}

// gather (CalcProp, Src/energy.F90:1259-1285) in particle order, then
// BDStep's forceo (fin_apply, UTotal's scaling)
template <int ORDER>
MFE_STAGE void p_gather(const SysDev &V, char *smem, int lds) {
  const int tid = threadIdx.x, NT = blockDim.x, np = V.np;
  const SpmeGeom g = V.sg;
  const int s1 = g.s[0], s2 = g.s[1], s3 = g.s[2];
  // particle i's field from the convolved mesh read through rd
  auto one = [&](int i, auto rd) {
    int mm[3];
    double fr[3];
    mesh_coords(V.ro + 3 * i, g, mm, fr);
    float v[3][ORDER], dv[3][ORDER];
    bspline<float, ORDER>((float)fr[0], v[0], dv[0]);
    bspline<float, ORDER>((float)fr[1], v[1], dv[1]);
    bspline<float, ORDER>((float)fr[2], v[2], dv[2]);
    Acc3<float> ac;
#pragma unroll
    for (int iz = 0; iz < ORDER; ++iz) {
      const int z = mm[2] - iz < 0 ? mm[2] - iz + s3 : mm[2] - iz;
      const float vz = v[2][iz], dz = dv[2][iz];
#pragma unroll
      for (int iy = 0; iy < ORDER; ++iy) {
        const int y = mm[1] - iy < 0 ? mm[1] - iy + s2 : mm[1] - iy;
        const int row = (z * s2 + y) * s1;
        float sx = 0.f, sdx = 0.f;
#pragma unroll
        for (int ix = 0; ix < ORDER; ++ix) {
          const int x = mm[0] - ix < 0 ? mm[0] - ix + s1 : mm[0] - ix;
          const float qv = rd(row + x);
          sx += v[0][ix] * qv;
          sdx += dv[0][ix] * qv;
        }
        ac.add(0, sdx * (v[1][iy] * vz));
        ac.add(1, sx * (dv[1][iy] * vz));
        ac.add(2, sx * (v[1][iy] * dz));
      }
    }
    const double qd = (double)(float)V.q[i];
    const double fld[3] = {-ac.get(0) * g.dkdr[0] * V.efp, -ac.get(1) * g.dkdr[1] * V.efp,
                           -ac.get(2) * g.dkdr[2] * V.efp};
    double ft[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) ft[m] = V.force[3 * i + m] + qd * fld[m];
#pragma unroll
    for (int m = 0; m < 3; ++m) V.force[3 * i + m] = ft[m];
    FinState st;
    fin_load(V, i, 0, st);
    const double es[3] = {0.0, 0.0, 0.0};
    fin_apply(V, i, ft, st, 0, 0.0, es);
  };
  const int M = s1 * s2 * s3;
  const float *__restrict__ mesh = (const float *)V.mesh_r;
  if ((size_t)M * sizeof(float) <= (size_t)lds) {
    // the convolved mesh staged in LDS (one pass over it), every stencil
    // read from there
    float *m_s = reinterpret_cast<float *>(smem);
    for (int k = tid; k < M; k += NT) m_s[k] = mesh[k];
    __syncthreads();
    for (int i = tid; i < np; i += NT) one(i, [&](int k) { return m_s[k]; });
    __syncthreads();
  } else {
    for (int i = tid; i < np; i += NT) one(i, [&](int k) { return mesh[k]; });
  }
}

// the gated list rebuild (the phases of k_rebuild with workgroup barriers)
MFE_STAGE void p_rebuild(const SysDev &V, int *wi) {
  const int tid = threadIdx.x, NT = blockDim.x;
  const int nc = V.cg.ncells;
  for (int k = tid; k < nc; k += NT) V.c_count[k] = 0;
  __syncthreads();
  ph_cell_count(V, tid, NT);
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  block_scan_counts(V.c_count, V.c_start, nc, wi);
  __syncthreads();
  ph_cell_scatter(V, tid, NT);
  __syncthreads();
  ph_cell_sort(V, tid, NT);
  __syncthreads();
  ph_sort_pos(V, tid, NT);
  __syncthreads();
  ph_nlist_build<true>(V, tid, NT);
  if (tid == 0) V.flags[2] += 1;
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
#undef MFE_STAGE

template <int ORDER>
__device__ void persist_replica(const SysDev &V, int nsteps, int energy, char *smem, int lds,
                                double (*wred)[40], double *wm, int *wi) {
  const int tid = threadIdx.x, nw = blockDim.x >> 6;
  const int lane = tid & 63, wave = tid >> 6;
  const int P = V.tab.nptpt;
  __shared__ PairRecF rec_s[MFE_MAX_PAIRS];
  __shared__ int pidx_s[MFE_MAX_TYPES * MFE_MAX_TYPES];
  __shared__ float4 off_s[125];
  __shared__ int flag_s;
  unsigned stepctr = (unsigned)V.flags[3];
  for (int step = 0; step < nsteps; ++step) {
    const int rebuild = p_move_skin(V, stepctr, wm, &flag_s);
    const int npt = V.tab.npt;
    size_t po = 0;
    if (pair_pos_in_lds(V, lds, po)) {
      if (npt <= 2) p_pair<2, true>(V, energy, smem, lds, wred, rec_s, pidx_s, off_s);
      else if (npt <= 4) p_pair<4, true>(V, energy, smem, lds, wred, rec_s, pidx_s, off_s);
      else p_pair<MFE_MAX_TYPES, true>(V, energy, smem, lds, wred, rec_s, pidx_s, off_s);
    } else {
      if (npt <= 2) p_pair<2, false>(V, energy, smem, lds, wred, rec_s, pidx_s, off_s);
      else if (npt <= 4) p_pair<4, false>(V, energy, smem, lds, wred, rec_s, pidx_s, off_s);
      else p_pair<MFE_MAX_TYPES, false>(V, energy, smem, lds, wred, rec_s, pidx_s, off_s);
    }
    p_spread<ORDER>(V, smem, lds);
    double esum = 0.0, vsum = 0.0;
    p_fftx_fwd(V, smem, lds);
    p_ffty(V, 0, smem, lds);
    p_fftconv(V, energy, smem, lds, &esum, &vsum);
    p_ffty(V, 1, smem, lds);
    p_fftx_inv(V, smem, lds);
    p_gather<ORDER>(V, smem, lds);
    // ---- reductions (ph_reduce's energy terms)
    for (int w = 32; w > 0; w >>= 1) {
      esum += __shfl_down(esum, w, 64);
      vsum += __shfl_down(vsum, w, 64);
    }
    if (lane == 0) {
      wred[wave][P + 1] = esum;
      wred[wave][P + 2] = vsum;
    }
    __syncthreads();
    if (tid == 0) {
      int *flags = V.flags;
      double *en = V.en;
      flags[3] = (int)(stepctr + 1u);
      for (int m = 0; m < 3; ++m) en[5 + m] = 0.0;
      if (energy) {
        // each component's wavefront partials in wavefront order
        auto total = [&](int k) {
          double sacc = wred[0][k];
          for (int w = 1; w < nw; ++w) sacc += wred[w][k];
          return sacc;
        };
        double twob0 = 0.0;
        for (int k = 0; k < P; ++k) {
          const double u = total(k);
          en[8 + k] = u;
          twob0 += u;
        }
        const double urec = total(P + 1) + V.self_energy + V.net_energy;
        const double virrec = total(P + 2) - 3.0 * V.net_energy;
        const double rec = V.efp * urec;
        en[0] = twob0 + rec;
        en[1] = twob0;
        en[2] = rec;
        en[3] = 0.0;
        en[4] = total(P) + V.efp * virrec;
        en[kEnAngle] = 0.0;
      }
      flags[1] = 0;
    }
    ++stepctr;
    __syncthreads();
    // ---- the gated list rebuild (SetVListMD after NList's test)
    if (rebuild) p_rebuild(V, wi);
  }
}

template <int ORDER>
__global__ void __launch_bounds__(kPersistThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_bd_persistent(PersistArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ int r_s;
  __shared__ double wred[kPersistThreads / 64][40];
  __shared__ double wm[kPersistThreads / 64];
  __shared__ int wi[kPersistThreads / 64];
  for (;;) {
    if (threadIdx.x == 0) r_s = atomicAdd(a.queue, 1);
    __syncthreads();
    const int rr = r_s;
    __syncthreads();
    if (rr >= a.nrep) return;
    const int r = a.order[rr];
    persist_replica<ORDER>(a.sys[r], a.nsteps, a.energy, smem, a.lds, wred, wm, wi);
  }
}

inline int nblocks(int n) { return (n + kBlock - 1) / kBlock; }

inline dim3 grid(int x, const Group &g, int y = 1) { return dim3(x > 0 ? x : 1, y, g.R); }

}  // namespace

template <typename R, int O>
const void *k_spread_ptr() {
  return (const void *)&k_spread<R, O>;
}
template <typename R, int O>
const void *k_gather_ptr(int mode, bool fin) {
  if (fin) return (const void *)&k_gather<R, O, GM_FORCE, true>;
  switch (mode) {
    case GM_POT: return (const void *)&k_gather<R, O, GM_POT, false>;
    case GM_FIELD: return (const void *)&k_gather<R, O, GM_FIELD, false>;
    case GM_POTONLY: return (const void *)&k_gather<R, O, GM_POTONLY, false>;
    default: return (const void *)&k_gather<R, O, GM_FORCE, false>;
  }
}

void philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  for (int k = 0; k < 4; ++k) out[k] = ctr[k];
  philox10(out, key[0], key[1]);
}

void launch_fill_int(int *p, int n, int v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, v);
}
void launch_fill_double(double *p, int n, double v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_double, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, v);
}

void launch_nlist_rebuild(const Group &g, bool gated) {
  if (gated && g.R == 1 && g.rebuild_bar && g.flags && g.ncu > 0) {
    hipLaunchKernelGGL(k_rebuild, dim3(g.ncu), dim3(1024), 0, g.stream, g.d, g.flags,
                       g.rebuild_bar, g.rebuild_sum, g.sync_spin >= 0 ? g.sync_spin : kGridSpin);
    return;
  }
  // gated kernels run on small grids (grid-stride loops): when the skin
  // test does not fire they cost one dispatch of a few hundred wavefronts
  const int gb = gated ? std::min(g.np_blocks, 256) : g.np_blocks;
  const int gc = gated ? std::min(g.cell_blocks, 256) : g.cell_blocks;
  const int gt = gated ? 1 : 0;
  hipLaunchKernelGGL(k_cell_count, grid(gb, g), dim3(kBlock), 0, g.stream, g.d, gt);
  hipLaunchKernelGGL(k_scan, grid(1, g), dim3(1024), 0, g.stream, g.d, 0, gt);
  hipLaunchKernelGGL(k_cell_scatter, grid(gb, g), dim3(kBlock), 0, g.stream, g.d, gt);
  hipLaunchKernelGGL(k_cell_sort, grid(gc, g), dim3(kBlock), 0, g.stream, g.d, gt);
  hipLaunchKernelGGL(k_nlist_build, grid(gb, g), dim3(kBlock), 0, g.stream, g.d, gt);
}

void launch_sort_pos(const Group &g) {
  hipLaunchKernelGGL(k_sort_pos, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d);
}

void launch_mc_place(const Group &g, int natm, const int *ia, const double *rtm,
                     double *ro_stage) {
  hipLaunchKernelGGL(k_mc_place, dim3((natm + 255) / 256), dim3(256), 0, g.stream, natm, ia, rtm,
                     ro_stage);
}

void launch_mc_kfac(const Group &g, int natm, const int *ia, const double *rtm, void *fac) {
  const int total = natm * 2 * (g.mc_s[0] + g.mc_s[1] + g.mc_s[2]);
  const dim3 gr(std::max(1, std::min((total + 255) / 256, 1024)), 1, g.R);
#define MFE_MC(R, O)                                                                         \
  case O:                                                                                    \
    hipLaunchKernelGGL((k_mc_kfac<R, O>), gr, dim3(256), 0, g.stream, g.d, natm, ia, rtm,     \
                       (double2 *)fac);                                                      \
    break;
  if (g.prec == MFE_PREC_MIXED) {
    switch (g.order) { MFE_MC(float, 3) MFE_MC(float, 4) MFE_MC(float, 5) MFE_MC(float, 6)
                       MFE_MC(float, 7) MFE_MC(float, 8) }
  } else {
    switch (g.order) { MFE_MC(double, 3) MFE_MC(double, 4) MFE_MC(double, 5) MFE_MC(double, 6)
                       MFE_MC(double, 7) MFE_MC(double, 8) }
  }
#undef MFE_MC
}

void launch_mc_kdelta(const Group &g, int natm, const void *fac, void *spec, bool apply,
                      double *part, double *out) {
  // the kept spectrum is FP64 on both precision paths
  const dim3 gr(kMcPairBlocks);
  hipLaunchKernelGGL(k_mc_kdelta<double>, gr, dim3(256), 0, g.stream, g.d, natm,
                     (const double2 *)fac, spec, apply ? 1 : 0, part);
  if (!apply)
    hipLaunchKernelGGL(k_mc_pair_sum, dim3(1), dim3(64), 0, g.stream, part, kMcPairBlocks, out);
}

void launch_mc_pair(const Group &g, int natm, const int *ia, const double *rtm, double *out,
                    double *part, bool all) {
  if (!all) {
    hipLaunchKernelGGL(k_mc_pair, grid(1, g), dim3(256), 0, g.stream, g.d, natm, ia, rtm, out,
                       part);
    return;
  }
  hipLaunchKernelGGL(k_mc_pair, dim3(kMcPairBlocks), dim3(256), 0, g.stream, g.d, natm, ia, rtm,
                     out, part);
  hipLaunchKernelGGL(k_mc_pair_sum, dim3(1), dim3(64), 0, g.stream, part, kMcPairBlocks, out);
}

void launch_pair(const Group &g, bool energy, bool sort_pos, bool skin) {
  if (sort_pos)
    hipLaunchKernelGGL(k_sort_pos, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d);
  const int pgrid = std::max(g.pair_blocks, 1);
  const dim3 gr = grid(pgrid + (skin ? (g.np_blocks + 1) / 2 : 0), g);
  const int e1 = energy ? 1 : 0, s1 = skin ? 1 : 0;
#define MFE_PAIR(NT, CF)                                                                  \
  hipLaunchKernelGGL((k_pair<NT, CF>), gr, dim3(kPairBlock), g.lds_pair, g.stream, g.d, e1, \
                     pgrid, s1)
#define MFE_PAIR_CF(NT) \
  do { if (g.pair_cf_lds) MFE_PAIR(NT, true); else MFE_PAIR(NT, false); } while (0)
  // the skin blocks ride on the first of the (one or two) launches; the
  // second has no blocks beyond pgrid
  const dim3 gr0 = grid(pgrid, g);
#define MFE_PAIRF(NT, CF, KN, P64, GR, SK)                                                 \
  hipLaunchKernelGGL((k_pair_f<NT, CF, KN, P64>), GR, dim3(kPairBlock), g.lds_pair, g.stream, \
                     g.d, e1, pgrid, SK)
#define MFE_PAIRF_LN(NT, CF, KN)                                         \
  do {                                                                   \
    if (g.pair_any32) MFE_PAIRF(NT, CF, KN, false, gr, s1);              \
    if (g.pair_any64) {                                                  \
      if (g.pair_any32) MFE_PAIRF(NT, CF, KN, true, gr0, 0);             \
      else MFE_PAIRF(NT, CF, KN, true, gr, s1);                          \
    }                                                                    \
  } while (0)
#define MFE_PAIRF_KN(NT, CF) \
  do { if (g.knots) MFE_PAIRF_LN(NT, CF, true); else MFE_PAIRF_LN(NT, CF, false); } while (0)
#define MFE_PAIRF_CF(NT) \
  do { if (g.pair_cf_lds) MFE_PAIRF_KN(NT, true); else MFE_PAIRF_KN(NT, false); } while (0)
  const int nt = g.pair_npt;
  if (g.prec == MFE_PREC_MIXED) {
    if (nt <= 3) MFE_PAIRF_CF(3);
    else if (nt <= 4) MFE_PAIRF_CF(4);
    else MFE_PAIRF_CF(8);
  } else {
    if (nt <= 3) MFE_PAIR_CF(3);
    else if (nt <= 4) MFE_PAIR_CF(4);
    else MFE_PAIR_CF(8);
  }
#undef MFE_PAIRF_CF
#undef MFE_PAIRF_KN
#undef MFE_PAIRF_LN
#undef MFE_PAIRF
#undef MFE_PAIR_CF
#undef MFE_PAIR
}

template <typename R>
static void spme_stages(const Group &g, bool energy, int first, int last, int gmode,
                        const FinArgs *fin) {
  for (int st = first; st <= last; ++st) {
    switch (st) {
      case SP_BIN:
        // counts, the last count block's scan, records
        hipLaunchKernelGGL(k_spme_count,
                           grid((g.np_blocks * kBlock + kCountThreads - 1) / kCountThreads, g),
                           dim3(kCountThreads), 0, g.stream, g.d);
        hipLaunchKernelGGL(k_spme_scatter<R>, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream,
                           g.d);
        if (last == SP_BIN)
          hipLaunchKernelGGL(k_spme_clear, grid((g.bricks + kBlock - 1) / kBlock, g), dim3(kBlock),
                             0, g.stream, g.d);
        break;
      case SP_SPREAD:
        switch (g.order) {
#define MFE_SPREAD(O)                                                                   \
  case O:                                                                               \
    hipLaunchKernelGGL((k_spread<R, O>), grid(g.bricks, g), dim3(kSpreadThreads),          \
                       g.lds_spread,                                                    \
                       g.stream, g.d);                                                  \
    break;
          MFE_SPREAD(3) MFE_SPREAD(4) MFE_SPREAD(5) MFE_SPREAD(6) MFE_SPREAD(7) MFE_SPREAD(8)
#undef MFE_SPREAD
        }
        break;
      case SP_FFTX: fft_x_forward(g); break;
      case SP_FFTY: fft_y(g, 0); break;
      case SP_CONV: fft_conv(g, energy); break;
      case SP_IFFTY: fft_y(g, 1); break;
      case SP_IFFTX: fft_x_inverse(g); break;
      case SP_GATHER: {
        const FinArgs fa = fin ? *fin : FinArgs{0, 0, 0, 0, 0.0};
        switch (g.order) {
#define MFE_GATHER(O)                                                                   \
  case O:                                                                               \
    if (fin)                                                                            \
      hipLaunchKernelGGL((k_gather<R, O, GM_FORCE, true>), grid(g.bricks + 1, g),       \
                         dim3(kGatherLanes), g.lds_gather, g.stream, g.d, fa);          \
    else if (gmode == GM_POT)                                                           \
      hipLaunchKernelGGL((k_gather<R, O, GM_POT, false>), grid(g.bricks, g),            \
                         dim3(kGatherLanes), g.lds_gather, g.stream, g.d, fa);          \
    else if (gmode == GM_FIELD)                                                         \
      hipLaunchKernelGGL((k_gather<R, O, GM_FIELD, false>), grid(g.bricks, g),          \
                         dim3(kGatherLanes), g.lds_gather, g.stream, g.d, fa);          \
    else if (gmode == GM_POTONLY)                                                       \
      hipLaunchKernelGGL((k_gather<R, O, GM_POTONLY, false>), grid(g.bricks, g),        \
                         dim3(kGatherLanes), g.lds_gather, g.stream, g.d, fa);          \
    else                                                                                \
      hipLaunchKernelGGL((k_gather<R, O, GM_FORCE, false>), grid(g.bricks, g),          \
                         dim3(kGatherLanes), g.lds_gather, g.stream, g.d, fa);          \
    break;
          MFE_GATHER(3) MFE_GATHER(4) MFE_GATHER(5) MFE_GATHER(6) MFE_GATHER(7) MFE_GATHER(8)
#undef MFE_GATHER
        }
        break;
      }
    }
  }
}

void launch_spme(const Group &g, bool energy, int first, int last, int gmode,
                 const FinArgs *fin) {
  if (g.prec == MFE_PREC_MIXED) spme_stages<float>(g, energy, first, last, gmode, fin);
  else spme_stages<double>(g, energy, first, last, gmode, fin);
}

void launch_std_ewald(const Group &g) {
  hipLaunchKernelGGL(k_ewald_sk, grid(g.kvec, g), dim3(kBlock), 0, g.stream, g.d, 1);
  hipLaunchKernelGGL(k_ewald_force, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d);
}

void launch_bonds(const Group &g) {
  hipLaunchKernelGGL(k_bonds, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d, 1);
}

void launch_surf(const Group &g) {
  hipLaunchKernelGGL(k_surf_sum, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d);
}

void launch_reduce(const Group &g, unsigned flags, bool skin, bool bd) {
  const int energy = (flags & (MFE_ENERGY | MFE_VIRIAL)) ? 1 : 0;
  hipLaunchKernelGGL(k_reduce, grid(1, g), dim3(kReduceThreads), 0, g.stream, g.d, energy,
                     skin ? 1 : 0, bd ? 1 : 0);
}

void launch_reduce_finalize(const Group &g, unsigned flags, bool skin, bool bd, int mode,
                            double dt, bool split) {
  const int energy = (flags & (MFE_ENERGY | MFE_VIRIAL)) ? 1 : 0;
  const int F = (g.np_blocks * kBlock + kReduceThreads - 1) / kReduceThreads;
  hipLaunchKernelGGL(k_reduce_finalize, grid(F + 1, g), dim3(kReduceThreads), 0, g.stream, g.d,
                     energy, skin ? 1 : 0, bd ? 1 : 0, mode, dt, split ? 1 : 0);
}

bool launch_finalize_rebuild(const Group &g, unsigned flags, bool bd, int mode, double dt,
                             bool split) {
  if (!(g.R == 1 && g.rebuild_bar && g.flags && g.ncu > 1)) return false;
  const int energy = (flags & (MFE_ENERGY | MFE_VIRIAL)) ? 1 : 0;
  hipLaunchKernelGGL(k_finalize_rebuild, dim3(g.ncu), dim3(kReduceThreads), 0, g.stream, g.d,
                     energy, bd ? 1 : 0, mode, dt, split ? 1 : 0, g.rebuild_bar, g.rebuild_sum,
                     g.sync_spin >= 0 ? g.sync_spin : kGridSpin);
  return true;
}

// the fused pair + gather launch (Group::fuse_pg)
void launch_pair_gather(const Group &g, bool energy, bool skin) {
  const int pgrid = std::max(g.pair_blocks, 1);
  const int npb = (pgrid + (skin ? (g.np_blocks + 1) / 2 : 0) + 7) / 8 * 8;
  const dim3 gr = grid(npb + g.bricks, g);
  const size_t lds = std::max(g.lds_pair, g.lds_gather);
  const int e1 = energy ? 1 : 0, s1 = skin ? 1 : 0;
  const int nt = g.pair_npt <= 3 ? 3 : (g.pair_npt <= 4 ? 4 : 8);
#define MFE_PG(NT, O)                                                                      \
  hipLaunchKernelGGL((k_pair_gather<NT, O>), gr, dim3(kPairBlock), lds, g.stream, g.d, e1, \
                     pgrid, s1, npb)
#define MFE_PG_NT(O)                \
  case O:                           \
    if (nt == 3) MFE_PG(3, O);      \
    else if (nt == 4) MFE_PG(4, O); \
    else MFE_PG(8, O);              \
    break;
  switch (g.order) {
    MFE_PG_NT(3) MFE_PG_NT(4) MFE_PG_NT(5) MFE_PG_NT(6) MFE_PG_NT(7) MFE_PG_NT(8)
  }
#undef MFE_PG_NT
#undef MFE_PG
}

void launch_finalize(const Group &g, int mode, double dt) {
  hipLaunchKernelGGL(k_finalize, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d, mode,
                     dt);
}

void launch_move(const Group &g, double dt, int bd) {
  hipLaunchKernelGGL(k_move, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d, dt, bd);
}

// flags[0..3] and en[0..nen) of every replica into one array each
__global__ void k_collect(const SysDev *__restrict__ sys, int *fl, double *en, int nen) {
  const SysDev &V = sys[blockIdx.z];
  for (int k = threadIdx.x; k < 4; k += blockDim.x) fl[4 * blockIdx.z + k] = V.flags[k];
  for (int k = threadIdx.x; k < nen; k += blockDim.x) en[(size_t)nen * blockIdx.z + k] = V.en[k];
}

// The persistent batched BD launch: blocks workgroups of kPersistThreads
// taking replicas from the queue (zeroed by the fill launch before it).
void launch_bd_persistent(const Group &g, const int *order, int nrep, int nsteps, int energy,
                          int *queue, int blocks, size_t lds) {
  PersistArgs a{g.d, order, nrep, nsteps, energy, queue, (int)lds};
  launch_fill_int(queue, 1, 0, g.stream);
  switch (g.order) {
#define MFE_PERSIST(O)                                                                      \
  case O:                                                                                  \
    hipLaunchKernelGGL(k_bd_persistent<O>, dim3(blocks), dim3(kPersistThreads), lds, g.stream, \
                       a);                                                                 \
    break;
    MFE_PERSIST(3) MFE_PERSIST(4) MFE_PERSIST(5) MFE_PERSIST(6) MFE_PERSIST(7) MFE_PERSIST(8)
#undef MFE_PERSIST
  }
}

// dynamic LDS above the default 64 KB for the persistent launch (the
// attribute only ever grows)
bool persist_allow_lds(int order, size_t bytes) {
  static size_t allowed[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (order < 3 || order > 8) return false;
  if (bytes <= 64 * 1024 || bytes <= allowed[order]) return true;
  const void *f = nullptr;
  switch (order) {
    case 3: f = (const void *)&k_bd_persistent<3>; break;
    case 4: f = (const void *)&k_bd_persistent<4>; break;
    case 5: f = (const void *)&k_bd_persistent<5>; break;
    case 6: f = (const void *)&k_bd_persistent<6>; break;
    case 7: f = (const void *)&k_bd_persistent<7>; break;
    default: f = (const void *)&k_bd_persistent<8>; break;
  }
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
      hipSuccess)
    return false;
  allowed[order] = bytes;
  return true;
}

void launch_collect(const Group &g, int *fl, double *en, int nen) {
  hipLaunchKernelGGL(k_collect, grid(1, g), dim3(64), 0, g.stream, g.d, fl, en, nen);
}

void launch_adjoint(const Group &g, int ncol, double *out) {
  // LDS of the knot accumulators: the largest replica's nptpt*nknot doubles
  const size_t lds = (size_t)std::max(ncol - 1, 1) * sizeof(double);
  hipLaunchKernelGGL(k_adjoint, grid(g.np_blocks, g), dim3(kBlock), lds, g.stream, g.d);
  hipLaunchKernelGGL(k_adjoint_q, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d);
  hipLaunchKernelGGL(k_adjoint_sum, grid(nblocks(ncol), g), dim3(kBlock), 0, g.stream, g.d,
                     out);
}

void launch_set_positions(const Group &g, const double *ro_new) {
  hipLaunchKernelGGL(k_set_positions, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d,
                     ro_new, 0, (const int *)nullptr, (const double *)nullptr);
}

void launch_set_positions_moved(const Group &g, int natm, const int *ia, const double *rtm) {
  hipLaunchKernelGGL(k_set_positions, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d,
                     (const double *)nullptr, natm, ia, rtm);
}

void launch_skin(const Group &g) {
  hipLaunchKernelGGL(k_skin_blocks, grid(g.np_blocks, g), dim3(kBlock), 0, g.stream, g.d);
}

}  // namespace mfe

// Bricks above 64 KB of LDS (meshes whose only divisors >= order-1 are
// large, e.g. 2*23) need the per-kernel dynamic LDS limit raised.
void mfe_allow_big_spread_lds(int prec, int bytes) {
  // the attribute only ever grows (per precision): a smaller brick set up
  // later must not take the room of an earlier, larger one
  static int allowed[2] = {0, 0};
  int &al = allowed[prec == MFE_PREC_MIXED ? 1 : 0];
  if (bytes <= al) return;
  al = bytes;
#define MFE_ATTR(R, O)                                                              \
  (void)hipFuncSetAttribute((const void *)mfe::k_spread_ptr<R, O>(),                 \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);        \
  for (int gm = 0; gm < 4; ++gm)                                                     \
    (void)hipFuncSetAttribute(mfe::k_gather_ptr<R, O>(gm, false),                     \
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);      \
  (void)hipFuncSetAttribute(mfe::k_gather_ptr<R, O>(mfe::GM_FORCE, true),             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);        \
  (void)hipFuncSetAttribute(mfe::k_fm_spread_ptr<R, O>(),                             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (prec == MFE_PREC_MIXED) {
    MFE_ATTR(float, 3) MFE_ATTR(float, 4) MFE_ATTR(float, 5)
    MFE_ATTR(float, 6) MFE_ATTR(float, 7) MFE_ATTR(float, 8)
  } else {
    MFE_ATTR(double, 3) MFE_ATTR(double, 4) MFE_ATTR(double, 5)
    MFE_ATTR(double, 6) MFE_ATTR(double, 7) MFE_ATTR(double, 8)
  }
#undef MFE_ATTR
}
