// HIP kernels for gfx950 (CDNA4): block-tile multigrid operators, FP64 Krylov
// vector kernels with two-stage deterministic reductions, and face kernels of
// the two-phase coefficient / projection / adjoint path.
//
// Data layout (every multigrid level): cells are block-major, BX x BY cells
// per block, j-major inside a block (the reference's cell order,
// src/mesh/topology/mesh.jl:28-30). A thread owns four consecutive cells of
// one block row; a 256-thread workgroup owns 256/(BX*BY/4) blocks, each with
// an LDS tile and a one-cell halo gathered from the four same-level neighbour
// blocks. Conductances of regular faces are stored like the cells (west and
// south face of each cell) plus the east column and north row of each block;
// rows next to coarse-fine interfaces or off-diagonal metric entries are
// fixed-width rows grouped per block whose columns are LDS slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "kernels.hpp"

namespace tpamr {

// ---------------------------------------------------------------------------
// helpers

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <typename T>
__device__ inline T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Sum over the workgroup; result valid in thread 0. `scratch` holds one
// value per wavefront.
template <typename T, int NT>
__device__ inline T block_sum(T v, T* scratch) {
  constexpr int NW = (NT + 63) / 64;
  v = wave_sum(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
      for (int q = 0; q < NW; ++q) s += scratch[q];
    __syncthreads();
    return s;
  }
}

template <typename T, int NT>
__device__ inline T block_max(T v, T* scratch) {
  constexpr int NW = (NT + 63) / 64;
  v = wave_max(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
      for (int q = 0; q < NW; ++q) s = fmax(s, scratch[q]);
    __syncthreads();
    return s;
  }
}

// Global fine-cell index -> owning coarse cell under the tentative transfer
// (identity block or one quadrant of a coarse block).
template <int BX, int BY>
__device__ inline int owner_of(int cell, const int* __restrict__ xfer) {
  constexpr int NB = BX * BY;
  const int blk = cell / NB, loc = cell - blk * NB;
  const int cb = xfer[2 * blk], mode = xfer[2 * blk + 1];
  if (mode < 0) return cb * NB + loc;
  const int i = loc % BX, j = loc / BX;
  const int ci = (mode & 1) * (BX / 2) + (i >> 1);
  const int cj = (mode >> 1) * (BY / 2) + (j >> 1);
  return cb * NB + cj * BX + ci;
}

// Logical block of workgroup `w`: consecutive blocks (same-level
// neighbours) go to the same XCD, whose L2 then serves the halo reads.
// Workgroups are dispatched to the 8 XCDs round-robin.
__device__ inline int xcd_block(int w, int G) {
  const int q = G >> 3, rem = G & 7, k = w & 7;
  return k * q + min(k, rem) + (w >> 3);
}

// Block group of tile workgroup `w` for a grid that is a multiple of 8*S:
// the XCD (w mod 8) takes runs of S consecutive groups spread over the whole
// block order. Neighbouring blocks in a run share the XCD's L2 for halo
// reads, and every XCD sees a similar mix of block levels and
// coarse-fine interfaces.
__device__ inline int xcd_chunk_group(int w, int S) {
  const int x = w & 7, s = w >> 3;
  return ((s / S) * 8 + x) * S + s % S;
}

#include "tile.hpp"

// ---------------------------------------------------------------------------
// coarsest level

// Coarsest level: x = inv * b with the dense inverse of the augmented
// (volume-constrained) operator, evaluated in FP64. The inverse is symmetric
// (so is the augmented operator), so its column-major storage is read row by
// row: one wavefront per row streams that row contiguously and sums with
// shuffles; four rows per workgroup.
template <typename T>
__global__ void __launch_bounds__(256) k_coarse_solve(int n, const double* __restrict__ inv, const T* __restrict__ b,
                                                     T* __restrict__ x, const int* __restrict__ done) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const double* row = inv + size_t(r) * n;
  double s0 = 0, s1 = 0;
  int c = lane;
  for (; c + 64 < n; c += 128) {
    s0 += row[c] * double(b[c]);
    s1 += row[c + 64] * double(b[c + 64]);
  }
  if (c < n) s0 += row[c] * double(b[c]);
  double s = s0 + s1;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) x[r] = T(s);
}

// ---------------------------------------------------------------------------
// FP64 outer Krylov kernels

// Sum (or max) of part[0..n) by one workgroup: 8 independent accumulators
// per thread keep many loads in flight.
template <bool MAX>
__device__ inline double wg_reduce(int n, const double* __restrict__ part, double* red) {
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int q = threadIdx.x;
  for (; q + 7 * 1024 < n; q += 8 * 1024) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double v = part[q + k * 1024];
      a[k] = MAX ? fmax(a[k], v) : a[k] + v;
    }
  }
  for (; q < n; q += 1024) a[0] = MAX ? fmax(a[0], part[q]) : a[0] + part[q];
  double v = a[0];
#pragma unroll
  for (int k = 1; k < 8; ++k) v = MAX ? fmax(v, a[k]) : v + a[k];
  return MAX ? block_max<double, 1024>(v, red) : block_sum<double, 1024>(v, red);
}

// Single-workgroup reduction of `nsets` sets of `n` partials (sum or max).
__global__ void __launch_bounds__(1024) k_reduce(int n, int nsets, const double* __restrict__ part,
                                                 double* __restrict__ out, int is_max) {
  __shared__ double red[16];
  for (int s = 0; s < nsets; ++s) {
    const double v = is_max ? wg_reduce<true>(n, part + size_t(s) * n, red) : wg_reduce<false>(n, part + size_t(s) * n, red);
    if (threadIdx.x == 0) out[s] = v;
    __syncthreads();
  }
}

// Two-stage deterministic reductions for the Krylov scalar steps: each of
// RED_WG workgroups reduces a fixed contiguous chunk of the per-block
// partials, publishes its value, and takes a ticket; the workgroup holding
// the last ticket reduces the RED_WG values in index order and finishes the
// scalar step. Nothing waits on another workgroup.
__device__ inline double chunk_sum(const double* __restrict__ part, int n, double* red) {
  const int ch = (n + RED_WG - 1) / RED_WG, lo = blockIdx.x * ch, hi = min(n, lo + ch);
  double a0 = 0, a1 = 0;
  int q = lo + threadIdx.x;
  for (; q + 256 < hi; q += 512) {
    a0 += part[q];
    a1 += part[q + 256];
  }
  if (q < hi) a0 += part[q];
  return block_sum<double, 256>(a0 + a1, red);
}

__device__ inline double chunk_max(const double* __restrict__ part, int n, double* red) {
  const int ch = (n + RED_WG - 1) / RED_WG, lo = blockIdx.x * ch, hi = min(n, lo + ch);
  double a = 0;
  for (int q = lo + threadIdx.x; q < hi; q += 256) a = fmax(a, part[q]);
  return block_max<double, 256>(a, red);
}

// Publish this workgroup's values (thread 0) and report whether it holds
// the last ticket; the ticket counter is left at zero for the next use.
template <int NS>
__device__ inline bool publish_last(const double* v, double* __restrict__ stage, int* ticket) {
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) __hip_atomic_store(stage + k * RED_WG + blockIdx.x, v[k], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    const int t = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == int(gridDim.x) - 1;
    if (last) __hip_atomic_store(ticket, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return last;
}

template <bool MAX>
__device__ inline double stage_reduce(const double* __restrict__ stage, double* red) {
  double v = 0;
  if (threadIdx.x < RED_WG) v = __hip_atomic_load(stage + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return MAX ? block_max<double, 256>(v, red) : block_sum<double, 256>(v, red);
}

// Krylov scalar step 1: pAp and the step length (Krylov.jl cg semantics).
// With `sums`, part also holds per-block sums of Ap (part[nb + b]) and the
// volume mean the residual update subtracts becomes the exact mean of
// r - alpha Ap (rmean holds the mean of r on entry).
__global__ void __launch_bounds__(256) k_cg_alpha(int nb, const double* __restrict__ part, double* __restrict__ stage,
                                                  CgState* st, int sums, double total_volume) {
  __shared__ double red[4];
  if (st->done) return;
  double v[2];
  v[0] = chunk_sum(part, nb, red);
  v[1] = sums ? chunk_sum(part + nb, nb, red) : 0.0;
  if (!publish_last<2>(v, stage, &st->ticket)) return;
  const double pAp = stage_reduce<false>(stage, red);
  const double sap = sums ? stage_reduce<false>(stage + RED_WG, red) : 0.0;
  if (threadIdx.x == 0) {
    st->pAp = pAp;
    const double eps = 2.220446049250313e-16;
    if (pAp <= eps * st->pnorm2 && fabs(pAp) <= eps * st->pnorm2) {
      st->done = 1;
      st->reason = 3;  // zero curvature
    } else {
      st->alpha = st->gamma / pAp;
      if (sums) st->rmean -= st->alpha * sap / total_volume;
    }
  }
}

// x += alpha p; r -= alpha Ap; emits the strong residual maximum partials
// and the V-cycle input on the finest level (b0 = r, x0 = w D^-1 r).
template <typename TV>
__global__ void k_cg_update_xr(int n, int nb_cells, const double* __restrict__ p, const double* __restrict__ Ap,
                               double* __restrict__ x, double* __restrict__ r, const double* __restrict__ volume,
                               TV* __restrict__ vb, TV* __restrict__ vx, const TV* __restrict__ dinv, double omega,
                               double* __restrict__ part, CgState* st, int init) {
  __shared__ double red[16];
  if (st->done) return;
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double q = 0;
  if (c < size_t(n)) {
    double rv;
    if (init) {
      rv = r[c] - Ap[c];  // r holds b on entry; Ap holds A x0
      r[c] = rv;
    } else {
      const double a = st->alpha;
      x[c] += a * p[c];
      // the volume mean of r (lagged by one iteration) is removed so the
      // recursive residual stays compatible: sum(A p) is zero only up to
      // round-off, and that drift cannot be reduced by the iteration
      rv = (r[c] - a * Ap[c]) - volume[c / nb_cells] * st->rmean;
      r[c] = rv;
    }
    const TV rt = TV(rv);
    vb[c] = rt;
    if (vx) vx[c] = TV(omega) * dinv[c] * rt;
    q = fabs(rv) / volume[c / nb_cells];
    if (q != q) q = __longlong_as_double(0x7ff8000000000000LL);
  }
  const double m = block_max<double, 256>(q, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Krylov scalar step 2: z mean removal, gamma, beta and the stopping tests.
__global__ void __launch_bounds__(256) k_cg_beta(int nb, int nmax, const double* __restrict__ part,
                                                 const double* __restrict__ maxpart, double* __restrict__ stage,
                                                 double total_volume, CgState* st, int init) {
  __shared__ double red[4];
  if (st->done) return;
  double v[4];
  v[0] = chunk_sum(part, nb, red);
  v[1] = chunk_sum(part + nb, nb, red);
  v[2] = chunk_sum(part + 2 * nb, nb, red);
  v[3] = chunk_max(maxpart, nmax, red);
  if (!publish_last<4>(v, stage, &st->ticket)) return;
  const double sz = stage_reduce<false>(stage, red);
  const double srz = stage_reduce<false>(stage + RED_WG, red);
  const double sr = stage_reduce<false>(stage + 2 * RED_WG, red);
  const double mx = stage_reduce<true>(stage + 3 * RED_WG, red);
  if (threadIdx.x != 0) return;
  const double m = sz / total_volume;
  st->rmean = sr / total_volume;
  const double g = srz - m * sr;
  st->zmean = m;
  st->strong = mx;
  if (init) {
    st->gamma = g;
    st->pnorm2 = g;
    st->rnorm = sqrt(g);
    st->beta = 0;
    st->iter = 0;
    st->rlast = mx;
    if (g < 0) {
      st->done = 1;
      st->reason = 5;  // indefinite preconditioner
    } else if (g == 0) {
      st->done = 1;
      st->reason = 1;
    }
    return;
  }
  const double rn = sqrt(g);
  st->rnorm = rn;
  // Krylov.jl's test on the unscaled residual norm (the refinement mode's
  // inner vectors carry the factor `scale`)
  const bool solved = rn / st->scale + 1.0 <= 1.0;
  // A replaced residual (FP64 b - A x, while the inner iteration's operator
  // is FP32) restarts the recurrence (p = z, as at a fresh start), or with
  // `keep` continues it (a reliable update: beta from the replaced residual).
  const bool replaced_now = st->replace != 0;
  const bool restart = replaced_now && !st->keep;
  if (!solved) {
    const double beta = restart ? 0.0 : g / st->gamma;
    st->pnorm2 = g + beta * beta * st->pnorm2;
    st->gamma = g;
    st->beta = beta;
  }
  st->iter += 1;
  if (replaced_now) {
    st->rlast = mx;
    st->replace = 0;
    st->replaced += 1;
  } else {
    st->rlast = fmax(st->rlast, mx);
  }
  const bool tired = st->iter >= st->itmax;
  const bool user = mx <= st->target;
  if (solved || tired || user) {
    st->done = 1;
    st->reason = solved ? 1 : (user ? 2 : 4);
  } else if (st->delta > 0 && mx <= st->delta * st->rlast) {
    // Refinement mode: the rounding of the FP32 correction and residual
    // bounds the true residual at about 1e-3 of the largest strong residual
    // since the last replacement, so the residual is replaced (and the
    // recurrence restarted) once it has fallen by delta below that value.
    st->replace = 1;
  }
}

// After a residual replacement: the strong maximum of the FP64 residual
// decides the stop (the Krylov callback's test) before the V-cycle runs;
// the iteration's step is counted here when it stops.
__global__ void __launch_bounds__(256) k_ir_check(int ng, const double* __restrict__ maxpart,
                                                  double* __restrict__ stage, CgState* st) {
  __shared__ double red[4];
  if (st->done || !st->replace) return;
  const double v = chunk_max(maxpart, ng, red);
  if (!publish_last<1>(&v, stage, &st->ticket)) return;
  const double mx = stage_reduce<true>(stage, red);
  if (threadIdx.x == 0 && mx <= st->target) {
    st->strong = mx;
    st->iter += 1;
    st->replaced += 1;
    st->done = 1;
    st->reason = 2;
  }
}

// p = (z - zmean) + beta p
template <typename TV>
__global__ void k_cg_update_p(int n, const TV* __restrict__ z, double* __restrict__ p, const CgState* st,
                              const int* __restrict__ done) {
  if (*done) return;
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= size_t(n)) return;
  p[c] = (double(z[c]) - st->zmean) + st->beta * p[c];
}

// ---------------------------------------------------------------------------
// Inner FP32 Krylov iteration of the mixed-precision refinement loop: the
// vectors e (correction), r (residual, also the finest level's V-cycle rhs),
// p and Ap are FP32; scalars, inner products and the update arithmetic are
// FP64. A thread owns four consecutive cells (n is a multiple of 4, cells
// of one block share a volume).

__device__ inline float4 ld_f4(const float* __restrict__ p) { return *reinterpret_cast<const float4*>(p); }
__device__ inline void st_f4(float* __restrict__ p, float a, float b, float c, float d) {
  *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
}

// init: r = scale * r64, e = 0. Otherwise e += alpha p and
// r = (r - alpha Ap) - V rmean, rmean being the exact volume mean of
// r - alpha Ap from the Ap sums of the matvec (k_cg_alpha); when the state
// requests a residual replacement and x64 is given, e then moves into the
// FP64 solution x64 (k_ir_correct's step, folded in).
// Both write the zero-start V-cycle iterate x0 = w D^-1 r and the per-
// workgroup maxima of |r| / V.
__global__ void __launch_bounds__(256) k_ir_update_xr(int n, int nb_cells, const float* __restrict__ p,
                                                      const float* __restrict__ Ap, float* __restrict__ e,
                                                      float* __restrict__ r, const double* __restrict__ volume,
                                                      float* __restrict__ vx, const float* __restrict__ dinv,
                                                      float omega, const double* __restrict__ r64,
                                                      double* __restrict__ part, CgState* st, int init,
                                                      double* __restrict__ x64) {
  __shared__ double red[4];
  if (st->done) return;
  const int c = 4 * (blockIdx.x * 256 + threadIdx.x);
  double q = 0;
  if (c < n) {
    const double V = volume[c / nb_cells];
    float rv[4];
    if (init) {
      const double scale = st->scale;
      const double2 a = *reinterpret_cast<const double2*>(r64 + c);
      const double2 b = *reinterpret_cast<const double2*>(r64 + c + 2);
      rv[0] = float(scale * a.x);
      rv[1] = float(scale * a.y);
      rv[2] = float(scale * b.x);
      rv[3] = float(scale * b.y);
      st_f4(e + c, 0.f, 0.f, 0.f, 0.f);
    } else {
      const double al = st->alpha, sh = V * st->rmean;
      const float4 pv = ld_f4(p + c), av = ld_f4(Ap + c), ev = ld_f4(e + c), ro = ld_f4(r + c);
      const float en[4] = {float(double(ev.x) + al * double(pv.x)), float(double(ev.y) + al * double(pv.y)),
                           float(double(ev.z) + al * double(pv.z)), float(double(ev.w) + al * double(pv.w))};
      if (x64 && st->replace) {
        // a residual replacement follows: the correction moves into the
        // FP64 solution now (x += e / scale, e = 0)
        const double sc = 1.0 / st->scale;
        double2 a = *reinterpret_cast<const double2*>(x64 + c);
        double2 b = *reinterpret_cast<const double2*>(x64 + c + 2);
        a.x += sc * double(en[0]);
        a.y += sc * double(en[1]);
        b.x += sc * double(en[2]);
        b.y += sc * double(en[3]);
        *reinterpret_cast<double2*>(x64 + c) = a;
        *reinterpret_cast<double2*>(x64 + c + 2) = b;
        st_f4(e + c, 0.f, 0.f, 0.f, 0.f);
      } else {
        st_f4(e + c, en[0], en[1], en[2], en[3]);
      }
      rv[0] = float((double(ro.x) - al * double(av.x)) - sh);
      rv[1] = float((double(ro.y) - al * double(av.y)) - sh);
      rv[2] = float((double(ro.z) - al * double(av.z)) - sh);
      rv[3] = float((double(ro.w) - al * double(av.w)) - sh);
    }
    st_f4(r + c, rv[0], rv[1], rv[2], rv[3]);
    if (vx) {
      const float4 dv = ld_f4(dinv + c);
      st_f4(vx + c, omega * dv.x * rv[0], omega * dv.y * rv[1], omega * dv.z * rv[2], omega * dv.w * rv[3]);
    }
    const float m = fmaxf(fmaxf(fabsf(rv[0]), fabsf(rv[1])), fmaxf(fabsf(rv[2]), fabsf(rv[3])));
    q = double(m) / V;
    if (q != q || rv[0] != rv[0] || rv[1] != rv[1] || rv[2] != rv[2] || rv[3] != rv[3])
      q = __longlong_as_double(0x7ff8000000000000LL);
  }
  const double m = block_max<double, 256>(q, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// p = (z - zmean) + beta p
__global__ void __launch_bounds__(256) k_ir_update_p(int n, const float* __restrict__ z, float* __restrict__ p,
                                                     const CgState* st, const int* __restrict__ done) {
  if (*done) return;
  const int c = 4 * (blockIdx.x * 256 + threadIdx.x);
  if (c >= n) return;
  const double zm = st->zmean, be = st->beta;
  const float4 zv = ld_f4(z + c), pv = ld_f4(p + c);
  st_f4(p + c, float((double(zv.x) - zm) + be * double(pv.x)), float((double(zv.y) - zm) + be * double(pv.y)),
        float((double(zv.z) - zm) + be * double(pv.z)), float((double(zv.w) - zm) + be * double(pv.w)));
}

// x += e / scale, e = 0 (the correction accumulated in FP32 since the last
// replacement moves into the FP64 solution): when the state requests a
// residual replacement, or always with `force`.
__global__ void __launch_bounds__(256) k_ir_correct(int n, float* __restrict__ e, double* __restrict__ x,
                                                    const CgState* st, int force) {
  if (!force && (st->done || !st->replace)) return;
  const int c = 4 * (blockIdx.x * 256 + threadIdx.x);
  if (c >= n) return;
  const double scale = 1.0 / st->scale;
  const float4 ev = ld_f4(e + c);
  st_f4(e + c, 0.f, 0.f, 0.f, 0.f);
  double2 a = *reinterpret_cast<const double2*>(x + c);
  double2 b = *reinterpret_cast<const double2*>(x + c + 2);
  a.x += scale * double(ev.x);
  a.y += scale * double(ev.y);
  b.x += scale * double(ev.z);
  b.y += scale * double(ev.w);
  *reinterpret_cast<double2*>(x + c) = a;
  *reinterpret_cast<double2*>(x + c + 2) = b;
}

__global__ void k_fill_f32(int n, float v, float* __restrict__ out) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < size_t(n)) out[c] = v;
}

// Partial sums of V z, r z and r per mesh block (used when the V-cycle does
// not end with a finest-level sweep).
template <typename T, int BX, int BY, typename TR>
__global__ void __launch_bounds__(BX* BY) k_zstats(const T* __restrict__ z, const TR* __restrict__ rz,
                                                  const double* __restrict__ volume, double* __restrict__ part,
                                                  const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[16];
  const int b = blockIdx.x;
  const size_t c = size_t(b) * (BX * BY) + threadIdx.x;
  const double zd = double(z[c]), rd = double(rz[c]);
  const double s_z = block_sum<double, BX * BY>(zd, red);
  const double s_rz = block_sum<double, BX * BY>(rd * zd, red);
  const double s_r = block_sum<double, BX * BY>(rd, red);
  if (threadIdx.x == 0) {
    const int nb = gridDim.x;
    part[b] = s_z * volume[b];
    part[nb + b] = s_rz;
    part[2 * nb + b] = s_r;
  }
}

// ---------------------------------------------------------------------------
// elementwise / face kernels

// `_projection_rhs!` (navier_stokes/projection.jl:130-156): rhs = -div / dt,
// plus w * inherited with w = (tau - dt) / (dt tau) when a shortened step
// keeps part of the inherited divergence (inherited = nullptr: none)
__global__ void k_projection_rhs(int n, const double* __restrict__ div, const double* __restrict__ inherited,
                                 double dt, double w, double* __restrict__ rhs) {
#pragma clang fp contract(off)
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= size_t(n)) return;
  const double v = -div[c] / dt;
  rhs[c] = inherited ? v + w * inherited[c] : v;
}

// y = a x + b y; with b = 0, y is only written (it may hold anything, NaN
// included, before); x may alias y
// open boundary faces after a projection (_project_prepared!,
// projection.jl:248-255): ub -= dt g_b with the boundary weighted gradient
// g_b = (-side * bgrad) * phi_cell (_elliptic_boundary_gradient!,
// elliptic.jl:1254-1263); closed faces have bgrad 0
__global__ void k_boundary_velocity(int nbf, const int* __restrict__ bcell, const signed char* __restrict__ bside,
                                    const double* __restrict__ bgrad, const double* __restrict__ phi, double dt,
                                    double* __restrict__ ub) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nbf) return;
  const double g = (-double(bside[f]) * bgrad[f]) * phi[bcell[f]];
  ub[f] = ub[f] - dt * g;
}

__global__ void k_axpby(int n, double a, const double* x, double b, double* y) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < size_t(n)) y[c] = b != 0.0 ? a * x[c] + b * y[c] : a * x[c];
}

// r = b - y with per-block strong maxima |r|/V
__global__ void k_residual_strong(int n, int nb_cells, const double* __restrict__ b, const double* __restrict__ y,
                                  double* __restrict__ r, const double* __restrict__ volume,
                                  double* __restrict__ part) {
  __shared__ double red[16];
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double q = 0, s2 = 0;
  if (c < size_t(n)) {
    const double rv = b[c] - y[c];
    r[c] = rv;
    const double V = volume[c / nb_cells];
    q = fabs(rv) / V;
    if (q != q) q = __longlong_as_double(0x7ff8000000000000LL);
    s2 = V * (rv / V) * (rv / V);
  }
  const double m = block_max<double, 256>(q, red);
  const double s = block_sum<double, 256>(s2, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = m;
    part[gridDim.x + blockIdx.x] = s;
  }
}

// per-block sums of V x (for the volume-weighted mean) and of x
__global__ void k_weighted_sums(int n, int nb_cells, const double* __restrict__ x, const double* __restrict__ volume,
                                double* __restrict__ part) {
  __shared__ double red[16];
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v = 0, a = 0, s = 0;
  if (c < size_t(n)) {
    v = volume[c / nb_cells] * x[c];
    a = fabs(x[c]);
    s = x[c];
  }
  v = block_sum<double, 256>(v, red);
  a = block_sum<double, 256>(a, red);
  s = block_sum<double, 256>(s, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v;
    part[gridDim.x + blockIdx.x] = a;
    part[2 * gridDim.x + blockIdx.x] = s;
  }
}

// Start of a solve: the integrated rhs b <- V b in place, and per-workgroup
// partial sums of |b|, b (integrated), V x and |x| (four sets).
__global__ void k_rhs_prepare(int n, int nb_cells, double* __restrict__ b, const double* __restrict__ x,
                              const double* __restrict__ volume, double* __restrict__ part) {
  __shared__ double red[16];
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double a = 0, s = 0, vx = 0, ax = 0;
  if (c < size_t(n)) {
    const double V = volume[c / nb_cells];
    const double bi = b[c] * V;
    b[c] = bi;
    a = fabs(bi);
    s = bi;
    vx = V * x[c];
    ax = fabs(x[c]);
  }
  a = block_sum<double, 256>(a, red);
  s = block_sum<double, 256>(s, red);
  vx = block_sum<double, 256>(vx, red);
  ax = block_sum<double, 256>(ax, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = s;
    part[2 * gridDim.x + blockIdx.x] = vx;
    part[3 * gridDim.x + blockIdx.x] = ax;
  }
}

// b -= V mean; with a zero solution the residual is r = b: per-workgroup
// maxima of |r| / V and sums of V (r / V)^2.
__global__ void k_shift_residual(int n, int nb_cells, double* __restrict__ b, double mean,
                                 const double* __restrict__ volume, double* __restrict__ r, double* __restrict__ part) {
  __shared__ double red[16];
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double q = 0, s2 = 0;
  if (c < size_t(n)) {
    const double V = volume[c / nb_cells];
    const double rv = b[c] - V * mean;
    b[c] = rv;
    r[c] = rv;
    q = fabs(rv) / V;
    if (q != q) q = __longlong_as_double(0x7ff8000000000000LL);
    s2 = V * (rv / V) * (rv / V);
  }
  const double m = block_max<double, 256>(q, red);
  const double s = block_sum<double, 256>(s2, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = m;
    part[gridDim.x + blockIdx.x] = s;
  }
}

// Haloed cell field: interior values (cell c of block b to position
// (i + h, j + h) of the block's (bx + 2h) x (by + 2h) array) ...
__global__ void k_halo_interior(int n, int bx, int by, int h, const double* __restrict__ cells,
                                double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int nb = bx * by, b = c / nb, l = c - b * nb, j = l / bx, i = l - j * bx;
  const int W = bx + 2 * h, H = by + 2 * h;
  out[size_t(b) * W * H + (j + h) * W + (i + h)] = cells[c];
}

// ... and the halo positions: the weighted donor sum in schedule order, or a
// physical entry (the nearest interior cell, or a prescribed value).
__global__ void k_halo_fill(int ne, const int* __restrict__ dest, const int* __restrict__ ptr,
                            const int* __restrict__ donor, const double* __restrict__ weight,
                            const int* __restrict__ phys, const double* __restrict__ cells,
                            const double* __restrict__ physical, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int p0 = ptr[e], p1 = ptr[e + 1], k = phys[e];
  double v;
  if (k >= 0) {
    v = physical ? physical[k] : cells[donor[p0]];
  } else {
    v = 0.0;
    for (int q = p0; q < p1; ++q) v += weight[q] * cells[donor[q]];
  }
  out[dest[e]] = v;
}

// x -= shift (scalar from device memory), optionally scaled by the cell volume
__global__ void k_shift(int n, int nb_cells, double* __restrict__ x, const double* __restrict__ shift, double scale,
                        const double* __restrict__ volume, int by_volume) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= size_t(n)) return;
  const double s = *shift * scale;
  x[c] -= by_volume ? volume[c / nb_cells] * s : s;
}

__global__ void k_scale_by_volume(int n, int nb_cells, const double* __restrict__ in, double* __restrict__ out,
                                  const double* __restrict__ volume, int divide) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= size_t(n)) return;
  const double V = volume[c / nb_cells];
  out[c] = divide ? in[c] / V : in[c] * V;
}

// K y on face vectors: diagonal plus off-diagonal CSR (symmetric)
__global__ void k_metric_apply(int nf, const double* __restrict__ kdiag, const int* __restrict__ optr,
                               const int* __restrict__ oface, const double* __restrict__ oval,
                               const double* __restrict__ in, double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double s = kdiag[f] * in[f];
  for (int q = optr[f]; q < optr[f + 1]; ++q) s += oval[q] * in[oface[q]];
  out[f] = s;
}

// face jump sqrt(k) |f| (p+ - p-), or sqrt(k) * j when p == nullptr
__global__ void k_face_jump(int nf, const int* __restrict__ fm, const int* __restrict__ fp,
                            const double* __restrict__ area, const double* __restrict__ sqk,
                            const double* __restrict__ p, const double* __restrict__ jumps, double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double w = sqk ? sqk[f] : 1.0;
  out[f] = p ? w * (area[f] * (p[fp[f]] - p[fm[f]])) : w * jumps[f];
}

__global__ void k_face_scale(int nf, const double* __restrict__ sqk, const double* __restrict__ in,
                             double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) out[f] = sqk[f] * in[f];
}

// mixture face coefficient kappa_f = 1 / (rho_l c_f + rho_g (1 - c_f))
__global__ void k_mixture(int nf, int nb_cells, const int* __restrict__ fm, const int* __restrict__ fp,
                          const signed char* __restrict__ axis, const double* __restrict__ spacing,
                          const double* __restrict__ c, double rho_l, double rho_g, double* __restrict__ kappa) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int m = fm[f], p = fp[f], a = axis[f];
  const double hm = spacing[2 * (m / nb_cells) + a], hp = spacing[2 * (p / nb_cells) + a];
  double cf = (hm * c[m] + hp * c[p]) / (hm + hp);
  // Julia's clamp: a NaN fraction stays NaN
  cf = cf > 1.0 ? 1.0 : (cf < 0.0 ? 0.0 : cf);
  kappa[f] = 1.0 / (rho_l * cf + rho_g * (1 - cf));
}

// Dense inverse of the augmented coarsest operator, Gauss-Jordan in place
// with partial pivoting (rows swapped, pivot = first largest |a_ik|). Step k:
// k_gj_pivot (one workgroup) picks and swaps the pivot row, scales it by
// 1 / a_kk (a_kk <- 1 / a_kk) and copies column k (the multipliers); then
// k_gj_update eliminates column k from every other row (a_ik <- -m_i / a_kk).
// The row swaps come back as column swaps in reverse order (k_gj_unpermute).
__global__ void __launch_bounds__(1024) k_gj_pivot(int N, int k, double* __restrict__ a, int* __restrict__ piv,
                                                   double* __restrict__ col, int* __restrict__ fail) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double best = -1.0;
  int bi = N;
  for (int i = k + t; i < N; i += 1024) {
    const double v = fabs(a[size_t(i) * N + k]);
    if (v > best) {
      best = v;
      bi = i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_down(best, off);
    const int oi = __shfl_down(bi, off);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  __shared__ double sb[16];
  __shared__ int si[16];
  __shared__ int sp;
  __shared__ double sv;
  if (lane == 0) {
    sb[w] = best;
    si[w] = bi;
  }
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < 16; ++q)
      if (sb[q] > sb[0] || (sb[q] == sb[0] && si[q] < si[0])) {
        sb[0] = sb[q];
        si[0] = si[q];
      }
    const bool ok = sb[0] > 0.0 && isfinite(sb[0]) && si[0] < N;
    sp = ok ? si[0] : k;
    sv = ok ? sb[0] : 0.0;
    piv[k] = sp;
    if (!ok) *fail = 1;
  }
  __syncthreads();
  if (!(sv > 0.0)) return;
  const int p = sp;
  double* rk = a + size_t(k) * N;
  if (p != k) {
    double* rp = a + size_t(p) * N;
    for (int j = t; j < N; j += 1024) {
      const double x = rk[j];
      rk[j] = rp[j];
      rp[j] = x;
    }
  }
  __syncthreads();
  const double inv = 1.0 / rk[k];
  __syncthreads();
  for (int i = t; i < N; i += 1024) col[i] = i == k ? 0.0 : a[size_t(i) * N + k];
  for (int j = t; j < N; j += 1024) rk[j] = j == k ? inv : rk[j] * inv;
}

// one row per blockIdx.y; rows with a zero multiplier are left alone
__global__ void k_gj_update(int N, int k, double* __restrict__ a, const double* __restrict__ col) {
  const int i = blockIdx.y;
  if (i == k) return;
  const double m = col[i];
  if (m == 0.0) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double* rk = a + size_t(k) * N;
  double* ri = a + size_t(i) * N;
  ri[j] = j == k ? -m * rk[k] : ri[j] - m * rk[j];
}

// Blocked Gauss-Jordan: the columns [c0, c0 + nb) form a panel. For each
// column k of the panel, k_gjb_pivot picks the pivot row among k..N-1,
// swaps rows k and p inside the panel's columns only, scales row k's panel
// entries by 1 / a_kk (a_kk <- 1 / a_kk) and copies column k (the
// multipliers); k_gjb_update eliminates column k from the panel's columns
// of every other row. The panel then holds the panel columns of the
// elimination matrix E (E A's panel columns are the identity's); the other
// columns receive the panel's row swaps (k_gjb_swap_rest) and the update
// A_rest <- A_rest + (E_panel - J) A_rest[panel rows] (J: the identity's panel
// columns) as two DGEMMs, from the copies k_gjb_prep makes.
__global__ void __launch_bounds__(1024) k_gjb_pivot(int N, int k, int c0, int nb, double* __restrict__ a,
                                                    int* __restrict__ piv, double* __restrict__ col,
                                                    int* __restrict__ fail) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double best = -1.0;
  int bi = N;
  for (int i = k + t; i < N; i += 1024) {
    const double v = fabs(a[size_t(i) * N + k]);
    if (v > best) {
      best = v;
      bi = i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_down(best, off);
    const int oi = __shfl_down(bi, off);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  __shared__ double sb[16];
  __shared__ int si[16];
  __shared__ int sp;
  __shared__ double sv;
  if (lane == 0) {
    sb[w] = best;
    si[w] = bi;
  }
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < 16; ++q)
      if (sb[q] > sb[0] || (sb[q] == sb[0] && si[q] < si[0])) {
        sb[0] = sb[q];
        si[0] = si[q];
      }
    const bool ok = sb[0] > 0.0 && isfinite(sb[0]) && si[0] < N;
    sp = ok ? si[0] : k;
    sv = ok ? sb[0] : 0.0;
    piv[k] = sp;
    if (!ok) *fail = 1;
  }
  __syncthreads();
  if (!(sv > 0.0)) return;
  const int p = sp;
  double* rk = a + size_t(k) * N;
  if (p != k) {
    double* rp = a + size_t(p) * N;
    for (int j = c0 + t; j < c0 + nb; j += 1024) {
      const double x = rk[j];
      rk[j] = rp[j];
      rp[j] = x;
    }
  }
  __syncthreads();
  const double inv = 1.0 / rk[k];
  __syncthreads();
  for (int i = t; i < N; i += 1024) col[i] = i == k ? 0.0 : a[size_t(i) * N + k];
  for (int j = c0 + t; j < c0 + nb; j += 1024) rk[j] = j == k ? inv : rk[j] * inv;
}

// one row per blockIdx.x, the panel's columns across the workgroup
__global__ void k_gjb_update(int N, int k, int c0, int nb, double* __restrict__ a, const double* __restrict__ col) {
  const int i = blockIdx.x;
  if (i == k) return;
  const double m = col[i];
  if (m == 0.0) return;
  const double* rk = a + size_t(k) * N;
  double* ri = a + size_t(i) * N;
  for (int j = c0 + threadIdx.x; j < c0 + nb; j += blockDim.x) ri[j] = j == k ? -m * rk[k] : ri[j] - m * rk[j];
}

// the panel's row swaps, in order, on every column outside the panel
__global__ void k_gjb_swap_rest(int N, int c0, int nb, double* __restrict__ a, const int* __restrict__ piv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N || (j >= c0 && j < c0 + nb)) return;
  for (int t = 0; t < nb; ++t) {
    const int k = c0 + t, p = piv[k];
    if (p != k) {
      const double x = a[size_t(k) * N + j];
      a[size_t(k) * N + j] = a[size_t(p) * N + j];
      a[size_t(p) * N + j] = x;
    }
  }
}

// R = rows c0 .. c0 + nb of a (nb x N, row major); M = E_panel - J (N x nb)
__global__ void k_gjb_prep(int N, int c0, int nb, const double* __restrict__ a, double* __restrict__ R,
                           double* __restrict__ M) {
  const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= size_t(N) * nb) return;
  const int t = int(q / N), j = int(q % N);
  R[q] = a[size_t(c0 + t) * N + j];
  const int i = int(q / nb), u = int(q % nb);
  M[q] = a[size_t(i) * N + c0 + u] - (i == c0 + u ? 1.0 : 0.0);
}

__global__ void k_gj_unpermute(int N, double* __restrict__ a, const int* __restrict__ piv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double* ri = a + size_t(i) * N;
  for (int k = N - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k) {
      const double x = ri[k];
      ri[k] = ri[p];
      ri[p] = x;
    }
  }
}

// top-left n x n block, column major
__global__ void k_gj_extract(int n, int N, const double* __restrict__ a, double* __restrict__ inv) {
  const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= size_t(n) * n) return;
  const size_t c = q / n, r = q % n;
  inv[q] = a[r * N + c];
}

// ---------------------------------------------------------------------------
// Surface-tension jumps with the face curvature of `_face_curvature`
// (src/two_phase/forces.jl:211-300, 345-378)

struct CurvMesh {
  const double* origin;   // 2 per block
  const double* spacing;  // 2 per block
  const double* fcx;      // face centres
  const double* fcy;
  int nblocks, bx, by;
  double extent[2];
  int periodic[2];
};

__device__ inline double wrapped_delta(double d, double extent, int periodic) {
#pragma clang fp contract(off)
  return periodic ? d - rint(d / extent) * extent : d;
}

// The LHF-first c(1-c) interpolant of the two cells' curvatures (NaN = none),
// else the c(1-c)-weighted extension of the cell curvatures within 3 h of the
// face centre over all blocks in order (rare: both cells pure); NaN when
// there is none. Rounded operation by operation like the reference.
__device__ double face_curvature(const CurvMesh& M, int f, int m, int p, const double* __restrict__ c,
                                 const double* __restrict__ kappa, const double* __restrict__ quality) {
#pragma clang fp contract(off)
  const double km = kappa[m], kp = kappa[p], cm = c[m], cp = c[p];
  const double wm = isfinite(km) ? cm * (1 - cm) : 0.0;
  const double wp = isfinite(kp) ? cp * (1 - cp) : 0.0;
  if (wm > 0 || wp > 0) {
    const double qm = wm > 0 ? quality[m] : 0.0;
    const double qp = wp > 0 ? quality[p] : 0.0;
    const double pref = fmax(qm, qp);
    double total = 0, weight = 0;
    if (wm > 0) {
      const double w = wm * (qm + (1 - pref) * (1 - qm));
      total += w * km;
      weight += w;
    }
    if (wp > 0) {
      const double w = wp * (qp + (1 - pref) * (1 - qp));
      total += w * kp;
      weight += w;
    }
    return total / weight;
  }
  const int nb = M.bx * M.by, bm = m / nb, bp = p / nb;
  const double local = fmax(fmax(M.spacing[2 * bm], M.spacing[2 * bm + 1]), fmax(M.spacing[2 * bp], M.spacing[2 * bp + 1]));
  const double radius = 3 * local;
  const double fx = M.fcx[f], fy = M.fcy[f];
  const double emax = fmax(M.extent[0], M.extent[1]);
  double total = 0, weight = 0;
  for (int b = 0; b < M.nblocks; ++b) {
    const double hx = M.spacing[2 * b], hy = M.spacing[2 * b + 1];
    const double ox = M.origin[2 * b], oy = M.origin[2 * b + 1];
    const double ex = hx * M.bx, ey = hy * M.by;
    const double hs = fmax(hx, hy);
    const double dx = fabs(wrapped_delta(ox + ex / 2 - fx, M.extent[0], M.periodic[0]));
    const double dy = fabs(wrapped_delta(oy + ey / 2 - fy, M.extent[1], M.periodic[1]));
    const double margin = 16 * 2.220446049250313e-16 *
                          fmax(fmax(fmax(emax, fabs(ox)), fmax(fabs(oy), fabs(fx))), fabs(fy));
    const double reach = radius + hs + margin;
    if (dx > ex / 2 + reach || dy > ey / 2 + reach) continue;
    for (int j = 0; j < M.by; ++j)
      for (int i = 0; i < M.bx; ++i) {
        const int cell = b * nb + j * M.bx + i;
        const double cc = c[cell], k = kappa[cell];
        const double sw = cc * (1 - cc);
        if (!(sw > 0 && isfinite(k))) continue;
        const double px = ox + (i + 0.5) * hx, py = oy + (j + 0.5) * hy;
        const double ddx = wrapped_delta(px - fx, M.extent[0], M.periodic[0]);
        const double ddy = wrapped_delta(py - fy, M.extent[1], M.periodic[1]);
        const double dist = hypot(ddx, ddy);
        if (!(dist <= radius + hs)) continue;
        const double gw = sw / fmax(dist * dist, (hs / 4) * (hs / 4));
        total += gw * k;
        weight += gw;
      }
  }
  return weight > 0 ? total / weight : __builtin_nan("");
}

__global__ void k_surface_jump_cells(int nf, const int* __restrict__ fm, const int* __restrict__ fp,
                                     const double* __restrict__ area, CurvMesh M, const double* __restrict__ c,
                                     const double* __restrict__ kappa, const double* __restrict__ quality,
                                     double sigma, int csf, double* __restrict__ jumps, int* __restrict__ unresolved) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int m = fm[f], p = fp[f];
  const double cm = c[m], cp = c[p];
  const double jump = csf ? cp - cm : double(cp > 0.5) - double(cm > 0.5);
  if (sigma == 0 || jump == 0) {
    jumps[f] = 0.0;
    return;
  }
  const double k = face_curvature(M, f, m, p, c, kappa, quality);
  if (k != k) {
    atomicAdd(unresolved, 1);
    jumps[f] = 0.0;
    return;
  }
  jumps[f] = sigma * k * jump * area[f];
}

// ---------------------------------------------------------------------------
// Coefficient refresh of one level on the device (the value pass of the
// host's refresh_coefficients, same operations in the same order, each
// product and sum rounded on its own)

// face coefficients of level l from level l - 1: k[f] = sum_q rw_q kprev[ridx_q]
__global__ void k_ref_restrict(int nf, const int* __restrict__ rptr, const int* __restrict__ ridx,
                               const double* __restrict__ rw, const double* __restrict__ kprev,
                               double* __restrict__ k) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double s = 0;
  for (int q = rptr[f]; q < rptr[f + 1]; ++q) s += rw[q] * kprev[ridx[q]];
  k[f] = s;
}

// regular faces: conductance k K_ff |f|^2 into its two slots of the device
// layout (dst code = index * 4 + array: 0 cw, 1 ce, 2 cs, 3 cn; -1: none)
__global__ void k_ref_regular(int nf, const double* __restrict__ k, const double* __restrict__ kdiag,
                              const double* __restrict__ area, const int* __restrict__ dst, double* cw, double* ce,
                              double* cs, double* cn) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int d0 = dst[2 * f], d1 = dst[2 * f + 1];
  if (d0 < 0) return;
  const double a = area[f];
  const double w = k[f] * kdiag[f] * a * a;
  double* arr[4] = {cw, ce, cs, cn};
  arr[d0 & 3][d0 >> 2] = w;
  arr[d1 & 3][d1 >> 2] = w;
}

// irregular rows, one thread per entry: the entry is the sum of its
// face-metric terms, in the host pass's order (kept in val_e, the entries in
// row order); the terms of non-regular faces (dst[2 f] < 0) also sum to the
// entry's ELL part (value and floor bound) when it has an ELL slot, and a
// diagonal entry's ELL bound part goes to its row's ell_bdiag
__global__ void k_ref_entries(int nnz, const int* __restrict__ term_ptr, const int* __restrict__ term_f,
                              const int* __restrict__ term_g, const double* __restrict__ term_k,
                              const double* __restrict__ term_s, const unsigned char* __restrict__ term_diag,
                              const double* __restrict__ k, const double* __restrict__ area,
                              const int* __restrict__ dst, const int* __restrict__ ell_pos,
                              const int* __restrict__ ell_dpos, double* __restrict__ ell_val,
                              double* __restrict__ ell_bound, double* __restrict__ ell_bdiag,
                              double* __restrict__ val_e) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double v = 0, vp = 0, bvp = 0;
  for (int q = term_ptr[e]; q < term_ptr[e + 1]; ++q) {
    const int f = term_f[q], g = term_g[q];
    const double kfg = term_k[q], af = area[f], ag = area[g];
    double core, babs;
    if (term_diag[q]) {
      core = k[f] * kfg * af * ag;
      babs = core;
    } else {
      const double sf = sqrt(k[f]), sg = sqrt(k[g]);
      core = (sf * kfg) * sg * af * ag;
      babs = (sf * fabs(kfg)) * sg * af * ag;
    }
    v += term_s[q] * core;
    if (dst[2 * f] < 0) {
      vp += term_s[q] * core;
      bvp += babs;
    }
  }
  const int pos = ell_pos[e];
  if (pos >= 0) {
    ell_val[pos] = vp;
    if (ell_bound) ell_bound[pos] = bvp;
  }
  if (ell_bdiag && ell_dpos[e] >= 0) ell_bdiag[ell_dpos[e]] = bvp;
  val_e[e] = v;
}

// open boundary faces, one thread per face: the conductance bw into the
// cell's slot beyond the side (dst code index * 4 + array)
__global__ void k_ref_boundary(int nopen, const double* __restrict__ bw, const int* __restrict__ dst, double* cw,
                               double* ce, double* cs, double* cn) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nopen) return;
  const int d = dst[k];
  double* arr[4] = {cw, ce, cs, cn};
  arr[d & 3][d >> 2] = bw[k];
}

// irregular rows, one thread per row: the row's diagonal (with the row's
// boundary diagonal bd when open sides are present) and the largest row
// ratio sum |a_rj| / a_rr (as the bits of a nonnegative double, whose order is
// the numeric order), summed in entry order
__global__ void k_ref_rows(int nrows, const int* __restrict__ irr_ptr, const int* __restrict__ diag_e,
                           const int* __restrict__ ell_row, const double* __restrict__ val_e,
                           double* __restrict__ ell_diag, double* __restrict__ irr_diag,
                           unsigned long long* __restrict__ ratio, const double* __restrict__ bd) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  // the largest row ratio of the wavefront, then one atomic per wavefront
  // (nonnegative doubles order as their bit patterns)
  double q = 0;
  if (r < nrows) {
    double rowsum = 0, diag = 0;
    const int de = diag_e[r];
    for (int e = irr_ptr[r]; e < irr_ptr[r + 1]; ++e) {
      double v = val_e[e];
      if (bd && e == de) v = v + bd[r];
      rowsum += fabs(v);
      if (e == de) diag = v;
    }
    irr_diag[r] = diag;
    ell_diag[ell_row[r]] = diag;
    if (diag > 0) q = rowsum / diag;
  }
  for (int off = 32; off > 0; off >>= 1) q = fmax(q, __shfl_xor(q, off, 64));
  if ((threadIdx.x & 63) == 0 && q > 0) atomicMax(ratio, (unsigned long long)__double_as_longlong(q));
}

// Jacobi diagonal inverse of every cell: the irregular row's diagonal, or the
// sum of the four conductances; flags a positive regular diagonal
__global__ void k_ref_dinv(int n, int bx, int by, const double* __restrict__ cw, const double* __restrict__ ce,
                           const double* __restrict__ cs, const double* __restrict__ cn,
                           const int* __restrict__ cell_irr, const double* __restrict__ irr_diag,
                           double* __restrict__ dinv, int* __restrict__ regular, const int* __restrict__ nbr,
                           unsigned long long* __restrict__ ratio) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  bool reg = false;
  double q = 0;
  if (c < n) {
    const int nb = bx * by, b = c / nb, qq = c - b * nb, i = qq % bx, j = qq / bx;
    double d;
    const int r = cell_irr[c];
    if (r >= 0) {
      d = irr_diag[r];
    } else {
      const double w = cw[c], e = i < bx - 1 ? cw[c + 1] : ce[b * by + j];
      const double s = cs[c], nn = j < by - 1 ? cs[c + bx] : cn[b * bx + i];
      d = w + e + s + nn;
      // a slot at the block edge without a neighbour block is a physical
      // side: with a positive conductance, an open boundary face
      bool ow = false, oe = false, os = false, on = false;
      if (nbr) {
        ow = i == 0 && nbr[4 * b] < 0;
        oe = i == bx - 1 && nbr[4 * b + 1] < 0;
        os = j == 0 && nbr[4 * b + 2] < 0;
        on = j == by - 1 && nbr[4 * b + 3] < 0;
      }
      const bool bnd = (ow && w > 0) || (oe && e > 0) || (os && s > 0) || (on && nn > 0);
      if (d > 0) {
        if (!bnd) {
          reg = true;
        } else {
          const double isum = (ow ? 0.0 : w) + (oe ? 0.0 : e) + (os ? 0.0 : s) + (on ? 0.0 : nn);
          q = (d + isum) / d;
        }
      }
    }
    dinv[c] = d > 0 ? 1.0 / d : 0.0;
  }
  // one store per wavefront that has a positive regular diagonal
  if (__any(reg) && (threadIdx.x & 63) == 0) *regular = 1;
  // the largest open-boundary row ratio of the wavefront, one atomic
  for (int off = 32; off > 0; off >>= 1) q = fmax(q, __shfl_xor(q, off, 64));
  if ((threadIdx.x & 63) == 0 && q > 0) atomicMax(ratio, (unsigned long long)__double_as_longlong(q));
}

// any face coefficient not finite or not positive
// face coefficients: bad[0] = 1 for a negative or non-finite one, bad[1] =
// 1 for a zero one (a split into components)
__global__ void k_ref_check(int nf, const double* __restrict__ k, int* __restrict__ bad) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double v = k[f];
  if (!isfinite(v) || v < 0) bad[0] = 1;
  else if (v == 0) bad[1] = 1;
}

// ---------------------------------------------------------------------------
// Pressure components: sums over each ungauged component in chunks of
// GAUGE_CHUNK_CELLS cells (`_chunked_slot_sums!`, threading.jl:202-283), then the
// projection of a right-hand side or the removal of volume means
// (hierarchy.jl:300-343, elliptic_solver.jl:130-190).

// one wavefront per (chunk, slot) pair: the sum of term(c) over the chunk's
// cells of that slot; mode 0: x, 1: V x, 2: |x|. Lanes take every 64th cell,
// then a fixed shuffle tree (deterministic)
template <typename T>
__global__ void __launch_bounds__(256) k_comp_partial(int npairs, int n, int nb_cells, const int* __restrict__ pair_chunk,
                                                      const int* __restrict__ pair_slot, const int* __restrict__ slot,
                                                      const T* __restrict__ x, const double* __restrict__ volume,
                                                      int mode, double* __restrict__ partial) {
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (w >= npairs) return;  // whole wavefronts
  const int c0 = pair_chunk[w] * GAUGE_CHUNK_CELLS, q = pair_slot[w];
  double s = 0;
  for (int i = lane; i < GAUGE_CHUNK_CELLS; i += 64) {
    const int c = c0 + i;
    if (c < n && slot[c] == q) {
      const double v = double(x[c]);
      s += mode == 0 ? v : (mode == 1 ? volume[c / nb_cells] * v : fabs(v));
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) partial[w] = s;
}

// per slot: its pairs' partials in chunk order; divided by the component's
// volume with `divide` (a mean)
__global__ void k_comp_total(int ncomp, const int* __restrict__ slot_pair_ptr, const int* __restrict__ slot_pairs,
                             const double* __restrict__ partial, const double* __restrict__ vol, int divide,
                             double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncomp) return;
  double s = 0;
  for (int p = slot_pair_ptr[q]; p < slot_pair_ptr[q + 1]; ++p) s += partial[slot_pairs[p]];
  out[q] = divide ? s / vol[q] : s;
}

// mode 0: x -= mean (remove volume means); 1: x -= V mean (project a rhs)
template <typename T>
__global__ void k_comp_shift(int n, int nb_cells, const int* __restrict__ slot, const double* __restrict__ mean,
                             const double* __restrict__ volume, int mode, T* __restrict__ x) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int q = slot[c];
  if (!q) return;
  const double m = mean[q - 1];
  x[c] = T(double(x[c]) - (mode ? volume[c / nb_cells] * m : m));
}

// the first sweep from zero: x = (w D^-1) b
template <typename T>
__global__ void k_zero_sweep(int n, double omega, const T* __restrict__ dinv, const T* __restrict__ b,
                             T* __restrict__ x) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  x[c] = (T(omega) * dinv[c]) * b[c];
}

__global__ void k_sqrt(int n, const double* __restrict__ in, double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) out[f] = sqrt(in[f]);
}

// gravity! potential (src/two_phase/forces.jl:545-556): g . x at the cell
// centre origin + (i + 1/2) h, moved by offset * h (the interface centroid of
// a mixed cell relative to its centre, in units of the cell spacing; nullptr:
// none). Products and sums round one at a time, as the reference's do.
__global__ void k_gravity_potential(int n, int bx, int nb_cells, const double* __restrict__ origin,
                                    const double* __restrict__ spacing, const double* __restrict__ offset, double gx,
                                    double gy, double* __restrict__ phi) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int b = c / nb_cells, l = c - b * nb_cells, i = l % bx, j = l / bx;
  const double hx = spacing[2 * b], hy = spacing[2 * b + 1];
  double x = origin[2 * b] + (i + 0.5) * hx;
  double y = origin[2 * b + 1] + (j + 0.5) * hy;
  if (offset) {
    x += offset[2 * c] * hx;
    y += offset[2 * c + 1] * hy;
  }
  phi[c] = gx * x + gy * y;
}

// gravity! jumps (forces.jl:560-576): rho_f |f| (phi+ - phi-) with the
// open-face phase mixture rho_f = rho_l a_l + rho_g a_g (open aperture 1)
__global__ void k_gravity_jump(int nf, const int* __restrict__ fm, const int* __restrict__ fp,
                               const double* __restrict__ area, const double* __restrict__ liquid,
                               const double* __restrict__ gas, double rho_l, double rho_g,
                               const double* __restrict__ phi, double* __restrict__ jumps) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double rho = rho_l * liquid[f] + rho_g * gas[f];
  jumps[f] = rho * area[f] * (phi[fp[f]] - phi[fm[f]]);
}

// CGFM (or CSF) surface-tension jump sigma kappa_f [H] |f|
__global__ void k_surface_jump(int nf, const int* __restrict__ fm, const int* __restrict__ fp,
                               const double* __restrict__ area, const double* __restrict__ c,
                               const double* __restrict__ curv, double sigma, int csf, double* __restrict__ jumps) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double cm = c[fm[f]], cp = c[fp[f]];
  const double jump = csf ? cp - cm : double(cp > 0.5) - double(cm > 0.5);
  const double k = curv[f];
  jumps[f] = (sigma == 0 || jump == 0 || k != k) ? 0.0 : sigma * k * jump * area[f];
}

// divergence: gather of oriented face fluxes in face order, / V
__global__ void k_divergence(int n, int nb_cells, const int* __restrict__ cptr, const int* __restrict__ cface,
                             const signed char* __restrict__ csign, const double* __restrict__ area,
                             const double* __restrict__ u, const int* __restrict__ bptr, const int* __restrict__ bface,
                             const signed char* __restrict__ bside, const double* __restrict__ barea,
                             const double* __restrict__ ub, const double* __restrict__ volume,
                             double* __restrict__ div) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double V = volume[c / nb_cells];
  double s = 0;
  for (int q = cptr[c]; q < cptr[c + 1]; ++q) {
    const int f = cface[q];
    const double flux = u[f] * area[f];
    s += csign[q] < 0 ? flux / V : -flux / V;
  }
  if (ub)
    for (int q = bptr[c]; q < bptr[c + 1]; ++q) {
      const int f = bface[q];
      s += (double(bside[f]) * ub[f] * barea[f]) / V;
    }
  div[c] = s;
}

// adjoint coefficient gradient scale [y_l (K y_p) + y_p (K y_l)] / (2 kappa); g may alias kyl
__global__ void k_adjoint_face(int nf, const double* __restrict__ yl, const double* __restrict__ kyp,
                               const double* __restrict__ yp,
                               const double* kyl, const double* __restrict__ kappa, double scale, double* g) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double v = kappa[f] > 0 ? scale * (yl[f] * kyp[f] + yp[f] * kyl[f]) / (2 * kappa[f]) : 0.0;
  g[f] = v;
}

__global__ void k_cast_d2f(int n, const double* __restrict__ in, float* __restrict__ out) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < size_t(n)) out[c] = float(in[c]);
}

__global__ void k_cast_f2d(int n, const float* __restrict__ in, double* __restrict__ out) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < size_t(n)) out[c] = double(in[c]);
}

__global__ void k_copy(int n, const double* __restrict__ in, double* __restrict__ out) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < size_t(n)) out[c] = in[c];
}

__global__ void k_fill(int n, double v, double* __restrict__ out) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < size_t(n)) out[c] = v;
}

__global__ void k_fill_int(int n, int v, int* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = v;
}

// ---------------------------------------------------------------------------
// launchers

static inline int cdiv(long a, long b) { return int((a + b - 1) / b); }

#define TPAMR_DISPATCH(BXV, CALL)                                  \
  switch (BXV) {                                                   \
    case 4: { constexpr int BX = 4, BY = 4; CALL; } break;         \
    case 8: { constexpr int BX = 8, BY = 8; CALL; } break;         \
    case 16: { constexpr int BX = 16, BY = 16; CALL; } break;      \
    case 32: { constexpr int BX = 32, BY = 32; CALL; } break;      \
    default: break;                                                \
  }

bool supported_block(int bx, int by) { return bx == by && (bx == 4 || bx == 8 || bx == 16 || bx == 32); }

// grid and workgroup of the block-tile kernels
template <int BX, int BY>
static inline dim3 tile_grid(int wg_grid) {
  return dim3(wg_grid);
}
template <int BX, int BY>
static inline dim3 tile_wg() {
  return dim3(Geo<BX, BY>::WG);
}
// dynamic LDS of a tile launch: one slot per block of the workgroup
template <typename T, int BX, int BY>
static inline unsigned tile_lds(const LevelView<T>& L) {
  return unsigned(Geo<BX, BY>::BPW * Slot4<T, BX, BY>::bytes(L.ext_cap));
}

template <typename T, int BX, int BY>
static void launch_sweep_t(hipStream_t s, const LevelView<T>& L, const T* xi, T* xo, const T* b, double w,
                           const int* done, bool fine, const ZStats& zs, const T* x0_dinv, double x0_w) {
  const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
  const unsigned sh = tile_lds<T, BX, BY>(L);
  const double* none = nullptr;
  if (zs.part && zs.r32) {
    hipLaunchKernelGGL((k_sweep<T, BX, BY, true, true, float>), g, t, sh, s, L, xi, xo, b, w, done, zs.r32, zs.part,
                       nullptr, 0.0);
  } else if (zs.part) {
    hipLaunchKernelGGL((k_sweep<T, BX, BY, true, true, double>), g, t, sh, s, L, xi, xo, b, w, done, zs.r64,
                       zs.part, nullptr, 0.0);
  } else if (x0_dinv) {
    if (fine)
      hipLaunchKernelGGL((k_sweep<T, BX, BY, true, false, double, true>), g, t, sh, s, L, xi, xo, b, w, done, none,
                         nullptr, x0_dinv, x0_w);
    else
      hipLaunchKernelGGL((k_sweep<T, BX, BY, false, false, double, true>), g, t, sh, s, L, xi, xo, b, w, done, none,
                         nullptr, x0_dinv, x0_w);
  } else if (fine) {
    hipLaunchKernelGGL((k_sweep<T, BX, BY, true, false>), g, t, sh, s, L, xi, xo, b, w, done, none, nullptr, nullptr,
                       0.0);
  } else {
    hipLaunchKernelGGL((k_sweep<T, BX, BY, false, false>), g, t, sh, s, L, xi, xo, b, w, done, none, nullptr, nullptr,
                       0.0);
  }
}

template <typename T>
void launch_sweep(hipStream_t s, const LevelView<T>& L, const T* xi, T* xo, const T* b, double w, const int* done,
                  bool fine, const ZStats& zs, const T* x0_dinv, double x0_w) {
  TPAMR_DISPATCH(L.bx, (launch_sweep_t<T, BX, BY>(s, L, xi, xo, b, w, done, fine, zs, x0_dinv, x0_w)));
}

template <typename T>
void launch_residual(hipStream_t s, const LevelView<T>& L, const T* x, const T* b, T* u, double wp, const int* done,
                     bool fine) {
  TPAMR_DISPATCH(L.bx, ({
                   const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
                   const unsigned sh = tile_lds<T, BX, BY>(L);
                   if (fine)
                     hipLaunchKernelGGL((k_residual<T, BX, BY, true>), g, t, sh, s, L, x, b, u, wp, done);
                   else
                     hipLaunchKernelGGL((k_residual<T, BX, BY, false>), g, t, sh, s, L, x, b, u, wp, done);
                 }));
}

template <typename T>
void launch_restrict(hipStream_t s, const LevelView<T>& L, const T* u, double wp, T* bc, T* xc, const T* dinv_c,
                     double wc, int zero_start, const int* done, bool fine) {
  TPAMR_DISPATCH(L.bx, ({
                   const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
                   const unsigned sh = tile_lds<T, BX, BY>(L);
                   if (fine)
                     hipLaunchKernelGGL((k_restrict<T, BX, BY, true>), g, t, sh, s, L, u, wp, bc, xc, dinv_c, wc,
                                        zero_start, done);
                   else
                     hipLaunchKernelGGL((k_restrict<T, BX, BY, false>), g, t, sh, s, L, u, wp, bc, xc, dinv_c, wc,
                                        zero_start, done);
                 }));
}

template <typename T>
void launch_prolong(hipStream_t s, const LevelView<T>& L, T* x, const T* xc, double wp, const int* done, bool fine) {
  TPAMR_DISPATCH(L.bx, ({
                   const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
                   const unsigned sh = tile_lds<T, BX, BY>(L);
                   if (fine)
                     hipLaunchKernelGGL((k_prolong<T, BX, BY, true>), g, t, sh, s, L, x, xc, wp, done);
                   else
                     hipLaunchKernelGGL((k_prolong<T, BX, BY, false>), g, t, sh, s, L, x, xc, wp, done);
                 }));
}

template <typename T>
void launch_coarse(hipStream_t s, int n, const double* inv, const T* b, T* x, const int* done) {
  hipLaunchKernelGGL((k_coarse_solve<T>), dim3(cdiv(n, 4)), dim3(256), 0, s, n, inv, b, x, done);
}

// Dynamic LDS a workgroup may use (the device limit less the tail kernel's
// static LDS of the coarse solve).
static int lds_budget() {
  static int budget = -1;
  if (budget < 0) {
    int dev = 0, v = 65536;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    budget = v - 16 * 65 * 8 - 256;
  }
  return budget;
}

template <typename T, int BX, int BY>
static void launch_vtail_t(hipStream_t s, const TailArgs<T>& A, const int* done) {
  int slot = 0;
  for (int l = 0; l + 1 < A.n; ++l) slot = std::max(slot, Slot4<T, BX, BY>::bytes(A.L[l].ext_cap));
  const int per_quarter = Geo<BX, BY>::BPW * slot;
  int nq = 4;
  while (nq > 1 && nq * per_quarter > lds_budget()) nq >>= 1;
  if (nq * per_quarter > lds_budget()) throw std::runtime_error("tail levels need more LDS than a workgroup has");
  const unsigned lds = unsigned(nq * per_quarter);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_vtail<T, BX, BY>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_budget());
    attr = true;
  }
  hipLaunchKernelGGL((k_vtail<T, BX, BY>), dim3(1), dim3(256 * nq), lds, s, A, done);
}

template <typename T>
void launch_vtail(hipStream_t s, const TailArgs<T>& A, const int* done) {
  TPAMR_DISPATCH(A.L[0].bx, (launch_vtail_t<T, BX, BY>(s, A, done)));
}
template void launch_vtail<float>(hipStream_t, const TailArgs<float>&, const int*);
template void launch_vtail<double>(hipStream_t, const TailArgs<double>&, const int*);

template <typename T>
void launch_matvec_dot(hipStream_t s, const LevelView<T>& L, const T* p, T* y, double* part, const int* done,
                       bool sums) {
  TPAMR_DISPATCH(L.bx, ({
                   const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
                   const unsigned sh = tile_lds<T, BX, BY>(L);
                   if (sums)
                     hipLaunchKernelGGL((k_matvec_dot<T, BX, BY, true>), g, t, sh, s, L, p, y, part, done);
                   else
                     hipLaunchKernelGGL((k_matvec_dot<T, BX, BY, false>), g, t, sh, s, L, p, y, part, done);
                 }));
}

void launch_floor_bound(hipStream_t s, const LevelView<double>& L, const double* x, const double* b, double* part) {
  TPAMR_DISPATCH(L.bx, ({
                   const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
                   const unsigned sh = tile_lds<double, BX, BY>(L);
                   hipLaunchKernelGGL((k_floor_bound<BX, BY>), g, t, sh, s, L, x, b, part);
                 }));
}

void launch_reduce(hipStream_t s, int n, int nsets, const double* part, double* out, bool is_max) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, n, nsets, part, out, int(is_max));
}

void launch_cg_alpha(hipStream_t s, int nb, const double* part, double* stage, CgState* st, bool sums,
                     double total_volume) {
  hipLaunchKernelGGL(k_cg_alpha, dim3(RED_WG), dim3(256), 0, s, nb, part, stage, st, int(sums), total_volume);
}

template <typename TV>
void launch_cg_update_xr(hipStream_t s, int n, int nbc, const double* p, const double* Ap, double* x, double* r,
                         const double* vol, TV* vb, TV* vx, const TV* dinv, double omega, double* part, CgState* st,
                         int init) {
  hipLaunchKernelGGL((k_cg_update_xr<TV>), dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, p, Ap, x, r, vol, vb, vx,
                     dinv, omega, part, st, init);
}

void launch_cg_beta(hipStream_t s, int nb, int nmax, const double* part, const double* maxpart, double* stage,
                    double total_volume, CgState* st, int init) {
  hipLaunchKernelGGL(k_cg_beta, dim3(RED_WG), dim3(256), 0, s, nb, nmax, part, maxpart, stage, total_volume, st,
                     init);
}

template <typename TV>
void launch_cg_update_p(hipStream_t s, int n, const TV* z, double* p, const CgState* st, const int* done) {
  hipLaunchKernelGGL((k_cg_update_p<TV>), dim3(cdiv(n, 256)), dim3(256), 0, s, n, z, p, st, done);
}

void launch_residual_strong(hipStream_t s, int n, int nbc, const double* b, const double* y, double* r,
                            const double* vol, double* part) {
  hipLaunchKernelGGL(k_residual_strong, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, b, y, r, vol, part);
}

void launch_weighted_sums(hipStream_t s, int n, int nbc, const double* x, const double* vol, double* part) {
  hipLaunchKernelGGL(k_weighted_sums, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, x, vol, part);
}

void launch_fill_halo(hipStream_t s, int n, int bx, int by, int h, int ne, const int* dest, const int* ptr,
                      const int* donor, const double* weight, const int* phys, const double* cells,
                      const double* physical, double* out) {
  hipLaunchKernelGGL(k_halo_interior, dim3(cdiv(n, 256)), dim3(256), 0, s, n, bx, by, h, cells, out);
  if (ne > 0)
    hipLaunchKernelGGL(k_halo_fill, dim3(cdiv(ne, 256)), dim3(256), 0, s, ne, dest, ptr, donor, weight, phys, cells,
                       physical, out);
}

void launch_rhs_prepare(hipStream_t s, int n, int nbc, double* b, const double* x, const double* vol, double* part) {
  hipLaunchKernelGGL(k_rhs_prepare, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, b, x, vol, part);
}

void launch_shift_residual(hipStream_t s, int n, int nbc, double* b, double mean, const double* vol, double* r,
                           double* part) {
  hipLaunchKernelGGL(k_shift_residual, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, b, mean, vol, r, part);
}

void launch_shift(hipStream_t s, int n, int nbc, double* x, const double* shift, double scale, const double* vol,
                  bool by_volume) {
  hipLaunchKernelGGL(k_shift, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, x, shift, scale, vol, int(by_volume));
}

void launch_scale_by_volume(hipStream_t s, int n, int nbc, const double* in, double* out, const double* vol,
                            bool divide) {
  hipLaunchKernelGGL(k_scale_by_volume, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, in, out, vol, int(divide));
}

void launch_metric_apply(hipStream_t s, int nf, const double* kd, const int* optr, const int* oface,
                         const double* oval, const double* in, double* out) {
  hipLaunchKernelGGL(k_metric_apply, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, kd, optr, oface, oval, in, out);
}

void launch_face_jump(hipStream_t s, int nf, const int* fm, const int* fp, const double* area, const double* sqk,
                      const double* p, const double* jumps, double* out) {
  hipLaunchKernelGGL(k_face_jump, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, fm, fp, area, sqk, p, jumps, out);
}

void launch_face_scale(hipStream_t s, int nf, const double* sqk, const double* in, double* out) {
  hipLaunchKernelGGL(k_face_scale, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, sqk, in, out);
}

void launch_mixture(hipStream_t s, int nf, int nbc, const int* fm, const int* fp, const signed char* axis,
                    const double* spacing, const double* c, double rl, double rg, double* kappa) {
  hipLaunchKernelGGL(k_mixture, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, nbc, fm, fp, axis, spacing, c, rl, rg,
                     kappa);
}

void launch_surface_jump_cells(hipStream_t s, int nf, const int* fm, const int* fp, const double* area,
                               const double* origin, const double* spacing, const double* fcx, const double* fcy,
                               int nblocks, int bx, int by, const double extent[2], const bool periodic[2],
                               const double* c, const double* kappa, const double* quality, double sigma, int csf,
                               double* jumps, int* unresolved) {
  CurvMesh M{origin, spacing, fcx, fcy, nblocks, bx, by, {extent[0], extent[1]}, {periodic[0], periodic[1]}};
  hipLaunchKernelGGL(k_surface_jump_cells, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, fm, fp, area, M, c, kappa,
                     quality, sigma, csf, jumps, unresolved);
}

void launch_ref_restrict(hipStream_t s, int nf, const int* rptr, const int* ridx, const double* rw,
                         const double* kprev, double* k) {
  if (nf > 0) hipLaunchKernelGGL(k_ref_restrict, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, rptr, ridx, rw, kprev, k);
}
void launch_ref_regular(hipStream_t s, int nf, const double* k, const double* kdiag, const double* area,
                        const int* dst, double* cw, double* ce, double* cs, double* cn) {
  if (nf > 0)
    hipLaunchKernelGGL(k_ref_regular, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, k, kdiag, area, dst, cw, ce, cs, cn);
}
void launch_ref_boundary(hipStream_t s, int nopen, const double* bw, const int* dst, double* cw, double* ce,
                         double* cs, double* cn) {
  if (nopen > 0)
    hipLaunchKernelGGL(k_ref_boundary, dim3(cdiv(nopen, 256)), dim3(256), 0, s, nopen, bw, dst, cw, ce, cs, cn);
}
void launch_ref_rows(hipStream_t s, int nrows, const int* irr_ptr, const int* term_ptr, const int* term_f,
                     const int* term_g, const double* term_k, const double* term_s, const unsigned char* term_diag,
                     const double* k, const double* area, const int* dst, const int* ell_pos, const int* ell_dpos,
                     const int* ell_row, const int* diag_e, double* ell_val, double* ell_bound, double* ell_bdiag,
                     double* ell_diag, double* irr_diag, double* val_e, int nnz, unsigned long long* ratio,
                     const double* bd) {
  if (nnz > 0)
    hipLaunchKernelGGL(k_ref_entries, dim3(cdiv(nnz, 256)), dim3(256), 0, s, nnz, term_ptr, term_f, term_g, term_k,
                       term_s, term_diag, k, area, dst, ell_pos, ell_dpos, ell_val, ell_bound, ell_bdiag, val_e);
  if (nrows > 0)
    hipLaunchKernelGGL(k_ref_rows, dim3(cdiv(nrows, 256)), dim3(256), 0, s, nrows, irr_ptr, diag_e, ell_row, val_e,
                       ell_diag, irr_diag, ratio, bd);
}
void launch_ref_dinv(hipStream_t s, int n, int bx, int by, const double* cw, const double* ce, const double* cs,
                     const double* cn, const int* cell_irr, const double* irr_diag, double* dinv, int* regular,
                     const int* nbr, unsigned long long* ratio) {
  hipLaunchKernelGGL(k_ref_dinv, dim3(cdiv(n, 256)), dim3(256), 0, s, n, bx, by, cw, ce, cs, cn, cell_irr, irr_diag,
                     dinv, regular, nbr, ratio);
}
template <typename T>
void launch_comp_sums(hipStream_t s, const GaugeView& G, int n, int nb_cells, const T* x, const double* volume,
                      int mode, int divide, double* out) {
  if (G.ncomp == 0) return;
  if (G.npairs > 0)
    hipLaunchKernelGGL(k_comp_partial<T>, dim3(cdiv(G.npairs, 4)), dim3(256), 0, s, G.npairs, n, nb_cells,
                       G.pair_chunk, G.pair_slot, G.slot, x, volume, mode, G.partial);
  hipLaunchKernelGGL(k_comp_total, dim3(cdiv(G.ncomp, 256)), dim3(256), 0, s, G.ncomp, G.slot_pair_ptr,
                     G.slot_pairs, G.partial, G.vol, divide, out);
}
template <typename T>
void launch_comp_shift(hipStream_t s, const GaugeView& G, int n, int nb_cells, const double* mean,
                       const double* volume, int mode, T* x) {
  if (G.ncomp == 0) return;
  hipLaunchKernelGGL(k_comp_shift<T>, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nb_cells, G.slot, mean, volume, mode,
                     x);
}
template <typename T>
void launch_zero_sweep(hipStream_t s, int n, double omega, const T* dinv, const T* b, T* x) {
  hipLaunchKernelGGL(k_zero_sweep<T>, dim3(cdiv(n, 256)), dim3(256), 0, s, n, omega, dinv, b, x);
}
template void launch_comp_sums<double>(hipStream_t, const GaugeView&, int, int, const double*, const double*, int,
                                       int, double*);
template void launch_comp_sums<float>(hipStream_t, const GaugeView&, int, int, const float*, const double*, int,
                                      int, double*);
template void launch_comp_shift<double>(hipStream_t, const GaugeView&, int, int, const double*, const double*, int,
                                        double*);
template void launch_comp_shift<float>(hipStream_t, const GaugeView&, int, int, const double*, const double*, int,
                                       float*);
template void launch_zero_sweep<double>(hipStream_t, int, double, const double*, const double*, double*);
template void launch_zero_sweep<float>(hipStream_t, int, double, const float*, const float*, float*);

void launch_ref_check(hipStream_t s, int nf, const double* k, int* bad) {
  if (nf > 0) hipLaunchKernelGGL(k_ref_check, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, k, bad);
}
void launch_sqrt(hipStream_t s, int n, const double* in, double* out) {
  if (n > 0) hipLaunchKernelGGL(k_sqrt, dim3(cdiv(n, 256)), dim3(256), 0, s, n, in, out);
}

// max |I - T|_ij of an N x N matrix as nonnegative-double bits (non-finite
// entries count as infinity)
__global__ void k_identity_residual(int N, const double* __restrict__ T, unsigned long long* __restrict__ out) {
  const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t NN = size_t(N) * N;
  double m = 0;
  for (size_t e = q; e < NN; e += size_t(gridDim.x) * blockDim.x) {
    const size_t i = e / N, j = e % N;
    double v = fabs((i == j ? 1.0 : 0.0) - T[e]);
    if (!(v <= 1.7976931348623157e308)) v = __builtin_inf();
    m = fmax(m, v);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

void launch_identity_residual(hipStream_t s, int N, const double* T, unsigned long long* out) {
  hipLaunchKernelGGL(k_identity_residual, dim3(1024), dim3(256), 0, s, N, T, out);
}

void launch_extract_block(hipStream_t s, int n, int N, const double* a, double* inv) {
  hipLaunchKernelGGL(k_gj_extract, dim3(cdiv(n * n, 256)), dim3(256), 0, s, n, N, a, inv);
}

void launch_gjb_panel(hipStream_t s, int N, int c0, int nb, double* a, int* piv, double* col, int* fail) {
  for (int k = c0; k < c0 + nb; ++k) {
    hipLaunchKernelGGL(k_gjb_pivot, dim3(1), dim3(1024), 0, s, N, k, c0, nb, a, piv, col, fail);
    hipLaunchKernelGGL(k_gjb_update, dim3(N), dim3(64), 0, s, N, k, c0, nb, a, col);
  }
  hipLaunchKernelGGL(k_gjb_swap_rest, dim3(cdiv(N, 256)), dim3(256), 0, s, N, c0, nb, a, piv);
}

void launch_gjb_prep(hipStream_t s, int N, int c0, int nb, const double* a, double* R, double* M) {
  const size_t q = size_t(N) * nb;
  hipLaunchKernelGGL(k_gjb_prep, dim3(unsigned((q + 255) / 256)), dim3(256), 0, s, N, c0, nb, a, R, M);
}

void launch_gj_finish(hipStream_t s, int N, double* a, const int* piv, int n, double* inv) {
  hipLaunchKernelGGL(k_gj_unpermute, dim3(cdiv(N, 256)), dim3(256), 0, s, N, a, piv);
  hipLaunchKernelGGL(k_gj_extract, dim3(unsigned((size_t(n) * n + 255) / 256)), dim3(256), 0, s, n, N, a, inv);
}

void launch_dense_inverse(hipStream_t s, int N, double* a, int* piv, double* col, int* fail, int n, double* inv) {
  for (int k = 0; k < N; ++k) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, s, N, k, a, piv, col, fail);
    hipLaunchKernelGGL(k_gj_update, dim3(cdiv(N, 256), N), dim3(256), 0, s, N, k, a, col);
  }
  hipLaunchKernelGGL(k_gj_unpermute, dim3(cdiv(N, 256)), dim3(256), 0, s, N, a, piv);
  hipLaunchKernelGGL(k_gj_extract, dim3(cdiv(n * n, 256)), dim3(256), 0, s, n, N, a, inv);
}

void launch_gravity_jumps(hipStream_t s, int n, int bx, int nbc, const double* origin, const double* spacing,
                          const double* offset, double gx, double gy, double* phi, int nf, const int* fm,
                          const int* fp, const double* area, const double* liquid, const double* gas, double rl,
                          double rg, double* jumps) {
  hipLaunchKernelGGL(k_gravity_potential, dim3(cdiv(n, 256)), dim3(256), 0, s, n, bx, nbc, origin, spacing, offset,
                     gx, gy, phi);
  hipLaunchKernelGGL(k_gravity_jump, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, fm, fp, area, liquid, gas, rl, rg, phi,
                     jumps);
}

void launch_surface_jump(hipStream_t s, int nf, const int* fm, const int* fp, const double* area, const double* c,
                         const double* curv, double sigma, int csf, double* jumps) {
  hipLaunchKernelGGL(k_surface_jump, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, fm, fp, area, c, curv, sigma, csf,
                     jumps);
}

void launch_divergence(hipStream_t s, int n, int nbc, const int* cptr, const int* cface, const signed char* csign,
                       const double* area, const double* u, const int* bptr, const int* bface,
                       const signed char* bside, const double* barea, const double* ub, const double* vol,
                       double* div) {
  hipLaunchKernelGGL(k_divergence, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nbc, cptr, cface, csign, area, u, bptr,
                     bface, bside, barea, ub, vol, div);
}

void launch_adjoint_face(hipStream_t s, int nf, const double* yl, const double* kyp, const double* yp,
                         const double* kyl, const double* kappa, double scale, double* g) {
  hipLaunchKernelGGL(k_adjoint_face, dim3(cdiv(nf, 256)), dim3(256), 0, s, nf, yl, kyp, yp, kyl, kappa, scale, g);
}

template <typename T>
void launch_zstats(hipStream_t s, int nblocks, int bx, const T* z, const ZStats& zs, const double* vol,
                   const int* done) {
  TPAMR_DISPATCH(bx, ({
                   if (zs.r32)
                     hipLaunchKernelGGL((k_zstats<T, BX, BY, float>), dim3(nblocks), dim3(BX * BY), 0, s, z, zs.r32,
                                        vol, zs.part, done);
                   else
                     hipLaunchKernelGGL((k_zstats<T, BX, BY, double>), dim3(nblocks), dim3(BX * BY), 0, s, z, zs.r64,
                                        vol, zs.part, done);
                 }));
}
template void launch_zstats<float>(hipStream_t, int, int, const float*, const ZStats&, const double*, const int*);
template void launch_zstats<double>(hipStream_t, int, int, const double*, const ZStats&, const double*, const int*);

int ir_groups(int n) { return cdiv(n, 1024); }

void launch_ir_update_xr(hipStream_t s, int n, int nbc, const float* p, const float* Ap, float* e, float* r,
                         const double* vol, float* vx, const float* dinv, double omega, const double* r64,
                         double* part, CgState* st, int init, double* x64) {
  hipLaunchKernelGGL(k_ir_update_xr, dim3(ir_groups(n)), dim3(256), 0, s, n, nbc, p, Ap, e, r, vol, vx, dinv,
                     float(omega), r64, part, st, init, x64);
}

void launch_ir_replace(hipStream_t s, const LevelView<double>& L, const double* x, const double* b, float* r32,
                       float* x0, const float* dinv, double omega, double* part, CgState* st) {
  TPAMR_DISPATCH(L.bx, ({
                   const dim3 g = tile_grid<BX, BY>(L.wg_grid), t = tile_wg<BX, BY>();
                   const unsigned sh = tile_lds<double, BX, BY>(L);
                   hipLaunchKernelGGL((k_ir_replace<BX, BY>), g, t, sh, s, L, x, b, r32, x0, dinv, float(omega),
                                      part, st);
                 }));
}

void launch_ir_check(hipStream_t s, int ng, const double* maxpart, double* stage, CgState* st) {
  hipLaunchKernelGGL(k_ir_check, dim3(RED_WG), dim3(256), 0, s, ng, maxpart, stage, st);
}

void launch_ir_update_p(hipStream_t s, int n, const float* z, float* p, const CgState* st, const int* done) {
  hipLaunchKernelGGL(k_ir_update_p, dim3(ir_groups(n)), dim3(256), 0, s, n, z, p, st, done);
}

void launch_ir_correct(hipStream_t s, int n, float* e, double* x, const CgState* st, bool force) {
  hipLaunchKernelGGL(k_ir_correct, dim3(ir_groups(n)), dim3(256), 0, s, n, e, x, st, int(force));
}

void launch_fill_f32(hipStream_t s, int n, float v, float* out) {
  hipLaunchKernelGGL(k_fill_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, n, v, out);
}

void launch_projection_rhs(hipStream_t s, int n, const double* div, const double* inherited, double dt, double w,
                           double* rhs) {
  hipLaunchKernelGGL(k_projection_rhs, dim3(cdiv(n, 256)), dim3(256), 0, s, n, div, inherited, dt, w, rhs);
}

void launch_boundary_velocity(hipStream_t s, int nbf, const int* bcell, const signed char* bside, const double* bgrad,
                              const double* phi, double dt, double* ub) {
  if (nbf > 0)
    hipLaunchKernelGGL(k_boundary_velocity, dim3(cdiv(nbf, 256)), dim3(256), 0, s, nbf, bcell, bside, bgrad, phi, dt,
                       ub);
}
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y) {
  hipLaunchKernelGGL(k_axpby, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, x, b, y);
}

void launch_cast_d2f(hipStream_t s, int n, const double* in, float* out) {
  hipLaunchKernelGGL(k_cast_d2f, dim3(cdiv(n, 256)), dim3(256), 0, s, n, in, out);
}
void launch_cast_f2d(hipStream_t s, int n, const float* in, double* out) {
  hipLaunchKernelGGL(k_cast_f2d, dim3(cdiv(n, 256)), dim3(256), 0, s, n, in, out);
}
void launch_copy(hipStream_t s, int n, const double* in, double* out) {
  hipLaunchKernelGGL(k_copy, dim3(cdiv(n, 256)), dim3(256), 0, s, n, in, out);
}
void launch_fill(hipStream_t s, int n, double v, double* out) {
  hipLaunchKernelGGL(k_fill, dim3(cdiv(n, 256)), dim3(256), 0, s, n, v, out);
}
void launch_fill_int(hipStream_t s, int n, int v, int* out) {
  hipLaunchKernelGGL(k_fill_int, dim3(cdiv(n, 256)), dim3(256), 0, s, n, v, out);
}

template void launch_sweep<float>(hipStream_t, const LevelView<float>&, const float*, float*, const float*, double,
                                  const int*, bool, const ZStats&, const float*, double);
template void launch_sweep<double>(hipStream_t, const LevelView<double>&, const double*, double*, const double*,
                                   double, const int*, bool, const ZStats&, const double*, double);
template void launch_matvec_dot<float>(hipStream_t, const LevelView<float>&, const float*, float*, double*,
                                       const int*, bool);
template void launch_matvec_dot<double>(hipStream_t, const LevelView<double>&, const double*, double*, double*,
                                        const int*, bool);
template void launch_residual<float>(hipStream_t, const LevelView<float>&, const float*, const float*, float*, double,
                                     const int*, bool);
template void launch_residual<double>(hipStream_t, const LevelView<double>&, const double*, const double*, double*,
                                      double, const int*, bool);
template void launch_restrict<float>(hipStream_t, const LevelView<float>&, const float*, double, float*, float*,
                                     const float*, double, int, const int*, bool);
template void launch_restrict<double>(hipStream_t, const LevelView<double>&, const double*, double, double*, double*,
                                      const double*, double, int, const int*, bool);
template void launch_prolong<float>(hipStream_t, const LevelView<float>&, float*, const float*, double, const int*,
                                    bool);
template void launch_prolong<double>(hipStream_t, const LevelView<double>&, double*, const double*, double,
                                     const int*, bool);
template void launch_coarse<float>(hipStream_t, int, const double*, const float*, float*, const int*);
template void launch_coarse<double>(hipStream_t, int, const double*, const double*, double*, const int*);
template void launch_cg_update_xr<float>(hipStream_t, int, int, const double*, const double*, double*, double*,
                                         const double*, float*, float*, const float*, double, double*, CgState*, int);
template void launch_cg_update_xr<double>(hipStream_t, int, int, const double*, const double*, double*, double*,
                                          const double*, double*, double*, const double*, double, double*, CgState*,
                                          int);
template void launch_cg_update_p<float>(hipStream_t, int, const float*, double*, const CgState*, const int*);
template void launch_cg_update_p<double>(hipStream_t, int, const double*, double*, const CgState*, const int*);

// Streaming pass for the HBM ceiling of a given access mix: out = sum of R
// input arrays, one float4 of each per thread (R = 1: a plain copy; R = 5:
// the Jacobi sweep's five reads per write).
template <int R>
__global__ void __launch_bounds__(256) k_stream(size_t n4, size_t stride4, const float4* __restrict__ in,
                                                float4* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n4) return;
  float4 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = in[size_t(r) * stride4 + i];
  float4 s = v[0];
#pragma unroll
  for (int r = 1; r < R; ++r) {
    s.x += v[r].x;
    s.y += v[r].y;
    s.z += v[r].z;
    s.w += v[r].w;
  }
  out[i] = s;
}

void launch_stream(hipStream_t s, int reads, size_t n4, const float* in, float* out) {
  const dim3 g(unsigned((n4 + 255) / 256)), t(256);
  const float4* a = reinterpret_cast<const float4*>(in);
  float4* o = reinterpret_cast<float4*>(out);
  if (reads == 1)
    hipLaunchKernelGGL(k_stream<1>, g, t, 0, s, n4, n4, a, o);
  else if (reads == 5)
    hipLaunchKernelGGL(k_stream<5>, g, t, 0, s, n4, n4, a, o);
  else
    throw std::invalid_argument("stream pass: 1 or 5 input arrays");
}

// Scratch of tpamr_solve_device_noalloc: the Krylov state and the reduction
// partials of one solve, carved at 256-byte offsets. Each GPU holds its own
// copy; host code reaches it only through the address that
// hipGetSymbolAddress returns for that device (the name itself is a
// host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch_address(int device, void** ptr) {
  static std::mutex mu;
  static std::vector<void*> addr;
  if (device < 0 || !ptr) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (size_t(device) >= addr.size()) addr.resize(size_t(device) + 1, nullptr);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

}  // namespace tpamr
