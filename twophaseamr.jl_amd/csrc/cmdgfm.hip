// Device kernels of the CMDGFM reduced-pressure/jump block solve: sliced-ELL
// products of the assembled system, BiCGSTAB vector updates and scalars, the
// block preconditioner's scatter / anchor / gather passes, the level-scheduled
// triangular solves of the jump factor, and the defect reductions.
//
// Every reduction is per-workgroup partials (grid-stride, at most CM_MAX_WG
// workgroups) folded by one workgroup in a fixed order, so results do not
// depend on timing. Contraction is off: each product and sum rounds as
// written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "cmdgfm.hpp"

namespace tpamr {
namespace {

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// maximum that keeps NaN (Julia's max)
__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return fmax(a, b);
}

// Workgroup partials of NS sets, one value per thread: sums, or maxima for
// set k when bit k of max_mask is set; part[blockIdx.x * NS + k].
template <int NS>
__device__ __forceinline__ void wg_partials(double (&v)[NS], int max_mask, double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[NS][CM_WG / 64];
  for (int s = 0; s < NS; ++s) {
    const bool mx = (max_mask >> s) & 1;
    for (int o = 32; o > 0; o >>= 1) {
      const double other = __shfl_xor(v[s], o, 64);
      v[s] = mx ? nanmax(v[s], other) : v[s] + other;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int s = 0; s < NS; ++s) red[s][w] = v[s];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int s = 0; s < NS; ++s) {
      const bool mx = (max_mask >> s) & 1;
      double acc = red[s][0];
      for (int k = 1; k < CM_WG / 64; ++k) acc = mx ? nanmax(acc, red[s][k]) : acc + red[s][k];
      part[blockIdx.x * NS + s] = acc;
    }
  }
}

__device__ __forceinline__ double row_dot(const int* __restrict__ ptr, const int* __restrict__ col,
                                          const double* __restrict__ val, const double* __restrict__ x, int i) {
#pragma clang fp contract(off)
  double acc = 0.0;
  for (int p = ptr[i]; p < ptr[i + 1]; ++p) acc = acc + val[p] * x[col[p]];
  return acc;
}

// Sliced ELL rows (CM_SLICE rows per slice, entries of a slice column-major,
// a row's entries in CSR order): y_i = rs_i sum_j a_ij (cs_j x_j) with the
// per-block scalings (pressure rows / columns before np, jump after).
__global__ void __launch_bounds__(CM_WG) k_cm_spmv(int n, int np, const int* __restrict__ sptr,
                                                   const int* __restrict__ len, const int* __restrict__ col,
                                                   const double* __restrict__ val, const double* __restrict__ x,
                                                   double pc, double jc, double pr, double jr,
                                                   double* __restrict__ y) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int base = sptr[i / CM_SLICE] + (i % CM_SLICE);
  const int m = len[i];
  double acc = 0.0;
  for (int e = 0; e < m; ++e) {
    const int q = base + e * CM_SLICE;
    const int j = col[q];
    acc = acc + val[q] * ((j < np ? pc : jc) * x[j]);
  }
  y[i] = (i < np ? pr : jr) * acc;
}

// r = b - A x over the sliced ELL rows
__global__ void __launch_bounds__(CM_WG) k_cm_residual(int n, const int* __restrict__ sptr,
                                                       const int* __restrict__ len, const int* __restrict__ col,
                                                       const double* __restrict__ val, const double* __restrict__ x,
                                                       const double* __restrict__ b, double* __restrict__ r) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int base = sptr[i / CM_SLICE] + (i % CM_SLICE);
  const int m = len[i];
  double acc = 0.0;
  for (int e = 0; e < m; ++e) {
    const int q = base + e * CM_SLICE;
    acc = acc + val[q] * x[col[q]];
  }
  r[i] = b[i] - acc;
}

template <int ND>
__global__ void __launch_bounds__(CM_WG) k_cm_dots(int n, const double* __restrict__ a0, const double* __restrict__ b0,
                                                   const double* __restrict__ a1, const double* __restrict__ b1,
                                                   double* __restrict__ part) {
#pragma clang fp contract(off)
  double v[ND];
  for (int k = 0; k < ND; ++k) v[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    v[0] = v[0] + a0[i] * b0[i];
    if constexpr (ND > 1) v[1] = v[1] + a1[i] * b1[i];
  }
  wg_partials<ND>(v, 0, part);
}

// Fixed-order fold of nsets partial sets: thread-strided sums, then a tree.
__device__ __forceinline__ void fold(int nparts, int nsets, int max_mask, const double* __restrict__ part,
                                     double (&red)[8][CM_WG]) {
#pragma clang fp contract(off)
  for (int s = 0; s < nsets; ++s) {
    const bool mx = (max_mask >> s) & 1;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += CM_WG) {
      const double v = part[i * nsets + s];
      acc = mx ? nanmax(acc, v) : acc + v;
    }
    red[s][threadIdx.x] = acc;
  }
  __syncthreads();
  for (int w = CM_WG / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int s = 0; s < nsets; ++s) {
        const bool mx = (max_mask >> s) & 1;
        red[s][threadIdx.x] = mx ? nanmax(red[s][threadIdx.x], red[s][threadIdx.x + w])
                                 : red[s][threadIdx.x] + red[s][threadIdx.x + w];
      }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(CM_WG) k_cm_scalars(int nparts, int nd, const double* __restrict__ part,
                                                      double* __restrict__ scal, int step) {
#pragma clang fp contract(off)
  __shared__ double red[8][CM_WG];
  fold(nparts, nd, 0, part, red);
  if (threadIdx.x != 0) return;
  const double s0 = red[0][0], s1 = nd > 1 ? red[1][0] : 0.0;
  if (step == CM_STEP_INIT) {
    scal[CM_RNORM2] = s0;
    scal[CM_NEXT_RHO] = s1;
    scal[CM_RHO] = 1.0;
    scal[CM_ALPHA] = 1.0;
    scal[CM_OMEGA] = 1.0;
    scal[CM_BETA] = 0.0;
  } else if (step == CM_STEP_ALPHA) {
    scal[CM_RHO] = scal[CM_NEXT_RHO];
    scal[CM_ALPHA] = scal[CM_RHO] / s0;
  } else if (step == CM_STEP_OMEGA) {
    scal[CM_OMEGA] = s0 / s1;
  } else {
    scal[CM_NEXT_RHO] = s0;
    scal[CM_BETA] = (s0 / scal[CM_RHO]) * (scal[CM_ALPHA] / scal[CM_OMEGA]);
    scal[CM_RNORM2] = s1;
  }
}

__global__ void __launch_bounds__(CM_WG) k_cm_update_s(int n, const double* __restrict__ r,
                                                       const double* __restrict__ v, const double* __restrict__ y,
                                                       double* __restrict__ sv, double* __restrict__ x,
                                                       const double* __restrict__ scal) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = scal[CM_ALPHA];
  sv[i] = r[i] - alpha * v[i];
  x[i] = x[i] + alpha * y[i];
}

__global__ void __launch_bounds__(CM_WG) k_cm_update_r(int n, const double* __restrict__ sv,
                                                       const double* __restrict__ t, const double* __restrict__ z,
                                                       double* __restrict__ r, double* __restrict__ x,
                                                       const double* __restrict__ scal) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double omega = scal[CM_OMEGA];
  x[i] = x[i] + omega * z[i];
  r[i] = sv[i] - omega * t[i];
}

__global__ void __launch_bounds__(CM_WG) k_cm_update_p(int n, const double* __restrict__ r,
                                                       const double* __restrict__ v, double* __restrict__ p,
                                                       const double* __restrict__ scal) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double omega = scal[CM_OMEGA], beta = scal[CM_BETA];
  const double q = p[i] - omega * v[i];
  p[i] = r[i] + beta * q;
}

// out = (rows) v: pressure entries times pr, jump entries times jr
__global__ void __launch_bounds__(CM_WG) k_cm_scale_rows(int n, int np, const double* __restrict__ v, double pr,
                                                         double jr, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (i < np ? pr : jr) * v[i];
}

// v = v / s
__global__ void __launch_bounds__(CM_WG) k_cm_divs(int n, double* v, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = v[i] / s;
}

// v / d: the preconditioner's input divided by its (pressure) row scaling
__global__ void __launch_bounds__(CM_WG) k_cm_scatter(int ncells, const int* __restrict__ role,
                                                      const double* __restrict__ v, double d,
                                                      double* __restrict__ full) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int k = role[c];
  full[c] = k >= 0 ? 0.0 + v[k] / d : 0.0;
}

__global__ void __launch_bounds__(CM_WG) k_cm_anchor_chunks(const int* __restrict__ chunk_ptr,
                                                            const int* __restrict__ list,
                                                            const double* __restrict__ v, double d,
                                                            double* __restrict__ part) {
#pragma clang fp contract(off)
  double acc[1] = {0.0};
  const int b = chunk_ptr[blockIdx.x], e = chunk_ptr[blockIdx.x + 1];
  for (int k = b + int(threadIdx.x); k < e; k += CM_WG) acc[0] = acc[0] + v[list[k]] / d;
  wg_partials<1>(acc, 0, part);
}

// one workgroup per anchor: its chunks' sums folded in a fixed order and
// subtracted from the anchor's (zero) entry
__global__ void __launch_bounds__(CM_WG) k_cm_anchor_totals(const int* __restrict__ anchor_cell,
                                                            const int* __restrict__ anchor_chunk_ptr,
                                                            const double* __restrict__ part,
                                                            double* __restrict__ full) {
#pragma clang fp contract(off)
  __shared__ double red[8][CM_WG];
  const int a = blockIdx.x;
  const int b = anchor_chunk_ptr[a];
  fold(anchor_chunk_ptr[a + 1] - b, 1, 0, part + b, red);
  if (threadIdx.x == 0) full[anchor_cell[a]] = full[anchor_cell[a]] - red[0][0];
}

__global__ void __launch_bounds__(CM_WG) k_cm_gather(int np, const int* __restrict__ idx,
                                                     const int* __restrict__ anchor, const double* __restrict__ z,
                                                     double d, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = (z[idx[i]] - z[anchor[i]]) / d;
}

// jr = -(C zp) + rhs / d; zp either a reduced vector or, with zfull, the
// reduced-pressure block's answer zfull[idx] - zfull[anchor]
__global__ void __launch_bounds__(CM_WG) k_cm_jump_rhs(int nx, const int* __restrict__ ptr,
                                                       const int* __restrict__ col, const double* __restrict__ val,
                                                       const double* __restrict__ zp,
                                                       const double* __restrict__ zfull,
                                                       const int* __restrict__ idx, const int* __restrict__ anchor,
                                                       const double* __restrict__ rhs, double d,
                                                       double* __restrict__ jr) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx) return;
  double cz = 0.0;
  if (ptr) {
    for (int p = ptr[k]; p < ptr[k + 1]; ++p) {
      const int q = col[p];
      const double z = zfull ? zfull[idx[q]] - zfull[anchor[q]] : zp[q];
      cz = cz + val[p] * z;
    }
  }
  jr[k] = -1.0 * cz + 1.0 * (rhs[k] / d);
}

// One workgroup: rows of one level are independent; a barrier separates
// levels, so every row reads finished values of the rows it depends on.
__global__ void __launch_bounds__(CM_ILU_WG) k_cm_ilu(int nx, const int* __restrict__ lptr,
                                                      const int* __restrict__ lcol, const double* __restrict__ lval,
                                                      int nllev, const int* __restrict__ llev,
                                                      const int* __restrict__ lorder, const int* __restrict__ uptr,
                                                      const int* __restrict__ ucol, const double* __restrict__ uval,
                                                      const double* __restrict__ udiag, int nulev,
                                                      const int* __restrict__ ulev, const int* __restrict__ uorder,
                                                      const double* __restrict__ b, double* y, double* x,
                                                      double d) {
#pragma clang fp contract(off)
  for (int l = 0; l < nllev; ++l) {
    for (int t = llev[l] + int(threadIdx.x); t < llev[l + 1]; t += CM_ILU_WG) {
      const int i = lorder[t];
      double acc = b[i];
      for (int p = lptr[i]; p < lptr[i + 1]; ++p) acc = acc - lval[p] * y[lcol[p]];
      y[i] = acc;
    }
    __syncthreads();
  }
  for (int l = 0; l < nulev; ++l) {
    for (int t = ulev[l] + int(threadIdx.x); t < ulev[l + 1]; t += CM_ILU_WG) {
      const int i = uorder[t];
      double acc = y[i];
      for (int p = uptr[i]; p < uptr[i + 1]; ++p) acc = acc - uval[p] * x[ucol[p]];
      x[i] = acc / udiag[i];
    }
    __syncthreads();
  }
  if (d != 1.0)
    for (int i = threadIdx.x; i < nx; i += CM_ILU_WG) x[i] = x[i] / d;
}

// The same solves with the iterate and the level schedules in LDS and the
// rows packed in level order at a fixed width W (entry k of packed row t at
// t W + k, len[t] of them, in the global kernel's order): every load of a
// row is independent of the others, and each thread fetches its row of the
// next level into registers before the barrier that ends the current one.
struct IluRow {
  int row, m;
  double rhs, diag;
  int col[CM_ILU_PF];
  double val[CM_ILU_PF];
};

__device__ __forceinline__ void ilu_fetch(int t, int w, const int* __restrict__ rows, const int* __restrict__ len,
                                          const int* __restrict__ col, const double* __restrict__ val,
                                          const double* __restrict__ rhs, const double* __restrict__ diag,
                                          IluRow& f) {
  f.row = rows[t];
  f.m = len[t];
  f.rhs = rhs ? rhs[t] : 0.0;
  f.diag = diag ? diag[t] : 1.0;
#pragma unroll
  for (int k = 0; k < CM_ILU_PF; ++k)
    if (k < w) {
      f.col[k] = col[size_t(t) * w + k];
      f.val[k] = val[size_t(t) * w + k];
    }
}

// one triangular solve over the level schedule lev (nlev levels, in LDS) of
// the packed rows; rhs_t: rhs by packed position (null: the LDS value of the
// row itself); diag: divisor by packed position (null: unit)
__device__ __forceinline__ void ilu_sweep(double* ys, int nlev, const int* lev, int w, const int* __restrict__ rows,
                                          const int* __restrict__ len, const int* __restrict__ col,
                                          const double* __restrict__ val, const double* __restrict__ rhs_t,
                                          const double* __restrict__ diag) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  IluRow cur, nxt;
  bool have = nlev > 0 && lev[0] + tid < lev[1];
  if (have) ilu_fetch(lev[0] + tid, w, rows, len, col, val, rhs_t, diag, cur);
  for (int l = 0; l < nlev; ++l) {
    bool nhave = false;
    if (l + 1 < nlev && lev[l + 1] + tid < lev[l + 2]) {
      ilu_fetch(lev[l + 1] + tid, w, rows, len, col, val, rhs_t, diag, nxt);
      nhave = true;
    }
    if (have) {
      double acc = rhs_t ? cur.rhs : ys[cur.row];
#pragma unroll
      for (int k = 0; k < CM_ILU_PF; ++k)
        if (k < cur.m) acc = acc - cur.val[k] * ys[cur.col[k]];
      ys[cur.row] = diag ? acc / cur.diag : acc;
    }
    // rows beyond one per thread in a wide level, straight from memory
    for (int t = lev[l] + tid + CM_ILU_WG; t < lev[l + 1]; t += CM_ILU_WG) {
      const int i = rows[t];
      double acc = rhs_t ? rhs_t[t] : ys[i];
      for (int k = 0; k < len[t]; ++k) acc = acc - val[size_t(t) * w + k] * ys[col[size_t(t) * w + k]];
      ys[i] = diag ? acc / diag[t] : acc;
    }
    __syncthreads();
    if (nhave) cur = nxt;
    have = nhave;
  }
}

__global__ void __launch_bounds__(CM_ILU_WG) k_cm_ilu_lds(int nx, int w, int nllev, const int* __restrict__ llev,
                                                          const int* __restrict__ lrows, const int* __restrict__ llen,
                                                          const int* __restrict__ lcol,
                                                          const double* __restrict__ lval, int nulev,
                                                          const int* __restrict__ ulev, const int* __restrict__ urows,
                                                          const int* __restrict__ ulen, const int* __restrict__ ucol,
                                                          const double* __restrict__ uval,
                                                          const double* __restrict__ udiag,
                                                          const double* __restrict__ b_t, double* __restrict__ x,
                                                          double d) {
#pragma clang fp contract(off)
  // LDS: the iterate (nx doubles), then both level schedules
  extern __shared__ double ys[];
  int* ll = reinterpret_cast<int*>(ys + nx);
  int* ul = ll + nllev + 1;
  for (int k = threadIdx.x; k <= nllev; k += CM_ILU_WG) ll[k] = llev[k];
  for (int k = threadIdx.x; k <= nulev; k += CM_ILU_WG) ul[k] = ulev[k];
  __syncthreads();
  ilu_sweep(ys, nllev, ll, w, lrows, llen, lcol, lval, b_t, nullptr);
  ilu_sweep(ys, nulev, ul, w, urows, ulen, ucol, uval, nullptr, udiag);
  for (int i = threadIdx.x; i < nx; i += CM_ILU_WG) x[i] = d != 1.0 ? ys[i] / d : ys[i];
}

// rhs in the forward solve's packed order: b_t[t] = b[lrows[t]]
__global__ void __launch_bounds__(CM_WG) k_cm_ilu_rhs(int nx, const int* __restrict__ lrows,
                                                      const double* __restrict__ b, double* __restrict__ b_t) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nx) b_t[t] = b[lrows[t]];
}

__global__ void __launch_bounds__(CM_WG) k_cm_check(int n, int np, const double* __restrict__ r,
                                                    const double* __restrict__ vol, const double* __restrict__ len,
                                                    double pr, double jr, const double* __restrict__ x,
                                                    double* __restrict__ part) {
#pragma clang fp contract(off)
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double ri = r[i];
    if (i < np) {
      if (vol[i] > 0.0) v[0] = nanmax(v[0], fabs(ri) / vol[i]);
    } else {
      v[1] = nanmax(v[1], fabs(ri) / len[i - np]);
    }
    v[2] = v[2] + ri * ri;
    const double q = (i < np ? pr : jr) * ri;
    v[3] = v[3] + q * q;
    if (x && !isfinite(x[i])) v[4] = 1.0;
    v[5] = nanmax(v[5], fabs(ri));
  }
  wg_partials<6>(v, 0b110011, part);
}

__global__ void __launch_bounds__(CM_WG) k_cm_omitted(int m, const int* __restrict__ ptr, const int* __restrict__ col,
                                                      const double* __restrict__ val, const double* __restrict__ x,
                                                      const double* __restrict__ rhs, const double* __restrict__ vol,
                                                      double* __restrict__ part) {
#pragma clang fp contract(off)
  double v[1] = {0.0};
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < m; a += gridDim.x * blockDim.x)
    if (vol[a] > 0.0) v[0] = nanmax(v[0], fabs(rhs[a] - row_dot(ptr, col, val, x, a)) / vol[a]);
  wg_partials<1>(v, 1, part);
}

__global__ void __launch_bounds__(CM_WG) k_cm_reduce(int nparts, int nsets, int max_mask,
                                                     const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[8][CM_WG];
  fold(nparts, nsets, max_mask, part, red);
  if (threadIdx.x == 0)
    for (int s = 0; s < nsets; ++s) out[s] = red[s][0];
}

__global__ void __launch_bounds__(CM_WG) k_cm_increment(int n, int np, double* __restrict__ x, double pc, double jc,
                                                        double scale, const double* __restrict__ inc,
                                                        int* __restrict__ flag) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double previous = x[i];
  const double updated = previous + (i < np ? pc : jc) * (scale * inc[i]);
  x[i] = updated;
  if (updated != previous) *flag = 1;
}

}  // namespace

int cm_parts(int n) { return std::max(1, std::min(cdiv(n, CM_WG), CM_MAX_WG)); }

void launch_cm_spmv(hipStream_t s, int n, int np, const int* sptr, const int* len, const int* col,
                    const double* val, const double* x, double pc, double jc, double pr, double jr, double* y) {
  if (n > 0)
    hipLaunchKernelGGL(k_cm_spmv, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, np, sptr, len, col, val, x, pc, jc, pr,
                       jr, y);
}
void launch_cm_residual(hipStream_t s, int n, const int* sptr, const int* len, const int* col, const double* val,
                        const double* x, const double* b, double* r) {
  if (n > 0)
    hipLaunchKernelGGL(k_cm_residual, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, sptr, len, col, val, x, b, r);
}
void launch_cm_dots(hipStream_t s, int n, int nd, const double* a0, const double* b0, const double* a1,
                    const double* b1, double* part) {
  if (nd == 1)
    hipLaunchKernelGGL(k_cm_dots<1>, dim3(cm_parts(n)), dim3(CM_WG), 0, s, n, a0, b0, a1, b1, part);
  else
    hipLaunchKernelGGL(k_cm_dots<2>, dim3(cm_parts(n)), dim3(CM_WG), 0, s, n, a0, b0, a1, b1, part);
}
void launch_cm_scalars(hipStream_t s, int nparts, int nd, const double* part, double* scal, int step) {
  hipLaunchKernelGGL(k_cm_scalars, dim3(1), dim3(CM_WG), 0, s, nparts, nd, part, scal, step);
}
void launch_cm_update_s(hipStream_t s, int n, const double* r, const double* v, const double* y, double* sv,
                        double* x, const double* scal) {
  if (n > 0) hipLaunchKernelGGL(k_cm_update_s, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, r, v, y, sv, x, scal);
}
void launch_cm_update_r(hipStream_t s, int n, const double* sv, const double* t, const double* z, double* r,
                        double* x, const double* scal) {
  if (n > 0) hipLaunchKernelGGL(k_cm_update_r, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, sv, t, z, r, x, scal);
}
void launch_cm_update_p(hipStream_t s, int n, const double* r, const double* v, double* p, const double* scal) {
  if (n > 0) hipLaunchKernelGGL(k_cm_update_p, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, r, v, p, scal);
}
void launch_cm_scale_rows(hipStream_t s, int n, int np, const double* v, double pr, double jr, double* out) {
  if (n > 0) hipLaunchKernelGGL(k_cm_scale_rows, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, np, v, pr, jr, out);
}
void launch_cm_divs(hipStream_t s, int n, double* v, double sc) {
  if (n > 0) hipLaunchKernelGGL(k_cm_divs, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, v, sc);
}
void launch_cm_scatter(hipStream_t s, int ncells, const int* role, const double* v, double d, double* full) {
  if (ncells > 0)
    hipLaunchKernelGGL(k_cm_scatter, dim3(cdiv(ncells, CM_WG)), dim3(CM_WG), 0, s, ncells, role, v, d, full);
}
void launch_cm_anchor_chunks(hipStream_t s, int nchunks, const int* chunk_ptr, const int* list, const double* v,
                             double d, double* part) {
  if (nchunks > 0)
    hipLaunchKernelGGL(k_cm_anchor_chunks, dim3(nchunks), dim3(CM_WG), 0, s, chunk_ptr, list, v, d, part);
}
void launch_cm_anchor_totals(hipStream_t s, int nanchors, const int* anchor_cell, const int* anchor_chunk_ptr,
                             const double* part, double* full) {
  if (nanchors > 0)
    hipLaunchKernelGGL(k_cm_anchor_totals, dim3(nanchors), dim3(CM_WG), 0, s, anchor_cell, anchor_chunk_ptr, part,
                       full);
}
void launch_cm_gather(hipStream_t s, int np, const int* idx, const int* anchor, const double* z, double d,
                      double* out) {
  if (np > 0) hipLaunchKernelGGL(k_cm_gather, dim3(cdiv(np, CM_WG)), dim3(CM_WG), 0, s, np, idx, anchor, z, d, out);
}
void launch_cm_jump_rhs(hipStream_t s, int nx, const int* ptr, const int* col, const double* val, const double* zp,
                        const double* zfull, const int* idx, const int* anchor, const double* rhs, double d,
                        double* jr) {
  if (nx > 0)
    hipLaunchKernelGGL(k_cm_jump_rhs, dim3(cdiv(nx, CM_WG)), dim3(CM_WG), 0, s, nx, ptr, col, val, zp, zfull, idx,
                       anchor, rhs, d, jr);
}
void launch_cm_ilu(hipStream_t s, int nx, const int* lptr, const int* lcol, const double* lval, int nllev,
                   const int* llev, const int* lorder, const int* uptr, const int* ucol, const double* uval,
                   const double* udiag, int nulev, const int* ulev, const int* uorder, const double* b, double* y,
                   double* x, double d) {
  if (nx > 0)
    hipLaunchKernelGGL(k_cm_ilu, dim3(1), dim3(CM_ILU_WG), 0, s, nx, lptr, lcol, lval, nllev, llev, lorder, uptr, ucol,
                       uval, udiag, nulev, ulev, uorder, b, y, x, d);
}
void launch_cm_ilu_lds(hipStream_t s, int nx, int w, int nllev, const int* llev, const int* lrows, const int* llen,
                       const int* lcol, const double* lval, int nulev, const int* ulev, const int* urows,
                       const int* ulen, const int* ucol, const double* uval, const double* udiag, const double* b,
                       double* b_t, double* x, double d) {
  if (nx <= 0) return;
  hipLaunchKernelGGL(k_cm_ilu_rhs, dim3(cdiv(nx, CM_WG)), dim3(CM_WG), 0, s, nx, lrows, b, b_t);
  const size_t lds = size_t(nx) * sizeof(double) + size_t(nllev + nulev + 2) * sizeof(int);
  hipLaunchKernelGGL(k_cm_ilu_lds, dim3(1), dim3(CM_ILU_WG), lds, s, nx, w, nllev, llev, lrows, llen, lcol, lval,
                     nulev, ulev, urows, ulen, ucol, uval, udiag, b_t, x, d);
}
void launch_cm_check(hipStream_t s, int n, int np, const double* r, const double* vol, const double* len, double pr,
                     double jr, const double* x, double* part) {
  hipLaunchKernelGGL(k_cm_check, dim3(cm_parts(n)), dim3(CM_WG), 0, s, n, np, r, vol, len, pr, jr, x, part);
}
void launch_cm_omitted(hipStream_t s, int m, const int* ptr, const int* col, const double* val, const double* x,
                       const double* rhs, const double* vol, double* part) {
  hipLaunchKernelGGL(k_cm_omitted, dim3(cm_parts(m)), dim3(CM_WG), 0, s, m, ptr, col, val, x, rhs, vol, part);
}
void launch_cm_reduce(hipStream_t s, int nparts, int nsets, int max_mask, const double* part, double* out) {
  hipLaunchKernelGGL(k_cm_reduce, dim3(1), dim3(CM_WG), 0, s, nparts, nsets, max_mask, part, out);
}
void launch_cm_increment(hipStream_t s, int n, int np, double* x, double pc, double jc, double scale,
                         const double* inc, int* flag) {
  if (n > 0)
    hipLaunchKernelGGL(k_cm_increment, dim3(cdiv(n, CM_WG)), dim3(CM_WG), 0, s, n, np, x, pc, jc, scale, inc, flag);
}

}  // namespace tpamr
