// Staggered momentum kernels of the one-velocity step on a uniform, fully
// periodic block mesh: Lax-Wendroff / upwind / centred transport of both
// velocity components, the staggered viscous operator, and the Krylov
// pieces of the implicit-midpoint viscous solve. Every kernel is one pass
// over the structured component grids (row-major, periodic wrap), one
// thread per unknown; all of them are HBM-bandwidth bound.
//
// Reference arithmetic (TwoPhaseAMR.jl): the dual flux of a dual face is
// (u_a + u_b) * area / 2 of its two primal flux faces, the Lax-Wendroff
// state ((1 - C) down + (1 + C) up) / 2 - dt (q- div- + q+ div+) / 4
// (staggered_schedules.jl:952-968), the tendency sum(sign * flux * state) / V
// over the unknown's dual faces (staggered_fluxes.jl:241-287). In one sweep
// an unknown has two active dual faces, and a sum of two terms does not
// depend on their order, so the transport equals the reference's bit for
// bit (contraction off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "staggered.hpp"

namespace tpamr {
namespace {

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Dual flux through the face between unknowns k and k + 1 of a line of
// component `comp` in sweep direction `sweep` (1: x lines at row `line`,
// 2: y lines at column `line`), from the advecting grids (U, V):
//   comp 0, sweep 1: normal face through cell (k, line):      (U(k, line) + U(k+1, line)) dy / 2
//   comp 0, sweep 2: transverse face, vertex (line, k+1):     (V(line-1, k+1) + V(line, k+1)) dx / 2
//   comp 1, sweep 2: normal face through cell (line, k):      (V(line, k) + V(line, k+1)) dx / 2
//   comp 1, sweep 1: transverse face, vertex (k+1, line):     (U(k+1, line-1) + U(k+1, line)) dy / 2
// (_uniform_dual_schedule, staggered_schedules.jl:396-458).
__device__ __forceinline__ double dual_flux(const StagGrid& g, int comp, int sweep, int line, int k,
                                            const double* __restrict__ U, const double* __restrict__ V) {
#pragma clang fp contract(off)
  const int nx = g.nx, ny = g.ny;
  if (sweep == 1) {
    if (comp == 0) return (U[line * nx + wrap(k, nx)] + U[line * nx + wrap(k + 1, nx)]) * g.dy / 2;
    const int c = wrap(k + 1, nx);
    return (U[wrap(line - 1, ny) * nx + c] + U[line * nx + c]) * g.dy / 2;
  }
  if (comp == 0) {
    const int r = wrap(k + 1, ny) * nx;
    return (V[r + wrap(line - 1, nx)] + V[r + line]) * g.dx / 2;
  }
  return (V[wrap(k, ny) * nx + line] + V[wrap(k + 1, ny) * nx + line]) * g.dx / 2;
}

__device__ __forceinline__ double lw_state(double qm, double qp, double flux, double divm, double divp, double dt,
                                           double area, double dist) {
#pragma clang fp contract(off)
  const double speed = flux / area;
  const double courant = fabs(dt * speed / dist);
  const double up = flux >= 0 ? qm : qp;
  const double down = flux >= 0 ? qp : qm;
  return ((1.0 - courant) * down + (1.0 + courant) * up) / 2 - dt * (qm * divm + qp * divp) / 4;
}

__device__ __forceinline__ double face_state(int scheme, double qm, double qp, double flux, double divm, double divp,
                                             double dt, double area, double dist) {
#pragma clang fp contract(off)
  if (scheme == STAG_UPWIND) return flux >= 0 ? qm : qp;
  if (scheme == STAG_CENTRED) return (qm + qp) / 2;
  return lw_state(qm, qp, flux, divm, divp, dt, area, dist);
}

__global__ void k_st_gather(int n, const int* __restrict__ fx, const int* __restrict__ fy,
                            const double* __restrict__ u, double* __restrict__ X, double* __restrict__ Y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  X[k] = u[fx[k]];
  Y[k] = u[fy[k]];
}

__global__ void k_st_scatter(int n, const int* __restrict__ fx, const int* __restrict__ fy,
                             const double* __restrict__ X, const double* __restrict__ Y, double* __restrict__ u) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  u[fx[k]] = X[k];
  u[fy[k]] = Y[k];
}

// One directional sweep (1: x, 2: y) of the split transport, or the unsplit
// stage (sweep 0). Threads [0, n) take component 0, [n, 2n) component 1.
__global__ void k_st_sweep(StagGrid g, int sweep, int scheme, double dt, const double* __restrict__ Xi,
                           const double* __restrict__ Yi, const double* __restrict__ U,
                           const double* __restrict__ V, double* __restrict__ Xo, double* __restrict__ Yo) {
#pragma clang fp contract(off)
  const int n = g.nx * g.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n) return;
  const int comp = t >= n ? 1 : 0;
  const int cell = t - comp * n;
  const int i = cell % g.nx, j = cell / g.nx;
  const double* Q = comp ? Yi : Xi;
  double tend = 0.0;
  if (sweep != 0) {
    // unknowns k0 - 1, k0, k0 + 1 of the sweep's line; fluxes of the dual
    // faces (k0-2|k0-1), (k0-1|k0), (k0|k0+1), (k0+1|k0+2)
    const bool xl = sweep == 1;
    const int line = xl ? j : i, k0 = xl ? i : j;
    const double area = xl ? g.dy : g.dx, dist = xl ? g.dx : g.dy;
    const double qm1 = xl ? Q[j * g.nx + wrap(i - 1, g.nx)] : Q[wrap(j - 1, g.ny) * g.nx + i];
    const double q0 = Q[cell];
    const double qp1 = xl ? Q[j * g.nx + wrap(i + 1, g.nx)] : Q[wrap(j + 1, g.ny) * g.nx + i];
    const double fm = dual_flux(g, comp, sweep, line, k0 - 1, U, V);
    const double f0 = dual_flux(g, comp, sweep, line, k0, U, V);
    double sm, s0;
    if (scheme == STAG_LW) {
      const double fmm = dual_flux(g, comp, sweep, line, k0 - 2, U, V);
      const double fp = dual_flux(g, comp, sweep, line, k0 + 1, U, V);
      // directional divergence of the advecting velocity per unknown
      const double dm1 = (fm - fmm) * g.inv_vol;
      const double d0 = (f0 - fm) * g.inv_vol;
      const double dp1 = (fp - f0) * g.inv_vol;
      sm = lw_state(qm1, q0, fm, dm1, d0, dt, area, dist);
      s0 = lw_state(q0, qp1, f0, d0, dp1, dt, area, dist);
    } else {
      sm = face_state(scheme, qm1, q0, fm, 0.0, 0.0, dt, area, dist);
      s0 = face_state(scheme, q0, qp1, f0, 0.0, 0.0, dt, area, dist);
    }
    tend = (f0 * s0 - fm * sm) * g.inv_vol;
  } else {
    // all four dual faces of the unknown, no predictor divergence
    double acc = 0.0;
    for (int sw = 1; sw <= 2; ++sw) {
      const bool xl = sw == 1;
      const int line = xl ? j : i, k0 = xl ? i : j;
      const double area = xl ? g.dy : g.dx, dist = xl ? g.dx : g.dy;
      const double qm1 = xl ? Q[j * g.nx + wrap(i - 1, g.nx)] : Q[wrap(j - 1, g.ny) * g.nx + i];
      const double q0 = Q[cell];
      const double qp1 = xl ? Q[j * g.nx + wrap(i + 1, g.nx)] : Q[wrap(j + 1, g.ny) * g.nx + i];
      const double fm = dual_flux(g, comp, sw, line, k0 - 1, U, V);
      const double f0 = dual_flux(g, comp, sw, line, k0, U, V);
      const double sm = face_state(scheme, qm1, q0, fm, 0.0, 0.0, dt, area, dist);
      const double s0 = face_state(scheme, q0, qp1, f0, 0.0, 0.0, dt, area, dist);
      acc = acc + f0 * s0;
      acc = acc - fm * sm;
    }
    tend = acc * g.inv_vol;
  }
  (comp ? Yo : Xo)[cell] = Q[cell] - dt * tend;
}

// rate of an unknown: sum of |dual flux| over its four dual faces / V;
// per-workgroup maxima (grid-stride over both components' unknowns)
template <int NS>
__device__ __forceinline__ void wg_partials(double (&v)[NS], int max_mask, double* __restrict__ part);

__global__ void __launch_bounds__(STAG_WG) k_st_rates(StagGrid g, const double* __restrict__ U,
                                                      const double* __restrict__ V, double* __restrict__ part) {
#pragma clang fp contract(off)
  const int n = g.nx * g.ny;
  double v[1] = {0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * n; t += gridDim.x * blockDim.x) {
    const int comp = t >= n ? 1 : 0;
    const int cell = t - comp * n;
    const int i = cell % g.nx, j = cell / g.nx;
    const double a = fabs(dual_flux(g, comp, 1, j, i - 1, U, V)) + fabs(dual_flux(g, comp, 1, j, i, U, V));
    const double b = fabs(dual_flux(g, comp, 2, i, j - 1, U, V)) + fabs(dual_flux(g, comp, 2, i, j, U, V));
    v[0] = fmax(v[0], (a + b) * g.inv_vol);
  }
  wg_partials<1>(v, 1, part);
}

// nu (L q) at one unknown: diagonal, then x neighbours, then y neighbours
__device__ __forceinline__ double lap(const StagGrid& g, const StagCoef& k, const double* __restrict__ Q, int i,
                                      int j) {
#pragma clang fp contract(off)
  const int c = j * g.nx + i;
  double acc = k.ld * Q[c];
  acc = acc + k.lx * Q[j * g.nx + wrap(i - 1, g.nx)];
  acc = acc + k.lx * Q[j * g.nx + wrap(i + 1, g.nx)];
  acc = acc + k.ly * Q[wrap(j - 1, g.ny) * g.nx + i];
  acc = acc + k.ly * Q[wrap(j + 1, g.ny) * g.nx + i];
  return acc;
}

__device__ __forceinline__ double shifted(const StagGrid& g, const StagCoef& k, const double* __restrict__ Q, int i,
                                          int j) {
#pragma clang fp contract(off)
  const int c = j * g.nx + i;
  double acc = k.md * Q[c];
  acc = acc + k.mx * Q[j * g.nx + wrap(i - 1, g.nx)];
  acc = acc + k.mx * Q[j * g.nx + wrap(i + 1, g.nx)];
  acc = acc + k.my * Q[wrap(j - 1, g.ny) * g.nx + i];
  acc = acc + k.my * Q[wrap(j + 1, g.ny) * g.nx + i];
  return acc;
}

__global__ void k_st_diffusion(StagGrid g, StagCoef k, double nu, const double* __restrict__ X,
                               const double* __restrict__ Y, double* __restrict__ oX, double* __restrict__ oY) {
#pragma clang fp contract(off)
  const int n = g.nx * g.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n) return;
  const int comp = t >= n ? 1 : 0;
  const int cell = t - comp * n;
  (comp ? oY : oX)[cell] = nu * lap(g, k, comp ? Y : X, cell % g.nx, cell / g.nx);
}

__global__ void k_st_midpoint_rhs(StagGrid g, StagCoef k, double nu, double dt, const double* __restrict__ c,
                                  const double* __restrict__ o, double* __restrict__ b) {
#pragma clang fp contract(off)
  const int n = g.nx * g.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double work = nu * lap(g, k, o, t % g.nx, t / g.nx);
  const double w = c[t] + dt * work / 2;
  b[t] = (g.dx * g.dy) * w;
}

// Workgroup partials: NS sets of one value per thread; sums or maxima per
// set (bit k of max_mask), lane 0 of wave 0 writes part[blockIdx.x * NS + k].
template <int NS>
__device__ __forceinline__ void wg_partials(double (&v)[NS], int max_mask, double* __restrict__ part) {
  __shared__ double red[NS][STAG_WG / 64];
  for (int s = 0; s < NS; ++s) {
    const bool mx = (max_mask >> s) & 1;
    for (int o = 32; o > 0; o >>= 1) {
      const double other = __shfl_xor(v[s], o, 64);
      v[s] = mx ? fmax(v[s], other) : v[s] + other;
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
      for (int k = 1; k < STAG_WG / 64; ++k) acc = mx ? fmax(acc, red[s][k]) : acc + red[s][k];
      part[blockIdx.x * NS + s] = acc;
    }
  }
}

__global__ void __launch_bounds__(STAG_WG) k_st_residual(StagGrid g, StagCoef k, const double* __restrict__ b,
                                                         const double* __restrict__ x, double* __restrict__ r,
                                                         double* __restrict__ part) {
#pragma clang fp contract(off)
  const int n = g.nx * g.ny;
  const double vol = g.dx * g.dy;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int i = t % g.nx, j = t / g.nx;
    const double rr = b[t] - shifted(g, k, x, i, j);
    r[t] = rr;
    const double strong = rr / vol;
    v[0] = fmax(v[0], fabs(strong));
    v[1] = v[1] + vol * (strong * strong);
    // round-off bound row: sum |M_ij| |x_j|
    double bound = fabs(k.md) * fabs(x[t]);
    bound = bound + fabs(k.mx) * fabs(x[j * g.nx + wrap(i - 1, g.nx)]);
    bound = bound + fabs(k.mx) * fabs(x[j * g.nx + wrap(i + 1, g.nx)]);
    bound = bound + fabs(k.my) * fabs(x[wrap(j - 1, g.ny) * g.nx + i]);
    bound = bound + fabs(k.my) * fabs(x[wrap(j + 1, g.ny) * g.nx + i]);
    v[2] = fmax(v[2], bound / vol);
    v[3] = fmax(v[3], fabs(b[t] / vol));
  }
  wg_partials<4>(v, 0b1101, part);
}

__global__ void __launch_bounds__(STAG_WG) k_st_cg_start(int n, StagCoef k, const double* __restrict__ r,
                                                         double* __restrict__ z, double* __restrict__ p,
                                                         double* __restrict__ part) {
#pragma clang fp contract(off)
  double v[1] = {0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const double zz = k.mdinv * r[t];
    z[t] = zz;
    p[t] = zz;
    v[0] = v[0] + r[t] * zz;
  }
  wg_partials<1>(v, 0, part);
}

__global__ void __launch_bounds__(STAG_WG) k_st_matvec_dot(StagGrid g, StagCoef k, const double* __restrict__ p,
                                                           double* __restrict__ Ap, double* __restrict__ part) {
#pragma clang fp contract(off)
  const int n = g.nx * g.ny;
  double v[1] = {0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const double a = shifted(g, k, p, t % g.nx, t / g.nx);
    Ap[t] = a;
    v[0] = v[0] + p[t] * a;
  }
  wg_partials<1>(v, 0, part);
}

__global__ void __launch_bounds__(STAG_WG) k_st_cg_update(int n, StagCoef k, double inv_vol, double* __restrict__ x,
                                                          double* __restrict__ r, double* __restrict__ z,
                                                          const double* __restrict__ p,
                                                          const double* __restrict__ Ap,
                                                          const double* __restrict__ scal,
                                                          double* __restrict__ part) {
#pragma clang fp contract(off)
  const double alpha = scal[1];
  double v[2] = {0.0, 0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    x[t] = x[t] + alpha * p[t];
    const double rr = r[t] - alpha * Ap[t];
    r[t] = rr;
    const double zz = k.mdinv * rr;
    z[t] = zz;
    v[0] = v[0] + rr * zz;
    v[1] = fmax(v[1], fabs(rr * inv_vol));
  }
  wg_partials<2>(v, 0b10, part);
}

__global__ void k_st_cg_direction(int n, const double* __restrict__ z, double* __restrict__ p,
                                  const double* __restrict__ scal) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = z[t] + scal[2] * p[t];
}

// one workgroup: fixed-order reduction of the partials, then the scalar step
__global__ void __launch_bounds__(STAG_WG) k_st_reduce(int nparts, int nsets, int max_mask,
                                                       const double* __restrict__ part, double* __restrict__ scal,
                                                       int step) {
#pragma clang fp contract(off)
  __shared__ double red[8][STAG_WG];
  for (int s = 0; s < nsets; ++s) {
    const bool mx = (max_mask >> s) & 1;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += STAG_WG) {
      const double v = part[i * nsets + s];
      acc = mx ? fmax(acc, v) : acc + v;
    }
    red[s][threadIdx.x] = acc;
  }
  __syncthreads();
  for (int w = STAG_WG / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int s = 0; s < nsets; ++s) {
        const bool mx = (max_mask >> s) & 1;
        red[s][threadIdx.x] = mx ? fmax(red[s][threadIdx.x], red[s][threadIdx.x + w])
                                 : red[s][threadIdx.x] + red[s][threadIdx.x + w];
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int s = 0; s < nsets; ++s) scal[4 + s] = red[s][0];
    const double s0 = red[0][0];
    if (step == 1) {
      scal[0] = s0;
    } else if (step == 2) {
      scal[1] = scal[0] / s0;
    } else if (step == 3) {
      scal[2] = s0 / scal[0];
      scal[0] = s0;
      scal[3] = red[1][0];
    } else if (step == 4) {
      scal[2] = s0 / scal[0];
      scal[0] = s0;
    }
  }
}

// ---------------------------------------------------------------------------
// Staggered tree multigrid (one component grid per level)

__device__ __forceinline__ double mg_apply(int nx, int ny, const StagCoef& k, const double* __restrict__ Q, int i,
                                           int j) {
#pragma clang fp contract(off)
  double acc = k.md * Q[j * nx + i];
  acc = acc + k.mx * Q[j * nx + wrap(i - 1, nx)];
  acc = acc + k.mx * Q[j * nx + wrap(i + 1, nx)];
  acc = acc + k.my * Q[wrap(j - 1, ny) * nx + i];
  acc = acc + k.my * Q[wrap(j + 1, ny) * nx + i];
  return acc;
}

__global__ void k_st_mg_smooth(int nx, int ny, StagCoef k, double w, const double* __restrict__ x_in,
                               const double* __restrict__ b, double* __restrict__ x_out) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny) return;
  if (!x_in) {
    x_out[t] = (w * k.mdinv) * b[t];
    return;
  }
  const double mx = mg_apply(nx, ny, k, x_in, t % nx, t / nx);
  x_out[t] = x_in[t] + (w * k.mdinv) * (b[t] - mx);
}

__global__ void k_st_mg_residual(int nx, int ny, StagCoef k, const double* __restrict__ b,
                                 const double* __restrict__ x, double* __restrict__ r) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny) return;
  r[t] = b[t] - mg_apply(nx, ny, k, x, t % nx, t / nx);
}

// coarse unknown (ic, jc) of component 0 gathers the fine x-faces of the two
// fine rows 2 jc, 2 jc + 1: the face on its line (weight 1) and the mid-line
// faces on either side (weight 1/2); component 1 likewise along columns
__global__ void k_st_mg_restrict(int ncx, int ncy, int comp, const double* __restrict__ rf,
                                 double* __restrict__ bc) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncx * ncy) return;
  const int ic = t % ncx, jc = t / ncx;
  const int nfx = 2 * ncx, nfy = 2 * ncy;
  double acc = 0.0;
  for (int s = 0; s < 2; ++s) {
    if (comp == 0) {
      const int jf = 2 * jc + s, i0 = 2 * ic;
      acc = acc + rf[jf * nfx + i0];
      acc = acc + 0.5 * rf[jf * nfx + i0 + 1];
      acc = acc + 0.5 * rf[jf * nfx + wrap(i0 - 1, nfx)];
    } else {
      const int ifc = 2 * ic + s, j0 = 2 * jc;
      acc = acc + rf[j0 * nfx + ifc];
      acc = acc + 0.5 * rf[(j0 + 1) * nfx + ifc];
      acc = acc + 0.5 * rf[wrap(j0 - 1, nfy) * nfx + ifc];
    }
  }
  bc[t] = acc;
}

__global__ void k_st_mg_prolong(int ncx, int ncy, int comp, const double* __restrict__ xc,
                                double* __restrict__ xf) {
#pragma clang fp contract(off)
  const int nfx = 2 * ncx, nfy = 2 * ncy;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nfx * nfy) return;
  const int i = t % nfx, j = t / nfx;
  double v;
  if (comp == 0) {
    const int jc = j >> 1;
    if ((i & 1) == 0) {
      v = xc[jc * ncx + (i >> 1)];
    } else {
      const int lo = i >> 1;
      v = 0.5 * xc[jc * ncx + lo] + 0.5 * xc[jc * ncx + wrap(lo + 1, ncx)];
    }
  } else {
    const int ic = i >> 1;
    if ((j & 1) == 0) {
      v = xc[(j >> 1) * ncx + ic];
    } else {
      const int lo = j >> 1;
      v = 0.5 * xc[lo * ncx + ic] + 0.5 * xc[wrap(lo + 1, ncy) * ncx + ic];
    }
  }
  xf[t] = xf[t] + v;
}

// M Q at (i, j) for Q given pointwise by q(index), in mg_apply's order
template <typename F>
__device__ __forceinline__ double mg_apply_f(int nx, int ny, const StagCoef& k, F q, int i, int j) {
#pragma clang fp contract(off)
  double acc = k.md * q(j * nx + i);
  acc = acc + k.mx * q(j * nx + wrap(i - 1, nx));
  acc = acc + k.mx * q(j * nx + wrap(i + 1, nx));
  acc = acc + k.my * q(wrap(j - 1, ny) * nx + i);
  acc = acc + k.my * q(wrap(j + 1, ny) * nx + i);
  return acc;
}

// the zero-start sweep and the next one in one pass: x1 = (w D^-1) b at the
// five points, x2 = x1 + (w D^-1)(b - M x1); the same operations as two
// k_st_mg_smooth launches
__global__ void k_st_mg_smooth2z(int nx, int ny, StagCoef k, double w, const double* __restrict__ b,
                                 double* __restrict__ x_out) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny) return;
  const double wd = w * k.mdinv;
  auto x1 = [&](int q) { return wd * b[q]; };
  const double mx = mg_apply_f(nx, ny, k, x1, t % nx, t / nx);
  x_out[t] = x1(t) + wd * (b[t] - mx);
}

// restriction of the fine residual b - M x computed on the fly (the fine
// residuals of k_st_mg_restrict's six faces)
__global__ void k_st_mg_restrict_res(int ncx, int ncy, int comp, StagCoef k, const double* __restrict__ bf,
                                     const double* __restrict__ xf, double* __restrict__ bc) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncx * ncy) return;
  const int ic = t % ncx, jc = t / ncx;
  const int nfx = 2 * ncx, nfy = 2 * ncy;
  auto res = [&](int i, int j) { return bf[j * nfx + i] - mg_apply(nfx, nfy, k, xf, i, j); };
  double acc = 0.0;
  for (int s = 0; s < 2; ++s) {
    if (comp == 0) {
      const int jf = 2 * jc + s, i0 = 2 * ic;
      acc = acc + res(i0, jf);
      acc = acc + 0.5 * res(i0 + 1, jf);
      acc = acc + 0.5 * res(wrap(i0 - 1, nfx), jf);
    } else {
      const int ifc = 2 * ic + s, j0 = 2 * jc;
      acc = acc + res(ifc, j0);
      acc = acc + 0.5 * res(ifc, j0 + 1);
      acc = acc + 0.5 * res(ifc, wrap(j0 - 1, nfy));
    }
  }
  bc[t] = acc;
}

// the prolonged correction of fine unknown (i, j), as k_st_mg_prolong adds it
__device__ __forceinline__ double prolonged(int ncx, int ncy, int comp, const double* __restrict__ xc, int i, int j) {
#pragma clang fp contract(off)
  if (comp == 0) {
    const int jc = j >> 1;
    if ((i & 1) == 0) return xc[jc * ncx + (i >> 1)];
    const int lo = i >> 1;
    return 0.5 * xc[jc * ncx + lo] + 0.5 * xc[jc * ncx + wrap(lo + 1, ncx)];
  }
  const int ic = i >> 1;
  if ((j & 1) == 0) return xc[(j >> 1) * ncx + ic];
  const int lo = j >> 1;
  return 0.5 * xc[lo * ncx + ic] + 0.5 * xc[wrap(lo + 1, ncy) * ncx + ic];
}

// prolongation and the first post-sweep in one pass: xp = x + P xc at the
// five points, out = xp + (w D^-1)(b - M xp)
__global__ void k_st_mg_prolong_smooth(int ncx, int ncy, int comp, StagCoef k, double w,
                                       const double* __restrict__ xc, const double* __restrict__ xf,
                                       const double* __restrict__ b, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int nfx = 2 * ncx, nfy = 2 * ncy;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nfx * nfy) return;
  auto xp = [&](int q) { return xf[q] + prolonged(ncx, ncy, comp, xc, q % nfx, q / nfx); };
  const double mx = mg_apply_f(nfx, nfy, k, xp, t % nfx, t / nfx);
  out[t] = xp(t) + (w * k.mdinv) * (b[t] - mx);
}

// one wavefront per row of the dense inverse
__global__ void __launch_bounds__(256) k_st_dense_solve(int n, const double* __restrict__ inv,
                                                        const double* __restrict__ b, double* __restrict__ x) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  double acc = 0.0;
  for (int c = lane; c < n; c += 64) acc += inv[size_t(row) * n + c] * b[c];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) x[row] = acc;
}

// x = A^-1 b for a symmetric circulant A on an nx x ny periodic grid, whose
// inverse is the periodic convolution with its first column g (row-major
// (j, i)): x[r] = sum_c g[(r_i - c_i) mod nx, (r_j - c_j) mod ny] b[c]. One
// wavefront per output cell, lanes striding along the grid rows.
__global__ void __launch_bounds__(256) k_st_circulant_solve(int nx, int ny, const double* __restrict__ g,
                                                            const double* __restrict__ b, double* __restrict__ x) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= nx * ny) return;
  const int ri = row % nx, rj = row / nx;
  double acc = 0.0;
  for (int cj = 0; cj < ny; ++cj) {
    const int dj = rj >= cj ? rj - cj : rj - cj + ny;
    const double* gr = g + size_t(dj) * nx;
    const double* br = b + size_t(cj) * nx;
    for (int ci = lane; ci < nx; ci += 64) {
      const int di = ri >= ci ? ri - ci : ri - ci + nx;
      acc += gr[di] * br[ci];
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) x[row] = acc;
}

__global__ void __launch_bounds__(STAG_WG) k_st_cg_update_xr(int n, double inv_vol, double* __restrict__ x,
                                                             double* __restrict__ r, const double* __restrict__ p,
                                                             const double* __restrict__ Ap,
                                                             const double* __restrict__ scal,
                                                             double* __restrict__ part) {
#pragma clang fp contract(off)
  const double alpha = scal[1];
  double v[1] = {0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    x[t] = x[t] + alpha * p[t];
    const double rr = r[t] - alpha * Ap[t];
    r[t] = rr;
    v[0] = fmax(v[0], fabs(rr * inv_vol));
  }
  wg_partials<1>(v, 1, part);
}

__global__ void __launch_bounds__(STAG_WG) k_st_dot(int n, const double* __restrict__ r, const double* __restrict__ z,
                                                    double* __restrict__ part) {
#pragma clang fp contract(off)
  double v[1] = {0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) v[0] = v[0] + r[t] * z[t];
  wg_partials<1>(v, 0, part);
}

inline int cdiv2(int a, int b) { return (a + b - 1) / b; }

}  // namespace

void launch_st_mg_smooth(hipStream_t s, int nx, int ny, const StagCoef& k, double w, const double* x_in,
                         const double* b, double* x_out) {
  hipLaunchKernelGGL(k_st_mg_smooth, dim3(cdiv2(nx * ny, 256)), dim3(256), 0, s, nx, ny, k, w, x_in, b, x_out);
}
void launch_st_mg_residual(hipStream_t s, int nx, int ny, const StagCoef& k, const double* b, const double* x,
                           double* r) {
  hipLaunchKernelGGL(k_st_mg_residual, dim3(cdiv2(nx * ny, 256)), dim3(256), 0, s, nx, ny, k, b, x, r);
}
void launch_st_mg_restrict(hipStream_t s, int ncx, int ncy, int comp, const double* r_f, double* b_c) {
  hipLaunchKernelGGL(k_st_mg_restrict, dim3(cdiv2(ncx * ncy, 256)), dim3(256), 0, s, ncx, ncy, comp, r_f, b_c);
}
void launch_st_mg_prolong(hipStream_t s, int ncx, int ncy, int comp, const double* x_c, double* x_f) {
  hipLaunchKernelGGL(k_st_mg_prolong, dim3(cdiv2(4 * ncx * ncy, 256)), dim3(256), 0, s, ncx, ncy, comp, x_c, x_f);
}
void launch_st_mg_smooth2z(hipStream_t s, int nx, int ny, const StagCoef& k, double w, const double* b,
                           double* x_out) {
  hipLaunchKernelGGL(k_st_mg_smooth2z, dim3(cdiv2(nx * ny, 256)), dim3(256), 0, s, nx, ny, k, w, b, x_out);
}
void launch_st_mg_restrict_res(hipStream_t s, int ncx, int ncy, int comp, const StagCoef& k, const double* b_f,
                               const double* x_f, double* b_c) {
  hipLaunchKernelGGL(k_st_mg_restrict_res, dim3(cdiv2(ncx * ncy, 256)), dim3(256), 0, s, ncx, ncy, comp, k, b_f, x_f,
                     b_c);
}
void launch_st_mg_prolong_smooth(hipStream_t s, int ncx, int ncy, int comp, const StagCoef& k, double w,
                                 const double* x_c, const double* x_f, const double* b, double* out) {
  hipLaunchKernelGGL(k_st_mg_prolong_smooth, dim3(cdiv2(4 * ncx * ncy, 256)), dim3(256), 0, s, ncx, ncy, comp, k, w,
                     x_c, x_f, b, out);
}
void launch_st_dense_solve(hipStream_t s, int n, const double* inv, const double* b, double* x) {
  hipLaunchKernelGGL(k_st_dense_solve, dim3(cdiv2(n, 4)), dim3(256), 0, s, n, inv, b, x);
}
void launch_st_circulant_solve(hipStream_t s, int nx, int ny, const double* g, const double* b, double* x) {
  hipLaunchKernelGGL(k_st_circulant_solve, dim3(cdiv2(nx * ny, 4)), dim3(256), 0, s, nx, ny, g, b, x);
}
void launch_st_cg_update_xr(hipStream_t s, int n, double inv_vol, double* x, double* r, const double* p,
                            const double* Ap, const double* scal, double* part) {
  hipLaunchKernelGGL(k_st_cg_update_xr, dim3(st_parts(n)), dim3(STAG_WG), 0, s, n, inv_vol, x, r, p, Ap, scal, part);
}
void launch_st_dot(hipStream_t s, int n, const double* r, const double* z, double* part) {
  hipLaunchKernelGGL(k_st_dot, dim3(st_parts(n)), dim3(STAG_WG), 0, s, n, r, z, part);
}

namespace {
inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

// workgroups of the reduction kernels (grid-stride beyond STAG_MAX_WG)
int st_parts(int n) { return std::min(cdiv(n, STAG_WG), STAG_MAX_WG); }

void launch_st_gather(hipStream_t s, int n, const int* fx, const int* fy, const double* u, double* X, double* Y) {
  hipLaunchKernelGGL(k_st_gather, dim3(cdiv(n, 256)), dim3(256), 0, s, n, fx, fy, u, X, Y);
}

void launch_st_scatter(hipStream_t s, int n, const int* fx, const int* fy, const double* X, const double* Y,
                       double* u) {
  hipLaunchKernelGGL(k_st_scatter, dim3(cdiv(n, 256)), dim3(256), 0, s, n, fx, fy, X, Y, u);
}

void launch_st_sweep(hipStream_t s, const StagGrid& g, int sweep, int scheme, double dt, const double* Xi,
                     const double* Yi, const double* U, const double* V, double* Xo, double* Yo) {
  const int n2 = 2 * g.nx * g.ny;
  hipLaunchKernelGGL(k_st_sweep, dim3(cdiv(n2, 256)), dim3(256), 0, s, g, sweep, scheme, dt, Xi, Yi, U, V, Xo, Yo);
}

void launch_st_rates(hipStream_t s, const StagGrid& g, const double* U, const double* V, double* part) {
  hipLaunchKernelGGL(k_st_rates, dim3(st_parts(2 * g.nx * g.ny)), dim3(STAG_WG), 0, s, g, U, V, part);
}

void launch_st_diffusion(hipStream_t s, const StagGrid& g, const StagCoef& k, double nu, const double* X,
                         const double* Y, double* oX, double* oY) {
  const int n2 = 2 * g.nx * g.ny;
  hipLaunchKernelGGL(k_st_diffusion, dim3(cdiv(n2, 256)), dim3(256), 0, s, g, k, nu, X, Y, oX, oY);
}

void launch_st_midpoint_rhs(hipStream_t s, const StagGrid& g, const StagCoef& k, double nu, double dt,
                            const double* c, const double* o, double* b) {
  const int n = g.nx * g.ny;
  hipLaunchKernelGGL(k_st_midpoint_rhs, dim3(cdiv(n, 256)), dim3(256), 0, s, g, k, nu, dt, c, o, b);
}

void launch_st_residual(hipStream_t s, const StagGrid& g, const StagCoef& k, const double* b, const double* x,
                        double* r, double* part) {
  const int n = g.nx * g.ny;
  hipLaunchKernelGGL(k_st_residual, dim3(st_parts(n)), dim3(STAG_WG), 0, s, g, k, b, x, r, part);
}

void launch_st_cg_start(hipStream_t s, int n, const StagCoef& k, const double* r, double* z, double* p,
                        double* part) {
  hipLaunchKernelGGL(k_st_cg_start, dim3(st_parts(n)), dim3(STAG_WG), 0, s, n, k, r, z, p, part);
}

void launch_st_matvec_dot(hipStream_t s, const StagGrid& g, const StagCoef& k, const double* p, double* Ap,
                          double* part) {
  const int n = g.nx * g.ny;
  hipLaunchKernelGGL(k_st_matvec_dot, dim3(st_parts(n)), dim3(STAG_WG), 0, s, g, k, p, Ap, part);
}

void launch_st_cg_update(hipStream_t s, int n, const StagCoef& k, double inv_vol, double* x, double* r, double* z,
                         const double* p, const double* Ap, const double* scal, double* part) {
  hipLaunchKernelGGL(k_st_cg_update, dim3(st_parts(n)), dim3(STAG_WG), 0, s, n, k, inv_vol, x, r, z, p, Ap, scal,
                     part);
}

void launch_st_cg_direction(hipStream_t s, int n, const double* z, double* p, const double* scal) {
  hipLaunchKernelGGL(k_st_cg_direction, dim3(cdiv(n, 256)), dim3(256), 0, s, n, z, p, scal);
}

void launch_st_reduce(hipStream_t s, int nparts, int nsets, int max_mask, const double* part, double* scal, int step) {
  hipLaunchKernelGGL(k_st_reduce, dim3(1), dim3(STAG_WG), 0, s, nparts, nsets, max_mask, part, scal, step);
}

}  // namespace tpamr
