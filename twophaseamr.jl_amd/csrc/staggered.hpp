// Staggered (face-centred) momentum operators of the one-velocity step on a
// uniform, fully periodic block mesh (staggered.hip).
//
// Both velocity components live on structured component grids of the mesh's
// nx x ny cell lattice, row-major (j * nx + i) with periodic wrap:
//   X(i, j): x-velocity of the x-face at the left of cell (i, j),
//   Y(i, j): y-velocity of the y-face below cell (i, j).
// A gather/scatter through per-unknown face indices moves values between
// the reference's face order and these grids.
#pragma once
#include <hip/hip_runtime.h>

namespace tpamr {

struct StagGrid {
  int nx, ny;
  double dx, dy;
  double inv_vol;  // 1 / (dx dy), the dual volume's inverse
};

// Coefficients of the staggered viscous operator (identical for both
// components on a uniform mesh): the Laplacian L = -V^-1 K (diagonal ld,
// x-neighbour lx, y-neighbour ly) and the shifted matrix M = h K + V of the
// implicit substep (md, mx, my) with its diagonal inverse.
struct StagCoef {
  double ld, lx, ly;
  double md, mx, my, mdinv;
};

// Advection schemes of the staggered transport (staggered_schemes.jl).
constexpr int STAG_UPWIND = 0, STAG_CENTRED = 1, STAG_LW = 2;

// Threads per workgroup of the reduction kernels, and their largest grid
// (grid-stride loops beyond it keep the partials few).
constexpr int STAG_WG = 256;
constexpr int STAG_MAX_WG = 1024;
// Largest coarsest level of the viscous staggered multigrid whose dense
// inverse is formed (Gauss-Jordan on the host); a larger one is solved by
// the periodic convolution with the first column of its inverse, up to
// STAG_CIRC_MAX unknowns.
constexpr int STAG_DENSE_MAX = 1024;
constexpr int STAG_CIRC_MAX = 65536;

void launch_st_gather(hipStream_t s, int n, const int* fx, const int* fy, const double* u, double* X, double* Y);
void launch_st_scatter(hipStream_t s, int n, const int* fx, const int* fy, const double* X, const double* Y,
                       double* u);
// One transport stage of both components: sweep 1 (x lines) or 2 (y lines)
// of the directionally split Lax-Wendroff transport, or sweep 0 (all dual
// faces, no predictor divergence) for the unsplit schemes; writes
// Q - dt * tendency(Q) for Q = Xi, Yi, transported by the face velocity on
// the grids (U, V).
void launch_st_sweep(hipStream_t s, const StagGrid& g, int sweep, int scheme, double dt, const double* Xi,
                     const double* Yi, const double* U, const double* V, double* Xo, double* Yo);
// Per-workgroup maxima (st_parts(2 n) partials) of the rate
// sum(|dual flux|) / V over both components' unknowns.
void launch_st_rates(hipStream_t s, const StagGrid& g, const double* U, const double* V, double* part);
// out = nu L q for both component grids.
void launch_st_diffusion(hipStream_t s, const StagGrid& g, const StagCoef& k, double nu, const double* X,
                         const double* Y, double* oX, double* oY);
// w = c + dt / 2 * (nu L o), then b = V w (the implicit substep's rhs).
void launch_st_midpoint_rhs(hipStream_t s, const StagGrid& g, const StagCoef& k, double nu, double dt,
                            const double* c, const double* o, double* b);
// Krylov pieces of the viscous solve on one component grid (n unknowns):
// r = b - M x with per-workgroup partials (max |r| / V, sum r^2 V^-1, and
// max of the round-off bound (|M||x|) / V and |b| / V).
void launch_st_residual(hipStream_t s, const StagGrid& g, const StagCoef& k, const double* b, const double* x,
                        double* r, double* part);
// z = D^-1 r, p = z, partial sums of r z
void launch_st_cg_start(hipStream_t s, int n, const StagCoef& k, const double* r, double* z, double* p,
                        double* part);
// Ap = M p, partial sums of p Ap
void launch_st_matvec_dot(hipStream_t s, const StagGrid& g, const StagCoef& k, const double* p, double* Ap,
                          double* part);
// x += alpha p, r -= alpha Ap, z = D^-1 r; partials of r z and max |r| / V
// (alpha from the scalars the previous reduction left in `scal`)
void launch_st_cg_update(hipStream_t s, int n, const StagCoef& k, double inv_vol, double* x, double* r, double* z,
                         const double* p, const double* Ap, const double* scal, double* part);
// p = z + beta p
void launch_st_cg_direction(hipStream_t s, int n, const double* z, double* p, const double* scal);
// Fixed-order reduction of `nparts` partials of `nsets` (<= 8) sets (set k
// of part i at part[i * nsets + k]); set k is a maximum when bit k of
// max_mask is set, else a sum. Then the Krylov scalar step `step`: 0 none;
// 1 gamma = s0; 2 alpha = gamma / s0 (s0 = pAp); 3 beta = s0 / gamma,
// gamma = s0, strong = s1; 4 beta = s0 / gamma, gamma = s0.
// scal: [0] gamma, [1] alpha, [2] beta, [3] strong, [4..] the reduced sets.
void launch_st_reduce(hipStream_t s, int nparts, int nsets, int max_mask, const double* part, double* scal, int step);
int st_parts(int n);

// Staggered tree multigrid of the viscous solve (one component grid per
// level; levels are nested periodic grids, the coarse one half as fine per
// axis). Level operator M = diagonal md, x-neighbours mx, y-neighbours my,
// inverse diagonal mdinv.
// x_out = x_in + (w D^-1)(b - M x_in); x_in = NULL: the zero start (w D^-1) b
void launch_st_mg_smooth(hipStream_t s, int nx, int ny, const StagCoef& k, double w, const double* x_in,
                         const double* b, double* x_out);
// r = b - M x
void launch_st_mg_residual(hipStream_t s, int nx, int ny, const StagCoef& k, const double* b, const double* x,
                           double* r);
// b_c = P' r_f and x_f += P x_c for component `comp` (the face prolongation:
// fine faces on coarse lines take the coarse face, mid-line faces the mean
// of the two bracketing coarse faces; two fine faces along each coarse face)
void launch_st_mg_restrict(hipStream_t s, int ncx, int ncy, int comp, const double* r_f, double* b_c);
void launch_st_mg_prolong(hipStream_t s, int ncx, int ncy, int comp, const double* x_c, double* x_f);
// x = inv b with a dense n x n row-major inverse
void launch_st_dense_solve(hipStream_t s, int n, const double* inv, const double* b, double* x);
// x = A^-1 b with a symmetric circulant A on a periodic nx x ny grid, given
// the first column g of A^-1 (nx * ny values, row-major)
void launch_st_circulant_solve(hipStream_t s, int nx, int ny, const double* g, const double* b, double* x);
// fused passes with the operations of the unfused sequence: the zero-start
// sweep and the next one; the restriction of the residual b_f - M x_f; the
// prolongation (x_f + P x_c, not stored) and the first post-sweep into out
void launch_st_mg_smooth2z(hipStream_t s, int nx, int ny, const StagCoef& k, double w, const double* b,
                           double* x_out);
void launch_st_mg_restrict_res(hipStream_t s, int ncx, int ncy, int comp, const StagCoef& k, const double* b_f,
                               const double* x_f, double* b_c);
void launch_st_mg_prolong_smooth(hipStream_t s, int ncx, int ncy, int comp, const StagCoef& k, double w,
                                 const double* x_c, const double* x_f, const double* b, double* out);
// x += alpha p, r -= alpha Ap; partial maxima of |r| / V (no preconditioning)
void launch_st_cg_update_xr(hipStream_t s, int n, double inv_vol, double* x, double* r, const double* p,
                            const double* Ap, const double* scal, double* part);
// partial sums of r z
void launch_st_dot(hipStream_t s, int n, const double* r, const double* z, double* part);

}  // namespace tpamr
