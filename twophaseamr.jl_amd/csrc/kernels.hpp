// Device-side views and kernel launchers (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace tpamr {

// Longest irregular row the kernels accept (entries of one row).
constexpr int ELL_MAX = 12;

// One multigrid level as the kernels see it (device pointers).
template <typename T>
struct LevelView {
  int nblocks, ncells, bx, by, ell_w;
  // tile-kernel grid: wg_grid workgroups (a multiple of 8 * wg_chunk); runs of
  // wg_chunk consecutive block groups are dealt to the 8 XCDs in turn
  int wg_chunk, wg_grid;
  int ext_cap;  // external-cell slots per block tile (most of any block of the level)
  // conductances: west and south face of every cell (block-major, like the
  // cells), east face of each block's last column, north face of its top row
  const T* cw;
  const T* ce;
  const T* cs;
  const T* cn;
  const T* ell_val;
  const T* ell_bound;  // floor-bound values (FP64 finest level only)
  const T* ell_diag;  // diagonal per irregular row, in ell_loc order
  const unsigned short* ell_ref;  // entry k of row x of block b at blk_irr[b] * ell_w + k * nrows_b + x
  const int* blk_irr;        // nblocks + 1 offsets of the irregular rows per block
  const unsigned* ell_loc;   // per row in a block's length order: local cell | length << 16
  const int* blk_ext;
  const int* ext_col;
  const int* nbr;
  const int* xfer;
  const double* volume;  // per block
};

// Device-resident Krylov state (Krylov.jl cg semantics). `ticket` counts
// finished workgroups of a two-stage reduction and is zero between kernels.
struct CgState {
  double gamma, pnorm2, alpha, beta, pAp, zmean, rnorm, strong, target, rmean;
  int iter, itmax, done, reason, ticket;
};

// Workgroups of the Krylov scalar-step reductions (stage buffer: 4 * RED_WG).
constexpr int RED_WG = 128;

bool supported_block(int bx, int by);

template <typename T>
void launch_sweep(hipStream_t s, const LevelView<T>& L, const T* xi, T* xo, const T* b, double w, const int* done,
                  bool fine, const double* r64, double* part);
template <typename T>
void launch_residual(hipStream_t s, const LevelView<T>& L, const T* x, const T* b, T* u, double wp, const int* done,
                     bool fine);
template <typename T>
void launch_restrict(hipStream_t s, const LevelView<T>& L, const T* u, double wp, T* bc, T* xc, const T* dinv_c,
                     double wc, int zero_start, const int* done, bool fine);
template <typename T>
void launch_prolong(hipStream_t s, const LevelView<T>& L, T* x, const T* xc, double wp, const int* done, bool fine);
template <typename T>
void launch_coarse(hipStream_t s, int n, const double* inv, const T* b, T* x, const int* done);

void launch_matvec_dot(hipStream_t s, const LevelView<double>& L, const double* p, double* y, double* part,
                       const int* done);
void launch_floor_bound(hipStream_t s, const LevelView<double>& L, const double* x, const double* b, double* part);
void launch_reduce(hipStream_t s, int n, int nsets, const double* part, double* out, bool is_max);
void launch_cg_alpha(hipStream_t s, int nb, const double* part, double* stage, CgState* st);
template <typename TV>
void launch_cg_update_xr(hipStream_t s, int n, int nbc, const double* p, const double* Ap, double* x, double* r,
                         const double* vol, TV* vb, TV* vx, const TV* dinv, double omega, double* part, CgState* st,
                         int init);
void launch_cg_beta(hipStream_t s, int nb, int nmax, const double* part, const double* maxpart, double* stage,
                    double total_volume, CgState* st, int init);
template <typename TV>
void launch_cg_update_p(hipStream_t s, int n, const TV* z, double* p, const CgState* st, const int* done);

void launch_residual_strong(hipStream_t s, int n, int nbc, const double* b, const double* y, double* r,
                            const double* vol, double* part);
void launch_weighted_sums(hipStream_t s, int n, int nbc, const double* x, const double* vol, double* part);
void launch_shift(hipStream_t s, int n, int nbc, double* x, const double* shift, double scale, const double* vol,
                  bool by_volume);
void launch_scale_by_volume(hipStream_t s, int n, int nbc, const double* in, double* out, const double* vol,
                            bool divide);
void launch_metric_apply(hipStream_t s, int nf, const double* kd, const int* optr, const int* oface,
                         const double* oval, const double* in, double* out);
void launch_face_jump(hipStream_t s, int nf, const int* fm, const int* fp, const double* area, const double* sqk,
                      const double* p, const double* jumps, double* out);
void launch_face_scale(hipStream_t s, int nf, const double* sqk, const double* in, double* out);
void launch_mixture(hipStream_t s, int nf, int nbc, const int* fm, const int* fp, const signed char* axis,
                    const double* spacing, const double* c, double rl, double rg, double* kappa);
void launch_surface_jump(hipStream_t s, int nf, const int* fm, const int* fp, const double* area, const double* c,
                         const double* curv, double sigma, int csf, double* jumps);
void launch_divergence(hipStream_t s, int n, int nbc, const int* cptr, const int* cface, const signed char* csign,
                       const double* area, const double* u, const int* bptr, const int* bface,
                       const signed char* bside, const double* barea, const double* ub, const double* vol,
                       double* div);
void launch_adjoint_face(hipStream_t s, int nf, const double* yl, const double* kyp, const double* yp,
                         const double* kyl, const double* kappa, double* g);
template <typename T>
void launch_zstats(hipStream_t s, int nblocks, int bx, const T* z, const double* r64, const double* vol, double* part,
                   const int* done);
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y);
void launch_cast_d2f(hipStream_t s, int n, const double* in, float* out);
void launch_cast_f2d(hipStream_t s, int n, const float* in, double* out);
void launch_copy(hipStream_t s, int n, const double* in, double* out);
void launch_fill(hipStream_t s, int n, double v, double* out);
void launch_fill_int(hipStream_t s, int n, int v, int* out);

}  // namespace tpamr
