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
  // irregular rows: the ELL part of the non-regular faces (off-diagonal
  // entries; their floor-bound values and the bound's diagonal part for the
  // FP64 finest level only) and the full diagonal per row, in ell_loc order
  const T* ell_val;
  const T* ell_bound;
  const T* ell_bdiag;
  const T* ell_diag;
  // irregular rows of block b (nr_b rows, ng = ceil(ell_w / 4) groups of four
  // entries per row): entry 4 g + t of row x at (blk_irr[b] * ng + g * nr_b + x) * 4 + t
  const unsigned short* ell_ref;
  const int* blk_irr;        // nblocks + 1 offsets of the irregular rows per block
  const int* blk_ng;         // per block: entry groups of its longest irregular row
  const unsigned* ell_loc;   // per row in a block's length order: local cell | length << 16
  const int* blk_ext;
  const int* ext_col;
  const int* nbr;
  const int* xfer;
  const double* volume;  // per block
};

// Device-resident Krylov state (Krylov.jl cg semantics). `ticket` counts
// finished workgroups of a two-stage reduction and is zero between kernels.
// The inner FP32 iteration of the refinement mode also keeps its FP32
// scaling (r32 = scale * r64), the largest strong residual since the last
// FP64 residual replacement (rlast), the replacement threshold (a
// replacement is requested once the strong residual falls to delta * rlast;
// delta = 0: never), the request flag and the number of replacements.
struct CgState {
  double gamma, pnorm2, alpha, beta, pAp, zmean, rnorm, strong, target, rmean;
  double scale, rlast, delta;
  int iter, itmax, done, reason, ticket, replace, replaced;
  int keep;  // refinement mode: a residual replacement keeps the recurrence (p, beta)
};

// Workgroups of the Krylov scalar-step reductions (stage buffer: 4 * RED_WG).
constexpr int RED_WG = 128;

// Cells per chunk of the component sums (as the host's plan).
constexpr int GAUGE_CHUNK_CELLS = 1024;

// A level's ungauged pressure components on the device (device pointers):
// slot per cell (0: gauged), component volumes, the chunked-sum plan of
// (chunk, slot) pairs and per slot its pairs in chunk order, and a partial
// per pair.
struct GaugeView {
  int ncomp = 0, npairs = 0;
  const int* slot = nullptr;
  const double* vol = nullptr;
  const int* pair_chunk = nullptr;
  const int* pair_slot = nullptr;
  const int* slot_pair_ptr = nullptr;
  const int* slot_pairs = nullptr;
  double* partial = nullptr;
};

// Krylov partial sums emitted by the V-cycle's last finest-level sweep:
// sums of V z, r z and r per block, with r the FP64 Krylov residual (r64) or
// the FP32 inner residual of the refinement loop (r32); part = nullptr: none.
struct ZStats {
  const double* r64 = nullptr;
  const float* r32 = nullptr;
  double* part = nullptr;
};

bool supported_block(int bx, int by);

template <typename T>
void launch_sweep(hipStream_t s, const LevelView<T>& L, const T* xi, T* xo, const T* b, double w, const int* done,
                  bool fine, const ZStats& zs, const T* x0_dinv = nullptr, double x0_w = 0.0);
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
// Coefficient refresh of one level (the host value pass, on the device):
// restricted face coefficients, regular conductances into the device layout
// (dst: 2 codes per face, index * 4 + array 0 cw / 1 ce / 2 cs / 3 cn, -1
// none), irregular rows from their face-metric terms (ratio: the largest row
// ratio as double bits), the Jacobi diagonal inverse (regular: set when a
// regular cell has a positive diagonal), a check of the coefficients (bad:
// set when one is not finite and positive)
void launch_ref_restrict(hipStream_t s, int nf, const int* rptr, const int* ridx, const double* rw,
                         const double* kprev, double* k);
void launch_ref_regular(hipStream_t s, int nf, const double* k, const double* kdiag, const double* area,
                        const int* dst, double* cw, double* ce, double* cs, double* cn);
// open boundary faces: bw into the cell's slot beyond the side
void launch_ref_boundary(hipStream_t s, int nopen, const double* bw, const int* dst, double* cw, double* ce,
                         double* cs, double* cn);
// irregular rows: full entries (val_e, diagonals, row ratios) and the ELL
// part of the non-regular faces (ell_val, ell_bound, ell_bdiag; bound arrays
// may be NULL); dst marks the regular faces; bd (NULL without open sides):
// per row the boundary diagonal added to its full diagonal
void launch_ref_rows(hipStream_t s, int nrows, const int* irr_ptr, const int* term_ptr, const int* term_f,
                     const int* term_g, const double* term_k, const double* term_s, const unsigned char* term_diag,
                     const double* k, const double* area, const int* dst, const int* ell_pos, const int* ell_dpos,
                     const int* ell_row, const int* diag_e, double* ell_val, double* ell_bound, double* ell_bdiag,
                     double* ell_diag, double* irr_diag, double* val_e, int nnz, unsigned long long* ratio,
                     const double* bd);
// nbr (NULL without open sides): regular cells with an open boundary slot add
// their row ratio to `ratio`
void launch_ref_dinv(hipStream_t s, int n, int bx, int by, const double* cw, const double* ce, const double* cs,
                     const double* cn, const int* cell_irr, const double* irr_diag, double* dinv, int* regular,
                     const int* nbr, unsigned long long* ratio);
// bad[0]: a negative or non-finite coefficient; bad[1]: a zero one
void launch_ref_check(hipStream_t s, int nf, const double* k, int* bad);
// component sums of x (mode 0: x, 1: V x, 2: |x|) into out[ncomp], divided
// by the component volumes with `divide`
template <typename T>
void launch_comp_sums(hipStream_t s, const GaugeView& G, int n, int nb_cells, const T* x, const double* volume,
                      int mode, int divide, double* out);
// x -= mean[slot] (mode 0) or x -= V mean[slot] (mode 1) on ungauged cells
template <typename T>
void launch_comp_shift(hipStream_t s, const GaugeView& G, int n, int nb_cells, const double* mean,
                       const double* volume, int mode, T* x);
// x = (w D^-1) b
template <typename T>
void launch_zero_sweep(hipStream_t s, int n, double omega, const T* dinv, const T* b, T* x);
void launch_sqrt(hipStream_t s, int n, const double* in, double* out);
// inverse of the dense N x N row-major matrix `a` (overwritten; Gauss-Jordan
// with partial pivoting, piv and col N-entry scratch); its top-left n x n
// block goes to `inv` column major; *fail is set on a zero or non-finite pivot
void launch_dense_inverse(hipStream_t s, int N, double* a, int* piv, double* col, int* fail, int n, double* inv);
// Blocked Gauss-Jordan pieces (panel of columns [c0, c0 + nb)): the panel's
// elimination with its pivots, then its row swaps on the other columns;
// R (nb x N) and M (N x nb) for the update of the other columns; the final
// column unpermutation and the n x n column-major block of the inverse.
void launch_gjb_panel(hipStream_t s, int N, int c0, int nb, double* a, int* piv, double* col, int* fail);
void launch_gjb_prep(hipStream_t s, int N, int c0, int nb, const double* a, double* R, double* M);
void launch_gj_finish(hipStream_t s, int N, double* a, const int* piv, int n, double* inv);
// *out = max(*out, max_ij |I - T|_ij) as the bits of a nonnegative double
void launch_identity_residual(hipStream_t s, int N, const double* T, unsigned long long* out);
// top-left n x n block of the row-major N x N `a`, column major
void launch_extract_block(hipStream_t s, int n, int N, const double* a, double* inv);

// The V-cycle of the small levels in one workgroup (k_vtail): levels
// 0..n-1 of the tail (the last one the coarsest), their vectors and weights.
// phase 0: the whole tail cycle; 1: the way down to the coarsest level's
// rhs; 2: the way up from the coarsest level's solution (a coarsest level of
// more than COARSE_INLINE cells is solved between the two by the
// many-workgroup coarse kernel).
constexpr int TAIL_MAX = 8;
constexpr int COARSE_INLINE = 256;
template <typename T>
struct TailArgs {
  LevelView<T> L[TAIL_MAX];
  T* x[TAIL_MAX];
  T* xs[TAIL_MAX];
  T* b[TAIL_MAX];
  T* u[TAIL_MAX];
  const T* dinv[TAIL_MAX];
  double omega[TAIL_MAX], omega_p[TAIL_MAX];
  const double* inv;
  int n, pre, post, coarse_n, phase;
};
template <typename T>
void launch_vtail(hipStream_t s, const TailArgs<T>& A, const int* done);

// y = A p with per-block partials of p.y (part[b]) and, with `sums`, of the
// sum of y (part[nblocks + b])
template <typename T>
void launch_matvec_dot(hipStream_t s, const LevelView<T>& L, const T* p, T* y, double* part, const int* done,
                       bool sums = false);
void launch_floor_bound(hipStream_t s, const LevelView<double>& L, const double* x, const double* b, double* part);
void launch_reduce(hipStream_t s, int n, int nsets, const double* part, double* out, bool is_max);
void launch_cg_alpha(hipStream_t s, int nb, const double* part, double* stage, CgState* st, bool sums = false,
                     double total_volume = 1.0);
// vx = NULL: the zero-start iterate x0 = w D^-1 r is not stored (the
// V-cycle's first sweep forms it)
template <typename TV>
void launch_cg_update_xr(hipStream_t s, int n, int nbc, const double* p, const double* Ap, double* x, double* r,
                         const double* vol, TV* vb, TV* vx, const TV* dinv, double omega, double* part, CgState* st,
                         int init);
void launch_cg_beta(hipStream_t s, int nb, int nmax, const double* part, const double* maxpart, double* stage,
                    double total_volume, CgState* st, int init);
template <typename TV>
void launch_cg_update_p(hipStream_t s, int n, const TV* z, double* p, const CgState* st, const int* done);

// Inner FP32 Krylov iteration of the mixed-precision refinement loop. Four
// cells per thread; the strong-residual maxima go to `part`, one per
// workgroup of ir_groups(n); vx = NULL: x0 is not stored.
int ir_groups(int n);
// x64 (not NULL): when a residual replacement is requested, the correction
// e moves into this FP64 solution (x += e / scale, e = 0) in the same pass
void launch_ir_update_xr(hipStream_t s, int n, int nbc, const float* p, const float* Ap, float* e, float* r,
                         const double* vol, float* vx, const float* dinv, double omega, const double* r64,
                         double* part, CgState* st, int init, double* x64 = nullptr);
void launch_ir_update_p(hipStream_t s, int n, const float* z, float* p, const CgState* st, const int* done);
// x += e / scale, e = 0: when a residual replacement is requested, or always
// with force
void launch_ir_correct(hipStream_t s, int n, float* e, double* x, const CgState* st, bool force);
// FP64 residual replacement (when requested): r32 = scale (b - A x) with the
// FP64 finest-level operator L64, the V-cycle's zero-start iterate and the
// strong-residual maxima (one per 1024 cells, like launch_ir_update_xr)
void launch_ir_replace(hipStream_t s, const LevelView<double>& L64, const double* x, const double* b, float* r32,
                       float* x0, const float* dinv, double omega, double* part, CgState* st);
// stop test on a replaced residual: done when its strong maximum (ng
// partials) meets the target
void launch_ir_check(hipStream_t s, int ng, const double* maxpart, double* stage, CgState* st);
void launch_fill_f32(hipStream_t s, int n, float v, float* out);

void launch_residual_strong(hipStream_t s, int n, int nbc, const double* b, const double* y, double* r,
                            const double* vol, double* part);
void launch_weighted_sums(hipStream_t s, int n, int nbc, const double* x, const double* vol, double* part);
// haloed cell field: interior copy and the ne halo entries of a schedule
void launch_fill_halo(hipStream_t s, int n, int bx, int by, int h, int ne, const int* dest, const int* ptr,
                      const int* donor, const double* weight, const int* phys, const double* cells,
                      const double* physical, double* out);
// b <- V b and partial sums of |b|, b, V x, |x| (four sets of ceil(n/256))
void launch_rhs_prepare(hipStream_t s, int n, int nbc, double* b, const double* x, const double* vol, double* part);
// b -= V mean, r = b, partial strong maxima and squared-norm sums of r
void launch_shift_residual(hipStream_t s, int n, int nbc, double* b, double mean, const double* vol, double* r,
                           double* part);
void launch_shift(hipStream_t s, int n, int nbc, double* x, const double* shift, double scale, const double* vol,
                  bool by_volume);
void launch_scale_by_volume(hipStream_t s, int n, int nbc, const double* in, double* out, const double* vol,
                            bool divide);
void launch_metric_apply(hipStream_t s, int nf, const double* kd, const int* optr, const int* oface,
                         const double* oval, const double* in, double* out);
// out = sqrt(k) |f| (p+ - p-) (p given) or sqrt(k) jumps; sqk = nullptr: unit weights
void launch_face_jump(hipStream_t s, int nf, const int* fm, const int* fp, const double* area, const double* sqk,
                      const double* p, const double* jumps, double* out);
// gravity potential of n cells (phi) and the nf face jumps rho_f |f| (phi+ - phi-)
void launch_gravity_jumps(hipStream_t s, int n, int bx, int nbc, const double* origin, const double* spacing,
                          const double* offset, double gx, double gy, double* phi, int nf, const int* fm,
                          const int* fp, const double* area, const double* liquid, const double* gas, double rl,
                          double rg, double* jumps);
void launch_face_scale(hipStream_t s, int nf, const double* sqk, const double* in, double* out);
void launch_mixture(hipStream_t s, int nf, int nbc, const int* fm, const int* fp, const signed char* axis,
                    const double* spacing, const double* c, double rl, double rg, double* kappa);
void launch_surface_jump(hipStream_t s, int nf, const int* fm, const int* fp, const double* area, const double* c,
                         const double* curv, double sigma, int csf, double* jumps);
// CGFM / CSF jumps with the face curvature of `_face_curvature` from cell
// curvatures (NaN = none) and their height-function coverage; counts the
// faces left without a curvature
void launch_surface_jump_cells(hipStream_t s, int nf, const int* fm, const int* fp, const double* area,
                               const double* origin, const double* spacing, const double* fcx, const double* fcy,
                               int nblocks, int bx, int by, const double extent[2], const bool periodic[2],
                               const double* c, const double* kappa, const double* quality, double sigma, int csf,
                               double* jumps, int* unresolved);
void launch_divergence(hipStream_t s, int n, int nbc, const int* cptr, const int* cface, const signed char* csign,
                       const double* area, const double* u, const int* bptr, const int* bface,
                       const signed char* bside, const double* barea, const double* ub, const double* vol,
                       double* div);
// g = scale [yl (K yp) + yp (K yl)] / (2 kappa) per face (g may alias kyl)
void launch_adjoint_face(hipStream_t s, int nf, const double* yl, const double* kyp, const double* yp,
                         const double* kyl, const double* kappa, double scale, double* g);
template <typename T>
void launch_zstats(hipStream_t s, int nblocks, int bx, const T* z, const ZStats& zs, const double* vol,
                   const int* done);
// projection rhs -div / dt (+ w inherited; inherited may be nullptr)
void launch_projection_rhs(hipStream_t s, int n, const double* div, const double* inherited, double dt, double w,
                           double* rhs);
// y = a x + b y (b = 0: y not read)
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y);
// ub -= dt (-side bgrad) phi_cell on every boundary face
void launch_boundary_velocity(hipStream_t s, int nbf, const int* bcell, const signed char* bside, const double* bgrad,
                              const double* phi, double dt, double* ub);
void launch_cast_d2f(hipStream_t s, int n, const double* in, float* out);
void launch_cast_f2d(hipStream_t s, int n, const float* in, double* out);
void launch_copy(hipStream_t s, int n, const double* in, double* out);
void launch_fill(hipStream_t s, int n, double v, double* out);
void launch_fill_int(hipStream_t s, int n, int v, int* out);

// out = sum of `reads` (1 or 5) consecutive input arrays of n4 float4 each
void launch_stream(hipStream_t s, int reads, size_t n4, const float* in, float* out);

// Device scratch of the allocation-free solve entry point: a file-scope
// array of this many bytes on every GPU the library runs on, held until the
// process ends.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(8) << 20;
// Its address on `device`, which must be the current device: looked up once
// per device with hipGetSymbolAddress; returns that call's status.
hipError_t noalloc_scratch_address(int device, void** ptr);

}  // namespace tpamr
