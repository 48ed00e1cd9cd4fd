// CMDGFM reduced-pressure/jump block solve on the device (cmdgfm.cpp,
// cmdgfm.hip): BiCGSTAB on an assembled two-velocity projection system whose
// first np unknowns are reduced pressures (cell minus its component's
// anchor) and whose last nx unknowns are mixed-face gradient jumps, right
// preconditioned by the block preconditioner (one V-cycle of the shared
// pressure hierarchy for the pressure block, the Crout ILU of the local jump
// Schur complement J - C D^-1 B for the jumps), inside the reference's
// residual-replacement loop.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

namespace tpamr {

// A HIP runtime failure inside the block solve.
struct CmdgfmDeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Threads per workgroup of the vector and reduction kernels, their largest
// grid, the one workgroup of the triangular solves, the retained entries one
// workgroup sums per anchor chunk, and the rows of one sliced-ELL slice.
constexpr int CM_WG = 256;
constexpr int CM_SLICE = 64;
// The jump factor's solves keep the iterate and the level schedules in LDS
// while they fit in CM_ILU_LDS bytes, and prefetch CM_ILU_PF entries of each
// row into registers.
constexpr int CM_ILU_LDS = 65536;
constexpr int CM_ILU_PF = 8;
constexpr int CM_MAX_WG = 1024;
constexpr int CM_ILU_WG = 1024;
constexpr int CM_ANCHOR_CHUNK = 4096;

// BiCGSTAB scalars on the device: scal[CM_RHO] ... (cmdgfm.hip)
enum CmScalar {
  CM_RHO = 0,
  CM_NEXT_RHO = 1,
  CM_ALPHA = 2,
  CM_OMEGA = 3,
  CM_BETA = 4,
  CM_RNORM2 = 5,
  CM_SCALARS = 8
};
// steps of the final reduction (launch_cm_scalars)
enum CmStep { CM_STEP_INIT = 0, CM_STEP_ALPHA = 1, CM_STEP_OMEGA = 2, CM_STEP_RHO = 3 };

int cm_parts(int n);

// Sliced ELL (CM_SLICE rows per slice; slice k's entries start at sptr[k],
// entry e of its row r at sptr[k] + e CM_SLICE + r; len[i] entries of row i
// in CSR order). y = rows (A (cols x)) with the per-block scalings: rows pr
// (pressure, i < np) or jr, columns pc (j < np) or jc.
void launch_cm_spmv(hipStream_t s, int n, int np, const int* sptr, const int* len, const int* col,
                    const double* val, const double* x, double pc, double jc, double pr, double jr, double* y);
// r = b - A x
void launch_cm_residual(hipStream_t s, int n, const int* sptr, const int* len, const int* col, const double* val,
                        const double* x, const double* b, double* r);
// per-workgroup partials of nd (1 or 2) dot products a_k . b_k
void launch_cm_dots(hipStream_t s, int n, int nd, const double* a0, const double* b0, const double* a1,
                    const double* b1, double* part);
// fixed-order final sum of the partials and the BiCGSTAB scalar step
void launch_cm_scalars(hipStream_t s, int nparts, int nd, const double* part, double* scal, int step);
// s = r - alpha v; x = x + alpha y
void launch_cm_update_s(hipStream_t s, int n, const double* r, const double* v, const double* y, double* sv,
                        double* x, const double* scal);
// x = x + omega z; r = s - omega t
void launch_cm_update_r(hipStream_t s, int n, const double* sv, const double* t, const double* z, double* r,
                        double* x, const double* scal);
// p = r + beta (p - omega v)
void launch_cm_update_p(hipStream_t s, int n, const double* r, const double* v, double* p, const double* scal);
// out = rows v (pr before np, jr after); v = v / sc
void launch_cm_scale_rows(hipStream_t s, int n, int np, const double* v, double pr, double jr, double* out);
void launch_cm_divs(hipStream_t s, int n, double* v, double sc);
// full-cell rhs of the reduced-pressure block from v / d (d the pressure
// row scaling): retained cells their value (role = local index), other
// cells zero; then per-chunk sums of the retained values in anchor order and
// each anchor's negated total (one workgroup per anchor)
void launch_cm_scatter(hipStream_t s, int ncells, const int* role, const double* v, double d, double* full);
void launch_cm_anchor_chunks(hipStream_t s, int nchunks, const int* chunk_ptr, const int* list, const double* v,
                             double d, double* part);
void launch_cm_anchor_totals(hipStream_t s, int nanchors, const int* anchor_cell, const int* anchor_chunk_ptr,
                             const double* part, double* full);
// out_i = (z[idx_i] - z[anchor_i]) / d
void launch_cm_gather(hipStream_t s, int np, const int* idx, const int* anchor, const double* z, double d,
                      double* out);
// jr = -(C zp) + rhs_j / d over the nx jump rows of C (ptr null: no C);
// with zfull, zp_q = zfull[idx_q] - zfull[anchor_q]
void launch_cm_jump_rhs(hipStream_t s, int nx, const int* ptr, const int* col, const double* val, const double* zp,
                        const double* zfull, const int* idx, const int* anchor, const double* rhs, double d,
                        double* jr);
// x = U^-1 L^-1 b by level-scheduled rows in one workgroup, then x /= d
void launch_cm_ilu(hipStream_t s, int nx, const int* lptr, const int* lcol, const double* lval, int nllev,
                   const int* llev, const int* lorder, const int* uptr, const int* ucol, const double* uval,
                   const double* udiag, int nulev, const int* ulev, const int* uorder, const double* b, double* y,
                   double* x, double d);
// the same with the iterate and level schedules in LDS (8 nx + 4 (nllev +
// nulev + 2) <= CM_ILU_LDS bytes) and the rows packed in level order at width
// w <= CM_ILU_PF (lrows / urows: row of each packed position, llen / ulen:
// its entries, entry k of position t at t w + k; udiag by packed position);
// b_t: nx doubles of scratch for the rhs in packed order
void launch_cm_ilu_lds(hipStream_t s, int nx, int w, int nllev, const int* llev, const int* lrows, const int* llen,
                       const int* lcol, const double* lval, int nulev, const int* ulev, const int* urows,
                       const int* ulen, const int* ucol, const double* uval, const double* udiag, const double* b,
                       double* b_t, double* x, double d);
// per-workgroup partials of the defects of a residual: [0] max |r_i| / vol_i
// over the pressure rows with vol > 0, [1] max |r_k| / len_k over the jump
// rows, [2] sum r^2, [3] sum (rows r)^2 (rows pr / jr), [4] 1 where x is not
// finite, [5] max |r_i|
void launch_cm_check(hipStream_t s, int n, int np, const double* r, const double* vol, const double* len, double pr,
                     double jr, const double* x, double* part);
// [0] max |rhs_a - (O x)_a| / vol_a over the omitted rows with vol > 0
void launch_cm_omitted(hipStream_t s, int m, const int* ptr, const int* col, const double* val, const double* x,
                       const double* rhs, const double* vol, double* part);
// final reduction: out[k] = sum or maximum (bit k of max_mask, NaN kept)
void launch_cm_reduce(hipStream_t s, int nparts, int nsets, int max_mask, const double* part, double* out);
// x += cols (scale inc) (cols pc before np, jc after); *flag = 1 when an
// entry changed
void launch_cm_increment(hipStream_t s, int n, int np, double* x, double pc, double jc, double scale,
                         const double* inc, int* flag);

// LinearSolverStats-like summary of one block solve (the solve fields of
// CMDGFMProjectionStats).
struct CmdgfmStats {
  int converged = 0, iterations = 0, krylov_runs = 0, initial_guess_used = 0;
  int pressure_unknowns = 0, jump_unknowns = 0;
  double initial_residual = 0, final_residual = 0;
  double pressure_defect = 0, jump_defect = 0, anchor_defect = 0;
  double asymmetry = 0, jump_diagonal_margin = 0;
};

// Host factor of the jump block: the Schur complement and its Crout ILU
// (strict L by rows, U by rows with its diagonal apart), with the level
// schedules of both triangular solves.
struct JumpFactor {
  int nx = 0;
  std::vector<int> sptr, scol;
  std::vector<double> sval;  // J - C D^-1 B
  std::vector<int> lptr, lcol, uptr, ucol;
  std::vector<double> lval, uval, udiag;
  std::vector<int> llev, lorder, ulev, uorder;
  // rows packed in level order at width pw (the longest row; 0 when it
  // exceeds CM_ILU_PF) for the LDS solves: lengths, entries of the forward
  // (L) and backward (U off the diagonal) rows, U's diagonal by position
  int pw = 0;
  std::vector<int> flen, fcol, blen, bcol;
  std::vector<double> fval, bval, bdiag;
};

// J - C D^-1 B of an assembled CSR system (rows sorted by column), D the
// diagonal of its first np x np block.
void cmdgfm_jump_schur(int n, int np, const int* ptr, const int* col, const double* val, JumpFactor& f);
// IncompleteLU.jl 0.2's Crout ILU of f's Schur complement with drop
// tolerance tau, then the level schedules.
void cmdgfm_crout_ilu(double tau, JumpFactor& f);

class CmdgfmSystem {
 public:
  // Enqueues z = V-cycle(r) of the shared pressure hierarchy on a full-cell
  // device vector (integrated r; the V-cycle may rewrite r with its own
  // values) and returns z's device address (full-cell FP64, valid until the
  // next call; defined up to a constant per pressure component).
  using Vcycle = std::function<const double*(double* r)>;
  // The V-cycle's own full-cell input buffer, when the rhs may be written
  // there directly (null: the system's own buffer, copied in by the V-cycle).
  using VcycleInput = std::function<double*()>;

  // Checks and factors the assembled system on the host; with a stream,
  // uploads it. Index arrays are 0-based; CSR column indices sorted per row.
  CmdgfmSystem(int ncells, int np, int nx, const int* ptr, const int* col, const double* val,
               const int* pressure_cells, const int* anchor_cells, const double* pressure_volumes,
               const double* jump_lengths, int nomitted, const int* optr, const int* ocol, const double* oval,
               const double* orhs, const double* ovol, double drop_tolerance, bool device, hipStream_t stream);
  ~CmdgfmSystem();
  CmdgfmSystem(const CmdgfmSystem&) = delete;
  CmdgfmSystem& operator=(const CmdgfmSystem&) = delete;

  int size() const;
  int pressure_size() const;
  int jump_size() const;
  const JumpFactor& factor() const;

  // The solve of _project_cmdgfm_assembled! on device vectors: rhs (n),
  // solution (n, overwritten), previous full-cell increment (ncells) or null.
  void solve(const double* rhs, double* solution, double dt, double density, double reference_length,
             double tolerance, int max_iterations, bool warm_start, const double* previous, bool preconditioned,
             const Vcycle& vcycle, const VcycleInput& input, CmdgfmStats& stats);

 private:
  struct Impl;
  std::unique_ptr<Impl> d_;
};

}  // namespace tpamr
