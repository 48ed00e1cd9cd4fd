// Host side of the CMDGFM reduced-pressure/jump block solve: checks and
// device layout of the assembled system, the jump Schur complement and its
// Crout ILU with level schedules, and the drivers of BiCGSTAB (Krylov.jl
// bicgstab! with a right preconditioner), the block preconditioner and the
// residual-replacement loop of _project_cmdgfm_assembled!
// (src/two_phase/projection/cmdgfm_projection.jl:2201-2318).
#include "cmdgfm.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>

namespace tpamr {
namespace {

#define CCHECK(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) throw CmdgfmDeviceError(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
struct Buf {
  T* p = nullptr;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (n) CCHECK(hipMalloc(&p, n * sizeof(T)));
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) CCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

const double EPS = std::numeric_limits<double>::epsilon();

// Julia's max (NaN wins)
double jmax(double a, double b) {
  if (std::isnan(a)) return a;
  if (std::isnan(b)) return b;
  return std::max(a, b);
}

void check_csr(const char* what, int rows, int cols, const int* ptr, const int* col) {
  if (!ptr || ptr[0] != 0) throw std::invalid_argument(std::string(what) + ": row pointers must start at 0");
  for (int i = 0; i < rows; ++i) {
    if (ptr[i + 1] < ptr[i]) throw std::invalid_argument(std::string(what) + ": decreasing row pointers");
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
      if (col[p] < 0 || col[p] >= cols) throw std::invalid_argument(std::string(what) + ": column out of range");
      if (p > ptr[i] && col[p] <= col[p - 1])
        throw std::invalid_argument(std::string(what) + ": columns must increase within a row");
    }
  }
}

// rows of the level schedule: level of every row, then rows by level
void level_schedule(const std::vector<int>& lev, std::vector<int>& ptr, std::vector<int>& order) {
  const int n = int(lev.size());
  int nl = 0;
  for (int v : lev) nl = std::max(nl, v + 1);
  ptr.assign(nl + 1, 0);
  for (int v : lev) ++ptr[v + 1];
  for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
  order.assign(n, 0);
  std::vector<int> cur(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < n; ++i) order[cur[lev[i]]++] = i;
}

}  // namespace

// J - C D^-1 B in the order of _cmdgfm_jump_schur! (cmdgfm_cache.jl:550-610):
// per jump row, the J entries, then for each C entry (by pressure column)
// the row of D^-1 B (by jump column), duplicates summed in that order.
void cmdgfm_jump_schur(int n, int np, const int* ptr, const int* col, const double* val, JumpFactor& f) {
  const int nx = n - np;
  f.nx = nx;
  std::vector<double> d(np, 0.0);
  for (int q = 0; q < np; ++q)
    for (int p = ptr[q]; p < ptr[q + 1]; ++p)
      if (col[p] == q) d[q] = val[p];
  for (int q = 0; q < np; ++q)
    if (d[q] == 0.0) throw std::invalid_argument("zero diagonal in the reduced pressure block");
  f.sptr.assign(1, 0);
  f.scol.clear();
  f.sval.clear();
  std::map<int, double> row;
  for (int k = 0; k < nx; ++k) {
    row.clear();
    const int r = np + k;
    for (int p = ptr[r]; p < ptr[r + 1]; ++p)
      if (col[p] >= np) row[col[p] - np] = val[p];
    for (int p = ptr[r]; p < ptr[r + 1]; ++p) {
      const int q = col[p];
      if (q >= np) break;
      const double c = val[p];
      for (int e = ptr[q]; e < ptr[q + 1]; ++e) {
        if (col[e] < np) continue;
        const double v = -c * (val[e] / d[q]);
        auto it = row.find(col[e] - np);
        if (it == row.end())
          row.emplace(col[e] - np, v);
        else
          it->second = it->second + v;
      }
    }
    for (const auto& kv : row) {
      f.scol.push_back(kv.first);
      f.sval.push_back(kv.second);
    }
    f.sptr.push_back(int(f.scol.size()));
  }
}

// IncompleteLU.jl 0.2 ilu(S; tau): row k of U and column k of L from the
// stored factors (U_row -= L[k,i] U[i,:] over i < k ascending, L_col -=
// U[i,k] L[:,i] likewise), kept when |value| >= tau (the diagonal of U
// always; L entries compared before the division by u_kk).
void cmdgfm_crout_ilu(double tau, JumpFactor& f) {
  const int n = f.nx;
  // S by columns (rows ascending)
  std::vector<int> cptr(n + 1, 0), crow(f.scol.size());
  std::vector<double> cval(f.scol.size());
  for (int c : f.scol) ++cptr[c + 1];
  for (int c = 0; c < n; ++c) cptr[c + 1] += cptr[c];
  {
    std::vector<int> cur(cptr.begin(), cptr.end() - 1);
    for (int r = 0; r < n; ++r)
      for (int p = f.sptr[r]; p < f.sptr[r + 1]; ++p) {
        const int q = cur[f.scol[p]]++;
        crow[q] = r;
        cval[q] = f.sval[p];
      }
  }
  std::vector<std::vector<std::pair<int, double>>> urow(n), ucolv(n), lcolv(n), lrowv(n);
  std::vector<double> uacc(n, 0.0), lacc(n, 0.0);
  std::vector<char> uon(n, 0), lon(n, 0);
  std::vector<int> ulist, llist;
  f.udiag.assign(n, 0.0);
  for (int k = 0; k < n; ++k) {
    ulist.clear();
    llist.clear();
    for (int p = f.sptr[k]; p < f.sptr[k + 1]; ++p) {
      const int j = f.scol[p];
      if (j < k) continue;
      if (!uon[j]) {
        uon[j] = 1;
        uacc[j] = 0.0;
        ulist.push_back(j);
      }
      uacc[j] = uacc[j] + f.sval[p];
    }
    for (int p = cptr[k]; p < cptr[k + 1]; ++p) {
      const int i = crow[p];
      if (i <= k) continue;
      if (!lon[i]) {
        lon[i] = 1;
        lacc[i] = 0.0;
        llist.push_back(i);
      }
      lacc[i] = lacc[i] + cval[p];
    }
    for (const auto& li : lrowv[k]) {
      const double lki = li.second;
      for (const auto& uj : urow[li.first]) {
        const int j = uj.first;
        if (j < k) continue;
        if (!uon[j]) {
          uon[j] = 1;
          uacc[j] = 0.0;
          ulist.push_back(j);
        }
        uacc[j] = uacc[j] - lki * uj.second;
      }
    }
    if (k < n - 1)
      for (const auto& ui : ucolv[k]) {
        const double uik = ui.second;
        for (const auto& lr : lcolv[ui.first]) {
          const int r = lr.first;
          if (r <= k) continue;
          if (!lon[r]) {
            lon[r] = 1;
            lacc[r] = 0.0;
            llist.push_back(r);
          }
          lacc[r] = lacc[r] - uik * lr.second;
        }
      }
    const double ukk = uon[k] ? uacc[k] : 0.0;
    if (ukk == 0.0) throw std::invalid_argument("zero pivot in the jump factor");
    std::sort(ulist.begin(), ulist.end());
    std::sort(llist.begin(), llist.end());
    for (int j : ulist) {
      const double v = uacc[j];
      uon[j] = 0;
      if (std::fabs(v) >= tau || j == k) {
        urow[k].emplace_back(j, v);
        if (j > k) ucolv[j].emplace_back(k, v);
      }
    }
    const double inv = 1.0 / ukk;
    for (int r : llist) {
      const double v = lacc[r];
      lon[r] = 0;
      if (std::fabs(v) >= tau) {
        lcolv[k].emplace_back(r, inv * v);
        lrowv[r].emplace_back(k, inv * v);
      }
    }
  }
  f.lptr.assign(1, 0);
  f.lcol.clear();
  f.lval.clear();
  f.uptr.assign(1, 0);
  f.ucol.clear();
  f.uval.clear();
  for (int i = 0; i < n; ++i) {
    for (const auto& e : lrowv[i]) {
      f.lcol.push_back(e.first);
      f.lval.push_back(e.second);
    }
    f.lptr.push_back(int(f.lcol.size()));
    for (const auto& e : urow[i]) {
      if (e.first == i) {
        f.udiag[i] = e.second;
      } else {
        f.ucol.push_back(e.first);
        f.uval.push_back(e.second);
      }
    }
    f.uptr.push_back(int(f.ucol.size()));
  }
  // level schedules: forward by L's dependencies, backward by U's
  std::vector<int> lev(n, 0);
  for (int i = 0; i < n; ++i)
    for (int p = f.lptr[i]; p < f.lptr[i + 1]; ++p) lev[i] = std::max(lev[i], lev[f.lcol[p]] + 1);
  level_schedule(lev, f.llev, f.lorder);
  lev.assign(n, 0);
  for (int i = n - 1; i >= 0; --i)
    for (int p = f.uptr[i]; p < f.uptr[i + 1]; ++p) lev[i] = std::max(lev[i], lev[f.ucol[p]] + 1);
  level_schedule(lev, f.ulev, f.uorder);
  // rows packed in level order at the width of the longest row
  int w = 0;
  for (int i = 0; i < n; ++i) w = std::max({w, f.lptr[i + 1] - f.lptr[i], f.uptr[i + 1] - f.uptr[i]});
  f.pw = w <= CM_ILU_PF ? std::max(w, 1) : 0;
  f.flen.clear();
  f.fcol.clear();
  f.fval.clear();
  f.blen.clear();
  f.bcol.clear();
  f.bval.clear();
  f.bdiag.clear();
  if (f.pw == 0) return;
  f.fcol.assign(size_t(n) * f.pw, 0);
  f.fval.assign(size_t(n) * f.pw, 0.0);
  f.bcol.assign(size_t(n) * f.pw, 0);
  f.bval.assign(size_t(n) * f.pw, 0.0);
  for (int t = 0; t < n; ++t) {
    const int i = f.lorder[t];
    f.flen.push_back(f.lptr[i + 1] - f.lptr[i]);
    for (int p = f.lptr[i]; p < f.lptr[i + 1]; ++p) {
      f.fcol[size_t(t) * f.pw + (p - f.lptr[i])] = f.lcol[p];
      f.fval[size_t(t) * f.pw + (p - f.lptr[i])] = f.lval[p];
    }
    const int j = f.uorder[t];
    f.blen.push_back(f.uptr[j + 1] - f.uptr[j]);
    for (int p = f.uptr[j]; p < f.uptr[j + 1]; ++p) {
      f.bcol[size_t(t) * f.pw + (p - f.uptr[j])] = f.ucol[p];
      f.bval[size_t(t) * f.pw + (p - f.uptr[j])] = f.uval[p];
    }
    f.bdiag.push_back(f.udiag[j]);
  }
}

struct CmdgfmSystem::Impl {
  hipStream_t s = nullptr;
  bool device = false;
  int ncells = 0, n = 0, np = 0, nx = 0, nomit = 0;
  int nanchors = 0, nchunks = 0;
  JumpFactor jf;
  double asym = 0, margin = 0;
  // device copies: the system as sliced ELL, its C block as CSR
  Buf<int> sptr, slen, scol, cptr, ccol, role, idx, anc, list, chunk_ptr, anchor_cell, anchor_chunk_ptr, optr, ocol;
  Buf<double> sval, cval, vol, len, oval, orhs, ovol;
  Buf<int> lptr, lcol, llev, lorder, uptr, ucol, ulev, uorder, flen, fcol, blen, bcol;
  Buf<double> lval, uval, udiag, fval, bval, bdiag, jbt;
  // work vectors
  Buf<double> corr, srhs, inc, kr, kp, kv, ks, kt, kyz, cand, res, full, jr, jy;
  Buf<double> part, apart, scal, out;
  Buf<int> flag;
  double host_out[8] = {};
  double host_scal[CM_SCALARS] = {};
  // Krylov scaling of the current solve: rows prow / jrow, columns pcol / jcol
  double prow = 1, pcol = 1, jrow = 1, jcol = 1;
  // the V-cycle's own input buffer for the current solve (null: full)
  double* vin_buf = nullptr;

  void read_out(int k) {
    CCHECK(hipMemcpyAsync(host_out, out.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    CCHECK(hipStreamSynchronize(s));
  }
  // defects and norms of residual r (and finiteness of x): [0] pressure,
  // [1] jump, [2] sum r^2, [3] sum (rows r)^2, [4] x not finite, [5] max |r|
  void check(const double* r, const double* x) {
    launch_cm_check(s, n, np, r, vol.p, len.p, prow, jrow, x, part.p);
    launch_cm_reduce(s, cm_parts(n), 6, 0b110011, part.p, out.p);
    read_out(6);
  }
  double anchor_defect(const double* x, double dt) {
    if (nomit == 0) return 0.0;
    launch_cm_omitted(s, nomit, optr.p, ocol.p, oval.p, x, orhs.p, ovol.p, part.p);
    launch_cm_reduce(s, cm_parts(nomit), 1, 1, part.p, out.p);
    read_out(1);
    return dt * dt * host_out[0];
  }
  void residual(const double* x, const double* b, double* r) {
    launch_cm_residual(s, n, sptr.p, slen.p, scol.p, sval.p, x, b, r);
  }

  // dst = N^-1 src of the scaled block preconditioner: the source divided by
  // the row scaling, the reduced-pressure block (V-cycle of the anchored
  // full-cell rhs, differences to the anchors), the jump rows through the
  // jump factor, each part divided by its column scaling
  void precondition(const double* src, double* dst, const Vcycle& vc) {
    const double* z = nullptr;
    if (np) {
      double* fb = vin_buf ? vin_buf : full.p;
      launch_cm_scatter(s, ncells, role.p, src, prow, fb);
      launch_cm_anchor_chunks(s, nchunks, chunk_ptr.p, list.p, src, prow, apart.p);
      launch_cm_anchor_totals(s, nanchors, anchor_cell.p, anchor_chunk_ptr.p, apart.p, fb);
      z = vc(fb);
      launch_cm_gather(s, np, idx.p, anc.p, z, pcol, dst);
    }
    if (nx) {
      launch_cm_jump_rhs(s, nx, np ? cptr.p : nullptr, ccol.p, cval.p, nullptr, z, idx.p, anc.p, src + np, jrow,
                         jr.p);
      ilu(dst + np, jcol);
    }
  }
  // x = ILU^-1 jr, then x /= d (the iterate in LDS while it fits)
  void ilu(double* x, double d) {
    const int nl = int(jf.llev.size()) - 1, nu = int(jf.ulev.size()) - 1;
    if (jf.pw > 0 && size_t(nx) * sizeof(double) + size_t(nl + nu + 2) * sizeof(int) <= size_t(CM_ILU_LDS))
      launch_cm_ilu_lds(s, nx, jf.pw, nl, llev.p, lorder.p, flen.p, fcol.p, fval.p, nu, ulev.p, uorder.p, blen.p,
                        bcol.p, bval.p, bdiag.p, jr.p, jbt.p, x, d);
    else
      launch_cm_ilu(s, nx, lptr.p, lcol.p, lval.p, nl, llev.p, lorder.p, uptr.p, ucol.p, uval.p, udiag.p, nu, ulev.p,
                    uorder.p, jr.p, jy.p, x, d);
  }

  // Krylov.jl bicgstab! (M = I, right preconditioner N when use_pc, zero start,
  // c = b, atol = 0) on the scaled operator; the solution into inc
  int bicgstab(const double* b, double rtol, int itmax, bool use_pc, const Vcycle& vc) {
    CCHECK(hipMemsetAsync(inc.p, 0, sizeof(double) * n, s));
    CCHECK(hipMemcpyAsync(kr.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    CCHECK(hipMemcpyAsync(kp.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const int P = cm_parts(n);
    launch_cm_dots(s, n, 2, kr.p, kr.p, b, kr.p, part.p);
    launch_cm_scalars(s, P, 2, part.p, scal.p, CM_STEP_INIT);
    CCHECK(hipMemcpyAsync(host_scal, scal.p, sizeof(host_scal), hipMemcpyDeviceToHost, s));
    CCHECK(hipStreamSynchronize(s));
    double rnorm = std::sqrt(host_scal[CM_RNORM2]);
    if (rnorm == 0) return 0;
    if (itmax == 0) itmax = 2 * n;
    const double eps_stop = 0.0 + rtol * rnorm;
    if (host_scal[CM_NEXT_RHO] == 0) return 0;
    bool solved = rnorm <= eps_stop, breakdown = false;
    int it = 0;
    while (!(solved || it >= itmax || breakdown)) {
      ++it;
      const double* y = kp.p;
      if (use_pc) {
        precondition(kp.p, kyz.p, vc);
        y = kyz.p;
      }
      launch_cm_spmv(s, n, np, sptr.p, slen.p, scol.p, sval.p, y, pcol, jcol, prow, jrow, kv.p);
      launch_cm_dots(s, n, 1, b, kv.p, nullptr, nullptr, part.p);
      launch_cm_scalars(s, P, 1, part.p, scal.p, CM_STEP_ALPHA);
      launch_cm_update_s(s, n, kr.p, kv.p, y, ks.p, inc.p, scal.p);
      const double* z = ks.p;
      if (use_pc) {
        precondition(ks.p, kyz.p, vc);
        z = kyz.p;
      }
      launch_cm_spmv(s, n, np, sptr.p, slen.p, scol.p, sval.p, z, pcol, jcol, prow, jrow, kt.p);
      launch_cm_dots(s, n, 2, kt.p, ks.p, kt.p, kt.p, part.p);
      launch_cm_scalars(s, P, 2, part.p, scal.p, CM_STEP_OMEGA);
      launch_cm_update_r(s, n, ks.p, kt.p, z, kr.p, inc.p, scal.p);
      launch_cm_dots(s, n, 2, b, kr.p, kr.p, kr.p, part.p);
      launch_cm_scalars(s, P, 2, part.p, scal.p, CM_STEP_RHO);
      launch_cm_update_p(s, n, kr.p, kv.p, kp.p, scal.p);
      CCHECK(hipMemcpyAsync(host_scal, scal.p, sizeof(host_scal), hipMemcpyDeviceToHost, s));
      CCHECK(hipStreamSynchronize(s));
      rnorm = std::sqrt(host_scal[CM_RNORM2]);
      const double alpha = host_scal[CM_ALPHA];
      const bool mach = rnorm + 1.0 <= 1.0;
      solved = rnorm <= eps_stop || mach;
      breakdown = alpha == 0 || std::isnan(alpha);
    }
    return it;
  }

  // _cmdgfm_initial_solution!: zero-pressure and previous-increment guesses
  // with the jump rows solved through the jump factor
  bool initial_solution(const double* rhs, double* x, double dt, const double* previous, bool use_pc) {
    check(rhs, nullptr);
    double best_defect = (dt * dt) * jmax(host_out[0], host_out[1]);
    double best_l2 = std::sqrt(host_out[3]);
    bool used = false;
    for (int use_previous = 0; use_previous < 2; ++use_previous) {
      CCHECK(hipMemsetAsync(cand.p, 0, sizeof(double) * n, s));
      if (use_previous) {
        if (!previous || np == 0) continue;
        launch_cm_gather(s, np, idx.p, anc.p, previous, 1.0, cand.p);
        check(cand.p, cand.p);
        // an all-zero increment is no guess, a nonfinite one is rejected
        if (host_out[5] == 0.0 || host_out[4] != 0.0) continue;
      }
      if (nx) {
        if (!use_pc) continue;
        launch_cm_jump_rhs(s, nx, np ? cptr.p : nullptr, ccol.p, cval.p, cand.p, nullptr, idx.p, anc.p, rhs + np,
                           1.0, jr.p);
        ilu(cand.p + np, 1.0);
      }
      residual(cand.p, rhs, res.p);
      check(res.p, cand.p);
      if (host_out[4] != 0.0) continue;
      const double defect = (dt * dt) * jmax(host_out[0], host_out[1]);
      const double l2 = std::sqrt(host_out[3]);
      if (std::isfinite(defect) && std::isfinite(l2) && defect < 0.95 * best_defect && l2 < best_l2) {
        CCHECK(hipMemcpyAsync(x, cand.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        best_defect = defect;
        best_l2 = l2;
        used = true;
      }
    }
    return used;
  }
};

CmdgfmSystem::CmdgfmSystem(int ncells, int np, int nx, const int* ptr, const int* col, const double* val,
                           const int* pressure_cells, const int* anchor_cells, const double* pressure_volumes,
                           const double* jump_lengths, int nomitted, const int* optr, const int* ocol,
                           const double* oval, const double* orhs, const double* ovol, double drop_tolerance,
                           bool device, hipStream_t stream)
    : d_(new Impl) {
  Impl& D = *d_;
  if (np < 0 || nx < 0 || nomitted < 0) throw std::invalid_argument("negative system size");
  if (!(drop_tolerance >= 0)) throw std::invalid_argument("drop_tolerance must be nonnegative");
  const int n = np + nx;
  D.s = stream;
  D.device = device;
  D.ncells = ncells;
  D.n = n;
  D.np = np;
  D.nx = nx;
  D.nomit = nomitted;
  if (n > 0 && (!ptr || !col || !val)) throw std::invalid_argument("missing system matrix");
  if (n > 0) check_csr("system matrix", n, n, ptr, col);
  if (np > 0 && (!pressure_cells || !anchor_cells || !pressure_volumes))
    throw std::invalid_argument("missing pressure indices, anchors or volumes");
  if (nx > 0 && !jump_lengths) throw std::invalid_argument("missing jump lengths");
  std::vector<int> role(ncells, -1);
  for (int i = 0; i < np; ++i) {
    const int c = pressure_cells[i];
    if (c < 0 || c >= ncells) throw std::invalid_argument("pressure index outside the mesh");
    if (role[c] >= 0) throw std::invalid_argument("pressure indices must be distinct");
    role[c] = i;
  }
  // anchors: slots in order of first use; retained entries grouped by
  // anchor, ascending, cut into chunks
  std::unordered_map<int, int> slot;
  std::vector<int> anchor_cell;
  std::vector<std::vector<int>> members;
  for (int i = 0; i < np; ++i) {
    const int a = anchor_cells[i];
    if (a < 0 || a >= ncells) throw std::invalid_argument("pressure anchor outside the mesh");
    if (role[a] >= 0) throw std::invalid_argument("a pressure anchor cannot be a retained pressure cell");
    auto it = slot.find(a);
    if (it == slot.end()) {
      it = slot.emplace(a, int(anchor_cell.size())).first;
      anchor_cell.push_back(a);
      members.emplace_back();
    }
    members[it->second].push_back(i);
  }
  std::vector<int> list, chunk_ptr(1, 0), anchor_chunk_ptr(1, 0);
  for (const auto& m : members) {
    for (size_t b = 0; b < m.size(); b += CM_ANCHOR_CHUNK) {
      const size_t e = std::min(m.size(), b + CM_ANCHOR_CHUNK);
      list.insert(list.end(), m.begin() + b, m.begin() + e);
      chunk_ptr.push_back(int(list.size()));
    }
    anchor_chunk_ptr.push_back(int(chunk_ptr.size()) - 1);
  }
  D.nanchors = int(anchor_cell.size());
  D.nchunks = int(chunk_ptr.size()) - 1;
  for (int i = 0; i < np; ++i)
    if (!std::isfinite(pressure_volumes[i])) throw std::invalid_argument("pressure volumes must be finite");
  for (int k = 0; k < nx; ++k)
    if (!(jump_lengths[k] > 0) || !std::isfinite(jump_lengths[k]))
      throw std::invalid_argument("jump lengths must be finite and positive");
  if (nomitted > 0) {
    if (!optr || !orhs || !ovol) throw std::invalid_argument("missing omitted rows");
    check_csr("omitted rows", nomitted, n, optr, ocol);
  }
  // C block (jump rows, pressure columns) and the jump factor
  std::vector<int> cptr(1, 0), ccol;
  std::vector<double> cval;
  for (int k = 0; k < nx; ++k) {
    for (int p = ptr[np + k]; p < ptr[np + k + 1] && col[p] < np; ++p) {
      ccol.push_back(col[p]);
      cval.push_back(val[p]);
    }
    cptr.push_back(int(ccol.size()));
  }
  if (nx > 0) {
    cmdgfm_jump_schur(n, np, ptr, col, val, D.jf);
    cmdgfm_crout_ilu(drop_tolerance, D.jf);
  }
  // _cmdgfm_matrix_asymmetry (max |A_ij - A_ji|) and the jump block's
  // diagonal-dominance margin
  double asym = 0.0;
  for (int i = 0; i < n; ++i)
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
      const int j = col[p];
      const int* b = col + ptr[j];
      const int* e = col + ptr[j + 1];
      const int* f = std::lower_bound(b, e, i);
      const double t = (f != e && *f == i) ? val[ptr[j] + int(f - b)] : 0.0;
      asym = jmax(asym, std::fabs(val[p] - t));
    }
  D.asym = asym;
  double margin = std::numeric_limits<double>::infinity();
  for (int k = 0; k < nx; ++k) {
    double dg = 0.0, off = 0.0;
    for (int p = ptr[np + k]; p < ptr[np + k + 1]; ++p) {
      if (col[p] < np) continue;
      if (col[p] == np + k)
        dg = dg + std::fabs(val[p]);
      else
        off = off + std::fabs(val[p]);
    }
    margin = std::min(margin, dg - off);
  }
  D.margin = margin;
  if (!device) return;
  auto up_i = [&](Buf<int>& b, const int* src, size_t k) {
    b.alloc(k);
    if (k) CCHECK(hipMemcpyAsync(b.p, src, k * sizeof(int), hipMemcpyHostToDevice, stream));
  };
  auto up_d = [&](Buf<double>& b, const double* src, size_t k) {
    b.alloc(k);
    if (k) CCHECK(hipMemcpyAsync(b.p, src, k * sizeof(double), hipMemcpyHostToDevice, stream));
  };
  // sliced ELL: CM_SLICE rows per slice, as wide as its longest row, entries
  // column-major (a wavefront reads one entry of 64 rows per load)
  {
    const int nsl = (n + CM_SLICE - 1) / CM_SLICE;
    std::vector<int> sp(nsl + 1, 0), sl(std::max(n, 1), 0);
    for (int k = 0; k < nsl; ++k) {
      int w = 0;
      for (int i = k * CM_SLICE; i < std::min(n, (k + 1) * CM_SLICE); ++i) w = std::max(w, ptr[i + 1] - ptr[i]);
      sp[k + 1] = sp[k] + w * CM_SLICE;
    }
    std::vector<int> sc(std::max(sp[nsl], 1), 0);
    std::vector<double> sv(std::max(sp[nsl], 1), 0.0);
    for (int i = 0; i < n; ++i) {
      sl[i] = ptr[i + 1] - ptr[i];
      const int base = sp[i / CM_SLICE] + i % CM_SLICE;
      for (int e = 0; e < sl[i]; ++e) {
        sc[base + e * CM_SLICE] = col[ptr[i] + e];
        sv[base + e * CM_SLICE] = val[ptr[i] + e];
      }
    }
    D.sptr.upload(sp, stream);
    D.slen.upload(sl, stream);
    D.scol.upload(sc, stream);
    D.sval.upload(sv, stream);
    CCHECK(hipStreamSynchronize(stream));
  }
  D.cptr.upload(cptr, stream);
  D.ccol.upload(ccol, stream);
  D.cval.upload(cval, stream);
  D.role.upload(role, stream);
  up_i(D.idx, pressure_cells, np);
  up_i(D.anc, anchor_cells, np);
  D.list.upload(list, stream);
  D.chunk_ptr.upload(chunk_ptr, stream);
  D.anchor_cell.upload(anchor_cell, stream);
  D.anchor_chunk_ptr.upload(anchor_chunk_ptr, stream);
  up_d(D.vol, pressure_volumes, np);
  up_d(D.len, jump_lengths, nx);
  if (nomitted > 0) {
    up_i(D.optr, optr, size_t(nomitted) + 1);
    up_i(D.ocol, ocol, optr[nomitted]);
    up_d(D.oval, oval, optr[nomitted]);
    up_d(D.orhs, orhs, nomitted);
    up_d(D.ovol, ovol, nomitted);
  }
  if (nx > 0) {
    const JumpFactor& f = D.jf;
    D.lptr.upload(f.lptr, stream);
    D.lcol.upload(f.lcol, stream);
    D.lval.upload(f.lval, stream);
    D.llev.upload(f.llev, stream);
    D.lorder.upload(f.lorder, stream);
    D.uptr.upload(f.uptr, stream);
    D.ucol.upload(f.ucol, stream);
    D.uval.upload(f.uval, stream);
    D.udiag.upload(f.udiag, stream);
    D.ulev.upload(f.ulev, stream);
    D.uorder.upload(f.uorder, stream);
    D.flen.upload(f.flen, stream);
    D.fcol.upload(f.fcol, stream);
    D.fval.upload(f.fval, stream);
    D.blen.upload(f.blen, stream);
    D.bcol.upload(f.bcol, stream);
    D.bval.upload(f.bval, stream);
    D.bdiag.upload(f.bdiag, stream);
  }
  for (Buf<double>* b : {&D.corr, &D.srhs, &D.inc, &D.kr, &D.kp, &D.kv, &D.ks, &D.kt, &D.kyz, &D.cand, &D.res})
    b->alloc(std::max(n, 1));
  D.full.alloc(std::max(ncells, 1));
  D.jr.alloc(std::max(nx, 1));
  D.jy.alloc(std::max(nx, 1));
  D.jbt.alloc(std::max(nx, 1));
  D.part.alloc(size_t(CM_MAX_WG) * 6);
  D.apart.alloc(std::max(D.nchunks, 1));
  D.scal.alloc(CM_SCALARS);
  D.out.alloc(8);
  D.flag.alloc(1);
  CCHECK(hipStreamSynchronize(stream));
}

CmdgfmSystem::~CmdgfmSystem() = default;

int CmdgfmSystem::size() const { return d_->n; }
int CmdgfmSystem::pressure_size() const { return d_->np; }
int CmdgfmSystem::jump_size() const { return d_->nx; }
const JumpFactor& CmdgfmSystem::factor() const { return d_->jf; }

void CmdgfmSystem::solve(const double* rhs, double* x, double dt, double density, double h, double tol,
                         int max_iterations, bool warm_start, const double* previous, bool use_pc, const Vcycle& vc,
                         const VcycleInput& vin, CmdgfmStats& st) {
  Impl& D = *d_;
  if (!D.device) throw std::invalid_argument("host-only block system");
  if (!(dt > 0)) throw std::invalid_argument("dt must be strictly positive");
  if (!(tol > 0)) throw std::invalid_argument("tolerance must be strictly positive");
  if (max_iterations <= 0) throw std::invalid_argument("max_iterations must be strictly positive");
  if (!(density > 0) || !(h > 0)) throw std::invalid_argument("density and reference length must be positive");
  const int n = D.n;
  hipStream_t s = D.s;
  st = CmdgfmStats{};
  st.pressure_unknowns = D.np;
  st.jump_unknowns = D.nx;
  st.asymmetry = D.asym;
  st.jump_diagonal_margin = D.margin;
  if (n == 0) {
    st.converged = 1;
    return;
  }
  // _cmdgfm_krylov_scaling!: pressure unknowns scale rho (L/dt)^2 (rows
  // (dt/L)^2), jumps L/dt^2 (rows dt^2/L)
  D.prow = (dt / h) * (dt / h);
  D.pcol = density * ((h / dt) * (h / dt));
  D.jrow = (dt * dt) / h;
  D.jcol = h / (dt * dt);
  D.vin_buf = vin ? vin() : nullptr;
  CCHECK(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  CCHECK(hipMemcpyAsync(D.corr.p, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  D.check(rhs, nullptr);
  const double initial = std::sqrt(D.host_out[2]);
  st.initial_residual = initial;
  int iterations = 0, runs = 0;
  bool used = false;
  if (initial > 0) {
    if (warm_start) {
      used = D.initial_solution(rhs, x, dt, previous, use_pc);
      D.residual(x, rhs, D.corr.p);
    }
    while (iterations < max_iterations) {
      D.check(D.corr.p, nullptr);
      const double pd = dt * dt * D.host_out[0], jd = dt * dt * D.host_out[1];
      const double rhs_norm = std::sqrt(D.host_out[3]);
      const double defect = jmax(jmax(pd, jd), D.anchor_defect(x, dt));
      if (!std::isfinite(defect) || defect <= tol) break;
      if (!std::isfinite(rhs_norm) || rhs_norm == 0) break;
      launch_cm_scale_rows(s, n, D.np, D.corr.p, D.prow, D.jrow, D.srhs.p);
      const double requested = std::min(tol, tol / std::max(defect, EPS));
      double rhs_scale = 1.0;
      if (rhs_norm < 1.0 && rhs_norm * requested < EPS) {
        launch_cm_divs(s, n, D.srhs.p, rhs_norm);
        rhs_scale = rhs_norm;
      }
      int its = D.bicgstab(D.srhs.p, requested, max_iterations - iterations, use_pc, vc);
      ++runs;
      iterations += its;
      D.check(D.inc.p, D.inc.p);
      bool finite = D.host_out[4] == 0.0;
      if (!finite && use_pc && iterations < max_iterations) {
        its = D.bicgstab(D.srhs.p, requested, max_iterations - iterations, false, vc);
        ++runs;
        iterations += its;
        D.check(D.inc.p, D.inc.p);
        finite = D.host_out[4] == 0.0;
      }
      if (!finite || its == 0) break;
      CCHECK(hipMemsetAsync(D.flag.p, 0, sizeof(int), s));
      launch_cm_increment(s, n, D.np, x, D.pcol, D.jcol, rhs_scale, D.inc.p, D.flag.p);
      int changed = 0;
      CCHECK(hipMemcpyAsync(&changed, D.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
      CCHECK(hipStreamSynchronize(s));
      if (!changed) break;
      D.residual(x, rhs, D.corr.p);
    }
  }
  D.check(D.corr.p, nullptr);
  st.final_residual = std::sqrt(D.host_out[2]);
  st.pressure_defect = dt * dt * D.host_out[0];
  st.jump_defect = dt * dt * D.host_out[1];
  st.anchor_defect = D.anchor_defect(x, dt);
  st.converged = st.pressure_defect <= tol && st.jump_defect <= tol;
  st.iterations = iterations;
  st.krylov_runs = runs;
  st.initial_guess_used = used;
}

}  // namespace tpamr
