// Runtime behind the C ABI: device memory of the hierarchy, the V-cycle
// launch sequence, the FP64 Krylov driver with the reference's stopping
// rules, and the face-level two-phase / projection / adjoint operations.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tpamr.h"
#include "hier.hpp"
#include "kernels.hpp"
#include "cmdgfm.hpp"
#include "momentum.hpp"

using namespace tpamr;

namespace {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                                  \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));           \
  } while (0)

const double EPS = std::numeric_limits<double>::epsilon();
constexpr int PIPE_MAX = 8;

template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  DevArray(DevArray&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevArray& operator=(DevArray&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    n = count;
    if (count) HIPCHECK(hipMalloc(&p, count * sizeof(T)));
  }
  template <typename S, typename A>
  void upload(const std::vector<S, A>& v, hipStream_t s) {
    alloc(v.size());
    if (v.empty()) return;
    if constexpr (std::is_same<S, T>::value) {
      HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHECK(hipStreamSynchronize(s));
    } else {
      std::vector<T> tmp(v.begin(), v.end());
      HIPCHECK(hipMemcpyAsync(p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHECK(hipStreamSynchronize(s));
    }
  }
};

struct DevTopo {
  int nblocks = 0, ncells = 0;
  DevArray<int> nbr, xfer, blk_ext, ext_col, blk_irr, blk_ng;
  DevArray<unsigned> ell_loc;
  DevArray<unsigned short> ell_ref;
  DevArray<double> volume;
};

// Device form of a level's coefficient refresh (uploaded once per topology):
// restriction from the finer level, per-face data, the regular faces' slots
// in the device layout, the irregular rows' terms, the level's face
// coefficients and, for levels without an FP64 copy, FP64 staging values.
struct DevPlan {
  int nfaces = 0, nrows = 0;
  DevArray<int> rptr, ridx, dst, irr_ptr, term_ptr, term_f, term_g, ell_pos, ell_dpos, ell_row, diag_e, cell_irr;
  DevArray<double> rw, kdiag, area, kappa, term_k, term_s, irr_diag, val_e;
  int nnz = 0;  // irregular entries
  DevArray<unsigned char> term_diag;
  DevArray<double> cw, ce, cs, cn, ell_val, ell_diag, dinv;
  // open boundary faces: conductance and destination (slot code or ELL
  // position), and every irregular row's boundary diagonal
  int nopen = 0;
  bool open = false;
  DevArray<int> bdst;
  DevArray<double> bw, irr_bd;
};

// A level's ungauged pressure components (uploaded after a host refresh of
// a hierarchy with several components): the arrays behind its GaugeView and
// the component means of the last projection or gauge.
struct DevGauge {
  int ncomp = 0;
  DevArray<int> slot, pair_chunk, pair_slot, slot_pair_ptr, slot_pairs;
  DevArray<double> vol, partial, mean;
  GaugeView view;
};

static_assert(GAUGE_CHUNK == GAUGE_CHUNK_CELLS, "host and device component chunks differ");

template <typename T>
struct DevLevel {
  DevArray<T> cw, ce, cs, cn, ell_val, ell_bound, ell_bdiag, ell_diag, dinv;
  DevArray<T> x, xs, b, u;
  double omega = 1, omega_p = 0;
  LevelView<T> view{};
};

}  // namespace

struct tpamr_ctx {
  int device = -1;
  int precision = TPAMR_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  MultigridConfig cfg;
  std::unique_ptr<Hierarchy> H;
  std::unique_ptr<TreeMesh> mesh_pending;  // before the first coefficients
  // the hierarchy of the previous mesh, whose coarse suffix the next build
  // may take over (same multigrid options)
  std::unique_ptr<Hierarchy> H_prev;
  bool have_coeff = false;
  // open (homogeneous Dirichlet) pressure sides, left / right / bottom / top,
  // and the finest level's boundary face coefficients: the caller's
  // (kb_user), or, with kb_extend, extended from the face coefficients of the
  // next coefficient update; kb_dirty: the device refresh plan's boundary
  // values are older than the hierarchy's
  // kb_pending: not yet applied to the hierarchy (the next coefficient
  // update applies it); bgrad_dirty: the device boundary gradient factors
  // are older than the hierarchy's coefficients
  bool open[4] = {false, false, false, false};
  std::vector<double> kb_user;
  bool kb_extend = true, kb_dirty = false, kb_pending = false, bgrad_dirty = false;
  bool any_open() const { return open[0] || open[1] || open[2] || open[3]; }
  // finest level's boundary faces on the device: cell, gradient factor
  // 2 kb / h (0 on closed faces), and the boundary velocities of a projection
  DevArray<int> bcell;
  DevArray<double> bgrad, fb;
  // components of every level (Hierarchy::general_gauge) and the last
  // solve's removed component means with their roots (1-based)
  std::vector<DevGauge> gauge;
  std::vector<int> last_roots;
  std::vector<double> last_means;
  // device data
  std::vector<DevTopo> topo;
  std::vector<DevLevel<double>> ld;  // FP64 levels (level 0 always, all in FP64 mode)
  std::vector<DevLevel<float>> lf;   // FP32 levels (FP32 mode)
  DevArray<double> coarse_inv, coarse_work, coarse_col;
  DevArray<int> coarse_piv, coarse_fail;
  // full inverse of the last coarse refresh (the start of the next one's
  // Newton-Schulz update), its identity residual, GEMM scratch
  DevArray<double> coarse_x, coarse_t, coarse_xn;
  DevArray<double> gj_R, gj_M;  // blocked Gauss-Jordan: panel rows, panel update factor
  DevArray<unsigned long long> coarse_res;
  bool coarse_x_valid = false;
  bool coarse_exact = false;  // TPAMR_COARSE_GJ=1: Gauss-Jordan at every refresh
  double coarse_gj_residual = 0;
  rocblas_handle blas = nullptr;
  // coefficient refresh on the device (after the first, host-built one);
  // TPAMR_HOST_REFRESH=1 in the environment keeps every refresh on the host
  std::vector<DevPlan> plan;
  bool plan_ready = false, host_refresh = false, host_stale = false;
  DevArray<unsigned long long> ref_ratio;
  DevArray<int> ref_flags;
  DevArray<double> x, r, p, Ap, b, tmp, tmp2;
  DevArray<double> part, maxpart, scal, stage;
  CgState* st_own = nullptr;  // device
  // Reduction scratch and Krylov state of the solve path: the context's own
  // buffers above, or (for the duration of tpamr_solve_device_noalloc) the
  // library's static device scratch.
  struct Reduce {
    double *part = nullptr, *maxpart = nullptr, *scal = nullptr, *stage = nullptr;
    CgState* st = nullptr;
  } rs;
  CgState* st_host = nullptr; // pinned
  CgState* st_ring = nullptr; // pinned, one per Krylov iteration in flight
  // face data (finest level)
  DevArray<int> fm, fp, kptr, kface, cptr, cface, bptr, bface;
  DevArray<signed char> faxis, csign, bside;
  DevArray<double> farea, kdiag, kval, kappa, sqk, spacing, barea;
  DevArray<double> origin, cell2;  // 2 per block: lower-left corner; 2 per cell: staging
  DevArray<double> fcx, fcy;       // face centres
  DevArray<int> counter;
  DevArray<double> f1, f2, f3, f4;
  // halo refresh schedule (built on first use for a halo width)
  HaloSchedule halo_host;
  int halo_dev = 0;  // halo width of the uploaded schedule (0: none)
  DevArray<int> h_dest, h_ptr, h_donor, h_phys;
  DevArray<double> h_weight, h_out, h_phys_values;
  // graph of one Krylov iteration (TPAMR_GRAPH=0 in the environment launches
  // the iteration's kernels one by one instead)
  bool use_graph = true;
  bool trace = false;  // TPAMR_TRACE=1: print the Krylov state of every iteration to stderr
  int graph_iters = 1;  // Krylov iterations per captured graph (TPAMR_GRAPH_ITERS)
  // V-cycle levels with at most this many blocks run in one workgroup
  // (TPAMR_TAIL_BLOCKS; 0: every level is launched on its own)
  int tail_blocks = 16;
  hipGraphExec_t graph = nullptr;
  hipGraphExec_t graph_ir = nullptr;  // one inner FP32 Krylov iteration
  // the first iteration of a Krylov run (its initial residual and V-cycle)
  hipGraphExec_t graph0 = nullptr, graph0_ir = nullptr;
  // TPAMR_FP32_REFINE: the inner FP32 Krylov iteration replaces its
  // residual by the FP64 residual b - A x once the strong residual has
  // fallen by ir_delta since the last replacement (TPAMR_IR_DELTA in the
  // environment)
  double ir_delta = 5e-3;
  bool ir_keep = false;  // TPAMR_IR_RESTART=0: replacements keep the recurrence
  DevArray<float> p32, Ap32, e32;  // inner Krylov vectors (TPAMR_FP32_REFINE)
  bool fp32() const { return precision != TPAMR_FP64; }
  // Krylov iterations in flight before the host reads a done flag
  // (TPAMR_PIPE in the environment, 2..PIPE_MAX)
  int pipe = 3;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evs[PIPE_MAX] = {};
  // the `_device` entry points wait for the work queued on the caller's
  // stream (tpamr_set_caller_stream; default the null stream) before they
  // touch its buffers
  hipStream_t caller = nullptr;
  hipEvent_t ev_caller = nullptr;
  // the library's static device scratch on this context's device
  void* noalloc = nullptr;
  // momentum operators of the one-velocity step (tpamr_set_momentum), the
  // step's old and transported velocities and a pressure staging vector,
  // and the last step's viscous solve statistics
  std::unique_ptr<Momentum> mom;
  DevArray<double> mold, mconv, mp, mforce;
  tpamr_stats vstats[2] = {};
  // assembled CMDGFM block system (tpamr_cmdgfm_set_system) and the device
  // staging of its host-array solve
  std::unique_ptr<CmdgfmSystem> cmd;
  DevArray<double> cm_rhs, cm_x, cm_prev, cm_z;
  const TreeMesh& mesh() const {
    if (H) return *H->levels[0]->mesh;
    if (mesh_pending) return *mesh_pending;
    throw std::logic_error("no mesh: call tpamr_set_mesh first");
  }
};

namespace {

struct StateError : std::logic_error {
  using std::logic_error::logic_error;
};
struct SizeError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct IncompatibleError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

template <class F>
int guard(tpamr_ctx* ctx, F&& f) {
  try {
    f();
    return TPAMR_OK;
  } catch (const HipError& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_EDEVICE;
  } catch (const MomentumDeviceError& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_EDEVICE;
  } catch (const CmdgfmDeviceError& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_EDEVICE;
  } catch (const StateError& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_ESTATE;
  } catch (const SizeError& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_ESIZE;
  } catch (const IncompatibleError& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_EINCOMPATIBLE;
  } catch (const std::invalid_argument& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_EINVAL;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return TPAMR_EINVAL;
  }
}

void need_device(const tpamr_ctx* c) {
  if (c->device < 0) throw StateError("host-only context: create the context on a HIP device to compute");
}

// Order the context stream after everything queued so far on the caller's
// stream, so device inputs written there are complete when the library
// reads them.
void order_after_caller(tpamr_ctx* c) {
  HIPCHECK(hipEventRecord(c->ev_caller, c->caller));
  HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_caller, 0));
}

void need_solver(const tpamr_ctx* c) {
  if (!c->H || !c->have_coeff) throw StateError("no operator: call tpamr_set_face_coefficients first");
}

void destroy_graph(tpamr_ctx* c);

// A new mesh or hierarchy invalidates everything derived from the old one's
// topology: the captured iteration, the device refresh plan (face slots,
// irregular-row terms, level sizes) and the previous coarse inverse that
// starts a Newton-Schulz update.
void reset_topology_state(tpamr_ctx* c) {
  destroy_graph(c);
  c->plan.clear();
  c->plan_ready = false;
  c->host_stale = false;
}

void destroy_graph(tpamr_ctx* c) {
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  for (hipGraphExec_t* g : {&c->graph0, &c->graph0_ir})
    if (*g) {
      (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  if (c->graph_ir) {
    (void)hipGraphExecDestroy(c->graph_ir);
    c->graph_ir = nullptr;
  }
}

template <typename T>
void upload_level_values(tpamr_ctx* c, DevLevel<T>& D, const LevelData& L, const DevTopo& Tp, bool with_bound) {
  hipStream_t s = c->stream;
  D.cw.upload(L.cw, s);
  D.ce.upload(L.ce, s);
  D.cs.upload(L.cs, s);
  D.cn.upload(L.cn, s);
  D.ell_val.upload(L.ell_val, s);
  if (with_bound) {
    D.ell_bound.upload(L.ell_bound, s);
    D.ell_bdiag.upload(L.ell_bdiag.empty() ? std::vector<double>(1, 0.0) : L.ell_bdiag, s);
  }
  D.ell_diag.upload(L.ell_diag, s);
  D.dinv.upload(L.dinv, s);
  D.omega = L.omega;
  D.omega_p = L.omega_p;
  D.x.alloc(L.ncells);
  D.xs.alloc(L.ncells);
  D.b.alloc(L.ncells);
  D.u.alloc(L.ncells);
  LevelView<T>& v = D.view;
  v.nblocks = L.nblocks;
  v.ncells = L.ncells;
  v.bx = L.mesh->bs[0];
  v.by = L.mesh->bs[1];
  {
    const int tpb = v.bx * v.by / 4, bpw = tpb >= 256 ? 1 : 256 / tpb;
    const int groups = (L.nblocks + bpw - 1) / bpw;
    v.wg_chunk = std::max(1, std::min(32, groups / 256));
    const int step = 8 * v.wg_chunk;
    v.wg_grid = (groups + step - 1) / step * step;
  }
  v.ext_cap = std::max(4, (L.ext_max + 3) & ~3);
  v.cw = D.cw.p;
  v.ce = D.ce.p;
  v.cs = D.cs.p;
  v.cn = D.cn.p;
  v.ell_w = L.ell_w;
  v.ell_val = D.ell_val.p;
  v.ell_bound = D.ell_bound.p;
  v.ell_bdiag = D.ell_bdiag.p;
  v.ell_diag = D.ell_diag.p;
  v.ell_ref = Tp.ell_ref.p;
  v.blk_irr = Tp.blk_irr.p;
  v.blk_ng = Tp.blk_ng.p;
  v.ell_loc = Tp.ell_loc.p;
  v.blk_ext = Tp.blk_ext.p;
  v.ext_col = Tp.ext_col.p;
  v.nbr = Tp.nbr.p;
  v.xfer = Tp.xfer.p;
  v.volume = Tp.volume.p;
}

void coarse_inverse_device(tpamr_ctx* c);

// Upload topology (once per hierarchy) and values (every refresh).
void upload_hierarchy(tpamr_ctx* c, bool topology) {
  Hierarchy& H = *c->H;
  hipStream_t s = c->stream;
  const int nl = int(H.levels.size());
  if (topology) {
    reset_topology_state(c);
    // the previous coarse inverse starts the next Newton-Schulz update only
    // when the coarsest level was taken over from the previous hierarchy
    if (H.reused_new < 0) c->coarse_x_valid = false;
    c->topo.clear();
    c->topo.resize(nl);
    for (int l = 0; l < nl; ++l) {
      const LevelData& L = *H.levels[l];
      DevTopo& T = c->topo[l];
      T.nblocks = L.nblocks;
      T.ncells = L.ncells;
      T.nbr.upload(L.nbr, s);
      T.xfer.upload(L.xfer.empty() ? std::vector<int>(2 * L.nblocks, -1) : L.xfer, s);
      if (L.ell_w > ELL_MAX) throw std::runtime_error("irregular row longer than the kernels support");
      T.ell_ref.upload(L.ell_ref.empty() ? std::vector<unsigned short>(1, 0) : L.ell_ref, s);
      T.blk_ext.upload(L.blk_ext, s);
      T.blk_irr.upload(L.blk_irr, s);
      {
        // a block's rows run in decreasing length: its first row is the longest
        std::vector<int> ngb(L.nblocks, 0);
        for (int b = 0; b < L.nblocks; ++b)
          if (L.blk_irr[b + 1] > L.blk_irr[b]) ngb[b] = int(((L.ell_loc[L.blk_irr[b]] >> 16) + 3) / 4);
        T.blk_ng.upload(ngb, s);
      }
      T.ell_loc.upload(L.ell_loc.empty() ? std::vector<uint32_t>(1, 0u) : L.ell_loc, s);
      T.ext_col.upload(L.ext_col.empty() ? std::vector<int>(1, 0) : L.ext_col, s);
      T.volume.upload(L.volume, s);
    }
    c->ld.clear();
    c->lf.clear();
    c->ld.resize(c->fp32() ? 1 : nl);
    if (c->fp32()) c->lf.resize(nl);
    // finest-level face data
    const TreeMesh& M = *H.levels[0]->mesh;
    const LevelData& L0 = *H.levels[0];
    c->fm.upload(M.fminus, s);
    c->fp.upload(M.fplus, s);
    std::vector<signed char> ax(M.faxis.begin(), M.faxis.end());
    c->faxis.upload(ax, s);
    c->farea.upload(M.farea, s);
    c->kdiag.upload(L0.K.diag, s);
    c->kptr.upload(L0.K.off_ptr, s);
    c->kface.upload(L0.K.off_face, s);
    c->kval.upload(L0.K.off_val, s);
    c->spacing.upload(M.bspacing, s);
    c->origin.upload(M.borigin, s);
    c->fcx.upload(M.fcx, s);
    c->fcy.upload(M.fcy, s);
    c->counter.alloc(1);
    // cell -> face incidence (face order) for the divergence gather
    const int N = M.ncells(), F = M.nfaces();
    std::vector<int> cnt(N + 1, 0);
    for (int f = 0; f < F; ++f) {
      cnt[M.fminus[f] + 1]++;
      cnt[M.fplus[f] + 1]++;
    }
    for (int q = 0; q < N; ++q) cnt[q + 1] += cnt[q];
    std::vector<int> cf(cnt[N]);
    std::vector<signed char> cs(cnt[N]);
    std::vector<int> next(cnt.begin(), cnt.end() - 1);
    for (int f = 0; f < F; ++f) {
      cf[next[M.fminus[f]]] = f;
      cs[next[M.fminus[f]]++] = -1;
      cf[next[M.fplus[f]]] = f;
      cs[next[M.fplus[f]]++] = 1;
    }
    c->cptr.upload(cnt, s);
    c->cface.upload(cf, s);
    c->csign.upload(cs, s);
    const int BF = int(M.bcell.size());
    std::vector<int> bcnt(N + 1, 0);
    for (int f = 0; f < BF; ++f) bcnt[M.bcell[f] + 1]++;
    for (int q = 0; q < N; ++q) bcnt[q + 1] += bcnt[q];
    std::vector<int> bf(std::max(BF, 1));
    std::vector<int> bnext(bcnt.begin(), bcnt.end() - 1);
    for (int f = 0; f < BF; ++f) bf[bnext[M.bcell[f]]++] = f;
    c->bptr.upload(bcnt, s);
    c->bface.upload(bf, s);
    {
      std::vector<int> bc(M.bcell.begin(), M.bcell.end());
      if (bc.empty()) bc.push_back(0);
      c->bcell.upload(bc, s);
      c->fb.alloc(std::max(BF, 1));
      c->bgrad_dirty = true;
    }
    std::vector<signed char> bsd(M.bside.begin(), M.bside.end());
    if (bsd.empty()) bsd.push_back(0);
    c->bside.upload(bsd, s);
    std::vector<double> ba(M.barea.begin(), M.barea.end());
    if (ba.empty()) ba.push_back(0);
    c->barea.upload(ba, s);
    // vectors and reduction scratch
    c->x.alloc(N);
    c->r.alloc(N);
    c->p.alloc(N);
    c->Ap.alloc(N);
    c->b.alloc(N);
    c->tmp.alloc(N);
    c->tmp2.alloc(N);
    c->halo_host = HaloSchedule{};
    c->halo_dev = 0;
    if (c->precision == TPAMR_FP32_REFINE) {
      c->p32.alloc(N);
      c->Ap32.alloc(N);
      c->e32.alloc(N);
    }
    c->f1.alloc(F);
    c->f2.alloc(F);
    c->f3.alloc(F);
    c->f4.alloc(F);
    const size_t nblk = size_t(L0.nblocks);
    const size_t ncb = (size_t(N) + 255) / 256;
    // per-block partials (3 sets: z statistics) or per-256-cell partials
    // (4 sets: the rhs preparation sums)
    c->part.alloc(std::max(3 * nblk, 4 * ncb) + 16);
    c->maxpart.alloc(std::max(nblk, ncb) + 16);
    c->scal.alloc(16);
    c->stage.alloc(4 * RED_WG);
    c->rs.part = c->part.p;
    c->rs.maxpart = c->maxpart.p;
    c->rs.scal = c->scal.p;
    c->rs.stage = c->stage.p;
    destroy_graph(c);
  }
  // kernel arguments of the captured iteration (weights, pointers) change
  destroy_graph(c);
  for (int l = 0; l < int(c->ld.size()); ++l) upload_level_values(c, c->ld[l], *H.levels[l], c->topo[l], l == 0);
  for (int l = 0; l < int(c->lf.size()); ++l) upload_level_values(c, c->lf[l], *H.levels[l], c->topo[l], false);
  coarse_inverse_device(c);
  c->kappa.upload(H.levels[0]->kappa, s);
  std::vector<double> sq(H.levels[0]->kappa.size());
  for (size_t f = 0; f < sq.size(); ++f) sq[f] = std::sqrt(H.levels[0]->kappa[f]);
  c->sqk.upload(sq, s);
}

// ---------------------------------------------------------------------------
// V-cycle launch sequence

// Enqueue one V-cycle. On entry the finest level's b holds the residual and
// its x holds the zero-start sweep (or zero). Returns the buffer holding z.
// x0_fused: the finest level's x is not stored; its first pre-sweep forms
// the zero-start iterate from b (pre >= 2).
template <typename T>
T* enqueue_vcycle(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, const int* done, const ZStats& zs,
                  bool x0_fused = false) {
  hipStream_t s = c->stream;
  const int nl = int(lv.size());
  const int pre = c->cfg.pre, post = c->cfg.post;
  std::vector<T*> cur(nl), oth(nl);
  for (int l = 0; l < nl; ++l) {
    cur[l] = lv[l].x.p;
    oth[l] = lv[l].xs.p;
  }
  if (nl == 1) {
    launch_coarse<T>(s, c->H->coarse_n, c->coarse_inv.p, lv[0].b.p, lv[0].x.p, done);
    if (zs.part) launch_zstats<T>(s, lv[0].view.nblocks, lv[0].view.bx, lv[0].x.p, zs, lv[0].view.volume, done);
    return lv[0].x.p;
  }
  // levels from `ts` on (at most tail_blocks blocks each) run in one
  // workgroup; ts = nl - 1: no tail
  int ts = nl - 1;
  for (int l = 1; l + 1 < nl; ++l)
    if (lv[l].view.nblocks <= c->tail_blocks && nl - l <= TAIL_MAX) {
      ts = l;
      break;
    }
  for (int l = 0; l < ts; ++l) {
    DevLevel<T>& L = lv[l];
    for (int q = 1; q < pre; ++q) {
      const bool zs0 = x0_fused && l == 0 && q == 1;
      launch_sweep<T>(s, L.view, cur[l], oth[l], L.b.p, L.omega, done, l == 0, ZStats{}, zs0 ? L.dinv.p : nullptr,
                      L.omega);
      std::swap(cur[l], oth[l]);
    }
    launch_residual<T>(s, L.view, cur[l], L.b.p, L.u.p, L.omega_p, done, l == 0);
    DevLevel<T>& C = lv[l + 1];
    const bool coarsest = l + 1 == nl - 1;
    launch_restrict<T>(s, L.view, L.u.p, L.omega_p, C.b.p, C.x.p, C.dinv.p, C.omega, (!coarsest && pre > 0) ? 1 : 0,
                       done, l == 0);
  }
  if (ts == nl - 1) {
    launch_coarse<T>(s, c->H->coarse_n, c->coarse_inv.p, lv[nl - 1].b.p, lv[nl - 1].x.p, done);
  } else {
    TailArgs<T> A{};
    A.n = nl - ts;
    for (int k = 0; k < A.n; ++k) {
      DevLevel<T>& L = lv[ts + k];
      A.L[k] = L.view;
      A.x[k] = L.x.p;
      A.xs[k] = L.xs.p;
      A.b[k] = L.b.p;
      A.u[k] = L.u.p;
      A.dinv[k] = L.dinv.p;
      A.omega[k] = L.omega;
      A.omega_p[k] = L.omega_p;
    }
    A.inv = c->coarse_inv.p;
    A.coarse_n = c->H->coarse_n;
    A.pre = pre;
    A.post = post;
    if (A.coarse_n <= COARSE_INLINE) {
      A.phase = 0;
      launch_vtail<T>(s, A, done);
    } else {
      A.phase = 1;
      launch_vtail<T>(s, A, done);
      launch_coarse<T>(s, A.coarse_n, c->coarse_inv.p, lv[nl - 1].b.p, lv[nl - 1].x.p, done);
      A.phase = 2;
      launch_vtail<T>(s, A, done);
    }
    // level ts's result: the buffer after its (pre - 1) + post sweeps
    if ((std::max(pre - 1, 0) + post) & 1) std::swap(cur[ts], oth[ts]);
  }
  for (int l = ts - 1; l >= 0; --l) {
    DevLevel<T>& L = lv[l];
    launch_prolong<T>(s, L.view, cur[l], cur[l + 1], L.omega_p, done, l == 0);
    const int sweeps = (l == 0 && zs.part && post == 0) ? 1 : post;
    for (int q = 0; q < sweeps; ++q) {
      const bool last = l == 0 && q == sweeps - 1 && zs.part;
      const double w = (post == 0) ? 0.0 : L.omega;
      launch_sweep<T>(s, L.view, cur[l], oth[l], L.b.p, w, done, l == 0, last ? zs : ZStats{});
      std::swap(cur[l], oth[l]);
    }
  }
  return cur[0];
}

// The V-cycle of a hierarchy with several components, literally as the
// reference runs it (cycles.jl:8-77): on every level with ungauged
// components the rhs is projected on entry and the component means are
// removed after every sweep and after the coarse correction; levels are
// launched one by one (no tail kernel) and the first sweep from zero runs
// after the projection. The Krylov statistics follow the last removal.
// Level l's component sums and shifts (a hierarchy with several
// components): remove each ungauged component's volume mean from x
// (`_multigrid_remove_component_means!`, hierarchy.jl:309-324), or project a
// rhs b (`_multigrid_project_component_rhs!`, :326-343).
template <typename T>
void gauge_remove(tpamr_ctx* c, int l, T* x) {
  DevGauge& G = c->gauge[l];
  if (!G.ncomp) return;
  const int N = c->topo[l].ncells, nb = c->H->levels[l]->mesh->nb();
  launch_comp_sums<T>(c->stream, G.view, N, nb, x, c->topo[l].volume.p, 1, 1, G.mean.p);
  launch_comp_shift<T>(c->stream, G.view, N, nb, G.mean.p, c->topo[l].volume.p, 0, x);
}
template <typename T>
void gauge_project(tpamr_ctx* c, int l, T* b) {
  DevGauge& G = c->gauge[l];
  if (!G.ncomp) return;
  const int N = c->topo[l].ncells, nb = c->H->levels[l]->mesh->nb();
  launch_comp_sums<T>(c->stream, G.view, N, nb, b, c->topo[l].volume.p, 0, 1, G.mean.p);
  launch_comp_shift<T>(c->stream, G.view, N, nb, G.mean.p, c->topo[l].volume.p, 1, b);
}

// The components of every level on the device (after a host refresh of a
// hierarchy with several components).
void upload_gauge(tpamr_ctx* c) {
  Hierarchy& H = *c->H;
  c->gauge.clear();
  c->gauge.resize(H.levels.size());
  if (!H.general_gauge) return;
  hipStream_t s = c->stream;
  for (size_t l = 0; l < H.levels.size(); ++l) {
    const LevelData& L = *H.levels[l];
    DevGauge& G = c->gauge[l];
    G.ncomp = L.ncomp;
    if (!G.ncomp) continue;
    if (L.comp_slot.size() != size_t(L.ncells) || L.slot_pair_ptr.size() != size_t(L.ncomp) + 1)
      throw std::logic_error("components without their reduction plan");
    G.slot.upload(L.comp_slot, s);
    G.vol.upload(L.comp_vol, s);
    G.pair_chunk.upload(L.pair_chunk, s);
    G.pair_slot.upload(L.pair_slot, s);
    G.slot_pair_ptr.upload(L.slot_pair_ptr, s);
    G.slot_pairs.upload(L.slot_pairs, s);
    G.partial.alloc(L.pair_slot.size());
    G.mean.alloc(L.ncomp);
    G.view.ncomp = L.ncomp;
    G.view.npairs = int(L.pair_slot.size());
    G.view.slot = G.slot.p;
    G.view.vol = G.vol.p;
    G.view.pair_chunk = G.pair_chunk.p;
    G.view.pair_slot = G.pair_slot.p;
    G.view.slot_pair_ptr = G.slot_pair_ptr.p;
    G.view.slot_pairs = G.slot_pairs.p;
    G.view.partial = G.partial.p;
  }
}

template <typename T>
void launch_fill_t(hipStream_t s, int n, T* x) {
  if constexpr (std::is_same<T, float>::value)
    launch_fill_f32(s, n, 0.f, x);
  else
    launch_fill(s, n, 0.0, x);
}

template <typename T>
T* enqueue_vcycle_general(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, const int* done, const ZStats& zs) {
  hipStream_t s = c->stream;
  const int nl = int(lv.size());
  const int pre = c->cfg.pre, post = c->cfg.post;
  std::vector<T*> cur(nl), oth(nl);
  for (int l = 0; l < nl; ++l) {
    cur[l] = lv[l].x.p;
    oth[l] = lv[l].xs.p;
  }
  for (int l = 0; l + 1 < nl; ++l) {
    DevLevel<T>& L = lv[l];
    const int N = L.view.ncells;
    gauge_project<T>(c, l, L.b.p);
    if (pre > 0) {
      launch_zero_sweep<T>(s, N, L.omega, L.dinv.p, L.b.p, cur[l]);
      gauge_remove<T>(c, l, cur[l]);
    } else {
      launch_fill_t<T>(s, N, cur[l]);
    }
    for (int q = 1; q < pre; ++q) {
      launch_sweep<T>(s, L.view, cur[l], oth[l], L.b.p, L.omega, done, l == 0, ZStats{});
      std::swap(cur[l], oth[l]);
      gauge_remove<T>(c, l, cur[l]);
    }
    launch_residual<T>(s, L.view, cur[l], L.b.p, L.u.p, L.omega_p, done, l == 0);
    DevLevel<T>& C = lv[l + 1];
    launch_restrict<T>(s, L.view, L.u.p, L.omega_p, C.b.p, C.x.p, C.dinv.p, C.omega, 0, done, l == 0);
  }
  gauge_project<T>(c, nl - 1, lv[nl - 1].b.p);
  launch_coarse<T>(s, c->H->coarse_n, c->coarse_inv.p, lv[nl - 1].b.p, lv[nl - 1].x.p, done);
  for (int l = nl - 2; l >= 0; --l) {
    DevLevel<T>& L = lv[l];
    launch_prolong<T>(s, L.view, cur[l], cur[l + 1], L.omega_p, done, l == 0);
    gauge_remove<T>(c, l, cur[l]);
    for (int q = 0; q < post; ++q) {
      launch_sweep<T>(s, L.view, cur[l], oth[l], L.b.p, L.omega, done, l == 0, ZStats{});
      std::swap(cur[l], oth[l]);
      gauge_remove<T>(c, l, cur[l]);
    }
  }
  if (zs.part) launch_zstats<T>(s, lv[0].view.nblocks, lv[0].view.bx, cur[0], zs, lv[0].view.volume, done);
  return cur[0];
}

template <typename T>
T* enqueue_vcycle_any(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, const int* done, const ZStats& zs,
                      bool x0_fused = false) {
  return c->H->general_gauge ? enqueue_vcycle_general<T>(c, lv, done, zs)
                             : enqueue_vcycle<T>(c, lv, done, zs, x0_fused);
}

// ---------------------------------------------------------------------------
// FP64 helpers on the finest level

LevelView<double>& fine64(tpamr_ctx* c) { return c->ld[0].view; }

void matvec(tpamr_ctx* c, const double* x, double* y) {
  launch_matvec_dot<double>(c->stream, fine64(c), x, y, nullptr, nullptr);
}

double reduce_host(tpamr_ctx* c, int n, int nsets, const double* part, bool is_max, double* out_sets = nullptr) {
  launch_reduce(c->stream, n, nsets, part, c->rs.scal, is_max);
  double h[16];
  HIPCHECK(hipMemcpyAsync(h, c->rs.scal, sizeof(double) * nsets, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (out_sets)
    for (int q = 0; q < nsets; ++q) out_sets[q] = h[q];
  return h[0];
}

int ncb(int n) { return (n + 255) / 256; }

// true residual r = b - A x; returns (l2, linf) of the strong residual
void true_residual(tpamr_ctx* c, double& l2, double& linf) {
  const int N = c->topo[0].ncells, nb = c->H->levels[0]->mesh->nb();
  matvec(c, c->x.p, c->Ap.p);
  launch_residual_strong(c->stream, N, nb, c->b.p, c->Ap.p, c->r.p, c->topo[0].volume.p, c->rs.part);
  double sets[2];
  reduce_host(c, ncb(N), 1, c->rs.part, true, sets);
  linf = sets[0];
  double s2[1];
  reduce_host(c, ncb(N), 1, c->rs.part + ncb(N), false, s2);
  l2 = std::sqrt(s2[0] / c->H->total_volume);
}

double roundoff_floor(tpamr_ctx* c) {
  const int nblk = c->topo[0].nblocks;
  launch_floor_bound(c->stream, fine64(c), c->x.p, c->b.p, c->rs.part);
  const double m = reduce_host(c, nblk, 1, c->rs.part, true);
  return 128 * EPS * std::max(1.0, m);
}

// The volume that divides the means the Krylov kernels remove (z, and the
// residual's compatibility drift): the total volume of a singular operator
// (one connected, ungauged component), infinity for one that open sides
// anchor, whose means are then 0 / inf = 0 and nothing is removed.
// A hierarchy with several components gauges them in the V-cycle itself,
// so the Krylov kernels remove nothing either.
double gauge_volume(const tpamr_ctx* c) {
  return (c->H->singular && !c->H->general_gauge) ? c->H->total_volume : std::numeric_limits<double>::infinity();
}

// subtract the volume-weighted mean (one connected, unanchored component;
// each ungauged component's own with several; nothing for an operator that
// open sides anchor)
void remove_mean(tpamr_ctx* c, double* v) {
  if (c->H->general_gauge) {
    gauge_remove<double>(c, 0, v);
    return;
  }
  if (!c->H->singular) return;
  const int N = c->topo[0].ncells, nb = c->H->levels[0]->mesh->nb();
  launch_weighted_sums(c->stream, N, nb, v, c->topo[0].volume.p, c->rs.part);
  launch_reduce(c->stream, ncb(N), 1, c->rs.part, c->rs.scal, false);
  launch_shift(c->stream, N, nb, v, c->rs.scal, 1.0 / c->H->total_volume, c->topo[0].volume.p, false);
}

double guarded_target(tpamr_ctx* c, double target, double initial, double current) {
  if (!(std::isfinite(initial) && std::isfinite(current) && target < current && current < initial)) return target;
  const double fl = roundoff_floor(c);
  return (std::isfinite(fl) && fl < initial / 100) ? std::max(target, fl) : target;
}

// The Krylov iterations leave the finest level's zero-start iterate to the
// V-cycle's first pre-sweep (it needs a second pre-sweep to fuse into, and
// the level-by-level V-cycle of several components projects b first).
bool x0_fused(const tpamr_ctx* c) { return c->cfg.pre >= 2 && !c->H->general_gauge; }

// enqueue z = M r for the Krylov loop: V-cycle input from r, z stats partials
template <typename T>
T* enqueue_precondition(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, const int* done, bool fused = false) {
  ZStats zs;
  zs.r64 = c->r.p;
  zs.part = c->rs.part;
  return enqueue_vcycle_any<T>(c, lv, done, zs, fused);
}

template <typename T>
void enqueue_iteration(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, bool init) {
  hipStream_t s = c->stream;
  const int N = c->topo[0].ncells, nb = c->H->levels[0]->mesh->nb(), nblk = c->topo[0].nblocks;
  const int* done = &c->rs.st->done;
  DevLevel<T>& L0 = lv[0];
  const double w0 = c->cfg.pre > 0 ? L0.omega : 0.0;
  const bool fused = x0_fused(c);
  T* x0 = fused ? nullptr : L0.x.p;
  if (init) {
    matvec(c, c->x.p, c->Ap.p);
    launch_copy(s, N, c->b.p, c->r.p);
    launch_cg_update_xr<T>(s, N, nb, nullptr, c->Ap.p, nullptr, c->r.p, c->topo[0].volume.p, L0.b.p, x0, L0.dinv.p,
                           w0, c->rs.maxpart, c->rs.st, 1);
  } else {
    launch_matvec_dot<double>(s, fine64(c), c->p.p, c->Ap.p, c->rs.part, done);
    launch_cg_alpha(s, nblk, c->rs.part, c->rs.stage, c->rs.st);
    launch_cg_update_xr<T>(s, N, nb, c->p.p, c->Ap.p, c->x.p, c->r.p, c->topo[0].volume.p, L0.b.p, x0, L0.dinv.p,
                           w0, c->rs.maxpart, c->rs.st, 0);
  }
  T* z = enqueue_precondition<T>(c, lv, done, fused);
  launch_cg_beta(s, nblk, ncb(N), c->rs.part, c->rs.maxpart, c->rs.stage, gauge_volume(c), c->rs.st, init ? 1 : 0);
  launch_cg_update_p<T>(s, N, z, c->p.p, c->rs.st, done);
}

// Capture `graph_iters` non-initial Krylov iterations into `graph`.
template <class Enq>
void capture_iteration(tpamr_ctx* c, hipGraphExec_t& graph, Enq&& enqueue, bool init = false) {
  if (graph) {
    (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
  hipGraph_t g;
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  if (init)
    enqueue(true);
  else
    for (int k = 0; k < c->graph_iters; ++k) enqueue(false);
  HIPCHECK(hipStreamEndCapture(c->stream, &g));
  HIPCHECK(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
  HIPCHECK(hipGraphDestroy(g));
}

// Run Krylov iterations until the device-side done flag (strong residual
// callback, iteration limit or breakdown). `enqueue(init)` enqueues one
// iteration; the state must already hold target and itmax. Returns the
// iteration count.
template <class Enq>
int run_iterations(tpamr_ctx* c, hipGraphExec_t& graph0, hipGraphExec_t& graph, int itmax, Enq&& enqueue) {
  hipStream_t s = c->stream;
  if (c->use_graph) {
    if (!graph0) capture_iteration(c, graph0, enqueue, true);
    HIPCHECK(hipGraphLaunch(graph0, s));
    if (!graph) capture_iteration(c, graph, enqueue);
  } else {
    enqueue(true);
  }
  // `pipe` iterations stay queued: the host reads the done flag of the
  // oldest one while the GPU runs the others; iterations queued after the
  // last one exit at once on the device-side done flag.
  const int P = c->pipe;
  const int per = c->use_graph ? c->graph_iters : 1;
  int launched = 0;
  for (;;) {
    if (c->use_graph)
      HIPCHECK(hipGraphLaunch(graph, s));
    else
      enqueue(false);
    HIPCHECK(hipMemcpyAsync(c->st_ring + launched % P, c->rs.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipEventRecord(c->evs[launched % P], s));
    ++launched;
    if (launched >= P) {
      const int q = (launched - P) % P;
      HIPCHECK(hipEventSynchronize(c->evs[q]));
      if (c->trace) {
        const CgState& t = c->st_ring[q];
        std::fprintf(stderr, "tpamr: iter %d strong %.6e gamma %.6e rlast %.6e replace %d replaced %d done %d\n",
                     t.iter, t.strong, t.gamma, t.rlast, t.replace, t.replaced, t.done);
      }
      if (c->st_ring[q].done) break;
    }
    if (launched * per > itmax + P * per) break;
  }
  HIPCHECK(hipStreamSynchronize(s));
  HIPCHECK(hipMemcpy(c->st_host, c->rs.st, sizeof(CgState), hipMemcpyDeviceToHost));
  return c->st_host->iter;
}

int reset_state(tpamr_ctx* c, double target, int itmax, double scale = 1.0, double delta = 0.0) {
  const int N = c->topo[0].ncells;
  CgState init{};
  init.target = target;
  init.itmax = itmax == 0 ? 2 * N : itmax;
  init.scale = scale;
  init.delta = delta;
  init.keep = c->ir_keep ? 1 : 0;
  init.rlast = 1.0;
  *c->st_host = init;
  HIPCHECK(hipMemcpyAsync(c->rs.st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
  return init.itmax;
}

// One run of Krylov CG (warm start from x) until the strong residual
// callback, the iteration limit, or breakdown. Returns the iteration count.
template <typename T>
int run_cg(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, double target, int itmax) {
  const int n = reset_state(c, target, itmax);
  launch_fill(c->stream, c->topo[0].ncells, 0.0, c->p.p);
  return run_iterations(c, c->graph0, c->graph, n, [&](bool init) { enqueue_iteration<T>(c, lv, init); });
}

// One inner FP32 Krylov iteration of the refinement mode: e (correction
// since the last replacement), r (the finest level's V-cycle rhs), p and Ap
// are FP32; pAp, the z statistics and every scalar are FP64. init: e = 0,
// r = scale * c->r. When the previous iteration requested it, the FP32
// correction moves into the FP64 solution c->x (in the Krylov update), r is
// replaced by scale * (b - A x) in FP64 and its strong norm decides the stop
// before the V-cycle (the replacement kernels exit at once otherwise).
void enqueue_ir_iteration(tpamr_ctx* c, std::vector<DevLevel<float>>& lv, bool init) {
  hipStream_t s = c->stream;
  const int N = c->topo[0].ncells, nb = c->H->levels[0]->mesh->nb(), nblk = c->topo[0].nblocks;
  const int* done = &c->rs.st->done;
  DevLevel<float>& L0 = lv[0];
  const double w0 = c->cfg.pre > 0 ? L0.omega : 0.0;
  if (!init) {
    launch_matvec_dot<float>(s, L0.view, c->p32.p, c->Ap32.p, c->rs.part, done, true);
    launch_cg_alpha(s, nblk, c->rs.part, c->rs.stage, c->rs.st, true, gauge_volume(c));
  }
  const bool fused = x0_fused(c);
  // a requested replacement first moves the FP32 correction into the FP64
  // solution (in the update's pass)
  launch_ir_update_xr(s, N, nb, c->p32.p, c->Ap32.p, c->e32.p, L0.b.p, c->topo[0].volume.p, fused ? nullptr : L0.x.p,
                      L0.dinv.p, w0, c->r.p, c->rs.maxpart, c->rs.st, init ? 1 : 0, init ? nullptr : c->x.p);
  if (!init) {
    launch_ir_replace(s, fine64(c), c->x.p, c->b.p, L0.b.p, fused ? nullptr : L0.x.p, L0.dinv.p, w0, c->rs.maxpart,
                      c->rs.st);
    launch_ir_check(s, ir_groups(N), c->rs.maxpart, c->rs.stage, c->rs.st);
  }
  ZStats zs;
  zs.r32 = L0.b.p;
  zs.part = c->rs.part;
  float* z = enqueue_vcycle_any<float>(c, lv, done, zs, fused);
  launch_cg_beta(s, nblk, ir_groups(N), c->rs.part, c->rs.maxpart, c->rs.stage, gauge_volume(c), c->rs.st,
                 init ? 1 : 0);
  launch_ir_update_p(s, N, z, c->p32.p, c->rs.st, done);
}

// One run of the refinement mode's Krylov iteration from the FP64 residual
// in c->r (scaled by `scale` into FP32) until the strong residual reaches
// `target` (scaled units). Returns the iteration count; the correction since
// the last replacement is left in c->e32.
int run_ir(tpamr_ctx* c, std::vector<DevLevel<float>>& lv, double target, double scale, int itmax) {
  const int n = reset_state(c, target, itmax, scale, c->ir_delta);
  launch_fill_f32(c->stream, c->topo[0].ncells, 0.f, c->p32.p);
  return run_iterations(c, c->graph0_ir, c->graph_ir, n, [&](bool init) { enqueue_ir_iteration(c, lv, init); });
}

template <typename T>
tpamr_stats solve_impl(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, double rtol, double atol, int max_cycles,
                       int compat) {
  hipStream_t s = c->stream;
  const int N = c->topo[0].ncells, nb = c->H->levels[0]->mesh->nb();
  const double* vol = c->topo[0].volume.p;
  tpamr_stats st{};
  HIPCHECK(hipEventRecord(c->ev0, s));
  // integrated rhs, its sums for the compatible projection, and whether the
  // initial solution is zero (one pass, one host synchronisation)
  launch_rhs_prepare(s, N, nb, c->b.p, c->x.p, vol, c->rs.part);
  double sums[4];
  {
    launch_reduce(s, ncb(N), 4, c->rs.part, c->rs.scal, false);
    HIPCHECK(hipMemcpyAsync(sums, c->rs.scal, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
  }
  const double scale = sums[0], total_sum = sums[1];
  const bool zero_start = sums[3] == 0.0;
  // sums of |b| and |x| are finite exactly when every entry is (device
  // inputs are not scanned on the host)
  if (!std::isfinite(scale) || !std::isfinite(total_sum)) throw std::invalid_argument("right-hand side must be finite");
  if (!std::isfinite(sums[3])) throw std::invalid_argument("initial solution must be finite");
  // the compatible projection (_project_component_rhs!,
  // elliptic_solver.jl:150-190) of a singular operator; an anchored one
  // takes the rhs as it is
  const bool singular = c->H->singular, general = singular && c->H->general_gauge;
  if (singular && !general && compat == 1 && std::fabs(total_sum) > 256 * EPS * std::max(1.0, scale))
    throw IncompatibleError("Poisson right-hand side has nonzero volume mean on component rooted at cell 1");
  const double mean = (singular && !general) ? total_sum / c->H->total_volume : 0.0;
  st.removed_rhs_mean = mean;
  c->last_roots.clear();
  c->last_means.clear();
  if (singular && !general) {
    c->last_roots.push_back(1);
    c->last_means.push_back(mean);
  }
  if (general) {
    // each ungauged component's own projection (_project_component_rhs!,
    // elliptic_solver.jl:150-190)
    const LevelData& L0 = *c->H->levels[0];
    DevGauge& G = c->gauge[0];
    const int K = G.ncomp;
    std::vector<double> sums(K), scales(K);
    launch_comp_sums<double>(s, G.view, N, nb, c->b.p, vol, 0, 0, G.mean.p);
    HIPCHECK(hipMemcpyAsync(sums.data(), G.mean.p, sizeof(double) * K, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    launch_comp_sums<double>(s, G.view, N, nb, c->b.p, vol, 2, 0, G.mean.p);
    HIPCHECK(hipMemcpyAsync(scales.data(), G.mean.p, sizeof(double) * K, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    double tsum = 0, tvol = 0;
    for (int q = 0; q < K; ++q) {
      if (compat == 1 && std::fabs(sums[q]) > 256 * EPS * std::max(1.0, scales[q]))
        throw IncompatibleError("Poisson right-hand side has nonzero volume mean on component rooted at cell " +
                                std::to_string(L0.comp_root[q] + 1));
      tsum += sums[q];
      tvol += L0.comp_vol[q];
    }
    for (int q = 0; q < K; ++q) {
      sums[q] /= L0.comp_vol[q];
      c->last_roots.push_back(L0.comp_root[q] + 1);
      c->last_means.push_back(sums[q]);
    }
    HIPCHECK(hipMemcpyAsync(G.mean.p, sums.data(), sizeof(double) * K, hipMemcpyHostToDevice, s));
    launch_comp_shift<double>(s, G.view, N, nb, G.mean.p, vol, 1, c->b.p);
    HIPCHECK(hipStreamSynchronize(s));
    st.removed_rhs_mean = tvol == 0 ? 0.0 : tsum / tvol;
  }
  double il2, ilinf, floor0;
  if (zero_start) {
    // b -= V mean; the residual of x = 0 is b and its round-off floor bound
    // reduces to max |b / V|
    launch_shift_residual(s, N, nb, c->b.p, mean, vol, c->r.p, c->rs.part);
    double sets[2];
    launch_reduce(s, ncb(N), 1, c->rs.part, c->rs.scal, true);
    launch_reduce(s, ncb(N), 1, c->rs.part + ncb(N), c->rs.scal + 1, false);
    HIPCHECK(hipMemcpyAsync(sets, c->rs.scal, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    ilinf = sets[0];
    il2 = std::sqrt(sets[1] / c->H->total_volume);
    floor0 = 128 * EPS * std::max(1.0, ilinf);
  } else {
    // b -= V * mean (scalar staged through device memory)
    if (singular && !general) {
      double h = mean;
      HIPCHECK(hipMemcpyAsync(c->rs.scal + 8, &h, sizeof(double), hipMemcpyHostToDevice, s));
      launch_shift(s, N, nb, c->b.p, c->rs.scal + 8, 1.0, vol, true);
    }
    remove_mean(c, c->x.p);
    true_residual(c, il2, ilinf);
    floor0 = roundoff_floor(c);
  }
  double target = std::max(atol, std::max(rtol * std::max(ilinf, EPS), floor0));
  // final residual norms and the guarded target of the last check
  double fl2 = il2, flinf = ilinf, ftarget = target;
  int iterations = 0;
  int reason = 0;
  int runs = 0, replaced = 0;
  constexpr bool fp32 = std::is_same<T, float>::value;
  if (ilinf > target && fp32 && c->precision == TPAMR_FP32_REFINE) {
    // FP32 Krylov CG with FP64 residual replacement; each run starts from
    // the FP64 residual scaled to a unit strong norm, and its end is checked
    // against the FP64 true residual (the reference's restart-and-verify
    // loop).
    int total = 0;
    double linf = ilinf;
    while (total < max_cycles) {
      ++runs;
      const double scale = 1.0 / linf;
      const int it = run_ir(c, c->lf, target * scale, scale, max_cycles - total);
      reason = c->st_host->reason;
      replaced += c->st_host->replaced;
      total += it;
      launch_ir_correct(s, N, c->e32.p, c->x.p, c->rs.st, true);
      remove_mean(c, c->x.p);
      true_residual(c, fl2, linf);
      flinf = linf;
      ftarget = guarded_target(c, target, ilinf, linf);
      if (linf <= ftarget) break;
      if (!std::isfinite(linf) || it == 0) break;
    }
    iterations = total;
  } else if (ilinf > target) {
    int total = 0;
    while (total < max_cycles) {
      ++runs;
      const int it = run_cg<T>(c, lv, target, max_cycles - total);
      reason = c->st_host->reason;
      total += it;
      remove_mean(c, c->x.p);
      true_residual(c, fl2, flinf);
      ftarget = guarded_target(c, target, ilinf, flinf);
      if (flinf <= ftarget) break;
      if (!std::isfinite(flinf) || it == 0) break;
    }
    iterations = total;
  }
  target = ftarget;
  HIPCHECK(hipEventRecord(c->ev1, s));
  HIPCHECK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  st.cycles = iterations;
  st.converged = std::isfinite(flinf) && std::isfinite(target) && flinf <= target;
  st.compatible = 1;
  st.reason = reason;
  st.initial_l2 = il2;
  st.initial_linf = ilinf;
  st.final_l2 = fl2;
  st.final_linf = flinf;
  st.target_linf = target;
  st.residual_scale = std::max(ilinf, EPS);
  st.solve_ms = ms;
  st.krylov_runs = runs;
  st.replacements = replaced;
  return st;
}

tpamr_stats solve_dispatch(tpamr_ctx* c, double rtol, double atol, int max_cycles, int compat) {
  if (!(rtol >= 0 && rtol < 1)) throw std::invalid_argument("rtol must lie in [0, 1)");
  if (!(atol >= 0)) throw std::invalid_argument("atol must be nonnegative");
  if (max_cycles < 0) throw std::invalid_argument("max_cycles must be nonnegative");
  if (compat != 0 && compat != 1) throw std::invalid_argument("compatibility must be :project or :error");
  if (c->cfg.pre != c->cfg.post)
    throw std::invalid_argument("CG requires equal pre- and post-smoothing sweep counts");
  if (c->fp32()) return solve_impl<float>(c, c->lf, rtol, atol, max_cycles, compat);
  return solve_impl<double>(c, c->ld, rtol, atol, max_cycles, compat);
}

// z = ldiv!(P, r) of the V-cycle preconditioner on device vectors
// (integrated r, rewritten with its own values; z with zero volume mean
// where the operator is singular), enqueued on the context stream
template <typename T>
void vcycle_into(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, double* r, double* zout) {
  hipStream_t s = c->stream;
  const int N = c->topo[0].ncells, nb = c->H->levels[0]->mesh->nb();
  CgState init{};
  init.scale = 1.0;
  *c->st_host = init;
  HIPCHECK(hipMemcpyAsync(c->rs.st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, s));
  DevLevel<T>& L0 = lv[0];
  const double w0 = c->cfg.pre > 0 ? L0.omega : 0.0;
  launch_fill(s, N, 0.0, c->Ap.p);
  launch_cg_update_xr<T>(s, N, nb, nullptr, c->Ap.p, nullptr, r, c->topo[0].volume.p, L0.b.p, L0.x.p,
                         L0.dinv.p, w0, c->rs.maxpart, c->rs.st, 1);
  T* z = enqueue_vcycle_any<T>(c, lv, &c->rs.st->done, ZStats{});
  if constexpr (std::is_same<T, float>::value)
    launch_cast_f2d(s, N, z, zout);
  else
    launch_copy(s, N, z, zout);
  remove_mean(c, zout);
}

// r in c->r (integrated); z into c->tmp
template <typename T>
void vcycle_impl(tpamr_ctx* c, std::vector<DevLevel<T>>& lv) {
  vcycle_into<T>(c, lv, c->r.p, c->tmp.p);
  HIPCHECK(hipStreamSynchronize(c->stream));
}

size_t upload_vec(tpamr_ctx* c, double* dst, const double* src, size_t n) {
  HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return n;
}

void download_vec(tpamr_ctx* c, double* dst, const double* src, size_t n) {
  HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}

void check_finite(const double* v, size_t n, const char* what) {
  for (size_t q = 0; q < n; ++q)
    if (!std::isfinite(v[q])) throw std::invalid_argument(std::string(what) + " must be finite");
}

// metric products on face vectors: out = sqrt(k) K (sqrt(k) in)
void weighted_metric(tpamr_ctx* c, const double* in_scaled, double* out) {
  const int F = c->mesh().nfaces();
  launch_metric_apply(c->stream, F, c->kdiag.p, c->kptr.p, c->kface.p, c->kval.p, in_scaled, c->f4.p);
  launch_face_scale(c->stream, F, c->sqk.p, c->f4.p, out);
}

// max |I - T| of the N x N product in c->coarse_t
double identity_residual(tpamr_ctx* c, int N) {
  hipStream_t s = c->stream;
  c->coarse_res.alloc(1);
  HIPCHECK(hipMemsetAsync(c->coarse_res.p, 0, sizeof(unsigned long long), s));
  launch_identity_residual(s, N, c->coarse_t.p, c->coarse_res.p);
  unsigned long long bits = 0;
  HIPCHECK(hipMemcpyAsync(&bits, c->coarse_res.p, sizeof(bits), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  double r;
  std::memcpy(&r, &bits, sizeof(r));
  return r;
}

void blas_check(rocblas_status st) {
  if (st != rocblas_status_success) throw HipError("rocBLAS call failed (status " + std::to_string(int(st)) + ")");
}

void blas_ready(tpamr_ctx* c) {
  if (!c->blas) {
    blas_check(rocblas_create_handle(&c->blas));
    blas_check(rocblas_set_pointer_mode(c->blas, rocblas_pointer_mode_host));
  }
  blas_check(rocblas_set_stream(c->blas, c->stream));
}

// T = A X for row-major N x N A (c->coarse_work) and X (c->coarse_x): the
// column-major product X^T A^T of the same buffers
void coarse_product(tpamr_ctx* c, int N) {
  const double one = 1.0, zero = 0.0;
  blas_check(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, N, N, N, &one, c->coarse_x.p, N,
                           c->coarse_work.p, N, &zero, c->coarse_t.p, N));
}

// Newton-Schulz update X <- X (2 I - A X) of the previous refresh's inverse,
// while its identity residual falls; true when the residual reaches the
// level of a Gauss-Jordan inverse (4x its residual, at least 1e-13).
bool coarse_newton_schulz(tpamr_ctx* c, int N) {
  hipStream_t s = c->stream;
  blas_ready(c);
  c->coarse_t.alloc(size_t(N) * N);
  c->coarse_xn.alloc(size_t(N) * N);
  const double accept = std::max(1e-13, 4 * c->coarse_gj_residual);
  double prev = 0;
  for (int it = 0; it < 6; ++it) {
    coarse_product(c, N);
    const double r = identity_residual(c, N);
    if (r <= accept) return true;
    if (!(r < 0.25) || (it > 0 && !(r < 0.25 * prev))) return false;
    prev = r;
    // X_n = 2 X - X T (row major) = 2 X^T - T^T X^T (column major)
    const double m1 = -1.0, two = 2.0;
    HIPCHECK(hipMemcpyAsync(c->coarse_xn.p, c->coarse_x.p, sizeof(double) * N * size_t(N), hipMemcpyDeviceToDevice,
                            s));
    blas_check(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, N, N, N, &m1, c->coarse_t.p, N,
                             c->coarse_x.p, N, &two, c->coarse_xn.p, N));
    std::swap(c->coarse_x.p, c->coarse_xn.p);
  }
  return false;
}

// Blocked Gauss-Jordan with partial pivoting of the N x N row-major matrix
// in c->coarse_x (in place; its n x n leading block of the inverse goes to
// c->coarse_inv, column major): per panel of GJ_PANEL columns, the panel's
// own elimination (two small kernels per column), its row swaps on the other
// columns, and their update A_rest += (E_panel - J) A_rest[panel rows] as two
// DGEMMs (the row-major blocks seen column major: C^T += R^T M^T).
void dense_inverse_blocked(tpamr_ctx* c, int N, int n) {
  constexpr int GJ_PANEL = 64;
  hipStream_t s = c->stream;
  blas_ready(c);
  c->gj_R.alloc(size_t(GJ_PANEL) * N);
  c->gj_M.alloc(size_t(N) * GJ_PANEL);
  double* a = c->coarse_x.p;
  const double one = 1.0;
  for (int c0 = 0; c0 < N; c0 += GJ_PANEL) {
    const int nb = std::min(GJ_PANEL, N - c0), c1 = c0 + nb;
    launch_gjb_panel(s, N, c0, nb, a, c->coarse_piv.p, c->coarse_col.p, c->coarse_fail.p);
    launch_gjb_prep(s, N, c0, nb, a, c->gj_R.p, c->gj_M.p);
    if (c0 > 0)
      blas_check(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c0, N, nb, &one, c->gj_R.p,
                               N, c->gj_M.p, nb, &one, a, N));
    if (c1 < N)
      blas_check(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, N - c1, N, nb, &one,
                               c->gj_R.p + c1, N, c->gj_M.p, nb, &one, a + c1, N));
  }
  launch_gj_finish(s, N, a, c->coarse_piv.p, n, c->coarse_inv.p);
}

// Dense coarse inverse of H.coarse_aug on the device into c->coarse_inv: a
// Newton-Schulz update of the previous refresh's inverse when it converges
// (the coefficients of consecutive steps differ little), else Gauss-Jordan.
void coarse_inverse_device(tpamr_ctx* c) {
  Hierarchy& H = *c->H;
  hipStream_t s = c->stream;
  const int n = H.coarse_n, N = H.coarse_N;
  const size_t NN = size_t(N) * N;
  c->coarse_work.upload(H.coarse_aug, s);
  c->coarse_inv.alloc(size_t(n) * n);
  if (c->coarse_x_valid && c->coarse_x.n == NN && !c->coarse_exact && coarse_newton_schulz(c, N)) {
    launch_extract_block(s, n, N, c->coarse_x.p, c->coarse_inv.p);
    return;
  }
  c->coarse_col.alloc(N);
  c->coarse_piv.alloc(N);
  c->coarse_fail.alloc(1);
  c->coarse_x.alloc(NN);
  // Gauss-Jordan on a copy; A stays in coarse_work for the residual
  HIPCHECK(hipMemcpyAsync(c->coarse_x.p, c->coarse_work.p, sizeof(double) * NN, hipMemcpyDeviceToDevice, s));
  HIPCHECK(hipMemsetAsync(c->coarse_fail.p, 0, sizeof(int), s));
  dense_inverse_blocked(c, N, n);
  int fail = 0;
  HIPCHECK(hipMemcpyAsync(&fail, c->coarse_fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  if (fail) {
    c->coarse_x_valid = false;
    throw std::runtime_error("singular coarse operator");
  }
  c->coarse_x_valid = true;
  if (!c->coarse_exact) {
    blas_ready(c);
    c->coarse_t.alloc(NN);
    coarse_product(c, N);
    c->coarse_gj_residual = identity_residual(c, N);
  }
}

void upload_plan(tpamr_ctx* c) {
  Hierarchy& H = *c->H;
  hipStream_t s = c->stream;
  const int nl = int(H.levels.size());
  c->plan.clear();
  c->plan.resize(nl);
  for (int l = 0; l < nl; ++l) {
    const LevelData& L = *H.levels[l];
    const TreeMesh& M = *L.mesh;
    DevPlan& P = c->plan[l];
    const int F = M.nfaces(), bx = M.bs[0], by = M.bs[1], nb = bx * by;
    P.nfaces = F;
    P.nrows = int(L.irr_local.size());
    if (l > 0) {
      P.rptr.upload(L.rptr, s);
      if (L.ridx.empty()) {
        P.ridx.upload(std::vector<int>(1, 0), s);
        P.rw.upload(std::vector<double>(1, 0.0), s);
      } else {
        P.ridx.upload(L.ridx, s);
        P.rw.upload(L.rw, s);
      }
    }
    P.kdiag.upload(L.K.diag, s);
    P.area.upload(M.farea, s);
    P.kappa.alloc(F);
    // regular faces: slot in cx / cy -> code index * 4 + array in the device layout
    std::vector<int> dst(2 * size_t(F), -1);
    auto code = [&](int axis, int64_t slot) {
      if (axis == 0) {
        const int64_t per = int64_t(bx + 1) * by, b = slot / per, rem = slot % per, j = rem / (bx + 1),
                      i = rem % (bx + 1);
        return i < bx ? int((b * nb + j * bx + i) * 4 + 0) : int((b * by + j) * 4 + 1);
      }
      const int64_t per = int64_t(bx) * (by + 1), b = slot / per, rem = slot % per, j = rem / bx, i = rem % bx;
      return j < by ? int((b * nb + j * bx + i) * 4 + 2) : int((b * bx + i) * 4 + 3);
    };
    for (int f = 0; f < F; ++f)
      if (L.face_regular[f]) {
        dst[2 * size_t(f)] = code(M.faxis[f], L.slot_a[f]);
        dst[2 * size_t(f) + 1] = code(M.faxis[f], L.slot_b[f]);
      }
    P.dst.upload(dst, s);
    // empty arrays upload one zero (kernels take valid pointers)
    auto up = [&](auto& dst, const auto& v, auto z) {
      if (v.empty()) dst.upload(std::vector<decltype(z)>(1, z), s);
      else dst.upload(v, s);
    };
    P.irr_ptr.upload(L.irr_ptr, s);
    P.term_ptr.upload(L.term_ptr, s);
    up(P.term_f, L.term_f, 0);
    up(P.term_g, L.term_g, 0);
    up(P.term_k, L.term_k, 0.0);
    up(P.term_s, L.term_s, 0.0);
    up(P.term_diag, L.term_diag, (unsigned char)0);
    up(P.ell_pos, L.ell_pos, 0);
    up(P.ell_dpos, L.ell_dpos, 0);
    up(P.ell_row, L.ell_row, 0);
    up(P.diag_e, L.irr_diag_e, 0);
    P.cell_irr.upload(L.cell_irr, s);
    P.irr_diag.alloc(std::max(P.nrows, 1));
    P.nnz = int(L.irr_col.size());
    P.val_e.alloc(std::max(P.nnz, 1));
    if (l >= int(c->ld.size())) {
      P.cw.upload(L.cw, s);
      P.ce.upload(L.ce, s);
      P.cs.upload(L.cs, s);
      P.cn.upload(L.cn, s);
      up(P.ell_val, L.ell_val, 0.0);
      up(P.ell_diag, L.ell_diag, 0.0);
      P.dinv.upload(L.dinv, s);
    }
    P.open = L.any_open();
    P.nopen = int(L.bopen.size());
    up(P.bdst, L.bdst, 0);
  }
  c->ref_ratio.alloc(nl);
  c->ref_flags.alloc(nl + 2);
  c->plan_ready = true;
  c->kb_dirty = true;
}

// The boundary values of every level (conductances and row boundary
// diagonals; they change with the boundary coefficients only) into the
// device refresh plan.
void upload_plan_boundary(tpamr_ctx* c) {
  Hierarchy& H = *c->H;
  for (size_t l = 0; l < H.levels.size(); ++l) {
    const LevelData& L = *H.levels[l];
    DevPlan& P = c->plan[l];
    if (!P.open) continue;
    if (L.bw.size() != L.bopen.size()) throw std::logic_error("open pressure sides without boundary coefficients");
    if (L.bw.empty()) P.bw.upload(std::vector<double>(1, 0.0), c->stream);
    else P.bw.upload(L.bw, c->stream);
    if (L.irr_bd.empty()) P.irr_bd.upload(std::vector<double>(1, 0.0), c->stream);
    else P.irr_bd.upload(L.irr_bd, c->stream);
  }
  c->kb_dirty = false;
}

// A coefficient refresh computed on the device (topology unchanged): the
// value pass of every level, FP32 copies, the finest level's sqrt(k), the
// Jacobi weights (one read-back) and the dense coarse inverse. The host keeps
// the coarsest level's values (for its dense matrix); the other levels' host
// values are refreshed only when asked for (host_stale). Returns false, with
// nothing changed, when a face coefficient is zero: the components of such
// an operator are labelled by the host refresh.
bool device_refresh(tpamr_ctx* c, const double* kappa, bool kappa_on_device = false) {
  Hierarchy& H = *c->H;
  hipStream_t s = c->stream;
  const int nl = int(H.levels.size());
  if (!c->plan_ready) upload_plan(c);
  if (int(c->plan.size()) != nl || c->plan[0].nfaces != c->mesh().nfaces() || int(c->topo.size()) != nl)
    throw StateError("device refresh plan does not match the current hierarchy");
  if (c->kb_dirty) upload_plan_boundary(c);
  DevPlan& P0 = c->plan[0];
  if (kappa_on_device)
    HIPCHECK(hipMemcpyAsync(P0.kappa.p, kappa, sizeof(double) * P0.nfaces, hipMemcpyDeviceToDevice, s));
  else
    upload_vec(c, P0.kappa.p, kappa, P0.nfaces);
  HIPCHECK(hipMemsetAsync(c->ref_ratio.p, 0, sizeof(unsigned long long) * nl, s));
  HIPCHECK(hipMemsetAsync(c->ref_flags.p, 0, sizeof(int) * (nl + 2), s));
  launch_ref_check(s, P0.nfaces, P0.kappa.p, c->ref_flags.p + nl);
  int bad[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(bad, c->ref_flags.p + nl, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  if (bad[0]) throw std::invalid_argument("face coefficients must be finite and nonnegative");
  if (bad[1]) return false;
  destroy_graph(c);
  for (int l = 0; l < nl; ++l) {
    const LevelData& L = *H.levels[l];
    DevPlan& P = c->plan[l];
    const int bx = L.mesh->bs[0], by = L.mesh->bs[1];
    if (l > 0) launch_ref_restrict(s, P.nfaces, P.rptr.p, P.ridx.p, P.rw.p, c->plan[l - 1].kappa.p, P.kappa.p);
    const bool has64 = l < int(c->ld.size());
    DevLevel<double>* D = has64 ? &c->ld[l] : nullptr;
    double* cw = has64 ? D->cw.p : P.cw.p;
    double* ce = has64 ? D->ce.p : P.ce.p;
    double* cs = has64 ? D->cs.p : P.cs.p;
    double* cn = has64 ? D->cn.p : P.cn.p;
    double* ev = has64 ? D->ell_val.p : P.ell_val.p;
    double* ed = has64 ? D->ell_diag.p : P.ell_diag.p;
    double* di = has64 ? D->dinv.p : P.dinv.p;
    double* eb = (has64 && l == 0) ? D->ell_bound.p : nullptr;
    double* ebd = (has64 && l == 0) ? D->ell_bdiag.p : nullptr;
    launch_fill(s, int(L.cw.size()), 0.0, cw);
    launch_fill(s, int(L.ce.size()), 0.0, ce);
    launch_fill(s, int(L.cs.size()), 0.0, cs);
    launch_fill(s, int(L.cn.size()), 0.0, cn);
    launch_ref_regular(s, P.nfaces, P.kappa.p, P.kdiag.p, P.area.p, P.dst.p, cw, ce, cs, cn);
    if (P.open) launch_ref_boundary(s, P.nopen, P.bw.p, P.bdst.p, cw, ce, cs, cn);
    launch_ref_rows(s, P.nrows, P.irr_ptr.p, P.term_ptr.p, P.term_f.p, P.term_g.p, P.term_k.p, P.term_s.p,
                    P.term_diag.p, P.kappa.p, P.area.p, P.dst.p, P.ell_pos.p, P.ell_dpos.p, P.ell_row.p, P.diag_e.p,
                    ev, eb, ebd, ed, P.irr_diag.p, P.val_e.p, P.nnz, c->ref_ratio.p + l, P.open ? P.irr_bd.p : nullptr);
    launch_ref_dinv(s, L.ncells, bx, by, cw, ce, cs, cn, P.cell_irr.p, P.irr_diag.p, di, c->ref_flags.p + l,
                    P.open ? c->topo[l].nbr.p : nullptr, c->ref_ratio.p + l);
    if (l < int(c->lf.size())) {
      DevLevel<float>& Fv = c->lf[l];
      launch_cast_d2f(s, int(L.cw.size()), cw, Fv.cw.p);
      launch_cast_d2f(s, int(L.ce.size()), ce, Fv.ce.p);
      launch_cast_d2f(s, int(L.cs.size()), cs, Fv.cs.p);
      launch_cast_d2f(s, int(L.cn.size()), cn, Fv.cn.p);
      if (!L.ell_val.empty()) launch_cast_d2f(s, int(L.ell_val.size()), ev, Fv.ell_val.p);
      if (!L.ell_diag.empty()) launch_cast_d2f(s, int(L.ell_diag.size()), ed, Fv.ell_diag.p);
      launch_cast_d2f(s, L.ncells, di, Fv.dinv.p);
    }
  }
  HIPCHECK(hipMemcpyAsync(c->kappa.p, P0.kappa.p, sizeof(double) * P0.nfaces, hipMemcpyDeviceToDevice, s));
  launch_sqrt(s, P0.nfaces, P0.kappa.p, c->sqk.p);
  // Jacobi weights: 4 / (3 max ratio), the ratio at least 2 with a regular row
  std::vector<unsigned long long> ratio(nl);
  std::vector<int> regular(nl);
  HIPCHECK(hipMemcpyAsync(ratio.data(), c->ref_ratio.p, sizeof(unsigned long long) * nl, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(regular.data(), c->ref_flags.p, sizeof(int) * nl, hipMemcpyDeviceToHost, s));
  // the coarsest level's face coefficients, for its dense matrix on the host
  std::vector<double> kc(c->plan[nl - 1].nfaces);
  HIPCHECK(hipMemcpyAsync(kc.data(), c->plan[nl - 1].kappa.p, sizeof(double) * kc.size(), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  for (int l = 0; l < nl; ++l) {
    double mr;
    std::memcpy(&mr, &ratio[l], sizeof(double));
    if (regular[l]) mr = std::max(mr, 2.0);
    double w = 1.0, wp = 0.0;
    if (mr != 0) {
      wp = 4.0 / (3.0 * mr);
      w = std::min(1.0, wp);
    }
    LevelData& L = *H.levels[l];
    L.omega = w;
    L.omega_p = wp;
    if (l < int(c->ld.size())) {
      c->ld[l].omega = w;
      c->ld[l].omega_p = wp;
    }
    if (l < int(c->lf.size())) {
      c->lf[l].omega = w;
      c->lf[l].omega_p = wp;
    }
  }
  // positive coefficients: one connected component per level
  connected_components(H);
  upload_gauge(c);
  refresh_coarsest(H, kc.data());
  coarse_inverse_device(c);
  c->host_stale = true;
  return true;
}

// Pending boundary coefficients of open sides into the hierarchy (before its
// value pass): the caller's, or extended from the face coefficients
// (`_boundary_material`); `rebuilt`: a new hierarchy needs them again.
void apply_boundary(tpamr_ctx* c, const double* kappa_host, bool rebuilt) {
  if (!c->any_open() || (!c->kb_pending && !rebuilt)) return;
  const TreeMesh& M = *c->H->levels[0]->mesh;
  std::vector<double> kb = c->kb_extend ? boundary_material(M, c->open, kappa_host) : c->kb_user;
  if (kb.size() != M.bcell.size()) throw SizeError("one boundary coefficient is required per boundary face");
  set_boundary_coefficients(*c->H, kb.data());
  c->kb_user = std::move(kb);
  c->kb_extend = false;
  c->kb_pending = false;
  c->kb_dirty = true;
  c->bgrad_dirty = true;
}

// The finest level's boundary gradient factors 2 kb / h (0 on closed faces)
// on the device.
void upload_boundary_gradient(tpamr_ctx* c) {
  if (!c->bgrad_dirty || c->device < 0 || c->topo.empty()) return;
  const LevelData& L0 = *c->H->levels[0];
  const TreeMesh& M = *L0.mesh;
  const int BF = int(M.bcell.size()), nb = M.nb();
  std::vector<double> g(std::max(BF, 1), 0.0);
  if (L0.any_open())
    for (int f = 0; f < BF; ++f)
      if (L0.open[2 * M.baxis[f] + (M.bside[f] > 0 ? 1 : 0)])
        g[f] = 2 * L0.kb[f] / M.bspacing[2 * (M.bcell[f] / nb) + M.baxis[f]];
  c->bgrad.upload(g, c->stream);
  c->bgrad_dirty = false;
}

// Host copies of the level values after device refreshes (for inspection).
void ensure_host_values(tpamr_ctx* c) {
  if (!c->host_stale) return;
  std::vector<double> k0(c->plan[0].nfaces);
  download_vec(c, k0.data(), c->kappa.p, k0.size());
  refresh_coefficients(*c->H, k0.data());
  c->host_stale = false;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int tpamr_create(tpamr_ctx** out, int device, int precision) {
  if (!out) return TPAMR_EINVAL;
  *out = nullptr;
  auto c = std::make_unique<tpamr_ctx>();
  if (precision != TPAMR_FP64 && precision != TPAMR_FP32_VCYCLE && precision != TPAMR_FP32_REFINE)
    return TPAMR_EINVAL;
  c->precision = precision;
  c->device = device;
  if (const char* g = std::getenv("TPAMR_GRAPH")) c->use_graph = std::strcmp(g, "0") != 0;
  if (const char* q = std::getenv("TPAMR_PIPE")) c->pipe = std::min(PIPE_MAX, std::max(2, std::atoi(q)));
  if (const char* q = std::getenv("TPAMR_TRACE")) c->trace = std::strcmp(q, "0") != 0;
  if (const char* q = std::getenv("TPAMR_HOST_REFRESH")) c->host_refresh = std::strcmp(q, "0") != 0;
  if (const char* q = std::getenv("TPAMR_COARSE_GJ")) c->coarse_exact = std::strcmp(q, "0") != 0;
  if (const char* q = std::getenv("TPAMR_GRAPH_ITERS")) c->graph_iters = std::min(16, std::max(1, std::atoi(q)));
  if (const char* q = std::getenv("TPAMR_TAIL_BLOCKS")) c->tail_blocks = std::max(0, std::atoi(q));
  if (const char* q = std::getenv("TPAMR_IR_RESTART")) c->ir_keep = std::strcmp(q, "0") == 0;
  if (const char* q = std::getenv("TPAMR_IR_DELTA")) {
    const double v = std::atof(q);
    if (v > 0 && v < 1) c->ir_delta = v;
  }
  if (device >= 0) {
    int rc = guard(c.get(), [&] {
      int n = 0;
      HIPCHECK(hipGetDeviceCount(&n));
      if (device >= n) throw HipError("device index out of range");
      HIPCHECK(hipSetDevice(device));
      HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      HIPCHECK(hipMalloc(&c->st_own, sizeof(CgState)));
      c->rs.st = c->st_own;
      HIPCHECK(hipHostMalloc(&c->st_host, sizeof(CgState), hipHostMallocDefault));
      HIPCHECK(hipHostMalloc(&c->st_ring, PIPE_MAX * sizeof(CgState), hipHostMallocDefault));
      HIPCHECK(hipEventCreate(&c->ev0));
      HIPCHECK(hipEventCreate(&c->ev1));
      for (int q = 0; q < PIPE_MAX; ++q) HIPCHECK(hipEventCreateWithFlags(&c->evs[q], hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&c->ev_caller, hipEventDisableTiming));
      HIPCHECK(noalloc_scratch_address(device, &c->noalloc));
    });
    if (rc != TPAMR_OK) {
      // keep the context so the caller can read the message
      *out = c.release();
      return rc;
    }
  }
  *out = c.release();
  return TPAMR_OK;
}

void tpamr_destroy(tpamr_ctx* c) {
  if (!c) return;
  if (c->device >= 0) {
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->mom.reset();
    c->cmd.reset();
    destroy_graph(c);
    if (c->blas) (void)rocblas_destroy_handle(c->blas);
    c->topo.clear();
    c->ld.clear();
    c->lf.clear();
    if (c->st_own) (void)hipFree(c->st_own);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->st_ring) (void)hipHostFree(c->st_ring);
    for (hipEvent_t e : {c->ev0, c->ev1, c->ev_caller})
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->evs)
      if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

const char* tpamr_last_error(const tpamr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int tpamr_set_caller_stream(tpamr_ctx* c, void* stream) {
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    c->caller = static_cast<hipStream_t>(stream);
  });
}

int tpamr_set_mesh(tpamr_ctx* c, const int base_mesh[2], const int block_size[2], const double origin[2],
                   const double extent[2], const int periodic[2], int max_level, int n_refined, const int* refined) {
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (!base_mesh || !block_size || !origin || !extent || !periodic) throw std::invalid_argument("null argument");
    if (n_refined < 0 || (n_refined > 0 && !refined)) throw std::invalid_argument("invalid refined key list");
    if (!supported_block(block_size[0], block_size[1]))
      throw std::invalid_argument("block_size must be square with 4, 8, 16 or 32 cells per side");
    for (int a = 0; a < 2; ++a)
      if (!std::isfinite(origin[a]) || !std::isfinite(extent[a]))
        throw std::invalid_argument("domain origin and extent must be finite");
    std::vector<BlockKey> keys(n_refined);
    for (int q = 0; q < n_refined; ++q) keys[q] = {refined[3 * q], refined[3 * q + 1], refined[3 * q + 2]};
    const bool per[2] = {periodic[0] != 0, periodic[1] != 0};
    if (c->device >= 0) HIPCHECK(hipSetDevice(c->device));
    reset_topology_state(c);
    c->mom.reset();
    c->cmd.reset();
    if (c->H) c->H_prev = std::move(c->H);
    c->have_coeff = false;
    c->mesh_pending = make_tree_mesh(base_mesh, block_size, origin, extent, per, max_level, keys);
    // open sides stay; their boundary coefficients on the new mesh are
    // extended from its first face coefficients unless given again
    c->kb_user.clear();
    c->kb_extend = true;
    c->kb_pending = c->any_open();
  });
}

int tpamr_mesh_sizes(const tpamr_ctx* cc, int* sizes) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !sizes) return TPAMR_EINVAL;
  return guard(c, [&] {
    const TreeMesh& M = c->mesh();
    sizes[0] = int(M.blocks.size());
    sizes[1] = M.ncells();
    sizes[2] = M.nfaces();
    sizes[3] = int(M.bcell.size());
    sizes[4] = M.max_leaf;
  });
}

int tpamr_mesh_blocks(const tpamr_ctx* cc, int* keys) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !keys) return TPAMR_EINVAL;
  return guard(c, [&] {
    const TreeMesh& M = c->mesh();
    for (size_t b = 0; b < M.blocks.size(); ++b) {
      keys[3 * b] = M.blocks[b].level;
      keys[3 * b + 1] = M.blocks[b].i;
      keys[3 * b + 2] = M.blocks[b].j;
    }
  });
}

int tpamr_mesh_faces(const tpamr_ctx* cc, int* minus, int* plus, int* axis, double* area) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    const TreeMesh& M = c->mesh();
    for (int f = 0; f < M.nfaces(); ++f) {
      if (minus) minus[f] = M.fminus[f];
      if (plus) plus[f] = M.fplus[f];
      if (axis) axis[f] = M.faxis[f];
      if (area) area[f] = M.farea[f];
    }
  });
}

int tpamr_mesh_boundary_faces(const tpamr_ctx* cc, int* cell, int* axis, int* side, double* area) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    const TreeMesh& M = c->mesh();
    for (size_t f = 0; f < M.bcell.size(); ++f) {
      if (cell) cell[f] = M.bcell[f];
      if (axis) axis[f] = M.baxis[f];
      if (side) side[f] = M.bside[f];
      if (area) area[f] = M.barea[f];
    }
  });
}

int tpamr_set_multigrid(tpamr_ctx* c, const tpamr_multigrid_options* o) {
  if (!c || !o) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (o->coarse_size <= 0) throw std::invalid_argument("coarse_size must be positive");
    if (o->pre_sweeps < 0) throw std::invalid_argument("pre_sweeps must be nonnegative");
    if (o->post_sweeps < 0) throw std::invalid_argument("post_sweeps must be nonnegative");
    if (o->max_levels <= 0) throw std::invalid_argument("max_levels must be positive");
    const bool topo_change = !c->H || o->coarse_size != c->cfg.coarse_size || o->max_levels != c->cfg.max_levels;
    c->cfg.coarse_size = o->coarse_size;
    c->cfg.pre = o->pre_sweeps;
    c->cfg.post = o->post_sweeps;
    c->cfg.max_levels = o->max_levels;
    destroy_graph(c);
    if (c->H && topo_change) {
      // rebuild the hierarchy at the next coefficient update (nothing is
      // reused across different options)
      reset_topology_state(c);
      c->mesh_pending = std::move(c->H->levels[0]->owned);
      c->H.reset();
      c->H_prev.reset();
      c->have_coeff = false;
    }
  });
}

int tpamr_set_face_coefficients(tpamr_ctx* c, const double* kappa) {
  if (!c || !kappa) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (c->device >= 0) HIPCHECK(hipSetDevice(c->device));
    bool topology = false;
    if (!c->H) {
      if (!c->mesh_pending) throw StateError("no mesh: call tpamr_set_mesh first");
      c->H = build_hierarchy(std::move(c->mesh_pending), c->cfg, c->H_prev.get(), c->open);
      c->H_prev.reset();
      topology = true;
    }
    apply_boundary(c, kappa, topology);
    const auto t0 = std::chrono::steady_clock::now();
    bool on_device = c->device >= 0 && !topology && !c->host_refresh;
    // zero coefficients (components) take the host refresh
    if (on_device && !device_refresh(c, kappa)) on_device = false;
    if (!on_device) {
      refresh_coefficients(*c->H, kappa);
      c->host_stale = false;
    }
    c->have_coeff = true;
    const auto t1 = std::chrono::steady_clock::now();
    if (!on_device && c->device >= 0) {
      upload_hierarchy(c, topology);
      upload_gauge(c);
    }
    upload_boundary_gradient(c);
    if (c->trace) {
      const auto t2 = std::chrono::steady_clock::now();
      std::fprintf(stderr, "tpamr refresh (%s): host values %.1f ms, device part %.1f ms\n",
                   on_device ? "device" : "host",
                   std::chrono::duration<double, std::milli>(t1 - t0).count(),
                   std::chrono::duration<double, std::milli>(t2 - t1).count());
    }
  });
}

int tpamr_set_face_coefficients_device(tpamr_ctx* c, const double* d_kappa) {
  if (!c || !d_kappa) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H || c->topo.empty())
      throw StateError("no hierarchy: call tpamr_set_face_coefficients once before device refreshes");
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    if (c->any_open() && c->kb_pending) {
      std::vector<double> k;
      if (c->kb_extend) {
        k.resize(c->mesh().nfaces());
        download_vec(c, k.data(), d_kappa, k.size());
      }
      apply_boundary(c, k.data(), false);
    }
    if (!device_refresh(c, d_kappa, true))
      throw std::invalid_argument(
          "zero face coefficients split the domain into components: refresh with tpamr_set_face_coefficients");
    upload_boundary_gradient(c);
  });
}

int tpamr_set_pressure_boundary(tpamr_ctx* c, const int open[4], const double* boundary_coefficient) {
  if (!c || !open) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (!c->H && !c->mesh_pending) throw StateError("no mesh: call tpamr_set_mesh first");
    const TreeMesh& M = c->mesh();
    const int BF = int(M.bcell.size());
    bool sides[4];
    for (int q = 0; q < 4; ++q) sides[q] = open[q] != 0;
    if (boundary_coefficient) {
      for (int f = 0; f < BF; ++f)
        if (!(std::isfinite(boundary_coefficient[f]) && boundary_coefficient[f] >= 0))
          throw std::invalid_argument("boundary coefficients must be finite and nonnegative");
      c->kb_user.assign(boundary_coefficient, boundary_coefficient + BF);
      c->kb_extend = false;
    } else {
      c->kb_user.clear();
      c->kb_extend = true;
    }
    c->kb_pending = true;
    const bool changed = !std::equal(sides, sides + 4, c->open);
    std::copy(sides, sides + 4, c->open);
    if (changed && c->H) {
      // the open sides shape the irregular rows: rebuild the hierarchy at the
      // next coefficient update (geometric_preconditioner.jl:494-496)
      if (c->device >= 0) HIPCHECK(hipSetDevice(c->device));
      reset_topology_state(c);
      c->mesh_pending = std::move(c->H->levels[0]->owned);
      c->H.reset();
      c->H_prev.reset();
      c->have_coeff = false;
    }
  });
}

int tpamr_level_boundary(const tpamr_ctx* cc, int level, double* boundary_coefficient, int* flags) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    if (level < 0 || level >= int(c->H->levels.size())) throw std::invalid_argument("level out of range");
    const LevelData& L = *c->H->levels[level];
    if (boundary_coefficient) {
      if (L.kb.size() == L.mesh->bcell.size())
        std::copy(L.kb.begin(), L.kb.end(), boundary_coefficient);
      else
        std::fill(boundary_coefficient, boundary_coefficient + L.mesh->bcell.size(), 0.0);
    }
    if (flags) {
      flags[0] = L.singular ? 1 : 0;
      flags[1] = c->H->coarse_N > c->H->coarse_n ? 1 : 0;
      flags[2] = int(L.bopen.size());
      flags[3] = int(L.mesh->bcell.size());
    }
  });
}

int tpamr_component_means(const tpamr_ctx* c, int* count, int* roots, double* means) {
  if (!c || !count) return TPAMR_EINVAL;
  *count = int(c->last_means.size());
  if (roots) std::copy(c->last_roots.begin(), c->last_roots.end(), roots);
  if (means) std::copy(c->last_means.begin(), c->last_means.end(), means);
  return TPAMR_OK;
}

int tpamr_level_components(const tpamr_ctx* cc, int level, int* count, int* slot, int* roots) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !count) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    if (level < 0 || level >= int(c->H->levels.size())) throw std::invalid_argument("level out of range");
    const LevelData& L = *c->H->levels[level];
    *count = L.ncomp;
    if (slot) {
      if (L.comp_slot.empty())
        std::fill(slot, slot + L.ncells, L.ncomp);
      else
        std::copy(L.comp_slot.begin(), L.comp_slot.end(), slot);
    }
    if (roots)
      for (int q = 0; q < L.ncomp; ++q) roots[q] = L.comp_root[q] + 1;
  });
}

int tpamr_num_levels(const tpamr_ctx* c) { return (c && c->H) ? int(c->H->levels.size()) : 0; }

int tpamr_hierarchy_reuse(const tpamr_ctx* c, int* new_start, int* old_start) {
  if (!c || !new_start || !old_start) return TPAMR_EINVAL;
  if (!c->H) return TPAMR_ESTATE;
  *new_start = c->H->reused_new;
  *old_start = c->H->reused_old;
  return TPAMR_OK;
}

int tpamr_level_info(const tpamr_ctx* cc, int level, int* ints, double* dbl) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    if (level < 0 || level >= int(c->H->levels.size())) throw std::invalid_argument("level out of range");
    const LevelData& L = *c->H->levels[level];
    if (ints) {
      ints[0] = L.ncells;
      ints[1] = L.mesh->nfaces();
      ints[2] = L.nblocks;
      ints[3] = int(L.irr_local.size());
      long nnz = 0;
      for (int b = 0; b < L.nblocks; ++b)
        for (int q = 0; q < L.mesh->nb(); ++q) {
          const int cidx = b * L.mesh->nb() + q;
          if (L.cell_irr[cidx] >= 0) nnz += L.irr_ptr[L.cell_irr[cidx] + 1] - L.irr_ptr[L.cell_irr[cidx]];
        }
      ints[4] = int(nnz);
      ints[5] = L.mesh->max_leaf;
    }
    if (dbl) {
      dbl[0] = L.omega;
      dbl[1] = L.omega_p;
    }
  });
}

int tpamr_level_matrix(const tpamr_ctx* cc, int level, int* nnz, int* rowptr, int* col, double* val) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !nnz) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    if (level < 0 || level >= int(c->H->levels.size())) throw std::invalid_argument("level out of range");
    ensure_host_values(c);
    const LevelData& L = *c->H->levels[level];
    const TreeMesh& M = *L.mesh;
    const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
    std::vector<int> rp(1, 0), cl;
    std::vector<double> vl;
    for (int b = 0; b < L.nblocks; ++b)
      for (int q = 0; q < nb; ++q) {
        const int cidx = b * nb + q;
        std::vector<std::pair<int, double>> row;
        if (L.cell_irr[cidx] >= 0) {
          const int r = L.cell_irr[cidx];
          for (int e = L.irr_ptr[r]; e < L.irr_ptr[r + 1]; ++e) row.push_back({L.irr_col[e], L.irr_val[e]});
        } else {
          const int i = q % bx, j = q / bx;
          const double* cx = &L.cx[size_t(b) * (bx + 1) * by];
          const double* cy = &L.cy[size_t(b) * bx * (by + 1)];
          const double w[4] = {cx[j * (bx + 1) + i], cx[j * (bx + 1) + i + 1], cy[j * bx + i], cy[(j + 1) * bx + i]};
          auto at = [&](int bb, int ii, int jj) { return bb * nb + jj * bx + ii; };
          const int nbc[4] = {
              i > 0 ? at(b, i - 1, j) : (L.nbr[4 * b] >= 0 ? at(L.nbr[4 * b], bx - 1, j) : -1),
              i < bx - 1 ? at(b, i + 1, j) : (L.nbr[4 * b + 1] >= 0 ? at(L.nbr[4 * b + 1], 0, j) : -1),
              j > 0 ? at(b, i, j - 1) : (L.nbr[4 * b + 2] >= 0 ? at(L.nbr[4 * b + 2], i, by - 1) : -1),
              j < by - 1 ? at(b, i, j + 1) : (L.nbr[4 * b + 3] >= 0 ? at(L.nbr[4 * b + 3], i, 0) : -1)};
          double d = (w[0] + w[1]) + (w[2] + w[3]);
          row.push_back({cidx, d});
          // a slot without a neighbour is an open boundary face (diagonal only)
          for (int sd = 0; sd < 4; ++sd)
            if (w[sd] != 0.0 && nbc[sd] >= 0) row.push_back({nbc[sd], -w[sd]});
        }
        std::sort(row.begin(), row.end());
        for (size_t e = 0; e < row.size(); ++e) {
          if (e > 0 && row[e].first == row[e - 1].first) {
            vl.back() += row[e].second;
            continue;
          }
          cl.push_back(row[e].first);
          vl.push_back(row[e].second);
        }
        rp.push_back(int(cl.size()));
      }
    *nnz = int(cl.size());
    if (rowptr) {
      std::copy(rp.begin(), rp.end(), rowptr);
      std::copy(cl.begin(), cl.end(), col);
      std::copy(vl.begin(), vl.end(), val);
    }
  });
}

int tpamr_level_transfer(const tpamr_ctx* cc, int level, int* xfer) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !xfer) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    if (level < 0 || level + 1 >= int(c->H->levels.size())) throw std::invalid_argument("level has no transfer");
    const LevelData& L = *c->H->levels[level];
    std::copy(L.xfer.begin(), L.xfer.end(), xfer);
  });
}

int tpamr_level_kappa(const tpamr_ctx* cc, int level, double* kappa) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !kappa) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    if (level < 0 || level >= int(c->H->levels.size())) throw std::invalid_argument("level out of range");
    ensure_host_values(c);
    const LevelData& L = *c->H->levels[level];
    std::copy(L.kappa.begin(), L.kappa.end(), kappa);
  });
}

int tpamr_solve_device(tpamr_ctx* c, const double* d_rhs, double* d_x, double rtol, double atol, int max_cycles,
                       int compatibility, tpamr_stats* stats) {
  if (!c || !d_rhs || !d_x) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    const int N = c->topo[0].ncells;
    HIPCHECK(hipMemcpyAsync(c->b.p, d_rhs, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->x.p, d_x, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    tpamr_stats st = solve_dispatch(c, rtol, atol, max_cycles, compatibility);
    HIPCHECK(hipMemcpyAsync(d_x, c->x.p, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (stats) *stats = st;
  });
}

int tpamr_solve_device_noalloc(tpamr_ctx* c, const double* d_rhs, double* d_x, double rtol, double atol,
                               int max_cycles, int compatibility, tpamr_stats* stats) {
  if (!c || !d_rhs || !d_x) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    if (!c->noalloc) throw StateError("no device scratch address on this context's device");
    // layout of the scratch: Krylov state, scalars, stage, partials (the
    // sizes of the context's own buffers)
    const size_t N = size_t(c->topo[0].ncells), nblk = size_t(c->topo[0].nblocks), nc = (N + 255) / 256;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t off = 0;
    const size_t o_st = off;
    off += up(sizeof(CgState));
    const size_t o_scal = off;
    off += up(16 * sizeof(double));
    const size_t o_stage = off;
    off += up(4 * RED_WG * sizeof(double));
    const size_t o_part = off;
    off += up((std::max(3 * nblk, 4 * nc) + 16) * sizeof(double));
    const size_t o_max = off;
    off += up((std::max(nblk, nc) + 16) * sizeof(double));
    if (off > NOALLOC_SCRATCH_BYTES)
      throw SizeError("mesh too large for the allocation-free solve: it needs " + std::to_string(off) +
                      " bytes of device scratch, the limit is " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                      " (about 5e7 cells); use tpamr_solve_device");
    order_after_caller(c);
    char* base = static_cast<char*>(c->noalloc);
    // the solve runs on the scratch and launches its iterations one by one
    // (no graph instantiation); the context's own buffers come back after
    struct Swap {
      tpamr_ctx* c;
      tpamr_ctx::Reduce saved;
      bool graph;
      ~Swap() {
        c->rs = saved;
        c->use_graph = graph;
      }
    } swap{c, c->rs, c->use_graph};
    c->rs.st = reinterpret_cast<CgState*>(base + o_st);
    c->rs.scal = reinterpret_cast<double*>(base + o_scal);
    c->rs.stage = reinterpret_cast<double*>(base + o_stage);
    c->rs.part = reinterpret_cast<double*>(base + o_part);
    c->rs.maxpart = reinterpret_cast<double*>(base + o_max);
    c->use_graph = false;
    HIPCHECK(hipMemcpyAsync(c->b.p, d_rhs, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->x.p, d_x, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    tpamr_stats st = solve_dispatch(c, rtol, atol, max_cycles, compatibility);
    HIPCHECK(hipMemcpyAsync(d_x, c->x.p, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (stats) *stats = st;
  });
}

int tpamr_solve(tpamr_ctx* c, const double* rhs, double* x, double rtol, double atol, int max_cycles,
                int compatibility, tpamr_stats* stats) {
  if (!c || !rhs || !x) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    const int N = c->topo[0].ncells;
    check_finite(rhs, N, "right-hand side");
    check_finite(x, N, "initial solution");
    upload_vec(c, c->b.p, rhs, N);
    upload_vec(c, c->x.p, x, N);
    tpamr_stats st = solve_dispatch(c, rtol, atol, max_cycles, compatibility);
    download_vec(c, x, c->x.p, N);
    if (stats) *stats = st;
  });
}

int tpamr_krylov(tpamr_ctx* c, const double* rhs, const double* x0, double target, int max_iterations, double* x,
                 int* iterations, int* reason) {
  if (!c || !rhs || !x0 || !x || !iterations) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    if (max_iterations < 0) throw std::invalid_argument("max_iterations must be nonnegative");
    if (!(target >= 0)) throw std::invalid_argument("target must be a nonnegative number");
    HIPCHECK(hipSetDevice(c->device));
    const int N = c->topo[0].ncells;
    check_finite(rhs, N, "right-hand side");
    check_finite(x0, N, "initial solution");
    upload_vec(c, c->b.p, rhs, N);
    upload_vec(c, c->x.p, x0, N);
    const int it = c->fp32() ? run_cg<float>(c, c->lf, target, max_iterations)
                                                     : run_cg<double>(c, c->ld, target, max_iterations);
    download_vec(c, x, c->x.p, N);
    *iterations = it;
    if (reason) *reason = c->st_host->reason;
  });
}

int tpamr_vcycle(tpamr_ctx* c, const double* r, double* z) {
  if (!c || !r || !z) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    const int N = c->topo[0].ncells;
    upload_vec(c, c->r.p, r, N);
    if (c->fp32())
      vcycle_impl<float>(c, c->lf);
    else
      vcycle_impl<double>(c, c->ld);
    download_vec(c, z, c->tmp.p, N);
  });
}

int tpamr_apply_device(tpamr_ctx* c, const double* d_x, double* d_y) {
  if (!c || !d_x || !d_y) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    // the tile kernels move four cells per 16-byte (FP32) or 2x16-byte
    // (FP64) access: misaligned caller buffers go through library buffers
    const bool aligned = ((reinterpret_cast<uintptr_t>(d_x) | reinterpret_cast<uintptr_t>(d_y)) & 31u) == 0;
    if (aligned) {
      matvec(c, d_x, d_y);
    } else {
      const size_t N = size_t(c->topo[0].ncells);
      HIPCHECK(hipMemcpyAsync(c->tmp.p, d_x, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      matvec(c, c->tmp.p, c->tmp2.p);
      HIPCHECK(hipMemcpyAsync(d_y, c->tmp2.p, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int tpamr_apply(tpamr_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    const int N = c->topo[0].ncells;
    upload_vec(c, c->tmp.p, x, N);
    matvec(c, c->tmp.p, c->tmp2.p);
    download_vec(c, y, c->tmp2.p, N);
  });
}

int tpamr_mixture_face_coefficients(tpamr_ctx* c, const double* fraction, double rl, double rg, double* kappa) {
  if (!c || !fraction || !kappa) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients once before two-phase refreshes");
    if (!(std::isfinite(rl) && rl > 0 && std::isfinite(rg) && rg > 0))
      throw std::invalid_argument("phase densities must be finite and positive");
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    upload_vec(c, c->tmp.p, fraction, N);
    launch_mixture(c->stream, F, M.nb(), c->fm.p, c->fp.p, c->faxis.p, c->spacing.p, c->tmp.p, rl, rg, c->f1.p);
    download_vec(c, kappa, c->f1.p, F);
  });
}

int tpamr_mixture_boundary_coefficients(tpamr_ctx* c, const double* fraction, double rl, double rg, double* kb) {
  if (!c || !fraction || !kb) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (!(std::isfinite(rl) && rl > 0 && std::isfinite(rg) && rg > 0))
      throw std::invalid_argument("phase densities must be finite and positive");
    const TreeMesh& M = c->mesh();
    check_finite(fraction, M.ncells(), "fraction");
    for (size_t f = 0; f < M.bcell.size(); ++f) {
      const double cf = fraction[M.bcell[f]];
      kb[f] = 1.0 / (rl * cf + rg * (1.0 - cf));
    }
  });
}

int tpamr_surface_tension_jumps(tpamr_ctx* c, const double* fraction, const double* curvature, double sigma,
                                int formulation, double* jumps) {
  if (!c || !fraction || !curvature || !jumps) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    if (!(std::isfinite(sigma) && sigma >= 0)) throw std::invalid_argument("sigma must be finite and nonnegative");
    if (formulation != 0 && formulation != 1) throw std::invalid_argument("formulation must be CSF() or CGFM()");
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    upload_vec(c, c->tmp.p, fraction, N);
    upload_vec(c, c->f2.p, curvature, F);
    launch_surface_jump(c->stream, F, c->fm.p, c->fp.p, c->farea.p, c->tmp.p, c->f2.p, sigma, formulation, c->f1.p);
    download_vec(c, jumps, c->f1.p, F);
  });
}

int tpamr_surface_tension_jumps_cells(tpamr_ctx* c, const double* fraction, const double* cell_curvature,
                                      const double* curvature_quality, double sigma, int formulation, double* jumps,
                                      int* unresolved) {
  if (!c || !fraction || !cell_curvature || !curvature_quality || !jumps) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    if (!(std::isfinite(sigma) && sigma >= 0)) throw std::invalid_argument("sigma must be finite and nonnegative");
    if (formulation != 0 && formulation != 1) throw std::invalid_argument("formulation must be CSF() or CGFM()");
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    if (c->cell2.n < size_t(N)) c->cell2.alloc(2 * size_t(N));
    upload_vec(c, c->tmp.p, fraction, N);
    upload_vec(c, c->tmp2.p, cell_curvature, N);
    upload_vec(c, c->cell2.p, curvature_quality, N);
    HIPCHECK(hipMemsetAsync(c->counter.p, 0, sizeof(int), c->stream));
    launch_surface_jump_cells(c->stream, F, c->fm.p, c->fp.p, c->farea.p, c->origin.p, c->spacing.p, c->fcx.p,
                              c->fcy.p, int(M.blocks.size()), M.bs[0], M.bs[1], M.extent, M.periodic, c->tmp.p,
                              c->tmp2.p, c->cell2.p, sigma, formulation, c->f1.p, c->counter.p);
    int n = 0;
    HIPCHECK(hipMemcpyAsync(&n, c->counter.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    download_vec(c, jumps, c->f1.p, F);
    if (unresolved) *unresolved = n;
  });
}

int tpamr_weighted_jump_force(tpamr_ctx* c, const double* jumps, double* accel) {
  if (!c || !jumps || !accel) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    const int F = c->mesh().nfaces();
    upload_vec(c, c->f1.p, jumps, F);
    launch_face_jump(c->stream, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, nullptr, c->f1.p, c->f2.p);
    weighted_metric(c, c->f2.p, c->f3.p);
    download_vec(c, accel, c->f3.p, F);
  });
}

namespace {
// gravity! jumps into c->f3 (liquid and gas apertures staged in f1, f2)
void gravity_checks(tpamr_ctx* c, double rl, double rg, double gx, double gy) {
  if (!(std::isfinite(rl) && rl > 0 && std::isfinite(rg) && rg > 0))
    throw std::invalid_argument("densities must be finite and positive");
  if (!(std::isfinite(gx) && std::isfinite(gy))) throw std::invalid_argument("gravity must be finite");
  const TreeMesh& M = c->mesh();
  if ((M.periodic[0] && gx != 0) || (M.periodic[1] && gy != 0))
    throw std::invalid_argument("gravity potential requires nonperiodic boundaries along gravity");
}

// Jumps of gravity! from device apertures and offsets (nullable) into `out`.
void launch_gravity(tpamr_ctx* c, const double* d_off, const double* d_liquid, const double* d_gas, double rl,
                    double rg, double gx, double gy, double* out) {
  const TreeMesh& M = c->mesh();
  launch_gravity_jumps(c->stream, M.ncells(), M.bs[0], M.nb(), c->origin.p, c->spacing.p, d_off, gx, gy, c->tmp.p,
                       M.nfaces(), c->fm.p, c->fp.p, c->farea.p, d_liquid, d_gas, rl, rg, out);
}

void gravity_jumps(tpamr_ctx* c, const double* offsets, const double* liquid, const double* gas, double rl,
                   double rg, double gx, double gy) {
  gravity_checks(c, rl, rg, gx, gy);
  const TreeMesh& M = c->mesh();
  const int N = M.ncells(), F = M.nfaces();
  check_finite(liquid, F, "liquid aperture");
  check_finite(gas, F, "gas aperture");
  upload_vec(c, c->f1.p, liquid, F);
  upload_vec(c, c->f2.p, gas, F);
  const double* doff = nullptr;
  if (offsets) {
    check_finite(offsets, 2 * size_t(N), "interface centroid offsets");
    if (c->cell2.n < 2 * size_t(N)) c->cell2.alloc(2 * size_t(N));
    upload_vec(c, c->cell2.p, offsets, 2 * size_t(N));
    doff = c->cell2.p;
  }
  launch_gravity(c, doff, c->f1.p, c->f2.p, rl, rg, gx, gy, c->f3.p);
}
}  // namespace

int tpamr_gravity_jumps(tpamr_ctx* c, const double* offsets, const double* liquid, const double* gas,
                        double rho_liquid, double rho_gas, double gx, double gy, double* jumps) {
  if (!c || !liquid || !gas || !jumps) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    HIPCHECK(hipSetDevice(c->device));
    gravity_jumps(c, offsets, liquid, gas, rho_liquid, rho_gas, gx, gy);
    download_vec(c, jumps, c->f3.p, c->mesh().nfaces());
  });
}

int tpamr_gravity(tpamr_ctx* c, const double* offsets, const double* liquid, const double* gas, double rho_liquid,
                  double rho_gas, double gx, double gy, double* accel) {
  if (!c || !liquid || !gas || !accel) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    gravity_jumps(c, offsets, liquid, gas, rho_liquid, rho_gas, gx, gy);
    const int F = c->mesh().nfaces();
    launch_face_jump(c->stream, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, nullptr, c->f3.p, c->f1.p);
    weighted_metric(c, c->f1.p, c->f2.p);
    download_vec(c, accel, c->f2.p, F);
  });
}

int tpamr_divergence(tpamr_ctx* c, const double* u, const double* ub, double* div) {
  if (!c || !u || !div) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces(), BF = int(M.bcell.size());
    upload_vec(c, c->f1.p, u, F);
    double* dub = nullptr;
    if (ub && BF > 0) {
      if (c->f2.n < size_t(BF)) throw SizeError("boundary face buffer too small");
      upload_vec(c, c->f2.p, ub, BF);
      dub = c->f2.p;
    }
    launch_divergence(c->stream, N, M.nb(), c->cptr.p, c->cface.p, c->csign.p, c->farea.p, c->f1.p, c->bptr.p,
                      c->bface.p, c->bside.p, c->barea.p, dub, c->topo[0].volume.p, c->tmp.p);
    download_vec(c, div, c->tmp.p, N);
  });
}

int tpamr_weighted_gradient(tpamr_ctx* c, const double* p, double* g) {
  if (!c || !p || !g) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    upload_vec(c, c->tmp.p, p, N);
    launch_face_jump(c->stream, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, c->tmp.p, nullptr, c->f1.p);
    weighted_metric(c, c->f1.p, c->f3.p);
    download_vec(c, g, c->f3.p, F);
  });
}

int tpamr_weighted_gradient_boundary(tpamr_ctx* c, const double* p, double* gb) {
  if (!c || !p || !gb) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_solver(c);
    const LevelData& L0 = *c->H->levels[0];
    const TreeMesh& M = *L0.mesh;
    const int BF = int(M.bcell.size()), nb = M.nb();
    check_finite(p, M.ncells(), "pressure");
    for (int f = 0; f < BF; ++f) {
      const bool open = L0.open[2 * M.baxis[f] + (M.bside[f] > 0 ? 1 : 0)];
      const double factor = open ? 2 * L0.kb[f] / M.bspacing[2 * (M.bcell[f] / nb) + M.baxis[f]] : 0.0;
      gb[f] = (-double(M.bside[f]) * factor) * p[M.bcell[f]];
    }
  });
}

namespace {
// copy into / out of the library's buffers: host memory (upload, download) or
// the caller's device memory
void copy_in(tpamr_ctx* c, double* dst, const double* src, size_t n, bool host) {
  if (host)
    upload_vec(c, dst, src, n);
  else
    HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}
void copy_out(tpamr_ctx* c, double* dst, const double* src, size_t n, bool host) {
  if (host)
    download_vec(c, dst, src, n);
  else
    HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}

void project_checks(double dt, double tol, double reference_dt) {
  if (!(dt > 0)) throw std::invalid_argument("dt must be strictly positive");
  if (!(std::isfinite(tol) && tol > 0 && tol < 1)) throw std::invalid_argument("tolerance must be finite and lie in (0, 1)");
  if (!(std::isfinite(reference_dt) && reference_dt >= dt))
    throw std::invalid_argument("reference dt must be finite and at least dt");
}

// The projection on the velocity in c->f1: rhs from its divergence (and the
// inherited divergence in c->tmp2 when `inherited`), the increment in c->x,
// f1 -= dt G x.
tpamr_stats project_core(tpamr_ctx* c, double dt, double tol, int max_cycles, bool inherited, double reference_dt,
                         bool boundary = false) {
  const TreeMesh& M = c->mesh();
  const int N = M.ncells(), F = M.nfaces(), nb = M.nb();
  hipStream_t s = c->stream;
  // `boundary`: the boundary face velocities in c->fb enter the divergence
  // and receive the open faces' correction
  launch_divergence(s, N, nb, c->cptr.p, c->cface.p, c->csign.p, c->farea.p, c->f1.p, c->bptr.p, c->bface.p,
                    c->bside.p, c->barea.p, boundary ? c->fb.p : nullptr, c->topo[0].volume.p, c->tmp.p);
  launch_projection_rhs(s, N, c->tmp.p, inherited ? c->tmp2.p : nullptr, dt, (reference_dt - dt) / (dt * reference_dt),
                        c->b.p);
  launch_fill(s, N, 0.0, c->x.p);
  tpamr_stats st = solve_dispatch(c, 0.0, tol / (dt * dt), max_cycles, 0);
  st.residual_scale = 1.0 / (dt * dt);
  // _projection_compatibility! (projection.jl:100-127): the largest removed
  // strong component mean, times dt^2, against the tolerance
  double mean = std::fabs(st.removed_rhs_mean);
  for (double m : c->last_means) mean = std::max(mean, std::fabs(m));
  st.compatible = (mean == 0 ? 0.0 : dt * dt * mean) <= tol;
  // u -= dt G phi
  launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, c->x.p, nullptr, c->f2.p);
  weighted_metric(c, c->f2.p, c->f3.p);
  launch_axpby(s, F, -dt, c->f3.p, 1.0, c->f1.p);
  if (boundary)
    launch_boundary_velocity(s, int(M.bcell.size()), c->bcell.p, c->bside.p, c->bgrad.p, c->x.p, dt * 1.0, c->fb.p);
  return st;
}

// _accumulate_pressure_increment! (navier_stokes/projection.jl:52-97) after
// p += phi: with several components each ungauged one's volume mean of the
// accumulated pressure is removed (cell order), the gauged ones keep theirs.
void gauge_pressure_host(const tpamr_ctx* c, double* p) {
  const LevelData& L0 = *c->H->levels[0];
  if (!L0.singular || L0.connected) return;
  const int N = L0.ncells, nb = L0.mesh->nb();
  std::vector<double> sums(L0.ncomp, 0.0);
  for (int q = 0; q < N; ++q)
    if (L0.comp_slot[q]) sums[L0.comp_slot[q] - 1] += L0.volume[q / nb] * p[q];
  for (int k = 0; k < L0.ncomp; ++k) sums[k] /= L0.comp_vol[k];
  for (int q = 0; q < N; ++q)
    if (L0.comp_slot[q]) p[q] = p[q] - sums[L0.comp_slot[q] - 1];
}

// _validate_projection_boundary (navier_stokes/projection.jl:37-49): the
// interior-face projections serve operators without open sides,
// tpamr_project_open those with them.
void projection_sides(const tpamr_ctx* c, bool open_api) {
  if (c->any_open() && !open_api)
    throw std::invalid_argument(
        "pressure operator and velocity boundary have different open sides; project with tpamr_project_open");
  if (!c->any_open() && open_api)
    throw std::invalid_argument(
        "pressure operator and velocity boundary have different open sides; set them with "
        "tpamr_set_pressure_boundary");
}
}  // namespace

int tpamr_project_ex(tpamr_ctx* c, double* u, double* p, double dt, double tol, int max_cycles,
                     const double* inherited_divergence, double reference_dt, double* phi, tpamr_stats* stats) {
  if (!c || !u || !p) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    project_checks(dt, tol, reference_dt);
    projection_sides(c, false);
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    check_finite(u, F, "velocity");
    upload_vec(c, c->f1.p, u, F);
    if (inherited_divergence) {
      check_finite(inherited_divergence, N, "inherited divergence");
      upload_vec(c, c->tmp2.p, inherited_divergence, N);
    }
    tpamr_stats st = project_core(c, dt, tol, max_cycles, inherited_divergence != nullptr, reference_dt);
    download_vec(c, u, c->f1.p, F);
    std::vector<double> ph(N);
    download_vec(c, ph.data(), c->x.p, N);
    for (int q = 0; q < N; ++q) p[q] += ph[q];
    gauge_pressure_host(c, p);
    if (phi) std::copy(ph.begin(), ph.end(), phi);
    if (stats) *stats = st;
  });
}

int tpamr_project_open(tpamr_ctx* c, double* u, double* ub, double* p, double dt, double tol, int max_cycles,
                       double* phi, tpamr_stats* stats) {
  if (!c || !u || !ub || !p) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    project_checks(dt, tol, dt);
    projection_sides(c, true);
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces(), BF = int(M.bcell.size());
    check_finite(u, F, "velocity");
    check_finite(ub, BF, "boundary velocity");
    upload_vec(c, c->f1.p, u, F);
    if (BF > 0) upload_vec(c, c->fb.p, ub, BF);
    tpamr_stats st = project_core(c, dt, tol, max_cycles, false, dt, true);
    download_vec(c, u, c->f1.p, F);
    if (BF > 0) download_vec(c, ub, c->fb.p, BF);
    std::vector<double> ph(N);
    download_vec(c, ph.data(), c->x.p, N);
    for (int q = 0; q < N; ++q) p[q] += ph[q];
    gauge_pressure_host(c, p);
    if (phi) std::copy(ph.begin(), ph.end(), phi);
    if (stats) *stats = st;
  });
}

int tpamr_project_device(tpamr_ctx* c, double* d_u, double* d_p, double dt, double tol, int max_cycles,
                         const double* d_inherited_divergence, double reference_dt, double* d_phi,
                         tpamr_stats* stats) {
  if (!c || !d_u) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    project_checks(dt, tol, reference_dt);
    projection_sides(c, false);
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    hipStream_t s = c->stream;
    copy_in(c, c->f1.p, d_u, F, false);
    if (d_inherited_divergence) copy_in(c, c->tmp2.p, d_inherited_divergence, N, false);
    tpamr_stats st = project_core(c, dt, tol, max_cycles, d_inherited_divergence != nullptr, reference_dt);
    copy_out(c, d_u, c->f1.p, F, false);
    if (d_p) {
      launch_axpby(s, N, 1.0, c->x.p, 1.0, d_p);
      // _accumulate_pressure_increment!: each ungauged component's own gauge
      // when the domain has several components
      const LevelData& L0 = *c->H->levels[0];
      if (L0.singular && !L0.connected) gauge_remove<double>(c, 0, d_p);
    }
    if (d_phi) copy_out(c, d_phi, c->x.p, N, false);
    HIPCHECK(hipStreamSynchronize(s));
    if (stats) *stats = st;
  });
}

int tpamr_apply_force_device(tpamr_ctx* c, const double* d_jumps, double dt, double* d_u) {
  if (!c || !d_jumps || !d_u) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    if (!std::isfinite(dt)) throw std::invalid_argument("dt must be finite");
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    const int F = c->mesh().nfaces();
    launch_face_jump(c->stream, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, nullptr, d_jumps, c->f1.p);
    weighted_metric(c, c->f1.p, c->f2.p);
    launch_axpby(c->stream, F, dt, c->f2.p, 1.0, d_u);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int tpamr_mixture_face_coefficients_device(tpamr_ctx* c, const double* d_fraction, double rl, double rg,
                                           double* d_kappa) {
  if (!c || !d_fraction || !d_kappa) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients once before two-phase refreshes");
    if (!(std::isfinite(rl) && rl > 0 && std::isfinite(rg) && rg > 0))
      throw std::invalid_argument("phase densities must be finite and positive");
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    const TreeMesh& M = c->mesh();
    launch_mixture(c->stream, M.nfaces(), M.nb(), c->fm.p, c->fp.p, c->faxis.p, c->spacing.p, d_fraction, rl, rg,
                   d_kappa);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int tpamr_gravity_jumps_device(tpamr_ctx* c, const double* d_offsets, const double* d_liquid, const double* d_gas,
                               double rho_liquid, double rho_gas, double gx, double gy, double* d_jumps) {
  if (!c || !d_liquid || !d_gas || !d_jumps) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    gravity_checks(c, rho_liquid, rho_gas, gx, gy);
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    launch_gravity(c, d_offsets, d_liquid, d_gas, rho_liquid, rho_gas, gx, gy, d_jumps);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int tpamr_divergence_device(tpamr_ctx* c, const double* d_u, double* d_div) {
  if (!c || !d_u || !d_div) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    const TreeMesh& M = c->mesh();
    launch_divergence(c->stream, M.ncells(), M.nb(), c->cptr.p, c->cface.p, c->csign.p, c->farea.p, d_u, c->bptr.p,
                      c->bface.p, c->bside.p, c->barea.p, nullptr, c->topo[0].volume.p, d_div);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int tpamr_pressure_predictor(tpamr_ctx* c, double* u, const double* p, double dt) {
  if (!c || !u || !p) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    if (!(dt > 0 && std::isfinite(dt))) throw std::invalid_argument("dt must be finite and strictly positive");
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    hipStream_t s = c->stream;
    upload_vec(c, c->f1.p, u, F);
    upload_vec(c, c->tmp.p, p, N);
    launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, c->tmp.p, nullptr, c->f2.p);
    weighted_metric(c, c->f2.p, c->f3.p);
    launch_axpby(s, F, -dt, c->f3.p, 1.0, c->f1.p);
    download_vec(c, u, c->f1.p, F);
  });
}

int tpamr_project(tpamr_ctx* c, double* u, double* p, double dt, double tol, int max_cycles, double* phi,
                  tpamr_stats* stats) {
  return tpamr_project_ex(c, u, p, dt, tol, max_cycles, nullptr, dt, phi, stats);
}

int tpamr_adjoint(tpamr_ctx* c, const double* w, const double* p, double rtol, int max_cycles, double* dJ_drhs,
                  double* dJ_dkappa, tpamr_stats* stats) {
  if (!c || !w || !p || !dJ_drhs || !dJ_dkappa) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    if (c->H->general_gauge) throw std::invalid_argument("the adjoint needs one connected pressure component");
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces(), nb = M.nb();
    hipStream_t s = c->stream;
    // A lambda = Q w: strong rhs w / V
    upload_vec(c, c->tmp.p, w, N);
    launch_scale_by_volume(s, N, nb, c->tmp.p, c->b.p, c->topo[0].volume.p, true);
    launch_fill(s, N, 0.0, c->x.p);
    tpamr_stats st = solve_dispatch(c, rtol, 0.0, max_cycles, 0);
    // dJ/drhs = V lambda
    launch_scale_by_volume(s, N, nb, c->x.p, c->tmp.p, c->topo[0].volume.p, false);
    download_vec(c, dJ_drhs, c->tmp.p, N);
    // y_l, y_p, K y
    upload_vec(c, c->tmp2.p, p, N);
    launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, c->x.p, nullptr, c->f1.p);
    launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, c->tmp2.p, nullptr, c->f2.p);
    launch_metric_apply(s, F, c->kdiag.p, c->kptr.p, c->kface.p, c->kval.p, c->f1.p, c->f3.p);
    launch_metric_apply(s, F, c->kdiag.p, c->kptr.p, c->kface.p, c->kval.p, c->f2.p, c->f4.p);
    // f3 = K y_l, f4 = K y_p
    launch_adjoint_face(s, F, c->f1.p, c->f4.p, c->f2.p, c->f3.p, c->kappa.p, -1.0, c->f1.p);
    download_vec(c, dJ_dkappa, c->f1.p, F);
    if (stats) *stats = st;
  });
}

namespace {
// Pullback of one projection step; inputs and outputs in host memory (host)
// or device memory
tpamr_stats project_adjoint_core(tpamr_ctx* c, const double* ubar, const double* phi, const double* jumps, double dt,
                                 double rtol, int max_cycles, double* ubar_star, double* dJ_drhs, double* dJ_dkappa,
                                 bool host) {
  if (!(dt > 0 && std::isfinite(dt))) throw std::invalid_argument("dt must be finite and strictly positive");
  projection_sides(c, false);
  if (c->H->general_gauge)
    throw std::invalid_argument("the projection adjoint needs one connected pressure component");
  const TreeMesh& M = c->mesh();
  const int N = M.ncells(), F = M.nfaces(), nb = M.nb();
  hipStream_t s = c->stream;
  if (host) check_finite(ubar, F, "velocity adjoint");
  copy_in(c, c->f1.p, ubar, F, host);
  // A lambda = J^T M ubar = -V div(M ubar), M = sqrt(k) K sqrt(k)
  launch_face_scale(s, F, c->sqk.p, c->f1.p, c->f2.p);
  weighted_metric(c, c->f2.p, c->f3.p);
  launch_divergence(s, N, nb, c->cptr.p, c->cface.p, c->csign.p, c->farea.p, c->f3.p, c->bptr.p, c->bface.p,
                    c->bside.p, c->barea.p, nullptr, c->topo[0].volume.p, c->tmp.p);
  launch_axpby(s, N, -1.0, c->tmp.p, 0.0, c->b.p);
  launch_fill(s, N, 0.0, c->x.p);
  tpamr_stats st = solve_dispatch(c, rtol, 0.0, max_cycles, 0);
  if (dJ_drhs) {
    // dJ/drhs = -dt V lambda
    launch_scale_by_volume(s, N, nb, c->x.p, c->tmp.p, c->topo[0].volume.p, false);
    launch_axpby(s, N, -dt, c->tmp.p, 0.0, c->tmp.p);
    copy_out(c, dJ_drhs, c->tmp.p, N, host);
  }
  // ubar* = ubar - J lambda, J x = |f| (x+ - x-)
  launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, nullptr, c->x.p, nullptr, c->f2.p);
  launch_axpby(s, F, -1.0, c->f2.p, 1.0, c->f1.p);
  copy_out(c, ubar_star, c->f1.p, F, host);
  if (dJ_dkappa) {
    // z = jumps - J phi; dJ/dk_f = dt [W_f (K Z)_f + Z_f (K W)_f] / (2 k_f)
    // with W = sqrt(k) ubar*, Z = sqrt(k) z
    if (host) check_finite(phi, N, "checkpointed pressure");
    copy_in(c, c->tmp2.p, phi, N, host);
    launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, nullptr, c->tmp2.p, nullptr, c->f3.p);
    if (jumps) {
      if (host) check_finite(jumps, F, "force jumps");
      copy_in(c, c->f2.p, jumps, F, host);
    } else {
      launch_fill(s, F, 0.0, c->f2.p);
    }
    launch_axpby(s, F, -1.0, c->f3.p, 1.0, c->f2.p);
    launch_face_scale(s, F, c->sqk.p, c->f2.p, c->f3.p);  // Z
    launch_face_scale(s, F, c->sqk.p, c->f1.p, c->f2.p);  // W
    launch_metric_apply(s, F, c->kdiag.p, c->kptr.p, c->kface.p, c->kval.p, c->f2.p, c->f1.p);
    launch_metric_apply(s, F, c->kdiag.p, c->kptr.p, c->kface.p, c->kval.p, c->f3.p, c->f4.p);
    launch_adjoint_face(s, F, c->f2.p, c->f4.p, c->f3.p, c->f1.p, c->kappa.p, dt, c->f1.p);
    copy_out(c, dJ_dkappa, c->f1.p, F, host);
  }
  HIPCHECK(hipStreamSynchronize(s));
  return st;
}
}  // namespace

int tpamr_project_adjoint(tpamr_ctx* c, const double* ubar, const double* phi, const double* jumps, double dt,
                          double rtol, int max_cycles, double* ubar_star, double* dJ_drhs, double* dJ_dkappa,
                          tpamr_stats* stats) {
  if (!c || !ubar || !ubar_star || (dJ_dkappa && !phi)) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    tpamr_stats st = project_adjoint_core(c, ubar, phi, jumps, dt, rtol, max_cycles, ubar_star, dJ_drhs, dJ_dkappa,
                                          true);
    if (stats) *stats = st;
  });
}

int tpamr_project_adjoint_device(tpamr_ctx* c, const double* d_ubar, const double* d_phi, const double* d_jumps,
                                 double dt, double rtol, int max_cycles, double* d_ubar_star, double* d_dJ_drhs,
                                 double* d_dJ_dkappa, tpamr_stats* stats) {
  if (!c || !d_ubar || !d_ubar_star || (d_dJ_dkappa && !d_phi)) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    tpamr_stats st = project_adjoint_core(c, d_ubar, d_phi, d_jumps, dt, rtol, max_cycles, d_ubar_star, d_dJ_drhs,
                                          d_dJ_dkappa, false);
    if (stats) *stats = st;
  });
}

int tpamr_halo_sizes(tpamr_ctx* c, int halo, int* entries, int* physical) {
  if (!c || !entries || !physical) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    if (c->halo_host.halo != halo) c->halo_host = halo_schedule(c->mesh(), halo);
    *entries = int(c->halo_host.dest.size());
    *physical = c->halo_host.nphysical;
  });
}

int tpamr_halo_schedule(tpamr_ctx* c, int halo, int* dest, int* ptr, int* donor, double* weight, int* phys) {
  if (!c || !dest || !ptr || !donor || !weight || !phys) return TPAMR_EINVAL;
  return guard(c, [&] {
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    if (c->halo_host.halo != halo) c->halo_host = halo_schedule(c->mesh(), halo);
    const HaloSchedule& S = c->halo_host;
    std::copy(S.dest.begin(), S.dest.end(), dest);
    std::copy(S.ptr.begin(), S.ptr.end(), ptr);
    std::copy(S.donor.begin(), S.donor.end(), donor);
    std::copy(S.weight.begin(), S.weight.end(), weight);
    std::copy(S.phys.begin(), S.phys.end(), phys);
  });
}

int tpamr_fill_halo(tpamr_ctx* c, int halo, const double* cells, const double* physical, double* haloed) {
  if (!c || !cells || !haloed) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (!c->H) throw StateError("no hierarchy: call tpamr_set_face_coefficients first");
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const TreeMesh& M = c->mesh();
    if (c->halo_dev != halo) {
      if (c->halo_host.halo != halo) c->halo_host = halo_schedule(M, halo);
      const HaloSchedule& S = c->halo_host;
      c->h_dest.upload(S.dest.empty() ? std::vector<int>(1, 0) : S.dest, s);
      c->h_ptr.upload(S.ptr, s);
      c->h_donor.upload(S.donor.empty() ? std::vector<int>(1, 0) : S.donor, s);
      c->h_weight.upload(S.weight.empty() ? std::vector<double>(1, 0.0) : S.weight, s);
      c->h_phys.upload(S.phys.empty() ? std::vector<int>(1, -1) : S.phys, s);
      c->halo_dev = halo;
    }
    const HaloSchedule& S = c->halo_host;
    const int N = M.ncells(), bx = M.bs[0], by = M.bs[1];
    const size_t total = size_t(M.blocks.size()) * (bx + 2 * halo) * (by + 2 * halo);
    c->h_out.alloc(total);
    upload_vec(c, c->tmp.p, cells, N);
    const double* dphys = nullptr;
    if (physical && S.nphysical > 0) {
      c->h_phys_values.alloc(S.nphysical);
      upload_vec(c, c->h_phys_values.p, physical, S.nphysical);
      dphys = c->h_phys_values.p;
    }
    launch_fill_halo(s, N, bx, by, halo, int(S.dest.size()), c->h_dest.p, c->h_ptr.p, c->h_donor.p, c->h_weight.p,
                     c->h_phys.p, c->tmp.p, dphys, c->h_out.p);
    download_vec(c, haloed, c->h_out.p, total);
  });
}

int tpamr_time_stream(tpamr_ctx* c, long long n, int reads, int reps, double* avg_ms, double* bytes) {
  if (!c || !avg_ms || reps <= 0) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    if (reads != 1 && reads != 5) throw std::invalid_argument("reads must be 1 or 5");
    if (n <= 0 || n % 4 != 0) throw std::invalid_argument("n must be a positive multiple of 4");
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    DevArray<float> in, out;
    in.alloc(size_t(n) * reads);
    out.alloc(size_t(n));
    HIPCHECK(hipMemsetAsync(in.p, 0, sizeof(float) * size_t(n) * reads, s));
    launch_stream(s, reads, size_t(n) / 4, in.p, out.p);  // first touch
    HIPCHECK(hipEventRecord(c->ev0, s));
    for (int q = 0; q < reps; ++q) launch_stream(s, reads, size_t(n) / 4, in.p, out.p);
    HIPCHECK(hipEventRecord(c->ev1, s));
    HIPCHECK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = double(ms) / reps;
    if (bytes) *bytes = double(n) * 4.0 * (reads + 1);
  });
}

int tpamr_time_smoother(tpamr_ctx* c, int reps, double* avg_ms, double* bytes) {
  if (!c || !avg_ms || reps <= 0) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_solver(c);
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int N = c->topo[0].ncells;
    const LevelData& L = *c->H->levels[0];
    const size_t nb = L.nblocks;
    const size_t bx = L.mesh->bs[0], by = L.mesh->bs[1];
    double elem = 8;
    HIPCHECK(hipEventRecord(c->ev0, s));
    if (c->fp32()) {
      DevLevel<float>& D = c->lf[0];
      for (int q = 0; q < reps; ++q)
        launch_sweep<float>(s, D.view, (q & 1) ? D.xs.p : D.x.p, (q & 1) ? D.x.p : D.xs.p, D.b.p, D.omega, nullptr,
                            true, ZStats{});
      elem = 4;
    } else {
      DevLevel<double>& D = c->ld[0];
      for (int q = 0; q < reps; ++q)
        launch_sweep<double>(s, D.view, (q & 1) ? D.xs.p : D.x.p, (q & 1) ? D.x.p : D.xs.p, D.b.p, D.omega, nullptr,
                             true, ZStats{});
    }
    HIPCHECK(hipEventRecord(c->ev1, s));
    HIPCHECK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = double(ms) / reps;
    if (bytes) {
      // compulsory traffic: x, b read, x' write, the two conductance
      // arrays (including the block-boundary faces), and the irregular rows
      // (entries, diagonal, cell/length word) with their external values
      const double cells = double(N) * 3 * elem;
      const double cond = double(nb) * ((bx + 1) * by + bx * (by + 1)) * elem;
      const double nnz = L.irr_ptr.empty() ? 0.0 : double(L.irr_ptr.back());
      const double irr = nnz * (elem + 2) + double(L.irr_local.size()) * (elem + 4) +
                         double(L.ext_col.size()) * (4 + elem);
      *bytes = cells + cond + irr;
    }
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// One-velocity momentum step (uniform, fully periodic meshes)

namespace {

Momentum& need_momentum(tpamr_ctx* c) {
  if (!c->mom) throw StateError("no momentum operators: call tpamr_set_momentum first");
  return *c->mom;
}

tpamr_stats viscous_stats(const ViscousStats& v) {
  tpamr_stats st{};
  st.cycles = v.cycles;
  st.converged = v.converged;
  st.compatible = 1;
  st.initial_l2 = v.initial_l2;
  st.initial_linf = v.initial_linf;
  st.final_l2 = v.final_l2;
  st.final_linf = v.final_linf;
  st.target_linf = v.target_linf;
  st.residual_scale = v.residual_scale;
  st.krylov_runs = 1;
  return st;
}

void momentum_buffers(tpamr_ctx* c) {
  const TreeMesh& M = c->mesh();
  c->mold.alloc(M.nfaces());
  c->mconv.alloc(M.nfaces());
  c->f1.alloc(M.nfaces());
  c->mp.alloc(M.ncells());
}

// the implicit viscous substep; a component that does not converge raises
// after dst and the statistics are written (implicit_diffusion.jl:953-963)
void implicit_checked(tpamr_ctx* c, double* dst, const double* conv, const double* old, double dt, double tol,
                      int max_cycles) {
  ViscousStats vs[2];
  c->mom->implicit(dst, conv, old, dt, tol, max_cycles, vs);
  c->vstats[0] = viscous_stats(vs[0]);
  c->vstats[1] = viscous_stats(vs[1]);
  for (int a = 0; a < 2; ++a)
    if (!vs[a].converged) {
      char msg[256];
      std::snprintf(msg, sizeof(msg),
                    "staggered diffusion solve did not converge (component %d: %d cycles, residual %.3e, target "
                    "%.3e; dt=%g, max_cycles=%d)",
                    a + 1, vs[a].cycles, vs[a].final_linf, vs[a].target_linf, dt, max_cycles);
      throw std::runtime_error(msg);
    }
}

// step! on the velocity in c->f1 and the device pressure dp
tpamr_stats step_core(tpamr_ctx* c, double* dp, double dt, const double* dforce, double tol, int max_cycles) {
  Momentum& Mo = need_momentum(c);
  const TreeMesh& M = c->mesh();
  const int N = M.ncells(), F = M.nfaces();
  hipStream_t s = c->stream;
  launch_copy(s, F, c->f1.p, c->mold.p);
  // _check_transport_timestep (single_phase/stepping.jl:95-103)
  const double bound = Mo.stable_timestep(c->mold.p, 1.0);
  if (!(dt <= bound * (1 + 1.0e-6 + 32 * EPS))) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "time step %.17g exceeds advective stability bound %.17g", dt, bound);
    throw std::invalid_argument(msg);
  }
  // transport with the old velocity frozen, body force, viscous substep
  launch_copy(s, F, c->mold.p, c->mconv.p);
  Mo.advect(c->mconv.p, c->mold.p, dt);
  if (dforce) launch_axpby(s, F, dt, dforce, 1.0, c->mconv.p);
  implicit_checked(c, c->f1.p, c->mconv.p, c->mold.p, dt, tol, max_cycles);
  // _pressure_predictor! with the old pressure, then the projection
  launch_face_jump(s, F, c->fm.p, c->fp.p, c->farea.p, c->sqk.p, dp, nullptr, c->f2.p);
  weighted_metric(c, c->f2.p, c->f3.p);
  launch_axpby(s, F, -dt, c->f3.p, 1.0, c->f1.p);
  tpamr_stats st = project_core(c, dt, tol, max_cycles, false, dt);
  launch_axpby(s, N, 1.0, c->x.p, 1.0, dp);
  const LevelData& L0 = *c->H->levels[0];
  if (L0.singular && !L0.connected) gauge_remove<double>(c, 0, dp);
  return st;
}

void step_checks(tpamr_ctx* c, double dt, double tol, int max_cycles) {
  need_device(c);
  need_solver(c);
  need_momentum(c);
  if (!(std::isfinite(dt) && dt > 0)) throw std::invalid_argument("dt must be finite and strictly positive");
  if (max_cycles < 0) throw std::invalid_argument("max_cycles must be nonnegative");
  project_checks(dt, tol, dt);
  projection_sides(c, false);
}

}  // namespace

extern "C" {

int tpamr_set_momentum(tpamr_ctx* c, double viscosity, int scheme) {
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    HIPCHECK(hipSetDevice(c->device));
    c->mom.reset();
    c->mom.reset(new Momentum(c->mesh(), viscosity, scheme, c->cfg, c->stream));
    momentum_buffers(c);
  });
}

int tpamr_momentum_sweep_forward(const tpamr_ctx* cc, int* forward) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !forward) return TPAMR_EINVAL;
  return guard(c, [&] { *forward = need_momentum(c).sweep_forward() ? 1 : 0; });
}

int tpamr_momentum_levels(const tpamr_ctx* cc, int* count, int* cells) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !count) return TPAMR_EINVAL;
  return guard(c, [&] {
    const std::vector<int> v = need_momentum(c).level_cells();
    *count = int(v.size());
    if (cells)
      for (size_t q = 0; q < v.size() && q < 64; ++q) cells[q] = v[q];
  });
}

int tpamr_advect(tpamr_ctx* c, double* field, const double* advecting, double dt) {
  if (!c || !field || !advecting) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    Momentum& Mo = need_momentum(c);
    if (!(dt >= 0)) throw std::invalid_argument("dt must be nonnegative");
    HIPCHECK(hipSetDevice(c->device));
    const int F = c->mesh().nfaces();
    upload_vec(c, c->mconv.p, field, F);
    const bool self = advecting == field;
    if (!self) upload_vec(c, c->mold.p, advecting, F);
    Mo.advect(c->mconv.p, self ? c->mconv.p : c->mold.p, dt);
    download_vec(c, field, c->mconv.p, F);
  });
}

int tpamr_diffusion(tpamr_ctx* c, const double* u, double* out) {
  if (!c || !u || !out) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    Momentum& Mo = need_momentum(c);
    HIPCHECK(hipSetDevice(c->device));
    const int F = c->mesh().nfaces();
    upload_vec(c, c->mold.p, u, F);
    Mo.diffusion(c->mold.p, c->mconv.p);
    download_vec(c, out, c->mconv.p, F);
  });
}

int tpamr_implicit_diffusion(tpamr_ctx* c, double* dst, const double* convected, const double* old, double dt,
                             double tolerance, int max_cycles, tpamr_stats stats[2]) {
  if (!c || !dst || !convected || !old) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    need_momentum(c);
    if (!(std::isfinite(dt) && dt > 0)) throw std::invalid_argument("dt must be finite and strictly positive");
    if (!(std::isfinite(tolerance) && tolerance > 0 && tolerance < 1))
      throw std::invalid_argument("tolerance must be finite and lie in (0, 1)");
    if (max_cycles < 0) throw std::invalid_argument("max_cycles must be nonnegative");
    HIPCHECK(hipSetDevice(c->device));
    const int F = c->mesh().nfaces();
    upload_vec(c, c->f1.p, dst, F);
    upload_vec(c, c->mconv.p, convected, F);
    upload_vec(c, c->mold.p, old, F);
    try {
      implicit_checked(c, c->f1.p, c->mconv.p, c->mold.p, dt, tolerance, max_cycles);
    } catch (const std::runtime_error& e) {
      download_vec(c, dst, c->f1.p, F);
      if (stats) {
        stats[0] = c->vstats[0];
        stats[1] = c->vstats[1];
      }
      throw std::invalid_argument(e.what());
    }
    download_vec(c, dst, c->f1.p, F);
    if (stats) {
      stats[0] = c->vstats[0];
      stats[1] = c->vstats[1];
    }
  });
}

int tpamr_stable_timestep(tpamr_ctx* c, const double* u, double cfl, double* dt_out) {
  if (!c || !u || !dt_out) return TPAMR_EINVAL;
  return guard(c, [&] {
    need_device(c);
    Momentum& Mo = need_momentum(c);
    HIPCHECK(hipSetDevice(c->device));
    upload_vec(c, c->mold.p, u, c->mesh().nfaces());
    *dt_out = Mo.stable_timestep(c->mold.p, cfl);
  });
}

int tpamr_step(tpamr_ctx* c, double* u, double* p, double dt, const double* force, double tolerance, int max_cycles,
               tpamr_stats* stats) {
  if (!c || !u || !p) return TPAMR_EINVAL;
  return guard(c, [&] {
    step_checks(c, dt, tolerance, max_cycles);
    HIPCHECK(hipSetDevice(c->device));
    const TreeMesh& M = c->mesh();
    const int N = M.ncells(), F = M.nfaces();
    check_finite(u, F, "velocity");
    check_finite(p, N, "pressure");
    upload_vec(c, c->f1.p, u, F);
    upload_vec(c, c->mp.p, p, N);
    // the force values have their own staging buffer (f2..f4 are the
    // projection's scratch)
    if (force) {
      check_finite(force, F, "force");
      c->mforce.alloc(F);
      upload_vec(c, c->mforce.p, force, F);
    }
    tpamr_stats st;
    try {
      st = step_core(c, c->mp.p, dt, force ? c->mforce.p : nullptr, tolerance, max_cycles);
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    download_vec(c, u, c->f1.p, F);
    download_vec(c, p, c->mp.p, N);
    if (stats) *stats = st;
  });
}

int tpamr_step_device(tpamr_ctx* c, double* d_u, double* d_p, double dt, const double* d_force, double tolerance,
                      int max_cycles, tpamr_stats* stats) {
  if (!c || !d_u || !d_p) return TPAMR_EINVAL;
  return guard(c, [&] {
    step_checks(c, dt, tolerance, max_cycles);
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    const int F = c->mesh().nfaces();
    copy_in(c, c->f1.p, d_u, F, false);
    tpamr_stats st;
    try {
      st = step_core(c, d_p, dt, d_force, tolerance, max_cycles);
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    copy_out(c, d_u, c->f1.p, F, false);
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (stats) *stats = st;
  });
}

int tpamr_step_viscous_stats(const tpamr_ctx* cc, tpamr_stats stats[2]) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !stats) return TPAMR_EINVAL;
  return guard(c, [&] {
    stats[0] = c->vstats[0];
    stats[1] = c->vstats[1];
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CMDGFM reduced-pressure/jump block solve on the context's pressure
// hierarchy (cmdgfm.hpp)

namespace {

CmdgfmSystem& need_cmdgfm(tpamr_ctx* c) {
  if (!c->cmd) throw StateError("no block system: call tpamr_cmdgfm_set_system first");
  return *c->cmd;
}

// One V-cycle of the shared hierarchy for the reduced-pressure block: the
// rhs goes straight into the finest level, whose zero-start iterate the
// first pre-sweep forms, and the answer stays where the V-cycle leaves it
// (its constant per component is irrelevant to the differences the block
// takes). Hierarchies with several components, or fewer than two
// pre-sweeps, take the preconditioner's full path.
template <typename T>
const double* cmdgfm_vcycle_t(tpamr_ctx* c, std::vector<DevLevel<T>>& lv, double* r, double* zout) {
  if (!x0_fused(c)) {
    vcycle_into<T>(c, lv, r, zout);
    return zout;
  }
  hipStream_t s = c->stream;
  const int N = c->topo[0].ncells;
  if constexpr (std::is_same<T, float>::value)
    launch_cast_d2f(s, N, r, lv[0].b.p);
  else if (r != lv[0].b.p)
    launch_copy(s, N, r, lv[0].b.p);
  T* z = enqueue_vcycle<T>(c, lv, &c->rs.st->done, ZStats{}, true);
  if constexpr (std::is_same<T, float>::value) {
    launch_cast_f2d(s, N, z, zout);
    return zout;
  } else {
    return z;
  }
}

CmdgfmSystem::Vcycle cmdgfm_vcycle(tpamr_ctx* c) {
  const int N = c->topo[0].ncells;
  c->cm_z.alloc(N);
  // the Krylov state's stop flag, which the V-cycle kernels read, is clear
  CgState init{};
  init.scale = 1.0;
  *c->st_host = init;
  HIPCHECK(hipMemcpyAsync(c->rs.st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
  return [c](double* r) -> const double* {
    if (c->fp32()) return cmdgfm_vcycle_t<float>(c, c->lf, r, c->cm_z.p);
    return cmdgfm_vcycle_t<double>(c, c->ld, r, c->cm_z.p);
  };
}

void cmdgfm_checks(tpamr_ctx* c, CmdgfmSystem& sys) {
  need_device(c);
  if (sys.pressure_size() > 0) need_solver(c);
}

void cmdgfm_run(tpamr_ctx* c, CmdgfmSystem& sys, const double* d_rhs, double* d_x, double dt, double density,
                double tolerance, int max_iterations, int warm_start, const double* d_prev, int preconditioned,
                tpamr_cmdgfm_stats* stats) {
  const TreeMesh& m = c->mesh();
  const double h = std::max(m.extent[0], m.extent[1]);
  CmdgfmStats st;
  // in FP64 with the fused zero-start sweep the block writes its rhs
  // straight into the finest level
  CmdgfmSystem::VcycleInput input;
  if (!c->fp32() && x0_fused(c)) input = [c]() { return c->ld[0].b.p; };
  sys.solve(d_rhs, d_x, dt, density, h, tolerance, max_iterations, warm_start != 0, d_prev, preconditioned != 0,
            cmdgfm_vcycle(c), input, st);
  if (stats) {
    stats->converged = st.converged;
    stats->iterations = st.iterations;
    stats->krylov_runs = st.krylov_runs;
    stats->initial_guess_used = st.initial_guess_used;
    stats->pressure_unknowns = st.pressure_unknowns;
    stats->jump_unknowns = st.jump_unknowns;
    stats->initial_residual = st.initial_residual;
    stats->final_residual = st.final_residual;
    stats->pressure_defect = st.pressure_defect;
    stats->jump_defect = st.jump_defect;
    stats->anchor_defect = st.anchor_defect;
    stats->asymmetry = st.asymmetry;
    stats->jump_diagonal_margin = st.jump_diagonal_margin;
  }
}

}  // namespace

extern "C" {

int tpamr_cmdgfm_set_system(tpamr_ctx* c, int np, int nx, const int* rowptr, const int* cols, const double* vals,
                            const int* pressure_cells, const int* anchor_cells, const double* pressure_volumes,
                            const double* jump_lengths, int nomitted, const int* omitted_rowptr,
                            const int* omitted_cols, const double* omitted_vals, const double* omitted_rhs,
                            const double* omitted_volumes, double drop_tolerance) {
  if (!c) return TPAMR_EINVAL;
  return guard(c, [&] {
    const int ncells = c->mesh().ncells();
    if (c->device >= 0) HIPCHECK(hipSetDevice(c->device));
    c->cmd.reset();
    c->cmd.reset(new CmdgfmSystem(ncells, np, nx, rowptr, cols, vals, pressure_cells, anchor_cells,
                                  pressure_volumes, jump_lengths, nomitted, omitted_rowptr, omitted_cols,
                                  omitted_vals, omitted_rhs, omitted_volumes, drop_tolerance, c->device >= 0,
                                  c->stream));
  });
}

int tpamr_cmdgfm_jump_factor(const tpamr_ctx* cc, int* sizes, int* sptr, int* scol, double* sval, int* lptr,
                             int* lcol, double* lval, int* uptr, int* ucol, double* uval, double* udiag) {
  auto* c = const_cast<tpamr_ctx*>(cc);
  if (!c || !sizes) return TPAMR_EINVAL;
  return guard(c, [&] {
    const JumpFactor& f = need_cmdgfm(c).factor();
    const int nx = need_cmdgfm(c).jump_size();
    sizes[0] = nx;
    sizes[1] = nx ? int(f.scol.size()) : 0;
    sizes[2] = nx ? int(f.lcol.size()) : 0;
    sizes[3] = nx ? int(f.ucol.size()) : 0;
    if (!nx) return;
    auto put_i = [](int* dst, const std::vector<int>& v) {
      if (dst) std::copy(v.begin(), v.end(), dst);
    };
    auto put_d = [](double* dst, const std::vector<double>& v) {
      if (dst) std::copy(v.begin(), v.end(), dst);
    };
    put_i(sptr, f.sptr);
    put_i(scol, f.scol);
    put_d(sval, f.sval);
    put_i(lptr, f.lptr);
    put_i(lcol, f.lcol);
    put_d(lval, f.lval);
    put_i(uptr, f.uptr);
    put_i(ucol, f.ucol);
    put_d(uval, f.uval);
    put_d(udiag, f.udiag);
  });
}

int tpamr_cmdgfm_solve(tpamr_ctx* c, const double* rhs, double* solution, double dt, double density,
                       double tolerance, int max_iterations, int warm_start, const double* previous_increment,
                       int preconditioned, tpamr_cmdgfm_stats* stats) {
  if (!c || !solution) return TPAMR_EINVAL;
  return guard(c, [&] {
    CmdgfmSystem& sys = need_cmdgfm(c);
    cmdgfm_checks(c, sys);
    const int n = sys.size();
    if (n > 0 && !rhs) throw std::invalid_argument("null rhs");
    HIPCHECK(hipSetDevice(c->device));
    const int N = c->mesh().ncells();
    c->cm_rhs.alloc(std::max(n, 1));
    c->cm_x.alloc(std::max(n, 1));
    if (n > 0) upload_vec(c, c->cm_rhs.p, rhs, n);
    const double* prev = nullptr;
    if (previous_increment) {
      c->cm_prev.alloc(N);
      upload_vec(c, c->cm_prev.p, previous_increment, N);
      prev = c->cm_prev.p;
    }
    cmdgfm_run(c, sys, c->cm_rhs.p, c->cm_x.p, dt, density, tolerance, max_iterations, warm_start, prev,
               preconditioned, stats);
    if (n > 0) download_vec(c, solution, c->cm_x.p, n);
  });
}

int tpamr_cmdgfm_solve_device(tpamr_ctx* c, const double* d_rhs, double* d_solution, double dt, double density,
                              double tolerance, int max_iterations, int warm_start, const double* d_previous,
                              int preconditioned, tpamr_cmdgfm_stats* stats) {
  if (!c || !d_solution) return TPAMR_EINVAL;
  return guard(c, [&] {
    CmdgfmSystem& sys = need_cmdgfm(c);
    cmdgfm_checks(c, sys);
    if (sys.size() > 0 && !d_rhs) throw std::invalid_argument("null rhs");
    HIPCHECK(hipSetDevice(c->device));
    order_after_caller(c);
    cmdgfm_run(c, sys, d_rhs, d_solution, dt, density, tolerance, max_iterations, warm_start, d_previous,
               preconditioned, stats);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
