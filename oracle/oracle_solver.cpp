// Oracle: elliptic pressure operator, tree geometric multigrid, Krylov CG with
// the reference's stopping rules, two-phase coefficients, projection and the
// pressure adjoint. TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <parallel/algorithm>
#include <omp.h>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

#include "oracle.hpp"

namespace orc {

static const double EPS = std::numeric_limits<double>::epsilon();

// `_chunked_slot_sums!` (src/threading.jl:202-283): items summed in index
// order within fixed chunks of 1024, chunk sums added in chunk order, so the
// result does not depend on the thread count (the chunks run on OpenMP
// threads). `slot(i)` is the 1-based slot of item i (0 skips it).
static constexpr long REDUCTION_CHUNK = 1024;

template <class Slot, class Term>
static std::vector<double> chunked_slot_sums(long n, size_t K, Slot&& slot, Term&& term) {
  std::vector<double> sums(K, 0.0);
  const long chunks = (n + REDUCTION_CHUNK - 1) / REDUCTION_CHUNK;
  if (chunks <= 1) {
    for (long i = 0; i < n; ++i) {
      const int k = slot(i);
      if (k) sums[k - 1] += term(i);
    }
    return sums;
  }
  std::vector<double> part(size_t(chunks) * K, 0.0);
#pragma omp parallel for schedule(static)
  for (long ch = 0; ch < chunks; ++ch) {
    double* pp = &part[size_t(ch) * K];
    const long hi = std::min(n, (ch + 1) * REDUCTION_CHUNK);
    for (long i = ch * REDUCTION_CHUNK; i < hi; ++i) {
      const int k = slot(i);
      if (k) pp[k - 1] += term(i);
    }
  }
  for (long ch = 0; ch < chunks; ++ch)
    for (size_t k = 0; k < K; ++k) sums[k] += part[size_t(ch) * K + k];
  return sums;
}


// COO -> CSC with duplicate summation in input order (SparseArrays `sparse`).
// Explicit zeros are kept: the reference keeps cancelled stiffness entries in
// the pattern (elliptic.jl:166-171).
static CSC coo_to_csc(int m, int n, const std::vector<int>& r, const std::vector<int>& c,
                      const std::vector<double>& v) {
  std::vector<int> order(r.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = int(q);
  __gnu_parallel::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return c[a] != c[b] ? c[a] < c[b] : r[a] < r[b];
  });
  CSC A;
  A.m = m;
  A.n = n;
  A.colptr.assign(n + 1, 0);
  size_t q = 0;
  for (int col = 0; col < n; ++col) {
    while (q < order.size() && c[order[q]] == col) {
      const int row = r[order[q]];
      double s = v[order[q]];
      ++q;
      while (q < order.size() && c[order[q]] == col && r[order[q]] == row) {
        s += v[order[q]];
        ++q;
      }
      A.rowval.push_back(row);
      A.nzval.push_back(s);
    }
    A.colptr[col + 1] = int(A.rowval.size());
  }
  return A;
}

static int csc_find(const CSC& A, int row, int col) {
  const int* b = A.rowval.data() + A.colptr[col];
  const int* e = A.rowval.data() + A.colptr[col + 1];
  const int* p = std::lower_bound(b, e, row);
  return (p != e && *p == row) ? int(p - A.rowval.data()) : -1;
}

// `_integrated_stiffness`, elliptic.jl:175-257: symmetric (A + A')/2 of
// J' M J with M = kappa K (diagonal K) or sqrt(kappa) K sqrt(kappa).
static CSC integrated_stiffness(const Mesh& M, const CSC& K, bool diag,
                                const std::vector<double>& kappa,
                                const std::vector<double>& sq) {
  const int N = M.ncells();
  std::vector<int> r, c;
  std::vector<double> v;
  r.reserve(4 * K.nnz() + N);
  c.reserve(4 * K.nnz() + N);
  v.reserve(4 * K.nnz() + N);
  for (int rf = 0; rf < K.n; ++rf)
    for (int p = K.colptr[rf]; p < K.colptr[rf + 1]; ++p) {
      const int lf = K.rowval[p];
      const double w = diag ? kappa[rf] * K.nzval[p] : (sq[lf] * K.nzval[p]) * sq[rf];
      const Face& L = M.faces[lf];
      const Face& R = M.faces[rf];
      const int lc[2] = {L.minus, L.plus}, rc[2] = {R.minus, R.plus};
      const double ls[2] = {-L.area, L.area}, rs[2] = {-R.area, R.area};
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          r.push_back(lc[a]);
          c.push_back(rc[b]);
          v.push_back(ls[a] * w * rs[b]);
        }
    }
  for (int cell = 0; cell < N; ++cell) {
    r.push_back(cell);
    c.push_back(cell);
    v.push_back(0.0);
  }
  CSC A = coo_to_csc(N, N, r, c, v);
  std::vector<double> sym(A.nnz());
  for (int col = 0; col < N; ++col)
    for (int p = A.colptr[col]; p < A.colptr[col + 1]; ++p) {
      const int mirror = csc_find(A, col, A.rowval[p]);
      if (mirror < 0) throw std::runtime_error("sparse topology is missing a required entry");
      const double val = (A.nzval[p] + A.nzval[mirror]) / 2;
      sym[p] = val == 0.0 ? 0.0 : val;
    }
  A.nzval = sym;
  return A;
}

// `_projection_component_labels!` (two_phase/projection/common.jl:53-87) and
// `_refresh_elliptic_components!` (elliptic_solver.jl:39-109) for an operator
// with all cells active: a component is gauged (has no slot) when one of its
// cells has a nonzero boundary diagonal (an open Dirichlet side), ungauged
// otherwise.
static Components components(const Mesh& M, const std::vector<double>& kappa,
                             const std::vector<double>& volumes, const std::vector<double>& bdiag) {
  const int N = M.ncells();
  std::vector<std::vector<int>> inc(N);
  for (int f = 0; f < int(M.faces.size()); ++f) {
    inc[M.faces[f].minus].push_back(f);
    if (M.faces[f].plus != M.faces[f].minus) inc[M.faces[f].plus].push_back(f);
  }
  Components C;
  C.labels.assign(N, -1);
  std::vector<int> stack;
  for (int root = 0; root < N; ++root) {
    if (!(volumes[root] > 0) || C.labels[root] >= 0) continue;
    C.labels[root] = root;
    stack.assign(1, root);
    while (!stack.empty()) {
      const int cell = stack.back();
      stack.pop_back();
      for (int f : inc[cell]) {
        if (kappa[f] == 0.0) continue;
        const int nb = cell == M.faces[f].minus ? M.faces[f].plus : M.faces[f].minus;
        if (C.labels[nb] >= 0 || !(volumes[nb] > 0)) continue;
        C.labels[nb] = root;
        stack.push_back(nb);
      }
    }
  }
  std::vector<double> vol(N, 0.0);
  std::vector<unsigned char> gauged(N, 0);
  for (int i = 0; i < N; ++i)
    if (C.labels[i] >= 0) {
      vol[C.labels[i]] += volumes[i];
      if (bdiag[i] != 0.0) gauged[C.labels[i]] = 1;
    }
  C.slots.assign(N, 0);
  for (int i = 0; i < N; ++i)
    if (C.labels[i] == i && !gauged[i]) {
      C.anchors.push_back(i);
      C.volumes.push_back(vol[i]);
      C.slots[i] = int(C.anchors.size());
    }
  for (int i = 0; i < N; ++i)
    if (C.labels[i] >= 0 && C.labels[i] != i) C.slots[i] = C.slots[C.labels[i]];
  return C;
}

// `_boundary_material` (elliptic.jl:90-131) for a vector of interior
// coefficients: every open boundary face takes the area-weighted mean of the
// interior faces on the opposite side of its cell along the same axis, summed
// in face order; closed faces get 0.
static std::vector<double> boundary_material(const Mesh& M, const std::vector<double>& kappa, const bool open[4]) {
  const int BF = int(M.bfaces.size());
  std::vector<double> values(BF, 0.0), weights(BF, 0.0);
  auto is_open = [&](const BFace& b) { return open[2 * b.axis + (b.side > 0 ? 1 : 0)]; };
  bool any = false;
  for (int s = 0; s < 4; ++s) any = any || open[s];
  if (!any) return values;
  std::map<std::tuple<int, int, int>, int> index;
  for (int f = 0; f < BF; ++f)
    if (is_open(M.bfaces[f])) index[{M.bfaces[f].cell, M.bfaces[f].axis, M.bfaces[f].side}] = f;
  for (int f = 0; f < int(M.faces.size()); ++f) {
    const Face& fa = M.faces[f];
    const std::pair<int, int> ends[2] = {{fa.minus, -1}, {fa.plus, 1}};
    for (const auto& e : ends) {
      auto it = index.find({e.first, fa.axis, e.second});
      if (it == index.end()) continue;
      values[it->second] += fa.area * kappa[f];
      weights[it->second] += fa.area;
    }
  }
  for (int f = 0; f < BF; ++f) {
    if (!is_open(M.bfaces[f])) continue;
    if (!(weights[f] > 0))
      throw std::invalid_argument(
          "cannot extend an interior coefficient to an open face without an opposite interior face; supply "
          "boundary_face_coefficient");
    values[f] /= weights[f];
  }
  return values;
}

// EllipticOperator / PoissonOperator (elliptic.jl:683-800, 969-980) with
// reaction 0, diffusion 1 and all cells active. `volumes` (NULL: the cells'
// geometric volumes) is the operator's `volumes` keyword (elliptic.jl:683-700);
// the stiffness does not depend on it. `open` (NULL: none) are the open
// pressure sides (left, right, bottom, top) of `velocity_boundary`
// (elliptic.jl:76-88) and `bcoef` (NULL: extended from `kappa`,
// `_boundary_material`) the boundary face coefficients; the boundary diagonal
// of `_elliptic_boundary_data!` (elliptic.jl:133-152) joins the stiffness
// diagonal (elliptic.jl:752-764).
static std::unique_ptr<Operator> make_operator(const Mesh* mesh, const std::vector<double>& kappa,
                                               const double* volumes = nullptr, const bool* open = nullptr,
                                               const double* bcoef = nullptr) {
  auto op = std::make_unique<Operator>();
  op->mesh = mesh;
  if (kappa.size() != mesh->faces.size())
    throw std::invalid_argument("one coefficient is required per interior face");
  for (double k : kappa) {
    if (!std::isfinite(k)) throw std::invalid_argument("face coefficients must be finite");
    if (!(k >= 0)) throw std::invalid_argument("face coefficients must be nonnegative");
  }
  op->kappa = kappa;
  op->sqrt_kappa.resize(kappa.size());
  for (size_t f = 0; f < kappa.size(); ++f) op->sqrt_kappa[f] = std::sqrt(kappa[f]);
  op->K = inverse_metric(*mesh, &op->K_diagonal);
  op->volumes.resize(mesh->ncells());
  const int nb = mesh->bs[0] * mesh->bs[1];
  for (int c = 0; c < mesh->ncells(); ++c) op->volumes[c] = volumes ? volumes[c] : mesh->blocks[c / nb].volume;
  for (double v : op->volumes)
    if (!(std::isfinite(v) && v >= 0)) throw std::invalid_argument("cell volumes must be finite and nonnegative");
  op->A = integrated_stiffness(*mesh, op->K, op->K_diagonal, op->kappa, op->sqrt_kappa);
  {
    const int N = mesh->ncells(), F = int(mesh->faces.size());
    op->inc_ptr.assign(N + 1, 0);
    for (int f = 0; f < F; ++f) {
      op->inc_ptr[mesh->faces[f].minus + 1]++;
      op->inc_ptr[mesh->faces[f].plus + 1]++;
    }
    for (int c = 0; c < N; ++c) op->inc_ptr[c + 1] += op->inc_ptr[c];
    op->inc_face.resize(op->inc_ptr[N]);
    op->inc_sign.resize(op->inc_ptr[N]);
    std::vector<int> next(op->inc_ptr.begin(), op->inc_ptr.end() - 1);
    for (int f = 0; f < F; ++f) {
      const int m = mesh->faces[f].minus, pl = mesh->faces[f].plus;
      op->inc_face[next[m]] = f;
      op->inc_sign[next[m]++] = -1;
      op->inc_face[next[pl]] = f;
      op->inc_sign[next[pl]++] = 1;
    }
  }
  // open pressure sides and the boundary diagonal
  const int BF = int(mesh->bfaces.size());
  for (int s = 0; s < 4; ++s) op->open[s] = open ? open[s] : false;
  if (bcoef) {
    op->bcoef.assign(bcoef, bcoef + BF);
  } else {
    op->bcoef = boundary_material(*mesh, op->kappa, op->open);
  }
  for (double k : op->bcoef)
    if (!(std::isfinite(k) && k >= 0)) throw std::invalid_argument("boundary coefficients must be finite and nonnegative");
  op->bgrad.assign(BF, 0.0);
  op->bdiag.assign(mesh->ncells(), 0.0);
  for (int f = 0; f < BF; ++f) {
    const BFace& b = mesh->bfaces[f];
    if (!op->open[2 * b.axis + (b.side > 0 ? 1 : 0)]) continue;
    const double factor = 2 * op->bcoef[f] / mesh->blocks[b.cell / nb].spacing[b.axis];
    op->bgrad[f] = factor;
    op->bdiag[b.cell] += b.area * 1.0 * factor;
  }
  for (int c = 0; c < mesh->ncells(); ++c) op->A.nzval[csc_find(op->A, c, c)] += op->bdiag[c];
  op->comp = components(*mesh, op->kappa, op->volumes, op->bdiag);
  op->singular = !op->comp.anchors.empty();
  return op;
}

// Factorized action `mul!(::FactorizedEllipticOperator)`, elliptic.jl:1125-1165
static void factorized_mul(const Operator& op, const double* x, double* y) {
  const Mesh& M = *op.mesh;
  const int F = int(M.faces.size());
  std::vector<double> fin(F), fout(F, 0.0);
#pragma omp parallel for schedule(static)
  for (int f = 0; f < F; ++f) {
    const Face& fa = M.faces[f];
    fin[f] = op.sqrt_kappa[f] * fa.area * (x[fa.plus] - x[fa.minus]);
  }
  if (op.K_diagonal) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; ++f) fout[f] = op.K.nzval[op.K.colptr[f]] * fin[f];
  } else {
    // CSR row product on the symmetric K (column r == row r)
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; ++f) {
      double s = 0;
      for (int p = op.K.colptr[f]; p < op.K.colptr[f + 1]; ++p) s += op.K.nzval[p] * fin[op.K.rowval[p]];
      fout[f] = s;
    }
  }
#pragma omp parallel for schedule(static)
  for (int f = 0; f < F; ++f) fin[f] = 1.0 * M.faces[f].area * op.sqrt_kappa[f] * fout[f];
  // cell gather in face order (the cell-face incidence order): the sums of
  // the face loop's scatter, term for term
  const int N = M.ncells();
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; ++c) {
    // (reaction V + diffusion B) x with reaction 0, diffusion 1
    double v = (0.0 * op.volumes[c] + 1.0 * op.bdiag[c]) * x[c];
    for (int q = op.inc_ptr[c]; q < op.inc_ptr[c + 1]; ++q) {
      const double flux = fin[op.inc_face[q]];
      v += op.inc_sign[q] < 0 ? -1.0 * flux : flux;
    }
    y[c] = v;
  }
}

static void csc_sym_mul(const CSC& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A.n; ++i) {
    double s = 0;
    for (int p = A.colptr[i]; p < A.colptr[i + 1]; ++p) s += A.nzval[p] * x[A.rowval[p]];
    y[i] = s;
  }
}

// `_jacobi_weight!`, hierarchy.jl:8-30
static double jacobi_weight(std::vector<double>& invdiag, const std::vector<double>& rowsum,
                            bool skip_nonpositive, bool allow_zero, bool clamp, double zero_weight) {
  double bound = 0;
  for (size_t i = 0; i < invdiag.size(); ++i) {
    const double d = invdiag[i];
    if (d == 0.0 && allow_zero) continue;
    if (!(d > 0)) {
      if (skip_nonpositive) continue;
      throw std::invalid_argument("elliptic matrix has a non-positive diagonal");
    }
    bound = std::max(bound, rowsum[i] / d);
    invdiag[i] = 1.0 / d;
  }
  if (bound == 0.0) return zero_weight;
  const double w = 4.0 / (3.0 * bound);
  return clamp ? std::min(1.0, w) : w;
}

// ---------------------------------------------------------------------------
// Tree coarsening, src/solvers/multigrid/tree_coarsening.jl

// `_all_leaf_pruned_mesh`, tree_coarsening.jl:30-57
static std::unique_ptr<Mesh> all_leaf_pruned(const Mesh& M) {
  std::vector<Key> retained;
  int maxl = -1;
  for (uint64_t pk : M.internal) {
    const int level = int(pk >> 56), i = int((pk >> 28) & 0xFFFFFFF), j = int(pk & 0xFFFFFFF);
    bool has_internal_child = false;
    for (int c = 0; c < 4 && !has_internal_child; ++c)
      has_internal_child = M.internal.count(kpack(level + 1, 2 * i + (c & 1), 2 * j + (c >> 1))) > 0;
    if (has_internal_child) {
      retained.push_back({level, i, j});
      maxl = std::max(maxl, level);
    }
  }
  const bool per[2] = {M.periodic[0], M.periodic[1]};
  return build_mesh(M.base, M.bs, M.origin, M.extent, per, maxl + 1, retained);
}

// `_coarsened_base_mesh`, tree_coarsening.jl:62-82
static std::unique_ptr<Mesh> coarsened_base(const Mesh& M) {
  if (M.layout[0] % 2 || M.layout[1] % 2) return nullptr;
  const int cb[2] = {M.base[0] / 2, M.base[1] / 2};
  for (int a = 0; a < 2; ++a)
    if (M.periodic[a] && cb[a] < 4) return nullptr;
  const bool per[2] = {M.periodic[0], M.periodic[1]};
  return build_mesh(cb, M.bs, M.origin, M.extent, per, 0, {});
}

static void common_lattice(const Mesh& f, const Mesh& c, long lat[2]) {
  for (int a = 0; a < 2; ++a)
    lat[a] = std::max(long(f.base[a]) << f.max_level, long(c.base[a]) << c.max_level);
}

static void cell_bounds(const Mesh& M, int cell, const long lat[2], long low[2], long scale[2]) {
  const int nb = M.bs[0] * M.bs[1];
  const int b = cell / nb, r = cell % nb;
  const int i = r % M.bs[0], j = r / M.bs[0];
  const Key& k = M.blocks[b].key;
  scale[0] = lat[0] / (long(M.base[0]) << k.level);
  scale[1] = lat[1] / (long(M.base[1]) << k.level);
  low[0] = (long(k.i) * M.bs[0] + i) * scale[0];
  low[1] = (long(k.j) * M.bs[1] + j) * scale[1];
}

// `_tree_cell_prolongation`, tree_coarsening.jl:228-252: owner coarse cell of
// each fine cell (the coarse cell containing the fine cell centre).
static std::vector<int> tentative_owners(const Mesh& fine, const Mesh& coarse) {
  long lat[2];
  common_lattice(fine, coarse, lat);
  std::unordered_map<uint64_t, int> leaf;
  for (int b = 0; b < int(coarse.blocks.size()); ++b)
    leaf[kpack(coarse.blocks[b].key.level, coarse.blocks[b].key.i, coarse.blocks[b].key.j)] = b;
  const int maxl = coarse.max_leaf_level();
  std::vector<int> owner(fine.ncells());
  for (int cell = 0; cell < fine.ncells(); ++cell) {
    long low[2], sc[2];
    cell_bounds(fine, cell, lat, low, sc);
    const long px = low[0] + sc[0] / 2, py = low[1] + sc[1] / 2;
    int found = -1;
    for (int l = 0; l <= maxl && found < 0; ++l) {
      const long bw0 = lat[0] / (long(coarse.layout[0]) << l), bw1 = lat[1] / (long(coarse.layout[1]) << l);
      auto it = leaf.find(kpack(l, int(px / bw0), int(py / bw1)));
      if (it == leaf.end()) continue;
      const int b = it->second;
      const long cs0 = lat[0] / ((long(coarse.layout[0]) * coarse.bs[0]) << l);
      const long cs1 = lat[1] / ((long(coarse.layout[1]) * coarse.bs[1]) << l);
      const int ci = int((px / cs0) % coarse.bs[0]), cj = int((py / cs1) % coarse.bs[1]);
      found = b * coarse.bs[0] * coarse.bs[1] + cj * coarse.bs[0] + ci;
    }
    if (found < 0) throw std::runtime_error("tentative transfer leaves a cell unowned");
    owner[cell] = found;
  }
  return owner;
}

struct FaceRestriction {
  std::vector<int> ptr, idx;
  std::vector<double> w;
};

// `_face_coefficient_restriction`, tree_coarsening.jl:347-398
static FaceRestriction face_restriction(const Mesh& coarse, const Mesh& fine) {
  long lat[2];
  common_lattice(fine, coarse, lat);
  std::map<std::pair<int, long>, std::vector<std::tuple<long, long, int>>> lines;
  for (int f = 0; f < int(fine.faces.size()); ++f) {
    long n, lo, hi;
    long ml[2], ms[2], pl[2], ps[2];
    const Face& fa = fine.faces[f];
    cell_bounds(fine, fa.minus, lat, ml, ms);
    cell_bounds(fine, fa.plus, lat, pl, ps);
    const int a = fa.axis, t = 1 - a;
    n = ml[a] + ms[a];
    if (fine.periodic[a]) n = ((n % lat[a]) + lat[a]) % lat[a];
    lo = std::max(ml[t], pl[t]);
    hi = std::min(ml[t] + ms[t], pl[t] + ps[t]);
    lines[{a, n}].emplace_back(lo, hi, f);
  }
  for (auto& kv : lines)
    std::stable_sort(kv.second.begin(), kv.second.end(),
                     [](const auto& x, const auto& y) { return std::get<0>(x) < std::get<0>(y); });
  FaceRestriction R;
  R.ptr.push_back(0);
  for (const Face& fa : coarse.faces) {
    long ml[2], ms[2], pl[2], ps[2];
    cell_bounds(coarse, fa.minus, lat, ml, ms);
    cell_bounds(coarse, fa.plus, lat, pl, ps);
    const int a = fa.axis, t = 1 - a;
    long n = ml[a] + ms[a];
    if (coarse.periodic[a]) n = ((n % lat[a]) + lat[a]) % lat[a];
    const long lo = std::max(ml[t], pl[t]), hi = std::min(ml[t] + ms[t], pl[t] + ps[t]);
    long covered = 0;
    auto it = lines.find({a, n});
    if (it != lines.end()) {
      for (const auto& seg : it->second) {
        const long flo = std::get<0>(seg), fhi = std::get<1>(seg);
        if (fhi <= lo) continue;
        if (flo >= hi) break;
        const long overlap = std::min(hi, fhi) - std::max(lo, flo);
        if (overlap > 0) {
          R.idx.push_back(std::get<2>(seg));
          R.w.push_back(double(overlap) / double(hi - lo));
          covered += overlap;
        }
      }
    }
    if (covered != hi - lo) throw std::runtime_error("tree hierarchy could not cover a coarse face with fine faces");
    R.ptr.push_back(int(R.idx.size()));
  }
  return R;
}

// `_face_coefficient_restriction(...; boundary = true)`
// (tree_coarsening.jl:347-398) with the boundary-face lattice geometry of
// mesh.jl:476-485: each coarse boundary face takes the overlap-weighted fine
// boundary faces on its line (axis, normal coordinate).
static FaceRestriction boundary_restriction(const Mesh& coarse, const Mesh& fine) {
  long lat[2];
  common_lattice(fine, coarse, lat);
  auto geom = [&](const Mesh& M, const BFace& b, long& n, long& lo, long& hi) {
    long low[2], sc[2];
    cell_bounds(M, b.cell, lat, low, sc);
    const int a = b.axis, t = 1 - a;
    n = low[a] + (b.side > 0 ? sc[a] : 0);
    lo = low[t];
    hi = low[t] + sc[t];
  };
  std::map<std::pair<int, long>, std::vector<std::tuple<long, long, int>>> lines;
  for (int f = 0; f < int(fine.bfaces.size()); ++f) {
    long n, lo, hi;
    geom(fine, fine.bfaces[f], n, lo, hi);
    lines[{fine.bfaces[f].axis, n}].emplace_back(lo, hi, f);
  }
  for (auto& kv : lines)
    std::stable_sort(kv.second.begin(), kv.second.end(),
                     [](const auto& x, const auto& y) { return std::get<0>(x) < std::get<0>(y); });
  FaceRestriction R;
  R.ptr.push_back(0);
  for (const BFace& b : coarse.bfaces) {
    long n, lo, hi;
    geom(coarse, b, n, lo, hi);
    long covered = 0;
    auto it = lines.find({b.axis, n});
    if (it != lines.end()) {
      for (const auto& seg : it->second) {
        const long flo = std::get<0>(seg), fhi = std::get<1>(seg);
        if (fhi <= lo) continue;
        if (flo >= hi) break;
        const long overlap = std::min(hi, fhi) - std::max(lo, flo);
        if (overlap > 0) {
          R.idx.push_back(std::get<2>(seg));
          R.w.push_back(double(overlap) / double(hi - lo));
          covered += overlap;
        }
      }
    }
    if (covered != hi - lo) throw std::runtime_error("tree hierarchy could not cover a coarse face with fine faces");
    R.ptr.push_back(int(R.idx.size()));
  }
  return R;
}

// `_smoothed_transfer_refresh` + `_refresh_smoothed_transfer!`,
// hierarchy.jl:631-721: P = (I - w_P D^-1 S) P_t on the pattern
// P_t union pattern(S) P_t.
static CSC smoothed_transfer(const CSC& S, const std::vector<int>& owner, int ncoarse) {
  const int nf = S.n;
  std::vector<int> r, c;
  std::vector<double> v;
  for (int row = 0; row < nf; ++row) {
    r.push_back(row);
    c.push_back(owner[row]);
    v.push_back(0.0);
  }
  for (int col = 0; col < nf; ++col)
    for (int p = S.colptr[col]; p < S.colptr[col + 1]; ++p) {
      r.push_back(S.rowval[p]);
      c.push_back(owner[col]);
      v.push_back(0.0);
    }
  CSC P = coo_to_csc(nf, ncoarse, r, c, v);
  std::vector<double> invd(nf), rowsum(nf);
  for (int col = 0; col < nf; ++col) {
    double s = 0, d = 0;
    for (int p = S.colptr[col]; p < S.colptr[col + 1]; ++p) {
      s += std::fabs(S.nzval[p]);
      if (S.rowval[p] == col) d = S.nzval[p];
    }
    rowsum[col] = s;
    invd[col] = d;
  }
  const double w = jacobi_weight(invd, rowsum, true, false, false, 0.0);
  for (int row = 0; row < nf; ++row) P.nzval[csc_find(P, row, owner[row])] = 1.0;
  // contributions in stiffness-pointer order
  for (int col = 0; col < nf; ++col)
    for (int p = S.colptr[col]; p < S.colptr[col + 1]; ++p) {
      const int row = S.rowval[p];
      const int q = csc_find(P, row, owner[col]);
      P.nzval[q] -= w * invd[row] * S.nzval[p];
    }
  return P;
}

// `_refresh_multigrid_components!` (hierarchy.jl:276-298): the level's
// ungauged components only
static void level_components(Level& L) {
  const Components& C = L.op->comp;
  const int N = int(L.op->volumes.size());
  L.comp_index.assign(N, 0);
  L.comp_volumes.clear();
  for (int cell = 0; cell < N; ++cell) {
    const int root = C.labels[cell];
    if (root < 0 || C.slots[root] == 0) continue;
    if (cell == root) {
      L.comp_volumes.push_back(C.volumes[C.slots[root] - 1]);
      L.comp_index[cell] = int(L.comp_volumes.size());
    } else {
      L.comp_index[cell] = L.comp_index[root];
    }
  }
}

static void lu_factor(std::vector<double>& a, std::vector<int>& piv, int n) {
  piv.resize(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[size_t(k) * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[size_t(i) * n + k]) > best) {
        best = std::fabs(a[size_t(i) * n + k]);
        p = i;
      }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[size_t(k) * n + j], a[size_t(p) * n + j]);
    const double d = a[size_t(k) * n + k];
    if (d == 0.0) throw std::runtime_error("singular coarse matrix");
    for (int i = k + 1; i < n; ++i) {
      const double l = a[size_t(i) * n + k] / d;
      a[size_t(i) * n + k] = l;
      if (l != 0.0)
        for (int j = k + 1; j < n; ++j) a[size_t(i) * n + j] -= l * a[size_t(k) * n + j];
    }
  }
}

static void lu_solve(const std::vector<double>& a, const std::vector<int>& piv, int n, double* x) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(x[k], x[piv[k]]);
  for (int i = 0; i < n; ++i) {
    double s = x[i];
    for (int j = 0; j < i; ++j) s -= a[size_t(i) * n + j] * x[j];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int j = i + 1; j < n; ++j) s -= a[size_t(i) * n + j] * x[j];
    x[i] = s / a[size_t(i) * n + i];
  }
}

// TreeGeometricMultigridPreconditioner constructor (pruning = :all, no
// reuse, no geometry), geometric_preconditioner.jl:171-442, with
// `_tree_next_coarse_mesh!` (tree_coarsening.jl:187-212).
Solver* build_solver(const Mesh& fine, const std::vector<double>& kappa, int coarse_size, int pre,
                     int post, int max_levels, bool use_mg, const double* volumes, const bool* open,
                     const double* bcoef) {
  auto S = new Solver();
  S->pre = pre;
  S->post = post;
  S->coarse_size = coarse_size;
  S->max_levels = max_levels;
  S->use_mg = use_mg;
  // coarse levels keep their geometric volumes (`_cell_volumes` of each
  // coarse mesh, geometric_preconditioner.jl:75-105)
  S->ops.push_back(make_operator(&fine, kappa, volumes, open, bcoef));
  std::vector<CSC> prolongations;
  if (use_mg) {
    int active_level = fine.max_leaf_level();
    const Mesh* current = &fine;
    for (;;) {
      std::unique_ptr<Mesh> coarse;
      while (current->ncells() > coarse_size && int(S->ops.size()) < max_levels) {
        std::unique_ptr<Mesh> cand = active_level == 0 ? coarsened_base(*current) : all_leaf_pruned(*current);
        if (!cand) break;
        if (cand->ncells() >= current->ncells()) break;
        active_level = cand->max_leaf_level();
        coarse = std::move(cand);
        break;
      }
      if (!coarse) break;
      const std::vector<int> owner = tentative_owners(*current, *coarse);
      prolongations.push_back(smoothed_transfer(S->ops.back()->A, owner, coarse->ncells()));
      const FaceRestriction R = face_restriction(*coarse, *current);
      std::vector<double> ck(coarse->faces.size());
      const std::vector<double>& fk = S->ops.back()->kappa;
      for (size_t f = 0; f < ck.size(); ++f) {
        double s = 0;
        for (int q = R.ptr[f]; q < R.ptr[f + 1]; ++q) s += R.w[q] * fk[R.idx[q]];
        ck[f] = s;
      }
      // coarse boundary coefficients (geometric_preconditioner.jl:249-251,
      // 283-287): overlap-weighted sums in restriction order
      const FaceRestriction RB = boundary_restriction(*coarse, *current);
      std::vector<double> cb(coarse->bfaces.size());
      const std::vector<double>& fb = S->ops.back()->bcoef;
      for (size_t f = 0; f < cb.size(); ++f) {
        double s = 0;
        for (int q = RB.ptr[f]; q < RB.ptr[f + 1]; ++q) s += RB.w[q] * fb[RB.idx[q]];
        cb[f] = s;
      }
      Mesh* cm = coarse.release();
      auto op = make_operator(cm, ck, nullptr, S->ops[0]->open, cb.data());
      op->owned_mesh.reset(cm);
      S->ops.push_back(std::move(op));
      current = cm;
    }
  }
  const int nlev = int(S->ops.size());
  S->levels.resize(nlev);
  for (int l = 0; l < nlev; ++l) {
    Level& L = S->levels[l];
    L.op = S->ops[l].get();
    const CSC& A = L.op->A;
    const int N = A.n;
    L.invdiag.assign(N, 0.0);
    std::vector<double> rowsum(N, 0.0);
    for (int col = 0; col < N; ++col)
      for (int p = A.colptr[col]; p < A.colptr[col + 1]; ++p) {
        rowsum[col] += std::fabs(A.nzval[p]);
        if (A.rowval[p] == col) L.invdiag[col] = A.nzval[p];
      }
    L.omega = jacobi_weight(L.invdiag, rowsum, false, L.op->singular, true, 1.0);
    if (l + 1 < nlev) L.P = prolongations[l];
    level_components(L);
    L.x.assign(N, 0.0);
    L.b.assign(N, 0.0);
    L.r.assign(N, 0.0);
    L.t.assign(N, 0.0);
  }
  // coarse factorization, geometric_preconditioner.jl:144-169, 416-430
  if (!use_mg) return S;
  Level& C = S->levels.back();
  const int n = int(C.op->volumes.size());
  const int k = int(C.comp_volumes.size());
  S->coarse_augmented = k > 0;
  S->coarse_n = n + k;
  S->lu.assign(size_t(n + k) * (n + k), 0.0);
  const CSC& A = C.op->A;
  for (int col = 0; col < n; ++col)
    for (int p = A.colptr[col]; p < A.colptr[col + 1]; ++p)
      S->lu[size_t(A.rowval[p]) * (n + k) + col] = A.nzval[p];
  for (int cell = 0; cell < n; ++cell)
    if (C.comp_index[cell]) {
      const int comp = n + C.comp_index[cell] - 1;
      S->lu[size_t(cell) * (n + k) + comp] = C.op->volumes[cell];
      S->lu[size_t(comp) * (n + k) + cell] = C.op->volumes[cell];
    }
  lu_factor(S->lu, S->piv, n + k);
  // fine factorized action incidence is implicit in factorized_mul
  return S;
}

// ---------------------------------------------------------------------------
// V-cycle, src/solvers/multigrid/cycles.jl:8-77 and hierarchy.jl:208-343

// `_multigrid_project_component_rhs!` (hierarchy.jl:326-343)
static void project_component_rhs(Level& L, std::vector<double>& b) {
  const size_t K = L.comp_volumes.size();
  const int* idx = L.comp_index.data();
  std::vector<double> s = chunked_slot_sums(long(b.size()), K, [&](long c) { return idx[c]; },
                                            [&](long c) { return b[c]; });
  for (size_t q = 0; q < K; ++q) s[q] /= L.comp_volumes[q];
  const double* V = L.op->volumes.data();
#pragma omp parallel for schedule(static)
  for (long c = 0; c < long(b.size()); ++c)
    if (idx[c]) b[c] -= V[c] * s[idx[c] - 1];
}

// `_multigrid_remove_component_means!` (hierarchy.jl:309-324)
static void remove_component_means(Level& L, std::vector<double>& x) {
  const size_t K = L.comp_volumes.size();
  const int* idx = L.comp_index.data();
  const double* V = L.op->volumes.data();
  std::vector<double> s = chunked_slot_sums(long(x.size()), K, [&](long c) { return idx[c]; },
                                            [&](long c) { return V[c] * x[c]; });
  for (size_t q = 0; q < K; ++q) s[q] /= L.comp_volumes[q];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < long(x.size()); ++c)
    if (idx[c]) x[c] -= s[idx[c] - 1];
}

static void smooth(Level& L, int sweeps, bool singular) {
  for (int s = 0; s < sweeps; ++s) {
    csc_sym_mul(L.op->A, L.x.data(), L.t.data());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(L.x.size()); ++i) L.x[i] += L.omega * L.invdiag[i] * (L.b[i] - L.t[i]);
    if (singular) remove_component_means(L, L.x);
  }
}

static void vcycle(Solver& S, int l) {
  Level& L = S.levels[l];
  const bool singular = !L.comp_volumes.empty();
  if (singular) project_component_rhs(L, L.b);
  if (l == int(S.levels.size()) - 1) {
    const int n = int(L.x.size());
    std::vector<double> rhs(S.coarse_n, 0.0);
    std::copy(L.b.begin(), L.b.end(), rhs.begin());
    lu_solve(S.lu, S.piv, S.coarse_n, rhs.data());
    std::copy(rhs.begin(), rhs.begin() + n, L.x.begin());
    return;
  }
  smooth(L, S.pre, singular);
  csc_sym_mul(L.op->A, L.x.data(), L.t.data());
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(L.r.size()); ++i) L.r[i] = L.b[i] - L.t[i];
  Level& C = S.levels[l + 1];
  // `_restrict!`: CSC column dots with muladd, plus zero
#pragma omp parallel for schedule(static)
  for (int col = 0; col < L.P.n; ++col) {
    double v = 0.0;
    for (int p = L.P.colptr[col]; p < L.P.colptr[col + 1]; ++p) v = std::fma(L.P.nzval[p], L.r[L.P.rowval[p]], v);
    C.b[col] = v + 0.0;
  }
  std::fill(C.x.begin(), C.x.end(), 0.0);
  vcycle(S, l + 1);
  // `_prolongate!`: rows in increasing column order with muladd
  if (L.prow_ptr.empty()) {
    std::vector<int> cnt(L.P.m + 1, 0);
    for (int p = 0; p < L.P.nnz(); ++p) cnt[L.P.rowval[p] + 1]++;
    for (int i = 0; i < L.P.m; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> cols(L.P.nnz()), ent(L.P.nnz()), next(cnt.begin(), cnt.end() - 1);
    for (int col = 0; col < L.P.n; ++col)
      for (int p = L.P.colptr[col]; p < L.P.colptr[col + 1]; ++p) {
        const int row = L.P.rowval[p];
        cols[next[row]] = col;
        ent[next[row]] = p;
        next[row]++;
      }
    L.prow_ptr = std::move(cnt);
    L.prow_col = std::move(cols);
    L.prow_ent = std::move(ent);
  }
#pragma omp parallel for schedule(static)
  for (int row = 0; row < L.P.m; ++row) {
    double v = 0.0;
    for (int q = L.prow_ptr[row]; q < L.prow_ptr[row + 1]; ++q)
      v = std::fma(L.P.nzval[L.prow_ent[q]], C.x[L.prow_col[q]], v);
    L.t[row] = v;
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(L.x.size()); ++i) L.x[i] += L.t[i];
  if (singular) remove_component_means(L, L.x);
  smooth(L, S.post, singular);
}

void apply_vcycle(Solver& S, const double* r, double* z) {
  Level& F = S.levels[0];
  std::copy(r, r + F.b.size(), F.b.begin());
  std::fill(F.x.begin(), F.x.end(), 0.0);
  vcycle(S, 0);
  std::copy(F.x.begin(), F.x.end(), z);
}

// ---------------------------------------------------------------------------
// Krylov driver, src/solvers/multigrid/elliptic_solver.jl:369-612 and
// cycles.jl:80-260. Krylov.jl 0.10 `cg!` (third-party; parity unpinned
// beyond iteration counts) is restated from its published algorithm.

// `_strong_residual_maximum` / `_strong_cells_maximum` (cycles.jl:226-260):
// max |r_i| / V_i over cells of positive volume; a NaN term is returned.
template <typename T>
static T strong_max(const std::vector<T>& r, const std::vector<T>& V) {
  const long n = long(r.size());
  const long chunks = (n + REDUCTION_CHUNK - 1) / REDUCTION_CHUNK;
  std::vector<T> part(std::max(chunks, 1L), T(0));
#pragma omp parallel for schedule(static)
  for (long ch = 0; ch < chunks; ++ch) {
    T m = 0;
    const long hi = std::min(n, (ch + 1) * REDUCTION_CHUNK);
    for (long i = ch * REDUCTION_CHUNK; i < hi; ++i) {
      if (!(V[i] > 0)) continue;
      const T q = std::fabs(r[i]) / V[i];
      if (std::isnan(q)) {
        m = q;
        break;
      }
      m = std::max(m, q);
    }
    part[ch] = m;
  }
  T m = 0;
  for (long ch = 0; ch < chunks; ++ch) {
    if (std::isnan(part[ch])) return part[ch];
    m = std::max(m, part[ch]);
  }
  return m;
}

// `_integrated_residual_norms` (cycles.jl:181-196)
template <typename T>
static void residual_norms(const std::vector<T>& r, const std::vector<T>& V, T& l2, T& linf) {
  T sq = 0, tv = 0;
  for (size_t i = 0; i < r.size(); ++i) {
    if (!(V[i] > 0)) continue;
    const T s = r[i] / V[i];
    sq += V[i] * s * s;
    tv += V[i];
  }
  l2 = std::sqrt(sq / tv);
  linf = strong_max(r, V);
}

// assembled `_roundoff_residual_floor!` (cycles.jl:80-101):
// 128 eps max(1, max_i (|A||x|)_i / V_i, max_i |b_i| / V_i)
template <typename T>
static T assembled_floor(int n, const int* colptr, const int* rowval, const T* nzval, const std::vector<T>& x,
                         const std::vector<T>& b, const std::vector<T>& V) {
  std::vector<T> t(n, T(0));
  for (int col = 0; col < n; ++col)
    for (int p = colptr[col]; p < colptr[col + 1]; ++p) t[rowval[p]] += std::fabs(nzval[p]) * std::fabs(x[col]);
  T scale = 1;
  for (int i = 0; i < n; ++i) {
    if (!(V[i] > 0)) continue;
    scale = std::max(scale, std::max(t[i] / V[i], std::fabs(b[i] / V[i])));
  }
  return T(128) * std::numeric_limits<T>::epsilon() * scale;
}

// factorized `_roundoff_residual_floor!`, cycles.jl:108-179
static double roundoff_floor(const Operator& op, const std::vector<double>& x,
                             const std::vector<double>& b) {
  const Mesh& M = *op.mesh;
  const int F = int(M.faces.size());
  std::vector<double> fin(F), fout(F);
#pragma omp parallel for schedule(static)
  for (int f = 0; f < F; ++f)
    fin[f] = op.sqrt_kappa[f] * M.faces[f].area * (std::fabs(x[M.faces[f].minus]) + std::fabs(x[M.faces[f].plus]));
#pragma omp parallel for schedule(static)
  for (int f = 0; f < F; ++f) {
    double s = 0;
    for (int p = op.K.colptr[f]; p < op.K.colptr[f + 1]; ++p) s += std::fabs(op.K.nzval[p]) * fin[op.K.rowval[p]];
    fout[f] = s;
  }
#pragma omp parallel for schedule(static)
  for (int f = 0; f < F; ++f) fin[f] = 1.0 * M.faces[f].area * op.sqrt_kappa[f] * fout[f];
  // each cell gathers its face bounds in incidence (face) order, the order of
  // the face loop's scatter (cycles.jl:108-112)
  const int N = M.ncells();
  std::vector<double> bound(N);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; ++c) {
    double v = (0.0 * op.volumes[c] + 1.0 * op.bdiag[c] + 0.0) * std::fabs(x[c]);
    for (int q = op.inc_ptr[c]; q < op.inc_ptr[c + 1]; ++q) v += fin[op.inc_face[q]];
    bound[c] = v;
  }
  double scale = 0;
  for (int c = 0; c < N; ++c) {
    const double V = op.volumes[c];
    if (!(V > 0)) continue;
    scale = std::max(scale, std::max(bound[c] / V, std::fabs(b[c]) / V));
  }
  return 128 * EPS * std::max(1.0, scale);
}

double solver_floor(Solver& S, const double* x, const double* b, bool assembled) {
  const Operator& op = *S.ops[0];
  const size_t n = op.volumes.size();
  std::vector<double> xv(x, x + n), bv(b, b + n);
  if (assembled)
    return assembled_floor<double>(int(n), op.A.colptr.data(), op.A.rowval.data(), op.A.nzval.data(), xv, bv,
                                   op.volumes);
  return roundoff_floor(op, xv, bv);
}

// `_remove_component_volume_means!` (elliptic_solver.jl:130-148)
static void remove_volume_means(const Operator& op, std::vector<double>& x) {
  const Components& C = op.comp;
  const int* slots = C.slots.data();
  const double* V = op.volumes.data();
  std::vector<double> s = chunked_slot_sums(long(x.size()), C.anchors.size(), [&](long c) { return slots[c]; },
                                            [&](long c) { return V[c] * x[c]; });
  for (size_t q = 0; q < s.size(); ++q) s[q] /= C.volumes[q];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < long(x.size()); ++c) {
    if (C.labels[c] < 0) x[c] = 0.0;
    else if (slots[c]) x[c] -= s[slots[c] - 1];
  }
}

// Krylov.jl's dot (a BLAS dot, whose summation order is the library's):
// index order. Fixed chunks instead would move plain CG on the techref
// composite 64^2 mesh from the published 780 iterations to 781.
static double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// Krylov.jl cg! with M = V-cycle (ldiv), warm start x0, atol = rtol = 0,
// itmax, and the strong-residual callback (elliptic_solver.jl:369-378).
static int krylov_cg(Solver& S, const std::vector<double>& b, std::vector<double>& x0,
                     double target, int itmax) {
  Operator& op = *S.ops[0];
  const std::vector<double>& V = op.volumes;
  const size_t n = b.size();
  std::vector<double> x(n, 0.0), r(n), p(n), Ap(n), z(n);
  factorized_mul(op, x0.data(), r.data());
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(n); ++i) r[i] = b[i] - r[i];
  auto precond = [&](const std::vector<double>& in, std::vector<double>& out) {
    if (S.use_mg) apply_vcycle(S, in.data(), out.data());
    else out = in;
  };
  precond(r, z);
  p = z;
  double gamma = dotv(r, z);
  if (gamma < 0) throw std::runtime_error("Preconditioner is not positive definite");
  double rnorm = std::sqrt(gamma);
  if (gamma == 0) {
  #pragma omp parallel for schedule(static)
  for (long i = 0; i < long(n); ++i) x0[i] = x[i] + x0[i];
    return 0;
  }
  int iter = 0;
  if (itmax == 0) itmax = 2 * int(n);
  double pnorm2 = gamma;
  const double eps_tol = 0.0;
  bool solved = rnorm <= eps_tol, tired = iter >= itmax, zero_curv = false, user_exit = false;
  while (!(solved || tired || zero_curv || user_exit)) {
    factorized_mul(op, p.data(), Ap.data());
    const double pAp = dotv(p, Ap);
    if (pAp <= EPS * pnorm2) {
      if (std::fabs(pAp) <= EPS * pnorm2) zero_curv = true;
    }
    if (zero_curv) continue;
    const double alpha = gamma / pAp;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
      x[i] += alpha * p[i];
      r[i] += -alpha * Ap[i];
    }
    precond(r, z);
    const double gamma_next = dotv(r, z);
    rnorm = std::sqrt(gamma_next);
    const bool mach = rnorm + 1.0 <= 1.0;
    solved = rnorm <= eps_tol || mach;
    if (!solved) {
      const double beta = gamma_next / gamma;
      pnorm2 = gamma_next + beta * beta * pnorm2;
      gamma = gamma_next;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < long(n); ++i) p[i] = z[i] + beta * p[i];
    }
    ++iter;
    tired = iter >= itmax;
    user_exit = strong_max(r, V) <= target;
  }
  for (long i = 0; i < long(n); ++i) x0[i] = x[i] + x0[i];
  return iter;
}

// `_guarded_roundoff_target!` (elliptic_solver.jl:454-473): the iterate's
// floor replaces the target only after progress (target < current <
// initial) and only while it is below initial / 100.
template <typename T, class Floor>
static T guarded_target(Floor&& floor, T target, T initial, T current) {
  if (!(std::isfinite(initial) && std::isfinite(current) && target < current && current < initial))
    return target;
  const T fl = floor();
  return (std::isfinite(fl) && fl < initial / 100) ? std::max(target, fl) : target;
}

template <typename T>
struct DriverStats {
  int cycles;
  bool converged;
  T initial_l2, initial_linf, final_l2, final_linf, target_linf, residual_scale;
};

// `_solve_krylov_with_stats!` (elliptic_solver.jl:550-612) with
// `_run_krylov_reliably!` (:487-526), independent of the operator and the
// Krylov method. `b` is the integrated (compatible) right-hand side and `x`
// the prepared initial iterate; `mul(x, y)` is the action y = A x,
// `floor(x)` the round-off floor of the iterate, `krylov(x, target, itmax)`
// one Krylov run that leaves Krylov.solution in `x` and returns
// Krylov.iteration_count, `gauge(x)` the component mean removal.
// `relative_to_rhs` selects the shifted-diffusion scale max |b_i| / V_i (atol
// unused) instead of the initial residual.
template <typename T, class Mul, class Floor, class Krylov, class Gauge>
static DriverStats<T> krylov_driver(const std::vector<T>& V, const std::vector<T>& b, std::vector<T>& x, T rtol,
                                    T atol, int max_cycles, bool relative_to_rhs, Mul&& mul, Floor&& floor,
                                    Krylov&& krylov, Gauge&& gauge) {
  const size_t n = V.size();
  const T eps = std::numeric_limits<T>::epsilon();
  std::vector<T> prod(n), res(n);
  auto true_residual = [&]() {
    mul(x, prod);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) res[i] = b[i] - prod[i];
  };
  true_residual();
  DriverStats<T> st{};
  residual_norms(res, V, st.initial_l2, st.initial_linf);
  const T ilinf = st.initial_linf;
  const T floor0 = floor(x);
  T target, scale;
  if (relative_to_rhs) {
    scale = strong_max(b, V);
    target = std::max(rtol * scale, floor0);
  } else {
    scale = std::max(ilinf, eps);
    target = std::max(atol, std::max(rtol * std::max(ilinf, eps), floor0));
  }
  int iterations = 0;
  if (ilinf > target) {
    int total = 0;
    while (total < max_cycles) {
      const int it = krylov(x, target, max_cycles - total);
      total += it;
      gauge(x);
      true_residual();
      const T final_linf = strong_max(res, V);
      const T eff = guarded_target([&] { return floor(x); }, target, ilinf, final_linf);
      if (final_linf <= eff) break;
      if (!std::isfinite(final_linf) || it == 0) break;
    }
    iterations = total;
  }
  gauge(x);
  true_residual();
  residual_norms(res, V, st.final_l2, st.final_linf);
  target = guarded_target([&] { return floor(x); }, target, ilinf, st.final_linf);
  st.cycles = iterations;
  st.converged = std::isfinite(st.final_linf) && std::isfinite(target) && st.final_linf <= target;
  st.target_linf = target;
  st.residual_scale = scale;
  return st;
}

// The driver above with a scripted Krylov method, as the reference's mock
// workspaces (test/solvers.jl:698-751): run k (1-based) leaves
// iterates[min(k, npasses)] as its solution and reports one iteration. The
// operator is an assembled CSC matrix with its assembled floor.
template <typename T>
static DriverStats<T> scripted_driver(int n, const int* colptr, const int* rowval, const std::vector<T>& nz,
                                      const std::vector<T>& V, const std::vector<T>& b, std::vector<T>& x, T rtol,
                                      T atol, int max_cycles, bool relative_to_rhs, int npasses,
                                      const std::vector<T>& iterates, int* passes) {
  int pass = 0;
  auto mul = [&](const std::vector<T>& in, std::vector<T>& out) {
    std::fill(out.begin(), out.end(), T(0));
    for (int col = 0; col < n; ++col)
      for (int p = colptr[col]; p < colptr[col + 1]; ++p) out[rowval[p]] += nz[p] * in[col];
  };
  auto floor = [&](const std::vector<T>& in) { return assembled_floor<T>(n, colptr, rowval, nz.data(), in, b, V); };
  auto krylov = [&](std::vector<T>& io, T, int) {
    ++pass;
    const int k = std::min(pass, npasses) - 1;
    std::copy(iterates.begin() + size_t(k) * n, iterates.begin() + size_t(k + 1) * n, io.begin());
    return 1;
  };
  auto gauge = [](std::vector<T>&) {};
  auto st = krylov_driver<T>(V, b, x, rtol, atol, max_cycles, relative_to_rhs, mul, floor, krylov, gauge);
  *passes = pass;
  return st;
}

double csc_floor(bool single, int n, const int* colptr, const int* rowval, const double* nzval,
                 const double* volumes, const double* x, const double* b) {
  auto run = [&](auto zero) {
    using T = decltype(zero);
    std::vector<T> nz(nzval, nzval + colptr[n]), V(volumes, volumes + n), xv(x, x + n), bv(b, b + n);
    return double(assembled_floor<T>(n, colptr, rowval, nz.data(), xv, bv, V));
  };
  return single ? run(0.0f) : run(0.0);
}

Stats scripted_solve(bool single, int n, const int* colptr, const int* rowval, const double* nzval,
                     const double* volumes, const double* b, double* x, double rtol, double atol, int max_cycles,
                     bool relative_to_rhs, int npasses, const double* iterates, int* passes) {
  if (npasses < 1) throw std::invalid_argument("at least one scripted iterate is required");
  auto run = [&](auto zero) {
    using T = decltype(zero);
    const int nnz = colptr[n];
    std::vector<T> nz(nzval, nzval + nnz), V(volumes, volumes + n), bv(b, b + n), xv(x, x + n),
        it(iterates, iterates + size_t(npasses) * n);
    auto d = scripted_driver<T>(n, colptr, rowval, nz, V, bv, xv, T(rtol), T(atol), max_cycles, relative_to_rhs,
                                npasses, it, passes);
    std::copy(xv.begin(), xv.end(), x);
    Stats st{};
    st.cycles = d.cycles;
    st.converged = d.converged;
    st.compatible = 1;
    st.initial_l2 = d.initial_l2;
    st.initial_linf = d.initial_linf;
    st.final_l2 = d.final_l2;
    st.final_linf = d.final_linf;
    st.target_linf = d.target_linf;
    st.residual_scale = d.residual_scale;
    st.removed_rhs_mean = 0;
    return st;
  };
  return single ? run(0.0f) : run(0.0);
}

// `solve!` (elliptic_solver.jl:628-675) with `_solve_krylov_with_stats!`
// (:550-612) and `_run_krylov_reliably!` (:487-526). `rhs` is the strong
// right-hand side; `x` holds the initial guess and receives the solution.
Stats solve(Solver& S, const double* rhs, double* xio, double rtol, double atol, int max_cycles,
            int compat_error, std::vector<double>* removed_means) {
  Operator& op = *S.ops[0];
  const std::vector<double>& V = op.volumes;
  const size_t n = V.size();
  std::vector<double> x(xio, xio + n), b(rhs, rhs + n);
  for (size_t i = 0; i < n; ++i) {
    if (V[i] == 0.0) {
      b[i] = 0;
      x[i] = 0;
    } else {
      b[i] *= V[i];
    }
  }
  double removed = 0;
  if (op.singular) {
    const Components& C = op.comp;
    const size_t K = C.anchors.size();
    // `_project_component_rhs!` (elliptic_solver.jl:150-190)
    const int* slots = C.slots.data();
    std::vector<double> sums = chunked_slot_sums(long(n), K, [&](long c) { return slots[c]; },
                                                 [&](long c) { return b[c]; });
    std::vector<double> scales = chunked_slot_sums(long(n), K, [&](long c) { return slots[c]; },
                                                   [&](long c) { return std::fabs(b[c]); });
    double total_sum = 0, total_vol = 0;
    for (size_t q = 0; q < K; ++q) {
      const double tol = 256 * EPS * std::max(1.0, scales[q]);
      if (compat_error && std::fabs(sums[q]) > tol)
        throw std::invalid_argument("Poisson right-hand side has nonzero volume mean");
      total_sum += sums[q];
      total_vol += C.volumes[q];
    }
    if (removed_means) removed_means->clear();
    for (size_t q = 0; q < K; ++q) {
      sums[q] /= C.volumes[q];
      if (removed_means) removed_means->push_back(sums[q]);
    }
    for (size_t c = 0; c < n; ++c)
      if (C.slots[c]) b[c] -= sums[C.slots[c] - 1] * V[c];
    removed = total_vol == 0 ? 0 : total_sum / total_vol;
    remove_volume_means(op, x);
  }
  auto mul = [&](const std::vector<double>& in, std::vector<double>& out) { factorized_mul(op, in.data(), out.data()); };
  auto floor = [&](const std::vector<double>& in) { return roundoff_floor(op, in, b); };
  auto krylov = [&](std::vector<double>& io, double target, int itmax) { return krylov_cg(S, b, io, target, itmax); };
  auto gauge = [&](std::vector<double>& io) {
    if (op.singular) remove_volume_means(op, io);
  };
  const auto d = krylov_driver<double>(V, b, x, rtol, atol, max_cycles, false, mul, floor, krylov, gauge);
  std::copy(x.begin(), x.end(), xio);
  Stats st{};
  st.cycles = d.cycles;
  st.converged = d.converged;
  st.compatible = 1;
  st.initial_l2 = d.initial_l2;
  st.initial_linf = d.initial_linf;
  st.final_l2 = d.final_l2;
  st.final_linf = d.final_linf;
  st.target_linf = d.target_linf;
  st.residual_scale = d.residual_scale;
  st.removed_rhs_mean = removed;
  return st;
}

// ---------------------------------------------------------------------------
// Two-phase coefficients and jumps, src/two_phase/forces.jl

// `_staggered_liquid_fraction` + `_mixture_density!` + `inv.(rho)`,
// forces.jl:23-36,69-84,131
// `_mixture_boundary_coefficient` (forces.jl:103-111): per boundary face
// inv(rho_l c + rho_g (1 - c)) with c the fraction of its cell (no embedded
// geometry: the conditional fraction is the fraction)
void mixture_boundary_coefficients(const Mesh& M, const double* c, double rho_l, double rho_g, double* kb) {
  if (!(std::isfinite(rho_l) && rho_l > 0 && std::isfinite(rho_g) && rho_g > 0))
    throw std::invalid_argument("phase densities must be finite and positive");
  for (size_t f = 0; f < M.bfaces.size(); ++f) {
    const double cf = c[M.bfaces[f].cell];
    kb[f] = 1.0 / (rho_l * cf + rho_g * (1.0 - cf));
  }
}

void mixture_coefficients(const Mesh& M, const double* c, double rho_l, double rho_g, double* kappa) {
  if (!(std::isfinite(rho_l) && rho_l > 0 && std::isfinite(rho_g) && rho_g > 0))
    throw std::invalid_argument("phase densities must be finite and positive");
  const int nb = M.bs[0] * M.bs[1];
  for (size_t f = 0; f < M.faces.size(); ++f) {
    const Face& fa = M.faces[f];
    const double hm = M.blocks[fa.minus / nb].spacing[fa.axis];
    const double hp = M.blocks[fa.plus / nb].spacing[fa.axis];
    double cf = (hm * c[fa.minus] + hp * c[fa.plus]) / (hm + hp);
    cf = std::min(std::max(cf, 0.0), 1.0);
    const double rho = rho_l * cf + rho_g * (1 - cf);
    kappa[f] = 1.0 / rho;
  }
}

// `_surface_tension_jumps!` with CGFM (forces.jl:312-314,345-378) and given
// face curvatures (NaN marks an unresolved face curvature).
void cgfm_jumps(const Mesh& M, const double* c, const double* curvature, double sigma, int csf,
                double* jumps) {
  if (!(std::isfinite(sigma) && sigma >= 0)) throw std::invalid_argument("sigma must be finite and nonnegative");
  for (size_t f = 0; f < M.faces.size(); ++f) {
    const Face& fa = M.faces[f];
    const double cm = c[fa.minus], cp = c[fa.plus];
    const double jump = csf ? cp - cm : double(cp > 0.5) - double(cm > 0.5);
    if (sigma == 0 || jump == 0) {
      jumps[f] = 0;
      continue;
    }
    const double k = curvature[f];
    if (std::isnan(k)) {
      jumps[f] = 0;
      continue;
    }
    jumps[f] = sigma * k * jump * fa.area;
  }
}

// `_wrapped_delta` (operators/gradient.jl:97-100): Julia's round is
// round-half-even, as rint in the default rounding mode.
static double wrapped_delta(double d, double extent, bool periodic) {
  return periodic ? d - std::rint(d / extent) * extent : d;
}

// `_face_curvature` (forces.jl:211-300): the LHF-first c(1-c) interpolant of
// the two adjacent cell curvatures (NaN curvature = none), else the
// c(1-c)-weighted extension of the cell curvatures within 3 h of the face
// centre (blocks and cells in order); NaN for `nothing`.
double face_curvature(const Mesh& M, const double* c, const double* kappa, const double* quality, int fi) {
  const Face& fa = M.faces[fi];
  const double km = kappa[fa.minus], kp = kappa[fa.plus];
  const double cm = c[fa.minus], cp = c[fa.plus];
  const double wm = std::isfinite(km) ? cm * (1 - cm) : 0.0;
  const double wp = std::isfinite(kp) ? cp * (1 - cp) : 0.0;
  if (wm > 0 || wp > 0) {
    const double qm = wm > 0 ? quality[fa.minus] : 0.0;
    const double qp = wp > 0 ? quality[fa.plus] : 0.0;
    const double pref = std::max(qm, qp);
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
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  const Block& bm = M.blocks[fa.minus / nb];
  const Block& bp = M.blocks[fa.plus / nb];
  const double local = std::max(std::max(bm.spacing[0], bm.spacing[1]), std::max(bp.spacing[0], bp.spacing[1]));
  const double radius = 3 * local;
  const double emax = std::max(M.extent[0], M.extent[1]);
  double total = 0, weight = 0;
  for (size_t b = 0; b < M.blocks.size(); ++b) {
    const Block& B = M.blocks[b];
    const double hs = std::max(B.spacing[0], B.spacing[1]);
    const double dx = std::fabs(wrapped_delta(B.origin[0] + B.extent[0] / 2 - fa.center[0], M.extent[0], M.periodic[0]));
    const double dy = std::fabs(wrapped_delta(B.origin[1] + B.extent[1] / 2 - fa.center[1], M.extent[1], M.periodic[1]));
    const double margin = 16 * 2.220446049250313e-16 *
                          std::max(std::max(std::max(emax, std::fabs(B.origin[0])), std::max(std::fabs(B.origin[1]),
                                   std::fabs(fa.center[0]))), std::fabs(fa.center[1]));
    const double reach = radius + hs + margin;
    if (dx > B.extent[0] / 2 + reach || dy > B.extent[1] / 2 + reach) continue;
    for (int j = 1; j <= by; ++j)
      for (int i = 1; i <= bx; ++i) {
        const size_t cell = b * nb + size_t(j - 1) * bx + (i - 1);
        const double cc = c[cell], k = kappa[cell];
        const double sw = cc * (1 - cc);
        if (!(sw > 0 && std::isfinite(k))) continue;
        const double px = B.origin[0] + (i - 0.5) * B.spacing[0];
        const double py = B.origin[1] + (j - 0.5) * B.spacing[1];
        const double ddx = wrapped_delta(px - fa.center[0], M.extent[0], M.periodic[0]);
        const double ddy = wrapped_delta(py - fa.center[1], M.extent[1], M.periodic[1]);
        const double dist = std::hypot(ddx, ddy);
        if (!(dist <= radius + hs)) continue;
        const double gw = sw / std::max(dist * dist, (hs / 4) * (hs / 4));
        total += gw * k;
        weight += gw;
      }
  }
  return weight > 0 ? total / weight : std::nan("");
}

// `_surface_tension_jumps!` (forces.jl:345-378) with the face curvature of
// `_face_curvature`; returns the number of unresolved faces.
int surface_tension_jumps_cells(const Mesh& M, const double* c, const double* kappa, const double* quality,
                                double sigma, int csf, double* jumps) {
  if (!(std::isfinite(sigma) && sigma >= 0)) throw std::invalid_argument("sigma must be finite and nonnegative");
  int unresolved = 0;
  for (size_t f = 0; f < M.faces.size(); ++f) {
    const Face& fa = M.faces[f];
    const double cm = c[fa.minus], cp = c[fa.plus];
    const double jump = csf ? cp - cm : double(cp > 0.5) - double(cm > 0.5);
    if (sigma == 0 || jump == 0) {
      jumps[f] = 0;
      continue;
    }
    const double k = face_curvature(M, c, kappa, quality, int(f));
    if (std::isnan(k)) {
      ++unresolved;
      jumps[f] = 0;
      continue;
    }
    jumps[f] = sigma * k * jump * fa.area;
  }
  return unresolved;
}

// `gravity!` jumps (forces.jl:528-580): potential g . x at the cell centre
// (`_leaf_cell_center`, mesh.jl:1142-1145), moved by the interface centroid
// offset (in cell spacings; offsets == nullptr: none), then per face
// rho |f| (phi+ - phi-) with rho = (rho_l a_l + rho_g a_g) / open, open = 1.
void gravity_jumps(const Mesh& M, const double* offsets, const double* liquid, const double* gas, double rho_l,
                   double rho_g, double gx, double gy, double* jumps) {
  if (!(std::isfinite(rho_l) && rho_l > 0 && std::isfinite(rho_g) && rho_g > 0))
    throw std::invalid_argument("densities must be finite and positive");
  if (!(std::isfinite(gx) && std::isfinite(gy))) throw std::invalid_argument("gravity must be finite");
  if ((M.periodic[0] && gx != 0) || (M.periodic[1] && gy != 0))
    throw std::invalid_argument("gravity potential requires nonperiodic boundaries along gravity");
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  std::vector<double> phi(M.ncells());
  for (size_t b = 0; b < M.blocks.size(); ++b) {
    const Block& B = M.blocks[b];
    for (int j = 1; j <= by; ++j)
      for (int i = 1; i <= bx; ++i) {
        const size_t cell = b * nb + size_t(j - 1) * bx + (i - 1);
        double x = B.origin[0] + (i - 0.5) * B.spacing[0];
        double y = B.origin[1] + (j - 0.5) * B.spacing[1];
        if (offsets) {
          x += offsets[2 * cell] * B.spacing[0];
          y += offsets[2 * cell + 1] * B.spacing[1];
        }
        phi[cell] = gx * x + gy * y;
      }
  }
  for (size_t f = 0; f < M.faces.size(); ++f) {
    const Face& fa = M.faces[f];
    const double open = 1.0;
    const double rho = (rho_l * liquid[f] + rho_g * gas[f]) / open;
    jumps[f] = rho * fa.area * (phi[fa.plus] - phi[fa.minus]);
  }
}

static void mul_metric(const Operator& op, const std::vector<double>& in, std::vector<double>& out) {
  const int F = int(in.size());
  out.assign(F, 0.0);
  if (op.K_diagonal) {
    for (int f = 0; f < F; ++f) out[f] = op.K.nzval[op.K.colptr[f]] * in[f];
  } else {
    for (int f = 0; f < F; ++f) {
      double s = 0;
      for (int p = op.K.colptr[f]; p < op.K.colptr[f + 1]; ++p) s += op.K.nzval[p] * in[op.K.rowval[p]];
      out[f] = s;
    }
  }
}

// `_weighted_jump_force!` (forces.jl:179-209, no geometry)
void weighted_jump_force(const Operator& op, const double* jumps, double* accel) {
  const int F = int(op.kappa.size());
  std::vector<double> in(F), out;
  for (int f = 0; f < F; ++f) in[f] = op.sqrt_kappa[f] * jumps[f];
  mul_metric(op, in, out);
  for (int f = 0; f < F; ++f) accel[f] = op.sqrt_kappa[f] * out[f];
}

// `weighted_gradient!` (elliptic.jl:1272-1297): interior faces; boundary
// faces of wall/periodic operators are zero.
void weighted_gradient(const Operator& op, const double* p, double* g) {
  const Mesh& M = *op.mesh;
  const int F = int(M.faces.size());
  std::vector<double> in(F), out;
  for (int f = 0; f < F; ++f)
    in[f] = op.sqrt_kappa[f] * (M.faces[f].area * (p[M.faces[f].plus] - p[M.faces[f].minus]));
  mul_metric(op, in, out);
  for (int f = 0; f < F; ++f) g[f] = op.sqrt_kappa[f] * out[f];
}

// `_elliptic_boundary_gradient!` (elliptic.jl:1254-1263): boundary faces
// -side * (2 k_b / h) * p_cell on open sides, zero on closed ones
void weighted_gradient_boundary(const Operator& op, const double* p, double* gb) {
  const Mesh& M = *op.mesh;
  for (size_t f = 0; f < M.bfaces.size(); ++f)
    gb[f] = -double(M.bfaces[f].side) * op.bgrad[f] * p[M.bfaces[f].cell];
}

// `divergence!`, operators/advection/cell_advection.jl:205-220
void divergence(const Mesh& M, const double* u, const double* ub, double* div) {
  const int nb = M.bs[0] * M.bs[1];
  for (int c = 0; c < M.ncells(); ++c) div[c] = 0.0;
  for (size_t f = 0; f < M.faces.size(); ++f) {
    const Face& fa = M.faces[f];
    const double flux = u[f] * fa.area;
    div[fa.minus] += flux / M.blocks[fa.minus / nb].volume;
    div[fa.plus] += -flux / M.blocks[fa.plus / nb].volume;
  }
  if (ub)
    for (size_t f = 0; f < M.bfaces.size(); ++f) {
      const BFace& fa = M.bfaces[f];
      const double flux = fa.side * ub[f] * fa.area;
      div[fa.cell] += flux / M.blocks[fa.cell / nb].volume;
    }
}

// `_project_prepared!` (navier_stokes/projection.jl:221-260) with one
// component: div -> rhs = -div/dt -> solve with rtol 0 and atol = tol/dt^2 ->
// u -= dt G phi -> p += phi. `ub` (NULL: none) are the boundary face
// velocities of an operator with open sides: they enter the divergence and
// receive the boundary gradient's correction.
// With `inherited` (a shortened step, `_projection_rhs!` projection.jl:130-156)
// the rhs keeps (tau - dt) / (dt tau) of the inherited divergence; the stats'
// `compatible` is `_projection_compatibility!` (:100-127).
Stats project(Solver& S, double* u, double* ub, double* p, double dt, double tol, int max_cycles,
              double* phi_out, const double* inherited, double reference_dt) {
  Operator& op = *S.ops[0];
  const Mesh& M = *op.mesh;
  const int N = M.ncells(), F = int(M.faces.size());
  if (!(std::isfinite(reference_dt) && reference_dt >= dt))
    throw std::invalid_argument("reference dt must be finite and at least dt");
  std::vector<double> div(N), phi(N, 0.0), g(F);
  divergence(M, u, ub, div.data());
  if (inherited) {
    const double w = (reference_dt - dt) / (dt * reference_dt);
    for (int c = 0; c < N; ++c) div[c] = -div[c] / dt + w * inherited[c];
  } else {
    for (int c = 0; c < N; ++c) div[c] = -div[c] / dt;
  }
  std::vector<double> means;
  Stats st = solve(S, div.data(), phi.data(), 0.0, tol / (dt * dt), max_cycles, 0, &means);
  st.residual_scale = 1.0 / (dt * dt);
  // `_projection_compatibility!`: the largest removed component mean
  double mean = std::fabs(st.removed_rhs_mean);
  for (double m : means) mean = std::max(mean, std::fabs(m));
  st.compatible = (mean == 0 ? 0.0 : dt * dt * mean) <= tol;
  weighted_gradient(op, phi.data(), g.data());
  const double scale = dt * 1.0;
  for (int f = 0; f < F; ++f) u[f] -= scale * g[f];
  if (ub) {
    std::vector<double> gb(M.bfaces.size());
    weighted_gradient_boundary(op, phi.data(), gb.data());
    for (size_t f = 0; f < gb.size(); ++f) ub[f] -= scale * gb[f];
  }
  for (int c = 0; c < N; ++c) p[c] += phi[c];
  // `_accumulate_pressure_increment!` (projection.jl:52-97): on a singular
  // operator with several components every ungauged component's volume mean
  // of the accumulated pressure is removed (sums in cell order)
  const Components& C = op.comp;
  int ncomponents = 0;
  for (int c = 0; c < N; ++c) ncomponents += C.labels[c] == c ? 1 : 0;
  if (op.singular && ncomponents > 1) {
    std::vector<double> sums(C.anchors.size(), 0.0);
    for (int c = 0; c < N; ++c)
      if (C.slots[c]) sums[C.slots[c] - 1] += op.volumes[c] * p[c];
    for (size_t q = 0; q < sums.size(); ++q) sums[q] /= C.volumes[q];
    for (int c = 0; c < N; ++c) {
      if (C.labels[c] < 0) p[c] = 0.0;
      else if (C.slots[c]) p[c] = p[c] - sums[C.slots[c] - 1];
    }
  }
  if (phi_out) std::copy(phi.begin(), phi.end(), phi_out);
  return st;
}

// ---------------------------------------------------------------------------
// Pressure adjoint (no reference counterpart; pinned by finite differences).
// Forward: p = solve(b, beta) with zero volume mean. J = w' p.
// dJ/db_i = V_i lambda_i where A lambda = Q w (Q w = w - V (1'w)/(1'V)),
// lambda with zero volume mean; dJ/dbeta_f = -[y_l,f (K y_p)_f +
// y_p,f (K y_l)_f] / (2 beta_f), y_x = sqrt(beta) |f| (x+ - x-).
Stats adjoint(Solver& S, const double* w, const double* p, double rtol, int max_cycles,
              double* dJ_db, double* dJ_dbeta, double* lambda_out) {
  Operator& op = *S.ops[0];
  const Mesh& M = *op.mesh;
  const int N = M.ncells(), F = int(M.faces.size());
  // solve() expects a strong RHS: pass w / V so that V * rhs = w
  std::vector<double> rhs(N), lam(N, 0.0);
  for (int c = 0; c < N; ++c) rhs[c] = w[c] / op.volumes[c];
  Stats st = solve(S, rhs.data(), lam.data(), rtol, 0.0, max_cycles, 0, nullptr);
  for (int c = 0; c < N; ++c) dJ_db[c] = op.volumes[c] * lam[c];
  std::vector<double> yl(F), yp(F), Kyl, Kyp;
  for (int f = 0; f < F; ++f) {
    const Face& fa = M.faces[f];
    yl[f] = op.sqrt_kappa[f] * fa.area * (lam[fa.plus] - lam[fa.minus]);
    yp[f] = op.sqrt_kappa[f] * fa.area * (p[fa.plus] - p[fa.minus]);
  }
  mul_metric(op, yl, Kyl);
  mul_metric(op, yp, Kyp);
  for (int f = 0; f < F; ++f)
    dJ_dbeta[f] = op.kappa[f] > 0 ? -(yl[f] * Kyp[f] + yp[f] * Kyl[f]) / (2 * op.kappa[f]) : 0.0;
  if (lambda_out) std::copy(lam.begin(), lam.end(), lambda_out);
  return st;
}

void apply_fine(Solver& S, const double* x, double* y) { factorized_mul(*S.ops[0], x, y); }
double jacobi_weight_public(std::vector<double>& invdiag, const std::vector<double>& rowsum, bool skip_nonpositive,
                            bool allow_zero, bool clamp, double zero_weight) {
  return jacobi_weight(invdiag, rowsum, skip_nonpositive, allow_zero, clamp, zero_weight);
}
void apply_level(Solver& S, int l, const double* x, double* y) { csc_sym_mul(S.levels[l].op->A, x, y); }

}  // namespace orc
