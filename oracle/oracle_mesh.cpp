// Oracle: block-quadtree mesh, face schedule and inverse face metric K.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <parallel/algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

#include "oracle.hpp"

namespace orc {

int Mesh::max_leaf_level() const {
  int m = 0;
  for (const Block& b : blocks) m = std::max(m, b.key.level);
  return m;
}

// RegionTrees child boundaries (origin + (c-1)*widths/2, widths/2), starting
// from the root rectangle of `_build_root_grid`, mesh.jl:738-757.
static void node_geometry(const Mesh& M, const Key& k, double o[2], double w[2]) {
  const int ri = k.i >> k.level, rj = k.j >> k.level;
  o[0] = M.origin[0] + ri * M.root_extent[0];
  o[1] = M.origin[1] + rj * M.root_extent[1];
  if (ri == 0) o[0] = M.origin[0];
  if (rj == 0) o[1] = M.origin[1];
  w[0] = M.root_extent[0];
  w[1] = M.root_extent[1];
  for (int l = 1; l <= k.level; ++l) {
    const int ci = (k.i >> (k.level - l)) & 1, cj = (k.j >> (k.level - l)) & 1;
    w[0] = w[0] / 2;
    w[1] = w[1] / 2;
    o[0] = o[0] + ci * w[0];
    o[1] = o[1] + cj * w[1];
  }
}

// `_balance!`, mesh.jl:673-714. The reference splits the coarser leaf of any
// face- or corner-touching pair whose levels differ by two or more, until no
// pair is left. Every such split is forced in any balanced refinement, so the
// fixed point is the unique minimal 2:1 refinement; this restatement finds it
// with whole passes over the leaves instead of restarting after each split.
static void balance(Mesh& M) {
  for (;;) {
    std::vector<Key> leaves;
    // collect leaves
    std::vector<Key> stack;
    for (int j = 0; j < M.layout[1]; ++j)
      for (int i = 0; i < M.layout[0]; ++i) stack.push_back({0, i, j});
    while (!stack.empty()) {
      Key k = stack.back();
      stack.pop_back();
      if (M.internal.count(kpack(k.level, k.i, k.j))) {
        for (int c = 0; c < 4; ++c)
          stack.push_back({k.level + 1, 2 * k.i + (c & 1), 2 * k.j + (c >> 1)});
      } else {
        leaves.push_back(k);
      }
    }
    std::vector<Key> split;
    for (const Key& L : leaves) {
      const int l1 = L.level + 1;
      const long px = long(M.layout[0]) << l1, py = long(M.layout[1]) << l1;
      bool need = false;
      for (int cx = 2 * L.i - 1; cx <= 2 * L.i + 2 && !need; ++cx) {
        for (int cy = 2 * L.j - 1; cy <= 2 * L.j + 2 && !need; ++cy) {
          const bool inside = cx >= 2 * L.i && cx <= 2 * L.i + 1 &&
                              cy >= 2 * L.j && cy <= 2 * L.j + 1;
          if (inside) continue;
          long x = cx, y = cy;
          if (x < 0 || x >= px) {
            if (!M.periodic[0]) continue;
            x = ((x % px) + px) % px;
          }
          if (y < 0 || y >= py) {
            if (!M.periodic[1]) continue;
            y = ((y % py) + py) % py;
          }
          if (M.internal.count(kpack(l1, int(x), int(y)))) need = true;
        }
      }
      if (need) split.push_back(L);
    }
    if (split.empty()) return;
    for (const Key& k : split) M.internal.insert(kpack(k.level, k.i, k.j));
  }
}

struct RawSide {
  long lo, hi;
  int cell;
};

// `_lattice_sides` + `_match_sides!` + `_push_boundary!`, mesh.jl:783-927.
static void match_axis(Mesh& M, int axis, int maxl) {
  const long lattice[2] = {long(M.base[0]) << maxl, long(M.base[1]) << maxl};
  const long lines = lattice[axis];
  const int nx = M.bs[0], ny = M.bs[1];
  std::vector<std::vector<RawSide>> lower(lines + 1), upper(lines + 1);
  int cell = 0;
  for (const Block& b : M.blocks) {
    const long scale = 1L << (maxl - b.key.level);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++cell) {
        const long xlo = long(b.key.i) * nx * scale + i * scale;
        const long ylo = long(b.key.j) * ny * scale + j * scale;
        long low, high, tlo, thi;
        if (axis == 0) {
          low = xlo; high = xlo + scale; tlo = ylo; thi = ylo + scale;
        } else {
          low = ylo; high = ylo + scale; tlo = xlo; thi = xlo + scale;
        }
        const long hc = M.periodic[axis] ? ((high % lines) + lines) % lines : high;
        const long lc = M.periodic[axis] ? ((low % lines) + lines) % lines : low;
        lower[hc].push_back({tlo, thi, cell});
        upper[lc].push_back({tlo, thi, cell});
      }
  }
  const int t = 1 - axis;
  const double normal_scale = M.extent[axis] / double(lattice[axis]);
  const double tangent_scale = M.extent[t] / double(lattice[t]);
  auto by_lo = [](const RawSide& a, const RawSide& b) { return a.lo < b.lo; };
  for (long line = 0; line <= lines; ++line) {
    auto& lw = lower[line];
    auto& up = upper[line];
    std::stable_sort(lw.begin(), lw.end(), by_lo);
    std::stable_sort(up.begin(), up.end(), by_lo);
    size_t li = 0, ui = 0;
    while (li < lw.size() && ui < up.size()) {
      const RawSide& ls = lw[li];
      const RawSide& us = up[ui];
      const long lo = std::max(ls.lo, us.lo), hi = std::min(ls.hi, us.hi);
      if (lo < hi) {
        Face f;
        f.minus = ls.cell;
        f.plus = us.cell;
        f.axis = axis;
        const double normal = M.origin[axis] + line * normal_scale;
        const double tangent = M.origin[t] + (lo + hi) * tangent_scale / 2;
        f.center[axis] = normal;
        f.center[t] = tangent;
        f.area = (hi - lo) * tangent_scale;
        M.faces.push_back(f);
      }
      const bool adv_l = ls.hi <= us.hi, adv_u = us.hi <= ls.hi;
      if (adv_l) ++li;
      if (adv_u) ++ui;
    }
    auto push_boundary = [&](const RawSide& s, int side) {
      BFace f;
      f.cell = s.cell;
      f.axis = axis;
      f.side = side;
      f.center[axis] = M.origin[axis] + line * M.extent[axis] / double(lattice[axis]);
      f.center[t] = M.origin[t] + (s.lo + s.hi) * M.extent[t] / (2.0 * double(lattice[t]));
      f.area = (s.hi - s.lo) * M.extent[t] / double(lattice[t]);
      M.bfaces.push_back(f);
    };
    for (; li < lw.size(); ++li) push_boundary(lw[li], 1);
    for (; ui < up.size(); ++ui) push_boundary(up[ui], -1);
  }
}

// BlockTreeMesh constructor, mesh.jl:1059-1121 (tree from the refined set,
// `_balance!`, `_leaf_blocks` sorted by (level, j, i), face schedules).
std::unique_ptr<Mesh> build_mesh(const int base[2], const int bs[2],
                                 const double origin[2], const double extent[2],
                                 const bool periodic[2], int max_level,
                                 const std::vector<Key>& refined) {
  auto M = std::make_unique<Mesh>();
  for (int a = 0; a < 2; ++a) {
    if (bs[a] < 2 || bs[a] % 2) throw std::invalid_argument("block dimensions must be even and at least two");
    if (base[a] < bs[a] || base[a] % bs[a]) throw std::invalid_argument("base mesh must be a multiple of the block size");
    if (!(extent[a] > 0)) throw std::invalid_argument("domain extents must be positive");
    M->base[a] = base[a];
    M->bs[a] = bs[a];
    M->origin[a] = origin[a];
    M->extent[a] = extent[a];
    M->periodic[a] = periodic[a];
    M->layout[a] = base[a] / bs[a];
    M->root_extent[a] = extent[a] / M->layout[a];
  }
  M->max_level = max_level;
  // `_build_tree`: refine requests below the ceiling; only nodes reachable
  // from a root through refined ancestors exist.
  std::unordered_set<uint64_t> req;
  for (const Key& k : refined)
    if (k.level < max_level) req.insert(kpack(k.level, k.i, k.j));
  std::vector<Key> stack;
  for (int j = 0; j < M->layout[1]; ++j)
    for (int i = 0; i < M->layout[0]; ++i) stack.push_back({0, i, j});
  while (!stack.empty()) {
    Key k = stack.back();
    stack.pop_back();
    if (k.level >= max_level || !req.count(kpack(k.level, k.i, k.j))) continue;
    M->internal.insert(kpack(k.level, k.i, k.j));
    for (int c = 0; c < 4; ++c)
      stack.push_back({k.level + 1, 2 * k.i + (c & 1), 2 * k.j + (c >> 1)});
  }
  balance(*M);
  // leaves
  std::vector<Key> leaves;
  for (int j = 0; j < M->layout[1]; ++j)
    for (int i = 0; i < M->layout[0]; ++i) stack.push_back({0, i, j});
  while (!stack.empty()) {
    Key k = stack.back();
    stack.pop_back();
    if (M->internal.count(kpack(k.level, k.i, k.j))) {
      for (int c = 0; c < 4; ++c)
        stack.push_back({k.level + 1, 2 * k.i + (c & 1), 2 * k.j + (c >> 1)});
    } else {
      leaves.push_back(k);
    }
  }
  std::sort(leaves.begin(), leaves.end(), [](const Key& a, const Key& b) {
    return std::make_tuple(a.level, a.j, a.i) < std::make_tuple(b.level, b.j, b.i);
  });
  for (const Key& k : leaves) {
    Block b;
    b.key = k;
    node_geometry(*M, k, b.origin, b.extent);
    b.spacing[0] = b.extent[0] / bs[0];
    b.spacing[1] = b.extent[1] / bs[1];
    b.volume = b.spacing[0] * b.spacing[1];
    M->blocks.push_back(b);
  }
  const int maxl = M->max_leaf_level();
  match_axis(*M, 0, maxl);
  match_axis(*M, 1, maxl);
  return M;
}

// ---------------------------------------------------------------------------
// Inverse face metric K, src/operators/gradient.jl:107-408.

static void cell_lattice_bounds(const Mesh& M, int cell, const long lat[2],
                                long low[2], long scale[2]) {
  const int nb = M.bs[0] * M.bs[1];
  const int b = cell / nb, r = cell % nb;
  const int i = r % M.bs[0], j = r / M.bs[0];
  const Key& k = M.blocks[b].key;
  scale[0] = lat[0] / (long(M.base[0]) << k.level);
  scale[1] = lat[1] / (long(M.base[1]) << k.level);
  low[0] = (long(k.i) * M.bs[0] + i) * scale[0];
  low[1] = (long(k.j) * M.bs[1] + j) * scale[1];
}

// `_face_lattice_geometry`, mesh.jl:463-476
static void face_lattice(const Mesh& M, const Face& f, const long lat[2], long& normal,
                         long& lo, long& hi) {
  long ml[2], ms[2], pl[2], ps[2];
  cell_lattice_bounds(M, f.minus, lat, ml, ms);
  cell_lattice_bounds(M, f.plus, lat, pl, ps);
  const int a = f.axis, t = 1 - a;
  normal = ml[a] + ms[a];
  if (M.periodic[a]) normal = ((normal % lat[a]) + lat[a]) % lat[a];
  lo = std::max(ml[t], pl[t]);
  hi = std::min(ml[t] + ms[t], pl[t] + ps[t]);
}

struct MFace {
  int index, axis;
  double area, d[2];
};

// `_face_displacement` + `_tangential_centre_offset`, gradient.jl:107-158
static void face_displacement(const Mesh& M, const Face& f, double d[2]) {
  const int nb = M.bs[0] * M.bs[1];
  const int a = f.axis, t = 1 - a;
  const double normal =
      (M.blocks[f.minus / nb].spacing[a] + M.blocks[f.plus / nb].spacing[a]) / 2;
  const long lat[2] = {long(M.base[0]) << M.max_level, long(M.base[1]) << M.max_level};
  long ml[2], ms[2], pl[2], ps[2];
  cell_lattice_bounds(M, f.minus, lat, ml, ms);
  cell_lattice_bounds(M, f.plus, lat, pl, ps);
  long offset = 2 * pl[t] + ps[t] - 2 * ml[t] - ms[t];
  if (M.periodic[t]) {
    const long period = 2 * lat[t];
    if (offset > lat[t]) offset -= period;
    if (offset < -lat[t]) offset += period;
  }
  const double spacing_t = M.extent[t] / double(lat[t]);
  const double tangential = double(offset) * (spacing_t / 2);
  d[a] = normal;
  d[t] = tangential;
}

// 2x2 solve with partial pivoting (the dense `\` of gradient.jl:219,222).
static void solve2(const double A[2][2], double* rhs, int nrhs) {
  // rhs is 2 x nrhs, row-major
  double a00 = A[0][0], a01 = A[0][1], a10 = A[1][0], a11 = A[1][1];
  bool swap = std::fabs(a10) > std::fabs(a00);
  if (swap) {
    std::swap(a00, a10);
    std::swap(a01, a11);
    for (int c = 0; c < nrhs; ++c) std::swap(rhs[c], rhs[nrhs + c]);
  }
  const double l = a10 / a00;
  const double u11 = a11 - l * a01;
  for (int c = 0; c < nrhs; ++c) {
    const double y1 = rhs[nrhs + c] - l * rhs[c];
    const double x1 = y1 / u11;
    const double x0 = (rhs[c] - a01 * x1) / a00;
    rhs[c] = x0;
    rhs[nrhs + c] = x1;
  }
}

// `_local_inverse_metric`, gradient.jl:194-232
static std::vector<double> local_inverse_metric(const std::vector<MFace>& F) {
  const int n = int(F.size());
  if (n <= 2) throw std::invalid_argument("a periodic vertex patch must contain at least three faces");
  std::vector<double> Mo(n * 2), No(n * 2, 0.0);
  for (int r = 0; r < n; ++r) {
    Mo[r * 2 + 0] = F[r].area * F[r].d[0];
    Mo[r * 2 + 1] = F[r].area * F[r].d[1];
    No[r * 2 + F[r].axis] = 0.5;
  }
  double MN[2][2] = {{0, 0}, {0, 0}};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      double s = 0;
      for (int r = 0; r < n; ++r) s += Mo[r * 2 + a] * No[r * 2 + b];
      MN[a][b] = s;
    }
  double S[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) S[a][b] = (MN[a][b] + MN[b][a]) / 2;
  // isposdef via Cholesky
  if (!(S[0][0] > 0) || !(S[1][1] - S[0][1] * S[0][1] / S[0][0] > 0))
    throw std::invalid_argument("degenerate vertex patch in gradient construction");
  // consistent = N * (S \ N')
  std::vector<double> X(2 * n);
  for (int r = 0; r < n; ++r) {
    X[0 * n + r] = No[r * 2 + 0];
    X[1 * n + r] = No[r * 2 + 1];
  }
  solve2(S, X.data(), n);
  std::vector<double> C(n * n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) C[r * n + c] = No[r * 2 + 0] * X[c] + No[r * 2 + 1] * X[n + c];
  double G[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      double s = 0;
      for (int r = 0; r < n; ++r) s += Mo[r * 2 + a] * Mo[r * 2 + b];
      G[a][b] = s;
    }
  std::vector<double> Y(2 * n);
  for (int r = 0; r < n; ++r) {
    Y[r] = Mo[r * 2 + 0];
    Y[n + r] = Mo[r * 2 + 1];
  }
  solve2(G, Y.data(), n);
  double trace = 0;
  for (int r = 0; r < n; ++r) trace += C[r * n + r];
  const double stab = trace / (n - 2);
  std::vector<double> L(n * n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      const double proj = (r == c ? 1.0 : 0.0) - (Mo[r * 2 + 0] * Y[c] + Mo[r * 2 + 1] * Y[n + c]);
      L[r * n + c] = C[r * n + c] + stab * proj;
    }
  std::vector<double> out(n * n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) out[r * n + c] = (L[r * n + c] + L[c * n + r]) / 2;
  return out;
}

using Descriptor = std::tuple<int, double, double, double>;

// `_assemble_inverse_metric` + `_cached_local_inverse_metric!` +
// `_append_boundary_patch_metric!` + `_metric_storage`, gradient.jl:241-393
CSC inverse_metric(const Mesh& M, bool* diagonal) {
  const int F = int(M.faces.size());
  const long lat[2] = {long(M.base[0]) << M.max_level, long(M.base[1]) << M.max_level};
  std::vector<MFace> mf(F);
  struct EP {
    long k1, k2;
    int ep;
  };
  std::vector<EP> eps(2 * F);
  for (int idx = 0; idx < F; ++idx) {
    const Face& f = M.faces[idx];
    double d[2];
    face_displacement(M, f, d);
    mf[idx] = {idx, f.axis, f.area, {d[0], d[1]}};
    long normal, lo, hi;
    face_lattice(M, f, lat, normal, lo, hi);
    for (int e = 0; e < 2; ++e) {
      const long tangent = e == 0 ? lo : hi;
      long i = f.axis == 0 ? normal : tangent;
      long j = f.axis == 0 ? tangent : normal;
      if (M.periodic[0]) i = ((i % lat[0]) + lat[0]) % lat[0];
      if (M.periodic[1]) j = ((j % lat[1]) + lat[1]) % lat[1];
      eps[2 * idx + e] = {i, j, 2 * idx + e};
    }
  }
  __gnu_parallel::stable_sort(eps.begin(), eps.end(), [](const EP& a, const EP& b) {
    return a.k1 != b.k1 ? a.k1 < b.k1 : a.k2 < b.k2;
  });
  std::vector<int> rows, cols;
  std::vector<double> vals;
  std::map<std::vector<Descriptor>, std::vector<double>> cache;
  size_t pos = 0;
  while (pos < eps.size()) {
    const long k1 = eps[pos].k1, k2 = eps[pos].k2;
    std::vector<MFace> faces;
    while (pos < eps.size() && eps[pos].k1 == k1 && eps[pos].k2 == k2) {
      faces.push_back(mf[eps[pos].ep >> 1]);
      ++pos;
    }
    const bool physical = (!M.periodic[0] && (k1 == 0 || k1 == lat[0])) ||
                          (!M.periodic[1] && (k2 == 0 || k2 == lat[1]));
    if (physical || faces.size() <= 2) {
      for (const MFace& f : faces) {
        const double nd = std::fabs(f.d[f.axis]);
        if (!(nd > 0)) throw std::invalid_argument("degenerate face beside a physical boundary");
        rows.push_back(f.index);
        cols.push_back(f.index);
        vals.push_back(1.0 / (2.0 * f.area * nd));
      }
      continue;
    }
    double scale = 0;
    for (const MFace& f : faces)
      scale = std::max(scale, std::max(f.area, std::max(std::fabs(f.d[0]), std::fabs(f.d[1]))));
    if (!(scale > 0)) throw std::invalid_argument("degenerate vertex-patch scale");
    auto desc = [scale](const MFace& f) {
      return Descriptor(f.axis, f.area / scale, f.d[0] / scale, f.d[1] / scale);
    };
    std::stable_sort(faces.begin(), faces.end(),
                     [&](const MFace& a, const MFace& b) { return desc(a) < desc(b); });
    std::vector<Descriptor> key;
    for (const MFace& f : faces) key.push_back(desc(f));
    auto it = cache.find(key);
    std::vector<double> normalized;
    if (it == cache.end()) {
      normalized = local_inverse_metric(faces);
      for (double& v : normalized) v = v * (scale * scale);
      cache.emplace(key, normalized);
    } else {
      normalized = it->second;
    }
    const int n = int(faces.size());
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < n; ++r) {
        rows.push_back(faces[r].index);
        cols.push_back(faces[c].index);
        vals.push_back(normalized[r * n + c] / (scale * scale));
      }
  }
  // sparse(rows, cols, vals): duplicates summed in input order; dropzeros!
  std::vector<int> order(rows.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = int(q);
  __gnu_parallel::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return cols[a] != cols[b] ? cols[a] < cols[b] : rows[a] < rows[b];
  });
  CSC K;
  K.m = K.n = F;
  K.colptr.assign(F + 1, 0);
  size_t q = 0;
  for (int c = 0; c < F; ++c) {
    while (q < order.size() && cols[order[q]] == c) {
      const int r = rows[order[q]];
      double s = vals[order[q]];
      ++q;
      while (q < order.size() && cols[order[q]] == c && rows[order[q]] == r) {
        s += vals[order[q]];
        ++q;
      }
      if (s != 0.0) {
        K.rowval.push_back(r);
        K.nzval.push_back(s);
      }
    }
    K.colptr[c + 1] = int(K.rowval.size());
  }
  bool diag = true;
  for (int c = 0; c < F && diag; ++c)
    for (int p = K.colptr[c]; p < K.colptr[c + 1]; ++p)
      if (K.rowval[p] != c) {
        diag = false;
        break;
      }
  if (diag) {
    for (int c = 0; c < F; ++c) {
      bool found = false;
      for (int p = K.colptr[c]; p < K.colptr[c + 1]; ++p) found |= K.nzval[p] > 0;
      if (!found) throw std::invalid_argument("gradient Riesz map has a non-positive diagonal");
    }
  }
  *diagonal = diag;
  return K;
}

// Halo refresh: `_build_halo_schedule` (src/mesh/topology/mesh.jl:944-1037),
// `_cellref_at_lattice` (mesh.jl:534-541, leaf lookup by tree descent), and
// `fill_halo!` (src/mesh/variables.jl:205-229) with `ZeroGradient` or, when
// `physical_values` is given, one prescribed value per physical entry in
// schedule order (the values a `Dirichlet` callback returns). Output: per
// block the haloed (bx + 2h) x (by + 2h) array with the x index fastest
// (`blockdata` of a CellField). Returns the number of physical entries.
int fill_halo(const Mesh& M, int h, const double* cells, const double* physical_values, double* out) {
  const int nx = M.bs[0], ny = M.bs[1], nb = nx * ny;
  const int W = nx + 2 * h, H = ny + 2 * h;
  const int L = M.max_leaf_level();
  std::unordered_map<uint64_t, int> index;
  for (size_t b = 0; b < M.blocks.size(); ++b) index[kpack(M.blocks[b].key.level, M.blocks[b].key.i, M.blocks[b].key.j)] = int(b);
  const long lattice[2] = {long(M.base[0]) << L, long(M.base[1]) << L};
  const long quarter[2] = {4 * lattice[0], 4 * lattice[1]};
  // leaf cell containing lattice cell p (finest-level units): descend from the root
  auto cell_at = [&](long px, long py) {
    for (int l = 0; l <= L; ++l) {
      const long cx = px >> (L - l), cy = py >> (L - l);
      auto it = index.find(kpack(l, int(cx / nx), int(cy / ny)));
      if (it != index.end()) return it->second * nb + int(cy % ny) * nx + int(cx % nx);
    }
    throw std::logic_error("halo sample outside every leaf");
  };
  auto wrap = [](long x, long period, bool periodic, bool& outside) {
    if (periodic) {
      outside = false;
      return ((x % period) + period) % period;
    }
    outside = x < 0 || x >= period;
    return std::min(std::max(x, 0L), period - 1);
  };
  int nphys = 0;
  for (size_t b = 0; b < M.blocks.size(); ++b) {
    const Block& B = M.blocks[b];
    double* o = out + b * size_t(W) * H;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) o[(i + h) + (j + h) * W] = cells[b * nb + j * nx + i];
    const long scale = 1L << (L - B.key.level);
    for (int j = 1 - h; j <= ny + h; ++j)
      for (int i = 1 - h; i <= nx + h; ++i) {
        if (i >= 1 && i <= nx && j >= 1 && j <= ny) continue;
        const long xb = 4 * (long(B.key.i) * nx + i - 1) * scale;
        const long yb = 4 * (long(B.key.j) * ny + j - 1) * scale;
        bool ox, oy;
        const long cx = wrap(xb + 2 * scale, quarter[0], M.periodic[0], ox);
        const long cy = wrap(yb + 2 * scale, quarter[1], M.periodic[1], oy);
        double v;
        if (ox || oy) {
          v = physical_values ? physical_values[nphys] : cells[cell_at(cx / 4, cy / 4)];
          ++nphys;
        } else {
          int donor[4];
          for (int smp = 0; smp < 4; ++smp) {
            const int qx = (smp % 2 == 0) ? 1 : 3, qy = smp < 2 ? 1 : 3;
            bool dummy;
            const long x = wrap(xb + qx * scale, quarter[0], M.periodic[0], dummy);
            const long y = wrap(yb + qy * scale, quarter[1], M.periodic[1], dummy);
            donor[smp] = cell_at(x / 4, y / 4);
          }
          v = 0.0;
          for (int d = 0; d < 4; ++d) {
            bool seen = false;
            for (int e = 0; e < d; ++e) seen |= donor[e] == donor[d];
            if (seen) continue;
            int mult = 0;
            for (int e = 0; e < 4; ++e) mult += donor[e] == donor[d];
            v += (double(mult) / 4) * cells[donor[d]];
          }
        }
        o[(i - 1 + h) + (j - 1 + h) * W] = v;
      }
  }
  return nphys;
}

}  // namespace orc
