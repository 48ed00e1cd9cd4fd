// Host-side topology and hierarchy construction (see hier.hpp).
//
// Reference behaviour followed here (TwoPhaseAMR.jl):
//   tree, 2:1 balance, leaf order, faces   src/mesh/topology/mesh.jl:673-927, 1059-1121
//   inverse face metric K                  src/operators/gradient.jl:107-408
//   integrated stiffness J' sqrt(k) K sqrt(k) J   src/operators/elliptic.jl:175-257
//   Jacobi weights                         src/solvers/multigrid/hierarchy.jl:8-30
//   coarse meshes, P_t, kappa restriction  src/solvers/multigrid/tree_coarsening.jl:30-82, 187-252, 347-398
//   smoothed transfer (applied matrix-free on the device)  hierarchy.jl:631-721
//   augmented coarse system                geometric_preconditioner.jl:144-169
#include "hier.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>

namespace tpamr {

namespace {

// TPAMR_TRACE=1: wall time of each setup phase to stderr.
struct PhaseTimer {
  bool on;
  std::chrono::steady_clock::time_point t;
  PhaseTimer() {
    const char* q = std::getenv("TPAMR_TRACE");
    on = q && std::strcmp(q, "0") != 0;
    t = std::chrono::steady_clock::now();
  }
  void mark(const char* what, long n = -1) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "tpamr setup: %-28s %9.1f ms", what, std::chrono::duration<double, std::milli>(now - t).count());
    if (n >= 0) std::fprintf(stderr, "  (%ld)", n);
    std::fprintf(stderr, "\n");
    t = now;
  }
};

inline long wrapl(long x, long p) { return ((x % p) + p) % p; }

// Host threads of the setup passes: the hardware threads, at most 16, or
// TPAMR_HOST_THREADS (1 to 64). Every pass splits its work by this count
// alone, and its results do not depend on it.
inline long host_threads() {
  static const long n = [] {
    if (const char* q = std::getenv("TPAMR_HOST_THREADS")) {
      const long v = std::atol(q);
      if (v >= 1) return std::min<long>(v, 64);
    }
    return std::max<long>(1, std::min<long>(16, std::thread::hardware_concurrency()));
  }();
  return n;
}

// Workers for n items: up to host_threads(), at least `grain` items each.
inline int worker_count(long n, long grain = 1 << 14) {
  return int(std::max<long>(1, std::min<long>(host_threads(), (n + grain - 1) / grain)));
}

// Run f(lo, hi, t) for t < nt over the contiguous chunks [n t / nt, n (t + 1)
// / nt); the split depends on n and nt only, so two passes with the same nt
// see the same chunks.
template <typename F>
void run_chunks(long n, int nt, F&& f) {
  if (nt <= 1) {
    f(0L, n, 0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; ++t) pool.emplace_back([&, t] { f(n * t / nt, n * (t + 1) / nt, t); });
  for (auto& th : pool) th.join();
}

template <typename F>
void parallel_for(long n, F&& f, long grain = 1 << 14) {
  run_chunks(n, worker_count(n, grain), f);
}

// Run f(k) for every task k < n on up to host_threads() threads, taking tasks
// in turn.
template <typename F>
void parallel_tasks(long n, F&& f) {
  const int nt = worker_count(n, 1);
  if (nt <= 1) {
    for (long k = 0; k < n; ++k) f(k);
    return;
  }
  std::atomic<long> next{0};
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      for (long k; (k = next.fetch_add(1)) < n;) f(k);
    });
  for (auto& th : pool) th.join();
}

// Sort v by the strict total order `less` in parallel: bucket(x) in [0, nbk)
// must not decrease along the order, so the elements are scattered to their
// buckets (per-thread counts keep the scatter deterministic) and every bucket
// is sorted on its own. The result is the one std::sort gives.
template <typename V, typename B, typename C>
void bucket_sort(V& v, int nbk, B bucket, C less) {
  const long n = long(v.size());
  const int nt = worker_count(n, 1 << 16);
  if (nt <= 1) {
    std::sort(v.begin(), v.end(), less);
    return;
  }
  std::vector<std::vector<long>> h(nt, std::vector<long>(nbk, 0));
  run_chunks(n, nt, [&](long a, long b, int w) {
    for (long q = a; q < b; ++q) h[w][bucket(v[q])]++;
  });
  std::vector<long> start(nbk + 1, 0);
  for (int k = 0; k < nbk; ++k) {
    long at = start[k];
    for (int w = 0; w < nt; ++w) {
      const long c = h[w][k];
      h[w][k] = at;
      at += c;
    }
    start[k + 1] = at;
  }
  hvec<typename V::value_type> out(n);
  run_chunks(n, nt, [&](long a, long b, int w) {
    for (long q = a; q < b; ++q) out[h[w][bucket(v[q])]++] = v[q];
  });
  parallel_tasks(nbk, [&](long k) { std::sort(out.begin() + start[k], out.begin() + start[k + 1], less); });
  run_chunks(n, nt, [&](long a, long b, int) { std::copy(out.begin() + a, out.begin() + b, v.begin() + a); });
}

void collect_leaves(const TreeMesh& M, std::vector<BlockKey>& leaves) {
  leaves.clear();
  std::vector<BlockKey> stack;
  for (int j = 0; j < M.layout[1]; ++j)
    for (int i = 0; i < M.layout[0]; ++i) stack.push_back({0, i, j});
  while (!stack.empty()) {
    BlockKey k = stack.back();
    stack.pop_back();
    if (M.internal.count(key64(k.level, k.i, k.j))) {
      for (int c = 0; c < 4; ++c) stack.push_back({k.level + 1, 2 * k.i + (c & 1), 2 * k.j + (c >> 1)});
    } else {
      leaves.push_back(k);
    }
  }
}

// Minimal 2:1 (face and corner) balanced refinement of the tree: a leaf is
// split whenever a level-(l+1) node touching it is itself refined.
void balance(TreeMesh& M) {
  std::vector<BlockKey> leaves, split;
  for (;;) {
    collect_leaves(M, leaves);
    split.clear();
    for (const BlockKey& L : leaves) {
      const int l1 = L.level + 1;
      const long px = long(M.layout[0]) << l1, py = long(M.layout[1]) << l1;
      bool need = false;
      for (int dy = -1; dy <= 2 && !need; ++dy)
        for (int dx = -1; dx <= 2 && !need; ++dx) {
          if ((dx == 0 || dx == 1) && (dy == 0 || dy == 1)) continue;
          long x = 2L * L.i + dx, y = 2L * L.j + dy;
          if (x < 0 || x >= px) {
            if (!M.periodic[0]) continue;
            x = wrapl(x, px);
          }
          if (y < 0 || y >= py) {
            if (!M.periodic[1]) continue;
            y = wrapl(y, py);
          }
          need = M.internal.count(key64(l1, int(x), int(y))) != 0;
        }
      if (need) split.push_back(L);
    }
    if (split.empty()) return;
    for (const BlockKey& k : split) M.internal.insert(key64(k.level, k.i, k.j));
  }
}

void block_geometry(const TreeMesh& M, const BlockKey& k, double* spacing) {
  double w0 = M.root_extent[0], w1 = M.root_extent[1];
  for (int l = 0; l < k.level; ++l) {
    w0 /= 2;
    w1 /= 2;
  }
  spacing[0] = w0 / M.bs[0];
  spacing[1] = w1 / M.bs[1];
}

// Lower-left corner of block k: the root rectangle's corner
// (`_build_root_grid`, src/mesh/topology/mesh.jl:736-757), then per level the
// child rule corner + c * width / 2, as `_leaf_blocks` (mesh.jl:759-773) reads
// it from the tree nodes.
void block_origin(const TreeMesh& M, const BlockKey& k, double* origin) {
  const int ri = k.i >> k.level, rj = k.j >> k.level;
  double o0 = ri == 0 ? M.origin[0] : M.origin[0] + ri * M.root_extent[0];
  double o1 = rj == 0 ? M.origin[1] : M.origin[1] + rj * M.root_extent[1];
  double w0 = M.root_extent[0], w1 = M.root_extent[1];
  for (int l = 1; l <= k.level; ++l) {
    w0 /= 2;
    w1 /= 2;
    o0 += ((k.i >> (k.level - l)) & 1) * w0;
    o1 += ((k.j >> (k.level - l)) & 1) * w1;
  }
  origin[0] = o0;
  origin[1] = o1;
}

// Face schedule in the reference order: axis-major, then lattice line, then
// increasing tangential position; boundary faces per line after the
// interior faces of that line. Per axis, every cell is bucketed by the line
// of its lower and of its upper side (per-thread histograms, so the buckets
// fill in parallel in cell order); each line's two buckets are sorted by
// tangential position and merged, once to count the faces of every line and
// once to write them at the line's offset.
void build_faces(TreeMesh& M) {
  const int maxl = M.max_leaf;
  const long lat[2] = {long(M.base[0]) << maxl, long(M.base[1]) << maxl};
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  const int nblk = int(M.blocks.size());
  const long ncell = long(nblk) * nb;
  struct Side {
    long lo, hi;
    int cell;
  };
  hvec<long> xlo(ncell), ylo(ncell), sc(ncell);
  parallel_for(nblk, [&](long b0, long b1, int) {
    for (long b = b0; b < b1; ++b) {
      const long s = 1L << (maxl - M.blocks[b].level);
      for (int j = 0; j < by; ++j)
        for (int i = 0; i < bx; ++i) {
          const long c = b * nb + j * bx + i;
          xlo[c] = (long(M.blocks[b].i) * bx + i) * s;
          ylo[c] = (long(M.blocks[b].j) * by + j) * s;
          sc[c] = s;
        }
    }
  }, 256);
  hvec<Side> lower(ncell), upper(ncell);
  // a line's merge yields at most |lower| + |upper| - 1 faces and leaves at
  // most that many boundary pieces: 2 ncell of each per axis
  for (auto* v : {&M.fminus, &M.fplus}) v->reserve(M.fminus.size() + 4 * size_t(ncell));
  M.faxis.reserve(M.faxis.size() + 4 * size_t(ncell));
  for (auto* v : {&M.farea, &M.fcx, &M.fcy}) v->reserve(v->size() + 4 * size_t(ncell));
  M.bcell.reserve(M.bcell.size() + 4 * size_t(ncell));
  for (auto* v : {&M.baxis, &M.bside}) v->reserve(v->size() + 4 * size_t(ncell));
  M.barea.reserve(M.barea.size() + 4 * size_t(ncell));
  for (int axis = 0; axis < 2; ++axis) {
    const int t = 1 - axis;
    const long lines = lat[axis];
    const hvec<long>& low = axis == 0 ? xlo : ylo;
    const hvec<long>& tlo = axis == 0 ? ylo : xlo;
    auto canon = [&](long v) { return M.periodic[axis] ? wrapl(v, lines) : v; };
    // per-thread bucket counts, then each thread's start in every bucket
    const int nt = worker_count(ncell);
    std::vector<std::vector<long>> lh(nt, std::vector<long>(lines + 1, 0)), uh(nt, std::vector<long>(lines + 1, 0));
    run_chunks(ncell, nt, [&](long c0, long c1, int w) {
      for (long c = c0; c < c1; ++c) {
        lh[w][canon(low[c] + sc[c])]++;
        uh[w][canon(low[c])]++;
      }
    });
    std::vector<long> lcount(lines + 2, 0), ucount(lines + 2, 0);
    for (long q = 0; q <= lines; ++q) {
      long ls = lcount[q], us = ucount[q];
      for (int w = 0; w < nt; ++w) {
        const long a = lh[w][q], b = uh[w][q];
        lh[w][q] = ls;
        uh[w][q] = us;
        ls += a;
        us += b;
      }
      lcount[q + 1] = ls;
      ucount[q + 1] = us;
    }
    run_chunks(ncell, nt, [&](long c0, long c1, int w) {
      for (long c = c0; c < c1; ++c) {
        const Side s{tlo[c], tlo[c] + sc[c], int(c)};
        lower[lh[w][canon(low[c] + sc[c])]++] = s;
        upper[uh[w][canon(low[c])]++] = s;
      }
    });
    lh.clear();
    uh.clear();
    const double nscale = M.extent[axis] / double(lat[axis]);
    const double tscale = M.extent[t] / double(lat[t]);
    auto by_lo = [](const Side& a, const Side& b) { return a.lo < b.lo; };
    // faces and boundary faces of each line: counted, then written
    std::vector<long> nf(lines + 2, 0), nbf(lines + 2, 0);
    const size_t f0 = M.fminus.size(), bf0 = M.bcell.size();
    auto merge = [&](long line, bool write) {
      Side* l0 = lower.data() + lcount[line];
      Side* l1 = lower.data() + lcount[line + 1];
      Side* u0 = upper.data() + ucount[line];
      Side* u1 = upper.data() + ucount[line + 1];
      if (!write) {
        std::sort(l0, l1, by_lo);
        std::sort(u0, u1, by_lo);
      }
      size_t fo = write ? f0 + nf[line] : 0, bo = write ? bf0 + nbf[line] : 0;
      long count = 0;
      Side *li = l0, *ui = u0;
      while (li < l1 && ui < u1) {
        const long lo = std::max(li->lo, ui->lo), hi = std::min(li->hi, ui->hi);
        if (lo < hi) {
          if (write) {
            M.fminus[fo] = li->cell;
            M.fplus[fo] = ui->cell;
            M.faxis[fo] = int8_t(axis);
            M.farea[fo] = (hi - lo) * tscale;
            const double nc = M.origin[axis] + line * nscale;
            const double tc = M.origin[t] + (lo + hi) * tscale / 2;
            M.fcx[fo] = axis == 0 ? nc : tc;
            M.fcy[fo] = axis == 0 ? tc : nc;
            ++fo;
          }
          ++count;
        }
        const long lh_ = li->hi, uh_ = ui->hi;
        if (lh_ <= uh_) ++li;
        if (uh_ <= lh_) ++ui;
      }
      if (!write) {
        nf[line + 1] = count;
        nbf[line + 1] = (l1 - li) + (u1 - ui);
        return;
      }
      for (; li < l1; ++li, ++bo) {
        M.bcell[bo] = li->cell;
        M.baxis[bo] = int8_t(axis);
        M.bside[bo] = 1;
        M.barea[bo] = (li->hi - li->lo) * M.extent[t] / double(lat[t]);
      }
      for (; ui < u1; ++ui, ++bo) {
        M.bcell[bo] = ui->cell;
        M.baxis[bo] = int8_t(axis);
        M.bside[bo] = -1;
        M.barea[bo] = (ui->hi - ui->lo) * M.extent[t] / double(lat[t]);
      }
    };
    parallel_for(lines + 1, [&](long q0, long q1, int) {
      for (long q = q0; q < q1; ++q) merge(q, false);
    }, 16);
    for (long q = 0; q <= lines; ++q) {
      nf[q + 1] += nf[q];
      nbf[q + 1] += nbf[q];
    }
    const size_t nfa = f0 + nf[lines + 1], nba = bf0 + nbf[lines + 1];
    M.fminus.resize(nfa);
    M.fplus.resize(nfa);
    M.faxis.resize(nfa);
    M.farea.resize(nfa);
    M.fcx.resize(nfa);
    M.fcy.resize(nfa);
    M.bcell.resize(nba);
    M.baxis.resize(nba);
    M.bside.resize(nba);
    M.barea.resize(nba);
    parallel_for(lines + 1, [&](long q0, long q1, int) {
      for (long q = q0; q < q1; ++q) merge(q, true);
    }, 16);
  }
}

}  // namespace

std::unique_ptr<TreeMesh> make_tree_mesh(const int base[2], const int bs[2], const double origin[2],
                                         const double extent[2], const bool periodic[2], int max_level,
                                         const std::vector<BlockKey>& refined) {
  auto M = std::make_unique<TreeMesh>();
  for (int a = 0; a < 2; ++a) {
    if (bs[a] < 2 || bs[a] % 2) throw std::invalid_argument("block dimensions must be even and at least two");
    if (base[a] < bs[a] || base[a] % bs[a])
      throw std::invalid_argument("base mesh dimensions must be integer multiples of block dimensions");
    if (!(extent[a] > 0)) throw std::invalid_argument("domain extents must be positive");
    M->base[a] = base[a];
    M->bs[a] = bs[a];
    M->layout[a] = base[a] / bs[a];
    M->origin[a] = origin[a];
    M->extent[a] = extent[a];
    M->root_extent[a] = extent[a] / M->layout[a];
    M->periodic[a] = periodic[a];
  }
  if (max_level < 0) throw std::invalid_argument("max_level must be nonnegative");
  M->max_level = max_level;
  std::unordered_set<uint64_t> req;
  req.reserve(refined.size() * 2 + 1);
  for (const BlockKey& k : refined)
    if (k.level < max_level) req.insert(key64(k.level, k.i, k.j));
  std::vector<BlockKey> stack;
  for (int j = 0; j < M->layout[1]; ++j)
    for (int i = 0; i < M->layout[0]; ++i) stack.push_back({0, i, j});
  while (!stack.empty()) {
    BlockKey k = stack.back();
    stack.pop_back();
    if (k.level >= max_level || !req.count(key64(k.level, k.i, k.j))) continue;
    M->internal.insert(key64(k.level, k.i, k.j));
    for (int c = 0; c < 4; ++c) stack.push_back({k.level + 1, 2 * k.i + (c & 1), 2 * k.j + (c >> 1)});
  }
  PhaseTimer pt;
  balance(*M);
  pt.mark("balance");
  collect_leaves(*M, M->blocks);
  std::sort(M->blocks.begin(), M->blocks.end(), [](const BlockKey& a, const BlockKey& b) {
    return std::make_tuple(a.level, a.j, a.i) < std::make_tuple(b.level, b.j, b.i);
  });
  M->block_of.reserve(M->blocks.size() * 2);
  M->bspacing.resize(2 * M->blocks.size());
  M->borigin.resize(2 * M->blocks.size());
  M->max_leaf = 0;
  for (int b = 0; b < int(M->blocks.size()); ++b) {
    const BlockKey& k = M->blocks[b];
    M->block_of[key64(k.level, k.i, k.j)] = b;
    block_geometry(*M, k, &M->bspacing[2 * b]);
    block_origin(*M, k, &M->borigin[2 * b]);
    M->max_leaf = std::max(M->max_leaf, k.level);
  }
  pt.mark("leaves and block geometry", long(M->blocks.size()));
  build_faces(*M);
  pt.mark("faces", long(M->nfaces()));
  return M;
}

// ---------------------------------------------------------------------------
// Inverse face metric.

namespace {

struct PatchFace {
  int face, axis;
  double area, d0, d1;
};

// Vertex-patch Riesz map (gradient.jl:194-232): L = N (sym M'N)^-1 N'
// + tr/(n-2) (I - M (M'M)^-1 M'), symmetrised.
std::vector<double> patch_metric(const std::vector<PatchFace>& F) {
  const int n = int(F.size());
  double mn00 = 0, mn01 = 0, mn10 = 0, mn11 = 0, g00 = 0, g01 = 0, g11 = 0;
  std::vector<double> m0(n), m1(n), n0(n), n1(n);
  for (int r = 0; r < n; ++r) {
    m0[r] = F[r].area * F[r].d0;
    m1[r] = F[r].area * F[r].d1;
    n0[r] = F[r].axis == 0 ? 0.5 : 0.0;
    n1[r] = F[r].axis == 1 ? 0.5 : 0.0;
    mn00 += m0[r] * n0[r];
    mn01 += m0[r] * n1[r];
    mn10 += m1[r] * n0[r];
    mn11 += m1[r] * n1[r];
    g00 += m0[r] * m0[r];
    g01 += m0[r] * m1[r];
    g11 += m1[r] * m1[r];
  }
  const double s00 = mn00, s11 = mn11, s01 = (mn01 + mn10) / 2;
  const double det_s = s00 * s11 - s01 * s01;
  if (!(s00 > 0) || !(det_s > 0)) throw std::invalid_argument("degenerate vertex patch in gradient construction");
  const double det_g = g00 * g11 - g01 * g01;
  if (!(det_g > 0)) throw std::invalid_argument("degenerate vertex patch in gradient construction");
  // inverses of the 2x2 symmetric matrices
  const double si00 = s11 / det_s, si11 = s00 / det_s, si01 = -s01 / det_s;
  const double gi00 = g11 / det_g, gi11 = g00 / det_g, gi01 = -g01 / det_g;
  std::vector<double> C(n * n), L(n * n);
  double trace = 0;
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      const double v = n0[r] * (si00 * n0[c] + si01 * n1[c]) + n1[r] * (si01 * n0[c] + si11 * n1[c]);
      C[r * n + c] = v;
      if (r == c) trace += v;
    }
  const double stab = trace / (n - 2);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      const double proj = (r == c ? 1.0 : 0.0) -
                          (m0[r] * (gi00 * m0[c] + gi01 * m1[c]) + m1[r] * (gi01 * m0[c] + gi11 * m1[c]));
      L[r * n + c] = C[r * n + c] + stab * proj;
    }
  std::vector<double> out(n * n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) out[r * n + c] = (L[r * n + c] + L[c * n + r]) / 2;
  return out;
}

}  // namespace

FaceMetric face_metric(const TreeMesh& M) {
  const int F = M.nfaces();
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  const int maxl = M.max_leaf;
  const long lat[2] = {long(M.base[0]) << maxl, long(M.base[1]) << maxl};
  const int nblk = int(M.blocks.size());
  // same-level leaf flags around each block: W E S N SW SE NW NE
  std::vector<std::array<uint8_t, 8>> flag(nblk);
  const int dxs[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, dys[8] = {0, 0, -1, 1, -1, -1, 1, 1};
  parallel_for(nblk, [&](long b0, long b1, int) {
    for (long b = b0; b < b1; ++b) {
      const BlockKey& k = M.blocks[b];
      const long px = long(M.layout[0]) << k.level, py = long(M.layout[1]) << k.level;
      for (int q = 0; q < 8; ++q) {
        long x = k.i + dxs[q], y = k.j + dys[q];
        bool ok = true;
        if (x < 0 || x >= px) {
          if (M.periodic[0]) x = wrapl(x, px);
          else ok = false;
        }
        if (y < 0 || y >= py) {
          if (M.periodic[1]) y = wrapl(y, py);
          else ok = false;
        }
        // outside a wall: the vertex is physical and handled before this test
        flag[b][q] = ok ? uint8_t(M.block_of.count(key64(k.level, int(x), int(y))) != 0) : uint8_t(1);
      }
    }
  }, 256);
  FaceMetric K;
  K.diag.resize(F);
  hvec<std::array<double, 2>> disp(F);
  // (vertex, face) pairs of the irregular vertices, per thread in face order
  const int nt = worker_count(F);
  std::vector<std::vector<std::pair<uint64_t, int>>> cand_t(nt);
  run_chunks(F, nt, [&](long fa, long fb, int w) {
    auto& cand = cand_t[w];
    for (long f = fa; f < fb; ++f) {
      const int m = M.fminus[f], p = M.fplus[f];
      const int a = M.faxis[f], t = 1 - a;
      const int bm = m / nb, bp = p / nb;
      const BlockKey& km = M.blocks[bm];
      const BlockKey& kp = M.blocks[bp];
      const double hm = M.bspacing[2 * bm + a], hp = M.bspacing[2 * bp + a];
      const long sm = 1L << (maxl - km.level), sp = 1L << (maxl - kp.level);
      const long mx = (long(km.i) * bx + (m % nb) % bx) * sm, my = (long(km.j) * by + (m % nb) / bx) * sm;
      const long qx = (long(kp.i) * bx + (p % nb) % bx) * sp, qy = (long(kp.j) * by + (p % nb) / bx) * sp;
      const long mlo[2] = {mx, my}, plo[2] = {qx, qy};
      // displacement: normal = mean width; tangential = lattice centre offset
      long offset = 2 * plo[t] + sp - 2 * mlo[t] - sm;
      if (M.periodic[t]) {
        if (offset > lat[t]) offset -= 2 * lat[t];
        if (offset < -lat[t]) offset += 2 * lat[t];
      }
      disp[f][a] = (hm + hp) / 2;
      disp[f][t] = double(offset) * ((M.extent[t] / double(lat[t])) / 2);
      K.diag[f] = 0.0;
      long normal = mlo[a] + sm;
      if (M.periodic[a]) normal = wrapl(normal, lat[a]);
      const long lo = std::max(mlo[t], plo[t]), hi = std::min(mlo[t] + sm, plo[t] + sp);
      const double simple = 1.0 / (2.0 * M.farea[f] * disp[f][a]);
      const bool same_level = km.level == kp.level;
      for (int e = 0; e < 2; ++e) {
        const long tv = e == 0 ? lo : hi;
        long vx = a == 0 ? normal : tv, vy = a == 0 ? tv : normal;
        if (M.periodic[0]) vx = wrapl(vx, lat[0]);
        if (M.periodic[1]) vy = wrapl(vy, lat[1]);
        const bool physical = (!M.periodic[0] && (vx == 0 || vx == lat[0])) ||
                              (!M.periodic[1] && (vy == 0 || vy == lat[1]));
        bool regular = physical;
        if (!regular && same_level) {
          // position of the vertex relative to the minus block, in its cells
          const long bx0 = long(km.i) * bx * sm, by0 = long(km.j) * by * sm;
          long lx = vx - bx0, ly = vy - by0;
          const long spanx = long(bx) * sm, spany = long(by) * sm;
          if (M.periodic[0]) {
            if (lx < 0) lx += lat[0];
            if (lx > spanx) lx -= lat[0];
          }
          if (M.periodic[1]) {
            if (ly < 0) ly += lat[1];
            if (ly > spany) ly -= lat[1];
          }
          const int ex = lx == 0 ? -1 : (lx == spanx ? 1 : 0);
          const int ey = ly == 0 ? -1 : (ly == spany ? 1 : 0);
          const auto& fl = flag[bm];
          if (ex == 0 && ey == 0) regular = true;
          else if (ey == 0) regular = fl[ex < 0 ? 0 : 1];
          else if (ex == 0) regular = fl[ey < 0 ? 2 : 3];
          else {
            const int side_x = ex < 0 ? 0 : 1, side_y = ey < 0 ? 2 : 3;
            const int corner = (ey < 0 ? 4 : 6) + (ex < 0 ? 0 : 1);
            regular = fl[side_x] && fl[side_y] && fl[corner];
          }
        }
        if (regular) {
          K.diag[f] += simple;
        } else {
          cand.push_back({(uint64_t(vx) << 32) | uint64_t(vy), int(f)});
        }
      }
    }
  });
  // vertex patches: every vertex's faces in increasing face order, vertices in
  // increasing key order
  std::vector<std::pair<uint64_t, int>> cand;
  {
    size_t total = 0;
    for (auto& v : cand_t) total += v.size();
    cand.reserve(total);
    for (auto& v : cand_t) {
      cand.insert(cand.end(), v.begin(), v.end());
      std::vector<std::pair<uint64_t, int>>().swap(v);
    }
  }
  {
    const int nbk = 64 * worker_count(long(cand.size()), 1 << 16);
    const uint64_t span = uint64_t(lat[0]) + 1;
    bucket_sort(cand, nbk, [&](const std::pair<uint64_t, int>& x) { return int((x.first >> 32) * nbk / span); },
                [](const std::pair<uint64_t, int>& x, const std::pair<uint64_t, int>& y) { return x < y; });
  }
  std::vector<size_t> vstart;
  for (size_t q = 0; q < cand.size(); ++q)
    if (q == 0 || cand[q].first != cand[q - 1].first) vstart.push_back(q);
  vstart.push_back(cand.size());
  const long nv = long(vstart.size()) - 1;
  // each face meets two vertices, so its diagonal collects two terms from
  // zero: the sum does not depend on the order in which patches are applied
  struct Entry {
    int f, g;
    double v;
  };
  const int pt_n = worker_count(nv, 256);
  std::vector<std::vector<std::pair<int, double>>> dg_t(pt_n);
  std::vector<std::vector<Entry>> off_t(pt_n);
  std::vector<std::string> err_t(pt_n);
  run_chunks(nv, pt_n, [&](long v0, long v1, int w) {
    try {
      std::vector<PatchFace> pf;
      for (long v = v0; v < v1; ++v) {
        const size_t q0 = vstart[v], q1 = vstart[v + 1];
        if (q1 - q0 <= 2) {
          for (size_t q = q0; q < q1; ++q) {
            const int f = cand[q].second;
            dg_t[w].push_back({f, 1.0 / (2.0 * M.farea[f] * disp[f][M.faxis[f]])});
          }
          continue;
        }
        pf.clear();
        for (size_t q = q0; q < q1; ++q) {
          const int f = cand[q].second;
          pf.push_back({f, M.faxis[f], M.farea[f], disp[f][0], disp[f][1]});
        }
        // a patch of four same-level faces with square geometry is the MAC patch
        bool mac = pf.size() == 4;
        if (mac) {
          int nax[2] = {0, 0};
          for (const auto& e : pf) {
            nax[e.axis]++;
            const double tang = e.axis == 0 ? e.d1 : e.d0;
            mac = mac && tang == 0.0 && M.block_level(M.fminus[e.face]) == M.block_level(M.fplus[e.face]) &&
                  M.block_level(M.fminus[e.face]) == M.block_level(M.fminus[pf[0].face]);
          }
          mac = mac && nax[0] == 2 && nax[1] == 2;
        }
        if (mac) {
          for (const auto& e : pf) dg_t[w].push_back({e.face, 1.0 / (2.0 * e.area * (e.axis == 0 ? e.d0 : e.d1))});
          continue;
        }
        const std::vector<double> L = patch_metric(pf);
        const int n = int(pf.size());
        for (int r = 0; r < n; ++r) {
          dg_t[w].push_back({pf[r].face, L[r * n + r]});
          for (int c = 0; c < n; ++c)
            if (c != r) off_t[w].push_back({pf[r].face, pf[c].face, L[r * n + c]});
        }
      }
    } catch (const std::exception& e) {
      err_t[w] = e.what();
    }
  });
  for (const auto& e : err_t)
    if (!e.empty()) throw std::invalid_argument(e);
  for (const auto& v : dg_t)
    for (const auto& d : v) K.diag[d.first] += d.second;
  std::vector<Entry> off;
  {
    size_t total = 0;
    for (auto& v : off_t) total += v.size();
    off.reserve(total);
    for (auto& v : off_t) off.insert(off.end(), v.begin(), v.end());
  }
  {
    const int nbk = 64 * worker_count(long(off.size()), 1 << 16);
    bucket_sort(off, nbk, [&](const Entry& x) { return int(long(x.f) * nbk / F); },
                [](const Entry& x, const Entry& y) { return std::tie(x.f, x.g, x.v) < std::tie(y.f, y.g, y.v); });
  }
  // merge duplicates (two faces meet in at most one patch, kept for safety)
  K.off_ptr.assign(F + 1, 0);
  for (size_t q = 0; q < off.size();) {
    const int f = off[q].f, g = off[q].g;
    double v = 0;
    bool first = true;
    for (; q < off.size() && off[q].f == f && off[q].g == g; ++q) {
      v = first ? off[q].v : v + off[q].v;
      first = false;
    }
    if (v != 0.0) {
      K.off_face.push_back(g);
      K.off_val.push_back(v);
      K.off_ptr[f + 1]++;
    }
  }
  for (int f = 0; f < F; ++f) K.off_ptr[f + 1] += K.off_ptr[f];
  std::atomic<bool> nonpositive{false};
  parallel_for(F, [&](long f0, long f1, int) {
    for (long f = f0; f < f1; ++f)
      if (!(K.diag[f] > 0)) nonpositive = true;
  });
  if (nonpositive) throw std::invalid_argument("gradient Riesz map has a non-positive diagonal");
  return K;
}

// ---------------------------------------------------------------------------
// Levels

namespace {

std::unique_ptr<TreeMesh> pruned_mesh(const TreeMesh& M) {
  std::vector<BlockKey> keep;
  int maxl = -1;
  for (uint64_t pk : M.internal) {
    const int l = int(pk >> 56), i = int((pk >> 28) & 0xFFFFFFF), j = int(pk & 0xFFFFFFF);
    bool child = false;
    for (int c = 0; c < 4 && !child; ++c)
      child = M.internal.count(key64(l + 1, 2 * i + (c & 1), 2 * j + (c >> 1))) != 0;
    if (child) {
      keep.push_back({l, i, j});
      maxl = std::max(maxl, l);
    }
  }
  return make_tree_mesh(M.base, M.bs, M.origin, M.extent, M.periodic, maxl + 1, keep);
}

std::unique_ptr<TreeMesh> agglomerated_mesh(const TreeMesh& M) {
  if (M.layout[0] % 2 || M.layout[1] % 2) return nullptr;
  const int cb[2] = {M.base[0] / 2, M.base[1] / 2};
  for (int a = 0; a < 2; ++a)
    if (M.periodic[a] && cb[a] < 4) return nullptr;
  return make_tree_mesh(cb, M.bs, M.origin, M.extent, M.periodic, 0, {});
}

// Topology-only parts of a level: slots, regularity, neighbours, patterns.
void level_topology(LevelData& L) {
  const TreeMesh& M = *L.mesh;
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  L.nblocks = int(M.blocks.size());
  L.ncells = M.ncells();
  PhaseTimer pt;
  L.K = face_metric(M);
  pt.mark("face metric", long(M.nfaces()));
  const int F = M.nfaces();
  L.face_regular.resize(F);
  L.slot_a.resize(F);
  L.slot_b.resize(F);
  parallel_for(F, [&](long f0, long f1, int) {
    for (long f = f0; f < f1; ++f) {
      const int m = M.fminus[f], p = M.fplus[f];
      const bool reg = M.block_level(m) == M.block_level(p) && L.K.off_ptr[f] == L.K.off_ptr[f + 1];
      L.face_regular[f] = reg;
      L.slot_a[f] = L.slot_b[f] = -1;
      if (!reg) continue;
      const int bm = m / nb, lm = m % nb, bp = p / nb, lp = p % nb;
      if (M.faxis[f] == 0) {
        L.slot_a[f] = int64_t(bm) * (bx + 1) * by + (lm / bx) * (bx + 1) + (lm % bx) + 1;
        L.slot_b[f] = int64_t(bp) * (bx + 1) * by + (lp / bx) * (bx + 1) + (lp % bx);
      } else {
        L.slot_a[f] = int64_t(bm) * bx * (by + 1) + (lm / bx + 1) * bx + (lm % bx);
        L.slot_b[f] = int64_t(bp) * bx * (by + 1) + (lp / bx) * bx + (lp % bx);
      }
    }
  });
  // neighbours
  L.nbr.assign(4 * L.nblocks, -1);
  parallel_for(L.nblocks, [&](long b0, long b1, int) {
    for (long b = b0; b < b1; ++b) {
      const BlockKey& k = M.blocks[b];
      const long px = long(M.layout[0]) << k.level, py = long(M.layout[1]) << k.level;
      const int dx[4] = {-1, 1, 0, 0}, dy[4] = {0, 0, -1, 1};
      for (int s = 0; s < 4; ++s) {
        long x = k.i + dx[s], y = k.j + dy[s];
        if (x < 0 || x >= px) {
          if (!M.periodic[0]) continue;
          x = wrapl(x, px);
        }
        if (y < 0 || y >= py) {
          if (!M.periodic[1]) continue;
          y = wrapl(y, py);
        }
        auto it = M.block_of.find(key64(k.level, int(x), int(y)));
        if (it != M.block_of.end()) L.nbr[4 * b + s] = it->second;
      }
    }
  }, 256);
  L.volume.resize(L.nblocks);
  for (int b = 0; b < L.nblocks; ++b) L.volume[b] = M.bspacing[2 * b] * M.bspacing[2 * b + 1];
  L.total_volume = 0;
  for (int b = 0; b < L.nblocks; ++b) L.total_volume += L.volume[b] * nb;
  // irregular cells: touching a non-regular face
  L.cell_irr.assign(L.ncells, -1);
  std::vector<uint8_t> irr(L.ncells, 0);
  for (int f = 0; f < F; ++f)
    if (!L.face_regular[f]) irr[M.fminus[f]] = irr[M.fplus[f]] = 1;
  L.blk_irr.assign(L.nblocks + 1, 0);
  L.irr_local.clear();
  for (int b = 0; b < L.nblocks; ++b) {
    for (int q = 0; q < nb; ++q)
      if (irr[b * nb + q]) {
        L.cell_irr[b * nb + q] = int(L.irr_local.size());
        L.irr_local.push_back(q);
      }
    L.blk_irr[b + 1] = int(L.irr_local.size());
  }
  // open boundary faces: the cell's conductance slot beyond the side (for
  // irregular cells too: the stencil holds their regular and open faces)
  L.bopen.clear();
  L.bslot.clear();
  L.bdst.clear();
  const int BF = int(M.bcell.size()), nrows = int(L.irr_local.size());
  std::vector<int> rcount(nrows + 1, 0);
  for (int f = 0; f < BF; ++f) {
    if (!L.open[2 * M.baxis[f] + (M.bside[f] > 0 ? 1 : 0)]) continue;
    const int c = M.bcell[f], b = c / nb, q = c % nb, i = q % bx, j = q / bx;
    L.bopen.push_back(f);
    if (L.cell_irr[c] >= 0) rcount[L.cell_irr[c] + 1]++;
    if (M.baxis[f] == 0) {
      const int ii = M.bside[f] > 0 ? bx : 0;
      L.bslot.push_back(int64_t(b) * (bx + 1) * by + int64_t(j) * (bx + 1) + ii);
      L.bdst.push_back(M.bside[f] > 0 ? (b * by + j) * 4 + 1 : c * 4 + 0);
    } else {
      const int jj = M.bside[f] > 0 ? by : 0;
      L.bslot.push_back(int64_t(b) * bx * (by + 1) + int64_t(jj) * bx + i);
      L.bdst.push_back(M.bside[f] > 0 ? (b * bx + i) * 4 + 3 : c * 4 + 2);
    }
  }
  for (int r = 0; r < nrows; ++r) rcount[r + 1] += rcount[r];
  L.irr_bptr = rcount;
  L.irr_bface.assign(rcount[nrows], 0);
  {
    std::vector<int> at(rcount.begin(), rcount.end() - 1);
    for (int k = 0; k < int(L.bopen.size()); ++k) {
      const int r = L.cell_irr[M.bcell[L.bopen[k]]];
      if (r >= 0) L.irr_bface[at[r]++] = k;
    }
  }
  pt.mark("level topology");
}

// Device form of the irregular rows (see LevelData): per block, the rows in
// decreasing length order with their entries in groups of four (entry 4g+t of
// row x of the block at (r0 * ng + g * nr + x) * 4 + t, ng = ceil(w / 4)),
// each entry's LDS slot in the block tile, and the cells outside the block and
// its same-level halo. Tile slots: interior cells j * bx + i, then the halo
// strips W (bx*by + j), E (+ by), S (+ 2 by + i), N (+ 2 by + bx + i), then
// the external cells (+ 2 (bx + by) + e). Padding entries: value 0, the row's
// own slot. The structure only: each entry's value position (ell_pos) and each
// row's position (ell_row) are kept for the value passes.
void build_ell(LevelData& L) {
  const TreeMesh& M = *L.mesh;
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  const int hw = nb, he = nb + by, hs = nb + 2 * by, hn = nb + 2 * by + bx, ext0 = nb + 2 * (bx + by);
  const int nrows = int(L.irr_local.size());
  // a row's ELL entries: its off-diagonal entries with a non-regular face's term
  std::vector<int> rlen(nrows, 0);
  for (int r = 0; r < nrows; ++r)
    for (int e = L.irr_ptr[r]; e < L.irr_ptr[r + 1]; ++e)
      if (L.ell_part[e] && e != L.irr_diag_e[r]) ++rlen[r];
  auto row_len = [&](int r) { return rlen[r]; };
  int width = 0;
  for (int r = 0; r < nrows; ++r) width = std::max(width, row_len(r));
  L.ell_w = width;
  const int ng = (width + 3) / 4, w4 = 4 * ng;
  L.ell_ref.assign(size_t(nrows) * w4, 0);
  L.ell_val.assign(size_t(nrows) * w4, 0.0);
  L.ell_bound.assign(size_t(nrows) * w4, 0.0);
  L.ell_loc.assign(nrows, 0u);
  L.ell_diag.assign(nrows, 0.0);
  L.ell_bdiag.assign(nrows, 0.0);
  L.ell_pos.assign(L.irr_col.size(), -1);
  L.ell_dpos.assign(L.irr_col.size(), -1);
  L.ell_row.assign(nrows, 0);
  L.ext_max = 0;
  L.blk_ext.assign(L.nblocks + 1, 0);
  L.ext_col.clear();
  // blocks in parallel chunks; each chunk's external cells are appended in
  // block order below
  const int nt = worker_count(L.nblocks, 256);
  std::vector<std::vector<int>> ext_t(nt), next_t(nt);
  std::vector<int> emax(nt, 0);
  std::vector<uint8_t> over(nt, 0);
  run_chunks(L.nblocks, nt, [&](long b0, long b1, int w) {
    std::unordered_map<int, int> ext;
    std::vector<int> order;
    for (long b = b0; b < b1; ++b) {
      ext.clear();
      const int* nbr = &L.nbr[4 * b];
      const int r0 = L.blk_irr[b], nr = L.blk_irr[b + 1] - r0;
      // rows of the block by decreasing length: the lanes reading group g of
      // the block's region are then a prefix
      order.resize(nr);
      for (int x = 0; x < nr; ++x) order[x] = r0 + x;
      std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return row_len(p) > row_len(q); });
      auto slot = [&](int x, int k) { return (size_t(r0) * ng + size_t(k / 4) * nr + x) * 4 + k % 4; };
      for (int x = 0; x < nr; ++x) {
        const int r = order[x];
        const int loc = L.irr_local[r];
        const int len = row_len(r);
        L.ell_loc[r0 + x] = uint32_t(loc) | (uint32_t(len) << 16);
        L.ell_row[r] = r0 + x;
        if (L.irr_diag_e[r] >= 0) L.ell_dpos[L.irr_diag_e[r]] = r0 + x;
        for (int k = 0; k < w4; ++k) L.ell_ref[slot(x, k)] = uint16_t(loc);
        int k = 0;
        for (int e = L.irr_ptr[r]; e < L.irr_ptr[r + 1]; ++e) {
          if (!L.ell_part[e] || e == L.irr_diag_e[r]) continue;
          const int col = L.irr_col[e];
          const int cb = col / nb, cl = col % nb, ci = cl % bx, cj = cl / bx;
          int ref;
          if (cb == b) ref = cl;
          else if (cb == nbr[0] && ci == bx - 1) ref = hw + cj;
          else if (cb == nbr[1] && ci == 0) ref = he + cj;
          else if (cb == nbr[2] && cj == by - 1) ref = hs + ci;
          else if (cb == nbr[3] && cj == 0) ref = hn + ci;
          else {
            auto it = ext.find(col);
            if (it == ext.end()) {
              it = ext.emplace(col, int(ext.size())).first;
              ext_t[w].push_back(col);
            }
            ref = ext0 + it->second;
          }
          const size_t at = slot(x, k);
          L.ell_ref[at] = uint16_t(ref);
          L.ell_pos[e] = int(at);
          ++k;
        }
      }
      // LDS slots of a block tile: up to max(nb, 32) external cells (4x4
      // tiles at coarse-fine interfaces reach beyond their own 16 cells)
      if (int(ext.size()) > std::max(nb, 32)) over[w] = 1;
      emax[w] = std::max(emax[w], int(ext.size()));
      next_t[w].push_back(int(ext.size()));
    }
  });
  for (uint8_t o : over)
    if (o) throw std::runtime_error("too many external couplings in one block");
  long b = 0;
  for (int w = 0; w < nt; ++w) {
    for (int n : next_t[w]) {
      L.blk_ext[b + 1] = L.blk_ext[b] + n;
      ++b;
    }
    L.ext_col.insert(L.ext_col.end(), ext_t[w].begin(), ext_t[w].end());
    L.ext_max = std::max(L.ext_max, emax[w]);
  }
}

// Structure of the irregular rows (built once per topology): the columns of
// every row (CSR, increasing column) and, per entry, the face-metric terms
// that sum to its value in the order the value pass adds them.
void level_pattern(LevelData& L) {
  const TreeMesh& M = *L.mesh;
  const int nb = M.nb();
  const int F = M.nfaces();
  const int nrows = int(L.irr_local.size());
  // faces per irregular row, in increasing face order (CSR)
  std::vector<int> inc_ptr(nrows + 1, 0), inc;
  {
    const int nt = worker_count(F);
    std::vector<std::vector<std::pair<int, int>>> pr(nt);
    run_chunks(F, nt, [&](long f0, long f1, int w) {
      for (long f = f0; f < f1; ++f) {
        const int m = M.fminus[f], p = M.fplus[f];
        if (L.cell_irr[m] >= 0) pr[w].push_back({L.cell_irr[m], int(f)});
        if (p != m && L.cell_irr[p] >= 0) pr[w].push_back({L.cell_irr[p], int(f)});
      }
    });
    for (const auto& v : pr)
      for (const auto& e : v) inc_ptr[e.first + 1]++;
    for (int r = 0; r < nrows; ++r) inc_ptr[r + 1] += inc_ptr[r];
    inc.resize(inc_ptr[nrows]);
    std::vector<int> at(inc_ptr.begin(), inc_ptr.end() - 1);
    for (const auto& v : pr)
      for (const auto& e : v) inc[at[e.first]++] = e.second;
  }
  std::vector<int> row_blk(nrows);
  for (int b = 0; b < L.nblocks; ++b)
    for (int r = L.blk_irr[b]; r < L.blk_irr[b + 1]; ++r) row_blk[r] = b;
  struct Term {
    int f, g;
    double k, s;
    bool diag;
  };
  // rows in parallel chunks; each chunk's columns and terms are appended in
  // row order below
  struct Part {
    std::vector<int> col, ncol, nterm, diag_e, f, g;
    std::vector<double> k, s;
    std::vector<uint8_t> dg;
  };
  const int nt = worker_count(nrows, 2048);
  std::vector<Part> part(nt);
  run_chunks(nrows, nt, [&](long r0, long r1, int w) {
    Part& P = part[w];
    // (column, (term index, 0)): sorted by column only, as the values were
    std::vector<std::pair<int, std::pair<double, double>>> row;
    std::vector<Term> terms;
    for (long r = r0; r < r1; ++r) {
      const int c = row_blk[r] * nb + L.irr_local[r];
      row.clear();
      terms.clear();
      for (int q = inc_ptr[r]; q < inc_ptr[r + 1]; ++q) {
        const int f = inc[q];
        const double sf = (M.fminus[f] == c) ? -1.0 : 1.0;
        auto add = [&](int g, double kfg, bool diag) {
          terms.push_back({f, g, kfg, sf * -1.0, diag});
          row.push_back({M.fminus[g], {double(terms.size() - 1), 0.0}});
          terms.push_back({f, g, kfg, sf * 1.0, diag});
          row.push_back({M.fplus[g], {double(terms.size() - 1), 0.0}});
        };
        add(f, L.K.diag[f], true);
        for (int o = L.K.off_ptr[f]; o < L.K.off_ptr[f + 1]; ++o) add(L.K.off_face[o], L.K.off_val[o], false);
      }
      std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      int de = -1, nc = 0;
      size_t q = 0;
      while (q < row.size()) {
        const int col = row[q].first;
        int n = 0;
        while (q < row.size() && row[q].first == col) {
          const Term& t = terms[size_t(row[q].second.first)];
          P.f.push_back(t.f);
          P.g.push_back(t.g);
          P.k.push_back(t.k);
          P.s.push_back(t.s);
          P.dg.push_back(t.diag);
          ++q;
          ++n;
        }
        if (col == c) de = int(P.col.size());
        P.col.push_back(col);
        P.nterm.push_back(n);
        ++nc;
      }
      P.ncol.push_back(nc);
      P.diag_e.push_back(de);
    }
  });
  // offsets of every chunk's rows, columns and terms, then a parallel copy
  std::vector<long> row_off(nt + 1, 0), col_off(nt + 1, 0), term_off(nt + 1, 0);
  for (int w = 0; w < nt; ++w) {
    row_off[w + 1] = row_off[w] + long(part[w].ncol.size());
    col_off[w + 1] = col_off[w] + long(part[w].col.size());
    term_off[w + 1] = term_off[w] + long(part[w].f.size());
  }
  const long ncols = col_off[nt], nterms = term_off[nt];
  L.irr_ptr.assign(nrows + 1, 0);
  L.irr_col.resize(ncols);
  L.irr_diag_e.assign(nrows, -1);
  L.term_ptr.resize(ncols + 1);
  L.term_ptr[0] = 0;
  L.term_f.resize(nterms);
  L.term_g.resize(nterms);
  L.term_k.resize(nterms);
  L.term_s.resize(nterms);
  L.term_diag.resize(nterms);
  parallel_tasks(nt, [&](long w) {
    Part& P = part[w];
    long cum = col_off[w];
    for (size_t x = 0; x < P.ncol.size(); ++x) {
      const long r = row_off[w] + long(x);
      cum += P.ncol[x];
      L.irr_ptr[r + 1] = int(cum);
      L.irr_diag_e[r] = P.diag_e[x] >= 0 ? int(col_off[w] + P.diag_e[x]) : -1;
    }
    long tcum = term_off[w];
    for (size_t e = 0; e < P.nterm.size(); ++e) {
      tcum += P.nterm[e];
      L.term_ptr[col_off[w] + long(e) + 1] = int(tcum);
    }
    std::copy(P.col.begin(), P.col.end(), L.irr_col.begin() + col_off[w]);
    std::copy(P.f.begin(), P.f.end(), L.term_f.begin() + term_off[w]);
    std::copy(P.g.begin(), P.g.end(), L.term_g.begin() + term_off[w]);
    std::copy(P.k.begin(), P.k.end(), L.term_k.begin() + term_off[w]);
    std::copy(P.s.begin(), P.s.end(), L.term_s.begin() + term_off[w]);
    std::copy(P.dg.begin(), P.dg.end(), L.term_diag.begin() + term_off[w]);
    P = Part();
  });
  L.irr_val.assign(L.irr_col.size(), 0.0);
  L.irr_bound.assign(L.irr_col.size(), 0.0);
  L.irr_diag.assign(nrows, 0.0);
  // entries with a term of a non-regular face (the others are regular faces'
  // conductances, which the stencil slots hold)
  L.ell_part.assign(L.irr_col.size(), 0);
  parallel_for(long(L.irr_col.size()), [&](long e0, long e1, int) {
    for (long e = e0; e < e1; ++e)
      for (int q = L.term_ptr[e]; q < L.term_ptr[e + 1]; ++q)
        if (!L.face_regular[L.term_f[q]]) {
          L.ell_part[e] = 1;
          break;
        }
  });
  build_ell(L);
  L.planned = true;
}

// Coefficient-dependent values of a level.
void level_values(LevelData& L) {
  const TreeMesh& M = *L.mesh;
  const int bx = M.bs[0], by = M.bs[1], nb = bx * by;
  const int F = M.nfaces();
  if (!L.planned) {
    PhaseTimer pt;
    level_pattern(L);
    pt.mark("irregular-row pattern", long(L.irr_col.size()));
  }
  L.cx.resize(size_t(L.nblocks) * (bx + 1) * by);
  L.cy.resize(size_t(L.nblocks) * bx * (by + 1));
  parallel_for(long(L.cx.size()), [&](long lo, long hi, int) { std::fill(&L.cx[lo], &L.cx[0] + hi, 0.0); });
  parallel_for(long(L.cy.size()), [&](long lo, long hi, int) { std::fill(&L.cy[lo], &L.cy[0] + hi, 0.0); });
  parallel_for(F, [&](long lo, long hi, int) {
    for (long f = lo; f < hi; ++f) {
      if (!L.face_regular[f]) continue;
      const double a = M.farea[f];
      const double w = L.kappa[f] * L.K.diag[f] * a * a;
      std::vector<double>& arr = M.faxis[f] == 0 ? L.cx : L.cy;
      arr[L.slot_a[f]] = w;
      arr[L.slot_b[f]] = w;
    }
  });
  // open boundary faces: their conductance in the cell's slot beyond the
  // side (the halo value there is zero)
  const bool opened = L.any_open();
  if (opened && L.bw.size() != L.bopen.size())
    throw std::logic_error("open pressure sides without boundary coefficients");
  for (size_t k = 0; k < L.bopen.size(); ++k)
    if (L.bslot[k] >= 0) (M.baxis[L.bopen[k]] == 0 ? L.cx : L.cy)[L.bslot[k]] = L.bw[k];
  // irregular rows: full rows of A (and of the floor bound |J|'|M||J|)
  const int nrows = int(L.irr_local.size());
  std::vector<double> ratio(host_threads(), 0.0);
  parallel_for(nrows, [&](long lo, long hi, int t) {
    double mr = 0;
    for (long r = lo; r < hi; ++r) {
      double rowsum = 0;
      const int de = L.irr_diag_e[r];
      for (int e = L.irr_ptr[r]; e < L.irr_ptr[r + 1]; ++e) {
        // the full entry (v, bound bv) and its non-regular faces' part (vp, bvp)
        double v = 0, bv = 0, vp = 0, bvp = 0;
        for (int q = L.term_ptr[e]; q < L.term_ptr[e + 1]; ++q) {
          const int f = L.term_f[q], g = L.term_g[q];
          const double kfg = L.term_k[q], af = M.farea[f], ag = M.farea[g];
          const bool diag = L.term_diag[q];
          const double sf = diag ? 0.0 : std::sqrt(L.kappa[f]), sg = diag ? 0.0 : std::sqrt(L.kappa[g]);
          const double core = diag ? L.kappa[f] * kfg * af * ag : (sf * kfg) * sg * af * ag;
          const double babs = diag ? core : (sf * std::fabs(kfg)) * sg * af * ag;
          v += L.term_s[q] * core;
          bv += babs;
          if (!L.face_regular[f]) {
            vp += L.term_s[q] * core;
            bvp += babs;
          }
        }
        if (opened && e == de) {
          // the boundary diagonal joins the row's full diagonal (and bound)
          v = v + L.irr_bd[r];
          bv = bv + L.irr_bd[r];
        }
        L.irr_val[e] = v;
        L.irr_bound[e] = bv;
        if (L.ell_pos[e] >= 0) {
          L.ell_val[L.ell_pos[e]] = vp;
          L.ell_bound[L.ell_pos[e]] = bvp;
        }
        if (L.ell_dpos[e] >= 0) L.ell_bdiag[L.ell_dpos[e]] = bvp;
        rowsum += std::fabs(v);
      }
      const double diag = de >= 0 ? L.irr_val[de] : 0.0;
      L.irr_diag[r] = diag;
      L.ell_diag[L.ell_row[r]] = diag;
      if (diag > 0) mr = std::max(mr, rowsum / diag);
    }
    ratio[t] = mr;
  });
  double max_ratio = *std::max_element(ratio.begin(), ratio.end());
  // device layout of the regular conductances, the diagonal and the Jacobi weights
  const size_t nbk = size_t(L.nblocks);
  // every entry is written below
  L.cw.resize(nbk * nb);
  L.cs.resize(nbk * nb);
  L.ce.resize(nbk * by);
  L.cn.resize(nbk * bx);
  L.dinv.resize(L.ncells);
  std::vector<uint8_t> regular(host_threads(), 0);
  // regular cells with a positive open-face conductance: their own row ratio
  // (diagonal plus the conductances of their interior faces, over the diagonal)
  std::vector<double> bratio(host_threads(), 0.0);
  parallel_for(long(nbk), [&](long lo, long hi, int t) {
    for (long b = lo; b < hi; ++b) {
      const double* cx = &L.cx[b * (bx + 1) * by];
      const double* cy = &L.cy[b * bx * (by + 1)];
      for (int j = 0; j < by; ++j) {
        for (int i = 0; i < bx; ++i) {
          L.cw[b * nb + j * bx + i] = cx[j * (bx + 1) + i];
          L.cs[b * nb + j * bx + i] = cy[j * bx + i];
        }
        L.ce[b * by + j] = cx[j * (bx + 1) + bx];
      }
      for (int i = 0; i < bx; ++i) L.cn[b * bx + i] = cy[by * bx + i];
      for (int q = 0; q < nb; ++q) {
        const long c = b * nb + q;
        double d;
        if (L.cell_irr[c] >= 0) {
          d = L.irr_diag[L.cell_irr[c]];
        } else {
          const int i = q % bx, j = q / bx;
          const double cwv = cx[j * (bx + 1) + i], cev = cx[j * (bx + 1) + i + 1], csv = cy[j * bx + i],
                       cnv = cy[(j + 1) * bx + i];
          d = cwv + cev + csv + cnv;
          const bool ow = i == 0 && L.nbr[4 * b] < 0, oe = i == bx - 1 && L.nbr[4 * b + 1] < 0,
                     os = j == 0 && L.nbr[4 * b + 2] < 0, on = j == by - 1 && L.nbr[4 * b + 3] < 0;
          const bool bnd = (ow && cwv > 0) || (oe && cev > 0) || (os && csv > 0) || (on && cnv > 0);
          if (d > 0) {
            if (!bnd) {
              regular[t] = 1;
            } else {
              const double isum = (ow ? 0.0 : cwv) + (oe ? 0.0 : cev) + (os ? 0.0 : csv) + (on ? 0.0 : cnv);
              bratio[t] = std::max(bratio[t], (d + isum) / d);
            }
          }
        }
        L.dinv[c] = d > 0 ? 1.0 / d : 0.0;
      }
    }
  });
  const bool any_regular = *std::max_element(regular.begin(), regular.end()) != 0;
  if (any_regular) max_ratio = std::max(max_ratio, 2.0);
  max_ratio = std::max(max_ratio, *std::max_element(bratio.begin(), bratio.end()));
  if (max_ratio == 0) {
    L.omega = 1.0;
    L.omega_p = 0.0;
  } else {
    const double w = 4.0 / (3.0 * max_ratio);
    L.omega = std::min(1.0, w);
    L.omega_p = w;
  }
}

void restriction_map(LevelData& C, const LevelData& Fl) {
  // Fine and coarse faces are both ordered by (axis, line, lo) on their own
  // lattices; map both onto the common lattice and merge.
  const TreeMesh& fm = *Fl.mesh;
  const TreeMesh& cm = *C.mesh;
  long lat[2];
  for (int a = 0; a < 2; ++a)
    lat[a] = std::max(long(fm.base[a]) << fm.max_leaf, long(cm.base[a]) << cm.max_leaf);
  auto geom = [&](const TreeMesh& M, int f, long& axis, long& normal, long& lo, long& hi) {
    const int nb = M.nb(), bx = M.bs[0], by = M.bs[1];
    const int m = M.fminus[f], p = M.fplus[f];
    const BlockKey& km = M.blocks[m / nb];
    const BlockKey& kp = M.blocks[p / nb];
    const long sm0 = lat[0] / (long(M.base[0]) << km.level), sm1 = lat[1] / (long(M.base[1]) << km.level);
    const long sp0 = lat[0] / (long(M.base[0]) << kp.level), sp1 = lat[1] / (long(M.base[1]) << kp.level);
    const long ml[2] = {(long(km.i) * bx + (m % nb) % bx) * sm0, (long(km.j) * by + (m % nb) / bx) * sm1};
    const long pl[2] = {(long(kp.i) * bx + (p % nb) % bx) * sp0, (long(kp.j) * by + (p % nb) / bx) * sp1};
    const long ms[2] = {sm0, sm1}, ps[2] = {sp0, sp1};
    const int a = M.faxis[f], t = 1 - a;
    axis = a;
    normal = ml[a] + ms[a];
    if (M.periodic[a]) normal = wrapl(normal, lat[a]);
    lo = std::max(ml[t], pl[t]);
    hi = std::min(ml[t] + ms[t], pl[t] + ps[t]);
  };
  const int FF = fm.nfaces(), CF = cm.nfaces();
  hvec<std::array<long, 4>> fg(FF);
  parallel_for(FF, [&](long f0, long f1, int) {
    for (long f = f0; f < f1; ++f) geom(fm, int(f), fg[f][0], fg[f][1], fg[f][2], fg[f][3]);
  });
  // Fine faces run in (axis, line, lo) order and cover each line with
  // disjoint segments, so (axis, line, hi) increases too: each chunk of
  // coarse faces finds its first fine segment by binary search and then
  // walks forward.
  const int nt = worker_count(CF, 4096);
  std::vector<std::vector<int>> idx_t(nt), cnt_t(nt);
  std::vector<std::vector<double>> w_t(nt);
  std::vector<uint8_t> bad(nt, 0);
  run_chunks(CF, nt, [&](long c0, long c1, int w) {
    auto& idx = idx_t[w];
    auto& wt = w_t[w];
    auto& cnt = cnt_t[w];
    long q = -1;
    for (long f = c0; f < c1; ++f) {
      long a, n, lo, hi;
      geom(cm, int(f), a, n, lo, hi);
      const auto key = std::make_tuple(a, n, lo);
      auto before = [&](const std::array<long, 4>& g) { return std::make_tuple(g[0], g[1], g[3]) <= key; };
      if (q < 0) q = std::partition_point(fg.begin(), fg.end(), before) - fg.begin();
      while (q < FF && before(fg[q])) ++q;
      long covered = 0;
      int k = 0;
      for (long r = q; r < FF && fg[r][0] == a && fg[r][1] == n && fg[r][2] < hi; ++r) {
        const long ov = std::min(hi, fg[r][3]) - std::max(lo, fg[r][2]);
        if (ov > 0) {
          idx.push_back(int(r));
          wt.push_back(double(ov) / double(hi - lo));
          covered += ov;
          ++k;
        }
      }
      if (covered != hi - lo) bad[w] = 1;
      cnt.push_back(k);
    }
  });
  for (uint8_t b : bad)
    if (b) throw std::runtime_error("tree hierarchy could not cover a coarse face with fine faces");
  // every chunk's coarse faces and entries at their offsets
  std::vector<long> face_off(nt + 1, 0), ent_off(nt + 1, 0);
  for (int w = 0; w < nt; ++w) {
    face_off[w + 1] = face_off[w] + long(cnt_t[w].size());
    ent_off[w + 1] = ent_off[w] + long(idx_t[w].size());
  }
  C.rptr.assign(CF + 1, 0);
  C.ridx.resize(ent_off[nt]);
  C.rw.resize(ent_off[nt]);
  parallel_tasks(nt, [&](long w) {
    long cum = ent_off[w];
    for (size_t x = 0; x < cnt_t[w].size(); ++x) {
      cum += cnt_t[w][x];
      C.rptr[face_off[w] + long(x) + 1] = int(cum);
    }
    std::copy(idx_t[w].begin(), idx_t[w].end(), C.ridx.begin() + ent_off[w]);
    std::copy(w_t[w].begin(), w_t[w].end(), C.rw.begin() + ent_off[w]);
  });
}

// Boundary-coefficient restriction from the finer level
// (`_face_coefficient_restriction(...; boundary = true)`,
// tree_coarsening.jl:347-398, with the boundary-face lattice geometry of
// mesh.jl:476-485): every coarse boundary face covers fine boundary faces on
// its line (axis, normal coordinate on the common lattice), weighted by the
// overlap fraction of its length.
void boundary_restriction_map(LevelData& C, const LevelData& Fl) {
  const TreeMesh& fm = *Fl.mesh;
  const TreeMesh& cm = *C.mesh;
  long lat[2];
  for (int a = 0; a < 2; ++a)
    lat[a] = std::max(long(fm.base[a]) << fm.max_leaf, long(cm.base[a]) << cm.max_leaf);
  auto geom = [&](const TreeMesh& M, int f, long& n, long& lo, long& hi) {
    const int nb = M.nb(), bx = M.bs[0], by = M.bs[1];
    const int c = M.bcell[f];
    const BlockKey& k = M.blocks[c / nb];
    const long sc[2] = {lat[0] / (long(M.base[0]) << k.level), lat[1] / (long(M.base[1]) << k.level)};
    const long low[2] = {(long(k.i) * bx + (c % nb) % bx) * sc[0], (long(k.j) * by + (c % nb) / bx) * sc[1]};
    const int a = M.baxis[f], t = 1 - a;
    n = low[a] + (M.bside[f] > 0 ? sc[a] : 0);
    lo = low[t];
    hi = low[t] + sc[t];
  };
  std::map<std::pair<int, long>, std::vector<std::array<long, 3>>> lines;
  for (int f = 0; f < int(fm.bcell.size()); ++f) {
    long n, lo, hi;
    geom(fm, f, n, lo, hi);
    lines[{int(fm.baxis[f]), n}].push_back({lo, hi, long(f)});
  }
  for (auto& kv : lines)
    std::stable_sort(kv.second.begin(), kv.second.end(), [](const auto& x, const auto& y) { return x[0] < y[0]; });
  C.brptr.assign(1, 0);
  C.bridx.clear();
  C.brw.clear();
  for (int f = 0; f < int(cm.bcell.size()); ++f) {
    long n, lo, hi;
    geom(cm, f, n, lo, hi);
    long covered = 0;
    auto it = lines.find({int(cm.baxis[f]), n});
    if (it != lines.end())
      for (const auto& seg : it->second) {
        if (seg[1] <= lo) continue;
        if (seg[0] >= hi) break;
        const long ov = std::min(hi, seg[1]) - std::max(lo, seg[0]);
        if (ov > 0) {
          C.bridx.push_back(int(seg[2]));
          C.brw.push_back(double(ov) / double(hi - lo));
          covered += ov;
        }
      }
    if (covered != hi - lo) throw std::runtime_error("tree hierarchy could not cover a coarse face with fine faces");
    C.brptr.push_back(int(C.bridx.size()));
  }
}

// Open-face conductances |f| 2 kb / h (the boundary diagonal terms of
// `_elliptic_boundary_data!`, elliptic.jl:133-152: factor 2 kb / h, then
// |f| * aperture 1 * factor), the irregular rows' boundary diagonals (sums
// in face order) and the level's singularity (no open face conducts).
void boundary_values(LevelData& L) {
  const TreeMesh& M = *L.mesh;
  const int nb = M.nb();
  L.bw.resize(L.bopen.size());
  bool gauged = false;
  for (size_t k = 0; k < L.bopen.size(); ++k) {
    const int f = L.bopen[k];
    const double h = M.bspacing[2 * (M.bcell[f] / nb) + M.baxis[f]];
    const double factor = 2 * L.kb[f] / h;
    L.bw[k] = M.barea[f] * 1.0 * factor;
    gauged = gauged || L.bw[k] != 0.0;
  }
  const int nrows = int(L.irr_local.size());
  L.irr_bd.assign(nrows, 0.0);
  for (int r = 0; r < nrows; ++r)
    for (int q = L.irr_bptr[r]; q < L.irr_bptr[r + 1]; ++q) L.irr_bd[r] += L.bw[L.irr_bface[q]];
  L.singular = !gauged;
}

// The level's pressure components (see LevelData): with `positive` every
// face coefficient of the level is positive, so the level is one connected
// component, gauged exactly when an open face conducts; otherwise the
// components are labelled over the faces with a nonzero coefficient (root =
// smallest cell, as the reference's depth-first labelling from increasing
// roots gives) and the chunked-sum plan is built.
void level_components(LevelData& L, bool positive) {
  const TreeMesh& M = *L.mesh;
  const int N = L.ncells, nb = M.nb();
  std::vector<uint8_t> anchored;
  bool any_anchor = false;
  for (size_t k = 0; k < L.bopen.size() && k < L.bw.size(); ++k)
    if (L.bw[k] != 0.0) any_anchor = true;
  L.pair_chunk.clear();
  L.pair_slot.clear();
  L.slot_pair_ptr.clear();
  L.slot_pairs.clear();
  if (positive) {
    L.connected = true;
    L.ncomp = any_anchor ? 0 : 1;
    L.comp_slot.clear();
    L.comp_root.assign(L.ncomp, 0);
    L.comp_vol.assign(L.ncomp, L.total_volume);
    L.singular = L.ncomp > 0;
    return;
  }
  // union-find, the smaller root wins
  std::vector<int> parent(N);
  for (int c = 0; c < N; ++c) parent[c] = c;
  auto find = [&](int c) {
    while (parent[c] != c) {
      parent[c] = parent[parent[c]];
      c = parent[c];
    }
    return c;
  };
  for (int f = 0; f < M.nfaces(); ++f) {
    if (L.kappa[f] == 0.0) continue;
    int a = find(M.fminus[f]), b = find(M.fplus[f]);
    if (a == b) continue;
    if (a > b) std::swap(a, b);
    parent[b] = a;
  }
  std::vector<int> root(N);
  for (int c = 0; c < N; ++c) root[c] = find(c);
  anchored.assign(N, 0);
  for (size_t k = 0; k < L.bopen.size() && k < L.bw.size(); ++k)
    if (L.bw[k] != 0.0) anchored[root[M.bcell[L.bopen[k]]]] = 1;
  int ncomponents = 0;
  std::vector<int> slot_of_root(N, 0);
  L.comp_root.clear();
  for (int c = 0; c < N; ++c)
    if (root[c] == c) {
      ++ncomponents;
      if (!anchored[c]) {
        L.comp_root.push_back(c);
        slot_of_root[c] = int(L.comp_root.size());
      }
    }
  L.ncomp = int(L.comp_root.size());
  L.connected = ncomponents == 1;
  L.singular = L.ncomp > 0;
  L.comp_slot.assign(N, 0);
  L.comp_vol.assign(L.ncomp, 0.0);
  for (int c = 0; c < N; ++c) {
    const int q = slot_of_root[root[c]];
    L.comp_slot[c] = q;
    if (q) L.comp_vol[q - 1] += L.volume[c / nb];
  }
  if (L.ncomp == 0) return;
  // chunked-sum plan: the slots present in every chunk, increasing
  const int nchunks = (N + GAUGE_CHUNK - 1) / GAUGE_CHUNK;
  std::vector<int> seen(L.ncomp + 1, -1), present;
  for (int ch = 0; ch < nchunks; ++ch) {
    present.clear();
    for (int c = ch * GAUGE_CHUNK; c < std::min(N, (ch + 1) * GAUGE_CHUNK); ++c) {
      const int q = L.comp_slot[c];
      if (q && seen[q] != ch) {
        seen[q] = ch;
        present.push_back(q);
      }
    }
    std::sort(present.begin(), present.end());
    for (int q : present) {
      L.pair_chunk.push_back(ch);
      L.pair_slot.push_back(q);
    }
  }
  L.slot_pair_ptr.assign(L.ncomp + 1, 0);
  for (int q : L.pair_slot) L.slot_pair_ptr[q]++;
  for (int q = 0; q < L.ncomp; ++q) L.slot_pair_ptr[q + 1] += L.slot_pair_ptr[q];
  L.slot_pairs.assign(L.pair_slot.size(), 0);
  std::vector<int> at(L.slot_pair_ptr.begin(), L.slot_pair_ptr.end() - 1);
  for (size_t p = 0; p < L.pair_slot.size(); ++p) L.slot_pairs[at[L.pair_slot[p] - 1]++] = int(p);
}

// The hierarchy's gauge flags from its levels' components.
void finish_gauge(Hierarchy& H) {
  H.singular = H.levels[0]->singular;
  H.general_gauge = false;
  for (const auto& L : H.levels) H.general_gauge = H.general_gauge || !L->simple_gauge();
}

void transfer_descriptors(LevelData& Fl, const LevelData& C, bool agglomerated) {
  const TreeMesh& fm = *Fl.mesh;
  const TreeMesh& cm = *C.mesh;
  Fl.xfer.assign(2 * Fl.nblocks, -1);
  for (int b = 0; b < Fl.nblocks; ++b) {
    const BlockKey& k = fm.blocks[b];
    if (agglomerated) {
      auto it = cm.block_of.find(key64(0, k.i >> 1, k.j >> 1));
      if (it == cm.block_of.end()) throw std::runtime_error("agglomerated root block missing");
      Fl.xfer[2 * b] = it->second;
      Fl.xfer[2 * b + 1] = (k.i & 1) + 2 * (k.j & 1);
      continue;
    }
    auto it = cm.block_of.find(key64(k.level, k.i, k.j));
    if (it != cm.block_of.end()) {
      Fl.xfer[2 * b] = it->second;
      Fl.xfer[2 * b + 1] = -1;
      continue;
    }
    if (k.level == 0) throw std::runtime_error("coarse mesh lost a root block");
    auto jt = cm.block_of.find(key64(k.level - 1, k.i >> 1, k.j >> 1));
    if (jt == cm.block_of.end()) throw std::runtime_error("coarse mesh is not a pruning of the fine mesh");
    Fl.xfer[2 * b] = jt->second;
    Fl.xfer[2 * b + 1] = (k.i & 1) + 2 * (k.j & 1);
  }
}

// Dense coarsest operator into H.coarse_aug, row-major N x N: A augmented
// with one volume constraint per ungauged component, [A C; C^T 0]
// (N = n + ncomp; C's column of a component holds its cells' volumes), or A
// itself when every component is gauged (geometric_preconditioner.jl:144-169,
// 415-430).
void coarse_matrix(Hierarchy& H) {
  LevelData& L = *H.levels.back();
  const TreeMesh& M = *L.mesh;
  const int n = L.ncells, nb = M.nb(), bx = M.bs[0], by = M.bs[1];
  const int N = n + L.ncomp;
  // The coarsest level is solved with a dense inverse. Tree coarsening ends
  // at coarse_size cells when the root layout agglomerates down to one
  // block; an odd root layout can stop far above that.
  if (n > COARSE_MAX)
    throw std::invalid_argument("coarsest multigrid level has " + std::to_string(n) +
                                " cells; the dense coarse solve supports at most " + std::to_string(COARSE_MAX) +
                                " (use a root-block layout that halves down to a few blocks)");
  H.coarse_n = n;
  H.coarse_N = N;
  std::vector<double>& a = H.coarse_aug;
  a.assign(size_t(N) * N, 0.0);
  for (int b = 0; b < L.nblocks; ++b)
    for (int q = 0; q < nb; ++q) {
      const int c = b * nb + q;
      if (L.cell_irr[c] >= 0) {
        const int r = L.cell_irr[c];
        for (int e = L.irr_ptr[r]; e < L.irr_ptr[r + 1]; ++e) a[size_t(c) * N + L.irr_col[e]] += L.irr_val[e];
      } else {
        const int i = q % bx, j = q / bx;
        const double* cx = &L.cx[size_t(b) * (bx + 1) * by];
        const double* cy = &L.cy[size_t(b) * bx * (by + 1)];
        const double w[4] = {cx[j * (bx + 1) + i], cx[j * (bx + 1) + i + 1], cy[j * bx + i], cy[(j + 1) * bx + i]};
        // neighbours across the four sides
        int nbc[4] = {-1, -1, -1, -1};
        auto cell_at = [&](int bb, int ii, int jj) { return bb * nb + jj * bx + ii; };
        nbc[0] = i > 0 ? cell_at(b, i - 1, j) : (L.nbr[4 * b] >= 0 ? cell_at(L.nbr[4 * b], bx - 1, j) : -1);
        nbc[1] = i < bx - 1 ? cell_at(b, i + 1, j) : (L.nbr[4 * b + 1] >= 0 ? cell_at(L.nbr[4 * b + 1], 0, j) : -1);
        nbc[2] = j > 0 ? cell_at(b, i, j - 1) : (L.nbr[4 * b + 2] >= 0 ? cell_at(L.nbr[4 * b + 2], i, by - 1) : -1);
        nbc[3] = j < by - 1 ? cell_at(b, i, j + 1) : (L.nbr[4 * b + 3] >= 0 ? cell_at(L.nbr[4 * b + 3], i, 0) : -1);
        // a regular cell's slot without a neighbour is an open boundary face:
        // its conductance is the boundary diagonal term
        for (int s = 0; s < 4; ++s) {
          if (w[s] == 0.0) continue;
          a[size_t(c) * N + c] += w[s];
          if (nbc[s] >= 0) a[size_t(c) * N + nbc[s]] -= w[s];
        }
      }
      const int slot = L.comp_slot.empty() ? L.ncomp : L.comp_slot[c];
      if (slot > 0) {
        const double V = L.volume[b];
        a[size_t(c) * N + n + slot - 1] = V;
        a[size_t(n + slot - 1) * N + c] = V;
      }
    }
}

}  // namespace

bool same_tree_topology(const TreeMesh& a, const TreeMesh& b) {
  for (int q = 0; q < 2; ++q)
    if (a.base[q] != b.base[q] || a.bs[q] != b.bs[q] || a.periodic[q] != b.periodic[q] ||
        !(a.origin[q] == b.origin[q]) || !(a.extent[q] == b.extent[q]))
      return false;
  if (a.blocks.size() != b.blocks.size() || a.nfaces() != b.nfaces() || a.bcell.size() != b.bcell.size())
    return false;
  for (size_t k = 0; k < a.blocks.size(); ++k)
    if (a.blocks[k].level != b.blocks[k].level || a.blocks[k].i != b.blocks[k].i || a.blocks[k].j != b.blocks[k].j)
      return false;
  return true;
}

std::unique_ptr<Hierarchy> build_hierarchy(std::unique_ptr<TreeMesh> fine, const MultigridConfig& cfg,
                                           Hierarchy* previous, const bool* open) {
  if (cfg.coarse_size <= 0) throw std::invalid_argument("coarse_size must be positive");
  if (cfg.pre < 0 || cfg.post < 0) throw std::invalid_argument("sweep counts must be nonnegative");
  if (cfg.max_levels <= 0) throw std::invalid_argument("max_levels must be positive");
  bool sides[4] = {false, false, false, false};
  for (int q = 0; q < 4 && open; ++q) sides[q] = open[q];
  auto H = std::make_unique<Hierarchy>();
  auto L0 = std::make_unique<LevelData>();
  L0->owned = std::move(fine);
  L0->mesh = L0->owned.get();
  std::copy(sides, sides + 4, L0->open);
  level_topology(*L0);
  H->levels.push_back(std::move(L0));
  std::vector<bool> agg;
  int active = H->levels[0]->mesh->max_leaf;
  for (;;) {
    const TreeMesh& cur = *H->levels.back()->mesh;
    if (!(cur.ncells() > cfg.coarse_size) || int(H->levels.size()) >= cfg.max_levels) break;
    const bool is_agg = active == 0;
    PhaseTimer pt;
    std::unique_ptr<TreeMesh> cm = is_agg ? agglomerated_mesh(cur) : pruned_mesh(cur);
    pt.mark("coarse mesh", cm ? long(cm->ncells()) : 0L);
    if (!cm || cm->ncells() >= cur.ncells()) break;
    active = cm->max_leaf;
    if (previous) {
      int match = -1;
      for (int l = 0; l < int(previous->levels.size()) && match < 0; ++l)
        if (previous->levels[l] && same_tree_topology(*previous->levels[l]->mesh, *cm) &&
            std::equal(sides, sides + 4, previous->levels[l]->open))
          match = l;
      const int suffix = match < 0 ? 0 : int(previous->levels.size()) - match;
      if (match >= 0 && int(H->levels.size()) + suffix <= cfg.max_levels) {
        // the matching level's restriction and the transfer into it come
        // from the new finer level; everything coarser is unchanged
        H->reused_new = int(H->levels.size());
        H->reused_old = match;
        std::unique_ptr<LevelData> L = std::move(previous->levels[match]);
        restriction_map(*L, *H->levels.back());
        boundary_restriction_map(*L, *H->levels.back());
        transfer_descriptors(*H->levels.back(), *L, is_agg);
        H->levels.push_back(std::move(L));
        for (int l = match + 1; l < int(previous->levels.size()); ++l)
          H->levels.push_back(std::move(previous->levels[l]));
        previous->levels.resize(match);
        break;
      }
    }
    auto L = std::make_unique<LevelData>();
    L->owned = std::move(cm);
    L->mesh = L->owned.get();
    std::copy(sides, sides + 4, L->open);
    level_topology(*L);
    pt.t = std::chrono::steady_clock::now();
    restriction_map(*L, *H->levels.back());
    boundary_restriction_map(*L, *H->levels.back());
    pt.mark("restriction map");
    transfer_descriptors(*H->levels.back(), *L, is_agg);
    pt.mark("transfer descriptors");
    H->levels.push_back(std::move(L));
  }
  const TreeMesh& M0 = *H->levels[0]->mesh;
  H->total_volume = 0;
  for (int b = 0; b < H->levels[0]->nblocks; ++b) H->total_volume += H->levels[0]->volume[b] * M0.nb();
  return H;
}

void set_boundary_coefficients(Hierarchy& H, const double* kb) {
  LevelData& L0 = *H.levels[0];
  const int BF = int(L0.mesh->bcell.size());
  for (int f = 0; f < BF; ++f)
    if (!(std::isfinite(kb[f]) && kb[f] >= 0))
      throw std::invalid_argument("boundary coefficients must be finite and nonnegative");
  L0.kb.assign(kb, kb + BF);
  for (size_t l = 0; l < H.levels.size(); ++l) {
    LevelData& L = *H.levels[l];
    if (l > 0) {
      // _restrict_face_coefficients! (tree_coarsening.jl:538-555)
      const LevelData& P = *H.levels[l - 1];
      L.kb.resize(L.mesh->bcell.size());
      for (size_t f = 0; f < L.kb.size(); ++f) {
        double v = 0;
        for (int q = L.brptr[f]; q < L.brptr[f + 1]; ++q) v += L.brw[q] * P.kb[L.bridx[q]];
        L.kb[f] = v;
      }
    }
    boundary_values(L);
  }
}

std::vector<double> boundary_material(const TreeMesh& m, const bool open[4], const double* kappa) {
  const int BF = int(m.bcell.size());
  std::vector<double> values(BF, 0.0), weights(BF, 0.0);
  auto is_open = [&](int f) { return open[2 * m.baxis[f] + (m.bside[f] > 0 ? 1 : 0)]; };
  // (cell, axis, side) of every open face
  std::unordered_map<uint64_t, int> index;
  auto key = [](int cell, int axis, int side) {
    return (uint64_t(uint32_t(cell)) << 2) | (uint64_t(axis) << 1) | uint64_t(side > 0 ? 1 : 0);
  };
  for (int f = 0; f < BF; ++f)
    if (is_open(f)) index[key(m.bcell[f], m.baxis[f], m.bside[f])] = f;
  if (index.empty()) return values;
  for (int f = 0; f < m.nfaces(); ++f) {
    const int ends[2][2] = {{m.fminus[f], -1}, {m.fplus[f], 1}};
    for (const auto& e : ends) {
      auto it = index.find(key(e[0], m.faxis[f], e[1]));
      if (it == index.end()) continue;
      values[it->second] += m.farea[f] * kappa[f];
      weights[it->second] += m.farea[f];
    }
  }
  for (int f = 0; f < BF; ++f) {
    if (!is_open(f)) continue;
    if (!(weights[f] > 0))
      throw std::invalid_argument(
          "cannot extend an interior coefficient to an open face without an opposite interior face; supply "
          "boundary_face_coefficient");
    values[f] /= weights[f];
  }
  return values;
}

void refresh_coarsest(Hierarchy& H, const double* kappa_coarsest) {
  LevelData& L = *H.levels.back();
  L.kappa.assign(kappa_coarsest, kappa_coarsest + L.mesh->nfaces());
  level_values(L);
  coarse_matrix(H);
}

void refresh_coefficients(Hierarchy& H, const double* kappa) {
  LevelData& L0 = *H.levels[0];
  const int F = L0.mesh->nfaces();
  bool positive = true;
  for (int f = 0; f < F; ++f) {
    if (!std::isfinite(kappa[f])) throw std::invalid_argument("face coefficients must be finite");
    if (!(kappa[f] >= 0)) throw std::invalid_argument("face coefficients must be nonnegative");
    positive = positive && kappa[f] > 0;
  }
  if (L0.any_open() && L0.kb.size() != L0.mesh->bcell.size())
    throw std::logic_error("open pressure sides need boundary coefficients before the coefficient refresh");
  L0.kappa.assign(kappa, kappa + F);
  for (size_t l = 0; l < H.levels.size(); ++l) {
    LevelData& L = *H.levels[l];
    if (l > 0) {
      const LevelData& P = *H.levels[l - 1];
      L.kappa.resize(L.mesh->nfaces());
      parallel_for(L.mesh->nfaces(), [&](long lo, long hi, int) {
        for (long f = lo; f < hi; ++f) {
          double s = 0;
          for (int q = L.rptr[f]; q < L.rptr[f + 1]; ++q) s += L.rw[q] * P.kappa[L.ridx[q]];
          L.kappa[f] = s;
        }
      });
    }
    level_values(L);
    level_components(L, positive);
  }
  finish_gauge(H);
  coarse_matrix(H);
}

void connected_components(Hierarchy& H) {
  for (auto& L : H.levels) level_components(*L, true);
  finish_gauge(H);
}

// Halo schedule of `_build_halo_schedule` (src/mesh/topology/mesh.jl:944-1037):
// a halo position of a block at level l covers a square of 2^(L-l) finest
// lattice cells (L the finest leaf level); its value is the average over the
// four quarter points of that square of the leaf cells containing them
// (same-level copy, coarse-to-fine injection, fine-to-coarse average; the
// periodic axes wrap). A position whose centre lies outside a non-periodic
// side is a physical entry fed by the leaf cell at the clamped centre.
HaloSchedule halo_schedule(const TreeMesh& m, int h) {
  if (h < 1) throw std::invalid_argument("halo must be positive");
  const int nx = m.bs[0], ny = m.bs[1], nb = nx * ny, W = nx + 2 * h, H = ny + 2 * h;
  int L = 0;
  for (const BlockKey& k : m.blocks) L = std::max(L, k.level);
  const long qp[2] = {4L * m.layout[0] * nx << L, 4L * m.layout[1] * ny << L};  // quarter-point periods
  // leaf cell of finest-lattice cell (px, py): the ancestor keys at every
  // level, finest first (a block's neighbours are within one level)
  auto leaf_cell = [&](long px, long py) {
    for (int l = L; l >= 0; --l) {
      const long cx = px >> (L - l), cy = py >> (L - l);
      auto it = m.block_of.find(key64(l, int(cx / nx), int(cy / ny)));
      if (it != m.block_of.end()) return it->second * nb + int(cy % ny) * nx + int(cx % nx);
    }
    throw std::logic_error("halo point outside the mesh");
  };
  auto fold = [&](long v, int axis, bool& out) {
    const long p = qp[axis];
    if (m.periodic[axis]) {
      out = false;
      v %= p;
      return v < 0 ? v + p : v;
    }
    out = v < 0 || v >= p;
    return v < 0 ? 0 : (v >= p ? p - 1 : v);
  };
  HaloSchedule S;
  S.halo = h;
  S.ptr.push_back(0);
  for (int b = 0; b < int(m.blocks.size()); ++b) {
    const BlockKey& k = m.blocks[b];
    const long sc = 1L << (L - k.level);
    for (int j = -h; j < ny + h; ++j)
      for (int i = -h; i < nx + h; ++i) {
        if (i >= 0 && i < nx && j >= 0 && j < ny) continue;
        S.dest.push_back(b * W * H + (j + h) * W + (i + h));
        const long x0 = 4 * (long(k.i) * nx + i) * sc, y0 = 4 * (long(k.j) * ny + j) * sc;
        bool ox, oy;
        const long cx = fold(x0 + 2 * sc, 0, ox), cy = fold(y0 + 2 * sc, 1, oy);
        if (ox || oy) {
          S.donor.push_back(leaf_cell(cx / 4, cy / 4));
          S.weight.push_back(1.0);
          S.phys.push_back(S.nphysical++);
        } else {
          // quarter points (1,1), (3,1), (1,3), (3,3); repeated cells merge
          // (first occurrence order) with weight multiplicity / 4
          int cells[4], mult[4] = {0, 0, 0, 0}, n = 0;
          for (int q = 0; q < 4; ++q) {
            bool o;
            const long px = fold(x0 + (q & 1 ? 3 : 1) * sc, 0, o), py = fold(y0 + (q & 2 ? 3 : 1) * sc, 1, o);
            const int c = leaf_cell(px / 4, py / 4);
            int at = 0;
            while (at < n && cells[at] != c) ++at;
            if (at == n) cells[n++] = c;
            ++mult[at];
          }
          for (int q = 0; q < n; ++q) {
            S.donor.push_back(cells[q]);
            S.weight.push_back(mult[q] / 4.0);
          }
          S.phys.push_back(-1);
        }
        S.ptr.push_back(int(S.donor.size()));
      }
  }
  return S;
}

}  // namespace tpamr
