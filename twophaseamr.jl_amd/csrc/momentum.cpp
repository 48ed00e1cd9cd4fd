// Host side of the one-velocity momentum operators on a uniform, fully
// periodic block mesh: the structured component grids, the transport
// stages, the viscous operator's coefficients and the Krylov driver of the
// implicit-midpoint viscous solve (the reference's _solve_krylov_with_stats!
// with relative_to_rhs, elliptic_solver.jl:550-627, around CG preconditioned
// by one V-cycle of the staggered tree multigrid,
// TreeStaggeredGeometricMultigridPreconditioner).
#include "momentum.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "staggered.hpp"

namespace tpamr {
namespace {

#define MCHECK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw MomentumDeviceError(std::string(#x) + ": " + hipGetErrorString(e_)); \
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
    if (n) MCHECK(hipMalloc(&p, n * sizeof(T)));
  }
};

const double EPS = std::numeric_limits<double>::epsilon();

}  // namespace

// One level of the viscous solve's staggered multigrid: an nx x ny periodic
// component grid, its operator, Jacobi weight and vectors.
struct StagLevel {
  int nx = 0, ny = 0;
  double hx = 0, hy = 0;
  StagCoef k{};
  double w = 1;
  Buf<double> x, xs, b, r;
};

namespace {
// The first column g = A^-1 e_0 of the inverse of a level's operator A =
// md I + mx (S_x + S_x^T) + my (S_y + S_y^T) on its periodic grid (S the
// periodic shifts; the coarse LU of the reference,
// implicit_diffusion.jl:343-347, solves the same system exactly). A is
// circulant with eigenvalues lambda = md + 2 mx cos(2 pi kx / nx) +
// 2 my cos(2 pi ky / ny), even in kx and ky, so
//   g[n, m] = 1 / (nx ny) sum_kx sum_ky cos(2 pi kx m / nx) cos(2 pi ky n / ny) / lambda
// (row-major (n, m) = (j, i)), summed as two separable products. Checked on
// the host: max |A g - e_0| must be round-off.
std::vector<double> circulant_inverse_column(const StagLevel& C) {
  const int nx = C.nx, ny = C.ny;
  const double two_pi = 2.0 * std::acos(-1.0);
  // cos(2 pi q / n) for q = 0..n-1; cos(2 pi k m / n) is entry (k m) mod n
  auto cos_table = [&](int n) {
    std::vector<double> t(n);
    for (int q = 0; q < n; ++q) t[q] = std::cos(two_pi * double(q) / n);
    return t;
  };
  const std::vector<double> cx = cos_table(nx), cy = cos_table(ny);
  std::vector<double> lam(size_t(nx) * ny);
  for (int ky = 0; ky < ny; ++ky)
    for (int kx = 0; kx < nx; ++kx) {
      const double l = C.k.md + 2.0 * C.k.mx * cx[kx] + 2.0 * C.k.my * cy[ky];
      if (!(l > 0)) throw std::invalid_argument("coarsest viscous operator is not positive definite");
      lam[size_t(ky) * nx + kx] = 1.0 / l;
    }
  // t[kx, n] = sum_ky cos(2 pi ky n / ny) / lambda(kx, ky)
  std::vector<double> t(size_t(nx) * ny, 0.0), g(size_t(nx) * ny, 0.0);
  for (int ky = 0; ky < ny; ++ky)
    for (int n = 0; n < ny; ++n) {
      const double c = cy[(int64_t(ky) * n) % ny];
      for (int kx = 0; kx < nx; ++kx) t[size_t(kx) * ny + n] += c * lam[size_t(ky) * nx + kx];
    }
  const double scale = 1.0 / (double(nx) * ny);
  for (int kx = 0; kx < nx; ++kx)
    for (int n = 0; n < ny; ++n) {
      const double tv = t[size_t(kx) * ny + n] * scale;
      for (int m = 0; m < nx; ++m) g[size_t(n) * nx + m] += cx[(int64_t(kx) * m) % nx] * tv;
    }
  double err = 0.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      auto at = [&](int ii, int jj) { return g[size_t((jj + ny) % ny) * nx + (ii + nx) % nx]; };
      const double ag = C.k.md * at(i, j) + C.k.mx * (at(i - 1, j) + at(i + 1, j)) +
                        C.k.my * (at(i, j - 1) + at(i, j + 1));
      err = std::max(err, std::fabs(ag - (i == 0 && j == 0 ? 1.0 : 0.0)));
    }
  if (!(err <= 1e-10)) throw std::runtime_error("circulant coarse inverse: residual " + std::to_string(err));
  return g;
}
}  // namespace

struct Momentum::Impl {
  hipStream_t s = nullptr;
  StagGrid g{};
  StagCoef k{};
  double nu = 0;
  int scheme = STAG_LW;
  bool forward = true;
  int n = 0, nfaces = 0;
  MultigridConfig mg;
  Buf<int> fx, fy;
  Buf<double> X, Y, U, V, X2, Y2;          // transported grids, advecting grids, stage outputs
  Buf<double> cx, cy, ox, oy;              // implicit substep: convected, old
  Buf<double> b, r, z, p, Ap;              // viscous Krylov vectors (one component)
  Buf<double> part, scal;
  double host_scal[12] = {};
  // staggered multigrid: levels (level 0 shares the Krylov grid size) and
  // the coarsest operator's inverse for the current shift (dense, or the
  // first column of the circulant inverse above STAG_DENSE_MAX)
  std::vector<std::unique_ptr<StagLevel>> lv;
  Buf<double> coarse_inv;
  double coarse_half = -1;
  int comp = 0;
  bool use_mg = true;

  // level operators for the shift h = dt nu / 2: the finest as the Krylov
  // operator V (I - h L), the coarser ones h K + V
  // (_refresh_staggered_component!, implicit_diffusion.jl:213-261); Jacobi
  // weights (_jacobi_weight!, hierarchy.jl:8-30); dense coarse inverse
  void set_shift(double half) {
    for (size_t l = 0; l < lv.size(); ++l) {
      StagLevel& L = *lv[l];
      const double hx = L.hx, hy = L.hy, vol = hx * hy;
      const double kx = ((hy * hx) * (1.0 / hx)) * (1.0 / hx);
      const double ky = ((hx * hy) * (1.0 / hy)) * (1.0 / hy);
      const double sd = ((kx + kx) + ky) + ky;
      if (l == 0) {
        const double iv = 1.0 / vol;
        const double ld = -iv * sd, lx = iv * kx, ly = iv * ky;
        L.k.md = vol * (1.0 - half * ld);
        L.k.mx = vol * (0.0 - half * lx);
        L.k.my = vol * (0.0 - half * ly);
      } else {
        L.k.md = half * sd + vol;
        L.k.mx = half * -kx;
        L.k.my = half * -ky;
      }
      L.k.mdinv = 1.0 / L.k.md;
      const double rowsum = (((std::fabs(L.k.md) + std::fabs(L.k.mx)) + std::fabs(L.k.mx)) + std::fabs(L.k.my)) +
                            std::fabs(L.k.my);
      const double bound = rowsum / L.k.md;
      L.w = std::min(1.0, 4.0 / (3.0 * bound));
    }
    k.md = lv[0]->k.md;
    k.mx = lv[0]->k.mx;
    k.my = lv[0]->k.my;
    k.mdinv = lv[0]->k.mdinv;
    coarse_half = half;
    if (!use_mg) return;
    const StagLevel& C = *lv.back();
    const int nc = C.nx * C.ny;
    if (nc > STAG_DENSE_MAX) {
      // the first column of the inverse of the circulant coarsest operator;
      // the coarse solve is the periodic convolution with it
      const std::vector<double> col = circulant_inverse_column(C);
      coarse_inv.alloc(col.size());
      MCHECK(hipMemcpyAsync(coarse_inv.p, col.data(), col.size() * sizeof(double), hipMemcpyHostToDevice, s));
      MCHECK(hipStreamSynchronize(s));
      return;
    }
    // dense inverse of the coarsest operator (Gauss-Jordan, partial pivoting)
    std::vector<double> a(size_t(nc) * nc, 0.0), inv(size_t(nc) * nc, 0.0);
    for (int j = 0; j < C.ny; ++j)
      for (int i = 0; i < C.nx; ++i) {
        const int r0 = j * C.nx + i;
        auto add = [&](int ii, int jj, double v) {
          const int c0 = ((jj + C.ny) % C.ny) * C.nx + ((ii + C.nx) % C.nx);
          a[size_t(r0) * nc + c0] += v;
        };
        add(i, j, C.k.md);
        add(i - 1, j, C.k.mx);
        add(i + 1, j, C.k.mx);
        add(i, j - 1, C.k.my);
        add(i, j + 1, C.k.my);
      }
    for (int q = 0; q < nc; ++q) inv[size_t(q) * nc + q] = 1.0;
    for (int col = 0; col < nc; ++col) {
      int piv = col;
      for (int q = col + 1; q < nc; ++q)
        if (std::fabs(a[size_t(q) * nc + col]) > std::fabs(a[size_t(piv) * nc + col])) piv = q;
      if (a[size_t(piv) * nc + col] == 0.0) throw std::invalid_argument("singular coarsest viscous operator");
      if (piv != col)
        for (int c2 = 0; c2 < nc; ++c2) {
          std::swap(a[size_t(piv) * nc + c2], a[size_t(col) * nc + c2]);
          std::swap(inv[size_t(piv) * nc + c2], inv[size_t(col) * nc + c2]);
        }
      const double d = 1.0 / a[size_t(col) * nc + col];
      for (int c2 = 0; c2 < nc; ++c2) {
        a[size_t(col) * nc + c2] *= d;
        inv[size_t(col) * nc + c2] *= d;
      }
      for (int q = 0; q < nc; ++q) {
        if (q == col) continue;
        const double f = a[size_t(q) * nc + col];
        if (f == 0.0) continue;
        for (int c2 = 0; c2 < nc; ++c2) {
          a[size_t(q) * nc + c2] -= f * a[size_t(col) * nc + c2];
          inv[size_t(q) * nc + c2] -= f * inv[size_t(col) * nc + c2];
        }
      }
    }
    coarse_inv.alloc(inv.size());
    MCHECK(hipMemcpyAsync(coarse_inv.p, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice, s));
    MCHECK(hipStreamSynchronize(s));
  }

  // z = V-cycle(rhs) of component `comp` (_vcycle!, cycles.jl:40-77): every
  // level ping-pongs between its own two buffers; the finest result is
  // copied to z
  void vcycle(const double* rhs, double* z) {
    const int nl = int(lv.size());
    if (nl == 1) {
      coarse_solve(rhs, z);
      return;
    }
    std::vector<double*> cur(nl), oth(nl);
    std::vector<const double*> bl(nl);
    for (int l = 0; l < nl; ++l) {
      cur[l] = lv[l]->x.p;
      oth[l] = lv[l]->xs.p;
      bl[l] = l ? lv[l]->b.p : rhs;
    }
    for (int l = 0; l + 1 < nl; ++l) {
      StagLevel& L = *lv[l];
      // the first sweep starts from zero (x = 0 + w D^-1 (b - 0)); with two
      // or more pre-sweeps the first two run as one pass
      int q = 0;
      if (mg.pre >= 2) {
        launch_st_mg_smooth2z(s, L.nx, L.ny, L.k, L.w, bl[l], oth[l]);
        std::swap(cur[l], oth[l]);
        q = 2;
      }
      for (; q < mg.pre; ++q) {
        launch_st_mg_smooth(s, L.nx, L.ny, L.k, L.w, q == 0 ? nullptr : cur[l], bl[l], oth[l]);
        std::swap(cur[l], oth[l]);
      }
      if (mg.pre == 0) launch_st_mg_smooth(s, L.nx, L.ny, L.k, 0.0, nullptr, bl[l], cur[l]);
      launch_st_mg_restrict_res(s, lv[l + 1]->nx, lv[l + 1]->ny, comp, L.k, bl[l], cur[l], lv[l + 1]->b.p);
    }
    coarse_solve(bl[nl - 1], cur[nl - 1]);
    bool in_z = false;
    for (int l = nl - 2; l >= 0; --l) {
      StagLevel& L = *lv[l];
      const int ncx = lv[l + 1]->nx, ncy = lv[l + 1]->ny;
      if (mg.post == 0) {
        launch_st_mg_prolong(s, ncx, ncy, comp, cur[l + 1], cur[l]);
        continue;
      }
      // the finest level's last sweep writes z itself
      for (int q = 0; q < mg.post; ++q) {
        double* out = (l == 0 && q == mg.post - 1) ? z : oth[l];
        if (q == 0)
          launch_st_mg_prolong_smooth(s, ncx, ncy, comp, L.k, L.w, cur[l + 1], cur[l], bl[l], out);
        else
          launch_st_mg_smooth(s, L.nx, L.ny, L.k, L.w, cur[l], bl[l], out);
        if (out == z) {
          in_z = true;
        } else {
          std::swap(cur[l], oth[l]);
        }
      }
    }
    if (!in_z) launch_copy_grid(cur[0], z);
  }
  // x = A^-1 b on the coarsest level: its dense inverse, or the periodic
  // convolution with the first column of its inverse above STAG_DENSE_MAX
  void coarse_solve(const double* b, double* x) {
    const StagLevel& C = *lv.back();
    if (C.nx * C.ny > STAG_DENSE_MAX)
      launch_st_circulant_solve(s, C.nx, C.ny, coarse_inv.p, b, x);
    else
      launch_st_dense_solve(s, C.nx * C.ny, coarse_inv.p, b, x);
  }
  void launch_copy_grid(const double* from, double* to) {
    MCHECK(hipMemcpyAsync(to, from, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  }

  void read_scal() {
    MCHECK(hipMemcpyAsync(host_scal, scal.p, sizeof(host_scal), hipMemcpyDeviceToHost, s));
    MCHECK(hipStreamSynchronize(s));
  }
  // the true residual r = b - M x and its statistics
  struct Res {
    double linf, l2, floor, bscale;
  };
  Res residual(const double* bb, double* x) {
    launch_st_residual(s, g, k, bb, x, r.p, part.p);
    launch_st_reduce(s, st_parts(n), 4, 0b1101, part.p, scal.p, 0);
    read_scal();
    Res q;
    q.linf = host_scal[4];
    q.l2 = std::sqrt(host_scal[5] / (double(n) * (g.dx * g.dy)));
    // _roundoff_residual_floor! (cycles.jl:78-101): 128 eps max(1, |M||x| / V, |b| / V)
    q.floor = 128.0 * EPS * std::max({1.0, host_scal[6], host_scal[7]});
    q.bscale = host_scal[7];
    return q;
  }
  // _guarded_roundoff_target! (elliptic_solver.jl:462-485) with the floor of
  // the current iterate
  static double guarded(double floor, double target, double initial, double current) {
    if (!(std::isfinite(initial) && std::isfinite(current) && target < current && current < initial)) return target;
    return (std::isfinite(floor) && floor < initial / 100) ? std::max(target, floor) : target;
  }
  // one Krylov run (Krylov.jl cg preconditioned by the staggered V-cycle)
  // from x, whose true residual is in r, until the strong recursive
  // residual reaches `target` or `itmax`
  void precondition(const double* rr, double* zz) {
    if (use_mg)
      vcycle(rr, zz);
    else
      launch_st_mg_smooth(s, g.nx, g.ny, k, 1.0, nullptr, rr, zz);
  }
  int cg_run(double* x, double target, int itmax) {
    precondition(r.p, z.p);
    launch_st_dot(s, n, r.p, z.p, part.p);
    launch_st_reduce(s, st_parts(n), 1, 0, part.p, scal.p, 1);
    launch_copy_grid(z.p, p.p);
    read_scal();
    if (host_scal[0] == 0.0) return 0;
    int it = 0;
    while (it < itmax) {
      launch_st_matvec_dot(s, g, k, p.p, Ap.p, part.p);
      launch_st_reduce(s, st_parts(n), 1, 0, part.p, scal.p, 2);
      launch_st_cg_update_xr(s, n, g.inv_vol, x, r.p, p.p, Ap.p, scal.p, part.p);
      launch_st_reduce(s, st_parts(n), 1, 1, part.p, scal.p, 0);
      read_scal();
      ++it;
      const double strong = host_scal[4];
      if (!std::isfinite(host_scal[1]) || strong <= target || !std::isfinite(strong)) break;
      precondition(r.p, z.p);
      launch_st_dot(s, n, r.p, z.p, part.p);
      launch_st_reduce(s, st_parts(n), 1, 0, part.p, scal.p, 4);
      launch_st_cg_direction(s, n, z.p, p.p, scal.p);
    }
    return it;
  }
  // _solve_krylov_with_stats! with relative_to_rhs (elliptic_solver.jl:550-627)
  ViscousStats solve(const double* bb, double* x, double tol, int max_cycles) {
    ViscousStats st;
    const Res r0 = residual(bb, x);
    const double target0 = std::max(tol * r0.bscale, r0.floor);
    Res cur = r0;
    int iterations = 0;
    if (r0.linf > target0) {
      // _run_krylov_reliably! (elliptic_solver.jl:487-526); every run starts
      // from the true residual of its iterate
      while (iterations < max_cycles) {
        const int its = cg_run(x, target0, max_cycles - iterations);
        iterations += its;
        cur = residual(bb, x);
        if (cur.linf <= guarded(cur.floor, target0, r0.linf, cur.linf)) break;
        if (!std::isfinite(cur.linf) || its == 0) break;
      }
    }
    const double target = guarded(cur.floor, target0, r0.linf, cur.linf);
    st.cycles = iterations;
    st.converged = std::isfinite(cur.linf) && std::isfinite(target) && cur.linf <= target;
    st.initial_l2 = r0.l2;
    st.initial_linf = r0.linf;
    st.final_l2 = cur.l2;
    st.final_linf = cur.linf;
    st.target_linf = target;
    st.residual_scale = r0.bscale;
    return st;
  }
};

Momentum::Momentum(const TreeMesh& m, double viscosity, int scheme, const MultigridConfig& mg, hipStream_t stream)
    : d_(new Impl) {
  Impl& D = *d_;
  if (!(std::isfinite(viscosity) && viscosity > 0))
    throw std::invalid_argument("constant viscosity must be strictly positive");
  if (scheme != STAG_UPWIND && scheme != STAG_CENTRED && scheme != STAG_LW)
    throw std::invalid_argument("unsupported staggered advection scheme");
  if (!(m.periodic[0] && m.periodic[1]) || !m.bcell.empty())
    throw std::invalid_argument("staggered advection currently requires periodic boundaries in both directions");
  if (m.blocks.empty()) throw std::invalid_argument("empty mesh");
  const int level = m.blocks[0].level;
  for (const BlockKey& b : m.blocks)
    if (b.level != level) throw std::invalid_argument("staggered advection currently requires a uniform tree level");
  D.s = stream;
  D.nu = viscosity;
  D.scheme = scheme;
  const int BX = m.bs[0], BY = m.bs[1], nb = m.nb();
  const int nx = m.layout[0] * BX << level, ny = m.layout[1] * BY << level;
  const int N = m.ncells(), F = m.nfaces();
  if (int64_t(nx) * ny != N) throw std::invalid_argument("uniform mesh lattice does not match its cells");
  // lattice position of every cell (block key and j-major local index)
  std::vector<int> cell_i(N), cell_j(N);
  for (int b = 0; b < int(m.blocks.size()); ++b)
    for (int jj = 0; jj < BY; ++jj)
      for (int ii = 0; ii < BX; ++ii) {
        const int c = b * nb + jj * BX + ii;
        const int i = m.blocks[b].i * BX + ii, j = m.blocks[b].j * BY + jj;
        cell_i[c] = i;
        cell_j[c] = j;
      }
  // each cell's minus x and y faces; every face joins lattice neighbours
  std::vector<int> fx(N, -1), fy(N, -1);
  for (int f = 0; f < F; ++f) {
    const int a = m.fminus[f], p = m.fplus[f];
    const int ax = m.faxis[f];
    const int di = (cell_i[p] - cell_i[a] + nx) % nx, dj = (cell_j[p] - cell_j[a] + ny) % ny;
    const bool ok = ax == 0 ? (di == 1 % nx && dj == 0) : (dj == 1 % ny && di == 0);
    if (!ok) throw std::invalid_argument("nonconforming periodic face schedule");
    int& slot = (ax == 0 ? fx : fy)[size_t(cell_j[p]) * nx + cell_i[p]];
    if (slot >= 0) throw std::invalid_argument("nonconforming periodic face schedule");
    slot = f;
  }
  for (int q = 0; q < N; ++q)
    if (fx[q] < 0 || fy[q] < 0) throw std::invalid_argument("incomplete periodic face schedule");
  D.n = N;
  D.nfaces = F;
  // the first block's cell spacing, as the reference's schedule takes it
  const double hx = m.bspacing[0], hy = m.bspacing[1];
  D.g.nx = nx;
  D.g.ny = ny;
  D.g.dx = hx;
  D.g.dy = hy;
  D.g.inv_vol = 1.0 / (hx * hy);
  // stiffness of one connection: (measure distance) (1/d) (1/d)
  // (_diffusion_stiffness with the two-point derivative), x connections:
  // measure dy, distance dx; y connections: measure dx, distance dy
  const double kx = ((hy * hx) * (1.0 / hx)) * (1.0 / hx);
  const double ky = ((hx * hy) * (1.0 / hy)) * (1.0 / hy);
  const double sd = ((kx + kx) + ky) + ky;
  const double iv = 1.0 / (hx * hy);
  D.k.ld = -iv * sd;
  D.k.lx = iv * kx;
  D.k.ly = iv * ky;
  D.fx.alloc(N);
  D.fy.alloc(N);
  MCHECK(hipMemcpyAsync(D.fx.p, fx.data(), sizeof(int) * N, hipMemcpyHostToDevice, stream));
  MCHECK(hipMemcpyAsync(D.fy.p, fy.data(), sizeof(int) * N, hipMemcpyHostToDevice, stream));
  for (Buf<double>* b : {&D.X, &D.Y, &D.U, &D.V, &D.X2, &D.Y2, &D.cx, &D.cy, &D.ox, &D.oy, &D.b, &D.r, &D.z, &D.p,
                         &D.Ap})
    b->alloc(N);
  D.part.alloc(size_t(st_parts(2 * N)) * 4);
  D.scal.alloc(12);
  // the staggered tree-coarsening plan of a uniform periodic mesh
  // (_tree_next_coarse_mesh! with _TreeCoarseningPlan{:staggered},
  // tree_coarsening.jl:145-201): prune leaf levels, then halve the base mesh
  // while the root layout is even and every periodic axis keeps four base
  // cells, down to coarse_size unknowns, max_levels, or no progress
  D.mg = mg;
  int base0 = m.base[0], base1 = m.base[1], lev = level;
  auto add_level = [&](int nxl, int nyl, double hxl, double hyl) {
    std::unique_ptr<StagLevel> L(new StagLevel);
    L->nx = nxl;
    L->ny = nyl;
    L->hx = hxl;
    L->hy = hyl;
    const size_t nl = size_t(nxl) * nyl;
    L->x.alloc(nl);
    L->xs.alloc(nl);
    L->r.alloc(nl);
    if (!D.lv.empty()) L->b.alloc(nl);
    D.lv.push_back(std::move(L));
  };
  add_level(nx, ny, hx, hy);
  while (int64_t(D.lv.back()->nx) * D.lv.back()->ny > mg.coarse_size && int(D.lv.size()) < mg.max_levels) {
    if (lev == 0) {
      if ((base0 / BX) % 2 || (base1 / BY) % 2) break;
      if (base0 / 2 < 4 || base1 / 2 < 4) break;
      base0 /= 2;
      base1 /= 2;
    } else {
      --lev;
    }
    const StagLevel& F = *D.lv.back();
    add_level(F.nx / 2, F.ny / 2, 2 * F.hx, 2 * F.hy);
  }
  // the coarsest level is solved exactly (dense inverse, or the convolution
  // with the first column of its inverse, nc^2 operations per solve); above
  // STAG_CIRC_MAX unknowns the Krylov solve takes the operator's diagonal as
  // preconditioner instead
  D.use_mg = int64_t(D.lv.back()->nx) * D.lv.back()->ny <= STAG_CIRC_MAX;
  MCHECK(hipStreamSynchronize(stream));
}

std::vector<int> Momentum::level_cells() const {
  std::vector<int> out;
  for (const auto& L : d_->lv) out.push_back(L->nx * L->ny);
  return out;
}

Momentum::~Momentum() = default;

int Momentum::nfaces() const { return d_->nfaces; }
double Momentum::viscosity() const { return d_->nu; }
int Momentum::scheme() const { return d_->scheme; }
bool Momentum::sweep_forward() const { return d_->forward; }

void Momentum::advect(double* field, const double* advecting, double dt) {
  Impl& D = *d_;
  if (!(dt >= 0)) throw std::invalid_argument("dt must be nonnegative");
  const bool self = advecting == field;
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, field, D.X.p, D.Y.p);
  if (!self) launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, advecting, D.U.p, D.V.p);
  const bool split = D.scheme == STAG_LW;
  int stages[2] = {0, 0};
  int nst = 1;
  if (split) {
    stages[0] = D.forward ? 1 : 2;
    stages[1] = D.forward ? 2 : 1;
    nst = 2;
  }
  double *xi = D.X.p, *yi = D.Y.p, *xo = D.X2.p, *yo = D.Y2.p;
  for (int q = 0; q < nst; ++q) {
    // the one-velocity shorthand transports each sweep with the published
    // intermediate
    const double* U = self ? xi : D.U.p;
    const double* V = self ? yi : D.V.p;
    launch_st_sweep(D.s, D.g, stages[q], D.scheme, dt, xi, yi, U, V, xo, yo);
    std::swap(xi, xo);
    std::swap(yi, yo);
  }
  launch_st_scatter(D.s, D.n, D.fx.p, D.fy.p, xi, yi, field);
  MCHECK(hipStreamSynchronize(D.s));
  if (split && dt != 0) D.forward = !D.forward;
}

void Momentum::diffusion(const double* u, double* out) {
  Impl& D = *d_;
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, u, D.X.p, D.Y.p);
  launch_st_diffusion(D.s, D.g, D.k, D.nu, D.X.p, D.Y.p, D.X2.p, D.Y2.p);
  launch_st_scatter(D.s, D.n, D.fx.p, D.fy.p, D.X2.p, D.Y2.p, out);
  MCHECK(hipStreamSynchronize(D.s));
}

void Momentum::implicit(double* dst, const double* convected, const double* old, double dt, double tol,
                        int max_cycles, ViscousStats stats[2]) {
  Impl& D = *d_;
  if (!(dt > 0)) throw std::invalid_argument("dt must be strictly positive");
  if (max_cycles < 0) throw std::invalid_argument("max_cycles must be nonnegative");
  // the shifted operators for h = dt nu / 2 (_update_staggered_elliptic!)
  const double half = dt * D.nu / 2;
  if (half != D.coarse_half) D.set_shift(half);
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, convected, D.cx.p, D.cy.p);
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, old, D.ox.p, D.oy.p);
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, dst, D.X.p, D.Y.p);
  for (int a = 0; a < 2; ++a) {
    D.comp = a;
    launch_st_midpoint_rhs(D.s, D.g, D.k, D.nu, dt, a ? D.cy.p : D.cx.p, a ? D.oy.p : D.ox.p, D.b.p);
    stats[a] = D.solve(D.b.p, a ? D.Y.p : D.X.p, tol, max_cycles);
  }
  launch_st_scatter(D.s, D.n, D.fx.p, D.fy.p, D.X.p, D.Y.p, dst);
  MCHECK(hipStreamSynchronize(D.s));
}

double Momentum::stable_timestep(const double* u, double cfl) {
  Impl& D = *d_;
  if (!(cfl > 0 && cfl <= 1)) throw std::invalid_argument("cfl must lie in (0, 1]");
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, u, D.U.p, D.V.p);
  launch_st_rates(D.s, D.g, D.U.p, D.V.p, D.part.p);
  launch_st_reduce(D.s, st_parts(2 * D.n), 1, 1, D.part.p, D.scal.p, 0);
  D.read_scal();
  const double rate = D.host_scal[4];
  return rate == 0 ? std::numeric_limits<double>::infinity() : cfl / rate;
}

}  // namespace tpamr
