// Host side of the one-velocity momentum operators on a uniform, fully
// periodic block mesh: the structured component grids, the transport
// stages, the viscous operator's coefficients and the Krylov driver of the
// implicit-midpoint viscous solve (the reference's _solve_krylov_with_stats!
// with relative_to_rhs, elliptic_solver.jl:550-627, around PCG with the
// shifted matrix's diagonal).
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

struct Momentum::Impl {
  hipStream_t s = nullptr;
  StagGrid g{};
  StagCoef k{};
  double nu = 0;
  int scheme = STAG_LW;
  bool forward = true;
  int n = 0, nfaces = 0;
  Buf<int> fx, fy;
  Buf<double> X, Y, U, V, X2, Y2;          // transported grids, advecting grids, stage outputs
  Buf<double> cx, cy, ox, oy;              // implicit substep: convected, old
  Buf<double> b, r, z, p, Ap;              // viscous Krylov vectors (one component)
  Buf<double> part, scal;
  double host_scal[12] = {};

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
  // one Krylov run (Krylov.jl cg with a diagonal preconditioner) from x,
  // whose true residual is in r, until the strong recursive residual
  // reaches `target` or `itmax`
  int cg_run(double* x, double target, int itmax) {
    launch_st_cg_start(s, n, k, r.p, z.p, p.p, part.p);
    launch_st_reduce(s, st_parts(n), 1, 0, part.p, scal.p, 1);
    read_scal();
    if (host_scal[0] == 0.0) return 0;
    int it = 0;
    while (it < itmax) {
      launch_st_matvec_dot(s, g, k, p.p, Ap.p, part.p);
      launch_st_reduce(s, st_parts(n), 1, 0, part.p, scal.p, 2);
      launch_st_cg_update(s, n, k, g.inv_vol, x, r.p, z.p, p.p, Ap.p, scal.p, part.p);
      launch_st_reduce(s, st_parts(n), 2, 0b10, part.p, scal.p, 3);
      read_scal();
      ++it;
      const double strong = host_scal[3];
      if (!std::isfinite(host_scal[1]) || strong <= target || !std::isfinite(strong)) break;
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

Momentum::Momentum(const TreeMesh& m, double viscosity, int scheme, hipStream_t stream) : d_(new Impl) {
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
  MCHECK(hipStreamSynchronize(stream));
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
  // the shifted matrix h K + V, h = dt nu / 2 (_update_staggered_elliptic!
  // with _staggered_stable_matrix, reaction 1)
  const double half = dt * D.nu / 2;
  const double hx = D.g.dx, hy = D.g.dy;
  const double kx = ((hy * hx) * (1.0 / hx)) * (1.0 / hx);
  const double ky = ((hx * hy) * (1.0 / hy)) * (1.0 / hy);
  const double sd = ((kx + kx) + ky) + ky;
  D.k.md = half * sd + hx * hy;
  D.k.mx = half * -kx;
  D.k.my = half * -ky;
  D.k.mdinv = 1.0 / D.k.md;
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, convected, D.cx.p, D.cy.p);
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, old, D.ox.p, D.oy.p);
  launch_st_gather(D.s, D.n, D.fx.p, D.fy.p, dst, D.X.p, D.Y.p);
  for (int a = 0; a < 2; ++a) {
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
