// Oracle C API for the Python test harness (ctypes). TEST INFRASTRUCTURE ONLY.
#include <cstring>
#include <omp.h>
#include <exception>
#include <string>

#include <stdexcept>

#include "oracle.hpp"

namespace orc {
Solver* build_solver(const Mesh& fine, const std::vector<double>& kappa, int coarse_size, int pre,
                     int post, int max_levels, bool use_mg, const double* volumes, const bool* open,
                     const double* bcoef);
void weighted_gradient_boundary(const Operator& op, const double* p, double* gb);
double solver_floor(Solver& S, const double* x, const double* b, bool assembled);
double csc_floor(bool single, int n, const int* colptr, const int* rowval, const double* nzval,
                 const double* volumes, const double* x, const double* b);
Stats scripted_solve(bool single, int n, const int* colptr, const int* rowval, const double* nzval,
                     const double* volumes, const double* b, double* x, double rtol, double atol, int max_cycles,
                     bool relative_to_rhs, int npasses, const double* iterates, int* passes);
void apply_vcycle(Solver& S, const double* r, double* z);
Stats solve(Solver& S, const double* rhs, double* xio, double rtol, double atol, int max_cycles,
            int compat_error, std::vector<double>* removed_means);
void mixture_coefficients(const Mesh& M, const double* c, double rho_l, double rho_g, double* kappa);
void mixture_boundary_coefficients(const Mesh& M, const double* c, double rho_l, double rho_g, double* kb);
double face_curvature(const Mesh& M, const double* c, const double* kappa, const double* quality, int fi);
int surface_tension_jumps_cells(const Mesh& M, const double* c, const double* kappa, const double* quality,
                                double sigma, int csf, double* jumps);
void gravity_jumps(const Mesh& M, const double* offsets, const double* liquid, const double* gas, double rho_l,
                   double rho_g, double gx, double gy, double* jumps);
void cgfm_jumps(const Mesh& M, const double* c, const double* curvature, double sigma, int csf,
                double* jumps);
void weighted_jump_force(const Operator& op, const double* jumps, double* accel);
void weighted_gradient(const Operator& op, const double* p, double* g);
void divergence(const Mesh& M, const double* u, const double* ub, double* div);
Stats project(Solver& S, double* u, double* ub, double* p, double dt, double tol, int max_cycles, double* phi_out,
              const double* inherited, double reference_dt);
Stats adjoint(Solver& S, const double* w, const double* p, double rtol, int max_cycles,
              double* dJ_db, double* dJ_dbeta, double* lambda_out);
void apply_fine(Solver& S, const double* x, double* y);
double jacobi_weight_public(std::vector<double>& invdiag, const std::vector<double>& rowsum, bool skip_nonpositive,
                            bool allow_zero, bool clamp, double zero_weight);
void apply_level(Solver& S, int l, const double* x, double* y);
}  // namespace orc

using namespace orc;

static thread_local std::string g_err;

#define GUARD(...)                  \
  try {                             \
    __VA_ARGS__;                    \
  } catch (const std::exception& e) { \
    g_err = e.what();               \
    return -1;                      \
  }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// OpenMP threads of the oracle's loops (results do not depend on the count)
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads() { return omp_get_max_threads(); }

// refined: n_refined triples (level, i, j) of blocks the refine predicate split
int orc_mesh_new(const int* base, const int* bs, const double* origin, const double* extent,
                 const int* periodic, int max_level, int n_refined, const int* refined, void** out) {
  GUARD({
    std::vector<Key> keys(n_refined);
    for (int q = 0; q < n_refined; ++q) keys[q] = {refined[3 * q], refined[3 * q + 1], refined[3 * q + 2]};
    const bool per[2] = {periodic[0] != 0, periodic[1] != 0};
    *out = build_mesh(base, bs, origin, extent, per, max_level, keys).release();
    return 0;
  })
}

void orc_mesh_free(void* m) { delete static_cast<Mesh*>(m); }

int orc_mesh_sizes(void* mp, int* out /* nblocks, ncells, nfaces, nbfaces, max_leaf */) {
  Mesh* m = static_cast<Mesh*>(mp);
  out[0] = int(m->blocks.size());
  out[1] = m->ncells();
  out[2] = int(m->faces.size());
  out[3] = int(m->bfaces.size());
  out[4] = m->max_leaf_level();
  return 0;
}

int orc_mesh_blocks(void* mp, int* keys, double* geom /* origin2 extent2 spacing2 volume */) {
  Mesh* m = static_cast<Mesh*>(mp);
  for (size_t b = 0; b < m->blocks.size(); ++b) {
    const Block& B = m->blocks[b];
    keys[3 * b] = B.key.level;
    keys[3 * b + 1] = B.key.i;
    keys[3 * b + 2] = B.key.j;
    if (geom) {
      double* g = geom + 7 * b;
      g[0] = B.origin[0]; g[1] = B.origin[1]; g[2] = B.extent[0]; g[3] = B.extent[1];
      g[4] = B.spacing[0]; g[5] = B.spacing[1]; g[6] = B.volume;
    }
  }
  return 0;
}

int orc_mesh_faces(void* mp, int* ints /* minus plus axis */, double* dbl /* cx cy area */) {
  Mesh* m = static_cast<Mesh*>(mp);
  for (size_t f = 0; f < m->faces.size(); ++f) {
    ints[3 * f] = m->faces[f].minus;
    ints[3 * f + 1] = m->faces[f].plus;
    ints[3 * f + 2] = m->faces[f].axis;
    dbl[3 * f] = m->faces[f].center[0];
    dbl[3 * f + 1] = m->faces[f].center[1];
    dbl[3 * f + 2] = m->faces[f].area;
  }
  return 0;
}

int orc_mesh_bfaces(void* mp, int* ints /* cell axis side */, double* dbl /* cx cy area */) {
  Mesh* m = static_cast<Mesh*>(mp);
  for (size_t f = 0; f < m->bfaces.size(); ++f) {
    ints[3 * f] = m->bfaces[f].cell;
    ints[3 * f + 1] = m->bfaces[f].axis;
    ints[3 * f + 2] = m->bfaces[f].side;
    dbl[3 * f] = m->bfaces[f].center[0];
    dbl[3 * f + 1] = m->bfaces[f].center[1];
    dbl[3 * f + 2] = m->bfaces[f].area;
  }
  return 0;
}

// inverse metric K as CSC; call with null arrays to query nnz
int orc_metric(void* mp, int* nnz, int* is_diag, int* colptr, int* rowval, double* nzval) {
  GUARD({
    Mesh* m = static_cast<Mesh*>(mp);
    bool diag;
    CSC K = inverse_metric(*m, &diag);
    *nnz = K.nnz();
    *is_diag = diag;
    if (colptr) {
      std::memcpy(colptr, K.colptr.data(), sizeof(int) * K.colptr.size());
      std::memcpy(rowval, K.rowval.data(), sizeof(int) * K.rowval.size());
      std::memcpy(nzval, K.nzval.data(), sizeof(double) * K.nzval.size());
    }
    return 0;
  })
}

int orc_solver_new(void* mp, const double* kappa, int coarse_size, int pre, int post, int max_levels,
                   int use_mg, void** out) {
  GUARD({
    Mesh* m = static_cast<Mesh*>(mp);
    std::vector<double> k(kappa, kappa + m->faces.size());
    *out = build_solver(*m, k, coarse_size, pre, post, max_levels, use_mg != 0, nullptr, nullptr, nullptr);
    return 0;
  })
}

// the same with per-cell volumes of the fine operator (NULL: geometric)
int orc_solver_new_ex(void* mp, const double* kappa, const double* volumes, int coarse_size, int pre, int post,
                      int max_levels, int use_mg, void** out) {
  GUARD({
    Mesh* m = static_cast<Mesh*>(mp);
    std::vector<double> k(kappa, kappa + m->faces.size());
    *out = build_solver(*m, k, coarse_size, pre, post, max_levels, use_mg != 0, volumes, nullptr, nullptr);
    return 0;
  })
}

// the same with open pressure sides open[4] (left, right, bottom, top) and
// boundary face coefficients (NULL: extended from kappa)
int orc_solver_new_open(void* mp, const double* kappa, const double* volumes, const int* open, const double* bcoef,
                        int coarse_size, int pre, int post, int max_levels, int use_mg, void** out) {
  GUARD({
    Mesh* m = static_cast<Mesh*>(mp);
    std::vector<double> k(kappa, kappa + m->faces.size());
    const bool ob[4] = {open[0] != 0, open[1] != 0, open[2] != 0, open[3] != 0};
    *out = build_solver(*m, k, coarse_size, pre, post, max_levels, use_mg != 0, volumes, ob, bcoef);
    return 0;
  })
}

// level l's component labels (0-based root cell, -1 inactive) and the
// ungauged components' roots (anchors, increasing; *count of them)
int orc_solver_components(void* sp, int l, int* labels, int* anchors, int* count) {
  GUARD({
    Solver* s = static_cast<Solver*>(sp);
    if (l < 0 || l >= int(s->levels.size())) throw std::invalid_argument("level out of range");
    const Components& C = s->levels[l].op->comp;
    if (labels) std::copy(C.labels.begin(), C.labels.end(), labels);
    if (anchors) std::copy(C.anchors.begin(), C.anchors.end(), anchors);
    *count = int(C.anchors.size());
    return 0;
  })
}

// level l's boundary face coefficients, boundary diagonal, and flags[2] =
// (singular, coarse augmented)
int orc_solver_level_boundary(void* sp, int l, double* bcoef, double* bdiag, int* flags) {
  GUARD({
    Solver* s = static_cast<Solver*>(sp);
    if (l < 0 || l >= int(s->levels.size())) throw std::invalid_argument("level out of range");
    const Operator& op = *s->levels[l].op;
    if (bcoef) std::copy(op.bcoef.begin(), op.bcoef.end(), bcoef);
    if (bdiag) std::copy(op.bdiag.begin(), op.bdiag.end(), bdiag);
    flags[0] = op.singular ? 1 : 0;
    flags[1] = s->coarse_augmented ? 1 : 0;
    return 0;
  })
}

// round-off residual floor of iterate x with integrated rhs b: the
// factorized operator's (assembled = 0) or the assembled matrix's
int orc_solver_floor(void* sp, const double* x, const double* b, int assembled, double* out) {
  GUARD({
    *out = solver_floor(*static_cast<Solver*>(sp), x, b, assembled != 0);
    return 0;
  })
}

// assembled `_roundoff_residual_floor!` of a CSC matrix (single = 1: Float32)
int orc_assembled_floor(int single, int n, const int* colptr, const int* rowval, const double* nzval,
                        const double* volumes, const double* x, const double* b, double* out) {
  GUARD({
    *out = csc_floor(single != 0, n, colptr, rowval, nzval, volumes, x, b);
    return 0;
  })
}

// the solve driver with scripted Krylov iterates on an assembled CSC matrix
// (single = 1: Float32 arithmetic); *passes receives the Krylov runs made
int orc_scripted_solve(int single, int n, const int* colptr, const int* rowval, const double* nzval,
                       const double* volumes, const double* b, double* x, double rtol, double atol, int max_cycles,
                       int relative_to_rhs, int npasses, const double* iterates, Stats* st, int* passes) {
  GUARD({
    *st = scripted_solve(single != 0, n, colptr, rowval, nzval, volumes, b, x, rtol, atol, max_cycles,
                         relative_to_rhs != 0, npasses, iterates, passes);
    return 0;
  })
}

void orc_solver_free(void* s) { delete static_cast<Solver*>(s); }

int orc_solver_nlevels(void* sp) { return int(static_cast<Solver*>(sp)->levels.size()); }

// per level: ncells, nfaces, nnz(A), nnz(P), omega
int orc_solver_level(void* sp, int l, int* ints, double* omega) {
  Solver* s = static_cast<Solver*>(sp);
  const Level& L = s->levels[l];
  ints[0] = int(L.op->volumes.size());
  ints[1] = int(L.op->kappa.size());
  ints[2] = L.op->A.nnz();
  ints[3] = L.P.nnz();
  ints[4] = L.op->K_diagonal;
  *omega = L.omega;
  return 0;
}

int orc_solver_level_matrix(void* sp, int l, int which, int* colptr, int* rowval, double* nzval) {
  Solver* s = static_cast<Solver*>(sp);
  const CSC& A = which == 0 ? s->levels[l].op->A : s->levels[l].P;
  std::memcpy(colptr, A.colptr.data(), sizeof(int) * A.colptr.size());
  std::memcpy(rowval, A.rowval.data(), sizeof(int) * A.rowval.size());
  std::memcpy(nzval, A.nzval.data(), sizeof(double) * A.nzval.size());
  return 0;
}

int orc_solver_level_kappa(void* sp, int l, double* kappa) {
  Solver* s = static_cast<Solver*>(sp);
  const auto& k = s->levels[l].op->kappa;
  std::memcpy(kappa, k.data(), sizeof(double) * k.size());
  return 0;
}

int orc_solver_level_mesh(void* sp, int l, void** mesh) {
  Solver* s = static_cast<Solver*>(sp);
  *mesh = const_cast<Mesh*>(s->levels[l].op->mesh);
  return 0;
}

int orc_apply(void* sp, const double* x, double* y) {
  GUARD({ apply_fine(*static_cast<Solver*>(sp), x, y); return 0; })
}

int orc_apply_level(void* sp, int l, const double* x, double* y) {
  GUARD({ apply_level(*static_cast<Solver*>(sp), l, x, y); return 0; })
}

int orc_vcycle(void* sp, const double* r, double* z) {
  GUARD({ apply_vcycle(*static_cast<Solver*>(sp), r, z); return 0; })
}

int orc_solve(void* sp, const double* rhs, double* x, double rtol, double atol, int max_cycles,
              int compat_error, Stats* st, double* removed_means) {
  GUARD({
    std::vector<double> rm;
    *st = solve(*static_cast<Solver*>(sp), rhs, x, rtol, atol, max_cycles, compat_error, &rm);
    if (removed_means)
      for (size_t q = 0; q < rm.size(); ++q) removed_means[q] = rm[q];
    return 0;
  })
}

int orc_mixture_boundary_coefficients(void* mp, const double* c, double rho_l, double rho_g, double* kb) {
  GUARD({ mixture_boundary_coefficients(*static_cast<Mesh*>(mp), c, rho_l, rho_g, kb); return 0; })
}

int orc_mixture_coefficients(void* mp, const double* c, double rho_l, double rho_g, double* kappa) {
  GUARD({ mixture_coefficients(*static_cast<Mesh*>(mp), c, rho_l, rho_g, kappa); return 0; })
}

int orc_cgfm_jumps(void* mp, const double* c, const double* curvature, double sigma, int csf, double* out) {
  GUARD({ cgfm_jumps(*static_cast<Mesh*>(mp), c, curvature, sigma, csf, out); return 0; })
}

int orc_face_curvature(void* mp, const double* c, const double* kappa, const double* quality, int n,
                       const int* faces, double* out) {
  GUARD({
    const Mesh& M = *static_cast<Mesh*>(mp);
    for (int q = 0; q < n; ++q) out[q] = face_curvature(M, c, kappa, quality, faces[q]);
    return 0;
  })
}

int orc_surface_tension_jumps_cells(void* mp, const double* c, const double* kappa, const double* quality,
                                    double sigma, int csf, double* jumps, int* unresolved) {
  GUARD({
    *unresolved = surface_tension_jumps_cells(*static_cast<Mesh*>(mp), c, kappa, quality, sigma, csf, jumps);
    return 0;
  })
}

int orc_gravity_jumps(void* mp, const double* offsets, const double* liquid, const double* gas, double rho_l,
                      double rho_g, double gx, double gy, double* jumps) {
  GUARD({ gravity_jumps(*static_cast<Mesh*>(mp), offsets, liquid, gas, rho_l, rho_g, gx, gy, jumps); return 0; })
}

int orc_weighted_jump_force(void* sp, const double* jumps, double* accel) {
  GUARD({ weighted_jump_force(*static_cast<Solver*>(sp)->ops[0], jumps, accel); return 0; })
}

int orc_weighted_gradient(void* sp, const double* p, double* g) {
  GUARD({ weighted_gradient(*static_cast<Solver*>(sp)->ops[0], p, g); return 0; })
}

int orc_fill_halo(void* mp, int halo, const double* cells, const double* physical, double* out, int* nphysical) {
  GUARD({
    if (halo < 1) throw std::invalid_argument("halo must be positive");
    *nphysical = fill_halo(*static_cast<Mesh*>(mp), halo, cells, physical, out);
    return 0;
  })
}

int orc_divergence(void* mp, const double* u, const double* ub, double* div) {
  GUARD({ divergence(*static_cast<Mesh*>(mp), u, ub, div); return 0; })
}

int orc_project(void* sp, double* u, double* p, double dt, double tol, int max_cycles, Stats* st,
                double* phi) {
  GUARD({ *st = project(*static_cast<Solver*>(sp), u, nullptr, p, dt, tol, max_cycles, phi, nullptr, dt); return 0; })
}

int orc_project_ex(void* sp, double* u, double* p, double dt, double tol, int max_cycles, Stats* st, double* phi,
                   const double* inherited, double reference_dt) {
  GUARD({
    *st = project(*static_cast<Solver*>(sp), u, nullptr, p, dt, tol, max_cycles, phi, inherited, reference_dt);
    return 0;
  })
}

// the projection with boundary face velocities ub (open sides)
int orc_project_open(void* sp, double* u, double* ub, double* p, double dt, double tol, int max_cycles, Stats* st,
                     double* phi) {
  GUARD({ *st = project(*static_cast<Solver*>(sp), u, ub, p, dt, tol, max_cycles, phi, nullptr, dt); return 0; })
}

int orc_weighted_gradient_boundary(void* sp, const double* p, double* gb) {
  GUARD({ weighted_gradient_boundary(*static_cast<Solver*>(sp)->ops[0], p, gb); return 0; })
}

int orc_adjoint(void* sp, const double* w, const double* p, double rtol, int max_cycles, double* dJ_db,
                double* dJ_dbeta, double* lambda, Stats* st) {
  GUARD({
    *st = adjoint(*static_cast<Solver*>(sp), w, p, rtol, max_cycles, dJ_db, dJ_dbeta, lambda);
    return 0;
  })
}

// `_jacobi_weight!` known answers (hierarchy.jl:8-30): diag is replaced in place
int orc_jacobi_weight(int n, double* diag, const double* rowsum, int skip_nonpositive, int allow_zero, int clamp,
                      double zero_weight, double* weight) {
  GUARD({
    std::vector<double> d(diag, diag + n), r(rowsum, rowsum + n);
    *weight = jacobi_weight_public(d, r, skip_nonpositive, allow_zero, clamp, zero_weight);
    std::copy(d.begin(), d.end(), diag);
    return 0;
  })
}

}  // extern "C"
