// Momentum parts of the one-velocity step on a uniform, fully periodic block
// mesh (momentum.cpp): the staggered transport workspace
// (StaggeredVelocityAdvectionWorkspace), the staggered viscous operator
// (StaggeredDiffusionOperator) and its implicit-midpoint solve, on the
// device. Vectors passed in and out are device memory in the reference's
// interior-face order.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <vector>

#include "hier.hpp"

namespace tpamr {

// A HIP runtime failure inside the momentum operators.
struct MomentumDeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// LinearSolverStats of one component's viscous solve (stats.jl fields).
struct ViscousStats {
  int cycles = 0, converged = 0;
  double initial_l2 = 0, initial_linf = 0, final_l2 = 0, final_linf = 0, target_linf = 0, residual_scale = 0;
};

class Momentum {
 public:
  // Throws std::invalid_argument unless the mesh is uniform and periodic in
  // both directions (the reference's _uniform_periodic_cell_faces). The
  // viscous solve's staggered multigrid takes the multigrid options.
  Momentum(const TreeMesh& mesh, double viscosity, int scheme, const MultigridConfig& mg, hipStream_t stream);
  ~Momentum();
  Momentum(const Momentum&) = delete;
  Momentum& operator=(const Momentum&) = delete;

  int nfaces() const;
  double viscosity() const;
  int scheme() const;
  bool sweep_forward() const;
  // advect!(field, advecting, dt) (staggered_update.jl:265-326); advecting
  // == field: the one-velocity shorthand
  void advect(double* field, const double* advecting, double dt);
  // diffusion! (staggered_diffusion.jl:880-900): out = nu L u
  void diffusion(const double* u, double* out);
  // implicit_midpoint_diffusion! (implicit_diffusion.jl:916-965); dst holds
  // the initial guess on entry
  void implicit(double* dst, const double* convected, const double* old, double dt, double tolerance,
                int max_cycles, ViscousStats stats[2]);
  // stable_timestep(velocity, workspace; cfl) (staggered_update.jl:426-455)
  double stable_timestep(const double* u, double cfl);
  // cells per level of the viscous solve's staggered hierarchy
  std::vector<int> level_cells() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> d_;
};

}  // namespace tpamr
