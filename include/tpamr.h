/* tpamr.h — C ABI of the MI355X (gfx950) pressure-projection solver for
 * TwoPhaseAMR.jl's tree geometric multigrid.
 *
 * Every pointer argument is caller-owned memory; the library copies inputs
 * and never retains a caller pointer. Host-pointer entry points accept plain
 * host arrays (Julia `Vector{Float64}` under GC.@preserve, numpy arrays).
 * The `_device` entry points take device pointers valid on the context's
 * device and run on the context's stream. That stream first waits for the
 * work queued so far on the caller's stream (tpamr_set_caller_stream; the
 * null stream by default), so inputs written there are complete before the
 * library reads them; each `_device` call returns when its outputs are
 * written.
 *
 * Cells are numbered in the reference's order (block-major, then j, then i;
 * src/mesh/topology/mesh.jl:28-30); interior and boundary faces in the
 * reference's face-schedule order (mesh.jl:874-927). All return codes are 0
 * on success and negative on failure; tpamr_last_error gives the message.
 * A context is single-threaded and not re-entrant, like the reference's
 * FactorizedEllipticOperator (src/operators/elliptic.jl:1004-1007).
 */
#ifndef TPAMR_H
#define TPAMR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tpamr_ctx tpamr_ctx;

enum {
  TPAMR_OK = 0,
  TPAMR_EINVAL = -1,       /* invalid argument (ArgumentError) */
  TPAMR_ESIZE = -2,        /* size mismatch (DimensionMismatch) */
  TPAMR_ENONFINITE = -3,   /* non-finite input */
  TPAMR_EDEVICE = -4,      /* HIP runtime or no device */
  TPAMR_ESTATE = -5,       /* call order (no mesh / no coefficients) */
  TPAMR_EINCOMPATIBLE = -6 /* compatibility = error and an incompatible RHS */
};

/* Precision of a context's solves.
 * TPAMR_FP64: FP64 Krylov CG preconditioned by an FP64 V-cycle.
 * TPAMR_FP32_VCYCLE: FP64 Krylov CG preconditioned by an FP32 V-cycle.
 * TPAMR_FP32_REFINE: FP64 iterative refinement around a Krylov CG whose
 *   vectors are FP32 (inner products and scalars FP64), preconditioned by the
 *   FP32 V-cycle: each time the strong residual has fallen to 5e-3 of its
 *   largest value since the last refinement step (the environment variable
 *   TPAMR_IR_DELTA sets another factor), the FP32 correction moves into the
 *   FP64 solution, the FP32 residual is replaced by the FP64 residual
 *   b - A x and the CG recurrence restarts from it, all on the device.
 *   The Krylov seam (tpamr_krylov) runs the TPAMR_FP32_VCYCLE iteration. */
enum { TPAMR_FP64 = 0, TPAMR_FP32_VCYCLE = 1, TPAMR_FP32_REFINE = 2 };

/* LinearSolverStats, src/solvers/multigrid/stats.jl:14-36 */
typedef struct {
  int cycles, converged, compatible, reason;
  double initial_l2, initial_linf, final_l2, final_linf, target_linf, residual_scale, removed_rhs_mean;
  double solve_ms; /* device time of the solve, HIP events on the context stream */
  int krylov_runs;  /* Krylov runs of the restart-and-verify loop */
  int replacements; /* FP64 residual replacements of the inner FP32 iteration (TPAMR_FP32_REFINE) */
} tpamr_stats;

/* MultigridOptions, src/solvers/multigrid/options.jl:1-52 */
typedef struct {
  int coarse_size, pre_sweeps, post_sweeps, max_levels;
} tpamr_multigrid_options;

/* Create a context on HIP device `device` (-1: host-only context that builds
 * and inspects hierarchies but cannot solve). Replaces the construction of
 * TreeGeometricMultigridPreconditioner + EllipticSolver
 * (src/solvers/multigrid/elliptic_solver.jl:282-302). */
int tpamr_create(tpamr_ctx** out, int device, int precision);
void tpamr_destroy(tpamr_ctx* ctx);
const char* tpamr_last_error(const tpamr_ctx* ctx);
/* The stream (a hipStream_t of the context's device; NULL = the null stream)
 * on which the caller produces the inputs of the `_device` entry points:
 * each of them orders the context's stream after the work queued on it at
 * the time of the call. No reference counterpart: the reference's arrays
 * live in host memory. */
int tpamr_set_caller_stream(tpamr_ctx* ctx, void* stream);

/* Block-quadtree mesh: BlockTreeMesh(base_mesh, block_size, origin, extent,
 * periodic, max_level, refine) (mesh.jl:1059-1121). `refined` lists
 * n_refined (level, i, j) triples of nodes the refine predicate split; the
 * library 2:1-balances the tree (mesh.jl:673-714). Passing the ancestors of
 * every leaf of an existing (balanced) BlockTreeMesh reproduces it. */
int tpamr_set_mesh(tpamr_ctx* ctx, const int base_mesh[2], const int block_size[2], const double origin[2],
                   const double extent[2], const int periodic[2], int max_level, int n_refined,
                   const int* refined);
/* sizes[5] = nblocks, ncells, nfaces, nboundary_faces, max leaf level */
int tpamr_mesh_sizes(const tpamr_ctx* ctx, int* sizes);
/* leaf keys (level,i,j) x nblocks in the reference block order */
int tpamr_mesh_blocks(const tpamr_ctx* ctx, int* keys);
/* interior faces: minus, plus (0-based cells), axis (0 = x), area */
int tpamr_mesh_faces(const tpamr_ctx* ctx, int* minus, int* plus, int* axis, double* area);
/* boundary faces: cell, axis, side (+1/-1), area */
int tpamr_mesh_boundary_faces(const tpamr_ctx* ctx, int* cell, int* axis, int* side, double* area);

/* MultigridOptions (defaults 64, 2, 2, 32). Takes effect at the next
 * tpamr_set_face_coefficients. */
int tpamr_set_multigrid(tpamr_ctx* ctx, const tpamr_multigrid_options* options);

/* Face coefficients kappa_f (one per interior face, reference order,
 * nonnegative): builds the hierarchy on first use, otherwise refreshes every
 * coefficient-dependent value (replaces _refresh_geometric_multigrid!,
 * src/solvers/multigrid/geometric_preconditioner.jl:483-594). Zero
 * coefficients split the domain into pressure components
 * (_refresh_elliptic_components!, elliptic_solver.jl:39-109): every component
 * without an open face is projected and gauged on its own, on every
 * multigrid level (the V-cycle then runs level by level); such a refresh
 * runs on the host. */
int tpamr_set_face_coefficients(tpamr_ctx* ctx, const double* kappa);

/* Open (homogeneous Dirichlet) pressure sides: PoissonOperator's
 * velocity_boundary and boundary_face_coefficient (src/operators/elliptic.jl:
 * 76-152, 683-800). open[4] = left, right, bottom, top (the Outflow and
 * ZeroGradientVelocity sides of VelocityBoundaryConditions; all zero: walls
 * and periodic sides only, the default). boundary_coefficient: one value per
 * boundary face (reference order, nonnegative), or NULL to extend the next
 * face coefficients to the open faces (_boundary_material: each open face
 * takes the area-weighted mean of the interior coefficients on the opposite
 * side of its cell). Every open face adds |f| 2 k_b / h to its cell's
 * diagonal; with one positive, the operator is nonsingular: solves take the
 * rhs without compatibility projection (removed_rhs_mean 0) and return no
 * gauge-shifted solution, and the coarsest level is solved without the
 * volume constraint. Takes effect at the next tpamr_set_face_coefficients
 * (or _device); changing the sides rebuilds the hierarchy there
 * (geometric_preconditioner.jl:494-496). A new mesh keeps the sides and
 * extends the coefficients again unless they are given anew. */
int tpamr_set_pressure_boundary(tpamr_ctx* ctx, const int open[4], const double* boundary_coefficient);
/* _elliptic_boundary_gradient! (elliptic.jl:1254-1263): per boundary face
 * -side (2 k_b / h) p_cell on open faces, 0 on closed ones (host arrays) */
int tpamr_weighted_gradient_boundary(tpamr_ctx* ctx, const double* p, double* gb);
/* _project_prepared! (navier_stokes/projection.jl:221-260) for an operator with
 * open sides: `ub` (one velocity per boundary face) enters the divergence and
 * its open faces receive -dt g_b; u (interior faces) and p as tpamr_project.
 * The interior-face projections (tpamr_project, _ex, _device and the
 * projection adjoints) reject an operator with open sides, and this one an
 * operator without them (_validate_projection_boundary, projection.jl:37-49). */
int tpamr_project_open(tpamr_ctx* ctx, double* u, double* ub, double* p, double dt, double tolerance, int max_cycles,
                       double* phi, tpamr_stats* stats);

/* The last solve's removed component means and their anchors
 * (LinearSolverStats.removed_rhs_component_means / _anchors, stats.jl):
 * *count components, roots 1-based smallest cells (NULL arrays: count
 * only). A connected singular operator reports one component rooted at 1;
 * an anchored one none. */
int tpamr_component_means(const tpamr_ctx* ctx, int* count, int* roots, double* means);

/* hierarchy inspection */
int tpamr_num_levels(const tpamr_ctx* ctx);
/* Level `level`'s ungauged components: *count, per cell its 1-based slot (0:
 * gauged; NULL: skipped) and per slot its root (1-based smallest cell; NULL:
 * skipped). */
int tpamr_level_components(const tpamr_ctx* ctx, int level, int* count, int* slot, int* roots);
/* Level `level`'s boundary face coefficients (restricted from the finest
 * level's, tree_coarsening.jl:347-398,538-555; NULL: skipped) and flags[4] =
 * (level singular, coarsest level solved with the volume constraint, open
 * boundary faces of the level, boundary faces of the level). */
int tpamr_level_boundary(const tpamr_ctx* ctx, int level, double* boundary_coefficient, int* flags);
/* Levels the last hierarchy build took over from the previous mesh's
 * hierarchy (a mesh replaced by tpamr_set_mesh, same multigrid options):
 * levels new_start.. of this hierarchy are old_start.. of that one, as the
 * reference's _tree_reuse_suffix! (tree_coarsening.jl:214-226) with
 * _same_tree_topology (:84-121); -1, -1 when nothing was reused. */
int tpamr_hierarchy_reuse(const tpamr_ctx* ctx, int* new_start, int* old_start);
/* ints[6] = ncells, nfaces, nblocks, irregular rows, nnz(A), max leaf level; dbl[2] = omega, omega_P */
int tpamr_level_info(const tpamr_ctx* ctx, int level, int* ints, double* dbl);
/* CSR of the level operator A (integrated, symmetric): pass NULL arrays to query nnz in *nnz */
int tpamr_level_matrix(const tpamr_ctx* ctx, int level, int* nnz, int* rowptr, int* col, double* val);
/* per fine block of `level`: coarse block and quadrant (-1: same block) of the tentative transfer */
int tpamr_level_transfer(const tpamr_ctx* ctx, int level, int* xfer);
int tpamr_level_kappa(const tpamr_ctx* ctx, int level, double* kappa);

/* solve!(solution, solver, rhs; rtol, atol, max_cycles, compatibility)
 * (src/solvers/multigrid/elliptic_solver.jl:628-675): strong rhs in, x holds
 * the initial guess and receives the zero-volume-mean solution.
 * compatibility: 0 = :project, 1 = :error. */
int tpamr_solve(tpamr_ctx* ctx, const double* rhs, double* x, double rtol, double atol, int max_cycles,
                int compatibility, tpamr_stats* stats);
int tpamr_solve_device(tpamr_ctx* ctx, const double* d_rhs, double* d_x, double rtol, double atol, int max_cycles,
                       int compatibility, tpamr_stats* stats);
/* tpamr_solve_device without any allocation during the call: no hipMalloc,
 * hipFree, graph instantiation or allocating library routine. No counterpart
 * in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it). The Krylov
 * state and the reduction partials live in the library's static device
 * scratch of 8 MiB per GPU (held until the process ends), which bounds the
 * finest level at about 5e7 cells: larger meshes are rejected with
 * TPAMR_ESIZE. The hierarchy and vectors are the context's, allocated at
 * set-up (tpamr_set_face_coefficients). Results equal tpamr_solve_device's
 * bit for bit. Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (two host threads, two contexts on one device) race. */
int tpamr_solve_device_noalloc(tpamr_ctx* ctx, const double* d_rhs, double* d_x, double rtol, double atol,
                               int max_cycles, int compatibility, tpamr_stats* stats);

/* One Krylov run at the plugin seam: replaces the nine-argument
 * TwoPhaseAMR._run_krylov!(workspace, matrix, rhs, x0, preconditioner,
 * target, volumes, max_iterations, options) (elliptic_solver.jl:393-452)
 * for a CG workspace: V-cycle-preconditioned CG on the integrated,
 * compatible `rhs` from `x0` until max_i |r_i|/V_i <= target (the
 * StrongResidualCallback), breakdown, or `max_iterations` (0: 2N). `x`
 * receives Krylov.solution, *iterations Krylov.iteration_count; *reason
 * (may be NULL): 1 solved, 2 callback, 3 zero curvature, 4 iteration limit,
 * 5 indefinite preconditioner. */
int tpamr_krylov(tpamr_ctx* ctx, const double* rhs, const double* x0, double target, int max_iterations, double* x,
                 int* iterations, int* reason);

/* ldiv!(z, P, r) of the V-cycle preconditioner (cycles.jl:262-283), integrated r */
int tpamr_vcycle(tpamr_ctx* ctx, const double* r, double* z);
/* integrated operator y = A x (FactorizedEllipticOperator mul!, elliptic.jl:1125-1165) */
int tpamr_apply(tpamr_ctx* ctx, const double* x, double* y);
int tpamr_apply_device(tpamr_ctx* ctx, const double* d_x, double* d_y);

/* two-phase coefficients, src/two_phase/forces.jl */
/* inv.(mixture_density) with the staggered liquid fraction (forces.jl:23-36,69-84,131) */
int tpamr_mixture_face_coefficients(tpamr_ctx* ctx, const double* fraction, double rho_liquid, double rho_gas,
                                    double* kappa);
/* _mixture_boundary_coefficient (forces.jl:103-111): per boundary face
 * inv(rho_l c + rho_g (1 - c)) of its cell's fraction c, the boundary
 * coefficients of the two-phase operator's open sides (host arrays; for
 * tpamr_set_pressure_boundary) */
int tpamr_mixture_boundary_coefficients(tpamr_ctx* ctx, const double* fraction, double rho_liquid, double rho_gas,
                                        double* boundary_coefficient);
/* _surface_tension_jumps! (forces.jl:345-378): formulation 0 = CGFM, 1 = CSF;
 * face_curvature per face (NaN = unresolved) */
int tpamr_surface_tension_jumps(tpamr_ctx* ctx, const double* fraction, const double* face_curvature, double sigma,
                                int formulation, double* jumps);
/* _surface_tension_jumps! with the face curvature of _face_curvature
 * (forces.jl:211-300,345-378): per cell the reconstruction's curvature (NaN =
 * none) and its height-function coverage in [0, 1]; the face takes the
 * LHF-first c(1-c) interpolant of its two cells, or, when both are pure, the
 * c(1-c)-weighted extension of the cell curvatures within 3 h of the face
 * centre. `unresolved` (may be NULL) receives the number of jump faces left
 * without a curvature (their jump is zero). */
int tpamr_surface_tension_jumps_cells(tpamr_ctx* ctx, const double* fraction, const double* cell_curvature,
                                      const double* curvature_quality, double sigma, int formulation, double* jumps,
                                      int* unresolved);
/* _weighted_jump_force! (forces.jl:179-209): accel = sqrt(k) K (sqrt(k) jumps) */
int tpamr_weighted_jump_force(tpamr_ctx* ctx, const double* jumps, double* accel);

/* gravity! (forces.jl:528-580) face jumps rho_f |f| (phi+ - phi-): potential
 * phi = gx x + gy y at the cell centres, moved in mixed cells to the
 * interface centroid; `offsets` (2 per cell, NULL = none) hold that centroid
 * relative to the cell centre in units of the cell spacing (the reference's
 * _interface_centroid of the reconstruction, zero in pure cells); rho_f =
 * rho_l a_l + rho_g a_g with the phase face apertures of the reconstruction
 * (_current_phase_face_apertures!, two_velocity/transport_geometry.jl:491-515),
 * one value per interior face each. Nonzero gravity along a periodic axis is
 * rejected. */
int tpamr_gravity_jumps(tpamr_ctx* ctx, const double* offsets, const double* liquid_aperture,
                        const double* gas_aperture, double rho_liquid, double rho_gas, double gx, double gy,
                        double* jumps);
/* gravity! face accelerations: the jumps above through _weighted_jump_force!
 * (forces.jl:179-209), accel = sqrt(k) K (sqrt(k) jumps) */
int tpamr_gravity(tpamr_ctx* ctx, const double* offsets, const double* liquid_aperture, const double* gas_aperture,
                  double rho_liquid, double rho_gas, double gx, double gy, double* accel);

/* divergence! (operators/advection/cell_advection.jl:205-220); u_boundary may be NULL */
int tpamr_divergence(tpamr_ctx* ctx, const double* u, const double* u_boundary, double* div);
/* weighted_gradient! interior faces (elliptic.jl:1272-1297) */
int tpamr_weighted_gradient(tpamr_ctx* ctx, const double* p, double* g);

/* Halo refresh of a cell field after the projection: fill_halo!(field,
 * ZeroGradient()) or fill_halo!(field, Dirichlet(f)) (src/mesh/variables.jl:205-229)
 * with the schedule of _build_halo_schedule (src/mesh/topology/mesh.jl:944-1037):
 * same-level copy, fine-to-coarse quarter average, coarse-to-fine injection,
 * periodic wrap, physical sides from the nearest interior cell. `cells` in
 * the reference cell order; `haloed` receives, per block, the
 * (bx + 2 halo) x (by + 2 halo) array of CellField.blocks[b] (x index
 * fastest). `physical` = NULL: zero-gradient physical halos; otherwise one
 * value per physical halo entry in schedule order (what a Dirichlet callback
 * returns at the entry centres). tpamr_halo_sizes reports the schedule's
 * entry and physical-entry counts. */
int tpamr_halo_sizes(tpamr_ctx* ctx, int halo, int* entries, int* physical);
/* The schedule itself (host-only contexts too): per entry the output offset
 * dest[e], donors donor[ptr[e] .. ptr[e+1]) with weights, and phys[e] (rank
 * among the physical entries, -1 otherwise); ptr has entries + 1 values. */
int tpamr_halo_schedule(tpamr_ctx* ctx, int halo, int* dest, int* ptr, int* donor, double* weight, int* phys);
int tpamr_fill_halo(tpamr_ctx* ctx, int halo, const double* cells, const double* physical, double* haloed);
/* _pressure_predictor! (single_phase/stepping.jl:144-168): u -= dt G p with
 * the weighted gradient of the pressure operator (interior faces) */
int tpamr_pressure_predictor(tpamr_ctx* ctx, double* u, const double* p, double dt);
/* _project_prepared! (navier_stokes/projection.jl:221-260) on a wall/periodic
 * domain: updates u (interior faces) and p, writes the increment phi */
int tpamr_project(tpamr_ctx* ctx, double* u, double* p, double dt, double tolerance, int max_cycles, double* phi,
                  tpamr_stats* stats);
/* The same with project!'s internal arguments _inherited_divergence and
 * _reference_dt (projection.jl:130-156,189-214): a step shortened from
 * reference_dt (>= dt) keeps (reference_dt - dt) / (dt reference_dt) times the
 * inherited divergence (per cell, strong; NULL: none) in its rhs. Both
 * entry points set stats->compatible as _projection_compatibility!
 * (projection.jl:100-127): dt^2 |removed rhs mean| <= tolerance. */
int tpamr_project_ex(tpamr_ctx* ctx, double* u, double* p, double dt, double tolerance, int max_cycles,
                     const double* inherited_divergence, double reference_dt, double* phi, tpamr_stats* stats);

/* Pressure adjoint (no reference counterpart): for J = w' p with p the
 * zero-mean solution of the current operator, dJ/drhs (strong rhs) and
 * dJ/dkappa per interior face (boundary coefficients held fixed); `p` is the
 * stored forward solution. The adjoints need one connected component. */
int tpamr_adjoint(tpamr_ctx* ctx, const double* w, const double* p, double rtol, int max_cycles, double* dJ_drhs,
                  double* dJ_dkappa, tpamr_stats* stats);

/* Device-memory twins for a time loop that keeps its fields in HBM (every
 * pointer is device memory of the context's GPU; the call returns when the
 * results are written):
 * - tpamr_set_face_coefficients_device: the coefficient refresh from device
 *   face coefficients (after one tpamr_set_face_coefficients);
 * - tpamr_apply_force_device: the force substep u += dt sqrt(k) K (sqrt(k)
 *   jumps) (_apply_phase_forces! with _weighted_jump_force!,
 *   two_phase/stepping.jl:181-201);
 * - tpamr_divergence_device: divergence! of interior face velocities;
 * - tpamr_project_device: tpamr_project_ex on device u (and p += phi when p
 *   is not NULL, phi copied out when not NULL);
 * - tpamr_mixture_face_coefficients_device: tpamr_mixture_face_coefficients
 *   (inv.(mixture_density), forces.jl:23-36,69-84,131) from a device cell
 *   fraction into device face coefficients;
 * - tpamr_gravity_jumps_device: tpamr_gravity_jumps (gravity!,
 *   forces.jl:528-580) from device apertures and offsets (NULL = none).
 * The twins do not scan their inputs for non-finite values; a non-finite
 * coefficient is rejected by the refresh that consumes it. */
int tpamr_set_face_coefficients_device(tpamr_ctx* ctx, const double* d_kappa);
int tpamr_mixture_face_coefficients_device(tpamr_ctx* ctx, const double* d_fraction, double rho_liquid,
                                           double rho_gas, double* d_kappa);
int tpamr_gravity_jumps_device(tpamr_ctx* ctx, const double* d_offsets, const double* d_liquid_aperture,
                               const double* d_gas_aperture, double rho_liquid, double rho_gas, double gx, double gy,
                               double* d_jumps);
int tpamr_apply_force_device(tpamr_ctx* ctx, const double* d_jumps, double dt, double* d_u);
int tpamr_divergence_device(tpamr_ctx* ctx, const double* d_u, double* d_div);
int tpamr_project_device(tpamr_ctx* ctx, double* d_u, double* d_p, double dt, double tolerance, int max_cycles,
                         const double* d_inherited_divergence, double reference_dt, double* d_phi,
                         tpamr_stats* stats);

/* One-velocity momentum step on a uniform mesh periodic in both directions
 * (the reference's uniform staggered schedule,
 * operators/advection/staggered_schedules.jl:226-264,396-458). Other meshes
 * are rejected with TPAMR_EINVAL. The context's pressure operator is the
 * step's (face coefficient 1 / rho). Velocities are interior-face vectors.
 *
 * tpamr_set_momentum: the momentum parts of NavierStokesSystem /
 *   NavierStokesWorkspace (navier_stokes/system.jl, workspace.jl): the
 *   staggered transport workspace (StaggeredVelocityAdvectionWorkspace,
 *   staggered_schedules.jl:925-934; its x/y sweep order starts forward) and
 *   the staggered viscous operator StaggeredDiffusionOperator(mesh;
 *   viscosity) (operators/staggered_diffusion.jl:765-811) with kinematic
 *   viscosity `viscosity`; scheme 0 Upwind, 1 Centred, 2 LaxWendroff (the
 *   default of AdvectionOptions, navier_stokes/options.jl:1-35). A new mesh
 *   (tpamr_set_mesh) drops them. */
int tpamr_set_momentum(tpamr_ctx* ctx, double viscosity, int scheme);
/* advect!(field, advecting, dt, workspace) (staggered_update.jl:265-326) in
 * place on `field`; advecting == field is the one-velocity shorthand
 * advect!(velocity, dt, workspace). Alternates the sweep order after every
 * split step with dt != 0. */
int tpamr_advect(tpamr_ctx* ctx, double* field, const double* advecting, double dt);
/* diffusion!(out, u, operator) (staggered_diffusion.jl:880-900): nu L u */
int tpamr_diffusion(tpamr_ctx* ctx, const double* u, double* out);
/* implicit_midpoint_diffusion!(dst, convected, old, operator, dt, ...)
 * (navier_stokes/implicit_diffusion.jl:916-965): solves
 * (I - dt nu L / 2) dst = convected + dt nu L old / 2 per component, from the
 * initial guess in dst, to the strong residual `tolerance` relative to the
 * right-hand side (relative_to_rhs, elliptic_solver.jl:550-627); stats (x,
 * then y component; may be NULL) as LinearSolverStats. The Krylov method is
 * CG preconditioned by one V-cycle of the staggered tree multigrid
 * (TreeStaggeredGeometricMultigridPreconditioner with the context's
 * multigrid options; a coarsest level above 1024 unknowns, which an odd
 * root layout can leave, falls back to the operator's diagonal). Raises
 * when a component does not converge. */
int tpamr_implicit_diffusion(tpamr_ctx* ctx, double* dst, const double* convected, const double* old, double dt,
                             double tolerance, int max_cycles, tpamr_stats stats[2]);
/* stable_timestep(velocity, advection workspace; cfl)
 * (staggered_update.jl:426-455) */
int tpamr_stable_timestep(tpamr_ctx* ctx, const double* u, double cfl, double* dt_out);
/* step!(velocity, pressure, solver, dt, workspace; force, tolerance,
 * max_cycles) (single_phase/stepping.jl:194-264) of one phase: the old
 * velocity checked against the advective bound at cfl 1 (times
 * 1 + 1e-6 + 32 eps), transported with itself frozen, the body-force face
 * values `force` (sampled at the midpoint time; NULL: none) added times dt,
 * the implicit-midpoint viscous substep, the pressure predictor with the old
 * pressure and the incremental projection (p += phi). `stats` receives the
 * projection's LinearSolverStats; tpamr_step_viscous_stats those of the
 * step's viscous solves. */
int tpamr_step(tpamr_ctx* ctx, double* u, double* p, double dt, const double* force, double tolerance,
               int max_cycles, tpamr_stats* stats);
int tpamr_step_device(tpamr_ctx* ctx, double* d_u, double* d_p, double dt, const double* d_force, double tolerance,
                      int max_cycles, tpamr_stats* stats);
int tpamr_step_viscous_stats(const tpamr_ctx* ctx, tpamr_stats stats[2]);
/* The transport's sweep order flag: 1 when the next split step sweeps x
 * first. */
int tpamr_momentum_sweep_forward(const tpamr_ctx* ctx, int* forward);
/* Cells per level of the viscous solve's staggered multigrid
 * (TreeStaggeredGeometricMultigridPreconditioner's levels,
 * implicit_diffusion.jl:419-541): *count levels, sizes in `cells` when not
 * NULL (at most 64 entries). */
int tpamr_momentum_levels(const tpamr_ctx* ctx, int* count, int* cells);

/* CMDGFM reduced-pressure/jump block solve (the two-velocity projection's
 * linear solve, src/two_phase/projection/cmdgfm_projection.jl). */

/* The solve fields of CMDGFMProjectionStats (cmdgfm_projection.jl:330-360):
 * converged means both the pressure-volume and the jump-row defects are
 * within the tolerance (the post-projection mixture and normal-velocity
 * defects need the interface geometry and are the caller's). */
typedef struct {
  int converged, iterations, krylov_runs, initial_guess_used;
  int pressure_unknowns, jump_unknowns;
  double initial_residual, final_residual;
  double pressure_defect, jump_defect, anchor_defect; /* dt^2-scaled defects */
  double asymmetry, jump_diagonal_margin;
} tpamr_cmdgfm_stats;

/* The assembled system of _assemble_cmdgfm_system!
 * (cmdgfm_projection.jl:1484-1523) on the context's mesh, whose pressure
 * operator (tpamr_set_face_coefficients) is the block's shared hierarchy
 * (assembled.shared_multigrid). CSR of the (np + nx)^2 matrix, columns
 * increasing per row: unknowns 0..np-1 are reduced pressures (cell
 * pressure_cells[i] minus anchor_cells[i], anchors not retained), np.. the
 * mixed-face jumps. pressure_volumes: the retained cells' volumes;
 * jump_lengths: _projection_face_spacing of each mixed face; the omitted
 * (anchor and isolated) rows of the full system over the same unknowns with
 * their rhs and volumes (for _cmdgfm_anchor_residual_defect, nomitted may
 * be 0). Forms the jump Schur complement J - C D^-1 B (_cmdgfm_jump_schur!,
 * cmdgfm_cache.jl:550-610) and its Crout ILU with drop_tolerance
 * (incomplete_lu_preconditioner, solvers/multigrid/cycles.jl:390-396; the
 * reference's default 1e-3, _CMDGFM_BLOCK_DEFAULTS). Works on a host-only
 * context (factor inspection); a new mesh drops the system. */
int tpamr_cmdgfm_set_system(tpamr_ctx* ctx, int np, int nx, const int* rowptr, const int* cols, const double* vals,
                            const int* pressure_cells, const int* anchor_cells, const double* pressure_volumes,
                            const double* jump_lengths, int nomitted, const int* omitted_rowptr,
                            const int* omitted_cols, const double* omitted_vals, const double* omitted_rhs,
                            const double* omitted_volumes, double drop_tolerance);
/* The jump factor of the current system: sizes = {nx, nnz(S), nnz(L),
 * nnz(U off the diagonal)}; S = J - C D^-1 B, L strictly lower (unit
 * diagonal), U strictly upper plus udiag, all CSR by rows. Arrays may be
 * NULL (sizes only). */
int tpamr_cmdgfm_jump_factor(const tpamr_ctx* ctx, int* sizes, int* sptr, int* scol, double* sval, int* lptr,
                             int* lcol, double* lval, int* uptr, int* ucol, double* uval, double* udiag);
/* The solve of _project_cmdgfm_assembled! (cmdgfm_projection.jl:
 * 2201-2318): Krylov scaling by the mesh's largest extent, `density` (the
 * phases' largest) and dt; residual replacement around Krylov.jl's
 * bicgstab! right-preconditioned by the CMDGFM block preconditioner
 * (preconditioned = 0: none), one unpreconditioned retry after a nonfinite
 * increment, at most max_iterations BiCGSTAB iterations in all; warm_start
 * tries the zero-pressure and the previous-increment guesses
 * (_cmdgfm_initial_solution!, previous_increment: a cell vector or NULL).
 * solution (np + nx) receives the increment. */
int tpamr_cmdgfm_solve(tpamr_ctx* ctx, const double* rhs, double* solution, double dt, double density,
                       double tolerance, int max_iterations, int warm_start, const double* previous_increment,
                       int preconditioned, tpamr_cmdgfm_stats* stats);
int tpamr_cmdgfm_solve_device(tpamr_ctx* ctx, const double* d_rhs, double* d_solution, double dt, double density,
                              double tolerance, int max_iterations, int warm_start, const double* d_previous,
                              int preconditioned, tpamr_cmdgfm_stats* stats);

/* Adjoint of one projection step (no reference counterpart). The forward
 * step is tpamr_project from u* = u + dt a (a = the step's force
 * accelerations sqrt(k) K (sqrt(k) jumps)): u' = u* - dt G phi with
 * A phi = V b, b = -div(u*) / dt. Given ubar = dJ/du' (interior faces) it
 * solves A lambda = J^T M ubar (J x = |f| (x+ - x-), M = sqrt(k) K sqrt(k))
 * with the current operator and writes
 *   ubar_star = dJ/du* = ubar - J lambda            (interior faces),
 *   dJ_drhs   = dJ/db  = -dt V lambda               (cells; NULL: skipped),
 *   dJ_dkappa = dJ/dk_f = dt [W_f (K Z)_f + Z_f (K W)_f] / (2 k_f)
 *               with W = sqrt(k) ubar_star, Z = sqrt(k) (jumps - J phi)
 *               (faces; NULL: skipped; needs the checkpointed phi of the
 *               step, `jumps` NULL = no force).
 * rtol is the relative tolerance of the adjoint solve. */
int tpamr_project_adjoint(tpamr_ctx* ctx, const double* ubar, const double* phi, const double* jumps, double dt,
                          double rtol, int max_cycles, double* ubar_star, double* dJ_drhs, double* dJ_dkappa,
                          tpamr_stats* stats);
/* The same on device memory (no reference counterpart). */
int tpamr_project_adjoint_device(tpamr_ctx* ctx, const double* d_ubar, const double* d_phi, const double* d_jumps,
                                 double dt, double rtol, int max_cycles, double* d_ubar_star, double* d_dJ_drhs,
                                 double* d_dJ_dkappa, tpamr_stats* stats);

/* Timing of the finest-level smoother kernel (HIP events, `reps` launches on
 * the context stream); ms per launch in *avg_ms and its algorithmic bytes
 * per launch in *bytes. */
int tpamr_time_smoother(tpamr_ctx* ctx, int reps, double* avg_ms, double* bytes);
/* HBM streaming rate of the smoother's access mix (no reference
 * counterpart): a float4 pass over `reads` (1: a copy; 5: the Jacobi
 * sweep's five reads per write) input arrays of n floats each into one
 * output, `reps` launches timed with HIP events on the context stream after
 * one untimed pass; ms per launch in *avg_ms, bytes read plus written per
 * launch in *bytes. Allocates (reads + 1) n floats for the call. */
int tpamr_time_stream(tpamr_ctx* ctx, long long n, int reads, int reps, double* avg_ms, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
