"""
    TPAMRShim

Julia side of the MI355X pressure solver: a Krylov workspace that TwoPhaseAMR.jl's
`LinearSolverOptions(krylov = factory)` seam accepts (src/solvers/multigrid/cycles.jl:417-432),
whose `_run_krylov!` method (src/solvers/multigrid/elliptic_solver.jl:393-452) runs
V-cycle-preconditioned CG on the GPU through `libtpamr_hip.so` (include/tpamr.h).

Two entry paths, both selected by the same option:

- `solve!` on an `EllipticSolver` whose Krylov workspace is a `GpuCgWorkspace` (the projection's
  call, navier_stokes/projection.jl:221-260) runs the whole solve on the GPU in the workspace's
  precision, including the FP32 refinement mode (`precision = :fp32_refine`): compatible RHS,
  Krylov runs, true residual, round-off floor, restart loop and stats (`tpamr_solve`).
- Any other caller of the nine-argument `_run_krylov!` seam keeps the reference's own driver
  around it (`_solve_krylov_with_stats!`, `_run_krylov_reliably!`); each Krylov run goes to the
  device with an FP32 V-cycle (FP64 V-cycle for `:fp64`).

Usage:

    using TwoPhaseAMR, TPAMRShim
    options = NavierStokesOptions(; pressure = LinearSolverOptions(
        krylov = TPAMRShim.gpu_krylov(; precision = :fp32_refine), preconditioner = nothing))

Loading the shim needs the built library at `ENV["TPAMR_LIB"]` or `../lib/libtpamr_hip.so`.
"""
module TPAMRShim

using Krylov
using SparseArrays
using TwoPhaseAMR

const LIB = get(ENV, "TPAMR_LIB", joinpath(@__DIR__, "..", "lib", "libtpamr_hip.so"))

const FP64 = Cint(0)
const FP32_VCYCLE = Cint(1)
const FP32_REFINE = Cint(2)

"""Mirror of `tpamr_stats` (include/tpamr.h), field for field; the first eleven
fields are `LinearSolverStats` (src/solvers/multigrid/stats.jl:14-36)."""
struct TpamrStats
    cycles::Cint
    converged::Cint
    compatible::Cint
    reason::Cint
    initial_l2::Float64
    initial_linf::Float64
    final_l2::Float64
    final_linf::Float64
    target_linf::Float64
    residual_scale::Float64
    removed_rhs_mean::Float64
    solve_ms::Float64
    krylov_runs::Cint
    replacements::Cint
end

function _check(ctx::Ptr{Cvoid}, rc::Cint)
    rc == 0 && return nothing
    msg = ctx == C_NULL ? "tpamr: error $rc" :
        unsafe_string(ccall((:tpamr_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    rc == -2 && throw(DimensionMismatch(msg))
    rc in (-1, -3, -6) && throw(ArgumentError(msg))
    throw(ErrorException(msg))
end

"""GPU CG workspace for the `krylov` seam. `m`, `n` are read by `_construct_krylov`."""
mutable struct GpuCgWorkspace
    ctx::Ptr{Cvoid}
    m::Int
    n::Int
    x::Vector{Float64}
    iterations::Int
    mesh_generation::UInt
    mesh_identity::Any
    coefficient::Vector{Float64}
    # open pressure sides and boundary coefficients last given to the context
    sides::Vector{Cint}
    boundary::Vector{Float64}
    # the momentum operators the context holds (kinematic viscosity, scheme;
    # NaN: none since the last mesh)
    momentum::Tuple{Float64, Cint}
end

Krylov.solution(ws::GpuCgWorkspace) = ws.x
Krylov.iteration_count(ws::GpuCgWorkspace) = ws.iterations

function _destroy!(ws::GpuCgWorkspace)
    ws.ctx == C_NULL || ccall((:tpamr_destroy, LIB), Cvoid, (Ptr{Cvoid},), ws.ctx)
    ws.ctx = C_NULL
    return nothing
end

"""
    gpu_krylov(; device = 0, precision = :fp32_vcycle, multigrid = MultigridOptions())

Factory `(matrix, rhs) -> GpuCgWorkspace` for `LinearSolverOptions(krylov = ...)`.
`precision`: `:fp64` (FP64 CG and V-cycle), `:fp32_vcycle` (FP64 CG, FP32 V-cycle) or
`:fp32_refine` (FP64 refinement around an FP32-vector CG with FP64 inner products; `solve!`
only, the `_run_krylov!` seam then runs the `:fp32_vcycle` iteration).
"""
function gpu_krylov(; device::Integer = 0, precision::Symbol = :fp32_vcycle,
        multigrid = TwoPhaseAMR.MultigridOptions())
    mode = precision === :fp64 ? FP64 : precision === :fp32_vcycle ? FP32_VCYCLE :
        precision === :fp32_refine ? FP32_REFINE :
        throw(ArgumentError("precision must be :fp64, :fp32_vcycle or :fp32_refine"))
    return function (matrix, rhs)
        handle = Ref{Ptr{Cvoid}}(C_NULL)
        _check(C_NULL, ccall((:tpamr_create, LIB), Cint, (Ref{Ptr{Cvoid}}, Cint, Cint), handle, device, mode))
        opts = Cint[multigrid.coarse_size, multigrid.pre_sweeps, multigrid.post_sweeps, multigrid.max_levels]
        _check(handle[], ccall((:tpamr_set_multigrid, LIB), Cint, (Ptr{Cvoid}, Ptr{Cint}), handle[], opts))
        n = length(rhs)
        ws = GpuCgWorkspace(handle[], n, n, zeros(n), 0, typemax(UInt), nothing, Float64[], Cint[0, 0, 0, 0],
            Float64[], (NaN, Cint(-1)))
        finalizer(_destroy!, ws)
        return ws
    end
end

# Refined nodes of the tree: every proper ancestor of every leaf. With the
# library's 2:1 balance this reproduces the mesh (tpamr_set_mesh).
function _refined_keys(mesh)
    seen = Set{NTuple{3, Int}}()
    for leaf in mesh.blocks
        level, i, j = leaf.key.level, leaf.key.i, leaf.key.j
        while level > 0
            level -= 1
            i >>= 1
            j >>= 1
            (level, i, j) in seen && break
            push!(seen, (level, i, j))
        end
    end
    keys = Vector{Cint}(undef, 3 * length(seen))
    for (q, (l, i, j)) in enumerate(seen)
        keys[3q - 2] = l
        keys[3q - 1] = i
        keys[3q] = j
    end
    return keys, length(seen)
end

function _sync!(ws::GpuCgWorkspace, operator)
    mesh = operator.mesh
    generation = mesh.topology_cache.generation
    if ws.mesh_identity !== mesh || ws.mesh_generation != generation
        keys, count = _refined_keys(mesh)
        base = Cint[mesh.base_mesh...]
        block = Cint[mesh.block_size...]
        periodic = Cint[mesh.periodic...]
        origin = Float64[mesh.origin...]
        extent = Float64[mesh.extent...]
        GC.@preserve keys base block periodic origin extent _check(ws.ctx, ccall((:tpamr_set_mesh, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}, Ptr{Cint}, Cint, Cint, Ptr{Cint}),
            ws.ctx, base, block, origin, extent, periodic, mesh.max_level, count, keys))
        ws.mesh_identity = mesh
        ws.mesh_generation = generation
        empty!(ws.coefficient)
        ws.momentum = (NaN, Cint(-1))   # tpamr_set_mesh drops the momentum operators
    end
    # open pressure sides and their boundary coefficients (elliptic.jl:76-152):
    # given before the face coefficients they shape
    sides = Cint[operator.boundary_sides...]
    boundary = operator.boundary_face_coefficient
    if ws.sides != sides || (any(!iszero, sides) && ws.boundary != boundary)
        GC.@preserve sides boundary _check(ws.ctx, ccall((:tpamr_set_pressure_boundary, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Cint}, Ptr{Float64}), ws.ctx, sides, boundary))
        ws.sides = sides
        ws.boundary = copy(boundary)
        empty!(ws.coefficient)
    end
    # two-phase steps rewrite the operator's coefficients in place
    # (two_phase/stepping.jl:79-110 with _replace_face_coefficient): compare values
    kappa = operator.face_coefficient
    if ws.coefficient != kappa
        GC.@preserve kappa _check(ws.ctx, ccall((:tpamr_set_face_coefficients, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}), ws.ctx, kappa))
        ws.coefficient = copy(kappa)
    end
    return ws
end

"""
    TwoPhaseAMR._resize_krylov_workspace!(ws::GpuCgWorkspace, number_of_unknowns)

After an AMR regrid `_reconfigure_elliptic_solver!` (elliptic_solver.jl:311-361) resizes the
Krylov workspace, or builds a new one through the factory when this returns `nothing`
(cycles.jl:465-480). The GPU workspace keeps its context: the next `_sync!` attaches the new mesh,
and the library's hierarchy build takes over the previous mesh's coarse levels whose tree topology
is unchanged (`tpamr_hierarchy_reuse`).
"""
function TwoPhaseAMR._resize_krylov_workspace!(ws::GpuCgWorkspace, number_of_unknowns::Int)
    ws.m = ws.n = number_of_unknowns
    resize!(ws.x, number_of_unknowns)
    fill!(ws.x, 0.0)
    ws.iterations = 0
    return ws
end

"""
    TwoPhaseAMR._run_krylov!(ws::GpuCgWorkspace, matrix, rhs, x0, preconditioner, target,
                             volumes, max_iterations, options)

The nine-argument seam method (elliptic_solver.jl:403-419): `rhs` is the integrated,
compatible right-hand side, `target` the strong L-infinity target of the
StrongResidualCallback. The GPU owns the V-cycle, so `preconditioner` is ignored
(configure `preconditioner = nothing`).
"""
function TwoPhaseAMR._run_krylov!(ws::GpuCgWorkspace, matrix, rhs, x0, preconditioner, target,
        volumes, max_iterations, options)
    _sync!(ws, matrix.operator)
    length(rhs) == ws.n || throw(DimensionMismatch("right-hand side length"))
    b = rhs isa Vector{Float64} ? rhs : Vector{Float64}(rhs)
    x_init = x0 isa Vector{Float64} ? x0 : Vector{Float64}(x0)
    iterations = Ref{Cint}(0)
    reason = Ref{Cint}(0)
    GC.@preserve b x_init _check(ws.ctx, ccall((:tpamr_krylov, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Cint, Ptr{Float64}, Ref{Cint}, Ref{Cint}),
        ws.ctx, b, x_init, Float64(target), Cint(max_iterations), ws.x, iterations, reason))
    ws.iterations = Int(iterations[])
    return ws
end

"""
    TwoPhaseAMR.solve!(solution, solver::EllipticSolver{Float64, O, A, GpuCgWorkspace}, rhs;
                       rtol = 1e-10, atol = 0, max_cycles = 100, compatibility = :project)

`solve!` (src/solvers/multigrid/elliptic_solver.jl:628-675) with the whole solve on the GPU:
the library takes the strong RHS and the initial pressure, integrates and projects the RHS,
runs the Krylov runs of `_run_krylov_reliably!` with the strong-residual callback, checks the
true residual with the guarded round-off floor and returns the zero-mean pressure, in the
precision of the workspace (`tpamr_solve`). The stats are those of
`_solve_krylov_with_stats!` (`cycles` counts Krylov iterations over every run). Open sides and
components split by zero coefficients are the library's; covered (inactive) cells are not.
"""
function TwoPhaseAMR.solve!(solution::TwoPhaseAMR.CellField{Float64},
        solver::TwoPhaseAMR.EllipticSolver{Float64, O, A, GpuCgWorkspace},
        right_hand_side::TwoPhaseAMR.CellField{Float64}; rtol = 1.0e-10, atol = 0,
        max_cycles = 100, compatibility = :project) where {O, A}
    solution.mesh === right_hand_side.mesh === solver.operator.mesh ||
        throw(ArgumentError("fields and multigrid solver must share a mesh"))
    0 <= rtol < 1 || throw(ArgumentError("rtol must lie in [0, 1)"))
    atol >= 0 || throw(ArgumentError("atol must be nonnegative"))
    max_cycles >= 0 || throw(ArgumentError("max_cycles must be nonnegative"))
    compatibility in (:project, :error) ||
        throw(ArgumentError("compatibility must be :project or :error"))
    ws = solver.krylov
    _sync!(ws, solver.operator)
    references = solver.operator.references
    TwoPhaseAMR._copy_from_field!(solver.solution, solution, references)
    TwoPhaseAMR._copy_from_field!(solver.right_hand_side, right_hand_side, references)
    x = solver.solution
    b = solver.right_hand_side
    st = Ref{TpamrStats}()
    GC.@preserve x b _check(ws.ctx, ccall((:tpamr_solve, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint, Cint, Ref{TpamrStats}),
        ws.ctx, b, x, Float64(rtol), Float64(atol), Cint(max_cycles),
        Cint(compatibility === :error ? 1 : 0), st))
    TwoPhaseAMR._copy_to_field!(solution, x, references)
    r = st[]
    stats = solver.stats
    stats.cycles = Int(r.cycles)
    stats.converged = r.converged != 0
    stats.compatible = true
    stats.initial_l2 = r.initial_l2
    stats.initial_linf = r.initial_linf
    stats.final_l2 = r.final_l2
    stats.final_linf = r.final_linf
    stats.target_linf = r.target_linf
    stats.residual_scale = r.residual_scale
    stats.removed_rhs_mean = r.removed_rhs_mean
    count = Ref{Cint}(0)
    _check(ws.ctx, ccall((:tpamr_component_means, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ptr{Cint}, Ptr{Float64}),
        ws.ctx, count, C_NULL, C_NULL))
    roots = Vector{Cint}(undef, count[])
    means = Vector{Float64}(undef, count[])
    _check(ws.ctx, ccall((:tpamr_component_means, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ptr{Cint}, Ptr{Float64}),
        ws.ctx, count, roots, means))
    resize!(stats.removed_rhs_component_means, length(means)) .= means
    resize!(stats.removed_rhs_component_anchors, length(roots)) .= roots
    stats.method = :cg
    stats.preconditioner = :tree_geometric
    return stats
end

"""
    gravity!(ws, acceleration, offsets, liquid, gas; density, gravity)

`gravity!` (src/two_phase/forces.jl:528-580) on the GPU: `offsets` holds, per cell,
the interface centroid relative to the cell centre in cell spacings (zero in pure
cells; the reconstruction's `_interface_centroid`), `liquid` and `gas` the phase
face apertures (`_current_phase_face_apertures!`). Writes the face accelerations.
"""
function gravity!(ws::GpuCgWorkspace, acceleration::Vector{Float64}, offsets::Vector{Float64},
        liquid::Vector{Float64}, gas::Vector{Float64}; density::NTuple{2}, gravity::NTuple{2})
    GC.@preserve acceleration offsets liquid gas _check(ws.ctx, ccall((:tpamr_gravity, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64, Ptr{Float64}),
        ws.ctx, offsets, liquid, gas, density[1], density[2], gravity[1], gravity[2], acceleration))
    return acceleration
end

"""
    project_pullback(ws, ubar, phi, jumps, dt; rtol = 1e-10, max_cycles = 200)

Pullback of one force-plus-projection step with the operator's current
coefficients (no reference counterpart): returns `(dJ/du, dJ/db, dJ/dκ, stats)`
for `ubar = dJ/du'`, the step's checkpointed pressure increment `phi` and its
force jumps (`nothing`: no force).
"""
function project_pullback(ws::GpuCgWorkspace, ubar::Vector{Float64}, phi::Vector{Float64}, jumps, dt::Real;
        rtol::Real = 1e-10, max_cycles::Integer = 200)
    ustar = similar(ubar)
    db = Vector{Float64}(undef, length(phi))
    dk = similar(ubar)
    st = Ref{TpamrStats}()
    j = jumps === nothing ? C_NULL : jumps
    GC.@preserve ubar phi jumps ustar db dk _check(ws.ctx, ccall((:tpamr_project_adjoint, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ref{TpamrStats}),
        ws.ctx, ubar, phi, j, Float64(dt), Float64(rtol), Cint(max_cycles), ustar, db, dk, st))
    return ustar, db, dk, st[]
end

# Device-memory twins: a time loop whose fields live in HBM (AMDGPU.jl arrays,
# passed as `pointer(a)`) runs every step without PCIe copies. Each call
# returns when its results are written, and first waits for the work queued on
# the caller's stream (`set_caller_stream!`; the null stream by default).

"""
    set_caller_stream!(ws, stream::Ptr{Cvoid})

The HIP stream on which the caller writes the device twins' inputs (for AMDGPU.jl,
`AMDGPU.stream().stream`): every later device call of `ws` orders itself after the work
queued there (`tpamr_set_caller_stream`).
"""
set_caller_stream!(ws::GpuCgWorkspace, stream::Ptr{Cvoid}) =
    _check(ws.ctx, ccall((:tpamr_set_caller_stream, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ws.ctx, stream))

"""
    step_device!(ws, d_u, d_fraction, d_liquid, d_gas, d_offsets, d_kappa, d_jumps, d_phi, dt;
                 density, gravity, tolerance = 1e-12, max_cycles = 200)

One `gravity!` plus `project!` step on device pointers: mixture coefficients
from the cell fraction (forces.jl:23-36,69-84,131) into `d_kappa`, the
coefficient refresh (`_refresh_pressure_substep!`, two_phase/stepping.jl:73-111),
gravity!'s jumps (forces.jl:528-580) into `d_jumps`, the force substep
(`_apply_phase_forces!`, stepping.jl:181-201) and the projection
(projection.jl:189-260) of `d_u` in place, the pressure increment into `d_phi`
(C_NULL: not kept). `d_offsets` may be C_NULL. Returns the solve stats. The
workspace needs its hierarchy first: one host-side refresh (`_sync!`, which the
Krylov seam runs) of the same mesh.
"""
function step_device!(ws::GpuCgWorkspace, d_u::Ptr{Float64}, d_fraction::Ptr{Float64}, d_liquid::Ptr{Float64},
        d_gas::Ptr{Float64}, d_offsets::Ptr{Float64}, d_kappa::Ptr{Float64}, d_jumps::Ptr{Float64},
        d_phi::Ptr{Float64}, dt::Real; density::NTuple{2}, gravity::NTuple{2}, tolerance::Real = 1e-12,
        max_cycles::Integer = 200)
    _check(ws.ctx, ccall((:tpamr_mixture_face_coefficients_device, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Ptr{Float64}), ws.ctx, d_fraction, density[1], density[2],
        d_kappa))
    _check(ws.ctx, ccall((:tpamr_set_face_coefficients_device, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), ws.ctx,
        d_kappa))
    _check(ws.ctx, ccall((:tpamr_gravity_jumps_device, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64, Ptr{Float64}),
        ws.ctx, d_offsets, d_liquid, d_gas, density[1], density[2], gravity[1], gravity[2], d_jumps))
    _check(ws.ctx, ccall((:tpamr_apply_force_device, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Float64}),
        ws.ctx, d_jumps, Float64(dt), d_u))
    st = Ref{TpamrStats}()
    _check(ws.ctx, ccall((:tpamr_project_device, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint, Ptr{Float64}, Float64, Ptr{Float64},
         Ref{TpamrStats}),
        ws.ctx, d_u, C_NULL, Float64(dt), Float64(tolerance), Cint(max_cycles), C_NULL, Float64(dt), d_phi, st))
    return st[]
end

# ---------------------------------------------------------------------------
# The one-velocity step on a uniform mesh periodic in both directions

"""
    momentum_step!(ws, velocity, pressure, dt; viscosity, scheme = 2, force = nothing,
                   tolerance = sqrt(eps()), max_cycles = 100)

`step!(velocity, pressure, solver, dt, workspace)` of one phase
(src/single_phase/stepping.jl:194-264) on the context of `ws`, whose mesh and
pressure face coefficients `1/rho` `_sync!` has set: `velocity` holds the
interior face velocities (the mesh's face order), `pressure` the cell
pressures (cell order), both updated in place; `viscosity` is the kinematic
viscosity mu/rho, `scheme` 0 Upwind, 1 Centred, 2 LaxWendroff; `force` the body
force's face values at the midpoint time. The momentum operators are set when
the viscosity, the scheme or the mesh changed. Returns the projection's
`TpamrStats`.
"""
function momentum_step!(ws::GpuCgWorkspace, velocity::Vector{Float64}, pressure::Vector{Float64}, dt::Real;
        viscosity::Real, scheme::Integer = 2, force = nothing, tolerance::Real = sqrt(eps()),
        max_cycles::Integer = 100)
    wanted = (Float64(viscosity), Cint(scheme))
    if ws.momentum !== wanted
        _check(ws.ctx, ccall((:tpamr_set_momentum, LIB), Cint, (Ptr{Cvoid}, Float64, Cint), ws.ctx, wanted...))
        ws.momentum = wanted
    end
    f = force === nothing ? C_NULL : force
    st = Ref{TpamrStats}()
    GC.@preserve velocity pressure force _check(ws.ctx, ccall((:tpamr_step, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Float64}, Float64, Cint, Ref{TpamrStats}),
        ws.ctx, velocity, pressure, Float64(dt), f, Float64(tolerance), Cint(max_cycles), st))
    return st[]
end

# ---------------------------------------------------------------------------
# The CMDGFM reduced-pressure/jump block solve

"""Mirror of `tpamr_cmdgfm_stats` (include/tpamr.h), field for field."""
struct TpamrCmdgfmStats
    converged::Cint
    iterations::Cint
    krylov_runs::Cint
    initial_guess_used::Cint
    pressure_unknowns::Cint
    jump_unknowns::Cint
    initial_residual::Float64
    final_residual::Float64
    pressure_defect::Float64
    jump_defect::Float64
    anchor_defect::Float64
    asymmetry::Float64
    jump_diagonal_margin::Float64
end

# rows of A (CSC) as 0-based CSR with increasing columns
function _csr0(A)
    T = SparseArrays.sparse(transpose(A))
    return Cint.(T.colptr .- 1), Cint.(SparseArrays.rowvals(T) .- 1), SparseArrays.nonzeros(T)
end

# the equations the reduced system omits (anchors, isolated cells), as
# _cmdgfm_anchor_residual_defect evaluates them (cmdgfm_projection.jl:1848-1893)
function _cmdgfm_omitted_rows(workspace, assembled, dt)
    op, cache = assembled.pressure_operator, workspace.pressure_cache
    omitted = [c for c in eachindex(op.volumes) if op.volumes[c] > 0 && iszero(cache.pressure_lookup[c])]
    row = Dict(c => k for (k, c) in pairs(omitted))
    mesh = op.mesh
    bx, by = mesh.block_size
    I, J, V = Int[], Int[], Float64[]
    for (k, f) in pairs(assembled.mixed_faces)
        face = mesh.faces[f]
        m = TwoPhaseAMR._cell_index(face.minus, bx, by)
        p = TwoPhaseAMR._cell_index(face.plus, bx, by)
        m == p && continue
        g = assembled.data.coupling[f] * face.area
        haskey(row, m) && (push!(I, row[m]); push!(J, k); push!(V, -g))
        haskey(row, p) && (push!(I, row[p]); push!(J, k); push!(V, g))
    end
    jumps = SparseArrays.sparse(I, J, V, length(omitted), length(assembled.mixed_faces))
    O = hcat(op.stiffness[omitted, assembled.pressure_indices], jumps)
    w, inherited = assembled.inherited_weight, assembled.inherited_cell_flux
    rhs = [cache.cell_flux[c] / dt - (iszero(w) || inherited === nothing ? 0.0 : w * inherited[c]) for c in omitted]
    return O, rhs, op.volumes[omitted]
end

"""
    cmdgfm_solve!(ws, workspace, assembled, solver, dt; tolerance, max_iterations,
                  warm_start = true, previous_increment = nothing, drop_tolerance = 1e-3)

The solve inside `_project_cmdgfm_assembled!` (src/two_phase/projection/
cmdgfm_projection.jl:2201-2318) on the GPU: the assembled system is handed over
(`tpamr_cmdgfm_set_system`, the jump Schur complement and its ILU formed by the
library), the Krylov scaling, the block preconditioner over the context's
pressure hierarchy (its coefficients those of `assembled.pressure_operator`),
BiCGSTAB and the replacement loop run on the device. Returns the solution
(reduced pressures, then jumps) and the `TpamrCmdgfmStats`; the reference's
recovery (pressure increment, velocity correction, mixture and normal-velocity
defects) follows unchanged.
"""
function cmdgfm_solve!(ws::GpuCgWorkspace, workspace, assembled, solver, dt::Real; tolerance::Real,
        max_iterations::Integer, warm_start::Bool = true, previous_increment = nothing,
        drop_tolerance::Real = 1.0e-3)
    op = assembled.pressure_operator
    np, nx = length(assembled.pressure_indices), length(assembled.mixed_faces)
    ptr, col, val = _csr0(assembled.matrix)
    O, orhs, ovol = _cmdgfm_omitted_rows(workspace, assembled, dt)
    optr, ocol, oval = _csr0(O)
    lengths = [TwoPhaseAMR._projection_face_spacing(op.mesh, f) for f in assembled.mixed_faces]
    cells = Cint.(assembled.pressure_indices .- 1)
    anchors = Cint.(assembled.pressure_anchors .- 1)
    volumes = op.volumes[assembled.pressure_indices]
    _check(ws.ctx, ccall((:tpamr_cmdgfm_set_system, LIB), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Ptr{Cint}, Ptr{Cint}, Ptr{Float64},
         Ptr{Float64}, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64),
        ws.ctx, np, nx, ptr, col, val, cells, anchors, volumes, lengths, size(O, 1), optr, ocol, oval, orhs,
        ovol, Float64(drop_tolerance)))
    x = zeros(np + nx)
    prev = previous_increment === nothing ? C_NULL : previous_increment
    st = Ref{TpamrCmdgfmStats}()
    rhs = assembled.right_hand_side
    GC.@preserve rhs x prev _check(ws.ctx, ccall((:tpamr_cmdgfm_solve, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Cint, Cint, Ptr{Float64}, Cint,
         Ref{TpamrCmdgfmStats}),
        ws.ctx, rhs, x, Float64(dt), Float64(maximum(solver.density)), Float64(tolerance), Cint(max_iterations),
        Cint(warm_start), prev, Cint(solver.preconditioner === :block), st))
    return x, st[]
end

end # module
