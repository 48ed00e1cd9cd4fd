"""CPU FP64 restatement of the CMDGFM reduced-pressure/jump block solve: the
Krylov driver of the two-velocity projection around its block preconditioner
(tree-multigrid reduced-pressure block, ILU of the local jump Schur
complement) on an assembled system.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker. The product
package never imports this module.

Reference files (TwoPhaseAMR.jl):
- reduced-pressure block ``CMDGFMReducedPressurePreconditioner`` ``ldiv!``
  (src/two_phase/projection/cmdgfm_projection.jl:374-402): retained cells
  take their value, each anchor collects minus the values of its component
  in index order, one V-cycle of the shared pressure hierarchy
  (``_apply_vcycle_preconditioner!``, src/solvers/multigrid/cycles.jl:
  262-283), differences to the anchors;
- block ``CMDGFMBlockPreconditioner`` ``ldiv!`` (cmdgfm_projection.jl:
  409-440): pressure block, then the jump rows' rhs minus ``C z_p`` through
  the jump factor;
- jump factor: the local Schur complement ``J - C D^-1 B``
  (``_cmdgfm_jump_schur!``, src/two_phase/projection/cmdgfm_cache.jl:
  540-610, D the diagonal of the reduced pressure block) factored by
  ``incomplete_lu_preconditioner`` (src/solvers/multigrid/cycles.jl:
  390-396), i.e. IncompleteLU.jl 0.2's Crout ILU with drop tolerance tau;
- Krylov scaling ``_cmdgfm_krylov_scaling!`` and the scaled operator and
  preconditioner (cmdgfm_projection.jl:131-180, 215-240);
- the solve loop of ``_project_cmdgfm_assembled!`` (cmdgfm_projection.jl:
  2201-2318): residual replacement with the defects
  ``_cmdgfm_residual_defects`` / ``_cmdgfm_anchor_residual_defect``
  (:1830-1893, src/two_phase/projection/common.jl:328-340), the requested
  relative tolerance, rhs normalisation, one unpreconditioned retry after a
  nonfinite increment, ``_cmdgfm_apply_increment!`` (:1951-1971);
- warm start ``_cmdgfm_initial_solution!`` (:1973-2040) and
  ``_cmdgfm_scaled_residual_norm`` (:1895-1908);
- ``_cmdgfm_matrix_asymmetry`` (cmdgfm_cache.jl:643-685) and
  ``_cmdgfm_jump_diagonal_margin`` (cmdgfm_projection.jl:2042-2061).

Third-party algorithms the reference calls (absent from the reference tree):
- Krylov.jl 0.10 ``bicgstab!`` (Project.toml: Krylov = "0.10") with a right
  preconditioner ``N`` applied by ``ldiv!``, ``M = I``, zero start, shadow
  residual ``c = b``, ``atol = 0``: van der Vorst's BiCGSTAB, stopping when
  ``||r_k|| <= atol + rtol ||b||`` or ``||r_k|| + 1 <= 1``, breakdown when
  ``alpha`` is zero or NaN, ``itmax = 0`` meaning ``2n``;
- IncompleteLU.jl 0.2 ``ilu(A; tau)`` (Project.toml: IncompleteLU = "0.2"):
  row k of U and column k of L are formed from the stored (already dropped)
  factors; an entry is kept when ``|value| >= tau`` (L entries compared
  before the division by ``u_kk``), the diagonal of U always; L is unit
  lower triangular.

Pins (tests/test_oracle_cmdgfm.py): BiCGSTAB iterates equal SciPy's
right-preconditioned BiCGSTAB step for step and converge to the direct
solution; the Crout ILU with tau = 0 is the exact LU; the Schur complement is
the reference test's dense formula (cmdgfm_projection.jl test:344-355); the
reduced-pressure block inverts the gauged pressure block exactly when the
V-cycle is replaced by an exact solve; the block preconditioner with exact
blocks makes BiCGSTAB converge in one iteration on a block-triangular
system.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

EPS = np.finfo(float).eps


# -- third-party restatements ------------------------------------------------

def crout_ilu(a, tau):
    """IncompleteLU.jl 0.2 ``ilu(A; tau)``: returns (L, U) as CSR, L strictly
    lower (unit diagonal implied), U upper with its diagonal."""
    a = sp.csr_matrix(a, dtype=np.float64)
    n = a.shape[0]
    ac = a.tocsc()
    urows = [dict() for _ in range(n)]   # U row i: {col: value}
    lcols = [dict() for _ in range(n)]   # L column k: {row: value}
    lrows = [dict() for _ in range(n)]   # L row i: {col k: value}, i > k
    ucols = [dict() for _ in range(n)]   # U column j: {row i: value}, i < j
    for k in range(n):
        # row k of A on and right of the diagonal, column k below it
        u = {}
        s, e = a.indptr[k], a.indptr[k + 1]
        for p in range(s, e):
            j = int(a.indices[p])
            if j >= k:
                u[j] = u.get(j, 0.0) + float(a.data[p])
        lc = {}
        s, e = ac.indptr[k], ac.indptr[k + 1]
        for p in range(s, e):
            i = int(ac.indices[p])
            if i > k:
                lc[i] = lc.get(i, 0.0) + float(ac.data[p])
        # U_row[k:] -= L[k, i] U[i, k:] for the stored L[k, i], i < k
        for i in sorted(lrows[k]):
            lki = lrows[k][i]
            for j, v in urows[i].items():
                if j >= k:
                    u[j] = u.get(j, 0.0) - lki * v
        # L_col[k+1:] -= U[i, k] L[k+1:, i] for the stored U[i, k], i < k
        if k < n - 1:
            for i in sorted(ucols[k]):
                uik = ucols[k][i]
                for r, v in lcols[i].items():
                    if r > k:
                        lc[r] = lc.get(r, 0.0) - uik * v
        ukk = u.get(k, 0.0)
        if ukk == 0.0:
            raise ZeroDivisionError(f"zero pivot in row {k}")
        for j, v in u.items():
            if abs(v) >= tau or j == k:
                urows[k][j] = v
                if j > k:
                    ucols[j][k] = v
        inv = 1.0 / ukk
        for r, v in lc.items():
            if abs(v) >= tau:
                lcols[k][r] = inv * v
                lrows[r][k] = inv * v
    def to_csr(rows):
        indptr = [0]
        idx, val = [], []
        for r in rows:
            for j in sorted(r):
                idx.append(j)
                val.append(r[j])
            indptr.append(len(idx))
        return sp.csr_matrix((np.asarray(val, np.float64), np.asarray(idx, np.int64), np.asarray(indptr)),
                             shape=(n, n))
    return to_csr(lrows), to_csr(urows)


def ilu_solve(lo, up, b):
    """ldiv! of the ILU factorization: unit-lower forward, upper backward."""
    n = lo.shape[0]
    y = np.array(b, np.float64)
    for i in range(n):
        s, e = lo.indptr[i], lo.indptr[i + 1]
        acc = y[i]
        for p in range(s, e):
            acc -= lo.data[p] * y[lo.indices[p]]
        y[i] = acc
    x = y
    for i in range(n - 1, -1, -1):
        s, e = up.indptr[i], up.indptr[i + 1]
        acc = x[i]
        d = 0.0
        for p in range(s, e):
            j = up.indices[p]
            if j == i:
                d = up.data[p]
            else:
                acc -= up.data[p] * x[j]
        x[i] = acc / d
    return x


def bicgstab(matvec, b, precondition=None, *, atol=0.0, rtol=np.sqrt(EPS), itmax=0):
    """Krylov.jl 0.10 ``bicgstab!`` with ``N = precondition`` (ldiv), M = I,
    zero start and c = b. Returns (x, iterations, status)."""
    b = np.asarray(b, np.float64)
    n = len(b)
    x = np.zeros(n)
    r = b.copy()
    p = r.copy()
    v = np.zeros(n)
    c = b
    rnorm = np.linalg.norm(r)
    if rnorm == 0:
        return x, 0, "zero residual"
    it = 0
    if itmax == 0:
        itmax = 2 * n
    eps_stop = atol + rtol * rnorm
    next_rho = float(c @ r)
    if next_rho == 0:
        return x, 0, "breakdown b'c = 0"
    solved = rnorm <= eps_stop
    tired = it >= itmax
    breakdown = False
    while not (solved or tired or breakdown):
        it += 1
        rho = next_rho
        y = p if precondition is None else precondition(p)
        v = matvec(y)
        alpha = rho / float(c @ v)
        s = r - alpha * v
        x = x + alpha * y
        z = s if precondition is None else precondition(s)
        t = matvec(z)
        omega = float(t @ s) / float(t @ t)
        x = x + omega * z
        r = s - omega * t
        next_rho = float(c @ r)
        beta = (next_rho / rho) * (alpha / omega)
        p = p - omega * v
        p = r + beta * p
        rnorm = np.linalg.norm(r)
        mach = rnorm + 1.0 <= 1.0
        solved = rnorm <= eps_stop or mach
        tired = it >= itmax
        breakdown = alpha == 0 or np.isnan(alpha)
    status = "solved" if solved else ("breakdown" if breakdown else "maximum iterations")
    return x, it, status


# -- the reference's block solve ---------------------------------------------

class BlockSystem:
    """The assembled CMDGFM system over reduced pressures (first np unknowns)
    and mixed-face jumps (last nx):

    ``matrix``: (np+nx)^2; ``pressure_indices``/``pressure_anchors``: cell of
    each reduced pressure and of its component's anchor; ``pressure_volumes``
    the retained cells' volumes; ``jump_lengths`` the defect length of each
    mixed face (``_projection_face_spacing``); the omitted (anchor and
    isolated) rows of the full system over the same unknowns, with their rhs
    and volumes; ``ncells`` of the pressure mesh; ``vcycle(r)`` the shared
    hierarchy's preconditioner on full-cell integrated vectors."""

    def __init__(self, matrix, np_, pressure_indices, pressure_anchors, pressure_volumes, jump_lengths,
                 omitted, omitted_rhs, omitted_volumes, ncells, vcycle, drop_tolerance=1e-3):
        self.a = sp.csr_matrix(matrix, dtype=np.float64)
        self.n = self.a.shape[0]
        self.np = int(np_)
        self.nx = self.n - self.np
        self.idx = np.asarray(pressure_indices, np.int64)
        self.anchor = np.asarray(pressure_anchors, np.int64)
        self.volumes = np.asarray(pressure_volumes, np.float64)
        self.lengths = np.asarray(jump_lengths, np.float64)
        self.omit = sp.csr_matrix(omitted, dtype=np.float64)
        self.omit_rhs = np.asarray(omitted_rhs, np.float64)
        self.omit_volumes = np.asarray(omitted_volumes, np.float64)
        self.ncells = int(ncells)
        self.vcycle = vcycle
        self.tau = float(drop_tolerance)
        npp = self.np
        self.c_block = self.a[npp:, :npp].tocsr()
        if self.nx:
            self.schur = jump_schur(self.a, npp)
            self.lo, self.up = crout_ilu(self.schur, self.tau)
        else:
            self.schur = None

    def reduced_pressure_ldiv(self, s):
        full = np.zeros(self.ncells)
        full[self.idx] = 0.0 + s
        for i in range(self.np):
            full[self.anchor[i]] -= s[i]
        z = self.vcycle(full)
        return z[self.idx] - z[self.anchor]

    def jump_ldiv(self, rhs):
        return ilu_solve(self.lo, self.up, rhs)

    def block_ldiv(self, s):
        npp = self.np
        out = np.zeros(self.n)
        if npp:
            out[:npp] = self.reduced_pressure_ldiv(s[:npp])
        if self.nx:
            jr = s[npp:].copy()
            if npp:
                jr = -1.0 * (self.c_block @ out[:npp]) + 1.0 * jr
            out[npp:] = self.jump_ldiv(jr)
        return out

    # defects (dimensionless volume and jump-row defects)
    def residual_defects(self, res, dt):
        npp = self.np
        v = self.volumes
        pos = v > 0
        pressure = float(np.max(np.abs(res[:npp][pos]) / v[pos])) if np.any(pos) else 0.0
        jump = float(np.max(np.abs(res[npp:]) / self.lengths)) if self.nx else 0.0
        return dt * dt * pressure, dt * dt * jump

    def anchor_defect(self, x, dt):
        if self.omit.shape[0] == 0:
            return 0.0
        prod = self.omit @ x
        pos = self.omit_volumes > 0
        if not np.any(pos):
            return 0.0
        return dt * dt * float(np.max(np.abs(self.omit_rhs[pos] - prod[pos]) / self.omit_volumes[pos]))

    def scaling(self, h, density, dt):
        rows = np.empty(self.n)
        cols = np.empty(self.n)
        rows[:self.np] = (dt / h) * (dt / h)
        cols[:self.np] = density * ((h / dt) * (h / dt))
        rows[self.np:] = (dt * dt) / h
        cols[self.np:] = h / (dt * dt)
        return rows, cols

    def scaled_residual_norm(self, res, h, dt):
        ps = (dt / h) * (dt / h)
        js = (dt * dt) / h
        val = 0.0
        for i, r in enumerate(res):
            val = float(np.hypot(val, (ps if i < self.np else js) * r))
        return val


def jump_schur(a, np_):
    """J - C D^-1 B of the assembled system (D: diagonal of the pressure
    block) in the order of _cmdgfm_jump_schur!: per jump row the J entries,
    then for each C entry the row of D^-1 B, duplicates summed in that
    order."""
    a = sp.csr_matrix(a)
    a.sort_indices()
    n = a.shape[0]
    nx = n - np_
    d = np.zeros(np_)
    for q in range(np_):
        for p in range(a.indptr[q], a.indptr[q + 1]):
            if a.indices[p] == q:
                d[q] = a.data[p]
    if np.any(d == 0):
        raise ZeroDivisionError("zero diagonal in the reduced pressure block")
    indptr, idx, val = [0], [], []
    for k in range(nx):
        r = np_ + k
        ent = {}
        span = range(a.indptr[r], a.indptr[r + 1])
        for p in span:
            if a.indices[p] >= np_:
                ent[int(a.indices[p]) - np_] = float(a.data[p])
        for p in span:
            q = int(a.indices[p])
            if q >= np_:
                continue
            c = float(a.data[p])
            for e in range(a.indptr[q], a.indptr[q + 1]):
                j = int(a.indices[e])
                if j < np_:
                    continue
                v = -c * (float(a.data[e]) / d[q])
                ent[j - np_] = ent[j - np_] + v if (j - np_) in ent else v
        for j in sorted(ent):
            idx.append(j)
            val.append(ent[j])
        indptr.append(len(idx))
    return sp.csr_matrix((np.asarray(val, np.float64), np.asarray(idx, np.int64), np.asarray(indptr)),
                         shape=(nx, nx))


def asymmetry(a):
    a = sp.csr_matrix(a)
    d = (a - a.T).tocoo()
    return float(np.max(np.abs(d.data))) if d.nnz else 0.0


def jump_diagonal_margin(j):
    """min over rows of |J_rr| - sum |J_rc| (c != r), each sum in column order."""
    if j is None or j.shape[0] == 0:
        return np.inf
    j = sp.csr_matrix(j)
    j.sort_indices()
    margin = np.inf
    for r in range(j.shape[0]):
        d = off = 0.0
        for p in range(j.indptr[r], j.indptr[r + 1]):
            if j.indices[p] == r:
                d += abs(float(j.data[p]))
            else:
                off += abs(float(j.data[p]))
        margin = min(margin, d - off)
    return float(margin)


def solve(system: BlockSystem, rhs, *, dt, density, reference_length, tolerance=np.sqrt(EPS),
          max_iterations=500, warm_start=True, previous_increment=None, preconditioned=True):
    """The solve part of ``_project_cmdgfm_assembled!``. Returns (solution,
    stats dict)."""
    s = system
    rhs = np.asarray(rhs, np.float64)
    n = s.n
    a = s.a
    initial_residual = float(np.linalg.norm(rhs))
    solution = np.zeros(n)
    iterations = 0
    runs = 0
    used = False
    correction = rhs.copy()
    if n and initial_residual > 0:
        rows, cols = s.scaling(reference_length, density, dt)
        if warm_start:
            solution, used = _initial_solution(s, rhs, dt, reference_length, previous_increment, preconditioned)
            correction = rhs - a @ solution

        def scaled_matvec(v):
            return rows * (a @ (cols * v))

        def scaled_pc(v):
            return s.block_ldiv(v / rows) / cols

        while iterations < max_iterations:
            pd, jd = s.residual_defects(correction, dt)
            defect = max(pd, jd, s.anchor_defect(solution, dt))
            if not np.isfinite(defect) or defect <= tolerance:
                break
            srhs = rows * correction
            rnorm = float(np.linalg.norm(srhs))
            if not np.isfinite(rnorm) or rnorm == 0:
                break
            requested = min(tolerance, tolerance / max(defect, EPS))
            rhs_scale = 1.0
            if rnorm < 1.0 and rnorm * requested < EPS:
                srhs = srhs / rnorm
                rhs_scale = rnorm
            inc, its, _ = bicgstab(scaled_matvec, srhs, scaled_pc if preconditioned else None, atol=0.0,
                                   rtol=requested, itmax=max_iterations - iterations)
            runs += 1
            iterations += its
            if not np.all(np.isfinite(inc)) and preconditioned and iterations < max_iterations:
                inc, its, _ = bicgstab(scaled_matvec, srhs, None, atol=0.0, rtol=requested,
                                       itmax=max_iterations - iterations)
                runs += 1
                iterations += its
            if not np.all(np.isfinite(inc)) or its == 0:
                break
            updated = solution + cols * (rhs_scale * inc)
            changed = bool(np.any(updated != solution))
            solution = updated
            if not changed:
                break
            correction = rhs - a @ solution
    final_residual = float(np.linalg.norm(correction))
    pd, jd = s.residual_defects(correction, dt)
    ad = s.anchor_defect(solution, dt)
    stats = dict(converged=bool(pd <= tolerance and jd <= tolerance), iterations=iterations, krylov_runs=runs,
                 initial_residual=initial_residual, final_residual=final_residual, pressure_defect=pd,
                 jump_defect=jd, anchor_defect=ad, asymmetry=asymmetry(a),
                 jump_diagonal_margin=jump_diagonal_margin(None if s.nx == 0 else a[s.np:, s.np:]),
                 pressure_unknowns=s.np, jump_unknowns=s.nx, initial_guess_used=used)
    return solution, stats


def _initial_solution(s: BlockSystem, rhs, dt, h, previous, preconditioned):
    npp = s.np
    best_defect = max(s.residual_defects(rhs, dt))
    best_l2 = s.scaled_residual_norm(rhs, h, dt)
    solution = np.zeros(s.n)
    used = False
    for use_previous in (False, True):
        cand = np.zeros(s.n)
        if use_previous:
            if previous is None:
                continue
            prev = np.asarray(previous, np.float64)
            cand[:npp] = prev[s.idx] - prev[s.anchor]
            if not np.any(cand[:npp] != 0):
                continue
        if not np.all(np.isfinite(cand[:npp])):
            continue
        if s.nx:
            if not preconditioned:
                continue
            jr = rhs[npp:].copy()
            jr = -1.0 * (s.c_block @ cand[:npp]) + 1.0 * jr
            cand[npp:] = s.jump_ldiv(jr)
        if not np.all(np.isfinite(cand)):
            continue
        res = rhs - s.a @ cand
        defect = max(s.residual_defects(res, dt))
        l2 = s.scaled_residual_norm(res, h, dt)
        if np.isfinite(defect) and np.isfinite(l2) and defect < 0.95 * best_defect and l2 < best_l2:
            solution = cand.copy()
            best_defect, best_l2 = defect, l2
            used = True
    return solution, used
