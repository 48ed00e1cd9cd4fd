"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg as the checker. The product package never imports
this module.

`build_refined_keys` restates the recursive refinement of ``_build_tree``
(TwoPhaseAMR.jl ``src/mesh/topology/mesh.jl:720-736``); the C++ side then
restates balancing, leaf order and the face schedule.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True, capture_output=True)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _LIB.orc_last_error.restype = C.c_char_p
    return _LIB


class Stats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("converged", C.c_int), ("compatible", C.c_int),
                ("pad", C.c_int), ("initial_l2", C.c_double), ("initial_linf", C.c_double),
                ("final_l2", C.c_double), ("final_linf", C.c_double),
                ("target_linf", C.c_double), ("residual_scale", C.c_double),
                ("removed_rhs_mean", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


def set_threads(n):
    """OpenMP threads of the oracle's loops; returns the previous count (the
    oracle's results do not depend on it)."""
    prev = lib().orc_get_threads()
    lib().orc_set_threads(int(n))
    return prev


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def build_refined_keys(base_mesh, block_size, origin, extent, max_level, refine):
    """Keys (level, i, j) for which ``refine(key, origin, extent)`` splits a node,
    visiting every root depth first like ``_build_tree``."""
    layout = (base_mesh[0] // block_size[0], base_mesh[1] // block_size[1])
    root_extent = (extent[0] / layout[0], extent[1] / layout[1])
    keys = []

    def visit(level, i, j, o, w):
        if level >= max_level:
            return
        if refine((level, i, j), o, w):
            keys.append((level, i, j))
            hw = (w[0] / 2, w[1] / 2)
            for cj in (0, 1):
                for ci in (0, 1):
                    visit(level + 1, 2 * i + ci, 2 * j + cj,
                          (o[0] + ci * hw[0], o[1] + cj * hw[1]), hw)

    for j in range(layout[1]):
        for i in range(layout[0]):
            o = (origin[0] if i == 0 else origin[0] + i * root_extent[0],
                 origin[1] if j == 0 else origin[1] + j * root_extent[1])
            visit(0, i, j, o, root_extent)
    return keys


class OMesh:
    def __init__(self, base_mesh=None, block_size=(16, 16), origin=(0.0, 0.0), extent=(1.0, 1.0),
                 periodic=(True, True), max_level=0, refine=None, refined_keys=None):
        L = lib()
        base_mesh = tuple(block_size) if base_mesh is None else tuple(base_mesh)
        if refined_keys is None:
            refined_keys = [] if refine is None else build_refined_keys(
                base_mesh, block_size, origin, extent, max_level, refine)
        keys = np.asarray(refined_keys, dtype=np.int32).reshape(-1, 3)
        base = np.asarray(base_mesh, np.int32)
        bs = np.asarray(block_size, np.int32)
        org = np.asarray(origin, np.float64)
        ext = np.asarray(extent, np.float64)
        per = np.asarray([int(bool(p)) for p in periodic], np.int32)
        out = C.c_void_p()
        _check(L.orc_mesh_new(_iptr(base), _iptr(bs), _dptr(org), _dptr(ext), _iptr(per),
                              int(max_level), len(keys), _iptr(np.ascontiguousarray(keys)),
                              C.byref(out)))
        self._h = out
        self._owned = True
        self.block_size = tuple(int(v) for v in block_size)
        self._load()

    @classmethod
    def _wrap(cls, handle):
        self = cls.__new__(cls)
        self._h = handle
        self._owned = False
        self.block_size = None
        self._load()
        return self

    def _load(self):
        L = lib()
        s = np.zeros(5, np.int32)
        L.orc_mesh_sizes(self._h, _iptr(s))
        self.nblocks, self.ncells, self.nfaces, self.nbfaces, self.max_leaf_level = map(int, s)
        self.block_keys = np.zeros((self.nblocks, 3), np.int32)
        self.block_geom = np.zeros((self.nblocks, 7))
        L.orc_mesh_blocks(self._h, _iptr(self.block_keys), _dptr(self.block_geom))
        fi = np.zeros((self.nfaces, 3), np.int32)
        fd = np.zeros((self.nfaces, 3))
        L.orc_mesh_faces(self._h, _iptr(fi), _dptr(fd))
        self.face_minus, self.face_plus, self.face_axis = fi[:, 0].copy(), fi[:, 1].copy(), fi[:, 2].copy()
        self.face_center, self.face_area = fd[:, :2].copy(), fd[:, 2].copy()
        bi = np.zeros((self.nbfaces, 3), np.int32)
        bd = np.zeros((self.nbfaces, 3))
        L.orc_mesh_bfaces(self._h, _iptr(bi), _dptr(bd))
        self.bface_cell, self.bface_axis, self.bface_side = bi[:, 0].copy(), bi[:, 1].copy(), bi[:, 2].copy()
        self.bface_center, self.bface_area = bd[:, :2].copy(), bd[:, 2].copy()
        nb = None
        self.volumes = np.repeat(self.block_geom[:, 6], self.cells_per_block) if self.nblocks else np.zeros(0)

    @property
    def cells_per_block(self):
        return self.ncells // max(self.nblocks, 1)

    def cell_centers(self):
        nxny = self.cells_per_block
        g = self.block_geom
        # block size from the cell count and the spacing/extent ratio
        nx = int(round(g[0, 2] / g[0, 4]))
        ny = nxny // nx
        ii, jj = np.meshgrid(np.arange(nx), np.arange(ny))
        ii = ii.ravel()
        jj = jj.ravel()
        x = g[:, 0:1] + (ii[None, :] + 0.5) * g[:, 4:5]
        y = g[:, 1:2] + (jj[None, :] + 0.5) * g[:, 5:6]
        return x.ravel(), y.ravel()

    def metric(self):
        L = lib()
        nnz = C.c_int()
        diag = C.c_int()
        _check(L.orc_metric(self._h, C.byref(nnz), C.byref(diag), None, None, None))
        colptr = np.zeros(self.nfaces + 1, np.int32)
        rowval = np.zeros(nnz.value, np.int32)
        nzval = np.zeros(nnz.value)
        _check(L.orc_metric(self._h, C.byref(nnz), C.byref(diag), _iptr(colptr), _iptr(rowval), _dptr(nzval)))
        import scipy.sparse as sp
        return sp.csc_matrix((nzval, rowval, colptr), shape=(self.nfaces, self.nfaces)), bool(diag.value)

    def mixture_coefficients(self, c, rho_l, rho_g):
        k = np.zeros(self.nfaces)
        _check(lib().orc_mixture_coefficients(self._h, _dptr(np.ascontiguousarray(c, np.float64)),
                                              C.c_double(rho_l), C.c_double(rho_g), _dptr(k)))
        return k

    def mixture_boundary_coefficients(self, c, rho_l, rho_g):
        kb = np.zeros(self.nbfaces)
        _check(lib().orc_mixture_boundary_coefficients(self._h, _dptr(np.ascontiguousarray(c, np.float64)),
                                                       C.c_double(rho_l), C.c_double(rho_g), _dptr(kb)))
        return kb

    def cgfm_jumps(self, c, curvature, sigma, csf=False):
        out = np.zeros(self.nfaces)
        curv = np.broadcast_to(np.asarray(curvature, np.float64), (self.nfaces,)).copy()
        _check(lib().orc_cgfm_jumps(self._h, _dptr(np.ascontiguousarray(c, np.float64)), _dptr(curv),
                                    C.c_double(sigma), int(csf), _dptr(out)))
        return out

    def face_curvature(self, c, kappa, quality, faces):
        """`_face_curvature` (src/two_phase/forces.jl:211-300) of the given
        faces; NaN for `nothing`."""
        faces = np.ascontiguousarray(faces, np.int32)
        out = np.zeros(len(faces))
        _check(lib().orc_face_curvature(self._h, _dptr(np.ascontiguousarray(c, np.float64)),
                                        _dptr(np.ascontiguousarray(kappa, np.float64)),
                                        _dptr(np.ascontiguousarray(quality, np.float64)), len(faces), _iptr(faces),
                                        _dptr(out)))
        return out

    def surface_tension_jumps_cells(self, c, kappa, quality, sigma, csf=False):
        """`_surface_tension_jumps!` with `_face_curvature`; returns (jumps,
        unresolved faces)."""
        out = np.zeros(self.nfaces)
        n = C.c_int()
        _check(lib().orc_surface_tension_jumps_cells(self._h, _dptr(np.ascontiguousarray(c, np.float64)),
                                                     _dptr(np.ascontiguousarray(kappa, np.float64)),
                                                     _dptr(np.ascontiguousarray(quality, np.float64)),
                                                     C.c_double(sigma), int(csf), _dptr(out), C.byref(n)))
        return out, n.value

    def gravity_jumps(self, liquid, gas, density, gravity, offsets=None):
        """gravity! face jumps (src/two_phase/forces.jl:528-580)."""
        out = np.zeros(self.nfaces)
        off = None if offsets is None else _dptr(np.ascontiguousarray(offsets, np.float64))
        _check(lib().orc_gravity_jumps(self._h, off, _dptr(np.ascontiguousarray(liquid, np.float64)),
                                       _dptr(np.ascontiguousarray(gas, np.float64)), C.c_double(density[0]),
                                       C.c_double(density[1]), C.c_double(gravity[0]), C.c_double(gravity[1]),
                                       _dptr(out)))
        return out

    def fill_halo(self, cells, halo=1, physical=None):
        """fill_halo! (src/mesh/variables.jl:205-229): per block the haloed
        array, shape (nblocks, by + 2 halo, bx + 2 halo); `physical`: one value
        per physical halo entry in schedule order (None: zero gradient)."""
        bx, by = self.block_size
        out = np.zeros((self.nblocks, by + 2 * halo, bx + 2 * halo))
        n = C.c_int()
        php = None if physical is None else _dptr(np.ascontiguousarray(physical, np.float64))
        _check(lib().orc_fill_halo(self._h, int(halo), _dptr(np.ascontiguousarray(cells, np.float64)), php,
                                   _dptr(out), C.byref(n)))
        self.physical_halo_entries = n.value
        return out

    def divergence(self, u, ub=None):
        out = np.zeros(self.ncells)
        ubp = None if ub is None else _dptr(np.ascontiguousarray(ub, np.float64))
        _check(lib().orc_divergence(self._h, _dptr(np.ascontiguousarray(u, np.float64)), ubp, _dptr(out)))
        return out

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            lib().orc_mesh_free(self._h)
            self._h = None


class OSolver:
    """EllipticSolver(PoissonOperator(mesh; face_coefficient)) with the tree
    geometric multigrid V-cycle (or plain CG with ``multigrid=False``)."""

    def __init__(self, mesh: OMesh, kappa=None, coarse_size=64, pre_sweeps=2, post_sweeps=2,
                 max_levels=32, multigrid=True, volumes=None, open_sides=None, boundary_coefficient=None):
        """``open_sides``: (left, right, bottom, top) open pressure sides of
        ``velocity_boundary`` (elliptic.jl:76-88); ``boundary_coefficient``:
        one value per boundary face (None: extended from ``kappa``)."""
        self.mesh = mesh
        k = np.ones(mesh.nfaces) if kappa is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(kappa, np.float64), (mesh.nfaces,)), np.float64)
        out = C.c_void_p()
        if open_sides is not None:
            ob = np.asarray([int(bool(v)) for v in open_sides], np.int32)
            assert ob.shape == (4,)
            v = None if volumes is None else np.ascontiguousarray(volumes, np.float64)
            bc = None if boundary_coefficient is None else np.ascontiguousarray(
                np.broadcast_to(np.asarray(boundary_coefficient, np.float64), (mesh.nbfaces,)), np.float64)
            self._keep = (v, bc)
            _check(lib().orc_solver_new_open(mesh._h, _dptr(k), None if v is None else _dptr(v), _iptr(ob),
                                             None if bc is None else _dptr(bc), coarse_size, pre_sweeps,
                                             post_sweeps, max_levels, int(multigrid), C.byref(out)))
        elif volumes is None:
            _check(lib().orc_solver_new(mesh._h, _dptr(k), coarse_size, pre_sweeps, post_sweeps,
                                        max_levels, int(multigrid), C.byref(out)))
        else:
            v = np.ascontiguousarray(volumes, np.float64)
            assert v.shape == (mesh.ncells,)
            self._volumes = v
            _check(lib().orc_solver_new_ex(mesh._h, _dptr(k), _dptr(v), coarse_size, pre_sweeps, post_sweeps,
                                           max_levels, int(multigrid), C.byref(out)))
        self._h = out
        self.nlevels = lib().orc_solver_nlevels(self._h)

    def level_info(self, l):
        ints = np.zeros(5, np.int32)
        om = C.c_double()
        lib().orc_solver_level(self._h, l, _iptr(ints), C.byref(om))
        return dict(ncells=int(ints[0]), nfaces=int(ints[1]), nnz_A=int(ints[2]), nnz_P=int(ints[3]),
                    K_diagonal=bool(ints[4]), omega=om.value)

    def level_matrix(self, l, which="A"):
        import scipy.sparse as sp
        info = self.level_info(l)
        if which == "A":
            m = n = info["ncells"]
            nnz = info["nnz_A"]
        else:
            m = info["ncells"]
            n = self.level_info(l + 1)["ncells"]
            nnz = info["nnz_P"]
        colptr = np.zeros(n + 1, np.int32)
        rowval = np.zeros(nnz, np.int32)
        nzval = np.zeros(nnz)
        lib().orc_solver_level_matrix(self._h, l, 0 if which == "A" else 1, _iptr(colptr), _iptr(rowval), _dptr(nzval))
        return sp.csc_matrix((nzval, rowval, colptr), shape=(m, n))

    def level_kappa(self, l):
        k = np.zeros(self.level_info(l)["nfaces"])
        lib().orc_solver_level_kappa(self._h, l, _dptr(k))
        return k

    def level_boundary(self, l):
        """Level l's boundary face coefficients and boundary diagonal, and
        whether its operator is singular and the coarse solve augmented."""
        m = self.level_mesh(l)
        bc = np.zeros(m.nbfaces)
        bd = np.zeros(m.ncells)
        fl = np.zeros(2, np.int32)
        _check(lib().orc_solver_level_boundary(self._h, l, _dptr(bc), _dptr(bd), _iptr(fl)))
        return dict(coefficient=bc, diagonal=bd, singular=bool(fl[0]), coarse_augmented=bool(fl[1]))

    def weighted_gradient_boundary(self, p):
        g = np.zeros(self.mesh.nbfaces)
        _check(lib().orc_weighted_gradient_boundary(self._h, _dptr(np.ascontiguousarray(p, np.float64)),
                                                    _dptr(g)))
        return g

    def project_open(self, u, ub, p, dt, tolerance=np.sqrt(np.finfo(float).eps), max_cycles=100):
        """The projection with boundary face velocities (open sides)."""
        u = np.array(u, np.float64)
        ub = np.array(ub, np.float64)
        p = np.array(p, np.float64)
        phi = np.zeros(self.mesh.ncells)
        st = Stats()
        _check(lib().orc_project_open(self._h, _dptr(u), _dptr(ub), _dptr(p), C.c_double(dt),
                                      C.c_double(tolerance), int(max_cycles), C.byref(st), _dptr(phi)))
        return u, ub, p, phi, st.as_dict()

    def level_mesh(self, l):
        h = C.c_void_p()
        lib().orc_solver_level_mesh(self._h, l, C.byref(h))
        return OMesh._wrap(h)

    def apply(self, x):
        y = np.zeros(self.mesh.ncells)
        _check(lib().orc_apply(self._h, _dptr(np.ascontiguousarray(x, np.float64)), _dptr(y)))
        return y

    def apply_level(self, l, x):
        y = np.zeros(len(x))
        _check(lib().orc_apply_level(self._h, l, _dptr(np.ascontiguousarray(x, np.float64)), _dptr(y)))
        return y

    def vcycle(self, r):
        z = np.zeros(self.mesh.ncells)
        _check(lib().orc_vcycle(self._h, _dptr(np.ascontiguousarray(r, np.float64)), _dptr(z)))
        return z

    def solve(self, rhs, x0=None, rtol=1e-10, atol=0.0, max_cycles=100, compatibility="project"):
        if compatibility not in ("project", "error"):
            raise ValueError("compatibility must be :project or :error")
        x = np.zeros(self.mesh.ncells) if x0 is None else np.array(x0, np.float64)
        st = Stats()
        rm = np.zeros(max(self.mesh.ncells, 1))
        _check(lib().orc_solve(self._h, _dptr(np.ascontiguousarray(rhs, np.float64)), _dptr(x),
                               C.c_double(rtol), C.c_double(atol), int(max_cycles),
                               int(compatibility == "error"), C.byref(st), _dptr(rm)))
        d = st.as_dict()
        _, anchors = self.components(0)
        d["removed_rhs_component_anchors"] = [int(a) + 1 for a in anchors]
        d["removed_rhs_component_means"] = [float(v) for v in rm[:len(anchors)]]
        return x, d

    def components(self, l=0):
        """Level l's component labels (0-based root cell per cell, -1 for an
        inactive cell) and the roots of its ungauged components."""
        n = self.level_info(l)["ncells"]
        labels = np.zeros(n, np.int32)
        cnt = C.c_int()
        _check(lib().orc_solver_components(self._h, l, _iptr(labels), None, C.byref(cnt)))
        anchors = np.zeros(max(cnt.value, 1), np.int32)
        _check(lib().orc_solver_components(self._h, l, None, _iptr(anchors), C.byref(cnt)))
        return labels, anchors[:cnt.value]

    def floor(self, x, b_integrated, assembled=False):
        """`_roundoff_residual_floor!` of iterate x with the integrated rhs:
        the factorized action's (cycles.jl:108-179) or, with `assembled`, the
        assembled matrix's (cycles.jl:80-101)."""
        out = C.c_double()
        _check(lib().orc_solver_floor(self._h, _dptr(np.ascontiguousarray(x, np.float64)),
                                      _dptr(np.ascontiguousarray(b_integrated, np.float64)), int(assembled),
                                      C.byref(out)))
        return out.value

    def weighted_gradient(self, p):
        g = np.zeros(self.mesh.nfaces)
        _check(lib().orc_weighted_gradient(self._h, _dptr(np.ascontiguousarray(p, np.float64)), _dptr(g)))
        return g

    def weighted_jump_force(self, jumps):
        a = np.zeros(self.mesh.nfaces)
        _check(lib().orc_weighted_jump_force(self._h, _dptr(np.ascontiguousarray(jumps, np.float64)), _dptr(a)))
        return a

    def project(self, u, p, dt, tolerance=np.sqrt(np.finfo(float).eps), max_cycles=100, inherited=None,
                reference_dt=None):
        u = np.array(u, np.float64)
        p = np.array(p, np.float64)
        phi = np.zeros(self.mesh.ncells)
        st = Stats()
        inh = None if inherited is None else _dptr(np.ascontiguousarray(inherited, np.float64))
        _check(lib().orc_project_ex(self._h, _dptr(u), _dptr(p), C.c_double(dt), C.c_double(tolerance),
                                    int(max_cycles), C.byref(st), _dptr(phi), inh,
                                    C.c_double(dt if reference_dt is None else reference_dt)))
        return u, p, phi, st.as_dict()

    def adjoint(self, w, p, rtol=1e-12, max_cycles=200):
        db = np.zeros(self.mesh.ncells)
        dk = np.zeros(self.mesh.nfaces)
        lam = np.zeros(self.mesh.ncells)
        st = Stats()
        _check(lib().orc_adjoint(self._h, _dptr(np.ascontiguousarray(w, np.float64)),
                                 _dptr(np.ascontiguousarray(p, np.float64)), C.c_double(rtol),
                                 int(max_cycles), _dptr(db), _dptr(dk), _dptr(lam), C.byref(st)))
        return db, dk, lam, st.as_dict()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_solver_free(self._h)
            self._h = None


def _csc(matrix):
    import scipy.sparse as sp
    A = sp.csc_matrix(np.asarray(matrix, np.float64))
    return (A.shape[0], np.ascontiguousarray(A.indptr, np.int32), np.ascontiguousarray(A.indices, np.int32),
            np.ascontiguousarray(A.data, np.float64))


def assembled_floor(matrix, volumes, x, b, single=False):
    """assembled `_roundoff_residual_floor!` (cycles.jl:80-101) of a dense matrix."""
    n, colptr, rowval, nz = _csc(matrix)
    out = C.c_double()
    _check(lib().orc_assembled_floor(int(single), n, _iptr(colptr), _iptr(rowval), _dptr(nz),
                                     _dptr(np.ascontiguousarray(volumes, np.float64)),
                                     _dptr(np.ascontiguousarray(x, np.float64)),
                                     _dptr(np.ascontiguousarray(b, np.float64)), C.byref(out)))
    return out.value


def scripted_solve(matrix, volumes, b, iterates, *, rtol, atol=0.0, max_cycles=100, relative_to_rhs=False,
                   x0=None, single=False):
    """`_solve_krylov_with_stats!` with a scripted Krylov method (the
    reference's mock workspaces, test/solvers.jl:698-751): Krylov run k leaves
    iterates[min(k, len)] as its solution and reports one iteration. `matrix`
    is a dense array (stored as CSC), `b` the integrated right-hand side;
    `single` runs the driver in Float32. Returns (x, stats, runs)."""
    n, colptr, rowval, nz = _csc(matrix)
    V = np.ascontiguousarray(volumes, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    its = np.ascontiguousarray(np.asarray(iterates, np.float64).reshape(-1, n))
    st = Stats()
    passes = C.c_int(0)
    _check(lib().orc_scripted_solve(int(single), n, _iptr(colptr), _iptr(rowval), _dptr(nz), _dptr(V), _dptr(bb),
                                    _dptr(x), C.c_double(rtol), C.c_double(atol), int(max_cycles),
                                    int(relative_to_rhs), its.shape[0], _dptr(its), C.byref(st), C.byref(passes)))
    return x, st.as_dict(), passes.value


def techref_noise(n):
    """``solv_noise`` (docs/techref/figures/solvers.jl:26) at 1-based indices."""
    idx = np.arange(1, n + 1, dtype=np.float64)
    v = np.sin(12.9898 * idx) * 43758.5453
    return v - np.floor(v) - 0.5


def jacobi_weight(diagonal, row_sums, skip_nonpositive=False, allow_zero_diagonal=False, clamp_weight=True,
                  zero_weight=1.0):
    """`_jacobi_weight!`: returns (weight, diagonal with positive entries inverted)."""
    d = np.array(diagonal, dtype=np.float64)
    r = np.ascontiguousarray(row_sums, dtype=np.float64)
    w = C.c_double()
    _check(lib().orc_jacobi_weight(len(d), _dptr(d), _dptr(r), int(skip_nonpositive), int(allow_zero_diagonal),
                                   int(clamp_weight), C.c_double(zero_weight), C.byref(w)))
    return w.value, d
