"""Host-only inspection of the multigrid hierarchy the library builds.

A host-only context (device -1) runs the C++ hierarchy builder without a GPU
so the level sizes, Jacobi weights, level operators and transfer
descriptors can be compared with the reference's published values. Nothing
here computes a solve.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import BlockTreeMesh
from .solver import MultigridOptions


class HostHierarchy:
    """`velocity_boundary` and `boundary_face_coefficient` as PoissonOperator's
    (open pressure sides; the boundary coefficients of a vector
    `face_coefficient` default to the library's extension)."""

    def __init__(self, mesh: BlockTreeMesh, face_coefficient=1.0, multigrid: MultigridOptions | None = None,
                 velocity_boundary=None, boundary_face_coefficient=None):
        from .solver import pressure_boundary_sides
        self.mesh = mesh
        self.ctx = L.Context(-1)
        mesh.attach(self.ctx)
        mg = multigrid or MultigridOptions()
        L.check(L.lib().tpamr_set_multigrid(self.ctx.h, C.byref(mg._c())), self.ctx.h)
        sides = pressure_boundary_sides(velocity_boundary)
        if any(sides):
            ob = np.asarray([int(v) for v in sides], np.int32)
            kb = None if boundary_face_coefficient is None else np.ascontiguousarray(
                np.broadcast_to(np.asarray(boundary_face_coefficient, np.float64), (mesh.nboundary_faces,)))
            L.check(L.lib().tpamr_set_pressure_boundary(self.ctx.h, L.ptr(ob), None if kb is None else L.ptr(kb)),
                    self.ctx.h)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(face_coefficient, np.float64), (mesh.nfaces,)))
        L.check(L.lib().tpamr_set_face_coefficients(self.ctx.h, L.ptr(k)), self.ctx.h)
        self.nlevels = L.lib().tpamr_num_levels(self.ctx.h)

    def level_components(self, level):
        """Per cell its 1-based ungauged-component slot (0: gauged) and per slot
        its root (1-based smallest cell)."""
        from .solver import _level_components
        return _level_components(self.ctx, level, self.level_info(level)["ncells"])

    def level_boundary(self, level):
        """Boundary face coefficients of the level and its flags (singular,
        coarse_augmented, open_faces)."""
        fl = np.zeros(4, np.int32)
        L.check(L.lib().tpamr_level_boundary(self.ctx.h, level, None, L.ptr(fl)), self.ctx.h)
        kb = np.zeros(max(int(fl[3]), 1))
        L.check(L.lib().tpamr_level_boundary(self.ctx.h, level, L.ptr(kb), L.ptr(fl)), self.ctx.h)
        return dict(coefficient=kb[:int(fl[3])], singular=bool(fl[0]), coarse_augmented=bool(fl[1]),
                    open_faces=int(fl[2]))

    def remesh(self, mesh: BlockTreeMesh, face_coefficient=1.0):
        """Attach a new mesh to the same context and build its hierarchy,
        taking over the coarse levels whose tree topology is unchanged
        (`_tree_reuse_suffix!`)."""
        self.mesh = mesh
        mesh.attach(self.ctx)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(face_coefficient, np.float64), (mesh.nfaces,)))
        L.check(L.lib().tpamr_set_face_coefficients(self.ctx.h, L.ptr(k)), self.ctx.h)
        self.nlevels = L.lib().tpamr_num_levels(self.ctx.h)

    def reuse(self):
        """(first reused level of this hierarchy, its level in the previous
        one), or (-1, -1)."""
        a, b = C.c_int(), C.c_int()
        L.check(L.lib().tpamr_hierarchy_reuse(self.ctx.h, C.byref(a), C.byref(b)), self.ctx.h)
        return a.value, b.value

    def level_info(self, level):
        ints = np.zeros(6, np.int32)
        dbl = np.zeros(2)
        L.check(L.lib().tpamr_level_info(self.ctx.h, level, L.ptr(ints), L.ptr(dbl)), self.ctx.h)
        return dict(ncells=int(ints[0]), nfaces=int(ints[1]), nblocks=int(ints[2]), irregular_rows=int(ints[3]),
                    irregular_nnz=int(ints[4]), max_leaf_level=int(ints[5]), omega=float(dbl[0]),
                    omega_p=float(dbl[1]))

    def sizes(self):
        return [self.level_info(l)["ncells"] for l in range(self.nlevels)]

    def level_matrix(self, level):
        import scipy.sparse as sp
        nnz = C.c_int()
        L.check(L.lib().tpamr_level_matrix(self.ctx.h, level, C.byref(nnz), None, None, None), self.ctx.h)
        n = self.level_info(level)["ncells"]
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        L.check(L.lib().tpamr_level_matrix(self.ctx.h, level, C.byref(nnz), L.ptr(rp), L.ptr(col), L.ptr(val)),
                self.ctx.h)
        return sp.csr_matrix((val, col, rp), shape=(n, n))

    def level_transfer(self, level):
        nb = self.level_info(level)["nblocks"]
        x = np.zeros(2 * nb, np.int32)
        L.check(L.lib().tpamr_level_transfer(self.ctx.h, level, L.ptr(x)), self.ctx.h)
        return x.reshape(-1, 2)

    def level_kappa(self, level):
        k = np.zeros(self.level_info(level)["nfaces"])
        L.check(L.lib().tpamr_level_kappa(self.ctx.h, level, L.ptr(k)), self.ctx.h)
        return k

    def tentative_owner(self, level):
        """Fine cell -> coarse cell of the tentative transfer P_t, from the
        block descriptors."""
        info = self.level_info(level)
        bx, by = self.mesh.block_size
        nb = bx * by
        xf = self.level_transfer(level)
        cells = np.arange(info["ncells"])
        blk, loc = cells // nb, cells % nb
        cb, mode = xf[blk, 0], xf[blk, 1]
        i, j = loc % bx, loc // bx
        quad = cb * nb + ((mode >> 1) * (by // 2) + (j >> 1)) * bx + (mode & 1) * (bx // 2) + (i >> 1)
        return np.where(mode < 0, cb * nb + loc, quad)

    def prolongation(self, level):
        """Dense-free smoothed prolongation P = (I - w_P D^-1 A) P_t assembled
        from the level operator and descriptors (for comparisons)."""
        import scipy.sparse as sp
        A = self.level_matrix(level)
        info = self.level_info(level)
        owner = self.tentative_owner(level)
        nc = self.level_info(level + 1)["ncells"]
        Pt = sp.csr_matrix((np.ones(len(owner)), (np.arange(len(owner)), owner)), shape=(len(owner), nc))
        d = A.diagonal()
        dinv = np.divide(1.0, d, out=np.zeros_like(d), where=d > 0)  # zero diagonal: tentative row
        return (Pt - info["omega_p"] * sp.diags(dinv) @ A @ Pt).tocsr()
