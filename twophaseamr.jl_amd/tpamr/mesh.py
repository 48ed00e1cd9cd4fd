"""BlockTreeMesh: the host-side description of a block quadtree.

Mirrors ``BlockTreeMesh(; base_mesh, block_size, origin, extent, periodic,
max_level, refine)`` of TwoPhaseAMR.jl (src/mesh/topology/mesh.jl:1040-1121).
The refine predicate is evaluated here, level by level, exactly as the
reference's ``_build_tree`` (mesh.jl:720-736) visits nodes: a node below
``max_level`` is split when ``refine(key, origin, extent)`` is true. The set
of split nodes is handed to the C++ library, which 2:1-balances the tree and
builds leaves, faces and the hierarchy.

``refine`` may be the reference's scalar form ``refine(key, origin, extent)``
or, for large meshes, a vectorised ``refine_many(level, i, j, ox, oy, ex, ey)``
returning a boolean array.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


class BlockKey(tuple):
    """Dyadic quadtree address (level, i, j), zero-based (mesh.jl:2-6)."""

    __slots__ = ()

    def __new__(cls, level, i, j):
        return super().__new__(cls, (int(level), int(i), int(j)))

    level = property(lambda self: self[0])
    i = property(lambda self: self[1])
    j = property(lambda self: self[2])


def _refined_keys(base_mesh, block_size, origin, extent, max_level, refine, refine_many):
    layout = (base_mesh[0] // block_size[0], base_mesh[1] // block_size[1])
    rex = (extent[0] / layout[0], extent[1] / layout[1])
    jj, ii = np.meshgrid(np.arange(layout[1]), np.arange(layout[0]), indexing="ij")
    lv = 0
    ci = ii.ravel().astype(np.int64)
    cj = jj.ravel().astype(np.int64)
    ox = np.where(ci == 0, float(origin[0]), origin[0] + ci * rex[0])
    oy = np.where(cj == 0, float(origin[1]), origin[1] + cj * rex[1])
    wx, wy = rex
    out = []
    while lv < max_level and len(ci) and (refine is not None or refine_many is not None):
        if refine_many is not None:
            split = np.asarray(refine_many(np.full(len(ci), lv), ci, cj, ox, oy,
                                           np.full(len(ci), wx), np.full(len(ci), wy)), dtype=bool)
        else:
            split = np.fromiter((bool(refine(BlockKey(lv, a, b), (x, y), (wx, wy)))
                                 for a, b, x, y in zip(ci.tolist(), cj.tolist(), ox.tolist(), oy.tolist())),
                                dtype=bool, count=len(ci))
        if not split.any():
            break
        si, sj, sx, sy = ci[split], cj[split], ox[split], oy[split]
        out.append(np.stack([np.full(len(si), lv), si, sj], axis=1))
        hx, hy = wx / 2, wy / 2
        # children in RegionTrees order; their origins are parent + c * half
        kids_i, kids_j, kids_x, kids_y = [], [], [], []
        for cy_ in (0, 1):
            for cx_ in (0, 1):
                kids_i.append(2 * si + cx_)
                kids_j.append(2 * sj + cy_)
                kids_x.append(sx + cx_ * hx)
                kids_y.append(sy + cy_ * hy)
        ci = np.concatenate(kids_i)
        cj = np.concatenate(kids_j)
        ox = np.concatenate(kids_x)
        oy = np.concatenate(kids_y)
        wx, wy = hx, hy
        lv += 1
    if not out:
        return np.zeros((0, 3), np.int32)
    return np.ascontiguousarray(np.concatenate(out), dtype=np.int32)


class BlockTreeMesh:
    """Serial 2-D forest of fixed-size blocks (see module docstring)."""

    def __init__(self, *, block_size=(16, 16), base_mesh=None, halo=1, origin=(0.0, 0.0), extent=(1.0, 1.0),
                 periodic=(True, True), max_level=0, refine=None, refine_many=None, refined_keys=None):
        base_mesh = tuple(block_size) if base_mesh is None else tuple(base_mesh)
        if not all(isinstance(v, (int, np.integer)) and not isinstance(v, bool) for v in base_mesh):
            raise ValueError("base mesh dimensions must be integers")
        nx, ny = (int(v) for v in block_size)
        if nx < 2 or ny < 2:
            raise ValueError("block dimensions must be at least two")
        if nx % 2 or ny % 2:
            raise ValueError("block dimensions must be even for dyadic transfer")
        if base_mesh[0] < nx or base_mesh[1] < ny:
            raise ValueError("base mesh dimensions must be at least the block dimensions")
        if base_mesh[0] % nx or base_mesh[1] % ny:
            raise ValueError("base mesh dimensions must be integer multiples of block dimensions")
        if halo < 1:
            raise ValueError("halo width must be positive")
        if max_level < 0:
            raise ValueError("max_level must be nonnegative")
        if not all(e > 0 for e in extent):
            raise ValueError("domain extents must be positive")
        self.base_mesh = (int(base_mesh[0]), int(base_mesh[1]))
        self.block_size = (nx, ny)
        self.halo = int(halo)
        self.origin = (float(origin[0]), float(origin[1]))
        self.extent = (float(extent[0]), float(extent[1]))
        self.periodic = (bool(periodic[0]), bool(periodic[1]))
        self.max_level = int(max_level)
        if refined_keys is None:
            refined_keys = _refined_keys(self.base_mesh, self.block_size, self.origin, self.extent,
                                         self.max_level, refine, refine_many)
        self.refined_keys = np.ascontiguousarray(np.asarray(refined_keys, dtype=np.int32).reshape(-1, 3))
        self._host = L.Context(-1)
        self.attach(self._host)
        self._load()

    def attach(self, ctx: "L.Context"):
        """Describe this mesh to a library context (tpamr_set_mesh)."""
        lib = L.lib()
        base = np.asarray(self.base_mesh, np.int32)
        bs = np.asarray(self.block_size, np.int32)
        org = np.asarray(self.origin, np.float64)
        ext = np.asarray(self.extent, np.float64)
        per = np.asarray([int(p) for p in self.periodic], np.int32)
        L.check(lib.tpamr_set_mesh(ctx.h, L.ptr(base), L.ptr(bs), L.ptr(org), L.ptr(ext), L.ptr(per),
                                   self.max_level, len(self.refined_keys), L.ptr(self.refined_keys)), ctx.h)

    def _load(self):
        lib = L.lib()
        h = self._host.h
        s = np.zeros(5, np.int32)
        L.check(lib.tpamr_mesh_sizes(h, L.ptr(s)), h)
        self.nblocks, self.ncells, self.nfaces, self.nboundary_faces, self.max_leaf_level = (int(v) for v in s)
        self.block_keys = np.zeros((self.nblocks, 3), np.int32)
        L.check(lib.tpamr_mesh_blocks(h, L.ptr(self.block_keys)), h)
        self.face_minus = np.zeros(self.nfaces, np.int32)
        self.face_plus = np.zeros(self.nfaces, np.int32)
        self.face_axis = np.zeros(self.nfaces, np.int32)
        self.face_area = np.zeros(self.nfaces)
        L.check(lib.tpamr_mesh_faces(h, L.ptr(self.face_minus), L.ptr(self.face_plus), L.ptr(self.face_axis),
                                     L.ptr(self.face_area)), h)
        nbf = self.nboundary_faces
        self.bface_cell = np.zeros(nbf, np.int32)
        self.bface_axis = np.zeros(nbf, np.int32)
        self.bface_side = np.zeros(nbf, np.int32)
        self.bface_area = np.zeros(nbf)
        L.check(lib.tpamr_mesh_boundary_faces(h, L.ptr(self.bface_cell), L.ptr(self.bface_axis),
                                              L.ptr(self.bface_side), L.ptr(self.bface_area)), h)
        layout = (self.base_mesh[0] // self.block_size[0], self.base_mesh[1] // self.block_size[1])
        lev = self.block_keys[:, 0].astype(np.float64)
        self.block_extent = np.stack([self.extent[0] / layout[0] / 2.0 ** lev,
                                      self.extent[1] / layout[1] / 2.0 ** lev], axis=1)
        self.block_spacing = self.block_extent / np.asarray(self.block_size, np.float64)
        self.block_volume = self.block_spacing[:, 0] * self.block_spacing[:, 1]

    @property
    def cells_per_block(self):
        return self.block_size[0] * self.block_size[1]

    @property
    def volumes(self):
        return np.repeat(self.block_volume, self.cells_per_block)

    def cell_centers(self):
        """Cell centres in the reference cell order."""
        nx, ny = self.block_size
        layout = (self.base_mesh[0] // nx, self.base_mesh[1] // ny)
        k = self.block_keys.astype(np.float64)
        ox = self.origin[0] + k[:, 1] * self.block_extent[:, 0]
        oy = self.origin[1] + k[:, 2] * self.block_extent[:, 1]
        jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        x = ox[:, None] + (ii.ravel()[None, :] + 0.5) * self.block_spacing[:, 0:1]
        y = oy[:, None] + (jj.ravel()[None, :] + 0.5) * self.block_spacing[:, 1:2]
        return x.ravel(), y.ravel()

    def cell_spacing(self):
        return (np.repeat(self.block_spacing[:, 0], self.cells_per_block),
                np.repeat(self.block_spacing[:, 1], self.cells_per_block))

    def __repr__(self):
        return (f"BlockTreeMesh(base_mesh={self.base_mesh}, block_size={self.block_size}, "
                f"blocks={self.nblocks}, cells={self.ncells}, faces={self.nfaces})")
