"""Synthetic workloads of the pressure-solver benchmarks (seeded, language
neutral: every random number comes from the splitmix64 stream below).

circles(levels=(5, 10), n=64): unit square, walls, 16x16 blocks, uniform
    level 5 (512^2 cells) refined to `levels[1]` within 4 finest cells of the
    union of `n` circles with centres ~U(0.1, 0.9)^2 and radii
    ~U(0.01, 0.05) (seed 20261203). Liquid fraction: clamp(sum of exact
    disk fractions, 0, 1). RHS: -div u of a random smooth wall-compatible
    velocity with 8 Fourier modes per axis and N(0, 1) amplitudes.
single_circle(): unit square, walls, 16x16 blocks, levels 3-7 near the
    circle r=0.15 at (0.5, 0.6) (about 60k cells).
taylor_green(): 2 pi periodic square, 128^2 cells in 16^2 blocks.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import BlockTreeMesh

SEED = 20261203
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed=SEED):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        # Box-Muller on two uniforms (first of the pair)
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def circle_set(n=64, seed=SEED):
    rng = SplitMix64(seed)
    cx, cy, r = [], [], []
    for _ in range(n):
        cx.append(0.1 + 0.8 * rng.uniform())
        cy.append(0.1 + 0.8 * rng.uniform())
        r.append(0.01 + 0.04 * rng.uniform())
    return np.array(cx), np.array(cy), np.array(r)


def _near_any_circle(ox, oy, ex, ey, cx, cy, r, band):
    """Vectorised `solv_near_circle` test against a set of circles."""
    hit = np.zeros(len(ox), dtype=bool)
    x0, x1, y0, y1 = ox, ox + ex, oy, oy + ey
    for k in range(len(cx)):
        nx_ = np.clip(cx[k], x0, x1) - cx[k]
        ny_ = np.clip(cy[k], y0, y1) - cy[k]
        nearest = np.hypot(nx_, ny_)
        fx = np.maximum(np.abs(x0 - cx[k]), np.abs(x1 - cx[k]))
        fy = np.maximum(np.abs(y0 - cy[k]), np.abs(y1 - cy[k]))
        farthest = np.hypot(fx, fy)
        hit |= (nearest - band <= r[k]) & (r[k] <= farthest + band)
    return hit


def circles_mesh(levels=(5, 10), n=64, seed=SEED, block=16):
    cx, cy, r = circle_set(n, seed)
    lo, hi = levels
    band = 4.0 / (block * (1 << hi))

    def refine_many(level, i, j, ox, oy, ex, ey):
        out = level < lo
        rest = ~out
        if rest.any():
            out[rest] = _near_any_circle(ox[rest], oy[rest], ex[rest], ey[rest], cx, cy, r, band)
        return out

    mesh = BlockTreeMesh(base_mesh=(block, block), block_size=(block, block), periodic=(False, False),
                         max_level=hi, refine_many=refine_many)
    return mesh, (cx, cy, r)


def single_circle_mesh(levels=(3, 7), centre=(0.5, 0.6), radius=0.15, block=16):
    lo, hi = levels
    band = 4.0 / (block * (1 << hi))
    cx, cy, r = np.array([centre[0]]), np.array([centre[1]]), np.array([radius])

    def refine_many(level, i, j, ox, oy, ex, ey):
        out = level < lo
        rest = ~out
        if rest.any():
            out[rest] = _near_any_circle(ox[rest], oy[rest], ex[rest], ey[rest], cx, cy, r, band)
        return out

    mesh = BlockTreeMesh(base_mesh=(block, block), block_size=(block, block), periodic=(False, False),
                         max_level=hi, refine_many=refine_many)
    return mesh, (cx, cy, r)


def taylor_green_mesh():
    return BlockTreeMesh(base_mesh=(128, 128), block_size=(16, 16), extent=(2 * math.pi, 2 * math.pi),
                         periodic=(True, True))


# exact disk / cell-rectangle areas ------------------------------------------

def _seg(X, r):
    X = np.clip(X, -r, r)
    return 0.5 * (X * np.sqrt(np.maximum(r * r - X * X, 0)) + r * r * np.arcsin(X / r))


def _quad_area(a, b, r):
    a, b = np.broadcast_arrays(np.minimum(a, r), b)
    c = np.sqrt(np.maximum(r * r - b * b, 0))

    def integ(lo, hi, two):
        lo = np.maximum(lo, -r)
        hi = np.maximum(np.minimum(hi, a), lo)
        return 2 * (_seg(hi, r) - _seg(lo, r)) if two else b * (hi - lo) + (_seg(hi, r) - _seg(lo, r))

    mr = -r * np.ones_like(a)
    pr = r * np.ones_like(a)
    full = 2 * (_seg(a, r) - _seg(mr, r))
    v_pos = integ(mr, -c, True) + integ(-c, c, False) + integ(c, pr, True)
    v_neg = integ(-c, c, False)
    return np.where(b >= r, full, np.where(b >= 0, v_pos, np.where(b > -r, v_neg, 0.0)))


def disk_fraction(x, y, dx, dy, centre, radius):
    cx, cy = centre
    a0, a1 = x - dx / 2 - cx, x + dx / 2 - cx
    b0, b1 = y - dy / 2 - cy, y + dy / 2 - cy
    area = (_quad_area(a1, b1, radius) - _quad_area(a0, b1, radius) - _quad_area(a1, b0, radius)
            + _quad_area(a0, b0, radius))
    return np.clip(area / (dx * dy), 0.0, 1.0)


def liquid_fraction(mesh: BlockTreeMesh, circles):
    """clamp(sum of exact disk fractions, 0, 1) per cell (blocks culled per circle)."""
    cx, cy, r = circles
    nb = mesh.cells_per_block
    k = mesh.block_keys.astype(np.float64)
    bx0 = mesh.origin[0] + k[:, 1] * mesh.block_extent[:, 0]
    by0 = mesh.origin[1] + k[:, 2] * mesh.block_extent[:, 1]
    bx1 = bx0 + mesh.block_extent[:, 0]
    by1 = by0 + mesh.block_extent[:, 1]
    nx, ny = mesh.block_size
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    c = np.zeros(mesh.ncells)
    for q in range(len(cx)):
        near = np.hypot(np.clip(cx[q], bx0, bx1) - cx[q], np.clip(cy[q], by0, by1) - cy[q]) < r[q]
        far = np.hypot(np.maximum(np.abs(bx0 - cx[q]), np.abs(bx1 - cx[q])),
                       np.maximum(np.abs(by0 - cy[q]), np.abs(by1 - cy[q]))) <= r[q]
        full = np.nonzero(far)[0]
        for b in full:
            c[b * nb:(b + 1) * nb] += 1.0
        cut = np.nonzero(near & ~far)[0]
        if len(cut) == 0:
            continue
        sx = mesh.block_spacing[cut, 0][:, None]
        sy = mesh.block_spacing[cut, 1][:, None]
        x = (bx0[cut][:, None] + (ii[None, :] + 0.5) * sx).ravel()
        y = (by0[cut][:, None] + (jj[None, :] + 0.5) * sy).ravel()
        dx = np.broadcast_to(sx, (len(cut), nb)).ravel()
        dy = np.broadcast_to(sy, (len(cut), nb)).ravel()
        idx = (cut[:, None] * nb + np.arange(nb)[None, :]).ravel()
        c[idx] += disk_fraction(x, y, dx, dy, (cx[q], cy[q]), r[q])
    return np.clip(c, 0.0, 1.0)


def face_centers(mesh: BlockTreeMesh):
    x, y = mesh.cell_centers()
    dx, dy = mesh.cell_spacing()
    m, p, ax = mesh.face_minus, mesh.face_plus, mesh.face_axis
    fx = np.where(ax == 0, x[m] + dx[m] / 2,
                  0.5 * (np.maximum(x[m] - dx[m] / 2, x[p] - dx[p] / 2) + np.minimum(x[m] + dx[m] / 2, x[p] + dx[p] / 2)))
    fy = np.where(ax == 1, y[m] + dy[m] / 2,
                  0.5 * (np.maximum(y[m] - dy[m] / 2, y[p] - dy[p] / 2) + np.minimum(y[m] + dy[m] / 2, y[p] + dy[p] / 2)))
    return fx, fy


def smooth_velocity(mesh: BlockTreeMesh, modes=8, seed=SEED + 1, device=None, chunk=1 << 22):
    """Face-normal velocity of u = sum a_kl sin(pi k x) cos(pi l y),
    v = sum b_kl cos(pi k x) sin(pi l y) (zero normal velocity on the walls),
    evaluated in float64 with torch on `device` (e.g. 'cuda:0') or numpy."""
    rng = SplitMix64(seed)
    a = np.array([[rng.normal() for _ in range(modes)] for _ in range(modes)])
    b = np.array([[rng.normal() for _ in range(modes)] for _ in range(modes)])
    fx, fy = face_centers(mesh)
    ax = mesh.face_axis
    u = np.zeros(mesh.nfaces)
    kk = np.arange(1, modes + 1) * math.pi
    if device is not None:
        import torch
        A = torch.as_tensor(a, device=device)
        B = torch.as_tensor(b, device=device)
        K = torch.as_tensor(kk, device=device)
        for s0 in range(0, mesh.nfaces, chunk):
            sl = slice(s0, min(mesh.nfaces, s0 + chunk))
            X = torch.as_tensor(fx[sl], device=device)[:, None] * K[None, :]
            Y = torch.as_tensor(fy[sl], device=device)[:, None] * K[None, :]
            ux = ((torch.sin(X) @ A) * torch.cos(Y)).sum(dim=1)
            uy = ((torch.cos(X) @ B) * torch.sin(Y)).sum(dim=1)
            axis = torch.as_tensor(ax[sl], device=device)
            u[sl] = torch.where(axis == 0, ux, uy).cpu().numpy()
        return u
    for s0 in range(0, mesh.nfaces, chunk):
        sl = slice(s0, min(mesh.nfaces, s0 + chunk))
        X = np.outer(fx[sl], kk)
        Y = np.outer(fy[sl], kk)
        ux = np.sum((np.sin(X) @ a) * np.cos(Y), axis=1)
        uy = np.sum((np.cos(X) @ b) * np.sin(Y), axis=1)
        u[sl] = np.where(ax[sl] == 0, ux, uy)
    return u
