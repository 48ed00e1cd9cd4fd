"""Mesh and right-hand-side fixtures shared by the tests.

The techref fixtures follow docs/techref/figures/solvers.jl:13-67 of
TwoPhaseAMR.jl (walls, 8x8 blocks, drop circle at (0.513, 0.468), radius
0.2371, refinement band 0.02, noise rhs frac(43758.5453 sin(12.9898 c)) - 1/2).
"""
import math

import numpy as np

DROP_CENTRE = (0.513, 0.468)
DROP_RADIUS = 0.2371
DROP_BAND = 0.02


def near_circle(origin, extent, centre=DROP_CENTRE, radius=DROP_RADIUS, band=DROP_BAND):
    """`solv_near_circle` (docs/techref/figures/solvers.jl:36-43)."""
    xs = (origin[0], origin[0] + extent[0])
    ys = (origin[1], origin[1] + extent[1])
    cx, cy = centre
    nearest = math.hypot(min(max(cx, xs[0]), xs[1]) - cx, min(max(cy, ys[0]), ys[1]) - cy)
    farthest = max(math.hypot(x - cx, y - cy) for x in xs for y in ys)
    return nearest - band <= radius <= farthest + band


def techref_noise(n):
    idx = np.arange(1, n + 1, dtype=np.float64)
    v = np.sin(12.9898 * idx) * 43758.5453
    return v - np.floor(v) - 0.5


MESHES = {
    "uniform64": dict(base_mesh=(64, 64), block_size=(8, 8), periodic=(False, False)),
    "uniform128": dict(base_mesh=(128, 128), block_size=(8, 8), periodic=(False, False)),
    "amr32": dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(False, False), max_level=2,
                  refine=lambda k, o, e: near_circle(o, e)),
    "amr64": dict(base_mesh=(64, 64), block_size=(8, 8), periodic=(False, False), max_level=2,
                  refine=lambda k, o, e: near_circle(o, e)),
    "periodic_tg": dict(base_mesh=(64, 64), block_size=(16, 16), periodic=(True, True),
                        extent=(2 * math.pi, 2 * math.pi)),
    "amr_mixed": dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(True, False), max_level=3,
                      refine=lambda k, o, e: near_circle(o, e, band=0.01)),
    "pruned_tree": dict(block_size=(8, 8), max_level=3,
                        refine=lambda k, o, e: k[0] == 0 or (k[0] < 3 and o[0] < 0.5 and o[1] < 0.5)),
    "circle16": dict(base_mesh=(16, 16), block_size=(16, 16), periodic=(False, False), max_level=4,
                     refine=lambda k, o, e: k[0] < 2 or near_circle(o, e, centre=(0.5, 0.6), radius=0.15,
                                                                     band=4.0 / (16 * 16))),
    # the smallest and the largest block tiles of the kernels
    "amr_block4": dict(base_mesh=(32, 32), block_size=(4, 4), periodic=(False, True), max_level=2,
                       refine=lambda k, o, e: near_circle(o, e)),
    "amr_block32": dict(base_mesh=(64, 64), block_size=(32, 32), periodic=(True, False), max_level=2,
                        refine=lambda k, o, e: k[0] == 0 and k[1] == 0 or (k[0] == 1 and o[1] < 0.3)),
}


def _seg(X, r):
    X = np.clip(X, -r, r)
    return 0.5 * (X * np.sqrt(np.maximum(r * r - X * X, 0)) + r * r * np.arcsin(X / r))


def _quad_area(a, b, r):
    """Area of {x^2 + y^2 <= r^2, x <= a, y <= b} (closed form)."""
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
    out = np.where(b >= r, full, np.where(b >= 0, v_pos, np.where(b > -r, v_neg, 0.0)))
    return out


def disk_fraction(x, y, dx, dy, centre, radius):
    """Exact liquid fraction of the cells [x +- dx/2] x [y +- dy/2] inside a disk."""
    cx, cy = centre
    a0, a1 = x - dx / 2 - cx, x + dx / 2 - cx
    b0, b1 = y - dy / 2 - cy, y + dy / 2 - cy
    area = (_quad_area(a1, b1, radius) - _quad_area(a0, b1, radius) - _quad_area(a1, b0, radius)
            + _quad_area(a0, b0, radius))
    return np.clip(area / (dx * dy), 0.0, 1.0)


# Open pressure sides (left, right, bottom, top): meshes, open sides, and the
# boundary coefficient kind ("material": 0.8 + 0.2 y + 0.1 x at the boundary
# face centres; "extended": the interior coefficients extended to the open
# faces).
OPEN_LR = (True, True, False, False)


def _kw(refine, periodic):
    return dict(base_mesh=(16, 8), block_size=(4, 4), extent=(2.0, 1.0), periodic=periodic, max_level=1,
                refine=refine)


OPEN_CASES = {
    "outflow_column": (_kw(lambda k, o, e: k[0] == 0 and k[1] == 3, (False, False)), OPEN_LR, "material"),
    "outflow_column_periodic_y": (_kw(lambda k, o, e: k[0] == 0 and k[1] == 3, (False, True)), OPEN_LR, "material"),
    "uniform": (_kw(lambda k, o, e: False, (False, False)), OPEN_LR, "material"),
    "corner_refined": (_kw(lambda k, o, e: k[0] == 0 and k[1] == 0 and k[2] == 0, (False, False)),
                       (True, True, True, True), "material"),
    "amr64_all_open": (dict(base_mesh=(64, 64), block_size=(8, 8), periodic=(False, False), max_level=2,
                            refine=lambda k, o, e: (o[0] < 0.3 and o[1] < 0.4) or (k[0] == 0 and o[0] > 0.7)),
                       (True, True, True, True), "extended"),
    "vertical_periodic_x": (dict(base_mesh=(16, 8), block_size=(4, 4), extent=(3.0, 1.0), periodic=(True, False),
                                 max_level=1, refine=lambda k, o, e: k[0] == 0 and k[1] == 1 and k[2] == 0),
                            (False, False, True, True), "material"),
}
