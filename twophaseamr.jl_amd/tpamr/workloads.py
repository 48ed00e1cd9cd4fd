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


# interfaces y = eta(x): cell fractions, phase face apertures, centroids -------

def _clamped_integral(x0, x1, l0, l1, hi):
    """Exact integral over [x0, x1] of clamp(l, 0, hi) for l linear from l0 at
    x0 to l1 at x1 (the clamp makes it piecewise linear: integrate between
    the points where l crosses 0 and hi)."""
    dl = l1 - l0
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = np.where(dl != 0, -l0 / dl, 2.0)
        t1 = np.where(dl != 0, (hi - l0) / dl, 2.0)
    ta = np.clip(np.minimum(t0, t1), 0.0, 1.0)
    tb = np.clip(np.maximum(t0, t1), 0.0, 1.0)
    total = 0.0
    for a, b in ((0.0, ta), (ta, tb), (tb, 1.0)):
        fa = np.clip(l0 + a * dl, 0.0, hi)
        fb = np.clip(l0 + b * dl, 0.0, hi)
        total = total + (b - a) * (fa + fb) / 2
    return total * (x1 - x0)


class InterfaceGeometry:
    """Interface data of the phase below (or above) a curve y = eta(x) on one
    mesh, with the curve taken as its tangent line at each cell centre for
    the cell fraction and the interface centroid, at the face position for
    vertical faces and as its chord over horizontal faces (all exact for a
    straight interface). Only cells that the tangent line crosses, and their
    faces, are computed in full; the mesh arrays are kept between calls."""

    def __init__(self, mesh: BlockTreeMesh):
        self.mesh = mesh
        self.x, self.y = mesh.cell_centers()
        self.hx, self.hy = mesh.cell_spacing()

    def __call__(self, eta, deta, liquid_below=True):
        """Returns (fraction per cell, liquid aperture per interior face, gas
        aperture, centroid offsets (ncells, 2) in cell spacings, zero in pure
        cells)."""
        mesh = self.mesh
        x, y, hx, hy = self.x, self.y, self.hx, self.hy
        ec_all = eta(x)
        s_all = deta(x)
        cut = np.abs(ec_all - y) <= hy / 2 + np.abs(s_all) * hx / 2
        below = (ec_all > y).astype(np.float64)
        ci = np.nonzero(cut)[0]
        xc, yc, ec, s, hxc, hyc = x[ci], y[ci], ec_all[ci], s_all[ci], hx[ci], hy[ci]
        x0, x1, y0 = xc - hxc / 2, xc + hxc / 2, yc - hyc / 2
        fb = _clamped_integral(x0, x1, ec - s * hxc / 2 - y0, ec + s * hxc / 2 - y0, hyc) / (hxc * hyc)
        below[ci] = np.clip(fb, 0.0, 1.0)
        frac = below if liquid_below else 1.0 - below
        # interface chord of the tangent line inside the cell: the x range
        # where the line lies between the bottom and the top of the cell
        with np.errstate(divide="ignore", invalid="ignore"):
            xa = np.where(s != 0, xc + (y0 - ec) / s, -np.inf)
            xb = np.where(s != 0, xc + (y0 + hyc - ec) / s, np.inf)
        lo = np.maximum(np.minimum(xa, xb), x0)
        hi = np.minimum(np.maximum(xa, xb), x1)
        mid = (lo + hi) / 2
        fc = frac[ci]
        mixed = (fc > 0) & (fc < 1) & (hi >= lo)
        off = np.zeros((mesh.ncells, 2))
        off[ci[mixed], 0] = (mid[mixed] - xc[mixed]) / hxc[mixed]
        off[ci[mixed], 1] = (ec[mixed] + s[mixed] * (mid[mixed] - xc[mixed]) - yc[mixed]) / hyc[mixed]
        np.clip(off, -0.5, 0.5, out=off)
        # faces: pure unless next to a crossed cell
        m, p = mesh.face_minus, mesh.face_plus
        lb = below[m].copy()
        fi = np.nonzero(cut[m] | cut[p])[0]
        fm, fp, ax = m[fi], p[fi], mesh.face_axis[fi]
        vert = ax == 0
        fx = np.where(vert, x[fm] + hx[fm] / 2,
                      0.5 * (np.maximum(x[fm] - hx[fm] / 2, x[fp] - hx[fp] / 2)
                             + np.minimum(x[fm] + hx[fm] / 2, x[fp] + hx[fp] / 2)))
        fy = np.where(~vert, y[fm] + hy[fm] / 2,
                      0.5 * (np.maximum(y[fm] - hy[fm] / 2, y[fp] - hy[fp] / 2)
                             + np.minimum(y[fm] + hy[fm] / 2, y[fp] + hy[fp] / 2)))
        ext = np.where(vert, np.minimum(hy[fm], hy[fp]), np.minimum(hx[fm], hx[fp]))
        a = np.empty(len(fi))
        a[vert] = np.clip((eta(fx[vert]) - (fy[vert] - ext[vert] / 2)) / ext[vert], 0.0, 1.0)
        h = ~vert
        e0 = eta(fx[h] - ext[h] / 2) - fy[h]
        e1 = eta(fx[h] + ext[h] / 2) - fy[h]
        # part of the face under the chord (phase below)
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.where(e1 != e0, -e0 / (e1 - e0), 0.0)
        t = np.clip(t, 0.0, 1.0)
        a[h] = np.where((e0 > 0) & (e1 > 0), 1.0,
                        np.where((e0 <= 0) & (e1 <= 0), 0.0, np.where(e1 > e0, 1.0 - t, t)))
        lb[fi] = a
        liquid = lb if liquid_below else 1.0 - lb
        return frac, liquid, 1.0 - liquid, off


def interface_geometry(mesh: BlockTreeMesh, eta, deta, liquid_below=True):
    """One-shot InterfaceGeometry(mesh)(eta, deta, liquid_below)."""
    return InterfaceGeometry(mesh)(eta, deta, liquid_below)


# Stokes-wave-like two-phase case ---------------------------------------------

WAVE_DENSITY = (1000.0, 1.2)
WAVE_GRAVITY = (0.0, -9.81)


def wave_surface(seed=SEED + 2):
    """eta(x, t) = 0.5 + cos(w t) (0.08 cos(pi x) + 0.01 (a2 cos(2 pi x) + a3
    cos(3 pi x))), a2, a3 ~ U(-1, 1) (seeded), w = sqrt(g pi tanh(pi / 2)) (the
    linear standing-wave frequency of the first mode at depth 1/2). Returns
    (eta, deta) as functions of (x, t)."""
    rng = SplitMix64(seed)
    a2 = 2 * rng.uniform() - 1
    a3 = 2 * rng.uniform() - 1
    w = math.sqrt(9.81 * math.pi * math.tanh(math.pi / 2))

    def eta(x, t=0.0):
        return 0.5 + math.cos(w * t) * (0.08 * np.cos(np.pi * x)
                                        + 0.01 * (a2 * np.cos(2 * np.pi * x) + a3 * np.cos(3 * np.pi * x)))

    def deta(x, t=0.0):
        return -math.cos(w * t) * np.pi * (0.08 * np.sin(np.pi * x)
                                           + 0.01 * (2 * a2 * np.sin(2 * np.pi * x) + 3 * a3 * np.sin(3 * np.pi * x)))

    return eta, deta


def wave_mesh(levels=(5, 11), block=32, band_cells=4, drift=0.0):
    """[0, 2] x [0, 1] with walls, block x block cells per block (two roots),
    uniform to levels[0] and refined to levels[1] within `band_cells` finest
    cells (plus `drift`, the largest interface displacement of the run) of
    the wave surface at t = 0."""
    eta, _ = wave_surface()
    lo, hi = levels
    band = band_cells / (block * (1 << hi)) + drift
    lip = 0.08 * math.pi + 0.01 * (2 * math.pi + 3 * math.pi)  # bound of |eta'|

    def refine_many(level, i, j, ox, oy, ex, ey):
        out = level < lo
        rest = ~out
        if rest.any():
            xs = ox[rest][:, None] + ex[rest][:, None] * np.linspace(0.0, 1.0, 17)[None, :]
            e = eta(xs)
            slack = lip * ex[rest] / 32 + band
            emin = e.min(axis=1) - slack
            emax = e.max(axis=1) + slack
            out[rest] = (emax >= oy[rest]) & (emin <= oy[rest] + ey[rest])
        return out

    return BlockTreeMesh(base_mesh=(2 * block, block), block_size=(block, block), extent=(2.0, 1.0),
                         periodic=(False, False), max_level=hi, refine_many=refine_many)


def wave_interfaces(mesh: BlockTreeMesh, steps, dt):
    """Yields, for n = 0 .. steps (time n dt), the interface data of the wave
    surface at that time: (cell fraction, liquid face aperture, gas face
    aperture, centroid offsets), one step at a time."""
    eta, deta = wave_surface()
    geom = InterfaceGeometry(mesh)
    for n in range(steps + 1):
        t = n * dt
        yield geom(lambda x: eta(x, t), lambda x: deta(x, t))


def wave_steps(solver, steps, dt, density=WAVE_DENSITY, gravity=WAVE_GRAVITY):
    """Per step n (time n dt): the mixture face coefficients of the surface
    at that time and the gravity! jumps, as ProjectionRun takes them, plus
    the final fraction (for the kinetic energy)."""
    from .physics import gravity_jumps, mixture_face_coefficients
    out = []
    frac = None
    for n, (frac, la, ga, off) in enumerate(wave_interfaces(solver.operator.mesh, steps, dt)):
        if n == steps:
            break
        kappa = mixture_face_coefficients(solver, frac, density)
        jumps = gravity_jumps(solver, la, ga, density=density, gravity=gravity, centroid_offsets=off)
        out.append((kappa, jumps))
    return out, frac


def cmdgfm_block_workload(n=1024, block=16, dt=0.02, rho_liquid=1000.0, rho_gas=1.0, centre=(0.5, 0.5),
                          radius=0.27, seed=SEED + 3):
    """A CMDGFM-shaped reduced-pressure/jump system on the unit square (walls,
    n^2 cells in block^2 blocks) with a liquid disk of `radius` at `centre`:
    pressure operator 1/rho with the harmonic face mean (rho_liquid inside,
    rho_gas outside), one anchor (cell 0), every face whose cells lie on
    opposite sides of the circle a jump unknown. Blocks: the pressure
    operator's stiffness restricted to the retained cells; B: -g / +g on the
    face's minus / plus cell (g = 0.3 u |S_mp|); C: the pressure difference
    across the face times u / (rho_liquid L); J: diagonal U(2, 4) plus
    U(-0.4, 0.4) couplings of faces sharing a cell (u ~ U(0.5, 1.5)); rhs of
    order-one volume and jump defects. The anchor's omitted row makes the
    full system consistent. Returns (mesh, face coefficients, dict of the
    CMDGFMBlockSystem arguments and rhs, dt, density)."""
    import scipy.sparse as sp

    from .hierarchy import HostHierarchy

    drawn = [0]

    def uniform(lo, hi, k):
        # the next k numbers of the splitmix64 stream of `seed`, vectorised
        step = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            z = np.uint64(seed & MASK64) + (np.arange(drawn[0] + 1, drawn[0] + k + 1, dtype=np.uint64) * step)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        drawn[0] += k
        return lo + (hi - lo) * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53)))

    mesh = BlockTreeMesh(base_mesh=(n, n), block_size=(block, block), periodic=(False, False))
    x, y = mesh.cell_centers()
    rho = np.where(np.hypot(x - centre[0], y - centre[1]) < radius, rho_liquid, rho_gas)
    m, p = mesh.face_minus, mesh.face_plus
    kappa = 2.0 / (rho[m] + rho[p])
    s = HostHierarchy(mesh, kappa).level_matrix(0)
    ncells = mesh.ncells
    idx = np.arange(1, ncells)
    anchor = np.zeros(ncells - 1, np.int64)
    phi = np.hypot(x - centre[0], y - centre[1]) - radius
    mixed = np.flatnonzero(np.sign(phi[m]) != np.sign(phi[p]))
    nx = len(mixed)
    fm, fp = m[mixed], p[mixed]
    smp = np.asarray(s[fm, fp]).ravel() if nx else np.zeros(0)
    g = 0.3 * np.abs(smp) * uniform(0.5, 1.5, nx)
    b_full = sp.csr_matrix((np.concatenate([-g, g]), (np.concatenate([fm, fp]), np.tile(np.arange(nx), 2))),
                           shape=(ncells, nx))
    cw = uniform(0.5, 1.5, nx) / rho_liquid
    local = np.concatenate([[-1], np.arange(ncells - 1)])
    rows = np.tile(np.arange(nx), 2)
    cols = np.concatenate([local[fm], local[fp]])
    vals = np.concatenate([-cw, cw])
    keep = cols >= 0
    c_blk = sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(nx, ncells - 1))
    # faces sharing a cell
    cell_face = np.concatenate([np.stack([fm, np.arange(nx)], 1), np.stack([fp, np.arange(nx)], 1)])
    cell_face = cell_face[np.lexsort((cell_face[:, 1], cell_face[:, 0]))]
    jr, jc = [np.arange(nx)], [np.arange(nx)]
    start = 0
    while start < len(cell_face):
        end = start
        while end < len(cell_face) and cell_face[end, 0] == cell_face[start, 0]:
            end += 1
        ks = cell_face[start:end, 1]
        if len(ks) > 1:
            a, b = np.meshgrid(ks, ks, indexing="ij")
            off = a != b
            jr.append(a[off])
            jc.append(b[off])
        start = end
    jr, jc = np.concatenate(jr), np.concatenate(jc)
    jv = np.concatenate([uniform(2.0, 4.0, nx), uniform(-0.4, 0.4, len(jr) - nx)])
    j_blk = sp.csr_matrix((jv, (jr, jc)), shape=(nx, nx))
    matrix = sp.bmat([[s[idx][:, idx], b_full[idx]], [c_blk, j_blk]], format="csr")
    matrix.sum_duplicates()
    matrix.sort_indices()
    vol = mesh.volumes
    rp = vol[idx] / (dt * dt) * (0.5 * np.sin(3.0 * x[idx] + 1.0) * np.cos(2.0 * y[idx]) +
                                 0.1 * uniform(-1.0, 1.0, ncells - 1)) / rho_liquid
    h = np.sqrt(vol)
    lengths = np.minimum(h[fm], h[fp])
    rj = lengths / (dt * dt) * uniform(-1.0, 1.0, nx)
    omitted = sp.hstack([s[[0]][:, idx], b_full[[0]]], format="csr")
    system = dict(matrix=matrix, pressure_indices=idx, pressure_anchors=anchor, pressure_volumes=vol[idx],
                  jump_lengths=lengths, omitted=omitted, omitted_rhs=np.array([-rp.sum()]),
                  omitted_volumes=vol[[0]])
    return mesh, kappa, system, np.concatenate([rp, rj]), dt, max(rho_liquid, rho_gas)
