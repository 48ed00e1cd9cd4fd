"""Assembled CMDGFM-shaped systems for the block-solve tests.

The reference assembles its reduced-pressure/jump system from the interface
geometry (``_assemble_cmdgfm_system!``, src/two_phase/projection/
cmdgfm_projection.jl:1484-1523). These fixtures build systems of the same
structure from a pressure operator and a circle of mixed faces:

- pressure rows: the level-0 stiffness S of the pressure operator restricted
  to the retained cells (every cell but one anchor per component, the
  reduced unknown being the cell's pressure minus its anchor's);
- jump columns of the pressure rows (B): each mixed face adds -g and +g to
  its minus and plus cells (the oriented face contributions of
  ``_cmdgfm_anchor_residual_defect``, :1848-1893);
- pressure columns of the jump rows (C): the pressure difference across the
  face, and the jump block J: a positive diagonal plus nonsymmetric
  couplings between mixed faces that share a cell;
- the omitted anchor rows over the same unknowns with the rhs that makes the
  full system consistent (S and B have zero column sums).

Magnitudes follow the reference's Krylov scaling (pressure unknowns
rho (L/dt)^2, jumps L/dt^2), so the scaled blocks are of order one.
"""
import numpy as np
import scipy.sparse as sp


def component_labels(s):
    """Connected components of the stiffness graph (zero couplings split)."""
    g = sp.csr_matrix(s)
    g = g - sp.diags(g.diagonal())
    g.eliminate_zeros()
    ncomp, labels = sp.csgraph.connected_components(g, directed=False)
    return ncomp, labels


def entries(s, rows, cols):
    """s[rows[k], cols[k]] as a flat array."""
    if len(rows) == 0:
        return np.zeros(0)
    return np.asarray(s[rows, cols]).ravel()


def build_system(om, osolver, *, dt=0.02, density=1000.0, centre=(0.5, 0.5), radius=0.27, seed=0, jumps=True):
    """Returns a dict with the system matrix (CSR), rhs and the arrays of
    CMDGFMBlockSystem / oracle_cmdgfm.BlockSystem."""
    import scipy.sparse.csgraph  # noqa: F401

    rng = np.random.default_rng(seed)
    s = sp.csr_matrix(osolver.level_matrix(0))
    ncells = om.ncells
    vol = om.volumes.copy()
    ncomp, labels = component_labels(s)
    anchors_of = np.array([int(np.flatnonzero(labels == c).min()) for c in range(ncomp)])
    is_anchor = np.zeros(ncells, bool)
    is_anchor[anchors_of] = True
    idx = np.flatnonzero(~is_anchor)
    anchor = anchors_of[labels[idx]]
    np_ = len(idx)
    h = 1.0
    cx, cy = om.cell_centers()
    if jumps:
        phi = np.hypot(cx - centre[0], cy - centre[1]) - radius
        m, p = om.face_minus, om.face_plus
        mixed = np.flatnonzero((np.sign(phi[m]) != np.sign(phi[p])) & (entries(s, m, p) != 0))
    else:
        mixed = np.zeros(0, np.int64)
    nx = len(mixed)
    fm, fp = om.face_minus[mixed], om.face_plus[mixed]
    # B: -g at the minus cell, +g at the plus cell (full cells x jumps)
    kf = -entries(s, fm, fp)
    g = 0.3 * kf * rng.uniform(0.5, 1.5, nx)
    rows_b = np.concatenate([fm, fp])
    cols_b = np.concatenate([np.arange(nx), np.arange(nx)])
    b_full = sp.csr_matrix((np.concatenate([-g, g]), (rows_b, cols_b)), shape=(ncells, nx))
    # C: pressure difference across the face (scaled to order one)
    local = -np.ones(ncells, np.int64)
    local[idx] = np.arange(np_)
    cw = rng.uniform(0.5, 1.5, nx) / (density * h)
    crow, ccol, cval = [], [], []
    for k in range(nx):
        for cell, sign in ((fm[k], -1.0), (fp[k], 1.0)):
            if local[cell] >= 0:
                crow.append(k)
                ccol.append(local[cell])
                cval.append(sign * cw[k])
    c_blk = sp.csr_matrix((cval, (crow, ccol)), shape=(nx, np_))
    # J: diagonal 3 +- 1, couplings of faces sharing a cell (nonsymmetric)
    jr, jc, jv = list(range(nx)), list(range(nx)), list(rng.uniform(2.0, 4.0, nx))
    by_cell = {}
    for k in range(nx):
        for cell in (fm[k], fp[k]):
            by_cell.setdefault(int(cell), []).append(k)
    for ks in by_cell.values():
        for a in ks:
            for b in ks:
                if a != b:
                    jr.append(a)
                    jc.append(b)
                    jv.append(rng.uniform(-0.4, 0.4))
    j_blk = sp.csr_matrix((jv, (jr, jc)), shape=(nx, nx))
    a_pp = s[idx][:, idx]
    b_blk = b_full[idx]
    matrix = sp.bmat([[a_pp, b_blk], [c_blk, j_blk]], format="csr")
    matrix.sum_duplicates()
    matrix.sort_indices()
    # rhs: pressure rows of volume-defect order one, jump rows of order one
    rp = vol[idx] / dt ** 2 * (0.5 * np.sin(3.0 * cx[idx] + 1.0) * np.cos(2.0 * cy[idx]) +
                               0.1 * rng.standard_normal(np_)) / density
    spacing = np.sqrt(vol)
    lengths = np.minimum(spacing[fm], spacing[fp])
    rj = lengths / dt ** 2 * rng.uniform(-1.0, 1.0, nx)
    rhs = np.concatenate([rp, rj])
    # omitted anchor rows: S and B columns over the retained cells and jumps
    omitted = sp.hstack([s[anchors_of][:, idx], b_full[anchors_of]], format="csr")
    omitted.sort_indices()
    omit_rhs = np.array([-rp[labels[idx] == c].sum() for c in range(ncomp)])
    return dict(matrix=matrix, rhs=rhs, np=np_, nx=nx, pressure_indices=idx, pressure_anchors=anchor,
                pressure_volumes=vol[idx], jump_lengths=lengths, omitted=omitted, omitted_rhs=omit_rhs,
                omitted_volumes=vol[anchors_of], ncells=ncells, mixed_faces=mixed, dt=dt, density=density,
                reference_length=h, components=ncomp)
