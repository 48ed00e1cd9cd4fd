"""GPU: tpamr_solve_device_noalloc, the allocation-free solve, against
tpamr_solve_device through the C ABI: bit for bit, in every precision mode,
on two calls in a row whose second mesh is smaller (the static device scratch
keeps the first call's contents), and after the context's own solves."""
import ctypes as C

import numpy as np
import pytest

from cases import MESHES, DROP_CENTRE, DROP_RADIUS, disk_fraction, techref_noise

pytestmark = pytest.mark.gpu

import tpamr  # noqa: E402
from tpamr import _lib as L  # noqa: E402


def _print(*a):
    print(*a, flush=True)


def _solver(name, precision):
    pm = tpamr.BlockTreeMesh(**MESHES[name])
    x, y = pm.cell_centers()
    dx, dy = pm.cell_spacing()
    c = disk_fraction(x, y, dx, dy, DROP_CENTRE, DROP_RADIUS)
    s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0), precision=precision)
    tpamr.refresh_two_phase(s, c, (1000.0, 1.0))
    s._sync()
    return pm, s


def _call(fn, s, db, dx):
    import torch
    L.order_after_torch(s.ctx.h, db.device)
    st = L.Stats()
    L.check(fn(s.ctx.h, C.c_void_p(db.data_ptr()), C.c_void_p(dx.data_ptr()), 1e-10, 0.0, 200, 0, C.byref(st)),
            s.ctx.h)
    torch.cuda.synchronize()
    return st, dx.cpu().numpy()


FIELDS = ("cycles", "converged", "initial_l2", "initial_linf", "final_l2", "final_linf", "target_linf",
          "removed_rhs_mean", "krylov_runs", "replacements")


@pytest.mark.parametrize("precision", ["fp64", "fp32_vcycle", "fp32_refine"])
def test_noalloc_solve_equals_solve_device(precision):
    import torch
    lib = L.lib()
    for name in ("amr64", "amr32"):  # the second call's mesh is smaller
        _print("case noalloc", precision, name)
        pm, s = _solver(name, precision)
        b = techref_noise(pm.ncells)
        x0 = 0.05 * np.sin(np.arange(pm.ncells))
        db = torch.tensor(b, dtype=torch.float64, device="cuda")
        ref_st, ref_x = _call(lib.tpamr_solve_device, s, db, torch.tensor(x0, dtype=torch.float64, device="cuda"))
        for rep in range(2):
            st, x = _call(lib.tpamr_solve_device_noalloc, s, db,
                          torch.tensor(x0, dtype=torch.float64, device="cuda"))
            assert st.converged
            assert np.array_equal(x, ref_x), (name, rep)
            for f in FIELDS:
                assert getattr(st, f) == getattr(ref_st, f), (name, rep, f)
        # the context's own graph-replayed solve is unchanged afterwards
        st2, x2 = _call(lib.tpamr_solve_device, s, db, torch.tensor(x0, dtype=torch.float64, device="cuda"))
        assert np.array_equal(x2, ref_x) and st2.cycles == ref_st.cycles
        s.close()
