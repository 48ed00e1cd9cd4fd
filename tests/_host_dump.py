"""Writes the host builder's outputs for one mesh to an .npz (run as a
subprocess by test_host_builder.py with different TPAMR_HOST_THREADS)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "twophaseamr.jl_amd"))
import tpamr  # noqa: E402
from tpamr import workloads as W  # noqa: E402


def main(path):
    pm = W.circles_mesh(levels=(3, 7), n=16)[0]
    k = np.exp(0.4 * np.sin(np.arange(1, pm.nfaces + 1)))
    h = tpamr.HostHierarchy(pm, k)
    out = {"keys": pm.block_keys}
    for a in ("face_minus", "face_plus", "face_axis", "face_area", "bface_cell", "bface_side"):
        out[a] = np.asarray(getattr(pm, a))
    for l in range(h.nlevels):
        A = h.level_matrix(l)
        info = h.level_info(l)
        out[f"{l}/indptr"], out[f"{l}/indices"], out[f"{l}/data"] = A.indptr, A.indices, A.data
        out[f"{l}/omega"] = np.array([info["omega"], info["omega_p"]])
        out[f"{l}/kappa"] = h.level_kappa(l)
        if l + 1 < h.nlevels:
            out[f"{l}/xfer"] = h.level_transfer(l)
    np.savez(path, **out)
    print(pm.ncells, h.nlevels)


if __name__ == "__main__":
    main(sys.argv[1])
