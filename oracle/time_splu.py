"""Timing of the oracle's direct solver (scipy splu, the reference's solver without MKL/MUMPS)
as the mesh grows: assembly, factorisation, fill and per-right-hand-side solve time.

    python -m oracle.time_splu 24x24x12 48x48x32     # core cells; 8 padding cells x1.3 per side

This is test infrastructure, like the rest of oracle/: it times the checker, not the product.
"""

import resource
import sys
import time

import numpy as np
from scipy.sparse.linalg import splu

from oracle import OracleDC, OracleMesh


def time_core(core, npad=8, nrhs=8):
    from simpeg_amd import workloads as W

    mesh = W.padded_mesh(core, npad=npad)
    om = OracleMesh(mesh.h, origin=mesh.origin)
    orc = OracleDC(om, [], bc_type="Robin")
    orc.set_model(np.full(mesh.n_cells, np.log(0.01)))
    t0 = time.perf_counter()
    A = orc.getA().tocsc()
    t1 = time.perf_counter()
    lu = splu(A)
    t2 = time.perf_counter()
    b = np.random.default_rng(0).standard_normal((A.shape[0], nrhs))
    x = lu.solve(b)
    t3 = time.perf_counter()
    return {
        "core": core, "nodes": mesh.n_nodes, "assembly_s": t1 - t0, "factor_s": t2 - t1,
        "lu_nnz": lu.L.nnz + lu.U.nnz, "solve_ms_per_rhs": 1e3 * (t3 - t2) / nrhs,
        "rel_residual": float(np.linalg.norm(A @ x - b) / np.linalg.norm(b)),
        "max_rss_gb": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6,
    }


if __name__ == "__main__":
    for arg in sys.argv[1:] or ["24x24x12"]:
        print(time_core(tuple(int(v) for v in arg.split("x"))), flush=True)
