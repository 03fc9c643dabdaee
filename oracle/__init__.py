"""CPU FP64 oracle for the 3D nodal DC-resistivity path.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``simpeg_amd`` product imports this
package.  It is imported by ``tests/`` (as the checker), by
``__graft_entry__.smoke()`` (as the checker) and by ``bench.py``'s
``cpu_baseline`` leg (as the CPU method being timed beside the GPU path).

It restates, in numpy/scipy, the arithmetic that
``simpeg.electromagnetics.static.resistivity.Simulation3DNodal`` performs when
neither Pardiso nor MUMPS is installed: the discretize tensor-mesh operators are
assembled as sparse matrices (kron products, the way discretize builds them),
A = G^T M_e(sigma) G + diag(AvgBC sigma) is factorised with
``scipy.sparse.linalg.splu`` (what ``pymatsolver.SolverLU`` wraps; reference
``simpeg/utils/solver_utils.py:39-44``), and fields / dpred / Jvec / Jtvec /
getJ / getJtJdiag follow ``resistivity/simulation.py:100-263, 523-644``.

Parity pinning: every function here is checked in ``tests/test_oracle.py``
against invariants and closed forms (see that file's docstring for the list).
The discretize routines themselves (``unpack_widths``, ``nodal_gradient``,
``get_edge_inner_product``, ``average_cell_to_face`` boundary extrapolation,
``get_interpolation_matrix``) are third-party code that is absent from the
reference tree; their restatements here are pinned by invariants, not by
bit-for-bit comparison ("parity unpinned" for exact summation order).
"""

from .dc_nodal import (  # noqa: F401
    OracleMesh,
    OracleExpMap,
    OracleIdentityMap,
    OracleDC,
    unpack_widths,
)
