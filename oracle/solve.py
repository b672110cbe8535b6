"""Linear solves (oracle; test infrastructure only).

* :func:`solve_augmented` -- the reference's default path: SuperLU on the
  Lagrange-augmented matrix (``JaxSSO/solver.py:176-218``, ``model.py:358-376``).
* :func:`solve_reduced` -- the accurate oracle: the equivalent SPD reduced
  system ``K_ff u_f = f_f`` (all prescribed displacements are zero), symmetric
  diagonal equilibration, SuperLU, then iterative refinement with residuals
  accumulated in x86 80-bit ``np.longdouble``.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import system


def solve_augmented(Ka, fa, ndof):
    """u = spsolve(K_aug, f_aug)[:ndof] -- SuperLU with COLAMD, as solver.py:195-209."""
    return spla.spsolve(Ka.tocsc(), fa)[:ndof]


def solve_augmented_transpose(Ka, g):
    """The reference adjoint's second factorisation, ``solve(K_aug^T, g)`` (solver.py:236)."""
    return spla.spsolve(Ka.T.tocsc(), g)


def solve_reduced(K, known_id, rhs, refine=4, return_info=False):
    """Solve K u = f with u[known] = 0 accurately.

    ``rhs`` may be (ndof,) or (ndof, m).  Returns u with zeros at ``known_id``.
    """
    ndof = K.shape[0]
    free = system.free_dofs(ndof, known_id)
    Kff = K[free][:, free].tocsc()
    d = np.abs(Kff.diagonal())
    d[d == 0] = 1.0
    s = 1.0 / np.sqrt(d)
    S = sp.diags(s)
    Ks = (S @ Kff @ S).tocsc()
    lu = spla.splu(Ks)
    b = np.asarray(rhs, dtype=float)
    vec = b.ndim == 1
    if vec:
        b = b[:, None]
    bf = b[free]
    Kff_ld = Kff.astype(np.longdouble)
    bf_ld = bf.astype(np.longdouble)
    y = lu.solve(bf * s[:, None])
    x = (s[:, None] * y).astype(np.longdouble)
    hist = []
    for _ in range(refine):
        r = bf_ld - Kff_ld @ x
        hist.append(float(np.max(np.linalg.norm(r.astype(float), axis=0) /
                                  np.maximum(np.linalg.norm(bf, axis=0), 1e-300))))
        dx = s[:, None] * lu.solve(s[:, None] * r.astype(float))
        x = x + dx.astype(np.longdouble)
    u = np.zeros((ndof, b.shape[1]))
    u[free] = x.astype(float)
    if vec:
        u = u[:, 0]
    if return_info:
        return u, hist
    return u
