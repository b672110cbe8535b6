"""Linear solvers -- the API of ``JaxSSO/solver.py``.

The reference offers a dense LU (``jax_dense_solve``, solver.py:75-93), a GPU
sparse direct solve (``jax_sparse_solve``, :102-168) and a host SuperLU through
``pure_callback`` (``sci_sparse_solve``, :176-250), each on the Lagrange-
augmented system.  Here every sparse solve is the GPU conjugate-gradient solver
of ``libjsso_hip.so`` (block-Jacobi scaled; FP64 by default, or FP32
recurrences with FP64 reliable updates when ``precision="mixed"`` -- faster on
shells, but slender beam chains need a tighter tolerance in that mode), and the
dense solve is an FP64 blocked Cholesky of the constrained system on the GPU.  All return ``u_aug = [u; mu]`` with the
Lagrange multipliers ``mu = f_c - (K u)_c`` (support reactions), as the
reference's augmented solve does.
"""
from __future__ import annotations

import torch

DEFAULT_PRECISION = "fp64"
DEFAULT_TOL = 1e-10
DEFAULT_MAXIT = 200000


def _finish(K_aug, f_aug, u):
    """u_aug = [u; mu] with the multipliers mu = f_c - (K u)_c of K_aug's own K."""
    ndof = K_aug.plan.ndof
    f_aug = torch.as_tensor(f_aug, dtype=torch.float64, device=u.device)
    known = torch.as_tensor(K_aug.known_id, device=u.device)
    Ku = K_aug.plan.spmv(u, K=K_aug.values)
    mu = f_aug[:ndof][known] - Ku[known]
    return torch.cat([u, mu])


def _cg(K_aug, f_aug, precision, tol, maxit):
    ndof = K_aug.plan.ndof
    f = torch.as_tensor(f_aug, dtype=torch.float64, device=K_aug.values.device)[:ndof].contiguous()
    crds, pb, pq = K_aug._inputs
    if crds is None:
        raise ValueError("this K_aug carries no assembly inputs (build it with K_func / K_aug_func)")
    K_aug.plan.assemble(crds, pb, pq)   # the plan may be shared with another evaluation
    u = K_aug.plan.solve(f, precision=precision, tol=tol, maxit=maxit)
    return _finish(K_aug, torch.as_tensor(f_aug, dtype=torch.float64, device=u.device), u)


def sci_sparse_solve(K_aug, f_aug, precision=None, tol=None, maxit=None):
    """GPU CG solve of the augmented system (replaces host SuperLU, solver.py:176-218)."""
    return _cg(K_aug, f_aug, precision or DEFAULT_PRECISION, tol or DEFAULT_TOL, maxit or DEFAULT_MAXIT)


def jax_sparse_solve(K_aug, f_aug, precision=None, tol=None, maxit=None):
    """GPU CG solve (replaces cuSolverSp spsolve, solver.py:102-125)."""
    return _cg(K_aug, f_aug, precision or DEFAULT_PRECISION, tol or DEFAULT_TOL, maxit or DEFAULT_MAXIT)


def jax_dense_solve(K_aug, f_aug, precision=None, tol=None, maxit=None):
    """Dense FP64 direct solve on the GPU (replaces the dense LU of K_aug, solver.py:75-93).

    ``jsso_dense_solve`` forms the constrained K of K_aug's own values as a dense
    matrix on the device and factorises it by a blocked Cholesky (hand-written
    HIP); memory grows as 8 ndof^2 bytes, so like the reference's dense path it
    is meant for models up to some tens of thousands of DOF.  ``precision``,
    ``tol`` and ``maxit`` are accepted for a uniform solver signature.
    """
    ndof = K_aug.plan.ndof
    f = torch.as_tensor(f_aug, dtype=torch.float64, device=K_aug.values.device)[:ndof].contiguous()
    u = K_aug.plan.dense_solve(f, K=K_aug.values)
    return _finish(K_aug, f_aug, u)
