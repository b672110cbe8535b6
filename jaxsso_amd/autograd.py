"""Differentiable FEA primitives (torch.autograd.Function) on the HIP kernels.

``displacement(plan, crds, prop_beam, prop_quad, f)`` returns u = K^-1 f and
backpropagates like the reference's custom VJP (``JaxSSO/solver.py:221-250``):
lam = K^-T u_bar (a second solve), then for every input p
    p_bar = -lam^T (dK/dp) u     (hand-written forward-mode adjoint kernel)
and f_bar = lam.

``compliance(plan, ...)`` returns C = 1/2 f.u (``JaxSSO/SSO_model.py:297-301``);
its adjoint is lam = u/2, so the backward needs no second solve.
"""
from __future__ import annotations

import torch

from .plan import DEFAULT_MAXIT, DEFAULT_TOL, FEPlan


class SolveOptions:
    """How the state solves run: ``method`` "sparse" (the GPU CG solver, with
    ``precision`` / ``tol`` / ``maxit``) or "dense" (the dense FP64 Cholesky of
    ``jsso_dense_solve``)."""

    def __init__(self, precision="fp64", tol=DEFAULT_TOL, maxit=DEFAULT_MAXIT, method="sparse"):
        if method not in ("sparse", "dense"):
            raise ValueError(f"unknown solve method {method!r} (sparse or dense)")
        self.precision = precision
        self.tol = tol
        self.maxit = maxit
        self.method = method


def _solve(plan, f, opts, transpose=False):
    if opts.method == "dense":
        return plan.dense_solve(f, transpose=transpose)
    if transpose:
        return plan.solve_transpose(f, precision=opts.precision, tol=opts.tol, maxit=opts.maxit)
    return plan.solve(f, precision=opts.precision, tol=opts.tol, maxit=opts.maxit)


def _grads(plan, crds, pb, pq, u, lam, scale, need, weight=None):
    """Adjoint gradients; with several load cases (u of shape (ndof, m)) the
    per-case gradients are weighted by ``weight`` (m,) and summed."""
    dc, db, dq = plan.adjoint(crds, pb, pq, u, lam=lam, lam_scale=scale)

    def fold(g, want):
        if not want or g is None:
            return None
        if weight is not None:
            g = g * weight
        return g.sum(-1) if u.dim() == 2 else g

    return fold(dc, need[0]), fold(db, need[1] and pb is not None), fold(dq, need[2] and pq is not None)


class _Displacement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, crds, prop_beam, prop_quad, f, plan: FEPlan, opts: SolveOptions):
        plan.assemble(crds, prop_beam, prop_quad)
        u = _solve(plan, f, opts)
        ctx.plan, ctx.opts = plan, opts
        ctx.save_for_backward(crds, prop_beam, prop_quad, u)
        return u

    @staticmethod
    def backward(ctx, u_bar):
        crds, pb, pq, u = ctx.saved_tensors
        plan, opts = ctx.plan, ctx.opts
        plan.ensure_assembled(crds, pb, pq)  # the plan may have been re-assembled since forward
        lam = _solve(plan, u_bar.contiguous(), opts, transpose=True)
        need = ctx.needs_input_grad
        dc, db, dq = _grads(plan, crds, pb, pq, u, lam, 0.0, need)
        f_bar = lam if need[3] else None
        return dc, db, dq, f_bar, None, None


class _Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, crds, prop_beam, prop_quad, f, plan: FEPlan, opts: SolveOptions):
        plan.assemble(crds, prop_beam, prop_quad)
        u = _solve(plan, f, opts)
        ctx.plan, ctx.opts = plan, opts
        ctx.save_for_backward(crds, prop_beam, prop_quad, u, f)
        C = 0.5 * (f * u).sum(0) if f.dim() == 2 else 0.5 * torch.dot(f, u)
        ctx.mark_non_differentiable(u)
        return C, u

    @staticmethod
    def backward(ctx, C_bar, _u_bar):
        crds, pb, pq, u, f = ctx.saved_tensors
        plan = ctx.plan
        need = ctx.needs_input_grad
        plan.ensure_assembled(crds, pb, pq)  # the plan may have been re-assembled since forward
        if plan.symmetric():
            dc, db, dq = _grads(plan, crds, pb, pq, u, None, 0.5, need, weight=C_bar)
            f_bar = (u * C_bar) if need[3] else None   # dC/df = K^-1 f = u
        else:
            # K not symmetric (members parallel to global Y, orthotropic quads): lam = K^-T f/2, as the
            # reference's custom VJP solves with K_aug^T (solver.py:236)
            lam = _solve(plan, 0.5 * f, ctx.opts, transpose=True)
            dc, db, dq = _grads(plan, crds, pb, pq, u, lam, 0.0, need, weight=C_bar)
            f_bar = ((0.5 * u + lam) * C_bar) if need[3] else None   # dC/df = (K^-1 + K^-T) f / 2
        return dc, db, dq, f_bar, None, None


def displacement(plan, crds, prop_beam, prop_quad, f, opts: SolveOptions | None = None):
    return _Displacement.apply(crds, prop_beam, prop_quad, f, plan, opts or SolveOptions())


def compliance(plan, crds, prop_beam, prop_quad, f, opts: SolveOptions | None = None):
    """Returns (C, u); only C is differentiable (through the lam = u/2 adjoint)."""
    return _Compliance.apply(crds, prop_beam, prop_quad, f, plan, opts or SolveOptions())
