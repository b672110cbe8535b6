"""Many structures (and load cases) evaluated together on one GPU.

:class:`StructureBatch` packs independent designs of uneven size into one
plan: node, element and DOF numbers are offset per structure, so K is block
diagonal over structures and one launch of every kernel (element stiffness,
assembly, each CG step, adjoint) advances all designs x load cases, each with
its own CG scalars and convergence flag.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .plan import DEFAULT_MAXIT, DEFAULT_TOL, FEPlan


class StructureBatch:
    """Pack ``structures`` (objects with ``crds``, ``cnct_beamcols``, ``prop_beamcols``,
    ``cnct_quads``, ``prop_quads``, ``known_id`` and ``load_cases()``)."""

    def __init__(self, structures, n_cases=None):
        self.structures = list(structures)
        n = [s.crds.shape[0] for s in self.structures]
        self.node_off = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
        cb, cq, kn, crds, pb, pq = [], [], [], [], [], []
        self.beam_off = [0]
        self.quad_off = [0]
        for s, o in zip(self.structures, self.node_off[:-1]):
            cb.append(np.asarray(s.cnct_beamcols, np.int64).reshape(-1, 2) + o)
            cq.append(np.asarray(s.cnct_quads, np.int64).reshape(-1, 4) + o)
            kn.append(np.asarray(s.known_id, np.int64) + 6 * o)
            crds.append(np.asarray(s.crds, float))
            pb.append(np.asarray(s.prop_beamcols, float).reshape(-1, 6))
            pq.append(np.asarray(s.prop_quads, float).reshape(-1, 5))
            self.beam_off.append(self.beam_off[-1] + len(cb[-1]))
            self.quad_off.append(self.quad_off[-1] + len(cq[-1]))
        if self.node_off[-1] >= 2**31 // 6:
            raise ValueError("batch too large for 32-bit DOF indices")
        self.plan = FEPlan(self.node_off.astype(np.int32), np.concatenate(cb), np.concatenate(cq), np.concatenate(kn))
        dev = nat.device()
        self.crds = torch.as_tensor(np.concatenate(crds), device=dev)
        self.prop_beam = torch.as_tensor(np.concatenate(pb), device=dev)
        self.prop_quad = torch.as_tensor(np.concatenate(pq), device=dev)
        cases = [s.load_cases() for s in self.structures]
        m = n_cases or max(c.shape[0] for c in cases)
        F = np.zeros((6 * int(self.node_off[-1]), m))
        for c, o in zip(cases, self.node_off[:-1]):
            k = min(c.shape[0], m)
            F[6 * o:6 * o + c.shape[1], :k] = c[:k].T
        self.n_cases = m
        self.f = torch.as_tensor(F, device=dev).contiguous()

    @property
    def n_structures(self):
        return len(self.structures)

    def assemble(self):
        self.plan.assemble(self.crds, self.prop_beam, self.prop_quad)

    def solve(self, precision="fp64", tol=DEFAULT_TOL, maxit=DEFAULT_MAXIT, assemble=True, preconditioner=None):
        if assemble:
            self.assemble()
        f = self.f[:, 0] if self.n_cases == 1 else self.f
        self.u = self.plan.solve(f, precision=precision, tol=tol, maxit=maxit, preconditioner=preconditioner)
        return self.u

    def compliance(self, u=None):
        """(n_structures, n_cases) compliance 1/2 f.u, reduced on the GPU in FP64
        (per-slice partial sums added in a fixed order)."""
        u = self.u if u is None else u
        return self.plan.compliance(self.f.reshape(u.shape), u, 0.5)

    def gradient(self, u=None):
        """Compliance gradients per structure and case: (dcrds, dprop_beam, dprop_quad),
        packed like the batch ([nodes|beams|quads] x [3|6|5] x cases)."""
        u = self.u if u is None else u
        if not self.plan.symmetric():
            # members parallel to global Y or quads with kx_mod != ky_mod: K is not symmetric, lam = K^-T f/2
            lam = self.plan.solve_transpose(0.5 * self.f.reshape(u.shape))
            return self.plan.adjoint(self.crds, self.prop_beam, self.prop_quad, u, lam=lam, lam_scale=0.0)
        return self.plan.adjoint(self.crds, self.prop_beam, self.prop_quad, u, lam=None, lam_scale=0.5)

    def compliance_and_gradient(self, precision="fp64", tol=DEFAULT_TOL, maxit=DEFAULT_MAXIT, preconditioner=None):
        u = self.solve(precision=precision, tol=tol, maxit=maxit, preconditioner=preconditioner)
        return self.compliance(u), self.gradient(u)

    def split_nodes(self, a):
        return [a[self.node_off[i]:self.node_off[i + 1]] for i in range(self.n_structures)]
