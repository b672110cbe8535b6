"""Assembly of the linear system -- the API of ``JaxSSO/assemblemodel.py``.

The reference returns ``jax.experimental.sparse.BCOO`` matrices.  Here the
stiffness lives on the GPU in 6x6 block-CSR form inside a plan; the functions
return :class:`BlockK` / :class:`AugmentedK` objects that the solvers of
:mod:`jaxsso_amd.solver` accept directly and that expand on request to the
reference's exact scalar pattern (``to_scipy``), explicit zeros included.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _native as nat
from .plan import FEPlan, blocks_to_csr

_PLANS: dict = {}


def _plan_for(n_node, cnct_b, cnct_q, known):
    cb = np.ascontiguousarray(np.asarray(cnct_b, dtype=np.int32).reshape(-1, 2))
    cq = np.ascontiguousarray(np.asarray(cnct_q, dtype=np.int32).reshape(-1, 4))
    kn = np.ascontiguousarray(np.asarray(known, dtype=np.int32).ravel())
    key = (int(n_node), cb.tobytes(), cq.tobytes(), kn.tobytes())
    plan = _PLANS.get(key)
    if plan is None:
        if len(_PLANS) > 16:
            _PLANS.clear()
        plan = FEPlan([0, int(n_node)], cb, cq, kn)
        _PLANS[key] = plan
    return plan


class BlockK:
    """Global stiffness K (``K_func``, assemblemodel.py:196-213) on the GPU."""

    def __init__(self, plan: FEPlan, values: torch.Tensor):
        self.plan = plan
        self.values = values            # (nnzb, 6, 6) FP64 device tensor
        self.shape = (plan.ndof, plan.ndof)

    def to_scipy(self) -> sp.csr_matrix:
        rowptr, colind = self.plan.pattern()
        return blocks_to_csr(rowptr, colind, self.values.detach().cpu().numpy())

    def todense(self):
        return self.to_scipy().toarray()

    def __matmul__(self, x):
        return self.plan.spmv(x)


class AugmentedK(BlockK):
    """``K_aug = [[K, V^T], [V, 0]]`` of ``assemblemodel.K_aug`` (assemblemodel.py:111-163).

    Held as K plus the constrained DOFs; the GPU solvers solve the equivalent
    constrained system (every prescribed displacement is zero).
    """

    def __init__(self, plan: FEPlan, values: torch.Tensor, known_id, crds, prop_beam, prop_quad):
        super().__init__(plan, values)
        self.known_id = np.asarray(known_id, dtype=np.int64)
        self.ncons = len(self.known_id)
        self.shape = (plan.ndof + self.ncons, plan.ndof + self.ncons)
        self._inputs = (crds, prop_beam, prop_quad)

    def to_scipy(self) -> sp.csr_matrix:
        K = super().to_scipy()
        nc, ndof = self.ncons, self.plan.ndof
        V = sp.csr_matrix((np.ones(nc), (np.arange(nc), self.known_id)), shape=(nc, ndof))
        Z = sp.csr_matrix((np.zeros(1), (np.zeros(1, np.int64), np.zeros(1, np.int64))), shape=(nc, nc))
        A = sp.bmat([[K, V.T], [V, Z]], format="csr")
        A.sort_indices()
        return A

    def __matmul__(self, x):
        x = torch.as_tensor(x, dtype=torch.float64, device=self.values.device)
        ndof = self.plan.ndof
        u, mu = x[:ndof], x[ndof:]
        known = torch.as_tensor(self.known_id, device=x.device)
        top = self.plan.spmv(u.contiguous())
        top = top.index_add(0, known, mu)
        return torch.cat([top, u[known]])


def _dev(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=nat.device()) if not isinstance(a, torch.Tensor) \
        else a.to(device=nat.device(), dtype=torch.float64)


def K_func(node_crds, ndof, n_beamcol, cnct_beamcols, prop_beamcols, n_quad, cnct_quads, prop_quads, known=()):
    """Global K (not augmented) for the given coordinates and properties."""
    n_node = int(ndof) // 6
    plan = _plan_for(n_node, cnct_beamcols if n_beamcol else np.zeros((0, 2)),
                     cnct_quads if n_quad else np.zeros((0, 4)), known)
    crds, pb, pq = _dev(node_crds), _dev(prop_beamcols), _dev(prop_quads)
    vals = plan.assemble(crds, pb, pq, export_K=True)
    return BlockK(plan, vals)


def K_aug_func(node_crds, ndof, known_dofs, n_beamcol, cnct_beamcols, prop_beamcols, n_quad, cnct_quads, prop_quads):
    """Augmented stiffness (``assemblemodel.py:215-229``)."""
    n_node = int(ndof) // 6
    plan = _plan_for(n_node, cnct_beamcols if n_beamcol else np.zeros((0, 2)),
                     cnct_quads if n_quad else np.zeros((0, 4)), known_dofs)
    crds, pb, pq = _dev(node_crds), _dev(prop_beamcols), _dev(prop_quads)
    vals = plan.assemble(crds, pb, pq, export_K=True)
    return AugmentedK(plan, vals, known_dofs, crds, pb, pq)


def f_aug(loads, ncons):
    """``[f; 0]`` (``assemblemodel.py:166-194``)."""
    f = _dev(loads)
    return torch.cat([f, torch.zeros(int(ncons), dtype=torch.float64, device=f.device)])


def f_aug_func(loads, known_dofs):
    return f_aug(loads, len(np.asarray(known_dofs).ravel()))


def model_K(model):
    return K_func(model.crds, model.ndof, model.n_beamcol, model.cnct_beamcols, model.prop_beamcols, model.n_quad,
                  model.cnct_quads, model.prop_quads, model.known_id)


def model_K_aug(model):
    return K_aug_func(model.crds, model.ndof, model.known_id, model.n_beamcol, model.cnct_beamcols,
                      model.prop_beamcols, model.n_quad, model.cnct_quads, model.prop_quads)


def model_f_aug(model):
    return f_aug_func(model.nodal_loads, model.known_id)
