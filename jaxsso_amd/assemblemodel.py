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
    """Global stiffness K (``K_func``, assemblemodel.py:196-213) on the GPU.

    ``nse`` is the entry count the reference's BCOO carries (one per element
    entry plus the (0,0) seed, before duplicates are summed); ``to_scipy()``
    gives the summed matrix with the reference's exact pattern."""

    def __init__(self, plan: FEPlan, values: torch.Tensor, inputs=None, seed=True):
        self.plan = plan
        self.values = values            # (nnzb, 6, 6) FP64 device tensor
        self.shape = (plan.ndof, plan.ndof)
        self.seed = seed                # K_func's explicit zero at (0, 0) (assemblemodel.py:202-204)
        self.nse = 144 * plan.n_beam + 576 * plan.n_quad + (1 if seed else 0)
        self._inputs = inputs           # (crds, prop_beam, prop_quad) it was assembled from

    def to_scipy(self) -> sp.csr_matrix:
        """Scalar CSR with the reference's pattern: every entry some element
        touches (explicit zeros kept), plus the (0,0) seed of K_func."""
        rowptr, colind = self.plan.pattern()
        vals = self.values.detach().cpu().numpy()
        used = np.zeros(self.plan.n_nodes, bool)
        used[self.plan.cnct_beam.ravel()] = True
        used[self.plan.cnct_quad.ravel()] = True
        if not used.all():
            # the plan keeps a diagonal block for every node; the reference has
            # no entries at nodes without elements
            br = np.repeat(np.arange(self.plan.n_nodes), np.diff(rowptr))
            keep = used[br] | (br != colind)
            counts = np.add.reduceat(keep.astype(np.int64), rowptr[:-1]) if len(keep) else np.zeros(0, np.int64)
            counts[np.diff(rowptr) == 0] = 0
            rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
            colind, vals = colind[keep], vals.reshape(-1, 6, 6)[keep]
        A = blocks_to_csr(rowptr, colind, vals)
        if self.seed and A.shape[0] and not used[0]:
            A = (A + sp.csr_matrix((np.zeros(1), ([0], [0])), shape=A.shape)).tocsr()
            A.sort_indices()
        return A

    def todense(self):
        return self.to_scipy().toarray()

    def __matmul__(self, x):
        """K x with this matrix's own values (the plan may since have assembled another K)."""
        return self.plan.spmv(x, K=self.values)


class AugmentedK(BlockK):
    """``K_aug = [[K, V^T], [V, 0]]`` of ``assemblemodel.K_aug`` (assemblemodel.py:111-163).

    Held as K plus the constrained DOFs; the GPU solvers solve the equivalent
    constrained system (every prescribed displacement is zero).
    """

    def __init__(self, plan: FEPlan, values: torch.Tensor, known_id, crds, prop_beam, prop_quad, k_nse=None):
        super().__init__(plan, values, (crds, prop_beam, prop_quad))
        self.known_id = np.asarray(known_id, dtype=np.int64)
        self.ncons = len(self.known_id)
        self.shape = (plan.ndof + self.ncons, plan.ndof + self.ncons)
        # bcoo_sum_duplicates(..., nse=k_nse + 2 ncons + 1) of assemblemodel.py:161-162
        self.nse = (self.nse if k_nse is None else int(k_nse)) + 2 * self.ncons + 1

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
        top = self.plan.spmv(u.contiguous(), K=self.values)
        top = top.index_add(0, known, mu)
        return torch.cat([top, u[known]])


def _dev(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=nat.device()) if not isinstance(a, torch.Tensor) \
        else a.to(device=nat.device(), dtype=torch.float64)


def K_func(node_crds, ndof, n_beamcol, cnct_beamcols, prop_beamcols, n_quad, cnct_quads, prop_quads, known=()):
    """Global K (not augmented) for the given coordinates and properties (assemblemodel.py:196-213)."""
    n_node = int(ndof) // 6
    plan = _plan_for(n_node, cnct_beamcols if n_beamcol else np.zeros((0, 2)),
                     cnct_quads if n_quad else np.zeros((0, 4)), known)
    crds, pb, pq = _dev(node_crds), _dev(prop_beamcols), _dev(prop_quads)
    vals = plan.assemble(crds, pb, pq, export_K=True)
    return BlockK(plan, vals, (crds, pb, pq))


def element_K(node_crds, ndof, cnct_beamcols, prop_beamcols, cnct_quads, prop_quads):
    """Global stiffness of the given elements without K_func's (0,0) seed: the
    per-type ``BeamCol.K_beamcol`` / ``Quad.K_quad`` (element.py:228-275,1190-1241)."""
    n_node = int(ndof) // 6
    plan = _plan_for(n_node, cnct_beamcols, cnct_quads, ())
    crds, pb, pq = _dev(node_crds), _dev(np.asarray(prop_beamcols, float).reshape(-1, 6)), \
        _dev(np.asarray(prop_quads, float).reshape(-1, 5))
    vals = plan.assemble(crds, pb, pq, export_K=True)
    return BlockK(plan, vals, (crds, pb, pq), seed=False)


def K_aug(known_dofs, K_BCOO, ndof, ncons, k_nse):
    """``[[K, V^T], [V, 0]]`` of a stiffness from :func:`K_func` (assemblemodel.py:111-163).

    ``K_BCOO`` is the :class:`BlockK` that ``K_func`` returns; ``k_nse`` is the
    reference's entry count of K (144 per beam + 576 per quad + 1), kept as
    ``nse = k_nse + 2 ncons + 1`` on the result.  The values are K_BCOO's own;
    the plan used for solves is the one of K's topology with these supports."""
    if not isinstance(K_BCOO, BlockK):
        raise TypeError("K_aug expects the BlockK returned by K_func")
    known = np.asarray(known_dofs, dtype=np.int64).ravel()
    if int(ncons) != len(known):
        raise ValueError(f"ncons={ncons} but {len(known)} known DOFs")
    if int(ndof) != K_BCOO.shape[0]:
        raise ValueError(f"ndof={ndof} but K is {K_BCOO.shape[0]} x {K_BCOO.shape[1]}")
    p = K_BCOO.plan
    plan = _plan_for(p.n_nodes, p.cnct_beam, p.cnct_quad, known)
    inputs = K_BCOO._inputs if K_BCOO._inputs is not None else (None, None, None)
    return AugmentedK(plan, K_BCOO.values, known, *inputs, k_nse=k_nse)


def K_aug_func(node_crds, ndof, known_dofs, n_beamcol, cnct_beamcols, prop_beamcols, n_quad, cnct_quads, prop_quads):
    """Augmented stiffness (``assemblemodel.py:215-229``)."""
    n_node = int(ndof) // 6
    plan = _plan_for(n_node, cnct_beamcols if n_beamcol else np.zeros((0, 2)),
                     cnct_quads if n_quad else np.zeros((0, 4)), known_dofs)
    crds, pb, pq = _dev(node_crds), _dev(prop_beamcols), _dev(prop_quads)
    vals = plan.assemble(crds, pb, pq, export_K=True)
    k_nse = 144 * int(n_beamcol) + 576 * int(n_quad) + 1    # assemblemodel.py:222-225
    return AugmentedK(plan, vals, known_dofs, crds, pb, pq, k_nse=k_nse)


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
