"""Maps from design variables to model parameters, used around the solver.

The reference builds these in its example notebooks rather than in the
library; they are provided here so that optimisation scripts port directly:

- ``linear_filter`` -- the hat filter of Examples/Shells_Mannheim_Multihalle_Shape.ipynb
  (:375-378): w_ij = max(0, (R - d_ij) / R) on the XY-projected node distance,
  rows normalised to sum 1.  Returned as a sparse CSR matrix (the notebook's
  dense n x n matrix does not scale past a few thousand nodes); the gradient
  maps back with the transpose, as the notebook's ``sens @ B_ij[design_nodes]``.
- ``DeviceLinearMap`` -- y = B x + c applied on the GPU (``jsso_csr_spmv``) as a
  differentiable torch function (the backward applies B^T, also on the GPU);
  ``SSO_model.set_node_filter`` puts the filter inside the parameter map.
- ``simp_modulus`` -- E = rho^p E0 (Examples/shells_topo_shape.ipynb), a torch
  expression, so dC/drho follows from dC/dE by autograd or by ``simp_chain``.
- ``load_quad_mesh_csv`` -- the CSV layout of Examples/Data/Mannheim_Quad
  (crd_x/y/z.csv, cnct.csv with 4 node ids per quad, bc_node.csv).
"""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp
import torch


def linear_filter(x, y, radius):
    """Row-normalised hat filter on XY distance, as a scipy CSR matrix (n, n)."""
    from scipy.spatial import cKDTree

    pts = np.stack([np.asarray(x, float), np.asarray(y, float)], 1)
    tree = cKDTree(pts)
    d = tree.sparse_distance_matrix(tree, radius, output_type="coo_matrix")
    w = np.maximum(0.0, (radius - d.data) / radius)
    W = sp.csr_matrix((w, (d.row, d.col)), shape=(len(pts), len(pts)))
    W.setdiag(1.0)   # d_ii = 0 (whether or not the distance matrix stored it)
    rs = np.asarray(W.sum(axis=1)).ravel()
    return sp.diags(1.0 / rs) @ W


class DeviceLinearMap:
    """y = B x + c on the GPU for a fixed sparse B (scipy sparse or dense, m x n)
    and an optional offset c (m,): the forward and its gradient map B^T run the
    hand-written CSR product ``jsso_csr_spmv``; ``__call__`` is differentiable
    with respect to x (torch autograd)."""

    def __init__(self, B, c=None):
        from . import _native as nat
        Bs = sp.csr_matrix(B, dtype=np.float64)
        Bs.sort_indices()
        Bt = Bs.T.tocsr()
        Bt.sort_indices()
        dev = nat.device()
        self.shape = Bs.shape

        def up(M):
            return (torch.as_tensor(M.indptr.astype(np.int32), device=dev),
                    torch.as_tensor(M.indices.astype(np.int32), device=dev),
                    torch.as_tensor(M.data, device=dev))

        self._B, self._Bt = up(Bs), up(Bt)
        self._c = None if c is None else torch.as_tensor(np.asarray(c, dtype=np.float64), device=dev).reshape(-1)
        if self._c is not None and self._c.numel() != self.shape[0]:
            raise ValueError(f"offset has {self._c.numel()} entries, B has {self.shape[0]} rows")

    @staticmethod
    def _spmv(M, x, c):
        from . import _native as nat
        rowptr, colind, vals = M
        y = torch.empty(rowptr.numel() - 1, dtype=torch.float64, device=x.device)
        nat.check(nat.lib().jsso_csr_spmv(y.numel(), nat.ptr(rowptr), nat.ptr(colind), nat.ptr(vals),
                                          nat.ptr(x), nat.ptr(c), nat.ptr(y), nat.stream_ptr()))
        return y

    def apply(self, x):
        """B x + c (no autograd)."""
        return self._spmv(self._B, x.contiguous(), self._c)

    def apply_transpose(self, g):
        """B^T g, the gradient map (no autograd)."""
        return self._spmv(self._Bt, g.contiguous(), None)

    def __call__(self, x):
        return _LinearMapFn.apply(x, self)


class _LinearMapFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, m):
        ctx.m = m
        return m.apply(x.to(torch.float64))

    @staticmethod
    def backward(ctx, g):
        return ctx.m.apply_transpose(g.to(torch.float64)), None


def simp_modulus(rho, E0, p=3.0):
    """E = rho^p * E0 (works on numpy arrays and on torch tensors)."""
    return rho ** p * E0


def simp_chain(dC_dE, rho, E0, p=3.0):
    """dC/drho from dC/dE for E = rho^p E0."""
    return dC_dE * p * rho ** (p - 1) * E0


def load_quad_mesh_csv(folder):
    """Read a quad mesh in the Examples/Data/Mannheim_Quad CSV layout.

    Returns (x, y, z, cnct (n_quad, 4) int, bc_nodes int): the coordinates as
    stored (the notebook then shifts x and y to start at 0), the connectivity
    read as a flat list of node ids, 4 per quad, and the supported nodes."""
    def col(name, dtype=float):
        return np.loadtxt(os.path.join(folder, name), delimiter=",", dtype=dtype, ndmin=1).ravel()

    x, y, z = col("crd_x.csv"), col("crd_y.csv"), col("crd_z.csv")
    flat = col("cnct.csv", dtype=float).astype(np.int64)
    n_quad = (flat.size + 1) // 4
    cnct = flat[:4 * n_quad].reshape(n_quad, 4)
    bc = col("bc_node.csv", dtype=float).astype(np.int64)
    return x, y, z, cnct, bc
