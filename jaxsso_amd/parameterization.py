"""Maps from design variables to model parameters, used around the solver.

The reference builds these in its example notebooks rather than in the
library; they are provided here so that optimisation scripts port directly:

- ``linear_filter`` -- the hat filter of Examples/Shells_Mannheim_Multihalle_Shape.ipynb
  (cells 10-13): w_ij = max(0, (R - d_ij) / R) on the XY-projected node distance,
  rows normalised to sum 1.  Returned as a sparse CSR matrix (the notebook's
  dense n x n matrix does not scale past a few thousand nodes); the gradient
  maps back with the transpose, as the notebook's ``sens @ B_ij[design_nodes]``.
- ``simp_modulus`` -- E = rho^p E0 (Examples/shells_topo_shape.ipynb), a torch
  expression, so dC/drho follows from dC/dE by autograd or by ``simp_chain``.
- ``load_quad_mesh_csv`` -- the CSV layout of Examples/Data/Mannheim_Quad
  (crd_x/y/z.csv, cnct.csv with 4 node ids per quad, bc_node.csv).
"""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp


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
