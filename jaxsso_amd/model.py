"""Finite-element model builder -- the ``Model`` API of ``JaxSSO/model.py``.

Same method names, argument order and conventions as the reference:
DOFs are ``6*tag + [ux, uy, uz, rx, ry, rz]`` (``model.py:183-214``); node
coordinates are taken in insertion order (``model.py:248-256``), so tags must
be ``0..n-1`` in insertion order (checked here; the reference silently mixes
orderings otherwise).  Arrays frozen by :meth:`model_ready` are host numpy
arrays; :meth:`solve` runs on the GPU (``libjsso_hip.so``) and stores ``u`` as
a device tensor.

Additions: vectorised builders (``add_nodes``, ``add_beamcols``, ``add_quads``,
``add_supports``, ``add_nodal_loads``) for large meshes, and
``Model.precision`` / ``tol`` / ``maxit`` solver settings.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from . import solver
from .element import BeamCol, Quad, Truss  # noqa: F401  (reference module namespace)


class Node:
    """A node (``model.py:15-30``)."""

    def __init__(self, nodeTag, X, Y, Z):
        self.nodeTag = nodeTag
        self.X, self.Y, self.Z = X, Y, Z


class Model:
    """The FE model yet to be analysed (``model.py:32-383``)."""

    def __init__(self):
        self.nodes = {}
        self.beamcols = {}
        self.quads = {}
        self.known_indices = []
        self.f = {}
        self.u = None
        self.precision = solver.DEFAULT_PRECISION
        self.tol = solver.DEFAULT_TOL
        self.maxit = solver.DEFAULT_MAXIT
        self._plan_key = None
        self._plan = None

    # -- builders (model.py:56-214) ---------------------------------------
    def add_node(self, nodeTag: int, X: float, Y: float, Z: float):
        self.nodes[nodeTag] = [X, Y, Z]

    def update_node(self, nodeTag: int, XYZ: int, value: float):
        try:
            self.nodes[nodeTag][XYZ] = value
        except (KeyError, IndexError):
            print("Node {} does not exist in the model".format(nodeTag))

    def add_beamcol(self, eleTag: int, i_nodeTag: int, j_nodeTag: int, E: float, G: float, Iy: float, Iz: float,
                    J: float, A: float):
        self.beamcols[eleTag] = BeamCol(eleTag, i_nodeTag, j_nodeTag, E, G, Iy, Iz, J, A)

    def add_truss(self, eleTag: int, i_nodeTag: int, j_nodeTag: int, E: float, A: float):
        """Accepted and ignored, as in the reference (``model.py:130-151``)."""
        pass

    def add_quad(self, eleTag: int, i_nodeTag: int, j_nodeTag: int, m_nodeTag: int, n_nodeTag: int, t: float,
                 E: float, nu: float, kx_mod=1.0, ky_mod=1.0):
        self.quads[eleTag] = Quad(eleTag, i_nodeTag, j_nodeTag, m_nodeTag, n_nodeTag, t, E, nu, kx_mod, ky_mod)

    def add_support(self, nodeTag, active_supports=(1, 1, 1, 1, 1, 1)):
        dofs = np.arange(nodeTag * 6, nodeTag * 6 + 6, dtype=np.int32)
        active = np.asarray(active_supports, dtype=np.int32)
        self.known_indices.extend(dofs[np.nonzero(active == 1)[0]].tolist())

    def add_nodal_load(self, nodeTag, nodal_load=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
        self.f[nodeTag] = list(nodal_load)

    # -- vectorised builders ------------------------------------------------
    def add_nodes(self, tags, X, Y, Z):
        for t, x, y, z in zip(np.asarray(tags).tolist(), np.asarray(X, float).tolist(),
                              np.asarray(Y, float).tolist(), np.asarray(Z, float).tolist()):
            self.nodes[t] = [x, y, z]

    def add_beamcols(self, tags, cnct, props):
        """cnct (n,2) node tags, props (n,6) [E,G,Iy,Iz,J,A]."""
        for t, c, p in zip(np.asarray(tags).tolist(), np.asarray(cnct).tolist(), np.asarray(props, float).tolist()):
            self.beamcols[t] = BeamCol(t, c[0], c[1], *p)

    def add_quads(self, tags, cnct, props):
        """cnct (n,4) node tags, props (n,3..5) [t,E,nu(,kx_mod,ky_mod)]."""
        props = np.asarray(props, float)
        if props.shape[1] == 3:
            props = np.concatenate([props, np.ones((len(props), 2))], 1)
        for t, c, p in zip(np.asarray(tags).tolist(), np.asarray(cnct).tolist(), props.tolist()):
            self.quads[t] = Quad(t, c[0], c[1], c[2], c[3], *p)

    def add_supports(self, node_tags, active_supports=(1, 1, 1, 1, 1, 1)):
        for t in np.asarray(node_tags).tolist():
            self.add_support(t, active_supports)

    def add_nodal_loads(self, node_tags, loads):
        for t, f in zip(np.asarray(node_tags).tolist(), np.asarray(loads, float).tolist()):
            self.f[t] = f

    # -- freeze (model.py:221-338) -----------------------------------------
    def model_ready(self):
        tags = list(self.nodes.keys())
        if tags != list(range(len(tags))):
            raise ValueError("node tags must be 0..n-1 in insertion order: DOFs come from the tag "
                             "(model.py:192) and coordinates from insertion order (model.py:248-256)")
        self.crds = self.get_node_crds()
        self.nodal_loads = self.get_loads()
        self.known_id, self.unknown_id = self.get_boundary_ids()
        self.ndof = self.get_dofs()
        self.n_beamcol = len(self.beamcols)
        self.cnct_beamcols = self.get_cnct_beamcols()
        self.prop_beamcols = self.get_beamcols_cross_prop()
        self.n_quad = len(self.quads)
        self.cnct_quads = self.get_cnct_quads()
        self.prop_quads = self.get_quads_cross_prop()

    def get_node_crds(self):
        if not self.nodes:
            return np.zeros((0, 3))
        return np.asarray(list(self.nodes.values()), dtype=float).reshape(-1, 3)

    def get_loads(self):
        load = np.zeros(self.get_dofs())
        for node, f in self.f.items():
            load[node * 6:node * 6 + 6] = f
        return load

    def get_boundary_ids(self):
        known_id = np.asarray(self.known_indices, dtype=np.int32).ravel()
        all_idx = np.arange(6 * len(self.nodes), dtype=np.int32)
        unknown_id = all_idx[~np.isin(all_idx, known_id)]
        return known_id, unknown_id

    def get_dofs(self):
        return len(self.nodes) * 6

    def get_cnct_beamcols(self):
        if not self.beamcols:
            return np.zeros((0, 2), np.int32)
        return np.asarray([[b.i_nodeTag, b.j_nodeTag] for b in self.beamcols.values()], dtype=np.int32)

    def get_cnct_quads(self):
        if not self.quads:
            return np.zeros((0, 4), np.int32)
        return np.asarray([[q.i_nodeTag, q.j_nodeTag, q.m_nodeTag, q.n_nodeTag] for q in self.quads.values()],
                          dtype=np.int32)

    def get_beamcols_cross_prop(self):
        if not self.beamcols:
            return np.zeros((0, 6))
        return np.asarray([[b.E, b.G, b.Iy, b.Iz, b.J, b.A] for b in self.beamcols.values()], dtype=float)

    def get_quads_cross_prop(self):
        if not self.quads:
            return np.zeros((0, 5))
        return np.asarray([[q.t, q.E, q.nu, q.kx_mod, q.ky_mod] for q in self.quads.values()], dtype=float)

    # -- analysis (model.py:340-383) ---------------------------------------
    def plan(self):
        """The device plan for this model's topology (built once, reused)."""
        from .plan import FEPlan
        key = (self.crds.shape[0], self.cnct_beamcols.tobytes(), self.cnct_quads.tobytes(), self.known_id.tobytes())
        if self._plan is None or self._plan_key != key:
            self._plan = FEPlan([0, self.crds.shape[0]], self.cnct_beamcols, self.cnct_quads, self.known_id)
            self._plan_key = key
        return self._plan

    def select_solver(self, which_solver, enforce_scipy_sparse):
        """'dense' -> dense direct solve; 'sparse' -> the GPU CG solver (the
        ``enforce_scipy_sparse`` flag is accepted; every sparse solve runs on the GPU)."""
        if which_solver == "dense":
            return solver.jax_dense_solve
        elif which_solver == "sparse":
            return solver.sci_sparse_solve if enforce_scipy_sparse else solver.jax_sparse_solve
        print("Please select the right solver: dense or sparse")
        return None

    def solve(self, which_solver="sparse", enforce_scipy_sparse=True):
        """u = K^-1 f with the supports applied; stored in ``self.u`` (device tensor)."""
        from . import assemblemodel
        K_aug = assemblemodel.model_K_aug(self)
        f_aug = assemblemodel.model_f_aug(self)
        fn = self.select_solver(which_solver, enforce_scipy_sparse)
        if fn is None:
            return
        self.u = fn(K_aug, f_aug, precision=self.precision, tol=self.tol, maxit=self.maxit)[:self.get_dofs()]

    def strain_energy(self):
        if self.u is not None:
            f = torch.as_tensor(self.nodal_loads, dtype=torch.float64, device=self.u.device)
            return 0.5 * torch.dot(f, self.u)
        print("Model has not been analyzed yet.")
        return None
