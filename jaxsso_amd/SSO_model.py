"""Design parameters, objective and sensitivities -- the API of ``JaxSSO/SSO_model.py``.

``SSO_model`` maps a flat parameter vector (node-coordinate parameters first,
then element-property parameters, ``SSO_model.py:185-222``) onto the model's
coordinates and properties, solves on the GPU and differentiates the objective
through the adjoint kernels.

Differences a user of the reference sees:
* the objective and ``params_u`` / ``params_c`` / ``helper_params_to_objective``
  are torch computations (differentiable, composable with outer torch
  autograd as the reference composes with outer JAX AD,
  ``Examples/Neural_Network_Topo_Shape.ipynb``); a user objective is written
  with torch ops on ``u``;
* ``value_grad_params`` / ``grad_params`` / ``params_to_objective`` return host
  values (a Python float and numpy arrays), like the concrete arrays the
  reference returns;
* the 'strain energy' objective uses the compliance adjoint lam = u/2 (no
  second solve); a user objective costs one extra solve, as in the reference.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import autograd as ag


class NodeParameter:
    """A node-coordinate design parameter (``SSO_model.py:33-58``); X_Y_Z: 0, 1, 2."""

    def __init__(self, nodetag, X_Y_Z=2, upper_bound=None, lower_bound=None):
        self.tag = nodetag
        self.XYZ = X_Y_Z


class ElementParameter:
    """An element-property design parameter (``SSO_model.py:60-90``).

    ele_type 0: BeamCol, prop_type 0..5 = E, G, Iy, Iz, J, A;
    ele_type 1: Quad, prop_type 0..4 = t, E, nu, kx_mod, ky_mod.
    """

    def __init__(self, eletag, ele_type=0, prop_type=0, upper_bound=None, lower_bound=None):
        self.tag = eletag
        self.type = ele_type
        self.prop = prop_type


class SSO_model:
    def __init__(self, model):
        self.model = model
        self.nodeparameters_tags = []
        self.nodeparameters_xyzs = []
        self.eleparameters_tags = []
        self.ele_types = []
        self.eleparameters_props = []
        self.objective = None
        self.objective_args = None
        self._strain_energy = False
        self.last_u = None
        self._node_filter = None

    # -- parameters (SSO_model.py:125-204) --------------------------------
    def add_nodeparameter(self, nodeparameter):
        self.nodeparameters_tags.append(nodeparameter.tag)
        self.nodeparameters_xyzs.append(nodeparameter.XYZ)

    def add_eleparameter(self, eleparameter):
        self.eleparameters_tags.append(eleparameter.tag)
        self.ele_types.append(eleparameter.type)
        self.eleparameters_props.append(eleparameter.prop)

    def initialize_nodeparameters_values(self):
        self.model.model_ready()
        crds = self.model.crds
        self.nodeparameters_values = np.asarray(crds[np.asarray(self.nodeparameters_tags, dtype=np.int64),
                                                     np.asarray(self.nodeparameters_xyzs, dtype=np.int64)],
                                                dtype=float)

    def initialize_eleparameters_values(self):
        self.model.model_ready()
        props = (self.model.prop_beamcols, self.model.prop_quads)
        self.eleparameters_values = np.asarray(
            [props[i][j, k] for (i, j, k) in zip(self.ele_types, self.eleparameters_tags, self.eleparameters_props)],
            dtype=float)
        types = np.asarray(self.ele_types, dtype=np.int64)
        self.parameters_bc = np.nonzero(types == 0)[0]
        self.parameters_quads = np.nonzero(types == 1)[0]

    def initialize_parameters_values(self):
        self.initialize_nodeparameters_values()
        self.initialize_eleparameters_values()
        self.n_node_params = len(self.nodeparameters_tags)
        self.n_ele_params = len(self.eleparameters_tags)
        self.n_bc_params = len(self.parameters_bc)
        self.n_quad_params = len(self.parameters_quads)
        self.parameter_values = np.concatenate([self.nodeparameters_values, self.eleparameters_values])
        self._index_cache = None

    def update_nodeparameter(self, params_values):
        v = np.asarray(_to_host(params_values), dtype=float)
        self.nodeparameters_values = v
        self.parameter_values = self.parameter_values.copy()
        self.parameter_values[:self.n_node_params] = v

    def update_eleparameter(self, params_values):
        v = np.asarray(_to_host(params_values), dtype=float)
        self.eleparameters_values = v
        self.parameter_values = self.parameter_values.copy()
        self.parameter_values[self.n_node_params:] = v

    def update_parameter(self, params_values):
        self.parameter_values = np.asarray(_to_host(params_values), dtype=float)

    # -- parameters -> model inputs (SSO_model.py:207-222) -----------------
    def _indices(self, dev):
        if self._index_cache is None or self._index_cache[0] != dev:
            tags = torch.as_tensor(np.asarray(self.nodeparameters_tags, dtype=np.int64), device=dev)
            xyz = torch.as_tensor(np.asarray(self.nodeparameters_xyzs, dtype=np.int64), device=dev)
            et = np.asarray(self.eleparameters_tags, dtype=np.int64)
            ep = np.asarray(self.eleparameters_props, dtype=np.int64)
            bc = torch.as_tensor(self.parameters_bc, device=dev)
            qd = torch.as_tensor(self.parameters_quads, device=dev)
            self._index_cache = (dev, tags, xyz,
                                 torch.as_tensor(et[self.parameters_bc], device=dev),
                                 torch.as_tensor(ep[self.parameters_bc], device=dev), bc,
                                 torch.as_tensor(et[self.parameters_quads], device=dev),
                                 torch.as_tensor(ep[self.parameters_quads], device=dev), qd)
        return self._index_cache

    def node_params_crds(self, nodeparameter_values):
        v = _to_dev(nodeparameter_values)
        ix = self._indices(v.device)
        crds = torch.as_tensor(self.model.crds, dtype=torch.float64, device=v.device)
        return crds.index_put((ix[1], ix[2]), v)

    def ele_params_props(self, eleparameter_values):
        v = _to_dev(eleparameter_values)
        ix = self._indices(v.device)
        pb = torch.as_tensor(self.model.prop_beamcols, dtype=torch.float64, device=v.device)
        pq = torch.as_tensor(self.model.prop_quads, dtype=torch.float64, device=v.device)
        if len(self.parameters_bc):
            pb = pb.index_put((ix[3], ix[4]), v[ix[5]])
        if len(self.parameters_quads):
            pq = pq.index_put((ix[6], ix[7]), v[ix[8]])
        return pb, pq

    def set_node_filter(self, B):
        """Filter the node parameters on the GPU before they reach the model, as
        the reference's shape-optimisation example does outside the library
        (``z = (B_ij @ p)[design_nodes]``, gradient ``(sens @ B_ij[design_nodes])[design_nodes]``,
        Examples/Shells_Mannheim_Multihalle_Shape.ipynb:454-461).

        ``B`` (n_node x n_node, scipy sparse or dense) acts on the full vector of
        node coordinates along the parameters' axis, with the parameters'
        current values at their nodes; every node parameter must use the same
        axis.  Parameter values then mean the unfiltered design variables x, and
        ``value_grad_params`` returns dObjective/dx.  ``None`` removes the filter."""
        if B is None:
            self._node_filter = None
            return
        from .parameterization import DeviceLinearMap
        import scipy.sparse as sp
        axes = set(int(a) for a in self.nodeparameters_xyzs)
        if len(axes) != 1:
            raise ValueError("a node filter needs all node parameters on one axis")
        axis = axes.pop()
        tags = np.asarray(self.nodeparameters_tags, dtype=np.int64)
        Bs = sp.csr_matrix(B, dtype=np.float64)
        n = Bs.shape[0]
        others = np.setdiff1d(np.arange(n), tags)
        base = np.asarray(self.model.crds, dtype=float)[:, axis]
        Bt = Bs[tags]
        self._node_filter = DeviceLinearMap(Bt[:, tags], Bt[:, others] @ base[others])

    def _inputs(self, parameter_values):
        p = _to_dev(parameter_values)
        if self.n_node_params > 0:
            v = p[:self.n_node_params]
            if self._node_filter is not None:
                v = self._node_filter(v)
            crds = self.node_params_crds(v)
        else:
            crds = torch.as_tensor(self.model.crds, dtype=torch.float64, device=p.device)
        if self.n_ele_params > 0:
            pb, pq = self.ele_params_props(p[self.n_node_params:])
        else:
            pb = torch.as_tensor(self.model.prop_beamcols, dtype=torch.float64, device=p.device)
            pq = torch.as_tensor(self.model.prop_quads, dtype=torch.float64, device=p.device)
        f = torch.as_tensor(self.model.nodal_loads, dtype=torch.float64, device=p.device)
        return crds, pb, pq, f

    def _opts(self, which_solver, enforce_scipy_sparse):
        """Solver choice as ``Model.select_solver`` makes it (model.py:340-356):
        'dense' -> the dense direct solve; 'sparse' -> the GPU CG solver whatever
        ``enforce_scipy_sparse`` says (both of the reference's sparse backends map
        to it)."""
        m = self.model
        if which_solver not in ("dense", "sparse"):
            raise ValueError("Please select the right solver: dense or sparse")
        return ag.SolveOptions(precision=getattr(m, "precision", "fp64"), tol=getattr(m, "tol", 1e-10),
                               maxit=getattr(m, "maxit", 200000), method=which_solver)

    # -- parameters -> results (SSO_model.py:224-270) ----------------------
    def params_u(self, parameter_values, which_solver="sparse", enforce_scipy_sparse=True):
        """Displacements u(params), differentiable (adjoint: one extra solve)."""
        crds, pb, pq, f = self._inputs(parameter_values)
        return ag.displacement(self.model.plan(), crds, pb, pq, f, self._opts(which_solver, enforce_scipy_sparse))

    def params_c(self, parameter_values, which_solver="sparse", enforce_scipy_sparse=True):
        """Compliance 1/2 f.u(params), differentiable (adjoint lam = u/2, no extra solve)."""
        crds, pb, pq, f = self._inputs(parameter_values)
        C, u = ag.compliance(self.model.plan(), crds, pb, pq, f, self._opts(which_solver, enforce_scipy_sparse))
        self.last_u = u
        return C

    def node_params_c(self, nodeparameter_values, which_solver="sparse", enforce_scipy_sparse=True):
        """Compliance as a function of the node parameters only."""
        p = torch.cat([_to_dev(nodeparameter_values),
                       _to_dev(self.parameter_values[self.n_node_params:])])
        return self.params_c(p, which_solver, enforce_scipy_sparse)

    # -- objective (SSO_model.py:275-339) -----------------------------------
    def set_objective(self, objective="strain energy", func=None, func_args=None):
        if objective == "strain energy":
            def SE(sso_model, u):
                f = torch.as_tensor(sso_model.model.nodal_loads, dtype=torch.float64, device=u.device)
                return 0.5 * torch.dot(f, u)
            self.objective = SE
            self._strain_energy = True
        if objective == "user":
            self.objective = func
            self._strain_energy = False
            if func_args is not None:
                self.objective_args = func_args

    def helper_params_to_objective(self, parameter_values, which_solver="sparse", enforce_scipy_sparse=True):
        """Objective(params) as a differentiable torch scalar."""
        if self._strain_energy:
            return self.params_c(parameter_values, which_solver, enforce_scipy_sparse)
        u = self.params_u(parameter_values, which_solver, enforce_scipy_sparse)
        self.last_u = u.detach()
        if self.objective_args is None:
            return self.objective(self, u)
        return self.objective(self, u, *self.objective_args)

    def params_to_objective(self, which_solver="sparse", enforce_scipy_sparse=True):
        with torch.no_grad():
            return float(self.helper_params_to_objective(self.parameter_values, which_solver, enforce_scipy_sparse))

    def grad_params(self, which_solver="sparse", enforce_scipy_sparse=True):
        return self.value_grad_params(which_solver, enforce_scipy_sparse)[1]

    def value_grad_params(self, which_solver="sparse", enforce_scipy_sparse=True):
        """(objective, d objective / d parameters) in parameter order (host values)."""
        p = _to_dev(self.parameter_values).clone().requires_grad_(True)
        val = self.helper_params_to_objective(p, which_solver, enforce_scipy_sparse)
        (g,) = torch.autograd.grad(val, p)
        return float(val.detach()), g.detach().cpu().numpy()


def _to_host(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _to_dev(x):
    if isinstance(x, torch.Tensor):
        return x.to(device=nat.device(), dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=nat.device())
