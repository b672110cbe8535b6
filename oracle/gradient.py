"""Compliance and its adjoint gradient (oracle; test infrastructure only).

The reference differentiates ``C = 1/2 f[:ndof].u`` (``JaxSSO/SSO_model.py:297-301``)
through ``sci_sparse_solve``'s custom VJP (``JaxSSO/solver.py:221-250``):
``lambda = solve(K_aug^T, dC/du)``, ``K_aug_bar = -lambda_row * u_col``, which the
transposes of ``bcoo_sum_duplicates`` and ``vmap(element_K_*)`` turn into

    dC/dp = - sum_e lambda_e^T (dK_e/dp) u_e .

Element derivatives here are complex-step derivatives of the restated element
routines (exact to rounding, no step-size error).  Members parallel to global Y
(the ``Cxz == 0`` branch, ``element.py:92-97``) get NaN coordinate derivatives,
as JAX's ``where`` + division by ``Cxz == 0`` produce.
"""
from __future__ import annotations

import numpy as np

from . import elements as el
from . import system
from . import solve as osolve

H = 1e-30


def _cs_contract(fn, crds_e, props_e, ue, le, which, chunk=1024, cols=None):
    """d/dp of -le^T K_e(p) ue for every input column of ``crds_e`` or ``props_e``
    (or only for the columns listed in ``cols``)."""
    n = crds_e.shape[0]
    base = crds_e if which == "crds" else props_e
    width = base.shape[1]
    cols = np.arange(width) if cols is None else np.asarray(cols)
    ndir = len(cols)
    out = np.zeros((n, ndir))
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        m = i1 - i0
        c = np.repeat(crds_e[i0:i1], ndir, axis=0).astype(complex)
        p = np.repeat(props_e[i0:i1], ndir, axis=0).astype(complex)
        pert = np.tile(np.eye(width)[cols], (m, 1)) * (1j * H)
        if which == "crds":
            c = c + pert
        else:
            p = p + pert
        dK = np.imag(fn(c, p)) / H                       # (m*ndir, d, d)
        u = np.repeat(ue[i0:i1], ndir, axis=0)
        lam = np.repeat(le[i0:i1], ndir, axis=0)
        out[i0:i1] = -np.einsum("ni,nij,nj->n", lam, dK, u).reshape(m, ndir)
    return out


def element_gradients(crds, cnct_b, prop_b, cnct_q, prop_q, u, lam):
    """Per-element derivatives of -lam^T K u.

    Returns dict with ``beam_crds`` (nb,6), ``beam_props`` (nb,6),
    ``quad_crds`` (nq,12), ``quad_props`` (nq,5).
    """
    res = {}
    crds = np.asarray(crds, dtype=float)
    if len(cnct_b):
        ce = el.gather_crds(crds, cnct_b)
        d = el.element_dofs(cnct_b)
        ue, le = u[d], lam[d]
        gc = _cs_contract(el.beam_Ke, ce, np.asarray(prop_b, float), ue, le, "crds")
        _, yflag = el.beam_dircos(ce)
        gc[yflag] = np.nan
        res["beam_crds"] = gc
        res["beam_props"] = _cs_contract(el.beam_Ke, ce, np.asarray(prop_b, float), ue, le, "props")
    if len(cnct_q):
        ce = el.gather_crds(crds, cnct_q)
        d = el.element_dofs(cnct_q)
        ue, le = u[d], lam[d]
        res["quad_crds"] = _cs_contract(el.quad_Ke, ce, np.asarray(prop_q, float), ue, le, "crds")
        res["quad_props"] = _cs_contract(el.quad_Ke, ce, np.asarray(prop_q, float), ue, le, "props")
    return res


def node_gradient(n_node, cnct_b, cnct_q, eg):
    """Scatter per-element coordinate derivatives to (n_node, 3)."""
    g = np.zeros((n_node, 3))
    for cnct, key in ((cnct_b, "beam_crds"), (cnct_q, "quad_crds")):
        if key in eg:
            c = np.asarray(cnct, dtype=np.int64)
            np.add.at(g, c.ravel(), eg[key].reshape(-1, 3))
    return g


def compliance_and_gradient(st, loads=None, faithful=False):
    """C = 1/2 f.u and dC/d(crds), dC/d(props) for a workload ``Structure``.

    faithful=False: accurate reduced solve, lambda = u/2 (K symmetric).
    faithful=True : the reference's sequence -- SuperLU on K_aug, then a second
    SuperLU on K_aug^T for lambda (solver.py:236).
    """
    f = st.nodal_loads if loads is None else loads
    K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
    if faithful:
        Ka, fa = system.augment(K, st.known_id, f)
        u = osolve.solve_augmented(Ka, fa, st.ndof)
        g = np.concatenate([0.5 * f, np.zeros(len(st.known_id))])
        lam = osolve.solve_augmented_transpose(Ka, g)[:st.ndof]
    else:
        u = osolve.solve_reduced(K, st.known_id, f)
        lam = 0.5 * u
    C = 0.5 * f @ u
    eg = element_gradients(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads,
                           st.prop_quads, u, lam)
    gn = node_gradient(st.n_node, st.cnct_beamcols, st.cnct_quads, eg)
    return dict(C=C, u=u, lam=lam, dcrds=gn, dprop_beam=eg.get("beam_props"),
                dprop_quad=eg.get("quad_props"))


def compliance(st, crds=None, prop_b=None, prop_q=None, loads=None):
    """C = 1/2 f.u for perturbed inputs (finite-difference checks)."""
    crds = st.crds if crds is None else crds
    prop_b = st.prop_beamcols if prop_b is None else prop_b
    prop_q = st.prop_quads if prop_q is None else prop_q
    f = st.nodal_loads if loads is None else loads
    K = system.assemble_K(crds, st.cnct_beamcols, prop_b, st.cnct_quads, prop_q)
    u = osolve.solve_reduced(K, st.known_id, f)
    return 0.5 * f @ u
