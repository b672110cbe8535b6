"""Global assembly and boundary conditions (oracle; test infrastructure only).

Restates ``JaxSSO/assemblemodel.py``: the element COO contributions
(``element.py:141-149,1097-1106``, ``meshgrid(dofs, dofs, indexing='ij')``),
the ``(0,0)`` zero seed (``assemblemodel.py:202-204``), duplicate summation
(``:160-162``), the augmented ``[[K, V^T], [V, 0]]`` (``:111-163``) and
``f_aug = [f; 0]`` (``:166-194``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import elements as el


def element_matrices(crds, cnct_b, prop_b, cnct_q, prop_q):
    """Per-element global stiffness (beams (nb,12,12), quads (nq,24,24))."""
    Kb = Kq = None
    if len(cnct_b):
        Kb = el.beam_Ke(el.gather_crds(crds, cnct_b), np.asarray(prop_b))
    if len(cnct_q):
        Kq = chunked(lambda c, p: el.quad_Ke(c, p), el.gather_crds(crds, cnct_q), np.asarray(prop_q))
    return Kb, Kq


def chunked(fn, crds_e, props, chunk=4096):
    out = [fn(crds_e[i:i + chunk], props[i:i + chunk]) for i in range(0, len(crds_e), chunk)]
    return np.concatenate(out) if out else None


def coo_entries(n_node, cnct_b, Kb, cnct_q, Kq):
    """COO (rows, cols, data) in the reference's order: seed, beams, quads."""
    rows = [np.zeros(1, np.int64)]
    cols = [np.zeros(1, np.int64)]
    data = [np.zeros(1)]
    for cnct, K in ((cnct_b, Kb), (cnct_q, Kq)):
        if K is None:
            continue
        d = el.element_dofs(cnct)
        m = d.shape[1]
        rows.append(np.repeat(d, m, axis=1).ravel())
        cols.append(np.tile(d, (1, m)).ravel())
        data.append(K.reshape(len(cnct), -1).ravel())
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(data)


def assemble_K(crds, cnct_b, prop_b, cnct_q, prop_q):
    """Global K (CSR, duplicates summed, explicit zeros kept), shape (6n, 6n)."""
    n = np.asarray(crds).shape[0]
    Kb, Kq = element_matrices(crds, cnct_b, prop_b, cnct_q, prop_q)
    r, c, d = coo_entries(n, cnct_b, Kb, cnct_q, Kq)
    K = sp.coo_matrix((d, (r, c)), shape=(6 * n, 6 * n)).tocsr()
    K.sort_indices()
    return K


def pattern_K(n_node, cnct_b, cnct_q):
    """Scalar sparsity pattern of K as a CSR with unit data (from connectivity alone)."""
    r, c, _ = coo_entries(n_node, cnct_b, None if not len(cnct_b) else np.zeros((len(cnct_b), 12, 12)),
                          cnct_q, None if not len(cnct_q) else np.zeros((len(cnct_q), 24, 24)))
    P = sp.coo_matrix((np.ones(len(r)), (r, c)), shape=(6 * n_node, 6 * n_node)).tocsr()
    P.sort_indices()
    P.data[:] = 1.0
    return P


def augment(K, known_id, loads):
    """``K_aug = [[K, V^T], [V, Z]]`` with Z holding an explicit (0,0) zero; ``f_aug = [f; 0]``."""
    known = np.asarray(known_id, dtype=np.int64)
    nc = len(known)
    ndof = K.shape[0]
    V = sp.csr_matrix((np.ones(nc), (np.arange(nc), known)), shape=(nc, ndof))
    Z = sp.csr_matrix((np.zeros(1), (np.zeros(1, np.int64), np.zeros(1, np.int64))), shape=(nc, nc))
    Ka = sp.bmat([[K, V.T], [V, Z]], format="csr")
    Ka.sort_indices()
    fa = np.concatenate([np.asarray(loads, dtype=float), np.zeros(nc)])
    return Ka, fa


def free_dofs(ndof, known_id):
    mask = np.ones(ndof, bool)
    mask[np.asarray(known_id, dtype=np.int64)] = False
    return np.nonzero(mask)[0]
