"""Element types of the model builder (``JaxSSO/element.py``).

``BeamCol``, ``Truss`` and ``Quad`` hold an element's attributes exactly as the
reference's classes do (``element.py:24-64, 361-382, 427-470``); their
stiffness routines run on the GPU through ``jsso_element_ke_beam`` /
``jsso_element_ke_quad`` (vectorised over a leading element axis).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat


def _as_batch(x, width):
    a = np.asarray(x, dtype=np.float64)
    single = a.ndim == 1
    return np.ascontiguousarray(a.reshape(-1, width)), single


class BeamCol:
    """12-DOF 3-D Euler-Bernoulli beam-column (``element.py:24-64``)."""

    def __init__(self, eleTag, i_nodeTag, j_nodeTag, E, G, Iy, Iz, J, A):
        self.dofs = 2 * 6
        self.eleTag = eleTag
        self.i_nodeTag = i_nodeTag
        self.j_nodeTag = j_nodeTag
        self.E, self.G, self.Iy, self.Iz, self.J, self.A = E, G, Iy, Iz, J, A

    @staticmethod
    def element_K_beamcol(crds, E, G, Iy, Iz, J, A):
        """Global stiffness (12,12) -- or (n,12,12) for batched inputs -- of
        ``BeamCol.element_K_beamcol`` (element.py:130-139), computed on the GPU."""
        c, single = _as_batch(crds, 6)
        props = np.stack(np.broadcast_arrays(*[np.asarray(v, np.float64).ravel() for v in (E, G, Iy, Iz, J, A)]), 1)
        props = np.ascontiguousarray(np.broadcast_to(props, (len(c), 6)))
        L = nat.lib()
        dev = nat.device()
        ct = torch.tensor(c, device=dev)
        pt = torch.tensor(props, device=dev)
        ke = torch.empty((len(c), 12, 12), dtype=torch.float64, device=dev)
        nat.check(L.jsso_element_ke_beam(len(c), nat.ptr(ct), nat.ptr(pt), nat.ptr(ke), nat.ptr(None), nat.stream_ptr()))
        return ke[0] if single else ke

    @staticmethod
    def element_K_beamcol_indices(i_nodeTag, j_nodeTag):
        """(144, 2) global (row, col) of the element's entries (element.py:141-149)."""
        dofs = np.concatenate([np.arange(6 * i_nodeTag, 6 * i_nodeTag + 6), np.arange(6 * j_nodeTag, 6 * j_nodeTag + 6)])
        r, c = np.meshgrid(dofs, dofs, indexing="ij")
        return np.stack([r.ravel(), c.ravel()], 1).astype(np.int32)


class Truss:
    """Truss element: attributes only, as in the reference (``element.py:361-382``),
    which never wires it into assembly (``model.py:130-151``)."""

    def __init__(self, eleTag, i_nodeTag, j_nodeTag, E, A):
        self.dofs = 2 * 6
        self.eleTag = eleTag
        self.i_nodeTag = i_nodeTag
        self.j_nodeTag = j_nodeTag
        self.E = E
        self.A = A


class Quad:
    """24-DOF MITC-4 flat shell (``element.py:427-470``)."""

    def __init__(self, eleTag, i_nodeTag, j_nodeTag, m_nodeTag, n_nodeTag, t, E, nu, kx_mod=1.0, ky_mod=1.0):
        self.dofs = 4 * 6
        self.eleTag = eleTag
        self.i_nodeTag = i_nodeTag
        self.j_nodeTag = j_nodeTag
        self.m_nodeTag = m_nodeTag
        self.n_nodeTag = n_nodeTag
        self.t, self.E, self.nu, self.kx_mod, self.ky_mod = t, E, nu, kx_mod, ky_mod

    @staticmethod
    def A(crds):
        """Surface area: two triangles (1,2,4) and (3,4,2) (``Quad.A``, element.py:471-487)."""
        c, single = _as_batch(crds, 12)
        X = [c[:, 3 * a:3 * a + 3] for a in range(4)]
        a1 = 0.5 * np.linalg.norm(np.cross(X[1] - X[0], X[3] - X[0]), axis=1)
        a2 = 0.5 * np.linalg.norm(np.cross(X[3] - X[2], X[1] - X[2]), axis=1)
        out = a1 + a2
        return out[0] if single else out

    @staticmethod
    def element_K_quad(crds, t, E, nu, kx_mod=1.0, ky_mod=1.0):
        """Global stiffness (24,24) -- or (n,24,24) -- of ``Quad.element_K_quad``
        (element.py:1073-1084), computed on the GPU."""
        c, single = _as_batch(crds, 12)
        props = np.stack(np.broadcast_arrays(*[np.asarray(v, np.float64).ravel() for v in (t, E, nu, kx_mod, ky_mod)]), 1)
        props = np.ascontiguousarray(np.broadcast_to(props, (len(c), 5)))
        L = nat.lib()
        dev = nat.device()
        ct = torch.tensor(c, device=dev)
        pt = torch.tensor(props, device=dev)
        ke = torch.empty((len(c), 24, 24), dtype=torch.float64, device=dev)
        nat.check(L.jsso_element_ke_quad(len(c), nat.ptr(ct), nat.ptr(pt), nat.ptr(ke), nat.stream_ptr()))
        return ke[0] if single else ke

    @staticmethod
    def element_K_quad_indices(i_nodeTag, j_nodeTag, m_nodeTag, n_nodeTag):
        """(576, 2) global (row, col) of the element's entries (element.py:1097-1106)."""
        dofs = np.concatenate([np.arange(6 * t, 6 * t + 6) for t in (i_nodeTag, j_nodeTag, m_nodeTag, n_nodeTag)])
        r, c = np.meshgrid(dofs, dofs, indexing="ij")
        return np.stack([r.ravel(), c.ravel()], 1).astype(np.int32)
