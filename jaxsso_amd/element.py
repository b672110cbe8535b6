"""Element types of the model builder (``JaxSSO/element.py``).

``BeamCol``, ``Truss`` and ``Quad`` hold an element's attributes exactly as the
reference's classes do (``element.py:24-64, 361-382, 427-470``).  Their static
routines -- frames, local and global stiffness, the MITC-4 Jacobian, B and
constitutive matrices, and the per-type global stiffness ``K_beamcol`` /
``K_quad`` -- run on the GPU through the C ABI (``jsso_element_*``,
``jsso_assemble``), vectorised over a leading element axis, and return device
tensors (the per-type global stiffness returns a ``BlockK``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat


def _as_batch(x, width):
    a = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    single = a.ndim == 1
    return np.ascontiguousarray(a.reshape(-1, width)), single


def _dev(a):
    return torch.as_tensor(np.array(a, dtype=np.float64, order="C"), device=nat.device())


def _props(n, *cols):
    p = np.stack(np.broadcast_arrays(*[np.asarray(v, np.float64).ravel() for v in cols]), 1)
    return np.ascontiguousarray(np.broadcast_to(p, (n, len(cols))))


def _gather(node_crds, cnct):
    """Element coordinates [x1,y1,z1,x2,...] (n_el, 3*nen) from node coordinates and connectivity."""
    c = node_crds.detach().cpu().numpy() if isinstance(node_crds, torch.Tensor) else np.asarray(node_crds, float)
    cn = np.asarray(cnct.detach().cpu().numpy() if isinstance(cnct, torch.Tensor) else cnct, dtype=np.int64)
    return np.ascontiguousarray(c.reshape(-1, 3)[cn].reshape(len(cn), -1))


def _host(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _beam_local(c, props, want_T, want_K):
    L = nat.lib()
    n = len(c)
    T = torch.empty((n, 12, 12), dtype=torch.float64, device=nat.device()) if want_T else None
    K = torch.empty((n, 12, 12), dtype=torch.float64, device=nat.device()) if want_K else None
    ct = _dev(c)
    pt = _dev(props) if props is not None else None
    nat.check(L.jsso_element_beam_local(n, nat.ptr(ct), nat.ptr(pt), nat.ptr(T), nat.ptr(K), nat.stream_ptr()))
    return T, K


def _quad_local(c, props, xy=False, normals=False, T=False, K=False):
    L = nat.lib()
    n = len(c)
    dev = nat.device()
    out = {"xy": torch.empty((n, 8), dtype=torch.float64, device=dev) if xy else None,
           "normals": torch.empty((n, 3, 3), dtype=torch.float64, device=dev) if normals else None,
           "T": torch.empty((n, 24, 24), dtype=torch.float64, device=dev) if T else None,
           "K": torch.empty((n, 24, 24), dtype=torch.float64, device=dev) if K else None}
    ct = _dev(c)
    pt = _dev(props) if props is not None else None
    nat.check(L.jsso_element_quad_local(n, nat.ptr(ct), nat.ptr(pt), nat.ptr(out["xy"]), nat.ptr(out["normals"]),
                                        nat.ptr(out["T"]), nat.ptr(out["K"]), nat.stream_ptr()))
    return out


def _quad_gauss(xy, props, r, s, which):
    """which: subset of J, Bk, Bg, Bm, Cb, Cs, Cm -> device tensors."""
    L = nat.lib()
    shapes = {"J": (2, 2), "Bk": (3, 12), "Bg": (2, 12), "Bm": (3, 8), "Cb": (3, 3), "Cs": (2, 2), "Cm": (3, 3)}
    n = len(xy) if xy is not None else len(props)
    out = {k: (torch.empty((n,) + shapes[k], dtype=torch.float64, device=nat.device()) if k in which else None)
           for k in shapes}
    xt = _dev(xy) if xy is not None else None
    pt = _dev(props) if props is not None else None
    nat.check(L.jsso_element_quad_gauss(n, nat.ptr(xt), nat.ptr(pt), float(r), float(s), nat.ptr(out["J"]),
                                        nat.ptr(out["Bk"]), nat.ptr(out["Bg"]), nat.ptr(out["Bm"]),
                                        nat.ptr(out["Cb"]), nat.ptr(out["Cs"]), nat.ptr(out["Cm"]),
                                        nat.stream_ptr()))
    return out


class BeamCol:
    """12-DOF 3-D Euler-Bernoulli beam-column (``element.py:24-64``)."""

    def __init__(self, eleTag, i_nodeTag, j_nodeTag, E, G, Iy, Iz, J, A):
        self.dofs = 2 * 6
        self.eleTag = eleTag
        self.i_nodeTag = i_nodeTag
        self.j_nodeTag = j_nodeTag
        self.E, self.G, self.Iy, self.Iz, self.J, self.A = E, G, Iy, Iz, J, A

    @staticmethod
    def T(crds):
        """(12,12) transformation matrix (or (n,12,12)) of ``BeamCol.T`` (element.py:66-105),
        computed on the GPU (``jsso_element_beam_local``)."""
        c, single = _as_batch(crds, 6)
        T, _ = _beam_local(c, None, True, False)
        return T[0] if single else T

    @staticmethod
    def K_local(crds, E, G, Iy, Iz, J, A):
        """(12,12) local stiffness (or (n,12,12)) of ``BeamCol.K_local`` (element.py:107-128), on the GPU."""
        c, single = _as_batch(crds, 6)
        _, K = _beam_local(c, _props(len(c), E, G, Iy, Iz, J, A), False, True)
        return K[0] if single else K

    @staticmethod
    def K_local_beamcol(node_crds, prop, cnct):
        """Local stiffness of every beam-column, (n_beamcol,12,12) (``element.py:157-197``)."""
        _, K = _beam_local(_gather(node_crds, cnct), np.ascontiguousarray(_host(prop), dtype=np.float64),
                           False, True)
        return K

    @staticmethod
    def T_beamcol(node_crds, cnct):
        """Transformation matrix of every beam-column, (n_beamcol,12,12) (``element.py:198-226``)."""
        T, _ = _beam_local(_gather(node_crds, cnct), None, True, False)
        return T

    @staticmethod
    def K_beamcol(node_crds, prop, cnct, ndof):
        """Global stiffness of the beam-columns alone (``BeamCol.K_beamcol``, element.py:228-275),
        as the :class:`~jaxsso_amd.assemblemodel.BlockK` the assembly kernels build."""
        from . import assemblemodel
        cn = np.asarray(_host(cnct), dtype=np.int32).reshape(-1, 2)
        return assemblemodel.element_K(node_crds, ndof, cn, _host(prop), np.zeros((0, 4), np.int32),
                                       np.zeros((0, 5)))

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
    def loc_crds(crds):
        """``Quad.loc_crds`` (element.py:488-539): (new_xys (8,), normal (3,), normal_tri1 (3,),
        normal_tri2 (3,)), or batched with a leading element axis; computed on the GPU."""
        c, single = _as_batch(crds, 12)
        o = _quad_local(c, None, xy=True, normals=True)
        xy, nv = o["xy"], o["normals"]
        if single:
            return xy[0], nv[0, 0], nv[0, 1], nv[0, 2]
        return xy, nv[:, 0], nv[:, 1], nv[:, 2]

    @staticmethod
    def T(crds):
        """(24,24) transformation matrix (or (n,24,24)) of ``Quad.T`` (element.py:643-696), on the GPU."""
        c, single = _as_batch(crds, 12)
        T = _quad_local(c, None, T=True)["T"]
        return T[0] if single else T

    @staticmethod
    def J(new_xys, r, s):
        """(2,2) Jacobian at (r, s) (``Quad.J``, element.py:698-709), on the GPU."""
        xy, single = _as_batch(new_xys, 8)
        J = _quad_gauss(xy, None, r, s, ("J",))["J"]
        return J[0] if single else J

    @staticmethod
    def B_kappa(new_xys, r, s, J=None):
        """(3,12) bending B at (r, s) (``Quad.B_kappa``, element.py:711-733); ``J`` is accepted
        for the reference's signature (the Jacobian is formed from ``new_xys``)."""
        xy, single = _as_batch(new_xys, 8)
        B = _quad_gauss(xy, None, r, s, ("Bk",))["Bk"]
        return B[0] if single else B

    @staticmethod
    def B_gamma_MITC4(new_xys, r, s, J=None):
        """(2,12) MITC-4 assumed transverse-shear B at (r, s) (element.py:735-801)."""
        xy, single = _as_batch(new_xys, 8)
        B = _quad_gauss(xy, None, r, s, ("Bg",))["Bg"]
        return B[0] if single else B

    @staticmethod
    def B_m(new_xys, r, s, J=None):
        """(3,8) membrane B at (r, s) (``Quad.B_m``, element.py:803-818)."""
        xy, single = _as_batch(new_xys, 8)
        B = _quad_gauss(xy, None, r, s, ("Bm",))["Bm"]
        return B[0] if single else B

    @staticmethod
    def Cb(nu, E, h):
        """(3,3) plate-bending constitutive matrix (``Quad.Cb``, element.py:820-834)."""
        p = _props(max(np.size(nu), np.size(E), np.size(h)), h, E, nu, 1.0, 1.0)
        C = _quad_gauss(None, p, 0.0, 0.0, ("Cb",))["Cb"]
        return C[0] if np.ndim(nu) == np.ndim(E) == np.ndim(h) == 0 else C

    @staticmethod
    def Cs(nu, E, h):
        """(2,2) transverse-shear constitutive matrix, k = 5/6 (``Quad.Cs``, element.py:836-850)."""
        p = _props(max(np.size(nu), np.size(E), np.size(h)), h, E, nu, 1.0, 1.0)
        C = _quad_gauss(None, p, 0.0, 0.0, ("Cs",))["Cs"]
        return C[0] if np.ndim(nu) == np.ndim(E) == np.ndim(h) == 0 else C

    @staticmethod
    def Cm(nu, E, kx_mod=1.0, ky_mod=1.0):
        """(3,3) plane-stress constitutive matrix with stiffness modifiers (``Quad.Cm``, element.py:852-875)."""
        n = max(np.size(nu), np.size(E), np.size(kx_mod), np.size(ky_mod))
        p = _props(n, 1.0, E, nu, kx_mod, ky_mod)
        C = _quad_gauss(None, p, 0.0, 0.0, ("Cm",))["Cm"]
        return C[0] if max(np.ndim(nu), np.ndim(E), np.ndim(kx_mod), np.ndim(ky_mod)) == 0 else C

    @staticmethod
    def element_K_quad_local(crds, t, E, nu, kx_mod=1.0, ky_mod=1.0):
        """(24,24) stiffness before the rotation, k_m + k_b (``Quad.element_K_quad_local``,
        element.py:1086-1095), on the GPU."""
        c, single = _as_batch(crds, 12)
        K = _quad_local(c, _props(len(c), t, E, nu, kx_mod, ky_mod), K=True)["K"]
        return K[0] if single else K

    @staticmethod
    def K_local_quad(node_crds, prop, cnct):
        """Local stiffness of every quad, (n_quad,24,24) (``element.py:1113-1156``)."""
        return _quad_local(_gather(node_crds, cnct), np.ascontiguousarray(_host(prop), dtype=np.float64),
                           K=True)["K"]

    @staticmethod
    def T_quad(node_crds, cnct):
        """Transformation matrix of every quad, (n_quad,24,24) (``element.py:1157-1189``)."""
        return _quad_local(_gather(node_crds, cnct), None, T=True)["T"]

    @staticmethod
    def K_quad(node_crds, prop, cnct, ndof):
        """Global stiffness of the quads alone (``Quad.K_quad``, element.py:1190-1241), as the
        :class:`~jaxsso_amd.assemblemodel.BlockK` the assembly kernels build."""
        from . import assemblemodel
        cn = np.asarray(_host(cnct), dtype=np.int32).reshape(-1, 4)
        return assemblemodel.element_K(node_crds, ndof, np.zeros((0, 2), np.int32), np.zeros((0, 6)), cn,
                                       _host(prop))

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
