"""Device-side finite-element plan: one or many structures packed together.

A :class:`FEPlan` owns a ``jsso_plan`` (libjsso_hip.so): the 6x6 block-CSR
pattern of K, the element-to-block contribution map and the node adjacency.
It is built once per topology (connectivity + supports) and then reused for
any coordinates / section properties, which is how ``SSO_model`` evaluates
many designs of one model.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from . import _native as nat

DEFAULT_TOL = 1e-10
DEFAULT_MAXIT = 200000
PRECONDITIONERS = {"block_jacobi": 0, "amg": 1}


class FEPlan:
    """Symbolic plan for ``n_sys`` structures packed into one node numbering.

    Parameters are host arrays: ``node_off`` (n_sys+1) node offsets,
    ``cnct_beam`` (nb,2) and ``cnct_quad`` (nq,4) global node ids, ``known_id``
    global constrained DOF ids.
    """

    def __init__(self, node_off, cnct_beam, cnct_quad, known_id):
        L = nat.load_library()
        self.node_off = np.ascontiguousarray(node_off, dtype=np.int32)
        self.cnct_beam = np.ascontiguousarray(np.asarray(cnct_beam, dtype=np.int32).reshape(-1, 2))
        self.cnct_quad = np.ascontiguousarray(np.asarray(cnct_quad, dtype=np.int32).reshape(-1, 4))
        self.known_id = np.ascontiguousarray(np.asarray(known_id, dtype=np.int32).ravel())
        self.n_sys = len(self.node_off) - 1
        self.n_nodes = int(self.node_off[-1])
        self.ndof = 6 * self.n_nodes
        self.n_beam = len(self.cnct_beam)
        self.n_quad = len(self.cnct_quad)
        h = ctypes.c_void_p()
        nat.check(L.jsso_plan_create(self.n_sys, nat.ptr(self.node_off), self.n_beam, nat.ptr(self.cnct_beam),
                                     self.n_quad, nat.ptr(self.cnct_quad), len(self.known_id),
                                     nat.ptr(self.known_id), ctypes.byref(h)))
        self._h = h
        sizes = np.zeros(6, np.int64)
        nat.check(L.jsso_plan_sizes(self._h, nat.ptr(sizes)))
        self.nnzb = int(sizes[1])
        self.preconditioner = "amg"   # the library's default
        self.last_info = None
        self._assembled_key = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and nat._lib is not None:
            nat._lib.jsso_plan_destroy(h)
            self._h = None

    # -- pattern ---------------------------------------------------------
    def pattern(self):
        """Host copy of the block pattern (rowptr (n+1,), colind (nnzb,)) as int32."""
        rowptr = np.zeros(self.n_nodes + 1, np.int32)
        colind = np.zeros(self.nnzb, np.int32)
        nat.check(nat.load_library().jsso_plan_pattern(self._h, nat.ptr(rowptr), nat.ptr(colind)))
        return rowptr, colind

    # -- numerics --------------------------------------------------------
    def assemble(self, crds: torch.Tensor, prop_beam: torch.Tensor | None, prop_quad: torch.Tensor | None,
                 export_K: bool = False):
        """Assemble K (FP64) for node coordinates (n,3) and properties; optionally return K values (nnzb,6,6)."""
        L = nat.lib()
        crds = _dev64(crds)
        pb = _dev64(prop_beam) if self.n_beam else None
        pq = _dev64(prop_quad) if self.n_quad else None
        assert crds.shape == (self.n_nodes, 3)
        Kout = torch.empty((self.nnzb, 6, 6), dtype=torch.float64, device=crds.device) if export_K else None
        nat.check(L.jsso_assemble(self._h, nat.ptr(crds), nat.ptr(pb), nat.ptr(pq), nat.ptr(Kout), nat.stream_ptr()))
        self._keep = (crds, pb, pq)
        return Kout

    def ensure_assembled(self, crds, prop_beam, prop_quad):
        """Assemble unless the plan's current assembly was made from these very tensors."""
        k = getattr(self, "_keep", None)
        if k is None or k[0] is not crds or k[1] is not prop_beam or k[2] is not prop_quad:
            self.assemble(crds, prop_beam, prop_quad)

    def symmetric(self) -> bool:
        """False when the last assembly holds beam-columns parallel to global Y (the
        reference's transformation for them is not orthogonal, element.py:92-94) or
        quads with kx_mod != ky_mod (non-symmetric Cm, element.py:852-875): K is not
        symmetric, solves run on the dense LU path and adjoints need K^T."""
        out = ctypes.c_int32(1)
        nat.check(nat.lib().jsso_plan_symmetric(self._h, ctypes.byref(out), nat.stream_ptr()))
        return bool(out.value)

    def solve_transpose(self, f: torch.Tensor, **kw) -> torch.Tensor:
        """K^T u = f (u = 0 at supports): the adjoint solve.  K is symmetric unless
        the assembly holds members parallel to global Y or orthotropic quads; then
        ``jsso_solve_transpose`` takes the dense LU of K^T or, for large plans
        without Y-parallel members, the defect correction (``nonsymmetric_route``)."""
        return self.solve(f, transpose=True, **kw)

    def nonsymmetric_route(self, mode: int = -1) -> int:
        """How non-symmetric plans are solved: 0 automatic (dense LU up to 16 384
        DOF or with members parallel to global Y, else defect correction on the
        symmetric part), 1 dense LU, 2 defect correction; -1 only reads it."""
        out = ctypes.c_int32(0)
        nat.check(nat.lib().jsso_plan_nonsymmetric_route(self._h, int(mode), ctypes.byref(out)))
        return int(out.value)

    def set_preconditioner(self, name: str) -> None:
        """Preconditioner of the batched precisions: 'amg' (smoothed-aggregation
        multigrid V-cycle, the default) or 'block_jacobi' (split block-Jacobi scaling)."""
        if name not in PRECONDITIONERS:
            raise ValueError(f"unknown preconditioner {name!r} ({sorted(PRECONDITIONERS)})")
        nat.check(nat.load_library().jsso_plan_set_preconditioner(self._h, PRECONDITIONERS[name]))
        self.preconditioner = name

    def element_tuning(self, set_epw: int = -1) -> int:
        """MITC-4 elements per workgroup of the assembly kernel (4, 8 or 16): 0 until
        the first assembly has timed the variants on this plan; set_epw 4/8/16
        forces one, 0 re-tunes at the next assembly."""
        out = ctypes.c_int32(0)
        nat.check(nat.load_library().jsso_plan_element_tuning(self._h, int(set_epw), ctypes.byref(out)))
        return int(out.value)

    def amg_reuse(self, mode: int = -1) -> dict:
        """Reuse of the AMG hierarchy across assemblies (optimisation loops): mode 0
        rebuilds it after every assembly (default), 1 keeps the hierarchy of an
        earlier design and refreshes only the level-0 smoothing bound, 2 rebuilds
        once at the next solve and then keeps it; -1 only reads.  Returns the mode
        and the counts of full set-ups and refreshes run so far."""
        out = np.zeros(3, np.int64)
        nat.check(nat.load_library().jsso_plan_amg_reuse(self._h, int(mode), nat.ptr(out)))
        return {"mode": int(out[0]), "full_setups": int(out[1]), "refreshes": int(out[2])}

    def amg_levels(self) -> list:
        """Block rows of every level of the plan's AMG hierarchy (built on the host)."""
        nodes = np.zeros(32, np.int32)
        nl = ctypes.c_int32(0)
        nat.check(nat.load_library().jsso_plan_amg_levels(self._h, nat.ptr(nodes), 32, ctypes.byref(nl)))
        return [int(v) for v in nodes[:nl.value]]

    def solve(self, f: torch.Tensor, precision: str = "fp64", tol: float = DEFAULT_TOL,
              maxit: int = DEFAULT_MAXIT, warn: bool = True, preconditioner: str | None = None,
              transpose: bool = False) -> torch.Tensor:
        """Solve K u = f (K^T u = f with ``transpose``; u = 0 at supports).  f: (ndof,)
        or (ndof, nrhs) FP64 on the device."""
        L = nat.lib()
        if preconditioner is not None and preconditioner != self.preconditioner:
            self.set_preconditioner(preconditioner)
        f = _dev64(f)
        vec = f.dim() == 1
        f2 = f.reshape(self.ndof, -1).contiguous()
        nrhs = f2.shape[1]
        u = torch.empty_like(f2)
        info = nat.Info()
        entry = L.jsso_solve_transpose if transpose else L.jsso_solve
        rc = entry(self._h, nat.ptr(f2), nat.ptr(u), nrhs, nat.PREC[precision], float(tol), int(maxit),
                   ctypes.byref(info), nat.stream_ptr())
        nat.check(rc, allow_not_converged=True)
        self.last_info = info.as_dict()
        if rc == nat.JSSO_NOT_CONVERGED and warn:
            warnings.warn(f"CG did not converge: {self.last_info}", RuntimeWarning, stacklevel=2)
        return u.reshape(f.shape) if vec else u

    def compliance(self, f: torch.Tensor, u: torch.Tensor, scale: float = 0.5) -> torch.Tensor:
        """scale * f.u per structure and load case, (n_sys, nrhs), reduced on the GPU
        in FP64 in a fixed order."""
        L = nat.lib()
        f2 = _dev64(f).reshape(self.ndof, -1).contiguous()
        u2 = _dev64(u).reshape(self.ndof, -1).contiguous()
        if f2.shape != u2.shape:
            raise ValueError(f"f {tuple(f2.shape)} and u {tuple(u2.shape)} differ")
        C = torch.empty((self.n_sys, f2.shape[1]), dtype=torch.float64, device=u2.device)
        nat.check(L.jsso_compliance(self._h, nat.ptr(f2), nat.ptr(u2), f2.shape[1], float(scale), nat.ptr(C),
                                    nat.stream_ptr()))
        return C

    def cg_tuning(self, precision: str = "mixed", nrhs: int = 1, set_wps: int = -1) -> int:
        """Waves per slice the batched CG kernel uses for (precision, nrhs): 0 until
        the first solve has tuned it.  set_wps 1/2/4 forces a variant, 0 re-tunes."""
        L = nat.lib()
        out = ctypes.c_int32(0)
        nat.check(L.jsso_plan_cg_tuning(self._h, nat.PREC[precision], int(nrhs), int(set_wps), ctypes.byref(out)))
        return int(out.value)

    def adjoint(self, crds, prop_beam, prop_quad, u, lam=None, lam_scale: float = 0.5, want=("crds", "props")):
        """g = -sum_e lam_e^T dK_e/dp u_e for coordinates and all element properties.

        lam=None uses lam = lam_scale * u (compliance: lam_scale = 1/2).
        Returns (dcrds (n,3[,nrhs]), dprop_beam (nb,6[,nrhs]), dprop_quad (nq,5[,nrhs])).
        """
        L = nat.lib()
        crds = _dev64(crds)
        pb = _dev64(prop_beam) if self.n_beam else None
        pq = _dev64(prop_quad) if self.n_quad else None
        u = _dev64(u)
        vec = u.dim() == 1
        u2 = u.reshape(self.ndof, -1).contiguous()
        nrhs = u2.shape[1]
        lam2 = None if lam is None else _dev64(lam).reshape(self.ndof, -1).contiguous()
        dev = crds.device
        dc = torch.empty((self.n_nodes, 3, nrhs), dtype=torch.float64, device=dev) if "crds" in want else None
        db = torch.empty((self.n_beam, 6, nrhs), dtype=torch.float64, device=dev) if ("props" in want and self.n_beam) else None
        dq = torch.empty((self.n_quad, 5, nrhs), dtype=torch.float64, device=dev) if ("props" in want and self.n_quad) else None
        nat.check(L.jsso_adjoint_grad(self._h, nat.ptr(crds), nat.ptr(pb), nat.ptr(pq), nat.ptr(u2), nat.ptr(lam2),
                                      float(lam_scale), nrhs, nat.ptr(dc), nat.ptr(db), nat.ptr(dq),
                                      nat.stream_ptr()))
        if vec:
            dc = None if dc is None else dc[..., 0]
            db = None if db is None else db[..., 0]
            dq = None if dq is None else dq[..., 0]
        return dc, db, dq

    def spmv(self, x: torch.Tensor, K: torch.Tensor | None = None) -> torch.Tensor:
        """y = K x with the unconstrained K: the block values ``K`` ((nnzb,6,6) in
        this plan's pattern) if given, else the plan's current assembly."""
        L = nat.lib()
        x = _dev64(x)
        y = torch.empty_like(x)
        Kv = None if K is None else _dev64(K)
        if Kv is not None and Kv.numel() != self.nnzb * 36:
            raise ValueError(f"K has {Kv.numel()} values, the pattern {self.nnzb * 36}")
        nat.check(L.jsso_spmv_k(self._h, nat.ptr(Kv), nat.ptr(x), nat.ptr(y), nat.stream_ptr()))
        return y

    def dense_solve(self, f: torch.Tensor, K: torch.Tensor | None = None, transpose: bool = False) -> torch.Tensor:
        """u with u = 0 at supports and K_ff u_f = f_f (K_ff^T with ``transpose``), by
        the dense FP64 factorisation of the constrained K on the device
        (``jsso_dense_solve``: Cholesky, or pivoted LU when the assembly holds
        members parallel to global Y).  ``K``: block values ((nnzb,6,6)) or None for
        the plan's current assembly; f: (ndof,) or (ndof, m)."""
        L = nat.lib()
        f = _dev64(f)
        vec = f.dim() == 1
        f2 = f.reshape(self.ndof, -1).contiguous()
        u = torch.empty_like(f2)
        Kv = None if K is None else _dev64(K)
        if Kv is not None and Kv.numel() != self.nnzb * 36:
            raise ValueError(f"K has {Kv.numel()} values, the pattern {self.nnzb * 36}")
        nat.check(L.jsso_dense_solve(self._h, nat.ptr(Kv), nat.ptr(f2), nat.ptr(u), f2.shape[1], int(bool(transpose)),
                                     nat.stream_ptr()))
        self.last_info = {"iterations": 0, "converged": self.n_sys * f2.shape[1], "method": "dense"}
        return u.reshape(f.shape) if vec else u

    def K_scalar_csr(self, Kvals: torch.Tensor):
        """Expand block values (nnzb,6,6) to a host scipy CSR matrix."""
        rowptr, colind = self.pattern()
        return blocks_to_csr(rowptr, colind, Kvals.detach().cpu().numpy())


class AmgKeepPolicy:
    """Keeps a plan's AMG hierarchy across the assemblies of an optimisation loop
    (``FEPlan.amg_reuse`` mode 1: after each assembly only the level-0 smoothing
    bound follows the new design) and rebuilds it once the CG iteration count has
    grown by more than ``growth`` over the count of the first solve after the
    last rebuild.  Call :meth:`after_solve` after every solve."""

    def __init__(self, plan: "FEPlan", growth: float = 1.25):
        self.plan = plan
        self.growth = float(growth)
        self.base = None
        self.rebuilds = 0
        plan.amg_reuse(2)          # build at the next solve, then keep

    def after_solve(self) -> bool:
        """True when this solve's iteration count triggers a rebuild at the next solve."""
        it = self.plan.last_info["iterations"]
        if self.base is None:
            self.base = it
            return False
        if it > self.growth * self.base:
            self.plan.amg_reuse(2)
            self.base = None
            self.rebuilds += 1
            return True
        return False


def blocks_to_csr(rowptr, colind, vals):
    """Scalar CSR (scipy.sparse.csr_matrix, explicit zeros kept) of a 6x6 block-CSR matrix."""
    import scipy.sparse as sp
    n = len(rowptr) - 1
    B = sp.bsr_matrix((np.asarray(vals, dtype=np.float64).reshape(-1, 6, 6), np.asarray(colind),
                       np.asarray(rowptr)), shape=(6 * n, 6 * n))
    return B.tocsr()


def _dev64(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t, dtype=np.float64))
    return t.to(device=nat.device(), dtype=torch.float64).contiguous()
