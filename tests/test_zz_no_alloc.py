"""jsso_adjoint_grad_noalloc (fixed device scratch, no allocation) against
jsso_adjoint_grad through the C ABI: two calls in a row, the second on a
smaller plan, results equal bit for bit (same kernels, same summation order);
a plan whose scratch need exceeds the fixed array is refused from the host."""
import ctypes

import numpy as np
import pytest
import torch

from jaxsso_amd import _native as nat
from jaxsso_amd import workloads as wl
from jaxsso_amd.batch import StructureBatch

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.jsso_adjoint_grad_noalloc
    P = ctypes.c_void_p
    f.argtypes = [P, P, P, P, P, P, ctypes.c_double, ctypes.c_int32, P, P, P, P]
    f.restype = ctypes.c_int
    return f


def _grads(b, u, fn):
    p = b.plan
    dev = u.device
    nrhs = u.shape[1] if u.dim() == 2 else 1
    dc = torch.full((p.n_nodes, 3, nrhs), float("nan"), dtype=torch.float64, device=dev)
    db = torch.full((p.n_beam, 6, nrhs), float("nan"), dtype=torch.float64, device=dev) if p.n_beam else None
    dq = torch.full((p.n_quad, 5, nrhs), float("nan"), dtype=torch.float64, device=dev) if p.n_quad else None
    pb = b.prop_beam.contiguous() if p.n_beam else None
    pq = b.prop_quad.contiguous() if p.n_quad else None
    rc = fn(p._h, nat.ptr(b.crds), nat.ptr(pb), nat.ptr(pq), nat.ptr(u.contiguous()), nat.ptr(None), 0.5, nrhs,
            nat.ptr(dc), nat.ptr(db), nat.ptr(dq), nat.stream_ptr())
    torch.cuda.synchronize()
    return rc, dc, db, dq


def test_noalloc_adjoint_matches_allocating_one():
    lib = nat.lib()
    na = _noalloc(lib)
    big = StructureBatch([wl.gridshell(12, cfg=3, design=1, half_span_cases=True), wl.cfg1(2)], n_cases=4)
    small = StructureBatch([wl.shell_pinned_edges(6, 6.0, 1.0, design=4)])
    for b in (big, small):
        u = b.solve(precision="fp64", tol=1e-10)
        rc0, *ref = _grads(b, u, lib.jsso_adjoint_grad)
        rc1, *got = _grads(b, u, na)
        assert rc0 == 0 and rc1 == 0, nat.last_error()
        for r, g in zip(ref, got):
            if r is None:
                continue
            assert torch.isfinite(r).all()
            assert torch.equal(r, g)


def test_noalloc_adjoint_refuses_a_plan_beyond_its_scratch():
    lib = nat.lib()
    na = _noalloc(lib)
    b = StructureBatch(wl.cfg4(52))          # 520,000 quads x 17 x 8 B > 64 MiB
    b.assemble()
    u = torch.zeros((b.plan.ndof, 1), dtype=torch.float64, device="cuda")
    rc, *_ = _grads(b, u, na)
    assert rc == -1          # JSSO_EINVAL, from the host-side size check
    assert "scratch" in nat.last_error()
    # the allocating entry point still serves it
    rc0, dc, _, dq = _grads(b, u, lib.jsso_adjoint_grad)
    assert rc0 == 0 and float(dc.abs().max()) == 0.0
