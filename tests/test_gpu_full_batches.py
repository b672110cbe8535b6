"""The two batched configurations at their full size on the GPU.

* 64 SIMP shells (100x100 MITC-4, E = rho^3 30 GPa): every design's compliance
  and gradient satisfy the exact invariants of the problem -- a rigid
  translation of all nodes leaves C unchanged (sum of dC/dx over the nodes is
  0 in each direction) and K is linear in E (sum_e E_e dC/dE_e = -C; for beams
  in E and G together); design 0's
  dC/drho and dC/dz against central finite differences of FP64 solves.
* 256 uneven beam gridshells x 4 load cases: the same invariants for every
  design and case (E of every member), and the oracle on 16 designs spread over
  the size range.
The oracle itself is too slow for all 64 SIMP designs; designs 0 and 1 are
compared with it in test_gpu_simp.py."""
import numpy as np
import pytest
import torch

from jaxsso_amd import workloads as wl
from jaxsso_amd.batch import StructureBatch
from oracle import gradient as og

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _invariants(b, C, dc, dE_times_E, case=None):
    """Per design (and case): |sum_n dC/dx_n| / (sum_n |dC/dx_n|) and
    |sum_e E dC/dE + C| / C."""
    worst_t, worst_h = 0.0, 0.0
    dcs = b.split_nodes(dc)
    for i in range(b.n_structures):
        g = dcs[i] if case is None else dcs[i][..., case]
        worst_t = max(worst_t, float(np.abs(g.sum(0)).max() / np.abs(g).sum(0).max()))
        Ci = C[i] if case is None else C[i, case]
        worst_h = max(worst_h, abs(dE_times_E[i] + Ci) / abs(Ci))
    return worst_t, worst_h


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_simp_full_batch_invariants(precision):
    designs = wl.cfg4(64)
    b = StructureBatch(designs)
    C, (dc, _, dq) = b.compliance_and_gradient(precision=precision, tol=1e-10)
    assert b.plan.last_info["converged"] == 64, b.plan.last_info
    C = C.cpu().numpy()[:, 0]
    dc = dc.cpu().numpy().reshape(-1, 3)
    dq = dq.cpu().numpy().reshape(-1, 5)
    E = b.prop_quad.cpu().numpy()[:, 1]
    sEdE = np.add.reduceat(E * dq[:, 1], np.asarray(b.quad_off[:-1]))
    t, h = _invariants(b, C, dc, sEdE)
    assert t < 1e-9 and h < 1e-9, (t, h)


def test_simp_design0_gradient_vs_central_differences():
    st = wl.simp_shell(0)
    b = StructureBatch([st])
    rho = st.rho
    C0, (dc, _, dq) = b.compliance_and_gradient(precision="fp64", tol=1e-11)
    drho = dq.cpu().numpy()[:, 1] * 3.0 * rho ** 2 * 30e9
    dz = dc.cpu().numpy()[:, 2]
    order = np.argsort(-np.abs(drho))
    elems = list(order[:3]) + list(np.random.default_rng(0).choice(st.n_quad, 3, replace=False))
    nodes = list(st.design_nodes[np.argsort(-np.abs(dz[st.design_nodes]))[:3]])
    pq0 = b.prop_quad.clone()
    cr0 = b.crds.clone()

    def C_of():
        C, _ = b.compliance_and_gradient(precision="fp64", tol=1e-11)
        return float(C[0, 0])

    # step 1e-3 rho: the truncation error (h^2/6 times the third derivative)
    # stays ~1e-7 relative, and the solves' compliance noise (~3e-12 C at the
    # attainable floor) divided by 2h stays under the bar for small gradients too
    for e in elems:
        h = 1e-3 * rho[e]
        vals = []
        for sgn in (1, -1):
            b.prop_quad.copy_(pq0)
            b.prop_quad[e, 1] = (rho[e] + sgn * h) ** 3 * 30e9
            vals.append(C_of())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - drho[e]) <= 2e-5 * abs(drho[e]) + 1e-7 * np.abs(drho).max(), (e, fd, drho[e])
    b.prop_quad.copy_(pq0)
    for n in nodes:
        h = 1e-4
        vals = []
        for sgn in (1, -1):
            b.crds.copy_(cr0)
            b.crds[n, 2] += sgn * h
            vals.append(C_of())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - dz[n]) <= 2e-5 * abs(dz[n]), (n, fd, dz[n])
    b.crds.copy_(cr0)


def test_gridshell_full_batch_invariants_and_oracle_sample():
    designs = wl.cfg3(256)
    b = StructureBatch(designs)
    C, (dc, db, _) = b.compliance_and_gradient(precision="mixed", tol=1e-10)
    assert b.plan.last_info["converged"] == 1024, b.plan.last_info
    C = C.cpu().numpy()
    dc = dc.cpu().numpy()
    db = db.cpu().numpy()
    E = b.prop_beam.cpu().numpy()[:, 0]
    G = b.prop_beam.cpu().numpy()[:, 1]
    starts = np.asarray(b.beam_off[:-1])
    for c in range(4):
        # a beam's K is linear in (E, G) together (torsion is G J / L)
        sEdE = np.add.reduceat(E * db[:, 0, c] + G * db[:, 1, c], starts)
        t, h = _invariants(b, C, dc, sEdE, case=c)
        assert t < 1e-9 and h < 1e-9, (c, t, h)
    sizes = np.array([d.n_node for d in designs])
    sample = np.argsort(sizes)[np.linspace(0, 255, 16).astype(int)]
    dcs = b.split_nodes(dc)
    for i in sample:
        st = designs[i]
        for c in range(4):
            ref = og.compliance_and_gradient(st, loads=st.load_cases()[c])
            assert abs(C[i, c] - ref["C"]) / ref["C"] < 1e-9, (i, c)
            assert rel(dcs[i][st.design_nodes, 2, c], ref["dcrds"][st.design_nodes, 2]) < 1e-8, (i, c)
