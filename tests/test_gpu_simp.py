"""SIMP shell designs (100x100 MITC-4, 61,206 DOF each, E = rho^3 * 30 GPa) as
one batch on the GPU: compliance, shape gradient and density gradient (chain
rule through E) against the FP64 oracle's fixture
(tests/golden/make_simp_fixture.py), and invariants for every design."""
import json
import os

import numpy as np
import pytest

from jaxsso_amd import workloads as wl
from jaxsso_amd.batch import StructureBatch

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "simp100_oracle.json")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("pc", ["amg", "block_jacobi"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_simp_batch_vs_oracle(precision, pc):
    with open(FIXTURE) as fh:
        fx = json.load(fh)
    designs = [wl.simp_shell(d) for d in range(4)]
    b = StructureBatch(designs)
    C, (dc, _, dq) = b.compliance_and_gradient(precision=precision, tol=1e-10, preconditioner=pc)
    info = b.plan.last_info
    assert info["converged"] == info["n_systems"] == 4, info
    C = C.cpu().numpy()[:, 0]
    u = b.u.cpu().numpy().ravel()
    dcs = b.split_nodes(dc.cpu().numpy().reshape(-1, 3))
    dqn = dq.cpu().numpy().reshape(-1, 5)
    for d, st in enumerate(designs):
        o = int(b.node_off[d])
        dqd = dqn[b.quad_off[d]:b.quad_off[d + 1]]
        drho = dqd[:, 1] * 3.0 * st.rho ** 2 * 30e9
        # K is linear in E (sum_e E_e dC/dE_e = -C) and a rigid translation of all
        # nodes changes no element
        assert abs(np.sum(st.prop_quads[:, 1] * dqd[:, 1]) + C[d]) / C[d] < 1e-9
        for a in range(3):
            assert abs(dcs[d][:, a].sum()) <= 1e-10 * np.abs(dcs[d][:, a]).sum()
        ref = [r for r in fx["designs"] if r["design"] == d]
        if not ref:
            continue
        r = ref[0]
        assert abs(C[d] - r["C"]) / r["C"] < 1e-9
        ud = u[6 * o:6 * o + st.ndof]
        assert rel(ud[r["u_index"]], r["u"]) < 1e-8
        assert rel(dcs[d][r["dz_node"], 2], r["dz"]) < 1e-8
        assert rel(drho[r["drho_elem"]], r["drho"]) < 1e-8
        assert abs(np.linalg.norm(drho) - r["norm_drho"]) / r["norm_drho"] < 1e-8
        assert abs(np.linalg.norm(dcs[d][st.design_nodes, 2]) - r["norm_dz_free"]) / r["norm_dz_free"] < 1e-8


@pytest.mark.parametrize("pc,precision", [("amg", "fp64"), ("amg", "mixed"), ("block_jacobi", "fp64")])
def test_slice_maps_give_the_same_bits(monkeypatch, pc, precision):
    """The solver kernels reach the XCDs in runs of slices: runs of up to 16 in a
    batch with a structure of 64 slices or more (each structure then spreads over
    all eight XCDs), else one contiguous eighth of the slices per XCD.  Every slice
    forms the same sums under either map, so JSSO_SLICE_SPREAD=0 / 1 (read when
    the plan is built) give the same solution bit for bit.  Four SIMP designs of
    160 slices each, the level-0 variant and the element kernel pinned."""
    designs = [wl.simp_shell(d) for d in range(4)]
    print(f"slice maps {pc} {precision}", flush=True)
    out = {}
    for sp in ("0", "1"):
        monkeypatch.setenv("JSSO_SLICE_SPREAD", sp)
        b = StructureBatch(designs)
        b.plan.set_preconditioner(pc)
        b.plan.cg_tuning(precision, 1, 4)
        b.plan.element_tuning(4)
        u = b.solve(precision=precision, tol=1e-10, preconditioner=pc).cpu().numpy()
        out[sp] = (u, dict(b.plan.last_info))
    assert out["0"][1]["converged"] == out["0"][1]["n_systems"] == 4, out["0"][1]
    assert out["0"][1]["iterations"] == out["1"][1]["iterations"]
    assert np.array_equal(out["0"][0], out["1"][0])
