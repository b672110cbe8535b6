"""The 200x200 MITC-4 shell (40,401 nodes, 40,000 quads, 242,406 DOF) on the
GPU: compliance and the full shape + thickness gradient against the FP64
oracle's fixture (tests/golden/make_shell200_fixture.py), and invariants
checked at full size."""
import json
import os

import numpy as np
import pytest

from jaxsso_amd import workloads as wl
from jaxsso_amd.batch import StructureBatch

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "shell200_oracle.json")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def shell200():
    with open(FIXTURE) as fh:
        fx = json.load(fh)
    st = wl.cfg2()
    return st, StructureBatch([st]), fx


@pytest.mark.parametrize("precision,pc", [("mixed", "amg"), ("fp64", "amg"), ("mixed", "block_jacobi"),
                                          ("fp64", "block_jacobi"), ("plain", None)])
def test_shell200_compliance_and_gradient_vs_oracle(shell200, precision, pc):
    st, b, fx = shell200
    C, (dc, _, dq) = b.compliance_and_gradient(precision=precision, tol=1e-10, preconditioner=pc)
    info = b.plan.last_info
    assert info["converged"] == info["n_systems"] == 1, info
    C = float(C[0, 0])
    u = b.u.cpu().numpy().ravel()
    dc = dc.cpu().numpy().reshape(st.n_node, 3)
    dq = dq.cpu().numpy().reshape(st.n_quad, 5)
    assert abs(C - fx["C"]) / fx["C"] < 1e-9
    assert rel(u[fx["u_index"]], fx["u"]) < 1e-8
    assert abs(np.linalg.norm(u) - fx["norm_u"]) / fx["norm_u"] < 1e-8
    dz = dc[:, 2]
    assert rel(dz[fx["dz_node"]], fx["dz"]) < 1e-8
    assert rel(dq[fx["dt_elem"], 0], fx["dt"]) < 1e-8
    assert abs(np.linalg.norm(dz[st.design_nodes]) - fx["norm_dz_free"]) / fx["norm_dz_free"] < 1e-8
    assert abs(np.linalg.norm(dq[:, 0]) - fx["norm_dt"]) / fx["norm_dt"] < 1e-8
    # a rigid translation of every node leaves every element matrix unchanged,
    # so the coordinate gradient sums to zero in each direction (for any u)
    for a in range(3):
        assert abs(dc[:, a].sum()) <= 1e-10 * np.abs(dc[:, a]).sum(), a
    # K is linear in E: sum_e E_e dC/dE_e = -1/2 u.K u = -C at the solution
    assert abs(np.sum(st.prop_quads[:, 1] * dq[:, 1]) + C) / C < 1e-9
