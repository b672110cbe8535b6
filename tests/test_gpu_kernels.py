"""GPU parity of the HIP kernels against the oracle, through the C ABI.

Element stiffness, assembly (bit-exact pattern, values to rounding), the
three solver paths and the adjoint kernel, on seeded inputs small enough for
the oracle to finish in seconds.
"""
import numpy as np
import pytest
import torch

from jaxsso_amd import workloads as wl
from jaxsso_amd import _native as nat
from jaxsso_amd.plan import FEPlan, blocks_to_csr
from oracle import elements as el
from oracle import gradient as og
from oracle import solve as osolve
from oracle import system

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def random_quads(n, seed, distort=0.25):
    """Quads: a square of random size, corners perturbed in-plane and out of plane, rotated."""
    rng = np.random.default_rng(seed)
    base = np.array([[0.5, 0.5, 0], [-0.5, 0.5, 0], [-0.5, -0.5, 0], [0.5, -0.5, 0.0]])
    base = base[[2, 3, 0, 1]]  # [BL, BR, TR, TL]
    out = np.zeros((n, 12))
    for e in range(n):
        h = rng.uniform(0.1, 3.0)
        X = base * h + rng.uniform(-distort, distort, size=(4, 3)) * h * np.array([1, 1, 0.3])
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        out[e] = (X @ q.T + rng.normal(size=3) * 5).ravel()
    props = np.stack([rng.uniform(0.05, 0.3, n), rng.uniform(1e9, 5e10, n), rng.uniform(0.0, 0.45, n),
                      rng.uniform(0.5, 1.5, n), rng.uniform(0.5, 1.5, n)], 1)
    return out, props


def random_beams(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, 3)) * 3
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = a + d * rng.uniform(0.2, 5.0, size=(n, 1))
    props = np.stack([rng.uniform(1e9, 3e11, n), rng.uniform(1e9, 1e11, n), rng.uniform(1e-6, 1e-4, n),
                      rng.uniform(1e-6, 1e-4, n), rng.uniform(1e-6, 1e-4, n), rng.uniform(1e-3, 1e-2, n)], 1)
    return np.concatenate([a, b], 1), props


def gpu_ke_quad(crds, props):
    L = nat.lib()
    c = torch.tensor(crds, dtype=torch.float64, device="cuda")
    p = torch.tensor(props, dtype=torch.float64, device="cuda")
    ke = torch.empty((len(crds), 24, 24), dtype=torch.float64, device="cuda")
    nat.check(L.jsso_element_ke_quad(len(crds), nat.ptr(c), nat.ptr(p), nat.ptr(ke), nat.stream_ptr()))
    return ke.cpu().numpy()


def gpu_ke_beam(crds, props):
    L = nat.lib()
    c = torch.tensor(crds, dtype=torch.float64, device="cuda")
    p = torch.tensor(props, dtype=torch.float64, device="cuda")
    ke = torch.empty((len(crds), 12, 12), dtype=torch.float64, device="cuda")
    fl = torch.empty(len(crds), dtype=torch.int32, device="cuda")
    nat.check(L.jsso_element_ke_beam(len(crds), nat.ptr(c), nat.ptr(p), nat.ptr(ke), nat.ptr(fl), nat.stream_ptr()))
    return ke.cpu().numpy(), fl.cpu().numpy()


def test_quad_ke_matches_oracle():
    crds, props = random_quads(500, 1)
    st = wl.cfg1()
    c1 = el.gather_crds(st.crds, st.cnct_quads[:300])
    crds = np.concatenate([crds, c1])
    props = np.concatenate([props, st.prop_quads[:300]])
    K = gpu_ke_quad(crds, props)
    R = el.quad_Ke(crds, props)
    err = np.abs(K - R).max(axis=(1, 2)) / np.abs(R).max(axis=(1, 2))
    assert err.max() < 1e-12, err.max()


def test_beam_ke_matches_oracle_including_y_parallel_branch():
    crds, props = random_beams(400, 2)
    crds[:5, 3] = crds[:5, 0]       # parallel to global Y: x and z equal
    crds[:5, 5] = crds[:5, 2]
    K, fl = gpu_ke_beam(crds, props)
    R = el.beam_Ke(crds, props)
    _, yflag = el.beam_dircos(crds)
    assert np.array_equal(fl.astype(bool), yflag) and yflag[:5].all()
    err = np.abs(K - R).max(axis=(1, 2)) / np.abs(R).max(axis=(1, 2))
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("case", ["cfg0", "cfg1", "barrel", "arch"])
def test_assembly_matches_oracle(case):
    st = getattr(wl, case)() if case.startswith("cfg") else getattr(wl, "notebook_" + case)()
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    Kv = plan.assemble(torch.tensor(st.crds), torch.tensor(st.prop_beamcols), torch.tensor(st.prop_quads),
                       export_K=True)
    A = plan.K_scalar_csr(Kv)
    R = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
    # y = K x through the C ABI
    x = np.random.default_rng(0).normal(size=st.ndof)
    y = plan.spmv(torch.tensor(x)).cpu().numpy()
    assert rel(y, R @ x) < 1e-13


@pytest.mark.parametrize("precision,pc", [("fp64", "block_jacobi"), ("mixed", "block_jacobi"), ("plain", None),
                                          ("fp64", "amg"), ("mixed", "amg")])
@pytest.mark.parametrize("case", ["cfg0", "cfg1", "barrel", "arch"])
def test_solve_matches_oracle(case, precision, pc):
    st = getattr(wl, case)() if case.startswith("cfg") else getattr(wl, "notebook_" + case)()
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    if pc:
        plan.set_preconditioner(pc)
    plan.assemble(torch.tensor(st.crds), torch.tensor(st.prop_beamcols), torch.tensor(st.prop_quads))
    # the slender arch amplifies the residual ~150x into u: a 1e-8 error bar
    # needs a 1e-11 scaled residual there
    tol = 1e-11 if case == "arch" else 1e-10
    u = plan.solve(torch.tensor(st.nodal_loads), precision=precision, tol=tol).cpu().numpy()
    K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
    ur = osolve.solve_reduced(K, st.known_id, st.nodal_loads)
    info = plan.last_info
    assert info["converged"] == info["n_systems"] == 1, info
    assert rel(u, ur) < 1e-8, (rel(u, ur), info)
    C, Cr = 0.5 * st.nodal_loads @ u, 0.5 * st.nodal_loads @ ur
    assert abs(C - Cr) / abs(Cr) < 1e-10
    assert np.all(u[st.known_id] == 0.0)


@pytest.mark.parametrize("case", ["cfg0", "cfg1", "barrel"])
def test_adjoint_kernel_matches_complex_step(case):
    """Feed the oracle's u to the HIP adjoint: per-node and per-element gradients
    match the oracle's complex-step element derivatives."""
    st = getattr(wl, case)() if case.startswith("cfg") else getattr(wl, "notebook_" + case)()
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
    u = osolve.solve_reduced(K, st.known_id, st.nodal_loads)
    eg = og.element_gradients(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads,
                              u, 0.5 * u)
    gref = og.node_gradient(st.n_node, st.cnct_beamcols, st.cnct_quads, eg)
    dc, db, dq = plan.adjoint(torch.tensor(st.crds), torch.tensor(st.prop_beamcols), torch.tensor(st.prop_quads),
                              torch.tensor(u), lam=None, lam_scale=0.5)
    scale = np.abs(gref).max()
    assert np.abs(dc.cpu().numpy() - gref).max() <= 1e-10 * scale
    if st.n_quad:
        r = eg["quad_props"]
        assert np.abs(dq.cpu().numpy() - r).max() <= 1e-10 * np.abs(r).max(axis=0).clip(1e-300).max()
    if st.n_beamcol:
        r = eg["beam_props"]
        g = db.cpu().numpy()
        for k in range(6):
            assert np.abs(g[:, k] - r[:, k]).max() <= 1e-10 * np.abs(r[:, k]).max()


def test_adjoint_general_lambda():
    """A general adjoint vector (not u/2) through the same kernel."""
    st = wl.notebook_barrel()
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    rng = np.random.default_rng(5)
    u = rng.normal(size=st.ndof)
    lam = rng.normal(size=st.ndof)
    eg = og.element_gradients(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads, u, lam)
    gref = og.node_gradient(st.n_node, st.cnct_beamcols, st.cnct_quads, eg)
    dc, _, dq = plan.adjoint(torch.tensor(st.crds), None, torch.tensor(st.prop_quads), torch.tensor(u),
                             lam=torch.tensor(lam))
    assert np.abs(dc.cpu().numpy() - gref).max() <= 1e-9 * np.abs(gref).max()
    assert np.abs(dq.cpu().numpy() - eg["quad_props"]).max() <= 1e-9 * np.abs(eg["quad_props"]).max()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("nrhs", [1, 2, 4])
def test_cg_kernel_variants_agree_with_oracle(precision, nrhs):
    """Every waves-per-slice variant of the batched CG kernel (forced through the
    tuning entry point) solves a batch of two uneven designs to the oracle."""
    designs = [wl.cfg1(0), wl.gridshell(9, 10.0, 2.0, 0.05, cfg=3, design=5, half_span_cases=True)]
    from jaxsso_amd.batch import StructureBatch
    b = StructureBatch(designs, n_cases=nrhs)
    b.plan.set_preconditioner("block_jacobi")
    refs = []
    for i, st in enumerate(designs):
        K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
        cases = st.load_cases()
        refs.append([osolve.solve_reduced(K, st.known_id, cases[c]) if c < cases.shape[0] else np.zeros(st.ndof)
                     for c in range(nrhs)])
    sols = {}
    for wps in (1, 2, 4):
        b.plan.cg_tuning(precision, nrhs, set_wps=wps)
        u = b.solve(precision=precision, tol=1e-10).cpu().numpy().reshape(-1, nrhs)
        info = b.plan.last_info
        assert info["converged"] == info["n_systems"] == 2 * nrhs, (wps, info)
        assert info["waves_per_slice"] in (1, 2, 4)
        for i, st in enumerate(designs):
            o = 6 * int(b.node_off[i])
            for c in range(nrhs):
                if np.any(refs[i][c]):
                    assert rel(u[o:o + st.ndof, c], refs[i][c]) < 1e-8, (wps, i, c)
        sols[wps] = u
    # the variants differ only in summation order
    assert rel(sols[2], sols[1]) < 1e-9 and rel(sols[4], sols[1]) < 1e-9


def test_cg_autotuner_records_a_choice():
    st = wl.cfg1()
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    plan.set_preconditioner("block_jacobi")
    plan.assemble(torch.tensor(st.crds), None, torch.tensor(st.prop_quads))
    assert plan.cg_tuning("mixed", 1) == 0
    u1 = plan.solve(torch.tensor(st.nodal_loads), precision="mixed").cpu().numpy()
    chosen = plan.cg_tuning("mixed", 1)
    assert chosen in (1, 2, 4) and plan.last_info["waves_per_slice"] == chosen
    # a second solve reuses the choice and repeats the first bit for bit
    u2 = plan.solve(torch.tensor(st.nodal_loads), precision="mixed").cpu().numpy()
    assert np.array_equal(u1, u2)
    assert plan.cg_tuning("fp64", 1) == 0


def test_compliance_reduction_matches_per_structure_dot_products():
    """jsso_compliance: scale * f.u per structure and load case, in a fixed order
    (two calls agree bit for bit)."""
    designs = [wl.cfg1(0), wl.gridshell(9, cfg=3, design=2, half_span_cases=True), wl.cfg0()]
    from jaxsso_amd.batch import StructureBatch
    b = StructureBatch(designs, n_cases=2)
    rng = np.random.default_rng(3)
    u = torch.tensor(rng.normal(size=b.f.shape), device=b.f.device)
    C1 = b.plan.compliance(b.f, u, 0.5).cpu().numpy()
    C2 = b.plan.compliance(b.f, u, 0.5).cpu().numpy()
    assert np.array_equal(C1, C2) and C1.shape == (3, 2)
    F, U = b.f.cpu().numpy(), u.cpu().numpy()
    for i in range(3):
        o0, o1 = 6 * int(b.node_off[i]), 6 * int(b.node_off[i + 1])
        ref = 0.5 * np.einsum("ij,ij->j", F[o0:o1], U[o0:o1])
        assert np.allclose(C1[i], ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_graph_and_eager_cg_agree_bitwise(precision, monkeypatch):
    """The CG chunks run as captured HIP graphs or as eager launches: the same
    kernels in the same order, so the same bits."""
    designs = [wl.cfg1(d) for d in range(3)]
    from jaxsso_amd.batch import StructureBatch
    b = StructureBatch(designs)
    b.plan.set_preconditioner("block_jacobi")
    b.plan.cg_tuning(precision, 1, set_wps=1)
    monkeypatch.setenv("JSSO_CG_GRAPH", "1")
    ug = b.solve(precision=precision, tol=1e-10).cpu().numpy()
    ig = dict(b.plan.last_info)
    monkeypatch.setenv("JSSO_CG_GRAPH", "0")
    ue = b.solve(precision=precision, tol=1e-10).cpu().numpy()
    assert ig["iterations"] == b.plan.last_info["iterations"] and ig["converged"] == 3
    assert np.array_equal(ug, ue)
    # a third solve reuses the captured graph
    monkeypatch.setenv("JSSO_CG_GRAPH", "1")
    assert np.array_equal(b.solve(precision=precision, tol=1e-10).cpu().numpy(), ug)
    # another tolerance on the same plan is honoured (the graphs are re-captured)
    b.solve(precision=precision, tol=1e-5)
    loose = dict(b.plan.last_info)
    assert loose["converged"] == 3 and loose["rel_residual"] <= 1e-5 and loose["iterations"] < ig["iterations"]
    assert np.array_equal(b.solve(precision=precision, tol=1e-10).cpu().numpy(), ug)
