"""Smoothed-aggregation AMG preconditioned CG (JSSO_PC_AMG) against the oracle,
through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from jaxsso_amd import workloads as wl
from jaxsso_amd.batch import StructureBatch
from jaxsso_amd.plan import FEPlan
from oracle import gradient as og
from oracle import solve as osolve
from oracle import system

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_u(st, f=None):
    K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
    return osolve.solve_reduced(K, st.known_id, st.nodal_loads if f is None else f)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("case", ["cfg0", "cfg1", "barrel", "arch"])
def test_amg_solve_matches_oracle(case, precision):
    st = getattr(wl, case)() if case.startswith("cfg") else getattr(wl, "notebook_" + case)()
    print(f"amg solve {case} {precision}", flush=True)
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    plan.set_preconditioner("amg")
    plan.assemble(torch.tensor(st.crds), torch.tensor(st.prop_beamcols), torch.tensor(st.prop_quads))
    tol = 1e-11 if case == "arch" else 1e-10
    u = plan.solve(torch.tensor(st.nodal_loads), precision=precision, tol=tol).cpu().numpy()
    info = plan.last_info
    ur = oracle_u(st)
    assert info["converged"] == info["n_systems"] == 1, info
    assert info["levels"] == len(plan.amg_levels()) >= 2, info
    assert rel(u, ur) < 1e-8, (rel(u, ur), info)
    C, Cr = 0.5 * st.nodal_loads @ u, 0.5 * st.nodal_loads @ ur
    assert abs(C - Cr) / abs(Cr) < 1e-10
    assert np.all(u[st.known_id] == 0.0)
    print(f"  iterations {info['iterations']}, levels {plan.amg_levels()}", flush=True)
    if case == "cfg1":
        # block-Jacobi CG needs ~2 900 (FP64) / ~3 700 (mixed) iterations here
        assert info["iterations"] <= 400, info


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("nrhs", [1, 2, 4])
def test_amg_batch_of_uneven_designs(precision, nrhs):
    designs = [wl.cfg1(0), wl.gridshell(9, 10.0, 2.0, 0.05, cfg=3, design=5, half_span_cases=True), wl.cfg0()]
    print(f"amg batch {precision} nrhs {nrhs}", flush=True)
    b = StructureBatch(designs, n_cases=nrhs)
    u = b.solve(precision=precision, tol=1e-10, preconditioner="amg").cpu().numpy().reshape(-1, nrhs)
    info = b.plan.last_info
    assert info["converged"] == info["n_systems"] == 3 * nrhs, info
    for i, st in enumerate(designs):
        cases = st.load_cases()
        o = 6 * int(b.node_off[i])
        for c in range(min(nrhs, cases.shape[0])):
            assert rel(u[o:o + st.ndof, c], oracle_u(st, cases[c])) < 1e-8, (i, c, info)


def test_amg_graph_and_eager_agree_bitwise(monkeypatch):
    designs = [wl.cfg1(d) for d in range(3)]
    b = StructureBatch(designs)
    monkeypatch.setenv("JSSO_CG_GRAPH", "1")
    ug = b.solve(precision="mixed", tol=1e-10, preconditioner="amg").cpu().numpy()
    ig = dict(b.plan.last_info)
    monkeypatch.setenv("JSSO_CG_GRAPH", "0")
    ue = b.solve(precision="mixed", tol=1e-10, preconditioner="amg").cpu().numpy()
    assert ig["converged"] == 3 and ig["iterations"] == b.plan.last_info["iterations"]
    assert np.array_equal(ug, ue)


@pytest.mark.parametrize("nrhs", [1, 4])
def test_amg_set_up_product_forms_agree_bitwise(monkeypatch, nrhs):
    """The AMG set-up products (P, W = A P, A_c = P^T W) come in a
    thread-per-block and a lane-per-entry form (chosen by the level's size); both
    add every entry's terms in the same order, so a hierarchy built with either
    form on every level (JSSO_AMG_ENTRY_PRODUCTS=0 / 1, read at each set-up)
    gives the same solution bit for bit.  Uneven shells with one and four load
    cases."""
    designs = [wl.cfg1(0), wl.cfg1(1), wl.shell_pinned_edges(60, 6.0, 1.0)]
    print(f"amg product forms nrhs {nrhs}", flush=True)
    out = {}
    for form in ("0", "1"):
        monkeypatch.setenv("JSSO_AMG_ENTRY_PRODUCTS", form)
        b = StructureBatch(designs, n_cases=nrhs)
        u = b.solve(precision="fp64", tol=1e-10, preconditioner="amg").cpu().numpy()
        out[form] = (u, dict(b.plan.last_info))
    assert out["0"][1]["converged"] == 3 * nrhs and out["0"][1]["iterations"] == out["1"][1]["iterations"]
    assert np.array_equal(out["0"][0], out["1"][0])


@pytest.mark.parametrize("maxit", [13, 17, 24])
def test_amg_stops_at_maxit_with_the_best_iterate(maxit):
    """An AMG solve cut off by maxit (a last chunk of five iterations, of one, or a
    full one) warns, reports exactly maxit iterations and no converged system, and
    returns its best iterate (smallest FP64 residual at a chunk end): a finite u
    whose relative error against the oracle lies between the converged solve's
    and the zero guess's.  (Before iteration 11 on these shells the FP64 residual
    has not yet dropped below the zero guess's, whose u = 0 is then the best.)"""
    designs = [wl.cfg1(0), wl.cfg1(1)]
    b = StructureBatch(designs)
    print(f"amg maxit {maxit}", flush=True)
    with pytest.warns(RuntimeWarning, match="did not converge"):
        u = b.solve(precision="fp64", tol=1e-14, maxit=maxit, preconditioner="amg").cpu().numpy()
    info = b.plan.last_info
    assert info["iterations"] == maxit and info["converged"] == 0, info
    assert np.all(np.isfinite(u))
    for i, st in enumerate(designs):
        o = 6 * int(b.node_off[i])
        e = rel(u[o:o + st.ndof], oracle_u(st))
        assert 1e-12 < e < 0.5, (i, e, info)


def test_amg_set_up_follows_the_assembly():
    """A new assembly on the same plan (other coordinates and thicknesses)
    rebuilds the numeric hierarchy before the next solve."""
    a, c = wl.cfg1(0), wl.cfg1(5)
    plan = FEPlan([0, a.n_node], a.cnct_beamcols, a.cnct_quads, a.known_id)
    plan.set_preconditioner("amg")
    for st in (a, c, a):
        plan.assemble(torch.tensor(st.crds), None, torch.tensor(st.prop_quads))
        u = plan.solve(torch.tensor(st.nodal_loads), precision="mixed", tol=1e-10).cpu().numpy()
        assert rel(u, oracle_u(st)) < 1e-8, plan.last_info


def test_amg_compliance_gradient_matches_oracle():
    st = wl.cfg1()
    b = StructureBatch([st])
    C, (dc, db, dq) = b.compliance_and_gradient(precision="mixed", tol=1e-10, preconditioner="amg")
    ref = og.compliance_and_gradient(st)
    assert abs(C.cpu().numpy()[0, 0] - ref["C"]) / ref["C"] < 1e-10
    assert rel(dc.cpu().numpy().reshape(ref["dcrds"].shape), ref["dcrds"]) < 1e-8
    assert rel(dq.cpu().numpy().reshape(ref["dprop_quad"].shape), ref["dprop_quad"]) < 1e-8


@pytest.mark.parametrize("nrhs", [1, 4])
def test_amg_unfused_coarse_levels_match_oracle(monkeypatch, nrhs):
    """JSSO_AMG_FUSE=0 (read when a plan builds its hierarchy) runs each
    structure's last sparse level with the per-level smoothing, restriction and
    prolongation kernels and the dense solve with k_mg_direct instead of the fused
    k_mg_tail: the same preconditioner up to FP32 rounding, so the same solution
    to the solver's tolerance and nearly the same iteration count.  Two cfg1 shells
    (three levels: the fused level is level 1) and a 60x60 pinned-edge shell (four
    levels: level 1 smoothed by the per-level kernels, level 2 fused)."""
    designs = [wl.cfg1(0), wl.cfg1(1), wl.shell_pinned_edges(60, 6.0, 1.0)]
    print(f"amg unfused nrhs {nrhs}", flush=True)
    its = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("JSSO_AMG_FUSE", fuse)
        b = StructureBatch(designs, n_cases=nrhs)
        u = b.solve(precision="mixed", tol=1e-10, preconditioner="amg").cpu().numpy().reshape(-1, nrhs)
        info = b.plan.last_info
        assert info["converged"] == info["n_systems"] == 3 * nrhs, (fuse, info)
        its[fuse] = info["iterations"]
        for i, st in enumerate(designs):
            o = 6 * int(b.node_off[i])
            cases = st.load_cases()
            for c in range(min(nrhs, cases.shape[0])):
                assert rel(u[o:o + st.ndof, c], oracle_u(st, cases[c])) < 1e-8, (fuse, i, c, info)
    assert abs(its["1"] - its["0"]) <= 8, its


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_amg_level0_wave_variants_agree(precision):
    """Every waves-per-slice variant of the AMG level-0 kernels (1, 2, 4) solves a
    batch of uneven shells to the oracle; the variants differ only in summation
    order (1e-9); an untuned plan times them on the live problem, records its
    choice, reports it in the solve info and repeats itself bit for bit."""
    designs = [wl.cfg1(0), wl.shell_pinned_edges(12, 8.0, 1.0, design=2)]
    b = StructureBatch(designs)
    assert b.plan.cg_tuning(precision, 1) == 0
    u0 = b.solve(precision=precision, tol=1e-10).cpu().numpy()
    chosen = b.plan.cg_tuning(precision, 1)
    assert chosen in (1, 2, 4) and b.plan.last_info["waves_per_slice"] == chosen
    u0b = b.solve(precision=precision, tol=1e-10).cpu().numpy()
    assert np.array_equal(u0, u0b)
    refs = []
    for st in designs:
        K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
        refs.append(osolve.solve_reduced(K, st.known_id, st.nodal_loads))
    sols = {}
    for wps in (1, 2, 4):
        b.plan.cg_tuning(precision, 1, set_wps=wps)
        u = b.solve(precision=precision, tol=1e-10).cpu().numpy()
        assert b.plan.last_info["converged"] == 2
        for i, st in enumerate(designs):
            o = 6 * int(b.node_off[i])
            assert rel(u[o:o + st.ndof], refs[i]) < 1e-8, (wps, i)
        sols[wps] = u
    assert rel(sols[2], sols[1]) < 1e-9 and rel(sols[4], sols[1]) < 1e-9


def test_element_kernel_tile_variants_agree():
    """The assembly's MITC-4 kernel at 4, 8 and 16 elements per workgroup writes
    the same K bit for bit; the first assembly records a timed choice."""
    st = wl.cfg1(3)
    from jaxsso_amd.plan import FEPlan
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    assert plan.element_tuning() == 0
    crds, pq = torch.tensor(st.crds), torch.tensor(st.prop_quads)
    K0 = plan.assemble(crds, None, pq, export_K=True).clone()
    assert plan.element_tuning() in (4, 8, 16)
    for epw in (4, 8, 16):
        plan.element_tuning(epw)
        K = plan.assemble(crds, None, pq, export_K=True)
        assert torch.equal(K, K0), epw


def test_amg_hierarchy_kept_across_assemblies():
    """jsso_plan_amg_reuse: a hierarchy built for one thickness field serves the
    next designs (only the level-0 smoothing bound is refreshed): the solution
    still matches the oracle, the counts show one full set-up and the refreshes,
    and a forced rebuild (mode 2) runs one more full set-up."""
    st = wl.cfg1(0)
    b = StructureBatch([st])
    info0 = None
    b.plan.amg_reuse(2)
    rng = np.random.default_rng(5)
    t0 = b.prop_quad[:, 0].clone()
    for k in range(3):
        if k:
            b.prop_quad[:, 0] = t0 * torch.tensor(rng.uniform(0.9, 1.1, st.n_quad), device="cuda")
        u = b.solve(precision="fp64", tol=1e-10).cpu().numpy()
        info0 = info0 or dict(b.plan.last_info)
        pq = b.prop_quad.cpu().numpy()
        K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, pq)
        ref = osolve.solve_reduced(K, st.known_id, st.nodal_loads)
        assert rel(u, ref) < 1e-8, k
        assert b.plan.last_info["converged"] == 1
    c = b.plan.amg_reuse(-1)
    assert c["mode"] == 1 and c["full_setups"] == 1 and c["refreshes"] == 2, c
    b.plan.amg_reuse(2)
    b.solve(precision="fp64", tol=1e-10)
    assert b.plan.amg_reuse(-1)["full_setups"] == 2
    b.plan.amg_reuse(0)
