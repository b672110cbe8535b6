"""Host-side logic without a GPU: the C-ABI library loads and exports every
symbol of include/jsso.h, the symbolic block pattern equals the reference's
scalar pattern, plan validation, and the Model builder."""
import ctypes
import os
import re

import numpy as np
import pytest

import jaxsso_amd.model as Model
from jaxsso_amd import _native as nat
from jaxsso_amd import workloads as wl
from jaxsso_amd.plan import FEPlan, blocks_to_csr
from oracle import system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = nat.load_library()
    header = open(os.path.join(ROOT, "include", "jsso.h")).read()
    declared = set(re.findall(r"^\s*int\s+(jsso_\w+)\s*\(", header, flags=re.M))
    assert declared == set(nat.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.jsso_version() == 1


def _pattern_equal(st, plan=None):
    plan = plan or FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    rp, ci = plan.pattern()
    A = blocks_to_csr(rp, ci, np.ones((len(ci), 6, 6)))
    P = system.pattern_K(st.n_node, st.cnct_beamcols, st.cnct_quads)
    return np.array_equal(A.indptr, P.indptr) and np.array_equal(A.indices, P.indices)


@pytest.mark.parametrize("name", ["cfg0", "cfg1", "notebook_arch", "notebook_barrel"])
def test_block_pattern_is_reference_pattern(name):
    assert _pattern_equal(getattr(wl, name)())


def test_block_pattern_mixed_elements_and_irregular_connectivity():
    rng = np.random.default_rng(3)
    st = wl.shell_pinned_edges(5, 5.0, 1.0)
    perm = rng.permutation(st.n_quad)
    st.cnct_quads = np.roll(st.cnct_quads[perm], 1, axis=1).astype(np.int32)   # element order, node order
    st.prop_quads = st.prop_quads[perm]
    st.cnct_beamcols = np.array([[0, 35], [7, 7 + 6], [20, 3]], np.int32)       # beams across the shell
    st.prop_beamcols = np.tile([2e11, 8e10, 1e-6, 1e-6, 2e-6, 1e-3], (3, 1))
    assert _pattern_equal(st)


def test_batched_plan_is_block_diagonal():
    a, b = wl.gridshell(3), wl.notebook_barrel()
    n = a.n_node
    plan = FEPlan([0, n, n + b.n_node], a.cnct_beamcols, b.cnct_quads + n,
                  np.concatenate([a.known_id, b.known_id + 6 * n]))
    rp, ci = plan.pattern()
    pa = FEPlan([0, n], a.cnct_beamcols, np.zeros((0, 4)), a.known_id).pattern()
    pb = FEPlan([0, b.n_node], np.zeros((0, 2)), b.cnct_quads, b.known_id).pattern()
    assert np.array_equal(rp[:n + 1], pa[0])
    assert np.array_equal(ci[:rp[n]], pa[1])
    assert np.array_equal(rp[n:] - rp[n], pb[0])
    assert np.array_equal(ci[rp[n]:] - n, pb[1])


@pytest.mark.parametrize("kwargs,msg", [
    (dict(known_id=[3, 3]), "listed twice"),
    (dict(known_id=[10 ** 6]), "out of range"),
    (dict(cnct_beam=[[0, 99]]), "out of range"),
    (dict(node_off=[0, 2, 4], cnct_beam=[[1, 2]]), "spans two structures"),
    (dict(node_off=[0, 0, 4]), "no nodes"),
])
def test_plan_validation(kwargs, msg):
    args = dict(node_off=[0, 4], cnct_beam=[[0, 1], [1, 2], [2, 3]], cnct_quad=np.zeros((0, 4)), known_id=[0, 1, 2])
    args.update(kwargs)
    with pytest.raises(nat.JssoError, match=msg):
        FEPlan(args["node_off"], args["cnct_beam"], args["cnct_quad"], args["known_id"])


def test_model_builder_matches_workload_arrays():
    st = wl.notebook_barrel()
    m = Model.Model()
    for i, (x, y, z) in enumerate(st.crds):
        m.add_node(i, x, y, z)
        if i not in set(st.design_nodes.tolist()):
            m.add_support(i, [1, 1, 1, 0, 0, 0])
        else:
            m.add_nodal_load(i, [0, 0, st.nodal_loads[6 * i + 2], 0, 0, 0])
    for e, c in enumerate(st.cnct_quads):
        m.add_quad(e, *[int(v) for v in c], *st.prop_quads[e, :3])
    m.add_truss(0, 0, 1, 1.0, 1.0)      # accepted and ignored, as the reference
    m.model_ready()
    assert np.array_equal(m.crds, st.crds)
    assert np.array_equal(m.nodal_loads, st.nodal_loads)
    assert np.array_equal(np.sort(m.known_id), np.sort(st.known_id))
    assert np.array_equal(m.cnct_quads, st.cnct_quads)
    assert np.array_equal(m.prop_quads, st.prop_quads)
    assert m.ndof == st.ndof and m.n_quad == st.n_quad and m.n_beamcol == 0
    assert len(m.unknown_id) + len(m.known_id) == m.ndof
    # the plan is built from the topology alone, on the host
    assert _pattern_equal(st, m.plan())


def test_model_rejects_out_of_order_tags():
    m = Model.Model()
    m.add_node(1, 0.0, 0.0, 0.0)
    m.add_node(0, 1.0, 0.0, 0.0)
    with pytest.raises(ValueError, match="insertion order"):
        m.model_ready()


def test_compute_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU visible")
    st = wl.cfg0()
    plan = FEPlan([0, st.n_node], st.cnct_beamcols, st.cnct_quads, st.known_id)
    with pytest.raises(nat.NativeUnavailable):
        plan.assemble(st.crds, st.prop_beamcols, st.prop_quads)


def test_quad_area_matches_reference_formula():
    from jaxsso_amd.element import Quad
    c = np.array([0, 0, 0, 2, 0, 0, 2, 3, 0, 0, 3, 0.0])
    assert abs(Quad.A(c) - 6.0) < 1e-14


def test_amg_hierarchy_is_per_structure_and_coarsens():
    """The symbolic AMG hierarchy (host only): every level is smaller than the one
    above, a batched plan's levels are the sums of its structures' levels (the
    aggregation never crosses structures), and the preconditioner setting is
    validated."""
    def plan_of(structs):
        off = np.cumsum([0] + [s.n_node for s in structs])
        cq = np.concatenate([s.cnct_quads + o for s, o in zip(structs, off)])
        kn = np.concatenate([s.known_id + 6 * o for s, o in zip(structs, off)])
        return FEPlan(off, np.zeros((0, 2)), cq, kn)

    single = [plan_of([wl.cfg1(d)]).amg_levels() for d in range(3)]
    assert single[0] == single[1] == single[2]
    lv = single[0]
    assert lv[0] == 1681 and all(a > b for a, b in zip(lv, lv[1:])) and lv[-1] <= 26
    batch = plan_of([wl.cfg1(d) for d in range(3)]).amg_levels()
    assert batch == [3 * v for v in lv]
    gs = wl.gridshell(24)
    g = FEPlan([0, gs.n_node], gs.cnct_beamcols, np.zeros((0, 4)), gs.known_id).amg_levels()
    assert g[0] == 625 and g[-1] <= 26 and all(a > b for a, b in zip(g, g[1:]))
    p = plan_of([wl.cfg1()])
    with pytest.raises(ValueError):
        p.set_preconditioner("ilu")
    p.set_preconditioner("amg")
    assert p.preconditioner == "amg"


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says so,
    before anything touches a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_trace_summary_averages_only_full_activity_launches(tmp_path, capsys):
    """bench.py --trace-summary: of the roofline kernel's launches in a kernel
    trace it drops the live tuner's (before the first solve) and averages, per
    solve, only the first full_iterations of its dir_launches.  Synthetic trace:
    3 tuner launches of 999 us, two solves of 6 launches each, the first 4 of
    each (every design running) 200 and 210 us, the rest 50 us, other kernels in
    between."""
    import json
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    kname = "k_pcg_dir<double, double, 1, 4>"
    rows, t = [], 1_000_000
    def add(name, us):
        nonlocal t
        rows.append((t, t + int(us * 1000), name))
        t += int(us * 1000) + 5000
    for _ in range(3):
        add(f"void jsso::{kname}(jsso::Slices)", 999.0)
    for solve in range(2):
        for k in range(6):
            add(f"void jsso::{kname}(jsso::Slices)", (200.0 if solve == 0 else 210.0) if k < 4 else 50.0)
            add("_ZN4jsso9k_pcg_preIdLi1ELb0ELi4EE", 120.0)
    d = tmp_path / "trace" / "host"
    d.mkdir(parents=True)
    with open(d / "run_kernel_trace.csv", "w") as fh:
        fh.write("Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for s0, s1, name in rows:
            fh.write(f"\"{name}\",{s0},{s1}\n")
    log = tmp_path / "bench.log"
    alg = 897189888
    log.write_text("# header\n" + json.dumps({"roofline": {"kernel": kname, "alg_bytes_per_launch": alg,
                                                           "avg_launch_us": 205.0},
                                              "solve_log": {"solves": [[6, 6, 4], [6, 6, 4]]}}) + "\n")
    bench.trace_summary(str(tmp_path / "trace"), str(log))
    out = capsys.readouterr().out
    line = [l for l in out.splitlines() if l.startswith("roofline_launches")][0]
    assert f"{kname}: 8 launches with every design running" in line
    assert "avg 205.00 us, min 200.00, max 210.00" in line
    assert f"-> {alg / 205e-6 / 1e9:.1f} GB/s" in line
    assert "the first 3: the tuner" in out


def test_traffic_summary_scales_by_the_copy_and_keeps_full_activity(tmp_path, capsys):
    """bench.py --traffic-summary: each counter (KiB per dispatch) is scaled by
    what it reads for the run's 1 GiB device copy, and only the roofline kernel's
    full-activity launches are averaged (tuner launches dropped).  Synthetic PMC
    output: FETCH_SIZE reads the copy as 512 MiB (factor 2), WRITE_SIZE as 1 GiB
    (factor 1); 2 tuner launches, then one solve of 4 launches, 2 of them with
    every design running."""
    import json
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    kname = "k_pcg_dir<double, double, 1, 4>"
    full = {"FETCH_SIZE": [400000.0, 600000.0], "WRITE_SIZE": [30000.0, 34000.0]}
    copy = {"FETCH_SIZE": 524288.0, "WRITE_SIZE": 1048576.0}
    for counter, sub in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
        d = tmp_path / sub / "host"
        d.mkdir(parents=True)
        lines = ["Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value"]
        did = 0
        def row(name, v):
            nonlocal did
            did += 1
            lines.append(f"{did},\"{name}\",{counter},{v}")
        row("__amd_rocclr_copyBuffer", copy[counter])
        for v in (9e6, 9e6) + tuple(full[counter]) + (10.0, 10.0):
            row(f"void jsso::{kname}(jsso::Slices)", v)
            row("void jsso::ell_assemble_kernel(long)", 1000.0)
        (d / "run_counter_collection.csv").write_text("\n".join(lines) + "\n")
    log = tmp_path / "bench.log"
    alg = 897189888
    log.write_text(json.dumps({"roofline": {"kernel": kname, "alg_bytes_per_launch": alg},
                               "solve_log": {"solves": [[4, 4, 2]]}}) + "\n")
    bench.traffic_summary(str(tmp_path / "fetch"), str(tmp_path / "write"), str(log))
    out = capsys.readouterr().out
    fetch = 500000.0 * 1024 * 2.0
    write = 32000.0 * 1024 * 1.0
    assert f"fetch_bytes_per_launch {kname}: {int(fetch)}" in out
    assert f"write_bytes_per_launch {kname}: {int(write)}" in out
    assert f"traffic / algorithmic bytes = {(fetch + write) / alg:.3f} over 2 launches" in out
    assert f"assembly_traffic ell_assemble_kernel: fetch {int(1000.0 * 1024 * 2)} write {int(1000.0 * 1024)}" in out
