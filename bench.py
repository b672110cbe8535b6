#!/usr/bin/env python3
"""Benchmark: compliance + full shape/thickness gradient evaluations per second.

Workload: the 40x40 MITC-4 shell (1,681 nodes, 1,600 quads, 10,086 DOF; four
fully fixed corners, -2 kN/node, t = 0.10 U(0.9, 1.1) per element), as a batch
of B seeded thickness designs (design 0 is the reference configuration).  One
step = assemble K for every design, build the numeric AMG hierarchy, solve
(flexible CG preconditioned by the FP32 smoothed-aggregation V-cycle, FP32
recurrences with FP64 residual replacement, to a 1e-10 scaled residual), reduce
the compliance in FP64 and run the adjoint kernel for d C / d(node coordinates)
and d C / d(element properties).  Inputs are resident on the GPU before timing
starts.

Also reported: the plain FP64 one-structure-per-launch GPU path, the batched
FP64 path, the block-Jacobi-preconditioned batched path, the error of the timed
path against the plain FP64 path and the FP64 oracle, the dominant kernel's
roofline (HIP events on the solver stream),
element stiffness + assembly throughput (elements/s), the oracle timed on this
host's cores, the 200x200 MITC-4 shell (242,406 DOF) as a single structure on
every solver path with its error against the committed oracle fixture, and two
more batches: 256 uneven beam gridshells x 4 load cases, and 64 SIMP shells.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--plain-designs", type=int, default=4)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-shell200", action="store_true", help="skip the 200x200 single-structure lines")
    ap.add_argument("--no-batches", action="store_true",
                    help="skip the gridshell-batch and SIMP-batch lines")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs every structure of a batch on a single "
              "MI355X and has no multi-GPU path (--gpus 1 only).", file=sys.stderr)
        sys.exit(2)

    import torch
    from jaxsso_amd import _native as nat
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch

    torch.cuda.init()
    L = nat.lib()
    designs = [wl.cfg1(d) for d in range(args.batch)]
    batch = StructureBatch(designs)

    def step(b, precision, tol):
        C, (dc, _, dq) = b.compliance_and_gradient(precision=precision, tol=tol)
        return C, dc, dq

    print(f"# batch of {args.batch} shells, {batch.plan.ndof} DOF total", flush=True)
    for _ in range(args.warmup):
        step(batch, args.precision, args.tol)
    torch.cuda.synchronize()
    L.jsso_timing_enable(8)
    L.jsso_timing_read(None, None, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        C, dc, dq = step(batch, args.precision, args.tol)
        iters.append(batch.plan.last_info["iterations"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.jsso_timing_enable(0)
    avg_ms = np.zeros(1)
    nsamp = np.zeros(1, np.int64)
    L.jsso_timing_read(nat.ptr(avg_ms), nat.ptr(nsamp), 1)
    info = dict(batch.plan.last_info)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = args.batch * args.steps / (t1 - t0)
    print(f"# timed: {ms_per_step:.1f} ms/step, {value:.1f} evals/s, CG iters {iters}, info {info}", flush=True)

    # -- dominant kernel roofline: k_pcg_dir, the direction and CG product of every
    # iteration (the stored upper blocks once in the solve's precision, their
    # indices; z and p_{k-1} read, p_k and q written once per DOF).  The timed
    # steps run the iterations as captured HIP graphs; one more solve with eager
    # launches brackets every k_pcg_dir launch with HIP events.
    p = batch.plan
    nnzb_off = p.nnzb - p.n_nodes
    s_m = 4 if args.precision == "mixed" else 8
    s_v = s_m
    ndof = p.ndof
    os.environ["JSSO_CG_GRAPH"] = "0"
    L.jsso_timing_read(None, None, 1)
    L.jsso_timing_enable(1)
    batch.solve(precision=args.precision, tol=args.tol, assemble=False)
    torch.cuda.synchronize()
    L.jsso_timing_enable(0)
    os.environ.pop("JSSO_CG_GRAPH")
    L.jsso_timing_read(nat.ptr(avg_ms), nat.ptr(nsamp), 1)
    # each off-diagonal block pair stored once (36 values), both halves indexed
    # (a column each, plus the value position of the transposed half)
    alg_bytes = (nnzb_off // 2) * 36 * s_m + nnzb_off * 4 + (nnzb_off // 2) * 4 + 4 * ndof * s_v
    achieved = alg_bytes / (avg_ms[0] * 1e-3) / 1e9 if avg_ms[0] > 0 else None
    kname = f"k_pcg_dir<{'float, float' if s_m == 4 else 'double, double'}, 1, 4>"
    traffic, traffic_src = _measured_traffic(args, kname)
    roofline = {"kernel": kname, "bound": "hbm",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "traffic_source": traffic_src,
                "avg_launch_us": avg_ms[0] * 1e3, "sampled_launches": int(nsamp[0]),
                "alg_bytes_per_launch": int(alg_bytes)}

    # -- element stiffness + assembly alone (elements/s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        batch.assemble()
    torch.cuda.synchronize()
    asm_ms = (time.perf_counter() - t0) * 1e3 / 5
    n_el = batch.plan.n_quad + batch.plan.n_beam

    # -- comparison paths on a few designs
    sub = StructureBatch(designs[:args.plain_designs])
    step(sub, "plain", args.tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Cp, dcp, dqp = step(sub, "plain", args.tol)
    torch.cuda.synchronize()
    plain_eps = args.plain_designs / (time.perf_counter() - t0)
    plain_iters = sub.plan.last_info["iterations"]
    step(batch, "fp64", args.tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Cf, dcf, dqf = step(batch, "fp64", args.tol)
    torch.cuda.synchronize()
    fp64_eps = args.batch / (time.perf_counter() - t0)
    fp64_iters = batch.plan.last_info["iterations"]
    # the same batch with the block-Jacobi preconditioner (fused one-kernel CG iteration)
    batch.plan.set_preconditioner("block_jacobi")
    step(batch, args.precision, args.tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step(batch, args.precision, args.tol)
    torch.cuda.synchronize()
    bj_eps = args.batch / (time.perf_counter() - t0)
    bj_iters = batch.plan.last_info["iterations"]
    batch.plan.set_preconditioner("amg")
    C, dc, dq = step(batch, args.precision, args.tol)     # the timed path's results for the error checks

    nd = args.plain_designs
    n_nodes_sub = sub.plan.n_nodes
    relerr = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    err_plain = {
        "C": relerr(C[:nd, 0], Cp[:, 0]),
        "grad_crds": relerr(dc[:n_nodes_sub], dcp),
        "grad_props": relerr(dq[:nd * 1600], dqp),
        "u": relerr(batch.u[:6 * n_nodes_sub], sub.u),
    }

    err_oracle = None
    cpu = None
    if not args.no_oracle_check or not args.no_cpu_baseline:
        import oracle.gradient as og
        st = designs[0]
        t0 = time.perf_counter()
        ref = og.compliance_and_gradient(st, faithful=True)
        cpu_s = time.perf_counter() - t0
        if not args.no_cpu_baseline:
            cpu = {"value": 1.0 / cpu_s, "unit": "evals/s", "cores": 1, "kind": "port",
                   "sample": "design 0 of the 40x40 shell: numpy/SciPy restatement of the reference "
                             "sequence (assembly, SuperLU on K_aug, second SuperLU on K_aug^T, "
                             "element VJP by complex step for every coordinate and property)",
                   "seconds": cpu_s, "cpu": _cpu_model()}
        if not args.no_oracle_check:
            acc = og.compliance_and_gradient(st)     # refined reduced solve (accurate FP64 oracle)
            u0 = batch.u[:6 * st.n_node].double().cpu().numpy()
            g0 = dc[:st.n_node].cpu().numpy()
            q0 = dq[:st.n_quad].cpu().numpy()
            rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
            err_oracle = {"C": abs(float(C[0, 0]) - acc["C"]) / acc["C"], "u": rel(u0, acc["u"]),
                          "grad_z_free": rel(g0[st.design_nodes, 2], acc["dcrds"][st.design_nodes, 2]),
                          "grad_crds": rel(g0, acc["dcrds"]), "grad_t": rel(q0[:, 0], acc["dprop_quad"][:, 0])}

    out = {
        "metric": "compliance+gradient evals/s at 1e-8 rel err vs plain FP64 GPU; HBM GB/s",
        "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": "MITC-4 shell 40x40 over 10 m (1,681 nodes, 1,600 quads, 10,086 DOF), "
                               f"batch of {args.batch} seeded thickness designs per step; compliance + "
                               "gradient w.r.t. all node coordinates and all element properties",
                   "batch": args.batch, "tol": args.tol, "precision": args.precision,
                   "preconditioner": "smoothed-aggregation AMG V-cycle (FP32)",
                   "cg_iterations": iters[-1], "parallelism": "batched structures, 1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_fp64_evals_per_s": plain_eps, "plain_fp64_cg_iterations": plain_iters,
        "fp64_batched_evals_per_s": fp64_eps, "fp64_batched_cg_iterations": fp64_iters,
        "block_jacobi_batched_evals_per_s": bj_eps, "block_jacobi_batched_cg_iterations": bj_iters,
        "amg_levels": batch.plan.amg_levels(),
        "speedup_vs_plain_fp64": value / plain_eps,
        "rel_err_vs_plain_fp64_gpu": err_plain, "rel_err_vs_fp64_oracle": err_oracle,
        "assembly_ms_per_step": asm_ms, "elements_per_s": n_el / (asm_ms * 1e-3),
        "shell200": None if args.no_shell200 else shell200(args.tol),
        "gridshell_batch": None if args.no_batches else gridshell_batch(args.tol),
        "simp_batch": None if args.no_batches else simp_batch(args.tol),
    }
    print(json.dumps(out), flush=True)


def shell200(tol):
    """The 200x200 MITC-4 shell (242,406 DOF) as one structure: compliance + full
    gradient per second for the mixed, batched-FP64 and plain FP64 paths, and the
    error against the committed FP64 oracle fixture."""
    import torch
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "tests", "golden", "shell200_oracle.json")) as fh:
        fx = json.load(fh)
    st = wl.cfg2()
    b = StructureBatch([st])
    res = {"workload": "MITC-4 shell 200x200 over 20 m (40,401 nodes, 40,000 quads, 242,406 DOF), "
                       "edges pinned, smooth Fourier thickness; one structure per evaluation"}
    rel = lambda a, r: float(np.linalg.norm(np.asarray(a) - np.asarray(r)) / np.linalg.norm(np.asarray(r)))
    for key, prec, pc in (("mixed", "mixed", "amg"), ("fp64", "fp64", "amg"),
                          ("mixed_block_jacobi", "mixed", "block_jacobi"), ("plain", "plain", None)):
        b.compliance_and_gradient(precision=prec, tol=tol, preconditioner=pc)     # warm-up (and kernel tuning)
        torch.cuda.synchronize()
        reps = 3 if pc == "amg" else 2
        t0 = time.perf_counter()
        for _ in range(reps):
            C, (dc, _, dq) = b.compliance_and_gradient(precision=prec, tol=tol, preconditioner=pc)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        dcn = dc.cpu().numpy().reshape(st.n_node, 3)
        dqn = dq.cpu().numpy().reshape(st.n_quad, 5)
        res[key] = {"preconditioner": pc or "block_jacobi (plain path)",
                     "evals_per_s": 1.0 / dt, "ms_per_eval": dt * 1e3,
                     "cg_iterations": b.plan.last_info["iterations"],
                     "waves_per_slice": b.plan.last_info.get("waves_per_slice"),
                     "rel_err_vs_fp64_oracle": {
                         "C": abs(float(C[0, 0]) - fx["C"]) / fx["C"],
                         "u_sample": rel(b.u.cpu().numpy().ravel()[fx["u_index"]], fx["u"]),
                         "grad_z_sample": rel(dcn[fx["dz_node"], 2], fx["dz"]),
                         "grad_t_sample": rel(dqn[fx["dt_elem"], 0], fx["dt"])}}
    res["speedup_mixed_vs_plain_fp64"] = res["mixed"]["evals_per_s"] / res["plain"]["evals_per_s"]
    res["speedup_mixed_vs_mixed_block_jacobi"] = (res["mixed"]["evals_per_s"] /
                                                  res["mixed_block_jacobi"]["evals_per_s"])
    return res


def _measured_traffic(args, kname):
    """HBM bytes per launch of the roofline kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a
    copy-kernel calibration); None for other workloads or kernels."""
    name = "amg_batch128_shell40_traffic.txt"
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    if args.batch != 128 or not os.path.exists(path):
        return None, None
    with open(path) as fh:
        for line in fh:
            if line.startswith(f"traffic_bytes_per_launch {kname}:"):
                return int(line.split(":")[1]), "profiles/" + name
    return None, None


def _timed(fn, reps=1):
    import torch
    fn()                                      # warm-up (and kernel tuning)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) / reps


def gridshell_batch(tol, n_designs=256):
    """256 beam gridshells of 8..32 bays (uneven sizes) with 4 load cases each in
    one plan: compliance and dC/dz per design and case (an evaluation = one
    design x one load case)."""
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    import oracle.gradient as og
    designs = wl.cfg3(n_designs)
    b = StructureBatch(designs)
    _, dt_bj = _timed(lambda: b.compliance_and_gradient(precision="mixed", tol=tol, preconditioner="block_jacobi"))
    bj_iters = b.plan.last_info["iterations"]
    (C, (dc, _, _)), dt = _timed(lambda: b.compliance_and_gradient(precision="mixed", tol=tol, preconditioner="amg"))
    st = designs[0]
    C = C.cpu().numpy()
    dz = b.split_nodes(dc.cpu().numpy())[0][:, 2, :]
    err_c, err_g = 0.0, 0.0
    for c in range(4):
        ref = og.compliance_and_gradient(st, loads=st.load_cases()[c])
        err_c = max(err_c, abs(C[0, c] - ref["C"]) / ref["C"])
        g, r = dz[st.design_nodes, c], ref["dcrds"][st.design_nodes, 2]
        err_g = max(err_g, float(np.linalg.norm(g - r) / np.linalg.norm(r)))
    return {"workload": f"{n_designs} beam gridshells, 8..32 bays (seeded), perimeter pinned, 4 load cases each, "
                        f"{b.plan.ndof} DOF in one plan", "precision": "mixed",
            "preconditioner": "amg", "evals_per_s": n_designs * 4 / dt, "ms_per_step": dt * 1e3,
            "cg_iterations": b.plan.last_info["iterations"], "converged": b.plan.last_info["converged"],
            "block_jacobi_evals_per_s": n_designs * 4 / dt_bj, "block_jacobi_cg_iterations": bj_iters,
            "rel_err_vs_fp64_oracle_design0": {"C": err_c, "grad_z_free": err_g}}


def simp_batch(tol, n_designs=64):
    """64 SIMP shells (100x100 MITC-4, 61,206 DOF each, E = rho^3 30 GPa) in one
    plan: compliance, dC/dz and dC/drho (chain rule through E) per design, AMG
    preconditioned; FP64 recurrences (with a 1e6 stiffness contrast the FP32
    ones need ~1.7x the iterations)."""
    import torch
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "tests", "golden", "simp100_oracle.json")) as fh:
        fx = json.load(fh)
    designs = wl.cfg4(n_designs)
    b = StructureBatch(designs)
    rho = torch.as_tensor(np.concatenate([d.rho for d in designs]), device=b.f.device)

    def step():
        C, (dc, _, dq) = b.compliance_and_gradient(precision="fp64", tol=tol)
        drho = dq.reshape(-1, 5)[:, 1] * 3.0 * rho ** 2 * 30e9
        return C, dc, drho

    (C, dc, drho), dt = _timed(step, reps=2)
    C = C.cpu().numpy()[:, 0]
    dc = dc.cpu().numpy().reshape(-1, 3)
    drho = drho.cpu().numpy()
    rel = lambda a, r: float(np.linalg.norm(np.asarray(a) - np.asarray(r)) / np.linalg.norm(np.asarray(r)))
    errs = {}
    for r in fx["designs"]:
        d = r["design"]
        if d >= n_designs:
            continue
        o = int(b.node_off[d])
        q0 = b.quad_off[d]
        errs[f"design{d}"] = {"C": abs(C[d] - r["C"]) / r["C"],
                              "grad_z_sample": rel(dc[o + np.asarray(r["dz_node"]), 2], r["dz"]),
                              "grad_rho_sample": rel(drho[q0 + np.asarray(r["drho_elem"])], r["drho"])}
    return {"workload": f"{n_designs} SIMP MITC-4 shells 100x100 over 15 m (10,201 nodes, 10,000 quads each), "
                        f"{b.plan.ndof} DOF in one plan", "precision": "fp64", "preconditioner": "amg",
            "evals_per_s": n_designs / dt, "ms_per_step": dt * 1e3,
            "cg_iterations": b.plan.last_info["iterations"], "converged": b.plan.last_info["converged"],
            "waves_per_slice": b.plan.last_info.get("waves_per_slice"),
            "rel_err_vs_fp64_oracle": errs}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical cpus)"
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
