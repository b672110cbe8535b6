#!/usr/bin/env python3
"""Benchmark: compliance + full gradient evaluations per second on one MI355X.

Headline workload -- the shape+topology SIMP batch: 64 seeded MITC-4 shell
designs of 100x100 quads over 15 m x 15 m (10,201 nodes, 10,000 quads, 61,206
DOF each; 3.92 M DOF in one plan), corners and mid-edge nodes pinned,
-1.5 kN/node, per-element density rho in [0.01, 1] with E = rho^3 30 GPa,
thickness 0.12 m, z from 6 Gaussian bumps.  One step evaluates all 64 designs:
element stiffness and assembly, the AMG numeric set-up, the solve (flexible CG
preconditioned by the smoothed-aggregation V-cycle, to a 1e-10 scaled
residual), the compliance in FP64 and the adjoint kernels for dC/d(node
coordinates) and dC/d(element properties), with dC/drho by the chain rule
through E.  Inputs are resident on the GPU before timing starts.

The JSON line carries:
* ``roofline``: the solve's dominant kernel (k_pcg_dir, the CG direction and
  product), algorithmic bytes per launch over its mean duration from HIP
  events on the solver stream, sampled inside the timed steps: every other
  iteration chunk is launched eagerly with its k_pcg_dir launches bracketed,
  and the samples of chunks that began with all designs still iterating (so
  that the launch processes the whole plan) make the average;
* ``assembly``: element-stiffness and gather kernels, same method;
* ``cpu_baseline``: the reference's sequence (assembly, SuperLU on K_aug, a
  second SuperLU on K_aug^T, element VJP) restated in numpy/SciPy, one design
  per single-threaded process on every host core (median warm time, first call
  separately);
* errors against the FP64 oracle (committed fixture, designs 0 and 1) and
  against the plain FP64 one-structure-per-launch GPU path;
* further lines: the 200x200 shell (one structure, 242,406 DOF) with its own
  roofline, 256 uneven beam gridshells x 4 load cases, the 40x40 shell.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))
TIMING_EVERY = 2        # one AMG iteration chunk in 2 runs eagerly with k_pcg_dir bracketed


def _rel(a, r):
    a = np.asarray(a, dtype=float)
    r = np.asarray(r, dtype=float)
    return float(np.linalg.norm(a - r) / np.linalg.norm(r))


def _slot(L, nat, name, reset=True):
    avg = np.zeros(1)
    n = np.zeros(1, np.int64)
    nat.check(L.jsso_timing_read_slot(nat.TIMING_SLOTS[name], nat.ptr(avg), nat.ptr(n), int(reset)))
    return float(avg[0]), int(n[0])


def _dir_bytes(plan, s_m, s_v):
    """Algorithmic bytes of one k_pcg_dir launch over the whole plan: the stored
    upper blocks once (36 values at s_m bytes), a column index per stored and
    per transposed position, a value position per transposed position, z and
    p_{k-1} read and p_k and q written (ndof values at s_v bytes each)."""
    off = plan.nnzb - plan.n_nodes
    return (off // 2) * 36 * s_m + off * 4 + (off // 2) * 4 + 4 * plan.ndof * s_v


SIZES = {"fp64": (8, 8, "double, double", "f64"), "mixed": (4, 4, "float, float", "f32+f64"),
         "mixed64": (4, 8, "float, double", "f32+f64")}


def _asm_bytes(plan):
    """Algorithmic bytes of the assembly kernels.
    element: node coordinates through the connectivity (12 doubles + 4 ints per
    quad, 6 + 2 per beam), properties, and the FP64 element blocks written;
    gather: every element block read once, its contribution index (4 B per
    block), the diagonal factors L^-1 and M^-1 written (72 doubles per node) and
    the stored upper blocks written in FP64, FP32 and FP16 (14 B per entry)."""
    nq, nb = plan.n_quad, plan.n_beam
    ke = (nq * 576 + nb * 144) * 8
    elem = nq * (12 * 8 + 16 + 5 * 8) + nb * (6 * 8 + 8 + 6 * 8) + ke
    off = plan.nnzb - plan.n_nodes
    gather = ke + (nq * 16 + nb * 4) * 4 + plan.n_nodes * 72 * 8 + (off // 2) * 36 * 14
    return elem, gather


def _measured_traffic(name, kname):
    """HBM bytes per launch of ``kname`` from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (counters scaled by a copy-kernel calibration)."""
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        for line in fh:
            if line.startswith(f"traffic_bytes_per_launch {kname}:"):
                return int(line.split(":")[1]), "profiles/" + name
    return None, None


def _assembly_traffic(name, kernels):
    """Sum of the committed per-launch HBM traffic (FETCH + WRITE) of ``kernels``."""
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    found = {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("assembly_traffic "):
                k, rest = line[len("assembly_traffic "):].split(":", 1)
                found[k.strip()] = int(rest.split("total")[1].split()[0])
    return sum(found[k] for k in kernels) if all(k in found for k in kernels) else None


def simp_headline(args, L, nat):
    """The SIMP batch: timed steps, roofline, assembly rates, errors."""
    import torch
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch

    designs = wl.cfg4(args.designs)
    b = StructureBatch(designs)
    rho = torch.as_tensor(np.concatenate([d.rho for d in designs]), device=b.f.device)

    solve_log = []

    def step(precision):
        C, (dc, _, dq) = b.compliance_and_gradient(precision=precision, tol=args.tol)
        drho = dq[:, 1] * 3.0 * rho ** 2 * 30e9
        if precision == args.precision:
            i = b.plan.last_info
            solve_log.append([i["iterations"], i["dir_launches"], i["full_iterations"]])
        return C, dc, drho

    print(f"# SIMP batch: {args.designs} shells, {b.plan.ndof} DOF in one plan", flush=True)
    if args.wps:
        b.plan.cg_tuning(args.precision, 1, args.wps)
    for _ in range(args.warmup):
        step(args.precision)
    torch.cuda.synchronize()
    L.jsso_timing_enable(TIMING_EVERY)
    for name in nat.TIMING_SLOTS:
        _slot(L, nat, name)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        C, dc, drho = step(args.precision)
        iters.append(b.plan.last_info["iterations"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.jsso_timing_enable(0)
    t = {name: _slot(L, nat, name) for name in nat.TIMING_SLOTS}
    info = dict(b.plan.last_info)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = args.designs * args.steps / (t1 - t0)
    print(f"# timed: {ms_per_step:.1f} ms/step, {value:.1f} evals/s, CG iterations {iters}, info {info}", flush=True)

    s_m, s_v, tname, _ = SIZES[args.precision]
    kname = f"k_pcg_dir<{tname}, 1, {info['waves_per_slice']}>"
    alg = _dir_bytes(b.plan, s_m, s_v)
    avg_ms, nsamp = t["solve"]
    achieved = alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    traffic, tsrc = _measured_traffic(f"simp64_shell100_{args.precision}_traffic.txt", kname)
    roofline = {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "traffic_source": tsrc, "avg_launch_us": avg_ms * 1e3, "sampled_launches": nsamp,
                "alg_bytes_per_launch": int(alg)}
    eb, gb = _asm_bytes(b.plan)
    (te, ne), (tg, ng) = t["element"], t["gather"]
    assembly = {"element_kernels_us": te * 1e3, "element_alg_bytes": eb,
                "element_GBps": eb / (te * 1e-3) / 1e9 if te > 0 else None,
                "gather_kernels_us": tg * 1e3, "gather_alg_bytes": gb,
                "gather_GBps": gb / (tg * 1e-3) / 1e9 if tg > 0 else None,
                "assembly_GBps": (eb + gb) / ((te + tg) * 1e-3) / 1e9 if te + tg > 0 else None,
                "assembly_frac_of_hbm": (eb + gb) / ((te + tg) * 1e-3) / 1e9 / HBM_PEAK_GBS if te + tg > 0 else None,
                "elements_per_s": b.plan.n_quad / ((te + tg) * 1e-3) if te + tg > 0 else None,
                "samples": ne}
    phases = {"element_ms": te, "gather_ms": tg, "amg_setup_ms": t["setup"][0], "adjoint_ms": t["adjoint"][0]}
    # FP64 rate of the element kernel and of the quad adjoint (operation counts
    # from the committed rocprofv3 SQ_INSTS_VALU_*_F64 summary of this workload)
    ff = "simp64_shell100_fp64_flops.txt"
    ke_name = f"ke_quad_kernel<0, {b.plan.element_tuning()}>"
    fe = _measured_flops(ff, ke_name)
    fa = [_measured_flops(ff, k) for k in ("adj_quad_fwd", "adj_quad_frame", "adj_quad_combine")]
    fa = sum(fa) if all(x is not None for x in fa) else None
    compute = {"fp64_vector_peak_TFLOPs": FP64_VECTOR_PEAK_TFLOPS, "source": f"profiles/{ff}" if fe else None,
               "element_kernel": ke_name, "element_flops_per_launch": fe,
               "element_TFLOPs": fe / (te * 1e-3) / 1e12 if fe and te > 0 else None,
               "adjoint_flops_per_step": fa,
               "adjoint_TFLOPs": fa / (t["adjoint"][0] * 1e-3) / 1e12 if fa and t["adjoint"][0] > 0 else None}
    for k in ("element", "adjoint"):
        v = compute[f"{k}_TFLOPs"]
        compute[f"{k}_frac_of_fp64_peak"] = v / FP64_VECTOR_PEAK_TFLOPS if v else None
    assembly["compute"] = compute
    tf = f"simp64_shell100_{args.precision}_traffic.txt"
    et = _assembly_traffic(tf, [ke_name])
    gt = _assembly_traffic(tf, ["diag_factor_kernel", "ell_assemble_kernel"])
    assembly["measured_traffic"] = {
        "source": f"profiles/{tf}" if et else None, "element_bytes": et, "gather_bytes": gt,
        "element_GBps": et / (te * 1e-3) / 1e9 if et and te > 0 else None,
        "gather_GBps": gt / (tg * 1e-3) / 1e9 if gt and tg > 0 else None}

    out = {"b": b, "value": value, "ms_per_step": ms_per_step, "iters": iters, "info": info,
           "roofline": roofline, "assembly": assembly, "phases_ms_per_step": phases,
           "solve_log": solve_log}
    if args.profile_run:
        return out
    C = C.cpu().numpy()[:, 0]
    dcn = dc.cpu().numpy().reshape(-1, 3)
    drn = drho.cpu().numpy()
    un = b.u.cpu().numpy()
    if not args.no_oracle_check:
        with open(os.path.join(HERE, "tests", "golden", "simp100_oracle.json")) as fh:
            fx = json.load(fh)
        errs = {}
        for r in fx["designs"]:
            d = r["design"]
            if d >= args.designs:
                continue
            o = int(b.node_off[d])
            q0 = b.quad_off[d]
            errs[f"design{d}"] = {"C": abs(C[d] - r["C"]) / r["C"],
                                  "u_sample": _rel(un[6 * o + np.asarray(r["u_index"])], r["u"]),
                                  "grad_z_sample": _rel(dcn[o + np.asarray(r["dz_node"]), 2], r["dz"]),
                                  "grad_rho_sample": _rel(drn[q0 + np.asarray(r["drho_elem"])], r["drho"])}
        out["rel_err_vs_fp64_oracle"] = errs

    # plain FP64 one-structure-per-launch GPU path on the first designs
    nd = args.plain_designs
    sub = StructureBatch(designs[:nd])
    rho_s = rho[:sub.plan.n_quad]
    sub.compliance_and_gradient(precision="plain", tol=args.tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Cp, (dcp, _, dqp) = sub.compliance_and_gradient(precision="plain", tol=args.tol)
    torch.cuda.synchronize()
    plain_eps = nd / (time.perf_counter() - t0)
    drp = (dqp[:, 1] * 3.0 * rho_s ** 2 * 30e9).cpu().numpy()
    no = int(sub.plan.n_nodes)
    out["plain_fp64"] = {"evals_per_s": plain_eps, "designs": nd,
                         "cg_iterations": sub.plan.last_info["iterations"],
                         "rel_err_headline_vs_plain": {
                             "C": _rel(C[:nd], Cp.cpu().numpy()[:, 0]),
                             "u": _rel(un[:6 * no], sub.u.cpu().numpy()),
                             "grad_crds": _rel(dcn[:no], dcp.cpu().numpy().reshape(-1, 3)),
                             "grad_rho": _rel(drn[:sub.plan.n_quad], drp)}}
    out["speedup_vs_plain_fp64"] = value / plain_eps

    # the other batched precisions on the same batch
    for other in sorted(SIZES):
        if other == args.precision:
            continue
        step(other)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(other)
        torch.cuda.synchronize()
        out[f"{other}_batched"] = {"evals_per_s": args.designs / (time.perf_counter() - t0),
                                   "cg_iterations": b.plan.last_info["iterations"],
                                   "converged": b.plan.last_info["converged"]}
    return out


# The 200x200 shell workload states its own accuracy: FP64 refinement to 1e-8
# relative residual (every path of that line solves to it).
SHELL200_TOL = 1e-8


def shell200(args, L, nat, profile=False):
    """The 200x200 MITC-4 shell (242,406 DOF) as one structure: evals/s on the
    AMG path (mixed and FP64) and the plain FP64 path, errors against the
    committed FP64 oracle fixture, and the mixed path's k_pcg_dir roofline
    (profile: the mixed path only, with its solve log).  Every path solves to
    the workload's FP64 relative residual of 1e-8 (SHELL200_TOL)."""
    import torch
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    with open(os.path.join(HERE, "tests", "golden", "shell200_oracle.json")) as fh:
        fx = json.load(fh)
    st = wl.cfg2()
    b = StructureBatch([st])
    res = {"workload": "MITC-4 shell 200x200 over 20 m (40,401 nodes, 40,000 quads, 242,406 DOF), edges pinned, "
                       "smooth Fourier thickness; one structure per evaluation", "tol": SHELL200_TOL}
    solve_log = []
    paths = (("mixed", "mixed", "amg", 10), ("fp64", "fp64", "amg", 5), ("plain", "plain", None, 1))
    for key, prec, pc, reps in paths[:1] if profile else paths:
        if args.wps and key == "mixed":
            b.plan.set_preconditioner("amg")
            b.plan.cg_tuning(prec, 1, args.wps)
        b.compliance_and_gradient(precision=prec, tol=SHELL200_TOL, preconditioner=pc)     # warm-up
        if key == "mixed":
            i = b.plan.last_info
            solve_log.append([i["iterations"], i["dir_launches"], i["full_iterations"]])
        torch.cuda.synchronize()
        if key == "mixed":
            L.jsso_timing_enable(TIMING_EVERY)
            for name in nat.TIMING_SLOTS:
                _slot(L, nat, name)
        t0 = time.perf_counter()
        for _ in range(reps):
            C, (dc, _, dq) = b.compliance_and_gradient(precision=prec, tol=SHELL200_TOL, preconditioner=pc)
            if key == "mixed":
                i = b.plan.last_info
                solve_log.append([i["iterations"], i["dir_launches"], i["full_iterations"]])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        dcn = dc.cpu().numpy().reshape(st.n_node, 3)
        dqn = dq.cpu().numpy().reshape(st.n_quad, 5)
        res[key] = {"preconditioner": pc or "block_jacobi (plain path)", "evals_per_s": 1.0 / dt,
                    "ms_per_eval": dt * 1e3, "cg_iterations": b.plan.last_info["iterations"],
                    "rel_err_vs_fp64_oracle": {
                        "C": abs(float(C[0, 0]) - fx["C"]) / fx["C"],
                        "u_sample": _rel(b.u.cpu().numpy().ravel()[fx["u_index"]], fx["u"]),
                        "grad_z_sample": _rel(dcn[fx["dz_node"], 2], fx["dz"]),
                        "grad_t_sample": _rel(dqn[fx["dt_elem"], 0], fx["dt"])}}
        if key == "mixed":
            L.jsso_timing_enable(0)
            t = {name: _slot(L, nat, name) for name in nat.TIMING_SLOTS}
            kname = f"k_pcg_dir<float, float, 1, {b.plan.last_info['waves_per_slice']}>"
            alg = _dir_bytes(b.plan, 4, 4)
            avg_ms, nsamp = t["solve"]
            ach = alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
            traffic, tsrc = _measured_traffic("shell200_mixed_traffic.txt", kname)
            res["roofline"] = {"kernel": kname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": ach / HBM_PEAK_GBS if ach else None, "traffic": traffic,
                               "traffic_source": tsrc, "avg_launch_us": avg_ms * 1e3, "sampled_launches": nsamp,
                               "alg_bytes_per_launch": int(alg),
                               "note": "working set (55 MB FP32 matrix) fits the 256 MB MALL"}
            eb, gb = _asm_bytes(b.plan)
            (te, _), (tg, _) = t["element"], t["gather"]
            res["assembly"] = {"element_kernels_us": te * 1e3, "gather_kernels_us": tg * 1e3,
                               "assembly_GBps": (eb + gb) / ((te + tg) * 1e-3) / 1e9 if te + tg > 0 else None}
            res["phases_ms_per_eval"] = {"element_ms": te, "gather_ms": tg, "amg_setup_ms": t["setup"][0],
                                         "adjoint_ms": t["adjoint"][0]}
    res["solve_log"] = {"fields": ["iterations", "dir_launches", "full_iterations"], "solves": solve_log}
    if not profile:
        res["speedup_mixed_vs_plain_fp64"] = res["mixed"]["evals_per_s"] / res["plain"]["evals_per_s"]
    return res


def simp_loop(args, L, nat, n_steps=8, step=0.05):
    """The SIMP batch inside an optimisation loop: every step evaluates the 64
    designs, then moves every element density by a normalised gradient step
    (rho <- clip(rho - step * g / max|g|, 0.01, 1) per design), so each step
    assembles and solves new designs.  The same sequence runs twice: with the AMG
    hierarchy rebuilt after every assembly, and with it kept across steps
    (AmgKeepPolicy: only the level-0 smoothing bound follows each design; a
    rebuild when the iterations grow by 25%).  Reported: evals/s, iterations per
    step, AMG set-up ms per step, rebuilds, and the largest relative difference
    of the compliances of the two runs."""
    import torch
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    from jaxsso_amd.plan import AmgKeepPolicy
    designs = wl.cfg4(args.designs)
    b = StructureBatch(designs)
    nd, nq = args.designs, designs[0].n_quad
    rho0 = torch.as_tensor(np.concatenate([d.rho for d in designs]), device=b.f.device)
    res = {"workload": f"SIMP batch ({nd} x 100x100 shells), {n_steps} normalised-gradient steps on rho "
                       f"(step {step}), FP64"}
    Cs = {}
    for keep in (False, True):
        rho = rho0.clone()
        b.plan.amg_reuse(0)
        b.prop_quad[:, 1] = rho ** 3 * 30e9
        b.compliance_and_gradient(precision="fp64", tol=args.tol)        # warm-up on the start design
        policy = AmgKeepPolicy(b.plan) if keep else None
        torch.cuda.synchronize()
        L.jsso_timing_enable(TIMING_EVERY)
        _slot(L, nat, "setup")
        iters, Cl = [], []
        t0 = time.perf_counter()
        for _ in range(n_steps):
            b.prop_quad[:, 1] = rho ** 3 * 30e9
            C, (_, _, dq) = b.compliance_and_gradient(precision="fp64", tol=args.tol)
            g = (dq[:, 1] * 3.0 * rho ** 2 * 30e9).reshape(nd, nq)
            gmax = g.abs().amax(1, keepdim=True).clamp_min(1e-300)
            rho = (rho.reshape(nd, nq) - step * g / gmax).clamp(0.01, 1.0).reshape(-1)
            iters.append(b.plan.last_info["iterations"])
            Cl.append(C[:, 0])
            if policy:
                policy.after_solve()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        L.jsso_timing_enable(0)
        setup_ms, _ = _slot(L, nat, "setup")
        Cs[keep] = torch.stack(Cl)
        key = "keep_hierarchy" if keep else "rebuild_every_step"
        res[key] = {"evals_per_s": nd * n_steps / dt, "ms_per_step": dt * 1e3 / n_steps, "cg_iterations": iters,
                    "amg_setup_ms_per_step": setup_ms, "amg": b.plan.amg_reuse(-1),
                    "rebuilds_triggered": policy.rebuilds if policy else None}
    b.plan.amg_reuse(0)
    res["max_rel_diff_C"] = float(((Cs[True] - Cs[False]).abs() / Cs[False].abs()).max())
    res["speedup_keep_vs_rebuild"] = res["keep_hierarchy"]["evals_per_s"] / res["rebuild_every_step"]["evals_per_s"]
    return res


def _timed(fn, reps=1):
    import torch
    fn()                                      # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) / reps


def gridshell_batch(args, n_designs=256):
    """256 beam gridshells of 8..32 bays (uneven sizes) with 4 load cases each in
    one plan: compliance and dC/dz per design and case (an evaluation = one
    design x one load case)."""
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    designs = wl.cfg3(n_designs)
    b = StructureBatch(designs)
    (C, (dc, _, _)), dt = _timed(lambda: b.compliance_and_gradient(precision="mixed", tol=args.tol), reps=3)
    res = {"workload": f"{n_designs} beam gridshells, 8..32 bays (seeded), perimeter pinned, 4 load cases each, "
                       f"{b.plan.ndof} DOF in one plan", "precision": "mixed", "preconditioner": "amg",
           "evals_per_s": n_designs * 4 / dt, "ms_per_step": dt * 1e3,
           "cg_iterations": b.plan.last_info["iterations"], "converged": b.plan.last_info["converged"]}
    if not args.no_oracle_check:
        import oracle.gradient as og
        st = designs[0]
        Cn = C.cpu().numpy()
        dz = b.split_nodes(dc.cpu().numpy())[0][:, 2, :]
        err_c, err_g = 0.0, 0.0
        for c in range(4):
            ref = og.compliance_and_gradient(st, loads=st.load_cases()[c])
            err_c = max(err_c, abs(Cn[0, c] - ref["C"]) / ref["C"])
            err_g = max(err_g, _rel(dz[st.design_nodes, c], ref["dcrds"][st.design_nodes, 2]))
        res["rel_err_vs_fp64_oracle_design0"] = {"C": err_c, "grad_z_free": err_g}
    return res


def shell40(args):
    """The 40x40 MITC-4 shell (10,086 DOF): one structure per evaluation, and a
    batch of 128 seeded thickness designs per launch."""
    from jaxsso_amd import workloads as wl
    from jaxsso_amd.batch import StructureBatch
    one = StructureBatch([wl.cfg1(0)])
    _, dt1 = _timed(lambda: one.compliance_and_gradient(precision="mixed", tol=args.tol), reps=10)
    batch = StructureBatch([wl.cfg1(d) for d in range(128)])
    _, dtb = _timed(lambda: batch.compliance_and_gradient(precision="mixed", tol=args.tol), reps=5)
    return {"workload": "MITC-4 shell 40x40 over 10 m (1,681 nodes, 1,600 quads, 10,086 DOF), four fixed corners",
            "single_evals_per_s": 1.0 / dt1, "single_cg_iterations": one.plan.last_info["iterations"],
            "batch128_evals_per_s": 128 / dtb, "batch128_cg_iterations": batch.plan.last_info["iterations"]}


def trace_summary(trace_dir, bench_log):
    """Per-kernel statistics of a rocprofv3 CSV kernel trace of ``bench.py
    --profile-run``, and the roofline kernel's launches that processed the whole
    plan: the first full_iterations launches of each headline solve (solve_log of
    the bench output; every solve launches the kernel dir_launches times)."""
    import csv
    import glob
    import re
    traces = sorted(glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True))
    if not traces:
        sys.exit(f"no *kernel_trace.csv under {trace_dir}")
    rows = []
    for tpath in traces:
        with open(tpath) as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    out = None
    with open(bench_log) as fh:
        for line in fh:
            if line.startswith("{"):
                out = json.loads(line)
    kname = out["roofline"]["kernel"]
    stats = {}
    for s0, s1, name in rows:
        st = stats.setdefault(name, [0, 0.0])
        st[0] += 1
        st[1] += (s1 - s0) / 1e3
    total = sum(v[1] for v in stats.values())
    print(f"# kernel trace: {len(rows)} dispatches, {total / 1e3:.2f} ms of kernel time")
    print(f"{'total_ms':>10} {'calls':>7} {'avg_us':>9}  kernel")
    for name, (n, us) in sorted(stats.items(), key=lambda kv: -kv[1][1])[:30]:
        short = re.sub(r"\s+", " ", name)[:120]
        print(f"{us / 1e3:10.2f} {n:7d} {us / n:9.2f}  {short}")
    d = [(s1 - s0) / 1e3 for s0, s1, name in rows if name.startswith(f"void jsso::{kname}")]
    log = out["solve_log"]["solves"]
    # the live tuner's timing launches (first solve on the plan, before its own
    # launches) are not part of any solve
    extra = len(d) - sum(l for _, l, _ in log)
    print(f"# {kname}: {len(d)} launches in the trace, {sum(l for _, l, _ in log)} expected from solve_log"
          f"{f' (the first {extra}: the tuner)' if extra > 0 else ''}")
    if extra > 0:
        d = d[extra:]
    full, pos = [], 0
    for iters, launches, full_it in log:
        full.extend(d[pos:pos + full_it])
        pos += launches
    if full:
        alg = out["roofline"]["alg_bytes_per_launch"]
        avg = sum(full) / len(full)
        print(f"roofline_launches {kname}: {len(full)} launches with every design running, "
              f"avg {avg:.2f} us, min {min(full):.2f}, max {max(full):.2f}; "
              f"{alg} algorithmic bytes -> {alg / (avg * 1e-6) / 1e9:.1f} GB/s "
              f"(bench.py events: {out['roofline']['avg_launch_us']:.2f} us)")
    # one iteration of the last solve with every design running (the kernels
    # from one roofline-kernel launch to the next) and the whole solve's GPU time
    starts = [(s0, s1) for s0, s1, name in rows if name.startswith(f"void jsso::{kname}")][extra if extra > 0 else 0:]
    iters, launches, full_it = log[-1]
    last = starts[len(starts) - launches:]
    if full_it >= 2 and len(last) == launches:
        a, b = last[full_it // 2][0], last[full_it // 2 + 1][0]
        print(f"# one full iteration (iteration {full_it // 2 + 1} of the last solve): {(b - a) / 1e3:.1f} us "
              f"from one {kname} start to the next")
        for s0, s1, name in rows:
            if a <= s0 < b:
                print(f"#   +{(s0 - a) / 1e3:7.1f} us  {(s1 - s0) / 1e3:7.1f} us  {re.sub(chr(32) + '+', ' ', name)[:90]}")
        print(f"# last solve: {iters} iterations ({full_it} with every design running), "
              f"{(last[-1][0] - last[0][0]) / 1e6:.2f} ms from its first to its last {kname} launch")


FP64_VECTOR_PEAK_TFLOPS = 78.6
FLOP_COUNTERS = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64")


def flops_summary(pmc_dir, bench_log):
    """FP64 operations per launch of the MITC-4 element kernel and of the quad
    adjoint kernels from a rocprofv3 --pmc run of the four SQ_INSTS_VALU_*_F64
    counters (wave instructions): 64 lanes per instruction, 2 per FMA (an upper
    bound where lanes are masked off)."""
    import csv
    import glob
    import re
    per = {}
    for path in sorted(glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True)):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] not in FLOP_COUNTERS:
                    continue
                d = per.setdefault(int(r["Dispatch_Id"]), {"name": re.sub(r"\(.*", "", r["Kernel_Name"])})
                d[r["Counter_Name"]] = float(r["Counter_Value"])
    agg = {}
    for d in per.values():
        fl = 64.0 * (d.get("SQ_INSTS_VALU_ADD_F64", 0) + d.get("SQ_INSTS_VALU_MUL_F64", 0) +
                     d.get("SQ_INSTS_VALU_TRANS_F64", 0) + 2.0 * d.get("SQ_INSTS_VALU_FMA_F64", 0))
        agg.setdefault(d["name"], []).append(fl)
    print("# FP64 lane operations per launch (64 x (ADD + MUL + TRANS + 2 FMA) wave instructions)")
    for name in sorted(agg):
        if name.startswith("void jsso::ke_quad_kernel") or name.startswith("jsso::adj_quad"):
            v = agg[name]
            print(f"flops_per_launch {name.replace('void jsso::', '').replace('jsso::', '')}: {int(sum(v) / len(v))}"
                  f"  # {len(v)} launches")


def _measured_flops(name, kname):
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(f"flops_per_launch {kname}:"):
                return int(line.split(":")[1].split("#")[0])
    return None


def traffic_summary(fetch_dir, write_dir, bench_log):
    """HBM bytes per launch of the roofline kernel from two rocprofv3 --pmc runs of
    ``bench.py --profile-run`` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
    each scaled by what the counter reads for the run's 1 GiB device copy; only
    the launches that processed the whole plan (solve_log) are averaged."""
    import csv
    import glob
    out = None
    with open(bench_log) as fh:
        for line in fh:
            if line.startswith("{"):
                out = json.loads(line)
    kname = out["roofline"]["kernel"]
    log = out["solve_log"]["solves"]
    res = {}
    asm = {}   # assembly kernels: counter -> name -> values per launch
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        vals, copies = [], []
        for path in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
            with open(path) as fh:
                for r in csv.DictReader(fh):
                    if r["Counter_Name"] != counter:
                        continue
                    v = float(r["Counter_Value"]) * 1024.0
                    name = r["Kernel_Name"]
                    if name.startswith(f"void jsso::{kname}"):
                        vals.append((int(r.get("Dispatch_Id", 0)), v))
                    elif "copyBuffer" in name or "copy" in name.lower():
                        copies.append(v)
                    for a in ("ke_quad_kernel<0, 4>", "ke_quad_kernel<0, 8>", "ke_quad_kernel<0, 16>",
                              "diag_factor_kernel", "ell_assemble_kernel"):
                        if a in name:
                            asm.setdefault(counter, {}).setdefault(a, []).append(v)
        vals.sort()
        v = [x for _, x in vals]
        extra = len(v) - sum(l for _, l, _ in log)    # the live tuner's launches come first
        if extra > 0:
            v = v[extra:]
        full, pos = [], 0
        for iters, launches, full_it in log:
            full.extend(v[pos:pos + full_it])
            pos += launches
        ref = max(copies) if copies else None
        factor = (1 << 30) / ref if ref else None
        res[counter] = (sum(full) / len(full) if full else None, factor, len(full), len(v))
        print(f"# {counter}: {len(v)} dispatches of {kname}, {len(full)} with every design running; "
              f"1 GiB copy reads {ref:.4g} B -> factor {factor:.4f}" if ref else f"# {counter}: no copy found")
    (f, ff, n, _), (w, wf, _, _) = res["FETCH_SIZE"], res["WRITE_SIZE"]
    if f is not None and w is not None and ff and wf:
        traffic = f * ff + w * wf
        alg = out["roofline"]["alg_bytes_per_launch"]
        print(f"fetch_bytes_per_launch {kname}: {int(f * ff)}")
        print(f"write_bytes_per_launch {kname}: {int(w * wf)}")
        print(f"traffic_bytes_per_launch {kname}: {int(traffic)}")
        print(f"# traffic / algorithmic bytes = {traffic / alg:.3f} over {n} launches")
        # assembly kernels (every launch covers the whole batch; the element kernel's
        # three variants are the autotuner's candidates)
        for a in sorted(asm.get("FETCH_SIZE", {})):
            fv, wv = asm["FETCH_SIZE"][a], asm.get("WRITE_SIZE", {}).get(a, [])
            if fv and wv:
                fb, wb = sum(fv) / len(fv) * ff, sum(wv) / len(wv) * wf
                print(f"assembly_traffic {a}: fetch {int(fb)} write {int(wb)} total {int(fb + wb)}"
                      f"  # {len(fv)} launches")


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical cpus)"
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--designs", type=int, default=64)
    ap.add_argument("--plain-designs", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--precision", default="fp64", choices=sorted(SIZES))
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the headline batch")
    ap.add_argument("--wps", type=int, choices=[0, 1, 2, 4], default=4,
                    help="level-0 waves per slice of the AMG kernels: 4 (default; the live tuner's choice on the "
                         "SIMP batch, pinned so that every run and profile times the same roofline kernel) or 0 "
                         "for the live tuner")
    ap.add_argument("--profile-run", nargs="?", const="simp", choices=["simp", "shell200"],
                    help="only the headline's warm-up and timed steps, or the 200x200 shell's mixed path "
                         "(for rocprofv3); implies --no-extras --no-cpu-baseline --no-oracle-check")
    ap.add_argument("--trace-summary", metavar="DIR",
                    help="summarise a rocprofv3 CSV kernel trace (with --bench-log) and exit")
    ap.add_argument("--bench-log", metavar="FILE", help="output of the profiled bench.py --profile-run")
    ap.add_argument("--flops-summary", metavar="DIR",
                    help="summarise a rocprofv3 --pmc run of the SQ_INSTS_VALU_*_F64 counters and exit")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) with --bench-log and exit")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.bench_log)
        return
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary, args.bench_log)
        return
    if args.flops_summary:
        flops_summary(args.flops_summary, args.bench_log)
        return
    if args.profile_run:
        args.no_extras = args.no_cpu_baseline = args.no_oracle_check = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs every structure of a batch on a single "
              "MI355X and has no multi-GPU path (--gpus 1 only).", file=sys.stderr)
        sys.exit(2)

    import torch
    from jaxsso_amd import _native as nat

    torch.cuda.init()
    L = nat.lib()
    if args.profile_run:
        # calibration for HBM counters: three device copies of 1 GiB
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.profile_run == "shell200":
        r = shell200(args, L, nat, profile=True)
        print(json.dumps({"workload": r["workload"], "roofline": r["roofline"], "solve_log": r["solve_log"],
                          "mixed": r["mixed"]}), flush=True)
        return
    h = simp_headline(args, L, nat)
    b = h.pop("b")
    extras = {}
    if not args.no_extras:
        print("# 200x200 shell", flush=True)
        extras["shell200"] = shell200(args, L, nat)
        print("# SIMP optimisation loop", flush=True)
        extras["simp_loop"] = simp_loop(args, L, nat)
        print("# gridshell batch", flush=True)
        extras["gridshell_batch"] = gridshell_batch(args)
        print("# 40x40 shell", flush=True)
        extras["shell40"] = shell40(args)
    cpu = None
    if not args.no_cpu_baseline:
        print(f"# CPU baseline: {args.cpu_workers} worker processes", flush=True)
        from oracle import baseline
        r = baseline.run_pool(4, args.cpu_workers, warm=1)
        cpu = {"value": r["evals_per_s"], "unit": "evals/s", "cores": args.cpu_workers, "kind": "port",
               "sample": f"designs 0..{args.cpu_workers - 1} of the SIMP batch, one per single-threaded process: "
                         "numpy/SciPy restatement of the reference sequence (assembly, SuperLU on K_aug, second "
                         "SuperLU on K_aug^T, element VJP for node z and E); value = workers / median warm "
                         "seconds per evaluation",
               "median_warm_s": r["median_warm_s"], "median_first_call_s": r["median_first_s"],
               "warm_runs": len(r["warm_s"]), "cpu": _cpu_model()}

    out = {
        "metric": "compliance+gradient evals/s at 1e-8 rel err vs plain FP64 GPU; HBM GB/s",
        "value": h["value"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": SIZES[args.precision][3], "data": "synthetic",
        "config": {"workload": f"SIMP shape+topology batch: {args.designs} seeded MITC-4 shells of 100x100 quads "
                               "over 15 m (10,201 nodes, 10,000 quads, 61,206 DOF each), E = rho^3 30 GPa, "
                               "corners + mid-edge nodes pinned, -1.5 kN/node; compliance + gradient w.r.t. "
                               "node coordinates and rho (through E) per design",
                   "batch": args.designs, "tol": args.tol, "precision": args.precision,
                   "preconditioner": "smoothed-aggregation AMG V-cycle (FP32 / FP16)",
                   "cg_iterations": h["iters"][-1], "parallelism": "batched structures, 1 GPU"},
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "assembly": h["assembly"],
        "phases_ms_per_step": h["phases_ms_per_step"],
        "amg_levels": b.plan.amg_levels(),
        "solve_log": {"fields": ["iterations", "dir_launches", "full_iterations"], "solves": h["solve_log"]},
        "rel_err_vs_fp64_oracle": h.get("rel_err_vs_fp64_oracle"),
        "plain_fp64": h.get("plain_fp64"), "speedup_vs_plain_fp64": h.get("speedup_vs_plain_fp64"),
    }
    for k in ("mixed_batched", "fp64_batched", "mixed64_batched"):
        if k in h:
            out[k] = h[k]
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
