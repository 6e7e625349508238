"""Benchmark: DG right-hand-side evaluations per second on one MI355X.

Default workload: linear acoustic wave, box [-1,1]^3 as 30^3 hexahedra x 6
tetrahedra (162,000 tets), P3, 4 fields of 3.24M nodes each, FP64, upwind
flux; 1000 timed RHS evaluations.  A "step" is one RHS evaluation
``rhs = L(q)`` with ``q`` already resident in HBM.

Prints one JSON line with the metric, the roofline position of the fused RHS
kernel and a CPU baseline (the oracle's OpenMP C restatement of the fused
RHS on the same workload, all host cores up to 16).  The kernel time is one
HIP event pair on the launch stream around all K back-to-back launches,
divided by K: it includes the gaps between consecutive launches.

    python bench.py [--steps 1000] [--warmup 20] [--workload wave_p3]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 (vector and matrix)


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="wave_p3")
    p.add_argument("--precision", default="fp64")
    p.add_argument("--variant", default="auto", help="FP64 kernel: auto | mfma | lds")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="approximate CPU time budget of the cpu_baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--adjoint-steps", type=int, default=200,
                   help="timed VJP evaluations for the adjoint/forward cost ratio (0: skip)")
    p.add_argument("--trajectory-adjoint", action="store_true",
                   help="instead of the RHS benchmark: gradient of a final-time misfit through "
                        "an LSERK4 trajectory (maxwell_p4: FP64, 100 steps, checkpoints every 10; "
                        "wave_p4_large: BF16x3 forward, FP32 adjoint, 200 steps, every 20, "
                        "error against FP64 forward + adjoint)")
    p.add_argument("--copy-probe", action="store_true",
                   help="also run a plain device copy of known size (HBM counter "
                        "calibration under rocprofv3 --pmc)")
    return p.parse_args()


def measured_traffic(workload: str, kernel: str) -> int | None:
    """HBM bytes per launch of ``kernel`` on ``workload`` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, calibrated against a
    copy of known size), or None when the summary has no row for it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fp64.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == workload and cells[1] == kernel:
            try:
                return int(float(cells[4]))
            except ValueError:
                return None
    return None


def cpu_baseline(workload, d, q, seconds: float):
    """Time the oracle's OpenMP C restatement of the fused RHS (oracle/dg_cpu.c,
    checked against the numpy oracle in tests/) on the same workload, with one
    thread per host core up to 16 (the box's CPU share), for about ``seconds``
    of wall time."""
    from oracle import dg_cpu

    threads = min(os.cpu_count() or 1, 16)
    op = dg_cpu.CpuOperator(d, workload.equation, workload.param, threads=threads)
    out = np.empty_like(q)
    op(q, out)                                   # warm-up (page-in, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        op(q, out)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 1000:
            break
    return {
        "value": n / el,
        "unit": "RHS evals/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"OpenMP C restatement of the fused RHS (oracle/dg_cpu.c), "
                   f"{workload.equation} P{workload.order} on the full {d.nelements}-tet "
                   f"mesh, {n} evals in {el:.1f} s, {threads} threads"),
    }


def trajectory_adjoint(w, plan, d, q, dev) -> dict:
    """Forward trajectory alone, then forward + reverse sweep, timed on the
    host around device synchronisations; the mixed wave workload also runs the
    FP64 forward + adjoint and reports the gradient's relative L2 error."""
    import numpy as np
    import torch

    from actx_dgfem_suite_amd.timestepping import LSERK4, cfl_timestep, misfit_gradient
    from actx_dgfem_suite_amd.workloads import initial_state, maxwell_target

    if w.name == "maxwell_p4":
        nsteps, every, fwd, adj = 100, 10, "fp64", torch.float64
        target = [torch.from_numpy(np.ascontiguousarray(f)).to(dev) for f in maxwell_target(d)]
    elif w.name == "wave_p4_large":
        nsteps, every, fwd, adj = 200, 20, "bf16x3", torch.float32
        target = [torch.from_numpy(np.ascontiguousarray(initial_state(w, d, seed=1)[0])).to(dev)]
    else:
        raise SystemExit(f"--trajectory-adjoint is defined for maxwell_p4 and wave_p4_large, "
                         f"not {w.name} (the Euler benchmark state, a uniform flow into walls, "
                         "forms a shock that an unlimited DG solution cannot follow)")
    speed = 1.0
    dt = cfl_timestep(d, speed)
    st = LSERK4(plan, dt, precision=fwd)
    qf = [t.clone() for t in q]
    st.run(qf, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st.run(qf, nsteps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    j, grad = misfit_gradient(st, q, target, nsteps, every, adjoint_dtype=adj)
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    out = {"workload": w.name, "steps": nsteps, "checkpoint_every": every,
           "forward_precision": fwd, "adjoint_dtype": str(adj).replace("torch.", ""),
           "dt": dt, "objective": j, "forward_s": t_fwd, "forward_plus_adjoint_s": t_all,
           "adjoint_over_forward": (t_all - t_fwd) / t_fwd}
    if fwd != "fp64":
        j64, g64 = misfit_gradient(LSERK4(plan, dt), q, target, nsteps, every)
        out["objective_fp64"] = j64
        out["grad_rel_l2_vs_fp64"] = [
            float(torch.linalg.norm(a.double() - b) / max(float(torch.linalg.norm(b)), 1e-300))
            for a, b in zip(grad, g64)]
    return out


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU path is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from actx_dgfem_suite_amd.perf_analysis import box_mesh_counts, float64_flops_on_mesh
    from actx_dgfem_suite_amd.plan import DGPlan
    from actx_dgfem_suite_amd.workloads import (
        WORKLOADS,
        algorithmic_bytes,
        algorithmic_flops,
        initial_state,
    )

    w = WORKLOADS[args.workload]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    d = w.discretization()
    plan = DGPlan(d, w.equation, w.param, device=dev, variant=args.variant)
    mixed = args.precision != "fp64"
    q = [torch.from_numpy(np.ascontiguousarray(f)).to(dev) for f in initial_state(w, d)]
    field_err = None
    if mixed:
        # per-field relative L2 of the mixed form against the FP64 form
        ref = plan.rhs(q)
        q = [t.float() for t in q]
        got = plan.rhs(q, precision=args.precision)
        torch.cuda.synchronize()
        # a field whose FP64 RHS is negligible (below 1e-3 of the whole RHS,
        # e.g. the transverse momentum of a uniform x-flow) is measured against
        # the whole RHS instead of its own round-off
        total = float(torch.sqrt(sum(torch.sum(r * r) for r in ref)))
        field_err = [float(torch.linalg.norm(g.double() - r)
                           / max(float(torch.linalg.norm(r)), 1e-3 * total))
                     for g, r in zip(got, ref)]
        del ref, got
    if args.trajectory_adjoint:
        print(json.dumps({"trajectory_adjoint": trajectory_adjoint(w, plan, d, q, dev)}),
              flush=True)
        return
    out = plan.empty_fields(torch.float32 if mixed else torch.float64)
    nf = len(q)
    k, npn, nfp = d.nelements, d.np, d.nfp
    print(f"[bench] {w.name}: K={k} Np={npn} fields={nf} setup done", file=sys.stderr,
          flush=True)

    for _ in range(args.warmup):
        plan.rhs(q, out, precision=args.precision)
    torch.cuda.synchronize()

    # One HIP event pair on the launch stream brackets the K launches: their
    # average duration, with the back-to-back launches' gaps and without the
    # cost of an event pair around every launch (about 4.5 us each here).
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.rhs(q, out, precision=args.precision)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps

    adjoint = None
    if not mixed and args.adjoint_steps > 0:
        # VJP of the same RHS, timed the same way (all launches of one
        # dg_rhs_vjp); cotangent = the forward output
        lam = [t.clone() for t in out]
        qv = q if w.equation == "euler" else None
        qbar = plan.empty_fields()
        for _ in range(5):
            plan.rhs_vjp(qv, lam, qbar)
        torch.cuda.synchronize()
        va = torch.cuda.Event(enable_timing=True)
        vb = torch.cuda.Event(enable_timing=True)
        va.record(stream)
        for _ in range(args.adjoint_steps):
            plan.rhs_vjp(qv, lam, qbar)
        vb.record(stream)
        torch.cuda.synchronize()
        vjp_ms = va.elapsed_time(vb) / args.adjoint_steps
        adjoint = {"vjp_us": vjp_ms * 1e3, "fwd_us": kernel_ms * 1e3,
                   "ratio": vjp_ms / kernel_ms, "steps": args.adjoint_steps,
                   "kernels": ("mass_apply_kernel + rhs_vjp_mfma_kernel" if w.equation == "euler"
                               else "mass_apply_kernel (R M^-1) + rhs_mfma_kernel with R M fused"
                               if plan.kernel_variant == "mfma"
                               else "mass_apply_kernel (R M^-1) + rhs_lds_kernel "
                                    "+ mass_apply_kernel (R M)")}
        del lam, qbar

    if args.copy_probe:
        # 1 GiB: larger than the 256 MB last-level cache, so the copy's reads
        # come from HBM (the calibration of the FETCH_SIZE counter)
        n = 1024 * 1024 * 1024 // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {n * 8} bytes read + written x3", file=sys.stderr)
        del src, dst

    b_alg = algorithmic_bytes(k, npn, nfp, nf, state_bytes=4 if mixed else 8)
    f_alg = algorithmic_flops(w.equation, k, npn, nfp, nf)
    # the reference's perf_analysis counting rule on this mesh
    rule_flops = float64_flops_on_mesh(w.equation, box_mesh_counts(w.order, w.nel_1d))
    t_kernel = kernel_ms * 1e-3
    achieved = b_alg / t_kernel / 1e9
    t_roof = max(b_alg / (HBM_PEAK_GBS * 1e9), f_alg / (FP64_PEAK_TFLOPS * 1e12))
    ms_per_step = wall / args.steps * 1e3

    kernel_name = (f"rhs_mixed_kernel<{w.equation},{args.precision},P{w.order}>" if mixed else
                   f"rhs_{plan.kernel_variant}_kernel<{w.equation},P{w.order}>")
    result = {
        "metric": "DG RHS evals/s",
        "value": args.steps / wall,
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not mixed else ("f32" if args.precision == "fp32" else "bf16x3/f32"),
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "equation": w.equation,
            "order": w.order,
            "mesh": f"[-1,1]^3, {w.nel_1d}^3 hex x 6 tets",
            "nelements": k,
            "dofs_per_field": k * npn,
            "fields": nf,
            "precision": args.precision,
            "variant": args.variant,
        },
        "gflops": f_alg / t_kernel / 1e9,
        "gflops_perf_analysis": rule_flops / t_kernel / 1e9,
        "gdofs_per_s": k * npn / t_kernel / 1e9,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(w.name, kernel_name),
            "traffic_source": "profiles/hbm_traffic_fp64.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
            "kernel": kernel_name,
            "kernel_us": kernel_ms * 1e3,
            "algorithmic_bytes": b_alg,
            "algorithmic_flops": f_alg,
            "combined_frac": t_roof / t_kernel,
        },
    }
    if field_err is not None:
        result["mixed_rel_l2_vs_fp64"] = field_err
    if adjoint is not None:
        result["adjoint"] = adjoint
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(w, d, initial_state(w, d), args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
