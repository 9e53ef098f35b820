"""Benchmark: FP64 residual + Jacobian assembly throughput (elements/s) on one MI355X.

Default workload: unit square, 1000 x 1000 order-2 (9-node) Lagrange quads, interior nodes jittered by
U(+-0.2 h) (seed 2), displacements U(+-1e-3) (seed 3), E = 70e9, nu = 0.3, t = 1, 3x3 Gauss points.
One step = one assembly of the residual and every BSR value through the resident plan (write-once
owner-computes kernel), inputs already resident in HBM. Prints ONE JSON line on stdout; progress goes
to stderr.

Roofline: algorithmic bytes per assembly B_alg = 16*nodes (coords) + 16*nodes (states) + 4*N*E (conn)
+ 16*nodes (residual) + 32*blocks (BSR values), divided by the kernel's average duration measured with
HIP events on the launch stream, against 8 TB/s.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-fem-project_amd"))

E_MOD, NU, THICK = 70e9, 0.3, 1.0
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "square1000-p2": dict(kind="square", n=1000, order=2, jitter=0.2, desc="unit square 1000x1000 order-2 quads, "
                          "3x3 Gauss, interior jitter U(+-0.2h) seed 2, u,v ~ U(+-1e-3) seed 3"),
    "annulus200x1250-p4": dict(kind="annulus", nr=200, nt=1250, order=4, desc="annulus r in [1,2], 200 radial x "
                               "1250 circumferential order-4 quads, 5x5 Gauss, smooth field + U(+-1e-5)"),
    "square64-p1": dict(kind="square", n=64, order=1, jitter=0.0, desc="unit square 64x64 order-1 quads, 2x2 Gauss, "
                        "u,v ~ U(+-1e-3) seed 1"),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def make_workload(F, name):
    w = WORKLOADS[name]
    if w["kind"] == "square":
        mesh = F.square_mesh(w["n"], w["order"], jitter=w["jitter"], seed=2)
        seed = 1 if name == "square64-p1" else 3
        u = F.fill_uniform(2 * mesh.num_nodes, -1e-3, 1e-3, seed)
    else:
        mesh = F.annulus_mesh(w["nr"], w["nt"], w["order"])
        u = F.smooth_field(mesh.coords, 1e-2, phase_seed=4, pert=1e-5, pert_seed=5)
    return mesh, u


def algorithmic_bytes(mesh, num_blocks, jac=True):
    nodes = mesh.num_nodes
    b = 16 * nodes + 16 * nodes + 4 * mesh.npe * mesh.num_elements + 16 * nodes
    return b + (32 * num_blocks if jac else 0)


def cpu_baseline(mesh, u, bmap, num_blocks, target_seconds):
    """Oracle (CPU restatement of the reference's assembly) timed on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = O.max_threads(fast=True)
    probe = min(mesh.num_elements, 2000)
    t = O.time_res_jac(mesh, u, E_MOD, NU, THICK, bmap, num_blocks, probe, threads)
    count = int(min(mesh.num_elements, max(probe, probe * target_seconds / max(t, 1e-6))))
    t = O.time_res_jac(mesh, u, E_MOD, NU, THICK, bmap, num_blocks, count, threads)
    t1 = O.time_res_jac(mesh, u, E_MOD, NU, THICK, bmap, num_blocks, min(count, max(200, count // threads)), 1)
    n1 = min(count, max(200, count // threads))
    return {
        "value": count / t,
        "unit": "elements/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {count} elements of the workload, residual + Jacobian (per-column forward tangent, "
                  f"FEKernels.h:317-508 restated), g++ -O3 -march=x86-64-v3 -fopenmp, {threads} threads, "
                  f"{t:.1f} s; 1 thread: {n1 / t1:.4g} elements/s",
        "host": platform.processor() or platform.machine(),
    }


def load_traffic(workload, path_name):
    try:
        with open(TRAFFIC_FILE) as f:
            data = json.load(f)
        entry = data.get(f"{workload}/{path_name}")
        return None if entry is None else float(entry["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="square1000-p2", choices=sorted(WORKLOADS))
    ap.add_argument("--path", default="plan", choices=["plan", "plan-det", "atomic"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X, and the assembly path is sized for "
              "one GPU's share (multi-GPU sharding is described, not run). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import fem_amd as F

    F.load()
    dev = torch.device("cuda:0")
    prec = F.FEM_PREC_FP64 if args.precision == "fp64" else F.FEM_PREC_MIXED
    log(f"building workload {args.workload}")
    mesh, u = make_workload(F, args.workload)
    pat = F.bsr_pattern(mesh)
    log(f"mesh: {mesh.num_elements} elements, {mesh.num_nodes} nodes, {pat.num_blocks} blocks")
    bmap = F.element_bcsr_map(mesh, pat)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    coords, states = d(mesh.coords), d(u)
    res = torch.zeros(2 * mesh.num_nodes, dtype=torch.float64, device=dev)
    vals = torch.zeros(4 * pat.num_blocks, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    plan = None
    if args.path.startswith("plan"):
        t0 = time.time()
        plan = F.Plan(mesh, pat, flags=F.FEM_PLAN_DETERMINISTIC if args.path == "plan-det" else 0)
        info = plan.info()
        log(f"plan built in {time.time() - t0:.1f} s: {info}")

        def step():
            plan.assemble(states, coords, E_MOD, NU, THICK, res=res, values=vals, precision=prec, stream=stream)
    else:
        conn, conn_ptr, dmap = d(mesh.conn), d(mesh.conn_ptr), d(bmap)
        info = {}

        def step():
            res.zero_()
            vals.zero_()
            F.jacobian(mesh.order, conn_ptr, conn, mesh.num_elements, states, coords, E_MOD, NU, THICK, dmap, res,
                       vals, precision=prec, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)]) * 1e-3
    ms_per_step = wall / args.steps * 1e3
    value = mesh.num_elements * args.steps / wall
    kernel_s = float(per.mean())
    b_alg = algorithmic_bytes(mesh, pat.num_blocks)
    achieved = b_alg / kernel_s / 1e9
    log(f"{ms_per_step:.4f} ms/step wall, kernel avg {kernel_s * 1e3:.4f} ms (min {per.min() * 1e3:.4f}), "
        f"{achieved:.0f} GB/s algorithmic")

    check = None
    if not args.no_check:
        # Size-independent properties on the full workload: R = K u, and the plan agrees with the
        # element-parallel atomic path (a different kernel and summation order).
        y = torch.zeros_like(res)
        F.bsr_spmv(d(pat.rowp), d(pat.cols), vals, states, y)
        r_ku = float((y - res).abs().max() / res.abs().max())
        check = {"R_vs_Ku": r_ku}
        if plan is not None:
            res2 = torch.zeros_like(res)
            vals2 = torch.zeros_like(vals)
            F.jacobian(mesh.order, d(mesh.conn_ptr), d(mesh.conn), mesh.num_elements, states, coords, E_MOD, NU,
                       THICK, d(bmap), res2, vals2, precision=F.FEM_PREC_FP64)
            torch.cuda.synchronize()
            check["values_vs_atomic_path"] = float((vals2 - vals).abs().max() / vals2.abs().max())
            check["residual_vs_atomic_path"] = float((res2 - res).abs().max() / res2.abs().max())
            del res2, vals2
        log(f"checks: {check}")

    cpu = None
    if not args.no_cpu_baseline:
        log("timing CPU baseline")
        cpu = cpu_baseline(mesh, u, bmap, pat.num_blocks, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.4g} elements/s on {cpu['cores']} threads")

    traffic = load_traffic(args.workload, args.path)
    out = {
        "metric": "res+Jac elements/s & % HBM roofline",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == F.FEM_PREC_FP64 else "f32+f64acc",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]['desc']}",
                   "elements": mesh.num_elements, "nodes": mesh.num_nodes, "bsr_blocks": pat.num_blocks,
                   "path": args.path, "precision": args.precision, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "owner_kernel" if plan is not None else "element_kernel",
                     "algorithmic_bytes": b_alg, "kernel_ms": kernel_s * 1e3},
        "cpu_baseline": cpu,
        "checks": check,
        "plan": info,
    }
    print(json.dumps(out), flush=True)
    if plan is not None:
        plan.close()


if __name__ == "__main__":
    main()
