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
import math
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
    "square500-p3": dict(kind="square", n=500, order=3, jitter=0.2, desc="unit square 500x500 order-3 quads, "
                         "4x4 Gauss, interior jitter U(+-0.2h) seed 2, u,v ~ U(+-1e-3) seed 3"),
}
POOL_WORKLOADS = {
    "batch256-mixed": "256 independent meshes (parameter seed 6, mesh i seed 6e6+i): order U{1..4}, elements "
                      "logU[1e3, 2e5], square or annulus 50/50, interior jitter +-20%, u,v ~ U(+-1e-3); mixed "
                      "precision, one persistent launch per order/output group, plans autotuned per mesh",
    "serve-poisson": "serving: Poisson arrivals over a pool of resident meshes (64 by default) (parameter seed 7): order "
                     "U{1..4}, elements logU[1e3, 1e6], square or annulus; each request residual-only or "
                     "residual+Jacobian with 50/50 odds; 200 GB HBM budget; mixed-precision served path vs "
                     "plain FP64 per-request launches",
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


def plan_kernel_name(path, order):
    """The kernel(s) a plan assembly launches: owner-computes parts (orders 1-2), or (orders 3-4) one
    element-once traversal that stores exclusive blocks and stages shared ones, then the gathers of the
    staged sums."""
    if path == "plan-owner" and order >= 3:
        return "owner_stream_kernel (FP64 row stage; mixed: FP32 MFMA stage)"
    if order >= 3:
        return "element_tiled_kernel+gather_shared"
    if path == "plan-det":
        return "owner_kernel"
    return "owner_stream_kernel"


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

# ---- mesh pools: the batched and served workloads ------------------------------------------------------
# Blocks per element of large meshes by order, and the reference's element->BSR map limit: its int32
# entries hold block offsets times 4, so a mesh served by the plain (map-based) path has < 2^31 / 4 blocks.
BLOCKS_PER_ELEMENT = {1: 9.1, 2: 64.1, 3: 225.5, 4: 577.0}
MAP_BLOCK_LIMIT = (2 ** 31 - 1) // 4


def mesh_pool(F, count, seed, e_max):
    """count meshes drawn from SplitMix64(seed): order U{1..4}, elements logU[1e3, e_max], square (n = round
    sqrt E) or annulus (n_r = round sqrt(E / 6.25), n_theta = round(E / n_r) made even) with 50/50 odds, interior
    jitter +-20% of the smaller adjacent knot gap (mesh seed = seed * 1e6 + i). Element counts are capped so the
    reference's int32 map can address every block."""
    u = F.fill_uniform(4 * count, 0.0, 1.0, seed)
    pool = []
    for i in range(count):
        order = 1 + min(3, int(4 * u[4 * i]))
        ne = int(round(math.exp(math.log(1e3) + u[4 * i + 1] * (math.log(e_max) - math.log(1e3)))))
        ne = min(ne, int(0.97 * MAP_BLOCK_LIMIT / BLOCKS_PER_ELEMENT[order]))
        mseed = seed * 10 ** 6 + i
        if u[4 * i + 2] < 0.5:
            n = max(1, round(math.sqrt(ne)))
            mesh = F.square_mesh(n, order, jitter=0.2, seed=mseed)
        else:
            nr = max(1, round(math.sqrt(ne / 6.25)))
            nt = max(4, round(ne / nr))
            nt += nt % 2
            mesh = F.annulus_mesh(nr, nt, order, jitter=0.2, seed=mseed)
        pool.append(mesh)
    return pool


def rel_err(torch, a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def bench_batch(args, F, torch):
    """256 meshes assembled by one batch call (one persistent launch per order/output group, graph replay)
    against one plan launch per mesh and against plain FP64 per-mesh atomic launches."""
    dev = torch.device("cuda:0")
    count = args.pool or 256
    prec = F.FEM_PREC_FP64 if args.precision == "fp64" else F.FEM_PREC_MIXED
    log(f"building {count} meshes")
    meshes = mesh_pool(F, count, 6, 2e5)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    items = []
    for i, m in enumerate(meshes):
        pat = F.bsr_pattern(m)
        u = F.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 6 * 10 ** 6 + i)
        it = dict(mesh=m, blocks=pat.num_blocks, plan=F.Plan(m, pat, flags=F.FEM_PLAN_AUTOTUNE), untuned=F.Plan(m, pat),
                  coords=d(m.coords), states=d(u),
                  res=torch.zeros(2 * m.num_nodes, dtype=torch.float64, device=dev),
                  vals=torch.zeros(4 * pat.num_blocks, dtype=torch.float64, device=dev),
                  conn=d(m.conn), conn_ptr=d(m.conn_ptr), map=d(F.element_bcsr_map(m, pat)))
        items.append(it)
        if (i + 1) % 32 == 0:
            log(f"  {i + 1} meshes ready")
    total_el = sum(m.num_elements for m in meshes)
    b_alg = sum(algorithmic_bytes(it["mesh"], it["blocks"]) for it in items)
    by_order = {p: sum(1 for m in meshes if m.order == p) for p in range(1, 5)}
    log(f"{total_el} elements in total, orders {by_order}")
    stream = torch.cuda.current_stream()
    graph = True
    batch = F.Batch(max_jobs=count, graph=True)
    try:
        for it in items:
            batch.add(it["plan"], it["states"], it["coords"], E_MOD, NU, THICK, res=it["res"], values=it["vals"])
    except F.FemError as e:
        log(f"graph batch unavailable ({e}); using launches")
        batch.close()
        graph = False
        batch = F.Batch(max_jobs=count, graph=False)
        for it in items:
            batch.add(it["plan"], it["states"], it["coords"], E_MOD, NU, THICK, res=it["res"], values=it["vals"])

    def time_steps(step, steps):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_batch = time_steps(lambda: batch.assemble(precision=prec, stream=stream), args.steps)
    log(f"batch (autotuned plans): {t_batch * 1e3:.3f} ms per assembly of all {count} meshes")
    scales, waves = {}, {}
    for it in items:
        inf = it["plan"].info()
        key = str(it["mesh"].order)
        scales.setdefault(key, {}).setdefault(str(inf["part_scale"]), 0)
        scales[key][str(inf["part_scale"])] += 1
        waves.setdefault(key, {}).setdefault(str(inf["waves_per_group"]), 0)
        waves[key][str(inf["waves_per_group"])] += 1
    log(f"element-once waves per workgroup chosen per order (plan assembly): {waves}")
    log(f"part scales chosen per order: {scales}")
    ubatch = F.Batch(max_jobs=count, graph=graph)
    for it in items:
        ubatch.add(it["untuned"], it["states"], it["coords"], E_MOD, NU, THICK, res=it["res"], values=it["vals"])
    t_untuned = time_steps(lambda: ubatch.assemble(precision=prec, stream=stream), args.steps)
    ubatch.close()
    log(f"batch (untuned plans): {t_untuned * 1e3:.3f} ms")

    def per_mesh_plans():
        for it in items:
            it["plan"].assemble(it["states"], it["coords"], E_MOD, NU, THICK, res=it["res"], values=it["vals"],
                                precision=prec, stream=stream)

    t_plans = time_steps(per_mesh_plans, max(1, args.steps // 2))
    log(f"one plan launch per mesh: {t_plans * 1e3:.3f} ms")

    def per_mesh_plain():
        for it in items:
            m = it["mesh"]
            it["res"].zero_()
            it["vals"].zero_()
            F.jacobian(m.order, it["conn_ptr"], it["conn"], m.num_elements, it["states"], it["coords"], E_MOD, NU,
                       THICK, it["map"], it["res"], it["vals"], precision=F.FEM_PREC_FP64, stream=stream)

    t_plain = time_steps(per_mesh_plain, max(1, args.steps // 5))
    log(f"plain FP64 atomic launch per mesh: {t_plain * 1e3:.3f} ms")

    check = None
    if not args.no_check:
        # Every mesh: the batch result against the plain FP64 result (the element-parallel atomic path).
        batch.assemble(precision=prec, stream=stream)
        torch.cuda.synchronize()
        got = [(it["res"].clone(), it["vals"].clone()) for it in items]
        per_mesh_plain()
        torch.cuda.synchronize()
        er = max(rel_err(torch, g[0], it["res"]) for g, it in zip(got, items))
        ev = max(rel_err(torch, g[1], it["vals"]) for g, it in zip(got, items))
        check = {"max_rel_err_residual_vs_fp64": er, "max_rel_err_values_vs_fp64": ev}
        log(f"checks: {check}")
        del got
    value = total_el / t_batch
    out = {
        "metric": "batched res+Jac elements/s",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == F.FEM_PREC_FP64 else "f32+f64acc",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {POOL_WORKLOADS[args.workload]}", "meshes": count,
                   "elements": total_el, "orders": by_order, "graph": graph, "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": b_alg / t_batch / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b_alg / t_batch / 1e9 / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": b_alg},
        "baselines": {"one_plan_launch_per_mesh_ms": t_plans * 1e3, "plain_fp64_atomic_per_mesh_ms": t_plain * 1e3,
                      "speedup_vs_plain": t_plain / t_batch, "untuned_plans_batch_ms": t_untuned * 1e3},
        "autotune": {"part_scales_by_order": scales, "part_scale_candidates": [1.0, 0.5, 0.25],
                     "waves_per_group_by_order": waves, "waves_candidates_orders_3_4": [1, 2, 4]},
        "cpu_baseline": None,
        "checks": check,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    for it in items:
        it["plan"].close()
        it["untuned"].close()


def request_stream(F, n, seed, pool_size):
    """Requests drawn from SplitMix64(seed): mesh U{0..pool-1}, residual+Jacobian with odds 1/2, and a unit-rate
    exponential inter-arrival gap (scaled by 1/lambda by the caller)."""
    u = F.fill_uniform(3 * n, 0.0, 1.0, seed)
    mesh = np.minimum((u[0::3] * pool_size).astype(np.int64), pool_size - 1)
    jac = u[1::3] < 0.5
    gaps = -np.log1p(-u[2::3])
    return mesh, jac, np.cumsum(gaps)


def wait_until(t):
    while True:
        d = t - time.monotonic()
        if d <= 0:
            return
        if d > 2e-3:
            time.sleep(d - 1e-3)


def bench_serve(args, F, torch):
    """Poisson arrivals over 64 resident meshes: the serving loop (mixed precision, continuous batching, graph
    replay) against plain FP64 per-request launches (FIFO, one request at a time, residual and Jacobian as
    separate element-parallel atomic launches). Latency = completion - scheduled arrival."""
    import threading
    import queue

    dev = torch.device("cuda:0")
    pool_size = args.pool or 64
    log(f"building {pool_size} meshes")
    meshes = mesh_pool(F, pool_size, 7, 1e6)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    srv = F.Server(budget_bytes=200 * 10 ** 9, max_batch_jobs=64, slots=3, precision=F.FEM_PREC_MIXED,
                   graphs=True)
    plain = []
    max_nodes = max_blocks = 0
    for i, m in enumerate(meshes):
        pat = F.bsr_pattern(m)
        mid = srv.add_mesh(m, E_MOD, NU, THICK)
        g = torch.Generator(device=dev)
        g.manual_seed(7 * 10 ** 6 + i)
        states = [torch.empty(2 * m.num_nodes, dtype=torch.float64, device=dev).uniform_(-1e-3, 1e-3, generator=g)
                  for _ in range(2)]
        plain.append(dict(mesh=m, mid=mid, blocks=pat.num_blocks, coords=d(m.coords), conn=d(m.conn),
                          conn_ptr=d(m.conn_ptr), map=d(F.element_bcsr_map(m, pat)), states=states))
        max_nodes = max(max_nodes, m.num_nodes)
        max_blocks = max(max_blocks, pat.num_blocks)
        if (i + 1) % 8 == 0:
            log(f"  {i + 1} meshes resident ({srv.stats()['bytes_in_use'] / 1e9:.1f} GB in the server)")
    orders = {p: sum(1 for m in meshes if m.order == p) for p in range(1, 5)}
    sizes = sorted(m.num_elements for m in meshes)
    log(f"pool: orders {orders}, elements min {sizes[0]} median {sizes[len(sizes) // 2]} max {sizes[-1]}")
    res_buf = torch.zeros(2 * max_nodes, dtype=torch.float64, device=dev)
    val_buf = torch.zeros(4 * max_blocks, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()

    def plain_one(k, jac, state):
        p = plain[k]
        m = p["mesh"]
        res = res_buf[:2 * m.num_nodes]
        res.zero_()
        F.residual(m.order, p["conn_ptr"], p["conn"], m.num_elements, p["states"][state], p["coords"], E_MOD, NU,
                   THICK, res, stream=stream)
        if jac:
            vals = val_buf[:4 * p["blocks"]]
            vals.zero_()
            F.jacobian(m.order, p["conn_ptr"], p["conn"], m.num_elements, p["states"][state], p["coords"], E_MOD,
                       NU, THICK, p["map"], None, vals, stream=stream)
        stream.synchronize()

    def run_plain(mesh_ids, jacs, arrivals):
        n = len(mesh_ids)
        lat = np.zeros(n)
        t0 = time.monotonic() + 0.05
        for i in range(n):
            target = t0 + arrivals[i]
            wait_until(target)
            plain_one(int(mesh_ids[i]), bool(jacs[i]), i % 2)
            lat[i] = time.monotonic() - target
            if (i + 1) % 2000 == 0:
                log(f"    plain {i + 1}/{n}")
        return lat, time.monotonic() - t0

    def run_served(mesh_ids, jacs, arrivals):
        n = len(mesh_ids)
        lat = np.zeros(n)
        q = queue.Queue()
        t0 = time.monotonic() + 0.05
        errors = []

        def reaper():
            try:
                for i in range(n):
                    rid = q.get()
                    times = srv.wait(rid)
                    lat[i] = times[2] - (t0 + arrivals[i])
                    srv.release(rid)
            except Exception as e:  # reported below
                errors.append(e)

        th = threading.Thread(target=reaper)
        th.start()
        for i in range(n):
            wait_until(t0 + arrivals[i])
            q.put(srv.submit(plain[int(mesh_ids[i])]["mid"], plain[int(mesh_ids[i])]["states"][i % 2],
                             bool(jacs[i])))
            if (i + 1) % 2000 == 0:
                log(f"    served {i + 1}/{n}")
        th.join()
        if errors:
            raise errors[0]
        return lat, time.monotonic() - t0

    n_main = args.requests
    mesh_ids, jacs, unit_arrivals = request_stream(F, max(n_main, args.side_requests), 7, pool_size)
    elems = np.array([m.num_elements for m in meshes])

    # Warm-up: every mesh once each way on both paths (graph capture, output-pool fill).
    for k in range(pool_size):
        for jac in (False, True):
            plain_one(k, jac, 0)
    warm = np.zeros(2 * pool_size)
    run_served(np.repeat(np.arange(pool_size), 2), np.tile([False, True], pool_size), warm)

    # Saturated throughput: the first S requests all present at t = 0, three runs of each path.
    sat = min(n_main, 1000)
    zero = np.zeros(sat)
    t_plain_runs, t_served_runs = [], []
    for _ in range(3):
        t_plain_runs.append(run_plain(mesh_ids[:sat], jacs[:sat], zero)[1])
        t_served_runs.append(run_served(mesh_ids[:sat], jacs[:sat], zero)[1])
    t_plain_sat, t_served_sat = float(np.median(t_plain_runs)), float(np.median(t_served_runs))
    sat_el = float(elems[mesh_ids[:sat]].sum())
    plain_rate = sat / t_plain_sat
    log(f"saturated: plain {[round(sat / t, 1) for t in t_plain_runs]} req/s, served "
        f"{[round(sat / t, 1) for t in t_served_runs]} req/s")

    loads = {}
    for load in args.serve_loads:
        n = n_main if load == 0.8 else args.side_requests
        if n <= 0:
            continue
        lam = load * plain_rate
        arr = unit_arrivals[:n] / lam
        log(f"load {load}: {n} requests at {lam:.1f} req/s")
        lp, tp = run_plain(mesh_ids[:n], jacs[:n], arr)
        ls, ts = run_served(mesh_ids[:n], jacs[:n], arr)
        el = float(elems[mesh_ids[:n]].sum())
        loads[str(load)] = {
            "requests": n, "arrival_rate_per_s": lam,
            "plain": {"p50_ms": float(np.percentile(lp, 50) * 1e3), "p99_ms": float(np.percentile(lp, 99) * 1e3),
                      "throughput_req_s": n / tp, "elements_per_s": el / tp},
            "served": {"p50_ms": float(np.percentile(ls, 50) * 1e3), "p99_ms": float(np.percentile(ls, 99) * 1e3),
                       "throughput_req_s": n / ts, "elements_per_s": el / ts},
        }
        log(f"  {loads[str(load)]}")

    check = None
    if not args.no_check:
        # Served (mixed) against plain FP64 for one residual+Jacobian request on each mesh whose BSR values
        # fit 640 MB (the comparison copies them to the host).
        er = ev = 0.0
        checked = 0
        for k, p in enumerate(plain):
            if p["blocks"] > 20_000_000:
                continue
            checked += 1
            rid = srv.submit(p["mid"], p["states"][0], True)
            r_host, v_host = srv.read(rid, p["mesh"].num_nodes, p["blocks"])
            srv.release(rid)
            plain_one(k, True, 0)
            r_ref = res_buf[:2 * p["mesh"].num_nodes]
            v_ref = val_buf[:4 * p["blocks"]]
            er = max(er, rel_err(torch, torch.from_numpy(r_host).to(dev), r_ref))
            ev = max(ev, rel_err(torch, torch.from_numpy(v_host).to(dev), v_ref))
        check = {"max_rel_err_residual_vs_fp64": er, "max_rel_err_values_vs_fp64": ev, "meshes_checked": checked}
        log(f"checks: {check}")
    st = srv.stats()
    main_load = loads["0.8"]
    out = {
        "metric": "served p50/p99 latency vs plain",
        "value": main_load["served"]["p99_ms"],
        "unit": "ms (p99 latency at 0.8x the plain path's saturated rate)",
        "n_gpus": 1,
        "steps": n_main,
        "warmup": 2 * pool_size,
        "ms_per_step": main_load["served"]["p50_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64acc",
        "data": "synthetic",
        "config": {"workload": f"serve-poisson: {POOL_WORKLOADS['serve-poisson']}", "pool": pool_size,
                   "orders": orders, "elements_min_median_max": [sizes[0], sizes[len(sizes) // 2], sizes[-1]],
                   "slots": 3, "max_batch_jobs": 64, "budget_bytes": 200 * 10 ** 9, "parallelism": "single GPU"},
        "saturated": {"requests": sat, "plain_req_s": plain_rate, "served_req_s": sat / t_served_sat,
                      "plain_elements_s": sat_el / t_plain_sat, "served_elements_s": sat_el / t_served_sat,
                      "plain_req_s_runs": [sat / t for t in t_plain_runs],
                      "served_req_s_runs": [sat / t for t in t_served_runs]},
        "loads": loads,
        "server": st,
        "cpu_baseline": None,
        "checks": check,
    }
    print(json.dumps(out), flush=True)
    srv.close()


def run_pool_workload(args, F, torch):
    if args.workload == "batch256-mixed":
        bench_batch(args, F, torch)
    else:
        bench_serve(args, F, torch)


def step_stats(per_s):
    """Per-step times (seconds) as the reference's harness reports them (verification/TestKernels.cpp:37-60):
    min / median / average / standard deviation, here in milliseconds."""
    ms = np.asarray(per_s) * 1e3
    return {"runs": int(ms.size), "min_ms": float(ms.min()), "median_ms": float(np.median(ms)),
            "avg_ms": float(ms.mean()), "std_ms": float(ms.std())}


def time_events(step, steps, warmup, stream, torch):
    """Warm-up, then `steps` timed calls bracketed by HIP events on `stream`; returns (wall seconds, per-step
    event seconds)."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, np.array([s.elapsed_time(e) for s, e in zip(starts, ends)]) * 1e-3


def reference_protocol(step, stream, torch, max_runs, max_seconds, prepare=None):
    """The reference's timing loop (verification/TestKernels.cpp:29-30, 228-287): one discarded run, then up to
    max_runs runs or max_seconds, outputs re-initialised outside the timed call (`prepare`)."""
    if prepare:
        prepare()
    step()
    torch.cuda.synchronize()
    per = []
    t_end = time.perf_counter() + max_seconds
    while len(per) < max_runs and time.perf_counter() < t_end:
        if prepare:
            prepare()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step()
        b.record(stream)
        b.synchronize()
        per.append(a.elapsed_time(b) * 1e-3)
    return step_stats(per)


def run_single(args, F, torch, workload, path, precision, strain="linear", autotune=False, cpu=False,
               ref_protocol=False, check=True):
    """One single-mesh assembly leg: builds the workload, times args.steps assemblies (inputs resident in HBM)
    and returns its record (throughput, per-step statistics, roofline, checks)."""
    dev = torch.device("cuda:0")
    prec = F.FEM_PREC_FP64 if precision == "fp64" else F.FEM_PREC_MIXED
    log(f"leg {workload} / {path} / {precision} / {strain}: building")
    mesh, u = make_workload(F, workload)
    pat = F.bsr_pattern(mesh)
    log(f"mesh: {mesh.num_elements} elements, {mesh.num_nodes} nodes, {pat.num_blocks} blocks")
    bmap = F.element_bcsr_map(mesh, pat)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    coords, states = d(mesh.coords), d(u)
    res = torch.zeros(2 * mesh.num_nodes, dtype=torch.float64, device=dev)
    vals = torch.zeros(4 * pat.num_blocks, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    plan = None
    kernel_only = None
    if path.startswith("plan"):
        t0 = time.time()
        flags = {"plan-det": F.FEM_PLAN_DETERMINISTIC, "plan-owner": F.FEM_PLAN_OWNER_ONLY}.get(path, 0)
        if autotune:
            flags |= F.FEM_PLAN_AUTOTUNE
        plan = F.Plan(mesh, pat, flags=flags)
        info = plan.info()
        log(f"plan built in {time.time() - t0:.1f} s: {info}")

        def step():
            plan.assemble(states, coords, E_MOD, NU, THICK, res=res, values=vals, precision=prec, stream=stream)
    else:
        conn, conn_ptr, dmap = d(mesh.conn), d(mesh.conn_ptr), d(bmap)
        info = {}

        def kernel():
            F.jacobian(mesh.order, conn_ptr, conn, mesh.num_elements, states, coords, E_MOD, NU, THICK, dmap, res,
                       vals, precision=prec, stream=stream, nonlinear=strain == "nonlinear", dropin=path == "dropin")

        def zero():
            res.zero_()
            vals.zero_()

        def step():
            zero()
            kernel()

    wall, per = time_events(step, args.steps, args.warmup, stream, torch)
    ms_per_step = wall / args.steps * 1e3
    kernel_s = float(per.mean())
    b_alg = algorithmic_bytes(mesh, pat.num_blocks)
    achieved = b_alg / kernel_s / 1e9
    log(f"{ms_per_step:.4f} ms/step wall, kernel avg {kernel_s * 1e3:.4f} ms (min {per.min() * 1e3:.4f}), "
        f"{achieved:.0f} GB/s algorithmic")
    protocol = None
    if ref_protocol:
        if plan is None:
            # the reference's number: the kernel alone, outputs zeroed outside the timed call (TestKernels.cpp:237,241)
            protocol = reference_protocol(kernel, stream, torch, args.ref_runs, args.ref_seconds, prepare=zero)
            kernel_only = protocol["avg_ms"]
        else:
            protocol = reference_protocol(step, stream, torch, args.ref_runs, args.ref_seconds)
        log(f"reference protocol: {protocol}")

    checks = None
    if check:
        # Size-independent properties on the full workload: R = K u, and the plan agrees with the
        # element-parallel atomic path (a different kernel and summation order). Full-size parity against the
        # oracle is a test (tests/test_gpu_fullsize.py).
        rowp_d, cols_d = d(pat.rowp), d(pat.cols)
        if strain == "nonlinear":
            # The tangent is the Hessian of the strain energy: symmetric, x^T K z = z^T K x.
            x = torch.rand_like(res) - 0.5
            z = torch.rand_like(res) - 0.5
            kx, kz = torch.zeros_like(res), torch.zeros_like(res)
            F.bsr_spmv(rowp_d, cols_d, vals, x, kx)
            F.bsr_spmv(rowp_d, cols_d, vals, z, kz)
            checks = {"K_symmetry": float((z @ kx - x @ kz).abs() / (x.norm() * kz.norm()))}
        else:
            y = torch.zeros_like(res)
            F.bsr_spmv(rowp_d, cols_d, vals, states, y)
            checks = {"R_vs_Ku": float((y - res).abs().max() / res.abs().max())}
        if plan is not None:
            res2 = torch.zeros_like(res)
            vals2 = torch.zeros_like(vals)
            F.jacobian(mesh.order, d(mesh.conn_ptr), d(mesh.conn), mesh.num_elements, states, coords, E_MOD, NU,
                       THICK, d(bmap), res2, vals2, precision=F.FEM_PREC_FP64)
            torch.cuda.synchronize()
            checks["values_vs_fp64_atomic_path"] = float((vals2 - vals).abs().max() / vals2.abs().max())
            checks["residual_vs_fp64_atomic_path"] = float((res2 - res).abs().max() / res2.abs().max())
            del res2, vals2
        log(f"checks: {checks}")

    cpu_rec = None
    if cpu:
        log("timing CPU baseline")
        cpu_rec = cpu_baseline(mesh, u, bmap, pat.num_blocks, args.cpu_seconds)
        log(f"cpu baseline {cpu_rec['value']:.4g} elements/s on {cpu_rec['cores']} threads")

    traffic = None if autotune else load_traffic(workload, path)
    if precision != "fp64" or strain != "linear":
        traffic = None
    rec = {
        "workload": workload, "path": path + ("+autotune" if autotune else ""), "precision": precision,
        "strain": strain, "elements": mesh.num_elements, "nodes": mesh.num_nodes, "bsr_blocks": pat.num_blocks,
        "elements_per_s": mesh.num_elements * args.steps / wall, "ms_per_step": ms_per_step,
        "steps": step_stats(per),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": plan_kernel_name(path, mesh.order) if plan is not None
                     else ("element_kernel_nl" if strain == "nonlinear" else
                           "owner_stream_kernel (accumulate)" if path == "dropin" and mesh.order <= 2
                           else "element_kernel"),
                     "algorithmic_bytes": b_alg, "kernel_ms": kernel_s * 1e3},
        "reference_protocol": protocol, "checks": checks, "plan": info, "cpu_baseline": cpu_rec,
        "dtype": "f64" if prec == F.FEM_PREC_FP64 else "f32+f64acc",
    }
    if kernel_only is not None:
        rec["kernel_only_elements_per_s"] = mesh.num_elements / (kernel_only * 1e-3)
    if plan is not None:
        plan.close()
    del res, vals, coords, states
    torch.cuda.empty_cache()
    return rec


def serve_leg(args, F, torch):
    """Compact serving leg of the default run: bench_serve on a smaller pool and request count."""
    import contextlib
    import io

    sub = argparse.Namespace(**vars(args))
    sub.pool, sub.requests, sub.side_requests, sub.serve_loads = args.serve_pool, args.serve_requests, 0, (0.8,)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        bench_serve(sub, F, torch)
    rec = json.loads(buf.getvalue().strip().splitlines()[-1])
    keep = ("config", "saturated", "loads", "server", "checks")
    return {k: rec[k] for k in keep}


def batch_leg(args, F, torch):
    """C4 leg of the default run: bench_batch (256 meshes, mixed precision, autotuned plans, graph replay)."""
    import contextlib
    import io

    sub = argparse.Namespace(**vars(args))
    sub.pool, sub.precision, sub.workload = 0, "mixed", "batch256-mixed"
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        bench_batch(sub, F, torch)
    rec = json.loads(buf.getvalue().strip().splitlines()[-1])
    keep = ("value", "unit", "ms_per_step", "config", "roofline", "baselines", "autotune", "checks")
    return {k: rec[k] for k in keep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="square1000-p2", choices=sorted(WORKLOADS) + sorted(POOL_WORKLOADS))
    ap.add_argument("--path", default="plan", choices=["plan", "plan-det", "plan-owner", "atomic", "dropin"],
                    help="plan-owner: orders 3-4 through the owner-computes parts (FEM_PLAN_OWNER_ONLY); atomic: the "
                         "element-parallel atomic kernels (fem_jacobian); dropin: the reference's entry point "
                         "(runJacobianKernel -> fem_jacobian_dropin, a cached plan in accumulate mode at orders 1-2)")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="default: fp64 for single meshes, mixed for batch256-mixed")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="default workload only: skip the extra legs (mixed, drop-in, annulus, serving)")
    ap.add_argument("--ref-runs", type=int, default=300, help="reference-protocol runs per leg (the reference: 1000)")
    ap.add_argument("--ref-seconds", type=float, default=2.0, help="reference-protocol time cap per leg (reference: 30)")
    ap.add_argument("--serve-pool", type=int, default=16, help="serving leg of the default run: resident meshes")
    ap.add_argument("--serve-requests", type=int, default=1500, help="serving leg of the default run: requests")
    ap.add_argument("--requests", type=int, default=10000, help="serve-poisson: requests at the main load")
    ap.add_argument("--side-requests", type=int, default=2000, help="serve-poisson: requests at 0.5 and 0.95 load")
    ap.add_argument("--pool", type=int, default=0, help="override the mesh-pool size (0: the workload's own)")
    ap.add_argument("--autotune", action="store_true", help="plan path: build the plan with FEM_PLAN_AUTOTUNE")
    ap.add_argument("--strain", default="linear", choices=["linear", "nonlinear"],
                    help="nonlinear: the Green-strain tangent (drop-in atomic path, FP64)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device twice (the known-size kernel the FETCH_SIZE / WRITE_SIZE "
                         "counters are scaled by) and exit")
    args = ap.parse_args()
    args.serve_loads = (0.8, 0.5, 0.95)
    if args.copy_calibration:
        import torch

        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 2}), flush=True)
        return
    if args.autotune and args.path != "plan":
        ap.error("--autotune applies to --path plan only")
    if args.strain == "nonlinear" and (args.path != "atomic" or args.precision == "mixed"):
        ap.error("--strain nonlinear runs on --path atomic in FP64")
    if args.precision is None:
        args.precision = "mixed" if args.workload == "batch256-mixed" else "fp64"
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X, and the assembly path is sized for "
              "one GPU's share (multi-GPU sharding is described, not run). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import fem_amd as F

    F.load()
    if args.workload in POOL_WORKLOADS:
        run_pool_workload(args, F, torch)
        return
    default_run = (args.workload == "square1000-p2" and args.path == "plan" and args.precision == "fp64"
                   and args.strain == "linear" and not args.autotune and not args.no_legs)
    main_rec = run_single(args, F, torch, args.workload, args.path, args.precision, strain=args.strain,
                          autotune=args.autotune, cpu=not args.no_cpu_baseline, ref_protocol=True,
                          check=not args.no_check)
    legs = None
    if default_run:
        # The rest of the metric "res+Jac elements/s & % HBM roofline; served p50/p99 latency vs plain":
        # the 1M-element mesh in mixed precision and through the reference's entry point (drop-in atomic), the
        # order-4 annulus in both precisions, and a compact serving run against plain per-request launches.
        legs = {}
        for wl, path, prec in (("square1000-p2", "plan", "mixed"), ("square1000-p2", "atomic", "fp64"),
                               ("square1000-p2", "dropin", "fp64"),
                               ("annulus200x1250-p4", "plan", "fp64"), ("annulus200x1250-p4", "plan", "mixed"),
                               ("annulus200x1250-p4", "atomic", "fp64"),
                               ("annulus200x1250-p4", "plan-owner", "mixed"), ("square64-p1", "plan", "fp64"),
                               ("square64-p1", "atomic", "fp64")):
            r = run_single(args, F, torch, wl, path, prec, ref_protocol=True, check=not args.no_check,
                           cpu=wl == "square64-p1" and not args.no_cpu_baseline)
            legs[f"{wl}/{path}/{prec}"] = r
        log("batch leg")
        legs["batch256-mixed"] = batch_leg(args, F, torch)
        torch.cuda.empty_cache()  # the serving leg's server takes its budget from free device memory
        log("serving leg")
        legs["serving"] = serve_leg(args, F, torch)
    r = main_rec
    out = {
        "metric": "res+Jac elements/s & % HBM roofline",
        "value": r["elements_per_s"],
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": r["dtype"],
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]['desc']}",
                   "elements": r["elements"], "nodes": r["nodes"], "bsr_blocks": r["bsr_blocks"],
                   "path": r["path"], "precision": args.precision, "strain": args.strain,
                   "parallelism": "single GPU"},
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu_baseline"],
        "step_stats": r["steps"],
        "reference_protocol": r["reference_protocol"],
        "checks": r["checks"],
        "plan": r["plan"],
        "legs": legs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
