"""GPU: the serving loop (fem_server_*) -- resident meshes, budget admission, continuous batching.

Checks: each served result against the oracle (FP64 server: 1e-12; mixed-precision server: 1e-5);
a request gives the same result alone and inside a burst of other requests (1e-14 normwise: sums of a
part use LDS atomics, so equality is to rounding); requests whose outputs exceed the budget wait until
earlier results are released; oversize meshes and requests are rejected with FEM_ERR_BUDGET; resident
meshes + outputs + staging stay inside the budget during a burst; destroying a server with queued and
unreleased requests returns; meshes added while requests run.
"""

import numpy as np
import pytest

import textbook

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0


@pytest.fixture(scope="module")
def pool(fem, orc):
    out = []
    specs = [("square", 1, 12), ("square", 2, 8), ("annulus", 3, 3), ("square", 4, 3), ("annulus", 2, 4)]
    for i, (kind, p, n) in enumerate(specs):
        m = (fem.square_mesh(n, p, jitter=0.2, seed=20 + i) if kind == "square"
             else fem.annulus_mesh(n, 4 * n, p, jitter=0.1, seed=20 + i))
        pat = fem.bsr_pattern(m)
        bmap = fem.element_bcsr_map(m, pat)
        states = [fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 1000 * i + k) for k in range(3)]
        refs = [orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks) for u in states]
        out.append(dict(mesh=m, pat=pat, states=states, refs=refs))
    return out


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)], ids=["fp64", "mixed"])
def test_server_results_match_oracle(fem, torch_gpu, pool, precision, tol):
    srv = fem.Server(budget_bytes=4 << 30, max_batch_jobs=16, slots=2, precision=precision, graphs=True)
    ids = [srv.add_mesh(c["mesh"], E, NU, T) for c in pool]
    dstates = [[_dev(torch_gpu, u) for u in c["states"]] for c in pool]
    torch_gpu.cuda.synchronize()
    reqs = []
    for rep in range(3):
        for i, c in enumerate(pool):
            with_jac = (rep + i) % 2 == 0
            reqs.append((srv.submit(ids[i], dstates[i][rep], with_jac), i, rep, with_jac))
    for rid, i, rep, with_jac in reqs:
        c = pool[i]
        res, vals = srv.read(rid, c["mesh"].num_nodes, c["pat"].num_blocks if with_jac else 0)
        ref_res, ref_vals = c["refs"][rep]
        assert textbook.normwise_rel(res, ref_res) < tol
        if with_jac:
            assert textbook.normwise_rel(vals, ref_vals) < tol
        t = srv.wait(rid)
        assert t[0] <= t[1] <= t[2]
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 1
    srv.close()


def test_same_alone_and_in_the_mix(fem, torch_gpu, pool):
    srv = fem.Server(budget_bytes=4 << 30, max_batch_jobs=16, slots=3, precision=1, graphs=True)
    ids = [srv.add_mesh(c["mesh"], E, NU, T) for c in pool]
    dstates = [[_dev(torch_gpu, u) for u in c["states"]] for c in pool]
    torch_gpu.cuda.synchronize()
    c = pool[1]
    # alone
    rid = srv.submit(ids[1], dstates[1][0], True)
    alone = srv.read(rid, c["mesh"].num_nodes, c["pat"].num_blocks)
    srv.release(rid)
    # in a burst of 30 other requests on every mesh
    burst = []
    for k in range(30):
        j = k % len(pool)
        burst.append(srv.submit(ids[j], dstates[j][k % 3], k % 2 == 0))
        if k == 13:
            rid = srv.submit(ids[1], dstates[1][0], True)
    mixed = srv.read(rid, c["mesh"].num_nodes, c["pat"].num_blocks)
    assert textbook.normwise_rel(mixed[0], alone[0]) < 1e-14
    assert textbook.normwise_rel(mixed[1], alone[1]) < 1e-14
    for b in burst:
        srv.wait(b)
        srv.release(b)
    srv.release(rid)
    assert srv.stats()["batches"] >= 2
    srv.close()


def test_budget_admission_and_rejection(fem, torch_gpu, pool):
    c = pool[1]
    m = c["mesh"]
    probe = fem.Server(budget_bytes=1 << 40, max_batch_jobs=4, slots=1, precision=1)
    mid = probe.add_mesh(m, E, NU, T)
    info = probe.mesh_info(mid)
    probe.close()
    res_bytes = 16 * info["nodes"]
    val_bytes = 32 * info["blocks"]
    out_bytes = res_bytes + val_bytes
    # a second resident mesh that does not fit is rejected
    small = fem.Server(budget_bytes=info["resident_bytes"] * 3 // 2, max_batch_jobs=4, slots=1, precision=1)
    small.add_mesh(m, E, NU, T)
    with pytest.raises(fem.FemError) as ei:
        small.add_mesh(m, E, NU, T)
    assert ei.value.status == -9
    small.close()
    # room for the mesh and exactly three requests' outputs
    srv = fem.Server(budget_bytes=info["resident_bytes"] + 3 * out_bytes + out_bytes // 2, max_batch_jobs=4,
                     slots=2, precision=1)
    mid = srv.add_mesh(m, E, NU, T)
    u = _dev(torch_gpu, c["states"][0])
    torch_gpu.cuda.synchronize()
    rids = [srv.submit(mid, u, True) for _ in range(6)]
    first = rids[:3]
    for r in first:
        srv.wait(r)
    # the other three cannot be admitted until outputs are released
    assert not any(srv.poll(r) for r in rids[3:])
    assert srv.stats()["bytes_in_use"] <= info["resident_bytes"] + 3 * out_bytes + out_bytes // 2
    for r in first:
        srv.release(r)
    for r in rids[3:]:
        res, vals = srv.read(r, m.num_nodes, c["pat"].num_blocks)
        assert textbook.normwise_rel(vals, c["refs"][0][1]) < 1e-5
        srv.release(r)
    srv.close()
    # a request whose outputs can never fit is rejected at submission; the residual alone fits
    tiny = fem.Server(budget_bytes=info["resident_bytes"] + res_bytes + val_bytes // 2, max_batch_jobs=4, slots=1,
                      precision=1)
    mid = tiny.add_mesh(m, E, NU, T)
    with pytest.raises(fem.FemError) as ei:
        tiny.submit(mid, u, True)
    assert ei.value.status == -9
    rid = tiny.submit(mid, u, False)
    res, _ = tiny.read(rid, m.num_nodes)
    assert textbook.normwise_rel(res, c["refs"][0][0]) < 1e-5
    tiny.release(rid)
    tiny.close()


def test_graph_batches_replay(fem, torch_gpu, pool):
    """A graph-mode batch replayed with changing compositions matches eager per-plan assembly."""
    torch = torch_gpu
    plans = [fem.Plan(c["mesh"], c["pat"]) for c in pool]
    coords = [_dev(torch, c["mesh"].coords) for c in pool]
    states = [[_dev(torch, u) for u in c["states"]] for c in pool]
    batch = fem.Batch(max_jobs=8, graph=True)
    for rep in range(3):
        batch.clear()
        outs = []
        for i, c in enumerate(pool):
            if (i + rep) % 3 == 2:
                continue  # composition changes between replays
            res = torch.zeros(2 * c["mesh"].num_nodes, dtype=torch.float64, device="cuda:0")
            vals = torch.zeros(4 * c["pat"].num_blocks, dtype=torch.float64, device="cuda:0")
            batch.add(plans[i], states[i][rep], coords[i], E, NU, T, res=res, values=vals)
            outs.append((i, res, vals))
        batch.assemble(precision=0)
        torch.cuda.synchronize()
        for i, res, vals in outs:
            er = torch.zeros_like(res)
            ev = torch.zeros_like(vals)
            plans[i].assemble(states[i][rep], coords[i], E, NU, T, res=er, values=ev)
            torch.cuda.synchronize()
            assert textbook.normwise_rel(res.cpu().numpy(), er.cpu().numpy()) < 1e-14
            assert textbook.normwise_rel(vals.cpu().numpy(), ev.cpu().numpy()) < 1e-14
            assert textbook.normwise_rel(vals.cpu().numpy(), pool[i]["refs"][rep][1]) < 1e-12
    batch.close()
    for p in plans:
        p.close()


def test_memory_within_budget_during_burst(fem, torch_gpu, pool):
    """Resident meshes + outputs + element-once staging stay inside the budget: the server's own accounting, and
    the device's free memory seen from outside while a burst (orders 1-4, Jacobians with staging) is served:
    after creation and mesh admission, serving takes nothing beyond the arena (the runtime's streams, queues and
    recorded graphs are allocated at creation, outside the budget)."""
    torch = torch_gpu
    dstates = [[_dev(torch, u) for u in c["states"]] for c in pool]
    torch.cuda.synchronize()
    budget = 1 << 30
    srv = fem.Server(budget_bytes=budget, max_batch_jobs=8, slots=3, precision=0, graphs=True)
    ids = [srv.add_mesh(c["mesh"], E, NU, T) for c in pool]
    # created and meshes resident: from here on, serving may take no device memory beyond the arena
    free0 = free_min = torch.cuda.mem_get_info()[0]
    rids = []
    for k in range(40):
        i = k % len(pool)
        rids.append((srv.submit(ids[i], dstates[i][k % 3], True), i, k % 3))
    for rid, i, rep in rids:
        free_min = min(free_min, torch.cuda.mem_get_info()[0])
        mem = srv.memory()
        assert mem["resident"] + mem["arena"] <= budget
        assert mem["arena_in_use"] <= mem["arena"] and mem["arena_peak"] <= mem["arena"]
        c = pool[i]
        res, vals = srv.read(rid, c["mesh"].num_nodes, c["pat"].num_blocks)
        assert textbook.normwise_rel(vals, c["refs"][rep][1]) < 1e-12
        assert textbook.normwise_rel(res, c["refs"][rep][0]) < 1e-12
        srv.release(rid)
    mem = srv.memory()
    assert mem["arena_in_use"] == 0 and mem["arena_peak"] > 0
    used = free0 - free_min
    # the arena (created at the first admission), up to the allocator's 2 MB rounding
    assert used <= mem["arena"] + (2 << 20), (used, mem)
    srv.close()


def test_destroy_returns_with_queued_and_unreleased_requests(fem, torch_gpu, pool):
    """Destroying a server whose queued requests wait for memory held by unreleased results returns: the queued
    requests fail and the batches in flight drain."""
    import threading

    c = pool[1]
    m = c["mesh"]
    probe = fem.Server(budget_bytes=1 << 40, max_batch_jobs=4, slots=1, precision=1)
    info = probe.mesh_info(probe.add_mesh(m, E, NU, T))
    probe.close()
    out_bytes = 16 * info["nodes"] + 32 * info["blocks"]
    srv = fem.Server(budget_bytes=info["resident_bytes"] + 2 * out_bytes + out_bytes // 2, max_batch_jobs=4,
                     slots=2, precision=1)
    mid = srv.add_mesh(m, E, NU, T)
    u = _dev(torch_gpu, c["states"][0])
    torch_gpu.cuda.synchronize()
    rids = [srv.submit(mid, u, True) for _ in range(5)]
    srv.wait(rids[0])
    srv.wait(rids[1])
    assert not srv.poll(rids[2])  # waits for memory that only a release would free
    done = threading.Event()
    th = threading.Thread(target=lambda: (srv.close(), done.set()), daemon=True)
    th.start()
    assert done.wait(30), "fem_server_destroy did not return"


def test_add_mesh_while_serving(fem, orc, torch_gpu, pool):
    """Meshes added while a burst is in flight: every result still matches the oracle."""
    torch = torch_gpu
    srv = fem.Server(budget_bytes=4 << 30, max_batch_jobs=8, slots=3, precision=0, graphs=True)
    srv.reserve(1 << 30)  # the other 3 GB stay open to meshes added while requests run
    ids = {0: srv.add_mesh(pool[0]["mesh"], E, NU, T)}
    dstates = [[_dev(torch, u) for u in c["states"]] for c in pool]
    torch.cuda.synchronize()
    rids = []
    for k in range(1, len(pool)):
        for _ in range(6):
            rids.append((srv.submit(ids[0], dstates[0][k % 3], True), 0, k % 3))
        ids[k] = srv.add_mesh(pool[k]["mesh"], E, NU, T)  # the mesh table grows while requests run
        rids.append((srv.submit(ids[k], dstates[k][0], True), k, 0))
    for rid, i, rep in rids:
        c = pool[i]
        res, vals = srv.read(rid, c["mesh"].num_nodes, c["pat"].num_blocks)
        assert textbook.normwise_rel(vals, c["refs"][rep][1]) < 1e-12
        assert textbook.normwise_rel(res, c["refs"][rep][0]) < 1e-12
        srv.release(rid)
    srv.close()
