"""GPU: batches of resident meshes (fem_batch_*) -- one persistent launch per element order and output set.

Each mesh's result in a mixed batch (orders 1..4, square and annulus, residual-only / Jacobian-only /
both) is checked against the oracle at 1e-12 (FP64) and against the same mesh assembled alone. Sums in a
part use LDS atomics, so "alone" and "in the mix" agree to rounding (1e-14 normwise), not bitwise.
"""

import numpy as np
import pytest

import textbook

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0


@pytest.fixture(scope="module")
def meshes(fem, orc):
    out = []
    specs = [("square", 1, 9), ("square", 2, 6), ("annulus", 3, 3), ("square", 4, 3), ("annulus", 2, 4),
             ("square", 1, 5), ("annulus", 4, 2), ("square", 3, 4)]
    for i, (kind, p, n) in enumerate(specs):
        if kind == "square":
            m = fem.square_mesh(n, p, jitter=0.2, seed=10 + i)
        else:
            m = fem.annulus_mesh(n, 4 * n, p, jitter=0.1, seed=10 + i)
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 100 + i)
        pat = fem.bsr_pattern(m)
        bmap = fem.element_bcsr_map(m, pat)
        res, vals = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks)
        out.append(dict(mesh=m, u=u, pat=pat, res=res, vals=vals))
    return out


def test_batch_mixed_orders_and_outputs(fem, torch_gpu, meshes):
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    batch = fem.Batch()
    plans, outs = [], []
    for i, c in enumerate(meshes):
        m = c["mesh"]
        plan = fem.Plan(m, c["pat"])
        coords, states = d(m.coords), d(c["u"])
        mode = i % 3  # 0: both, 1: residual only, 2: values only
        res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0") if mode != 2 else None
        vals = torch.full((4 * c["pat"].num_blocks,), float("nan"), dtype=torch.float64,
                          device="cuda:0") if mode != 1 else None
        batch.add(plan, states, coords, E, NU, T, res=res, values=vals)
        plans.append((plan, coords, states))
        outs.append((res, vals))
    batch.assemble()
    torch.cuda.synchronize()
    for c, (res, vals), (plan, coords, states) in zip(meshes, outs, plans):
        if res is not None:
            r = res.cpu().numpy()
            assert np.isfinite(r).all()
            assert textbook.normwise_rel(r, c["res"]) < 1e-12
            alone = torch.zeros_like(res)
            plan.assemble(states, coords, E, NU, T, res=alone)
            torch.cuda.synchronize()
            assert textbook.normwise_rel(r, alone.cpu().numpy()) < 1e-14
        if vals is not None:
            v = vals.cpu().numpy()
            assert np.isfinite(v).all()
            assert textbook.normwise_rel(v, c["vals"]) < 1e-12
            alone = torch.zeros_like(vals)
            plan.assemble(states, coords, E, NU, T, values=alone)
            torch.cuda.synchronize()
            assert textbook.normwise_rel(v, alone.cpu().numpy()) < 1e-14
    # Reuse: clear and refill with mixed precision.
    batch.clear()
    for c, (res, vals), (plan, coords, states) in zip(meshes, outs, plans):
        batch.add(plan, states, coords, E, NU, T, res=res, values=vals)
    batch.assemble(precision=fem.FEM_PREC_MIXED)
    torch.cuda.synchronize()
    for c, (res, vals), _ in zip(meshes, outs, plans):
        if res is not None:
            assert textbook.normwise_rel(res.cpu().numpy(), c["res"]) < 1e-5
        if vals is not None:
            assert textbook.normwise_rel(vals.cpu().numpy(), c["vals"]) < 1e-5
    batch.close()
    for plan, _, _ in plans:
        plan.close()


def test_batch_deterministic_and_residual_only_plans(fem, torch_gpu, meshes):
    """Batches without graphs run jobs of deterministic and residual-only plans as the plan's own assembly
    after the groups: their outputs equal the plan assembled alone (deterministic plans bit for bit) and the
    oracle; graph batches reject such jobs."""
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    nan = lambda n: torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")  # noqa: E731
    batch = fem.Batch()
    jobs = []
    for i, c in enumerate(meshes[:6]):
        m = c["mesh"]
        flags = [fem.FEM_PLAN_DETERMINISTIC, fem.FEM_PLAN_RESIDUAL_ONLY, 0][i % 3]
        print(f"batch with special plans: mesh {i}, order {m.order}, flags {flags}", flush=True)
        plan = fem.Plan(m, c["pat"], flags=flags)
        coords, states = d(m.coords), d(c["u"])
        res = nan(2 * m.num_nodes)
        vals = None if flags == fem.FEM_PLAN_RESIDUAL_ONLY else nan(4 * c["pat"].num_blocks)
        if flags == fem.FEM_PLAN_RESIDUAL_ONLY:
            with pytest.raises(fem.FemError):
                batch.add(plan, states, coords, E, NU, T, res=res, values=nan(4 * c["pat"].num_blocks))
        batch.add(plan, states, coords, E, NU, T, res=res, values=vals)
        jobs.append((c, plan, flags, coords, states, res, vals))
    batch.assemble()
    torch.cuda.synchronize()
    for c, plan, flags, coords, states, res, vals in jobs:
        r = res.cpu().numpy()
        assert textbook.normwise_rel(r, c["res"]) < 1e-12
        r1 = nan(r.size)
        v1 = None if vals is None else nan(vals.numel())
        plan.assemble(states, coords, E, NU, T, res=r1, values=v1)
        torch.cuda.synchronize()
        if vals is not None:
            v = vals.cpu().numpy()
            assert textbook.normwise_rel(v, c["vals"]) < 1e-12
        if flags == fem.FEM_PLAN_DETERMINISTIC:
            assert np.array_equal(r, r1.cpu().numpy()) and np.array_equal(v, v1.cpu().numpy())
    graph = fem.Batch(graph=True)
    c, plan, _, coords, states, res, vals = jobs[0]
    with pytest.raises(fem.FemError):
        graph.add(plan, states, coords, E, NU, T, res=res, values=vals)
    graph.close()
    batch.close()
    for job in jobs:
        job[1].close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_element_once_groups(fem, torch_gpu, meshes, graph):
    """Jacobian jobs of orders 3-4 run in the batch's element-once groups, each with its own staging in the
    batch's arena, also two jobs on one plan. LINEAR strain: K does not depend on the states and
    R(2u) = 2 R(u). Replayed after clear and refill (graph batches read their tables from device memory)."""
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    high = [c for c in meshes if c["mesh"].order >= 3]
    plans = [fem.Plan(c["mesh"], c["pat"]) for c in high]
    batch = fem.Batch(graph=graph)
    for rnd in range(2):
        batch.clear()
        jobs = []
        for i, (c, plan) in enumerate(zip(high, plans)):
            m = c["mesh"]
            coords = d(m.coords)
            for k, scale in enumerate((1.0, 2.0)):  # k = 1: the same plan again
                states = d(scale * c["u"])
                with_res = (i + k + rnd) % 2 == 0
                res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64,
                                 device="cuda:0") if with_res else None
                vals = torch.full((4 * c["pat"].num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
                batch.add(plan, states, coords, E, NU, T, res=res, values=vals)
                jobs.append((c, scale, res, vals, states, coords))
        batch.assemble()
        torch.cuda.synchronize()
        for c, scale, res, vals, _, _ in jobs:
            assert textbook.normwise_rel(vals.cpu().numpy(), c["vals"]) < 1e-12
            if res is not None:
                assert textbook.normwise_rel(res.cpu().numpy(), scale * c["res"]) < 1e-12
    # the element-once results are bitwise those of the plan assembled alone
    for (c, scale, res, vals, states, coords), plan in zip(jobs[::2], plans):
        alone = torch.zeros_like(vals)
        plan.assemble(states, coords, E, NU, T, values=alone)
        torch.cuda.synchronize()
        assert torch.equal(alone, vals)
    batch.close()
    for plan in plans:
        plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_element_once_many_units(fem, orc, torch_gpu, graph):
    """Element-once jobs with more units (four wavefronts' elements) than the persistent grid of 8
    workgroups per CU on 256 CUs (each workgroup loops), and element counts that leave wavefronts of the
    last unit without elements (16900 = 8 * 2112 + 4 elements at order 4, two per wavefront; 33124 =
    16 * 2070 + 4 at order 3, four per wavefront). Against the oracle and, bitwise, the plan assembled
    alone."""
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    cases = []
    for p, n in ((4, 130), (3, 182)):
        m = fem.square_mesh(n, p, jitter=0.2, seed=p)
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 40 + p)
        pat = fem.bsr_pattern(m)
        res_ref, vals_ref = orc.jacobian(m, u, E, NU, T, fem.element_bcsr_map(m, pat), pat.num_blocks)
        cases.append((m, pat, fem.Plan(m, pat), d(u), d(m.coords), res_ref, vals_ref))
    batch = fem.Batch(graph=graph)
    outs = []
    for m, pat, plan, states, coords, _, _ in cases:
        res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0")
        vals = torch.full((4 * pat.num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
        batch.add(plan, states, coords, E, NU, T, res=res, values=vals)
        outs.append((res, vals))
    for _ in range(2):
        batch.assemble()
    torch.cuda.synchronize()
    for (m, pat, plan, states, coords, res_ref, vals_ref), (res, vals) in zip(cases, outs):
        assert textbook.normwise_rel(res.cpu().numpy(), res_ref) < 1e-12
        assert textbook.normwise_rel(vals.cpu().numpy(), vals_ref) < 1e-12
        r1 = torch.full_like(res, float("nan"))
        v1 = torch.full_like(vals, float("nan"))
        plan.assemble(states, coords, E, NU, T, res=r1, values=v1)
        torch.cuda.synchronize()
        assert torch.equal(r1, res) and torch.equal(v1, vals)
        plan.close()
    batch.close()


def test_batches_sharing_plans_run_concurrently(fem, torch_gpu, meshes):
    """Two batches holding the same order 3-4 plans, assembled on two streams without synchronisation in
    between (as the serving loop's slots do): each batch stages into its own arena, so both results are the
    oracle's (states u and 3u: K unchanged, R scaled)."""
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    high = [c for c in meshes if c["mesh"].order >= 3]
    plans = [fem.Plan(c["mesh"], c["pat"]) for c in high]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    batches, outs = [], []
    for k, scale in enumerate((1.0, 3.0)):
        b = fem.Batch(graph=bool(k))
        for c, plan in zip(high, plans):
            m = c["mesh"]
            res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0")
            vals = torch.full((4 * c["pat"].num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
            b.add(plan, d(scale * c["u"]), d(m.coords), E, NU, T, res=res, values=vals)
            outs.append((c, scale, res, vals))
        batches.append(b)
    torch.cuda.synchronize()
    for _ in range(3):
        for b, st in zip(batches, streams):
            b.assemble(stream=st)
    torch.cuda.synchronize()
    for c, scale, res, vals in outs:
        assert textbook.normwise_rel(vals.cpu().numpy(), c["vals"]) < 1e-12
        assert textbook.normwise_rel(res.cpu().numpy(), scale * c["res"]) < 1e-12
    for b in batches:
        b.close()
    for plan in plans:
        plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_staging_regrows(fem, torch_gpu, meshes, graph):
    """A batch's own element-once staging arena grows when a later composition needs more: one order 3-4 job,
    then every order 3-4 job three times over (several times the staging). Results against the oracle after
    the regrow, on replays with eager work in between."""
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    high = [c for c in meshes if c["mesh"].order >= 3]
    plans = [fem.Plan(c["mesh"], c["pat"]) for c in high]
    batch = fem.Batch(max_jobs=16, graph=graph)
    for rep, compo in enumerate(([0], list(range(len(high))) * 3)):
        batch.clear()
        outs = []
        for k in compo:
            c = high[k]
            m = c["mesh"]
            res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0")
            vals = torch.full((4 * c["pat"].num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
            batch.add(plans[k], d(c["u"]), d(m.coords), E, NU, T, res=res, values=vals)
            outs.append((c, res, vals))
        for _ in range(2):
            batch.assemble()
            torch.cuda.synchronize()
            _ = torch.ones(1024, device="cuda:0").sum().item()  # eager work between replays
            for c, res, vals in outs:
                assert textbook.normwise_rel(res.cpu().numpy(), c["res"]) < 1e-12
                assert textbook.normwise_rel(vals.cpu().numpy(), c["vals"]) < 1e-12
    batch.close()
    for p in plans:
        p.close()
