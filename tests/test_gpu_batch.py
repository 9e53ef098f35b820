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


def test_batch_rejects_deterministic_plans(fem, torch_gpu, meshes):
    c = meshes[1]
    plan = fem.Plan(c["mesh"], c["pat"], flags=fem.FEM_PLAN_DETERMINISTIC)
    batch = fem.Batch()
    t = torch_gpu.zeros(8, dtype=torch_gpu.float64, device="cuda:0")
    with pytest.raises(fem.FemError):
        batch.add(plan, t, t, E, NU, T, res=t)
    batch.close()
    plan.close()
