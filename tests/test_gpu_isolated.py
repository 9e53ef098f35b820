"""Meshes holding nodes that no element references, through every plan route. A BDF file can carry such nodes.
The reference takes its matrix pattern from TACS (verification/TestKernels.cpp:113-115) and fills its element
map from the elements only (verification/TACSHelpers.h:138-182), so nothing is ever added to those nodes' rows;
this library's pattern gives each such node its diagonal block. Plans write those blocks and residual entries as
zeros (accumulating plans leave them as they were); every other entry matches the oracle.

The isolated nodes are scattered through the numbering and placed inside the domain, so that they fall
between the element nodes both in id and in the partition's Morton order."""

import numpy as np
import pytest

import textbook

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0
TOL = 1e-12


def isolated_mesh(fem, order, extra, seed):
    base = fem.square_mesh(5, order, jitter=0.2, seed=order)
    rng = np.random.default_rng(seed)
    n0 = base.num_nodes
    perm = rng.permutation(n0 + extra).astype(np.int32)  # new id of old node i; ids n0.. are the isolated ones
    coords = np.empty((n0 + extra, 2))
    coords[perm[:n0]] = base.coords
    coords[perm[n0:]] = rng.random((extra, 2))
    conn = np.ascontiguousarray(perm[base.conn])
    return fem.Mesh(order, coords, conn, "square 5x5 + isolated nodes"), np.sort(perm[n0:])


def diagonal_slots(pat, nodes):
    return np.array([pat.rowp[n] + int(np.searchsorted(pat.cols[pat.rowp[n]:pat.rowp[n + 1]], n)) for n in nodes])


@pytest.fixture(scope="module")
def cases(fem, orc):
    out = {}
    for order in (1, 2, 3, 4):
        m, iso = isolated_mesh(fem, order, 9, 100 + order)
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 7)
        pat = fem.bsr_pattern(m)
        bmap = fem.element_bcsr_map(m, pat)
        res, vals = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks)
        assert np.all(res.reshape(-1, 2)[iso] == 0)
        out[order] = dict(mesh=m, iso=iso, u=u, pat=pat, bmap=bmap, res=res, vals=vals,
                          diag=diagonal_slots(pat, iso))
    return out


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def filled(torch, n, value):
    return torch.full((n,), value, dtype=torch.float64, device="cuda:0")


def check(c, r, v, base=0.0):
    """Isolated entries hold `base` exactly; the rest is base + the oracle's sums."""
    if r is not None:
        assert np.all(r.reshape(-1, 2)[c["iso"]] == base)
        assert textbook.normwise_rel(r - base, c["res"]) < TOL
    if v is not None:
        assert np.all(v.reshape(-1, 4)[c["diag"]] == base)
        assert textbook.normwise_rel(v - base, c["vals"]) < TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("route", ["default", "deterministic", "owner-only", "residual-only"])
def test_plan_routes(fem, torch_gpu, cases, order, route):
    if route == "owner-only" and order < 3:
        pytest.skip("the owner-only flag selects a route at orders 3-4 only")
    torch = torch_gpu
    c = cases[order]
    m, pat = c["mesh"], c["pat"]
    flags = {"default": 0, "deterministic": fem.FEM_PLAN_DETERMINISTIC, "owner-only": fem.FEM_PLAN_OWNER_ONLY,
             "residual-only": fem.FEM_PLAN_RESIDUAL_ONLY}[route]
    print(f"isolated nodes: order {order}, {route}, {m.num_nodes} nodes", flush=True)
    plan = fem.Plan(m, pat, flags=flags)
    states, coords = dev(torch, c["u"]), dev(torch, m.coords)
    res = filled(torch, 2 * m.num_nodes, float("nan"))
    vals = None if route == "residual-only" else filled(torch, 4 * pat.num_blocks, float("nan"))
    plan.assemble(states, coords, E, NU, T, res=res, values=vals)
    torch.cuda.synchronize()
    check(c, res.cpu().numpy(), None if vals is None else vals.cpu().numpy())
    if route == "default":
        # accumulating assembly onto ones: isolated entries stay 1
        res1 = filled(torch, 2 * m.num_nodes, 1.0)
        vals1 = filled(torch, 4 * pat.num_blocks, 1.0)
        plan.assemble(states, coords, E, NU, T, res=res1, values=vals1, accumulate=True)
        torch.cuda.synchronize()
        check(c, res1.cpu().numpy(), vals1.cpu().numpy(), base=1.0)
    plan.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_batch(fem, torch_gpu, cases, order):
    torch = torch_gpu
    c = cases[order]
    m, pat = c["mesh"], c["pat"]
    print(f"isolated nodes: batch, order {order}", flush=True)
    plan = fem.Plan(m, pat)
    batch = fem.Batch(4)
    states, coords = dev(torch, c["u"]), dev(torch, m.coords)
    res = filled(torch, 2 * m.num_nodes, float("nan"))
    vals = filled(torch, 4 * pat.num_blocks, float("nan"))
    batch.add(plan, states, coords, E, NU, T, res=res, values=vals)
    batch.assemble()
    torch.cuda.synchronize()
    check(c, res.cpu().numpy(), vals.cpu().numpy())
    batch.close()
    plan.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_dropin(fem, torch_gpu, cases, order):
    """runJacobianKernel's entry point on zeroed outputs (the reference's protocol)."""
    torch = torch_gpu
    c = cases[order]
    m, pat = c["mesh"], c["pat"]
    print(f"isolated nodes: drop-in, order {order}", flush=True)
    fem.dropin_cache_clear()
    res = filled(torch, 2 * m.num_nodes, 0.0)
    vals = filled(torch, 4 * pat.num_blocks, 0.0)
    fem.jacobian(order, dev(torch, m.conn_ptr), dev(torch, m.conn), m.num_elements, dev(torch, c["u"]),
                 dev(torch, m.coords), E, NU, T, dev(torch, c["bmap"]), res, vals, dropin=True)
    torch.cuda.synchronize()
    check(c, res.cpu().numpy(), vals.cpu().numpy())
    fem.dropin_cache_clear()
