"""GPU: the reference-interface Jacobian backed by a cached resident plan (fem_jacobian_dropin) and plan
assembly in accumulate mode (fem_plan_assemble_accumulate), against the oracle.

The reference's kernels add into caller-zeroed outputs (element/FEKernels.h:1125-1225); here outputs are
prefilled with random values X, so a call must leave X + oracle (FP64: 1e-12 normwise against the result), and a
second call X + 2 oracle. Checked: orders 1-2 are served by a plan; the cache notices connectivity and map
rewritten in place (node ids relabelled: a stale plan would write the old layout); a caller matrix of another
layout and orders 3-4 run the atomic kernels; accumulate mode on owner-route and element-once plans of orders 3-4
(with wavefronts left partly empty)."""

import numpy as np
import pytest

import textbook

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0
TOL = 1e-12


def d(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def case(fem, orc, mesh, seed):
    u = fem.fill_uniform(2 * mesh.num_nodes, -1e-3, 1e-3, seed)
    pat = fem.bsr_pattern(mesh)
    bmap = fem.element_bcsr_map(mesh, pat)
    res, vals = orc.jacobian(mesh, u, E, NU, T, bmap, pat.num_blocks)
    return u, pat, bmap, res, vals


def call(fem, torch, mesh, conn_d, cp_d, u_d, x_d, map_d, res, vals):
    fem.jacobian(mesh.order, cp_d, conn_d, mesh.num_elements, u_d, x_d, E, NU, T, map_d, res, vals, dropin=True)
    torch.cuda.synchronize()


@pytest.mark.parametrize("kind,order", [("square", 1), ("square", 2), ("annulus", 1), ("annulus", 2)])
def test_dropin_plan_accumulates(fem, orc, torch_gpu, kind, order):
    torch = torch_gpu
    fem.dropin_cache_clear()
    m = (fem.square_mesh(30, order, jitter=0.2, seed=order) if kind == "square"
         else fem.annulus_mesh(6, 40, order, jitter=0.1, seed=order))
    u, pat, bmap, rref, vref = case(fem, orc, m, 7)
    rng = np.random.default_rng(order)
    X_r = rng.uniform(-1, 1, rref.size) * np.abs(rref).max()
    X_v = rng.uniform(-1, 1, vref.size) * np.abs(vref).max()
    res, vals = d(torch, X_r), d(torch, X_v)
    cp_d, conn_d, map_d, u_d, x_d = d(torch, m.conn_ptr), d(torch, m.conn), d(torch, bmap), d(torch, u), d(torch, m.coords)
    call(fem, torch, m, conn_d, cp_d, u_d, x_d, map_d, res, vals)
    assert fem.dropin_cache_info(order, cp_d, conn_d, m.num_elements, map_d) == (True, True)
    assert textbook.normwise_rel(res.cpu().numpy(), X_r + rref) < TOL
    assert textbook.normwise_rel(vals.cpu().numpy(), X_v + vref) < TOL
    call(fem, torch, m, conn_d, cp_d, u_d, x_d, map_d, res, vals)  # served from the cache
    assert textbook.normwise_rel(res.cpu().numpy(), X_r + 2 * rref) < TOL
    assert textbook.normwise_rel(vals.cpu().numpy(), X_v + 2 * vref) < TOL
    # values only (residual NULL, as the reference allows)
    vals2 = torch.zeros_like(vals)
    fem.jacobian(order, cp_d, conn_d, m.num_elements, u_d, x_d, E, NU, T, map_d, None, vals2, dropin=True)
    torch.cuda.synchronize()
    assert textbook.normwise_rel(vals2.cpu().numpy(), vref) < TOL
    fem.dropin_cache_clear()


def test_dropin_notices_arrays_rewritten_in_place(fem, orc, torch_gpu):
    """Same device pointers, node ids relabelled in place: the cached plan must be rebuilt."""
    torch = torch_gpu
    fem.dropin_cache_clear()
    m = fem.square_mesh(24, 2, jitter=0.2, seed=3)
    u, pat, bmap, rref, vref = case(fem, orc, m, 8)
    perm = np.random.default_rng(5).permutation(m.num_nodes).astype(np.int32)  # old id -> new id
    m2 = fem.Mesh(order=2, coords=np.empty_like(m.coords), conn=perm[m.conn])
    m2.coords[perm] = m.coords
    u2 = np.empty_like(u).reshape(-1, 2)
    u2[perm] = u.reshape(-1, 2)
    u2 = u2.reshape(-1)
    pat2 = fem.bsr_pattern(m2)
    bmap2 = fem.element_bcsr_map(m2, pat2)
    rref2, vref2 = orc.jacobian(m2, u2, E, NU, T, bmap2, pat2.num_blocks)
    assert pat2.num_blocks == pat.num_blocks
    cp_d, conn_d, map_d = d(torch, m.conn_ptr), d(torch, m.conn), d(torch, bmap)
    res = torch.zeros(rref.size, dtype=torch.float64, device="cuda:0")
    vals = torch.zeros(vref.size, dtype=torch.float64, device="cuda:0")
    call(fem, torch, m, conn_d, cp_d, d(torch, u), d(torch, m.coords), map_d, res, vals)
    assert textbook.normwise_rel(vals.cpu().numpy(), vref) < TOL
    conn_d.copy_(d(torch, m2.conn))  # rewritten in place
    map_d.copy_(d(torch, bmap2))
    res.zero_()
    vals.zero_()
    call(fem, torch, m2, conn_d, cp_d, d(torch, u2), d(torch, m2.coords), map_d, res, vals)
    assert fem.dropin_cache_info(2, cp_d, conn_d, m.num_elements, map_d) == (True, True)
    assert textbook.normwise_rel(res.cpu().numpy(), rref2) < TOL
    assert textbook.normwise_rel(vals.cpu().numpy(), vref2) < TOL
    fem.dropin_cache_clear()


def test_dropin_other_layout_uses_atomic_kernels(fem, orc, torch_gpu):
    """A caller matrix whose pattern holds extra blocks (each node also coupled to the next node id): the map
    differs from this library's, so the call runs the atomic kernels into the caller's layout."""
    torch = torch_gpu
    fem.dropin_cache_clear()
    m = fem.square_mesh(12, 2, jitter=0.2, seed=4)
    u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 9)
    pat = fem.bsr_pattern(m)
    rows = [set(pat.cols[pat.rowp[r]:pat.rowp[r + 1]].tolist()) | {(r + 1) % m.num_nodes} for r in range(m.num_nodes)]
    rowp = np.zeros(m.num_nodes + 1, dtype=np.int32)
    rowp[1:] = np.cumsum([len(r) for r in rows])
    cols = np.concatenate([np.array(sorted(r), dtype=np.int32) for r in rows])
    fat = fem.Pattern(rowp=rowp, cols=cols)
    bmap = fem.element_bcsr_map(m, fat)
    rref, vref = orc.jacobian(m, u, E, NU, T, bmap, fat.num_blocks)
    cp_d, conn_d, map_d = d(torch, m.conn_ptr), d(torch, m.conn), d(torch, bmap)
    res = torch.zeros(rref.size, dtype=torch.float64, device="cuda:0")
    vals = torch.zeros(vref.size, dtype=torch.float64, device="cuda:0")
    call(fem, torch, m, conn_d, cp_d, d(torch, u), d(torch, m.coords), map_d, res, vals)
    assert fem.dropin_cache_info(2, cp_d, conn_d, m.num_elements, map_d) == (True, False)
    assert textbook.normwise_rel(res.cpu().numpy(), rref) < TOL
    assert textbook.normwise_rel(vals.cpu().numpy(), vref) < TOL
    fem.dropin_cache_clear()


@pytest.mark.parametrize("kind,order", [("square", 3), ("annulus", 3), ("square", 4), ("annulus", 4)])
def test_element_once_accumulate_and_dropin_orders_3_4(fem, orc, torch_gpu, kind, order):
    """Orders 3-4: element-once plans in accumulate mode (element counts that leave the last wavefront partly
    empty, so the emission's repeating lanes must not add twice), and the drop-in, which runs the atomic kernels
    there."""
    torch = torch_gpu
    fem.dropin_cache_clear()
    m = (fem.square_mesh(7, order, jitter=0.2, seed=order) if kind == "square"
         else fem.annulus_mesh(3, 13, order, jitter=0.1, seed=order))
    u, pat, bmap, rref, vref = case(fem, orc, m, 10)
    rng = np.random.default_rng(order)
    X_r = rng.uniform(-1, 1, rref.size) * np.abs(rref).max()
    X_v = rng.uniform(-1, 1, vref.size) * np.abs(vref).max()
    cp_d, conn_d, map_d, u_d, x_d = d(torch, m.conn_ptr), d(torch, m.conn), d(torch, bmap), d(torch, u), d(torch, m.coords)
    plan = fem.Plan(m, pat)
    res, vals = d(torch, X_r), d(torch, X_v)
    for k in (1, 2):
        plan.assemble(u_d, x_d, E, NU, T, res=res, values=vals, accumulate=True)
        torch.cuda.synchronize()
        assert textbook.normwise_rel(res.cpu().numpy(), X_r + k * rref) < TOL
        assert textbook.normwise_rel(vals.cpu().numpy(), X_v + k * vref) < TOL
    plan.close()
    res, vals = d(torch, X_r), d(torch, X_v)
    call(fem, torch, m, conn_d, cp_d, u_d, x_d, map_d, res, vals)
    assert fem.dropin_cache_info(order, cp_d, conn_d, m.num_elements, map_d) == (False, False)
    assert textbook.normwise_rel(res.cpu().numpy(), X_r + rref) < TOL
    assert textbook.normwise_rel(vals.cpu().numpy(), X_v + vref) < TOL


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)], ids=["fp64", "mixed"])
def test_plan_accumulate_owner_route(fem, orc, torch_gpu, precision, tol):
    torch = torch_gpu
    m = fem.annulus_mesh(3, 12, 3, jitter=0.1, seed=2)
    u, pat, bmap, rref, vref = case(fem, orc, m, 11)
    X_r = np.linspace(-1, 1, rref.size) * np.abs(rref).max()
    X_v = np.linspace(1, -1, vref.size) * np.abs(vref).max()
    res, vals = d(torch, X_r), d(torch, X_v)
    plan = fem.Plan(m, pat, flags=fem.FEM_PLAN_OWNER_ONLY)
    plan.assemble(d(torch, u), d(torch, m.coords), E, NU, T, res=res, values=vals, precision=precision,
                  accumulate=True)
    torch.cuda.synchronize()
    assert textbook.normwise_rel(res.cpu().numpy(), X_r + rref) < tol
    assert textbook.normwise_rel(vals.cpu().numpy(), X_v + vref) < tol
    plan.close()
    once = fem.Plan(m, pat)  # element-once Jacobian, accumulate mode
    res, vals = d(torch, X_r), d(torch, X_v)
    once.assemble(d(torch, u), d(torch, m.coords), E, NU, T, res=res, values=vals, precision=precision,
                  accumulate=True)
    torch.cuda.synchronize()
    assert textbook.normwise_rel(res.cpu().numpy(), X_r + rref) < tol
    assert textbook.normwise_rel(vals.cpu().numpy(), X_v + vref) < tol
    once.close()
