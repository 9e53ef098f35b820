"""GPU parity: every device path of libfem_amd against the oracle (the CPU restatement of the reference's
assembly, element/FEKernels.h:317-508) on identical seeded meshes and displacements.

Tolerances (normwise relative, max|x - ref| / max|ref|):
  FP64  <= 1e-12 for the residual and the BSR values;
  MIXED (FP32 basis/stress/B^T D B, FP64 geometry and accumulation) <= 1e-5.
Deterministic paths (colour path, deterministic plan) must be bitwise reproducible run to run.
"""

import numpy as np
import pytest

import textbook

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0
FP64_TOL = 1e-12
MIXED_TOL = 1e-5


def build_case(fem, orc, kind, order):
    if kind == "square":
        m = fem.square_mesh(7, order, jitter=0.2, seed=2)  # 49 elements: not a multiple of any batch size
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 3)
    else:
        m = fem.annulus_mesh(3, 12, order, jitter=0.1, seed=4)
        u = fem.smooth_field(m.coords, 1e-2, phase_seed=4, pert=1e-5, pert_seed=5)
    pat = fem.bsr_pattern(m)
    bmap = fem.element_bcsr_map(m, pat)
    res, vals = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks)
    return dict(mesh=m, u=u, pat=pat, bmap=bmap, res=res, vals=vals)


CASES = [(k, p) for p in (1, 2, 3, 4) for k in ("square", "annulus")]


@pytest.fixture(scope="module")
def cases(fem, orc):
    return {c: build_case(fem, orc, *c) for c in CASES}


class Dev:
    def __init__(self, torch, case):
        self.torch = torch
        m = case["mesh"]
        d = self.d
        self.coords = d(m.coords)
        self.states = d(case["u"])
        self.conn = d(m.conn)
        self.conn_ptr = d(m.conn_ptr)
        self.bmap = d(case["bmap"])
        self.rowp = d(case["pat"].rowp)
        self.cols = d(case["pat"].cols)
        self.n_res = 2 * m.num_nodes
        self.n_val = 4 * case["pat"].num_blocks

    def d(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")

    def zeros(self, n):
        return self.torch.zeros(n, dtype=self.torch.float64, device="cuda:0")

    def nan(self, n):
        return self.torch.full((n,), float("nan"), dtype=self.torch.float64, device="cuda:0")


def host(t, torch):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("case_id", CASES, ids=[f"{k}-p{p}" for k, p in CASES])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_dropin_atomic_path(fem, torch_gpu, cases, case_id, precision):
    c = cases[case_id]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    tol = FP64_TOL if precision == 0 else MIXED_TOL
    res = D.zeros(D.n_res)
    fem.residual(m.order, D.conn_ptr, D.conn, m.num_elements, D.states, D.coords, E, NU, T, res, precision)
    assert textbook.normwise_rel(host(res, torch_gpu), c["res"]) < tol
    vals = D.zeros(D.n_val)
    res2 = D.zeros(D.n_res)
    fem.jacobian(m.order, D.conn_ptr, D.conn, m.num_elements, D.states, D.coords, E, NU, T, D.bmap, res2, vals,
                 precision)
    assert textbook.normwise_rel(host(vals, torch_gpu), c["vals"]) < tol
    assert textbook.normwise_rel(host(res2, torch_gpu), c["res"]) < tol
    vals3 = D.zeros(D.n_val)
    fem.jacobian(m.order, D.conn_ptr, D.conn, m.num_elements, D.states, D.coords, E, NU, T, D.bmap, None, vals3,
                 precision)
    assert textbook.normwise_rel(host(vals3, torch_gpu), c["vals"]) < tol


@pytest.mark.parametrize("case_id", CASES, ids=[f"{k}-p{p}" for k, p in CASES])
def test_colour_path_parity_and_bitwise_determinism(fem, torch_gpu, cases, case_id):
    c = cases[case_id]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    colour, nc = fem.element_colouring(m)
    elems, ptr = fem.colour_lists(colour, nc)
    delems = D.d(elems)
    outs = []
    for _ in range(2):
        vals = D.zeros(D.n_val)
        res = D.zeros(D.n_res)
        fem.jacobian_coloured(m.order, D.conn_ptr, D.conn, m.num_elements, D.states, D.coords, E, NU, T, D.bmap,
                              delems, ptr, res, vals)
        outs.append((host(res, torch_gpu), host(vals, torch_gpu)))
    assert textbook.normwise_rel(outs[0][1], c["vals"]) < FP64_TOL
    assert textbook.normwise_rel(outs[0][0], c["res"]) < FP64_TOL
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("case_id", CASES, ids=[f"{k}-p{p}" for k, p in CASES])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("det", [False, True], ids=["atomicLDS", "det"])
def test_plan_path(fem, torch_gpu, cases, case_id, precision, det):
    c = cases[case_id]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    tol = FP64_TOL if precision == 0 else MIXED_TOL
    plan = fem.Plan(m, c["pat"], flags=fem.FEM_PLAN_DETERMINISTIC if det else 0)
    # Outputs start as NaN: the plan must overwrite every entry (write-once, no zero-fill needed).
    res = D.nan(D.n_res)
    vals = D.nan(D.n_val)
    plan.assemble(D.states, D.coords, E, NU, T, res=res, values=vals, precision=precision)
    r, v = host(res, torch_gpu), host(vals, torch_gpu)
    assert np.isfinite(r).all() and np.isfinite(v).all()
    assert textbook.normwise_rel(v, c["vals"]) < tol
    assert textbook.normwise_rel(r, c["res"]) < tol
    # residual alone and values alone
    res_only = D.nan(D.n_res)
    plan.assemble(D.states, D.coords, E, NU, T, res=res_only, precision=precision)
    assert textbook.normwise_rel(host(res_only, torch_gpu), c["res"]) < tol
    vals_only = D.nan(D.n_val)
    plan.assemble(D.states, D.coords, E, NU, T, values=vals_only, precision=precision)
    assert textbook.normwise_rel(host(vals_only, torch_gpu), c["vals"]) < tol
    if det:
        again = D.nan(D.n_val)
        res_again = D.nan(D.n_res)
        plan.assemble(D.states, D.coords, E, NU, T, res=res_again, values=again, precision=precision)
        assert np.array_equal(host(again, torch_gpu), v) and np.array_equal(host(res_again, torch_gpu), r)
    plan.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_residual_only_plan(fem, torch_gpu, cases, order):
    c = cases[("annulus", order)]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    plan = fem.Plan(m, c["pat"], flags=fem.FEM_PLAN_RESIDUAL_ONLY)
    res = D.nan(D.n_res)
    plan.assemble(D.states, D.coords, E, NU, T, res=res)
    assert textbook.normwise_rel(host(res, torch_gpu), c["res"]) < FP64_TOL
    with pytest.raises(fem.FemError):
        plan.assemble(D.states, D.coords, E, NU, T, values=D.zeros(D.n_val))
    plan.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_residual_equals_k_times_u_on_device(fem, torch_gpu, cases, order):
    c = cases[("square", order)]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    plan = fem.Plan(m, c["pat"])
    res = D.zeros(D.n_res)
    vals = D.zeros(D.n_val)
    plan.assemble(D.states, D.coords, E, NU, T, res=res, values=vals)
    y = D.zeros(D.n_res)
    fem.bsr_spmv(D.rowp, D.cols, vals, D.states, y)
    assert textbook.normwise_rel(host(y, torch_gpu), host(res, torch_gpu)) < 1e-13
    plan.close()


def test_empty_and_single_element(fem, orc, torch_gpu):
    m = fem.square_mesh(1, 3, jitter=0.0)
    pat = fem.bsr_pattern(m)
    bmap = fem.element_bcsr_map(m, pat)
    u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 9)
    res_ref, vals_ref = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks)
    c = dict(mesh=m, u=u, pat=pat, bmap=bmap)
    D = Dev(torch_gpu, c)
    # numElements = 0: a no-op that succeeds
    res = D.zeros(D.n_res)
    fem.residual(3, D.conn_ptr, D.conn, 0, D.states, D.coords, E, NU, T, res)
    assert not host(res, torch_gpu).any()
    vals = D.zeros(D.n_val)
    fem.jacobian(3, D.conn_ptr, D.conn, 1, D.states, D.coords, E, NU, T, D.bmap, res, vals)
    assert textbook.normwise_rel(host(vals, torch_gpu), vals_ref) < FP64_TOL
    assert textbook.normwise_rel(host(res, torch_gpu), res_ref) < FP64_TOL
    plan = fem.Plan(m, pat)
    v2 = D.nan(D.n_val)
    plan.assemble(D.states, D.coords, E, NU, T, values=v2)
    assert textbook.normwise_rel(host(v2, torch_gpu), vals_ref) < FP64_TOL
    plan.close()


def test_elapsed_time_reported(fem, torch_gpu, cases):
    c = cases[("square", 2)]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    res = D.zeros(D.n_res)
    sec = fem.residual(2, D.conn_ptr, D.conn, m.num_elements, D.states, D.coords, E, NU, T, res, timed=True)
    assert 0 < sec < 1.0


def test_cpp_dropin_demo_runs(fem, orc, tmp_path):
    """The C++ program calling runResidualKernel / runJacobianKernel as the reference's harness does: its written
    outputs (the reference's .csv / .mtx formats) against the oracle on the same generated mesh."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-fem-project_amd",
                       "fem_dropin_demo")
    for order in (1, 2, 3, 4):
        prefix = os.path.join(str(tmp_path), f"p{order}")
        out = subprocess.run([exe, str(order), "12", prefix], capture_output=True, text=True, timeout=120)
        print(out.stdout)
        assert out.returncode == 0, out.stdout + out.stderr
        m = fem.square_mesh(12, order, jitter=0.2, seed=2)
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 3)
        pat = fem.bsr_pattern(m)
        rref, vref = orc.jacobian(m, u, E, NU, T, fem.element_bcsr_map(m, pat), pat.num_blocks)
        res = np.loadtxt(prefix + "_residual.csv")
        mtx = np.loadtxt(prefix + "_jacobian.mtx", skiprows=1)
        i, j = mtx[:, 0].astype(np.int64), mtx[:, 1].astype(np.int64)
        br, bc = i // 2, j // 2
        slot = np.array([pat.rowp[r] + np.searchsorted(pat.cols[pat.rowp[r]:pat.rowp[r + 1]], c)
                         for r, c in zip(br, bc)])
        got = np.zeros_like(vref)
        got[4 * slot + 2 * (i % 2) + (j % 2)] = mtx[:, 2]
        assert len(mtx) == vref.size
        assert textbook.normwise_rel(res, rref) < FP64_TOL
        assert textbook.normwise_rel(got, vref) < FP64_TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_autotuned_plan_matches_oracle(fem, orc, torch_gpu, order):
    """FEM_PLAN_AUTOTUNE keeps the fastest of three part sizes; whichever it keeps, results are the oracle's."""
    torch = torch_gpu
    n = {1: 12, 2: 8, 3: 5, 4: 4}[order]
    m = fem.square_mesh(n, order, jitter=0.2, seed=70 + order)
    u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 71 + order)
    pat = fem.bsr_pattern(m)
    ref_res, ref_vals = orc.jacobian(m, u, E, NU, T, fem.element_bcsr_map(m, pat), pat.num_blocks)
    plan = fem.Plan(m, pat, flags=fem.FEM_PLAN_AUTOTUNE)
    info = plan.info()
    assert info["part_scale"] in (1.0, 0.5, 0.25)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0")
    vals = torch.full((4 * pat.num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
    plan.assemble(d(u), d(m.coords), E, NU, T, res=res, values=vals)
    torch.cuda.synchronize()
    assert textbook.normwise_rel(res.cpu().numpy(), ref_res) < 1e-12
    assert textbook.normwise_rel(vals.cpu().numpy(), ref_vals) < 1e-12
    plan.close()


@pytest.mark.parametrize("case_id", [c for c in CASES if c[1] >= 3], ids=[f"{k}-p{p}" for k, p in CASES if p >= 3])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("det", [False, True], ids=["atomicLDS", "det"])
def test_owner_route_orders_3_4(fem, torch_gpu, cases, case_id, precision, det):
    """Orders 3-4 through the owner-computes parts (FEM_PLAN_OWNER_ONLY: the route a plan takes when its
    element-once codes would overflow an int): FP64 row-restricted Gram stage, mixed FP32 MFMA Gram stage.
    NaN-prefilled outputs, against the oracle."""
    c = cases[case_id]
    m = c["mesh"]
    D = Dev(torch_gpu, c)
    tol = FP64_TOL if precision == 0 else MIXED_TOL
    flags = fem.FEM_PLAN_OWNER_ONLY | (fem.FEM_PLAN_DETERMINISTIC if det else 0)
    plan = fem.Plan(m, c["pat"], flags=flags)
    res = D.nan(D.n_res)
    vals = D.nan(D.n_val)
    plan.assemble(D.states, D.coords, E, NU, T, res=res, values=vals, precision=precision)
    r, v = host(res, torch_gpu), host(vals, torch_gpu)
    assert np.isfinite(r).all() and np.isfinite(v).all()
    assert textbook.normwise_rel(v, c["vals"]) < tol
    assert textbook.normwise_rel(r, c["res"]) < tol
    vals_only = D.nan(D.n_val)
    plan.assemble(D.states, D.coords, E, NU, T, values=vals_only, precision=precision)
    assert textbook.normwise_rel(host(vals_only, torch_gpu), c["vals"]) < tol
    plan.close()
