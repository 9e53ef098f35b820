"""GPU parity at the full benchmark sizes: the resident plan path against the oracle (the CPU restatement of the
reference's assembly, element/FEKernels.h:317-508, run threaded on the host) on the whole mesh, as the
reference's harness compares full assembled outputs (verification/TestKernels.cpp:302-327).

Workloads (the benchmark's own generators and seeds):
  * unit square 1000 x 1000 order 2 (1M elements, 64M BSR blocks), interior jitter U(+-0.2h) seed 2,
    u,v ~ U(+-1e-3) seed 3;
  * annulus r in [1, 2], 200 x 1250 order 4 (250k elements, 144M blocks), smooth field (phases seed 4) plus
    U(+-1e-5) (seed 5).

Bars, normwise relative (max|x - ref| / max|ref|), for the residual and the BSR values separately:
  FP64   <= 1e-12;
  mixed  <= 1e-5 (FP32 Gram sums; geometry, stresses, residual and all accumulation in FP64).
Outputs are NaN-prefilled, so an entry the path never writes fails the comparison. The cancellation ratio of the
residual, max(|K||u|) / max|K u| (R = K u exactly for linear strain), is printed beside its error.
"""

import os

import numpy as np
import pytest

import textbook

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0
FP64_TOL = 1e-12
MIXED_TOL = 1e-5


def oracle_threads():
    return max(1, min(16, os.cpu_count() or 1))


def abs_spmv(rowp, cols, vals, x):
    """|K| |x| for a BSR matrix whose rows are all non-empty (every row holds its diagonal block)."""
    A = np.abs(vals.reshape(-1, 2, 2))
    prod = np.einsum("kij,kj->ki", A, np.abs(x.reshape(-1, 2))[cols])
    return np.add.reduceat(prod, rowp[:-1].astype(np.int64), axis=0).reshape(-1)


def build(fem, orc, name):
    if name == "square1000-p2":
        m = fem.square_mesh(1000, 2, jitter=0.2, seed=2)
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 3)
    else:
        m = fem.annulus_mesh(200, 1250, 4)
        u = fem.smooth_field(m.coords, 1e-2, phase_seed=4, pert=1e-5, pert_seed=5)
    pat = fem.bsr_pattern(m)
    bmap = fem.element_bcsr_map(m, pat)
    res, vals = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks, threads=oracle_threads())
    del bmap
    kappa = float(abs_spmv(pat.rowp, pat.cols, vals, u).max() / np.abs(res).max())
    return dict(mesh=m, u=u, pat=pat, res=res, vals=vals, kappa=kappa)


@pytest.fixture(scope="module")
def fullsize(fem, orc):
    return {}


def case(fem, orc, store, name):
    if name not in store:
        store.clear()  # one full-size case resident on the host at a time
        store[name] = build(fem, orc, name)
    return store[name]


@pytest.mark.parametrize("name", ["square1000-p2", "annulus200x1250-p4"])
@pytest.mark.parametrize("precision,tol", [(0, FP64_TOL), (1, MIXED_TOL)], ids=["fp64", "mixed"])
def test_plan_matches_oracle_full_size(fem, orc, torch_gpu, fullsize, name, precision, tol):
    torch = torch_gpu
    c = case(fem, orc, fullsize, name)
    m = c["mesh"]
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0")
    vals = torch.full((4 * c["pat"].num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
    plan = fem.Plan(m, c["pat"])
    plan.assemble(d(c["u"]), d(m.coords), E, NU, T, res=res, values=vals, precision=precision)
    torch.cuda.synchronize()
    plan.close()
    e_res = textbook.normwise_rel(res.cpu().numpy(), c["res"])
    hv = vals.cpu().numpy()
    del vals
    e_val = textbook.normwise_rel(hv, c["vals"])
    print(f"{name} precision={precision}: residual {e_res:.3e} (cancellation ratio {c['kappa']:.3g}), "
          f"values {e_val:.3e}")
    assert e_res <= tol
    assert e_val <= tol
