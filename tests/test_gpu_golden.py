"""GPU: the HIP paths against the committed golden fixtures (tests/golden/*.npz).

FP64: 1e-12 normwise relative for the residual and the BSR values; mixed precision (FP32 arithmetic, FP64
geometry and accumulation): 1e-5. Paths: element-parallel FP64-atomic drop-in, element-colouring
deterministic, resident plan (LDS atomics) and deterministic plan.
"""

import glob
import os

import numpy as np
import pytest

import textbook
from test_golden import load_case, regenerate

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def check(c, res, vals, tol):
    assert textbook.normwise_rel(res, c["residual"]) < tol
    assert textbook.normwise_rel(vals, c["values"]) < tol


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
@pytest.mark.parametrize("path_kind", ["atomic", "coloured", "plan", "plan-det"])
@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)], ids=["fp64", "mixed"])
def test_hip_matches_fixture(fem, torch_gpu, path, path_kind, precision, tol):
    torch = torch_gpu
    c = load_case(path)
    m, u = regenerate(fem, c)
    E, nu, t = (float(v) for v in c["material"])
    pat = fem.bsr_pattern(m)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    coords, states = d(m.coords), d(u)
    res = torch.full((2 * m.num_nodes,), float("nan"), dtype=torch.float64, device="cuda:0")
    vals = torch.full((4 * pat.num_blocks,), float("nan"), dtype=torch.float64, device="cuda:0")
    if path_kind.startswith("plan"):
        plan = fem.Plan(m, pat, flags=fem.FEM_PLAN_DETERMINISTIC if path_kind == "plan-det" else 0)
        plan.assemble(states, coords, E, nu, t, res=res, values=vals, precision=precision)
        torch.cuda.synchronize()
        plan.close()
    else:
        res.zero_()
        vals.zero_()
        bmap = d(fem.element_bcsr_map(m, pat))
        if path_kind == "atomic":
            fem.jacobian(m.order, d(m.conn_ptr), d(m.conn), m.num_elements, states, coords, E, nu, t, bmap, res, vals,
                         precision=precision)
        else:
            colour, nc = fem.element_colouring(m)
            elems, ptr = fem.colour_lists(colour, nc)
            fem.jacobian_coloured(m.order, d(m.conn_ptr), d(m.conn), m.num_elements, states, coords, E, nu, t, bmap,
                                  d(elems), ptr, res, vals, precision=precision)
        torch.cuda.synchronize()
    check(c, res.cpu().numpy(), vals.cpu().numpy(), tol)
