"""GPU parity of the NONLINEAR (Green) strain drop-in paths (fem_residual_nonlinear, fem_jacobian_nonlinear)
against the oracle's NONLINEAR restatement of element/FEKernels.h:286-508, FP64 at <= 1e-12 normwise
relative, on displacement fields whose gradients reach tens of percent (so the quadratic strain terms
matter), plus the rigid-rotation closed form on the device and the FP64-only contract."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E, NU, T = 70e9, 0.3, 1.0
TOL = 1e-12


def rel(x, ref):
    return float(np.abs(x - ref).max() / np.abs(ref).max())


def case(fem, kind, order):
    if kind == "square":
        m = fem.square_mesh(6, order, jitter=0.2, seed=order)
    else:
        m = fem.annulus_mesh(3, 10, order, jitter=0.1, seed=order)
    u = fem.smooth_field(m.coords, 5e-2, phase_seed=order, pert=1e-3, pert_seed=7)
    return m, u


@pytest.mark.parametrize("kind", ["square", "annulus"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_nonlinear_matches_oracle(fem, orc, torch_gpu, kind, order):
    torch = torch_gpu
    m, u = case(fem, kind, order)
    pat = fem.bsr_pattern(m)
    bmap = fem.element_bcsr_map(m, pat)
    r_ref = orc.residual(m, u, E, NU, T, nonlinear=True)
    rj_ref, v_ref = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks, nonlinear=True)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    cp, conn, st, xy, mp = d(m.conn_ptr), d(m.conn.reshape(-1)), d(u), d(m.coords), d(bmap)
    res = torch.zeros(2 * m.num_nodes, dtype=torch.float64, device="cuda:0")
    fem.residual(order, cp, conn, m.num_elements, st, xy, E, NU, T, res, nonlinear=True)
    res2 = torch.zeros_like(res)
    vals = torch.zeros(4 * pat.num_blocks, dtype=torch.float64, device="cuda:0")
    fem.jacobian(order, cp, conn, m.num_elements, st, xy, E, NU, T, mp, res2, vals, nonlinear=True)
    torch.cuda.synchronize()
    assert rel(res.cpu().numpy(), r_ref) <= TOL
    assert rel(res2.cpu().numpy(), rj_ref) <= TOL
    assert rel(vals.cpu().numpy(), v_ref) <= TOL
    # the NONLINEAR tangent differs from the LINEAR stiffness at this displacement
    _, v_lin = orc.jacobian(m, u, E, NU, T, bmap, pat.num_blocks)
    assert rel(v_ref, v_lin) > 1e-3


def test_nonlinear_rigid_rotation_on_device(fem, torch_gpu):
    torch = torch_gpu
    m = fem.square_mesh(5, 3, jitter=0.2, seed=3)
    th = 0.5
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    u = (m.coords @ (R - np.eye(2)).T).reshape(-1)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    res = torch.zeros(2 * m.num_nodes, dtype=torch.float64, device="cuda:0")
    fem.residual(3, d(m.conn_ptr), d(m.conn.reshape(-1)), m.num_elements, d(u), d(m.coords), E, NU, T, res,
                 nonlinear=True)
    torch.cuda.synchronize()
    assert float(res.abs().max()) <= 1e-11 * E * T * np.abs(u).max()


def test_nonlinear_is_fp64_only(fem, torch_gpu):
    torch = torch_gpu
    m = fem.square_mesh(2, 1)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    res = torch.zeros(2 * m.num_nodes, dtype=torch.float64, device="cuda:0")
    with pytest.raises(fem.FemError) as e:
        fem.residual(1, d(m.conn_ptr), d(m.conn.reshape(-1)), m.num_elements, d(np.zeros(2 * m.num_nodes)),
                     d(m.coords), E, NU, T, res, precision=fem.FEM_PREC_MIXED, nonlinear=True)
    assert e.value.status == -7
    # no elements: nothing to do
    fem.residual(1, d(m.conn_ptr), d(m.conn.reshape(-1)), 0, d(np.zeros(2 * m.num_nodes)), d(m.coords), E, NU, T,
                 res, nonlinear=True)
    torch.cuda.synchronize()
    assert float(res.abs().max()) == 0.0
