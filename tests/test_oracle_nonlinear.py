"""The oracle's NONLINEAR-strain assembly (the reference's kernels with GreenStrainType::NONLINEAR,
element/FEKernels.h:286-315, 317-508), pinned by what the mathematics fixes:
- zero displacement: zero residual and the LINEAR stiffness;
- a rigid rotation: zero Green strain, so zero residual (the LINEAR residual of the same field is not zero);
- a homogeneous deformation F of the unit square: interior residuals vanish (patch test) and the nodal
  moments sum_a x_a R_a, sum_a y_a R_a equal t * area * P, with P = F S the first Piola-Kirchhoff stress
  of the Saint Venant-Kirchhoff material S = 2 mu E + lambda tr(E) I, E = (F^T F - I) / 2;
- the stiffness is the derivative of the residual (central finite differences) and is symmetric;
- small displacements reduce to the LINEAR residual.
"""

import numpy as np
import pytest

E, NU, T = 70e9, 0.3, 1.0
MU = 0.5 * E / (1 + NU)
LAM = 2 * MU * NU / (1 - NU)


def pk1(F):
    Eg = 0.5 * (F.T @ F - np.eye(2))
    S = 2 * MU * Eg + LAM * np.trace(Eg) * np.eye(2)
    return F @ S


def elem_coords(fem, order, seed):
    m = fem.square_mesh(1, order, jitter=0.0)
    rng = np.random.default_rng(seed)
    x = m.coords[m.conn[0]].copy()
    x += 0.05 * rng.uniform(-1, 1, x.shape)  # a distorted element
    return x


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_zero_displacement(fem, orc, order):
    x = elem_coords(fem, order, order)
    u = np.zeros_like(x)
    r_nl, k_nl = orc.element(order, x, u, E, NU, T, nonlinear=True)
    r_l, k_l = orc.element(order, x, u, E, NU, T)
    assert np.all(r_nl == 0)
    assert np.abs(k_nl - k_l).max() <= 1e-14 * np.abs(k_l).max()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rigid_rotation_has_no_residual(fem, orc, order):
    mesh = fem.square_mesh(3, order, jitter=0.2, seed=order)
    th = 0.4
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    u = (mesh.coords @ (R - np.eye(2)).T).reshape(-1)
    r_nl = orc.residual(mesh, u, E, NU, T, nonlinear=True)
    r_l = orc.residual(mesh, u, E, NU, T)
    scale = E * T * np.abs(u).max()  # stress x length scale of the field
    assert np.abs(r_nl).max() <= 1e-11 * scale
    assert np.abs(r_l).max() >= 1e-3 * scale  # the linear strain of a rotation is not zero


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_homogeneous_deformation_closed_form(fem, orc, order):
    mesh = fem.square_mesh(3, order, jitter=0.2, seed=10 + order)
    F = np.array([[1.1, 0.2], [-0.05, 0.95]])
    u = (mesh.coords @ (F - np.eye(2)).T).reshape(-1)
    r = orc.residual(mesh, u, E, NU, T, nonlinear=True).reshape(-1, 2)
    P = pk1(F)
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    tol = 1e-12 * np.abs(P).max()
    # sum_a x_a R_a,i = t * area * P_i0, sum_a y_a R_a,i = t * area * P_i1 (area 1)
    assert np.abs(x @ r - T * P[:, 0]).max() <= tol
    assert np.abs(y @ r - T * P[:, 1]).max() <= tol
    interior = (x > 1e-12) & (x < 1 - 1e-12) & (y > 1e-12) & (y < 1 - 1e-12)
    assert np.abs(r[interior]).max() <= tol


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_tangent_is_residual_derivative_and_symmetric(fem, orc, order):
    x = elem_coords(fem, order, 20 + order)
    rng = np.random.default_rng(30 + order)
    u = 0.05 * rng.uniform(-1, 1, x.shape)  # displacement gradients of a few percent: nonlinear regime
    r0, K = orc.element(order, x, u, E, NU, T, nonlinear=True)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    h = 1e-6
    uf = u.reshape(-1)
    for j in rng.choice(uf.size, size=min(6, uf.size), replace=False):
        up, um = uf.copy(), uf.copy()
        up[j] += h
        um[j] -= h
        rp, _ = orc.element(order, x, up, E, NU, T, nonlinear=True)
        rm, _ = orc.element(order, x, um, E, NU, T, nonlinear=True)
        fd = (rp - rm) / (2 * h)
        assert np.abs(fd - K[:, j]).max() <= 1e-6 * np.abs(K).max(), j


def test_small_strain_limit(fem, orc):
    mesh = fem.square_mesh(4, 2, jitter=0.2, seed=3)
    u = fem.fill_uniform(2 * mesh.num_nodes, -1e-9, 1e-9, 4)
    r_nl = orc.residual(mesh, u, E, NU, T, nonlinear=True)
    r_l = orc.residual(mesh, u, E, NU, T)
    assert np.abs(r_nl - r_l).max() <= 1e-6 * np.abs(r_l).max()
