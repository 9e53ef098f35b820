"""Pins the oracle's element residual/stiffness (restating element/FEKernels.h:229-508) to the textbook.

Checks: the unit-square 4-node plane-stress stiffness of Andreassen et al. (2011), "Efficient topology
optimization in MATLAB using 88 lines of code"; an independent B^T D B element (tests/textbook.py);
symmetry, the rigid-body null space and positive definiteness on its complement; R_e = K_e u_e; the
faithful per-column form against the oracle's Gram form.
"""

import numpy as np
import pytest

import textbook

E, NU, T = 70e9, 0.3, 1.0


def ke_88line(nu):
    k = [1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8, -1 / 4 + nu / 12, -1 / 8 - nu / 8,
         nu / 6, 1 / 8 - 3 * nu / 8]
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
           [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6], [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    return np.array([[k[i] for i in row] for row in idx]) / (1 - nu**2)


def element_nodes(order, rng=None, distort=0.0, curved=False):
    """Node coords of one element on [0,1]^2 at cosine knots (a = ay*(p+1)+ax), optionally perturbed."""
    knots = 0.5 * (1 + textbook.cosine_knots(order))
    xy = np.array([[knots[ax], knots[ay]] for ay in range(order + 1) for ax in range(order + 1)])
    if distort and rng is not None:
        xy = xy + distort * (rng.random(xy.shape) - 0.5) / order
    if curved:
        xy = xy + 0.05 * np.stack([np.sin(np.pi * xy[:, 1]), np.sin(np.pi * xy[:, 0])], axis=1)
    return xy


def test_unit_square_q4_matches_88_line_stiffness(orc):
    xy = element_nodes(1)  # (0,0) (1,0) (0,1) (1,1)
    _, K = orc.element(1, xy, np.zeros(8), 1.0, NU, 1.0)
    # 88-line node order is counter-clockwise from the lower left: (0,0) (1,0) (1,1) (0,1).
    perm_nodes = [0, 1, 3, 2]
    perm = np.array([[2 * a, 2 * a + 1] for a in perm_nodes]).reshape(-1)
    np.testing.assert_allclose(K[np.ix_(perm, perm)], ke_88line(NU), atol=1e-14)
    ev = np.linalg.eigvalsh(K)
    np.testing.assert_allclose(ev, [0, 0, 0, 0.4945054945, 0.4945054945, 0.7692307692, 0.7692307692, 1.4285714286],
                               atol=1e-9)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("curved", [False, True])
def test_element_matches_textbook_bdb(orc, order, curved):
    rng = np.random.default_rng(order)
    xy = element_nodes(order, rng, distort=0.3, curved=curved)
    u = rng.uniform(-1e-3, 1e-3, xy.shape)
    res, K = orc.element(order, xy, u, E, NU, T)
    K_ref = textbook.element_stiffness(order, xy, E, NU, T)
    assert textbook.normwise_rel(K, K_ref) < 1e-12
    assert textbook.normwise_rel(res, K_ref @ u.reshape(-1)) < 1e-12


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_faithful_and_gram_forms_agree(orc, order):
    rng = np.random.default_rng(10 + order)
    xy = element_nodes(order, rng, distort=0.4, curved=True)
    u = rng.uniform(-1e-3, 1e-3, xy.shape)
    r1, K1 = orc.element(order, xy, u, E, NU, T, gram=False)
    r2, K2 = orc.element(order, xy, u, E, NU, T, gram=True)
    assert textbook.normwise_rel(K1, K2) < 1e-14
    assert textbook.normwise_rel(r1, r2) < 1e-14


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_element_symmetric_rigid_modes_and_definite(orc, order):
    rng = np.random.default_rng(20 + order)
    xy = element_nodes(order, rng, distort=0.3)
    _, K = orc.element(order, xy, np.zeros_like(xy), E, NU, T)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-13 * scale
    modes = [np.tile([1.0, 0.0], len(xy)), np.tile([0.0, 1.0], len(xy)),
             np.stack([-xy[:, 1], xy[:, 0]], axis=1).reshape(-1)]
    for m in modes:
        assert np.abs(K @ m).max() < 1e-11 * scale
    ev = np.linalg.eigvalsh(K / scale)
    assert np.all(np.abs(ev[:3]) < 1e-11) and ev[3] > 1e-6


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_residual_is_stiffness_times_states(orc, order):
    rng = np.random.default_rng(30 + order)
    xy = element_nodes(order, rng, distort=0.3, curved=True)
    u = rng.uniform(-1e-3, 1e-3, xy.shape)
    res, K = orc.element(order, xy, u, E, NU, T)
    assert textbook.normwise_rel(res, K @ u.reshape(-1)) < 1e-13


def test_signed_detj_weight(orc):
    # Mirrored element (detJ < 0): the reference weights with the signed detJ, so K flips sign.
    xy = element_nodes(1)
    _, K = orc.element(1, xy, np.zeros(8), E, NU, T)
    mirrored = xy * np.array([-1.0, 1.0])
    _, Km = orc.element(1, mirrored, np.zeros(8), E, NU, T)
    S = np.diag(np.tile([-1.0, 1.0], 4))
    np.testing.assert_allclose(Km, -(S @ K @ S), rtol=1e-13, atol=1e-3)
