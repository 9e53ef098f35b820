"""Pins the oracle's global assembly (element/FEKernels.h:317-508 restated) with global invariants and an
independent dense textbook assembler."""

import numpy as np
import pytest

import textbook

E, NU, T = 70e9, 0.3, 1.0


def meshes(fem):
    yield fem.square_mesh(3, 1, jitter=0.2, seed=2)
    yield fem.square_mesh(3, 2, jitter=0.2, seed=2)
    yield fem.square_mesh(2, 3, jitter=0.15, seed=4)
    yield fem.annulus_mesh(2, 6, 4, jitter=0.1, seed=6)
    yield fem.annulus_mesh(2, 8, 2)


@pytest.fixture(scope="module")
def cases(fem, orc):
    out = []
    for m in meshes(fem):
        pat = fem.bsr_pattern(m)
        mp = fem.element_bcsr_map(m, pat)
        u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 3)
        res, vals = orc.jacobian(m, u, E, NU, T, mp, pat.num_blocks)
        out.append((m, pat, mp, u, res, vals))
    return out


def test_assembly_matches_dense_textbook(cases):
    for m, pat, mp, u, res, vals in cases:
        K = textbook.assemble_dense(m.order, m.coords, m.conn, E, NU, T)
        assert textbook.normwise_rel(vals, textbook.dense_to_bsr_values(K, pat.rowp, pat.cols)) < 1e-12, m.kind
        assert textbook.normwise_rel(res, K @ u) < 1e-12, m.kind
        # no stiffness outside the pattern
        Kb = textbook.bsr_to_dense(m.num_nodes, pat.rowp, pat.cols, vals)
        assert np.abs(Kb - K).max() <= 1e-12 * np.abs(K).max()


def test_residual_only_matches_jacobian_residual(orc, cases):
    for m, pat, mp, u, res, vals in cases:
        r = orc.residual(m, u, E, NU, T)
        assert textbook.normwise_rel(r, res) < 1e-15


def test_global_symmetry_rigid_modes_and_r_equals_ku(cases):
    for m, pat, mp, u, res, vals in cases:
        K = textbook.bsr_to_dense(m.num_nodes, pat.rowp, pat.cols, vals)
        s = np.abs(K).max()
        assert np.abs(K - K.T).max() < 1e-13 * s
        x, y = m.coords[:, 0], m.coords[:, 1]
        for mode in (np.stack([np.ones_like(x), 0 * x], 1), np.stack([0 * x, np.ones_like(x)], 1),
                     np.stack([-y, x], 1)):
            assert np.abs(K @ mode.reshape(-1)).max() < 1e-11 * s
        assert textbook.normwise_rel(textbook.bsr_spmv(pat.rowp, pat.cols, vals, u), res) < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_patch_test_linear_field(fem, orc, order):
    # Linear displacement on a distorted mesh: constant stress, interior residual vanishes.
    m = fem.square_mesh(3, order, jitter=0.2, seed=7)
    x, y = m.coords[:, 0], m.coords[:, 1]
    u = np.stack([1e-3 * x + 2e-3 * y, -5e-4 * x + 1e-3 * y], 1).reshape(-1)
    r = orc.residual(m, u, E, NU, T).reshape(-1, 2)
    mx = order * 3 + 1
    jy, jx = np.divmod(np.arange(m.num_nodes), mx)
    interior = (jx > 0) & (jx < mx - 1) & (jy > 0) & (jy < mx - 1)
    assert np.abs(r[interior]).max() < 1e-9 * np.abs(r).max()


@pytest.mark.parametrize("order", [1, 2, 4])
def test_uniform_strain_energy(fem, orc, order):
    # u = (eps x, 0) on the unit square: u^T K u = E t A eps^2 / (1 - nu^2).
    m = fem.square_mesh(3, order, jitter=0.2, seed=8)
    pat = fem.bsr_pattern(m)
    mp = fem.element_bcsr_map(m, pat)
    eps = 1e-3
    u = np.stack([eps * m.coords[:, 0], 0 * m.coords[:, 0]], 1).reshape(-1)
    _, vals = orc.jacobian(m, u, E, NU, T, mp, pat.num_blocks, with_residual=False)
    energy = u @ textbook.bsr_spmv(pat.rowp, pat.cols, vals, u)
    assert abs(energy / (E * T * 1.0 * eps**2 / (1 - NU**2)) - 1) < 1e-12


def test_threaded_oracle_matches_serial(fem, orc):
    m = fem.square_mesh(8, 2, jitter=0.2, seed=2)
    pat = fem.bsr_pattern(m)
    mp = fem.element_bcsr_map(m, pat)
    u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, 3)
    r1, v1 = orc.jacobian(m, u, E, NU, T, mp, pat.num_blocks, threads=1)
    r4, v4 = orc.jacobian(m, u, E, NU, T, mp, pat.num_blocks, threads=4)
    assert textbook.normwise_rel(v4, v1) < 1e-15 and textbook.normwise_rel(r4, r1) < 1e-15
