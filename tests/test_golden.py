"""Committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle).

The fixtures are pinned here by what they do not share with the oracle: the mesh generator and BSR pattern
builder of libfem_amd reproduce their inputs exactly, and an independent numpy textbook assembly (Gauss-Legendre
from numpy, B^T D B per element, dense global matrix) reproduces their residual and BSR values to 1e-12
normwise. The oracle reproduces them to rounding (1e-14), so a change to either shows up here.
"""

import glob
import os

import numpy as np
import pytest

import textbook

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_case(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def regenerate(fem, c):
    kind, order = str(c["kind"]), int(c["order"])
    size, jitter, seed = [int(v) for v in c["size"]], float(c["jitter"]), int(c["mesh_seed"])
    if kind == "square":
        m = fem.square_mesh(size[0], order, jitter=jitter, seed=seed, ny=size[1])
    else:
        m = fem.annulus_mesh(size[0], size[1], order, jitter=jitter, seed=seed)
    u = fem.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, int(c["state_seed"]))
    return m, u


def test_fixtures_present():
    names = {os.path.basename(f) for f in FILES}
    assert "c1_square_p1_64x64.npz" in names and len(names) >= 6
    assert {int(load_case(f)["order"]) for f in FILES} == {1, 2, 3, 4}


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_generator_and_pattern_reproduce_fixture(fem, path):
    c = load_case(path)
    m, u = regenerate(fem, c)
    pat = fem.bsr_pattern(m)
    assert m.num_nodes == int(c["num_nodes"]) and m.num_elements == int(c["num_elements"])
    assert np.array_equal(pat.rowp, c["rowp"])
    assert np.array_equal(m.coords.reshape(-1), c["coords"].reshape(-1))
    assert np.array_equal(m.conn, c["conn"])
    assert np.array_equal(u, c["states"])
    assert np.array_equal(pat.cols, c["cols"])


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_textbook_assembly_matches_fixture(fem, path):
    c = load_case(path)
    m, u = regenerate(fem, c)
    E, nu, t = (float(v) for v in c["material"])
    pat = fem.bsr_pattern(m)
    if m.num_nodes > 1000:
        # 64x64 order 1: the BSR values by a sparse textbook assembly, the residual as K u through them
        vals = textbook.assemble_bsr(m.order, m.coords, m.conn, pat.rowp, pat.cols, E, nu, t)
        assert textbook.normwise_rel(vals, c["values"]) < 1e-12
        assert textbook.normwise_rel(textbook.bsr_spmv(pat.rowp, pat.cols, vals, u), c["residual"]) < 1e-12
        return
    K = textbook.assemble_dense(m.order, m.coords, m.conn, E, nu, t)
    assert textbook.normwise_rel(textbook.dense_to_bsr_values(K, pat.rowp, pat.cols), c["values"]) < 1e-12
    assert textbook.normwise_rel(K @ u, c["residual"]) < 1e-12


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_oracle_reproduces_fixture(fem, orc, path):
    c = load_case(path)
    m, u = regenerate(fem, c)
    E, nu, t = (float(v) for v in c["material"])
    rowp, cols = orc.bsr_pattern(m)
    bmap, misses = orc.element_bcsr_map(m, rowp, cols)
    assert misses == 0
    res, vals = orc.jacobian(m, u, E, nu, t, bmap, int(rowp[-1]))
    assert textbook.normwise_rel(res, c["residual"]) < 1e-14
    assert textbook.normwise_rel(vals, c["values"]) < 1e-14
