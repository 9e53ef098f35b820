"""BDF mesh files (fem_read_bdf / fem_write_bdf, host C ABI): the GRID and CQUADn cards the reference's meshes
use (read through TACS, verification/TACSHelpers.h:86-114; laid out by fixGmshBDF, meshes/GmshUtils.py:118-196).

Pinned by: round trips of the generator's own meshes (orders 1-4, square and annulus) through the writer, bit
for bit; hand-written cards in each accepted layout (fixed small field with blank-field and '+'-marker
continuation lines, GRID* large field, free field, implicit-exponent reals, comments, non-contiguous ids) with
their expected arrays written out; Nastran's CQUAD4 / CQUAD9 node order converted to the tensor-product order
and checked by the assembled element (the oracle's K_e of the converted element equals that of the same
element given in tensor order); and the error paths."""

import os

import numpy as np
import pytest


def write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


@pytest.mark.parametrize("kind,order", [("square", 1), ("square", 2), ("annulus", 3), ("square", 4), ("annulus", 4)])
def test_round_trip(fem, tmp_path, kind, order):
    m = (fem.square_mesh(5, order, jitter=0.2, seed=order) if kind == "square"
         else fem.annulus_mesh(2, 8, order, jitter=0.1, seed=order))
    path = os.path.join(str(tmp_path), "m.bdf")
    fem.write_bdf(path, m)
    r, ids = fem.read_bdf(path)
    assert r.order == m.order
    assert np.array_equal(r.coords, m.coords)  # "%.17g" reals read back exactly
    assert np.array_equal(r.conn, m.conn)
    assert np.array_equal(ids, np.arange(1, m.num_nodes + 1))
    # element cards: nine 8-character fields per line, continuation lines opening with a blank field
    lines = open(path).read().splitlines()
    quad = [i for i, l in enumerate(lines) if l.startswith("CQUAD")]
    assert len(quad) == m.num_elements and lines[quad[0]].startswith(f"CQUAD{m.npe}")
    if m.npe + 3 > 9:
        assert lines[quad[0] + 1].startswith(" " * 8)
        assert len(lines[quad[0]]) == 72


FIXED = """$ hand-written cards in the accepted layouts
BEGIN BULK
GRID    10              0.      0.      0.
GRID    11              .5      0.
GRID*   12                              1.0                     0.0
*       0.
GRID,13,0,0.,5.-1,0.
GRID    14      0       .5      .5      0.
GRID    15              1.      .5+0
GRID    20              0.      1.      0.
GRID    21              .5      1.D0
GRID,22,,1.0E0,1.
CQUAD9  7       1       10      11      12      13      14      15
        20      21      22
ENDDATA
GRID    99      0       9.      9.
"""

GMSH_PLUS = """BEGIN BULK
GRID,1,0,0.,0.,0.
GRID,2,0,1.,0.,0.
GRID,3,0,0.,1.,0.
GRID,4,0,1.,1.,0.
GRID,5,0,.5,0.,0.
GRID,6,0,1.,.5,0.
GRID,7,0,.5,1.,0.
GRID,8,0,0.,.5,0.
GRID,9,0,.5,.5,0.
CQUAD9  1       1       1       2       4       3       5       6       +E1
+E1     7       8       9
ENDDATA
"""


def test_fixed_large_and_free_fields(fem, tmp_path):
    m, ids = fem.read_bdf(write(tmp_path, "fixed.bdf", FIXED))
    assert m.order == 2 and m.num_elements == 1
    assert ids.tolist() == [10, 11, 12, 13, 14, 15, 20, 21, 22]
    expect = [[0, 0], [.5, 0], [1, 0], [0, .5], [.5, .5], [1, .5], [0, 1], [.5, 1], [1, 1]]
    assert np.array_equal(m.coords, np.array(expect, dtype=np.float64))
    assert m.conn.tolist() == [list(range(9))]


def test_plus_marker_continuation_and_nastran_order(fem, orc, tmp_path):
    path = write(tmp_path, "gmsh.bdf", GMSH_PLUS)
    raw, _ = fem.read_bdf(path)
    assert raw.conn.tolist() == [[0, 1, 3, 2, 4, 5, 6, 7, 8]]  # card order
    tensor, _ = fem.read_bdf(path, nastran_order=True)
    # corners CCW (G1-G4), mid-sides G5 (G1-G2), G6 (G2-G3), G7 (G3-G4), G8 (G4-G1), centre G9 -> a = ay*3 + ax
    assert tensor.conn.tolist() == [[0, 4, 1, 7, 8, 5, 2, 6, 3]]
    lx = tensor.coords[tensor.conn[0]]
    grid = np.array([[x, y] for y in (0, .5, 1) for x in (0, .5, 1)])
    assert np.array_equal(lx, grid)  # the converted element's nodes sit at the tensor-product positions
    u = np.linspace(-1e-3, 1e-3, 18)
    r1, k1 = orc.element(2, lx, u, 70e9, .3, 1.)
    r2, k2 = orc.element(2, grid, u, 70e9, .3, 1.)
    assert np.array_equal(k1, k2) and np.array_equal(r1, r2)


def test_cquad4_nastran_order(fem, tmp_path):
    text = "GRID,1,,0.,0.\nGRID,2,,2.,0.\nGRID,3,,2.,1.\nGRID,4,,0.,1.\nCQUAD4,1,1,1,2,3,4\n"
    m, _ = fem.read_bdf(write(tmp_path, "q4.bdf", text), nastran_order=True)
    assert m.order == 1 and m.conn.tolist() == [[0, 1, 3, 2]]


@pytest.mark.parametrize("text,status", [
    ("GRID,1,,0.,0.\nCQUAD4,1,1,1,2,3,4\n", -3),                          # element node without a GRID card
    ("GRID,1,,0.,0.,1.\nCQUAD4,1,1,1,1,1,1\n", -7),                       # X3 != 0
    ("GRID,1,1,0.,0.\nCQUAD4,1,1,1,1,1,1\n", -7),                         # coordinate system 1
    ("GRID,1,,0.,0.\nCQUAD4,1,1,1,1,1\n", -3),                            # three node fields
    ("GRID,1,,0.,0.\nGRID,1,,1.,0.\nCQUAD4,1,1,1,1,1,1\n", -3),           # duplicate GRID id
    ("GRID,1,,0.,0.\nCQUAD4,1,1,1,1,1,1\nCQUAD9,2,1,1,1,1,1,1,1,1,1,1\n", -7),  # two element orders
    ("GRID,1,,abc,0.\nCQUAD4,1,1,1,1,1,1\n", -3),                         # not a real
    ("$ nothing\n", -3),
])
def test_errors(fem, tmp_path, text, status):
    with pytest.raises(fem.FemError) as ei:
        fem.read_bdf(write(tmp_path, "bad.bdf", text))
    assert ei.value.status == status


def test_missing_file(fem):
    with pytest.raises(fem.FemError) as ei:
        fem.read_bdf("/nonexistent/mesh.bdf")
    assert ei.value.status == -3
