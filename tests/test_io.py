"""Output files in the reference's formats (verification/TACSHelpers.h:234-307): a vector as one "% .17g"
value per line, a BSR matrix as a "rows cols nnz" header plus 0-based "row col value" lines, with the
optional leading-rows / leading-columns limits. Checked line by line against Python's printf-style
formatting and by reading the matrix back into a dense array."""

import numpy as np
import pytest


def test_write_array_matches_printf_format(fem, tmp_path):
    v = np.array([0.0, -0.0, 1.0, -1.5e-300, 123456789.123456789, np.pi, -np.e * 1e17, np.inf, -np.inf])
    path = tmp_path / "res.csv"
    fem.write_array(v, str(path))
    lines = path.read_text().splitlines()
    assert lines == ["% .17g" % x for x in v]
    back = np.array([float(x) for x in lines])
    assert np.array_equal(back, v)  # 17 significant digits round-trip every double


def test_write_array_empty(fem, tmp_path):
    path = tmp_path / "empty.csv"
    fem.write_array(np.zeros(0), str(path))
    assert path.read_text() == ""


def dense_from_bsr(pattern, values, bsize=2):
    n = pattern.rowp.size - 1
    a = np.zeros((n * bsize, n * bsize))
    for i in range(n):
        for k in range(pattern.rowp[i], pattern.rowp[i + 1]):
            j = pattern.cols[k]
            a[i * bsize:(i + 1) * bsize, j * bsize:(j + 1) * bsize] = values[4 * k:4 * k + 4].reshape(2, 2)
    return a


def read_matrix(path):
    lines = path.read_text().splitlines()
    rows, cols, nnz = (int(x) for x in lines[0].split())
    entries = [line.split() for line in lines[1:]]
    assert len(entries) == nnz
    a = np.zeros((rows, cols))
    for r, c, v in entries:
        a[int(r), int(c)] = float(v)
    return (rows, cols, nnz), entries, a


def test_write_bsr_matrix_round_trip(fem, tmp_path):
    mesh = fem.square_mesh(3, 2, jitter=0.2, seed=1)
    pat = fem.bsr_pattern(mesh)
    vals = np.random.default_rng(3).standard_normal(4 * pat.num_blocks)
    path = tmp_path / "jac.mtx"
    fem.write_bsr_matrix(pat, vals, str(path))
    (rows, cols, nnz), entries, a = read_matrix(path)
    n = mesh.num_nodes
    assert (rows, cols, nnz) == (2 * n, 2 * n, 4 * pat.num_blocks)
    assert np.array_equal(a, dense_from_bsr(pat, vals))
    # line format: block by block, each block row-major, values as "% .17g"
    i0, j0 = 0, int(pat.cols[0])
    first = ["%d %d % .17g" % (2 * i0 + r, 2 * j0 + c, vals[2 * r + c]) for r in range(2) for c in range(2)]
    assert [" ".join(e) for e in entries[:4]] == [" ".join(f.split()) for f in first]
    assert path.read_text().splitlines()[1:5] == first


@pytest.mark.parametrize("max_rows,max_cols", [(5, -1), (-1, 7), (6, 4), (1000, 1000)])
def test_write_bsr_matrix_limits(fem, tmp_path, max_rows, max_cols):
    mesh = fem.square_mesh(3, 1)
    pat = fem.bsr_pattern(mesh)
    vals = np.arange(1, 4 * pat.num_blocks + 1, dtype=np.float64)
    path = tmp_path / "jac_part.mtx"
    fem.write_bsr_matrix(pat, vals, str(path), max_rows=max_rows, max_cols=max_cols)
    (rows, cols, nnz), _, a = read_matrix(path)
    n = mesh.num_nodes
    nr = min(max_rows, n) if max_rows > 0 else n
    nc = min(max_cols, n) if max_cols > 0 else n
    assert (rows, cols) == (2 * nr, 2 * nc)
    full = dense_from_bsr(pat, vals)
    assert np.array_equal(a, full[:2 * nr, :2 * nc])
    blocks = sum(1 for i in range(nr) for k in range(pat.rowp[i], pat.rowp[i + 1]) if pat.cols[k] < nc)
    assert nnz == 4 * blocks


def test_write_errors(fem, tmp_path):
    with pytest.raises(fem.FemError) as ei:
        fem.write_array(np.ones(3), str(tmp_path / "missing_dir" / "x.csv"))
    assert ei.value.status == -3
