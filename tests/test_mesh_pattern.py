"""Host set-up: synthetic meshes, SplitMix64 draws, the TACS-style BSR pattern, the element->BSR map
(TACS-free restatement of verification/TACSHelpers.h:138-182) and element colouring."""

import numpy as np
import pytest

import textbook

SPLITMIX_SEED_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                         16408922859458223821]


def test_splitmix64_known_answer(fem):
    # Published SplitMix64 outputs for seed 1234567, through u01 = (x >> 11) * 2^-53.
    got = fem.fill_uniform(5, 0.0, 1.0, 1234567)
    want = np.array([(v >> 11) * 2.0**-53 for v in SPLITMIX_SEED_1234567])
    assert np.array_equal(got, want)
    lo, hi = -1e-3, 1e-3
    got = fem.fill_uniform(5, lo, hi, 1234567)
    assert np.array_equal(got, lo + (hi - lo) * want)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_square_mesh_layout(fem, order):
    n = 5
    m = fem.square_mesh(n, order)
    mx = order * n + 1
    assert m.num_nodes == mx * mx and m.num_elements == n * n
    knots = (np.arange(n)[:, None] + 0.5 * (1 + textbook.cosine_knots(order))[None, :]) / n
    xs = np.unique(np.round(knots.reshape(-1), 15))
    np.testing.assert_allclose(np.sort(np.unique(np.round(m.coords[:, 0], 15))), xs, atol=1e-15)
    # node id = jy*mx + jx, element e = ey*n + ex, local a = ay*(p+1) + ax
    e = 2 * n + 3
    ex, ey = 3, 2
    for ay in range(order + 1):
        for ax in range(order + 1):
            assert m.conn[e, ay * (order + 1) + ax] == (ey * order + ay) * mx + ex * order + ax
    assert fem.min_detj(m) > 0


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_square_jitter_bounded_and_boundary_fixed(fem, order):
    n = 6
    a = fem.square_mesh(n, order)
    b = fem.square_mesh(n, order, jitter=0.2, seed=2)
    mx = order * n + 1
    d = np.abs(b.coords - a.coords).reshape(mx, mx, 2)
    assert d[0].max() == 0 and d[-1].max() == 0 and d[:, 0].max() == 0 and d[:, -1].max() == 0
    gaps = np.diff(np.unique(a.coords[:, 0]))
    # each interior node moves at most 0.2 x its smaller adjacent knot gap, per axis
    assert d.max() <= 0.2 * gaps.max() + 1e-15 and d.max() >= 0.2 * gaps.min() * 0.5
    assert fem.min_detj(b) > 0
    # Same seed -> same mesh.
    c = fem.square_mesh(n, order, jitter=0.2, seed=2)
    assert np.array_equal(b.coords, c.coords)


def test_c2_jitter_draw_order(fem):
    # Interior nodes only, node-id order, dx then dy, amplitude 0.2*h for order 2 (h uniform).
    n, order = 4, 2
    m = fem.square_mesh(n, order, jitter=0.2, seed=2)
    base = fem.square_mesh(n, order)
    mx = order * n + 1
    h = 1.0 / (order * n)
    interior = [(jy, jx) for jy in range(1, mx - 1) for jx in range(1, mx - 1)]
    draws = fem.fill_uniform(2 * len(interior), -1.0, 1.0, 2)
    for k, (jy, jx) in enumerate(interior):
        node = jy * mx + jx
        np.testing.assert_allclose(m.coords[node] - base.coords[node], 0.2 * h * draws[2 * k: 2 * k + 2], atol=1e-17)


@pytest.mark.parametrize("order", [1, 2, 4])
def test_annulus_mesh(fem, order):
    nr, nt = 3, 10
    m = fem.annulus_mesh(nr, nt, order)
    assert m.num_nodes == (order * nr + 1) * order * nt and m.num_elements == nr * nt
    r = np.hypot(m.coords[:, 0], m.coords[:, 1])
    assert abs(r.min() - 1.0) < 1e-14 and abs(r.max() - 2.0) < 1e-14
    assert fem.min_detj(m) > 0
    # Periodic seam: the last element ring reuses the first node column.
    assert set(m.conn[(nt - 1) * nr, -(order + 1):]) == set(range(order + 1))


def closed_form_blocks_square(n, p):
    """Block count of the pattern of an n x n order-p square mesh."""
    m = p * n + 1
    tot = 0
    for j in range(m):
        # nodes coupled along one axis: span of the (one or two) elements containing knot j
        lo = (j - 1) // p * p if j % p == 0 and j > 0 else j // p * p
        hi = (j // p + 1) * p if j % p != 0 or j == 0 else min(j + p, m - 1)
        if j % p == 0:
            lo = max(j - p, 0)
            hi = min(j + p, m - 1)
        tot += hi - lo + 1
    return tot * tot


@pytest.mark.parametrize("order,n", [(1, 64), (2, 7), (3, 5), (4, 4)])
def test_pattern_matches_bruteforce_and_closed_form(fem, orc, order, n):
    m = fem.square_mesh(n, order)
    pat = fem.bsr_pattern(m)
    rowp, cols = orc.bsr_pattern(m)
    assert np.array_equal(pat.rowp, rowp) and np.array_equal(pat.cols, cols)
    assert pat.num_blocks == closed_form_blocks_square(n, order)
    if (order, n) == (1, 64):
        assert pat.num_blocks == 37249 and m.num_nodes == 4225  # the 64x64 order-1 CPU configuration
    for r in range(m.num_nodes):
        row = pat.cols[pat.rowp[r]: pat.rowp[r + 1]]
        assert np.all(np.diff(row) > 0) and r in row


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_map_matches_linear_search_restatement(fem, orc, order):
    m = fem.annulus_mesh(2, 8, order, jitter=0.1, seed=5)
    pat = fem.bsr_pattern(m)
    mp = fem.element_bcsr_map(m, pat)
    mo, misses = orc.element_bcsr_map(m, pat.rowp, pat.cols)
    assert misses == 0 and np.array_equal(mp, mo)
    # Invariant: the block at map[e,i,j] is (conn[e,i], conn[e,j]).
    N = m.npe
    slots = mp.reshape(m.num_elements, N, N) // 4
    rows_of_slot = np.repeat(np.arange(m.num_nodes), np.diff(pat.rowp))
    assert np.array_equal(rows_of_slot[slots], np.broadcast_to(m.conn[:, :, None], slots.shape))
    assert np.array_equal(pat.cols[slots], np.broadcast_to(m.conn[:, None, :], slots.shape))


def test_map_miss_is_an_error(fem):
    m = fem.square_mesh(3, 1)
    pat = fem.bsr_pattern(m)
    bad_cols = pat.cols.copy()
    bad_cols[pat.rowp[5]] = bad_cols[pat.rowp[5] + 1]  # drop one column of row 5 (duplicates another)
    with pytest.raises(fem.FemError) as ei:
        fem.element_bcsr_map(m, fem.Pattern(pat.rowp, bad_cols))
    assert ei.value.status == -5


@pytest.mark.parametrize("order", [1, 2, 3])
def test_colouring_valid(fem, order):
    m = fem.square_mesh(6, order, jitter=0.1, seed=1)
    colour, nc = fem.element_colouring(m)
    assert nc == 4  # structured quads, greedy in element order
    for c in range(nc):
        els = np.where(colour == c)[0]
        nodes = m.conn[els].reshape(-1)
        assert len(np.unique(nodes)) == len(nodes)
    order_, ptr = fem.colour_lists(colour, nc)
    assert ptr[-1] == m.num_elements and np.array_equal(np.sort(order_), np.arange(m.num_elements))


def test_invalid_inputs_are_errors(fem):
    with pytest.raises(fem.FemError) as ei:
        fem.square_mesh(3, 5)
    assert ei.value.status == -1
    with pytest.raises(fem.FemError):
        fem.square_mesh(0, 1)
    m = fem.square_mesh(2, 1)
    m.conn[0, 0] = 10_000  # node id out of range
    with pytest.raises(fem.FemError) as ei:
        fem.bsr_pattern(m)
    assert ei.value.status == -3


def test_smooth_field_reference_phases(fem):
    m = fem.square_mesh(4, 2)
    u = fem.smooth_field(m.coords, 1e-2, phase_seed=0)
    x, y = m.coords[:, 0], m.coords[:, 1]
    # TACSHelpers.h:192-196 with zero phases
    np.testing.assert_allclose(u[0::2], 1e-2 * (np.sin(2 * np.pi * x) + np.sin(2 * np.pi * y)), atol=1e-17)
    np.testing.assert_allclose(u[1::2], 1e-2 * (np.cos(2 * np.pi * x) + np.cos(2 * np.pi * y)), atol=1e-17)
