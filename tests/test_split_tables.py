"""The element-once tables behind the order 3-4 plans (fem_split_tables, the host builder fem_plan_create uses),
checked on the host: assembling random element contributions through the tables (exclusive outputs stored,
the others staged and summed in slot order) reproduces a plain scatter-add in element order bit for bit, and
writes every BSR block and residual entry exactly once. Meshes: every order, square and annulus, shuffled
node and element numbering, a high-valence fan, nodes that no element references (their outputs are
gathered from no contributions: zero)."""

import numpy as np
import pytest


def permuted(fem, mesh, seed):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(mesh.num_nodes).astype(np.int32)
    conn = perm[mesh.conn][rng.permutation(mesh.num_elements)]
    coords = np.empty_like(mesh.coords)
    coords[perm] = mesh.coords
    return fem.Mesh(mesh.order, coords, np.ascontiguousarray(conn), mesh.kind + " permuted")


def fan(fem, k):
    conn = np.array([[0, 1 + 3 * i, 2 + 3 * i, 3 + 3 * i] for i in range(k)], dtype=np.int32)
    return fem.Mesh(1, np.random.default_rng(k).random((1 + 3 * k, 2)), conn, f"fan of {k}")


def with_isolated(fem, mesh, extra, seed):
    """`extra` nodes that no element references, scattered through the numbering."""
    rng = np.random.default_rng(seed)
    n0 = mesh.num_nodes
    perm = rng.permutation(n0 + extra).astype(np.int32)
    coords = np.zeros((n0 + extra, 2))
    coords[perm[:n0]] = mesh.coords
    return fem.Mesh(mesh.order, coords, np.ascontiguousarray(perm[mesh.conn]), mesh.kind + " + isolated nodes")


def meshes(fem):
    out = []
    for p in (1, 2, 3, 4):
        out.append(fem.square_mesh(5, p, jitter=0.2, seed=p))
        out.append(fem.annulus_mesh(2, 8, p, jitter=0.1, seed=p))
        out.append(permuted(fem, fem.square_mesh(4, p), 10 + p))
    out.append(fan(fem, 40))
    out.append(fem.square_mesh(1, 4))
    out.append(with_isolated(fem, fem.square_mesh(3, 2), 6, 21))
    out.append(with_isolated(fem, fem.square_mesh(3, 4), 6, 22))
    return out


def split_assemble(t, contrib, outputs, width):
    """Stores exclusive contributions, stages the rest, then sums each gathered output's slots in order."""
    code, dst, ptr, slots = t
    out = np.full((outputs, width), np.nan)
    stage = np.full((slots, width), np.nan)
    excl = code < outputs
    out[code[excl]] = contrib[excl]
    stage[code[~excl] - outputs] = contrib[~excl]
    for i, d in enumerate(dst):
        s = np.zeros(width)
        for k in range(ptr[i], ptr[i + 1]):
            s = s + stage[k]
        out[d] = s
    return out, excl


@pytest.mark.parametrize("idx", range(16))
def test_split_tables_reproduce_scatter_add(fem, idx):
    mesh = meshes(fem)[idx]
    pat = fem.bsr_pattern(mesh)
    t = fem.split_tables(mesh, pat)
    E, N, nb, nn = mesh.num_elements, mesh.npe, pat.num_blocks, mesh.num_nodes
    rng = np.random.default_rng(idx)

    # BSR blocks: destinations from the TACS-style element map (offsets / 4)
    dest = fem.element_bcsr_map(mesh, pat) // 4
    contrib = rng.standard_normal((E * N * N, 4))
    ref = np.zeros((nb, 4))
    np.add.at(ref, dest, contrib)
    out, excl = split_assemble((t["code"], t["gdst"], t["gptr"], t["slots"]), contrib, nb, 4)
    assert np.array_equal(out, ref)  # every block written, sums in the same order
    counts = np.bincount(dest, minlength=nb)
    assert np.array_equal(np.sort(t["gdst"]), np.flatnonzero(counts != 1))
    assert t["slots"] == int(counts[counts != 1].sum())
    assert np.array_equal(excl, counts[dest] == 1)
    assert np.array_equal(t["code"][excl], dest[excl])
    assert np.all(np.diff(t["gptr"]) == counts[t["gdst"]])

    # residual entries: destinations are the element nodes
    rdest = mesh.conn.reshape(-1)
    rcontrib = rng.standard_normal((E * N, 2))
    rref = np.zeros((nn, 2))
    np.add.at(rref, rdest, rcontrib)
    rout, _ = split_assemble((t["rcode"], t["rdst"], t["rptr"], t["rslots"]), rcontrib, nn, 2)
    assert np.array_equal(rout, rref)
    assert t["rslots"] == int(np.bincount(rdest, minlength=nn)[np.bincount(rdest, minlength=nn) != 1].sum())


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [1, 3, 7])
def test_split_tables_structured_counts(fem, p, n):
    """Closed form on an n x n square of order p. A block is shared iff both nodes lie on one interior edge:
    each of the 2n(n-1) interior edges has (p+1)^2 node pairs formed by its 2 elements, except that the
    (v, v) block of a vertex collects one contribution per element at v (4 inside, 2 on a side, 1 at a
    corner). Residual entries: interior vertices 4 contributions, side vertices and nodes inside interior
    edges 2, everything else 1."""
    mesh = fem.square_mesh(n, p)
    t = fem.split_tables(mesh, fem.bsr_pattern(mesh))
    e_int, v_int, v_side = 2 * n * (n - 1), (n - 1) ** 2, 4 * (n - 1)
    pair_blocks = e_int * ((p + 1) ** 2 - 2)
    assert t["slots"] == 2 * pair_blocks + 4 * v_int + 2 * v_side
    assert len(t["gdst"]) == pair_blocks + v_int + v_side
    assert t["rslots"] == 4 * v_int + 2 * v_side + 2 * (p - 1) * e_int
    assert len(t["rdst"]) == v_int + v_side + (p - 1) * e_int


def test_split_tables_errors(fem):
    mesh = fem.square_mesh(2, 3)
    pat = fem.bsr_pattern(mesh)
    diag = fem.Pattern(np.arange(mesh.num_nodes + 1, dtype=np.int32), np.arange(mesh.num_nodes, dtype=np.int32))
    with pytest.raises(fem.FemError) as e:
        fem.split_tables(mesh, diag)
    assert e.value.status == -5
    bad = fem.Mesh(3, mesh.coords, np.ascontiguousarray(mesh.conn[:, ::-1] * 0 + mesh.num_nodes), "bad ids")
    with pytest.raises(fem.FemError):
        fem.split_tables(bad, pat)
