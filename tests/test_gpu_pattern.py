"""GPU: the device-built BSR pattern and element map are identical to the host builders (and, on small
meshes, to a brute-force numpy pattern), through the C ABI (fem_bsr_pattern_device,
fem_element_bcsr_map_device). Integer results: bit-exact equality. Error paths (invalid node id, short cols
buffer, map miss, a node beyond the per-row sort capacity) come back as status codes."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def brute_pattern(mesh):
    rows = [set() for _ in range(mesh.num_nodes)]
    for en in mesh.conn:
        for a in en:
            rows[a].update(int(b) for b in en)
    for n, r in enumerate(rows):
        if not r:
            r.add(n)  # isolated node: diagonal block only
    rowp = np.zeros(mesh.num_nodes + 1, dtype=np.int64)
    rowp[1:] = np.cumsum([len(r) for r in rows])
    cols = np.concatenate([np.array(sorted(r), dtype=np.int64) for r in rows])
    return rowp, cols


def device_pattern(fem, torch, mesh):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    cp, conn = d(mesh.conn_ptr), d(mesh.conn.reshape(-1))
    rowp, cols = fem.bsr_pattern_device(mesh.num_nodes, mesh.num_elements, mesh.npe, cp, conn)
    bmap = fem.element_bcsr_map_device(cp, conn, mesh.num_elements, mesh.npe, mesh.num_nodes, rowp, cols)
    torch.cuda.synchronize()
    return rowp.cpu().numpy(), cols.cpu().numpy(), bmap.cpu().numpy()


def permuted(fem, mesh, seed):
    """Same mesh with shuffled node ids and element order (irregular node numbering)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(mesh.num_nodes).astype(np.int32)
    conn = perm[mesh.conn][rng.permutation(mesh.num_elements)]
    coords = np.empty_like(mesh.coords)
    coords[perm] = mesh.coords
    return fem.Mesh(mesh.order, coords, np.ascontiguousarray(conn), mesh.kind + " permuted")


def with_isolated(fem, mesh, extra):
    coords = np.vstack([mesh.coords, np.zeros((extra, 2))])
    return fem.Mesh(mesh.order, coords, mesh.conn.copy(), mesh.kind + " + isolated nodes")


def fan_mesh(fem, k):
    """k order-1 quads sharing node 0 (node 0's row gathers 4k node references)."""
    conn = np.array([[0, 1 + 3 * i, 2 + 3 * i, 3 + 3 * i] for i in range(k)], dtype=np.int32)
    coords = np.random.default_rng(k).random((1 + 3 * k, 2))
    return fem.Mesh(1, coords, conn, f"fan of {k}")


def meshes(fem):
    out = []
    for p in (1, 2, 3, 4):
        out.append(fem.square_mesh(6, p, jitter=0.2, seed=p))
        out.append(fem.annulus_mesh(2, 9, p, jitter=0.1, seed=p))
    out.append(permuted(fem, fem.square_mesh(5, 2, jitter=0.2, seed=9), 11))
    out.append(permuted(fem, fem.annulus_mesh(3, 8, 4), 12))
    out.append(with_isolated(fem, fem.square_mesh(4, 3), 7))
    out.append(fan_mesh(fem, 200))  # 800 references at node 0: within one wavefront's sort
    out.append(fem.square_mesh(1, 1))
    return out


def test_device_pattern_matches_host_and_brute_force(fem, torch_gpu):
    torch = torch_gpu
    for mesh in meshes(fem):
        pat = fem.bsr_pattern(mesh)
        hmap = fem.element_bcsr_map(mesh, pat)
        rowp, cols, dmap = device_pattern(fem, torch, mesh)
        assert np.array_equal(rowp, pat.rowp), mesh.kind
        assert np.array_equal(cols, pat.cols), mesh.kind
        assert np.array_equal(dmap, hmap), mesh.kind
        brow, bcol = brute_pattern(mesh)
        assert np.array_equal(rowp, brow) and np.array_equal(cols, bcol), mesh.kind


@pytest.mark.parametrize("n,p", [(300, 2), (60, 4)])
def test_device_pattern_matches_host_large(fem, torch_gpu, n, p):
    torch = torch_gpu
    mesh = fem.square_mesh(n, p, jitter=0.2, seed=5)
    pat = fem.bsr_pattern(mesh)
    rowp, cols, dmap = device_pattern(fem, torch, mesh)
    assert np.array_equal(rowp, pat.rowp)
    assert np.array_equal(cols, pat.cols)
    assert np.array_equal(dmap, fem.element_bcsr_map(mesh, pat))


def test_device_pattern_empty_mesh(fem, torch_gpu):
    torch = torch_gpu
    mesh = fem.Mesh(2, np.zeros((3, 2)), np.zeros((0, 9), dtype=np.int32), "no elements")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    rowp, cols = fem.bsr_pattern_device(3, 0, 9, d(np.zeros(1, np.int32)), d(np.zeros(1, np.int32)))
    assert rowp.cpu().tolist() == [0, 1, 2, 3]
    assert cols.cpu().tolist() == [0, 1, 2]


def test_device_pattern_errors(fem, torch_gpu):
    torch = torch_gpu
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    mesh = fem.square_mesh(3, 1)
    cp = d(mesh.conn_ptr)
    bad = mesh.conn.reshape(-1).copy()
    bad[5] = mesh.num_nodes  # out of range
    with pytest.raises(fem.FemError) as e:
        fem.bsr_pattern_device(mesh.num_nodes, mesh.num_elements, 4, cp, d(bad))
    assert e.value.status == -3

    conn = d(mesh.conn.reshape(-1))
    rowp = torch.empty(mesh.num_nodes + 1, dtype=torch.int32, device="cuda:0")
    cols = torch.empty(3, dtype=torch.int32, device="cuda:0")
    lib = fem.load()
    nb = ctypes.c_int64(0)
    rc = lib.fem_bsr_pattern_device(mesh.num_nodes, mesh.num_elements, 4, fem._dptr(cp), fem._dptr(conn),
                                    fem._dptr(rowp), fem._dptr(cols), 3, ctypes.byref(nb), fem._stream())
    assert rc == -3 and nb.value == fem.bsr_pattern(mesh).num_blocks

    # map against a diagonal-only pattern: a miss
    diag = torch.arange(mesh.num_nodes + 1, dtype=torch.int32, device="cuda:0")
    with pytest.raises(fem.FemError) as e:
        fem.element_bcsr_map_device(cp, conn, mesh.num_elements, 4, mesh.num_nodes, diag,
                                    torch.arange(mesh.num_nodes, dtype=torch.int32, device="cuda:0"))
    assert e.value.status == -5

    fan = fan_mesh(fem, 300)  # 1200 references at node 0: beyond one wavefront's sort
    with pytest.raises(fem.FemError) as e:
        fem.bsr_pattern_device(fan.num_nodes, fan.num_elements, 4, d(fan.conn_ptr), d(fan.conn.reshape(-1)))
    assert e.value.status == -7
