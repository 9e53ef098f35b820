"""GPU: fem_bsr_pattern_device_noalloc (fixed per-GPU scratch, no device allocation) through the C ABI against
fem_bsr_pattern_device on the same meshes, bit-exact (integer arrays), on two calls in a row with the second mesh
smaller than the first (the scratch keeps the first call's contents), and its size limit: a mesh whose scratch
exceeds the fixed array is rejected with FEM_ERR_INVALID_ARGUMENT before any device work."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _patterns(fem, torch, mesh):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    cp, conn = d(mesh.conn_ptr), d(mesh.conn.reshape(-1))
    ref = fem.bsr_pattern_device(mesh.num_nodes, mesh.num_elements, mesh.npe, cp, conn)
    got = fem.bsr_pattern_device(mesh.num_nodes, mesh.num_elements, mesh.npe, cp, conn, noalloc=True)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in ref], [t.cpu().numpy() for t in got]


def test_noalloc_pattern_matches_allocating_builder(fem, torch_gpu):
    big = fem.square_mesh(300, 2, jitter=0.2, seed=5)   # 361 201 nodes, 810 000 node references
    small = fem.annulus_mesh(6, 40, 3, jitter=0.1, seed=6)
    for mesh in (big, small):
        (rowp_ref, cols_ref), (rowp, cols) = _patterns(fem, torch_gpu, mesh)
        assert np.array_equal(rowp, rowp_ref)
        assert np.array_equal(cols, cols_ref)
        host = fem.bsr_pattern(mesh)
        assert np.array_equal(rowp, host.rowp) and np.array_equal(cols, host.cols)


def test_noalloc_pattern_rejects_meshes_beyond_its_scratch(fem, torch_gpu):
    torch = torch_gpu
    mesh = fem.square_mesh(1000, 2)  # 4 004 001 nodes, 9M node references: about 84 MB of scratch
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    cp, conn = d(mesh.conn_ptr), d(mesh.conn.reshape(-1))
    rowp = torch.empty(mesh.num_nodes + 1, dtype=torch.int32, device="cuda:0")
    nb = ctypes.c_int64(-1)
    lib = fem.load()
    rc = lib.fem_bsr_pattern_device_noalloc(mesh.num_nodes, mesh.num_elements, mesh.npe, cp.data_ptr(),
                                            conn.data_ptr(), rowp.data_ptr(), None, 0, ctypes.byref(nb), None)
    assert rc == -3 and nb.value == -1
    assert b"64" in lib.fem_last_error() or b"scratch" in lib.fem_last_error()
