"""ctypes binding of libfem_amd.so (the C ABI declared in include/fem_amd.h).

Host helpers take and return numpy arrays; device entry points take torch tensors on a ROCm device
(torch is only the device-memory and stream plumbing) and launch on torch's current stream.

The library is the only compute path: if libfem_amd.so is missing, `load()` raises. There is no CPU
fallback for the device entry points.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfem_amd.so")

FEM_OK = 0
FEM_PREC_FP64 = 0
FEM_PREC_MIXED = 1
FEM_PLAN_DETERMINISTIC = 1
FEM_PLAN_RESIDUAL_ONLY = 2
FEM_PLAN_AUTOTUNE = 4
FEM_PLAN_OWNER_ONLY = 8
FEM_BATCH_GRAPH = 1

STATUS_NAMES = {
    0: "FEM_OK",
    -1: "FEM_ERR_INVALID_ORDER",
    -2: "FEM_ERR_NULL_POINTER",
    -3: "FEM_ERR_INVALID_ARGUMENT",
    -4: "FEM_ERR_HIP",
    -5: "FEM_ERR_MAP_MISS",
    -6: "FEM_ERR_OUT_OF_MEMORY",
    -7: "FEM_ERR_UNSUPPORTED",
    -8: "FEM_ERR_BAD_GEOMETRY",
    -9: "FEM_ERR_BUDGET",
}

# Every symbol include/fem_amd.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_PI64 = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "fem_last_error": ([], ctypes.c_char_p),
    "fem_version": ([], _I),
    "fem_basis_table": ([_I, _P, _P, _P, _P], _I),
    "fem_square_mesh_size": ([_I, _I, _I, _PI64, _PI64], _I),
    "fem_square_mesh": ([_I, _I, _I, _D, _U64, _P, _P], _I),
    "fem_annulus_mesh_size": ([_I, _I, _I, _PI64, _PI64], _I),
    "fem_annulus_mesh": ([_I, _I, _I, _D, _D, _D, _U64, _P, _P], _I),
    "fem_fill_uniform": ([_I64, _D, _D, _U64, _P], _I),
    "fem_fill_smooth_field": ([_I64, _P, _D, _U64, _D, _U64, _P], _I),
    "fem_min_detj": ([_I, _I, _P, _P, _P, ctypes.POINTER(_D)], _I),
    "fem_bsr_pattern_rowp": ([_I, _I, _I, _P, _P, _P], _I),
    "fem_bsr_pattern_cols": ([_I, _I, _I, _P, _P, _P, _P], _I),
    "fem_element_bcsr_map": ([_P, _P, _I, _I, _I, _P, _P, _I, _P], _I),
    "fem_split_tables": ([_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _PI64, _PI64, _PI64, _PI64], _I),
    "fem_bsr_pattern_device": ([_I, _I, _I, _P, _P, _P, _P, _I64, _PI64, _P], _I),
    "fem_bsr_pattern_device_noalloc": ([_I, _I, _I, _P, _P, _P, _P, _I64, _PI64, _P], _I),
    "fem_write_array": ([_P, _I64, ctypes.c_char_p], _I),
    "fem_read_bdf_size": ([ctypes.c_char_p, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "fem_read_bdf": ([ctypes.c_char_p, _I, _P, _P, _P, _P], _I),
    "fem_write_bdf": ([ctypes.c_char_p, _I, _I, _I, _P, _P, _P], _I),
    "fem_write_bsr_matrix": ([_I, _I, _I, _P, _P, _P, _I, _I, ctypes.c_char_p], _I),
    "fem_element_bcsr_map_device": ([_P, _P, _I, _I, _I, _P, _P, _I, _P, _P], _I),
    "fem_element_colouring": ([_I, _I, _I, _P, _P, _P, ctypes.POINTER(_I)], _I),
    "fem_residual": ([_I, _P, _P, _I, _P, _P, _D, _D, _D, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_jacobian": ([_I, _P, _P, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_jacobian_dropin": ([_I, _P, _P, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_dropin_cache_clear": ([], _I),
    "fem_dropin_cache_info": ([_I, _P, _P, _I, _P, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "fem_residual_nonlinear": ([_I, _P, _P, _I, _P, _P, _D, _D, _D, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_jacobian_nonlinear": ([_I, _P, _P, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_jacobian_coloured": (
        [_I, _P, _P, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I, _P, _P, _I, _P, ctypes.POINTER(_D)],
        _I,
    ),
    "fem_plan_create": ([_I, _I, _I, _P, _P, _P, _P, _P, _I, ctypes.POINTER(_P)], _I),
    "fem_plan_info": ([_P, _PI64, _PI64, _PI64, _PI64], _I),
    "fem_plan_tuning": ([_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], _I),
    "fem_plan_tuning_shape": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "fem_plan_assemble": ([_P, _P, _P, _D, _D, _D, _P, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_plan_assemble_accumulate": ([_P, _P, _P, _D, _D, _D, _P, _P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_plan_destroy": ([_P], _I),
    "fem_bsr_spmv": ([_I, _P, _P, _P, _P, _P, _P], _I),
    "fem_batch_create": ([_I, _I, ctypes.POINTER(_P)], _I),
    "fem_batch_clear": ([_P], _I),
    "fem_batch_add": ([_P, _P, _P, _P, _D, _D, _D, _P, _P], _I),
    "fem_batch_assemble": ([_P, _I, _P, ctypes.POINTER(_D)], _I),
    "fem_batch_destroy": ([_P], _I),
    "fem_server_create": ([_P, ctypes.POINTER(_P)], _I),
    "fem_server_add_mesh": ([_P, _I, _I, _I, _P, _P, _P, _D, _D, _D, ctypes.POINTER(_I)], _I),
    "fem_server_mesh_info": ([_P, _I, _PI64, _PI64, _PI64], _I),
    "fem_server_submit": ([_P, _I, _P, _I, _PI64], _I),
    "fem_server_poll": ([_P, _I64, ctypes.POINTER(_I)], _I),
    "fem_server_wait": ([_P, _I64, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_D)], _I),
    "fem_server_read": ([_P, _I64, _P, _P], _I),
    "fem_server_release": ([_P, _I64], _I),
    "fem_server_stats": ([_P, _PI64, _PI64, _PI64], _I),
    "fem_server_memory": ([_P, _PI64, _PI64, _PI64, _PI64], _I),
    "fem_server_reserve": ([_P, _I64], _I),
    "fem_server_destroy": ([_P], _I),
}

_lib = None


class FemError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libfem_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    # One HIP runtime per process: torch bundles its own libamdhip64 (SONAME libamdhip64.so.7). Loading
    # torch first makes this library's libamdhip64.so.7 dependency resolve to that same copy instead of
    # pulling a second runtime from /opt/rocm, which would not share devices or streams with torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != FEM_OK:
        msg = load().fem_last_error().decode()
        raise FemError(status, msg)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def _dptr(t) -> Optional[int]:
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def nodes_per_element(order: int) -> int:
    return (order + 1) ** 2


# ---- host: element tables, meshes, fields --------------------------------------------------------
def basis_table(order: int):
    n = order + 1
    phi = np.zeros(n * n)
    dphi = np.zeros(n * n)
    xg = np.zeros(n)
    wg = np.zeros(n)
    _check(load().fem_basis_table(order, _ptr(phi), _ptr(dphi), _ptr(xg), _ptr(wg)))
    return phi.reshape(n, n), dphi.reshape(n, n), xg, wg


@dataclass
class Mesh:
    order: int
    coords: np.ndarray  # float64 [numNodes, 2]
    conn: np.ndarray  # int32 [numElements, N]
    kind: str = ""

    @property
    def num_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def num_elements(self) -> int:
        return int(self.conn.shape[0])

    @property
    def npe(self) -> int:
        return nodes_per_element(self.order)

    @property
    def conn_ptr(self) -> np.ndarray:
        return (np.arange(self.num_elements, dtype=np.int64) * self.npe).astype(np.int32)


def square_mesh(n: int, order: int, jitter: float = 0.0, seed: int = 0, ny: Optional[int] = None) -> Mesh:
    ny = n if ny is None else ny
    lib = load()
    nn = ctypes.c_int64()
    ne = ctypes.c_int64()
    _check(lib.fem_square_mesh_size(n, ny, order, ctypes.byref(nn), ctypes.byref(ne)))
    coords = np.zeros((nn.value, 2))
    conn = np.zeros((ne.value, nodes_per_element(order)), dtype=np.int32)
    _check(lib.fem_square_mesh(n, ny, order, jitter, seed, _ptr(coords), _ptr(conn)))
    return Mesh(order, coords, conn, f"square {n}x{ny} p={order}")


def annulus_mesh(nr: int, ntheta: int, order: int, r_in: float = 1.0, r_out: float = 2.0, jitter: float = 0.0,
                 seed: int = 0) -> Mesh:
    lib = load()
    nn = ctypes.c_int64()
    ne = ctypes.c_int64()
    _check(lib.fem_annulus_mesh_size(nr, ntheta, order, ctypes.byref(nn), ctypes.byref(ne)))
    coords = np.zeros((nn.value, 2))
    conn = np.zeros((ne.value, nodes_per_element(order)), dtype=np.int32)
    _check(lib.fem_annulus_mesh(nr, ntheta, order, r_in, r_out, jitter, seed, _ptr(coords), _ptr(conn)))
    return Mesh(order, coords, conn, f"annulus {nr}x{ntheta} p={order}")


def fill_uniform(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    out = np.zeros(n)
    _check(load().fem_fill_uniform(n, lo, hi, seed, _ptr(out)))
    return out


def smooth_field(coords: np.ndarray, amplitude: float = 1e-2, phase_seed: int = 0, pert: float = 0.0,
                 pert_seed: int = 0) -> np.ndarray:
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    out = np.zeros(coords.shape[0] * 2)
    _check(load().fem_fill_smooth_field(coords.shape[0], _ptr(coords), amplitude, phase_seed, pert, pert_seed,
                                        _ptr(out)))
    return out


def min_detj(mesh: Mesh) -> float:
    v = ctypes.c_double()
    cp = mesh.conn_ptr
    _check(load().fem_min_detj(mesh.order, mesh.num_elements, _ptr(cp), _ptr(mesh.conn), _ptr(mesh.coords),
                               ctypes.byref(v)))
    return v.value


# ---- host: pattern, map, colouring ---------------------------------------------------------------
@dataclass
class Pattern:
    rowp: np.ndarray  # int32 [numNodes+1]
    cols: np.ndarray  # int32 [nnzb]

    @property
    def num_blocks(self) -> int:
        return int(self.rowp[-1])


def bsr_pattern(mesh: Mesh) -> Pattern:
    lib = load()
    cp = mesh.conn_ptr
    rowp = np.zeros(mesh.num_nodes + 1, dtype=np.int32)
    _check(lib.fem_bsr_pattern_rowp(mesh.num_nodes, mesh.num_elements, mesh.npe, _ptr(cp), _ptr(mesh.conn),
                                    _ptr(rowp)))
    cols = np.zeros(int(rowp[-1]), dtype=np.int32)
    _check(lib.fem_bsr_pattern_cols(mesh.num_nodes, mesh.num_elements, mesh.npe, _ptr(cp), _ptr(mesh.conn),
                                    _ptr(rowp), _ptr(cols)))
    return Pattern(rowp, cols)


def element_bcsr_map(mesh: Mesh, pattern: Pattern, bsize: int = 2) -> np.ndarray:
    cp = mesh.conn_ptr
    out = np.zeros(mesh.num_elements * mesh.npe * mesh.npe, dtype=np.int32)
    _check(load().fem_element_bcsr_map(_ptr(cp), _ptr(mesh.conn), mesh.num_elements, mesh.npe, mesh.num_nodes,
                                       _ptr(pattern.rowp), _ptr(pattern.cols), bsize, _ptr(out)))
    return out


def split_tables(mesh: Mesh, pattern: Pattern) -> dict:
    """The element-once tables a plan builds for Jacobian assemblies (fem_split_tables), on the host:
    block codes [E][N][N], residual codes [E][N], and the gather lists of the non-exclusive outputs."""
    E, N, nn, nb = mesh.num_elements, mesh.npe, mesh.num_nodes, pattern.num_blocks
    code = np.zeros(E * N * N, dtype=np.int32)
    rcode = np.zeros(E * N, dtype=np.int32)
    gdst = np.zeros(max(nb, 1), dtype=np.int32)
    gptr = np.zeros(nb + 1, dtype=np.int32)
    rdst = np.zeros(max(nn, 1), dtype=np.int32)
    rptr = np.zeros(nn + 1, dtype=np.int32)
    counts = [ctypes.c_int64() for _ in range(4)]
    _check(load().fem_split_tables(mesh.order, nn, E, _ptr(mesh.conn_ptr), _ptr(mesh.conn), _ptr(pattern.rowp),
                                   _ptr(pattern.cols), _ptr(code), _ptr(rcode), _ptr(gdst), _ptr(gptr), _ptr(rdst),
                                   _ptr(rptr), *[ctypes.byref(c) for c in counts]))
    ng, nrg = counts[1].value, counts[3].value
    return dict(code=code, rcode=rcode, gdst=gdst[:ng], gptr=gptr[:ng + 1], rdst=rdst[:nrg], rptr=rptr[:nrg + 1],
                slots=counts[0].value, rslots=counts[2].value)


def write_array(values: np.ndarray, path: str) -> None:
    """The reference's vector file: one value per line, "% .17g"."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    _check(load().fem_write_array(_ptr(v), v.size, os.fsencode(path)))


def write_bsr_matrix(pattern: "Pattern", values: np.ndarray, path: str, bsize: int = 2, ncols: Optional[int] = None,
                     max_rows: int = -1, max_cols: int = -1) -> None:
    """The reference's matrix file: "rows cols nnz", then 0-based "row col value" lines."""
    nrows = pattern.rowp.size - 1
    v = np.ascontiguousarray(values, dtype=np.float64)
    _check(load().fem_write_bsr_matrix(nrows, nrows if ncols is None else ncols, bsize, _ptr(pattern.rowp),
                                       _ptr(pattern.cols), _ptr(v), max_rows, max_cols, os.fsencode(path)))


FEM_BDF_NASTRAN_ORDER = 1


def read_bdf(path: str, nastran_order: bool = False):
    """Mesh from a BDF file (GRID and CQUADn cards): returns (Mesh, GRID ids of the nodes)."""
    lib = load()
    flags = FEM_BDF_NASTRAN_ORDER if nastran_order else 0
    order, nn, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib.fem_read_bdf_size(path.encode(), flags, ctypes.byref(order), ctypes.byref(nn), ctypes.byref(ne)))
    npe = (order.value + 1) ** 2
    coords = np.zeros((nn.value, 2))
    conn_ptr = np.zeros(ne.value + 1, dtype=np.int32)
    conn = np.zeros((ne.value, npe), dtype=np.int32)
    ids = np.zeros(nn.value, dtype=np.int64)
    _check(lib.fem_read_bdf(path.encode(), flags, _ptr(coords), _ptr(conn_ptr), _ptr(conn), _ptr(ids)))
    return Mesh(order=order.value, coords=coords, conn=conn), ids


def write_bdf(path: str, mesh) -> None:
    """Writes the mesh as GRID (free field, full precision) and CQUADn cards (tensor-product node order)."""
    cp = mesh.conn_ptr
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
    _check(load().fem_write_bdf(path.encode(), mesh.order, mesh.num_nodes, mesh.num_elements, _ptr(cp), _ptr(conn),
                                _ptr(coords)))


def bsr_pattern_device(num_nodes, num_elements, npe, conn_ptr, conn, stream=None, noalloc=False):
    """Device-built pattern from device arrays (torch int32 tensors): returns (rowp, cols) torch tensors on the
    same device, identical to bsr_pattern's arrays. noalloc: through fem_bsr_pattern_device_noalloc (the
    library's fixed per-GPU scratch instead of allocations)."""
    import torch

    rowp = torch.empty(num_nodes + 1, dtype=torch.int32, device=conn.device)
    nb = ctypes.c_int64(0)
    lib = load()
    fn = lib.fem_bsr_pattern_device_noalloc if noalloc else lib.fem_bsr_pattern_device
    _check(fn(num_nodes, num_elements, npe, _dptr(conn_ptr), _dptr(conn), _dptr(rowp), None, 0, ctypes.byref(nb),
              _stream(stream)))
    cols = torch.empty(max(nb.value, 1), dtype=torch.int32, device=conn.device)
    _check(fn(num_nodes, num_elements, npe, _dptr(conn_ptr), _dptr(conn), _dptr(rowp), _dptr(cols), cols.numel(),
              ctypes.byref(nb), _stream(stream)))
    return rowp, cols[: nb.value]


def element_bcsr_map_device(conn_ptr, conn, num_elements, npe, num_rows, rowp, cols, bsize=2, stream=None):
    """Device-built element->BSR map (torch int32 tensor of num_elements * npe^2 entries)."""
    import torch

    out = torch.empty(max(num_elements * npe * npe, 1), dtype=torch.int32, device=conn.device)
    _check(load().fem_element_bcsr_map_device(_dptr(conn_ptr), _dptr(conn), num_elements, npe, num_rows, _dptr(rowp),
                                              _dptr(cols), bsize, _dptr(out), _stream(stream)))
    return out[: num_elements * npe * npe]


def element_colouring(mesh: Mesh):
    cp = mesh.conn_ptr
    colour = np.zeros(mesh.num_elements, dtype=np.int32)
    nc = ctypes.c_int()
    _check(load().fem_element_colouring(mesh.num_nodes, mesh.num_elements, mesh.npe, _ptr(cp), _ptr(mesh.conn),
                                        _ptr(colour), ctypes.byref(nc)))
    return colour, nc.value


def colour_lists(colour: np.ndarray, num_colours: int):
    """Element ids grouped by colour and the numColours+1 offsets."""
    order = np.argsort(colour, kind="stable").astype(np.int32)
    counts = np.bincount(colour, minlength=num_colours)
    ptr = np.zeros(num_colours + 1, dtype=np.int32)
    ptr[1:] = np.cumsum(counts)
    return order, ptr


# ---- device ----------------------------------------------------------------------------------------
def _elapsed_arg(timed: bool):
    return ctypes.byref(ctypes.c_double()) if timed else None


def residual(order, conn_ptr, conn, num_elements, states, coords, E, nu, t, out, precision=FEM_PREC_FP64,
             stream=None, timed=False, nonlinear=False):
    """out += R (device tensors). Returns seconds when timed. nonlinear: the NONLINEAR (Green) strain."""
    sec = ctypes.c_double()
    fn = load().fem_residual_nonlinear if nonlinear else load().fem_residual
    _check(fn(order, _dptr(conn_ptr), _dptr(conn), num_elements, _dptr(states), _dptr(coords), E, nu, t, _dptr(out),
              precision, _stream(stream), ctypes.byref(sec) if timed else None))
    return sec.value if timed else None


def jacobian(order, conn_ptr, conn, num_elements, states, coords, E, nu, t, bcsr_map, res, values,
             precision=FEM_PREC_FP64, stream=None, timed=False, nonlinear=False, dropin=False):
    """values += K through the map; res (optional) += R. nonlinear: the NONLINEAR (Green) strain's tangent.
    dropin: through fem_jacobian_dropin (cached resident plan in accumulate mode where the map allows)."""
    sec = ctypes.c_double()
    if nonlinear and dropin:
        raise ValueError("the drop-in plan cache is LINEAR only")
    fn = (load().fem_jacobian_nonlinear if nonlinear else
          load().fem_jacobian_dropin if dropin else load().fem_jacobian)
    _check(fn(order, _dptr(conn_ptr), _dptr(conn), num_elements, _dptr(states), _dptr(coords), E, nu, t,
              _dptr(bcsr_map), _dptr(res), _dptr(values), precision, _stream(stream),
              ctypes.byref(sec) if timed else None))
    return sec.value if timed else None


def dropin_cache_clear():
    _check(load().fem_dropin_cache_clear())


def dropin_cache_info(order, conn_ptr, conn, num_elements, bcsr_map):
    """(cached, served by a plan) for a drop-in combination."""
    c, u = ctypes.c_int(), ctypes.c_int()
    _check(load().fem_dropin_cache_info(order, _dptr(conn_ptr), _dptr(conn), num_elements, _dptr(bcsr_map),
                                        ctypes.byref(c), ctypes.byref(u)))
    return bool(c.value), bool(u.value)


def jacobian_coloured(order, conn_ptr, conn, num_elements, states, coords, E, nu, t, bcsr_map, colour_elems,
                      colour_ptr: np.ndarray, res, values, precision=FEM_PREC_FP64, stream=None, timed=False):
    sec = ctypes.c_double()
    colour_ptr = np.ascontiguousarray(colour_ptr, dtype=np.int32)
    _check(load().fem_jacobian_coloured(order, _dptr(conn_ptr), _dptr(conn), num_elements, _dptr(states),
                                        _dptr(coords), E, nu, t, _dptr(bcsr_map), _dptr(colour_elems),
                                        _ptr(colour_ptr), len(colour_ptr) - 1, _dptr(res), _dptr(values), precision,
                                        _stream(stream), ctypes.byref(sec) if timed else None))
    return sec.value if timed else None


def bsr_spmv(pattern_rowp, pattern_cols, values, x, y, stream=None):
    n = pattern_rowp.numel() - 1
    _check(load().fem_bsr_spmv(n, _dptr(pattern_rowp), _dptr(pattern_cols), _dptr(values), _dptr(x), _dptr(y),
                               _stream(stream)))


class Plan:
    """Resident owner-computes assembly plan (fem_plan_*)."""

    def __init__(self, mesh: Mesh, pattern: Pattern, flags: int = 0, morton: bool = True):
        self._lib = load()
        handle = ctypes.c_void_p()
        cp = mesh.conn_ptr
        _check(self._lib.fem_plan_create(mesh.order, mesh.num_nodes, mesh.num_elements, _ptr(cp), _ptr(mesh.conn),
                                         _ptr(pattern.rowp), _ptr(pattern.cols),
                                         _ptr(mesh.coords) if morton else None, flags, ctypes.byref(handle)))
        self.handle = handle
        self.order = mesh.order
        self.flags = flags

    def info(self):
        vals = [ctypes.c_int64() for _ in range(4)]
        _check(self._lib.fem_plan_info(self.handle, *[ctypes.byref(v) for v in vals]))
        out = dict(zip(["parts", "blocks", "element_visits", "device_bytes"], [v.value for v in vals]))
        scale, sec = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.fem_plan_tuning(self.handle, ctypes.byref(scale), ctypes.byref(sec)))
        out["part_scale"] = scale.value
        shape = [ctypes.c_int() for _ in range(3)]
        _check(self._lib.fem_plan_tuning_shape(self.handle, *[ctypes.byref(v) for v in shape]))
        out["waves_per_group"], out["elements_per_wave"], out["lanes_per_element"] = [v.value for v in shape]
        return out

    def assemble(self, states, coords, E, nu, t, res=None, values=None, precision=FEM_PREC_FP64, stream=None,
                 timed=False, accumulate=False):
        """Writes (accumulate: adds to) the residual and / or BSR values."""
        sec = ctypes.c_double()
        fn = self._lib.fem_plan_assemble_accumulate if accumulate else self._lib.fem_plan_assemble
        _check(fn(self.handle, _dptr(states), _dptr(coords), E, nu, t, _dptr(res), _dptr(values), precision,
                  _stream(stream), ctypes.byref(sec) if timed else None))
        return sec.value if timed else None

    def close(self):
        if self.handle:
            self._lib.fem_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many assemblies on resident plans in one call (fem_batch_*): one persistent launch per element
    order and output set."""

    def __init__(self, max_jobs: int = 256, graph: bool = False):
        self._lib = load()
        handle = ctypes.c_void_p()
        _check(self._lib.fem_batch_create(max_jobs, FEM_BATCH_GRAPH if graph else 0, ctypes.byref(handle)))
        self.handle = handle
        self._keep = []

    def clear(self):
        _check(self._lib.fem_batch_clear(self.handle))
        self._keep = []

    def add(self, plan: Plan, states, coords, E, nu, t, res=None, values=None):
        _check(self._lib.fem_batch_add(self.handle, plan.handle, _dptr(states), _dptr(coords), E, nu, t, _dptr(res),
                                       _dptr(values)))
        self._keep.append((plan, states, coords, res, values))

    def assemble(self, precision=FEM_PREC_FP64, stream=None, timed=False):
        sec = ctypes.c_double()
        _check(self._lib.fem_batch_assemble(self.handle, precision, _stream(stream),
                                            ctypes.byref(sec) if timed else None))
        return sec.value if timed else None

    def close(self):
        if self.handle:
            self._lib.fem_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbmBudgetBytes", ctypes.c_int64), ("maxBatchJobs", ctypes.c_int), ("numSlots", ctypes.c_int),
                ("precision", ctypes.c_int), ("useGraphs", ctypes.c_int)]


class Server:
    """Serving loop (fem_server_*): resident meshes, budget admission, continuous batching."""

    def __init__(self, budget_bytes: int, max_batch_jobs: int = 64, slots: int = 2, precision: int = FEM_PREC_MIXED,
                 graphs: bool = True):
        self._lib = load()
        cfg = ServerConfig(int(budget_bytes), max_batch_jobs, slots, precision, 1 if graphs else 0)
        handle = ctypes.c_void_p()
        _check(self._lib.fem_server_create(ctypes.byref(cfg), ctypes.byref(handle)))
        self.handle = handle
        self.meshes = []

    def add_mesh(self, mesh: Mesh, E: float, nu: float, t: float) -> int:
        mid = ctypes.c_int()
        cp = mesh.conn_ptr
        _check(self._lib.fem_server_add_mesh(self.handle, mesh.order, mesh.num_nodes, mesh.num_elements, _ptr(cp),
                                             _ptr(mesh.conn), _ptr(np.ascontiguousarray(mesh.coords)), E, nu, t,
                                             ctypes.byref(mid)))
        self.meshes.append(mesh)
        return mid.value

    def mesh_info(self, mesh_id: int):
        v = [ctypes.c_int64() for _ in range(3)]
        _check(self._lib.fem_server_mesh_info(self.handle, mesh_id, *[ctypes.byref(x) for x in v]))
        return dict(zip(["nodes", "blocks", "resident_bytes"], [x.value for x in v]))

    def submit(self, mesh_id: int, states, with_jacobian: bool) -> int:
        rid = ctypes.c_int64()
        _check(self._lib.fem_server_submit(self.handle, mesh_id, _dptr(states), 1 if with_jacobian else 0,
                                           ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int):
        times = (ctypes.c_double * 3)()
        _check(self._lib.fem_server_wait(self.handle, rid, None, None, times))
        return tuple(times)

    def poll(self, rid: int) -> bool:
        done = ctypes.c_int()
        _check(self._lib.fem_server_poll(self.handle, rid, ctypes.byref(done)))
        return bool(done.value)

    def read(self, rid: int, num_nodes: int, num_blocks: int = 0):
        res = np.zeros(2 * num_nodes)
        vals = np.zeros(4 * num_blocks) if num_blocks else None
        _check(self._lib.fem_server_read(self.handle, rid, _ptr(res), _ptr(vals)))
        return res, vals

    def release(self, rid: int):
        _check(self._lib.fem_server_release(self.handle, rid))

    def stats(self):
        v = [ctypes.c_int64() for _ in range(3)]
        _check(self._lib.fem_server_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(["batches", "completed", "bytes_in_use"], [x.value for x in v]))

    def reserve(self, arena_bytes: int):
        """Creates the output arena now with arena_bytes (the rest of the budget stays open to meshes)."""
        _check(self._lib.fem_server_reserve(self.handle, int(arena_bytes)))

    def memory(self):
        """Device bytes of the server: resident meshes, the arena, arena bytes held now and their peak."""
        v = [ctypes.c_int64() for _ in range(4)]
        _check(self._lib.fem_server_memory(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(["resident", "arena", "arena_in_use", "arena_peak"], [x.value for x in v]))

    def close(self):
        if self.handle:
            self._lib.fem_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
