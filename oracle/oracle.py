"""ORACLE — test infrastructure only (loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, never by the product). ctypes loader for oracle/oracle.cpp, the CPU restatement of the
reference's assembly; see the header of oracle.cpp for the reference lines each function follows.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PARITY_LIB = os.path.join(HERE, "_build", "liboracle.so")
FAST_LIB = os.path.join(HERE, "_build", "liboracle_fast.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SIGS = {
    "oracle_gauss": ([_I, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)], _I),
    "oracle_lagrange1d": ([_I, _D, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)], _I),
    "oracle_element": ([_I, _P, _P, _D, _D, _D, _I, _P, _P], _I),
    "oracle_residual": ([_I, _P, _P, _I, _I, _P, _P, _D, _D, _D, _P, _I], _I),
    "oracle_jacobian": ([_I, _P, _P, _I, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I, _I], _I),
    "oracle_bsr_pattern": ([_I, _I, _I, _P, _P, _P, _P], _I),
    "oracle_element_bcsr_map": ([_P, _P, _I, _I, _P, _P, _I, _P], _I),
    "oracle_time_res_jac": ([_I, _P, _P, _I, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I], _D),
    "oracle_max_threads": ([], _I),
    "oracle_element_nl": ([_I, _P, _P, _D, _D, _D, _P, _P], _I),
    "oracle_residual_nl": ([_I, _P, _P, _I, _I, _P, _P, _D, _D, _D, _P, _I], _I),
    "oracle_jacobian_nl": ([_I, _P, _P, _I, _I, _P, _P, _D, _D, _D, _P, _P, _P, _I], _I),
}
_libs: dict = {}


def load(fast: bool = False) -> ctypes.CDLL:
    path = FAST_LIB if fast else PARITY_LIB
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _libs[path] = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def gauss(order: int, i: int):
    x = ctypes.c_double()
    w = ctypes.c_double()
    assert load().oracle_gauss(order, i, ctypes.byref(x), ctypes.byref(w)) == 0
    return x.value, w.value


def lagrange1d(order: int, x: float, k: int):
    n = ctypes.c_double()
    d = ctypes.c_double()
    assert load().oracle_lagrange1d(order, x, k, ctypes.byref(n), ctypes.byref(d)) == 0
    return n.value, d.value


def element(order: int, lx: np.ndarray, lu: np.ndarray, E: float, nu: float, t: float, gram: bool = False,
            nonlinear: bool = False):
    """Residual (2N) and row-major stiffness (2N x 2N) of one element from nodal coords/states (nonlinear:
    the NONLINEAR Green strain, whose stiffness is the tangent at lu)."""
    N = (order + 1) ** 2
    lx = np.ascontiguousarray(lx, dtype=np.float64).reshape(-1)
    lu = np.ascontiguousarray(lu, dtype=np.float64).reshape(-1)
    res = np.zeros(2 * N)
    K = np.zeros((2 * N, 2 * N))
    if nonlinear:
        assert load().oracle_element_nl(order, _p(lx), _p(lu), E, nu, t, _p(res), _p(K)) == 0
    else:
        assert load().oracle_element(order, _p(lx), _p(lu), E, nu, t, int(gram), _p(res), _p(K)) == 0
    return res, K


def residual(mesh, states, E, nu, t, threads: int = 1, first: int = 0, count=None, nonlinear: bool = False):
    count = mesh.num_elements - first if count is None else count
    out = np.zeros(mesh.num_nodes * 2)
    cp = mesh.conn_ptr
    fn = load().oracle_residual_nl if nonlinear else load().oracle_residual
    assert fn(mesh.order, _p(cp), _p(mesh.conn), first, count, _p(states), _p(mesh.coords), E, nu, t, _p(out),
              threads) == 0
    return out


def jacobian(mesh, states, E, nu, t, bcsr_map, num_blocks, with_residual=True, gram=False, threads: int = 1,
             nonlinear: bool = False):
    values = np.zeros(num_blocks * 4)
    res = np.zeros(mesh.num_nodes * 2) if with_residual else None
    cp = mesh.conn_ptr
    if nonlinear:
        assert load().oracle_jacobian_nl(mesh.order, _p(cp), _p(mesh.conn), 0, mesh.num_elements, _p(states),
                                         _p(mesh.coords), E, nu, t, _p(bcsr_map), _p(res), _p(values), threads) == 0
        return res, values
    assert load().oracle_jacobian(mesh.order, _p(cp), _p(mesh.conn), 0, mesh.num_elements, _p(states),
                                  _p(mesh.coords), E, nu, t, _p(bcsr_map), _p(res), _p(values), int(gram),
                                  threads) == 0
    return res, values


def bsr_pattern(mesh):
    cp = mesh.conn_ptr
    rowp = np.zeros(mesh.num_nodes + 1, dtype=np.int32)
    lib = load()
    lib.oracle_bsr_pattern(mesh.num_nodes, mesh.num_elements, mesh.npe, _p(cp), _p(mesh.conn), _p(rowp), None)
    cols = np.zeros(int(rowp[-1]), dtype=np.int32)
    lib.oracle_bsr_pattern(mesh.num_nodes, mesh.num_elements, mesh.npe, _p(cp), _p(mesh.conn), _p(rowp), _p(cols))
    return rowp, cols


def element_bcsr_map(mesh, rowp, cols, bsize: int = 2):
    cp = mesh.conn_ptr
    out = np.zeros(mesh.num_elements * mesh.npe * mesh.npe, dtype=np.int32)
    misses = load().oracle_element_bcsr_map(_p(cp), _p(mesh.conn), mesh.num_elements, mesh.npe, _p(rowp), _p(cols),
                                            bsize, _p(out))
    return out, misses


def time_res_jac(mesh, states, E, nu, t, bcsr_map, num_blocks, count: int, threads: int):
    """Seconds for residual + Jacobian over the first `count` elements (timing build)."""
    lib = load(fast=True)
    res = np.zeros(mesh.num_nodes * 2)
    vals = np.zeros(num_blocks * 4)
    cp = mesh.conn_ptr
    return lib.oracle_time_res_jac(mesh.order, _p(cp), _p(mesh.conn), 0, count, _p(states), _p(mesh.coords), E, nu,
                                   t, _p(bcsr_map), _p(res), _p(vals), threads)


def max_threads(fast: bool = True) -> int:
    return load(fast).oracle_max_threads()
