"""The C ABI library loads, exports every symbol include/fem_amd.h declares, and rejects bad input on the
host side (no GPU needed for any of these calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fem_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("fem_residual", "fem_jacobian", "fem_jacobian_coloured", "fem_plan_create", "fem_plan_assemble",
                 "fem_element_bcsr_map", "fem_bsr_pattern_rowp", "fem_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(fem):
    lib = fem.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) == set(fem.SIGNATURES)


def test_version_and_error_string(fem):
    lib = fem.load()
    assert lib.fem_version() >= 100
    with pytest.raises(fem.FemError):
        fem.basis_table(9)
    assert b"order" in lib.fem_last_error()


def test_device_entry_points_validate_before_launch(fem):
    lib = fem.load()
    sec = ctypes.c_double()
    # invalid order: the reference silently no-ops (FEKernels.h:1092-1093); here it is an error
    assert lib.fem_residual(5, None, None, 0, None, None, 70e9, 0.3, 1.0, None, 0, None, ctypes.byref(sec)) == -1
    assert lib.fem_jacobian(0, None, None, 10, None, None, 70e9, 0.3, 1.0, None, None, None, 0, None, None) == -1
    # null pointers with elements present
    assert lib.fem_residual(2, None, None, 10, None, None, 70e9, 0.3, 1.0, None, 0, None, None) == -2
    # unknown precision mode
    assert lib.fem_residual(2, None, None, 0, None, None, 70e9, 0.3, 1.0, None, 7, None, None) == -3
    # plan handle checks
    assert lib.fem_plan_assemble(None, None, None, 70e9, 0.3, 1.0, None, None, 0, None, None) == -2
    assert lib.fem_plan_destroy(None) == 0


def test_plan_create_rejects_bad_meshes_on_host(fem):
    lib = fem.load()
    m = fem.square_mesh(2, 2)
    pat = fem.bsr_pattern(m)
    handle = ctypes.c_void_p()
    cp = m.conn_ptr
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    # order outside 1..4
    assert lib.fem_plan_create(6, m.num_nodes, m.num_elements, p(cp), p(m.conn), p(pat.rowp), p(pat.cols), None, 0,
                               ctypes.byref(handle)) == -1
    # node id out of range
    bad = m.conn.copy()
    bad[1, 3] = m.num_nodes + 5
    assert lib.fem_plan_create(2, m.num_nodes, m.num_elements, p(cp), p(bad), p(pat.rowp), p(pat.cols), None, 0,
                               ctypes.byref(handle)) == -3
    # pattern missing a block
    cols = pat.cols.copy()
    cols[pat.rowp[4]] = cols[pat.rowp[4] + 1]
    assert lib.fem_plan_create(2, m.num_nodes, m.num_elements, p(cp), p(m.conn), p(pat.rowp), p(cols), None, 0,
                               ctypes.byref(handle)) == -5


def test_plan_autotune_rejects_other_flags_and_missing_coords(fem):
    lib = fem.load()
    m = fem.square_mesh(2, 2)
    pat = fem.bsr_pattern(m)
    handle = ctypes.c_void_p()
    cp = m.conn_ptr
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    for extra in (fem.FEM_PLAN_DETERMINISTIC, fem.FEM_PLAN_RESIDUAL_ONLY):
        assert lib.fem_plan_create(2, m.num_nodes, m.num_elements, p(cp), p(m.conn), p(pat.rowp), p(pat.cols),
                                   p(m.coords), fem.FEM_PLAN_AUTOTUNE | extra, ctypes.byref(handle)) == -3
    assert lib.fem_plan_create(2, m.num_nodes, m.num_elements, p(cp), p(m.conn), p(pat.rowp), p(pat.cols), None,
                               fem.FEM_PLAN_AUTOTUNE, ctypes.byref(handle)) == -2
    assert not handle.value


def test_cpp_dropin_header_compiles_against_library():
    # The C++ drop-in demo (include/FEKernels.h + libfem_amd.so) is built by build().
    assert os.path.exists(os.path.join(ROOT, "gpu-fem-project_amd", "fem_dropin_demo"))
