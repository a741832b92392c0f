"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and rejects bad arguments without touching a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluid_projection.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\*|int)\s+(fp_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("fp_ctx_create", "fp_grid_create", "fp_project", "fp_grid_destroy", "fp_ctx_destroy", "fp_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol(fpl):
    lib = fpl.load_library()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", fpl.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fp_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_abi_version_and_defaults(fpl):
    lib = fpl.load_library()
    assert lib.fp_abi_version() == 3
    o = fpl.default_options()
    assert o.relative_tolerance == 1e-6
    assert o.pre_sweeps == o.post_sweeps == 2
    assert o.precision == fpl.PRECISIONS["fp16"]
    assert o.solver == fpl.SOLVERS["mgpcg"] and o.open_top == 0
    assert o.full_cycles == o.v_cycles == o.lcp_sweeps == 0  # the CM11a schedule's defaults
    # the structs match the header's layout
    assert C.sizeof(fpl.Options) == 80 and C.sizeof(fpl.Report) == 104


def test_grid_bytes_and_layout_rejection(fpl):
    lib = fpl.load_library()
    b = C.c_uint64()
    d = fpl.GridDesc(0, 64, 64, 64, 0, 0, 1 / 64, None)
    assert lib.fp_grid_bytes(C.byref(d), 1, C.byref(b)) == 0
    assert b.value > 64 ** 3 * 40
    d = fpl.GridDesc(0, 1, 64, 64, 0, 0, 1 / 64, None)
    assert lib.fp_grid_bytes(C.byref(d), 1, C.byref(b)) == -1
    assert b"dims" in lib.fp_last_error()
    d = fpl.GridDesc(0, 64, 64, 64, 0, 0, -1.0, None)
    assert lib.fp_grid_bytes(C.byref(d), 1, C.byref(b)) == -1


def test_null_arguments_are_errors(fpl):
    lib = fpl.load_library()
    assert lib.fp_ctx_create(0, 0, None) == -1
    assert lib.fp_project(None, None, None, None, None) == -1
    assert lib.fp_grid_destroy(None) == 0
    assert lib.fp_ctx_destroy(None) == 0


def test_no_device_is_a_clean_error(fpl):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    lib = fpl.load_library()
    h = C.c_void_p()
    assert lib.fp_ctx_create(0, 0, C.byref(h)) == -2
    assert b"device" in lib.fp_last_error()
