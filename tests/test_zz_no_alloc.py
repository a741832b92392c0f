"""fp_project_noalloc: the projection entry point that allocates nothing.

It is checked against fp_project on the same grid and inputs through the C
ABI, to the parity bar of integer work (the same kernels in the same order:
bitwise equal), on two calls in a row, the second on a smaller grid: the
scratch keeps the first call's tables and state, which the second must not
see."""
import ctypes as C

import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


def noalloc(fpl, g, s, precision):
    lib = fpl.load_library()
    ins, outs, res, keep = g._prepare(s.phi, s.u, s.v, s.w, s.dt, s.rho)
    o = fpl.default_options(precision=precision, relative_tolerance=1e-6)
    rep = fpl.Report()
    rc = lib.fp_project_noalloc(g._g, C.byref(ins), C.byref(outs), C.byref(o), C.byref(rep))
    assert rc == 0, lib.fp_last_error()
    del keep
    return res, rep.as_dict()


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_noalloc_matches_fp_project_twice(fpl, precision):
    big = scenes.dam_spheres_scene(96, seed=4)
    small = scenes.dam_break_scene(40, 5)
    for s in (big, small):
        g = fpl.UniformPressureProjection(s.shape, s.h)
        try:
            ref = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, relative_tolerance=1e-6)
            (p, u, v, w), rep = noalloc(fpl, g, s, precision)
            assert rep["flags"] == 0 and rep["relative_residual"] <= 1e-6, rep
            assert rep["iterations"] == ref.report["iterations"]
            assert rep["relative_residual"] == ref.report["relative_residual"]
            for a, b in ((p, ref.pressure), (u, ref.u), (v, ref.v), (w, ref.w)):
                assert np.array_equal(a, b)
        finally:
            g.destroy()


def test_noalloc_rejects_tall_grids(fpl):
    sc = scenes.tall_random_scene(8, 8, 24, 8, seed=1)
    g = fpl.TallCellPressureProjection.from_layout(sc.layout)
    try:
        lib = fpl.load_library()
        ins, outs, res, keep = g._prepare(sc.phi, sc.u, sc.v, sc.w, sc.dt, sc.rho)
        o = fpl.default_options()
        rep = fpl.Report()
        assert lib.fp_project_noalloc(g._g, C.byref(ins), C.byref(outs), C.byref(o), C.byref(rep)) == -1
        assert b"uniform grids only" in lib.fp_last_error()
    finally:
        g.destroy()
