"""Tall-cell remesh on the GPU (fp_tall_remesh) against the CPU restatement.

Bar: bit-exact (bases, phi, u, v, w, changed-column count) -- the device and
the restatement run the same FP64 operations in the same order without
contraction.  Also: device tensors updated in place give the host result;
the grid projects with its new layout exactly as a grid created on the new
bases does; a second remesh changes nothing; refused forms.
"""
import ctypes as C

import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd import tall_layout as tl
import tall_remesh as orm
from remesh_cases import dropped_scene, moved_scene

pytestmark = pytest.mark.gpu


def oracle_remesh(lay, phi, u, v, w, lh=16, ah=8):
    return orm.remesh(lay.bases, phi, u, v, w, lay.fine_ny, lay.band, lay.max_delta, lay.h, lh, ah)


def assert_same(got, want):
    nb, phi, u, v, w, moved = want
    print(f"changed columns {moved} of {nb.size}", flush=True)
    assert got[4] == moved
    for a, b, name in zip(got[:4], (phi, u, v, w), ("phi", "u", "v", "w")):
        a = np.asarray(a)
        assert a.shape == b.shape, name
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, f"{name}: {bad.size} values differ, first at {np.unravel_index(bad[0], a.shape)}"


CASES = {
    "random_drop": lambda: moved_scene(seed=0, drop=18),
    "random_rise": lambda: moved_scene(seed=2, drop=-12),
    "delta5": lambda: moved_scene(nx=9, nz=14, ny=80, band=20, seed=3, drop=25, max_delta=5),
    "linear": lambda: moved_scene(seed=4, drop=20, linear=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_remesh_bit_exact(fpl, name):
    lay, phi, u, v, w = CASES[name]()
    print(f"case {name}: {lay.nx}x{lay.nz} columns, {lay.fine_ny} rows, band {lay.band}", flush=True)
    want = oracle_remesh(lay, phi, u, v, w)
    g = fpl.TallCellPressureProjection.from_layout(lay)
    try:
        got = g.remesh(phi, u, v, w)
        np.testing.assert_array_equal(g.bases, want[0])
        assert_same(got, want)
    finally:
        g.destroy()


def test_remesh_halos_and_dam(fpl):
    sc = scenes.tall_dam_scene(24, 48, band=16, fill=(0.5, 0.6, 0.45), seed=8)
    lay = sc.layout
    phi = dropped_scene(sc, 9, columns=12)
    want = oracle_remesh(lay, phi, sc.u, sc.v, sc.w, lh=5, ah=3)
    g = fpl.TallCellPressureProjection.from_layout(lay)
    try:
        got = g.remesh(phi, sc.u, sc.v, sc.w, liquid_halo=5, air_halo=3)
        np.testing.assert_array_equal(g.bases, want[0])
        assert_same(got, want)
    finally:
        g.destroy()


def test_device_in_place_matches_host(fpl):
    import torch
    lay, phi, u, v, w = moved_scene(seed=5, drop=22)
    g = fpl.TallCellPressureProjection.from_layout(lay)
    h = fpl.TallCellPressureProjection.from_layout(lay)
    try:
        want = h.remesh(phi, u, v, w)
        t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (phi, u, v, w)]
        got = g.remesh(*t, out=t)
        assert got[4] == want[4]
        for a, b in zip(t, want[:4]):
            np.testing.assert_array_equal(a.cpu().numpy(), b)
        np.testing.assert_array_equal(g.bases, h.bases)
    finally:
        g.destroy()
        h.destroy()


def test_projection_uses_the_new_layout(fpl):
    """After a remesh the grid projects exactly as a fresh grid built on the
    new bases (its warm start cleared, its plans rebuilt)."""
    lay, phi, u, v, w = moved_scene(nx=16, nz=12, ny=64, band=16, seed=6, drop=14)
    g = fpl.TallCellPressureProjection.from_layout(lay)
    try:
        # a projection first, so that a warm start and plans of the old layout exist
        g.project(phi, u, v, w, 1.0 / 120.0, 1000.0)
        phi1, u1, v1, w1, moved = g.remesh(phi, u, v, w)
        assert moved > 0
        new = tl.TallCellLayout(g.bases, lay.fine_ny, lay.band, lay.h, lay.max_delta)
        new.validate()
        f = fpl.TallCellPressureProjection.from_layout(new)
        try:
            a = g.project(phi1, u1, v1, w1, 1.0 / 120.0, 1000.0)
            b = f.project(phi1, u1, v1, w1, 1.0 / 120.0, 1000.0)
            assert a.report["flags"] == b.report["flags"] == 0, a.report
            for x, y in zip((a.pressure, a.u, a.v, a.w), (b.pressure, b.u, b.v, b.w)):
                np.testing.assert_array_equal(x, y)
        finally:
            f.destroy()
    finally:
        g.destroy()


def test_second_remesh_is_identity(fpl):
    lay, phi, u, v, w = moved_scene(seed=8, drop=28)
    g = fpl.TallCellPressureProjection.from_layout(lay)
    try:
        first = g.remesh(phi, u, v, w)
        assert first[4] > 0
        bases = g.bases.copy()
        second = g.remesh(*first[:4])
        assert second[4] == 0
        np.testing.assert_array_equal(g.bases, bases)
        for a, b in zip(first[:4], second[:4]):
            np.testing.assert_array_equal(a, b)
    finally:
        g.destroy()


def test_refused_forms(fpl):
    lib = fpl.load_library()
    u = fpl.UniformPressureProjection((8, 8, 8), 0.1)
    z = np.zeros(16, np.float32)
    args = fpl.TallRemeshArgs(0, 0, 0, 0, z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                              None, None, None, None, None)
    try:
        assert lib.fp_tall_remesh(u.handle, C.byref(args), None) == fpl.ERR_INVALID_ARGUMENT
    finally:
        u.destroy()
    # no tall cell to re-lay: fine_ny - band < 2
    g = fpl.TallCellPressureProjection(np.full((6, 5), 2, np.int32), 13, 12, 0.1)
    try:
        with pytest.raises(fpl.FluidProjectionError) as e:
            s = g.shapes()
            g.remesh(*[np.zeros(x, np.float32) for x in s])
        assert e.value.code == fpl.ERR_LAYOUT
    finally:
        g.destroy()


def test_full_size_wave(fpl):
    """The 512^2-column, 512-row, 32-row-band wave with its surface 4 rows
    lower everywhere: every column moves; bit-exact, then idempotent."""
    sc = scenes.tall_wave_scene(512, 512, 32, seed=0)
    lay = sc.layout
    phi = dropped_scene(sc, 4)
    print("full-size oracle remesh", flush=True)
    want = oracle_remesh(lay, phi, sc.u, sc.v, sc.w)
    g = fpl.TallCellPressureProjection.from_layout(lay)
    try:
        got = g.remesh(phi, sc.u, sc.v, sc.w)
        np.testing.assert_array_equal(g.bases, want[0])
        assert_same(got, want)
        again = g.remesh(*got[:4])
        assert again[4] == 0
    finally:
        g.destroy()
