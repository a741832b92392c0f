"""Velocity extrapolation on the GPU (fp_extrapolate_velocity) against the
CPU restatement (orc_extrapolate).

Bar: bit-exact -- u, v, w and the report (band passes, faces still active,
hierarchy levels, faces known after the band, open faces left unknown) --
since both run the same FP32 operations in the same order without
contraction.  Cases: flat pools (closed and open top), a droplet over a pool,
a single liquid cell, no liquid and no air, non-cubic and odd dims, density
input, a pass ceiling that stops an unconverged front, the dam + spheres
generator at 64^3, 128^3 and its full 384^3 size; device tensors updated in
place; refused forms.
"""
import ctypes as C

import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


def pool(shape, depth, h):
    return np.fromfunction(lambda i, j, k: (j + 0.5) * h - depth, shape).astype(np.float32)


def fields(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((nx + 1, ny, nz)).astype(np.float32),
            rng.standard_normal((nx, ny + 1, nz)).astype(np.float32),
            rng.standard_normal((nx, ny, nz + 1)).astype(np.float32))


def droplet(n):
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.minimum(Y - 0.3, np.sqrt((X - 0.45) ** 2 + (Y - 0.7) ** 2 + (Z - 0.55) ** 2) - 0.12).astype(np.float32)


def single(shape):
    phi = np.ones(shape, np.float32)
    phi[shape[0] // 2, shape[1] // 3, shape[2] // 2] = -1.0
    return phi


def noisy(shape, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / max(shape)
    base = pool(shape, 0.45, h)
    return (base + rng.uniform(-3, 3, size=shape).astype(np.float32) * h).astype(np.float32)


CASES = {
    "pool16": dict(shape=(16, 16, 16), phi=lambda s: pool(s, 0.4, 1 / 16)),
    "pool16_open_top": dict(shape=(16, 16, 16), phi=lambda s: pool(s, 0.4, 1 / 16), open_top=1),
    "droplet24": dict(shape=(24, 24, 24), phi=lambda s: droplet(24)),
    "single": dict(shape=(12, 13, 11), phi=single),
    "all_liquid": dict(shape=(8, 9, 10), phi=lambda s: -np.ones(s, np.float32)),
    "no_liquid": dict(shape=(8, 9, 10), phi=lambda s: np.ones(s, np.float32)),
    "noisy_odd": dict(shape=(21, 14, 17), phi=lambda s: noisy(s, 3)),
    "thin": dict(shape=(32, 5, 24), phi=lambda s: pool(s, 0.05, 1 / 32)),
    "ceiling": dict(shape=(20, 20, 20), phi=lambda s: droplet(20), max_passes=2),
    "density": dict(shape=(18, 16, 14), phi=lambda s: noisy(s, 4), density=True),
}


def run_case(fpl, oracle, name):
    c = CASES[name]
    shape = c["shape"]
    h = 1.0 / max(shape)
    phi = c["phi"](shape)
    u, v, w = fields(shape, 7)
    kw = dict(open_top=c.get("open_top", 0), max_passes=c.get("max_passes", 0))
    print(f"case {name}: {shape}", flush=True)
    g = fpl.UniformPressureProjection(shape, h)
    try:
        if c.get("density"):
            dens = (0.5 - phi / h).astype(np.float32)
            got = g.extrapolate(u, v, w, density=dens, **kw)
            want = oracle.extrapolate(u, v, w, h, density=dens, **kw)
        else:
            got = g.extrapolate(u, v, w, phi=phi, **kw)
            want = oracle.extrapolate(u, v, w, h, phi=phi, **kw)
    finally:
        g.destroy()
    return got, want


def assert_same(got, want):
    for a, b, name in zip(got[:3], want[:3], "uvw"):
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, f"{name}: {bad.size} faces differ, first at {np.unravel_index(bad[0], a.shape)}"
    r, o = got[3], want[3]
    assert (r["passes"], r["active_left"], r["levels"], r["band_faces"], r["unknown_faces"]) == \
        (o["passes"], o["active_left"], o["levels"], o["band_known"], o["unknown"]), (r, o)


@pytest.mark.parametrize("name", list(CASES))
def test_extrapolation_bit_exact(fpl, oracle, name):
    got, want = run_case(fpl, oracle, name)
    assert_same(got, want)
    if name == "ceiling":
        assert got[3]["passes"] == 2 and got[3]["active_left"] > 0
    if name == "no_liquid":
        assert got[3]["band_faces"] == 0 and all((a == 0).all() for a in got[:3])


@pytest.mark.parametrize("n", [64, 128])
def test_dam_spheres_bit_exact(fpl, oracle, n):
    s = scenes.dam_spheres_scene(n)
    print(f"dam + spheres {n}^3", flush=True)
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        got = g.extrapolate(s.u, s.v, s.w, phi=s.phi)
    finally:
        g.destroy()
    want = oracle.extrapolate(s.u, s.v, s.w, s.h, phi=s.phi)
    assert_same(got, want)
    assert got[3]["active_left"] == 0 and got[3]["unknown_faces"] == 0


def test_device_in_place_matches_host(fpl):
    import torch
    s = scenes.dam_spheres_scene(48)
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        want = g.extrapolate(s.u, s.v, s.w, phi=s.phi)
        t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (s.phi, s.u, s.v, s.w)]
        got = g.extrapolate(t[1], t[2], t[3], phi=t[0], out=t[1:])
        for a, b in zip(t[1:], want[:3]):
            np.testing.assert_array_equal(a.cpu().numpy(), b)
        assert got[3]["band_faces"] == want[3]["band_faces"]
    finally:
        g.destroy()


def test_refused_forms(fpl):
    lib = fpl.load_library()
    z = np.zeros(64, np.float32)
    g = fpl.UniformPressureProjection((8, 8, 8), 0.1)
    try:
        args = fpl.ExtrapolateArgs(0, 0, 17, 0, z.ctypes.data, None, *([z.ctypes.data] * 6))
        assert lib.fp_extrapolate_velocity(g.handle, C.byref(args), None) == fpl.ERR_INVALID_ARGUMENT
        args = fpl.ExtrapolateArgs(0, 0, 0, 0, None, None, *([z.ctypes.data] * 6))
        assert lib.fp_extrapolate_velocity(g.handle, C.byref(args), None) == fpl.ERR_INVALID_ARGUMENT
    finally:
        g.destroy()
    bases = np.full((6, 5), 2, np.int32)
    t = fpl.TallCellPressureProjection(bases, 14, 12, 0.1)
    try:
        args = fpl.ExtrapolateArgs(0, 0, 0, 0, *([z.ctypes.data] * 8))
        assert lib.fp_extrapolate_velocity(t.handle, C.byref(args), None) == fpl.ERR_INVALID_ARGUMENT
    finally:
        t.destroy()


def test_full_size_dam_spheres(fpl, oracle):
    """The 384^3 dam + 8 spheres grid: bit-exact against the oracle."""
    s = scenes.dam_spheres_scene(384)
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        got = g.extrapolate(s.u, s.v, s.w, phi=s.phi)
        print(f"GPU done: {got[3]}", flush=True)
    finally:
        g.destroy()
    want = oracle.extrapolate(s.u, s.v, s.w, s.h, phi=s.phi)
    assert_same(got, want)
