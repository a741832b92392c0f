"""Velocity advection and body forces on the GPU (fp_advect_velocity)
against the CPU restatement (orc_advect_velocity).

Bar: bit-exact u, v, w -- both run the same FP32 operations in the same
order without contraction (the MacCormack prediction is extrapolated by the
library's and the oracle's bit-identical extrapolations).  Cases: both
schemes; gravity only; viscosity with no-slip walls; an open top; density
input; large steps (several RK2 pieces); a caller-given transport field; the
dam + spheres generator at 64^3 and 128^3; device tensors in place; refused
forms.
"""
import ctypes as C

import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


def run(fpl, oracle, s, **kw):
    g = fpl.UniformPressureProjection(s.phi.shape, s.h)
    try:
        got = g.advect(s.u, s.v, s.w, kw.pop("dt"), phi=s.phi, **kw)
    finally:
        g.destroy()
    return got


def assert_same(got, want):
    for a, b, name in zip(got[:3], want, "uvw"):
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, f"{name}: {bad.size} faces differ, first at {np.unravel_index(bad[0], a.shape)}"


CASES = {
    "sl": dict(dt=1 / 30),
    "maccormack": dict(dt=1 / 30, scheme="maccormack"),
    "gravity_only_zero_velocity": dict(dt=1 / 60, zero=True),
    "viscous_no_slip": dict(dt=1 / 60, viscosity=1e-3, no_slip=True),
    "viscous_maccormack": dict(dt=1 / 60, viscosity=1e-3, scheme="maccormack"),
    "open_top": dict(dt=1 / 30, open_top=True),
    "open_top_maccormack": dict(dt=1 / 30, open_top=True, scheme="maccormack"),
    "large_step": dict(dt=0.2, scheme="maccormack"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_advection_bit_exact(fpl, oracle, name):
    kw = dict(CASES[name])
    s = scenes.dam_spheres_scene(32, seed=3)
    if kw.pop("zero", False):
        s.u[:] = 0
        s.v[:] = 0
        s.w[:] = 0
    print(f"case {name}", flush=True)
    got = run(fpl, oracle, s, **kw)
    scheme = kw.get("scheme", "semi-lagrangian")
    want = oracle.advect_velocity(s.u, s.v, s.w, s.h, kw["dt"], phi=s.phi, gravity=-9.80665,
                                  nu=kw.get("viscosity", 0.0), scheme=scheme, open_top=int(kw.get("open_top", 0)),
                                  no_slip=int(kw.get("no_slip", 0)))
    assert_same(got, want)


def test_density_input_and_given_transport(fpl, oracle):
    s = scenes.dam_break_scene(24, 2)
    dens = (0.5 - s.phi / s.h).astype(np.float32)
    t = oracle.extrapolate(s.u, s.v, s.w, s.h, density=dens)[:3]
    g = fpl.UniformPressureProjection(s.phi.shape, s.h)
    try:
        got = g.advect(s.u, s.v, s.w, 1 / 30, density=dens, transport=t, scheme="maccormack")
    finally:
        g.destroy()
    want = oracle.advect_velocity(s.u, s.v, s.w, s.h, 1 / 30, density=dens, transport=t, gravity=-9.80665,
                                  scheme="maccormack")
    assert_same(got, want)


@pytest.mark.parametrize("n,scheme", [(64, "semi-lagrangian"), (64, "maccormack"), (128, "maccormack")])
def test_dam_spheres_bit_exact(fpl, oracle, n, scheme):
    s = scenes.dam_spheres_scene(n)
    print(f"dam + spheres {n}^3 {scheme}", flush=True)
    got = run(fpl, oracle, s, dt=1 / 60, scheme=scheme, viscosity=1e-6)
    want = oracle.advect_velocity(s.u, s.v, s.w, s.h, 1 / 60, phi=s.phi, gravity=-9.80665, nu=1e-6, scheme=scheme)
    assert_same(got, want)


@pytest.mark.parametrize("shape", [(9, 7, 2), (6, 5, 2), (5, 6, 3)])
@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_thin_and_odd_grids_bit_exact(fpl, oracle, shape, scheme):
    """Grids two cells deep in z (the z-paired corner loads clamp at both ends
    of every row) and odd extents, with fast random velocities that trace into
    every wall, closed and open top."""
    rng = np.random.default_rng(7)
    nx, ny, nz = shape
    h = 1.0 / max(shape)
    y = (np.arange(ny) + 0.5) * h
    phi = np.broadcast_to((y - 0.5 * ny * h)[None, :, None], shape).astype(np.float32).copy()
    u = rng.uniform(-3, 3, (nx + 1, ny, nz)).astype(np.float32)
    v = rng.uniform(-3, 3, (nx, ny + 1, nz)).astype(np.float32)
    w = rng.uniform(-3, 3, (nx, ny, nz + 1)).astype(np.float32)
    for open_top in (False, True):
        print(f"shape {shape} {scheme} open_top {open_top}", flush=True)
        g = fpl.UniformPressureProjection(shape, h)
        try:
            got = g.advect(u, v, w, 0.2, phi=phi, scheme=scheme, open_top=open_top)
        finally:
            g.destroy()
        want = oracle.advect_velocity(u, v, w, h, 0.2, phi=phi, gravity=-9.80665, scheme=scheme,
                                      open_top=int(open_top))
        assert_same(got, want)


def test_device_in_place_matches_host(fpl):
    import torch
    s = scenes.dam_spheres_scene(40)
    g = fpl.UniformPressureProjection(s.phi.shape, s.h)
    try:
        want = g.advect(s.u, s.v, s.w, 1 / 30, phi=s.phi, scheme="maccormack")
        t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (s.phi, s.u, s.v, s.w)]
        g.advect(t[1], t[2], t[3], 1 / 30, phi=t[0], scheme="maccormack", out=t[1:])
        for a, b in zip(t[1:], want[:3]):
            np.testing.assert_array_equal(a.cpu().numpy(), b)
    finally:
        g.destroy()


def test_refused_forms(fpl):
    lib = fpl.load_library()
    z = np.zeros(4096, np.float32)
    p = z.ctypes.data
    g = fpl.UniformPressureProjection((8, 8, 8), 0.1)
    try:
        bad_scheme = fpl.AdvectArgs(0, 7, 0, 0, 0.01, -9.8, 0.0, p, None, p, p, p, None, None, None, p, p, p)
        assert lib.fp_advect_velocity(g.handle, C.byref(bad_scheme), None) == fpl.ERR_INVALID_ARGUMENT
        bad_dt = fpl.AdvectArgs(0, 0, 0, 0, 0.0, -9.8, 0.0, p, None, p, p, p, None, None, None, p, p, p)
        assert lib.fp_advect_velocity(g.handle, C.byref(bad_dt), None) == fpl.ERR_INVALID_ARGUMENT
        partial = fpl.AdvectArgs(0, 0, 0, 0, 0.01, -9.8, 0.0, p, None, p, p, p, p, None, None, p, p, p)
        assert lib.fp_advect_velocity(g.handle, C.byref(partial), None) == fpl.ERR_INVALID_ARGUMENT
    finally:
        g.destroy()


@pytest.mark.parametrize("cells,scheme", [(0.3, "semi-lagrangian"), (4.0, "semi-lagrangian"), (2.5, "maccormack")])
def test_full_size_translation_is_exact(fpl, cells, scheme):
    """At the 384^3 size of the large workload, where the oracle is too slow:
    a constant transport velocity U along x carries a field linear in x
    exactly (exact RK2 departure points of a constant field, in pieces of at
    most 1.5 cells; exact trilinear interpolation of a linear field; the
    MacCormack correction vanishes), so interior faces hold a + b (x - U dt)
    and u stays U.  All liquid, no gravity."""
    import torch
    n = 384
    h = 1.0 / n
    U = 0.8
    dt = cells * h / U
    dev = torch.device("cuda:0")
    phi = torch.full((n, n, n), -1.0, device=dev)
    u = torch.full((n + 1, n, n), U, device=dev)
    u[0] = 0.0
    u[-1] = 0.0
    xs = (torch.arange(n, device=dev, dtype=torch.float64) + 0.5) * h
    v = (0.02 + 0.05 * xs).float()[:, None, None].expand(n, n + 1, n).contiguous()
    w = torch.zeros((n, n, n + 1), device=dev)
    print(f"384^3 translation {cells} cells {scheme}", flush=True)
    g = fpl.UniformPressureProjection((n, n, n), h)
    try:
        uo, vo, wo, _ = g.advect(u, v, w, dt, phi=phi, transport=(u, v, w), gravity=0.0, scheme=scheme)
        torch.cuda.synchronize()
    finally:
        g.destroy()
    lo = int(np.ceil(cells)) + 3
    hi = n - 3 - (int(np.ceil(cells)) if scheme == "maccormack" else 0)
    want = (0.02 + 0.05 * (xs[lo:hi] - U * dt)).float()
    got = vo[lo:hi, 3:-3, 3:-3]
    assert float((got - want[:, None, None]).abs().max()) <= 2e-6
    assert float((uo[lo:hi, 3:-3, 3:-3] - U).abs().max()) <= 2e-6
    assert float(wo[:, :, 1:-1].abs().max()) == 0.0


def test_budget_refusal_leaves_the_grid_usable(fpl, oracle):
    """A context whose HBM budget holds the grid but not the advection state:
    fp_advect_velocity fails with FP_ERR_BUDGET and allocates nothing; the grid
    still projects, and a context with room advects bit-exactly."""
    s = scenes.dam_spheres_scene(32, seed=1)
    # what the grid and one host-memory projection hold (state and staging)
    probe = fpl.Context(0, 0)
    g = fpl.UniformPressureProjection(s.phi.shape, s.h, context=probe)
    g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho)
    grid_bytes = probe.bytes_in_use()
    g.destroy()
    probe.destroy()
    ctx = fpl.Context(0, grid_bytes + 4096)
    g = fpl.UniformPressureProjection(s.phi.shape, s.h, context=ctx)
    try:
        assert g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho).report["flags"] == 0
        with pytest.raises(fpl.FluidProjectionError) as e:
            g.advect(s.u, s.v, s.w, 1 / 30, phi=s.phi)
        assert e.value.code == fpl.ERR_BUDGET
        assert ctx.bytes_in_use() == grid_bytes
        assert g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho).report["flags"] == 0
    finally:
        g.destroy()
        ctx.destroy()
    got = run(fpl, oracle, s, dt=1 / 30)
    assert_same(got, oracle.advect_velocity(s.u, s.v, s.w, s.h, 1 / 30, phi=s.phi, gravity=-9.80665))
