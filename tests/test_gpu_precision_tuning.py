"""Automatic FP16 -> BF16 fallback, launch-shape tuning and the matrix-core
coarsest solve, on the GPU through the C ABI.

Bars: an FP16 multigrid that overflows (forced from the host by shifting
the V-cycle input normalisation by 2^24, past FP16's range) is re-solved in
BF16 and reported with FP_FLAG_PRECISION_FALLBACK only, reaching the
tolerance and matching the FP64 oracle within 1e-5 (fp_project and
fp_submit); every launch shape a grid can be tuned to gives the same
solution (pressure within 1e-5 of the default shape's); the V-cycle, whose
coarsest solve runs on the matrix cores, stays within its parity bar."""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu

FALLBACK = 0x20


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def scene_and_oracle(oracle):
    s = scenes.dam_spheres_scene(64, seed=2)
    ref = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
    return s, ref


def test_fp16_overflow_falls_back_to_bf16(fpl, scene_and_oracle):
    s, (op, (ou, ov, ow), ob, orep) = scene_and_oracle
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        normal = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16")
        assert normal.report["flags"] == 0
        for how in ("project", "submit"):
            kw = dict(precision="fp16", relative_tolerance=1e-6, precision_test_shift=24)
            if how == "project":
                res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, **kw)
            else:
                res = g.submit(s.phi, s.u, s.v, s.w, s.dt, s.rho, **kw).wait(60000)
            rep = res.report
            print(how, rep)
            assert rep["flags"] == FALLBACK, (how, rep["flag_names"])
            assert rep["relative_residual"] <= 1e-6
            assert rel_l2(res.pressure, op) < 1e-5
            assert rel_l2(res.v, ov) < 1e-5
        # FP32 and BF16 never fall back: the shift is inside their range
        for prec in ("fp32", "bf16"):
            r = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=prec, precision_test_shift=24)
            assert r.report["flags"] == 0, (prec, r.report["flag_names"])
    finally:
        g.destroy()


@pytest.mark.parametrize("dims", [(96, 96, 96), (48, 40, 24)])
def test_every_tuned_shape_gives_the_same_solution(fpl, oracle, dims):
    from fluid_amd import autotune
    nx, ny, nz = dims
    if nx == ny == nz:
        s = scenes.dam_spheres_scene(nx, seed=3)
        phi, u, v, w, h, dt, rho = s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho
    else:
        h = 1.0 / nx
        c = lambda m: (np.arange(m) + 0.5) * h
        phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                             (0, 0, 0), (0.6, 0.5, 0.55)).astype(np.float32)
        rng = np.random.default_rng(5)
        mu, mv, mw = scenes.fluid_face_masks(phi)
        u = np.where(mu, rng.uniform(-0.3, 0.3, mu.shape), 0).astype(np.float32)
        v = np.where(mv, rng.uniform(-0.3, 0.3, mv.shape) - 0.08, 0).astype(np.float32)
        w = np.where(mw, rng.uniform(-0.3, 0.3, mw.shape), 0).astype(np.float32)
        dt, rho = 1 / 120, 998.2
    op, _, _, _ = oracle.uniform_projection(phi, u, v, w, h, dt, rho, rel_tol=1e-12)
    g = fpl.UniformPressureProjection(dims, h)
    try:
        base = g.project(phi, u, v, w, dt, rho, precision="fp16")
        for cfg in autotune.candidates(dims):
            g.set_tuning(tile_z=cfg["tile_z"], groups=cfg["groups"], pairs=cfg["pairs"])
            r = g.project(phi, u, v, w, dt, rho, precision="fp16", pre_sweeps=cfg["sweeps"],
                          post_sweeps=cfg["sweeps"])
            assert r.report["flags"] == 0 and r.report["relative_residual"] <= 1e-6, cfg
            assert rel_l2(r.pressure, op) < 1e-5, cfg
            if cfg["sweeps"] == 2:
                assert rel_l2(r.pressure, base.pressure) < 1e-5, cfg
            if cfg["sweeps"] == 2 and cfg["tile_z"] == (32 if nz >= 32 else 16) and cfg["groups"] == 8192:
                assert np.array_equal(r.pressure, base.pressure), cfg  # pairs or not: the same bits
        g.set_tuning(tile_z=0, groups=0)
        again = g.project(phi, u, v, w, dt, rho, precision="fp16")
        assert np.array_equal(again.pressure, base.pressure)
    finally:
        g.destroy()


def test_tuning_rejects_bad_shapes(fpl):
    g = fpl.UniformPressureProjection((32, 32, 16), 1 / 32)
    try:
        for kw in (dict(tile_z=64), dict(tile_z=32), dict(groups=100), dict(groups=1 << 20)):
            with pytest.raises(fpl.FluidProjectionError):
                g.set_tuning(**kw)
    finally:
        g.destroy()


@pytest.mark.parametrize("dims", [(6, 5, 4), (12, 10, 8), (37, 37, 37)])
def test_vcycle_with_matrix_core_coarsest_solve(fpl, oracle, dims):
    """Small lattices whose V-cycle is dominated by the dense coarsest solve
    (two levels at 6 x 5 x 4): the FP32 V-cycle, its coarsest product on the
    matrix cores, matches the oracle's FP64 V-cycle within 1e-5."""
    nx, ny, nz = dims
    h = 1.0 / nx
    c = lambda m: (np.arange(m) + 0.5) * h
    phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                         (0, 0, 0), (0.7, 0.6, 0.8)).astype(np.float32)
    r = np.where(phi < 0, np.random.default_rng(9).standard_normal(dims), 0).astype(np.float32)
    g = fpl.UniformPressureProjection(dims, h)
    try:
        z = g.probe_vcycle(phi, r, precision="fp32")
        oz = oracle.vcycle(phi, r.astype(np.float64))
        assert rel_l2(z, oz) < 1e-5, rel_l2(z, oz)
    finally:
        g.destroy()


def test_persistent_sweeps_give_the_same_bits(fpl):
    """Sweeps as persistent launches (a fixed grid walking each stage's table)
    compute exactly what one workgroup per tile computes, for one grid and
    for a batch of uneven grids."""
    ss = [scenes.dam_spheres_scene(96, seed=1), scenes.dam_break_scene(40, 2), scenes.dam_break_scene(64, 3)]
    grids = [fpl.UniformPressureProjection(s.shape, s.h) for s in ss]
    try:
        reqs = [dict(phi=s.phi, u=s.u, v=s.v, w=s.w, dt=s.dt, rho=s.rho) for s in ss]
        for prec in ("fp16", "fp32"):
            base = fpl.project_batch(grids, reqs, precision=prec)
            pers = fpl.project_batch(grids, reqs, precision=prec, persistent=1)
            one = grids[0].project(ss[0].phi, ss[0].u, ss[0].v, ss[0].w, ss[0].dt, ss[0].rho, precision=prec,
                                   persistent=1)
            for a, b in zip(base, pers):
                assert a.report["flags"] == 0 and b.report["iterations"] == a.report["iterations"]
                assert np.array_equal(a.pressure, b.pressure) and np.array_equal(a.u, b.u)
            assert np.array_equal(one.pressure, base[0].pressure)
    finally:
        for g in grids:
            g.destroy()
