"""Finest-level sweeps fused in pairs (fp_tuning.sweep_pairs: one launch
reads each plane once for two red-black sweeps) against one launch per
sweep, on the GPU through the C ABI.  Bar: the same bits -- the V-cycle output, the pressure, the
velocities and the iteration counts -- for every sweep count, tile width,
odd and even rows, FP16 and BF16, batches and persistent launches."""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


def box_case(dims, seed):
    nx, ny, nz = dims
    h = 1.0 / nx
    c = lambda m: (np.arange(m) + 0.5) * h
    phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                         (0, 0, 0), (0.7, 0.55, 0.8)).astype(np.float32)
    rng = np.random.default_rng(seed)
    mu, mv, mw = scenes.fluid_face_masks(phi)
    u = np.where(mu, rng.uniform(-0.3, 0.3, mu.shape), 0).astype(np.float32)
    v = np.where(mv, rng.uniform(-0.3, 0.3, mv.shape) - 0.08, 0).astype(np.float32)
    w = np.where(mw, rng.uniform(-0.3, 0.3, mw.shape), 0).astype(np.float32)
    return phi, u, v, w, h


@pytest.mark.parametrize("dims,tile", [((64, 64, 64), 32), ((48, 40, 24), 16), ((37, 29, 33), 16),
                                       ((40, 36, 35), 32)])
@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_vcycle_bits(fpl, dims, tile, precision):
    phi, _, _, _, h = box_case(dims, 1)
    r = np.where(phi < 0, np.random.default_rng(4).standard_normal(dims), 0).astype(np.float32)
    g = fpl.UniformPressureProjection(dims, h)
    try:
        for sweeps in (1, 2, 3, 4):
            out = []
            for pairs in (1, 0):
                g.set_tuning(tile_z=tile, pairs=pairs)
                out.append(g.probe_vcycle(phi, r, precision=precision, pre_sweeps=sweeps, post_sweeps=sweeps))
            assert np.array_equal(out[0], out[1]), (sweeps, np.abs(out[0] - out[1]).max())
            assert np.isfinite(out[0]).all() and np.abs(out[0]).max() > 0
    finally:
        g.destroy()


@pytest.mark.parametrize("persistent", [0, 1])
def test_projection_bits(fpl, persistent):
    ss = [scenes.dam_spheres_scene(96, seed=1), scenes.dam_break_scene(40, 2), scenes.dam_break_scene(64, 3)]
    grids = [fpl.UniformPressureProjection(s.shape, s.h) for s in ss]
    try:
        reqs = [dict(phi=s.phi, u=s.u, v=s.v, w=s.w, dt=s.dt, rho=s.rho) for s in ss]
        for prec in ("fp16", "bf16"):
            res = []
            for pairs in (1, 0):
                for g in grids:
                    g.set_tuning(pairs=pairs)
                res.append(fpl.project_batch(grids, reqs, precision=prec, persistent=persistent))
            for a, b in zip(*res):
                assert a.report["flags"] == 0 and a.report["iterations"] == b.report["iterations"]
                assert a.report["relative_residual"] <= 1e-6
                for k in ("pressure", "u", "v", "w"):
                    assert np.array_equal(getattr(a, k), getattr(b, k)), (prec, k)
            grids[0].set_tuning(pairs=1)
            one = grids[0].project(ss[0].phi, ss[0].u, ss[0].v, ss[0].w, ss[0].dt, ss[0].rho, precision=prec,
                                   persistent=persistent)
            assert np.array_equal(one.pressure, res[1][0].pressure)
    finally:
        for g in grids:
            g.destroy()


def test_pairs_argument_checked(fpl):
    g = fpl.UniformPressureProjection((32, 32, 32), 1 / 32)
    try:
        with pytest.raises(fpl.FluidProjectionError):
            g.set_tuning(pairs=2)
    finally:
        g.destroy()
