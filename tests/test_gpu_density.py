"""The surface-density input (fp_inputs.density): the reference's
pressurePhi chain phi = -(density - 1/2) h (webgpu-uniform-reference.wgsl.ts:
215-242, open fraction 1) on the device, for both uniform solvers.  Bar: the
same bits as the projection of that level set given directly (the
conversion is the same FP32 expression on host and device)."""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["mgpcg", "cm11a"])
def test_density_equals_its_level_set(fpl, solver):
    s = scenes.dam_break_scene(32, 4)
    rng = np.random.default_rng(2)
    # a smeared surface: fractions in [0, 1] around the dam's interface
    density = np.clip(0.5 - s.phi / s.h + rng.uniform(-0.2, 0.2, s.phi.shape), 0, 1).astype(np.float32)
    phi = (-(density - np.float32(0.5)) * np.float32(s.h)).astype(np.float32)
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        a = g.project(phi, s.u, s.v, s.w, s.dt, s.rho, solver=solver)
        b = g.project(None, s.u, s.v, s.w, s.dt, s.rho, density=density, solver=solver)
        assert a.report["flags"] == 0 and a.report["liquid_samples"] == int((density > 0.5).sum())
        for k in ("pressure", "u", "v", "w"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
    finally:
        g.destroy()


def test_density_refused_on_tall_cells(fpl):
    nx = nz = 16
    ny, band = 32, 8
    bases = np.full((nx, nz), ny - band, np.int32)
    g = fpl.TallCellPressureProjection(bases, ny, band, 1 / 16)
    try:
        shapes = g.shapes()
        arrs = [np.zeros(sh, np.float32) for sh in shapes]
        with pytest.raises(fpl.FluidProjectionError) as e:
            g.project(None, *arrs[1:], 1 / 120, 998.2, density=np.zeros(shapes[0], np.float32))
        assert e.value.code == fpl.ERR_UNSUPPORTED
    finally:
        g.destroy()
