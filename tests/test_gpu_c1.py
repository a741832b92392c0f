"""The large-grid workload (uniform MAC grid, dam [0,0.5]x[0,0.7]x[0,1] plus 8
seeded spheres, Gaussian face velocities, dt = 1/240) against the FP64
oracle: at 128^3 in every precision, and at its full 384^3 size, where the
hierarchy is semi-coarsened to 3 cells (384 is not a power of two), the
cache-resident levels march in 2-plane chunks and the V-cycle bottom runs in
one workgroup with its levels in LDS.

Bars: flags 0, FP64 relative residual <= 1e-6, pressure and face velocities
within 1e-5 relative L2 of the oracle's FP64 MGPCG solve, max |div u| of the
FP32 output no larger than the oracle's at the same tolerance (128^3), and a
multigrid-like iteration count (<= 14) at 384^3.
"""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def c1_128(oracle):
    s = scenes.dam_spheres_scene(128, seed=0)
    ref = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
    _, vel6, _, _ = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-6)
    div6 = oracle.max_divergence(s.phi, *(a.astype(np.float32) for a in vel6), s.h)
    return s, ref, div6


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp32"])
def test_c1_128_matches_oracle(fpl, oracle, c1_128, precision):
    s, (op, (ou, ov, ow), ob, orep), div6 = c1_128
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, relative_tolerance=1e-6)
        rep = res.report
        print(precision, rep)
        assert rep["flags"] == 0, rep["flag_names"]
        assert rep["relative_residual"] <= 1e-6
        assert rel_l2(res.pressure, op) < 1e-5
        for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
            assert rel_l2(a, b) < 1e-5
        gdiv = oracle.max_divergence(s.phi, res.u, res.v, res.w, s.h)
        print("max div gpu", gdiv, "oracle at 1e-6", div6)
        assert gdiv <= div6
    finally:
        g.destroy()


def test_c1_384_matches_oracle(fpl, oracle):
    s = scenes.dam_spheres_scene(384, seed=0)
    dims = oracle.plan_levels(*s.shape)[0]
    assert dims[-1].tolist() == [3, 3, 3] and len(dims) == 8
    op, (ou, ov, ow), ob, orep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-10)
    assert orep["relative_residual"] <= 1e-10
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        for precision in ("fp16", "fp32"):
            res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, relative_tolerance=1e-6)
            rep = res.report
            print(precision, rep)
            assert rep["flags"] == 0, rep["flag_names"]
            assert rep["relative_residual"] <= 1e-6
            assert rep["iterations"] <= 14
            assert rel_l2(res.pressure, op) < 1e-5
            for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
                assert rel_l2(a, b) < 1e-5
    finally:
        g.destroy()
