"""Unanchored pressure (a grid liquid everywhere) and host-forced failure
paths on the GPU, through the C ABI, against the FP64 oracle.

A liquid component with no air face has a pure-Neumann pressure: A is
singular with the constants as its null space.  The reference removes the
component's mean from b and from the iterate
(lib/methods/adaptive-mass/sparse-atlas-composite-projection.ts:1086-1146,
1194-1211); with closed walls as the only solid boundary, the only such
component is a domain liquid everywhere (any other component has a neighbour
outside it, which is air).  Bars: FP_FLAG_UNANCHORED_COMPONENT set, the FP64
relative residual reached, pressure (zero mean) and velocities within 1e-5
relative L2 of the oracle's nullspace-projected solve.  A liquid droplet or
an air bubble is anchored (no flag).  Non-finite inputs end the solve with
FP_FLAG_NONFINITE instead of an error or a hang.
"""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu

UNANCHORED = 0x10
NONFINITE = 0x1
NONCONVERGENCE = 0x8


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def random_faces(n, seed, shape=None):
    nx, ny, nz = shape or (n, n, n)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, (nx + 1, ny, nz)).astype(np.float32)
    v = rng.uniform(-1, 1, (nx, ny + 1, nz)).astype(np.float32)
    w = rng.uniform(-1, 1, (nx, ny, nz + 1)).astype(np.float32)
    return u, v, w


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("dims", [(24, 24, 24), (20, 12, 16)])
def test_full_box_is_unanchored(fpl, oracle, dims, precision):
    nx, ny, nz = dims
    h = 1.0 / nx
    phi = np.full(dims, -h, np.float32)
    u, v, w = random_faces(nx, 11, dims)
    op, (ou, ov, ow), ob, orep = oracle.uniform_projection(phi, u, v, w, h, 0.01, 1000.0, rel_tol=1e-11,
                                                           solver="jacobi")
    assert orep["unanchored"] == 1
    g = fpl.UniformPressureProjection(dims, h)
    try:
        res = g.project(phi, u, v, w, 0.01, 1000.0, precision=precision, relative_tolerance=1e-6)
        rep = res.report
        print(precision, dims, rep)
        assert rep["flags"] == UNANCHORED, rep["flag_names"]
        assert rep["relative_residual"] <= 1e-6
        assert abs(float(res.pressure.astype(np.float64).mean())) < 1e-6 * np.abs(op).max()
        assert rel_l2(res.pressure, op) < 1e-5
        for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
            assert rel_l2(a, b) < 1e-5
        odiv = oracle.max_divergence(phi, *(x.astype(np.float32) for x in oracle.uniform_projection(
            phi, u, v, w, h, 0.01, 1000.0, rel_tol=1e-6, solver="jacobi")[1]), h)
        assert rep["max_div_after"] <= odiv * 1.0001 + 1e-6
    finally:
        g.destroy()


def test_full_box_inconsistent_target_divergence(fpl, oracle):
    """A net target divergence cannot be met in a closed full box: its mean
    is removed with b's (the oracle projects b onto the range the same way)."""
    n = 16
    h = 1.0 / n
    phi = np.full((n, n, n), -h, np.float32)
    u, v, w = random_faces(n, 12)
    q = np.random.default_rng(13).uniform(0.5, 1.5, (n, n, n)).astype(np.float32)
    op, (ou, ov, ow), ob, orep = oracle.uniform_projection(phi, u, v, w, h, 0.01, 1000.0, rel_tol=1e-11,
                                                           solver="jacobi", q=q)
    g = fpl.UniformPressureProjection((n, n, n), h)
    try:
        res = g.project(phi, u, v, w, 0.01, 1000.0, target_divergence=q, precision="fp32",
                        relative_tolerance=1e-7)
        assert res.report["flags"] == UNANCHORED, res.report["flag_names"]
        assert res.report["relative_residual"] <= 1e-7
        assert rel_l2(res.pressure, op) < 1e-5
        for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
            assert rel_l2(a, b) < 1e-5
    finally:
        g.destroy()


def test_bubble_and_droplet_are_anchored(fpl, oracle):
    """A full box with one small air bubble, and a pool with a droplet
    floating above it: every component touches air, so no flag, and the
    coarse levels (where the bubble averages away) keep an SPD inverse."""
    n = 32
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    bubble = np.full((n, n, n), -h, np.float32)
    bubble[15:17, 20:22, 9:11] = 0.5 * h
    pool = np.minimum(Y - 0.3, np.sqrt((X - 0.5) ** 2 + (Y - 0.7) ** 2 + (Z - 0.45) ** 2) - 0.12)
    for name, phi in (("bubble", bubble), ("droplet", pool.astype(np.float32))):
        u, v, w = random_faces(n, 21)
        op, (ou, ov, ow), ob, orep = oracle.uniform_projection(phi, u, v, w, h, 0.01, 1000.0, rel_tol=1e-12)
        assert orep["unanchored"] == 0
        g = fpl.UniformPressureProjection((n, n, n), h)
        try:
            for precision in ("fp32", "fp16"):
                res = g.project(phi, u, v, w, 0.01, 1000.0, precision=precision, relative_tolerance=1e-6)
                print(name, precision, res.report)
                assert res.report["flags"] == 0, (name, precision, res.report["flag_names"])
                assert res.report["relative_residual"] <= 1e-6
                assert rel_l2(res.pressure, op) < 1e-5, name
                assert rel_l2(res.v, ov) < 1e-5, name
        finally:
            g.destroy()


def test_tall_full_column_grid_is_unanchored(fpl, oracle):
    """Tall cells whose bands reach the lid, liquid everywhere: no air row
    anywhere, so the pressure is unanchored (compared up to its mean)."""
    sc = scenes.tall_random_scene(6, 5, 14, 12, seed=7)
    L = sc.layout
    assert int((L.bases + L.band).min()) >= L.fine_ny
    phi = np.where(L.active(), -L.h, 5 * L.h).astype(np.float32)
    t = oracle.Tall(L.bases, phi, L.fine_ny, L.band, L.h, L.max_delta)
    unk = t.unknowns().astype(bool)
    op, (ou, ov, ow), ob, orep = t.projection(sc.u, sc.v, sc.w, sc.dt, sc.rho, rel_tol=1e-11)
    g = fpl.TallCellPressureProjection.from_layout(L)
    try:
        res = g.project(phi, sc.u, sc.v, sc.w, sc.dt, sc.rho, relative_tolerance=1e-7)
        rep = res.report
        print(rep)
        assert rep["flags"] == UNANCHORED, rep["flag_names"]
        assert rep["relative_residual"] <= 1e-7
        p = res.pressure.astype(np.float64)
        assert abs(p[unk].mean()) < 1e-6 * np.abs(p[unk]).max()
        ref = np.where(unk, op - op[unk].mean(), 0.0)
        assert rel_l2(p, ref) < 1e-5
        for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
            assert rel_l2(a, b) < 1e-5
    finally:
        g.destroy()


@pytest.mark.parametrize("what", ["u_nan", "v_inf", "q_nan"])
def test_nonfinite_input_is_flagged(fpl, what):
    """A NaN or infinite input velocity (or target divergence) on a liquid
    face makes b non-finite: the solve stops at its first FP64 residual with
    NONFINITE (and NONCONVERGENCE), returns FP_OK and the grid stays usable."""
    s = scenes.dam_break_scene(32, 0)
    u, v, w = s.u.copy(), s.v.copy(), s.w.copy()
    q = None
    liq = np.argwhere(s.phi < 0)[len(np.argwhere(s.phi < 0)) // 2]
    i, j, k = (int(x) for x in liq)
    if what == "u_nan":
        u[i + 1, j, k] = np.nan
    elif what == "v_inf":
        v[i, j + 1, k] = np.inf
    else:
        q = np.zeros(s.shape, np.float32)
        q[i, j, k] = np.nan
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        res = g.project(s.phi, u, v, w, s.dt, s.rho, target_divergence=q, relative_tolerance=1e-6)
        assert res.report["flags"] & NONFINITE, res.report
        assert res.report["flags"] & NONCONVERGENCE, res.report
        assert res.report["iterations"] == 0
        # the same grid then projects finite inputs normally
        ok = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, relative_tolerance=1e-6)
        assert ok.report["flags"] == 0 and ok.report["relative_residual"] <= 1e-6
    finally:
        g.destroy()


def test_tall_nonfinite_input_is_flagged(fpl):
    sc = scenes.tall_random_scene(16, 12, 40, 8, seed=5)
    v = sc.v.copy()
    unk = np.argwhere(sc.phi < 0)
    x, py, z = (int(a) for a in unk[len(unk) // 2])
    v[x, py, z] = np.nan
    g = fpl.TallCellPressureProjection.from_layout(sc.layout)
    try:
        res = g.project(sc.phi, sc.u, v, sc.w, sc.dt, sc.rho, relative_tolerance=1e-6)
        assert res.report["flags"] & NONFINITE, res.report
        ok = g.project(sc.phi, sc.u, sc.v, sc.w, sc.dt, sc.rho, relative_tolerance=1e-6)
        assert ok.report["flags"] == 0 and ok.report["relative_residual"] <= 1e-6
    finally:
        g.destroy()
