"""GPU parity of the uniform projection (HIP library through the C ABI)
against the FP64 CPU oracle on identical seeded inputs.

Bars (stated per test): integer maps bit-exact; FP64 rhs/operator to 1e-12
relative; pressure and projected face velocities within 1e-5 relative L2 of
an FP64 oracle solve; FP64 relative residual <= the requested tolerance; max
|div u| after projection no larger than the oracle's at the same tolerance
(both measured on FP32-rounded velocities)."""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp32", "fp16", "bf16"]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def c0():
    return scenes.dam_break_scene(64, 0)


@pytest.fixture(scope="module")
def c0_oracle(oracle, c0):
    s = c0
    p, vel, b, rep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
    return p, vel, b, rep


def oracle_maxdiv_fp32(oracle, s, tol):
    p, (uo, vo, wo), b, rep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=tol)
    return oracle.max_divergence(s.phi, uo.astype(np.float32), vo.astype(np.float32), wo.astype(np.float32), s.h)


@pytest.mark.parametrize("dims", [(64, 64, 64), (384, 384, 384), (37, 37, 37), (128, 128, 8), (48, 32, 40),
                                  (100, 100, 100), (2, 2, 2)])
def test_level_plan_bit_exact(fpl, oracle, dims):
    g = fpl.UniformPressureProjection(dims, 1.0 / dims[0])
    d, s = g.levels()
    od, os_, _ = oracle.plan_levels(*dims)
    assert d.tolist() == od.tolist()
    assert s.tolist() == os_.tolist()
    g.destroy()


@pytest.mark.parametrize("scene", ["dam64", "spheres128", "dam37", "dam48x32x40"])
def test_coarse_level_sets_and_labels(fpl, oracle, scene):
    """Coarse level sets (FP32, D4-ordered sums of the eight children) and the
    labels derived from them are bit-exact on every level, with no mask."""
    if scene == "dam48x32x40":
        nx, ny, nz = 48, 32, 40
        c = lambda m: (np.arange(m) + 0.5) / nx
        phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                             (0, 0, 0), (0.55, 0.5, 0.6)).astype(np.float32)
    else:
        phi = {"dam64": lambda: scenes.dam_break_scene(64, 0).phi,
               "spheres128": lambda: scenes.dam_spheres_scene(128, seed=0).phi,
               "dam37": lambda: scenes.dam_break_scene(37, 0).phi}[scene]()
    g = fpl.UniformPressureProjection(phi.shape, 1.0 / phi.shape[0])
    dims, _ = g.levels()
    for level in range(1, len(dims)):
        gp = g.probe_level_phi(phi, level)
        op = oracle.level_phi(phi, level)
        assert gp.shape == op.shape
        assert np.array_equal(gp.astype(np.float64), op), (level, np.abs(gp - op).max())
        assert np.array_equal(gp < 0, op < 0)
    g.destroy()


@pytest.mark.parametrize("dims", [(64, 64, 64), (48, 32, 40), (37, 37, 37), (384, 384, 384), (128, 128, 8)])
def test_transfer_taps_bit_exact(fpl, oracle, dims):
    """Prolongation taps (and the restriction's, their transpose) as the
    kernels evaluate them equal the oracle's taps1d on every level and axis."""
    g = fpl.UniformPressureProjection(dims, 1.0 / dims[0])
    ldims, strides = g.levels()
    for level in range(1, len(ldims)):
        for axis in range(3):
            nf, nc, st = int(ldims[level - 1][axis]), int(ldims[level][axis]), int(strides[level][axis])
            (pc, pw), (rf, rw) = g.probe_taps(level, axis)
            oi, ow = oracle.taps1d(nf, nc, st)
            assert np.array_equal(pc, oi), (level, axis)
            assert np.array_equal(pw.astype(np.float64), ow), (level, axis)
            # restriction rows: the transpose of the prolongation taps
            want = [dict() for _ in range(nc)]
            for f in range(nf):
                for q in range(2):
                    if oi[f, q] >= 0:
                        want[oi[f, q]][f] = ow[f, q]
            for c in range(nc):
                got = {int(rf[c, q]): float(rw[c, q]) for q in range(4) if rf[c, q] >= 0}
                assert got == want[c], (level, axis, c, got, want[c])
    g.destroy()


def test_rhs_matches_oracle(fpl, oracle, c0):
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    b = g.probe_rhs(s.phi, s.u, s.v, s.w, s.dt, s.rho)
    ob = oracle.rhs(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho)
    assert np.abs(b - ob).max() <= 1e-12 * np.abs(ob).max()
    q = np.random.default_rng(3).uniform(-1, 1, s.shape).astype(np.float32)
    bq = g.probe_rhs(s.phi, s.u, s.v, s.w, s.dt, s.rho, target_divergence=q)
    obq = oracle.rhs(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, q)
    assert np.abs(bq - obq).max() <= 1e-12 * np.abs(obq).max()
    g.destroy()


def test_operator_matches_oracle(fpl, oracle, c0):
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    p = np.random.default_rng(1).standard_normal(s.shape)
    ap = g.probe_apply(s.phi, p)
    oap = oracle.apply(s.phi, p)
    assert np.abs(ap - oap).max() <= 1e-12 * np.abs(oap).max()
    g.destroy()


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("bf16", 3e-2), ("fp16", 1e-2)])
def test_vcycle_matches_oracle(fpl, oracle, c0, precision, tol):
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    r = oracle.rhs(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho).astype(np.float32)
    z = g.probe_vcycle(s.phi, r, precision=precision)
    oz = oracle.vcycle(s.phi, r.astype(np.float64))
    assert rel_l2(z, oz) < tol, rel_l2(z, oz)
    g.destroy()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_projection_c0_parity(fpl, oracle, c0, c0_oracle, precision):
    s = c0
    op, (ou, ov, ow), ob, orep = c0_oracle
    g = fpl.UniformPressureProjection(s.shape, s.h)
    res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, relative_tolerance=1e-6)
    rep = res.report
    print(precision, rep)
    assert rep["flags"] == 0, rep["flag_names"]
    assert rep["relative_residual"] <= 1e-6
    assert rel_l2(res.pressure, op) < 1e-5
    for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
        assert rel_l2(a, b) < 1e-5
    # divergence: the GPU's FP32 output against the oracle's at the same tolerance
    gdiv = oracle.max_divergence(s.phi, res.u, res.v, res.w, s.h)
    assert gdiv == pytest.approx(rep["max_div_after"], rel=1e-3)
    odiv = oracle_maxdiv_fp32(oracle, s, 1e-6)
    print("max div gpu", gdiv, "oracle at the same tolerance", odiv)
    assert gdiv <= odiv
    assert rep["max_div_before"] == pytest.approx(orep["max_div_before"], rel=1e-5)
    assert rep["liquid_samples"] == int((s.phi < 0).sum())
    g.destroy()


def test_tight_tolerance_reaches_fp64_accuracy(fpl, c0, c0_oracle):
    s = c0
    op = c0_oracle[0]
    g = fpl.UniformPressureProjection(s.shape, s.h)
    res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16", relative_tolerance=1e-10)
    assert res.report["relative_residual"] <= 1e-10
    assert rel_l2(res.pressure, op) < 1e-7  # FP32 output rounding
    g.destroy()


def test_hydrostatic_column_gpu(fpl):
    n, H = 32, 0.53
    h = 1.0 / n
    y = (np.arange(n) + 0.5) * h
    phi = np.broadcast_to((y - H)[None, :, None], (n, n, n)).astype(np.float32).copy()
    dt, rho, gacc = 1 / 60, 998.2, 9.81
    mu, mv, mw = scenes.fluid_face_masks(phi)
    u = np.zeros((n + 1, n, n), np.float32)
    v = np.where(mv, np.float32(-gacc * dt), 0).astype(np.float32)
    w = np.zeros((n, n, n + 1), np.float32)
    g = fpl.UniformPressureProjection((n, n, n), h)
    res = g.project(phi, u, v, w, dt, rho, relative_tolerance=1e-9)
    exact = np.where(phi < 0, rho * gacc * (H - y)[None, :, None], 0.0)
    assert np.abs(res.pressure - exact).max() < 2e-6 * exact.max()
    assert np.abs(res.v).max() < 1e-5
    g.destroy()


def test_empty_domain(fpl):
    """All air: nothing to solve, zero pressure and velocities (the all-liquid
    domain is covered by test_gpu_nullspace_faults.py)."""
    n = 16
    g = fpl.UniformPressureProjection((n, n, n), 1 / n)
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, (n + 1, n, n)).astype(np.float32)
    v = rng.uniform(-1, 1, (n, n + 1, n)).astype(np.float32)
    w = rng.uniform(-1, 1, (n, n, n + 1)).astype(np.float32)
    air = np.ones((n, n, n), np.float32)
    res = g.project(air, u, v, w, 0.01, 1000.0)
    assert res.report["flags"] == 0
    assert np.abs(res.pressure).max() == 0 and np.abs(res.u).max() == 0
    g.destroy()


@pytest.mark.parametrize("dims", [(37, 37, 37), (48, 32, 40), (5, 7, 3)])
def test_uneven_grids(fpl, oracle, dims):
    n = dims[0]
    rng = np.random.default_rng(sum(dims))
    nx, ny, nz = dims
    h = 1.0 / n
    c = lambda m: (np.arange(m) + 0.5) * h
    phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                         (0, 0, 0), (0.55, 0.45 * ny * h * 1.3, 0.6)).astype(np.float32)
    mu, mv, mw = scenes.fluid_face_masks(phi)
    u = np.where(mu, rng.uniform(-0.3, 0.3, mu.shape), 0).astype(np.float32)
    v = np.where(mv, rng.uniform(-0.3, 0.3, mv.shape) - 0.08, 0).astype(np.float32)
    w = np.where(mw, rng.uniform(-0.3, 0.3, mw.shape), 0).astype(np.float32)
    op, (ou, ov, ow), ob, orep = oracle.uniform_projection(phi, u, v, w, h, 1 / 120, 998.2, rel_tol=1e-12)
    g = fpl.UniformPressureProjection(dims, h)
    res = g.project(phi, u, v, w, 1 / 120, 998.2, relative_tolerance=1e-7)
    assert res.report["flags"] == 0, res.report
    assert res.report["relative_residual"] <= 1e-7
    assert rel_l2(res.pressure, op) < 1e-5
    assert rel_l2(res.v, ov) < 1e-5
    g.destroy()


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_bottom_in_lds_and_in_l2(fpl, oracle, precision):
    """32 x 32 x 28: the one-workgroup bottom of the V-cycle starts at the
    16 x 16 x 14 level (4172 cells down to the coarsest).  Its FP16 arrays fit
    the LDS budget (41.7 KB), its FP32 ones (66.8 KB) do not, so the two
    precisions run the LDS-resident and the L2-resident bottom kernels; level 0
    (nz = 28) sweeps in 16-wide tiles clamped to the liquid hulls."""
    dims = (32, 32, 28)
    od, _, _ = oracle.plan_levels(*dims)
    assert [int(np.prod(x)) for x in od] == [28672, 3584, 448, 112, 28]
    rng = np.random.default_rng(7)
    nx, ny, nz = dims
    h = 1.0 / nx
    c = lambda m: (np.arange(m) + 0.5) * h
    phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                         (0, 0, 0), (0.6, 0.55, 0.5)).astype(np.float32)
    mu, mv, mw = scenes.fluid_face_masks(phi)
    u = np.where(mu, rng.uniform(-0.3, 0.3, mu.shape), 0).astype(np.float32)
    v = np.where(mv, rng.uniform(-0.3, 0.3, mv.shape) - 0.08, 0).astype(np.float32)
    w = np.where(mw, rng.uniform(-0.3, 0.3, mw.shape), 0).astype(np.float32)
    op, (ou, ov, ow), ob, orep = oracle.uniform_projection(phi, u, v, w, h, 1 / 120, 998.2, rel_tol=1e-12)
    g = fpl.UniformPressureProjection(dims, h)
    try:
        res = g.project(phi, u, v, w, 1 / 120, 998.2, precision=precision, relative_tolerance=1e-7)
        assert res.report["flags"] == 0, res.report
        assert res.report["relative_residual"] <= 1e-7
        assert rel_l2(res.pressure, op) < 1e-5
        for a, b in ((res.u, ou), (res.v, ov), (res.w, ow)):
            assert rel_l2(a, b) < 1e-5
    finally:
        g.destroy()


def test_warm_start_is_same_solution_fewer_iterations(fpl, c0, c0_oracle):
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    cold = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, relative_tolerance=1e-7)
    warm = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, relative_tolerance=1e-7, warm_start=1)
    assert warm.report["iterations"] < cold.report["iterations"]
    assert rel_l2(warm.pressure, c0_oracle[0]) < 1e-5
    g.destroy()


def test_device_memory_path_matches_host(fpl, c0):
    import torch
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    host = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho)
    t = [torch.from_numpy(a).cuda() for a in (s.phi, s.u, s.v, s.w)]
    dev = g.project(*t, s.dt, s.rho)
    torch.cuda.synchronize()
    assert np.array_equal(dev.pressure.cpu().numpy(), host.pressure)
    assert np.array_equal(dev.u.cpu().numpy(), host.u)
    g.destroy()


def test_repeatable_bitwise(fpl, c0):
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    a = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho)
    b = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho)
    assert np.array_equal(a.pressure, b.pressure) and np.array_equal(a.w, b.w)
    g.destroy()


@pytest.mark.parametrize("first_fill", [(0.9, 0.8, 0.9), (0.2, 0.3, 0.25)])
def test_moving_liquid_leaves_no_trace(fpl, first_fill):
    """The hull-clamped passes skip the air planes of each 32 x 8 tile and
    leave some vectors there as the previous projection wrote them.  A grid
    that first projected a larger (or smaller) liquid region must give the
    same bits as a fresh grid on the next scene."""
    n = 64
    prev = scenes.dam_break_scene(n, seed=41, fill=first_fill)
    cur = scenes.dam_break_scene(n, seed=42, fill=(0.5, 0.6, 0.45))
    used = fpl.UniformPressureProjection(cur.shape, cur.h)
    fresh = fpl.UniformPressureProjection(cur.shape, cur.h)
    try:
        used.project(prev.phi, prev.u, prev.v, prev.w, prev.dt, prev.rho)
        a = used.project(cur.phi, cur.u, cur.v, cur.w, cur.dt, cur.rho)
        b = fresh.project(cur.phi, cur.u, cur.v, cur.w, cur.dt, cur.rho)
        assert a.report["iterations"] == b.report["iterations"]
        for x, y in ((a.pressure, b.pressure), (a.u, b.u), (a.v, b.v), (a.w, b.w)):
            assert np.array_equal(x, y)
    finally:
        used.destroy()
        fresh.destroy()


def test_iteration_cap_reports_nonconvergence(fpl, c0):
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, max_iterations=1, max_outer=1, relative_tolerance=1e-9)
    assert "NONCONVERGENCE" in res.report["flag_names"]
    assert np.all(np.isfinite(res.pressure))
    g.destroy()


def test_budget_is_enforced(fpl):
    ctx = fpl.Context(0, hbm_budget_bytes=1 << 20)
    with pytest.raises(fpl.FluidProjectionError) as e:
        fpl.UniformPressureProjection((64, 64, 64), 1 / 64, context=ctx)
    assert e.value.code == -5
    ctx.destroy()


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_graph_replay_matches_plain_launches(fpl, c0, precision):
    """Each pair of iterations replayed as a hipGraph gives the same bits as
    plain launches, also across several replays of the same graph."""
    s = c0
    g = fpl.UniformPressureProjection(s.shape, s.h)
    plain = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, graphs=0)
    for _ in range(3):
        gr = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, graphs=1)
        assert gr.report["iterations"] == plain.report["iterations"]
        assert np.array_equal(gr.pressure, plain.pressure)
        assert np.array_equal(gr.u, plain.u) and np.array_equal(gr.w, plain.w)
    g.destroy()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_iteration_counts_stay_multigrid_like(fpl, precision):
    """The preconditioner must keep MGPCG's grid-independent iteration counts:
    a broken storage path still converges (CG is robust) but needs tens of
    times more iterations, which parity checks alone do not see."""
    counts = {}
    for n in (32, 64, 96):
        s = scenes.dam_spheres_scene(n, seed=1)
        g = fpl.UniformPressureProjection(s.shape, s.h)
        res = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, relative_tolerance=1e-6)
        g.destroy()
        assert res.report["flags"] == 0
        counts[n] = res.report["iterations"]
    print(precision, counts)
    assert max(counts.values()) <= 14, counts
