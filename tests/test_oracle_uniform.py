"""Pins the uniform-grid oracle against closed forms, invariants and brute force.

Each check is chosen so that a plausible mistake (a dropped term, a wrong sign
or index, a transposed operand) fails it:
  * theta ........ exhaustive table of the cm12GhostFluidTheta cases and clamps
  * A ............ symmetry + positive definiteness + LAPACK solve on small grids
  * rhs/gradient . discrete adjoint identity sum p.b(u) = s sum_f u_f (G p)_f
                   (sparse-atlas-cm12-transport.test.ts:45-80 pattern)
  * solve ........ exact hydrostatic column (closed form), Jacobi-PCG == MGPCG
                   == LAPACK, energy identity (sparse-atlas-composite-projection.ts:1535-1563)
  * planner ...... hand-computed ladders of planUniformCM11aHierarchy
"""
import numpy as np
import pytest

from fluid_amd import scenes


# --------------------------------------------------------------------------- theta
@pytest.mark.parametrize("pl,pa,expect", [
    (-0.5, 0.5, 0.5), (-0.1, 0.9, 0.1), (-0.9, 0.1, 0.9), (-0.01, 0.99, 0.05),  # clamp at 0.05
    (-1.0, 0.0, 1.0), (-2.0, 2.0, 0.5), (-1e-12, 1e-12, 0.05), (-4e-10, 4e-10, 0.4), (0.0, 0.0, 0.05),
])
def test_ghost_theta_closed_form(oracle, pl, pa, expect):
    assert oracle.ghost_theta(pl, pa, 1e-9) == pytest.approx(expect, rel=1e-15, abs=1e-15)


def test_ghost_theta_symmetric_in_sign(oracle):
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, b = rng.uniform(-1, 0), rng.uniform(0, 1)
        t = oracle.ghost_theta(a, b)
        assert t == oracle.ghost_theta(-a, -b)
        assert 0.05 <= t <= 1.0
        assert t == pytest.approx(min(1, max(0.05, abs(a) / (abs(a) + abs(b)))), rel=1e-15)


# --------------------------------------------------------------------------- planner
def test_level_plans_match_hand_computed(oracle):
    d, s, ext = oracle.plan_levels(64, 64, 64)
    assert d.tolist() == [[64] * 3, [32] * 3, [16] * 3, [8] * 3, [4] * 3, [2] * 3] and not ext
    d, s, ext = oracle.plan_levels(384, 384, 384)   # 384 is not a power of two: semi-coarsened
    assert [r[0] for r in d.tolist()] == [384, 192, 96, 48, 24, 12, 6, 3] and not ext
    d, s, ext = oracle.plan_levels(128, 128, 8)     # lockstep stalls, semi-coarsening reaches 2x2x2
    assert d.tolist()[-1] == [2, 2, 2] and not ext
    assert oracle.plan_levels_reference(37, 37, 37) is None   # the reference rejects odd edges
    d, s, ext = oracle.plan_levels(37, 37, 37)
    assert [r[0] for r in d.tolist()] == [37, 19, 10, 5, 3] and ext
    assert s.tolist()[1] == [2, 2, 2] and s.tolist()[0] == [1, 1, 1]
    assert oracle.plan_levels(1, 8, 8) is None


# --------------------------------------------------------------------------- operator
def _small_scene(n=6, seed=3):
    rng = np.random.default_rng(seed)
    s = scenes.dam_break_scene(n, seed, fill=(0.6, 0.55, 0.5))
    return s, rng


def _dense_A(oracle, phi):
    n = phi.size
    liq = np.flatnonzero(phi.ravel() < 0)
    A = np.zeros((len(liq), len(liq)))
    for c, idx in enumerate(liq):
        e = np.zeros(n)
        e[idx] = 1.0
        col = oracle.apply(phi, e.reshape(phi.shape)).ravel()
        A[:, c] = col[liq]
    return A, liq


def test_operator_symmetric_positive_definite(oracle):
    s, _ = _small_scene(6)
    A, liq = _dense_A(oracle, s.phi)
    assert len(liq) > 10
    assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))
    assert np.linalg.eigvalsh(A).min() > 0


def test_operator_rows_are_diagonally_dominant_with_theta(oracle):
    s, _ = _small_scene(6)
    A, liq = _dense_A(oracle, s.phi)
    diag = np.diag(A)
    off = np.abs(A).sum(1) - np.abs(diag)
    assert np.all(diag >= off - 1e-12)
    assert np.any(diag > off + 1e-6)  # free-surface rows: 1/theta terms


def test_adjoint_identity(oracle):
    """sum_c p_c b_c(u) = (rho h/dt) sum_f u_f (p+ - p-) over faces with a liquid side."""
    s, rng = _small_scene(7, 11)
    phi = s.phi
    u = rng.standard_normal(s.u.shape).astype(np.float32)
    v = rng.standard_normal(s.v.shape).astype(np.float32)
    w = rng.standard_normal(s.w.shape).astype(np.float32)
    p = rng.standard_normal(phi.shape) * (phi < 0)
    h, dt, rho = s.h, s.dt, s.rho
    b = oracle.rhs(phi, u, v, w, h, dt, rho)
    lhs = float(np.sum(p * b))
    # gradient of p across every interior face with at least one liquid side
    pl = np.where(phi < 0, p, 0.0)
    liq = phi < 0
    rhs_sum = 0.0
    for arr, ax in ((u, 0), (v, 1), (w, 2)):
        sl_p = [slice(None)] * 3
        sl_m = [slice(None)] * 3
        sl_p[ax] = slice(1, None)
        sl_m[ax] = slice(None, -1)
        g = pl[tuple(sl_p)] - pl[tuple(sl_m)]
        act = liq[tuple(sl_p)] | liq[tuple(sl_m)]
        fa = [slice(None)] * 3
        fa[ax] = slice(1, -1)
        rhs_sum += float(np.sum(arr[tuple(fa)].astype(np.float64) * g * act))
    rhs_sum *= rho * h / dt
    assert lhs == pytest.approx(rhs_sum, rel=1e-12, abs=1e-9)


def test_solvers_agree_with_lapack(oracle):
    s, _ = _small_scene(8, 4)
    b = oracle.rhs(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho)
    A, liq = _dense_A(oracle, s.phi)
    ref = np.zeros(s.phi.size)
    ref[liq] = np.linalg.solve(A, b.ravel()[liq])
    pj, rj = oracle.jacobi_pcg(s.phi, b, rel_tol=1e-13, abs_tol=0)
    pm, rm = oracle.mgpcg(s.phi, b, rel_tol=1e-13)
    assert rj["status"] == 0 and rm["status"] == 0
    scale = np.abs(ref).max()
    assert np.abs(pj.ravel() - ref).max() < 1e-9 * scale
    assert np.abs(pm.ravel() - ref).max() < 1e-9 * scale


def test_hydrostatic_column_is_exact(oracle):
    """Tank of depth H: v = -g dt on fluid faces -> p = rho g (H - y), u' = 0."""
    n, H = 16, 0.53
    h = 1.0 / n
    y = (np.arange(n) + 0.5) * h
    phi = np.broadcast_to((y - H)[None, :, None], (n, n, n)).astype(np.float32).copy()
    dt, rho, g = 1 / 60, 998.2, 9.81
    mu, mv, mw = scenes.fluid_face_masks(phi)
    u = np.zeros((n + 1, n, n), np.float32)
    v = np.where(mv, np.float32(-g * dt), 0).astype(np.float32)
    w = np.zeros((n, n, n + 1), np.float32)
    p, (uo, vo, wo), b, rep = oracle.uniform_projection(phi, u, v, w, h, dt, rho, rel_tol=1e-13)
    exact = np.where(phi < 0, rho * g * (H - y)[None, :, None], 0.0)
    # v was rounded to float32; compare at that precision
    assert np.abs(p - exact).max() < 1e-6 * exact.max()
    assert np.abs(vo).max() < 1e-6 and np.abs(uo).max() < 1e-9 and np.abs(wo).max() < 1e-9


def test_energy_identity_and_divergence(oracle):
    s = scenes.dam_break_scene(24, 7)
    p, (uo, vo, wo), b, rep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
    wx, wy, wz = oracle.face_weights(s.phi)
    ke = 0.0
    for a_in, a_out, wgt in ((s.u, uo, wx), (s.v, vo, wy), (s.w, wo, wz)):
        act = wgt > 0
        theta = np.where(act, 1.0 / np.where(act, wgt, 1.0), 0.0)
        before = a_in.astype(np.float64)
        corr = before - a_out
        ke += 0.5 * np.sum(theta * act * (before ** 2 - a_out ** 2 - corr ** 2))
    ke_scale = 0.5 * sum(float(np.sum(a.astype(np.float64) ** 2)) for a in (s.u, s.v, s.w))
    assert abs(ke) < 1e-9 * ke_scale
    assert rep["relative_residual"] < 1e-12
    assert rep["max_div_after"] < 1e-8 * rep["max_div_before"] + 1e-9


def test_mirror_symmetry(oracle):
    """Mirroring the scene in x mirrors p and negates/mirrors u."""
    s = scenes.dam_break_scene(20, 2, fill=(0.5, 0.6, 0.45))
    phi = s.phi
    phim = phi[::-1].copy()
    um = -s.u[::-1].copy()
    vm = s.v[::-1].copy()
    wm = s.w[::-1].copy()
    p1, _, _, _ = oracle.uniform_projection(phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
    p2, _, _, _ = oracle.uniform_projection(phim, um, vm, wm, s.h, s.dt, s.rho, rel_tol=1e-12)
    assert np.abs(p1[::-1] - p2).max() < 1e-8 * np.abs(p1).max()


def test_unanchored_component_uses_nullspace(oracle):
    """A closed box completely full of liquid is a pure-Neumann problem."""
    n = 8
    phi = -np.ones((n, n, n), np.float32)
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, (n + 1, n, n)).astype(np.float32)
    v = rng.uniform(-1, 1, (n, n + 1, n)).astype(np.float32)
    w = rng.uniform(-1, 1, (n, n, n + 1)).astype(np.float32)
    comp, anch = oracle.components(phi)
    assert len(anch) == 1 and not anch[0]
    p, (uo, vo, wo), b, rep = oracle.uniform_projection(phi, u, v, w, 1 / n, 0.01, 1000.0, rel_tol=1e-11,
                                                        solver="jacobi")
    assert rep["unanchored"] == 1
    assert abs(p.mean()) < 1e-9 * np.abs(p).max()
    assert rep["max_div_after"] < 1e-7 * rep["max_div_before"]


def test_vcycle_is_symmetric_positive(oracle):
    """The preconditioner must be a fixed SPD map for CG (M = M^T > 0)."""
    s = scenes.dam_break_scene(12, 9)
    liq = np.flatnonzero(s.phi.ravel() < 0)
    rng = np.random.default_rng(1)
    x = np.zeros(s.phi.size)
    y = np.zeros(s.phi.size)
    x[liq] = rng.standard_normal(len(liq))
    y[liq] = rng.standard_normal(len(liq))
    Mx = oracle.vcycle(s.phi, x.reshape(s.phi.shape)).ravel()
    My = oracle.vcycle(s.phi, y.reshape(s.phi.shape)).ravel()
    assert np.dot(y, Mx) == pytest.approx(np.dot(x, My), rel=1e-10)
    assert np.dot(x, Mx) > 0 and np.dot(y, My) > 0
    # linear
    Mxy = oracle.vcycle(s.phi, (x - 0.375 * y).reshape(s.phi.shape)).ravel()
    assert np.abs(Mxy - (Mx - 0.375 * My)).max() < 1e-10 * np.abs(Mx).max()


def test_level_phi_preserves_bubbles(oracle):
    """CM11a Eq. 15-16: on the C finest coarse levels a mixed-sign block keeps
    the mean of its non-negative children (stays air); deeper it averages."""
    n = 16
    phi = -np.ones((n, n, n), np.float32) * 0.1
    phi[4:6, 4:6, 4:6] = 0.01  # a 2x2x2 bubble that straddles coarse blocks
    phi[4, 4, 4] = 0.05
    l1 = oracle.level_phi(phi, 1, preserve=2)
    assert l1[2, 2, 2] == pytest.approx((0.05 + 7 * 0.01) / 8)
    l1n = oracle.level_phi(phi, 1, preserve=0)
    assert l1n[2, 2, 2] == pytest.approx((0.05 + 7 * 0.01) / 8)  # all children positive: plain mean
    phi[5, 5, 5] = -0.2
    l1 = oracle.level_phi(phi, 1, preserve=2)
    assert l1[2, 2, 2] == pytest.approx((0.05 + 6 * 0.01) / 7)
    l1n = oracle.level_phi(phi, 1, preserve=0)
    assert l1n[2, 2, 2] == pytest.approx((0.05 + 6 * 0.01 - 0.2) / 8)


def test_level_phi_d4_pairing_and_children(oracle):
    """The coarse level set sums the eight children in FP32 with the pairing
    ((c0+c5)+(c1+c4)) + ((c2+c7)+(c3+c6)) of mgD4Sum8, corner bit 0 = x,
    bit 1 = y, bit 2 = z: with c0 = 2^24, c5 = -2^24 and six ones that sum is
    6 (a left-to-right sum rounds the ones away and gives 2), so the plain mean
    is 0.75; children that are exact binary fractions average exactly."""
    n = 16
    phi = np.full((n, n, n), 0.5, np.float32)
    phi[4:6, 6:8, 2:4] = 1.0
    phi[4, 6, 2] = 2.0 ** 24   # corner 0 of coarse cell (2, 3, 1)
    phi[5, 6, 3] = -2.0 ** 24  # corner 5: x + 1, z + 1
    l1 = oracle.level_phi(phi, 1, preserve=0)
    assert l1[2, 3, 1] == 0.75
    # swapping corners 5 and 6 (x+1, z+1 -> y+1, z+1) pairs 2^24 with a one
    phi[5, 6, 3] = 1.0
    phi[4, 7, 3] = -2.0 ** 24
    l1 = oracle.level_phi(phi, 1, preserve=0)
    assert l1[2, 3, 1] == np.float32(np.float32(2.0 ** 24 + 1.0) + np.float32(2.0)
                                     + np.float32(np.float32(1.0 - 2.0 ** 24) + np.float32(2.0))) / 8
    # exact binary fractions: the mean of children 1/8 .. 8/8 is 9/16
    phi = np.full((n, n, n), 0.5, np.float32)
    for o in range(8):
        phi[6 + (o & 1), 8 + ((o >> 1) & 1), 10 + ((o >> 2) & 1)] = (o + 1) / 8
    assert oracle.level_phi(phi, 1, preserve=2)[3, 4, 5] == 9 / 16


def test_level_phi_odd_axis_repeats_last_child(oracle):
    """An odd lattice (37, ceil-halved to 19) averages the last coarse cell's
    one child along that axis with itself."""
    n = 37
    phi = np.full((n, n, n), 0.25, np.float32)
    phi[36, 0:2, 0:2] = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    l1 = oracle.level_phi(phi, 1, preserve=0)
    assert l1[18, 0, 0] == 2.5


def test_taps1d_reproduce_linear_fields(oracle):
    """Cell-centred linear prolongation: weights 3/4 + 1/4 sum to 1, a tap
    outside the coarse grid is dropped and the other renormalised to 1, and a
    linear field x is reproduced exactly at interior fine centres."""
    nf, nc = 16, 8
    idx, wt = oracle.taps1d(nf, nc, 2)
    assert np.allclose(wt.sum(axis=1), 1.0)
    coarse = 2.0 * np.arange(nc) + 1.0  # coarse centres in fine-cell units
    for f in range(1, nf - 1):
        val = sum(wt[f, q] * coarse[idx[f, q]] for q in range(2) if idx[f, q] >= 0)
        assert val == f + 0.5
    assert idx[0].tolist() == [0, -1] and wt[0].tolist() == [1.0, 0.0]
    assert idx[nf - 1].tolist() == [nc - 1, -1]
    i1, w1 = oracle.taps1d(5, 5, 1)
    assert i1[:, 0].tolist() == list(range(5)) and np.all(w1[:, 0] == 1.0)
    # odd fine axis (37 -> 19): fine 36 lies between coarse 18 and 17
    i3, w3 = oracle.taps1d(37, 19, 2)
    assert i3[36].tolist() == [18, 17] and w3[36].tolist() == [0.75, 0.25]
