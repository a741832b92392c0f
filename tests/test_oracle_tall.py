"""Pins the restricted tall-cell oracle (oracle/fp_oracle.c, orc_tall_*).

The reference holds no tests or vectors for the tall-cell path (its GPU solver
was deleted; only the layout half, lib/core/tall-cell-grid.ts, remains), so
the oracle is pinned by:
  * layout ...... hand-computed chooseTallCellBase / Jacobi-limiter cases and
                  bit-exact agreement of the C restatement with the numpy one
  * operator .... A == E^T D^T W D E assembled independently on the expanded
                  fine grid (tests/tall_brute.py), and == E^T A_u E through
                  the pinned uniform oracle when every deep tall cell is
                  single-phase; symmetric positive definite
  * rhs/subtract  == the brute-force expansion, and the identity
                  rhs(subtract(p)) == b - A p for any p
  * reduction ... base-2 columns whose band covers the domain reproduce the
                  uniform projection of the same grid
  * solve ....... exact hydrostatic columns (closed form with the ghost-fluid
                  surface), relative residual and post-projection divergence
"""
import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd import tall_layout as tl

import tall_brute


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def bind(oracle, sc):
    L = sc.layout
    return oracle.Tall(L.bases, sc.phi, L.fine_ny, L.band, L.h, L.max_delta)


# --------------------------------------------------------------------------- layout
def test_choose_base_hand_computed(oracle):
    # high preset: regular 32, liquid halo 16, air halo 8
    # lowest 300, highest 301: lower = 301+1+8-32 = 278, upper = 300+1-16 = 285 -> round(281.5) = 282
    assert oracle.tall_choose_base(300, 301, 512) == 282
    assert int(tl.choose_base(300, 301, 512)) == 282
    # halos cannot both hold (surface spans 20 rows): the air halo wins -> lower
    assert oracle.tall_choose_base(100, 120, 512) == 120 + 1 + 8 - 32
    # clamps: at most fineNy - regularLayers, at least 2
    assert oracle.tall_choose_base(600, 601, 512) == 480
    assert oracle.tall_choose_base(3, 4, 512) == 2
    # a domain too shallow for a restricted column keeps base 0
    assert oracle.tall_choose_base(5, 6, 33) == 0
    # Math.round of a negative half rounds up: (-19 + -18)/2 = -18.5 -> -18 -> clamped 2
    assert oracle.tall_choose_base(0, 1, 512) == 2


def test_choose_base_c_equals_numpy(oracle):
    rng = np.random.default_rng(3)
    for ny, reg, lh, ah in ((512, 32, 16, 8), (64, 12, 8, 8), (200, 40, 24, 8), (40, 32, 16, 8)):
        lo = rng.integers(-1, ny, size=400)
        hi = lo + rng.integers(0, 30, size=400)
        got = tl.choose_base(lo, hi, ny, reg, lh, ah)
        ref = [oracle.tall_choose_base(int(a), int(b), ny, reg, lh, ah) for a, b in zip(lo, hi)]
        assert np.array_equal(got, np.array(ref, np.int32))


def test_limiter_is_jacobi(oracle):
    # one Jacobi pass reads the previous pass only: [2, 30, 30] -> [2, 6, 30]
    b = np.array([[2], [30], [30]], np.int32)
    assert oracle.tall_limit_bases(b, 4, 1).ravel().tolist() == [2, 6, 30]
    assert tl.limit_bases(b, 4, 1).ravel().tolist() == [2, 6, 30]
    assert oracle.tall_limit_bases(b, 4, 2).ravel().tolist() == [2, 6, 10]
    rng = np.random.default_rng(11)
    for _ in range(20):
        raw = rng.integers(2, 200, size=(rng.integers(1, 12), rng.integers(1, 12))).astype(np.int32)
        d = int(rng.integers(1, 6))
        assert np.array_equal(oracle.tall_limit_bases(raw, d), tl.limit_bases(raw, d))


def test_reference_packed_index():
    # the reference's x + nx (py + P z) equals this library's C-order index with x <-> z swapped
    nx, P, nz = 5, 7, 3
    a = np.arange(nx * P * nz).reshape(nz, P, nx)  # reference memory read as (z, py, x)
    x, py, z = np.meshgrid(np.arange(nx), np.arange(P), np.arange(nz), indexing="ij")
    assert np.array_equal(a[z, py, x], tl.reference_packed_index(x, py, z, nx, P))


def test_sample_rows():
    L = tl.TallCellLayout(np.array([[2, 5], [9, 3]], np.int32), 12, 4, 1 / 12)
    rows = L.sample_rows()
    assert rows[0, :, 1].tolist() == [0, 4, 5, 6, 7, 8]
    assert rows[1, :, 0].tolist() == [0, 8, 9, 10, 11, 12]
    assert L.active()[1, :, 0].tolist() == [True, True, True, True, True, False]


# --------------------------------------------------------------------------- operator
LAYOUTS = [
    dict(nx=5, nz=4, fine_ny=24, band=6, seed=1),
    dict(nx=4, nz=6, fine_ny=30, band=5, seed=2, max_delta=3),
    dict(nx=6, nz=3, fine_ny=20, band=8, seed=3, deep_mixed=True),
    dict(nx=3, nz=3, fine_ny=14, band=12, seed=4),          # band reaches the lid
]


@pytest.fixture(scope="module", params=range(len(LAYOUTS)))
def case(request, oracle):
    sc = scenes.tall_random_scene(**LAYOUTS[request.param])
    return sc, bind(oracle, sc), tall_brute.Expanded(sc.layout, sc.phi)


def test_layouts_are_valid(case):
    sc, t, ex = case
    sc.layout.validate()
    assert t.check() == 0
    assert np.array_equal(t.unknowns().ravel(), ex.unknown)


def test_operator_equals_brute_force(case):
    sc, t, ex = case
    rng = np.random.default_rng(5)
    for _ in range(3):
        p = rng.standard_normal(sc.phi.shape) * t.unknowns()
        assert rel(t.apply(p), ex.apply(p)) < 1e-12


def test_operator_through_uniform_oracle(oracle, case):
    """E^T A_u E with the pinned uniform operator on the expanded grid: equal
    whenever the fine liquid mask is the expansion of the sample mask (no
    mixed tall cell deeper than two rows)."""
    sc, t, ex = case
    bot, top = sc.phi[:, 0, :] < 0, sc.phi[:, 1, :] < 0
    if np.any((bot != top) & (sc.layout.bases > 2)):
        pytest.skip("mixed deep tall cells: the sample-level mask is the definition")
    rng = np.random.default_rng(6)
    p = rng.standard_normal(sc.phi.shape) * t.unknowns()
    pf = (ex.E @ p.ravel()).reshape(ex.shape)
    au = oracle.apply(ex.phi_fine.astype(np.float32), pf)
    got = (ex.E.T @ au.ravel()).reshape(sc.phi.shape) * t.unknowns()
    # the uniform oracle reads phi in float32: interpolated rows round there
    assert rel(t.apply(p), got) < 1e-6


def test_operator_spd(case):
    sc, t, ex = case
    u = ex.unknown
    A = ex.A[u][:, u].toarray()
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > 0
    assert np.allclose(np.diag(A), t.diagonal().ravel()[u], rtol=1e-12)


def test_rhs_and_subtract_equal_brute_force(case):
    sc, t, ex = case
    b = t.rhs(sc.u, sc.v, sc.w, sc.dt, sc.rho)
    assert rel(b, ex.rhs(sc.u, sc.v, sc.w, sc.dt, sc.rho)) < 1e-12
    rng = np.random.default_rng(7)
    p = rng.standard_normal(sc.phi.shape) * t.unknowns() * 50.0
    got = t.subtract(p, sc.u, sc.v, sc.w, sc.dt, sc.rho)
    ref = ex.subtract(p, sc.u, sc.v, sc.w, sc.dt, sc.rho)
    for a, r in zip(got, ref):
        assert rel(a, r) < 1e-12


def test_rhs_of_subtracted_field(case):
    """b(u - s W G p) == b(u) - A p for any p: the slot projection keeps the
    reduced divergence exact."""
    sc, t, ex = case
    rng = np.random.default_rng(8)
    p = rng.standard_normal(sc.phi.shape) * t.unknowns() * 30.0
    b = t.rhs(sc.u, sc.v, sc.w, sc.dt, sc.rho)
    uo, vo, wo = t.subtract(p, sc.u, sc.v, sc.w, sc.dt, sc.rho)
    b2 = t.rhs(uo, vo, wo, sc.dt, sc.rho)
    assert rel(b2, (b - t.apply(p)) * t.unknowns()) < 1e-11


def test_projection_converges(case):
    sc, t, ex = case
    p, (uo, vo, wo), b, rep = t.projection(sc.u, sc.v, sc.w, sc.dt, sc.rho, rel_tol=1e-12)
    assert rep["status"] == 0
    assert rep["relative_residual"] < 1e-11
    assert rep["max_div_after"] < 1e-9 * max(1.0, rep["max_div_before"])
    # the dense solve agrees
    u = ex.unknown
    x = np.linalg.solve(ex.A[u][:, u].toarray(), b.ravel()[u])
    assert rel(p.ravel()[u], x) < 1e-9


# --------------------------------------------------------------------------- reductions
def test_base_two_band_to_lid_is_uniform(oracle):
    """Base 2 everywhere with the band reaching the lid is a re-indexed
    uniform grid: pressures equal the uniform oracle's, and fine velocities
    (slots 0/1 are the mean and slope of two rows) equal its projection."""
    nx, ny, nz = 6, 10, 5
    B = ny - 2
    rng = np.random.default_rng(9)
    h = 1.0 / ny
    uni = scenes.dam_break_scene(n=ny, seed=4, fill=(0.7, 0.6, 0.5))
    phi_u = uni.phi[:nx, :, :nz].copy()
    L = tl.TallCellLayout(np.full((nx, nz), 2, np.int32), ny, B, h)
    phi_t = np.ascontiguousarray(phi_u)  # rows 0..ny-1 are py 0..ny-1
    t = oracle.Tall(L.bases, phi_t, ny, B, h, L.max_delta)
    ex = tall_brute.Expanded(L, phi_t)
    # random fine velocities on liquid-touching interior faces, as slots
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # the uniform oracle reads float32
    fu = f32(rng.standard_normal((nx + 1, ny, nz))); fu[0] = fu[-1] = 0
    fv = f32(rng.standard_normal((nx, ny + 1, nz)) - 9.81 * uni.dt); fv[:, 0] = fv[:, -1] = 0
    fw = f32(rng.standard_normal((nx, ny, nz + 1))); fw[:, :, 0] = fw[:, :, -1] = 0
    f = np.concatenate([fu[1:-1].ravel(), fv[:, 1:-1].ravel(), fw[:, :, 1:-1].ravel()])
    us, vs, ws = ex.project_slots(f)
    assert np.allclose(ex.expand_velocity(us, vs, ws), f, atol=1e-12)  # lossless for m = 2
    p, (uo, vo, wo), b, rep = t.projection(us, vs, ws, uni.dt, uni.rho, rel_tol=1e-13)
    pu, (uu, vu, wu), bu, repu = oracle.uniform_projection(phi_u, fu.astype(np.float32), fv.astype(np.float32),
                                                           fw.astype(np.float32), h, uni.dt, uni.rho,
                                                           rel_tol=1e-13, solver="pcg")
    assert rel(b, bu) < 1e-12
    assert rel(p, pu) < 1e-10
    fo = ex.expand_velocity(uo, vo, wo)
    fu_o = np.concatenate([uu[1:-1].ravel(), vu[:, 1:-1].ravel(), wu[:, :, 1:-1].ravel()])
    assert rel(fo, fu_o) < 1e-10


def test_hydrostatic_columns(oracle):
    """Flat surface, velocity = gravity impulse: p = rho g h (theta + J - y)
    on every fine row (linear, so the tall cells represent it exactly) and
    the projected velocity vanishes."""
    nx, nz, ny, B = 5, 4, 40, 8
    h = 1.0 / ny
    s_height = 0.61 * ny  # surface in cells (row J = 23 liquid, theta from phi)
    bases = np.full((nx, nz), 20, np.int32)
    bases[1, 2] = 17
    bases[3, 1] = 22
    L = tl.TallCellLayout(bases, ny, B, h)
    yc = L.sample_heights()
    phi = np.where(L.active(), np.clip((yc - s_height) * h, -5 * h, 5 * h), 5 * h).astype(np.float32)
    dt, rho, g = 1 / 120, 998.2, 9.81
    fu, fv, fw = scenes.tall_fluid_slots(L, phi)
    u = np.zeros(fu.shape)
    w = np.zeros(fw.shape)
    v = np.where(fv, -g * dt, 0.0)
    v[:, 0, :] = np.where(fv[:, 0, :], -g * dt, 0.0)
    t = oracle.Tall(bases, phi, ny, B, h)
    p, (uo, vo, wo), b, rep = t.projection(u, v, w, dt, rho, rel_tol=1e-14)
    rows = L.sample_rows()
    J = int(np.floor(s_height - 0.5))
    phl = float(np.float32((J + 0.5 - s_height) * h))  # the float32 samples theta is taken from
    pha = float(np.float32((J + 1.5 - s_height) * h))
    theta = min(1.0, max(0.05, abs(phl) / (abs(phl) + abs(pha))))
    exact = rho * g * h * (theta + J - rows)
    liq = t.unknowns()
    assert rel(p[liq], exact[liq]) < 1e-10
    assert np.abs(vo).max() < 1e-9 and np.abs(uo).max() < 1e-9 and np.abs(wo).max() < 1e-9


def test_matrix_export_equals_apply(case):
    sc, t, ex = case
    A = t.matrix()
    assert abs(A - ex.A).max() < 1e-12 * abs(ex.A).max()
