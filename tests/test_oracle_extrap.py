"""Pins of the CPU restatement of the uniform velocity extrapolation
(oracle/fp_oracle_extrap.c, orc_extrapolate).

What fixes each check, independently of the restatement's own formulas:
  * Godunov distances around a single liquid cell equal the textbook values of
    the first-order Eikonal update: h next to a source, (1 + 1/sqrt 2) h on the
    2D diagonal, 2 h two steps away; nothing beyond the two-cell band is known;
  * a numpy re-evaluation of the Godunov update and of the upwind extension on
    the returned band state: every updated face sits at the fixed point;
  * a flat pool with power-of-two h: the band copies the value below exactly
    (the upwind weights are powers of two), the far field holds convex
    combinations of known values (within their range, antisymmetric about
    the mean for a linear profile), a constant field stays constant, source
    faces are returned unchanged;
  * mirror symmetry: mirrored inputs give mirrored outputs bit for bit;
  * the hierarchy leaves exactly the horizontal components of the top row of
    a 16^3 pool unknown (written 0): the face-centred restriction of a
    horizontal component reads the wall face at the coarsest levels and the
    vertical-only corresponding-cell fallback does not apply to it.
"""
from __future__ import annotations

import numpy as np
import pytest

INF = 65504.0


def pool(n=16, depth=0.4, ny=None, nz=None):
    ny = ny or n
    nz = nz or n
    h = 1.0 / n
    phi = np.fromfunction(lambda i, j, k: (j + 0.5) * h - depth, (n, ny, nz)).astype(np.float32)
    return phi, h


def fields(shape, seed=0):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((nx + 1, ny, nz)).astype(np.float32),
            rng.standard_normal((nx, ny + 1, nz)).astype(np.float32),
            rng.standard_normal((nx, ny, nz + 1)).astype(np.float32))


def source_mask(phi, c, open_top=False):
    """Positive faces (lattice of cells) of component c that are sources."""
    liq = phi < 0
    nx, ny, nz = phi.shape
    shifted = np.zeros_like(liq)
    sl = [slice(None)] * 3
    sl[c] = slice(0, -1)
    src = [slice(None)] * 3
    src[c] = slice(1, None)
    shifted[tuple(sl)] = liq[tuple(src)]
    open_ = np.ones_like(liq)
    last = [slice(None)] * 3
    last[c] = -1
    open_[tuple(last)] = bool(open_top and c == 1)
    return open_ & (liq | shifted), open_


def face_view(a, c):
    """The positive-face lattice view (nx, ny, nz) of a face array."""
    sl = [slice(None)] * 3
    sl[c] = slice(1, None)
    return a[tuple(sl)]


def test_single_cell_distances(oracle):
    n, h = 12, 1.0 / 16
    phi = np.full((n, n, n), 1.0, np.float32)
    p0 = (5, 6, 4)
    phi[p0] = -1.0
    u, v, w = fields(phi.shape)
    _, _, _, rep, (dist, val, known) = oracle.extrapolate(u, v, w, h, phi=phi, band=True)
    assert rep["active_left"] == 0
    for c in range(3):
        a, b = [ax for ax in range(3) if ax != c]
        d = dist[c]
        x0 = np.array(p0)

        def at(off):
            q = x0.copy()
            for ax, k in off.items():
                q[ax] += k
            return float(d[tuple(q)])

        assert at({}) == 0.0 and at({c: -1}) == 0.0          # the two faces of the liquid cell
        assert at({a: 1}) == pytest.approx(h, rel=1e-6)
        assert at({b: -1}) == pytest.approx(h, rel=1e-6)
        assert at({a: 1, b: 1}) == pytest.approx((1 + 1 / np.sqrt(2)) * h, rel=1e-6)
        assert at({a: 2}) == pytest.approx(2 * h, rel=1e-6)
        assert at({a: 3}) == INF and at({a: 2, b: 1}) == INF  # beyond the two-cell band
        assert (d[d < INF] <= 2 * h * (1 + 1e-6)).all()


def numpy_upwind(dist, val, known, open_, h, solved):
    """First-order upwind extension (CM11b Eq. 7) of every face from its
    neighbours' distances and values, ties between the two sides averaged (FP64)."""
    D = np.where(open_, dist.astype(np.float64), INF)
    V = np.where(open_ & known, val.astype(np.float64), 0.0)
    Dp = np.pad(D, 1, constant_values=INF)
    Vp = np.pad(V, 1)
    n = D.shape
    num = np.zeros(n)
    den = np.zeros(n)
    for ax in range(3):
        lo = [slice(1, 1 + n[0]), slice(1, 1 + n[1]), slice(1, 1 + n[2])]
        hi = list(lo)
        lo[ax] = slice(0, n[ax])
        hi[ax] = slice(2, 2 + n[ax])
        dl, dh = Dp[tuple(lo)], Dp[tuple(hi)]
        mn = np.minimum(dl, dh)
        tie_l = np.isclose(dl, mn, rtol=1e-6, atol=0)
        tie_h = np.isclose(dh, mn, rtol=1e-6, atol=0)
        mv = np.where(tie_l, Vp[tuple(lo)], 0) + np.where(tie_h, Vp[tuple(hi)], 0)
        mc = tie_l.astype(float) + tie_h.astype(float)
        use = (mn < solved * (1 - 1e-6)) & (mc > 0)
        wgt = np.where(use, (solved - mn) / h ** 2, 0.0)
        num += wgt * np.where(mc > 0, mv / np.maximum(mc, 1), 0)
        den += wgt
    return np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)


def numpy_godunov(dist, open_, h):
    """First-order Godunov update of every face from its 6 neighbours (FP64)."""
    D = np.where(open_, dist.astype(np.float64), INF)
    pad = np.pad(D, 1, constant_values=INF)
    n = D.shape
    mins = []
    for ax in range(3):
        lo = [slice(1, 1 + n[0]), slice(1, 1 + n[1]), slice(1, 1 + n[2])]
        hi = list(lo)
        lo[ax] = slice(0, n[ax])
        hi[ax] = slice(2, 2 + n[ax])
        mins.append(np.minimum(pad[tuple(lo)], pad[tuple(hi)]))
    m = np.sort(np.stack(mins), axis=0)
    root = m[0] + h
    for k in (2, 3):
        use = (m[k - 1] < INF / 2) & (root > m[k - 1])
        mm = m[:k]
        a = k / h ** 2
        b = mm.sum(0) / h ** 2
        cc = -1 + (mm ** 2).sum(0) / h ** 2
        r = (b + np.sqrt(np.maximum(b * b - a * cc, 0))) / a
        root = np.where(use, r, root)
    return np.where(m[0] < INF / 2, root, INF)


@pytest.mark.parametrize("scene", ["pool", "spheres"])
def test_band_state_is_the_fim_fixed_point(oracle, scene):
    n = 24
    h = 1.0 / n
    if scene == "pool":
        phi, h = pool(n, 0.45)
    else:
        x = (np.arange(n) + 0.5) * h
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        phi = np.minimum(Y - 0.3, np.sqrt((X - 0.5) ** 2 + (Y - 0.65) ** 2 + (Z - 0.45) ** 2) - 0.12)
        phi = phi.astype(np.float32)
    u, v, w = fields(phi.shape, 1)
    _, _, _, rep, (dist, val, known) = oracle.extrapolate(u, v, w, h, phi=phi, band=True)
    assert rep["active_left"] == 0 and rep["passes"] >= 2
    for c in range(3):
        src, open_ = source_mask(phi, c)
        kn = (known >> c & 1).astype(bool)
        band = kn & ~src
        assert band.any()
        g = numpy_godunov(dist[c], open_, h)
        # every band face is converged: its distance is the update of its neighbours
        np.testing.assert_allclose(dist[c][band], g[band], rtol=2e-6, atol=0)
        assert (dist[c][band] <= 2 * h * (1 + 1e-6)).all()
        # and its value the upwind extension of its neighbours' values
        ext = numpy_upwind(dist[c], val[c], kn, open_, h, dist[c].astype(np.float64))
        np.testing.assert_allclose(val[c][band], ext[band], rtol=1e-5, atol=1e-6)
        # faces that stayed unknown and touch the band would update above the band
        near = ~kn & open_ & (g <= 2 * h * (1 - 1e-5))
        assert not near.any()
        # sources keep their input value
        np.testing.assert_array_equal(val[c][src], face_view((u, v, w)[c], c)[src])


def test_pool_band_copies_and_convex_far_field(oracle):
    n = 16
    phi, h = pool(n, 0.4)
    xs = (np.arange(n) + 0.5) * h
    u = np.zeros((n + 1, n, n), np.float32)
    w = np.zeros((n, n, n + 1), np.float32)
    v = np.broadcast_to((0.25 + 2.0 * xs)[:, None, None], (n, n + 1, n)).astype(np.float32).copy()
    uo, vo, wo, rep = oracle.extrapolate(u, v, w, h, phi=phi)
    # the band rows (first two air rows) are bit-exact copies of the value below
    top_liquid = int(np.max(np.nonzero(phi[0, :, 0] < 0)))
    for r in (top_liquid + 1, top_liquid + 2):
        np.testing.assert_array_equal(vo[:, r + 1, :], v[:, top_liquid + 1, :])
    # every far-field value is a convex combination of known ones: per x slice
    # bounded by the source range, and mirror-paired about the box centre
    lo, hi = float(v.min()), float(v.max())
    inner = vo[:, 1:-1, :]
    assert (inner >= lo - 1e-6).all() and (inner <= hi + 1e-6).all()
    mean = 0.25 + 2.0 * 0.5
    np.testing.assert_allclose(inner + inner[::-1], 2 * mean, rtol=0, atol=2e-6)
    assert vo[:, 0, :].max() == 0 and vo[:, -1, :].max() == 0  # walls


def test_constant_field_stays_constant(oracle):
    n = 16
    phi, h = pool(n, 0.3)
    u = np.full((n + 1, n, n), 0.75, np.float32)
    v = np.full((n, n + 1, n), -1.25, np.float32)
    w = np.full((n, n, n + 1), 0.5, np.float32)
    uo, vo, wo, rep = oracle.extrapolate(u, v, w, h, phi=phi)
    for a, c, val in ((uo, 0, 0.75), (vo, 1, -1.25), (wo, 2, 0.5)):
        lat = face_view(a, c)
        _, open_ = source_mask(phi, c)
        known = open_ & (lat != 0)
        np.testing.assert_allclose(lat[known], val, rtol=1e-6)


def test_mirror_symmetry_bitwise(oracle):
    n = 16
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    # symmetric in x and z about the box centre
    phi = np.minimum(Y - 0.25, np.sqrt((X - 0.5) ** 2 + (Y - 0.6) ** 2 + (Z - 0.5) ** 2) - 0.15).astype(np.float32)
    rng = np.random.default_rng(3)
    u, v, w = fields(phi.shape, 3)
    # make the inputs mirror-consistent in x: u odd (face i <-> n - i), v, w even
    u = 0.5 * (u - u[::-1])
    v = 0.5 * (v + v[::-1])
    w = 0.5 * (w + w[::-1])
    u, v, w = (np.ascontiguousarray(a, np.float32) for a in (u, v, w))
    uo, vo, wo, _ = oracle.extrapolate(u, v, w, h, phi=phi)
    np.testing.assert_array_equal(uo, -uo[::-1])
    np.testing.assert_array_equal(vo, vo[::-1])
    np.testing.assert_array_equal(wo, wo[::-1])


def test_hierarchy_coverage_of_a_pool(oracle):
    n = 16
    phi, h = pool(n, 0.4)
    u, v, w = fields(phi.shape, 4)
    uo, vo, wo, rep = oracle.extrapolate(u, v, w, h, phi=phi)
    assert rep["levels"] == 4 and rep["unknown"] == 2 * (n - 1) * n
    assert (face_view(uo, 0)[:-1, -1, :] == 0).all() and (face_view(wo, 2)[:, -1, :-1] == 0).all()
    assert (face_view(uo, 0)[:-1, :-1, :] != 0).all() and (face_view(vo, 1)[:, :-1, :] != 0).all()


def test_open_top_and_density_input(oracle):
    n = 16
    phi, h = pool(n, 0.4)
    u, v, w = fields(phi.shape, 5)
    closed = oracle.extrapolate(u, v, w, h, phi=phi)
    opened = oracle.extrapolate(u, v, w, h, phi=phi, open_top=1)
    assert (closed[1][:, -1, :] == 0).all() and (opened[1][:, -1, :] != 0).all()
    np.testing.assert_array_equal(closed[0], opened[0])
    dens = (0.5 - phi / h).astype(np.float32)  # liquid iff density > 1/2
    by_density = oracle.extrapolate(u, v, w, h, density=dens)
    for a, b in zip(closed[:3], by_density[:3]):
        np.testing.assert_array_equal(a, b)


def test_pass_ceiling_reports_active_faces(oracle):
    n = 16
    phi, h = pool(n, 0.4)
    u, v, w = fields(phi.shape, 6)
    rep = oracle.extrapolate(u, v, w, h, phi=phi, max_passes=1)[3]
    assert rep["passes"] == 1 and rep["active_left"] > 0
