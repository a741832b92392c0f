"""Pins of the CPU restatement of the uniform velocity advection and body
forces (oracle/fp_oracle_advect.c, orc_advect_velocity).

What fixes each check, independently of the restatement's formulas:
  * gravity alone: zero velocities gain g dt exactly on the +y faces whose
    lower or upper cell holds liquid above 1e-5 occupancy (occupancy
    1/2 - phi / 4h recomputed in numpy), nothing elsewhere;
  * translation: with a constant transport velocity U along x and a field
    linear in x, the departure point is exact (RK2 of a constant field, in
    pieces of at most 1.5 cells) and trilinear interpolation of a linear field
    is exact, so interior faces hold a + b (x - U dt) -- also for U dt of four
    cells (three pieces) and for MacCormack, whose correction vanishes there;
  * viscosity: with a zero transport field the result is dt nu times the
    discrete Laplacian of the current field (a quadratic in x: 2 / h^2);
  * boundedness: every result before forces lies within the range of the
    transported field (convex combinations, or the MacCormack limiter);
  * the MacCormack limiter on a step, worked by hand: the correction kept
    where it stays within the donors (0.72), the prediction kept where it
    would undershoot them (0, not -0.12);
  * mirror symmetry: mirrored inputs give mirrored outputs up to rounding
    (the interpolation weights 1 - f and f of mirrored positions round
    differently, so not bit for bit; the viscosity's positive-face stencil
    is not mirror-symmetric at walls in the reference either, so nu = 0).
"""
from __future__ import annotations

import numpy as np
import pytest

G = -9.81


def pool(n=16, depth=0.5):
    h = 1.0 / n
    return np.fromfunction(lambda i, j, k: (j + 0.5) * h - depth, (n, n, n)).astype(np.float32), h


def zeros(n):
    return (np.zeros((n + 1, n, n), np.float32), np.zeros((n, n + 1, n), np.float32),
            np.zeros((n, n, n + 1), np.float32))


def test_gravity_alone(oracle):
    n = 16
    phi, h = pool(n, 0.45)
    u, v, w = zeros(n)
    dt = 1.0 / 60.0
    uo, vo, wo = oracle.advect_velocity(u, v, w, h, dt, phi=phi, transport=(u, v, w), gravity=G)
    occ = np.clip(0.5 - phi.astype(np.float32) / np.float32(4 * h), 0, 1)
    lower = occ > 1e-5
    upper = np.zeros_like(lower)
    upper[:, :-1, :] = lower[:, 1:, :]
    want = np.zeros_like(v)
    want[:, 1:, :] = np.where(lower | upper, np.float32(G) * np.float32(dt), 0.0)
    np.testing.assert_array_equal(vo, want)
    assert not uo.any() and not wo.any()


@pytest.mark.parametrize("cells,scheme", [(0.3, "semi-lagrangian"), (4.0, "semi-lagrangian"),
                                          (0.3, "maccormack"), (2.5, "maccormack")])
def test_translation_is_exact(oracle, cells, scheme):
    n = 24
    h = 1.0 / n
    phi = -np.ones((n, n, n), np.float32)  # all liquid: the extrapolation keeps every value
    U = 0.8
    dt = cells * h / U
    u = np.full((n + 1, n, n), U, np.float32)
    u[0] = 0.0
    u[-1] = 0.0
    xs = (np.arange(n) + 0.5) * h
    # v moves the trace along y too: small, so that no trace reaches the y walls
    v = np.broadcast_to((0.02 + 0.05 * xs)[:, None, None], (n, n + 1, n)).astype(np.float32).copy()
    w = np.zeros((n, n, n + 1), np.float32)
    uo, vo, wo = oracle.advect_velocity(u, v, w, h, dt, phi=phi, transport=(u, v, w), gravity=0.0, scheme=scheme)
    lo = int(np.ceil(cells)) + 3
    # MacCormack's reverse trace reaches U dt downstream as well
    hi = n - 3 - (int(np.ceil(cells)) if scheme == "maccormack" else 0)
    inner = slice(lo, hi)
    want = 0.02 + 0.05 * (xs[inner] - U * dt)
    got = vo[inner, 3:-3, 3:-3]
    np.testing.assert_allclose(got, np.broadcast_to(want[:, None, None], got.shape), rtol=0, atol=2e-6)
    # u itself translates as the constant it is
    np.testing.assert_allclose(uo[lo:hi, 3:-3, 3:-3], U, rtol=0, atol=2e-6)


def test_viscosity_is_dt_nu_laplacian(oracle):
    n = 16
    h = 1.0 / n
    phi = -np.ones((n, n, n), np.float32)
    u, v, w = zeros(n)
    xs = (np.arange(n) + 0.5)
    v = np.broadcast_to((xs * xs)[:, None, None], (n, n + 1, n)).astype(np.float32).copy()
    dt, nu = 0.01, 1e-3
    t = zeros(n)
    uo, vo, wo = oracle.advect_velocity(u, v, w, h, dt, phi=phi, transport=t, gravity=0.0, nu=nu)
    want = np.float32(dt) * np.float32(nu) * np.float32(2.0 / (h * h))
    np.testing.assert_allclose(vo[2:-2, 1:-1, :], want, rtol=1e-5)
    assert not uo[1:-1].any() and not wo[:, :, 1:-1].any()


@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_bounded_by_the_transported_field(oracle, scheme):
    n = 20
    phi, h = pool(n, 0.55)
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, (n + 1, n, n)).astype(np.float32)
    v = rng.uniform(-1, 1, (n, n + 1, n)).astype(np.float32)
    w = rng.uniform(-1, 1, (n, n, n + 1)).astype(np.float32)
    t = oracle.extrapolate(u, v, w, h, phi=phi)[:3]
    uo, vo, wo = oracle.advect_velocity(u, v, w, h, 1.0 / 30.0, phi=phi, transport=t, gravity=0.0, scheme=scheme)
    for out, tr, cur, c in ((uo, t[0], u, 0), (vo, t[1], v, 1), (wo, t[2], w, 2)):
        inner = [slice(None)] * 3
        inner[c] = slice(1, -1)
        lo = min(float(tr.min()), float(cur.min())) - 1e-6
        hi = max(float(tr.max()), float(cur.max())) + 1e-6
        assert lo <= float(out[tuple(inner)].min()) and float(out[tuple(inner)].max()) <= hi


@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_mirror_symmetry(oracle, scheme):
    n = 16
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.minimum(Y - 0.3, np.sqrt((X - 0.5) ** 2 + (Y - 0.6) ** 2 + (Z - 0.4) ** 2) - 0.15).astype(np.float32)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, (n + 1, n, n)).astype(np.float32)
    v = rng.uniform(-1, 1, (n, n + 1, n)).astype(np.float32)
    w = rng.uniform(-1, 1, (n, n, n + 1)).astype(np.float32)
    u = np.ascontiguousarray(0.5 * (u - u[::-1]), np.float32)
    v = np.ascontiguousarray(0.5 * (v + v[::-1]), np.float32)
    w = np.ascontiguousarray(0.5 * (w + w[::-1]), np.float32)
    u[0] = 0.0
    u[-1] = 0.0
    uo, vo, wo = oracle.advect_velocity(u, v, w, h, 1.0 / 30.0, phi=phi, scheme=scheme)
    # MacCormack's limiter is a threshold: a rounding-level difference at it
    # picks the prediction on one side and the correction on the other, at a
    # handful of faces
    allowed = 0 if scheme == "semi-lagrangian" else 4
    for a, mirrored in ((uo, -uo[::-1]), (vo, vo[::-1]), (wo, wo[::-1])):
        assert int((np.abs(a - mirrored) > 1e-5).sum()) <= allowed


def test_maccormack_limiter_on_a_step(oracle):
    """A step in v (0 below x = 12 cells, 1 above) carried along x by a
    constant u of 0.4 cells per step, worked by hand.  The prediction
    (semi-Lagrangian) is 0.6 at the face x = 12.5 (departure 12.1, between the
    0 at 11.5 and the 1 at 12.5) and 0 at 11.5.  Its reverse trace (+0.4
    cells) gives 0.24 at 11.5 and 0.76 at 12.5, so the unlimited correction
    pred + (orig - rev) / 2 is -0.12 at 11.5 (below every donor: the limiter
    keeps the prediction, 0) and 0.6 + 0.12 = 0.72 at 12.5 (inside its donors
    0 and 1: kept).  A missing limiter fails the [0, 1] bound; one that always
    reverts fails the 0.72."""
    n = 24
    h = 1.0 / n
    phi = -np.ones((n, n, n), np.float32)
    U = 0.8
    dt = 0.4 * h / U
    u = np.full((n + 1, n, n), U, np.float32)
    u[0] = 0.0
    u[-1] = 0.0
    v = np.zeros((n, n + 1, n), np.float32)
    v[12:] = 1.0
    w = np.zeros((n, n, n + 1), np.float32)
    t = (u, v, w)
    sl = oracle.advect_velocity(u, v, w, h, dt, phi=phi, transport=t, gravity=0.0)[1]
    mc = oracle.advect_velocity(u, v, w, h, dt, phi=phi, transport=t, gravity=0.0, scheme="maccormack")[1]
    inner = (slice(3, n - 3), slice(3, n - 2), slice(3, n - 3))
    col_sl, col_mc = sl[inner][:, 5, 5], mc[inner][:, 5, 5]
    np.testing.assert_allclose(col_sl[12 - 3], 0.6, atol=1e-6)
    np.testing.assert_array_equal(col_sl[11 - 3], 0.0)
    assert float(mc[inner].min()) >= 0.0 and float(mc[inner].max()) <= 1.0
    np.testing.assert_array_equal(col_mc[11 - 3], 0.0)
    np.testing.assert_allclose(col_mc[12 - 3], 0.72, atol=1e-6)
    np.testing.assert_array_equal(col_mc[: 11 - 3], 0.0)
    np.testing.assert_array_equal(col_mc[13 - 3:], 1.0)
