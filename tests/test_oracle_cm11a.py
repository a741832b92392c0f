"""Pins of the CM11a separating-wall restatement (oracle/fp_oracle_cm11a.c)
by what the mathematics fixes, on the CPU.

* No active constraint (pressure above p_min everywhere): a halo unknown has
  one coupling, to its interior neighbour, and b = 0, so it copies that
  neighbour and no flux crosses the wall; the interior equations are then the
  Neumann-wall ghost-fluid system of the FP64 oracle (orc_mgpcg), whose
  pressure the CM11a cycles must reach.
* Separating walls: the converged result satisfies the LCP's
  complementarity conditions and equals a brute-force projected Gauss-Seidel
  solve (FP64, numpy, thousands of sweeps) of the same assembled system.
* The reference's schedule (3 full cycles + 4 V-cycles, 6 + 6 sweeps)
  reaches the same solutions closely on these grids.
"""
import numpy as np
import pytest

from fluid_amd import scenes


def hydrostatic(n_or_dims, seed=0):
    dims = (n_or_dims,) * 3 if isinstance(n_or_dims, int) else tuple(n_or_dims)
    nx, ny, nz = dims
    h = 1.0 / nx
    c = lambda m: (np.arange(m) + 0.5) * h
    phi = scenes.box_sdf(c(nx)[:, None, None], c(ny)[None, :, None], c(nz)[None, None, :],
                         (0, 0, 0), (0.4 * nx * h, 0.6 * ny * h, 0.4 * nz * h)).astype(np.float32)
    dt, rho = 1 / 120, 998.2
    mu, mv, mw = scenes.fluid_face_masks(phi)
    u = np.zeros((nx + 1, ny, nz), np.float32)
    v = np.where(mv, -9.81 * dt, 0).astype(np.float32)
    v[:, 0] = 0
    v[:, -1] = 0
    w = np.zeros((nx, ny, nz + 1), np.float32)
    return phi, u, v, w, h, dt, rho


def separating(n, speed=1.0):
    """The dam moving away from the x = 0 wall: suction there in the Neumann
    solution, so the separating wall takes p = 0 (active constraints)."""
    phi, u, v, w, h, dt, rho = hydrostatic(n)
    mu, _, _ = scenes.fluid_face_masks(phi)
    u = np.where(mu, speed, 0).astype(np.float32)
    u[0] = 0
    u[-1] = 0
    return phi, u, v, w, h, dt, rho


def brute_pgs(coef, b, pmin, sweeps):
    """FP64 projected red-black Gauss-Seidel on the assembled padded system."""
    D = b.shape
    liq = coef[..., 3] > 0.5
    a = np.zeros(D + (6,))
    for ax in range(3):
        cp = coef[..., ax].astype(np.float64)
        a[..., 2 * ax + 1] = cp
        lo = [slice(None)] * 3
        lo[ax] = slice(1, None)
        hi = [slice(None)] * 3
        hi[ax] = slice(0, -1)
        a[tuple(lo) + (2 * ax,)] = cp[tuple(hi)]
    diag = a.sum(-1)
    X, Y, Z = np.meshgrid(*[np.arange(d) for d in D], indexing="ij")
    col = (X + Y + Z) & 1
    bb = b.astype(np.float64)
    pm = pmin.astype(np.float64)

    def nsum(p):
        s = np.zeros(D)
        pl = np.where(liq, p, 0.0)
        for ax in range(3):
            lo = [slice(None)] * 3
            lo[ax] = slice(1, None)
            hi = [slice(None)] * 3
            hi[ax] = slice(0, -1)
            s[tuple(hi)] += a[tuple(hi) + (2 * ax + 1,)] * pl[tuple(lo)]
            s[tuple(lo)] += a[tuple(lo) + (2 * ax,)] * pl[tuple(hi)]
        return s

    p = np.zeros(D)
    for _ in range(sweeps):
        for c in (0, 1):
            new = np.where(diag > 0, (nsum(p) + bb) / np.where(diag > 0, diag, 1.0), 0.0)
            p = np.where((col == c) & liq, np.maximum(new, pm), p)
    r = np.where(liq, bb - (diag * p - nsum(p)), 0.0)
    return p, r, liq


@pytest.mark.parametrize("dims", [16, (12, 10, 8), (12, 16, 8), (24, 20, 16)])
def test_no_active_constraint_is_the_neumann_solution(oracle, dims):
    phi, u, v, w, h, dt, rho = hydrostatic(dims)
    P, (uo, vo, wo), rep = oracle.cm11a_project(phi, u, v, w, h, dt, rho, v_cycles=40)
    op, (ou, ov, ow), _, _ = oracle.uniform_projection(phi, u, v, w, h, dt, rho, rel_tol=1e-12)
    liq = phi < 0
    pi = P[1:-1, 1:-1, 1:-1]
    assert pi[liq].min() > 0.0
    assert np.linalg.norm((pi - op)[liq]) / np.linalg.norm(op[liq]) < 2e-5  # FP32 arithmetic
    # a halo unknown beside liquid copies its neighbour: no flux through the wall
    assert np.allclose(P[0, 1:-1, 1:-1][liq[0]], pi[0][liq[0]], rtol=2e-5)
    assert np.allclose(P[1:-1, 0, 1:-1][liq[:, 0]], pi[:, 0][liq[:, 0]], rtol=2e-5)
    # velocities: the Neumann oracle's on interior faces, walls unchanged at 0
    assert np.abs(uo - ou).max() < 2e-4 and np.abs(vo - ov).max() < 2e-4 and np.abs(wo - ow).max() < 2e-4
    assert rep["coarse_failed"] == 0
    # the reference's schedule gets close too
    P2, _, rep2 = oracle.cm11a_project(phi, u, v, w, h, dt, rho)
    assert np.linalg.norm((P2[1:-1, 1:-1, 1:-1] - op)[liq]) / np.linalg.norm(op[liq]) < 1e-4


@pytest.mark.parametrize("open_top", [0, 1])
def test_separating_wall_matches_brute_force_lcp(oracle, open_top):
    phi, u, v, w, h, dt, rho = separating(12)
    coef, b, pmin = oracle.cm11a_system(phi, u, v, w, h, dt, rho, open_top)
    pb, rb, liq = brute_pgs(coef, b, pmin, 5000)
    # the brute force itself satisfies complementarity
    free = liq & (pb > pmin + 1e-6)
    active = liq & ~free
    assert active.sum() > 10, "the scene must activate separating walls"
    scale = np.abs(b).max()
    assert np.abs(rb[free]).max() < 1e-9 * scale
    assert rb[active].max() <= 1e-9 * scale and (pb[liq] >= pmin[liq] - 1e-9).all()
    P, _, rep = oracle.cm11a_project(phi, u, v, w, h, dt, rho, v_cycles=60, open_top=open_top)
    assert np.abs(P - pb)[liq].max() / np.abs(pb[liq]).max() < 2e-5
    assert (P[liq] >= pmin[liq]).all()
    P2, _, _ = oracle.cm11a_project(phi, u, v, w, h, dt, rho, open_top=open_top)
    assert np.abs(P2 - pb)[liq].max() / np.abs(pb[liq]).max() < 1e-3


def test_open_top_is_air(oracle):
    """A liquid-filled box open at the top, velocities of one gravity step:
    the hydrostatic closed form, p falling by rho g h per cell up to the open
    top, where the liquid meets air."""
    nx = ny = nz = 8
    h, dt, rho, g = 1 / nx, 1 / 120, 998.2, 9.81
    phi = np.full((nx, ny, nz), -0.5 * h, np.float32)  # all liquid
    v = np.zeros((nx, ny + 1, nz), np.float32)
    v[:, 1:-1] = -g * dt
    u = np.zeros((nx + 1, ny, nz), np.float32)
    w = np.zeros((nx, ny, nz + 1), np.float32)
    P, (uo, vo, wo), rep = oracle.cm11a_project(phi, u, v, w, h, dt, rho, v_cycles=40, open_top=1)
    pi = P[1:-1, 1:-1, 1:-1].astype(np.float64)
    assert np.allclose(np.diff(pi, axis=1), -rho * g * h, rtol=1e-4)
    assert np.abs(pi[:, -1]).max() < 1e-4 * pi[:, 0].max()
    # hydrostatic balance: the interior faces come to rest
    assert np.abs(vo[:, 1:-1]).max() < 1e-5


def test_lattices_without_exact_halving_are_rejected(oracle):
    phi, u, v, w, h, dt, rho = hydrostatic((9, 14, 11))
    with pytest.raises(RuntimeError):
        oracle.cm11a_project(phi, u, v, w, h, dt, rho)


def test_target_divergence_enters_the_rhs(oracle):
    phi, u, v, w, h, dt, rho = hydrostatic(12)
    q = np.where(phi < 0, 0.3, 0).astype(np.float32)
    coef, b0, _ = oracle.cm11a_system(phi, u, v, w, h, dt, rho)
    P0, _, _ = oracle.cm11a_project(phi, u, v, w, h, dt, rho)
    P1, _, _ = oracle.cm11a_project(phi, u, v, w, h, dt, rho, q=q)
    assert not np.array_equal(P0, P1)
