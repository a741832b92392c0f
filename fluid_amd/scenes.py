"""Seeded synthetic projection inputs (the scene generators of the workloads).

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` in a fixed,
documented order, so the same seed gives bit-identical inputs on every host.

Conventions (shared by the library, the tests and the benchmark):
  * unit-cube domain [0,1]^3 split into n^3 cubic cells of size h = 1/n, with
    closed walls on all six faces;
  * level set phi at cell centres ((i+1/2)h, (j+1/2)h, (k+1/2)h), liquid iff
    phi < 0 (mgLiquid, lib/methods/uniform/webgpu-uniform-pressure-multigrid.wgsl.ts:80);
  * MAC face velocities u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1), C order,
    z fastest; wall faces are zero;
  * a "fluid face" is an interior face with at least one liquid neighbour; the
    gravity impulse and the velocity noise are applied to fluid faces only.

The dam-break box level set is the exact box signed distance of
lib/core/tall-cell-grid.ts:108-115 (``boxSignedDistance``); sphere volumes are
unioned by taking the minimum of signed distances as in
lib/core/tall-cell-grid.ts:124-127.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

RHO_WATER = 998.2  # lib/core/default-scene.json fluid density (kg/m^3)
GRAVITY = 9.81


@dataclass
class UniformScene:
    """Inputs of one uniform-grid projection."""

    phi: np.ndarray  # (nx, ny, nz) float32
    u: np.ndarray    # (nx+1, ny, nz) float32
    v: np.ndarray    # (nx, ny+1, nz) float32
    w: np.ndarray    # (nx, ny, nz+1) float32
    h: float
    dt: float
    rho: float = RHO_WATER
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.phi.shape

    @property
    def cells(self) -> int:
        return int(self.phi.size)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def box_sdf(px, py, pz, lo, hi):
    """Signed distance to the axis-aligned box [lo, hi] (negative inside)."""
    cx, cy, cz = [(a + b) * 0.5 for a, b in zip(lo, hi)]
    hx, hy, hz = [(b - a) * 0.5 for a, b in zip(lo, hi)]
    qx = np.abs(px - cx) - hx
    qy = np.abs(py - cy) - hy
    qz = np.abs(pz - cz) - hz
    outside = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2)
    inside = np.minimum(np.maximum(np.maximum(qx, qy), qz), 0)
    return outside + inside


def _phi_field(n: int, box_lo, box_hi, spheres=(), slab: int = 16) -> np.ndarray:
    h = 1.0 / n
    c = (np.arange(n, dtype=np.float64) + 0.5) * h
    phi = np.empty((n, n, n), np.float32)
    py = c[None, :, None]
    pz = c[None, None, :]
    for x0 in range(0, n, slab):
        px = c[x0:x0 + slab, None, None]
        d = box_sdf(px, py, pz, box_lo, box_hi)
        for (sx, sy, sz, r) in spheres:
            ds = np.sqrt((px - sx) ** 2 + (py - sy) ** 2 + (pz - sz) ** 2) - r
            d = np.minimum(d, ds)
        phi[x0:x0 + slab] = d
    return phi


def fluid_face_masks(phi: np.ndarray):
    """Boolean masks of interior faces with at least one liquid neighbour."""
    liq = phi < 0
    nx, ny, nz = phi.shape
    mu = np.zeros((nx + 1, ny, nz), bool)
    mu[1:nx] = liq[:-1] | liq[1:]
    mv = np.zeros((nx, ny + 1, nz), bool)
    mv[:, 1:ny] = liq[:, :-1] | liq[:, 1:]
    mw = np.zeros((nx, ny, nz + 1), bool)
    mw[:, :, 1:nz] = liq[:, :, :-1] | liq[:, :, 1:]
    return mu, mv, mw


def _velocities(rng, phi, dt, noise: str, sigma: float):
    nx, ny, nz = phi.shape
    mu, mv, mw = fluid_face_masks(phi)
    out = []
    for shape, mask, grav in (((nx + 1, ny, nz), mu, 0.0), ((nx, ny + 1, nz), mv, -GRAVITY * dt),
                              ((nx, ny, nz + 1), mw, 0.0)):
        if noise == "uniform":
            a = rng.uniform(-sigma, sigma, size=shape).astype(np.float32)
        else:
            a = rng.standard_normal(size=shape, dtype=np.float32) * np.float32(sigma)
        a += np.float32(grav)
        a[~mask] = 0.0
        out.append(np.ascontiguousarray(a))
    return out


def dam_break_scene(n: int = 64, seed: int = 0, fill=(0.4, 0.6, 0.4), dt: float = 1.0 / 120.0,
                    noise: float = 0.25, name: str = "dam_break") -> UniformScene:
    """Corner dam break: liquid box [0,fill] in a closed unit box.

    Velocity = gravity impulse -9.81*dt on v plus U[-noise, noise] on fluid
    faces, drawn for u, v, w in that order from PCG64(seed).
    With n=64 and fill (0.4, 0.6, 0.4) this is the 64^3 CPU-runnable workload.
    """
    rng = _rng(seed)
    phi = _phi_field(n, (0.0, 0.0, 0.0), tuple(fill))
    u, v, w = _velocities(rng, phi, dt, "uniform", noise)
    return UniformScene(phi, u, v, w, 1.0 / n, dt, name=name, meta={"n": n, "seed": seed, "fill": tuple(fill)})


def dam_spheres_scene(n: int = 384, seed: int = 0, dt: float = 1.0 / 240.0, sigma: float = 0.5,
                      spheres: int = 8) -> UniformScene:
    """Large single grid: dam [0,0.5]x[0,0.7]x[0,1] unioned with `spheres` balls.

    Draw order from PCG64(seed): radii U[0.03,0.08] (spheres values), centres
    (x, y, z) per sphere with x,z ~ U[0,1], y ~ U[0.5,1]; then N(0, sigma^2)
    face velocities for u, v, w; plus the gravity impulse -9.81*dt on v.
    """
    rng = _rng(seed)
    radii = rng.uniform(0.03, 0.08, size=spheres)
    centres = np.stack([rng.uniform(0.0, 1.0, spheres), rng.uniform(0.5, 1.0, spheres),
                        rng.uniform(0.0, 1.0, spheres)], axis=1)
    sph = [(float(cx), float(cy), float(cz), float(r)) for (cx, cy, cz), r in zip(centres, radii)]
    phi = _phi_field(n, (0.0, 0.0, 0.0), (0.5, 0.7, 1.0), sph)
    u, v, w = _velocities(rng, phi, dt, "gaussian", sigma)
    return UniformScene(phi, u, v, w, 1.0 / n, dt, name="dam_spheres",
                        meta={"n": n, "seed": seed, "spheres": sph})


def batch_sizes(count: int = 96, seed: int = 0, lo: int = 32, hi: int = 192):
    """Edge lengths n = round(exp(U[ln lo, ln hi])) drawn from PCG64(seed)."""
    rng = _rng(seed)
    return [int(round(math.exp(x))) for x in rng.uniform(math.log(lo), math.log(hi), size=count)]


def batch_scenes(count: int = 96, seed: int = 0, lo: int = 32, hi: int = 192, max_edge=None):
    """Uneven batch: cubic grids with random dam fills U[0.2,0.8] per axis.

    Grid g uses PCG64(seed + 1 + g) for its fill fractions (3 draws) and then
    its velocities exactly as `dam_break_scene` (dt = 1/120, U[-0.25,0.25]).
    """
    sizes = batch_sizes(count, seed, lo, hi)
    scenes = []
    for g, n in enumerate(sizes):
        if max_edge is not None:
            n = min(n, max_edge)
        rng = _rng(seed + 1 + g)
        fill = tuple(float(x) for x in rng.uniform(0.2, 0.8, size=3))
        phi = _phi_field(n, (0.0, 0.0, 0.0), fill)
        u, v, w = _velocities(rng, phi, 1.0 / 120.0, "uniform", 0.25)
        scenes.append(UniformScene(phi, u, v, w, 1.0 / n, 1.0 / 120.0, name=f"batch{g}",
                                   meta={"n": n, "fill": fill, "index": g}))
    return scenes
