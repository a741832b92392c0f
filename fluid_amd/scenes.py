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


def scene_lattice(scene: dict):
    """(nx, ny, nz, h) of a SceneDescription: voxelDomain.finestCellSize_m over
    the container, at least 8 cells per axis (the sizing of
    createTallCellLayout, lib/core/tall-cell-grid.ts:311-314).  The projection
    needs cubic cells: other containers raise ValueError."""
    c = scene["container"]
    size = scene["voxelDomain"]["finestCellSize_m"]

    def cells(extent):  # JavaScript Math.round: halves round up
        return max(8, int(math.floor(extent / size + 0.5)))

    nx, ny, nz = cells(c["width_m"]), cells(c["height_m"]), cells(c["depth_m"])
    hx, hy, hz = c["width_m"] / nx, c["height_m"] / ny, c["depth_m"] / nz
    if abs(hy - hx) > 1e-6 * hx or abs(hz - hx) > 1e-6 * hx:
        raise ValueError(f"the projection needs cubic cells (container {c['width_m']} x {c['height_m']} x "
                         f"{c['depth_m']} m at {size} m gives {hx}, {hy}, {hz} m)")
    return nx, ny, nz, hx


def dam_break_box(scene: dict):
    """Normalised reservoir box (min, max) of a dam-break scene
    (sceneDamBreakBox, lib/core/initial-fluid.ts:16-63): the fill-derived
    square footprint of height max(0.92, fill) unless authored dimensions or
    an origin are given."""
    c, f = scene["container"], scene["fluid"]
    dims = f.get("initialDamBreakDimensions_m")
    if dims:
        size = (dims["x"] / c["width_m"], dims["y"] / c["height_m"], dims["z"] / c["depth_m"])
    else:
        fill = max(0.0, min(1.0, c["fillFraction"]))
        height = 0.0 if fill == 0 else max(0.92, fill)
        foot = 0.0 if fill == 0 else math.sqrt(fill / height)
        size = (foot, height, foot)
    o = f.get("initialDamBreakOrigin_m")
    lo = (o["x"] / c["width_m"], o["y"] / c["height_m"], o["z"] / c["depth_m"]) if o else (0.0, 0.0, 0.0)
    return lo, tuple(a + b for a, b in zip(lo, size))


def scene_from_description(scene) -> UniformScene:
    """The inputs of a uniform grid for a SceneDescription
    (lib/core/model.ts:52-130; a dict, JSON text or a path to a JSON file such
    as lib/core/default-scene.json), built by the same rules as the Node
    plugin's createSolverAsync: the lattice of scene_lattice, the level set of
    the dam-break reservoir box (signed distance in metres) or of the
    tank-fill level at cell centres (i + 1/2) h, the fluid at rest,
    dt = numerics.fixedDt_s, rho = fluid.density_kg_m3.  ``meta`` carries
    what the step needs: kinematic viscosity (dynamicViscosity / density),
    gravity, open top, no-slip walls, the solver tolerance and iteration cap.
    Box containers with the dam-break or tank-fill initial condition only
    (ValueError otherwise)."""
    import json
    import os
    if isinstance(scene, (str, os.PathLike)):
        text = str(scene)
        if not text.lstrip().startswith("{"):
            with open(scene) as fh:
                text = fh.read()
        scene = json.loads(text)
    for key in ("container", "voxelDomain", "fluid", "numerics"):
        if key not in scene:
            raise ValueError(f"a SceneDescription needs {key}")
    c, f, num = scene["container"], scene["fluid"], scene["numerics"]
    if c.get("shape", "box") != "box":
        raise ValueError("box containers only")
    ic = f["initialCondition"]
    if ic not in ("dam-break", "tank-fill"):
        raise ValueError(f"unsupported initial condition {ic}")
    nx, ny, nz, h = scene_lattice(scene)
    x = (np.arange(nx) + 0.5) * h
    y = (np.arange(ny) + 0.5) * h
    z = (np.arange(nz) + 0.5) * h
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    if ic == "dam-break":
        blo, bhi = dam_break_box(scene)
        ext = (c["width_m"], c["height_m"], c["depth_m"])
        lo = [b * e for b, e in zip(blo, ext)]
        hi = [b * e for b, e in zip(bhi, ext)]
        q = [np.abs(P - (lo[a] + hi[a]) / 2) - (hi[a] - lo[a]) / 2 for a, P in enumerate((X, Y, Z))]
        outside = np.sqrt(np.maximum(q[0], 0) ** 2 + np.maximum(q[1], 0) ** 2 + np.maximum(q[2], 0) ** 2)
        phi = outside + np.minimum(np.maximum(np.maximum(q[0], q[1]), q[2]), 0)
    else:
        phi = Y - c["fillFraction"] * c["height_m"]
    rho = float(f["density_kg_m3"])
    meta = {"nu": float(f["dynamicViscosity_Pa_s"]) / rho, "gravity": float(f["gravity_m_s2"]["y"]),
            "open_top": c.get("top") == "open", "no_slip": c.get("fluidWallMode") == "no-slip",
            "relative_tolerance": float(num["pressureRelativeTolerance"]),
            "max_iterations": max(1, min(1000, int(num["pressureMaxIterations"]))),
            "scene_id": scene.get("sceneId", "")}
    return UniformScene(phi=phi.astype(np.float32), u=np.zeros((nx + 1, ny, nz), np.float32),
                        v=np.zeros((nx, ny + 1, nz), np.float32), w=np.zeros((nx, ny, nz + 1), np.float32),
                        h=h, dt=float(num["fixedDt_s"]), rho=rho, name=scene.get("sceneId", "scene"), meta=meta)


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


# --------------------------------------------------------------------------
# Restricted tall-cell scenes
# --------------------------------------------------------------------------
@dataclass
class TallScene:
    """Inputs of one tall-cell projection (see fluid_amd.tall_layout)."""

    layout: object   # TallCellLayout
    phi: np.ndarray  # (nx, P, nz) float32 samples
    u: np.ndarray    # (nx+1, S, nz) float32 slots
    v: np.ndarray    # (nx, P, nz)
    w: np.ndarray    # (nx, S, nz+1)
    dt: float
    rho: float = RHO_WATER
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def h(self):
        return self.layout.h

    @property
    def samples(self) -> int:
        return int(self.phi.size)


def tall_row_phi(layout, phi, y):
    """phi of every column at fine row(s) y ((nx, nz) or broadcastable):
    linear between the tall samples below the base, the band sample inside
    the band, +5h above it (representedInitialPhiVolumeCells,
    lib/core/tall-cell-grid.ts:228-238)."""
    b = np.asarray(layout.bases, np.int64)
    y = np.broadcast_to(np.asarray(y, np.int64), b.shape)
    bot = phi[:, 0, :].astype(np.float64)
    top = phi[:, 1, :].astype(np.float64)
    k = y / np.maximum(b - 1, 1).astype(np.float64)
    tall = bot + (top - bot) * k
    py = np.clip(y - b + 2, 2, layout.packed_ny - 1)
    band = np.take_along_axis(phi, py[:, None, :], axis=1)[:, 0, :].astype(np.float64)
    in_band = (y >= b) & (y - b < layout.band) & (y < layout.fine_ny)
    out = np.where(y < b, tall, np.where(in_band, band, 5.0 * layout.h))
    return out


def tall_fluid_slots(layout, phi):
    """Slots with an interior face that has a liquid side."""
    from . import tall_layout as tl
    mu, mv, mw = layout.slot_masks()
    b = np.asarray(layout.bases, np.int64)
    liq = phi < 0
    fv = np.zeros_like(mv)
    fv[:, 0, :] = liq[:, 0, :] | liq[:, 1, :]
    fv[:, 1, :] = liq[:, 1, :] | liq[:, 2, :] if layout.packed_ny > 2 else liq[:, 1, :]
    for s in range(2, layout.band + 2):
        upper = liq[:, s + 1, :] if s + 1 < layout.packed_ny else np.zeros_like(liq[:, s, :])
        fv[:, s, :] = liq[:, s, :] | upper
    fu = np.zeros_like(mu)
    fw = np.zeros_like(mw)
    for arr, sl_a, sl_n, ax in ((fu, (slice(0, -1), slice(None)), (slice(1, None), slice(None)), 0),
                                (fw, (slice(None), slice(0, -1)), (slice(None), slice(1, None)), 2)):
        ba, bn = b[sl_a], b[sl_n]
        m = np.minimum(ba, bn)
        tall_liq = (liq[:, 0, :] | liq[:, 1, :])
        for s in range(layout.slots):
            if s < 2:
                f = tall_liq[sl_a] | tall_liq[sl_n]
            else:
                y = m + s - 2
                pa = tall_row_phi(layout, phi, np.pad(y, ((0, 1), (0, 0)) if ax == 0 else ((0, 0), (0, 1)),
                                                      constant_values=0))[sl_a]
                pn = tall_row_phi(layout, phi, np.pad(y, ((1, 0), (0, 0)) if ax == 0 else ((0, 0), (1, 0)),
                                                      constant_values=0))[sl_n]
                f = (pa < 0) | (pn < 0)
            if ax == 0:
                arr[1:-1, s, :] = f
            else:
                arr[:, s, 1:-1] = f
    return fu & mu, fv & mv, fw & mw


def _tall_velocities(rng, layout, phi, dt, noise: str, sigma: float):
    """Slot velocities: noise on fluid slots (slope slots per fine row, i.e.
    scaled by 2/m so the end rows deviate by about sigma), gravity on v."""
    fu, fv, fw = tall_fluid_slots(layout, phi)
    b = np.asarray(layout.bases, np.float64)
    out = []
    for shape, mask, grav, ax in ((fu.shape, fu, 0.0, 0), (fv.shape, fv, -GRAVITY * dt, 1),
                                  (fw.shape, fw, 0.0, 2)):
        if noise == "uniform":
            a = rng.uniform(-sigma, sigma, size=shape).astype(np.float32)
        else:
            a = rng.standard_normal(size=shape, dtype=np.float32) * np.float32(sigma)
        if ax == 0:
            m = np.minimum(b[:-1], b[1:])
            a[1:-1, 1, :] *= (2.0 / m).astype(np.float32)
        elif ax == 2:
            m = np.minimum(b[:, :-1], b[:, 1:])
            a[:, 1, 1:-1] *= (2.0 / m).astype(np.float32)
        else:
            a += np.float32(grav)
        a[~mask] = 0.0
        out.append(np.ascontiguousarray(a))
    return out


def tall_wave_scene(n: int = 512, fine_ny: int = 512, band: int = 32, seed: int = 0, dt: float = 1.0 / 120.0,
                    sigma: float = 0.3, max_delta: int = 4) -> TallScene:
    """Restricted tall-cell grid with a wavy free surface.

    Domain 1 m tall, n x n columns of h = 1/fine_ny.  Surface height
    s(x, z) = 0.6 + 0.05 sin(2 pi 3 x) cos(2 pi 2 z) + U[-0.01, 0.01] at the
    column centres (noise drawn first from PCG64(seed), shape (n, n)); liquid
    below.  phi = (sample height - s), clamped to +-5h (buildInitialPhi's
    narrow band, lib/core/tall-cell-grid.ts:171-190).  Bases: the surface
    cells of each column (last wet row and the dry row above it,
    isInitialSurfaceCell, tall-cell-grid.ts:296-307) through choose_base with
    regularLayers = band, halos 16/8, then limit_bases with delta 4.
    Velocities: N(0, sigma^2) on fluid slots, u, v, w in that order, plus the
    gravity impulse on v.
    """
    from . import tall_layout as tl
    rng = _rng(seed)
    h = 1.0 / fine_ny
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    noise = rng.uniform(-0.01, 0.01, size=(n, n))
    surf = 0.6 + 0.05 * np.sin(2 * np.pi * 3 * c)[:, None] * np.cos(2 * np.pi * 2 * c)[None, :] + noise
    # last wet row J: (J + 0.5) h <= s
    J = np.floor(surf / h - 0.5).astype(np.int64)
    J = np.clip(J, -1, fine_ny - 1)
    lowest = np.where(J >= 0, J, J + 1)
    highest = np.where(J + 1 < fine_ny, J + 1, J)
    # halos clamped to the band as createTallCellLayout does (tall-cell-grid.ts:348)
    raw = tl.choose_base(lowest, highest, fine_ny, band, min(tl.LIQUID_HALO, band), min(tl.AIR_HALO, band))
    bases = tl.limit_bases(raw, max_delta)
    layout = tl.TallCellLayout(bases, fine_ny, band, h, max_delta)
    yc = layout.sample_heights() * h
    phi = np.clip(yc - surf[:, None, :], -5 * h, 5 * h)
    phi = np.where(layout.active(), phi, 5 * h).astype(np.float32)
    u, v, w = _tall_velocities(rng, layout, phi, dt, "gaussian", sigma)
    return TallScene(layout, phi, u, v, w, dt, name="tall_wave", meta={"n": n, "fine_ny": fine_ny, "band": band,
                                                                       "seed": seed})


def tall_dam_scene(n: int = 256, fine_ny: int | None = None, band: int = 32, seed: int = 0,
                   fill=(0.4, 0.6, 0.4), dt: float = 1.0 / 120.0, noise: float = 0.25,
                   max_delta: int = 4) -> TallScene:
    """Corner dam break on a restricted tall-cell grid (n x n columns).

    Liquid box [0, fill] in the unit cube (h = 1/fine_ny vertically, 1/n
    across); phi is the box signed distance at the sample centres clamped
    to +-5h.  Wet columns take choose_base of their surface rows; dry
    columns take the maximum base (tall-cell-grid.ts:380-389) before the
    limiter.  Velocities U[-noise, noise] on fluid slots plus gravity on v,
    drawn from PCG64(seed).
    """
    from . import tall_layout as tl
    rng = _rng(seed)
    ny = fine_ny or n
    h = 1.0 / ny
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    inside = (c[:, None] <= fill[0]) & (c[None, :] <= fill[2])
    J = np.floor(fill[1] * ny - 0.5).astype(np.int64)  # last wet row: (J + 0.5) h <= fill_y
    J = min(max(J, -1), ny - 1)
    max_base = max(0, ny - band)
    if max_base < 2:
        raise ValueError(f"a {band}-row band leaves no tall cell in {ny} rows")
    wet_base = tl.choose_base(np.full((n, n), J), np.full((n, n), min(J + 1, ny - 1)), ny, band,
                              min(tl.LIQUID_HALO, band), min(tl.AIR_HALO, band))
    dry_base = min(2, max_base) if max_base <= 3 else max_base
    raw = np.where(inside & (J >= 0), wet_base, max(2, dry_base)).astype(np.int32)
    bases = tl.limit_bases(raw, max_delta)
    layout = tl.TallCellLayout(bases, ny, band, h, max_delta)
    yc = layout.sample_heights() * h
    px = c[:, None, None]
    pz = c[None, None, :]
    phi = box_sdf(px, yc, pz, (0.0, 0.0, 0.0), tuple(fill))
    phi = np.where(layout.active(), np.clip(phi, -5 * h, 5 * h), 5 * h).astype(np.float32)
    u, v, w = _tall_velocities(rng, layout, phi, dt, "uniform", noise)
    return TallScene(layout, phi, u, v, w, dt, name="tall_dam", meta={"n": n, "fine_ny": ny, "band": band,
                                                                      "fill": tuple(fill), "seed": seed})


def tall_random_scene(nx: int, nz: int, fine_ny: int, band: int, seed: int = 0, max_delta: int = 4,
                      surface=(0.3, 0.7), deep_mixed: bool = False, dt: float = 1.0 / 120.0) -> TallScene:
    """Small random layouts for the oracle and parity tests: a random rough
    surface per column (liquid below), bases through choose_base with halos
    scaled to the band, optional air pockets, Gaussian slot velocities."""
    from . import tall_layout as tl
    rng = _rng(seed)
    h = 1.0 / fine_ny
    surf = rng.uniform(surface[0], surface[1], size=(nx, nz))
    J = np.clip(np.floor(surf / h - 0.5).astype(np.int64), 0, fine_ny - 1)
    lh, ah = max(1, band // 3), max(1, band // 4)
    raw = tl.choose_base(J, np.minimum(J + 1, fine_ny - 1), fine_ny, band, lh, ah)
    raw = np.maximum(raw, 2)
    bases = tl.limit_bases(raw, max_delta)
    layout = tl.TallCellLayout(bases, fine_ny, band, h, max_delta)
    yc = layout.sample_heights() * h
    phi = yc - surf[:, None, :] + rng.uniform(-0.3, 0.3, size=(nx, band + 2, nz)) * h
    if deep_mixed:
        # flip some tall tops to air: mixed tall cells deeper than two rows
        flip = rng.uniform(size=(nx, nz)) < 0.2
        phi[:, 1, :] = np.where(flip, np.abs(phi[:, 1, :]) + 0.5 * h, phi[:, 1, :])
    phi = np.where(layout.active(), np.clip(phi, -5 * h, 5 * h), 5 * h).astype(np.float32)
    u, v, w = _tall_velocities(rng, layout, phi, dt, "gaussian", 0.3)
    return TallScene(layout, phi, u, v, w, dt, name="tall_random", meta={"seed": seed})
