"""Inputs of the tall-cell remesh tests: a layout whose phi describes a
moved surface, so that the remesh changes bases."""
from __future__ import annotations

import numpy as np

from fluid_amd import scenes


def moved_scene(nx=12, nz=10, ny=96, band=24, seed=0, drop=18, max_delta=4, linear=False):
    """A tall_random_scene layout whose phi describes a surface moved by
    `drop` rows over half the columns (so bases change), plus its slots."""
    sc = scenes.tall_random_scene(nx, nz, ny, band, seed=seed, max_delta=max_delta)
    lay = sc.layout
    h = lay.h
    rng = np.random.default_rng(seed + 100)
    surf = rng.uniform(0.45, 0.6, size=(nx, nz))
    surf[: nx // 2] -= drop * h
    yc = lay.sample_heights() * h
    phi = yc - surf[:, None, :]
    if not linear:
        phi = phi + rng.uniform(-0.2, 0.2, size=phi.shape) * h
        phi = np.clip(phi, -5 * h, 5 * h)
    phi = np.where(lay.active(), phi, 5 * h).astype(np.float32)
    return lay, phi, sc.u, sc.v, sc.w


def dropped_scene(sc, rows, columns=None):
    """phi of a tall scene with its surface `rows` fine rows lower (over the
    first `columns` x-slices; all when None), in the scene's own layout.
    Samples clamped at -5 h (deeper than the narrow band) stay liquid; the
    others move up by `rows` h and are clamped to +-5 h again, as the scene
    generators clamp them."""
    lay = sc.layout
    h = lay.h
    phi = sc.phi.astype(np.float64)
    moved = np.where(phi <= -5 * h * (1 - 1e-6), phi, phi + rows * h)
    if columns is not None:
        moved[columns:] = phi[columns:]
    moved = np.where(lay.active(), np.clip(moved, -5 * h, 5 * h), 5 * h)
    return moved.astype(np.float32)
