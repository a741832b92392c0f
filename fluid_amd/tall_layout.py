"""Restricted tall-cell layout: column bases, packed samples and face slots.

Host-side restatement of the layout half of lib/core/tall-cell-grid.ts:
  * ``choose_base``      chooseTallCellBase (tall-cell-grid.ts:61-82)
  * ``limit_bases``      limitNeighboringTallCellBases (tall-cell-grid.ts:84-106)
  * ``sample_rows``      the packed-y -> fine-row map of buildInitialPhi
                         (tall-cell-grid.ts:171-190): py 0 is the tall bottom
                         (fine row 0), py 1 the tall top (row base-1), py >= 2
                         the cubic cell base + py - 2, active while < fineNy.

Array conventions of this library (C order, z fastest):
  * column bases (nx, nz) int32, 2 <= base <= fine_ny;
  * samples (nx, P, nz), P = band + 2 -- the reference's packed index
    x + nx * (py + P * z) is the same memory with the x and z axis names
    swapped (``reference_packed_index``), as for the uniform grids;
  * face slots u (nx+1, S, nz), v (nx, P, nz), w (nx, S, nz+1),
    S = band + 2 + max_delta.  v slot 0 is the velocity of the base-1 faces
    inside the tall cell, slot 1 the face tall top -> band row 0, slot 1+k the
    face band k-1 -> k, slot band+1 the face band top -> air.  u/w slots 0 and
    1 are the mean and the per-row slope (about row (m-1)/2) of an affine
    velocity over fine rows 0..m-1, m the smaller base of the two columns;
    slot 2+j is the single fine row m+j, up to the higher column's band top.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# tallCellSettings.high (tall-cell-grid.ts:52-59)
REGULAR_LAYERS = 32
LIQUID_HALO = 16
AIR_HALO = 8
MAX_NEIGHBOR_DELTA = 4


def _js_round(x):
    """Math.round: halves round towards +infinity."""
    return np.floor(np.asarray(x, np.float64) + 0.5).astype(np.int64)


def choose_base(lowest, highest, fine_ny: int, regular_layers: int = REGULAR_LAYERS,
                liquid_halo: int = LIQUID_HALO, air_halo: int = AIR_HALO):
    """Tall-cell height per column from its lowest/highest surface rows
    (chooseTallCellBase, tall-cell-grid.ts:61-82); vectorised over columns."""
    lowest = np.asarray(lowest, np.int64)
    highest = np.asarray(highest, np.int64)
    maximum_base = max(0, fine_ny - regular_layers)
    lower = highest + 1 + air_halo - regular_layers
    upper = lowest + 1 - liquid_halo
    desired = np.where(lower <= upper, _js_round((lower + upper) / 2.0), lower)
    clamped = np.clip(desired, 0, maximum_base)
    if maximum_base < 2:
        return np.zeros_like(clamped, dtype=np.int32)
    return np.maximum(2, clamped).astype(np.int32)


def limit_bases(bases, max_delta: int = MAX_NEIGHBOR_DELTA, passes: int | None = None):
    """Jacobi passes of base <= min(4-neighbour) + max_delta
    (limitNeighboringTallCellBases, tall-cell-grid.ts:84-106; the layout
    builder runs nx + nz passes, tall-cell-grid.ts:395)."""
    cur = np.array(bases, dtype=np.int64)
    nx, nz = cur.shape
    if passes is None:
        passes = nx + nz
    for _ in range(passes):
        up = cur.copy()
        up[1:, :] = np.minimum(up[1:, :], cur[:-1, :] + max_delta)
        up[:-1, :] = np.minimum(up[:-1, :], cur[1:, :] + max_delta)
        up[:, 1:] = np.minimum(up[:, 1:], cur[:, :-1] + max_delta)
        up[:, :-1] = np.minimum(up[:, :-1], cur[:, 1:] + max_delta)
        if np.array_equal(up, cur):
            break
        cur = up
    return cur.astype(np.int32)


def sample_rows(bases, band: int):
    """Fine row of every packed sample (nx, P, nz) and its activity."""
    b = np.asarray(bases, np.int64)[:, None, :]
    py = np.arange(band + 2, dtype=np.int64)[None, :, None]
    rows = np.where(py == 0, 0, np.where(py == 1, b - 1, b + py - 2))
    return rows


def reference_packed_index(x, py, z, nx: int, packed_ny: int):
    """The reference's packed sample index x + nx * (py + packedNy * z)."""
    return np.asarray(x) + nx * (np.asarray(py) + packed_ny * np.asarray(z))


@dataclass
class TallCellLayout:
    """Column bases of one restricted tall-cell grid."""

    bases: np.ndarray     # (nx, nz) int32
    fine_ny: int
    band: int
    h: float
    max_delta: int = MAX_NEIGHBOR_DELTA

    @property
    def nx(self) -> int:
        return int(self.bases.shape[0])

    @property
    def nz(self) -> int:
        return int(self.bases.shape[1])

    @property
    def packed_ny(self) -> int:
        return self.band + 2

    @property
    def slots(self) -> int:
        return self.band + 2 + self.max_delta

    @property
    def samples(self) -> int:
        return self.nx * self.packed_ny * self.nz

    @property
    def equivalent_cells(self) -> int:
        return self.nx * self.fine_ny * self.nz

    def face_shapes(self):
        return ((self.nx + 1, self.slots, self.nz), (self.nx, self.packed_ny, self.nz),
                (self.nx, self.slots, self.nz + 1))

    def validate(self):
        b = np.asarray(self.bases)
        if b.ndim != 2 or b.min() < 2 or b.max() > self.fine_ny:
            raise ValueError("tall-cell bases must lie in [2, fine_ny]")
        if (np.abs(np.diff(b, axis=0)).max(initial=0) > self.max_delta
                or np.abs(np.diff(b, axis=1)).max(initial=0) > self.max_delta):
            raise ValueError("neighbouring tall-cell bases differ by more than max_delta")

    def sample_rows(self):
        return sample_rows(self.bases, self.band)

    def active(self):
        rows = self.sample_rows()
        py = np.arange(self.packed_ny)[None, :, None]
        return (py < 2) | (rows < self.fine_ny)

    def sample_heights(self):
        """Cell-centre height (in cells) of every sample, as buildInitialPhi."""
        b = np.asarray(self.bases, np.float64)[:, None, :]
        py = np.arange(self.packed_ny, dtype=np.float64)[None, :, None]
        return np.where(py == 0, 0.5, np.where(py == 1, np.maximum(0.5, b - 0.5), b + py - 1.5))

    def slot_masks(self):
        """Slots that carry an interior fine face (others are walls / unused)."""
        nx, nz, B, ny = self.nx, self.nz, self.band, self.fine_ny
        b = np.asarray(self.bases, np.int64)
        su, sv, sw = self.face_shapes()
        mv = np.zeros(sv, bool)
        # v slot 0 exists when base >= 2 (always), slot s >= 1 is the face at
        # lower row base + s - 2, interior while that row + 1 < fine_ny
        mv[:, 0, :] = b >= 2
        for s in range(1, B + 2):
            mv[:, s, :] = b + s - 1 < ny
        mu = np.zeros(su, bool)
        mw = np.zeros(sw, bool)
        for arr, bm, bn in ((mu, b[:-1, :], b[1:, :]), (mw, b[:, :-1], b[:, 1:])):
            m = np.minimum(bm, bn)
            top = np.minimum(ny, np.maximum(bm, bn) + B)
            view = arr[1:-1, :, :] if arr is mu else arr[:, :, 1:-1]
            for s in range(self.slots):
                ok = (m + (s - 2) < top) if s >= 2 else np.ones_like(m, bool)
                if arr is mu:
                    view[:, s, :] = ok
                else:
                    view[:, s, :] = ok
        return mu, mv, mw
