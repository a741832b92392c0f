"""Pins of the CPU restatement of the tall-cell remesh (oracle/tall_remesh.py).

Each check is fixed by something other than the restatement's own formulas:
the base by the closed form of the neighbour limit (the largest layout below
the surface-chosen bases whose neighbours differ by at most D is
min_j raw_j + D |i - j|_1) over surface rows found on an independent dense
expansion of the samples (fluid_amd.scenes.tall_row_phi); the tall samples by
numpy's least-squares line (np.polyfit) through the dense column; linear
columns and affine faces by their exact reproduction; v slot 0 by the mean
of the faces it replaces; and a second remesh by idempotence.
"""
from __future__ import annotations

import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd import tall_layout as tl
import tall_remesh as orm
from remesh_cases import moved_scene

LH, AH = 16, 8


def dense_phi(lay, phi):
    """(nx, ny, nz) phi of every fine row, by the independent expansion."""
    return np.stack([scenes.tall_row_phi(lay, phi, y) for y in range(lay.fine_ny)], axis=1)


def closed_form_limit(raw, d):
    nx, nz = raw.shape
    x, z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    out = np.empty_like(raw)
    for i in range(nx):
        for k in range(nz):
            out[i, k] = np.min(raw + d * (np.abs(x - i) + np.abs(z - k)))
    return out


def run(lay, phi, u, v, w, lh=LH, ah=AH):
    return orm.remesh(lay.bases, phi, u, v, w, lay.fine_ny, lay.band, lay.max_delta, lay.h, lh, ah)


@pytest.mark.parametrize("seed,drop", [(0, 18), (1, 30), (2, -12)])
def test_bases_match_closed_form(seed, drop):
    lay, phi, u, v, w = moved_scene(seed=seed, drop=drop)
    nb = run(lay, phi, u, v, w)[0]
    dense = dense_phi(lay, phi) < 0
    ny, band = lay.fine_ny, lay.band
    below = np.concatenate([dense[:, :1], dense[:, :-1]], axis=1)
    above = np.concatenate([dense[:, 1:], np.zeros_like(dense[:, :1])], axis=1)
    surf = (dense != below) | (dense != above)
    rows = np.arange(ny)[None, :, None]
    lowest = np.where(surf, rows, ny).min(axis=1)
    highest = np.where(surf, rows, -1).max(axis=1)
    raw = tl.choose_base(lowest, highest, ny, band, min(LH, band), min(AH, band)).astype(np.int64)
    maximum_base = ny - band
    raw = np.where(highest >= 0, raw, maximum_base)
    want = closed_form_limit(raw, lay.max_delta)
    np.testing.assert_array_equal(nb, want)
    assert nb.min() >= 2 and nb.max() <= maximum_base
    assert np.abs(np.diff(nb, axis=0)).max() <= lay.max_delta
    assert np.abs(np.diff(nb, axis=1)).max() <= lay.max_delta
    assert (nb != lay.bases).any(), "the scene must move some bases"


def test_tall_samples_are_the_least_squares_line():
    lay, phi, u, v, w = moved_scene(seed=3, drop=25)
    nb, phi_o, *_ = run(lay, phi, u, v, w)
    dense = dense_phi(lay, phi)
    moved = nb != lay.bases
    for x, z in zip(*np.nonzero(moved)):
        m = int(nb[x, z])
        rows = np.arange(m)
        c1, c0 = np.polyfit(rows, dense[x, :m, z], 1)
        scale = np.abs(dense[x, :m, z]).max() + lay.h
        assert abs(phi_o[x, 0, z] - c0) <= 1e-5 * scale
        assert abs(phi_o[x, 1, z] - (c0 + c1 * (m - 1))) <= 1e-5 * scale
        # band samples: the old dense value at their fine row
        for py in range(2, lay.packed_ny):
            y = m + py - 2
            want = dense[x, y, z] if y < lay.fine_ny else 5 * lay.h
            assert phi_o[x, py, z] == np.float32(want)


def test_linear_columns_are_reproduced():
    lay, phi, u, v, w = moved_scene(seed=4, drop=20, linear=True)
    nb, phi_o, *_ = run(lay, phi, u, v, w)
    new = tl.TallCellLayout(nb, lay.fine_ny, lay.band, lay.h, lay.max_delta)
    rows = new.sample_rows()
    # the generating line, recovered from the old tall cell's two samples
    b = lay.bases[:, None, :].astype(np.float64)
    slope = (phi[:, 1, :].astype(np.float64) - phi[:, 0, :])[:, None, :] / (b - 1)
    want = phi[:, :1, :].astype(np.float64) + slope * rows
    # rows above the old band were air in the old layout (5 h)
    want = np.where(rows < b + lay.band, want, 5 * lay.h)
    act = new.active()
    np.testing.assert_allclose(phi_o[act], want[act], rtol=0, atol=2e-6)
    assert (phi_o[~act] == np.float32(5 * lay.h)).all()


def test_v_slot0_is_the_mean_of_the_replaced_faces():
    lay, phi, u, v, w = moved_scene(seed=5, drop=22)
    nb, _, _, v_o, _, _ = run(lay, phi, u, v, w)
    ny, P = lay.fine_ny, lay.packed_ny

    def face(col_v, b, f):
        if f <= 0 or f >= ny:
            return 0.0
        if f < b:
            return float(col_v[0])
        s = f - b + 1
        return float(col_v[s]) if s < P else 0.0

    for x in range(lay.nx):
        for z in range(lay.nz):
            bo, bn = int(lay.bases[x, z]), int(nb[x, z])
            if bo == bn:
                assert (v_o[x, :, z] == v[x, :, z]).all()
                continue
            old = [face(v[x, :, z], bo, f) for f in range(ny + 1)]
            assert v_o[x, 0, z] == pytest.approx(np.mean(old[1:bn]), rel=1e-6, abs=1e-7)
            for s in range(1, P):
                assert v_o[x, s, z] == np.float32(old[bn + s - 1]) if bn + s - 1 <= ny else v_o[x, s, z] == 0


@pytest.mark.parametrize("axis", [0, 1])
def test_affine_faces_are_reproduced(axis):
    lay, phi, u, v, w = moved_scene(seed=6, drop=26)
    b = lay.bases.astype(np.int64)
    ny, B, S = lay.fine_ny, lay.band, lay.slots
    rng = np.random.default_rng(7)
    shape = u.shape if axis == 0 else w.shape
    alpha = rng.uniform(-1, 1, size=(shape[0], shape[2]))
    beta = rng.uniform(-0.01, 0.01, size=(shape[0], shape[2]))

    def pair(bases, i, k):
        return (bases[i - 1, k], bases[i, k]) if axis == 0 else (bases[i, k - 1], bases[i, k])

    def slots(bases, old=None):
        """Slots of the affine profile; with `old`, rows above the old
        layout's band top hold 0 (the old layout did not represent them)."""
        a = np.zeros(shape, np.float32)
        for i in range(shape[0]):
            for k in range(shape[2]):
                if (axis == 0 and (i == 0 or i == lay.nx)) or (axis == 1 and (k == 0 or k == lay.nz)):
                    continue
                ba, bb = pair(bases, i, k)
                m, top = min(ba, bb), min(ny, max(ba, bb) + B)
                if old is not None:
                    top = min(top, ny, max(pair(old, i, k)) + B)
                a[i, 0, k] = alpha[i, k] + beta[i, k] * 0.5 * (m - 1)
                a[i, 1, k] = beta[i, k]
                for s in range(2, S):
                    y = m + s - 2
                    a[i, s, k] = alpha[i, k] + beta[i, k] * y if y < top else 0.0
        return a

    a = slots(b)
    args = (a, v, w) if axis == 0 else (u, v, a)
    nb, _, u_o, _, w_o, _ = run(lay, phi, *args)
    got = u_o if axis == 0 else w_o
    np.testing.assert_allclose(got, slots(nb.astype(np.int64), b), rtol=0, atol=3e-6)


def test_second_remesh_changes_nothing():
    lay, phi, u, v, w = moved_scene(seed=8, drop=28)
    nb, phi1, u1, v1, w1, moved = run(lay, phi, u, v, w)
    assert moved > 0
    lay2 = tl.TallCellLayout(nb, lay.fine_ny, lay.band, lay.h, lay.max_delta)
    nb2, phi2, u2, v2, w2, moved2 = run(lay2, phi1, u1, v1, w1)
    assert moved2 == 0
    np.testing.assert_array_equal(nb2, nb)
    for a, b in ((phi1, phi2), (u1, u2), (v1, v2), (w1, w2)):
        np.testing.assert_array_equal(a, b)


def test_dry_and_flooded_columns():
    """A dry column takes the highest base (or 2 under a 3-row limit); a
    column wet to the lid has no surface below it either."""
    lay, phi, u, v, w = moved_scene(nx=6, nz=5, seed=9)
    h = lay.h
    phi = phi.copy()
    phi[0, :, 0] = 5 * h          # dry column
    nb = run(lay, phi, u, v, w)[0]
    raw = orm.raw_bases(phi, lay.bases.astype(np.int64), lay.fine_ny, lay.band, h)
    assert raw[0, 0] == lay.fine_ny - lay.band
    np.testing.assert_array_equal(nb, closed_form_limit(raw, lay.max_delta))
