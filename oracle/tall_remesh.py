"""CPU restatement of the tall-cell remesh (fp_tall_remesh), numpy, FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
remesh; never the measured or shipped path.  Every sum runs in the device's
order (fine rows bottom to top) with the same FP64 operations and no
contraction, so the results are bit-identical to the library's.

What it restates:
  * the phi of fine row y of a column with base b: linear between the tall
    bottom (row 0) and top (row b-1) samples inside the tall cell, the band
    sample above it, 5 h beyond the band -- representedInitialPhiVolumeCells,
    lib/core/tall-cell-grid.ts:220-238;
  * surface rows: a row whose wetness (phi < 0) differs from the row below
    (row 0: itself) or above (the lid: dry) -- isInitialSurfaceCell,
    tall-cell-grid.ts:263-279;
  * the base: chooseTallCellBase (tall-cell-grid.ts:61-82) with halos
    min(halo, band) (:327), the dry-column rule (:378-386), then
    limitNeighboringTallCellBases (:84-106) until no base changes (the layout
    builder's nx + nz passes reach the same fixed point, :395);
  * the transfer of docs/papers/tallCells.txt:240-252: band samples copy the
    old value at their fine row, a tall cell's bottom/top samples are the
    least-squares line through its rows; v slot 0 the mean of the faces
    inside the tall cell; u/w slots 0/1 the least-squares affine fit over the
    rows shared by both columns' tall cells, slot 2+j the old value at row
    m+j below the higher band top.  Unchanged columns (faces) are copied.

Layout (z fastest): bases (nx, nz); phi, v (nx, P, nz), P = band + 2;
u (nx+1, S, nz), w (nx, S, nz+1), S = P + max_delta.
"""
from __future__ import annotations

import numpy as np

F64 = np.float64


def column_phi(phi, b, y, air):
    """phi of fine row y (array per column) of columns with bases b; phi (nx, P, nz)."""
    P = phi.shape[1]
    bottom = phi[:, 0, :].astype(F64)
    top = phi[:, 1, :].astype(F64)
    y = np.broadcast_to(np.asarray(y, np.int64), b.shape)
    t = np.minimum(1.0, np.maximum(0.0, y.astype(F64) / np.maximum(b - 1, 1).astype(F64)))
    inside = bottom + (top - bottom) * t
    py = 2 + y - b
    ok = (py >= 0) & (py < P)
    band = np.take_along_axis(phi, np.clip(py, 0, P - 1)[:, None, :], axis=1)[:, 0, :].astype(F64)
    return np.where(y < b, inside, np.where(ok, band, F64(air)))


def column_v(v, b, f, ny):
    """v at fine face f (between rows f-1 and f) of columns with bases b."""
    P = v.shape[1]
    f = np.broadcast_to(np.asarray(f, np.int64), b.shape)
    s = f - b + 1
    slot = np.take_along_axis(v, np.clip(s, 0, P - 1)[:, None, :], axis=1)[:, 0, :].astype(F64)
    tall = v[:, 0, :].astype(F64)
    val = np.where(f < b, tall, np.where(s < P, slot, 0.0))
    return np.where((f <= 0) | (f >= ny), 0.0, val)


def face_value(a, m, y, S):
    """Slots a (nfx, S, nfz) of faces whose affine part covers rows 0..m-1."""
    y = np.broadcast_to(np.asarray(y, np.int64), m.shape)
    mean = a[:, 0, :].astype(F64)
    slope = a[:, 1, :].astype(F64)
    aff = mean + slope * (y.astype(F64) - 0.5 * (m - 1).astype(F64))
    s = 2 + y - m
    slot = np.take_along_axis(a, np.clip(s, 0, S - 1)[:, None, :], axis=1)[:, 0, :].astype(F64)
    return np.where(y < m, aff, np.where(s < S, slot, 0.0))


def js_round_half(lower, upper):
    """Math.round((lower + upper) / 2) on integers."""
    return np.floor((lower + upper).astype(F64) / 2.0 + 0.5).astype(np.int64)


def raw_bases(phi, bases, ny, band, h, liquid_halo=16, air_halo=8):
    """Surface-chosen base per column before the neighbour limit."""
    b = np.asarray(bases, np.int64)
    air = 5.0 * h
    lowest = np.full(b.shape, ny, np.int64)
    highest = np.full(b.shape, -1, np.int64)
    wet = [column_phi(phi, b, y, air) < 0.0 for y in range(ny)]
    for y in range(ny):
        under = wet[y] if y == 0 else wet[y - 1]
        above = wet[y + 1] if y + 1 < ny else np.zeros(b.shape, bool)
        surf = (wet[y] != under) | (wet[y] != above)
        lowest = np.where(surf, np.minimum(lowest, y), lowest)
        highest = np.where(surf, np.maximum(highest, y), highest)
    maximum_base = max(0, ny - band)
    lh, ah = min(liquid_halo, band), min(air_halo, band)
    lower = highest + 1 + ah - band
    upper = lowest + 1 - lh
    desired = np.where(lower <= upper, js_round_half(lower, upper), lower)
    chosen = np.maximum(0, np.minimum(maximum_base, desired))
    chosen = np.maximum(2, chosen) if maximum_base >= 2 else np.zeros_like(chosen)
    dry = min(2, maximum_base) if maximum_base <= 3 else maximum_base
    surface = np.where(highest >= 0, chosen, dry)
    return np.maximum(2, surface) if maximum_base >= 2 else np.zeros_like(surface)


def limit(raw, max_delta):
    """Jacobi passes of base <= min(4-neighbour) + max_delta to the fixed point."""
    cur = np.array(raw, np.int64)
    while True:
        up = cur.copy()
        up[1:, :] = np.minimum(up[1:, :], cur[:-1, :] + max_delta)
        up[:-1, :] = np.minimum(up[:-1, :], cur[1:, :] + max_delta)
        up[:, 1:] = np.minimum(up[:, 1:], cur[:, :-1] + max_delta)
        up[:, :-1] = np.minimum(up[:, :-1], cur[:, 1:] + max_delta)
        if np.array_equal(up, cur):
            return cur
        cur = up


def _columns(phi, v, ob, nb, ny, air):
    nx, P, nz = phi.shape
    phi_o = phi.copy()
    v_o = v.copy()
    moved = ob != nb
    mid = 0.5 * (nb - 1).astype(F64)
    sf = np.zeros(nb.shape)
    stf = np.zeros(nb.shape)
    stt = np.zeros(nb.shape)
    for y in range(int(nb.max())):
        use = y < nb
        f = column_phi(phi, ob, y, air)
        t = F64(y) - mid
        sf = np.where(use, sf + f, sf)
        stf = np.where(use, stf + t * f, stf)
        stt = np.where(use, stt + t * t, stt)
    mean = sf / nb.astype(F64)
    slope = np.where(stt > 0.0, stf / np.where(stt > 0.0, stt, 1.0), 0.0)
    new = np.empty_like(phi)
    new[:, 0, :] = (mean - slope * mid).astype(np.float32)
    new[:, 1, :] = (mean + slope * mid).astype(np.float32)
    for py in range(2, P):
        y = nb + py - 2
        new[:, py, :] = np.where(y < ny, column_phi(phi, ob, y, air), F64(air)).astype(np.float32)
    sv = np.zeros(nb.shape)
    for f in range(1, int(nb.max())):
        sv = np.where(f < nb, sv + column_v(v, ob, f, ny), sv)
    nv = np.empty_like(v)
    nv[:, 0, :] = (sv / np.maximum(nb - 1, 1).astype(F64)).astype(np.float32)
    for s in range(1, P):
        nv[:, s, :] = column_v(v, ob, nb + s - 1, ny).astype(np.float32)
    phi_o[:, :, :] = np.where(moved[:, None, :], new, phi)
    v_o[:, :, :] = np.where(moved[:, None, :], nv, v)
    return phi_o, v_o, int(moved.sum())


def _faces(a, axis, ob, nb, ny, band):
    nfx, S, nfz = a.shape
    nx, nz = ob.shape
    x = np.arange(nfx)[:, None]
    z = np.arange(nfz)[None, :]
    if axis == 0:
        wall = (x == 0) | (x == nx)
        xa, za, xn, zn = np.clip(x - 1, 0, nx - 1), z, np.clip(x, 0, nx - 1), z
    else:
        wall = (z == 0) | (z == nz)
        xa, za, xn, zn = x, np.clip(z - 1, 0, nz - 1), x, np.clip(z, 0, nz - 1)
    wall = np.broadcast_to(wall, (nfx, nfz))
    oa, on = ob[xa, za], ob[xn, zn]
    na, nn = nb[xa, za], nb[xn, zn]
    keep = wall | ((oa == na) & (on == nn))
    mo = np.minimum(oa, on)
    mn = np.minimum(na, nn)
    top = np.minimum(ny, np.maximum(na, nn) + band)
    mid = 0.5 * (mn - 1).astype(F64)
    sf = np.zeros(mn.shape)
    stf = np.zeros(mn.shape)
    stt = np.zeros(mn.shape)
    for y in range(int(mn.max())):
        use = y < mn
        f = face_value(a, mo, y, S)
        t = F64(y) - mid
        sf = np.where(use, sf + f, sf)
        stf = np.where(use, stf + t * f, stf)
        stt = np.where(use, stt + t * t, stt)
    new = np.empty_like(a)
    new[:, 0, :] = (sf / mn.astype(F64)).astype(np.float32)
    new[:, 1, :] = np.where(stt > 0.0, stf / np.where(stt > 0.0, stt, 1.0), 0.0).astype(np.float32)
    for s in range(2, S):
        y = mn + s - 2
        new[:, s, :] = np.where(y < top, face_value(a, mo, y, S), 0.0).astype(np.float32)
    return np.where(keep[:, None, :], a, new)


def remesh(bases, phi, u, v, w, ny, band, max_delta, h, liquid_halo=16, air_halo=8):
    """(new bases, phi, u, v, w, changed columns) of one remesh."""
    ob = np.asarray(bases, np.int64)
    phi = np.asarray(phi, np.float32)
    nb = limit(raw_bases(phi, ob, ny, band, h, liquid_halo, air_halo), max_delta)
    air = 5.0 * h
    phi_o, v_o, moved = _columns(phi, np.asarray(v, np.float32), ob, nb, ny, air)
    u_o = _faces(np.asarray(u, np.float32), 0, ob, nb, ny, band)
    w_o = _faces(np.asarray(w, np.float32), 1, ob, nb, ny, band)
    return nb.astype(np.int32), phi_o, u_o, v_o, w_o, moved
