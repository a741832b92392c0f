"""Brute-force tall-cell operators built from the expanded fine grid (tests only).

Independent of the oracle's face walk: the fine grid is materialised, the
prolongation E (fine cells x samples) is a sparse matrix, D is the interior
face-difference matrix of the fine grid and W the ghost-fluid face weights of
the expanded phi, so A = E^T D^T W D E, b = (rho h/dt) E^T D^T u_fine and the
subtracted fine velocities are projected back onto the slots by least
squares.  Sizes are small (a few thousand fine cells).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

THETA_EPS = 1e-9


def face_weight(pa, pb):
    pa = np.asarray(pa, np.float64)
    pb = np.asarray(pb, np.float64)
    la, lb = pa < 0, pb < 0
    liq = np.where(la, pa, pb)
    air = np.where(la, pb, pa)
    theta = np.clip(np.abs(liq) / np.maximum(np.abs(liq) + np.abs(air), THETA_EPS), 0.05, 1.0)
    return np.where(la & lb, 1.0, np.where(la | lb, 1.0 / theta, 0.0))


class Expanded:
    def __init__(self, layout, phi):
        self.L = layout
        nx, nz, ny, B, P = layout.nx, layout.nz, layout.fine_ny, layout.band, layout.packed_ny
        self.shape = (nx, ny, nz)
        b = np.asarray(layout.bases, np.int64)
        rows, cols, vals = [], [], []
        phif = np.full(self.shape, 5.0 * layout.h)
        phi = np.asarray(phi, np.float32)

        def sidx(x, py, z):
            return (x * P + py) * nz + z

        for x in range(nx):
            for z in range(nz):
                bb = int(b[x, z])
                pb, pt = float(phi[x, 0, z]), float(phi[x, 1, z])
                for y in range(ny):
                    f = (x * ny + y) * nz + z
                    if y < bb:
                        k = y / (bb - 1)
                        if y == 0:
                            rows.append(f); cols.append(sidx(x, 0, z)); vals.append(1.0)
                            phif[x, y, z] = pb
                        elif y == bb - 1:
                            rows.append(f); cols.append(sidx(x, 1, z)); vals.append(1.0)
                            phif[x, y, z] = pt
                        else:
                            rows += [f, f]; cols += [sidx(x, 0, z), sidx(x, 1, z)]; vals += [1.0 - k, k]
                            phif[x, y, z] = pb + (pt - pb) * k
                    elif y - bb < B:
                        rows.append(f); cols.append(sidx(x, y - bb + 2, z)); vals.append(1.0)
                        phif[x, y, z] = float(phi[x, y - bb + 2, z])
        self.E = sp.csr_matrix((vals, (rows, cols)), shape=(nx * ny * nz, nx * P * nz))
        self.phi_fine = phif
        # unknown samples: liquid and active
        self.unknown = ((phi < 0) & layout.active()).ravel()
        # interior faces of the fine grid, in the order x, y, z
        fa, fb, self.faces = [], [], []
        idx = np.arange(nx * ny * nz).reshape(self.shape)
        for ax in range(3):
            sl_a = [slice(None)] * 3
            sl_b = [slice(None)] * 3
            sl_a[ax] = slice(0, -1)
            sl_b[ax] = slice(1, None)
            a = idx[tuple(sl_a)].ravel()
            bidx = idx[tuple(sl_b)].ravel()
            fa.append(a)
            fb.append(bidx)
            self.faces.append(idx[tuple(sl_a)].shape)
        self.fa = np.concatenate(fa)
        self.fb = np.concatenate(fb)
        nf = self.fa.size
        self.D = sp.csr_matrix((np.concatenate([-np.ones(nf), np.ones(nf)]),
                                (np.concatenate([np.arange(nf)] * 2), np.concatenate([self.fa, self.fb]))),
                               shape=(nf, nx * ny * nz))
        pf = phif.ravel()
        self.w = face_weight(pf[self.fa], pf[self.fb])
        M = sp.diags(self.unknown.astype(np.float64))
        self.G = self.D @ self.E @ M          # face gradients of (masked) sample pressures
        self.A = (self.G.T @ sp.diags(self.w) @ self.G).tocsr()

    # ---- velocities ---------------------------------------------------------
    def expand_velocity(self, u, v, w):
        """Fine interior-face velocities (D's face order) of slot fields."""
        L = self.L
        nx, ny, nz = self.shape
        B = L.band
        b = np.asarray(L.bases, np.int64)
        out = []
        # x faces (nx-1, ny, nz)
        fx = np.zeros((nx - 1, ny, nz))
        for x in range(nx - 1):
            for z in range(nz):
                m, M = min(b[x, z], b[x + 1, z]), max(b[x, z], b[x + 1, z])
                for y in range(min(ny, M + B)):
                    fx[x, y, z] = (u[x + 1, 0, z] + (y - 0.5 * (m - 1)) * u[x + 1, 1, z]) if y < m \
                        else u[x + 1, 2 + y - m, z]
        fy = np.zeros((nx, ny - 1, nz))
        for x in range(nx):
            for z in range(nz):
                bb = b[x, z]
                for y in range(min(ny - 1, bb + B)):
                    fy[x, y, z] = v[x, 0, z] if y < bb - 1 else v[x, y - bb + 2, z]
        fz = np.zeros((nx, ny, nz - 1))
        for x in range(nx):
            for z in range(nz - 1):
                m, M = min(b[x, z], b[x, z + 1]), max(b[x, z], b[x, z + 1])
                for y in range(min(ny, M + B)):
                    fz[x, y, z] = (w[x, 0, z + 1] + (y - 0.5 * (m - 1)) * w[x, 1, z + 1]) if y < m \
                        else w[x, 2 + y - m, z + 1]
        return np.concatenate([fx.ravel(), fy.ravel(), fz.ravel()])

    def project_slots(self, f):
        """Least-squares slot values of fine interior-face velocities."""
        L = self.L
        nx, ny, nz = self.shape
        B = L.band
        b = np.asarray(L.bases, np.int64)
        su, sv, sw = L.face_shapes()
        u, v, w = np.zeros(su), np.zeros(sv), np.zeros(sw)
        n0 = (nx - 1) * ny * nz
        n1 = n0 + nx * (ny - 1) * nz
        fx = f[:n0].reshape(nx - 1, ny, nz)
        fy = f[n0:n1].reshape(nx, ny - 1, nz)
        fz = f[n1:].reshape(nx, ny, nz - 1)
        for x in range(nx - 1):
            for z in range(nz):
                m, M = min(b[x, z], b[x + 1, z]), max(b[x, z], b[x + 1, z])
                t = np.arange(m) - 0.5 * (m - 1)
                u[x + 1, 0, z] = fx[x, :m, z].mean()
                u[x + 1, 1, z] = (t * fx[x, :m, z]).sum() / (t * t).sum()
                for y in range(m, min(ny, M + B)):
                    u[x + 1, 2 + y - m, z] = fx[x, y, z]
        for x in range(nx):
            for z in range(nz):
                bb = b[x, z]
                v[x, 0, z] = fy[x, :bb - 1, z].mean()
                for y in range(bb - 1, min(ny - 1, bb + B)):
                    v[x, y - bb + 2, z] = fy[x, y, z]
        for x in range(nx):
            for z in range(nz - 1):
                m, M = min(b[x, z], b[x, z + 1]), max(b[x, z], b[x, z + 1])
                t = np.arange(m) - 0.5 * (m - 1)
                w[x, 0, z + 1] = fz[x, :m, z].mean()
                w[x, 1, z + 1] = (t * fz[x, :m, z]).sum() / (t * t).sum()
                for y in range(m, min(ny, M + B)):
                    w[x, 2 + y - m, z + 1] = fz[x, y, z]
        return u, v, w

    # ---- operators ----------------------------------------------------------
    def apply(self, p):
        return (self.A @ np.asarray(p, np.float64).ravel()).reshape(self.L.bases.shape[0], -1, self.L.bases.shape[1])

    def rhs(self, u, v, w, dt, rho):
        s = rho * self.L.h / dt
        f = self.expand_velocity(u, v, w)
        b = s * (self.G.T @ f)
        return b.reshape(self.L.nx, self.L.packed_ny, self.L.nz)

    def subtract(self, p, u, v, w, dt, rho):
        s = dt / (rho * self.L.h)
        f = self.expand_velocity(u, v, w)
        g = self.G @ np.asarray(p, np.float64).ravel()
        fn = f - s * self.w * g
        # a slot is live when one of its faces has a liquid side or an unknown sample in its expansion
        live_face = ((self.w != 0.0) | (np.asarray(abs(self.G).sum(axis=1)).ravel() > 0)).astype(np.float64)
        lu, lv, lw = self.project_slots(live_face)
        lu[:, 1, :] = lu[:, 0, :]
        lw[:, 1, :] = lw[:, 0, :]
        return tuple(np.where(l > 0, a, 0.0) for a, l in zip(self.project_slots(fn), (lu, lv, lw)))
