"""ctypes wrapper of the FP64 CPU oracle (oracle/fp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  It is the checker, never the measured or
shipped path.  Arrays are numpy, C order, z fastest:
cells (nx, ny, nz); u (nx+1, ny, nz); v (nx, ny+1, nz); w (nx, ny, nz+1).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfp_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class SolveReport(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("components", C.c_int32),
        ("unanchored", C.c_int32),
        ("rhs_norm", C.c_double),
        ("relative_residual", C.c_double),
        ("max_residual", C.c_double),
    ]


class MGOptions(C.Structure):
    _fields_ = [("pre_sweeps", C.c_int32), ("post_sweeps", C.c_int32), ("phi_preserve_levels", C.c_int32)]


class CM11aOptions(C.Structure):
    _fields_ = [("full_cycles", C.c_int32), ("v_cycles", C.c_int32), ("pre", C.c_int32), ("post", C.c_int32),
                ("constraint_levels", C.c_int32), ("preserve", C.c_int32), ("coarse_cap", C.c_int32),
                ("coarse_tol", C.c_float), ("open_top", C.c_int32)]


class CM11aReport(C.Structure):
    _fields_ = [("levels", C.c_int32), ("coarse_sweeps", C.c_int32), ("coarse_failed", C.c_int32),
                ("coarse_solves", C.c_int32), ("lcp_residual", C.c_float), ("coarse_residual", C.c_float)]


class TallLayout(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("band", C.c_int32),
                ("max_delta", C.c_int32), ("bases", C.c_void_p), ("phi", C.c_void_p), ("h", C.c_double)]


class ExtrapReport(C.Structure):
    _fields_ = [("passes", C.c_int32), ("active_left", C.c_int32), ("levels", C.c_int32), ("reserved", C.c_int32),
                ("band_known", C.c_int64), ("unknown", C.c_int64)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i = C.c_int
        d = C.c_double
        L.orc_ghost_theta.restype = d
        L.orc_ghost_theta.argtypes = [d, d, d]
        L.orc_dot_kahan.restype = d
        L.orc_dot_kahan.argtypes = [_dp, _dp, C.c_size_t]
        L.orc_dot_neumaier.restype = d
        L.orc_dot_neumaier.argtypes = [_dp, _dp, C.c_size_t]
        L.orc_plan_levels.restype = i
        L.orc_set_threads.restype = i
        L.orc_set_threads.argtypes = [i]
        L.orc_taps1d.restype = i
        L.orc_taps1d.argtypes = [i, i, i, C.c_void_p, C.c_void_p]
        L.orc_plan_levels.argtypes = [i, i, i, _ip, _ip, C.POINTER(i)]
        L.orc_plan_levels_reference.restype = i
        L.orc_plan_levels_reference.argtypes = [i, i, i, _ip]
        L.orc_face_weights.argtypes = [i, i, i, _fp, _dp, _dp, _dp]
        L.orc_rhs.argtypes = [i, i, i, d, d, d, _fp, _fp, _fp, _fp, C.c_void_p, _dp]
        L.orc_apply.argtypes = [i, i, i, _fp, _dp, _dp, _dp, _dp, _dp]
        L.orc_diagonal.argtypes = [i, i, i, _fp, _dp, _dp, _dp, _dp]
        L.orc_components.restype = i
        L.orc_components.argtypes = [i, i, i, _fp, _dp, _dp, _dp, _ip, _u8]
        L.orc_project.argtypes = [i, i, i, d, d, d, _fp, _dp, _dp, _dp, _dp, _fp, _fp, _fp, _dp, _dp, _dp]
        L.orc_max_divergence.restype = d
        L.orc_max_divergence.argtypes = [i, i, i, d, _fp, _dp, _dp, _dp, C.c_void_p]
        L.orc_jacobi_pcg.restype = i
        L.orc_jacobi_pcg.argtypes = [i, i, i, _fp, _dp, d, d, i, _dp, C.POINTER(SolveReport)]
        L.orc_mgpcg.restype = i
        L.orc_mgpcg.argtypes = [i, i, i, _fp, _dp, C.POINTER(MGOptions), d, i, _dp, C.POINTER(SolveReport)]
        L.orc_vcycle_apply.restype = i
        L.orc_vcycle_apply.argtypes = [i, i, i, _fp, C.POINTER(MGOptions), _dp, _dp]
        L.orc_level_phi.restype = i
        L.orc_level_phi.argtypes = [i, i, i, _fp, i, i, _dp]
        L.orc_dense_pcg.restype = i
        L.orc_dense_pcg.argtypes = [i, _dp, _dp, C.c_void_p, i, i, d, d, i, _dp,
                                    C.POINTER(i), C.POINTER(i), C.POINTER(d)]
        L.orc_tall_choose_base.restype = i
        L.orc_tall_choose_base.argtypes = [i, i, i, i, i, i]
        L.orc_tall_limit_bases.argtypes = [_ip, i, i, i, i]
        tp = C.POINTER(TallLayout)
        L.orc_tall_check.restype = i
        L.orc_tall_check.argtypes = [tp]
        L.orc_tall_unknowns.argtypes = [tp, _u8]
        L.orc_tall_apply.restype = i
        L.orc_tall_apply.argtypes = [tp, _dp, _dp]
        L.orc_tall_rhs.restype = i
        L.orc_tall_rhs.argtypes = [tp, _dp, _dp, _dp, d, d, _dp]
        L.orc_tall_subtract.restype = i
        L.orc_tall_subtract.argtypes = [tp, _dp, _dp, _dp, _dp, d, d, _dp, _dp, _dp]
        L.orc_tall_diagonal.restype = i
        L.orc_tall_diagonal.argtypes = [tp, _dp]
        L.orc_tall_solve.restype = i
        L.orc_tall_solve.argtypes = [tp, _dp, d, i, _dp, C.POINTER(i), C.POINTER(d)]
        L.orc_tall_max_divergence.restype = d
        L.orc_tall_max_divergence.argtypes = [tp, _dp, _dp, _dp]
        L.orc_tall_coo.restype = C.c_int64
        L.orc_tall_coo.argtypes = [tp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_cm11a_default_options.argtypes = [C.POINTER(CM11aOptions)]
        L.orc_cm11a_project.restype = i
        L.orc_cm11a_project.argtypes = [i, i, i, d, _fp, _fp, _fp, _fp, C.c_void_p, d, d, C.POINTER(CM11aOptions),
                                        _fp, _fp, _fp, _fp, C.POINTER(CM11aReport)]
        L.orc_cm11a_system.restype = i
        L.orc_cm11a_system.argtypes = [i, i, i, d, _fp, _fp, _fp, _fp, d, d, i, _fp, _fp, _fp]
        vp = C.c_void_p
        L.orc_advect_velocity.restype = i
        L.orc_advect_velocity.argtypes = [i, i, i, d, d, d, d, i, i, i, vp, vp, _fp, _fp, _fp, _fp, _fp, _fp,
                                          _fp, _fp, _fp]
        L.orc_extrapolate.restype = i
        L.orc_extrapolate.argtypes = [i, i, i, d, i, i, vp, vp, _fp, _fp, _fp, _fp, _fp, _fp, vp, vp, vp,
                                      C.POINTER(ExtrapReport)]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ghost_theta(liquid_phi: float, air_phi: float, eps: float = 1e-9) -> float:
    return lib().orc_ghost_theta(float(liquid_phi), float(air_phi), float(eps))


def plan_levels(nx, ny, nz):
    """(dims[L,3], strides[L,3], extended) or None when no plan exists."""
    dims = np.zeros(16 * 3, np.int32)
    strides = np.zeros(16 * 3, np.int32)
    ext = C.c_int(0)
    L = lib().orc_plan_levels(nx, ny, nz, dims, strides, C.byref(ext))
    if L < 0:
        return None
    return dims[: 3 * L].reshape(L, 3), strides[: 3 * L].reshape(L, 3), bool(ext.value)


def plan_levels_reference(nx, ny, nz):
    dims = np.zeros(16 * 3, np.int32)
    L = lib().orc_plan_levels_reference(nx, ny, nz, dims)
    if L < 0:
        return None
    return dims[: 3 * L].reshape(L, 3)


def face_weights(phi):
    nx, ny, nz = phi.shape
    wx = np.zeros((nx + 1, ny, nz))
    wy = np.zeros((nx, ny + 1, nz))
    wz = np.zeros((nx, ny, nz + 1))
    lib().orc_face_weights(nx, ny, nz, _f32(phi), wx, wy, wz)
    return wx, wy, wz


def rhs(phi, u, v, w, h, dt, rho, q=None):
    nx, ny, nz = phi.shape
    b = np.zeros((nx, ny, nz))
    qa = _f32(q) if q is not None else None
    lib().orc_rhs(nx, ny, nz, h, dt, rho, _f32(phi), _f32(u), _f32(v), _f32(w),
                  qa.ctypes.data if qa is not None else None, b)
    return b


def apply(phi, p, weights=None):
    nx, ny, nz = phi.shape
    wx, wy, wz = weights if weights is not None else face_weights(phi)
    out = np.zeros((nx, ny, nz))
    lib().orc_apply(nx, ny, nz, _f32(phi), wx, wy, wz, _f64(p), out)
    return out


def diagonal(phi, weights=None):
    nx, ny, nz = phi.shape
    wx, wy, wz = weights if weights is not None else face_weights(phi)
    out = np.zeros((nx, ny, nz))
    lib().orc_diagonal(nx, ny, nz, _f32(phi), wx, wy, wz, out)
    return out


def components(phi, weights=None):
    nx, ny, nz = phi.shape
    wx, wy, wz = weights if weights is not None else face_weights(phi)
    comp = np.zeros((nx, ny, nz), np.int32)
    anchored = np.zeros(nx * ny * nz + 1, np.uint8)
    n = lib().orc_components(nx, ny, nz, _f32(phi), wx, wy, wz, comp, anchored)
    return comp, anchored[:n].astype(bool)


def project(phi, p, u, v, w, h, dt, rho, weights=None):
    nx, ny, nz = phi.shape
    wx, wy, wz = weights if weights is not None else face_weights(phi)
    uo = np.zeros((nx + 1, ny, nz))
    vo = np.zeros((nx, ny + 1, nz))
    wo = np.zeros((nx, ny, nz + 1))
    lib().orc_project(nx, ny, nz, h, dt, rho, _f32(phi), wx, wy, wz, _f64(p),
                      _f32(u), _f32(v), _f32(w), uo, vo, wo)
    return uo, vo, wo


def max_divergence(phi, u, v, w, h, q=None):
    nx, ny, nz = phi.shape
    qa = _f32(q) if q is not None else None
    return lib().orc_max_divergence(nx, ny, nz, h, _f32(phi), _f64(u), _f64(v), _f64(w),
                                    qa.ctypes.data if qa is not None else None)


def _report(rep):
    return {f: getattr(rep, f) for f, _ in SolveReport._fields_}


def jacobi_pcg(phi, b, rel_tol=1e-10, abs_tol=1e-12, max_iter=None, p0=None):
    nx, ny, nz = phi.shape
    p = _f64(p0).copy() if p0 is not None else np.zeros((nx, ny, nz))
    rep = SolveReport()
    if max_iter is None:
        max_iter = max(64, 4 * nx * ny * nz)
    status = lib().orc_jacobi_pcg(nx, ny, nz, _f32(phi), _f64(b), rel_tol, abs_tol, int(max_iter), p,
                                  C.byref(rep))
    out = _report(rep)
    out["status"] = status
    return p, out


def mgpcg(phi, b, rel_tol=1e-10, max_iter=200, pre=2, post=2, preserve=2, p0=None):
    nx, ny, nz = phi.shape
    p = _f64(p0).copy() if p0 is not None else np.zeros((nx, ny, nz))
    rep = SolveReport()
    opt = MGOptions(pre, post, preserve)
    status = lib().orc_mgpcg(nx, ny, nz, _f32(phi), _f64(b), C.byref(opt), rel_tol, int(max_iter), p,
                             C.byref(rep))
    out = _report(rep)
    out["status"] = status
    return p, out


def vcycle(phi, r, pre=2, post=2, preserve=2):
    nx, ny, nz = phi.shape
    z = np.zeros((nx, ny, nz))
    opt = MGOptions(pre, post, preserve)
    if lib().orc_vcycle_apply(nx, ny, nz, _f32(phi), C.byref(opt), _f64(r), z) != 0:
        raise ValueError("no multigrid plan")
    return z


def set_threads(n: int = 0) -> int:
    """OpenMP threads of the oracle's stencil loops (n > 0 sets it); returns
    the count in effect."""
    return int(lib().orc_set_threads(int(n)))


def taps1d(nf, nc, stride):
    """Prolongation taps of every fine index of an axis: (coarse index (nf, 2),
    weight (nf, 2)), index -1 and weight 0 where a tap is unused."""
    L = lib()
    idx = np.full((nf, 2), -1, np.int32)
    wt = np.zeros((nf, 2))
    i2 = (C.c_int * 2)()
    w2 = (C.c_double * 2)()
    for f in range(nf):
        n = L.orc_taps1d(f, stride, nc, i2, w2)
        for q in range(n):
            idx[f, q] = i2[q]
            wt[f, q] = w2[q]
    return idx, wt


def level_phi(phi, level, preserve=2):
    nx, ny, nz = phi.shape
    plan = plan_levels(nx, ny, nz)
    dims = plan[0][level]
    out = np.zeros(tuple(int(x) for x in dims))
    if lib().orc_level_phi(nx, ny, nz, _f32(phi), preserve, level, out) != 0:
        raise ValueError("bad level")
    return out


def dense_pcg(A, b, x0=None, jacobi=True, max_iter=None, rel_tol=1e-10, abs_tol=0.0, validate=True):
    A = _f64(A)
    n = A.shape[0]
    x = np.zeros(n)
    it, red, rn = C.c_int(0), C.c_int(0), C.c_double(0)
    x0a = _f64(x0) if x0 is not None else None
    status = lib().orc_dense_pcg(n, A, _f64(b), x0a.ctypes.data if x0a is not None else None,
                                 int(jacobi), int(max_iter if max_iter is not None else n), rel_tol, abs_tol,
                                 int(validate), x, C.byref(it), C.byref(red), C.byref(rn))
    return x, {"status": status, "iterations": it.value, "reductions": red.value, "residual_norm": rn.value}


def tall_choose_base(lowest, highest, fine_ny, regular_layers=32, liquid_halo=16, air_halo=8):
    return lib().orc_tall_choose_base(lowest, highest, fine_ny, regular_layers, liquid_halo, air_halo)


def tall_limit_bases(bases, max_delta=4, passes=None):
    b = np.ascontiguousarray(bases, dtype=np.int32).copy()
    nx, nz = b.shape
    if passes is None:
        passes = nx + nz
    lib().orc_tall_limit_bases(b, nx, nz, max_delta, passes)
    return b


def uniform_projection(phi, u, v, w, h, dt, rho, rel_tol=1e-10, solver="mgpcg", q=None, p0=None):
    """Full FP64 projection: rhs -> solve -> gradient subtract -> diagnostics."""
    weights = face_weights(phi)
    b = rhs(phi, u, v, w, h, dt, rho, q)
    if solver == "mgpcg":
        p, rep = mgpcg(phi, b, rel_tol=rel_tol, p0=p0)
    else:
        p, rep = jacobi_pcg(phi, b, rel_tol=rel_tol, abs_tol=0.0, p0=p0)
    uo, vo, wo = project(phi, p, u, v, w, h, dt, rho, weights)
    rep["max_div_before"] = max_divergence(phi, u.astype(np.float64) * _wall_mask(u, 0),
                                           v.astype(np.float64) * _wall_mask(v, 1),
                                           w.astype(np.float64) * _wall_mask(w, 2), h, q)
    rep["max_div_after"] = max_divergence(phi, uo, vo, wo, h, q)
    return p, (uo, vo, wo), b, rep


def _wall_mask(a, axis):
    m = np.ones(a.shape)
    sl = [slice(None)] * 3
    sl[axis] = 0
    m[tuple(sl)] = 0
    sl[axis] = -1
    m[tuple(sl)] = 0
    return m


# --------------------------------------------------------------------------
# Restricted tall-cell projection.  Samples (nx, B+2, nz): py 0 tall bottom
# (fine row 0), py 1 tall top (row base-1), py 2+k band row base+k.  Faces:
# v (nx, B+2, nz), u (nx+1, S, nz), w (nx, S, nz+1), S = B+2+max_delta.
# --------------------------------------------------------------------------
class Tall:
    """A tall-cell layout bound for the oracle calls (keeps its arrays alive)."""

    def __init__(self, bases, phi, ny, band, h, max_delta=4):
        self.bases = np.ascontiguousarray(bases, dtype=np.int32)
        self.phi = np.ascontiguousarray(phi, dtype=np.float32)
        nx, nz = self.bases.shape
        assert self.phi.shape == (nx, band + 2, nz), self.phi.shape
        self.nx, self.ny, self.nz, self.band, self.max_delta, self.h = nx, int(ny), nz, int(band), int(max_delta), h
        self.S = self.band + 2 + self.max_delta
        self.P = self.band + 2
        self._c = TallLayout(nx, self.ny, nz, self.band, self.max_delta, self.bases.ctypes.data,
                             self.phi.ctypes.data, float(h))

    @property
    def ref(self):
        return C.byref(self._c)

    def shapes(self):
        return (self.nx + 1, self.S, self.nz), (self.nx, self.P, self.nz), (self.nx, self.S, self.nz + 1)

    def check(self):
        return lib().orc_tall_check(self.ref)

    def unknowns(self):
        m = np.zeros((self.nx, self.P, self.nz), np.uint8)
        lib().orc_tall_unknowns(self.ref, m)
        return m.astype(bool)

    def _rc(self, rc):
        if rc != 0:
            raise ValueError(f"tall-cell oracle error {rc}")

    def apply(self, p):
        out = np.zeros((self.nx, self.P, self.nz))
        self._rc(lib().orc_tall_apply(self.ref, _f64(p), out))
        return out

    def diagonal(self):
        out = np.zeros((self.nx, self.P, self.nz))
        self._rc(lib().orc_tall_diagonal(self.ref, out))
        return out

    def rhs(self, u, v, w, dt, rho):
        out = np.zeros((self.nx, self.P, self.nz))
        self._rc(lib().orc_tall_rhs(self.ref, _f64(u), _f64(v), _f64(w), dt, rho, out))
        return out

    def subtract(self, p, u, v, w, dt, rho):
        su, sv, sw = self.shapes()
        uo, vo, wo = np.zeros(su), np.zeros(sv), np.zeros(sw)
        self._rc(lib().orc_tall_subtract(self.ref, _f64(p), _f64(u), _f64(v), _f64(w), dt, rho, uo, vo, wo))
        return uo, vo, wo

    def solve(self, b, rel_tol=1e-12, max_iter=20000):
        p = np.zeros((self.nx, self.P, self.nz))
        it, rel = C.c_int(0), C.c_double(0)
        rc = lib().orc_tall_solve(self.ref, _f64(b), rel_tol, int(max_iter), p, C.byref(it), C.byref(rel))
        return p, {"status": rc, "iterations": it.value, "relative_residual": rel.value}

    def max_divergence(self, u, v, w):
        return lib().orc_tall_max_divergence(self.ref, _f64(u), _f64(v), _f64(w))

    def matrix(self):
        """The operator as a scipy CSR matrix over all samples (zero rows and
        columns off the unknowns)."""
        import scipy.sparse as sp
        cnt = lib().orc_tall_coo(self.ref, None, None, None, 0)
        if cnt < 0:
            raise ValueError(f"tall-cell oracle error {-cnt}")
        r = np.zeros(cnt, np.int64)
        c = np.zeros(cnt, np.int64)
        v = np.zeros(cnt)
        lib().orc_tall_coo(self.ref, r.ctypes.data, c.ctypes.data, v.ctypes.data, cnt)
        n = self.nx * self.P * self.nz
        return sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()

    def projection(self, u, v, w, dt, rho, rel_tol=1e-12):
        """FP64 projection: rhs -> Jacobi-PCG -> subtract -> divergence."""
        b = self.rhs(u, v, w, dt, rho)
        p, rep = self.solve(b, rel_tol)
        uo, vo, wo = self.subtract(p, u, v, w, dt, rho)
        rep["max_div_before"] = self.max_divergence(u, v, w)
        rep["max_div_after"] = self.max_divergence(uo, vo, wo)
        return p, (uo, vo, wo), b, rep


def cm11a_options(**kw):
    """The reference's CM11a schedule (webgpu-uniform-pressure-multigrid.ts:5-19),
    with overrides."""
    o = CM11aOptions()
    lib().orc_cm11a_default_options(C.byref(o))
    for k, val in kw.items():
        setattr(o, k, val)
    return o


def cm11a_project(phi, u, v, w, h, dt, rho, q=None, **opts):
    """CM11a separating-wall projection in FP32 (fp_oracle_cm11a.c): the
    halo-padded finest pressure (nx+2, ny+2, nz+2), the projected
    velocities and the report."""
    phi = _f32(phi)
    u, v, w = _f32(u), _f32(v), _f32(w)
    nx, ny, nz = phi.shape
    qq = _f32(q) if q is not None else None
    p = np.zeros((nx + 2, ny + 2, nz + 2), np.float32)
    uo, vo, wo = np.zeros_like(u), np.zeros_like(v), np.zeros_like(w)
    rep = CM11aReport()
    o = cm11a_options(**opts)
    rc = lib().orc_cm11a_project(nx, ny, nz, h, phi, u, v, w, qq.ctypes.data if qq is not None else None, dt, rho,
                                 C.byref(o), p, uo, vo, wo, C.byref(rep))
    if rc <= 0:
        raise RuntimeError(f"orc_cm11a_project failed ({rc})")
    return p, (uo, vo, wo), {k: getattr(rep, k) for k, _ in CM11aReport._fields_}


def cm11a_system(phi, u, v, w, h, dt, rho, open_top=0):
    """The finest level's assembled CM11a LCP: coefficients (padded, 4 per
    cell: +x, +y, +z, liquid flag), b and p_min."""
    phi = _f32(phi)
    u, v, w = _f32(u), _f32(v), _f32(w)
    nx, ny, nz = phi.shape
    D = (nx + 2, ny + 2, nz + 2)
    coef = np.zeros(D + (4,), np.float32)
    b = np.zeros(D, np.float32)
    pmin = np.zeros(D, np.float32)
    lib().orc_cm11a_system(nx, ny, nz, h, phi, u, v, w, dt, rho, open_top, coef, b, pmin)
    return coef, b, pmin


def extrapolate(u, v, w, h, phi=None, density=None, open_top=0, max_passes=0, band=False):
    """Velocity extrapolation of the uniform solver (fp_oracle_extrap.c, FP32):
    returns (u, v, w, report dict) and, with band=True, the narrow-band state
    (dist (3, nx, ny, nz), value (3, nx, ny, nz), known bits (nx, ny, nz))."""
    field = phi if phi is not None else density
    nx, ny, nz = field.shape
    f = _f32(field)
    u, v, w = _f32(u), _f32(v), _f32(w)
    uo, vo, wo = np.zeros_like(u), np.zeros_like(v), np.zeros_like(w)
    bd = bv = bk = None
    if band:
        bd = np.zeros((3, nx, ny, nz), np.float32)
        bv = np.zeros((3, nx, ny, nz), np.float32)
        bk = np.zeros((nx, ny, nz), np.uint8)
    rep = ExtrapReport()
    rc = lib().orc_extrapolate(nx, ny, nz, h, int(open_top), int(max_passes),
                               f.ctypes.data if phi is not None else None,
                               f.ctypes.data if phi is None else None, u, v, w, uo, vo, wo,
                               bd.ctypes.data if band else None, bv.ctypes.data if band else None,
                               bk.ctypes.data if band else None, C.byref(rep))
    if rc != 0:
        raise RuntimeError(f"orc_extrapolate failed ({rc})")
    r = {k: getattr(rep, k) for k, _ in ExtrapReport._fields_ if k != "reserved"}
    if band:
        return uo, vo, wo, r, (bd, bv, bk)
    return uo, vo, wo, r


def advect_velocity(u, v, w, h, dt, phi=None, density=None, transport=None, gravity=-9.81, nu=0.0,
                    scheme="semi-lagrangian", open_top=0, no_slip=0):
    """Velocity advection and body forces of the uniform solver
    (fp_oracle_advect.c, FP32).  `transport` is the extrapolated (u, v, w); by
    default the oracle's own extrapolation of (u, v, w)."""
    field = phi if phi is not None else density
    nx, ny, nz = field.shape
    f = _f32(field)
    u, v, w = _f32(u), _f32(v), _f32(w)
    if transport is None:
        transport = extrapolate(u, v, w, h, phi=phi, density=density, open_top=open_top)[:3]
    tu, tv, tw = (_f32(a) for a in transport)
    uo, vo, wo = np.zeros_like(u), np.zeros_like(v), np.zeros_like(w)
    sch = {"semi-lagrangian": 0, "maccormack": 1}[scheme]
    rc = lib().orc_advect_velocity(nx, ny, nz, h, dt, gravity, nu, sch, int(open_top), int(no_slip),
                                   f.ctypes.data if phi is not None else None,
                                   f.ctypes.data if phi is None else None, u, v, w, tu, tv, tw, uo, vo, wo)
    if rc != 0:
        raise RuntimeError(f"orc_advect_velocity failed ({rc})")
    return uo, vo, wo
