"""Host-side mirror of the reference's projection interface over the C ABI.

``UniformPressureProjection`` follows the surface of the reference's
``WebGPUUniformPressureMultigrid`` (lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:228-607):
construction plans the hierarchy and raises ``ValueError`` (the reference's
``RangeError``, :251-256) for an unplannable lattice; ``project`` encodes one
projection (``encode`` :327-364 followed by the ``project`` kernel,
webgpu-uniform-reference.wgsl.ts:1052-1099); ``diagnostics`` holds the decoded
report (``readStats``, webgpu-uniform-reference.ts:1358-1421); ``destroy`` is
idempotent (:595-606).

The GPU library must be present: there is no CPU fallback.  Arrays may be
numpy arrays (host) or torch CUDA tensors (device, zero-copy).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libfluid_projection.so")

FP_OK = 0
FLAG_NAMES = {
    0x1: "NONFINITE",
    0x2: "NONPOSITIVE_CURVATURE",
    0x4: "NONPOSITIVE_PRECONDITIONER",
    0x8: "NONCONVERGENCE",
    0x10: "UNANCHORED_COMPONENT",
    0x20: "PRECISION_FALLBACK",
    0x40: "COARSE_SINGULAR",
}
LAYOUT_UNIFORM, LAYOUT_TALL_CELL = 0, 1
MEMORY_HOST, MEMORY_DEVICE = 0, 1
# fp_* return codes
(ERR_INVALID_ARGUMENT, ERR_DEVICE, ERR_OUT_OF_MEMORY, ERR_LAYOUT, ERR_BUDGET, ERR_STATE,
 ERR_UNSUPPORTED) = (-1, -2, -3, -4, -5, -6, -7)
PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2}
SOLVERS = {"mgpcg": 0, "cm11a": 1}
KERNELS = {"none": 0, "smooth_fine": 1, "residual_fine": 2, "cg_direction": 3, "cg_update": 4,
           "true_residual": 5, "prolong_fine": 6, "restrict_fine": 7}

EXPORTED_SYMBOLS = [
    "fp_last_error", "fp_abi_version", "fp_ctx_create", "fp_ctx_destroy", "fp_ctx_bytes_in_use",
    "fp_default_options", "fp_grid_bytes", "fp_grid_create", "fp_grid_destroy", "fp_grid_levels",
    "fp_grid_set_tuning",
    "fp_grid_face_counts", "fp_project", "fp_project_batch", "fp_project_noalloc", "fp_submit", "fp_poll",
    "fp_probe_apply",
    "fp_probe_rhs", "fp_probe_vcycle", "fp_probe_level_phi", "fp_probe_tall_layout", "fp_probe_taps",
    "fp_tall_remesh", "fp_extrapolate_velocity", "fp_advect_velocity",
]
ADVECT_SCHEMES = {"semi-lagrangian": 0, "maccormack": 1}
PENDING = 1


class GridDesc(C.Structure):
    _fields_ = [("layout", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("band", C.c_int32), ("max_delta", C.c_int32), ("h", C.c_double),
                ("column_bases", C.c_void_p)]


class Inputs(C.Structure):
    _fields_ = [("memory", C.c_int32), ("stream_order", C.c_int32), ("phi", C.c_void_p), ("u", C.c_void_p),
                ("v", C.c_void_p), ("w", C.c_void_p), ("target_divergence", C.c_void_p),
                ("dt", C.c_double), ("rho", C.c_double), ("stream", C.c_void_p), ("density", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("sweep_tile_z", C.c_int32), ("sweep_groups", C.c_int32), ("sweep_pairs", C.c_int32),
                ("reserved", C.c_int32)]


class TallRemeshArgs(C.Structure):
    _fields_ = [("memory", C.c_int32), ("liquid_halo", C.c_int32), ("air_halo", C.c_int32),
                ("reserved", C.c_int32), ("phi", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("w", C.c_void_p), ("phi_out", C.c_void_p), ("u_out", C.c_void_p), ("v_out", C.c_void_p),
                ("w_out", C.c_void_p), ("column_bases_out", C.c_void_p)]


class ExtrapolateArgs(C.Structure):
    _fields_ = [("memory", C.c_int32), ("open_top", C.c_int32), ("max_passes", C.c_int32), ("reserved", C.c_int32),
                ("phi", C.c_void_p), ("density", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("w", C.c_void_p), ("u_out", C.c_void_p), ("v_out", C.c_void_p), ("w_out", C.c_void_p)]


class ExtrapolateReport(C.Structure):
    _fields_ = [("passes", C.c_int32), ("active_left", C.c_int32), ("levels", C.c_int32), ("reserved", C.c_int32),
                ("band_faces", C.c_int64), ("unknown_faces", C.c_int64), ("device_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class AdvectArgs(C.Structure):
    _fields_ = [("memory", C.c_int32), ("scheme", C.c_int32), ("open_top", C.c_int32), ("no_slip", C.c_int32),
                ("dt", C.c_double), ("gravity", C.c_double), ("viscosity", C.c_double),
                ("phi", C.c_void_p), ("density", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("w", C.c_void_p), ("tu", C.c_void_p), ("tv", C.c_void_p), ("tw", C.c_void_p),
                ("u_out", C.c_void_p), ("v_out", C.c_void_p), ("w_out", C.c_void_p)]


class AdvectReport(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("reserved", C.c_int32 * 4)]


class Outputs(C.Structure):
    _fields_ = [("pressure", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("relative_tolerance", C.c_double), ("max_iterations", C.c_int32), ("max_outer", C.c_int32),
                ("precision", C.c_int32), ("pre_sweeps", C.c_int32), ("post_sweeps", C.c_int32),
                ("warm_start", C.c_int32), ("phi_preserve_levels", C.c_int32), ("check_every", C.c_int32),
                ("timed_kernel", C.c_int32), ("graphs", C.c_int32), ("precision_test_shift", C.c_int32),
                ("persistent", C.c_int32), ("solver", C.c_int32), ("open_top", C.c_int32),
                ("full_cycles", C.c_int32), ("v_cycles", C.c_int32), ("lcp_sweeps", C.c_int32)]


class Report(C.Structure):
    _fields_ = [("flags", C.c_int32), ("iterations", C.c_int32), ("outer_iterations", C.c_int32),
                ("levels", C.c_int32), ("relative_residual", C.c_double), ("rhs_norm", C.c_double),
                ("max_div_before", C.c_double), ("max_div_after", C.c_double), ("device_ms", C.c_double),
                ("liquid_samples", C.c_int64), ("samples", C.c_int64), ("timed_kernel_ms", C.c_double),
                ("timed_kernel_launches", C.c_int32), ("coarse_sweeps", C.c_int32),
                ("lcp_residual", C.c_double), ("coarse_failures", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["flag_names"] = [n for b, n in FLAG_NAMES.items() if self.flags & b]
        return d


class FluidProjectionError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
    # One HIP runtime per process: when PyTorch is present its libamdhip64
    # (SONAME libamdhip64.so.7) must be the one the library binds to, so torch
    # is loaded first and the dynamic loader reuses its runtime for ours.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    vp = C.c_void_p
    L.fp_last_error.restype = C.c_char_p
    L.fp_abi_version.restype = C.c_int
    L.fp_ctx_create.argtypes = [C.c_int32, C.c_uint64, C.POINTER(vp)]
    L.fp_ctx_destroy.argtypes = [vp]
    L.fp_ctx_bytes_in_use.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.fp_default_options.argtypes = [C.POINTER(Options)]
    L.fp_grid_bytes.argtypes = [C.POINTER(GridDesc), C.c_int32, C.POINTER(C.c_uint64)]
    L.fp_grid_create.argtypes = [vp, C.POINTER(GridDesc), C.POINTER(vp)]
    L.fp_grid_destroy.argtypes = [vp]
    L.fp_grid_levels.argtypes = [vp, C.c_int32, vp, vp, C.POINTER(C.c_int32)]
    L.fp_grid_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.fp_grid_face_counts.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.fp_project.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Outputs), C.POINTER(Options), C.POINTER(Report)]
    L.fp_project_batch.argtypes = [vp, C.c_int32, C.POINTER(Inputs), C.POINTER(Outputs), C.POINTER(Options),
                                   C.POINTER(Report)]
    L.fp_project_noalloc.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Outputs), C.POINTER(Options),
                                     C.POINTER(Report)]
    L.fp_submit.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Outputs), C.POINTER(Options), C.POINTER(C.c_uint64)]
    L.fp_poll.argtypes = [vp, C.c_uint64, C.c_int32, C.POINTER(Report), C.POINTER(C.c_double)]
    L.fp_probe_apply.argtypes = [vp, vp, vp, vp]
    L.fp_probe_rhs.argtypes = [vp, C.POINTER(Inputs), vp]
    L.fp_probe_vcycle.argtypes = [vp, vp, C.POINTER(Options), vp, vp]
    L.fp_probe_level_phi.argtypes = [vp, vp, C.c_int32, C.c_int32, vp]
    L.fp_probe_tall_layout.argtypes = [vp, vp, vp, vp]
    L.fp_probe_taps.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp]
    L.fp_tall_remesh.argtypes = [vp, C.POINTER(TallRemeshArgs), C.POINTER(C.c_int64)]
    L.fp_extrapolate_velocity.argtypes = [vp, C.POINTER(ExtrapolateArgs), C.POINTER(ExtrapolateReport)]
    L.fp_advect_velocity.argtypes = [vp, C.POINTER(AdvectArgs), C.POINTER(AdvectReport)]
    for name in EXPORTED_SYMBOLS:
        if name not in ("fp_last_error", "fp_abi_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != FP_OK:
        raise FluidProjectionError(rc, _lib.fp_last_error().decode(errors="replace"))


def default_options(**kw) -> Options:
    o = Options()
    _check(load_library().fp_default_options(C.byref(o)))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        if k == "timed_kernel" and isinstance(v, str):
            v = KERNELS[v]
        if k == "solver" and isinstance(v, str):
            v = SOLVERS[v]
        setattr(o, k, v)
    return o


class Context:
    """One device, one stream, an HBM budget (fp_ctx_*)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        L = load_library()
        self._h = C.c_void_p()
        _check(L.fp_ctx_create(device, hbm_budget_bytes, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def bytes_in_use(self) -> int:
        b = C.c_uint64()
        _check(_lib.fp_ctx_bytes_in_use(self._h, C.byref(b)))
        return b.value

    def destroy(self):
        if self._h:
            _check(_lib.fp_ctx_destroy(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


_default_ctx = None


def grid_bytes(dims, precision: str = "fp16") -> int:
    """Device bytes a uniform grid of these dims holds (fp_grid_bytes)."""
    nx, ny, nz = (int(x) for x in dims)
    b = C.c_uint64()
    d = GridDesc(LAYOUT_UNIFORM, nx, ny, nz, 0, 0, 1.0 / nx, None)
    _check(load_library().fp_grid_bytes(C.byref(d), PRECISIONS[precision], C.byref(b)))
    return int(b.value)


def tall_grid_bytes(nx: int, fine_ny: int, nz: int, band: int, max_delta: int = 4) -> int:
    """Device bytes a tall-cell grid of this description would hold."""
    b = C.c_uint64()
    bases = np.full((nx, nz), 2, np.int32)
    d = GridDesc(LAYOUT_TALL_CELL, nx, fine_ny, nz, band, max_delta, 1.0 / fine_ny, bases.ctypes.data)
    _check(load_library().fp_grid_bytes(C.byref(d), PRECISIONS["fp32"], C.byref(b)))
    return int(b.value)


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _ptr(a, dtype):
    """(pointer, memory kind, keepalive) for a numpy array or a torch CUDA tensor."""
    if a is None:
        return None, None, None
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        import torch
        want = torch.float32 if dtype == np.float32 else torch.float64
        if a.dtype != want or not a.is_contiguous():
            raise TypeError(f"tensor must be contiguous {want}")
        return a.data_ptr(), (MEMORY_DEVICE if a.is_cuda else MEMORY_HOST), a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr.ctypes.data, MEMORY_HOST, arr


@dataclass
class ProjectionResult:
    pressure: object
    u: object
    v: object
    w: object
    report: dict


class UniformPressureProjection:
    """One resident uniform MAC grid (nx, ny, nz) of cubic cells of size h."""

    def __init__(self, dims, spacing, context: Context | None = None):
        nx, ny, nz = (int(x) for x in dims)
        h = float(spacing if np.isscalar(spacing) else spacing[0])
        if not np.isscalar(spacing) and not all(abs(float(s) - h) <= 1e-12 * h for s in spacing):
            raise ValueError("cells must be cubic (equal spacing on every axis)")
        self.ctx = context or default_context()
        self.dims = (nx, ny, nz)
        self.h = h
        desc = GridDesc(LAYOUT_UNIFORM, nx, ny, nz, 0, 0, h, None)
        self._g = C.c_void_p()
        rc = load_library().fp_grid_create(self.ctx.handle, C.byref(desc), C.byref(self._g))
        if rc == -1 or rc == -4:
            raise ValueError(_lib.fp_last_error().decode())
        _check(rc)
        self.diagnostics = None

    @property
    def handle(self):
        return self._g

    def shapes(self):
        """Shapes of (phi / pressure, u, v, w)."""
        nx, ny, nz = self.dims
        return [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]

    @property
    def level_count(self) -> int:
        return len(self.levels()[0])

    def levels(self):
        dims = np.zeros(48, np.int32)
        strides = np.zeros(48, np.int32)
        L = C.c_int32()
        _check(_lib.fp_grid_levels(self._g, 16, dims.ctypes.data, strides.ctypes.data, C.byref(L)))
        return dims[: 3 * L.value].reshape(-1, 3), strides[: 3 * L.value].reshape(-1, 3)

    def _prepare(self, phi, u, v, w, dt, rho, target_divergence=None, out=None, density=None):
        """C input/output records for one projection, the result arrays and
        the objects that must stay alive until the call returns.  With
        `density` (the surface density per cell) phi may be None."""
        if phi is None and density is None:
            raise ValueError("give phi or density")
        dens = _ptr(density, np.float32)
        parts = [_ptr(phi, np.float32), _ptr(u, np.float32), _ptr(v, np.float32), _ptr(w, np.float32),
                 _ptr(target_divergence, np.float32), dens]
        phi = phi if phi is not None else density
        kinds = {p[1] for p in parts if p[1] is not None}
        if len(kinds) != 1:
            raise TypeError("all arrays must live in the same memory (host or device)")
        mem = kinds.pop()
        shapes = self.shapes()
        for arr, shp in zip((phi, u, v, w), shapes):
            if tuple(arr.shape) != shp:
                raise ValueError(f"array shape {tuple(arr.shape)} != {shp}")
        ins = Inputs(mem, 0, parts[0][0], parts[1][0], parts[2][0], parts[3][0], parts[4][0], float(dt), float(rho),
                     None, dens[0])
        if mem == MEMORY_DEVICE:
            # the library's work starts after what torch has queued on the
            # current stream (producers of the inputs, earlier users of `out`)
            import torch
            ins.stream_order = 1
            ins.stream = torch.cuda.current_stream(phi.device).cuda_stream  # 0: the null stream
        if mem == MEMORY_HOST:
            p = np.zeros(shapes[0], np.float32)
            uo = np.zeros(shapes[1], np.float32)
            vo = np.zeros(shapes[2], np.float32)
            wo = np.zeros(shapes[3], np.float32)
            outs = Outputs(p.ctypes.data, uo.ctypes.data, vo.ctypes.data, wo.ctypes.data)
        else:
            import torch
            dev = phi.device
            if out is None:
                out = (torch.empty(shapes[0], device=dev), torch.empty(shapes[1], device=dev),
                       torch.empty(shapes[2], device=dev), torch.empty(shapes[3], device=dev))
            p, uo, vo, wo = out
            outs = Outputs(p.data_ptr() if p is not None else None, uo.data_ptr(), vo.data_ptr(), wo.data_ptr())
        return ins, outs, (p, uo, vo, wo), [q[2] for q in parts]

    def project(self, phi, u, v, w, dt, rho, target_divergence=None, out=None, options: Options | None = None,
                density=None, **opts) -> ProjectionResult:
        """One projection.  Host arrays in -> host arrays out; CUDA tensors in ->
        results written into `out` tensors (or fresh tensors) on the device.
        `density` (uniform grids) replaces phi by the reference's surface
        density: phi = -(density - 1/2) h."""
        o = options or default_options(**opts)
        ins, outs, res, keep = self._prepare(phi, u, v, w, dt, rho, target_divergence, out, density)
        rep = Report()
        _check(_lib.fp_project(self._g, C.byref(ins), C.byref(outs), C.byref(o), C.byref(rep)))
        del keep
        self.diagnostics = rep.as_dict()
        return ProjectionResult(*res, self.diagnostics)

    def submit(self, phi, u, v, w, dt, rho, target_divergence=None, out=None, options: Options | None = None,
               **opts) -> "Ticket":
        """Queue one projection on the context's serving worker (fp_submit);
        the arrays stay referenced by the ticket until it completes."""
        o = options or default_options(**opts)
        ins, outs, res, keep = self._prepare(phi, u, v, w, dt, rho, target_divergence, out)
        t = C.c_uint64()
        _check(_lib.fp_submit(self._g, C.byref(ins), C.byref(outs), C.byref(o), C.byref(t)))
        return Ticket(self, t.value, res, (keep, ins, outs, o))

    # ---- parity probes --------------------------------------------------------
    def probe_apply(self, phi, p):
        phi = np.ascontiguousarray(phi, np.float32)
        p = np.ascontiguousarray(p, np.float64)
        out = np.zeros(self.shapes()[0], np.float64)
        _check(_lib.fp_probe_apply(self._g, phi.ctypes.data, p.ctypes.data, out.ctypes.data))
        return out

    def probe_rhs(self, phi, u, v, w, dt, rho, target_divergence=None):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (phi, u, v, w)]
        q = np.ascontiguousarray(target_divergence, np.float32) if target_divergence is not None else None
        ins = Inputs(MEMORY_HOST, 0, arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data,
                     arrs[3].ctypes.data, q.ctypes.data if q is not None else None, float(dt), float(rho))
        b = np.zeros(self.shapes()[0], np.float64)
        _check(_lib.fp_probe_rhs(self._g, C.byref(ins), b.ctypes.data))
        return b

    def probe_vcycle(self, phi, r, options: Options | None = None, **opts):
        o = options or default_options(**opts)
        phi = np.ascontiguousarray(phi, np.float32)
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros(self.dims, np.float32)
        _check(_lib.fp_probe_vcycle(self._g, phi.ctypes.data, C.byref(o), r.ctypes.data, z.ctypes.data))
        return z

    def probe_level_phi(self, phi, level, preserve=2):
        dims, _ = self.levels()
        out = np.zeros(tuple(int(x) for x in dims[level]), np.float32)
        phi = np.ascontiguousarray(phi, np.float32)
        _check(_lib.fp_probe_level_phi(self._g, phi.ctypes.data, level, preserve, out.ctypes.data))
        return out

    def extrapolate(self, u, v, w, phi=None, density=None, open_top: bool = False, max_passes: int = 0, out=None):
        """Velocity extrapolation into the air (fp_extrapolate_velocity): the
        narrow band of Jeong-Ross-Whitaker passes, then the hierarchy fill.
        Returns (u, v, w, report dict): host arrays for host inputs, CUDA
        tensors (written into `out` when given; it may be the inputs) for CUDA
        tensors."""
        if (phi is None) == (density is None):
            raise ValueError("give exactly one of phi and density")
        field = phi if phi is not None else density
        parts = [_ptr(a, np.float32) for a in (field, u, v, w)]
        kinds = {q[1] for q in parts}
        if len(kinds) != 1:
            raise TypeError("all arrays must live in the same memory (host or device)")
        mem = kinds.pop()
        shapes = self.shapes()
        for arr, shp in zip((field, u, v, w), shapes):
            if tuple(arr.shape) != shp:
                raise ValueError(f"array shape {tuple(arr.shape)} != {shp}")
        if mem == MEMORY_HOST:
            res = [np.zeros(shp, np.float32) for shp in shapes[1:]]
            optr = [r.ctypes.data for r in res]
        else:
            import torch
            res = list(out) if out is not None else [torch.empty(shp, device=field.device) for shp in shapes[1:]]
            for r, shp in zip(res, shapes[1:]):
                if tuple(r.shape) != shp or r.dtype != torch.float32 or not r.is_contiguous():
                    raise ValueError("out must hold contiguous float32 tensors of the face shapes")
            optr = [r.data_ptr() for r in res]
            torch.cuda.current_stream(field.device).synchronize()
        args = ExtrapolateArgs(mem, int(bool(open_top)), int(max_passes), 0,
                               parts[0][0] if phi is not None else None, parts[0][0] if phi is None else None,
                               parts[1][0], parts[2][0], parts[3][0], *optr)
        rep = ExtrapolateReport()
        _check(_lib.fp_extrapolate_velocity(self._g, C.byref(args), C.byref(rep)))
        return (*res, rep.as_dict())

    def advect(self, u, v, w, dt, phi=None, density=None, transport=None, gravity: float = -9.80665,
               viscosity: float = 0.0, scheme: str = "semi-lagrangian", open_top: bool = False,
               no_slip: bool = False, out=None):
        """Velocity advection and body forces (fp_advect_velocity): u, v, w
        traced back through `transport` (their extrapolation; computed by the
        library when None), semi-Lagrangian or bounded MacCormack, then
        gravity along y and viscosity.  Returns (u, v, w, report dict)."""
        if (phi is None) == (density is None):
            raise ValueError("give exactly one of phi and density")
        field = phi if phi is not None else density
        arrays = [field, u, v, w] + (list(transport) if transport is not None else [])
        parts = [_ptr(a, np.float32) for a in arrays]
        kinds = {q[1] for q in parts}
        if len(kinds) != 1:
            raise TypeError("all arrays must live in the same memory (host or device)")
        mem = kinds.pop()
        shapes = self.shapes()
        for arr, shp in zip(arrays, shapes + shapes[1:]):
            if tuple(arr.shape) != shp:
                raise ValueError(f"array shape {tuple(arr.shape)} != {shp}")
        if mem == MEMORY_HOST:
            res = [np.zeros(shp, np.float32) for shp in shapes[1:]]
            optr = [r.ctypes.data for r in res]
        else:
            import torch
            res = list(out) if out is not None else [torch.empty(shp, device=field.device) for shp in shapes[1:]]
            for r, shp in zip(res, shapes[1:]):
                if tuple(r.shape) != shp or r.dtype != torch.float32 or not r.is_contiguous():
                    raise ValueError("out must hold contiguous float32 tensors of the face shapes")
            optr = [r.data_ptr() for r in res]
            torch.cuda.current_stream(field.device).synchronize()
        tr = [p[0] for p in parts[4:]] if transport is not None else [None, None, None]
        args = AdvectArgs(mem, ADVECT_SCHEMES[scheme], int(bool(open_top)), int(bool(no_slip)), float(dt),
                          float(gravity), float(viscosity), parts[0][0] if phi is not None else None,
                          parts[0][0] if phi is None else None, parts[1][0], parts[2][0], parts[3][0], *tr, *optr)
        rep = AdvectReport()
        _check(_lib.fp_advect_velocity(self._g, C.byref(args), C.byref(rep)))
        return (*res, {"device_ms": rep.device_ms})

    def set_tuning(self, tile_z: int = 0, groups: int = 0, pairs: int = 0):
        """Sweep tile width along z (16 or 32) and workgroups per sweep of the
        fused levels (fp_grid_set_tuning); 0 restores the automatic choice.
        pairs 1 runs the finest level's sweeps two to a launch (FP16/BF16
        multigrid; the same bits as one launch per sweep)."""
        _check(_lib.fp_grid_set_tuning(self._g, C.byref(Tuning(int(tile_z), int(groups), int(pairs)))))

    def probe_taps(self, level, axis):
        """Transfer taps between level-1 and level along axis, as the kernels
        evaluate them: (prolong coarse index (nf, 2), weight (nf, 2)),
        (restrict fine index (nc, 4), weight (nc, 4)); index -1 where unused."""
        dims, _ = self.levels()
        nf, nc = int(dims[level - 1][axis]), int(dims[level][axis])
        pc = np.zeros((nf, 2), np.int32)
        pw = np.zeros((nf, 2), np.float32)
        rf = np.zeros((nc, 4), np.int32)
        rw = np.zeros((nc, 4), np.float32)
        _check(_lib.fp_probe_taps(self._g, level, axis, pc.ctypes.data, pw.ctypes.data, rf.ctypes.data,
                                  rw.ctypes.data))
        return (pc, pw), (rf, rw)

    def destroy(self):
        if self._g:
            _check(_lib.fp_grid_destroy(self._g))
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class TallCellPressureProjection(UniformPressureProjection):
    """One resident restricted tall-cell grid: nx x nz columns of fine_ny rows
    of cubic cells h, each column a tall cell of `bases[x, z]` rows under a
    band of `band` cubic cells (fluid_amd.tall_layout for the sample and slot
    conventions).  project() takes phi on the samples (nx, band+2, nz) and
    the slot velocities u (nx+1, S, nz), v (nx, band+2, nz), w (nx, S, nz+1)."""

    def __init__(self, bases, fine_ny, band, spacing, max_delta: int = 4, context: Context | None = None):
        bases = np.ascontiguousarray(bases, np.int32)
        if bases.ndim != 2:
            raise ValueError("bases must be (nx, nz)")
        nx, nz = bases.shape
        self.ctx = context or default_context()
        self.bases = bases
        self.fine_ny, self.band, self.max_delta = int(fine_ny), int(band), int(max_delta)
        self.h = float(spacing)
        self.dims = (nx, self.fine_ny, nz)
        desc = GridDesc(LAYOUT_TALL_CELL, nx, self.fine_ny, nz, self.band, self.max_delta, self.h,
                        bases.ctypes.data)
        self._g = C.c_void_p()
        rc = load_library().fp_grid_create(self.ctx.handle, C.byref(desc), C.byref(self._g))
        if rc == -1 or rc == -4:
            raise ValueError(_lib.fp_last_error().decode())
        _check(rc)
        self.diagnostics = None

    @classmethod
    def from_layout(cls, layout, context: Context | None = None):
        return cls(layout.bases, layout.fine_ny, layout.band, layout.h, layout.max_delta, context)

    def shapes(self):
        nx, _, nz = self.dims
        P, S = self.band + 2, self.band + 2 + self.max_delta
        return [(nx, P, nz), (nx + 1, S, nz), (nx, P, nz), (nx, S, nz + 1)]

    def probe_vcycle(self, *a, **k):
        raise FluidProjectionError(-1, "the V-cycle probe covers uniform grids")

    def probe_level_phi(self, *a, **k):
        raise FluidProjectionError(-1, "the level-set probe covers uniform grids")

    def remesh(self, phi, u, v, w, liquid_halo: int = 0, air_halo: int = 0, out=None):
        """Re-lay the grid out around the surface of phi and move phi and the
        slot velocities to the new layout on the device (fp_tall_remesh).
        Returns (phi, u, v, w, changed_columns) in the new layout -- host
        arrays for host inputs, CUDA tensors (written into `out` when given;
        it may be the inputs themselves) for CUDA tensors -- and updates
        self.bases.  Halos 0 take tallCellSettings.high (16 liquid, 8 air)."""
        parts = [_ptr(a, np.float32) for a in (phi, u, v, w)]
        kinds = {q[1] for q in parts}
        if len(kinds) != 1:
            raise TypeError("all arrays must live in the same memory (host or device)")
        mem = kinds.pop()
        shapes = self.shapes()
        for arr, shp in zip((phi, u, v, w), shapes):
            if tuple(arr.shape) != shp:
                raise ValueError(f"array shape {tuple(arr.shape)} != {shp}")
        if mem == MEMORY_HOST:
            res = [np.zeros(shp, np.float32) for shp in shapes]
            optr = [r.ctypes.data for r in res]
        else:
            import torch
            res = list(out) if out is not None else [torch.empty(shp, device=phi.device) for shp in shapes]
            for r, shp in zip(res, shapes):
                if tuple(r.shape) != shp or r.dtype != torch.float32 or not r.is_contiguous():
                    raise ValueError("out must hold contiguous float32 tensors of the input shapes")
            optr = [r.data_ptr() for r in res]
            # the library's stream is ordered after the caller's work by a device sync
            torch.cuda.current_stream(phi.device).synchronize()
        bases = np.zeros(self.bases.shape, np.int32)
        args = TallRemeshArgs(mem, int(liquid_halo), int(air_halo), 0, parts[0][0], parts[1][0], parts[2][0],
                              parts[3][0], *optr, bases.ctypes.data)
        changed = C.c_int64(0)
        _check(_lib.fp_tall_remesh(self._g, C.byref(args), C.byref(changed)))
        self.bases = bases
        return (*res, int(changed.value))

    def probe_layout(self, phi):
        """(fine row, unknown flag) of every sample, computed on the device."""
        phi = np.ascontiguousarray(phi, np.float32)
        rows = np.zeros(self.shapes()[0], np.int32)
        unk = np.zeros(self.shapes()[0], np.uint8)
        _check(_lib.fp_probe_tall_layout(self._g, phi.ctypes.data, rows.ctypes.data, unk.ctypes.data))
        return rows, unk.astype(bool)


def project_scene(scene, **opts) -> ProjectionResult:
    """Convenience: one projection of a fluid_amd.scenes.UniformScene."""
    proj = UniformPressureProjection(scene.shape, scene.h)
    try:
        return proj.project(scene.phi, scene.u, scene.v, scene.w, scene.dt, scene.rho, **opts)
    finally:
        proj.destroy()


def project_batch(grids, requests, options: Options | None = None, **opts):
    """Independent projections of several grids (one context) in one batched
    solve (fp_project_batch).  requests: per grid a dict with phi, u, v, w, dt,
    rho and optionally target_divergence / out.  Returns a ProjectionResult
    per grid; each equals what the grid's projection alone returns."""
    if len(grids) != len(requests) or not grids:
        raise ValueError("one request per grid")
    o = options or default_options(**opts)
    n = len(grids)
    ins_arr = (Inputs * n)()
    outs_arr = (Outputs * n)()
    results, keep = [], []
    for i, (g, rq) in enumerate(zip(grids, requests)):
        ins, outs, res, k = g._prepare(rq["phi"], rq["u"], rq["v"], rq["w"], rq["dt"], rq["rho"],
                                       rq.get("target_divergence"), rq.get("out"))
        ins_arr[i] = ins
        outs_arr[i] = outs
        results.append(res)
        keep.append(k)
    reps = (Report * n)()
    handles = (C.c_void_p * n)(*[g._g.value for g in grids])
    _check(_lib.fp_project_batch(C.cast(handles, C.c_void_p), n, ins_arr, outs_arr, C.byref(o), reps))
    del keep
    out = []
    for g, res, rep in zip(grids, results, reps):
        g.diagnostics = rep.as_dict()
        out.append(ProjectionResult(*res, g.diagnostics))
    return out


class Ticket:
    """A submitted projection (fp_submit); wait() returns its result."""

    def __init__(self, grid, ticket, res, keep):
        self.grid = grid
        self.ticket = ticket
        self._res = res
        self._keep = keep
        self.result = None
        self.latency_ms = None

    def wait(self, timeout_ms: int = -1):
        """The ProjectionResult once done, else None after timeout_ms."""
        if self.result is not None:
            return self.result
        rep = Report()
        lat = C.c_double()
        rc = _lib.fp_poll(self.grid.ctx.handle, C.c_uint64(self.ticket), int(timeout_ms), C.byref(rep), C.byref(lat))
        if rc == PENDING:
            return None
        _check(rc)
        self.latency_ms = lat.value
        self.grid.diagnostics = rep.as_dict()
        self.result = ProjectionResult(*self._res, self.grid.diagnostics)
        self._keep = None
        return self.result
