"""pfsht — Python host-side mirror of PartialFSHT.jl's operator API over libpfsht_hip.so.

    plan = Plan(theta, phi, lmax=..., tol=1e-10, ntrans=1, method="auto")
    mul(f, plan, alm)          # mul!(f, plan, alm)        src/Plan.jl:137-141
    mul(alm, plan.H, f)        # mul!(alm, plan', f)       src/Plan.jl:143-148
    f = plan @ alm ; alm = plan.H @ f                      # Base.:* (src/Plan.jl:190-199)

Arrays are the reference's column-major layouts (src/Layout.jl:13-56). Host arrays are numpy
complex128 (Fortran order for alm); device arrays are torch complex128 CUDA tensors whose memory
is the column-major layout (the helpers `alm_device`/`field_device` make such tensors). Shape
errors raise ValueError (DimensionMismatch), bad arguments ValueError (ArgumentError). Every call
goes to the HIP library; there is no CPU fallback, and importing raises when the library is
missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["Plan", "AdjointPlan", "Rings", "mul", "coeff_size", "sph_mode", "zeros_alm",
           "PFSHTError", "DimensionMismatch", "ArgumentError", "lib", "library_path"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# PFSHT_HIP_LIB names another build of the library (the Julia extension reads the same variable)
_LIB = os.environ.get("PFSHT_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libpfsht_hip.so")


class PFSHTError(RuntimeError):
    pass


class DimensionMismatch(ValueError):
    pass


class ArgumentError(ValueError):
    pass


class GeometryRefused(ArgumentError):
    pass


class InfoC(C.Structure):
    _fields_ = [("method", C.c_int32), ("kind", C.c_int32), ("flip", C.c_int32),
                ("lmax", C.c_int32), ("ntrans", C.c_int32), ("precision", C.c_int32),
                ("nth_half", C.c_int32), ("nth", C.c_int32), ("nph", C.c_int32),
                ("mmax", C.c_int32), ("es_w", C.c_int32), ("es_n1", C.c_int32),
                ("es_n2", C.c_int32), ("npts", C.c_int64), ("ring_work", C.c_int64),
                ("device_bytes", C.c_int64), ("tol", C.c_double), ("es_beta", C.c_double),
                ("es_rows", C.c_int32), ("nboxes", C.c_int32), ("spin", C.c_int32),
                ("seg_blocks", C.c_int32), ("adj_tiles", C.c_int32), ("mfma_tiles", C.c_int32),
                ("grid_fft", C.c_int32)]


class RingShapeC(C.Structure):
    _fields_ = [("seg_blocks", C.c_int32), ("adj_tiles", C.c_int32), ("mfma_tiles", C.c_int32),
                ("reserved", C.c_int32), ("ms", C.c_double)]


class PlanOptsC(C.Structure):
    _fields_ = [("method", C.c_int32), ("precision", C.c_int32), ("spin", C.c_int32), ("maxboxes", C.c_int32),
                ("oversample_theta", C.c_double), ("oversample_phi", C.c_double), ("epsapo_max", C.c_double),
                ("packed", C.c_int32), ("tune", C.c_int32), ("grid_fft", C.c_int32)]


class BoxC(C.Structure):
    _fields_ = [("thlo", C.c_double), ("thhi", C.c_double), ("phistart", C.c_double), ("phiext", C.c_double),
                ("touches_north", C.c_int32), ("touches_south", C.c_int32), ("fsky", C.c_double),
                ("dl", C.c_double)]


class SolveInfoC(C.Structure):
    _fields_ = [("shannon", C.c_double), ("dl", C.c_double), ("relres", C.c_double), ("iters", C.c_int32),
                ("mode", C.c_int32)]


SOLVE_MODES = {"cg": 0, "wiener": 1, "minnorm": 2}

NSTAGES = 9
STAGES = ["ring", "pack", "phi_fft", "dfs", "grid_fft", "deconv", "fine_fft", "nufft", "ring_reduce"]
METHODS = {"auto": 0, "direct": 1, "stretch": 2, "multibox": 3}
KINDS = {0: "pole", 1: "interior", 2: "full", -1: "direct", -2: "multibox"}
# the grid stage of a stretched apply: LDS kernels with fused FFTs, or rocFFT plans (PFSHT_GRID_FFT_*)
GRID_FFT = {"auto": 0, "lds": 1, "library": 2}
WHERE_HOST, WHERE_DEVICE = 0, 1

EXPORTS = {
    "pfsht_last_error": (C.c_char_p, []),
    "pfsht_version": (C.c_char_p, []),
    "pfsht_plan_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_double,
                                    C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "pfsht_plan_create_spin": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_double,
                                         C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "pfsht_plan_opts_default": (None, [C.c_void_p]),
    "pfsht_plan_create_opts": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_int32,
                                         C.c_void_p, C.POINTER(C.c_void_p)]),
    "pfsht_plan_estimate_bytes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_int32,
                                            C.c_void_p, C.POINTER(C.c_int64)]),
    "pfsht_tune_get": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                 C.POINTER(C.c_int32)]),
    "pfsht_tune_put": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "pfsht_tune_load": (C.c_int, [C.c_char_p]),
    "pfsht_tune_save": (C.c_int, [C.c_char_p]),
    "pfsht_tune_clear": (None, []),
    "pfsht_tune_size": (C.c_int64, []),
    "pfsht_plan_tune_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "pfsht_plan_tune_key": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "pfsht_group_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pfsht_group_destroy": (None, [C.c_void_p]),
    "pfsht_group_synth": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pfsht_group_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pfsht_synth_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_adjoint_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_synth_cols": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pfsht_adjoint_cols": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pfsht_plan_destroy": (None, [C.c_void_p]),
    "pfsht_plan_info": (C.c_int, [C.c_void_p, C.POINTER(InfoC)]),
    "pfsht_synth": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_rings_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                     C.c_int32, C.POINTER(C.c_void_p)]),
    "pfsht_spin_rings_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                          C.c_int32, C.POINTER(C.c_void_p)]),
    "pfsht_rings_destroy": (None, [C.c_void_p]),
    "pfsht_rings_work": (C.c_int64, [C.c_void_p]),
    "pfsht_rings_mmax": (C.c_int32, [C.c_void_p]),
    "pfsht_rings_from_alm": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_alm_from_rings": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_alm_from_rings_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pfsht_plan_rings": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "pfsht_plan_nufft": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "pfsht_stage_times": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]),
    "pfsht_bounding_box": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(BoxC)]),
    "pfsht_partition_boxes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_double,
                                        C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]),
    "pfsht_apply_resolution": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "pfsht_pseudo_spectrum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                        C.c_void_p]),
    "pfsht_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_void_p,
                              C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(SolveInfoC)]),
    "pfsht_server_opts_default": (None, [C.c_void_p]),
    "pfsht_server_create_ex": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "pfsht_server_create": (C.c_int, [C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]),
    "pfsht_server_destroy": (None, [C.c_void_p]),
    "pfsht_server_plan": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                    C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_int32)]),
    "pfsht_server_submit": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_int64)]),
    "pfsht_server_query": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "pfsht_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "pfsht_server_get_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pfsht_server_epoch": (C.c_int, [C.c_void_p]),
    "pfsht_server_times": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
}

_lib = None


def library_path() -> str:
    return _LIB


def lib():
    """Load libpfsht_hip.so. Raises when it has not been built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError(f"{_LIB} is missing: build it with partialfsht.jl_amd/build.py")
        L = C.CDLL(_LIB)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc == 0:
        return
    msg = lib().pfsht_last_error().decode()
    if rc == 1:
        raise DimensionMismatch(msg)
    if rc == 3:
        raise GeometryRefused(msg)
    if rc in (2, 6):
        raise ArgumentError(msg)
    raise PFSHTError(f"error {rc}: {msg}")


def coeff_size(lmax: int):
    """(lmax+1, 2lmax+1) — src/Layout.jl:13."""
    return (int(lmax) + 1, 2 * int(lmax) + 1)


def sph_mode(l: int, m: int):
    """0-based (row, column) of (l, m) — src/Layout.jl:20."""
    am = abs(int(m))
    return (int(l) - am, 2 * am + (1 if m >= 0 else 0) - 1)


def spin_coeff_index(l: int, m: int, lmax: int):
    """spin_coeff_index(l, m, lmax) — src/Layout.jl:28-29, 0-based: (l, m + lmax)."""
    return (int(l), int(m) + int(lmax))


def zeros_alm(lmax: int, ntrans: int = 1):
    shape = coeff_size(lmax) if ntrans == 1 else (*coeff_size(lmax), ntrans)
    return np.zeros(shape, np.complex128, order="F")


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _column_major(x) -> bool:
    """True when a torch tensor's memory is its column-major (Fortran) layout: stride 1 along the
    first axis and each later stride the product of the earlier sizes (size-1 axes are free)."""
    want = 1
    for n, s in zip(x.shape, x.stride()):
        if n != 1 and s != want:
            return False
        want *= n
    return True


def _is_c64(x) -> bool:
    """True for a complex64 (ComplexF32) array or tensor."""
    if _is_torch(x):
        import torch
        return x.dtype == torch.complex64
    return isinstance(x, np.ndarray) and x.dtype == np.complex64


def _buf(x, count: int, what: str, c64: bool = False):
    """(pointer, where) of a complex128 buffer (complex64 when c64) of `count` elements in
    column-major order. A torch tensor must be 1-D and contiguous, or hold its column-major layout (a
    row-major 2-D or 3-D tensor would be read transposed, so it is refused)."""
    want = "complex64" if c64 else "complex128"
    if _is_torch(x):
        import torch
        if x.dtype != (torch.complex64 if c64 else torch.complex128):
            raise ArgumentError(f"{what} must be {want}")
        if x.dim() <= 1:
            if not x.is_contiguous():
                raise ArgumentError(f"{what} must be contiguous")
        elif not _column_major(x):
            raise ArgumentError(f"{what} is a {x.dim()}-D tensor with strides {tuple(x.stride())}: pass a flat "
                                f"tensor or one whose memory is column-major (e.g. t.T.contiguous().T)")
        if x.numel() != count:
            raise DimensionMismatch(f"{what} has {x.numel()} elements, expected {count}")
        return C.c_void_p(x.data_ptr()), (WHERE_DEVICE if x.is_cuda else WHERE_HOST)
    if not isinstance(x, np.ndarray) or x.dtype != (np.complex64 if c64 else np.complex128):
        raise ArgumentError(f"{what} must be a {want} numpy array or torch tensor")
    if x.size != count:
        raise DimensionMismatch(f"{what} has shape {x.shape}, expected {count} elements")
    if not (x.flags.f_contiguous or x.ndim <= 1):
        raise ArgumentError(f"{what} must be Fortran-ordered (column-major)")
    return C.c_void_p(x.ctypes.data), WHERE_HOST


def _assert_coeffs(alm, L: int, B: int):
    """src/Layout.jl:35-45."""
    shp = tuple(alm.shape)
    want = coeff_size(L)
    if B == 1 and len(shp) == 2:
        ok = shp == want
    else:
        ok = len(shp) == 3 and shp == (*want, B)
    if _is_torch(alm) and alm.dim() == 1:
        ok = alm.numel() == want[0] * want[1] * B
    if not ok:
        raise DimensionMismatch(f"coefficient array is {shp}, expected {(*want, B) if B > 1 else want}")


def _assert_field(f, M: int, B: int):
    """src/Layout.jl:47-56."""
    shp = tuple(f.shape)
    if B == 1 and len(shp) == 1:
        ok = shp[0] == M
    else:
        ok = len(shp) == 2 and shp == (M, B)
    if _is_torch(f) and f.dim() == 1:
        ok = f.numel() == M * B
    if not ok:
        raise DimensionMismatch(f"field is {shp}, expected {(M, B) if B > 1 else (M,)}")


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        except ImportError:
            pass
        return C.c_void_p(0)
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


def bounding_box(theta, phi) -> dict:
    """bounding_box(θ, φ) (src/Box.jl:43-73) with box_fsky and spectral_resolution: host arrays."""
    th = np.ascontiguousarray(theta, np.float64)
    ph = np.ascontiguousarray(phi, np.float64)
    if th.shape != ph.shape:
        raise DimensionMismatch("theta and phi must have the same length")
    out = BoxC()
    _check(lib().pfsht_bounding_box(th.ctypes.data, ph.ctypes.data, len(th), C.byref(out)))
    d = {k: getattr(out, k) for k, _ in BoxC._fields_}
    d["touches_north"], d["touches_south"] = bool(d["touches_north"]), bool(d["touches_south"])
    return d


def partition_boxes(theta, phi, lmax: int, ntrans: int = 1, tol: float = 1e-10, maxboxes: int = 8):
    """partition_boxes(θ, φ, costfn; maxboxes) with the stretched leaf's counted cost
    (src/Box.jl:159-235, src/Cost.jl:193): a list of index arrays, in the reference's order."""
    th = np.ascontiguousarray(theta, np.float64)
    ph = np.ascontiguousarray(phi, np.float64)
    if th.shape != ph.shape:
        raise DimensionMismatch("theta and phi must have the same length")
    gid = np.zeros(len(th), np.int32)
    ng = C.c_int32()
    _check(lib().pfsht_partition_boxes(th.ctypes.data, ph.ctypes.data, len(th), int(lmax), int(ntrans),
                                       float(tol), int(maxboxes), gid.ctypes.data, C.byref(ng)))
    return [np.flatnonzero(gid == g) for g in range(ng.value)]


def plan_estimate_bytes(theta, phi, lmax: int, tol: float = 1e-10, ntrans: int = 1, method: str = "auto",
                        precision: str = "fp64", spin: int = 0, packed: bool = False) -> int:
    """Device bytes a Plan with these arguments would take (an upper estimate), without building it."""
    th = np.ascontiguousarray(theta, np.float64)
    ph = np.ascontiguousarray(phi, np.float64)
    if th.shape != ph.shape or th.ndim != 1:
        raise DimensionMismatch("theta and phi must have the same length")
    o = PlanOptsC(method=METHODS[method], precision={"fp64": 0, "mixed": 1}[precision], spin=int(spin),
                  packed=int(bool(packed)))
    out = C.c_int64()
    _check(lib().pfsht_plan_estimate_bytes(th.ctypes.data, ph.ctypes.data, len(th), int(lmax), float(tol),
                                           int(ntrans), C.byref(o), C.byref(out)))
    return int(out.value)


def tune_get(arch: str, lmax: int, nrings: int, mmax: int, ntrans: int):
    """The ring-pass shape cached under the key, or None (pfsht_tune_get; host only)."""
    sh = RingShapeC()
    found = C.c_int32()
    _check(lib().pfsht_tune_get(arch.encode(), int(lmax), int(nrings), int(mmax), int(ntrans), C.byref(sh),
                                C.byref(found)))
    if not found.value:
        return None
    return {k: getattr(sh, k) for k in ("seg_blocks", "adj_tiles", "mfma_tiles", "ms")}


def tune_put(arch: str, lmax: int, nrings: int, mmax: int, ntrans: int, seg_blocks: int = 0, adj_tiles: int = 0,
             mfma_tiles: int = 0, ms: float = 0.0):
    """Record a ring-pass shape under the key (pfsht_tune_put; host only)."""
    sh = RingShapeC(seg_blocks=int(seg_blocks), adj_tiles=int(adj_tiles), mfma_tiles=int(mfma_tiles), ms=float(ms))
    _check(lib().pfsht_tune_put(arch.encode(), int(lmax), int(nrings), int(mmax), int(ntrans), C.byref(sh)))


def tune_load(path: str):
    _check(lib().pfsht_tune_load(str(path).encode()))


def tune_save(path: str):
    _check(lib().pfsht_tune_save(str(path).encode()))


def tune_clear():
    lib().pfsht_tune_clear()


def tune_size() -> int:
    return int(lib().pfsht_tune_size())


def _weights(w, M: int, where: int):
    """(pointer, keep-alive) of npts float64 weights on the same side as the other arrays."""
    if w is None:
        return C.c_void_p(0), None
    if _is_torch(w):
        import torch
        if w.dtype != torch.float64 or not w.is_contiguous() or w.numel() != M:
            raise ArgumentError("w must be a contiguous float64 tensor of npts entries")
        if (WHERE_DEVICE if w.is_cuda else WHERE_HOST) != where:
            raise ArgumentError("w must live where the other arrays do")
        return C.c_void_p(w.data_ptr()), w
    if where != WHERE_HOST:
        raise ArgumentError("w must live where the other arrays do")
    wa = np.ascontiguousarray(w, np.float64)
    if wa.size != M:
        raise DimensionMismatch(f"weights have {wa.size} entries, field has {M}")
    return C.c_void_p(wa.ctypes.data), wa


def _is_f32(x) -> bool:
    if _is_torch(x):
        import torch
        return x.dtype == torch.float32
    return np.asarray(x).dtype == np.float32


class Plan:
    """Plan(theta, phi; lmax, tol=1e-10, ntrans=1, method) — src/Plan.jl:84-116.

    As in the reference, the element type follows the nodes (T = float(eltype(theta))): float32
    theta and phi make a Float32 plan, whose `@` products return complex64 (gpu/runtests.jl:122-139).
    Its default precision is "mixed" (FP64 recurrence, FP32 grids) at tol >= 1e-6, else "fp64". Any
    plan applies to complex128 arrays, and to complex64 ones through pfsht_synth_f32 /
    pfsht_adjoint_f32."""

    def __init__(self, theta, phi, lmax: int, tol: float = 1e-10, ntrans: int = 1,
                 method: str = "auto", precision: str | None = None, spin: int = 0, maxboxes: int = 8,
                 oversample_theta: float = 1.0, oversample_phi: float = 1.0, epsapo_max: float = 1.0,
                 packed: bool = False, tune: int = 0, grid_fft: str = "auto"):
        self.float32 = _is_f32(theta) and _is_f32(phi)
        if precision is None:
            precision = "mixed" if self.float32 and tol >= 1e-6 else "fp64"
        if _is_torch(theta):
            theta = theta.detach().cpu().numpy()
        if _is_torch(phi):
            phi = phi.detach().cpu().numpy()
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
        ph = np.ascontiguousarray(np.asarray(phi, dtype=np.float64))
        if th.shape != ph.shape or th.ndim != 1:
            raise DimensionMismatch("theta and phi must have the same length")
        if lmax < 0:
            raise ArgumentError("lmax must be nonnegative")
        if ntrans < 1:
            raise ArgumentError("ntrans must be >= 1")
        if spin != 0 and abs(spin) > lmax:
            raise ArgumentError(f"(|spin| = {abs(spin)}) > lmax = {lmax}")
        self.lmax, self.ntrans, self.tol, self.npts = int(lmax), int(ntrans), float(tol), len(th)
        self.spin = int(spin)
        if precision not in ("fp64", "mixed"):
            raise ArgumentError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        self.precision = precision
        if grid_fft not in GRID_FFT:
            raise ArgumentError(f"grid_fft must be one of {sorted(GRID_FFT)}, got {grid_fft!r}")
        h = C.c_void_p()
        o = PlanOptsC(method=METHODS[method], precision={"fp64": 0, "mixed": 1}[precision], spin=self.spin,
                      maxboxes=int(maxboxes), oversample_theta=float(oversample_theta),
                      oversample_phi=float(oversample_phi), epsapo_max=float(epsapo_max), packed=int(bool(packed)),
                      tune=int(tune), grid_fft=GRID_FFT[grid_fft])
        _check(lib().pfsht_plan_create_opts(th.ctypes.data_as(C.c_void_p), ph.ctypes.data_as(C.c_void_p),
                                            len(th), self.lmax, self.tol, self.ntrans, C.byref(o), C.byref(h)))
        self._h = h
        inf = InfoC()
        _check(lib().pfsht_plan_info(self._h, C.byref(inf)))
        self.info = {f: getattr(inf, f) for f, _ in InfoC._fields_}
        self.info["kind_name"] = KINDS.get(inf.kind, "direct")
        self.theta, self.phi = th, ph

    def close(self):
        if getattr(self, "_h", None):
            lib().pfsht_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        alg = {"direct": "DirectAlgorithm", "stretch": "StretchAlgorithm",
               "multibox": "MultiBoxAlgorithm"}[self.algorithm]
        return (f"pfsht.Plan(lmax={self.lmax}, npts={self.npts}, ntrans={self.ntrans}, spin={self.spin}, "
                f"algorithm={alg}, kind={self.info['kind_name']})")

    def tune_table(self):
        """The ring-pass shapes timed when the plan was built (empty when its shape came from the
        cache or the default): dicts with seg_blocks, adj_tiles, mfma_tiles and ms."""
        rows = (RingShapeC * 16)()
        n = C.c_int32()
        _check(lib().pfsht_plan_tune_table(self._h, rows, 16, C.byref(n)))
        return [{k: getattr(rows[i], k) for k in ("seg_blocks", "adj_tiles", "mfma_tiles", "ms")} for i in range(n.value)]

    def tune_key(self):
        """(arch, lmax, nrings, mmax, ntrans): the key the plan's ring shape is cached under."""
        arch = C.create_string_buffer(128)
        v = [C.c_int32() for _ in range(4)]
        _check(lib().pfsht_plan_tune_key(self._h, arch, 128, *[C.byref(x) for x in v]))
        return (arch.value.decode(), *[x.value for x in v])

    @property
    def algorithm(self) -> str:
        """algorithm(plan) — src/Plan.jl:32-36: "direct", "stretch" or "multibox"."""
        return {1: "direct", 2: "stretch", 3: "multibox"}[self.info["method"]]

    @property
    def H(self) -> "AdjointPlan":
        """plan' — src/Plan.jl:172."""
        return AdjointPlan(self)

    @property
    def ring_work(self) -> int:
        return int(self.info["ring_work"])

    def synth(self, f, alm, stream=None):
        """mul!(f, plan, alm): overwrites f. complex64 alm and f take the ComplexF32 path."""
        L, B, M = self.lmax, self.ntrans, self.npts
        _assert_coeffs(alm, L, B)
        _assert_field(f, M, B)
        c64 = _is_c64(alm)
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm", c64)
        pf, wf = _buf(f, M * B, "f", c64)
        if wa != wf:
            raise ArgumentError("alm and f must both be host or both be device arrays")
        fn = lib().pfsht_synth_f32 if c64 else lib().pfsht_synth
        _check(fn(self._h, pa, pf, wa, _stream_ptr(stream) if wa else C.c_void_p(0)))
        return f

    def adjoint(self, alm, f, stream=None):
        """mul!(alm, plan', f): overwrites alm, padding included. complex64 arrays take the
        ComplexF32 path."""
        L, B, M = self.lmax, self.ntrans, self.npts
        _assert_coeffs(alm, L, B)
        _assert_field(f, M, B)
        c64 = _is_c64(f)
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm", c64)
        pf, wf = _buf(f, M * B, "f", c64)
        if wa != wf:
            raise ArgumentError("alm and f must both be host or both be device arrays")
        fn = lib().pfsht_adjoint_f32 if c64 else lib().pfsht_adjoint
        _check(fn(self._h, pf, pa, wa, _stream_ptr(stream) if wa else C.c_void_p(0)))
        return alm

    def _cols(self, ins, outs, nin: int, nout: int, stream):
        nb = len(ins)
        if len(outs) != nb:
            raise DimensionMismatch(f"{nb} input columns but {len(outs)} output columns")
        ptrs_in = (C.c_void_p * max(nb, 1))()
        ptrs_out = (C.c_void_p * max(nb, 1))()
        for b, (x, y) in enumerate(zip(ins, outs)):
            px, wx = _buf(x, nin, "input column")
            py, wy = _buf(y, nout, "output column")
            if wx != WHERE_DEVICE or wy != WHERE_DEVICE:
                raise ArgumentError("column applies take device (CUDA tensor) columns")
            ptrs_in[b], ptrs_out[b] = px, py
        return nb, ptrs_in, ptrs_out

    def synth_cols(self, f_cols, alm_cols, stream=None):
        """mul!(f, plan, alm) over len(alm_cols) <= ntrans transforms whose columns are separate device
        tensors (each (lmax+1)(2lmax+1) coefficients in; npts values out)."""
        L, M = self.lmax, self.npts
        nb, pi, po = self._cols(alm_cols, f_cols, (L + 1) * (2 * L + 1), M, stream)
        _check(lib().pfsht_synth_cols(self._h, nb, pi, po, _stream_ptr(stream)))
        return f_cols

    def adjoint_cols(self, alm_cols, f_cols, stream=None):
        """mul!(alm, plan', f) over len(f_cols) <= ntrans transforms in separate device tensors."""
        L, M = self.lmax, self.npts
        nb, pi, po = self._cols(f_cols, alm_cols, M, (L + 1) * (2 * L + 1), stream)
        _check(lib().pfsht_adjoint_cols(self._h, nb, pi, po, _stream_ptr(stream)))
        return alm_cols

    def __matmul__(self, alm):
        """plan * alm (src/Plan.jl:190-193): a new field, host or device like alm."""
        B = self.ntrans
        c64 = _is_c64(alm)
        if _is_torch(alm):
            import torch
            f = torch.empty(self.npts * B, dtype=torch.complex64 if c64 else torch.complex128, device=alm.device)
            return self.synth(f, alm)
        f = np.zeros((self.npts,) if B == 1 else (self.npts, B), np.complex64 if c64 else np.complex128, order="F")
        return self.synth(f, alm)

    def set_timing(self, on: bool = True):
        _check(lib().pfsht_set_timing(self._h, 1 if on else 0))

    def stage_times(self, direction: str = "synth", reset: bool = False):
        ms = np.zeros(NSTAGES)
        cnt = np.zeros(NSTAGES, np.int64)
        _check(lib().pfsht_stage_times(self._h, 0 if direction == "synth" else 1,
                                       ms.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p),
                                       1 if reset else 0))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(STAGES)}

    def nufft(self, typ: int, out, inp, stream=None):
        """_nufft_exec!(nufft_typ, inp, out) on the plan's stretched points (src/NUFFT.jl:17-25):
        type 2 (iflag +1) modes -> f, type 1 (iflag -1) f -> modes; modes are nth x nph x ntrans,
        column-major, FFT order."""
        nm = self.info["nth"] * self.info["nph"] * self.ntrans
        nf = self.npts * self.ntrans
        pi, wi = _buf(inp, nm if typ == 2 else nf, "in")
        po, wo = _buf(out, nf if typ == 2 else nm, "out")
        if wi != wo:
            raise ArgumentError("in and out must both be host or both be device arrays")
        _check(lib().pfsht_plan_nufft(self._h, int(typ), pi, po, wi, _stream_ptr(stream) if wi else C.c_void_p(0)))
        return out

    # -- analysis (src/Analysis.jl) ---------------------------------------------------------------
    def _new_alm(self, like):
        L, B = self.lmax, self.ntrans
        if _is_torch(like):
            import torch
            return torch.zeros((L + 1) * (2 * L + 1) * B, dtype=torch.complex128, device=like.device)
        return np.zeros((L + 1, 2 * L + 1) if B == 1 else (L + 1, 2 * L + 1, B), np.complex128, order="F")

    def apply_resolution(self, alm, w=None, out=None, stream=None):
        """apply_resolution!(out, plan, alm, w) (src/Analysis.jl:96-104): out = A' (w .* (A alm))."""
        L, B = self.lmax, self.ntrans
        _assert_coeffs(alm, L, B)
        out = self._new_alm(alm) if out is None else out
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm")
        po, wo = _buf(out, (L + 1) * (2 * L + 1) * B, "out")
        if wa != wo:
            raise ArgumentError("alm and out must both be host or both be device arrays")
        pw, keep = _weights(w, self.npts, wa)
        _check(lib().pfsht_apply_resolution(self._h, pa, pw, po, wa, _stream_ptr(stream) if wa else C.c_void_p(0)))
        return out

    def pseudo_spectrum(self, f, w=None, batch: int = 1, stream=None):
        """pseudo_spectrum!(C, plan, f, w; batch) (src/Analysis.jl:286-308): C_l of A' (w .* f)."""
        M, B, L = self.npts, self.ntrans, self.lmax
        _assert_field(f, M, B)
        pf, wf = _buf(f, M * B, "f")
        pw, keep = _weights(w, M, wf)
        if wf == WHERE_DEVICE:
            import torch
            Cl = torch.zeros(L + 1, dtype=torch.float64, device=f.device)
            pc = C.c_void_p(Cl.data_ptr())
        else:
            Cl = np.zeros(L + 1)
            pc = C.c_void_p(Cl.ctypes.data)
        _check(lib().pfsht_pseudo_spectrum(self._h, pf, pw, int(batch), pc, wf,
                                           _stream_ptr(stream) if wf else C.c_void_p(0)))
        return Cl

    def solve(self, f, mode: str = "cg", rtol: float = 1e-6, maxiter: int = 500, prior=None, w=None,
              alm=None, stream=None):
        """solve!(alm, plan, f; mode, rtol, maxiter, prior, w) (src/Analysis.jl:375-436).
        Returns (alm, info) with info = {shannon, dl, iters, relres, mode}."""
        if mode not in SOLVE_MODES:
            raise ArgumentError(f"mode must be cg, wiener or minnorm, got {mode}")
        M, B, L = self.npts, self.ntrans, self.lmax
        _assert_field(f, M, B)
        pf, wf = _buf(f, M * B, "f")
        alm = self._new_alm(f) if alm is None else alm
        _assert_coeffs(alm, L, B)
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm")
        if wa != wf:
            raise ArgumentError("f and alm must both be host or both be device arrays")
        pw, keep = _weights(w, M, wf)
        pp = C.c_void_p(0)
        if prior is not None:
            pr = np.ascontiguousarray(prior, np.float64)
            if pr.size < L + 1:
                raise DimensionMismatch(f"prior must cover l = 0:{L}, got {pr.size} entries")
            pp = C.c_void_p(pr.ctypes.data)
        inf = SolveInfoC()
        _check(lib().pfsht_solve(self._h, pf, pw, SOLVE_MODES[mode], float(rtol), int(maxiter), pp, pa, wf,
                                 _stream_ptr(stream) if wf else C.c_void_p(0), C.byref(inf)))
        info = {"shannon": inf.shannon, "dl": inf.dl, "iters": int(inf.iters), "relres": inf.relres, "mode": mode}
        return alm, info

    def rings(self) -> "Rings":
        h = C.c_void_p()
        _check(lib().pfsht_plan_rings(self._h, C.byref(h)))
        return Rings._borrow(h, self.lmax, self.info["nth_half"], self.ntrans)


class AdjointPlan:
    """AdjointPlan — src/Plan.jl:168-173."""

    def __init__(self, parent: Plan):
        self.parent = parent

    @property
    def H(self) -> Plan:
        return self.parent

    def __matmul__(self, f):
        p = self.parent
        if _is_torch(f):
            import torch
            alm = torch.zeros((p.lmax + 1) * (2 * p.lmax + 1) * p.ntrans,
                              dtype=torch.complex64 if _is_c64(f) else torch.complex128, device=f.device)
            return p.adjoint(alm, f)
        alm = zeros_alm(p.lmax, p.ntrans)
        if _is_c64(f):
            alm = alm.astype(np.complex64, order="F")
        return p.adjoint(alm, f)


class ServerStatsC(C.Structure):
    _fields_ = [("hits", C.c_int64), ("misses", C.c_int64), ("evictions", C.c_int64),
                ("rejected", C.c_int64), ("submitted", C.c_int64), ("resident_bytes", C.c_int64),
                ("budget_bytes", C.c_int64), ("plans", C.c_int64), ("launches", C.c_int64),
                ("packed_transforms", C.c_int64), ("graph_launches", C.c_int64), ("graph_failures", C.c_int64),
                ("failed", C.c_int64), ("queued", C.c_int64)]


class ServerOptsC(C.Structure):
    _fields_ = [("nstreams", C.c_int32), ("depth", C.c_int32), ("pack", C.c_int32), ("graphs", C.c_int32),
                ("hbm_budget_bytes", C.c_int64)]


class Server:
    """Serving runtime over pfsht_server_* (include/pfsht.h): plans cached in HBM under integer keys
    with an HBM budget and LRU eviction; requests queued and packed by a dispatcher thread into column
    applies of their plan (continuous batching), replayed from hipGraphs. Inputs and outputs are device
    (torch) tensors holding the key's ntrans transforms in the reference's layouts; they are kept
    alive here until their request has completed."""

    def __init__(self, nstreams: int = 4, hbm_budget_bytes: int = 200 * 10**9, depth: int = 2, pack: int = 8,
                 graphs: bool = True):
        o = ServerOptsC(nstreams=int(nstreams), depth=int(depth), pack=int(pack), graphs=int(bool(graphs)),
                        hbm_budget_bytes=int(hbm_budget_bytes))
        h = C.c_void_p()
        _check(lib().pfsht_server_create_ex(C.byref(o), C.byref(h)))
        self._h = h
        self._shape = {}   # key -> (lmax, npts, ntrans)
        self._live = {}    # ticket -> (input, output): kept alive until the request completes

    def close(self):
        if getattr(self, "_h", None):
            lib().pfsht_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, key: int, theta, phi, lmax: int, tol: float = 1e-10, ntrans: int = 1,
             method: str = "auto"):
        """The plan cached under key (built on a miss). Returns (info dict, hit)."""
        th = np.ascontiguousarray(theta, np.float64)
        ph = np.ascontiguousarray(phi, np.float64)
        if th.shape != ph.shape or th.ndim != 1:
            raise DimensionMismatch("theta and phi must have the same length")
        hp = C.c_void_p()
        hit = C.c_int32()
        _check(lib().pfsht_server_plan(self._h, C.c_uint64(key), th.ctypes.data, ph.ctypes.data, len(th),
                                       int(lmax), float(tol), int(ntrans), METHODS[method], C.byref(hp),
                                       C.byref(hit)))
        inf = InfoC()
        _check(lib().pfsht_plan_info(hp, C.byref(inf)))
        self._shape[int(key)] = (int(lmax), len(th), int(ntrans))
        return {k: getattr(inf, k) for k, _ in InfoC._fields_}, bool(hit.value)

    def submit(self, key: int, direction: str, inp, out) -> int:
        """Queue mul!(out, plan, inp) ("synth") or mul!(out, plan', inp) ("adjoint"); returns a ticket."""
        if direction not in ("synth", "adjoint"):
            raise ArgumentError(f"direction must be 'synth' or 'adjoint', got {direction!r}")
        if int(key) not in self._shape:
            raise ArgumentError(f"no plan was registered under key {key}")
        L, M, B = self._shape[int(key)]
        na, nf = (L + 1) * (2 * L + 1) * B, M * B
        pi_, wi = _buf(inp, na if direction == "synth" else nf, "input")
        po, wo = _buf(out, nf if direction == "synth" else na, "output")
        if wi != WHERE_DEVICE or wo != WHERE_DEVICE:
            raise ArgumentError("server requests take device (CUDA) tensors")
        t = C.c_int64()
        _check(lib().pfsht_server_submit(self._h, C.c_uint64(key), 0 if direction == "synth" else 1, pi_, po,
                                         C.byref(t)))
        self._live[t.value] = (inp, out)
        if len(self._live) > 4096:
            self._prune()
        return t.value

    def _prune(self):
        for t in list(self._live):
            if self.query(t)[0]:
                self._live.pop(t, None)

    def query(self, ticket: int):
        """(done, device_ms) of a ticket, without waiting."""
        done = C.c_int32()
        ms = C.c_double(-1.0)
        _check(lib().pfsht_server_query(self._h, int(ticket), C.byref(done), C.byref(ms)))
        if done.value:
            self._live.pop(int(ticket), None)
        return bool(done.value), ms.value

    def wait(self, ticket: int) -> float:
        ms = C.c_double(-1.0)
        _check(lib().pfsht_server_wait(self._h, int(ticket), C.byref(ms)))
        self._live.pop(int(ticket), None)
        return ms.value

    def epoch(self):
        """Record the device-clock reference of times() and wait for it."""
        _check(lib().pfsht_server_epoch(self._h))

    def times(self, ticket: int):
        """(start_ms, end_ms) of a ticket's launch after the epoch, on the device clock (waits for it)."""
        a, b = C.c_double(), C.c_double()
        _check(lib().pfsht_server_times(self._h, int(ticket), C.byref(a), C.byref(b)))
        self._live.pop(int(ticket), None)
        return a.value, b.value

    def stats(self) -> dict:
        st = ServerStatsC()
        _check(lib().pfsht_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStatsC._fields_}


class Group:
    """Many single-transform plans applied as one (pfsht_group_*): their ring passes in one launch
    over a (plan, order, ring tile) work list, their other stages concurrently, each direction replayed
    from a hipGraph. synth(fs, alms) / adjoint(alms, fs) take one device tensor per member."""

    def __init__(self, plans):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p._h for p in self.plans])
        h = C.c_void_p()
        _check(lib().pfsht_group_create(arr, len(self.plans), C.byref(h)))
        self._h = h
        self._live = []

    def close(self):
        if getattr(self, "_h", None):
            lib().pfsht_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _tables(self, ins, outs, synth: bool):
        n = len(self.plans)
        if len(ins) != n or len(outs) != n:
            raise DimensionMismatch(f"the group has {n} plans: pass {n} inputs and {n} outputs")
        pi = (C.c_void_p * n)()
        po = (C.c_void_p * n)()
        for i, (p, x, y) in enumerate(zip(self.plans, ins, outs)):
            na, nf = (p.lmax + 1) * (2 * p.lmax + 1), p.npts
            px, wx = _buf(x, na if synth else nf, "input")
            py, wy = _buf(y, nf if synth else na, "output")
            if wx != WHERE_DEVICE or wy != WHERE_DEVICE:
                raise ArgumentError("group applies take device (CUDA) tensors")
            pi[i], po[i] = px, py
        self._live = [ins, outs]  # kept alive while the call may still run
        return pi, po

    def synth(self, fs, alms, stream=None):
        pi, po = self._tables(alms, fs, True)
        _check(lib().pfsht_group_synth(self._h, pi, po, _stream_ptr(stream)))
        return fs

    def adjoint(self, alms, fs, stream=None):
        pi, po = self._tables(fs, alms, False)
        _check(lib().pfsht_group_adjoint(self._h, pi, po, _stream_ptr(stream)))
        return alms


def mul(out, op, x, stream=None):
    """mul!(out, op, x) for op a Plan or an AdjointPlan."""
    if isinstance(op, Plan):
        return op.synth(out, x, stream)
    if isinstance(op, AdjointPlan):
        return op.parent.adjoint(out, x, stream)
    raise ArgumentError("op must be a Plan or plan.H")


class Rings:
    """RingLegendre(theta, lmax, eps; ntrans, reflect) — src/Legendre.jl:194-255; with spin != 0,
    SpinRings(theta, lmax, spin, eps; ntrans, reflect) — src/Spin.jl:192-270 (coefficients in the
    (l, m + lmax) layout)."""

    def __init__(self, theta, lmax: int, eps: float, ntrans: int = 1, reflect: bool = False, spin: int = 0):
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
        h = C.c_void_p()
        if spin:
            _check(lib().pfsht_spin_rings_create(th.ctypes.data_as(C.c_void_p), len(th), int(lmax), int(spin),
                                                 float(eps), int(ntrans), 1 if reflect else 0, C.byref(h)))
        else:
            _check(lib().pfsht_rings_create(th.ctypes.data_as(C.c_void_p), len(th), int(lmax), float(eps),
                                            int(ntrans), 1 if reflect else 0, C.byref(h)))
        self._h, self._owned = h, True
        self.lmax, self.n, self.ntrans = int(lmax), len(th), int(ntrans)

    @classmethod
    def _borrow(cls, h, lmax, n, ntrans):
        r = cls.__new__(cls)
        r._h, r._owned = h, False
        r.lmax, r.n, r.ntrans = int(lmax), int(n), int(ntrans)
        return r

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "_h", None):
            lib().pfsht_rings_destroy(self._h)
            self._h = None

    @property
    def work(self) -> int:
        return int(lib().pfsht_rings_work(self._h))

    @property
    def mmax(self) -> int:
        return int(lib().pfsht_rings_mmax(self._h))

    def rings_from_alm(self, G, alm, stream=None):
        L, B = self.lmax, self.ntrans
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm")
        pg, wg = _buf(G, self.n * (2 * L + 1) * B, "G")
        if wa != wg:
            raise ArgumentError("alm and G must both be host or both be device arrays")
        _check(lib().pfsht_rings_from_alm(self._h, pa, pg, wa, _stream_ptr(stream) if wa else C.c_void_p(0)))
        return G

    def alm_from_rings(self, alm, G, stream=None):
        L, B = self.lmax, self.ntrans
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm")
        pg, wg = _buf(G, self.n * (2 * L + 1) * B, "G")
        if wa != wg:
            raise ArgumentError("alm and G must both be host or both be device arrays")
        _check(lib().pfsht_alm_from_rings(self._h, pg, pa, wa, _stream_ptr(stream) if wa else C.c_void_p(0)))
        return alm

    def alm_from_rings_noalloc(self, alm, G, stream=None):
        """alm_from_rings! on device tensors with the library's static scratch (allocates nothing)."""
        L, B = self.lmax, self.ntrans
        pa, wa = _buf(alm, (L + 1) * (2 * L + 1) * B, "alm")
        pg, wg = _buf(G, self.n * (2 * L + 1) * B, "G")
        if wa != WHERE_DEVICE or wg != WHERE_DEVICE:
            raise ArgumentError("the allocation-free adjoint takes device (CUDA) tensors")
        _check(lib().pfsht_alm_from_rings_noalloc(self._h, pg, pa, _stream_ptr(stream)))
        return alm


# --- device helpers: torch tensors holding column-major data ------------------------------------
def to_device(a: np.ndarray, device="cuda"):
    """Copy a column-major numpy array into a flat contiguous torch CUDA tensor (same memory order)."""
    import torch
    flat = np.asarray(a, dtype=np.complex128).ravel(order="F")
    return torch.from_numpy(flat.copy()).to(device)


def from_device(t, shape) -> np.ndarray:
    """Flat device tensor -> column-major numpy array of `shape`."""
    return t.detach().cpu().numpy().reshape(shape, order="F")
