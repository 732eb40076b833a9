"""ctypes front end of the CPU oracle (oracle/pfsht_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the reference results are compared against. The product library never
imports this module.

Arrays follow the reference's column-major layouts (Layout.jl:13-20; Stretch.jl:215-220):
numpy arrays here are Fortran-ordered complex128, alm of shape (L+1, 2L+1[, B]) and f of shape
(M[, B]).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libpfsht_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class GeomC(C.Structure):
    _fields_ = [("kind", C.c_int), ("lmax", C.c_int), ("K", C.c_int),
                ("eps", C.c_double), ("beta", C.c_double), ("h", C.c_double),
                ("thlo", C.c_double), ("thhi", C.c_double), ("epsapo", C.c_double),
                ("period", C.c_double), ("Theta1", C.c_double), ("sigma", C.c_double),
                ("ThetaT", C.c_double), ("invZ", C.c_double)]


class PlanInfoC(C.Structure):
    _fields_ = [("kind", C.c_int), ("flip", C.c_int), ("lmax", C.c_int), ("ntrans", C.c_int),
                ("method", C.c_int), ("nth_half", C.c_int), ("nth", C.c_int), ("nph", C.c_int),
                ("mmax", C.c_int), ("ring_work", C.c_int64), ("npts", C.c_int64),
                ("nufft_kind", C.c_int), ("es_w", C.c_int), ("es_n1", C.c_int),
                ("es_n2", C.c_int), ("es_beta", C.c_double)]


KIND_NAMES = {0: "pole", 1: "interior", 2: "full", -1: "direct"}

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_I64 = C.c_int64


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "or_last_error": (C.c_char_p, []),
            "or_sph_mode": (_I64, [_I, _I, _I]),
            "or_polar_pad_factor": (_D, []),
            "or_polar_mmax": (_I, [_D, _I, _D]),
            "or_cost_stretch": (_D, [_D, _D, _I64, _I, _I, _D]),
            "or_cost_direct": (_D, [_I64, _I, _I]),
            "or_sectorial_log2c": (None, [_I, _P]),
            "or_nested_ranges": (_I, [_P, _I, _I, _P, _P]),
            "or_retained_triples": (_I64, [_P, _I, _I]),
            "or_alf": (_D, [_I, _I, _D]),
            "or_rings_create": (_P, [_P, _I, _I, _D, _I, _I]),
            "or_rings_destroy": (None, [_P]),
            "or_rings_work": (_I64, [_P]),
            "or_rings_mmax": (_I, [_P]),
            "or_rings_tables": (None, [_P, _P, _P, _P, _P, _P]),
            "or_rings_from_alm": (_I, [_P, _P, _P, _I]),
            "or_alm_from_rings": (_I, [_P, _P, _P, _I]),
            "or_stretch_geometry": (_I, [_D, _D, _I, _D, _D, C.POINTER(GeomC)]),
            "or_es_cdf": (_D, [_D, _D]),
            "or_stretch_window": (_D, [C.POINTER(GeomC), _D]),
            "or_theta_to_stretch": (_D, [C.POINTER(GeomC), _D]),
            "or_stretch_to_theta": (_D, [C.POINTER(GeomC), _D]),
            "or_next_smooth": (_I64, [_I64]),
            "or_next_smooth_even": (_I64, [_I64]),
            "or_es_fine_size": (_I, [_I, _I]),
            "or_dfs_double": (None, [_P, _P, _I, _I, _I, _I]),
            "or_dfs_fold": (None, [_P, _P, _I, _I, _I, _I]),
            "or_fft1": (None, [_P, _I, _I]),
            "or_fft2": (None, [_P, _I, _I, _I, _I, _I]),
            "or_nufft_direct": (_I, [_I, _I, _I, _I, _I, _P, _P, _I64, _P, _P]),
            "or_nufft_es": (_I, [_I, _I, _I, _I, _I, _P, _P, _I64, _D, _P, _P, _I]),
            "or_es_params": (None, [_D, _P, _P]),
            "or_plan_create": (_P, [_P, _P, _I64, _I, _D, _I, _I, _I, _I]),
            "or_plan_destroy": (None, [_P]),
            "or_plan_get_info": (None, [_P, C.POINTER(PlanInfoC)]),
            "or_synth": (_I, [_P, _P, _P]),
            "or_adjoint": (_I, [_P, _P, _P]),
            "or_plan_last_G": (_I, [_P, _P]),
            "or_plan_last_Fhat": (_I, [_P, _P]),
            "or_plan_last_times": (None, [_P, _P]),
            "or_plan_points": (_I, [_P, _P, _P]),
            "or_wigner_d": (_D, [_I, _I, _I, _D]),
            "or_sylm": (None, [_I, _I, _I, _D, _D, _P, _P]),
            "or_spin_rings_create": (_P, [_P, _I, _I, _I, _D, _I, _I]),
            "or_spin_rings_destroy": (None, [_P]),
            "or_spin_rings_mmax": (_I, [_P]),
            "or_spin_rings_work": (_I64, [_P]),
            "or_spin_rings_from_alm": (_I, [_P, _P, _P, _I]),
            "or_spin_alm_from_rings": (_I, [_P, _P, _P, _I]),
            "or_plan_create_spin": (_P, [_P, _P, _I64, _I, _D, _I, _I, _I, _I, _I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _err() -> str:
    return lib().or_last_error().decode()


def _c128(a, shape=None) -> np.ndarray:
    a = np.asarray(a, dtype=np.complex128)
    if shape is not None:
        a = a.reshape(shape, order="F")
    return np.asfortranarray(a)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# --- Layout.jl -------------------------------------------------------------------------------
def coeff_size(L: int):
    return (L + 1, 2 * L + 1)


def sph_mode(l: int, m: int):
    """0-based (row, col) of (l, m) in the (L+1) x (2L+1) array (Layout.jl:20)."""
    am = abs(m)
    return (l - am, 2 * am + (1 if m >= 0 else 0) - 1)


def spin_coeff_index(l: int, m: int, L: int):
    """0-based (row, col) of a spin coefficient in the dense (l, m + L) layout (Layout.jl:28-29)."""
    return (l, m + L)


def spin_valid_mask(L: int, s: int) -> np.ndarray:
    """Slots of the spin layout the transform reads and writes: |m| <= l and l >= |s|."""
    l = np.arange(L + 1)[:, None]
    m = np.arange(-L, L + 1)[None, :]
    return (np.abs(m) <= l) & (l >= abs(s))


# --- Spin.jl ---------------------------------------------------------------------------------
def wigner_d(l: int, m: int, n: int, beta: float) -> float:
    """wigner_d(l, m, n, beta) (Spin.jl:61-137)."""
    return lib().or_wigner_d(int(l), int(m), int(n), float(beta))


def sylm(s: int, l: int, m: int, theta: float, phi: float) -> complex:
    """sYlm(s, l, m, theta, phi) = N_l d^l_{m,-s}(theta) e^{i m phi} (Spin.jl:144-149)."""
    re, im = C.c_double(), C.c_double()
    lib().or_sylm(int(s), int(l), int(m), float(theta), float(phi), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


class SpinRings:
    """SpinRings (Spin.jl:170-270) with spin_rings_from_alm!/spin_alm_from_rings! (:316-334);
    coefficients in the dense (l, m + L) layout."""

    def __init__(self, theta, L: int, spin: int, eps: float, ntrans: int = 1, reflect: bool = False):
        self.theta = _f64(theta)
        self.L, self.s, self.B, self.n = L, spin, ntrans, len(self.theta)
        self.h = lib().or_spin_rings_create(_ptr(self.theta), self.n, L, int(spin), float(eps), ntrans,
                                            int(reflect))
        if not self.h:
            raise ValueError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_spin_rings_destroy(self.h)
            self.h = None

    @property
    def work(self) -> int:
        return int(lib().or_spin_rings_work(self.h))

    @property
    def mmax(self) -> int:
        return int(lib().or_spin_rings_mmax(self.h))

    def rings_from_alm(self, alm, nthreads: int = 1) -> np.ndarray:
        L, B = self.L, self.B
        a = _c128(alm, (L + 1, 2 * L + 1, B))
        G = np.zeros((self.n, 2 * L + 1, B), np.complex128, order="F")
        lib().or_spin_rings_from_alm(self.h, _ptr(a), _ptr(G), nthreads)
        return G

    def alm_from_rings(self, G, nthreads: int = 1) -> np.ndarray:
        L, B = self.L, self.B
        g = _c128(G, (self.n, 2 * L + 1, B))
        a = np.zeros((L + 1, 2 * L + 1, B), np.complex128, order="F")
        lib().or_spin_alm_from_rings(self.h, _ptr(g), _ptr(a), nthreads)
        return a[:, :, 0] if B == 1 else a


# --- Legendre.jl -----------------------------------------------------------------------------
def polar_mmax(theta: float, L: int, eps: float) -> int:
    return lib().or_polar_mmax(float(theta), int(L), float(eps))


def sectorial_log2c(L: int) -> np.ndarray:
    out = np.zeros(L + 1)
    lib().or_sectorial_log2c(L, _ptr(out))
    return out


def nested_ranges(mlim, L: int):
    ml = np.ascontiguousarray(mlim, dtype=np.int32)
    jlo = np.zeros(L + 1, np.int32)
    jhi = np.zeros(L + 1, np.int32)
    if lib().or_nested_ranges(_ptr(ml), len(ml), L, _ptr(jlo), _ptr(jhi)) != 0:
        raise ValueError(_err())
    return jlo, jhi


def retained_triples(mlim, L: int) -> int:
    ml = np.ascontiguousarray(mlim, dtype=np.int32)
    return int(lib().or_retained_triples(_ptr(ml), len(ml), L))


def alf(l: int, m: int, theta: float) -> float:
    return lib().or_alf(int(l), int(m), float(theta))


class Rings:
    """RingLegendre (Legendre.jl:194-255) with rings_from_alm!/alm_from_rings!."""

    def __init__(self, theta, L: int, eps: float, ntrans: int = 1, reflect: bool = False):
        self.theta = _f64(theta)
        self.L, self.B, self.n = L, ntrans, len(self.theta)
        self.h = lib().or_rings_create(_ptr(self.theta), self.n, L, float(eps), ntrans, int(reflect))
        if not self.h:
            raise ValueError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_rings_destroy(self.h)
            self.h = None

    @property
    def work(self) -> int:
        return int(lib().or_rings_work(self.h))

    @property
    def mmax(self) -> int:
        return int(lib().or_rings_mmax(self.h))

    def tables(self):
        n, L = self.n, self.L
        cs, l2s = np.zeros(n), np.zeros(n)
        ml = np.zeros(n, np.int32)
        jlo, jhi = np.zeros(L + 1, np.int32), np.zeros(L + 1, np.int32)
        lib().or_rings_tables(self.h, _ptr(cs), _ptr(l2s), _ptr(ml), _ptr(jlo), _ptr(jhi))
        return dict(costh=cs, log2sin=l2s, mlim=ml, jlo=jlo, jhi=jhi)

    def rings_from_alm(self, alm, nthreads: int = 1) -> np.ndarray:
        L, B = self.L, self.B
        a = _c128(alm, (L + 1, 2 * L + 1, B))
        G = np.zeros((self.n, 2 * L + 1, B), np.complex128, order="F")
        lib().or_rings_from_alm(self.h, _ptr(a), _ptr(G), nthreads)
        return G

    def alm_from_rings(self, G, nthreads: int = 1) -> np.ndarray:
        L, B = self.L, self.B
        g = _c128(G, (self.n, 2 * L + 1, B))
        a = np.zeros((L + 1, 2 * L + 1, B), np.complex128, order="F")
        lib().or_alm_from_rings(self.h, _ptr(g), _ptr(a), nthreads)
        return a


# --- Windows.jl / Stretch.jl -----------------------------------------------------------------
def stretch_geometry(thlo, thhi, L, eps, epsapo_max=1.0):
    """Returns a GeomC, or None when degenerate; raises ValueError on a bad argument."""
    g = GeomC()
    rc = lib().or_stretch_geometry(float(thlo), float(thhi), int(L), float(eps),
                                   float(epsapo_max), C.byref(g))
    if rc < 0:
        raise ValueError(_err())
    return None if rc == 1 else g


def es_beta(eps: float) -> float:
    return abs(np.log(eps))


def es_cdf(beta, u):
    return lib().or_es_cdf(float(beta), float(u))


def stretch_window(g, Theta):
    return lib().or_stretch_window(C.byref(g), float(Theta))


def theta_to_stretch(g, theta):
    return lib().or_theta_to_stretch(C.byref(g), float(theta))


def stretch_to_theta(g, Theta):
    return lib().or_stretch_to_theta(C.byref(g), float(Theta))


def next_smooth(n):
    return int(lib().or_next_smooth(int(n)))


def next_smooth_even(n):
    return int(lib().or_next_smooth_even(int(n)))


def es_fine_size(n, w):
    """Fine-grid length of the ES NUFFT for a length-n axis at kernel width w."""
    return int(lib().or_es_fine_size(int(n), int(w)))


# --- DFS.jl, FFT, NUFFT.jl -------------------------------------------------------------------
def dfs_double(F, s=0):
    F = _c128(F)
    if F.ndim == 2:
        F = F[:, :, None]
    nth, nph, B = F.shape
    F = np.asfortranarray(F)
    Ft = np.zeros((2 * nth, nph, B), np.complex128, order="F")
    lib().or_dfs_double(_ptr(F), _ptr(Ft), nth, nph, B, s)
    return Ft


def dfs_fold(Ft, s=0):
    Ft = _c128(Ft)
    if Ft.ndim == 2:
        Ft = Ft[:, :, None]
    n2, nph, B = Ft.shape
    Ft = np.asfortranarray(Ft)
    F = np.zeros((n2 // 2, nph, B), np.complex128, order="F")
    lib().or_dfs_fold(_ptr(Ft), _ptr(F), n2 // 2, nph, B, s)
    return F


def fft1(x, sign):
    x = np.array(x, dtype=np.complex128)
    lib().or_fft1(_ptr(x), len(x), sign)
    return x


def fft2(x, sign, nthreads=1):
    x = np.array(x, dtype=np.complex128, order="F")
    if x.ndim == 2:
        x = x[:, :, None]
    x = np.asfortranarray(x)
    n1, n2, B = x.shape
    lib().or_fft2(_ptr(x), n1, n2, B, sign, nthreads)
    return x


def es_params(tol):
    w = C.c_int()
    b = C.c_double()
    lib().or_es_params(float(tol), C.byref(w), C.byref(b))
    return w.value, b.value


def nufft_direct(typ, n1, n2, iflag, x, y, inp, B=1):
    x, y = _f64(x), _f64(y)
    M = len(x)
    if typ == 2:
        a = _c128(inp, (n1, n2, B))
        out = np.zeros((M, B), np.complex128, order="F")
    else:
        a = _c128(inp, (M, B))
        out = np.zeros((n1, n2, B), np.complex128, order="F")
    if lib().or_nufft_direct(typ, n1, n2, iflag, B, _ptr(x), _ptr(y), M, _ptr(a), _ptr(out)) != 0:
        raise ValueError(_err())
    return out


def nufft_es(typ, n1, n2, iflag, x, y, inp, tol, B=1, nthreads=1):
    x, y = _f64(x), _f64(y)
    M = len(x)
    if typ == 2:
        a = _c128(inp, (n1, n2, B))
        out = np.zeros((M, B), np.complex128, order="F")
    else:
        a = _c128(inp, (M, B))
        out = np.zeros((n1, n2, B), np.complex128, order="F")
    if lib().or_nufft_es(typ, n1, n2, iflag, B, _ptr(x), _ptr(y), M, float(tol), _ptr(a),
                         _ptr(out), nthreads) != 0:
        raise ValueError(_err())
    return out


# --- Plan.jl ---------------------------------------------------------------------------------
METHODS = {"auto": 0, "direct": 1, "stretch": 2}
NUFFTS = {"es": 0, "direct": 1}


class Plan:
    """Plan(theta, phi; lmax, tol, ntrans, method) with synth (mul!(f, plan, alm)) and
    adjoint (mul!(alm, plan', f)) — Plan.jl:84-148."""

    def __init__(self, theta, phi, lmax, tol=1e-10, ntrans=1, method="auto", nufft="es",
                 nthreads=1, spin=0):
        self.theta, self.phi = _f64(theta), _f64(phi)
        if len(self.theta) != len(self.phi):
            raise ValueError("theta and phi must have the same length")
        self.L, self.B, self.M, self.tol = int(lmax), int(ntrans), len(self.theta), float(tol)
        self.spin = int(spin)
        self.h = lib().or_plan_create_spin(_ptr(self.theta), _ptr(self.phi), self.M, self.L,
                                           self.tol, self.B, self.spin, METHODS[method], NUFFTS[nufft],
                                           int(nthreads))
        if not self.h:
            raise ValueError(_err())
        inf = PlanInfoC()
        lib().or_plan_get_info(self.h, C.byref(inf))
        self.info = {f: getattr(inf, f) for f, _ in PlanInfoC._fields_}
        self.info["kind_name"] = KIND_NAMES.get(inf.kind, "?")
        self.method = "direct" if inf.method == 1 else "stretch"

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_plan_destroy(self.h)
            self.h = None

    def synth(self, alm) -> np.ndarray:
        L, B = self.L, self.B
        a = _c128(alm, (L + 1, 2 * L + 1, B))
        f = np.zeros((self.M, B), np.complex128, order="F")
        if lib().or_synth(self.h, _ptr(a), _ptr(f)) != 0:
            raise ValueError(_err())
        return f[:, 0] if B == 1 else f

    def adjoint(self, f) -> np.ndarray:
        L, B = self.L, self.B
        ff = _c128(f, (self.M, B))
        a = np.zeros((L + 1, 2 * L + 1, B), np.complex128, order="F")
        if lib().or_adjoint(self.h, _ptr(ff), _ptr(a)) != 0:
            raise ValueError(_err())
        return a[:, :, 0] if B == 1 else a

    def last_G(self):
        n = self.info["nth_half"] if self.method == "stretch" else self.M
        G = np.zeros((n, 2 * self.L + 1, self.B), np.complex128, order="F")
        lib().or_plan_last_G(self.h, _ptr(G))
        return G

    def last_Fhat(self):
        F = np.zeros((self.info["nth"], self.info["nph"], self.B), np.complex128, order="F")
        lib().or_plan_last_Fhat(self.h, _ptr(F))
        return F

    def points(self):
        T = np.zeros(self.M)
        p = np.zeros(self.M)
        if lib().or_plan_points(self.h, _ptr(T), _ptr(p)) != 0:
            raise ValueError("not a stretched plan")
        return T, p

    def last_times(self):
        t = np.zeros(3)
        lib().or_plan_last_times(self.h, _ptr(t))
        return t


def rand_alm(L: int, rng, ntrans: int = 1, spectrum=None) -> np.ndarray:
    """Random coefficients in the valid modes only, padding zero (test/runtests.jl:24-36).
    spectrum(l) scales the per-mode standard deviation; a_l0 is kept real when spectrum is
    given, as in the synthetic workloads."""
    a = np.zeros((L + 1, 2 * L + 1, ntrans), np.complex128, order="F")
    for b in range(ntrans):
        for l in range(L + 1):
            s = 1.0 if spectrum is None else spectrum(l)
            ms = np.arange(-l, l + 1)
            re = rng.standard_normal(len(ms)) * s
            im = rng.standard_normal(len(ms)) * s
            if spectrum is not None:
                im[ms == 0] = 0.0
            for m, r, i in zip(ms, re, im):
                rr, cc = sph_mode(l, int(m))
                a[rr, cc, b] = r + 1j * i
    return a[:, :, 0] if ntrans == 1 else a


def rand_alm_fast(L: int, rng, ntrans: int = 1, cl_power: float | None = None) -> np.ndarray:
    """Vectorised valid-mode fill. With cl_power p, Re/Im ~ N(0, C_l/2) with C_l = (l+1)^-p and
    a_l0 real; otherwise Re/Im ~ N(0, 1)."""
    rows = np.arange(L + 1)[:, None]
    cols = np.arange(2 * L + 1)[None, :]
    am = (cols + 1) // 2
    l = rows + am
    valid = l <= L
    out = np.zeros((L + 1, 2 * L + 1, ntrans), np.complex128, order="F")
    for b in range(ntrans):
        re = rng.standard_normal((L + 1, 2 * L + 1))
        im = rng.standard_normal((L + 1, 2 * L + 1))
        z = re + 1j * im
        if cl_power is not None:
            z = z * np.sqrt(0.5 * (l + 1.0) ** (-cl_power))
            z[:, 0] = z[:, 0].real
        z[~valid] = 0
        out[:, :, b] = z
    return out[:, :, 0] if ntrans == 1 else out


def valid_mask(L: int) -> np.ndarray:
    rows = np.arange(L + 1)[:, None]
    cols = np.arange(2 * L + 1)[None, :]
    return rows + (cols + 1) // 2 <= L


# ------------------------------------------------------------------------------------------------
# Bounding box and analysis (src/Box.jl, src/Analysis.jl), restated in numpy over an oracle Plan.
# ------------------------------------------------------------------------------------------------
def bounding_box(theta, phi) -> dict:
    """bounding_box (Box.jl:43-73) with the circular extent found by sorting (the reference's
    bucket scan, Box.jl:246-296, returns the same largest gap), box_fsky (:106-108) and
    spectral_resolution (:148-153)."""
    th = np.asarray(theta, np.float64)
    ph = np.asarray(phi, np.float64)
    thlo = max(float(th.min()), 0.0)
    thhi = min(float(th.max()), np.pi)
    ptol = np.sqrt(np.finfo(np.float64).eps)
    north, south = thlo <= ptol, thhi >= np.pi - ptol
    if north or south:
        start, ext = 0.0, 2 * np.pi
    else:
        ps = np.mod(ph, 2 * np.pi)
        if len(ps) == 1:
            start, ext = float(ps[0]), 0.0
        else:
            s = np.sort(ps)
            d = np.diff(s)
            k = int(np.argmax(d)) if len(d) else 0
            gap, at = (float(d[k]), float(s[k + 1])) if len(d) and d[k] > 0 else (0.0, float(s[-1]))
            wrap = 2 * np.pi - s[-1] + s[0]
            start = at
            if wrap >= gap:
                gap, start = wrap, float(s[0])
            ext = 2 * np.pi - gap
            if ext <= 0:
                start, ext = float(s[0]), 0.0
            elif ext >= 2 * np.pi:
                start, ext = 0.0, 2 * np.pi
    fsky = (np.cos(thlo) - np.cos(thhi)) * ext / (4 * np.pi)
    e = np.sqrt(np.finfo(np.float64).eps)
    mid = min(max(0.5 * (thlo + thhi), e), np.pi - e)
    lam = max(thhi - thlo, ext * np.sin(mid))
    dl = np.inf if lam <= 0 else np.pi / lam
    return dict(thlo=thlo, thhi=thhi, phistart=start, phiext=ext, touches_north=north, touches_south=south,
                fsky=fsky, dl=dl)


def _weight(f, w):
    return f if w is None else (f * (w if f.ndim == 1 else w[:, None]))


def apply_resolution(plan: "Plan", alm, w=None) -> np.ndarray:
    """apply_resolution! (Analysis.jl:96-104): A' (w .* (A alm))."""
    return plan.adjoint(np.asfortranarray(_weight(plan.synth(alm), w)))


def pseudo_spectrum(plan: "Plan", f, w=None, batch: int = 1) -> np.ndarray:
    """pseudo_spectrum! (Analysis.jl:286-308)."""
    a = plan.adjoint(np.asfortranarray(_weight(np.asarray(f), w)))
    if a.ndim == 3:
        a = a[:, :, batch - 1]
    L = plan.L
    C = np.zeros(L + 1)
    for l in range(L + 1):
        acc = abs(a[l, 0]) ** 2
        for m in range(1, l + 1):
            acc += abs(a[l - m, 2 * m]) ** 2 + abs(a[l - m, 2 * m - 1]) ** 2
        C[l] = acc / (2 * l + 1)
    return C


def solve(plan: "Plan", f, mode="cg", rtol=1e-6, maxiter=500, prior=None, w=None):
    """solve! (Analysis.jl:375-436): CG on the normal equations, optional Wiener prior."""
    L = plan.L
    fw = np.asfortranarray(_weight(np.asarray(f), w))
    rhs = plan.adjoint(fw)
    reg = None
    if mode == "wiener":
        reg = np.array([1.0 / c if c > 0 else 0.0 for c in prior[:L + 1]])
    tol = min(rtol, 1e-12) if mode == "minnorm" else rtol
    mask = valid_mask(L)
    lgrid = np.arange(L + 1)[:, None] + (np.arange(2 * L + 1)[None, :] + 1) // 2
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy()
    rsold = np.vdot(r, r)
    rhsn = np.sqrt(abs(np.vdot(rhs, rhs)))
    if rhsn == 0:
        return x, dict(iters=0, relres=0.0)
    iters, rel = 0, 1.0
    for i in range(1, maxiter + 1):
        iters = i
        Ap = plan.adjoint(np.asfortranarray(_weight(plan.synth(p), w)))
        if reg is not None:
            s = np.where(mask, reg[np.minimum(lgrid, L)], 0.0)
            Ap = Ap + (s[..., None] * p if p.ndim == 3 else s * p)
        pAp = np.vdot(p, Ap)
        alpha = 0.0 if pAp == 0 else rsold / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        rsnew = np.vdot(r, r)
        rel = np.sqrt(abs(rsnew)) / rhsn
        if rel <= tol:
            break
        p = r + (rsnew / rsold) * p
        rsold = rsnew
    return np.asfortranarray(x), dict(iters=iters, relres=float(rel))


# ------------------------------------------------------------------------------------------------
# Multi-box (src/Box.jl:159-374, src/MultiBox.jl, src/Cost.jl:174-201), restated over oracle plans.
# ------------------------------------------------------------------------------------------------
def _max_gap(v):
    """_max_gap (Box.jl:257-296) by sorting: (lo, hi, gap, at) with `at` the upper edge of the first
    largest gap in increasing order."""
    s = np.sort(v)
    lo, hi = float(s[0]), float(s[-1])
    if len(s) < 2 or hi <= lo:
        return lo, hi, 0.0, lo
    d = np.diff(s)
    k = int(np.argmax(d))
    return lo, hi, float(d[k]), float(s[k + 1])


def _group_box(th, ph):
    """_group_box (Box.jl:298-319): (thlo, thhi) unclamped; the longitude arc is not needed by the cost."""
    return float(th.min()), float(th.max())


def partition_boxes(theta, phi, L, B=1, tol=1e-10, maxboxes=8):
    """partition_boxes_costed (Box.jl:173-235) with costfn = cost_stretch (Cost.jl:193)."""
    th = np.asarray(theta, np.float64)
    ph = np.mod(np.asarray(phi, np.float64), 2 * np.pi)

    def cost(g):
        lo, hi = _group_box(th[g], ph[g])
        return lib().or_cost_stretch(lo, hi, len(g), L, B, float(tol))

    groups = [np.arange(len(th))]
    costs = [cost(groups[0])]
    e = np.sqrt(np.finfo(np.float64).eps)
    while len(groups) < maxboxes:
        best_gain, best_at, best = 0.0, -1, None
        for gi, g in enumerate(groups):
            if len(g) < 2:
                continue
            tlo, thi, gt, at_t = _max_gap(th[g])
            _, _, gp, at_p = _max_gap(ph[g])
            wphi = gp * np.sin(min(max((tlo + thi) / 2, e), np.pi - e))
            if gt >= wphi:
                if not gt > 0:
                    continue
                mask = th[g] < at_t
            else:
                if not gp > 0:
                    continue
                mask = ph[g] < at_p
            ga, gb = g[mask], g[~mask]
            if len(ga) == 0 or len(gb) == 0:
                continue
            gain = costs[gi] - (cost(ga) + cost(gb))
            if gain > best_gain:
                best_gain, best_at, best = gain, gi, (ga, gb)
        if best_at < 0:
            break
        del groups[best_at]
        del costs[best_at]
        for g in best:
            groups.append(g)
            costs.append(cost(g))
    return groups, costs


def select_method(theta, phi, L, tol=1e-10, B=1):
    """select_method (Cost.jl:174-201) without the NUFSHT extension: the cheapest of Direct, the
    stretched leaf over the bounding box, and the multi-box leaf when the partition has more than
    one box. Returns "direct", "stretch" or "multibox"."""
    th = np.asarray(theta, np.float64)
    n = len(th)
    best, bestc = "direct", lib().or_cost_direct(n, L, B)
    cs = lib().or_cost_stretch(max(float(th.min()), 0.0), min(float(th.max()), np.pi), n, L, B, float(tol))
    if cs < bestc:
        best, bestc = "stretch", cs
    groups, costs = partition_boxes(th, phi, L, B, tol)
    if len(groups) > 1 and sum(costs) < bestc:
        best = "multibox"
    return best


class MultiBox:
    """MultiBoxLeaf (MultiBox.jl:20-144) over oracle stretched plans: a disjoint union in synthesis,
    the sum over boxes (in box order) in the adjoint."""

    def __init__(self, theta, phi, lmax, tol=1e-10, ntrans=1, maxboxes=8, nthreads=1, spin=0):
        th, ph = np.asarray(theta, np.float64), np.asarray(phi, np.float64)
        self.L, self.B, self.M = lmax, ntrans, len(th)
        self.groups, _ = partition_boxes(th, ph, lmax, ntrans, tol, maxboxes)
        self.plans = [Plan(th[g], ph[g], lmax, tol=tol, ntrans=ntrans, method="stretch", nufft="es",
                           nthreads=nthreads, spin=spin) for g in self.groups]

    def synth(self, alm):
        f = np.zeros((self.M, self.B), np.complex128, order="F")
        for g, p in zip(self.groups, self.plans):
            f[g, :] = p.synth(alm).reshape(len(g), self.B, order="F")
        return f[:, 0] if self.B == 1 else f

    def adjoint(self, f):
        f = np.asarray(f).reshape(self.M, self.B, order="F")
        alm = None
        for g, p in zip(self.groups, self.plans):
            a = p.adjoint(np.asfortranarray(f[g, :] if self.B > 1 else f[g, 0]))
            alm = a if alm is None else alm + a
        return alm
