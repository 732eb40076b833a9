"""The C-ABI library loads and exports every function include/pfsht.h declares; host-side
validation errors come back as codes with messages. No device calls are made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfsht.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(pfsht_[a-z_0-9]+)\s*\(", src, re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared_functions()
    assert "pfsht_plan_create" in names and "pfsht_synth" in names and "pfsht_adjoint" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    import pfsht
    lib = C.CDLL(pfsht.library_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly the declared set
    assert sorted(pfsht.EXPORTS) == declared_functions()


def test_version_and_null_argument_errors():
    import pfsht
    L = pfsht.lib()
    assert b"gfx950" in L.pfsht_version()
    out = C.c_void_p()
    rc = L.pfsht_plan_create(None, None, 10, 8, 1e-8, 1, 0, 0, C.byref(out))
    assert rc == 2 and b"NULL" in L.pfsht_last_error()
    assert L.pfsht_synth(None, None, None, 0, None) == 2
    assert L.pfsht_plan_info(None, None) == 2


def test_host_validation_before_device():
    # argument checks happen before any device work (src/Plan.jl:98-102)
    import pfsht
    L = pfsht.lib()
    th = np.zeros(4)
    out = C.c_void_p()
    p = th.ctypes.data_as(C.c_void_p)
    assert L.pfsht_plan_create(p, p, 4, -1, 1e-8, 1, 0, 0, C.byref(out)) == 2      # lmax < 0
    assert L.pfsht_plan_create(p, p, 4, 8, 1e-8, 0, 0, 0, C.byref(out)) == 2       # ntrans < 1
    assert L.pfsht_plan_create(p, p, 0, 8, 1e-8, 1, 0, 0, C.byref(out)) == 2       # no points
    assert L.pfsht_plan_create(p, p, 4, 8, 2.0, 1, 0, 0, C.byref(out)) == 2        # tol out of range
    assert L.pfsht_plan_create(p, p, 4, 8, 1e-8, 1, 0, 7, C.byref(out)) == 2       # unknown precision
    assert L.pfsht_plan_create(p, p, 4, 8, 1e-8, 1, 0, 1, C.byref(out)) == 2       # mixed needs tol >= 1e-6
    with pytest.raises(pfsht.DimensionMismatch):
        pfsht.Plan(np.zeros(3), np.zeros(4), lmax=8)


def test_server_argument_errors():
    # serving runtime: argument checks come back as codes before any device work
    import pfsht
    L = pfsht.lib()
    out = C.c_void_p()
    assert L.pfsht_server_create(0, 1 << 30, C.byref(out)) == 2      # no streams
    assert L.pfsht_server_create(65, 1 << 30, C.byref(out)) == 2     # too many streams
    assert L.pfsht_server_create(2, 0, C.byref(out)) == 2            # no budget
    assert L.pfsht_server_create(2, 1 << 30, None) == 2
    t = C.c_int64()
    assert L.pfsht_server_submit(None, 0, 0, None, None, C.byref(t)) == 2
    done = C.c_int32()
    assert L.pfsht_server_query(None, 0, C.byref(done), None) == 2
    assert L.pfsht_server_get_stats(None, None) == 2


def test_torch_layout_checked():
    # a 2-D / 3-D torch tensor is read as column-major memory: a row-major one is refused instead of
    # being read transposed; 1-D tensors and tensors holding the column-major layout are accepted
    import torch
    import pfsht
    L, B = 4, 3
    n = (L + 1) * (2 * L + 1)
    rowmajor = torch.zeros((L + 1, 2 * L + 1), dtype=torch.complex128)
    with pytest.raises(pfsht.ArgumentError):
        pfsht._buf(rowmajor, n, "alm")
    colmajor = torch.zeros((2 * L + 1, L + 1), dtype=torch.complex128).T
    assert tuple(colmajor.shape) == (L + 1, 2 * L + 1)
    pfsht._buf(colmajor, n, "alm")
    pfsht._buf(torch.zeros(n, dtype=torch.complex128), n, "alm")
    cube = torch.zeros((B, 2 * L + 1, L + 1), dtype=torch.complex128).permute(2, 1, 0)
    pfsht._buf(cube, n * B, "alm")
    with pytest.raises(pfsht.ArgumentError):
        pfsht._buf(torch.zeros((L + 1, 2 * L + 1, B), dtype=torch.complex128), n * B, "alm")
    with pytest.raises(pfsht.ArgumentError):
        pfsht._buf(torch.zeros(2 * n, dtype=torch.complex128)[::2], n, "alm")  # strided 1-D
    field = torch.zeros((7, B), dtype=torch.complex128)  # M x B row-major
    with pytest.raises(pfsht.ArgumentError):
        pfsht._buf(field, 7 * B, "f")


def test_plan_estimate_on_host():
    # the size estimate runs on the host only: larger for more transforms, finite for a refused
    # stretch (the Direct leaf's size), argument errors as codes
    import pfsht
    rng = np.random.default_rng(3)
    ts = np.deg2rad(20)
    th = np.arccos(rng.uniform(np.cos(ts), 1.0, 5000))
    ph = rng.uniform(0, 2 * np.pi, 5000)
    b1 = pfsht.plan_estimate_bytes(th, ph, 256, 1e-8, 1, "stretch")
    b8 = pfsht.plan_estimate_bytes(th, ph, 256, 1e-8, 8, "stretch")
    bm = pfsht.plan_estimate_bytes(th, ph, 256, 1e-5, 8, "stretch", "mixed")
    assert 0 < b1 < b8 and bm < b8
    t5 = np.arccos(rng.uniform(np.cos(np.deg2rad(5)), 1.0, 2000))
    assert pfsht.plan_estimate_bytes(t5, ph[:2000], 256, 1e-8, 1, "auto") > 0
    with pytest.raises(pfsht.ArgumentError):
        pfsht.plan_estimate_bytes(th, ph, 256, 1e-8, 0, "stretch")
    with pytest.raises(pfsht.ArgumentError):
        pfsht.plan_estimate_bytes(th, ph, 256, 1e-8, 1, "stretch", "mixed")  # mixed needs tol >= 1e-6
