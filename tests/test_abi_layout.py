"""The structs of include/pfsht.h against their ctypes mirrors in pfsht/__init__.py: the header is
compiled by the host C compiler, which reports each struct's size and every field's offset; the
mirror must agree field for field (a field added to one side only would shift every later field
and corrupt the caller's memory silently). CPU only."""
import ctypes as C
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfsht.h")

PAIRS = [("pfsht_info", "InfoC"), ("pfsht_ring_shape", "RingShapeC"), ("pfsht_plan_opts", "PlanOptsC"),
         ("pfsht_box", "BoxC"), ("pfsht_solve_info", "SolveInfoC"), ("pfsht_server_stats", "ServerStatsC"),
         ("pfsht_server_opts", "ServerOptsC")]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    import pfsht
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {",
             '  printf("{");']
    first = True
    for cname, pyname in PAIRS:
        fields = [f for f, _ in getattr(pfsht, pyname)._fields_]
        parts = [f'\\"size\\": %zu'] + [f'\\"{f}\\": %zu' for f in fields]
        args = [f"sizeof({cname})"] + [f"offsetof({cname}, {f})" for f in fields]
        sep = "" if first else ", "
        first = False
        lines.append(f'  printf("{sep}\\"{cname}\\": {{{", ".join(parts)}}}", {", ".join(args)});')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run([cc, "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


@pytest.mark.parametrize("cname,pyname", PAIRS)
def test_struct_layout_matches_header(c_layout, cname, pyname):
    import pfsht
    py = getattr(pfsht, pyname)
    got = c_layout[cname]
    assert C.sizeof(py) == got["size"], (cname, C.sizeof(py), got["size"])
    for f, _ in py._fields_:
        assert getattr(py, f).offset == got[f], (cname, f)


def test_grid_fft_option_values():
    import pfsht
    src = open(HEADER).read()
    for name, v in pfsht.GRID_FFT.items():
        assert f"PFSHT_GRID_FFT_{ {'auto': 'AUTO', 'lds': 'LDS', 'library': 'LIBRARY'}[name] } = {v}" in src
    with pytest.raises(pfsht.ArgumentError):
        pfsht.Plan([0.1, 0.2], [0.0, 1.0], lmax=8, grid_fft="fftw")
