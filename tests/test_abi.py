"""The C-ABI library loads and exports every entry point the header declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from sources_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "sphg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sphg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    declared = header_symbols()
    assert declared, "no declarations found"
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(abi.EXPORTED_SYMBOLS) == declared


def test_struct_layouts_match_the_header():
    # sizes computed by the C compiler from include/sphg.h
    src = ('#include "sphg.h"\n#include <stdio.h>\n#include <stddef.h>\n'
           'int main(){printf("%zu %zu %zu %zu\\n", sizeof(sphg_params), sizeof(sphg_phase_view),'
           ' offsetof(sphg_params, dt), offsetof(sphg_phase_view, dwdx));return 0;}\n')
    exe = "/tmp/sphg_layout_check"
    r = subprocess.run(["gcc", "-x", "c", "-", "-I", os.path.join(ROOT, "include"), "-o", exe],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [C.sizeof(abi.Params), C.sizeof(abi.PhaseView),
                                      abi.Params.dt.offset, abi.PhaseView.dwdx.offset]


def test_strerror_names_reference_stops():
    lib = abi.load_library()
    assert b"Too many interactions" in lib.sphg_strerror(abi.SPHG_ERR_TOO_MANY_INTERACTIONS)
    assert b"out of range" in lib.sphg_strerror(abi.SPHG_ERR_OUT_OF_RANGE)


def test_missing_library_fails_loudly(tmp_path):
    # no CPU fallback: without the HIP library the product path refuses to run
    with pytest.raises(RuntimeError, match="missing"):
        abi.load_library(str(tmp_path / "libsphg.so"))
