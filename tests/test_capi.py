"""The C ABI library loads and exports every symbol include/naja_amd.h declares (CPU)."""
import ctypes
import os
import re

import naja_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "naja_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(naja_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    assert declared_symbols() == sorted(naja_amd.EXPORTED_SYMBOLS)


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(naja_amd.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_version_and_error_calls_need_no_gpu():
    lib = naja_amd.load_library()
    assert lib.naja_version() >= 100
    assert isinstance(lib.naja_last_error(), bytes)


def test_params_struct_matches_header_layout(tmp_path):
    """ctypes' naja_chr_params mirrors the C struct field for field (offsetof from gcc)."""
    import subprocess

    fields = [f[0] for f in naja_amd.ChrParams._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%zu ", offsetof(naja_chr_params, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "naja_amd.h"\n'
                   f'int main(void){{{body}printf("%zu", sizeof(naja_chr_params));return 0;}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(naja_amd.ChrParams, f).offset for f in fields] + [ctypes.sizeof(naja_amd.ChrParams)]
    assert got == want
