"""The ctypes mirrors of the C-ABI structs match the header's layout: a
small C program compiled by gcc against include/plsqr_hip.h prints sizeof
and offsetof of every field, compared with the ctypes Structure (no GPU)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

# C type name -> ctypes class name in the package
STRUCTS = {
    "plsqr_system": "SystemC",
    "plsqr_gen_params": "GenParamsC",
    "plsqr_grid": "GridC",
    "plsqr_solve_opts": "SolveOptsC",
    "plsqr_solve_out": "SolveOutC",
    "plsqr_batch_tuning": "BatchTuningC",
    "plsqr_damping_spec": "DampingSpecC",
}
# ctypes field names that differ from the C member names
RENAME = {("plsqr_gen_params", "lambda_"): "lambda"}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_ctypes_structs_match_header(P, tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "plsqr_hip.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(P._lib, pyname)
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            member = RENAME.get((cname, fname), fname)
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {member}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c_vals = {}
    for ln in out:
        if ln:
            cname, key, val = ln.split()
            c_vals[(cname, key)] = int(val)
    for cname, pyname in STRUCTS.items():
        cls = getattr(P._lib, pyname)
        assert c_vals[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert c_vals[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
