"""The ctypes mirrors of include/smgpu.h's structs (smgpu/_lib.py) match the C
layout: a C program built with gcc from the header prints every struct's size and
every field's offset, and each must equal ctypes' (a field added on one side only,
or a padding mismatch, would shift every field after it)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from smgpu import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {
    "smgpu_problem_info": _lib.ProblemInfo,
    "smgpu_solve_params": _lib.SolveParams,
    "smgpu_solve_stats": _lib.SolveStats,
    "smgpu_server_config": _lib.ServerConfig,
    "smgpu_request_info": _lib.RequestInfo,
    "smgpu_server_stats": _lib.ServerStats,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "smgpu.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
