"""The ctypes mirror of include/dtwj.h: every structure the Python side passes through the
C ABI has the size and field offsets the C compiler gives it.  A C program built with gcc
against the header prints sizeof / offsetof; no GPU and no library call are involved."""
import ctypes
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, load_pkg

STRUCTS = {
    "dtwj_options": "Options",
    "dtwj_result": "Result",
    "dtwj_series": "Series",
    "dtwj_mlp": "MLP",
    "dtwj_proposal_report": "ProposalReport",
    "dtwj_gen_spec": "GenSpec",
    "dtwj_plan_info": "PlanInfo",
}


def _c_layout(tmp_path):
    d = load_pkg()
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dtwj.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in getattr(d, pyname)._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, field, cname, field))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cname, key, val = line.split()
        got[(cname, key)] = int(val)
    return d, got


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_structures_match_the_header(tmp_path):
    d, c = _c_layout(tmp_path)
    for cname, pyname in STRUCTS.items():
        cls = getattr(d, pyname)
        assert ctypes.sizeof(cls) == c[(cname, "size")], cname
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == c[(cname, field)], (cname, field)
