"""The ctypes mirrors of the C ABI's structs have the header's layout: every field's
offset and every struct's size, as gcc lays out include/rmpcd_capi.h, on the CPU."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from rmpcdmd_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {"rmpcd_box_t": capi.Box, "rmpcd_particles_t": capi.Particles, "rmpcd_cells_t": capi.Cells,
           "rmpcd_lj_t": capi.LJ, "rmpcd_rigid_body_t": capi.RigidBody, "rmpcd_dimer_config_t": capi.DimerConfig,
           "rmpcd_dimer_stats_t": capi.DimerStats, "rmpcd_program_config_t": capi.ProgramConfig,
           "rmpcd_program_stats_t": capi.ProgramStats}


def _c_fields(cls):
    # the header's field names: ctypes names with a leading underscore (kept-alive host
    # arrays) map to the same C name without it
    return [(name, name.lstrip("_")) for name, *_ in cls._fields_ if not name.endswith("pad_")]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_struct_layout(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "rmpcd_capi.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for py, c in _c_fields(cls):
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, py, cname, c))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for py, _ in _c_fields(cls):
            assert got[(cname, py)] == getattr(cls, py).offset, (cname, py)
