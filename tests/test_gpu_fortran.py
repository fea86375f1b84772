"""The Fortran module API (rmpcdmd_amd/fortran, amdflang) driving the HIP library:
the driver programs built by build() run on the GPU and check their own invariants
(test_cell_system_0's sort invariants, SRD momentum and temperature, dimer momentum
and bond length)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rmpcdmd_amd", "fortran", "bin")


@pytest.mark.parametrize("prog", ["fortran_cells", "fortran_fluid", "fortran_dimer", "fortran_walls",
                                  "fortran_bodies"])
def test_fortran_program(gpu, prog):
    exe = os.path.join(BIN, prog)
    assert os.path.exists(exe), "Fortran programs are not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "%s ok" % prog in r.stdout
