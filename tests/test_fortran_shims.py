"""The host-side link shims of the Fortran module API, on the CPU:

* the tester, hdf5 and h5md_module stand-ins (interfaces of the reference's absent
  submodules fortran_tester and fortran_h5md and of HDF5), particle_system_t%init_from_file
  reading what h5md_module wrote, profile_t, and the tester's own bookkeeping, through the
  fortran_shims program;
* the reference's own Fortran tests and the n_colloids_pbc driver, compiled unmodified from
  the reference checkout against this library's modules and linked with it (skipped where
  the checkout is absent).  The binaries stay in rmpcdmd_amd/fortran/refbuild, which is
  neither committed nor sent to a GPU; they need a GPU to run.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORTRAN = os.path.join(REPO, "rmpcdmd_amd", "fortran")
BIN = os.path.join(FORTRAN, "bin")
FMOD = os.path.join(REPO, "rmpcdmd_amd", "_build", "fmod")
REFERENCE = "/root/reference"
AMDFLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")


def test_shim_program(tmp_path):
    exe = os.path.join(BIN, "fortran_shims")
    if not os.path.exists(exe):
        pytest.skip("fortran_shims is not built")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    # the second tester planted four failures in six checks; the main one passed all twelve
    assert "tester: 4 error(s) for 6 test(s)" in r.stdout
    assert "tester: 0 error(s) for 12 test(s)" in r.stdout
    lines = (tmp_path / "shim_check.h5.txt").read_text().splitlines()
    assert lines[0] == "H5MD-TEXT 1"
    assert "group /particles/fluid" in lines
    i = lines.index("attribute /particles/fluid/box boundary string 3")
    assert lines[i + 1:i + 4] == ["periodic"] * 3
    i = lines.index("fixed /particles/fluid/box/edges double 1 3")
    assert np.allclose([float(v) for v in lines[i + 1:i + 4]], [8, 5, 5])
    # time element: three frames with their steps and times
    frames = [ln.split() for ln in lines if ln.startswith("frame /particles/fluid/position ")]
    assert [int(f[2]) for f in frames] == [1, 2, 3]
    assert np.allclose([float(f[3]) for f in frames], [0.1, 0.2, 0.3])
    # linear element: steps offset + k * 10, times k * 0.5
    frames = [ln.split() for ln in lines if ln.startswith("frame /observables/n ")]
    assert [int(f[2]) for f in frames] == [0, 10, 20, 30]
    assert np.allclose([float(f[3]) for f in frames], [0, 0.5, 1.0, 1.5])
    # frame 2 holds x + 2 in Fortran order, doubles written to 17 digits
    i = lines.index(next(ln for ln in lines if ln.startswith("frame /particles/fluid/position 2 ")))
    x = np.array([k * 0.5 + 1 / 3 for k in range(1, 16)])
    assert np.array_equal(np.array([float(v) for v in lines[i + 1:i + 16]]), x + 2)


REF_UNITS = [
    ("test", "test_cell_system_0"), ("test", "test_neighbor_list_0"), ("test", "test_neighbor_list_1"),
    ("test", "test_md_0"), ("test", "test_common_0"), ("test", "test_common_1"),
    ("test", "test_elastic_network_0"), ("test", "setup_simple_fluid"), ("test", "setup_bulk_decay"),
    ("test", "setup_thermal_fluid"), ("test", "setup_simple_colloids"), ("test", "setup_single_colloid"),
    ("test", "setup_simple_rattle"), ("test", "setup_single_catalytic_fixed_sphere"),
    ("programs", "n_colloids_pbc"),
]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "test")), reason="reference checkout absent")
@pytest.mark.parametrize("where,name", REF_UNITS, ids=[n for _, n in REF_UNITS])
def test_reference_source_links(where, name):
    """the unmodified reference source compiles against the module API and links with
    librmpcd_fortran and librmpcd_amd"""
    if shutil.which(AMDFLANG) is None and not os.path.exists(AMDFLANG):
        pytest.skip("amdflang absent")
    if not os.path.isdir(FMOD) or not os.path.exists(os.path.join(FORTRAN, "librmpcd_fortran.so")):
        pytest.skip("the Fortran module API is not built")
    out = os.path.join(FORTRAN, "refbuild")
    os.makedirs(out, exist_ok=True)
    src = os.path.join(REFERENCE, where, name + ".f90")
    cmd = [AMDFLANG, "-fopenmp", "-I", FMOD, src, "-o", os.path.join(out, name), "-L" + FORTRAN, "-lrmpcd_fortran",
           "-L" + os.path.join(REPO, "rmpcdmd_amd"), "-lrmpcd_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=out)
    assert r.returncode == 0, r.stderr[-2000:]
    assert os.path.exists(os.path.join(out, name))
