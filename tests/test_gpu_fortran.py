"""The Fortran module API (rmpcdmd_amd/fortran, amdflang) driving the HIP library:
the driver programs built by build() run on the GPU and check their own invariants
(test_cell_system_0's sort invariants, SRD momentum and temperature, dimer momentum
and bond length)."""
import os
import subprocess

import numpy as np

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


def test_single_dimer_pbc_driver(gpu, tmp_path):
    """the single_dimer_pbc driver on the reference's test/dimer_test.parameters (L=24,
    N_MD=125, 20 + 200 loops), FP64 form: the observables table has a row per sampled
    loop, the temperature holds at T, the reference's CI bars on internal energy and
    centre-of-mass velocity hold, and the colloid mass is the reference's printed
    1130.9733867645264 (doc/run.rst:103; the driver's single-precision pi)"""
    exe = os.path.join(BIN, "single_dimer_pbc")
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dimer_test.parameters")
    out = tmp_path / "dimer.txt"
    print("single_dimer_pbc on dimer_test.parameters", flush=True)
    r = subprocess.run([exe, fixture, str(out), "42"], capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    head = [ln for ln in out.read_text().splitlines() if ln.startswith("#")]
    mass = [float(x) for ln in head if ln.startswith("# mass") for x in ln.split()[2:]]
    assert mass and all(abs(m / 1130.9733867645264 - 1) < 1e-13 for m in mass)
    rows = np.loadtxt(out)
    assert rows.shape == (201, 11)
    assert (rows[:, 0] == np.arange(20, 221)).all()
    T = rows[:, 1]
    assert abs(T.mean() - 1) < 0.03 and np.all(np.abs(T - 1) < 0.1)
    # the reference's own CI bars on this run (test/dimer_test.py:13-22): the internal
    # energy (column e_total) fluctuates with a standard deviation below 0.1, and the
    # centre-of-mass velocity stays within 1e-15 of its mean
    e_int = rows[:, 4] - rows[:, 4].mean()
    print("internal energy std", e_int.std())
    assert e_int.std() < 0.1
    v_com = rows[:, 5:8] - rows[:, 5:8].mean(axis=0)
    print("max |v_com - mean|", np.max(np.abs(v_com)))
    assert np.max(np.abs(v_com)) < 1e-15
    assert np.max(np.abs(rows[:, 5:8])) < 1e-3
