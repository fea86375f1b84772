"""ParseText-compatible reading of the reference's parameter files (the drivers' PTparse /
PTread_*), without a GPU: fortran_params reads tests/golden/dimer_test.parameters (the
reference's test/dimer_test.parameters) and prints what single_dimer_pbc asks for; every
value must equal what a plain reading of the same file gives.  A missing name stops the
program with a message, as ParseText does."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "rmpcdmd_amd", "fortran", "bin", "fortran_params")
FIXTURE = os.path.join(REPO, "tests", "golden", "dimer_test.parameters")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran programs not built")


def plain(path):
    out = {}
    for line in open(path):
        line = line.split("#")[0]
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        out[k.strip()] = v.split()
    return out


def as_value(tokens):
    vals = []
    for t in tokens:
        if t in ("T", "F", ".true.", ".false."):
            vals.append(t in ("T", ".true."))
        else:
            vals.append(float(t))
    return vals


def test_reads_the_reference_parameter_file():
    want = plain(FIXTURE)
    r = subprocess.run([EXE, FIXTURE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    seen = 0
    for line in r.stdout.splitlines():
        name, *tokens = line.split()
        assert as_value(tokens) == as_value(want[name]), name
        seen += 1
    assert seen == 21


def test_missing_name_stops(tmp_path):
    text = "\n".join("# d removed" if ln.startswith("d =") else ln for ln in open(FIXTURE).read().splitlines())
    f = tmp_path / "no_d.parameters"
    f.write_text(text)
    r = subprocess.run([EXE, str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "parameter d not found" in r.stdout + r.stderr
