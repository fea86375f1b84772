"""Writes the golden vectors under tests/golden/ from the reference's own Fortran modules
compiled from source into oracle/_ref/ (``make -C oracle ref``; needs the reference
checkout, so it runs only where that exists).  The fixtures are data: inputs and the
reference's outputs.

    python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")

# bit counts per axis: the L=(8,5,5) setup, L=(120,30,30) of test_cell_system_0 is
# (7,5,5) and sampled, cubes 24..64 give (5,5,5)/(6,6,6), plus uneven shapes
HILBERT_FULL = [(1, 1, 1), (2, 3, 1), (3, 3, 3), (4, 2, 3), (1, 4, 2), (3, 1, 4), (5, 5, 5),
                (2, 5, 3)]
HILBERT_SAMPLED = [(6, 6, 6), (7, 5, 5), (7, 7, 7)]


def run_hilbert(M, mode=0):
    out = subprocess.run([os.path.join(REF, "hilbert_drv"), *map(str, M), str(mode)],
                         capture_output=True, text=True, check=True).stdout.split("\n")
    n = 2 ** sum(M) if mode == 0 else 2 ** (3 * M[0])
    keys = np.array([int(v) for v in out[:n]], dtype=np.int32)
    inv = np.array([[int(t) for t in line.split()] for line in out[n:2 * n]], dtype=np.int32)
    return keys, inv


def hilbert():
    data = {}
    for M in HILBERT_FULL:
        keys, inv = run_hilbert(M)
        tag = "%d_%d_%d" % M
        data["keys_" + tag] = keys
        data["inv_" + tag] = inv
    rng = np.random.default_rng(20261202)
    for M in HILBERT_SAMPLED:
        keys, inv = run_hilbert(M)
        tag = "%d_%d_%d" % M
        sel = rng.choice(keys.size, 4096, replace=False).astype(np.int32)
        data["skeys_" + tag] = np.stack([sel, keys[sel]], axis=1)
        data["sinv_" + tag] = np.concatenate([sel[:, None], inv[sel]], axis=1)
    for m in (1, 2, 3):
        keys, inv = run_hilbert((m, m, m), mode=1)
        data["ukeys_%d" % m] = keys
        data["uinv_%d" % m] = inv
    np.savez_compressed(os.path.join(HERE, "hilbert_ref.npz"), **data)


def interaction():
    rng = np.random.default_rng(7)
    n = 400
    sigma = rng.uniform(0.5, 5.0, n)
    eps = rng.uniform(0.1, 2.0, n)
    r = sigma * rng.uniform(0.8, 2 ** (1 / 6), n)
    d = rng.normal(size=(n, 3))
    d = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    r_sq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inp = np.column_stack([d, r_sq, eps, sigma])
    lines = [str(n)] + [" ".join("%016x" % v for v in row.view(np.uint64)) for row in inp]
    out = subprocess.run([os.path.join(REF, "interaction_drv")], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True).stdout.split()
    res = np.array([int(v, 16) for v in out], dtype=np.uint64).view(np.float64).reshape(n, 5)
    np.savez_compressed(os.path.join(HERE, "interaction_ref.npz"), inputs=inp, outputs=res)


if __name__ == "__main__":
    if not os.path.exists(os.path.join(REF, "hilbert_drv")):
        sys.exit("oracle/_ref not built (make -C oracle ref)")
    hilbert()
    interaction()
    print("golden vectors written to", HERE)
