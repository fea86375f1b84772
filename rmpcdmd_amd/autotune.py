"""Launch-shape autotuner of the fused single-dimer program.

For a workload (box size, replica count or the mixed-size batch, form) it times a few
collision cycles for each candidate launch shape and keeps the fastest, one parameter
at a time (coordinate descent, two passes):
    stream_blocks  stream-role workgroups per replica of the step kernel (a cap: the grid
                   takes no more than its 16-byte accesses need)
    sort_blocks    workgroups of the key, place and gather passes of the collision sort
    extra_sort_blocks  the same for the sorts that may follow each MD step
    sort_order     the sort passes' units: a contiguous range per workgroup (1) or
                   grouped by XCD (2)
    groups         replica groups run as concurrent branches of the captured cycle
    collide_cells  cells (threads) per workgroup of the collision kernel (its LDS tile)
The result is stored in a JSON table keyed by the workload; bench.py applies the entry
of its workload when the table exists.

    python -m rmpcdmd_amd.autotune [--L 64] [--replicas 16] [--mixed-L] [--form fp32]
                                   [--out rmpcdmd_amd/tuning.json]
"""
from __future__ import annotations

import argparse
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
CANDIDATES = {"stream_blocks": [1024, 2048, 2560, 4096, 5120, 8192, 16384], "sort_blocks": [2048, 4096, 8192, 16384, 32768],
              "extra_sort_blocks": [256, 512, 1024, 2048, 4096], "sort_order": [1, 2], "groups": [1, 2, 3, 4],
              "collide_cells": [64, 128, 256]}


def workload_key(form, L, replicas, mixed=False):
    return "%s mixed-L replicas=%d" % (form, replicas) if mixed else "%s L=%d replicas=%d" % (form, L, replicas)


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(form, L, replicas, mixed=False, path=TABLE):
    """the tuned launch shape of a workload, or {} (the library defaults)"""
    e = load_table(path).get(workload_key(form, L, replicas, mixed))
    return {k: int(e[k]) for k in CANDIDATES if e and k in e}


def time_shape(make, shape, cycles=4, warmup=2):
    import torch
    sim = make(**shape)
    sim.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(cycles)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / cycles
    sim.close()
    return ms


def tune(make, passes=2, log=print):
    shape = {k: 0 for k in CANDIDATES}
    best = time_shape(make, shape)
    log("default shape: %.3f ms per cycle" % best)
    seen = {}
    for _ in range(passes):
        for name, values in CANDIDATES.items():
            for v in values:
                trial = dict(shape, **{name: v})
                key = tuple(sorted(trial.items()))
                if key not in seen:
                    seen[key] = time_shape(make, trial)
                    log("  %s: %.3f ms per cycle" % (trial, seen[key]))
                if seen[key] < best:
                    best, shape = seen[key], trial
    return shape, best, {"%s" % dict(k): v for k, v in seen.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--replicas", type=int, default=16)
    ap.add_argument("--mixed-L", action="store_true", help="128 replicas with L from {24, 32, 40, 48}")
    ap.add_argument("--form", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--n-md", type=int, default=50)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    import torch

    from . import capi
    from .sim import SingleDimer
    torch.zeros(1, device="cuda")
    capi.call("rmpcd_init", 0)
    form = capi.FP32 if a.form == "fp32" else capi.FP64
    if a.mixed_L:
        import numpy as np
        sizes = [int(v) for v in np.array([24, 32, 40, 48])[np.random.default_rng(20261202).integers(0, 4, 128)]]

        def make(**shape):
            return SingleDimer(n_md=a.n_md, form=form, replica_L=[(q, q, q) for q in sizes], **shape)
        key = workload_key(a.form, 0, len(sizes), mixed=True)
    else:
        def make(**shape):
            return SingleDimer(L=(a.L,) * 3, n_md=a.n_md, form=form, n_replicas=a.replicas, **shape)
        key = workload_key(a.form, a.L, a.replicas)
    shape, best, trials = tune(make)
    table = load_table(a.out)
    entry = {k: v for k, v in shape.items() if v}
    entry.update(ms_per_cycle=best, trials=trials)
    table[key] = entry
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({key: entry}))


if __name__ == "__main__":
    main()
