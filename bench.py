"""Throughput of the single-dimer nanomotor time step (programs/single_dimer_pbc.f90) on
one MI355X.

A step is one collision cycle: N_MD velocity-Verlet steps of the solvent and the dimer
(forces, RATTLE, catalytic conversion, skin-triggered re-sorts) followed by the grid
shift, the cell sort, the neighbour-list rebuild and the SRD collision.  The metric
counts solvent particle-steps: N_solvent x N_MD per cycle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fp32|fp64]
                    [--mode fused|reference] [--L 64]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "solvent particle-steps/s, FP32-mixed vs FP64 at stated error; % of HBM peak"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="collision cycles timed")
    ap.add_argument("--warmup", type=int, default=1, help="untimed collision cycles")
    ap.add_argument("--form", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--mode", choices=["fused", "reference"], default="fused")
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--n-md", type=int, default=50)
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    return ap.parse_args()


def run_gpu(form, mode, L, n_md, steps, warmup):
    import torch

    from rmpcdmd_amd import capi
    from rmpcdmd_amd.sim import SingleDimer

    f = capi.FP32 if form == "fp32" else capi.FP64
    m = capi.SIM_FUSED if mode == "fused" else capi.SIM_REFERENCE
    sim = SingleDimer(L=(L, L, L), n_md=n_md, form=f, mode=m)
    n = sim.n_solvent
    if warmup:
        sim.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = sim.state()[4]
    sec = t1 - t0
    out = dict(n=n, seconds=sec, ps_per_s=n * n_md * steps / sec, ms_per_cycle=1e3 * sec / steps,
               extra_sorts=int(st.n_extra_sorting), cycles=int(st.cycles))
    sim.close()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X, so --gpus %d "
              "is not supported (replicas would shard one subset per GPU with no communication)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    from rmpcdmd_amd import capi
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.zeros(1, device="cuda")
    capi.call("rmpcd_init", 0)
    main_run = run_gpu(a.form, a.mode, a.L, a.n_md, a.steps, a.warmup)
    compare = None
    if not a.no_compare and a.form == "fp32":
        compare = run_gpu("fp64", a.mode, a.L, a.n_md, a.steps, a.warmup)
    line = {
        "metric": METRIC,
        "value": main_run["ps_per_s"],
        "unit": "solvent particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_run["ms_per_cycle"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.form == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "single_dimer_pbc: dimer nanomotor (sigma_C=sigma_N=2, d=4.5) in a periodic "
                        "L=%d^3 box, rho=10, tau=0.5, N_MD=%d, T=1; a step is one collision cycle"
                        % (a.L, a.n_md),
            "n_solvent": main_run["n"],
            "form": "fp32-mixed (FP32 solvent, FP64 colloids/forces/RATTLE)" if a.form == "fp32" else "fp64",
            "mode": a.mode,
            "extra_sorts": main_run["extra_sorts"],
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    if compare:
        line["fp64_form"] = {"value": compare["ps_per_s"], "ms_per_step": compare["ms_per_cycle"],
                             "mixed_over_fp64": main_run["ps_per_s"] / compare["ps_per_s"]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
