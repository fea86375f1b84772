"""Throughput of the single-dimer nanomotor time step (programs/single_dimer_pbc.f90) on
one MI355X.

A step is one collision cycle: N_MD velocity-Verlet steps of the solvent and the dimer
(forces, RATTLE, catalytic conversion, skin-triggered re-sorts) followed by the grid
shift, the cell sort, the neighbour-list rebuild and the SRD collision.  The metric
counts solvent particle-steps: N_solvent x N_MD per cycle, summed over replicas.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fp32|fp64]
                    [--mode fused|reference] [--L 64] [--replicas R] [--no-cpu]

The headline runs R independent L=64 systems (default R=16) in every launch, FP32-mixed,
with the FP64 form and a single system alongside.  The JSON line carries the roofline
of the dominant kernel, k_fz_step (the streaming pass over all solvent, 9 words per
particle: read position and velocity, write position; the list members' force work
rides along), timed with HIP event-record nodes inside the captured graph during the
timed cycles, and the CPU baseline: the oracle (FP64 C with OpenMP) running one cycle of
the single-system workload on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "solvent particle-steps/s, FP32-mixed vs FP64 at stated error; % of HBM peak"
HBM_PEAK_GBS = 8000.0
PROFILE_STRIDE = 5  # event nodes around every 5th streaming launch of the captured cycle


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="collision cycles timed")
    ap.add_argument("--warmup", type=int, default=2, help="untimed collision cycles")
    ap.add_argument("--form", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--mode", choices=["fused", "reference"], default="fused")
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--replicas", type=int, default=16, help="independent L^3 systems in every launch")
    ap.add_argument("--n-md", type=int, default=50)
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-single", action="store_true", help="skip the single-replica run")
    ap.add_argument("--no-lazy", action="store_true", help="skip the lazy-streaming run")
    ap.add_argument("--no-bodies", action="store_true", help="skip the linked-colloid routines")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the wall, forced-flow, thermostat and energy routines")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the 128-replica run with L drawn from {24, 32, 40, 48}")
    ap.add_argument("--no-colloids", action="store_true",
                    help="skip the n_colloids_pbc workload (L=128, 1000 colloids, lazy streaming)")
    ap.add_argument("--colloids-only", action="store_true",
                    help="run only the n_colloids_pbc workload (for profiling it)")
    ap.add_argument("--colloids-cycles", type=int, default=4, help="timed cycles of the n_colloids_pbc workload")
    ap.add_argument("--colloids-forms", default="fp32,fp64", help="forms of the n_colloids_pbc workload")
    ap.add_argument("--mixed-only", action="store_true",
                    help="run only the 128-replica mixed-L batch (for profiling it)")
    ap.add_argument("--groups", type=int, default=0,
                    help="replica groups of the fused graph instead of the tuned number (1 for traffic "
                         "counters of whole-batch launches)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="first copy MB megabytes device to device once (the reference dispatch for "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def mixed_sizes():
    """128 replicas with L drawn uniformly from {24, 32, 40, 48} (seeded)"""
    from rmpcdmd_amd.sim import mixed_L_sizes
    return mixed_L_sizes()


GROUPS = 0  # --groups: replica groups instead of the tuned number (0: tuned)


def run_gpu(form, mode, L, n_md, steps, warmup, replicas, profile, sizes=None, lazy=0):
    import torch

    from rmpcdmd_amd import capi
    from rmpcdmd_amd.sim import SingleDimer

    f = capi.FP32 if form == "fp32" else capi.FP64
    m = capi.SIM_FUSED if mode == "fused" else capi.SIM_REFERENCE
    from rmpcdmd_amd.autotune import lookup
    if sizes is not None:
        replicas = len(sizes)
        shape = lookup(form, 0, replicas, mixed=True) if mode == "fused" else {}
        if GROUPS and mode == "fused":
            shape = dict(shape, groups=GROUPS)
        sim = SingleDimer(n_md=n_md, form=f, mode=m, replica_L=[(q, q, q) for q in sizes], **shape)
    else:
        shape = lookup(form, L, replicas) if mode == "fused" else {}
        if GROUPS and mode == "fused":
            shape = dict(shape, groups=GROUPS)
        sim = SingleDimer(L=(L, L, L), n_md=n_md, form=f, mode=m, n_replicas=replicas, lazy=lazy, **shape)
    st0 = sim.state()[4]
    n_total = int(st0.n_solvent_total)
    out = {}
    if warmup:
        sim.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    out.update(n=n_total, seconds=sec, ps_per_s=n_total * n_md * steps / sec, ms_per_cycle=1e3 * sec / steps,
               launch_shape=shape or "defaults")
    extra = 0
    for r in range(replicas):
        extra += int(sim.state(r)[4].n_extra_sorting)
    out["extra_sorts"] = extra
    out["cycles"] = warmup + steps
    if profile and mode == "fused":
        # the streaming kernel's duration: `steps` more cycles of the same systems, with event
        # nodes around every PROFILE_STRIDE-th streaming launch of the cycle graph.  The nodes
        # cost about 0.4 ms per cycle, so the value above is timed without them.
        try:
            capi.call("rmpcd_dimer_profile", sim._h, PROFILE_STRIDE)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            sim.run(steps)
            torch.cuda.synchronize()
            out["profiled_ms_per_cycle"] = 1e3 * (time.perf_counter() - t2) / steps
            ms, k = C.c_double(), C.c_int32()
            capi.call("rmpcd_dimer_stream_time", sim._h, C.byref(ms), C.byref(k))
            out["stream_ms"] = ms.value
            out["stream_samples"] = k.value
        except capi.RmpcdError as e:
            out["profile_error"] = str(e)
    sim.close()
    return out


def host_cpu():
    """the host's CPU model and the physical cores this process may run on (lscpu), and the
    threads the CPU baseline uses: one per physical core, capped at OMP_NUM_THREADS when
    the environment sets it (a host shared by several GPUs grants each job a share)"""
    import subprocess
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    model, cores = "unknown", None
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        m = re.search(r"^Model name:\s*(.+)$", txt, re.M)
        model = m.group(1).strip() if m else model
        topo = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        core_of = {}
        for ln in topo.splitlines():
            if ln and not ln.startswith("#"):
                c, core, sock = ln.split(",")[:3]
                core_of[int(c)] = (core, sock)
        cores = len({core_of[c] for c in allowed if c in core_of}) or None
        host_cores = len(set(core_of.values()))
    except (OSError, subprocess.SubprocessError, ValueError):
        host_cores = None
    cores = cores or len(allowed)
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(cores, int(share)) if share.isdigit() and int(share) > 0 else cores
    return {"model": model, "physical_cores_allowed": cores, "physical_cores_host": host_cores,
            "threads": threads, "omp_num_threads_env": share or None}


def cpu_baseline(L, n_md, min_seconds=10.0):
    """the oracle (FP64 C with OpenMP) on this host's cores, in a child process with one
    thread per physical core bound close (OMP_PROC_BIND=close, OMP_PLACES=cores):
    collision cycles of one system of the same workload, repeated until about
    min_seconds of CPU work have run"""
    import subprocess
    cpu = host_cpu()
    env = dict(os.environ, OMP_NUM_THREADS=str(cpu["threads"]), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", str(L), str(n_md),
                        str(cpu["threads"]), str(min_seconds)], env=env, capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0:
        return {"error": "CPU baseline child exited with %d: %s" % (r.returncode, r.stderr[-400:])}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out.update({"cores": cpu["threads"], "cpu_model": cpu["model"],
                "physical_cores_allowed": cpu["physical_cores_allowed"],
                "physical_cores_host": cpu["physical_cores_host"], "omp_proc_bind": "close",
                "omp_places": "cores"})
    return out


def cpu_baseline_child(L, n_md, threads, min_seconds, max_cycles=40):
    from oracle import oracle as O
    from oracle.dimer import TimingDraws, synthetic_dimer
    O.lib().orc_set_threads(threads)
    drv = synthetic_dimer((L, L, L), seed=1, tau=0.5, n_md=n_md)
    draws = TimingDraws(2)
    t0 = time.perf_counter()
    cycles = 0
    while cycles < max_cycles and (cycles == 0 or time.perf_counter() - t0 < min_seconds):
        drv.cycle(draws)
        cycles += 1
    sec = time.perf_counter() - t0
    print(json.dumps({
        "value": drv.s.n * n_md * cycles / sec, "unit": "solvent particle-steps/s", "kind": "port",
        "sample": "%d collision cycles (each %d MD steps + grid shift, sort, list rebuild, SRD) of the L=%d "
                  "dimer workload, %d solvent, on the FP64 C/OpenMP oracle in %.1f s; change_species draws "
                  "only for the particles the reference draws for" % (cycles, n_md, L, drv.s.n, sec)}))


def measured_traffic(form, L, replicas):
    """HBM bytes per launch of the streaming kernel, from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration), when one matches"""
    import glob
    key = "%s L=%d replicas=%d" % (form, L, replicas) if L else "%s mixed-L replicas=%d" % (form, replicas)
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.txt"))):
        for line in open(path):
            m = re.match(r"\s*traffic_bytes_per_launch\[(.*)\]\s*=\s*([0-9.eE+]+)", line)
            if m and m.group(1) == key:
                return float(m.group(2))
    return None


def linked_bodies(with_cpu):
    """the linked-colloid routines on a Janus-like shell: 300 beads on a sphere of radius
    4.5, linked to the beads closer than 1.8 mean spacings (i < j, by i then j, as the
    reference's single_body driver builds them; 1202 links in 155 levels); one MD step of
    dt = 0.01 from the constrained shape with unit thermal velocities, RATTLE to 1e-8.
    GPU times are per call, wall clock, with the device state reset between calls."""
    import numpy as np
    import torch

    from rmpcdmd_amd import capi
    from rmpcdmd_amd.system import ParticleSet
    n, R = 300, 4.5
    k = np.arange(n) + 0.5
    th, ph = np.arccos(1 - 2 * k / n), np.pi * (1 + 5 ** 0.5) * k
    r = R * np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], 1)
    i, j = np.triu_indices(n, 1)
    near = np.linalg.norm(r[i] - r[j], axis=1) < 1.8 * np.sqrt(4 * np.pi * R * R / n)
    links = np.ascontiguousarray(np.stack([i[near], j[near]], 1).astype(np.int32))
    d = np.linalg.norm(r[links[:, 0]] - r[links[:, 1]], axis=1)
    rng = np.random.default_rng(5)
    edges = np.array([40.0] * 3)
    pr = r + 20.0
    vel = rng.normal(size=(n, 3))
    vel -= vel.mean(0)
    pos = pr + 0.01 * vel
    mass = np.array([1.0])
    g = ParticleSet(n, form=0, device="cuda")
    g.load(pos=pos, vel=vel, force=rng.normal(size=(n, 3)), force_old=np.zeros((n, 3)), image=np.zeros((n, 3), np.int32),
           id=np.arange(n, dtype=np.int32), species=np.ones(n, np.int32), flags=np.zeros(n, np.int32))
    p0, v0 = g.pos.clone(), g.vel.clone()
    rt0 = torch.as_tensor(np.ascontiguousarray(pr.T)).cuda()
    rt = rt0.clone()
    lk, dd = links.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p)
    m = (C.c_double * 1)(1.0)
    e3 = capi.d3(edges)
    it = C.c_int32(0)

    def timed(fn, reps=5):
        best = 1e30
        for _ in range(reps):
            g.pos.copy_(p0)
            g.vel.copy_(v0)
            rt.copy_(rt0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        return best * 1e3

    def rpos():
        capi.call("rmpcd_rattle_body_pos", g.struct(), C.c_void_p(rt.data_ptr()), lk, dd, len(d), m, 1, 0.01, e3,
                  1e-8, C.byref(it))

    ms_pos = timed(rpos)
    sweeps_pos = it.value
    rpos()
    p1, v1 = g.pos.clone(), g.vel.clone()

    def rvel():
        g.pos.copy_(p1)
        g.vel.copy_(v1)
        capi.call("rmpcd_rattle_body_vel", g.struct(), lk, dd, len(d), m, 1, 0.01, e3, 1e-8, C.byref(it))

    ms_vel = timed(rvel)
    sweeps_vel = it.value
    e = C.c_double(0)
    ms_el = timed(lambda: capi.call("rmpcd_elastic_network", g.struct(), lk, dd, len(d), 30.0, e3, C.byref(e)))
    body = torch.zeros(C.sizeof(capi.RigidBody), dtype=torch.uint8, device="cuda")
    rows = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    capi.call("rmpcd_rigid_body_init", C.c_void_p(body.data_ptr()), g.struct(), m, 1, e3, 0, n - 1,
              C.c_void_p(rows.data_ptr()))
    bp = C.c_void_p(body.data_ptr())

    def rigid_step():
        capi.call("rmpcd_rigid_body_compute_force_torque", bp, g.struct())
        capi.call("rmpcd_rigid_body_vv1", bp, g.struct(), 0.01, 1e-12, e3, None)
        capi.call("rmpcd_rigid_body_vv2", bp, g.struct(), 0.01)

    ms_rigid = timed(rigid_step)
    out = {"beads": n, "links": int(len(d)), "rattle_pos": {"ms": ms_pos, "sweeps": sweeps_pos},
           "rattle_vel": {"ms": ms_vel, "sweeps": sweeps_vel}, "elastic_network_ms": ms_el,
           "rigid_body_step_ms": ms_rigid}
    if with_cpu:
        from oracle import oracle as O
        sp = np.ones(n, np.int32)
        t = time.perf_counter()
        w_pos, w_vel, w_pr = pos.copy(), vel.copy(), pr.copy()
        O.rattle_body_pos(w_pos, w_vel, w_pr, sp, mass, links, d, 0.01, edges, 1e-8)
        cpu_pos = (time.perf_counter() - t) * 1e3
        t = time.perf_counter()
        O.rattle_body_vel(w_pos, w_vel, sp, mass, links, d, 0.01, edges, 1e-8)
        cpu_vel = (time.perf_counter() - t) * 1e3
        out["cpu_baseline"] = {"kind": "port", "cores": 1, "rattle_pos_ms": cpu_pos, "rattle_vel_ms": cpu_vel,
                               "sample": "the oracle (C, one thread) on the same shell and step"}
    return out


def mixed_batch(a):
    """128 single-dimer replicas with L drawn from {24, 32, 40, 48} (67,496,960 solvent, a
    working set far above the 256 MB Infinity Cache), with the roofline of its streaming
    kernel (event nodes in the captured cycle, as for the headline) and the whole step
    against B_ps"""
    sizes = mixed_sizes()
    mx = run_gpu(a.form, a.mode, 0, a.n_md, max(2, a.steps // 2), 1, 0, profile=a.mode == "fused", sizes=sizes)
    out = {"value": mx["ps_per_s"], "ms_per_step": mx["ms_per_cycle"], "n_solvent": mx["n"],
           "replicas": len(sizes), "L": "128 replicas, L uniform in {24,32,40,48}, seeded"}
    w = 4 if a.form == "fp32" else 8
    S = 38 if a.form == "fp32" else 68
    n_sorts = (mx["extra_sorts"] / max(1, mx["cycles"] * len(sizes)) + 1) / a.n_md
    bps = 9 * w + n_sorts * (2 * S + 16 / 10)
    out["step_bytes_per_particle_step"] = bps
    out["step_frac_of_hbm_peak"] = bps * mx["ps_per_s"] / 1e9 / HBM_PEAK_GBS
    if mx.get("stream_ms", 0) > 0:
        alg = 9 * w * mx["n"]
        ach = alg / (mx["stream_ms"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "kernel": "k_fz_step", "algorithmic_bytes_per_launch": alg,
                           "avg_launch_ms": mx["stream_ms"], "launches_timed": mx["stream_samples"],
                           "traffic": measured_traffic(a.form, 0, len(sizes))}
    return out


def n_colloids(cycles, warmup=1, forms=("fp32", "fp64")):
    """programs/n_colloids_pbc.f90's production loop at L=128, rho=10, 1000 colloids
    sigma=2.5 (20,316,937 solvent), tau=0.5, N_MD=50, as the graph-captured program (one
    hipGraph per collision cycle).  Lazy streaming: the solvent of the cells near the
    colloids (fraction f_md) moves every step, the rest at the sorts, so the algorithmic
    bytes per particle-step are f_md 9w + (n_sort / n_dt)(2S + 16/rho + 9w).  The FP32-mixed
    form and the FP64 form start from the same state and draws; their difference after the
    timed cycles is the FP32 error at this box size."""
    import numpy as np
    import torch

    from rmpcdmd_amd import capi
    from rmpcdmd_amd.programs import PROG_COLLOIDS, Program
    out = {"workload": "n_colloids_pbc production loop: L=128^3, rho=10, 1000 colloids sigma=2.5 (colloid-colloid "
                       "LJ cut 6 sigma), skin 0.8, tau=0.5, N_MD=50, T=1, lazy cell-wise streaming; a step is one "
                       "collision cycle"}
    runs = {}
    for form in forms:
        f = capi.FP32 if form == "fp32" else capi.FP64
        t0 = time.perf_counter()
        prog = Program(PROG_COLLOIDS, form=f, seed=20261202, graph=1, L=(128, 128, 128), n_md=50, tau=0.5,
                       sigma=2.5, n_colloids=1000, T=1.0)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        prog.run(warmup)
        torch.cuda.synchronize()
        e0 = int(prog.state()[4].n_extra_sorting)
        temps = []
        fmd = []
        t_run = 0.0
        for c in range(cycles):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            prog.run(1)
            torch.cuda.synchronize()
            t_run += time.perf_counter() - t1
            st = prog.state()[4]
            temps.append(float(st.temperature))
            d = prog.download_cells()
            fmd.append(float(d["cell_count"][d["is_md"] != 0].sum()) / float(st.n_solvent))
        st = prog.state()[4]
        n = int(st.n_solvent)
        extra = int(st.n_extra_sorting) - e0
        w = 4 if form == "fp32" else 8
        S = 38 if form == "fp32" else 68
        f_md = float(np.mean(fmd))
        n_sort = 1.0 + extra / cycles
        bps = f_md * 9 * w + (n_sort / 50.0) * (2 * S + 16.0 / 10 + 9 * w)
        value = n * 50 * cycles / t_run
        col = prog.download_colloids()
        runs[form] = {"value": value, "ms_per_step": 1e3 * t_run / cycles, "n_solvent": n, "setup_s": setup,
                      "f_md": f_md, "sorts_per_cycle": n_sort, "extra_sorts": extra,
                      "bytes_per_particle_step": bps, "effective_GBs": bps * value / 1e9,
                      "frac_of_hbm_peak": bps * value / 1e9 / HBM_PEAK_GBS, "temperature": temps,
                      "_colloid_pos": col}
        prog.close()
    if "fp32" in runs and "fp64" in runs:
        a, b = runs["fp32"], runs["fp64"]
        ta, tb = np.array(a["temperature"]), np.array(b["temperature"])
        out["fp32_vs_fp64"] = {
            "temperature_rel_max": float(np.abs(ta - tb).max() / np.abs(tb).max()),
            "colloid_pos_abs_max": float(np.abs(a["_colloid_pos"] - b["_colloid_pos"]).max()),
            "cycles": cycles + warmup}
    for form in runs:
        runs[form].pop("_colloid_pos")
        out[form] = runs[form]
    return out


def mpcd_variants(with_cpu, L=64, reps=20):
    """the wall, forced-flow, thermostat and energy routines (mpcd.f90:144-160, 193-337,
    459-511, 646-700, 748-801) on one L^3 box at rho = 10 (walls: z walls, periodic x, y),
    each form: mean time per call over `reps` calls (HIP events on the library's stream),
    algorithmic bytes per call and the fraction of the HBM peak.  Per particle:
    SRD-like cell passes read and write the velocities and read the species (6w + s),
    MPC-AT also reads the id (Threefry normals keyed by particle), the endothermic
    reaction also writes species, the forced-flow stream reads position and velocity and
    writes the position (9w); per cell: count and start (8 B)."""
    import numpy as np
    import torch

    from rmpcdmd_amd import capi
    from rmpcdmd_amd.system import CellSet, ParticleSet
    n = 10 * L ** 3
    out = {"L": L, "n": n}
    for name, form in (("fp32", capi.FP32), ("fp64", capi.FP64)):
        w, sb = (4, 1) if form == capi.FP32 else (8, 4)
        gen = torch.Generator(device="cuda").manual_seed(11)
        res = {}
        for walls in (False, True):
            src, dst = ParticleSet(n, form), ParticleSet(n, form)
            src.pos.copy_((torch.rand(3, n, generator=gen, device="cuda", dtype=torch.float64) * L).to(src.pos.dtype))
            src.vel.copy_(torch.randn(3, n, generator=gen, device="cuda", dtype=torch.float64).to(src.vel.dtype))
            sp = 1 + (torch.rand(n, generator=gen, device="cuda") < 0.3).to(src.species.dtype)
            src.species.copy_(sp)
            cs = CellSet((L, L, L), 1.0, walls)
            if walls:
                cs.box.bc[0] = 1  # periodic x and y, bounce-back z (RMPCD_*_BC)
                cs.box.bc[1] = 1
                cs.box.bc[2] = 2
                cs.set_origin((-0.3, -0.6, 0.0))
            else:
                cs.set_origin((-0.3, -0.6, -0.9))
            i2i = torch.zeros(n, dtype=torch.int32, device="cuda")
            capi.call("rmpcd_sort", cs.boxp(), src.struct(), dst.struct(), C.c_void_p(i2i.data_ptr()), cs.cellsp(), 0)
            cs.is_reac.fill_(1)
            nc = cs.box.ncells
            if walls:
                calls = {
                    "wall_mpcd_step": (lambda k: capi.call("rmpcd_wall_mpcd_step", cs.boxp(), cs.cellsp(), dst.struct(),
                                                           0, None, None, 0, 1.0, 1, 0, 0.0, None, None, None, 7, k),
                                       (6 * w + sb) * n + 8 * nc),
                    "mpcd_stream_xforce_yzwall": (lambda k: capi.call("rmpcd_mpcd_stream_xforce_yzwall", cs.boxp(),
                                                                      dst.struct(), 1e-4, 0.01), 9 * w * n),
                }
            else:
                calls = {
                    "mpcat_step": (lambda k: capi.call("rmpcd_mpcat_step", cs.boxp(), cs.cellsp(), dst.struct(), None, 7,
                                                       k, 1.0, 1), (6 * w + sb + 4) * n + 8 * nc),
                    "rescale_cells": (lambda k: capi.call("rmpcd_rescale_cells", cs.boxp(), cs.cellsp(), dst.struct(),
                                                          1.0), (6 * w + sb) * n + 8 * nc),
                    "bulk_reaction_endothermic": (lambda k: capi.call("rmpcd_bulk_reaction_endothermic", cs.boxp(),
                                                                      cs.cellsp(), dst.struct(), 2, 1, 1e-3, 0.5, 0.01,
                                                                      None, 7, k), (6 * w + 2 * sb) * n + 9 * nc),
                }
            for key, (fn, alg) in calls.items():
                fn(0)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for k in range(reps):
                    fn(k + 1)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                gbs = alg / (ms * 1e-3) / 1e9
                res[key] = {"ms": ms, "algorithmic_bytes": alg, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
            del src, dst, cs
        out[name] = res
    if with_cpu:
        from oracle import oracle as O
        Lc = 32
        rng = np.random.default_rng(3)
        c = O.cells_init((Lc, Lc, Lc))
        O.set_origin(c, (-0.3, -0.6, -0.9))
        ps = O.PState(10 * Lc ** 3)
        ps.pos[:] = rng.random((ps.n, 3)) * Lc
        ps.vel[:] = rng.normal(size=(ps.n, 3))
        ps.species[:] = 1 + (rng.random(ps.n) < 0.3)
        srt, _, cnt, st = O.sort(c, ps)
        cpu = {}
        t = time.perf_counter()
        O.rescale_cells(c, cnt, st, srt.vel, srt.species, 1.0)
        cpu["rescale_cells_ns_per_particle"] = (time.perf_counter() - t) / ps.n * 1e9
        g = rng.normal(size=srt.vel.shape)
        t = time.perf_counter()
        O.mpcat_step(c, cnt, st, srt.vel, srt.species, g, 1.0, hydro=True)
        cpu["mpcat_step_ns_per_particle"] = (time.perf_counter() - t) / ps.n * 1e9
        out["cpu_baseline"] = {"kind": "port", "cores": 1, "sample": "the oracle (C, one thread) on an L=32 box "
                               "(327,680 particles), one call each", **cpu}
    return out


def main():
    global GROUPS
    a = parse()
    GROUPS = a.groups
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
    if a.copy_calibration:
        nbytes = a.copy_calibration << 20
        src, dst = C.c_void_p(), C.c_void_p()
        capi.call("rmpcd_malloc", C.byref(src), nbytes)
        capi.call("rmpcd_malloc", C.byref(dst), nbytes)
        capi.call("rmpcd_memcpy", dst, src, nbytes, 2)
        capi.call("rmpcd_free", src)
        capi.call("rmpcd_free", dst)
    if a.colloids_only:
        print(json.dumps({"n_colloids_L128": n_colloids(a.colloids_cycles, forms=tuple(a.colloids_forms.split(",")))}))
        return
    if a.mixed_only:
        print(json.dumps({"batched_mixed_L": mixed_batch(a)}))
        return
    main_run = run_gpu(a.form, a.mode, a.L, a.n_md, a.steps, a.warmup, a.replicas, profile=True)
    compare = None
    if not a.no_compare and a.form == "fp32":
        compare = run_gpu("fp64", a.mode, a.L, a.n_md, a.steps, a.warmup, a.replicas, profile=False)
    w = 4 if a.form == "fp32" else 8
    roofline = None
    if "stream_ms" in main_run and main_run["stream_ms"] > 0:
        alg = 9 * w * main_run["n"]
        achieved = alg / (main_run["stream_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(a.form, a.L, a.replicas),
                    "kernel": "k_fz_step", "algorithmic_bytes_per_launch": alg,
                    "avg_launch_ms": main_run["stream_ms"], "launches_timed": main_run["stream_samples"],
                    "timed_over": "%d more cycles after the value's, with event nodes (%.3f ms per cycle)"
                                  % (a.steps, main_run["profiled_ms_per_cycle"])}
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
                        "L=%d^3 box, rho=10, tau=0.5, N_MD=%d, T=1, x%d replica(s); a step is one collision "
                        "cycle" % (a.L, a.n_md, a.replicas),
            "n_solvent": main_run["n"],
            "replicas": a.replicas,
            "form": "fp32-mixed (FP32 solvent, FP64 colloids/forces/RATTLE)" if a.form == "fp32" else "fp64",
            "mode": a.mode,
            "extra_sorts": main_run["extra_sorts"],
            "launch_shape": main_run["launch_shape"],
        },
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if "profile_error" in main_run:
        line["roofline_error"] = main_run["profile_error"]
    if compare:
        line["fp64_form"] = {"value": compare["ps_per_s"], "ms_per_step": compare["ms_per_cycle"],
                             "mixed_over_fp64": main_run["ps_per_s"] / compare["ps_per_s"]}
    if not a.no_lazy and a.mode == "fused":
        # the same trajectory bit for bit, but the particles outside the neighbour lists are
        # moved only when a sort or an exact skin test reads them: reported apart from the
        # eager value
        lz = run_gpu(a.form, a.mode, a.L, a.n_md, a.steps, a.warmup, a.replicas, profile=False, lazy=1)
        line["lazy_streaming"] = {"value": lz["ps_per_s"], "ms_per_step": lz["ms_per_cycle"],
                                  "over_eager": lz["ps_per_s"] / main_run["ps_per_s"]}
    if not a.no_mixed:
        line["batched_mixed_L"] = mixed_batch(a)
    if not a.no_single and a.replicas > 1:
        single = run_gpu(a.form, a.mode, a.L, a.n_md, a.steps, a.warmup, 1, profile=False)
        line["single_replica"] = {"value": single["ps_per_s"], "ms_per_step": single["ms_per_cycle"],
                                  "n_solvent": single["n"]}
    # the whole step's algorithmic bytes per particle-step: 9w per MD step plus the sorts
    n_sorts = (main_run["extra_sorts"] / max(1, main_run["cycles"] * a.replicas) + 1) / a.n_md
    S = 38 if a.form == "fp32" else 68
    line["step_bytes_per_particle_step"] = 9 * w + n_sorts * (2 * S + 16 / 10)
    line["step_effective_GBs"] = line["step_bytes_per_particle_step"] * main_run["ps_per_s"] / 1e9
    line["step_frac_of_hbm_peak"] = line["step_effective_GBs"] / HBM_PEAK_GBS
    if not a.no_colloids:
        line["n_colloids_L128"] = n_colloids(a.colloids_cycles)
    if not a.no_bodies:
        line["linked_bodies"] = linked_bodies(not a.no_cpu)
    if not a.no_variants:
        line["mpcd_variants"] = mpcd_variants(not a.no_cpu)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.L, a.n_md)
    print(json.dumps(line))


if __name__ == "__main__":
    if len(sys.argv) == 6 and sys.argv[1] == "--cpu-baseline-child":
        cpu_baseline_child(int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]), float(sys.argv[5]))
    else:
        main()
