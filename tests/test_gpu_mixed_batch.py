"""The batched single-dimer workload at its full size: 128 independent systems with box
sides drawn from {24, 32, 40, 48} (rho = 10, tau = 0.5, N_MD = 50; 67.5M solvent in all)
in one launch per kernel, programs/single_dimer_pbc.f90:303-445 for every replica.

* The FP32-mixed batch and the FP64 batch run from the same initial states and draws;
  per replica: temperature within 0.1 % of the FP64 form, momentum conserved (FP64 to
  dimer_test.py's 1e-15 per particle, FP32 to its rounding level), the dimer bond held
  at d by RATTLE, ids a permutation, and (FP64) cell counts equal to the oracle's count
  of the downloaded positions.
* The two largest box sizes (L = 40 and L = 48) of a batch run one collision cycle each
  against the oracle's loop on the same Threefry draws: integer state bit-exact, FP64
  within 1e-12.
"""
import ctypes as C

import numpy as np
import pytest

from rmpcdmd_amd import capi
from rmpcdmd_amd.sim import SingleDimer, mixed_L_sizes

pytestmark = pytest.mark.gpu

KW = dict(n_md=50, tau=0.5, seed=4242, mode=capi.SIM_FUSED)


def _temperature(sim, r):
    sol, col, box, cells, st = sim.state(r)
    te = C.c_double()
    capi.call("rmpcd_compute_temperature", C.byref(box), C.byref(cells), C.byref(sol), C.byref(te))
    return te.value


def _close(a, b, rtol=1e-12):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and bool((np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b))).all())


def _momentum(s, m):
    return (s["vel"].astype(np.float64).sum(axis=0) + m[0] * s["colloid_vel"][0]
            + m[1] * s["colloid_vel"][1])


def _bond(s, L):
    r = s["colloid_pos"][0] - s["colloid_pos"][1]
    r -= np.round(r / L) * L
    return float(np.sqrt(r @ r))


def test_full_mixed_L_batch(gpu):
    from oracle import oracle as O
    sizes = mixed_L_sizes()
    assert len(sizes) == 128 and set(sizes) == {24, 32, 40, 48}
    rl = [(q, q, q) for q in sizes]
    print("128 replicas:", {q: sizes.count(q) for q in (24, 32, 40, 48)}, flush=True)
    s64 = SingleDimer(form=capi.FP64, replica_L=rl, **KW)
    s32 = SingleDimer(form=capi.FP32, replica_L=rl, **KW)
    m = np.array(s64.state()[4].mass[:])
    p0 = {}
    for r in range(len(sizes)):
        a = s32.download(r)
        b = s64.download(r)
        assert a["n_solvent"] == b["n_solvent"]
        p0[r] = (_momentum(a, m), _momentum(b, m))
    n_total = sum(int(s64.state(r)[4].n_solvent) for r in range(len(sizes)))
    print("solvent in all: %d" % n_total, flush=True)
    assert n_total > 60_000_000
    cycles = 2
    worst = {"temperature": 0.0, "p64": 0.0, "p32": 0.0, "bond64": 0.0, "bond32": 0.0}
    for cycle in range(cycles):
        s64.run(1)
        s32.run(1)
        for r in range(len(sizes)):
            t64, t32 = _temperature(s64, r), _temperature(s32, r)
            worst["temperature"] = max(worst["temperature"], abs(t32 - t64) / t64)
        print("cycle %d done, worst temperature so far %.2e" % (cycle, worst["temperature"]), flush=True)
    for r, q in enumerate(sizes):
        a = s32.download(r)
        b = s64.download(r)
        msg = "replica %d (L=%d)" % (r, q)
        n = b["n_solvent"]
        for s in (a, b):
            assert s["cycles"] == cycles and s["md_steps"] == cycles * 50, msg
            assert s["error"] == 0, msg
            assert (np.sort(s["id"]) == np.arange(n)).all(), msg
            assert int(s["cell_count"].sum()) == n, msg
            assert (np.diff(s["cell_start"]) == s["cell_count"][:-1]).all(), msg
        # the cell of every FP64 position (the oracle's count) is the cell it is stored in
        c = O.cells_init((q, q, q))
        O.set_origin(c, b["origin"])
        cnt, st = O.count_particles(c, np.ascontiguousarray(b["pos"]))
        assert (cnt == b["cell_count"]).all() and (st == b["cell_start"]).all(), msg
        worst["p64"] = max(worst["p64"], float(np.abs(_momentum(b, m) - p0[r][1]).max()) / (n + m.sum()))
        worst["p32"] = max(worst["p32"], float(np.abs(_momentum(a, m) - p0[r][0]).max()) / (n + m.sum()))
        worst["bond64"] = max(worst["bond64"], abs(_bond(b, q) - 4.5))
        worst["bond32"] = max(worst["bond32"], abs(_bond(a, q) - 4.5))
        assert (b["species"] == 2).sum() > 0, msg
    print("128-replica batch, FP32-mixed against FP64 after %d cycles:" % cycles, worst)
    assert worst["temperature"] < 1e-3
    # dimer_test.py: centre-of-mass velocity drift below 1e-15 per particle in FP64
    assert worst["p64"] < 1e-15
    assert worst["p32"] < 1e-6
    assert worst["bond64"] < 1e-9 and worst["bond32"] < 1e-9
    s64.close()
    s32.close()


def test_large_boxes_of_a_batch_against_the_oracle(gpu):
    """one collision cycle of the L=40 and L=48 systems of a batch, each against the oracle's
    loop with that replica's Threefry stream (key seed + replica)"""
    from oracle.dimer import from_download
    from oracle.threefry import ThreefryDraws
    rl = [(40, 40, 40), (48, 48, 48)]
    sim = SingleDimer(form=capi.FP64, replica_L=rl, **KW)
    orcs = []
    for r, L in enumerate(rl):
        st = sim.download(r)
        o = from_download(st, L, 4.5, 0.5, 50)
        assert (o.nl == st["nlist"]).all()
        orcs.append(o)
    sim.run(1)
    for r, L in enumerate(rl):
        o = orcs[r]
        o.cycle(ThreefryDraws(KW["seed"] + r))
        g = sim.download(r)
        msg = "replica %d (L=%d), %d solvent" % (r, L[0], g["n_solvent"])
        print(msg, "extra sorts", g["n_extra_sorting"], flush=True)
        assert g["n_extra_sorting"] == o.n_extra and g["md_steps"] == o.md_steps, msg
        assert (g["id"] == o.s.id).all(), msg
        assert (g["species"] == o.s.species).all(), msg
        assert (g["flags"] == o.s.flags).all(), msg
        assert (g["image"] == o.s.image).all(), msg
        assert (g["colloid_image"] == o.col.image).all(), msg
        assert (g["cell_count"] == o.cnt).all() and (g["cell_start"] == o.st).all(), msg
        assert (g["nlist"] == o.nl).all(), msg
        for j in range(2):
            assert (g["list"][j, :g["nlist"][j]] == o.lst[j, :o.nl[j]]).all(), msg
        assert _close(g["origin"], o.c.origin[:]), msg
        for k in ("pos", "vel", "force", "force_old"):
            assert _close(g[k], getattr(o.s, k)), "%s %s" % (k, msg)
            assert _close(g["colloid_" + k], getattr(o.col, k)), "colloid %s %s" % (k, msg)
        assert _close(g["e1"], o.e1) and _close(g["e2"], o.e2), msg
    sim.close()
