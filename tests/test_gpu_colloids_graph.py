"""The graph-captured n_colloids_pbc production loop (programs/n_colloids_pbc.f90:286-346;
rmpcd_colloids.hip): every decision on the device, one hipGraph per collision cycle.  It
must reproduce the reference call sequence of the same program (integer state bit for
bit: sort order, ids, species, flags, images, cell counts, lists, the number of extra
sorts; FP64 to 1e-12, the solvent forces of particles in several lists being summed by
atomics in either mode), for an even and an odd N_MD (the graph of each starting force
parity), with many extra sorts (small skin), and the oracle's loop; a list overflow stops
it with the reference's message."""
import numpy as np
import pytest

from rmpcdmd_amd import capi
from rmpcdmd_amd.programs import PROG_COLLOIDS, Program

pytestmark = pytest.mark.gpu

INT_KEYS = ("id", "species", "flags", "image", "cell_count", "cell_start", "is_md", "nlist", "colloid_image")
REAL_KEYS = ("pos", "vel", "force", "force_old", "pos_old", "colloid_pos", "colloid_vel", "colloid_force",
             "colloid_force_old", "colloid_pos_old")


def close(a, b, rtol):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and bool((np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b))).all())


def compare(a, b, msg, rtol):
    for k in INT_KEYS:
        assert a[k].tobytes() == b[k].tobytes(), "%s %s" % (k, msg)
    for j in range(a["nlist"].size):
        n = a["nlist"][j]
        assert (a["list"][j, :n] == b["list"][j, :n]).all(), "list %d %s" % (j, msg)
    for k in REAL_KEYS:
        assert close(a[k], b[k], rtol), "%s %s" % (k, msg)
    for k in ("md_steps", "cycles", "n_extra_sorting", "last_sort"):
        assert a[k] == b[k], "%s %s: %s %s" % (k, msg, a[k], b[k])
    assert close(a["origin"], b["origin"], 0.0), msg
    for k in ("e1", "e2", "temperature"):
        assert close(a[k], b[k], rtol), "%s %s" % (k, msg)


@pytest.mark.parametrize("form,n_md,skin", [(capi.FP64, 10, 0.0), (capi.FP64, 5, 0.0), (capi.FP32, 10, 0.0),
                                            (capi.FP64, 10, 0.3), (capi.FP32, 7, 0.3)],
                         ids=["fp64", "fp64-odd-nmd", "fp32", "fp64-small-skin", "fp32-odd-small-skin"])
def test_graph_equals_reference_sequence(gpu, form, n_md, skin):
    kw = dict(L=(20, 20, 20), n_md=n_md, tau=0.2, sigma=1.5, n_colloids=12, T=1.0, skin=skin)
    ref = Program(PROG_COLLOIDS, form=form, seed=31, **kw)
    gra = Program(PROG_COLLOIDS, form=form, seed=31, graph=1, **kw)
    rtol = 1e-12 if form == capi.FP64 else 1e-5
    for cycle in range(3):
        ref.run(1)
        gra.run(1)
        compare(gra.download(), ref.download(), "after cycle %d" % cycle, rtol)
    # two cycles in one run: the graph is replayed (and, for an odd N_MD, both parities)
    ref.run(2)
    gra.run(2)
    a, b = gra.download(), ref.download()
    compare(a, b, "after a run of two cycles", rtol)
    print("extra sorts:", a["n_extra_sorting"], "md slots cells:", int(a["is_md"].sum()))
    if skin:
        assert a["n_extra_sorting"] > 5
    ref.close()
    gra.close()


@pytest.mark.parametrize("rho", [12, 16])
def test_graph_dense_cells(gpu, rho):
    """denser cells: the gather works on tiles of 128 cells with 1536 slots in LDS; at
    rho = 12 about half of the tiles hold more and rank (and collide) from memory, at
    rho = 16 all do"""
    kw = dict(L=(20, 20, 20), rho=rho, n_md=10, tau=0.2, sigma=1.5, n_colloids=12, T=1.0, skin=0.3)
    ref = Program(PROG_COLLOIDS, form=capi.FP64, seed=41, **kw)
    gra = Program(PROG_COLLOIDS, form=capi.FP64, seed=41, graph=1, **kw)
    for cycle in range(2):
        ref.run(1)
        gra.run(1)
        compare(gra.download(), ref.download(), "rho %d, after cycle %d" % (rho, cycle), 1e-12)
    assert gra.download()["n_extra_sorting"] > 0
    ref.close()
    gra.close()


def test_graph_matches_oracle(gpu):
    from oracle.programs import ColloidsOracle
    L = (20, 20, 20)
    kw = dict(L=L, n_md=10, tau=0.2, sigma=1.5, n_colloids=12, T=1.0)
    prog = Program(PROG_COLLOIDS, form=capi.FP64, seed=13, graph=1, **kw)
    st = prog.download()
    orc = ColloidsOracle(st, L, 0.2, 10, 13, sigma=1.5)
    for cycle in range(3):
        prog.run(1)
        orc.cycle()
        g = prog.download()
        msg = "after cycle %d" % cycle
        assert (g["id"] == orc.s.id).all(), msg
        assert (g["flags"] == orc.s.flags).all(), msg
        assert (g["cell_count"] == orc.cnt).all() and (g["cell_start"] == orc.st).all(), msg
        for k in ("pos", "vel", "force", "force_old"):
            assert close(g[k], getattr(orc.s, k), 1e-12), "%s %s" % (k, msg)
            assert close(g["colloid_" + k], getattr(orc.col, k), 1e-12), "colloid %s %s" % (k, msg)
        assert (g["nlist"] == orc.nl).all(), msg
        assert g["n_extra_sorting"] == orc.n_extra and g["last_sort"] == orc.last_sort, msg
        assert close(g["temperature"], orc.temperature, 1e-12), msg
    prog.close()


def test_graph_list_overflow_stops(gpu):
    kw = dict(L=(20, 20, 20), n_md=10, tau=0.2, sigma=1.5, n_colloids=12, T=1.0)
    probe = Program(PROG_COLLOIDS, form=capi.FP64, seed=37, **kw)
    n0 = int(probe.download()["nlist"].max())
    probe.close()
    prog = Program(PROG_COLLOIDS, form=capi.FP64, seed=37, graph=1, nmax_list=n0, **kw)
    with pytest.raises(capi.RmpcdError) as e:
        prog.run(6)
    assert e.value.code == capi.RMPCD_ERR_OVERFLOW
    assert "maximum of neighbor list reached" in e.value.msg
    g = prog.download()
    assert np.isfinite(g["pos"]).all() and (g["nlist"] <= n0).all()
    assert g["md_steps"] < 60
    prog.close()


def test_fp32_error_at_L128(gpu):
    """the n_colloids_pbc workload at its size (L=128, 1000 colloids sigma=2.5, 20.3M
    solvent): the FP32-mixed form against the FP64 form from the same state and draws,
    temperature within 0.1 % every cycle, and the FP32 form's total momentum kept to its
    rounding level (FP32 positions at L=128 carry an ulp of 2^-17)"""
    kw = dict(L=(128, 128, 128), n_md=50, tau=0.5, sigma=2.5, n_colloids=1000, T=1.0, graph=1)
    s64 = Program(PROG_COLLOIDS, form=capi.FP64, seed=20261202, **kw)
    s32 = Program(PROG_COLLOIDS, form=capi.FP32, seed=20261202, **kw)
    mass = float(s64.state()[4].mass)

    def momentum(p):
        sol, col, box, cells, st = p.state()
        d = p.download()
        return d["vel"].astype(np.float64).sum(axis=0) + mass * d["colloid_vel"].sum(axis=0), int(st.n_solvent)

    p0, n = momentum(s32)
    worst = {"temperature": 0.0, "colloid_pos": 0.0}
    for cycle in range(2):
        s64.run(1)
        s32.run(1)
        t64, t32 = s64.state()[4].temperature, s32.state()[4].temperature
        worst["temperature"] = max(worst["temperature"], abs(t32 - t64) / t64)
        d = np.abs(s32.download_colloids() - s64.download_colloids()).max()
        worst["colloid_pos"] = max(worst["colloid_pos"], float(d))
        print("cycle %d: T64 %.9f T32 %.9f, colloid positions differ by %.3e" % (cycle, t64, t32, d), flush=True)
    p1, _ = momentum(s32)
    drift = float(np.abs(p1 - p0).max()) / (n + 1000 * mass)
    print("FP32 vs FP64 at L=128:", worst, "momentum drift per particle %.2e" % drift)
    assert worst["temperature"] < 1e-3
    assert drift < 1e-6
    s64.close()
    s32.close()
