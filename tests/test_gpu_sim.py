"""The single-dimer program on the GPU (programs/single_dimer_pbc.f90): the fused,
graph-captured step must reproduce the reference call sequence (same draws, same sort
decisions; integer state bit for bit, floating point to the FP64 bar of 1e-12 relative,
as the colloid force sums run in another order), for one and for several replicas of
uneven size; both must match the oracle's loop; plus the invariants of dimer_test.py
(momentum conservation) and RATTLE."""
import numpy as np
import pytest

from rmpcdmd_amd import capi
from rmpcdmd_amd.sim import SingleDimer

pytestmark = pytest.mark.gpu

KEYS = ("pos", "vel", "force", "force_old", "image", "id", "species", "flags", "colloid_pos",
        "colloid_vel", "colloid_force", "colloid_force_old", "colloid_image", "cell_count", "cell_start",
        "nlist", "origin")


def _close(a, b, rtol=1e-12):
    """|a - b| <= rtol * max(1, |b|) elementwise (the FP64 bar of the whole-step check)"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and bool((np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b))).all())


def compare(a, b, replica_msg="", rtol=1e-12):
    for k in KEYS:
        assert a[k].shape == b[k].shape, k
        if a[k].dtype.kind == "f":
            assert _close(a[k], b[k], rtol), "%s differs %s" % (k, replica_msg)
        else:
            assert a[k].tobytes() == b[k].tobytes(), "%s differs %s" % (k, replica_msg)
    for i in range(2):
        n = a["nlist"][i]
        assert (a["list"][i, :n] == b["list"][i, :n]).all()
    for k in ("md_steps", "cycles", "n_extra_sorting"):
        assert a[k] == b[k], k
    assert _close(a["e1"], b["e1"], rtol) and _close(a["e2"], b["e2"], rtol)


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_fused_equals_reference_sequence(gpu, form):
    kw = dict(L=(16, 16, 16), n_md=10, tau=0.1, form=form, seed=77)
    ref = SingleDimer(mode=capi.SIM_REFERENCE, **kw)
    fus = SingleDimer(mode=capi.SIM_FUSED, **kw)
    for cycle in range(3):
        ref.run(1)
        fus.run(1)
        a, b = ref.download(), fus.download()
        # FP32: the solvent state is single precision, the colloids FP64
        compare(a, b, "after cycle %d" % cycle, 1e-12 if form == capi.FP64 else 1e-5)
    assert a["n_extra_sorting"] >= 1  # the first step re-sorts (pos_old is the set-up order)
    ref.close()
    fus.close()


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_fused_list_overflow_stops(gpu, form):
    """a list longer than its capacity stops the replica at that MD step with the
    reference's message (neighbor_list.f90:135-137) instead of running on with a
    truncated list: the state stays readable, later runs change nothing, and the other
    replica of the batch keeps going (unless it overflows too).  The capacity is the
    longest set-up list, so one of the rebuilds that the small skin asks for exceeds it."""
    kw = dict(L=(16, 16, 16), n_md=10, tau=0.1, form=form, seed=91, skin=0.6)
    probe = SingleDimer(mode=capi.SIM_FUSED, n_replicas=2, **kw)
    n0 = max(int(probe.download(r)["nlist"].max()) for r in range(2))
    probe.close()
    print("list capacity", n0, flush=True)
    sim = SingleDimer(mode=capi.SIM_FUSED, nmax_list=n0, n_replicas=2, **kw)
    with pytest.raises(capi.RmpcdError) as e:
        sim.run(8)
    assert e.value.code == capi.RMPCD_ERR_OVERFLOW
    assert "maximum of neighbor list reached" in e.value.msg
    d = [sim.download(r) for r in range(2)]
    stopped = [r for r in range(2) if d[r]["error"] == capi.RMPCD_ERR_OVERFLOW]
    assert stopped, [x["error"] for x in d]
    for r in range(2):
        assert np.isfinite(d[r]["pos"]).all() and np.isfinite(d[r]["vel"]).all()
        assert (d[r]["nlist"] <= n0).all()
        if r in stopped:
            assert d[r]["md_steps"] < 8 * 10
        else:
            assert d[r]["md_steps"] == 8 * 10 and d[r]["cycles"] == 8
    with pytest.raises(capi.RmpcdError):
        sim.run(1)
    again = [sim.download(r) for r in range(2)]
    for r in stopped:
        for k in ("pos", "vel", "species", "id"):
            assert again[r][k].tobytes() == d[r][k].tobytes(), k
        assert again[r]["md_steps"] == d[r]["md_steps"]
    sim.close()


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_fused_dense_cells(gpu, form):
    """rho = 40: cells hold more particles than the gather and rank passes stage on either
    side of their tiles, so the ranks of those cells read the rest from memory, and the
    collision blocks overflow their LDS tiles; the fused program must still equal the
    reference call sequence, and its odd-grid run, its run with a contiguous unit range per
    workgroup (instead of XCD-grouped units) and its runs with 64 and 256 cells per
    collision workgroup the tuned one bit for bit"""
    kw = dict(L=(12, 12, 12), rho=40, n_md=10, tau=0.1, form=form, seed=123)
    ref = SingleDimer(mode=capi.SIM_REFERENCE, **kw)
    fus = SingleDimer(mode=capi.SIM_FUSED, **kw)
    odd = SingleDimer(mode=capi.SIM_FUSED, sort_blocks=96, extra_sort_blocks=7, **kw)
    flat = SingleDimer(mode=capi.SIM_FUSED, sort_order=1, **kw)
    tiles = [SingleDimer(mode=capi.SIM_FUSED, collide_cells=ct, **kw) for ct in (64, 256)]
    for cycle in range(2):
        ref.run(1)
        fus.run(1)
        odd.run(1)
        flat.run(1)
        for t in tiles:
            t.run(1)
        a, b, d, e = ref.download(), fus.download(), odd.download(), flat.download()
        assert (a["cell_count"] > 32).sum() > 100  # cells larger than the 32-slot halo
        compare(a, b, "after cycle %d" % cycle, 1e-12 if form == capi.FP64 else 1e-5)
        compare(b, d, "odd grids after cycle %d" % cycle, rtol=0.0)
        compare(b, e, "contiguous unit ranges after cycle %d" % cycle, rtol=0.0)
        for t, ct in zip(tiles, (64, 256)):
            compare(b, t.download(), "%d cells per collision workgroup after cycle %d" % (ct, cycle), rtol=0.0)
    ref.close()
    fus.close()
    odd.close()
    flat.close()
    for t in tiles:
        t.close()


def test_fused_replicas_of_uneven_size(gpu):
    Ls = [(12, 12, 12), (16, 12, 14), (12, 16, 12)]
    kw = dict(n_md=10, tau=0.1, form=capi.FP64, seed=5, replica_L=Ls)
    ref = SingleDimer(mode=capi.SIM_REFERENCE, **kw)
    fus = SingleDimer(mode=capi.SIM_FUSED, **kw)
    ref.run(2)
    fus.run(2)
    for r in range(len(Ls)):
        compare(ref.download(r), fus.download(r), "replica %d" % r)
    # replicas are independent systems with their own seeds
    a0, a2 = fus.download(0), fus.download(2)
    assert not np.array_equal(a0["vel"][:100], a2["vel"][:100])
    ref.close()
    fus.close()


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_fused_streaming_without_vectors(gpu, form):
    """a replica whose particle count is not a multiple of 4 (13^3 cells at rho = 10) streams
    one particle per thread and sizes the streaming grid for that; the replica before it
    keeps its 16-byte accesses"""
    Ls = [(12, 12, 12), (13, 13, 13)]
    kw = dict(n_md=10, tau=0.1, form=form, seed=9, replica_L=Ls)
    ref = SingleDimer(mode=capi.SIM_REFERENCE, **kw)
    fus = SingleDimer(mode=capi.SIM_FUSED, **kw)
    ref.run(2)
    fus.run(2)
    for r in range(len(Ls)):
        compare(ref.download(r), fus.download(r), "replica %d" % r, 1e-12 if form == capi.FP64 else 1e-5)
    assert fus.download(1)["pos"].shape[0] % 4 != 0
    ref.close()
    fus.close()


def test_dimer_invariants(gpu):
    sim = SingleDimer(L=(20, 20, 20), form=capi.FP64, mode=capi.SIM_FUSED, n_md=25, tau=0.25, seed=3)
    st = sim.state()[4]
    m = np.array(st.mass[:])

    def momentum(s):
        return s["vel"].sum(axis=0) + m[0] * s["colloid_vel"][0] + m[1] * s["colloid_vel"][1]

    s0 = sim.download()
    p0 = momentum(s0)
    sim.run(4)
    s1 = sim.download()
    p1 = momentum(s1)
    # dimer_test.py: centre-of-mass velocity drift below 1e-15 per particle (FP64)
    assert np.abs(p1 - p0).max() / (s1["n_solvent"] + m.sum()) < 1e-15
    r = s1["colloid_pos"][0] - s1["colloid_pos"][1]
    r -= np.round(r / 20.0) * 20.0
    assert abs(np.sqrt(r @ r) - 4.5) < 1e-9
    assert s1["cycles"] == 4 and s1["md_steps"] == 100
    # the catalytic sphere converted some A into B
    assert (s1["species"] == 2).sum() > 0
    sim.close()


@pytest.mark.parametrize("mode", [capi.SIM_REFERENCE, capi.SIM_FUSED])
@pytest.mark.parametrize("bulk", [0, 1])
def test_sequence_matches_oracle(gpu, bulk, mode):
    """the whole single_dimer_pbc cycle (MD steps with sort decisions, forces, RATTLE,
    change_species; grid shift, sort, SRD, refuel or bulk reaction) on the GPU against
    the oracle's loop consuming the same Threefry draws. Integer state (sort order, ids,
    species, flags, images, cell counts, lists) bit-exact; FP64 within 1e-12 relative:
    the colloid force sums are tree-reduced on the GPU, everything else keeps the
    reference's evaluation order."""
    from oracle.dimer import from_download
    from oracle.threefry import ThreefryDraws
    L = (12, 12, 12)
    kw = dict(L=L, n_md=10, tau=0.1, d=4.5, sigma_C=2.0, sigma_N=2.0, bulk_rmpcd=bulk, bulk_rate=0.5)
    sim = SingleDimer(mode=mode, form=capi.FP64, seed=41, **kw)
    st = sim.download()
    orc = from_download(st, L, 4.5, 0.1, 10, bulk_rmpcd=bool(bulk), bulk_rate=0.5)
    assert (orc.nl == st["nlist"]).all()
    for j in range(2):
        assert (orc.lst[j, :orc.nl[j]] == st["list"][j, :st["nlist"][j]]).all()
    draws = ThreefryDraws(41)
    for cycle in range(3):
        sim.run(1)
        orc.cycle(draws)
        g = sim.download()
        msg = "after cycle %d" % cycle
        assert g["n_extra_sorting"] == orc.n_extra and g["md_steps"] == orc.md_steps, msg
        assert (g["id"] == orc.s.id).all(), msg
        assert (g["species"] == orc.s.species).all(), msg
        assert (g["flags"] == orc.s.flags).all(), msg
        assert (g["image"] == orc.s.image).all(), msg
        assert (g["colloid_image"] == orc.col.image).all(), msg
        assert (g["cell_count"] == orc.cnt).all() and (g["cell_start"] == orc.st).all(), msg
        assert (g["nlist"] == orc.nl).all(), msg
        for j in range(2):
            assert (g["list"][j, :g["nlist"][j]] == orc.lst[j, :orc.nl[j]]).all(), msg
        assert _close(g["origin"], orc.c.origin[:]), msg
        for k in ("pos", "vel", "force", "force_old"):
            assert _close(g[k], getattr(orc.s, k)), "%s %s" % (k, msg)
            assert _close(g["colloid_" + k], getattr(orc.col, k)), "colloid %s %s" % (k, msg)
        assert _close(g["e1"], orc.e1) and _close(g["e2"], orc.e2), msg
    if bulk:
        assert (g["species"] == 2).sum() > 0
    sim.close()


def test_batched_replicas_of_mixed_sizes(gpu):
    """replicas of unequal L (the batched configuration draws L from {24, 32, 40, 48}),
    FP64, tau=0.5 and N_MD=50 as in the dimer runs: the fused batch against the reference
    call sequence replica by replica, and replica 0 of the batch against the same system
    run on its own"""
    Ls = [(24, 24, 24), (32, 32, 32), (24, 24, 24), (32, 32, 32)]
    kw = dict(n_md=50, tau=0.5, form=capi.FP64, seed=17)
    ref = SingleDimer(mode=capi.SIM_REFERENCE, replica_L=Ls, **kw)
    fus = SingleDimer(mode=capi.SIM_FUSED, replica_L=Ls, **kw)
    one = SingleDimer(mode=capi.SIM_FUSED, L=Ls[0], **kw)
    # the batch as one graph branch, and as three branches of unequal replica counts (the
    # groups split the particle count evenly): the same trajectories bit for bit
    g1 = SingleDimer(mode=capi.SIM_FUSED, replica_L=Ls, groups=1, **kw)
    g3 = SingleDimer(mode=capi.SIM_FUSED, replica_L=Ls, groups=3, **kw)
    ref.run(1)
    fus.run(1)
    one.run(1)
    g1.run(1)
    g3.run(1)
    for r in range(len(Ls)):
        compare(ref.download(r), fus.download(r), "replica %d" % r)
        compare(fus.download(r), g1.download(r), "one group, replica %d" % r, rtol=0.0)
        compare(fus.download(r), g3.download(r), "three groups, replica %d" % r, rtol=0.0)
    compare(fus.download(0), one.download(0), "replica 0 alone")
    ref.close()
    fus.close()
    one.close()
    g1.close()
    g3.close()


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_lazy_streaming_equals_eager(gpu, form):
    """lazy mode replays the per-step additions of the particles outside the lists when a
    sort reads them: the trajectory must be the eager one bit for bit, for every replica"""
    Ls = [(16, 16, 16), (16, 12, 14)]
    kw = dict(n_md=20, tau=0.4, form=form, seed=23, replica_L=Ls, mode=capi.SIM_FUSED)
    eager = SingleDimer(**kw)
    lazy = SingleDimer(lazy=1, **kw)
    for cycle in range(3):
        eager.run(1)
        lazy.run(1)
        for r in range(len(Ls)):
            a, b = eager.download(r), lazy.download(r)
            compare(a, b, "replica %d cycle %d" % (r, cycle), rtol=0.0)
            assert a["pos_old"].tobytes() == b["pos_old"].tobytes()
    assert a["n_extra_sorting"] >= 1
    eager.close()
    lazy.close()


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_many_extra_sorts_per_cycle(gpu, form):
    """a small skin (0.6) makes the skin test fire several times per collision cycle, so
    extra sorts follow each other before the collision sort (the fused mode keeps the
    particles in place at an extra sort and carries the order in perm/inv): the fused batch
    must still equal the reference call sequence, and lazy streaming the eager one"""
    Ls = [(16, 16, 16), (20, 16, 12)]
    kw = dict(n_md=50, tau=0.5, form=form, seed=31, replica_L=Ls, skin=0.6)
    ref = SingleDimer(mode=capi.SIM_REFERENCE, **kw)
    fus = SingleDimer(mode=capi.SIM_FUSED, **kw)
    lazy = SingleDimer(mode=capi.SIM_FUSED, lazy=1, **kw)
    # odd grids for the sort passes (units are split over whatever grid there is): the
    # same trajectory bit for bit
    odd = SingleDimer(mode=capi.SIM_FUSED, sort_blocks=96, extra_sort_blocks=7, **kw)
    # a contiguous unit range per workgroup instead of XCD-grouped units, in one graph
    # branch instead of one per replica: the same again
    flat = SingleDimer(mode=capi.SIM_FUSED, sort_order=1, sort_blocks=64, extra_sort_blocks=16, groups=1, **kw)
    for cycle in range(2):
        ref.run(1)
        fus.run(1)
        lazy.run(1)
        odd.run(1)
        flat.run(1)
        for r in range(len(Ls)):
            a, b, c, d = ref.download(r), fus.download(r), lazy.download(r), odd.download(r)
            compare(a, b, "replica %d cycle %d" % (r, cycle), 1e-12 if form == capi.FP64 else 1e-5)
            compare(b, c, "lazy replica %d cycle %d" % (r, cycle), rtol=0.0)
            compare(b, d, "odd grids replica %d cycle %d" % (r, cycle), rtol=0.0)
            compare(b, flat.download(r), "contiguous unit ranges replica %d cycle %d" % (r, cycle), rtol=0.0)
    for r in range(len(Ls)):
        assert fus.download(r)["n_extra_sorting"] >= 6, "replica %d" % r
    ref.close()
    fus.close()
    lazy.close()
    odd.close()
    flat.close()
