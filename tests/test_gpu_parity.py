"""Parity of the HIP path (through the C ABI) with the oracle, function by function,
FP64 form.  Integer/index outputs must be bit-exact; floating-point outputs are held
to 1e-12 relative (most come out bit-exact: the library is built without FMA
contraction and keeps the reference's evaluation order)."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rmpcdmd_amd import capi
from rmpcdmd_amd.system import CellSet, NeighborList, ParticleSet, lj_params

from _helpers import cells_to_device, same_bits, solvent_state, sorted_state, to_device

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def ptr(t):
    return C.c_void_p(t.data_ptr())


def close(a, b, rtol=RTOL, scale=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.maximum(np.abs(b), 1.0) if scale is None else scale
    return np.all(np.abs(a - b) <= rtol * s)


# ---- cells ------------------------------------------------------------------------
@pytest.mark.parametrize("L", [(8, 5, 5), (32, 32, 32), (24, 40, 48), (120, 30, 30), (1, 20, 5)])
def test_hilbert_table(gpu, L):
    cs = CellSet(L)
    keys = torch.zeros(int(np.prod(L)), dtype=torch.int32, device="cuda")
    capi.call("rmpcd_hilbert_table", cs.boxp(), ptr(keys))
    torch.cuda.synchronize()
    got = keys.cpu().numpy()
    M = tuple(cs.box.M)
    idx = np.random.default_rng(1).choice(got.size, min(got.size, 3000), replace=False)
    for t in idx:
        p = (t % L[0], (t // L[0]) % L[1], t // (L[0] * L[1]))
        assert got[t] == O.compact_p_to_h(p, M)
    assert len(np.unique(got)) == got.size


@pytest.mark.parametrize("L,rho,seed", [((8, 5, 5), 10, 1), ((32, 32, 32), 10, 2), ((24, 40, 48), 3, 3)])
def test_count_particles(gpu, L, rho, seed):
    ps = solvent_state(L, rho, seed)
    c = O.cells_init(L)
    O.set_origin(c, np.random.default_rng(seed).random(3) - 1)
    cnt, st = O.count_particles(c, ps.pos)
    d = to_device(ps)
    cs = cells_to_device(c)
    capi.call("rmpcd_count_particles", cs.boxp(), d.struct(), cs.cellsp())
    torch.cuda.synchronize()
    assert same_bits(cs.cell_count.cpu().numpy(), cnt)
    assert same_bits(cs.cell_start.cpu().numpy(), st)
    # test_cell_system_0 invariants
    assert cs.cell_start[0].item() == 0 and cs.cell_count.sum().item() == ps.n


def test_cartesian_indices_boundaries(gpu):
    # positions on cell faces and at the wrap p == L (cell_system.f90:170-173)
    L = (8, 5, 5)
    c = O.cells_init(L)
    O.set_origin(c, (-0.25, -1.0, -0.5))
    xs = np.array([[0, 0, 0], [7.75, 4.0, 4.5], [7.999999999999999, 4.999999999, 4.4999],
                   [3.0, 2.0, 1.5], [0.75, 0.0, 0.5]], np.float64)
    ps = O.PState(len(xs))
    ps.pos[:] = xs
    d = to_device(ps)
    cs = cells_to_device(c)
    out = torch.zeros(3, len(xs), dtype=torch.int32, device="cuda")
    capi.call("rmpcd_cartesian_indices", cs.boxp(), d.struct(), ptr(out))
    torch.cuda.synchronize()
    got = out.cpu().numpy().T
    for i, x in enumerate(xs):
        assert tuple(got[i]) == O.cartesian_indices(c, x)


@pytest.mark.parametrize("L,rho,seed,wrap", [((8, 5, 5), 10, 4, 0), ((32, 32, 32), 10, 5, 0),
                                             ((24, 40, 48), 5, 6, 1), ((16, 16, 16), 10, 7, 1)])
def test_sort(gpu, L, rho, seed, wrap):
    ps = solvent_state(L, rho, seed)
    rng = np.random.default_rng(seed)
    if wrap:  # particles that left the box since the last wrap
        ps.pos[:] += rng.uniform(-1.5, 1.5, size=ps.pos.shape)
    c = O.cells_init(L)
    O.set_origin(c, rng.random(3) - 1)
    ref_in = ps.copy()
    if wrap:
        O.apply_pbc(ref_in.pos, ref_in.image, c.edges)
    srt, i2i, cnt, st = O.sort(c, ref_in)
    src = to_device(ps)
    dst = ParticleSet(ps.n)
    i2i_d = torch.zeros(ps.n, dtype=torch.int32, device="cuda")
    cs = cells_to_device(c)
    capi.call("rmpcd_sort", cs.boxp(), src.struct(), dst.struct(), ptr(i2i_d), cs.cellsp(), wrap)
    torch.cuda.synchronize()
    assert same_bits(cs.cell_count.cpu().numpy(), cnt)
    assert same_bits(cs.cell_start.cpu().numpy(), st)
    for name in ("pos", "vel", "force", "force_old", "image", "id", "species", "flags"):
        assert same_bits(dst.host(name), getattr(srt, name)), name
    assert same_bits(i2i_d.cpu().numpy(), i2i)


def test_sort_outside_box_and_empty_cells(gpu):
    # sparse particles (most cells empty) and a few outside [0,L): the reference keys
    # them through the low bits of p; the GPU must agree and stay in bounds
    L = (12, 7, 9)
    ps = solvent_state(L, 0.2, 11)
    ps.pos[:5] = [[-0.3, 1, 1], [12.6, 3, 3], [5, -0.9, 8.95], [11.99, 6.99, 9.7], [3, 3, -1.2]]
    c = O.cells_init(L)
    O.set_origin(c, (-0.1, -0.9, -0.5))
    srt, i2i, cnt, st = O.sort(c, ps)
    src = to_device(ps)
    dst = ParticleSet(ps.n)
    i2i_d = torch.zeros(ps.n, dtype=torch.int32, device="cuda")
    cs = cells_to_device(c)
    capi.call("rmpcd_sort", cs.boxp(), src.struct(), dst.struct(), ptr(i2i_d), cs.cellsp(), 0)
    torch.cuda.synchronize()
    assert same_bits(cs.cell_count.cpu().numpy(), cnt)
    assert same_bits(dst.host("pos"), srt.pos)
    assert same_bits(dst.host("id"), srt.id)


def test_maximum_displacement(gpu):
    ps = solvent_state((10, 10, 10), 10, 12)
    old = ps.pos.copy()
    ps.pos[:] += np.random.default_rng(3).normal(size=ps.pos.shape) * 0.05
    ps.species[::7] = -1  # excluded (species < 0)
    want = O.maximum_displacement(ps.pos, old, ps.species)
    d = to_device(ps)
    old_d = torch.as_tensor(np.ascontiguousarray(old.T)).cuda()
    r = C.c_double()
    capi.call("rmpcd_maximum_displacement", d.struct(), ptr(old_d), C.byref(r))
    assert r.value == want


def test_cell_maximum_displacement(gpu):
    c, srt, cnt, st = sorted_state((10, 10, 10), 10, 13)
    old = srt.pos.copy()
    srt.pos[:] += np.random.default_rng(4).normal(size=srt.pos.shape) * 0.05
    is_md = (np.random.default_rng(5).random(c.N) < 0.1).astype(np.uint8)
    c.max_v = 0.3
    want = O.cell_maximum_displacement(c, cnt, st, is_md, srt.pos, old, srt.species, 0.2)
    d = to_device(srt)
    cs = cells_to_device(c, cnt, st)
    cs.is_md.copy_(torch.as_tensor(is_md))
    old_d = torch.as_tensor(np.ascontiguousarray(old.T)).cuda()
    r = C.c_double()
    capi.call("rmpcd_cell_maximum_displacement", cs.boxp(), cs.cellsp(), d.struct(), ptr(old_d), 0.2,
              C.byref(r))
    assert r.value == want


# ---- mpcd -------------------------------------------------------------------------
def test_rand_sphere(gpu):
    rng = np.random.default_rng(21)
    n_items, stride = 5000, 8
    u = rng.random((n_items, stride))
    ud = torch.as_tensor(u).cuda()
    axes = torch.zeros(3, n_items, dtype=torch.float64, device="cuda")
    used = torch.zeros(n_items, dtype=torch.int32, device="cuda")
    capi.call("rmpcd_rand_sphere", ptr(ud), stride, n_items, ptr(axes), ptr(used))
    torch.cuda.synchronize()
    ax = axes.cpu().numpy().T
    us = used.cpu().numpy()
    for i in range(0, n_items, 7):
        n, k = O.rand_sphere(u[i])
        assert us[i] == k
        if k > 0:
            assert same_bits(ax[i], n)


@pytest.mark.parametrize("alpha", [None, 2.0])
@pytest.mark.parametrize("L,frac2", [((8, 5, 5), 0.0), ((20, 20, 20), 0.3), ((24, 40, 48), 0.0)])
def test_simple_mpcd_step(gpu, L, frac2, alpha):
    c, srt, cnt, st = sorted_state(L, 10, 14, frac2=frac2)
    srt.species[::13] = 0  # species 0 is left out of the mean but still rotated
    u = np.random.default_rng(15).random(8 * c.N + 64)
    axes, used = O.srd_axes_from_stream(c, cnt, st, srt.species, u)
    assert used > 0
    d = to_device(srt)
    cs = cells_to_device(c, cnt, st)
    ax_d = torch.as_tensor(np.ascontiguousarray(axes.T)).cuda()
    want = srt.vel.copy()
    O.simple_mpcd_step(c, cnt, st, want, srt.species, axes, alpha)
    capi.call("rmpcd_simple_mpcd_step", cs.boxp(), cs.cellsp(), d.struct(), ptr(ax_d), 0, 0,
              0 if alpha is None else 1, 0.0 if alpha is None else alpha, 0, 1.0, 1)
    torch.cuda.synchronize()
    assert close(d.host("vel"), want)


def test_srd_threefry_conserves_momentum(gpu):
    c, srt, cnt, st = sorted_state((16, 16, 16), 10, 16)
    d = to_device(srt)
    cs = cells_to_device(c, cnt, st)
    p0 = d.vel.sum(dim=1).cpu().numpy()
    e0 = (d.vel ** 2).sum().item()
    capi.call("rmpcd_simple_mpcd_step", cs.boxp(), cs.cellsp(), d.struct(), None, 1234, 7, 0, 0.0, 0, 1.0, 1)
    torch.cuda.synchronize()
    v = d.vel.cpu().numpy()
    assert np.allclose(v.sum(axis=1), p0, rtol=0, atol=1e-9)
    assert abs((v ** 2).sum() - e0) < 1e-9 * e0
    assert not np.allclose(v, srt.vel.T)


def test_compute_temperature(gpu):
    c, srt, cnt, st = sorted_state((20, 20, 20), 10, 17, frac2=0.2)
    srt.species[::11] = 0
    want = O.compute_temperature(c, cnt, st, srt.vel, srt.species)
    d = to_device(srt)
    cs = cells_to_device(c, cnt, st)
    te = C.c_double()
    capi.call("rmpcd_compute_temperature", cs.boxp(), cs.cellsp(), d.struct(), C.byref(te))
    assert abs(te.value - want) <= RTOL * abs(want)


@pytest.mark.parametrize("form", [capi.FP64, capi.FP32])
def test_compute_temperature_profile(gpu, form):
    """compute_temperature with its tz profile: the temperature and the bins of the
    oracle, binning into the caller's (accumulated, not cleared) arrays"""
    c, srt, cnt, st = sorted_state((12, 10, 16), 10, 19, origin=(-0.2, -0.7, -0.35), frac2=0.2)
    srt.species[::9] = 0
    d = to_device(srt, form=form)
    if form == capi.FP32:
        srt.vel[:] = d.host("vel").astype(np.float64)
    want_te = O.compute_temperature(c, cnt, st, srt.vel, srt.species)
    xmin, dx, nb = -0.5, 1.5, 12
    want_d, want_c = O.temperature_profile(c, cnt, st, srt.vel, srt.species, xmin, dx, nb)
    cs = cells_to_device(c, cnt, st)
    data = torch.full((nb,), 0.25, dtype=torch.float64, device="cuda")
    count = torch.full((nb,), 3, dtype=torch.int32, device="cuda")
    te = C.c_double()
    capi.call("rmpcd_compute_temperature_profile", cs.boxp(), cs.cellsp(), d.struct(), C.byref(te), xmin, dx,
              nb, ptr(data), ptr(count))
    tol = RTOL if form == capi.FP64 else 1e-5
    assert abs(te.value - want_te) <= tol * abs(want_te)
    assert (count.cpu().numpy() == want_c + 3).all()
    assert close(data.cpu().numpy(), want_d + 0.25, rtol=tol)
    assert want_c.sum() > 0


def test_stream_periodic(gpu):
    ps = solvent_state((8, 5, 5), 10, 18)
    c = O.cells_init((8, 5, 5))
    d = to_device(ps)
    cs = cells_to_device(c)
    want_pos, want_im = ps.pos.copy(), ps.image.copy()
    O.mpcd_stream_periodic(c, want_pos, ps.vel, want_im, 0.7)
    capi.call("rmpcd_mpcd_stream_periodic", cs.boxp(), d.struct(), 0.7)
    torch.cuda.synchronize()
    assert same_bits(d.host("pos"), want_pos)
    assert same_bits(d.host("image"), want_im)


def test_bulk_reaction(gpu):
    c, srt, cnt, st = sorted_state((16, 16, 16), 10, 19, frac2=0.5)
    draws = np.random.default_rng(20).random(c.N) * 0.1
    is_reac = (np.random.default_rng(21).random(c.N) < 0.8).astype(np.uint8)
    want = srt.species.copy()
    O.bulk_reaction(c, cnt, st, is_reac, want, 2, 1, 0.01, 1.0, draws)
    assert (want != srt.species).sum() > 10
    d = to_device(srt)
    cs = cells_to_device(c, cnt, st)
    cs.is_reac.copy_(torch.as_tensor(is_reac))
    dr = torch.as_tensor(draws).cuda()
    capi.call("rmpcd_bulk_reaction", cs.boxp(), cs.cellsp(), d.struct(), 2, 1, 0.01, 1.0, ptr(dr), 0, 0)
    torch.cuda.synchronize()
    assert same_bits(d.host("species"), want)


def test_stream_zwall(gpu):
    ps = solvent_state((6, 6, 6), 10, 22)
    ps.vel[:] *= 3
    c = O.cells_init((6, 6, 6), has_walls=True)
    d = to_device(ps)
    cs = cells_to_device(c)
    want = ps.copy()
    O.mpcd_stream_nogravity_zwall(c, want.pos, want.vel, want.force, want.flags, 0.5)
    capi.call("rmpcd_mpcd_stream_nogravity_zwall", cs.boxp(), d.struct(), 0.5)
    torch.cuda.synchronize()
    for name in ("pos", "vel", "flags"):
        assert same_bits(d.host(name), getattr(want, name)), name


# ---- md ---------------------------------------------------------------------------
def test_md_pos_vel_pbc(gpu):
    ps = solvent_state((10, 7, 9), 10, 23)
    ps.flags[::3] |= 1 << O.WALL_BIT
    d = to_device(ps)
    want = ps.copy()
    O.md_pos(want.pos, want.vel, want.force, 0.013)
    capi.call("rmpcd_md_pos", d.struct(), 0.013)
    O.md_vel(want.vel, want.force, want.force_old, want.flags, 0.013)
    capi.call("rmpcd_md_vel", d.struct(), 0.013)
    want.pos[:] += 9.3
    d.pos += 9.3
    torch.cuda.synchronize()
    O.apply_pbc(want.pos, want.image, (10, 7, 9))
    capi.call("rmpcd_apply_pbc", d.struct(), capi.d3((10, 7, 9)))
    torch.cuda.synchronize()
    for name in ("pos", "vel", "flags", "image"):
        assert same_bits(d.host(name), getattr(want, name)), name


@pytest.mark.parametrize("axis,bc", [(0, 0), (1, O.BOUNCE_BACK_BC), (1, O.SPECULAR_BC), (2, O.BOUNCE_BACK_BC)])
def test_cell_md(gpu, axis, bc):
    L = (1, 20, 5) if axis == 1 else (6, 6, 6)
    c, srt, cnt, st = sorted_state(L, 10, 24)
    if axis:
        c.bc[axis] = bc
        srt.vel[:] *= 10
    is_md = (np.random.default_rng(25).random(c.N) < 0.5).astype(np.uint8)
    srt.flags[::5] |= 1 << O.WALL_BIT
    d = to_device(srt)
    cs = cells_to_device(c, cnt, st)
    cs.is_md.copy_(torch.as_tensor(is_md))
    want = srt.copy()
    for flag in (True, False):
        if axis:
            O.cell_md_pos_wall(c, cnt, st, is_md, want.pos, want.vel, want.force, 0.3, flag, axis)
            capi.call("rmpcd_cell_md_pos_wall", cs.boxp(), cs.cellsp(), d.struct(), 0.3, int(flag), axis)
        else:
            O.cell_md_pos(c, cnt, st, is_md, want.pos, want.vel, want.force, 0.3, flag)
            capi.call("rmpcd_cell_md_pos", cs.boxp(), cs.cellsp(), d.struct(), 0.3, int(flag))
        O.cell_md_vel(c, cnt, st, is_md, want.vel, want.force, want.force_old, want.flags, 0.3, flag)
        capi.call("rmpcd_cell_md_vel", cs.boxp(), cs.cellsp(), d.struct(), 0.3, int(flag))
    torch.cuda.synchronize()
    for name in ("pos", "vel", "flags"):
        assert same_bits(d.host(name), getattr(want, name)), name


def colloid_pair(L, seed):
    rng = np.random.default_rng(seed)
    cp = O.PState(2)
    cp.pos[0] = np.asarray(L) / 2
    u = rng.normal(size=3)
    cp.pos[1] = cp.pos[0] + 4.5 * u / np.linalg.norm(u)
    cp.vel[:] = rng.normal(size=(2, 3)) * 0.01
    cp.force[:] = rng.normal(size=(2, 3)) * 5
    cp.force_old[:] = rng.normal(size=(2, 3)) * 5
    cp.species[:] = [1, 2]
    return cp


def test_rattle_and_colloid_verlet(gpu):
    L = (32.0, 32.0, 32.0)
    cp = colloid_pair(L, 26)
    mass = np.array([335.10321638291124, 335.10321638291124])
    mbp = mass.copy()
    want = cp.copy()
    rattle = cp.pos.copy()
    O.colloid_md_pos(want.pos, want.vel, want.force, mbp, 0.01)
    O.rattle_dimer_pos(want.pos, want.vel, rattle, want.species, mass, 4.5, 0.01, L)
    O.colloid_md_vel(want.vel, want.force, want.force_old, mbp, 0.01)
    O.rattle_dimer_vel(want.pos, want.vel, want.species, mass, 4.5, 0.01, L)
    d = to_device(cp)
    rat = torch.as_tensor(np.ascontiguousarray(rattle.T)).cuda()
    mb = torch.as_tensor(mbp).cuda()
    m = (C.c_double * 2)(*mass)
    capi.call("rmpcd_colloid_md_pos", d.struct(), ptr(mb), 0.01)
    capi.call("rmpcd_rattle_dimer_pos", d.struct(), ptr(rat), m, 2, 4.5, 0.01, capi.d3(L))
    capi.call("rmpcd_colloid_md_vel", d.struct(), ptr(mb), 0.01)
    capi.call("rmpcd_rattle_dimer_vel", d.struct(), m, 2, 4.5, 0.01, capi.d3(L))
    torch.cuda.synchronize()
    assert same_bits(d.host("pos"), want.pos)
    assert same_bits(d.host("vel"), want.vel)
    r = want.pos[0] - want.pos[1]
    assert abs(np.sqrt(r @ r) - 4.5) < 1e-10


# ---- neighbour list and forces --------------------------------------------------------
def dimer_setup(L, rho, seed, sigma=(2.0, 2.0), frac2=0.2):
    c, srt, cnt, st = sorted_state(L, rho, seed, frac2=frac2)
    cp = colloid_pair(L, seed + 1)
    eps = np.array([[1.0, 1.0], [1.0, 0.5]])
    sig = np.array([[sigma[0], sigma[1]], [sigma[0], sigma[1]]])
    cut = sig * 2 ** (1 / 6)
    return c, srt, cnt, st, cp, eps, sig, cut


@pytest.mark.parametrize("L,rho,with_lj", [((16, 16, 16), 10, True), ((24, 20, 28), 10, False),
                                           ((32, 32, 32), 10, True)])
def test_update_list(gpu, L, rho, with_lj):
    c, srt, cnt, st, cp, eps, sig, cut = dimer_setup(L, rho, 27)
    radius = cut.max() + 2.0
    stencil = O.make_stencil(1.0, radius)
    lj_o = O.lj_params(eps, sig, cut) if with_lj else None
    nmax = 4000
    lst, nl, is_md, is_reac = O.update_list(c, cnt, st, cp.pos, cp.species, srt.pos, srt.species, radius,
                                             stencil, lj_o, nmax)
    d = to_device(srt)
    cd = to_device(cp)
    cs = cells_to_device(c, cnt, st)
    nlst = NeighborList(2, nmax)
    st_h = nlst.make_stencil(1.0, radius)
    assert same_bits(st_h, stencil)
    lj = lj_params(eps, sig, cut) if with_lj else None
    capi.call("rmpcd_update_list", cs.boxp(), cs.cellsp(), cd.struct(), d.struct(), radius,
              ptr(nlst.stencil), nlst.n_stencil, C.byref(lj) if lj else None, nmax, ptr(nlst.list),
              ptr(nlst.n))
    torch.cuda.synchronize()
    assert same_bits(nlst.n.cpu().numpy(), nl)
    got = nlst.list.cpu().numpy()
    for i in range(2):
        assert same_bits(got[i, :nl[i]], lst[i, :nl[i]])
    assert same_bits(cs.is_md.cpu().numpy(), is_md)
    assert same_bits(cs.is_reac.cpu().numpy(), is_reac)


def test_update_list_overflow(gpu):
    c, srt, cnt, st, cp, eps, sig, cut = dimer_setup((16, 16, 16), 10, 28)
    d = to_device(srt)
    cd = to_device(cp)
    cs = cells_to_device(c, cnt, st)
    nlst = NeighborList(2, 50)
    nlst.make_stencil(1.0, 4.0)
    with pytest.raises(capi.RmpcdError, match="maximum of neighbor list reached"):
        capi.call("rmpcd_update_list", cs.boxp(), cs.cellsp(), cd.struct(), d.struct(), 4.0,
                  ptr(nlst.stencil), nlst.n_stencil, None, 50, ptr(nlst.list), ptr(nlst.n))
    assert nlst.n.max().item() <= 50


def test_compute_force_and_helpers(gpu):
    L = (20, 20, 20)
    c, srt, cnt, st, cp, eps, sig, cut = dimer_setup(L, 10, 29)
    # pull a few solvent particles inside the colloid cutoffs
    rng = np.random.default_rng(30)
    for k in range(40):
        u = rng.normal(size=3)
        srt.pos[k * 50] = cp.pos[k % 2] + u / np.linalg.norm(u) * rng.uniform(1.8, 2.3)
    srt.pos[:] %= np.asarray(L, float)
    srt.force[:] = 0
    srt.flags[:] &= ~((1 << O.REAC_BIT))
    O.apply_pbc(srt.pos, srt.image, L)
    srt2, i2i, cnt, st = O.sort(c, srt)
    srt = srt2
    radius = cut.max() + 2.0
    stencil = O.make_stencil(1.0, radius)
    lj_o = O.lj_params(eps, sig, cut)
    nmax = 4000
    lst, nl, is_md, is_reac = O.update_list(c, cnt, st, cp.pos, cp.species, srt.pos, srt.species, radius,
                                             stencil, lj_o, nmax)
    d = to_device(srt)
    cd = to_device(cp)
    want_s, want_c = srt.copy(), cp.copy()
    e_o = O.compute_force(want_c.pos, want_c.species, want_c.force, want_s.pos, want_s.species,
                          want_s.force, want_s.flags, lst, nl, L, lj_o)
    assert e_o > 0
    lj = lj_params(eps, sig, cut)
    ld = torch.as_tensor(lst).cuda()
    nd = torch.as_tensor(nl).cuda()
    e = C.c_double()
    capi.call("rmpcd_compute_force", cd.struct(), d.struct(), ptr(ld), ptr(nd), nmax, capi.d3(L),
              C.byref(lj), C.byref(e))
    assert abs(e.value - e_o) <= RTOL * abs(e_o)
    assert same_bits(d.host("flags"), want_s.flags)
    # the solvent forces in the reference's colloid order (one launch per colloid), also for
    # the particles in both lists
    in_both = np.intersect1d(lst[0, :nl[0]], lst[1, :nl[1]])
    assert in_both.size > 0
    assert same_bits(d.host("force"), want_s.force)
    scale = np.abs(want_c.force - cp.force).sum() + 1.0
    assert close(cd.host("force"), want_c.force, scale=scale)
    # colloid-colloid (n2) with overlapping pair
    cc = colloid_pair(L, 31)
    cc.pos[1] = cc.pos[0] + np.array([2.1, 0.3, -0.2])
    ceps = np.array([[1.0, 1.0], [1.0, 1.0]])
    csig = np.array([[4.0, 4.0], [4.0, 4.0]])
    cl_o = O.lj_params(ceps, csig, csig * 2 ** (1 / 6))
    want = cc.copy()
    e2_o = O.compute_force_n2(want.pos, want.species, want.force, L, cl_o)
    ccd = to_device(cc)
    cl = lj_params(ceps, csig, csig * 2 ** (1 / 6))
    e2 = C.c_double()
    capi.call("rmpcd_compute_force_n2", ccd.struct(), capi.d3(L), C.byref(cl), C.byref(e2))
    assert e2.value == e2_o
    assert same_bits(ccd.host("force"), want.force)
    # catalytic helpers on the force-updated state
    O.flag_particles_nl(lst[0], nl[0], cp.pos[0], want_s.pos, want_s.species, want_s.flags, L,
                        lj_o.cut_sq[0])
    capi.call("rmpcd_flag_particles_nl", d.struct(), cd.struct(), ptr(ld), ptr(nd), nmax, capi.d3(L),
              lj_o.cut_sq[0])
    torch.cuda.synchronize()
    assert same_bits(d.host("flags"), want_s.flags)
    nflag = (want_s.flags & (1 << O.REAC_BIT)).astype(bool).sum()
    assert nflag > 0
    # move the colloids away so flagged particles fall outside both cutoffs
    moved = cp.pos + 3.0
    cd.pos.copy_(torch.as_tensor(np.ascontiguousarray(moved.T)))
    draws = np.random.default_rng(32).random(srt.n)
    O.change_species(moved, want_s.pos, want_s.species, want_s.flags, L, lj_o.cut_sq[0],
                     lj_o.cut_sq[2], 0.5, draws)
    dr = torch.as_tensor(draws).cuda()
    capi.call("rmpcd_change_species", d.struct(), cd.struct(), capi.d3(L), lj_o.cut_sq[0],
              lj_o.cut_sq[2], 0.5, ptr(dr), 0, 0)
    torch.cuda.synchronize()
    assert same_bits(d.host("species"), want_s.species)
    assert same_bits(d.host("flags"), want_s.flags)
    O.refuel(moved, want_s.pos, want_s.species, L, 20)
    capi.call("rmpcd_refuel", d.struct(), cd.struct(), capi.d3(L), 20)
    torch.cuda.synchronize()
    assert same_bits(d.host("species"), want_s.species)


def test_threefry_device_matches_host(gpu):
    out = torch.zeros(1000, dtype=torch.float64, device="cuda")
    capi.call("rmpcd_threefry_uniform", 99, 4, 17, 1000, ptr(out))
    torch.cuda.synchronize()
    o = (C.c_uint64 * 2)()
    for i in (0, 1, 500, 999):
        capi.lib().rmpcd_threefry2x64_20((C.c_uint64 * 2)(17, i), (C.c_uint64 * 2)(99, 4), o)
        assert out[i].item() == (o[0] >> 11) * 2.0 ** -53
    v = out.cpu().numpy()
    assert abs(v.mean() - 0.5) < 0.05 and v.min() >= 0 and v.max() < 1
