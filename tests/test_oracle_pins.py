"""Known answers for the oracle routines that the golden vectors do not cover: periodic
wrapping, the is_md selection of the cell-wise Verlet halves and of the cell-wise
displacement, and the species-index order, sign and energy of the LJ force routines.
Each case is a closed form worked out by hand (or a brute force / finite difference
written from the mathematics), chosen so that a flipped sign, a swapped index or an
inverted selection in the oracle fails it.  CPU only."""
import numpy as np

from oracle import oracle as O


def _cells(L, rho, seed):
    rng = np.random.default_rng(seed)
    c = O.cells_init(L)
    O.set_origin(c, rng.random(3) - 1)
    n = int(rho * np.prod(L))
    ps = O.PState(n)
    ps.pos[:] = rng.random((n, 3)) * L
    ps.vel[:] = rng.normal(size=(n, 3))
    ps.force[:] = rng.normal(size=(n, 3))
    ps.force_old[:] = rng.normal(size=(n, 3))
    srt, _, cnt, st = O.sort(c, ps)
    return c, srt, cnt, st, rng


def _cell_of(cnt, st):
    """cell index of every sorted slot"""
    return np.repeat(np.arange(cnt.size), cnt)


# ---- apply_pbc (md.f90:202-217) --------------------------------------------------------
def test_apply_pbc_known_images():
    # x = x0 + k*edge with x0 in [0, edge) on a 1/64 grid: every value is exact in FP64,
    # so the wrapped position is x0 bit for bit and the image count grows by exactly k
    rng = np.random.default_rng(1)
    edges = np.array([20.0, 16.0, 12.5])
    n = 500
    k = rng.integers(-4, 6, size=(n, 3))
    x0 = np.floor(rng.random((n, 3)) * edges * 64) / 64
    pos = x0 + k * edges
    im0 = rng.integers(-3, 4, size=(n, 3)).astype(np.int32)
    image = im0.copy()
    O.apply_pbc(pos, image, edges)
    assert (pos == x0).all()
    assert (image == im0 + k).all()
    # the wrap points themselves: x = edge goes to 0 with one image up, x = 0 stays
    pos = np.array([[20.0, 0.0, -12.5]])
    image = np.zeros((1, 3), np.int32)
    O.apply_pbc(pos, image, edges)
    assert pos.tolist() == [[0.0, 0.0, 0.0]] and image.tolist() == [[1, 0, -1]]
    # a second call changes nothing (the wrapped state is a fixed point)
    p1, i1 = pos.copy(), image.copy()
    O.apply_pbc(pos, image, edges)
    assert (pos == p1).all() and (image == i1).all()


# ---- cell_md_pos / cell_md_vel (md.f90:79-102, 239-263) -----------------------------------
def test_cell_md_pos_selects_cells_by_md_flag():
    c, srt, cnt, st, rng = _cells((6, 5, 7), 6, 2)
    is_md = (rng.random(c.N) < 0.5).astype(np.uint8)
    cell = _cell_of(cnt, st)
    dt = 0.125
    for md_flag in (True, False):
        pos = srt.pos.copy()
        O.cell_md_pos(c, cnt, st, is_md, pos, srt.vel, srt.force, dt, md_flag)
        sel = (is_md[cell] != 0) == md_flag
        assert 0 < sel.sum() < srt.n
        assert (pos[~sel] == srt.pos[~sel]).all()
        want = srt.pos + dt * srt.vel + dt * dt / 2 * srt.force
        assert np.allclose(pos[sel], want[sel], rtol=0, atol=1e-13)
        assert (np.abs(pos[sel] - srt.pos[sel]).sum(axis=1) > 0).all()


def test_cell_md_vel_selects_cells_by_md_flag():
    c, srt, cnt, st, rng = _cells((6, 5, 7), 6, 3)
    is_md = (rng.random(c.N) < 0.5).astype(np.uint8)
    cell = _cell_of(cnt, st)
    wall = 1 << O.WALL_BIT
    flags0 = np.where(rng.random(srt.n) < 0.2, wall | 1, 1).astype(np.int32)
    dt = 0.25
    for md_flag in (True, False):
        vel, flags = srt.vel.copy(), flags0.copy()
        O.cell_md_vel(c, cnt, st, is_md, vel, srt.force, srt.force_old, flags, dt, md_flag)
        sel = (is_md[cell] != 0) == md_flag
        walled = (flags0 & wall) != 0
        # unselected cells: velocity and flags untouched
        assert (vel[~sel] == srt.vel[~sel]).all() and (flags[~sel] == flags0[~sel]).all()
        # selected and walled: the velocity is kept and only the wall bit is cleared
        m = sel & walled
        assert m.any() and (vel[m] == srt.vel[m]).all() and (flags[m] == 1).all()
        # selected, not walled: v + dt (f + f_old) / 2
        m = sel & ~walled
        want = srt.vel + dt * (srt.force + srt.force_old) / 2
        assert m.any() and np.allclose(vel[m], want[m], rtol=0, atol=1e-13)
        assert (flags[m] == 1).all()


# ---- cell_maximum_displacement (particle_system.f90:380-410) ------------------------------
def test_cell_maximum_displacement_filter_and_floor():
    c, srt, cnt, st, rng = _cells((5, 5, 5), 4, 4)
    cell = _cell_of(cnt, st)
    is_md = np.zeros(c.N, np.uint8)
    occupied = np.nonzero(cnt >= 2)[0]
    a, b = occupied[3], occupied[10]
    is_md[a] = 1
    pos_old = srt.pos.copy()
    pos = srt.pos.copy()
    sp = srt.species.copy()
    ja, jb = st[a], st[b]
    pos[ja] += [0.3, 0.0, 0.4]            # md cell: displacement 0.5
    pos[jb] += [5.0, 0.0, 0.0]            # not an md cell: ignored
    pos[ja + 1] += [0.0, 1.2, 0.0]        # md cell but species < 0: ignored
    sp[ja + 1] = -1
    assert cell[ja] == a and cell[ja + 1] == a and cell[jb] == b
    c.max_v = 0.0
    assert O.cell_maximum_displacement(c, cnt, st, is_md, pos, pos_old, sp, 0.1) == 0.5
    # the floor max_v * delta_t wins when it is larger
    c.max_v = 8.0
    assert O.cell_maximum_displacement(c, cnt, st, is_md, pos, pos_old, sp, 0.125) == 1.0
    # species 0 counts (the reference tests species >= 0)
    sp[ja + 1] = 0
    c.max_v = 0.0
    assert abs(O.cell_maximum_displacement(c, cnt, st, is_md, pos, pos_old, sp, 0.1) - 1.2) < 1e-15
    # with the second cell marked md the large displacement counts
    is_md[b] = 1
    assert O.cell_maximum_displacement(c, cnt, st, is_md, pos, pos_old, sp, 0.1) == 5.0


# ---- LJ force routines -------------------------------------------------------------------
def _lj_closed(r, eps, sig, cut):
    """(|F|, E) of a pair at distance r from the textbook LJ with the energy shifted to
    zero at the cutoff; F > 0 means repulsive"""
    s6 = (sig / r) ** 6
    s6c = (sig / cut) ** 6
    return 24 * eps * s6 * (2 * s6 - 1) / r, 4 * eps * (s6 * s6 - s6) - 4 * eps * (s6c * s6c - s6c)


def test_compute_force_index_order_sign_and_flags():
    # solvent species s2 and colloid species s1 use the (s2, s1) entry of the matrices
    # (neighbor_list.f90:204-208); the pair sits at r = 1.5 = sigma(2,1), inside
    # cut(2,1) = 2.0 and outside cut(1,2) = 1.2, with eps(2,1) = 2 and eps(1,2) = 0.5
    eps = np.array([[1.0, 0.5], [2.0, 4.0]])
    sig = np.array([[1.0, 1.1], [1.5, 1.3]])
    cut = np.array([[1.3, 1.2], [2.0, 1.6]])
    lj = O.lj_params(eps, sig, cut)
    L = (10.0, 10.0, 10.0)
    col = np.array([[5.0, 5.0, 5.0], [1.0, 1.0, 1.0]])
    csp = np.array([1, 2], np.int32)
    sol = np.array([[5.0, 6.5, 5.0],      # species 2 near colloid 0 (species 1)
                    [1.0, 1.0, 2.15],     # species 1 near colloid 1 (species 2): r = 1.15 < cut(1,2) = 1.2
                    [9.9, 1.0, 1.0]])     # species 2 near colloid 1 through the x boundary: r = 1.1
    ssp = np.array([2, 1, 2], np.int32)
    lst = np.array([[0, 2], [1, 2]], np.int32)
    nl = np.array([1, 2], np.int32)
    f1 = np.zeros((2, 3))
    f2 = np.zeros((3, 3))
    fl = np.array([1 << O.MD_BIT, 0, 0], np.int32)
    e = O.compute_force(col, csp, f1, sol, ssp, f2, fl, lst, nl, L, lj)
    F0, E0 = _lj_closed(1.5, 2.0, 1.5, 2.0)       # (2,1)
    F1, E1 = _lj_closed(1.15, 0.5, 1.1, 1.2)      # (1,2)
    F2, E2 = _lj_closed(1.1, 4.0, 1.3, 1.6)       # (2,2)
    assert F0 > 0 and F1 > 0 and F2 > 0
    # the solvent is pushed away from the colloid (+y, +z, -x through the boundary)
    assert np.allclose(f2[0], [0, F0, 0], rtol=1e-13)
    assert np.allclose(f2[1], [0, 0, F1], rtol=1e-13)
    assert np.allclose(f2[2], [-F2, 0, 0], rtol=1e-13)
    assert np.allclose(f1[0], -f2[0], rtol=1e-13) and np.allclose(f1[1], -(f2[1] + f2[2]), rtol=1e-13)
    assert abs(e - (E0 + E1 + E2)) < 1e-12 * (abs(E0) + abs(E1) + abs(E2))
    # MD set on every interacting particle, CATALYZED only next to the species-1 colloid,
    # and the old MD bit rolled into PAST_MD first
    MD, PAST, CAT = 1 << O.MD_BIT, 1 << O.PAST_MD_BIT, 1 << O.CATALYZED_BIT
    assert fl.tolist() == [MD | PAST | CAT, MD, MD]


def test_compute_force_outside_cutoff_and_flag_roll():
    eps, sig = np.array([[1.0]]), np.array([[1.0]])
    lj = O.lj_params(eps, sig, sig * 2 ** (1 / 6))
    L = (8.0, 8.0, 8.0)
    col = np.array([[4.0, 4.0, 4.0]])
    sol = np.array([[4.0, 4.0, 4.0 + 2 ** (1 / 6) + 1e-9], [4.0, 4.0, 3.0]])
    fl = np.array([1 << O.MD_BIT, 1 << O.PAST_MD_BIT], np.int32)
    f2 = np.zeros((2, 3))
    f1 = np.zeros((1, 3))
    e = O.compute_force(col, np.array([1], np.int32), f1, sol, np.array([1, 1], np.int32), f2, fl,
                        np.array([[0, 1]], np.int32), np.array([2], np.int32), L, lj)
    F, E = _lj_closed(1.0, 1.0, 1.0, 2 ** (1 / 6))
    assert (f2[0] == 0).all() and np.allclose(f2[1], [0, 0, -F], rtol=1e-13)
    assert abs(e - E) < 1e-13 and E > 0
    MD, PAST, CAT = 1 << O.MD_BIT, 1 << O.PAST_MD_BIT, 1 << O.CATALYZED_BIT
    assert fl.tolist() == [PAST, MD | CAT]


def test_compute_force_n2_index_order():
    # colloid pairs i < j: the cut-off is cut(s_i, s_j), force and energy use (s_j, s_i)
    # (neighbor_list.f90:243-258)
    eps = np.array([[1.0, 0.25], [3.0, 2.0]])
    sig = np.array([[1.0, 1.5], [2.0, 1.2]])
    cut = np.array([[1.5, 3.0], [1.0, 1.4]])
    lj = O.lj_params(eps, sig, cut)
    L = (20.0, 20.0, 20.0)
    pos = np.array([[10.0, 10.0, 10.0], [12.0, 10.0, 10.0]])
    # (s_i, s_j) = (1, 2): cut(1,2) = 3 admits r = 2; force from eps(2,1) = 3, sigma(2,1) = 2
    f = np.zeros((2, 3))
    e = O.compute_force_n2(pos, np.array([1, 2], np.int32), f, L, lj)
    F, _ = _lj_closed(2.0, 3.0, 2.0, 1.0)
    assert F == 36.0
    assert np.allclose(f, [[-F, 0, 0], [F, 0, 0]], rtol=1e-14)
    s6c = (2.0 / 1.0) ** 6  # cut_energy(2,1) with cut(2,1) = 1
    want = 0.0 - 4 * 3.0 * (s6c * s6c - s6c)
    assert abs(e - want) < 1e-13 * abs(want)
    # (s_i, s_j) = (2, 1): cut(2,1) = 1 excludes the pair
    f = np.zeros((2, 3))
    e = O.compute_force_n2(pos, np.array([2, 1], np.int32), f, L, lj)
    assert (f == 0).all() and e == 0


def test_compute_force_n2_energy_gradient():
    # two species with a symmetric cut-off matrix: the energy is a brute-force sum over
    # pairs, and the force on every colloid is minus its central-difference gradient
    rng = np.random.default_rng(5)
    eps = np.array([[1.0, 0.7], [0.7, 1.6]])
    sig = np.array([[2.0, 2.4], [2.4, 2.8]])
    cut = 2.5 * sig
    lj = O.lj_params(eps, sig, cut)
    L = np.array([16.0, 14.0, 18.0])
    n = 12
    pos = np.empty((n, 3))
    k = 0
    while k < n:  # no pair closer than 1.9 (keeps the gradient well conditioned)
        x = rng.random(3) * L
        d = pos[:k] - x
        d -= L * np.round(d / L)
        if k == 0 or (d * d).sum(axis=1).min() > 1.9 ** 2:
            pos[k] = x
            k += 1
    sp = rng.integers(1, 3, size=n).astype(np.int32)

    def energy(p):
        e = 0.0
        for i in range(n):
            for j in range(i + 1, n):
                d = p[i] - p[j]
                d -= L * np.round(d / L)
                r = np.sqrt(d @ d)
                a, b = sp[i] - 1, sp[j] - 1
                if r <= cut[a, b]:
                    e += _lj_closed(r, eps[b, a], sig[b, a], cut[b, a])[1]
        return e

    f = np.zeros((n, 3))
    e = O.compute_force_n2(pos, sp, f, L, lj)
    assert abs(e - energy(pos)) < 1e-10 * max(1.0, abs(e))
    h = 1e-6
    for i in range(n):
        for k in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += h
            p2[i, k] -= h
            g = (energy(p1) - energy(p2)) / (2 * h)
            assert abs(f[i, k] + g) < 1e-5 * max(1.0, abs(g)), (i, k, f[i, k], g)
