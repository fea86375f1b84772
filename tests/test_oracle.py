"""The oracle pinned: against the reference's own modules compiled from source (golden
vectors in tests/golden/, and oracle/_ref where it is built), and against what the
mathematics fixes (closed forms, invariants, brute force).  CPU only."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref")


# ---- hilbert: golden vectors from the reference's hilbert.f90 ------------------------
def test_hilbert_golden():
    z = np.load(os.path.join(GOLD, "hilbert_ref.npz"))
    for name in z.files:
        if name.startswith("keys_"):
            M = tuple(int(v) for v in name[5:].split("_"))
            keys, inv = z[name], z["inv_" + name[5:]]
            idx = 0
            for k in range(2 ** M[2]):
                for j in range(2 ** M[1]):
                    for i in range(2 ** M[0]):
                        assert O.compact_p_to_h((i, j, k), M) == keys[idx]
                        idx += 1
            for h in range(keys.size):
                assert O.compact_h_to_p(h, M) == tuple(inv[h])
        elif name.startswith("skeys_"):
            M = tuple(int(v) for v in name[6:].split("_"))
            nx, ny = 2 ** M[0], 2 ** M[1]
            for t, h in z[name]:
                p = (t % nx, (t // nx) % ny, t // (nx * ny))
                assert O.compact_p_to_h(p, M) == h
            for row in z["sinv_" + name[6:]]:
                assert O.compact_h_to_p(row[0], M) == tuple(row[1:])
        elif name.startswith("ukeys_"):
            m = int(name[6:])
            keys, inv = z[name], z["uinv_%d" % m]
            n = 2 ** m
            for t in range(keys.size):
                assert O.p_to_h((t % n, (t // n) % n, t // (n * n)), m) == keys[t]
                assert O.h_to_p(t, m) == tuple(inv[t])


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "hilbert_drv")),
                    reason="oracle/_ref not built (needs the reference sources)")
def test_hilbert_against_ref_binary():
    for M in [(2, 2, 4), (6, 3, 2), (4, 4, 4)]:
        out = subprocess.run([os.path.join(REF, "hilbert_drv"), *map(str, M), "0"],
                             capture_output=True, text=True, check=True).stdout.split("\n")
        n = 2 ** sum(M)
        keys = [int(v) for v in out[:n]]
        idx = 0
        for k in range(2 ** M[2]):
            for j in range(2 ** M[1]):
                for i in range(2 ** M[0]):
                    assert O.compact_p_to_h((i, j, k), M) == keys[idx]
                    idx += 1


def test_hilbert_math():
    # bijection onto [0, 2^sum M); consecutive keys are face neighbours on cubes
    for M in [(3, 3, 3), (2, 4, 3)]:
        seen = {}
        for k in range(2 ** M[2]):
            for j in range(2 ** M[1]):
                for i in range(2 ** M[0]):
                    h = O.compact_p_to_h((i, j, k), M)
                    assert 0 <= h < 2 ** sum(M) and h not in seen
                    seen[h] = (i, j, k)
                    assert O.compact_h_to_p(h, M) == (i, j, k)
        if M[0] == M[1] == M[2]:
            for h in range(1, 2 ** sum(M)):
                a, b = np.array(seen[h - 1]), np.array(seen[h])
                assert np.abs(a - b).sum() == 1


# ---- LJ: golden vectors from the reference's interaction.f90 -------------------------
def test_lj_golden_bit_exact():
    z = np.load(os.path.join(GOLD, "interaction_ref.npz"))
    for row, o in zip(z["inputs"], z["outputs"]):
        f = O.lj_force(row[:3], row[3], row[4], row[5])
        e = O.lj_energy(row[3], row[4], row[5], 0.0)
        lj = O.lj_params([[row[4]]], [[row[5]]], [[np.sqrt(row[3])]])
        got = np.array([*f, e, lj.cut_energy[0]])
        assert got.view(np.uint64).tolist() == o.view(np.uint64).tolist()


def test_lj_math():
    # F = -dE/dr (central difference), WCA energy zero at the cutoff, sizes checked
    eps, sig = 1.3, 2.0
    for r in (1.9, 2.05, 2.2):
        h = 1e-6
        dE = (O.lj_energy((r + h) ** 2, eps, sig) - O.lj_energy((r - h) ** 2, eps, sig)) / (2 * h)
        f = O.lj_force([r, 0, 0], r * r, eps, sig)
        assert abs(f[0] + dE) < 1e-5 * abs(dE)
    cut = sig * 2 ** (1 / 6)
    lj = O.lj_params([[eps]], [[sig]], [[cut]])
    assert abs(O.lj_energy(cut * cut, eps, sig, lj.cut_energy[0])) < 1e-12
    with pytest.raises(ValueError, match="unequal sizes"):
        O.lj_params(np.ones((5, 4)), np.ones((5, 4)), np.ones((5, 4)))


# ---- cells and sort ----------------------------------------------------------------
def test_cells_init_padding():
    c = O.cells_init((120, 30, 30))
    assert tuple(c.M) == (7, 5, 5) and c.N == 2 ** 17
    c = O.cells_init((8, 5, 5), has_walls=True)
    assert tuple(c.L) == (8, 5, 6) and tuple(c.M) == (3, 3, 3)
    c = O.cells_init((1, 1, 1))
    assert tuple(c.M) == (1, 1, 1) and c.N == 8


def test_cell_system_invariants():
    # test_cell_system_0.f90:37-69 at a smaller N
    rng = np.random.default_rng(0)
    L = (120, 30, 30)
    n = 20000
    c = O.cells_init(L)
    ps = O.PState(n)
    ps.pos[:] = rng.random((n, 3)) * L
    srt, i2i, cnt, st = O.sort(c, ps)
    assert st[0] == 0 and cnt.sum() == n and st[-1] + cnt[-1] == n
    assert (st[1:] == np.cumsum(cnt)[:-1]).all()
    keys = np.array([O.compact_p_to_h(O.cartesian_indices(c, x), c.M) for x in srt.pos])
    assert (np.diff(keys) >= 0).all()
    for j in rng.choice(n, 300, replace=False):
        assert O.compact_h_to_p(keys[j], c.M) == O.cartesian_indices(c, srt.pos[j])
    # stable: within a cell the original order is kept; id_to_idx inverts the permutation
    for h in np.unique(keys)[:200]:
        ids = srt.id[keys == h]
        assert (np.diff(ids) > 0).all()
    assert (i2i[srt.id] == np.arange(n)).all()
    # a multi-threaded count gives the same counts
    cnt4, st4 = O.count_particles(c, ps.pos, nthreads=4)
    assert (cnt4 == cnt).all() and (st4 == st).all()


def test_cartesian_wrap_and_shift():
    c = O.cells_init((8, 5, 5))
    O.set_origin(c, (-0.25, -1.0, -0.5))
    assert O.cartesian_indices(c, (7.75, 4.0, 4.5)) == (0, 0, 0)
    assert O.cartesian_indices(c, (0.0, 0.0, 0.0)) == (0, 1, 0)
    c.has_walls = 1
    assert O.cartesian_indices(c, (0.2, 0.2, 4.6))[2] == 5


def test_maximum_displacement_closed_form():
    # test_particle_system_0.f90:29-47
    rng = np.random.default_rng(1)
    n = 3000
    pos = rng.random((n, 3)) * (1, 20, 5)
    old = pos.copy()
    sp = np.ones(n, np.int32)
    i = 1234
    x = rng.random(3)
    pos[i] += x
    assert abs(O.maximum_displacement(pos, old, sp) - np.sqrt(x @ x)) < 1e-14
    assert O.maximum_displacement(pos, old, sp) <= np.sqrt(3)
    pos[i] -= x
    assert O.maximum_displacement(pos, old, sp) < 1e-15
    sp[i] = -1
    pos[i] += 5
    assert O.maximum_displacement(pos, old, sp) < 1e-15


# ---- mpcd -----------------------------------------------------------------------------
def test_rand_sphere_math():
    rng = np.random.default_rng(2)
    ns = []
    for _ in range(4000):
        u = rng.random(16)
        n, k = O.rand_sphere(u)
        assert k >= 2 and k % 2 == 0
        assert abs(n @ n - 1) < 1e-14
        a, b = 2 * u[k - 2] - 1, 2 * u[k - 1] - 1
        assert a * a + b * b < 1
        for j in range(0, k - 2, 2):  # rejected pairs lie outside the unit disc
            assert (2 * u[j] - 1) ** 2 + (2 * u[j + 1] - 1) ** 2 >= 1
        ns.append(n)
    ns = np.array(ns)
    # uniform on the sphere: zero mean, <n_i n_j> = delta_ij / 3
    assert np.abs(ns.mean(axis=0)).max() < 0.05
    assert np.abs(ns.T @ ns / len(ns) - np.eye(3) / 3).max() < 0.03


def _cells_state(L, rho, seed, frac2=0.0):
    rng = np.random.default_rng(seed)
    c = O.cells_init(L)
    O.set_origin(c, rng.random(3) - 1)
    n = int(rho * np.prod(L))
    ps = O.PState(n)
    ps.pos[:] = rng.random((n, 3)) * L
    ps.vel[:] = rng.normal(size=(n, 3))
    if frac2:
        ps.species[rng.random(n) < frac2] = 2
    srt, i2i, cnt, st = O.sort(c, ps)
    return c, srt, cnt, st


def test_srd_rotation_math():
    # per-cell momentum and energy are conserved; the applied matrix is the rotation by
    # -alpha about the axis (mpcd.f90:163-171 builds R^T): w perpendicular to n goes to
    # -n x w for the default alpha = 90 degrees
    c, srt, cnt, st = _cells_state((6, 6, 6), 10, 3)
    rng = np.random.default_rng(4)
    axes = rng.normal(size=(c.N, 3))
    axes /= np.linalg.norm(axes, axis=1)[:, None]
    v = srt.vel.copy()
    O.simple_mpcd_step(c, cnt, st, v, srt.species, axes)
    for ci in np.nonzero(cnt > 1)[0]:
        s, n = st[ci], cnt[ci]
        u0 = srt.vel[s:s + n].mean(axis=0)
        assert np.allclose(v[s:s + n].sum(axis=0), srt.vel[s:s + n].sum(axis=0), atol=1e-12)
        w0 = srt.vel[s:s + n] - u0
        w1 = v[s:s + n] - u0
        ax = axes[ci]
        assert np.allclose(w1 @ ax, w0 @ ax, atol=1e-12)  # axis component kept
        wp = w0 - np.outer(w0 @ ax, ax)
        w1p = w1 - np.outer(w1 @ ax, ax)
        assert np.allclose(w1p, -np.cross(ax, wp), atol=1e-12)
    # a given alpha rotates by -alpha
    v2 = srt.vel.copy()
    O.simple_mpcd_step(c, cnt, st, v2, srt.species, axes, alpha=0.7)
    ci = int(np.nonzero(cnt > 2)[0][0])
    s, n = st[ci], cnt[ci]
    u0 = srt.vel[s:s + n].mean(axis=0)
    ax = axes[ci]
    w0 = srt.vel[s] - u0
    wp = w0 - (w0 @ ax) * ax
    want = (w0 @ ax) * ax + np.cos(0.7) * wp - np.sin(0.7) * np.cross(ax, wp)
    assert np.allclose(v2[s] - u0, want, atol=1e-12)


def test_srd_species_quirks():
    # mean over species > 0 only; every particle of the cell is rotated; cells with a
    # single counted particle draw nothing (mpcd.f90:135-141, 173-175)
    c, srt, cnt, st = _cells_state((4, 4, 4), 10, 5)
    ci = int(np.nonzero(cnt > 3)[0][0])
    s, n = st[ci], cnt[ci]
    srt.species[s] = 0
    u = np.random.default_rng(6).random(8 * c.N)
    axes, used = O.srd_axes_from_stream(c, cnt, st, srt.species, u)
    v = srt.vel.copy()
    O.simple_mpcd_step(c, cnt, st, v, srt.species, axes)
    u_eff = srt.vel[s + 1:s + n].mean(axis=0)
    assert np.allclose(v[s + 1:s + n].mean(axis=0), u_eff, atol=1e-12)
    assert not np.allclose(v[s], srt.vel[s])
    srt.species[s:s + n] = 0
    srt.species[s] = 1
    axes2, used2 = O.srd_axes_from_stream(c, cnt, st, srt.species, u)
    assert np.all(axes2[ci] == 0)


def test_temperature_closed_form():
    # equipartition: T -> kT for Maxwell-Boltzmann velocities; and the formula's
    # quirks (mean over n, count before the n_effective test)
    c, srt, cnt, st = _cells_state((12, 12, 12), 10, 7)
    T = O.compute_temperature(c, cnt, st, srt.vel, srt.species)
    assert abs(T - 1.0) < 0.02
    ci = int(np.nonzero(cnt > 3)[0][0])
    s, n = st[ci], cnt[ci]
    sp = srt.species.copy()
    sp[:] = 0
    sp[s:s + n] = 1
    sp[s] = 0
    vv = srt.vel[s:s + n]
    u = vv[1:].sum(axis=0) / n
    k = ((vv[1:] - u) ** 2).sum() / (3 * (n - 2))
    ncount = (cnt > 1).sum()
    assert abs(O.compute_temperature(c, cnt, st, srt.vel, sp) - k / ncount) < 1e-12


def test_temperature_profile_pins():
    """the tz profile: its bins hold, per z layer of cells, the cells with n_eff > 1 and
    the sum of their kinetic terms.  Pinned by a route that does not decode Hilbert keys:
    each particle's cell from floor(x - origin), cells grouped by their particles'
    coordinates; the total over bins covering the box equals te * count of
    compute_temperature; bins outside the range are dropped."""
    L = (6, 5, 8)
    c, srt, cnt, st = _cells_state(L, 4, 12)
    O.set_origin(c, (-0.3, -0.6, -0.45))
    srt2, i2i, cnt, st = O.sort(c, srt)
    sp = srt2.species.copy()
    sp[::7] = 0
    te = O.compute_temperature(c, cnt, st, srt2.vel, sp)
    # z bins of width 1 from the bottom of the shifted box, plus one bin beyond its top
    xmin, dx, nb = c.origin[2], 1.0, L[2] + 1
    data, count = O.temperature_profile(c, cnt, st, srt2.vel, sp, xmin, dx, nb)
    # brute force: the cell of each particle from its position, per-cell kinetic terms
    p = np.floor(srt2.pos - np.array(c.origin[:])).astype(int) % np.array(L)
    want_d, want_c = np.zeros(nb), np.zeros(nb, int)
    counted = 0
    cells = {}
    for i in range(srt2.n):
        cells.setdefault(tuple(p[i]), []).append(i)
    for (x, y, z), idx in cells.items():
        if len(idx) <= 1:
            continue
        counted += 1
        idx = np.array(idx)
        keep = idx[sp[idx] > 0]
        if keep.size <= 1:
            continue
        u = srt2.vel[keep].sum(axis=0) / len(idx)
        kin = ((srt2.vel[keep] - u) ** 2).sum() / (3 * (keep.size - 1))
        want_d[z] += kin
        want_c[z] += 1
    assert (count == want_c).all()
    assert np.allclose(data, want_d, rtol=1e-12, atol=1e-12)
    assert count[-1] == 0 and data[-1] == 0
    assert abs(data.sum() - te * counted) <= 1e-12 * abs(te * counted)
    # a range that starts at the middle of the box keeps the upper half only
    d2, c2 = O.temperature_profile(c, cnt, st, srt2.vel, sp, xmin + 4, dx, nb)
    assert (c2[:4] == want_c[4:8]).all() and c2[4:].sum() == 0


def test_stream_periodic_closed_form():
    c = O.cells_init((8, 5, 5))
    pos = np.array([[7.9, 4.9, 0.1], [0.1, 0.2, 0.3]])
    vel = np.array([[1.0, 1.0, -1.0], [-3.0, 10.0, 0.0]])
    im = np.zeros((2, 3), np.int32)
    O.mpcd_stream_periodic(c, pos, vel, im, 0.5)
    assert np.allclose(pos, [[0.4, 0.4, 4.6], [6.6, 0.2, 0.3]])
    assert (im == [[1, 1, -1], [-1, 1, 0]]).all()


def test_bulk_reaction_closed_form():
    # one draw per eligible cell (even with no 'from' particle), the last 'from'
    # particle converts when u < 1 - exp(-k n_from tau)
    c, srt, cnt, st = _cells_state((4, 4, 4), 10, 8, frac2=0.5)
    is_reac = np.ones(c.N, np.uint8)
    draws = np.full(c.N, 0.5)
    sp = srt.species.copy()
    O.bulk_reaction(c, cnt, st, is_reac, sp, 2, 1, 1.0, 1.0, draws)
    for ci in np.nonzero(cnt > 1)[0]:
        s, n = st[ci], cnt[ci]
        nf = (srt.species[s:s + n] == 2).sum()
        changed = np.nonzero(sp[s:s + n] != srt.species[s:s + n])[0]
        p = 1 - np.exp(-1.0 * nf)
        if 0.5 < p:
            assert list(changed) == [np.nonzero(srt.species[s:s + n] == 2)[0][-1]]
        else:
            assert len(changed) == 0
    sp2 = srt.species.copy()
    O.bulk_reaction(c, cnt, st, np.zeros(c.N, np.uint8), sp2, 2, 1, 1.0, 1.0, draws * 0)
    assert (sp2 == srt.species).all()


def test_mpcat_thermostat_invariants():
    # MPC-AT (mpcd.f90:144-160): in a drawing cell the species > 0 velocities become
    # sqrt(T) times the normals, shifted (hydro) so that the cell's species > 0 momentum is
    # unchanged; species <= 0 particles and cells with fewer than two counted particles
    # keep their velocities; without hydro the velocities are the scaled normals
    c, srt, cnt, st = _cells_state((5, 4, 6), 10, 21)
    ci = int(np.nonzero(cnt > 3)[0][0])
    s0 = st[ci]
    srt.species[s0] = 0
    cj = int(np.nonzero(cnt == 2)[0][0]) if (cnt == 2).any() else None
    if cj is not None:
        srt.species[st[cj]] = 0  # a cell with a single counted particle draws nothing
    rng = np.random.default_rng(22)
    g = rng.normal(size=srt.vel.shape)
    T = 1.7
    v = srt.vel.copy()
    O.mpcat_step(c, cnt, st, v, srt.species, g, T, hydro=True)
    for ci2 in np.nonzero(cnt > 1)[0]:
        s, n = st[ci2], cnt[ci2]
        on = srt.species[s:s + n] > 0
        if on.sum() <= 1:
            assert (v[s:s + n] == srt.vel[s:s + n]).all()
            continue
        assert np.allclose(v[s:s + n][on].sum(axis=0), srt.vel[s:s + n][on].sum(axis=0), atol=1e-11)
        d = v[s:s + n][on] - np.sqrt(T) * g[s:s + n][on]
        assert np.allclose(d, d[0], atol=1e-12)  # one shift per cell
        assert (v[s:s + n][~on] == srt.vel[s:s + n][~on]).all()
    v2 = srt.vel.copy()
    O.mpcat_step(c, cnt, st, v2, srt.species, g, T, hydro=False)
    s, n = st[ci], cnt[ci]
    on = srt.species[s:s + n] > 0
    assert (v2[s:s + n][on] == np.sqrt(T) * g[s:s + n][on]).all()
    # the bulk temperature follows T
    c3, srt3, cnt3, st3 = _cells_state((8, 8, 8), 10, 23)
    g3 = np.random.default_rng(24).normal(size=srt3.vel.shape)
    O.mpcat_step(c3, cnt3, st3, srt3.vel, srt3.species, g3, 0.5, hydro=True)
    assert abs(O.compute_temperature(c3, cnt3, st3, srt3.vel, srt3.species) / 0.5 - 1) < 0.03


def test_rescale_cells_invariants():
    # rescale_cells (mpcd.f90:748-801): per drawing cell the species > 0 momentum is kept,
    # the directions about the mean are kept and their kinetic energy becomes
    # 3 (n_eff - 1) T / 2; so compute_temperature returns T exactly for a cell set with
    # species > 0 only
    c, srt, cnt, st = _cells_state((5, 5, 4), 10, 25)
    T = 0.37
    v = srt.vel.copy()
    O.rescale_cells(c, cnt, st, v, srt.species, T)
    for ci in np.nonzero(cnt > 1)[0]:
        s, n = st[ci], cnt[ci]
        u0 = srt.vel[s:s + n].mean(axis=0)
        assert np.allclose(v[s:s + n].mean(axis=0), u0, atol=1e-12)
        w0, w1 = srt.vel[s:s + n] - u0, v[s:s + n] - u0
        assert np.isclose((w1 * w1).sum(), 3 * (n - 1) * T, rtol=1e-12)
        scale = np.sqrt((w1 * w1).sum() / (w0 * w0).sum())
        assert np.allclose(w1, scale * w0, atol=1e-12)
    assert np.isclose(O.compute_temperature(c, cnt, st, v, srt.species), T, rtol=1e-12)


def test_bulk_reaction_endothermic_invariants():
    # bulk_reaction_endothermic (mpcd.f90:646-700): a reacting cell converts its last
    # 'from' particle and loses exactly delta_u of kinetic energy about its centre of
    # mass, keeping its momentum; cells below delta_u or without 'from' do not react
    c, srt, cnt, st = _cells_state((4, 5, 4), 10, 26, frac2=0.5)
    is_reac = np.ones(c.N, np.uint8)
    is_reac[0] = 0
    du = 2.0
    for draw, expect in ((0.0, True), (1.0, False)):
        v, sp = srt.vel.copy(), srt.species.copy()
        O.bulk_reaction_endothermic(c, cnt, st, is_reac, v, sp, 2, 1, 0.3, 0.5, du, np.full(c.N, draw))
        reacted = 0
        for ci in np.nonzero(cnt > 1)[0]:
            s, n = st[ci], cnt[ci]
            com = srt.vel[s:s + n].mean(axis=0)
            k0 = 0.5 * ((srt.vel[s:s + n] - com) ** 2).sum()
            nf = (srt.species[s:s + n] == 2).sum()
            if expect and is_reac[ci] and nf > 0 and k0 > du:
                reacted += 1
                k1 = 0.5 * ((v[s:s + n] - com) ** 2).sum()
                assert np.isclose(k0 - k1, du, rtol=1e-10, atol=1e-10)
                assert np.allclose(v[s:s + n].sum(axis=0), srt.vel[s:s + n].sum(axis=0), atol=1e-11)
                ch = np.nonzero(sp[s:s + n] != srt.species[s:s + n])[0]
                assert list(ch) == [np.nonzero(srt.species[s:s + n] == 2)[0][-1]]
                assert sp[s + ch[0]] == 1
            else:
                assert (v[s:s + n] == srt.vel[s:s + n]).all() and (sp[s:s + n] == srt.species[s:s + n]).all()
        assert reacted > 0 or not expect


def _flight(x0, v0, g, dt, L, bc):
    """the forced flight between y/z walls restated from the physics (independent of the
    oracle's code): straight lines in y and z, uniform acceleration along x, the earliest
    wall crossing handled first; periodic axes are crossed freely (unwrapped)"""
    x, v, t = np.array(x0, float), np.array(v0, float), dt
    gv = np.array([g, 0.0, 0.0])
    for _ in range(64):
        best, dim = None, None
        for d in (1, 2):
            if bc[d] == O.PERIODIC_BC or v[d] == 0:
                continue
            tc = (L[d] - x[d]) / v[d] if v[d] > 0 else -x[d] / v[d]
            if 0 < tc <= t and (best is None or tc < best):
                best, dim = tc, d
        if best is None:
            break
        x = x + v * best + gv * best * best / 2
        v = v + gv * best
        t -= best
        if bc[dim] == O.BOUNCE_BACK_BC:
            v = -v
        else:
            v[dim] = -v[dim]
    return x + v * t + gv * t * t / 2, v + gv * t


@pytest.mark.parametrize("bc", [(1, 2, 2), (1, 3, 3), (1, 1, 2), (1, 3, 2)])
def test_yzwall_stream_against_the_physics(bc):
    # mpcd_stream_xforce_yzwall (mpcd.f90:459-511): particles that stay inside move by
    # v dt + g dt^2/2 and keep their velocity (bit for bit); the others follow the flight
    # with reflections (bounce-back reverses v, specular the normal component) and get the
    # WALL bit
    L = (6.0, 5.0, 4.0)
    c = O.cells_init((6, 5, 4), has_walls=True)
    for k in range(3):
        c.bc[k] = bc[k]
    rng = np.random.default_rng(40)
    n = 4000
    pos = rng.random((n, 3)) * L
    vel = rng.normal(size=(n, 3))
    flags = np.zeros(n, np.int32)
    dt, g = 0.3, 0.7
    p1, v1 = pos.copy(), vel.copy()
    O.mpcd_stream_xforce_yzwall(c, p1, v1, flags, dt, g)
    hit = (flags & (1 << 3)) != 0
    assert 50 < hit.sum() < n
    free = ~hit
    want = (pos[free] + vel[free] * dt) + np.array([g, 0, 0]) * (dt * dt) / 2
    assert (p1[free] == want).all() and (v1[free] == vel[free]).all()
    for i in np.nonzero(hit)[0]:
        x, v = _flight(pos[i], vel[i], g, dt, L, bc)
        assert np.allclose(p1[i], x, atol=1e-9) and np.allclose(v1[i], v, atol=1e-9), (i, p1[i], x, v1[i], v)


def test_wall_mpcd_step_invariants():
    # wall_mpcd_step (mpcd.f90:193-337): SRD about a cell mean that counts every particle,
    # plus, in the first and last z layer, wall_n - n virtual particles of total velocity
    # sqrt((wall_n - n) T_wall) g; the rotation keeps each velocity's distance from that
    # mean and its axis component.  Thermostat: sqrt(T) times the normals, shifted to the
    # mean (keep_cell_v) or not.
    Lc = (6, 5, 10)
    c = O.cells_init(Lc, has_walls=True)
    O.set_origin(c, (-0.3, -0.6, -0.5))  # half-filled wall layers
    rng = np.random.default_rng(41)
    n = 10 * 6 * 5 * 10
    ps = O.PState(n)
    ps.pos[:] = rng.random((n, 3)) * (6, 5, 10)
    ps.vel[:] = rng.normal(size=(n, 3))
    ps.id[:] = np.arange(n, dtype=np.int32)
    srt, _, cnt, st = O.sort(c, ps)
    axes = rng.normal(size=(c.N, 3))
    axes /= np.linalg.norm(axes, axis=1)[:, None]
    vnorm = rng.normal(size=(c.N, 3))
    pnorm = rng.normal(size=(n, 3))
    wt, wn = (0.8, 1.3), (14, 12)
    v = srt.vel.copy()
    O.wall_mpcd_step(c, cnt, st, v, axes=axes, vnorm=vnorm, wall_temperature=wt, wall_n=wn)
    layers = set()
    for ci in np.nonzero(cnt > 1)[0]:
        s, m = st[ci], cnt[ci]
        pz = O.compact_h_to_p(int(ci), list(c.M))[2]
        w = 0 if pz == 0 else 1 if pz == c.L[2] - 1 else -1
        u = srt.vel[s:s + m].sum(axis=0)
        if w >= 0 and m < wn[w]:
            layers.add(w)
            u = (u + vnorm[ci] * np.sqrt((wn[w] - m) * wt[w])) / wn[w]
        else:
            u = u / m
        d0, d1 = srt.vel[s:s + m] - u, v[s:s + m] - u
        assert np.allclose(np.linalg.norm(d1, axis=1), np.linalg.norm(d0, axis=1), atol=1e-12)
        assert np.allclose(d1 @ axes[ci], d0 @ axes[ci], atol=1e-12)
    assert layers == {0, 1}
    v2 = srt.vel.copy()
    O.wall_mpcd_step(c, cnt, st, v2, vnorm=vnorm, pnorm=pnorm, wall_temperature=wt, wall_n=wn, thermostat=True,
                     bulk_temperature=0.6)
    v3 = srt.vel.copy()
    O.wall_mpcd_step(c, cnt, st, v3, pnorm=pnorm, thermostat=True, bulk_temperature=0.6, keep_cell_v=False)
    for ci in np.nonzero(cnt > 1)[0][:300]:
        s, m = st[ci], cnt[ci]
        d = v2[s:s + m] - np.sqrt(0.6) * pnorm[s:s + m]
        assert np.allclose(d, d[0], atol=1e-12)
        assert (v3[s:s + m] == np.sqrt(0.6) * pnorm[s:s + m]).all()


def test_zwall_stream_reflection():
    c = O.cells_init((5, 5, 5), has_walls=True)
    pos = np.array([[1, 1, 0.1], [1, 1, 4.9], [1, 1, 2.0]])
    vel = np.array([[0.3, 0, -1.0], [0, 0.2, 1.0], [0, 0, 1.0]])
    f = np.zeros((3, 3))
    fl = np.zeros(3, np.int32)
    p0, v0 = pos.copy(), vel.copy()
    O.mpcd_stream_nogravity_zwall(c, pos, vel, f, fl, 0.5)
    assert np.allclose(vel[:2], -v0[:2]) and np.allclose(vel[2], v0[2])
    assert np.allclose(pos[0], [1 - 0.3 * 0.3, 1, 0.1 + 0.3]) and 0 < pos[1, 2] < 5
    assert list(fl) == [8, 8, 0]


# ---- md -------------------------------------------------------------------------------
def test_wall_md_invariants():
    # test_md_0.f90:31-103: |v| preserved, 0 < y < L; bounce-back v -> -v, specular flips v_y
    for bc in (O.BOUNCE_BACK_BC, O.SPECULAR_BC):
        c, srt, cnt, st = _cells_state((1, 20, 5), 20, 9)
        c.bc[1] = bc
        rng = np.random.default_rng(10)
        srt.vel[:] = (rng.random((srt.n, 3)) - 0.5) * 2 * np.array([1, 20, 5])
        x0, v0 = srt.pos.copy(), srt.vel.copy()
        will = (x0[:, 1] + v0[:, 1] < 0) | (x0[:, 1] + v0[:, 1] > 20)
        is_md = np.zeros(c.N, np.uint8)
        srt.force[:] = 0
        O.cell_md_pos_wall(c, cnt, st, is_md, srt.pos, srt.vel, srt.force, 1.0, False, 1)
        assert np.allclose(np.linalg.norm(srt.vel, axis=1), np.linalg.norm(v0, axis=1))
        assert (srt.pos[:, 1] > 0).all() and (srt.pos[:, 1] < 20).all()
        if bc == O.BOUNCE_BACK_BC:
            assert np.allclose(srt.vel[will], -v0[will])
        else:
            assert np.allclose(srt.vel[will][:, [0, 2]], v0[will][:, [0, 2]])
            assert np.allclose(srt.vel[will][:, 1], -v0[will][:, 1])


def test_verlet_closed_form():
    pos = np.array([[1.0, 2.0, 3.0]])
    vel = np.array([[0.5, -1.0, 2.0]])
    f = np.array([[2.0, 0.0, -4.0]])
    fo = np.array([[0.0, 2.0, 0.0]])
    fl = np.array([0], np.int32)
    O.md_pos(pos, vel, f, 0.1)
    assert np.allclose(pos, [[1.06, 1.9, 3.18]])
    O.md_vel(vel, f, fo, fl, 0.1)
    assert np.allclose(vel, [[0.6, -0.9, 1.8]])
    fl[0] = 1 << O.WALL_BIT
    O.md_vel(vel, f, fo, fl, 0.1)
    assert np.allclose(vel, [[0.6, -0.9, 1.8]]) and fl[0] == 0


def test_rattle_constraint():
    # after rattle_dimer_pos |r12| = d; after rattle_dimer_vel (v1 - v2).r12 = 0
    rng = np.random.default_rng(11)
    L = (32.0, 32.0, 32.0)
    for _ in range(20):
        pos = np.array([[16.0, 16, 16], [20.5, 16, 16]]) + rng.normal(size=(2, 3)) * 0.01
        old = pos.copy()
        pos += rng.normal(size=(2, 3)) * 0.02
        vel = rng.normal(size=(2, 3)) * 0.1
        sp = np.array([1, 2], np.int32)
        mass = np.array([300.0, 450.0])
        O.rattle_dimer_pos(pos, vel, old, sp, mass, 4.5, 0.01, L)
        r = pos[0] - pos[1]
        assert abs(np.sqrt(r @ r) - 4.5) < 1e-12
        O.rattle_dimer_vel(pos, vel, sp, mass, 4.5, 0.01, L)
        assert abs((vel[0] - vel[1]) @ r) < 1e-12


# ---- neighbour list and forces ---------------------------------------------------------
def test_stencil_closed_form():
    for cut in (1.5, 3.6, 4.245, 6.0):
        st = O.make_stencil(1.0, cut)
        m = int(np.floor(cut)) + 1
        want = []
        for i in range(-m, m + 1):
            for j in range(-m, m + 1):
                for k in range(-m, m + 1):
                    d = np.maximum(0, np.abs([i, j, k]) - 1)
                    if np.sqrt((d * d).sum()) < cut:
                        want.append((i, j, k))
        assert [tuple(r) for r in st] == want


def _nl_state(L, rho, seed, n_col):
    rng = np.random.default_rng(seed)
    c = O.cells_init(L)
    n = int(rho * np.prod(L))
    ps = O.PState(n)
    ps.pos[:] = rng.random((n, 3)) * L
    srt, i2i, cnt, st = O.sort(c, ps)
    col = O.PState(n_col)
    col.pos[:] = rng.random((n_col, 3)) * L
    return c, srt, cnt, st, col


def test_neighbor_list_brute_force():
    # test_neighbor_list_0.f90:47-81: every solvent particle within the cutoff is listed
    L = (20, 10, 35)
    c, srt, cnt, st, col = _nl_state(L, 5, 12, 6)
    cutoff = 3.6
    stl = O.make_stencil(1.0, cutoff)
    lst, nl, is_md, is_reac = O.update_list(c, cnt, st, col.pos, col.species, srt.pos, srt.species,
                                             cutoff, stl, None, 2000)
    for i in range(col.n):
        d = np.array([O.rel_pos(col.pos[i], y, L) for y in srt.pos])
        inside = set(np.nonzero((d * d).sum(axis=1) < cutoff ** 2)[0])
        assert set(lst[i, :nl[i]]) == inside
    # the reference's negative test (WILL_FAIL): a list built with 0.8*cutoff misses some
    lst2, nl2, _, _ = O.update_list(c, cnt, st, col.pos, col.species, srt.pos, srt.species,
                                     0.8 * cutoff, O.make_stencil(1.0, 0.8 * cutoff), None, 2000)
    d = np.array([O.rel_pos(col.pos[0], y, L) for y in srt.pos])
    inside = set(np.nonzero((d * d).sum(axis=1) < cutoff ** 2)[0])
    assert not inside <= set(lst2[0, :nl2[0]])


def test_force_newton_and_energy():
    L = (16, 16, 16)
    c, srt, cnt, st, col = _nl_state(L, 10, 13, 2)
    col.species[:] = [1, 2]
    eps = np.array([[1.0, 0.5]])
    sig = np.array([[2.0, 2.5]])
    cut = sig * 2 ** (1 / 6)
    lj = O.lj_params(eps, sig, cut)
    rad = cut.max() + 1.0
    lst, nl, _, _ = O.update_list(c, cnt, st, col.pos, col.species, srt.pos, srt.species, rad,
                                   O.make_stencil(1.0, rad), lj, 3000)
    f2 = np.zeros_like(srt.pos)
    f1 = np.zeros_like(col.pos)
    fl = np.zeros(srt.n, np.int32)
    e = O.compute_force(col.pos, col.species, f1, srt.pos, srt.species, f2, fl, lst, nl, L, lj)
    assert np.allclose(f1.sum(axis=0) + f2.sum(axis=0), 0, atol=1e-9 * np.abs(f2).sum())
    # brute-force energy and flag bits over all pairs
    e_bf = 0.0
    near = np.zeros((2, srt.n), bool)
    for i in range(2):
        for j in range(srt.n):
            d = O.rel_pos(col.pos[i], srt.pos[j], L)
            r2 = d @ d
            p = i  # (solvent species 1, colloid species i+1)
            if r2 <= lj.cut_sq[p]:
                e_bf += O.lj_energy(r2, lj.eps[p], lj.sigma[p], lj.cut_energy[p])
                near[i, j] = True
    # MD on every particle inside a cut-off; CATALYZED iff inside the species-1 colloid's
    assert near[0].any() and near[1].any()
    assert ((fl >> O.MD_BIT & 1) == (near[0] | near[1])).all()
    assert ((fl >> O.CATALYZED_BIT & 1) == near[0]).all()
    assert abs(e - e_bf) < 1e-10 * max(1, abs(e_bf))


def test_force_n2_newton():
    rng = np.random.default_rng(14)
    L = (30.0, 30.0, 30.0)
    n = 40
    pos = rng.random((n, 3)) * 30
    sp = np.ones(n, np.int32)
    f = np.zeros((n, 3))
    sig = 5.0
    lj = O.lj_params([[1.0]], [[sig]], [[sig * 2 ** (1 / 6)]])
    O.compute_force_n2(pos, sp, f, L, lj)
    assert np.allclose(f.sum(axis=0), 0, atol=1e-9 * max(1, np.abs(f).sum()))


def test_catalytic_helpers():
    L = (20.0, 20.0, 20.0)
    cpos = np.array([[10.0, 10, 10], [14.5, 10, 10]])
    pos = np.array([[11.0, 10, 10], [10, 10, 13.5], [10, 10, 1.0], [1.0, 1, 1], [14.5, 11, 10]])
    sp = np.array([1, 1, 2, 2, 1], np.int32)
    fl = np.zeros(5, np.int32)
    lst = np.array([0, 1, 4], np.int32)
    O.flag_particles_nl(lst, 3, cpos[0], pos, sp, fl, L, 2.5 ** 2)
    assert list(fl >> O.REAC_BIT & 1) == [1, 0, 0, 0, 0]
    pos[0] = [10, 17.0, 10]
    O.change_species(cpos, pos, sp, fl, L, 2.5 ** 2, 2.5 ** 2, 0.5, np.array([0.4, 0, 0, 0, 0]))
    assert sp[0] == 2 and fl[0] == 0
    O.refuel(cpos, pos, sp, L, 20)  # far = (20*0.45)^2 = 81 in single precision
    assert list(sp) == [2, 1, 2, 1, 1]


def test_numpy_threefry_known_answers():
    # Random123 kat_vectors, threefry2x64 with 20 rounds: zero input, and the digits of pi
    from oracle import threefry as T
    a, b = T.threefry2x64(0, 0, 0, 0)
    assert (int(a), int(b)) == (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)
    a, b = T.threefry2x64(0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0, 0x082EFA98EC4E6C89)
    assert (int(a), int(b)) == (0x263C7D30BB0F0AF1, 0x56BE8361D3311526)


def test_numpy_threefry_matches_library_and_sphere():
    import ctypes as C
    from oracle import threefry as T
    from rmpcdmd_amd import capi
    rng = np.random.default_rng(3)
    v = rng.integers(0, 2 ** 63, size=(50, 4), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    a, b = T.threefry2x64(v[:, 0], v[:, 1], v[:, 2], v[:, 3])
    for i in range(50):
        o = (C.c_uint64 * 2)()
        capi.lib().rmpcd_threefry2x64_20((C.c_uint64 * 2)(int(v[i, 0]), int(v[i, 1])),
                                         (C.c_uint64 * 2)(int(v[i, 2]), int(v[i, 3])), o)
        assert (o[0], o[1]) == (int(a[i]), int(b[i]))
    ax = T.srd_axes(7, 3, 4096)
    assert np.allclose((ax * ax).sum(axis=1), 1.0, atol=1e-14)
    assert np.abs(ax.mean(axis=0)).max() < 0.05  # uniform on the sphere
    assert abs((ax[:, 2] ** 2).mean() - 1 / 3) < 0.03
    # the shift draws are uniforms in [0, 1)
    u = np.array([T.shift_uniforms(5, s) for s in range(200)])
    assert (u >= 0).all() and (u < 1).all() and abs(u.mean() - 0.5) < 0.05


def _program_state(L, rho, nc, seed, sigma):
    """a synthetic download-style state (what the device programs hand to the oracle
    loops): sorted solvent outside the colloid cutoffs, colloids apart"""
    rng = np.random.default_rng(seed)
    c = O.cells_init(L)
    n = rho * int(np.prod(L))
    ps = O.PState(n)
    cut = sigma * 2 ** (1 / 6)
    cp = np.array(L, float) / 2 + np.zeros((nc, 3))
    if nc > 1:
        cp = rng.random((nc, 3)) * np.array(L)
    pos = rng.random((n, 3)) * np.array(L)
    for _ in range(50):
        bad = np.zeros(n, bool)
        for j in range(nc):
            d = pos - cp[j]
            d -= np.round(d / np.array(L)) * np.array(L)
            bad |= (d * d).sum(axis=1) < cut * cut
        if not bad.any():
            break
        pos[bad] = rng.random((bad.sum(), 3)) * np.array(L)
    ps.pos[:] = pos
    ps.vel[:] = rng.normal(size=(n, 3))
    ps.vel -= ps.vel.mean(axis=0)
    srt, _, cnt, st = O.sort(c, ps)
    d = dict(n_solvent=n, pos=srt.pos, vel=srt.vel, force=np.zeros((n, 3)), force_old=np.zeros((n, 3)),
             image=srt.image, id=srt.id, species=srt.species, flags=srt.flags, pos_old=srt.pos.copy(),
             cell_count=cnt, cell_start=st, origin=np.zeros(3), md_steps=0, cycles=0, n_extra_sorting=0,
             e1=0.0, e2=0.0, temperature=0.0, n_colloids=nc, nmax=6000, mass=50.0, last_sort=0)
    if nc:
        for k in ("vel", "force", "force_old"):
            d["colloid_" + k] = np.zeros((nc, 3))
        d["colloid_pos"] = cp
        d["colloid_image"] = np.zeros((nc, 3), np.int32)
        d["colloid_pos_old"] = cp.copy()
    return d


def test_oracle_program_loops_conserve():
    """the oracle's driver loops (the checker of the device programs): the fluid loop keeps
    the total momentum and the energy of SRD + free streaming; the many-colloid loop keeps
    the total momentum (pairwise LJ, SRD); the fixed-sphere loop converts A to B at the
    sphere and keeps every particle in the box bookkeeping (id set unchanged)"""
    from oracle.programs import ColloidsOracle, FluidOracle, SphereOracle
    st = _program_state((8, 6, 10), 10, 0, 1, 1.0)
    f = FluidOracle(st, (8, 6, 10), 0.1, 2, 3)
    p0, e0 = f.s.vel.sum(axis=0), (f.s.vel ** 2).sum()
    for _ in range(3):
        f.cycle()
    assert np.allclose(f.s.vel.sum(axis=0), p0, atol=1e-9)
    assert abs((f.s.vel ** 2).sum() - e0) < 1e-9 * e0
    assert sorted(f.s.id.tolist()) == list(range(f.s.n))
    st = _program_state((12, 12, 12), 10, 1, 2, 2.0)
    s = SphereOracle(st, (12, 12, 12), 0.5, 10, 4, sigma=2.0, bulk_rate=0.01)
    for _ in range(2):
        s.cycle()
    assert sorted(s.s.id.tolist()) == list(range(s.s.n))
    assert (s.s.species == 2).sum() > 0
    st = _program_state((16, 16, 16), 10, 4, 3, 1.5)
    c = ColloidsOracle(st, (16, 16, 16), 0.2, 10, 5, sigma=1.5)
    mass = float(st["mass"])
    p0 = c.s.vel.sum(axis=0) + mass * c.col.vel.sum(axis=0)
    for _ in range(2):
        c.cycle()
    p1 = c.s.vel.sum(axis=0) + mass * c.col.vel.sum(axis=0)
    assert np.abs(p1 - p0).max() / c.s.n < 1e-12
    assert c.last_sort == c.md_steps
