"""The single_dimer_pbc time loop restated over the C oracle -- TEST INFRASTRUCTURE ONLY
(the whole-step parity checker and the CPU baseline of bench.py).

Follows programs/single_dimer_pbc.f90: the MD step (:306-383), the collision cycle
(:386-443) and the helpers flag_particles_nl / change_species / refuel (:493-565).
Random draws are injected through a provider object (see ``NumpyDraws``):
    shift(cycle) -> 3 uniforms                      (cell_system.f90:159-161)
    srd_axes(cycle, cells, cnt, st, species) -> dense (ncells, 3) axes
    change(md_step, ids) -> one uniform per slot     (single_dimer_pbc.f90:533)
    bulk(cycle, ncells) -> one uniform per cell      (mpcd.f90:636)
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


class NumpyDraws:
    """uniform draws from numpy, consumed in the reference's single-thread order"""

    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)

    def shift(self, cycle):
        return self.rng.random(3)

    def srd_axes(self, cycle, c, cnt, st, species):
        u = self.rng.random(4 * c.N + 64)
        axes, used = O.srd_axes_from_stream(c, cnt, st, species, u)
        if used < 0:
            u = self.rng.random(16 * c.N + 64)
            axes, used = O.srd_axes_from_stream(c, cnt, st, species, u)
        return axes

    def change(self, md_step, ids):
        return self.rng.random(len(ids))

    def bulk(self, cycle, ncells):
        return self.rng.random(ncells)


class TimingDraws(NumpyDraws):
    """for timing loops: as NumpyDraws, except that change_species draws inline, and only
    for the particles the reference draws for (single_dimer_pbc.f90:533), from a
    counter-based stream (``change_seed``) instead of a dense array per MD step"""

    def __init__(self, seed):
        super().__init__(seed)
        self.change_seed = int(seed) * 0x9E3779B97F4A7C15 % (1 << 64)


class DimerOracle:
    """state and loop of single_dimer_pbc on the oracle (FP64, 0-based)"""

    def __init__(self, L, solvent, colloids, pos_old, colloid_pos_old, mass, d, tau, n_md, sigma_C=2.0,
                 sigma_N=2.0, eps_C=(1.0, 1.0), eps_N=(1.0, 1.0), eps_CC=1.0, eps_NC=1.0, eps_NN=1.0,
                 prob=1.0, bulk_rmpcd=False, bulk_rate=0.01, skin=2.0, nmax=4000, origin=(0, 0, 0)):
        self.L = tuple(int(v) for v in L)
        self.c = O.cells_init(self.L)
        O.set_origin(self.c, origin)
        self.edges = tuple(float(v) for v in self.L)
        self.s = solvent
        self.col = colloids
        self.pos_old = pos_old
        self.cpos_old = colloid_pos_old
        self.mass = np.asarray(mass, np.float64)
        self.d, self.tau, self.n_md = d, tau, n_md
        self.dt = tau / n_md
        self.skin = skin
        self.prob, self.bulk_rmpcd, self.bulk_rate = prob, bulk_rmpcd, bulk_rate
        sig = np.array([[sigma_C, sigma_N], [sigma_C, sigma_N]])
        eps = np.array([[eps_C[0], eps_N[0]], [eps_C[1], eps_N[1]]])
        self.lj_sc = O.lj_params(eps, sig, sig * 2 ** (1 / 6))
        csig = np.array([[2 * sigma_C, sigma_C + sigma_N], [sigma_C + sigma_N, 2 * sigma_N]])
        ceps = np.array([[eps_CC, eps_NC], [eps_NC, eps_NN]])
        self.lj_cc = O.lj_params(ceps, csig, csig * 2 ** (1 / 6))
        self.max_cut = float(np.max(sig * 2 ** (1 / 6)))
        self.radius = self.max_cut + skin
        self.stencil = O.make_stencil(1.0, self.radius)
        self.nmax = nmax
        self.cnt, self.st = O.count_particles(self.c, self.s.pos)
        self.lst = self.nl = None
        self.md_steps = 0
        self.cycles = 0
        self.n_extra = 0
        self.e1 = self.e2 = 0.0

    # -- pieces -----------------------------------------------------------------
    def sort(self):
        srt, i2i, cnt, st = O.sort(self.c, self.s)
        self.s, self.cnt, self.st = srt, cnt, st

    def update_list(self):
        self.lst, self.nl, self.is_md, self.is_reac = O.update_list(
            self.c, self.cnt, self.st, self.col.pos, self.col.species, self.s.pos, self.s.species, self.radius,
            self.stencil, self.lj_sc, self.nmax)

    def md_step(self, draws):
        """single_dimer_pbc.f90:306-383"""
        dt, s, col = self.dt, self.s, self.col
        O.md_pos(s.pos, s.vel, s.force, dt)
        pos_rattle = col.pos.copy()
        O.colloid_md_pos(col.pos, col.vel, col.force, self.mass, dt)
        O.rattle_dimer_pos(col.pos, col.vel, pos_rattle, col.species, self.mass, self.d, dt, self.edges)
        so_max = O.maximum_displacement(s.pos, self.pos_old, s.species)
        co_max = O.maximum_displacement(col.pos, self.cpos_old, col.species)
        if co_max >= self.skin * 0.1 or so_max >= self.skin * 0.89:
            O.apply_pbc(s.pos, s.image, self.edges)
            O.apply_pbc(col.pos, col.image, self.edges)
            self.sort()
            s = self.s
            self.update_list()
            self.pos_old = s.pos.copy()
            self.cpos_old = col.pos.copy()
            self.n_extra += 1
        s.force, s.force_old = s.force_old, s.force
        col.force, col.force_old = col.force_old, col.force
        s.force[:] = 0
        col.force[:] = 0
        self.e1 = O.compute_force(col.pos, col.species, col.force, s.pos, s.species, s.force, s.flags, self.lst,
                                  self.nl, self.edges, self.lj_sc)
        self.e2 = O.compute_force_n2(col.pos, col.species, col.force, self.edges, self.lj_cc)
        O.md_vel(s.vel, s.force, s.force_old, s.flags, dt)
        O.colloid_md_vel(col.vel, col.force, col.force_old, self.mass, dt)
        O.rattle_dimer_vel(col.pos, col.vel, col.species, self.mass, self.d, dt, self.edges)
        O.flag_particles_nl(self.lst[0], self.nl[0], col.pos[0], s.pos, s.species, s.flags, self.edges,
                            self.lj_sc.cut_sq[0])
        if hasattr(draws, "change_seed"):
            O.change_species_seeded(col.pos, s.pos, s.species, s.flags, self.edges, self.lj_sc.cut_sq[0],
                                    self.lj_sc.cut_sq[2], self.prob, draws.change_seed, self.md_steps, s.id)
        else:
            if getattr(draws, "per_id", False):
                # a draw that depends on (md step, id) alone: made for the flagged particles only
                u = np.zeros(s.n)
                cand = np.flatnonzero(s.flags & (1 << O.REAC_BIT))
                if cand.size:
                    u[cand] = draws.change(self.md_steps, s.id[cand])
            else:
                u = np.ascontiguousarray(draws.change(self.md_steps, s.id))
            O.change_species(col.pos, s.pos, s.species, s.flags, self.edges, self.lj_sc.cut_sq[0],
                             self.lj_sc.cut_sq[2], self.prob, u)
        self.md_steps += 1

    def collide(self, draws):
        """single_dimer_pbc.f90:405-428"""
        u = draws.shift(self.cycles)
        for k in range(3):
            self.c.origin[k] = float(u[k]) - 1  # random_shift
        O.apply_pbc(self.s.pos, self.s.image, self.edges)
        O.apply_pbc(self.col.pos, self.col.image, self.edges)
        self.sort()
        self.update_list()
        self.pos_old = self.s.pos.copy()
        self.cpos_old = self.col.pos.copy()
        axes = draws.srd_axes(self.cycles, self.c, self.cnt, self.st, self.s.species)
        O.simple_mpcd_step(self.c, self.cnt, self.st, self.s.vel, self.s.species, np.ascontiguousarray(axes))
        if self.bulk_rmpcd:
            O.bulk_reaction(self.c, self.cnt, self.st, self.is_reac, self.s.species, 2, 1, self.bulk_rate,
                            self.tau, np.ascontiguousarray(draws.bulk(self.cycles, self.c.N)))
        else:
            O.refuel(self.col.pos, self.s.pos, self.s.species, self.edges, self.L[0])
        self.cycles += 1

    def cycle(self, draws):
        for _ in range(self.n_md):
            self.md_step(draws)
        self.collide(draws)


def from_download(st, L, d, tau, n_md, sigma_C=2.0, sigma_N=2.0, **kw):
    """a DimerOracle holding a state downloaded from the device program (FP64 form):
    st is rmpcdmd_amd.sim.SingleDimer.download()'s dict. The neighbour lists and the
    cell counts are recomputed by the oracle (and can be compared with st's)."""
    n = int(st["n_solvent"])
    s = PState_from(st, "", n)
    col = PState_from(st, "colloid_", 2)
    col.species[:] = [1, 2]
    drv = DimerOracle(L, s, col, np.array(st["pos_old"], np.float64), np.array(st["colloid_pos_old"]),
                      st["mass"], d, tau, n_md, sigma_C=sigma_C, sigma_N=sigma_N, nmax=int(st["nmax"]),
                      origin=st["origin"], **kw)
    drv.update_list()
    drv.md_steps, drv.cycles, drv.n_extra = int(st["md_steps"]), int(st["cycles"]), int(st["n_extra_sorting"])
    drv.e1, drv.e2 = float(st["e1"]), float(st["e2"])
    return drv


def PState_from(st, prefix, n):
    p = O.PState(n)
    for k in ("pos", "vel", "force", "force_old", "image"):
        getattr(p, k)[:] = st[prefix + k]
    if not prefix:
        p.id[:] = st["id"]
        p.species[:] = st["species"]
        p.flags[:] = st["flags"]
    return p


def synthetic_dimer(L, rho=10, seed=0, sigma=2.0, d=4.5, T=1.0, **kw):
    """a set-up like single_dimer_pbc.f90:240-296 with numpy draws: Maxwell-Boltzmann
    velocities without net momentum, solvent placed outside the colloid cutoffs, sorted,
    first forces and force_old = force"""
    rng = np.random.default_rng(seed)
    L = tuple(int(v) for v in L)
    n = rho * L[0] * L[1] * L[2]
    s = O.PState(n)
    s.vel[:] = rng.normal(size=(n, 3)) * np.sqrt(T)
    s.vel -= s.vel.mean(axis=0)
    s.force[:] = 0
    s.force_old[:] = 0
    s.image[:] = 0
    s.flags[:] = 0
    col = O.PState(2)
    col.species[:] = [1, 2]
    col.vel[:] = 0
    col.pos[0] = np.asarray(L, float) / 2
    col.pos[1] = col.pos[0]
    col.pos[1, 0] += d
    cut = sigma * 2 ** (1 / 6)
    pos = rng.random((n, 3)) * L
    for _ in range(100):
        bad = np.zeros(n, bool)
        for j in range(2):
            dd = pos - col.pos[j]
            dd -= np.round(dd / np.asarray(L)) * np.asarray(L)
            bad |= (dd * dd).sum(axis=1) < cut * cut
        if not bad.any():
            break
        pos[bad] = rng.random((bad.sum(), 3)) * L
    s.pos[:] = pos
    pi_f = float(np.float32(3.14159265))
    m = (((rho * sigma ** 3) * 4) * pi_f) / 3
    drv = DimerOracle(L, s, col, s.pos.copy(), np.zeros((2, 3)), (m, m), d, sigma_C=sigma, sigma_N=sigma, **kw)
    unsorted = drv.s.pos.copy()
    drv.sort()
    drv.pos_old = unsorted
    drv.update_list()
    drv.e1 = O.compute_force(col.pos, col.species, col.force, drv.s.pos, drv.s.species, drv.s.force,
                             drv.s.flags, drv.lst, drv.nl, drv.edges, drv.lj_sc)
    drv.e2 = O.compute_force_n2(col.pos, col.species, col.force, drv.edges, drv.lj_cc)
    drv.s.force_old[:] = drv.s.force
    col.force_old[:] = col.force
    return drv
