"""The driver loops of setup_simple_fluid, setup_single_catalytic_fixed_sphere and
n_colloids_pbc restated over the C oracle -- TEST INFRASTRUCTURE ONLY (the whole-loop
parity checker of the device programs).

Each oracle starts from a state downloaded from the device program
(rmpcdmd_amd.programs.Program.download(), FP64 form) and consumes the device program's
Threefry draws (oracle.threefry), so that both advance the same trajectory:
    FluidOracle     test/setup_simple_fluid.f90:108-118
    SphereOracle    test/setup_single_catalytic_fixed_sphere.f90:188-253, helpers :311-349
    ColloidsOracle  programs/n_colloids_pbc.f90:227-346 (thermostat phase and production)
"""
from __future__ import annotations

import numpy as np

from . import oracle as O
from . import threefry as T


def _pstate(st, prefix, n):
    p = O.PState(n)
    for k in ("pos", "vel", "force", "force_old", "image"):
        getattr(p, k)[:] = st[prefix + k]
    if prefix:
        p.species[:] = 1
    else:
        p.id[:] = st["id"]
        p.species[:] = st["species"]
        p.flags[:] = st["flags"]
    return p


class _Loop:
    def __init__(self, st, L, tau, n_md, seed):
        self.L = tuple(int(v) for v in L)
        self.edges = tuple(float(v) for v in self.L)
        self.c = O.cells_init(self.L)
        O.set_origin(self.c, st["origin"])
        self.s = _pstate(st, "", int(st["n_solvent"]))
        self.pos_old = np.array(st["pos_old"], np.float64)
        # the cells of the last sort (its origin may have been replaced by the next shift)
        self.cnt = np.array(st["cell_count"], np.int32)
        self.st = np.array(st["cell_start"], np.int32)
        self.tau, self.n_md, self.dt = tau, n_md, tau / n_md
        self.seed = int(seed)
        self.md_steps, self.cycles = int(st["md_steps"]), int(st["cycles"])
        self.n_extra = int(st["n_extra_sorting"])
        self.e1, self.e2, self.temperature = float(st["e1"]), float(st["e2"]), float(st["temperature"])

    def sort(self):
        srt, i2i, cnt, st = O.sort(self.c, self.s)
        self.s, self.cnt, self.st = srt, cnt, st

    def shift(self, s):
        u = T.shift_uniforms(self.seed, s)
        for k in range(3):
            self.c.origin[k] = float(u[k]) - 1  # random_shift, cell_system.f90:153-163

    def srd(self):
        axes = T.srd_axes(self.seed, self.cycles, int(self.c.N))
        O.simple_mpcd_step(self.c, self.cnt, self.st, self.s.vel, self.s.species, np.ascontiguousarray(axes))


class FluidOracle(_Loop):
    """setup_simple_fluid.f90:108-118: SRD, streaming (n_md sub-steps of tau/n_md),
    grid shift, sort"""

    def cycle(self):
        self.srd()
        for _ in range(self.n_md):
            O.mpcd_stream_periodic(self.c, self.s.pos, self.s.vel, self.s.image, self.dt)
        self.shift(self.cycles + 1)
        self.sort()
        self.md_steps += self.n_md
        self.cycles += 1


class _WithColloids(_Loop):
    def __init__(self, st, L, tau, n_md, seed, lj, radius, nmax):
        super().__init__(st, L, tau, n_md, seed)
        nc = int(st["n_colloids"])
        self.col = _pstate(st, "colloid_", nc)
        self.cpos_old = np.array(st["colloid_pos_old"], np.float64)
        self.lj = lj
        self.radius = radius
        self.nmax = int(nmax)
        self.stencil = O.make_stencil(1.0, radius)
        self.update_list()

    def update_list(self):
        self.lst, self.nl, self.is_md, self.is_reac = O.update_list(
            self.c, self.cnt, self.st, self.col.pos, self.col.species, self.s.pos, self.s.species, self.radius,
            self.stencil, None, self.nmax)

    def swap_forces(self, zero_colloid):
        s, col = self.s, self.col
        s.force, s.force_old = s.force_old, s.force
        col.force, col.force_old = col.force_old, col.force
        s.force[:] = 0
        if zero_colloid:
            col.force[:] = 0
        self.e1 = O.compute_force(col.pos, col.species, col.force, s.pos, s.species, s.force, s.flags, self.lst,
                                  self.nl, self.edges, self.lj)


class SphereOracle(_WithColloids):
    """the fixed catalytic sphere (setup_single_catalytic_fixed_sphere.f90:188-253)"""

    def __init__(self, st, L, tau, n_md, seed, sigma=4.0, epsilon=1.0, prob=1.0, bulk_rate=0.01, skin=1.5):
        sc = sigma * 2 ** (1 / 6)
        lj = O.lj_params([[epsilon], [epsilon]], [[sigma], [sigma]], [[sc], [sc]])
        self.skin, self.prob, self.bulk_rate = skin, prob, bulk_rate
        super().__init__(st, L, tau, n_md, seed, lj, sc + skin, st["nmax"])

    def md_step(self):
        s, col = self.s, self.col
        O.md_pos(s.pos, s.vel, s.force, self.dt)
        if O.maximum_displacement(s.pos, self.pos_old, s.species) >= self.skin:
            O.apply_pbc(s.pos, s.image, self.edges)
            self.sort()
            self.update_list()
            self.pos_old = self.s.pos.copy()
            self.n_extra += 1
        self.swap_forces(zero_colloid=False)  # the driver zeroes the solvent forces only
        s = self.s
        O.md_vel(s.vel, s.force, s.force_old, s.flags, self.dt)
        n0 = int(self.nl[0])
        ids = s.id[self.lst[0, :n0]].astype(np.uint64)
        a, _ = T.threefry2x64(self.md_steps, ids, self.seed, T.CHANGE)
        draws = np.ascontiguousarray(T.u01(a)) if n0 else np.zeros(1)
        O.flag_particles_nl_prob(self.lst[0], n0, col.pos[0], s.pos, s.species, s.flags, self.edges,
                                 self.lj.cut_sq[0], self.prob, draws)
        O.change_species_sphere(col.pos[0], s.pos, s.species, s.flags, self.edges, self.lj.cut_sq[0])
        self.md_steps += 1

    def collide(self):
        self.temperature = O.compute_temperature(self.c, self.cnt, self.st, self.s.vel, self.s.species)
        self.shift(self.cycles)
        unsorted = self.s.pos.copy()  # the sort's switch leaves pos_old with these
        self.sort()
        self.pos_old = unsorted
        self.update_list()
        self.srd()
        a, _ = T.threefry2x64(self.cycles, np.arange(self.c.N, dtype=np.uint64), self.seed, T.REACTION)
        O.bulk_reaction(self.c, self.cnt, self.st, self.is_reac, self.s.species, 2, 1, self.bulk_rate, self.tau,
                        np.ascontiguousarray(T.u01(a)))
        self.cycles += 1

    def cycle(self):
        for _ in range(self.n_md):
            self.md_step()
        self.collide()


class ColloidsOracle(_WithColloids):
    """many colloids with lazy streaming (n_colloids_pbc.f90:286-346)"""

    def __init__(self, st, L, tau, n_md, seed, sigma=2.5, epsilon=1.0, epsilon_colloids=1.0, skin=0.8, T=1.0,
                 n_thermo=0, T_final=None, n_loop=0, thermostat=False, hydro=True):
        self.T0, self.n_thermo, self.n_loop = T, n_thermo, n_loop
        self.T_final = T if T_final is None else T_final
        self.thermostat, self.hydro = thermostat, hydro
        sc = sigma * 2 ** (1 / 6)
        lj = O.lj_params([[epsilon]], [[sigma]], [[sc]])
        self.lj_cc = O.lj_params([[epsilon_colloids]], [[2 * sigma]], [[2 * (3 * sigma)]])
        self.skin = skin
        self.mass = np.full(int(st["n_colloids"]), float(st["mass"]))
        self.last_sort = int(st["last_sort"])
        super().__init__(st, L, tau, n_md, seed, lj, sc + skin, st["nmax"])

    def catch_up(self, k):
        """cell_md_pos/vel with md_flag false over k steps: the cells streamed lazily"""
        s = self.s
        O.cell_md_pos(self.c, self.cnt, self.st, self.is_md, s.pos, s.vel, s.force, k * self.dt, False)
        O.cell_md_vel(self.c, self.cnt, self.st, self.is_md, s.vel, s.force, s.force_old, s.flags, k * self.dt,
                      False)

    def resort(self):
        O.apply_pbc(self.s.pos, self.s.image, self.edges)
        O.apply_pbc(self.col.pos, self.col.image, self.edges)
        self.sort()
        self.update_list()
        self.pos_old = self.s.pos.copy()
        self.cpos_old = self.col.pos.copy()

    def md_step(self):
        s, col, dt = self.s, self.col, self.dt
        O.cell_md_pos(self.c, self.cnt, self.st, self.is_md, s.pos, s.vel, s.force, dt, True)
        O.colloid_md_pos(col.pos, col.vel, col.force, self.mass, dt)
        k1 = self.md_steps + 1 - self.last_sort
        so_max = O.cell_maximum_displacement(self.c, self.cnt, self.st, self.is_md, s.pos, self.pos_old, s.species,
                                             dt * k1)
        co_max = O.maximum_displacement(col.pos, self.cpos_old, col.species)
        if co_max >= self.skin * 0.1 or so_max >= self.skin * 0.9:
            self.catch_up(k1)
            self.resort()
            self.last_sort = self.md_steps + 1
            self.n_extra += 1
        self.swap_forces(zero_colloid=True)
        s = self.s
        self.e2 = O.compute_force_n2(col.pos, col.species, col.force, self.edges, self.lj_cc)
        O.cell_md_vel(self.c, self.cnt, self.st, self.is_md, s.vel, s.force, s.force_old, s.flags, dt, True)
        O.colloid_md_vel(col.vel, col.force, col.force_old, self.mass, dt)
        self.md_steps += 1

    def mpcat(self, temp):
        normals = T.normals3(self.seed, self.cycles, self.s.id)
        O.mpcat_step(self.c, self.cnt, self.st, self.s.vel, self.s.species, np.ascontiguousarray(normals), temp,
                     self.hydro)

    def collide(self):
        self.shift(self.cycles)
        self.catch_up(self.md_steps - self.last_sort)
        self.resort()
        self.last_sort = self.md_steps
        if self.cycles < self.n_thermo:
            # thermostat phase (n_colloids_pbc.f90:276-279)
            self.mpcat(self.T0)
            if (self.cycles + 1) % 10 == 0:
                O.rescale_cells(self.c, self.cnt, self.st, self.s.vel, self.s.species, self.T0)
        else:
            # production loop i with the temperature ramp (:334-336)
            i = self.cycles - self.n_thermo + 1
            temp = self.T0 + (i - 1) * (self.T_final - self.T0) / (self.n_loop - 1) if self.n_loop > 1 else self.T0
            if self.thermostat:
                self.mpcat(temp)
            else:
                self.srd()
        self.temperature = O.compute_temperature(self.c, self.cnt, self.st, self.s.vel, self.s.species)
        self.cycles += 1

    def cycle(self):
        for _ in range(self.n_md):
            self.md_step()
        self.collide()
