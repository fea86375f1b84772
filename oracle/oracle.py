"""ctypes front-end of the C oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- never by the product package.  Arrays follow
the oracle's conventions: (3, N) column-major vectors are numpy arrays of shape (N, 3)
in C order (same memory), all indices 0-based.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

REAC_BIT, MD_BIT, WALL_BIT, PAST_MD_BIT, CATALYZED_BIT = 1, 2, 3, 4, 6
PERIODIC_BC, BOUNCE_BACK_BC, SPECULAR_BC = 1, 2, 3


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)


class Cells(C.Structure):
    _fields_ = [("L", C.c_int * 3), ("M", C.c_int * 3), ("N", C.c_int), ("a", C.c_double),
                ("edges", C.c_double * 3), ("origin", C.c_double * 3),
                ("has_walls", C.c_int), ("bc", C.c_int * 3), ("max_v", C.c_double)]


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("pos", C.c_void_p), ("vel", C.c_void_p),
                ("force", C.c_void_p), ("force_old", C.c_void_p), ("image", C.c_void_p),
                ("id", C.c_void_p), ("species", C.c_void_p), ("flags", C.c_void_p)]


class LJ(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("eps", C.c_double * 16),
                ("sigma", C.c_double * 16), ("cut", C.c_double * 16),
                ("cut_sq", C.c_double * 16), ("cut_energy", C.c_double * 16)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_maximum_displacement.restype = C.c_double
        _lib.orc_cell_maximum_displacement.restype = C.c_double
        _lib.orc_compute_temperature.restype = C.c_double
        _lib.orc_compute_force.restype = C.c_double
        _lib.orc_compute_force_n2.restype = C.c_double
        _lib.orc_lj_energy.restype = C.c_double
        _lib.orc_lj_energy.argtypes = [C.c_double] * 4
        _lib.orc_srd_axes_from_stream.restype = C.c_int64
        _lib.orc_elastic_network.restype = C.c_double
        for name in ("orc_md_pos", "orc_md_vel", "orc_mpcd_stream_periodic",
                     "orc_mpcd_stream_nogravity_zwall", "orc_mpcd_stream_xforce_yzwall", "orc_wall_mpcd_step"):
            getattr(_lib, name).restype = None
    return _lib


def P(a):
    """pointer of a contiguous numpy array (or None)"""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


D = C.c_double
I64 = C.c_int64


# ---- hilbert ---------------------------------------------------------------------
def compact_p_to_h(p, M):
    pa = (C.c_int * 3)(*[int(v) for v in p])
    ma = (C.c_int * 3)(*[int(v) for v in M])
    return lib().orc_compact_p_to_h(pa, ma)


def compact_h_to_p(h, M):
    ma = (C.c_int * 3)(*[int(v) for v in M])
    out = (C.c_int * 3)()
    lib().orc_compact_h_to_p(int(h), ma, out)
    return tuple(out)


def p_to_h(p, m):
    pa = (C.c_int * 3)(*[int(v) for v in p])
    return lib().orc_p_to_h(pa, int(m))


def h_to_p(h, m):
    out = (C.c_int * 3)()
    lib().orc_h_to_p(int(h), int(m), out)
    return tuple(out)


# ---- cells -----------------------------------------------------------------------
def cells_init(L, a=1.0, has_walls=False):
    c = Cells()
    la = (C.c_int * 3)(*[int(v) for v in L])
    lib().orc_cells_init(C.byref(c), la, D(a), int(bool(has_walls)))
    return c


def set_origin(c, origin):
    for k in range(3):
        c.origin[k] = float(origin[k])


def cartesian_indices(c, x):
    xa = (C.c_double * 3)(*[float(v) for v in x])
    out = (C.c_int * 3)()
    lib().orc_cartesian_indices(C.byref(c), xa, out)
    return tuple(out)


def count_particles(c, pos, nthreads=0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    cnt = np.zeros(c.N, np.int32)
    st = np.zeros(c.N, np.int32)
    lib().orc_count_particles(C.byref(c), I64(pos.shape[0]), P(pos), P(cnt), P(st), int(nthreads))
    return cnt, st


class PState:
    """numpy particle state in the oracle layout (N,3) C-order == Fortran (3,N)"""

    def __init__(self, n):
        self.n = n
        self.pos = np.zeros((n, 3))
        self.vel = np.zeros((n, 3))
        self.force = np.zeros((n, 3))
        self.force_old = np.zeros((n, 3))
        self.image = np.zeros((n, 3), np.int32)
        self.id = np.arange(n, dtype=np.int32)
        self.species = np.ones(n, np.int32)
        self.flags = np.zeros(n, np.int32)

    def copy(self):
        o = PState(self.n)
        for k in ("pos", "vel", "force", "force_old", "image", "id", "species", "flags"):
            setattr(o, k, getattr(self, k).copy())
        return o

    def as_struct(self):
        return Particles(self.n, P(self.pos), P(self.vel), P(self.force), P(self.force_old),
                         P(self.image), P(self.id), P(self.species), P(self.flags))


def sort(c, ps, nthreads=0):
    """particle_system%sort: returns (sorted PState, id_to_idx, cell_count, cell_start);
    the unsorted input plays the reference's pos_old after the pointer switch."""
    out = PState(ps.n)
    i2i = np.zeros(ps.n, np.int32)
    cnt = np.zeros(c.N, np.int32)
    st = np.zeros(c.N, np.int32)
    s_in, s_out = ps.as_struct(), out.as_struct()
    lib().orc_sort(C.byref(c), C.byref(s_in), C.byref(s_out), P(i2i), P(cnt), P(st), int(nthreads))
    return out, i2i, cnt, st


def maximum_displacement(pos, pos_old, species):
    return lib().orc_maximum_displacement(I64(pos.shape[0]), P(pos), P(pos_old), P(species))


def cell_maximum_displacement(c, cnt, st, is_md, pos, pos_old, species, delta_t):
    return lib().orc_cell_maximum_displacement(C.byref(c), P(cnt), P(st), P(is_md), P(pos),
                                               P(pos_old), P(species), D(delta_t))


# ---- mpcd ------------------------------------------------------------------------
def rand_sphere(u):
    u = np.ascontiguousarray(u, np.float64)
    out = (C.c_double * 3)()
    used = lib().orc_rand_sphere(P(u), int(u.size), out)
    return np.array(out[:]), used


def srd_axes_from_stream(c, cnt, st, species, u):
    u = np.ascontiguousarray(u, np.float64)
    axes = np.zeros((c.N, 3))
    used = lib().orc_srd_axes_from_stream(C.byref(c), P(cnt), P(st), P(species), P(u),
                                          I64(u.size), P(axes))
    return axes, used


def simple_mpcd_step(c, cnt, st, vel, species, axes, alpha=None):
    lib().orc_simple_mpcd_step(C.byref(c), P(cnt), P(st), P(vel), P(species), P(axes),
                               0 if alpha is None else 1, D(0.0 if alpha is None else alpha))


def compute_temperature(c, cnt, st, vel, species):
    return lib().orc_compute_temperature(C.byref(c), P(cnt), P(st), P(vel), P(species))


def temperature_profile(c, cnt, st, vel, species, xmin, dx, n_bins):
    """the tz profile of compute_temperature: (data, count) over n_bins bins from zero"""
    data = np.zeros(n_bins)
    count = np.zeros(n_bins, np.int32)
    lib().orc_temperature_profile(C.byref(c), P(cnt), P(st), P(vel), P(species), D(xmin), D(dx), int(n_bins),
                                  P(data), P(count))
    return data, count


def mpcd_stream_periodic(c, pos, vel, image, dt):
    lib().orc_mpcd_stream_periodic(C.byref(c), I64(pos.shape[0]), P(pos), P(vel), P(image), D(dt))


def bulk_reaction(c, cnt, st, is_reac, species, frm, to, rate, tau, draws):
    lib().orc_bulk_reaction(C.byref(c), P(cnt), P(st), P(is_reac), P(species), int(frm), int(to),
                            D(rate), D(tau), P(draws))


def mpcat_step(c, cnt, st, vel, species, normals, T, hydro=True):
    """the MPC-AT branch of simple_mpcd_step (mpcd.f90:144-160); normals: (n, 3) per slot"""
    lib().orc_mpcat_step(C.byref(c), P(cnt), P(st), P(vel), P(species), P(normals), D(T), 1 if hydro else 0)


def rescale_cells(c, cnt, st, vel, species, T):
    """rescale_cells (mpcd.f90:748-801)"""
    lib().orc_rescale_cells(C.byref(c), P(cnt), P(st), P(vel), P(species), D(T))


def bulk_reaction_endothermic(c, cnt, st, is_reac, vel, species, frm, to, rate, tau, delta_u, draws):
    """bulk_reaction_endothermic (mpcd.f90:646-700); draws: one per cell"""
    lib().orc_bulk_reaction_endothermic(C.byref(c), P(cnt), P(st), P(is_reac), P(vel), P(species), int(frm),
                                        int(to), D(rate), D(tau), D(delta_u), P(draws))


def wall_mpcd_step(c, cnt, st, vel, axes=None, vnorm=None, pnorm=None, wall_temperature=None, wall_n=None,
                   thermostat=False, bulk_temperature=1.0, keep_cell_v=True, alpha=None):
    """wall_mpcd_step (mpcd.f90:193-337); axes, vnorm: (ncells, 3); pnorm: (n, 3)"""
    given = wall_temperature is not None
    wt = (C.c_double * 2)(*(wall_temperature if given else (0.0, 0.0)))
    wn = (C.c_int32 * 2)(*(wall_n if given else (0, 0)))
    lib().orc_wall_mpcd_step(C.byref(c), P(cnt), P(st), P(vel), None, 1 if given else 0, wt, wn,
                             1 if thermostat else 0, D(bulk_temperature), 1 if keep_cell_v else 0,
                             0 if alpha is None else 1, D(0.0 if alpha is None else alpha), P(axes), P(vnorm),
                             P(pnorm))


def mpcd_stream_xforce_yzwall(c, pos, vel, flags, dt, g):
    """mpcd_stream_xforce_yzwall (mpcd.f90:459-511); y/z boundary conditions from c.bc"""
    lib().orc_mpcd_stream_xforce_yzwall(C.byref(c), I64(pos.shape[0]), P(pos), P(vel), P(flags), D(dt), D(g))


def mpcd_stream_nogravity_zwall(c, pos, vel, force, flags, dt):
    lib().orc_mpcd_stream_nogravity_zwall(C.byref(c), I64(pos.shape[0]), P(pos), P(vel),
                                          P(force), P(flags), D(dt))


# ---- md --------------------------------------------------------------------------
def md_pos(pos, vel, force, dt):
    lib().orc_md_pos(I64(pos.shape[0]), P(pos), P(vel), P(force), D(dt))


def md_vel(vel, force, force_old, flags, dt):
    lib().orc_md_vel(I64(vel.shape[0]), P(vel), P(force), P(force_old), P(flags), D(dt))


def apply_pbc(pos, image, edges):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_apply_pbc(I64(pos.shape[0]), P(pos), P(image), ea)


def cell_md_pos(c, cnt, st, is_md, pos, vel, force, dt, md_flag):
    lib().orc_cell_md_pos(C.byref(c), P(cnt), P(st), P(is_md), P(pos), P(vel), P(force), D(dt),
                          int(bool(md_flag)))


def cell_md_vel(c, cnt, st, is_md, vel, force, force_old, flags, dt, md_flag):
    lib().orc_cell_md_vel(C.byref(c), P(cnt), P(st), P(is_md), P(vel), P(force), P(force_old),
                          P(flags), D(dt), int(bool(md_flag)))


def cell_md_pos_wall(c, cnt, st, is_md, pos, vel, force, dt, md_flag, axis):
    lib().orc_cell_md_pos_wall(C.byref(c), P(cnt), P(st), P(is_md), P(pos), P(vel), P(force),
                               D(dt), int(bool(md_flag)), int(axis))


def rattle_dimer_pos(pos, vel, pos_rattle, species, mass, d, dt, edges):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_rattle_dimer_pos(P(pos), P(vel), P(pos_rattle), P(species), P(mass), D(d), D(dt), ea)


def rattle_dimer_vel(pos, vel, species, mass, d, dt, edges):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_rattle_dimer_vel(P(pos), P(vel), P(species), P(mass), D(d), D(dt), ea)


class Body(C.Structure):
    """rigid_body_t (md.f90:39-57) with zero-based inclusive i_start/i_stop; pos_body
    points into the numpy array kept in .body_rows"""
    _fields_ = [("i_start", C.c_int32), ("i_stop", C.c_int32), ("mass", C.c_double),
                ("pos", C.c_double * 3), ("vel", C.c_double * 3), ("force", C.c_double * 3),
                ("q", C.c_double * 4), ("L", C.c_double * 3), ("torque", C.c_double * 3),
                ("I_body", C.c_double * 3), ("omega_body", C.c_double * 3),
                ("pos_body", C.POINTER(C.c_double))]


def _e3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def rattle_body_pos(pos, vel, pos_rattle, species, mass, links, distances, dt, edges, precision):
    """md.f90:321-408 on (n, 3) arrays in place; links (n_links, 2) zero-based int32;
    returns the sweep counter at exit"""
    return lib().orc_rattle_body_pos(I64(pos.shape[0]), P(pos), P(vel), P(pos_rattle), P(species), P(mass),
                                     P(links), P(distances), int(distances.shape[0]), D(dt), _e3(edges),
                                     D(precision))


def rattle_body_vel(pos, vel, species, mass, links, distances, dt, edges, precision):
    """md.f90:410-480 on (n, 3) arrays in place; returns the sweep counter at exit"""
    return lib().orc_rattle_body_vel(I64(pos.shape[0]), P(pos), P(vel), P(species), P(mass), P(links),
                                     P(distances), int(distances.shape[0]), D(dt), _e3(edges), D(precision))


def elastic_network(pos, force, links, distances, k, edges):
    """md.f90:516-544: adds to force in place, returns the energy"""
    return lib().orc_elastic_network(P(pos), P(force), P(links), P(distances), int(distances.shape[0]), D(k),
                                     _e3(edges))


def rigid_body_init(pos, vel, image, species, mass, edges, i_start, i_stop):
    """md.f90:547-588; returns a Body"""
    b = Body()
    b.body_rows = np.zeros((i_stop - i_start + 1, 3))
    b.pos_body = b.body_rows.ctypes.data_as(C.POINTER(C.c_double))
    lib().orc_rigid_body_init(C.byref(b), P(pos), P(vel), P(image), P(species), P(mass), _e3(edges),
                              int(i_start), int(i_stop))
    return b


def rigid_body_compute_force_torque(b, force):
    lib().orc_rigid_body_compute_force_torque(C.byref(b), P(force))


def rigid_body_vv1(b, pos, dt, threshold, edges):
    """md.f90:613-673; returns the iteration counter at exit"""
    return lib().orc_rigid_body_vv1(C.byref(b), P(pos), D(dt), D(threshold), _e3(edges))


def rigid_body_vv2(b, vel, dt):
    lib().orc_rigid_body_vv2(C.byref(b), P(vel), D(dt))


def colloid_md_pos(pos, vel, force, mass_by_particle, dt):
    lib().orc_colloid_md_pos(I64(pos.shape[0]), P(pos), P(vel), P(force), P(mass_by_particle), D(dt))


def colloid_md_vel(vel, force, force_old, mass_by_particle, dt):
    lib().orc_colloid_md_vel(I64(vel.shape[0]), P(vel), P(force), P(force_old),
                             P(mass_by_particle), D(dt))


def rel_pos(x, y, L):
    xa = (C.c_double * 3)(*[float(v) for v in x])
    ya = (C.c_double * 3)(*[float(v) for v in y])
    la = (C.c_double * 3)(*[float(v) for v in L])
    out = (C.c_double * 3)()
    lib().orc_rel_pos(xa, ya, la, out)
    return np.array(out[:])


# ---- interaction -----------------------------------------------------------------
def lj_params(eps, sigma, cut):
    """eps/sigma/cut: 2-D arrays indexed (solvent species, colloid species), 1-based
    in the reference; stored column-major as the reference does."""
    eps = np.asarray(eps, np.float64)
    n1, n2 = eps.shape
    lj = LJ()
    e = np.ascontiguousarray(eps.T.reshape(-1))
    s = np.ascontiguousarray(np.asarray(sigma, np.float64).T.reshape(-1))
    ct = np.ascontiguousarray(np.asarray(cut, np.float64).T.reshape(-1))
    rc = lib().orc_lj_params_init(C.byref(lj), n1, n2, P(e), P(s), P(ct))
    if rc != 0:
        raise ValueError("unequal sizes for arguments to lj_params_init")
    return lj


def lj_force(d, r_sq, eps, sigma):
    da = (C.c_double * 3)(*[float(v) for v in d])
    out = (C.c_double * 3)()
    lib().orc_lj_force(da, D(r_sq), D(eps), D(sigma), out)
    return np.array(out[:])


def lj_energy(r_sq, eps, sigma, cut_energy=0.0):
    return lib().orc_lj_energy(r_sq, eps, sigma, cut_energy)


# ---- neighbour list --------------------------------------------------------------
def make_stencil(a, cut):
    n = lib().orc_make_stencil(D(a), D(cut), None)
    st = np.zeros((n, 3), np.int32)
    lib().orc_make_stencil(D(a), D(cut), P(st))
    return st


def update_list(c, cnt, st, pos1, species1, pos2, species2, radius, stencil, lj, nmax):
    n1 = pos1.shape[0]
    lst = np.zeros((n1, nmax), np.int32)
    nl = np.zeros(n1, np.int32)
    is_md = np.zeros(c.N, np.uint8)
    is_reac = np.zeros(c.N, np.uint8)
    rc = lib().orc_update_list(C.byref(c), P(cnt), P(st), I64(n1), P(pos1), P(species1), P(pos2),
                               P(species2), D(radius), P(stencil), int(stencil.shape[0]),
                               None if lj is None else C.byref(lj), int(nmax), P(lst), P(nl),
                               P(is_md), P(is_reac))
    if rc != 0:
        raise RuntimeError("maximum of neighbor list reached")
    return lst, nl, is_md, is_reac


def compute_force(pos1, species1, force1, pos2, species2, force2, flags2, lst, nl, L, lj):
    la = (C.c_double * 3)(*[float(v) for v in L])
    return lib().orc_compute_force(I64(pos1.shape[0]), P(pos1), P(species1), P(force1),
                                   I64(pos2.shape[0]), P(pos2), P(species2), P(force2), P(flags2),
                                   P(lst), P(nl), int(lst.shape[1]), la, C.byref(lj))


def compute_force_n2(pos, species, force, L, lj):
    la = (C.c_double * 3)(*[float(v) for v in L])
    return lib().orc_compute_force_n2(I64(pos.shape[0]), P(pos), P(species), P(force), la,
                                      C.byref(lj))


# ---- driver helpers ----------------------------------------------------------------
def flag_particles_nl(list1, n1, cpos, pos, species, flags, edges, cut_sq_c):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_flag_particles_nl(P(list1), int(n1), P(cpos), P(pos), P(species), P(flags), ea,
                                D(cut_sq_c))


def change_species(cpos, pos, species, flags, edges, cut_sq_c, cut_sq_n, prob, draws):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_change_species(I64(pos.shape[0]), P(cpos), P(pos), P(species), P(flags), ea,
                             D(cut_sq_c), D(cut_sq_n), D(prob), P(draws))


def change_species_seeded(cpos, pos, species, flags, edges, cut_sq_c, cut_sq_n, prob, seed, step, ids):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_change_species_seeded(I64(pos.shape[0]), P(cpos), P(pos), P(species), P(flags), ea,
                                    D(cut_sq_c), D(cut_sq_n), D(prob), C.c_uint64(seed), I64(step), P(ids))


def refuel(cpos, pos, species, edges, L1):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_refuel(I64(pos.shape[0]), P(cpos), P(pos), P(species), ea, int(L1))


def flag_particles_nl_prob(list1, n1, cpos, pos, species, flags, edges, cut_sq_c, prob, draws):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_flag_particles_nl_prob(P(list1), int(n1), P(cpos), P(pos), P(species), P(flags), ea,
                                     D(cut_sq_c), D(prob), P(draws))


def change_species_sphere(cpos, pos, species, flags, edges, cut_sq_c):
    ea = (C.c_double * 3)(*[float(v) for v in edges])
    lib().orc_change_species_sphere(I64(pos.shape[0]), P(cpos), P(pos), P(species), P(flags), ea,
                                    D(cut_sq_c))
