"""The driver loops other than the dimer's (test/setup_simple_fluid.f90,
test/setup_single_catalytic_fixed_sphere.f90, programs/n_colloids_pbc.f90) as
device-resident programs behind the C ABI (rmpcd_program_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import FP32, FP64, PROG_COLLOIDS, PROG_FLUID, PROG_SPHERE

DEFAULTS = {
    # setup_simple_fluid: plain SRD fluid, L=32, rho=10
    PROG_FLUID: dict(L=(32, 32, 32), rho=10, n_md=1, T=1.0 / 3.0, tau=0.1, sigma=1.0, epsilon=1.0,
                     epsilon_colloids=1.0, n_colloids=0, prob=1.0, bulk_rate=0.0, skin=0.0),
    # setup_single_catalytic_fixed_sphere: one catalytic sphere sigma=4 at the box centre,
    # bulk B->A at 0.01
    PROG_SPHERE: dict(L=(32, 32, 32), rho=10, n_md=10, T=1.0 / 3.0, tau=0.1, sigma=4.0, epsilon=1.0,
                      epsilon_colloids=1.0, n_colloids=1, prob=1.0, bulk_rate=0.01, skin=1.5),
    # n_colloids_pbc: 1000 colloids sigma=2.5 in a periodic L=128 box
    PROG_COLLOIDS: dict(L=(128, 128, 128), rho=10, n_md=50, T=1.0, tau=0.5, sigma=2.5, epsilon=1.0,
                        epsilon_colloids=1.0, n_colloids=1000, prob=1.0, bulk_rate=0.0, skin=0.8),
}


def make_config(kind, form=FP64, seed=1, nmax_list=0, n_thermo=0, T_final=None, n_loop=0, thermostat=0, hydro=1,
                graph=0, **kw):
    p = dict(DEFAULTS[kind])
    unknown = set(kw) - set(p)
    if unknown:
        raise TypeError("unknown program parameters: %s" % sorted(unknown))
    p.update(kw)
    c = capi.ProgramConfig()
    c.kind = int(kind)
    c.L = (C.c_int32 * 3)(*p["L"])
    c.rho, c.n_md, c.form = int(p["rho"]), int(p["n_md"]), int(form)
    c.n_colloids, c.nmax_list = int(p["n_colloids"]), int(nmax_list)
    for k in ("T", "tau", "sigma", "epsilon", "epsilon_colloids", "prob", "bulk_rate", "skin"):
        setattr(c, k, float(p[k]))
    c.seed = int(seed)
    c.n_thermo, c.n_loop, c.thermostat, c.hydro = int(n_thermo), int(n_loop), int(thermostat), int(hydro)
    c.graph = int(graph)
    c.T_final = float(p["T"] if T_final is None else T_final)
    return c


class Program:
    def __init__(self, kind, **kw):
        self.kind = kind
        self.config = make_config(kind, **kw)
        self._h = C.c_void_p()
        capi.call("rmpcd_program_create", C.byref(self.config), C.byref(self._h))

    def run(self, n_cycles):
        capi.call("rmpcd_program_run", self._h, int(n_cycles))

    def state(self):
        sol, col = capi.Particles(), capi.Particles()
        box, cells, st = capi.Box(), capi.Cells(), capi.ProgramStats()
        capi.call("rmpcd_program_state", self._h, C.byref(sol), C.byref(col), C.byref(box), C.byref(cells),
                  C.byref(st))
        return sol, col, box, cells, st

    def download(self):
        """host copies of the state: vectors as (n, 3) arrays, scalars as (n,) arrays"""
        sol, col, box, cells, st = self.state()
        n, nc = sol.n, col.n
        real = np.float64 if sol.form == FP64 else np.float32
        small = np.int32 if sol.form == FP64 else np.uint8

        def get(ptr, count, dtype):
            a = np.empty(count, dtype=dtype)
            if count:
                capi.call("rmpcd_memcpy", a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes, 1)
            return a

        out = {}
        for name in ("pos", "vel", "force", "force_old"):
            out[name] = get(getattr(sol, name), 3 * n, real).reshape(3, n).T.copy()
        out["pos_old"] = get(st.pos_old, 3 * n, real).reshape(3, n).T.copy()
        out["image"] = get(sol.image, 3 * n, np.int32).reshape(3, n).T.copy()
        out["id"] = get(sol.id, n, np.int32)
        out["species"] = get(sol.species, n, small)
        out["flags"] = get(sol.flags, n, small)
        if nc:
            for name in ("pos", "vel", "force", "force_old"):
                out["colloid_" + name] = get(getattr(col, name), 3 * nc, np.float64).reshape(3, nc).T.copy()
            out["colloid_pos_old"] = get(st.colloid_pos_old, 3 * nc, np.float64).reshape(3, nc).T.copy()
            out["colloid_image"] = get(col.image, 3 * nc, np.int32).reshape(3, nc).T.copy()
            out["list"] = get(st.list, nc * st.nmax, np.int32).reshape(nc, st.nmax)
            out["nlist"] = get(st.nlist, nc, np.int32)
        out["cell_count"] = get(cells.cell_count, box.ncells, np.int32)
        out["cell_start"] = get(cells.cell_start, box.ncells, np.int32)
        out["is_md"] = get(cells.is_md, box.ncells, np.uint8)
        out["origin"] = np.array(box.origin[:])
        for k in ("n_solvent", "n_colloids", "md_steps", "cycles", "n_extra_sorting", "e1", "e2", "temperature",
                  "dt", "mass", "nmax", "last_sort"):
            out[k] = getattr(st, k)
        return out

    def download_cells(self):
        """host copies of the cell arrays (count, start, is_md) only"""
        sol, col, box, cells, st = self.state()
        out = {}
        for name, ptr, dt in (("cell_count", cells.cell_count, np.int32), ("cell_start", cells.cell_start, np.int32),
                              ("is_md", cells.is_md, np.uint8)):
            a = np.empty(box.ncells, dtype=dt)
            capi.call("rmpcd_memcpy", a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes, 1)
            out[name] = a
        return out

    def download_colloids(self):
        """host copy of the colloid positions, (n, 3)"""
        sol, col, box, cells, st = self.state()
        a = np.empty(3 * col.n, dtype=np.float64)
        capi.call("rmpcd_memcpy", a.ctypes.data_as(C.c_void_p), C.c_void_p(col.pos), a.nbytes, 1)
        return a.reshape(3, col.n).T.copy()

    def close(self):
        if self._h:
            capi.lib().rmpcd_program_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Program", "make_config", "PROG_FLUID", "PROG_SPHERE", "PROG_COLLOIDS", "FP32", "FP64"]
