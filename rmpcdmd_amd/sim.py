"""The single-dimer nanomotor program (programs/single_dimer_pbc.f90) as a
device-resident simulation behind the C ABI (rmpcd_dimer_*)."""
from __future__ import annotations

import ctypes as C

from . import capi
from .capi import FP32, FP64, SIM_FUSED, SIM_REFERENCE

# the dimer of the nanomotor experiments: sigma = 2, d = 4.5, tau = 0.5, N_MD = 50
DIMER_DEFAULTS = dict(L=(48, 48, 48), rho=10, n_md=50, T=1.0, tau=0.5, sigma_C=2.0, sigma_N=2.0,
                      d=4.5, epsilon_C=(1.0, 1.0), epsilon_N=(1.0, 1.0), epsilon_CC=1.0,
                      epsilon_NC=1.0, epsilon_NN=1.0, prob=1.0, bulk_rmpcd=0, bulk_rate=0.01,
                      skin=2.0, nmax_list=0, seed=20261202, form=FP32, mode=SIM_FUSED, n_replicas=1,
                      replica_L=None, stream_blocks=0, sort_blocks=0, lazy=0, extra_sort_blocks=0,
                      sort_order=0, groups=0, collide_cells=0)


def mixed_L_sizes(n=128, seed=20261202):
    """box sides of the batched single-dimer workload: n replicas with L drawn uniformly
    from {24, 32, 40, 48} by a seeded generator"""
    import numpy as np
    return [int(v) for v in np.array([24, 32, 40, 48])[np.random.default_rng(seed).integers(0, 4, n)]]


def make_config(**kw):
    p = dict(DIMER_DEFAULTS)
    unknown = set(kw) - set(p)
    if unknown:
        raise TypeError("unknown dimer parameters: %s" % sorted(unknown))
    p.update(kw)
    c = capi.DimerConfig()
    c.L = (C.c_int32 * 3)(*p["L"])
    for k in ("rho", "n_md", "form", "mode", "bulk_rmpcd", "nmax_list", "n_replicas", "stream_blocks",
              "sort_blocks", "lazy", "extra_sort_blocks", "sort_order", "groups", "collide_cells"):
        setattr(c, k, int(p[k]))
    for k in ("T", "tau", "sigma_C", "sigma_N", "d", "epsilon_CC", "epsilon_NC", "epsilon_NN", "prob",
              "bulk_rate", "skin"):
        setattr(c, k, float(p[k]))
    c.epsilon_C = (C.c_double * 2)(*p["epsilon_C"])
    c.epsilon_N = (C.c_double * 2)(*p["epsilon_N"])
    c.seed = int(p["seed"])
    if p["replica_L"] is not None:
        rl = [int(v) for row in p["replica_L"] for v in row]
        c.n_replicas = len(rl) // 3
        c._replica_L = (C.c_int32 * len(rl))(*rl)  # kept alive with the config
        c.replica_L = C.cast(c._replica_L, C.c_void_p)
    return c


class SingleDimer:
    def __init__(self, **kw):
        self.config = make_config(**kw)
        self._h = C.c_void_p()
        capi.call("rmpcd_dimer_create", C.byref(self.config), C.byref(self._h))

    def run(self, n_cycles):
        capi.call("rmpcd_dimer_run", self._h, int(n_cycles))

    def state(self, replica=0):
        sol, col = capi.Particles(), capi.Particles()
        box, cells, st = capi.Box(), capi.Cells(), capi.DimerStats()
        capi.call("rmpcd_dimer_state", self._h, int(replica), C.byref(sol), C.byref(col), C.byref(box),
                  C.byref(cells), C.byref(st))
        return sol, col, box, cells, st

    @property
    def n_solvent(self):
        return self.state()[4].n_solvent

    def download(self, replica=0):
        """host copies of replica's state: solvent and colloid vectors as (n, 3) arrays
        (the reference's (3, N) layout), scalars as (n,) arrays, plus the counters"""
        import numpy as np
        sol, col, box, cells, st = self.state(replica)
        n = sol.n
        real = np.float64 if sol.form == FP64 else np.float32
        small = np.int32 if sol.form == FP64 else np.uint8

        def get(ptr, count, dtype):
            a = np.empty(count, dtype=dtype)
            capi.call("rmpcd_memcpy", a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes, 1)
            return a

        out = {}
        for name in ("pos", "vel", "force", "force_old"):
            out[name] = get(getattr(sol, name), 3 * n, real).reshape(3, n).T.copy()
        out["image"] = get(sol.image, 3 * n, np.int32).reshape(3, n).T.copy()
        out["id"] = get(sol.id, n, np.int32)
        out["species"] = get(sol.species, n, small)
        out["flags"] = get(sol.flags, n, small)
        for name in ("pos", "vel", "force", "force_old"):
            out["colloid_" + name] = get(getattr(col, name), 6, np.float64).reshape(3, 2).T.copy()
        out["colloid_image"] = get(col.image, 6, np.int32).reshape(3, 2).T.copy()
        out["cell_count"] = get(cells.cell_count, box.ncells, np.int32)
        out["cell_start"] = get(cells.cell_start, box.ncells, np.int32)
        out["pos_old"] = get(st.pos_old, 3 * n, real).reshape(3, n).T.copy()
        out["colloid_pos_old"] = get(st.colloid_pos_old, 6, np.float64).reshape(3, 2).T.copy()
        out["list"] = get(st.list, 2 * st.nmax, np.int32).reshape(2, st.nmax)
        out["nlist"] = get(st.nlist, 2, np.int32)
        out["origin"] = np.array(box.origin[:])
        for k in ("md_steps", "cycles", "n_extra_sorting", "e1", "e2", "n_solvent", "nmax", "dt", "error"):
            out[k] = getattr(st, k)
        out["mass"] = np.array(st.mass[:])
        return out

    def close(self):
        if self._h:
            capi.lib().rmpcd_dimer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
