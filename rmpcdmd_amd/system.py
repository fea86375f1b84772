"""Device-resident particle and cell sets for the C ABI (torch provides the HBM
allocations; the arithmetic is in librmpcd_amd.so).

Names follow the reference types: ``ParticleSet`` holds the arrays of
particle_system_t (src/particle_system.f90:22-77) in the library layout (three planes
per vector, 0-based ids), ``CellSet`` those of cell_system_t
(src/cell_system.f90:20-41), ``NeighborList`` those of neighbor_list_t
(src/neighbor_list.f90:23-34).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import FP32, FP64


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class ParticleSet:
    """one buffer of a particle_system_t (the reference double-buffers by pointer swap;
    callers keep two ParticleSets and swap them after a sort)"""

    def __init__(self, n, form=FP64, device="cuda"):
        self.n = int(n)
        self.form = form
        real = torch.float64 if form == FP64 else torch.float32
        small = torch.int32 if form == FP64 else torch.uint8
        self.pos = torch.zeros(3, n, dtype=real, device=device)
        self.vel = torch.zeros(3, n, dtype=real, device=device)
        self.force = torch.zeros(3, n, dtype=real, device=device)
        self.force_old = torch.zeros(3, n, dtype=real, device=device)
        self.image = torch.zeros(3, n, dtype=torch.int32, device=device)
        self.id = torch.arange(n, dtype=torch.int32, device=device)
        self.species = torch.ones(n, dtype=small, device=device)
        self.flags = torch.zeros(n, dtype=small, device=device)
        self._s = None

    def struct(self):
        self._s = capi.Particles(self.n, self.form, 0, _ptr(self.pos), _ptr(self.vel),
                                 _ptr(self.force), _ptr(self.force_old), _ptr(self.image),
                                 _ptr(self.id), _ptr(self.species), _ptr(self.flags))
        return C.byref(self._s)

    # conversions to/from the reference (N,3) = Fortran (3,N) layout
    def load(self, pos=None, vel=None, force=None, force_old=None, image=None, id=None,
             species=None, flags=None):
        for name, arr in (("pos", pos), ("vel", vel), ("force", force),
                          ("force_old", force_old), ("image", image)):
            if arr is not None:
                t = getattr(self, name)
                t.copy_(torch.as_tensor(np.ascontiguousarray(np.asarray(arr).T)).to(t.dtype))
        for name, arr in (("id", id), ("species", species), ("flags", flags)):
            if arr is not None:
                t = getattr(self, name)
                t.copy_(torch.as_tensor(np.asarray(arr)).to(t.dtype))
        return self

    def host(self, name):
        t = getattr(self, name).cpu().numpy()
        if t.ndim == 2:
            return np.ascontiguousarray(t.T)
        return t


class CellSet:
    def __init__(self, L, a=1.0, has_walls=False, device="cuda"):
        self.box = capi.Box()
        capi.call("rmpcd_box_init", C.byref(self.box), capi.i3(L), float(a), int(bool(has_walls)))
        nc = self.box.ncells
        self.cell_count = torch.zeros(nc, dtype=torch.int32, device=device)
        self.cell_start = torch.zeros(nc, dtype=torch.int32, device=device)
        self.is_md = torch.zeros(nc, dtype=torch.uint8, device=device)
        self.is_reac = torch.ones(nc, dtype=torch.uint8, device=device)
        self._c = capi.Cells(_ptr(self.cell_count), _ptr(self.cell_start), _ptr(self.is_md),
                             _ptr(self.is_reac))

    @property
    def ncells(self):
        return self.box.ncells

    def boxp(self):
        return C.byref(self.box)

    def cellsp(self):
        return C.byref(self._c)

    def set_origin(self, origin):
        for k in range(3):
            self.box.origin[k] = float(origin[k])


class NeighborList:
    def __init__(self, npoints, nmax, device="cuda"):
        self.npoints = int(npoints)
        self.nmax = int(nmax)
        self.list = torch.zeros(npoints, nmax, dtype=torch.int32, device=device)
        self.n = torch.zeros(npoints, dtype=torch.int32, device=device)
        self.stencil = None
        self.n_stencil = 0

    def make_stencil(self, a, cut, device="cuda"):
        n = capi.lib().rmpcd_make_stencil(float(a), float(cut), None, 0)
        buf = (C.c_int32 * (3 * n))()
        capi.lib().rmpcd_make_stencil(float(a), float(cut), buf, n)
        st = np.frombuffer(buf, dtype=np.int32).reshape(n, 3).copy()
        self.stencil = torch.as_tensor(st.reshape(-1)).to(device)
        self.n_stencil = n
        return st


def lj_params(eps, sigma, cut):
    """lj_params_t from 2-D (solvent species, colloid species) arrays"""
    eps = np.asarray(eps, np.float64)
    n1, n2 = eps.shape
    lj = capi.LJ()
    e = np.ascontiguousarray(eps.T.reshape(-1))
    s = np.ascontiguousarray(np.asarray(sigma, np.float64).T.reshape(-1))
    c = np.ascontiguousarray(np.asarray(cut, np.float64).T.reshape(-1))
    dp = C.POINTER(C.c_double)
    capi.call("rmpcd_lj_params_init", C.byref(lj), n1, n2, e.ctypes.data_as(dp),
              s.ctypes.data_as(dp), c.ctypes.data_as(dp))
    return lj
