"""shared fixtures for the parity tests: seeded synthetic states in the oracle layout
and their device copies"""
import numpy as np

from oracle import oracle as O


def solvent_state(L, rho, seed, n_species=1, frac2=0.0):
    """uniform positions in [0,L), normal velocities, zero forces, species 1 (and 2)"""
    rng = np.random.default_rng(seed)
    L = np.asarray(L, np.float64)
    n = int(rho * np.prod(L))
    ps = O.PState(n)
    ps.pos[:] = rng.random((n, 3)) * L
    ps.vel[:] = rng.normal(size=(n, 3))
    ps.force[:] = rng.normal(size=(n, 3)) * 0.1
    ps.force_old[:] = rng.normal(size=(n, 3)) * 0.1
    ps.image[:] = rng.integers(-3, 4, size=(n, 3))
    ps.species[:] = 1
    if frac2 > 0:
        ps.species[rng.random(n) < frac2] = 2
    ps.flags[:] = rng.integers(0, 128, size=n)
    ps.id[:] = rng.permutation(n).astype(np.int32)
    return ps


def sorted_state(L, rho, seed, origin=(-0.5, -0.5, -0.5), frac2=0.0):
    """a solvent state already cell-sorted by the oracle, with its cells"""
    c = O.cells_init(L)
    O.set_origin(c, origin)
    ps = solvent_state(L, rho, seed, frac2=frac2)
    ps.id[:] = np.arange(ps.n, dtype=np.int32)
    srt, i2i, cnt, st = O.sort(c, ps)
    return c, srt, cnt, st


def to_device(ps, form=0, device="cuda"):
    from rmpcdmd_amd.system import ParticleSet
    d = ParticleSet(ps.n, form=form, device=device)
    d.load(pos=ps.pos, vel=ps.vel, force=ps.force, force_old=ps.force_old, image=ps.image,
           id=ps.id, species=ps.species, flags=ps.flags)
    return d


def cells_to_device(c, cnt=None, st=None, device="cuda"):
    import torch
    from rmpcdmd_amd.system import CellSet
    cs = CellSet(tuple(c.L[k] - (1 if (k == 2 and c.has_walls) else 0) for k in range(3)),
                 c.a, bool(c.has_walls), device=device)
    cs.set_origin(c.origin)
    for k in range(3):
        cs.box.bc[k] = c.bc[k]
    cs.box.max_v = c.max_v
    if cnt is not None:
        cs.cell_count.copy_(torch.as_tensor(cnt))
        cs.cell_start.copy_(torch.as_tensor(st))
    return cs


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()
