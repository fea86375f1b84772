"""Threefry-2x64-20 in numpy and the draw streams of the device program -- TEST
INFRASTRUCTURE ONLY (the whole-step parity checker).

The reference's generator module (random_module, Threefry via C) is an empty submodule in
the reference checkout, so its streams are not reproducible. The device program keys its
own counter-based Threefry-2x64-20 (Salmon et al., "Parallel random numbers: as easy as
1, 2, 3", SC'11; Random123's rotation constants and key-schedule parity) as follows, and
this module restates it so that the oracle's time loop can consume the same draws:

    u01(x) = (x >> 11) * 2^-53
    key = (seed + replica, stream)
    grid shift of collision s    : ctr (s, 0) -> u0 = u01(a), u1 = u01(b); ctr (s, 1) -> u2 = u01(a)
    SRD axis of cell c, cycle s  : Marsaglia pairs k = 0, 1, ... from ctr (s, c + k 2^40)
    MPC-AT normals, id, cycle s  : Box-Muller on ctr (2s, id) and (2s + 1, id), stream THERMO
    change_species, md step t    : ctr (t, id) -> u01(a)
    bulk_reaction, cell c        : ctr (s, c) -> u01(a)
"""
from __future__ import annotations

import numpy as np

MASK = (1 << 64) - 1
ROT = (16, 42, 12, 31, 16, 32, 24, 21)
PARITY = 0x1BD11BDAA9FC1A22

SRD, REACTION, CHANGE, UNIFORM, SHIFT, INIT, PLACE, THERMO = 1, 2, 3, 4, 5, 6, 7, 8


def _rotl(x, r):
    return (x << np.uint64(r)) | (x >> np.uint64(64 - r))


def threefry2x64(c0, c1, k0, k1):
    """vectorised Threefry-2x64 with 20 rounds; arguments broadcast as uint64 arrays"""
    with np.errstate(over="ignore"):
        c0 = np.asarray(c0, np.uint64)
        c1 = np.asarray(c1, np.uint64)
        k0 = np.asarray(k0, np.uint64)
        k1 = np.asarray(k1, np.uint64)
        ks = (k0, k1, k0 ^ k1 ^ np.uint64(PARITY))
        x0 = c0 + ks[0]
        x1 = c1 + ks[1]
        for r in range(20):
            x0 = x0 + x1
            x1 = _rotl(x1, ROT[r % 8])
            x1 = x1 ^ x0
            if r % 4 == 3:
                s = (r + 1) // 4
                x0 = x0 + ks[s % 3]
                x1 = x1 + ks[(s + 1) % 3] + np.uint64(s)
        return x0, x1


def u01(x):
    return (np.asarray(x, np.uint64) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def shift_uniforms(seed, s):
    a, b = threefry2x64(s, 0, seed, SHIFT)
    c, _ = threefry2x64(s, 1, seed, SHIFT)
    return np.array([u01(a), u01(b), u01(c)], np.float64)


def normals3(seed, step, ids):
    """the three normals of each particle id for the MPC-AT draw of step: Box-Muller on
    the uniforms of counters (2 step, id) and (2 step + 1, id), stream THERMO"""
    ids = np.asarray(ids).astype(np.uint64)
    a, b = threefry2x64(2 * step, ids, seed, THERMO)
    c, d = threefry2x64(2 * step + 1, ids, seed, THERMO)
    u1, u2, u3, u4 = 1.0 - u01(a), u01(b), 1.0 - u01(c), u01(d)
    r1 = np.sqrt(-2.0 * np.log(u1))
    r2 = np.sqrt(-2.0 * np.log(u3))
    return np.stack([r1 * np.cos(2 * np.pi * u2), r1 * np.sin(2 * np.pi * u2), r2 * np.cos(2 * np.pi * u4)], axis=1)


def srd_axes(seed, step, ncells):
    """dense (ncells, 3) unit vectors: Marsaglia (1972) from threefry pairs"""
    cells = np.arange(ncells, dtype=np.uint64)
    n1 = np.zeros(ncells)
    n2 = np.zeros(ncells)
    s = np.full(ncells, 2.0)
    todo = np.arange(ncells)
    for k in range(64):
        if todo.size == 0:
            break
        a, b = threefry2x64(step, cells[todo] + np.uint64(k << 40), seed, SRD)
        m1 = 2 * u01(a) - 1
        m2 = 2 * u01(b) - 1
        ss = m1 * m1 + m2 * m2
        n1[todo], n2[todo], s[todo] = m1, m2, ss
        todo = todo[ss >= 1.0]
    alpha = 2.0 * np.sqrt(1.0 - s)
    return np.stack([n1 * alpha, n2 * alpha, 1.0 - 2.0 * s], axis=1)


class ThreefryDraws:
    """the device program's draws, for the oracle's time loop (oracle.dimer.DimerOracle)"""

    per_id = True  # change(md_step, ids) draws from (md_step, id) alone

    def __init__(self, seed):
        self.seed = int(seed)

    def shift(self, cycle):
        return shift_uniforms(self.seed, cycle)

    def srd_axes(self, cycle, c, cnt, st, species):
        return srd_axes(self.seed, cycle, int(c.N))

    def change(self, md_step, ids):
        a, _ = threefry2x64(md_step, np.asarray(ids).astype(np.uint64), self.seed, CHANGE)
        return u01(a)

    def bulk(self, cycle, ncells):
        a, _ = threefry2x64(cycle, np.arange(ncells, dtype=np.uint64), self.seed, REACTION)
        return u01(a)
