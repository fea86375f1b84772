"""The C ABI library without a GPU: it loads, exports every symbol include/rmpcd_capi.h
declares, and its host-only entry points agree with the golden vectors and the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from rmpcdmd_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "rmpcd_capi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rmpcd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_box_init_matches_cell_system_init():
    for L, walls in [((8, 5, 5), 0), ((120, 30, 30), 0), ((24, 40, 48), 0), ((8, 5, 8), 1), ((1, 1, 1), 0)]:
        b = capi.Box()
        capi.call("rmpcd_box_init", C.byref(b), capi.i3(L), 1.0, walls)
        c = O.cells_init(L, 1.0, walls)
        assert tuple(b.L) == tuple(c.L) and tuple(b.M) == tuple(c.M) and b.ncells == c.N
        assert tuple(b.edges) == tuple(c.edges)
    b = capi.Box()
    with pytest.raises(capi.RmpcdError):
        capi.call("rmpcd_box_init", C.byref(b), capi.i3((0, 5, 5)), 1.0, 0)
    capi.call("rmpcd_random_shift", C.byref(b), capi.d3((0.25, 0.5, 0.999)))
    assert tuple(b.origin) == (-0.75, -0.5, 0.999 - 1)


def test_host_hilbert_matches_golden():
    z = np.load(os.path.join(HERE, "golden", "hilbert_ref.npz"))
    for name in ("keys_3_3_3", "keys_2_5_3", "keys_4_2_3"):
        M = tuple(int(v) for v in name[5:].split("_"))
        keys, inv = z[name], z["inv_" + name[5:]]
        idx = 0
        Ma = capi.i3(M)
        for k in range(2 ** M[2]):
            for j in range(2 ** M[1]):
                for i in range(2 ** M[0]):
                    assert capi.lib().rmpcd_compact_p_to_h(capi.i3((i, j, k)), Ma) == keys[idx]
                    idx += 1
        out = (C.c_int32 * 3)()
        for h in range(keys.size):
            capi.lib().rmpcd_compact_h_to_p(h, Ma, out)
            assert tuple(out) == tuple(inv[h])


def test_lj_params_init_matches_golden():
    z = np.load(os.path.join(HERE, "golden", "interaction_ref.npz"))
    dp = C.POINTER(C.c_double)
    for row, o in list(zip(z["inputs"], z["outputs"]))[:100]:
        lj = capi.LJ()
        e, s, c = (np.array([v]) for v in (row[4], row[5], np.sqrt(row[3])))
        capi.call("rmpcd_lj_params_init", C.byref(lj), 1, 1, e.ctypes.data_as(dp), s.ctypes.data_as(dp),
                  c.ctypes.data_as(dp))
        assert np.float64(lj.cut_energy[0]).view(np.uint64) == o[4].view(np.uint64)
    lj = capi.LJ()
    big = np.ones(20)
    with pytest.raises(capi.RmpcdError, match="unequal sizes for arguments to lj_params_init"):
        capi.call("rmpcd_lj_params_init", C.byref(lj), 5, 4, big.ctypes.data_as(dp), big.ctypes.data_as(dp),
                  big.ctypes.data_as(dp))


def test_make_stencil_matches_oracle():
    for cut in (1.0, 3.6, 4.245, 5.99, 7.5):
        n = capi.lib().rmpcd_make_stencil(1.0, cut, None, 0)
        buf = (C.c_int32 * (3 * n))()
        assert capi.lib().rmpcd_make_stencil(1.0, cut, buf, n) == n
        got = np.frombuffer(buf, np.int32).reshape(n, 3)
        assert (got == O.make_stencil(1.0, cut)).all()


def test_threefry_known_answer():
    # Threefry-2x64-20 with zero counter and key: the first known-answer vector of the
    # Random123 distribution (kat_vectors)
    out = (C.c_uint64 * 3)()
    capi.lib().rmpcd_threefry2x64_20((C.c_uint64 * 2)(0, 0), (C.c_uint64 * 2)(0, 0), out)
    assert (out[0], out[1]) == (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)


def test_sort_refuses_aliased_buffers():
    # rmpcd_sort is out of place: a destination that shares any array with the source is
    # refused on the host, before anything reaches the device (no GPU needed)
    b = capi.Box()
    capi.call("rmpcd_box_init", C.byref(b), capi.i3((4, 4, 4)), 1.0, 0)
    names = ("pos", "vel", "force", "force_old", "image", "id", "species", "flags")

    def fake(base):
        p = capi.Particles()
        p.n, p.form = 640, 0
        for i, k in enumerate(names):
            setattr(p, k, base + 0x10000 * i)
        return p

    cells = capi.Cells(0x900000, 0xA00000, 0xB00000, 0xC00000)
    i2i = C.c_void_p(0xD00000)
    src = fake(0x100000)
    for k in names:
        dst = fake(0x500000)
        setattr(dst, k, getattr(src, k))
        with pytest.raises(capi.RmpcdError) as e:
            capi.call("rmpcd_sort", C.byref(b), C.byref(src), C.byref(dst), i2i, C.byref(cells), 0)
        assert e.value.code == capi.RMPCD_ERR_ARG and "share an array" in e.value.msg
