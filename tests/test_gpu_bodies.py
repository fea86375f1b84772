"""Linked colloids on the GPU through the C ABI against the oracle: rattle_body_pos and
rattle_body_vel (md.f90:321-480), elastic_network (md.f90:516-544) and the rigid body
(md.f90:547-700).  The kernels walk the links and particles in the reference's order
with the oracle's arithmetic, so positions, velocities, forces, the body's state and the
iteration counts are compared bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rmpcdmd_amd import capi

from _helpers import same_bits, to_device

pytestmark = pytest.mark.gpu

EDGES = np.array([16.0, 12.0, 20.0])


def ptr(t):
    return C.c_void_p(t.data_ptr())


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def host(a):
    """host arrays for the link list (the library schedules it on the device)"""
    return a.ctypes.data_as(C.c_void_p)


def sphere_body(n, radius, centre, seed, f=1.8):
    """n beads on a Fibonacci sphere, each linked to the beads closer than f times the
    mean spacing, listed as the reference's drivers build them (i < j, by i then j)"""
    k = np.arange(n) + 0.5
    th = np.arccos(1 - 2 * k / n)
    ph = np.pi * (1 + 5 ** 0.5) * k
    r = radius * np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], 1)
    i, j = np.triu_indices(n, 1)
    near = np.linalg.norm(r[i] - r[j], axis=1) < f * np.sqrt(4 * np.pi * radius ** 2 / n)
    links = np.ascontiguousarray(np.stack([i[near], j[near]], 1).astype(np.int32))
    d = np.linalg.norm(r[links[:, 0]] - r[links[:, 1]], axis=1)
    rng = np.random.default_rng(seed)
    ps = O.PState(n)
    ps.pos[:] = r + np.asarray(centre)
    ps.vel[:] = rng.normal(size=(n, 3))
    ps.force[:] = rng.normal(size=(n, 3))
    ps.species[:] = rng.integers(1, 3, size=n)
    return ps, links, d, rng


def fold(x, edges=EDGES):
    return x - np.floor(x / edges) * edges


# up to 900 beads the body and its schedule fit in LDS (12 sweeps overlapped for 300
# beads, 1 or 2 for 900); 1400 beads (5800 links) run from global memory
@pytest.mark.parametrize("n,radius,centre,edges", [
    (12, 2.0, (8.0, 6.0, 10.0), EDGES),
    (60, 2.0, (15.5, 0.3, 19.6), EDGES),
    (300, 4.5, (3.0, 6.0, 10.0), EDGES * 2),
    (900, 8.5, (39.0, 20.0, 1.0), np.array([40.0, 40.0, 40.0])),
    (1400, 10.5, (1.0, 20.0, 39.5), np.array([44.0, 44.0, 44.0])),
])
def test_rattle_body_pos_vel(gpu, n, radius, centre, edges):
    ps, links, d, rng = sphere_body(n, radius, centre, 11 + n)
    mass = np.array([1.0, 2.5])
    rattle = fold(ps.pos, edges)
    ps.pos[:] = fold(ps.pos + 0.01 * rng.normal(size=ps.pos.shape), edges)
    want, want_r = ps.copy(), rattle.copy()
    it_p = O.rattle_body_pos(want.pos, want.vel, want_r, want.species, mass, links, d, 0.01, edges, 1e-11)
    g = to_device(ps)
    rt, lk, dd = dev(rattle.T), host(links), host(d)
    it = C.c_int32(0)
    m = (C.c_double * 2)(*mass)
    print(f"rattle_body_pos n={n} links={len(d)}", flush=True)
    capi.call("rmpcd_rattle_body_pos", g.struct(), ptr(rt), lk, dd, len(d), m, 2, 0.01,
              capi.d3(edges), 1e-11, C.byref(it))
    assert it.value == it_p and 1 < it_p
    assert same_bits(g.host("pos"), want.pos)
    assert same_bits(g.host("vel"), want.vel)
    assert same_bits(rt.cpu().numpy().T, want_r)
    it_v = O.rattle_body_vel(want.pos, want.vel, want.species, mass, links, d, 0.01, edges, 1e-12)
    capi.call("rmpcd_rattle_body_vel", g.struct(), lk, dd, len(d), m, 2, 0.01, capi.d3(edges), 1e-12,
              C.byref(it))
    torch.cuda.synchronize()
    assert it.value == it_v
    assert same_bits(g.host("pos"), want.pos)
    assert same_bits(g.host("vel"), want.vel)


def test_rattle_body_without_links(gpu):
    """no links: the first sweep passes, and the positions still go through the
    reference's unfold next to particle 1 and fold back (md.f90:343-353, 400-402)"""
    ps, _, _, rng = sphere_body(12, 2.0, (15.9, 0.1, 10.0), 7)
    ps.pos[:] = fold(ps.pos)
    links = np.zeros((0, 2), np.int32)
    d = np.zeros(0)
    mass = np.array([1.0, 2.0])
    want = ps.copy()
    assert O.rattle_body_vel(want.pos, want.vel, want.species, mass, links, d, 0.01, EDGES, 1e-12) == 1
    g = to_device(ps)
    it = C.c_int32(0)
    m = (C.c_double * 2)(*mass)
    capi.call("rmpcd_rattle_body_vel", g.struct(), None, None, 0, m, 2, 0.01, capi.d3(EDGES), 1e-12, C.byref(it))
    assert it.value == 1
    assert same_bits(g.host("pos"), want.pos)
    assert same_bits(g.host("vel"), want.vel)
    e = C.c_double(1.0)
    capi.call("rmpcd_elastic_network", g.struct(), None, None, 0, 3.0, capi.d3(EDGES), C.byref(e))
    assert e.value == 0.0


def test_body_link_out_of_range(gpu):
    ps, links, d, _ = sphere_body(12, 2.0, (8.0, 6.0, 10.0), 5)
    links[3, 1] = 12
    g = to_device(ps)
    m = (C.c_double * 2)(1.0, 1.0)
    lk, dd = host(links), host(d)
    e = C.c_double(0)
    with pytest.raises(capi.RmpcdError, match="outside"):
        capi.call("rmpcd_rattle_body_vel", g.struct(), lk, dd, len(d), m, 2, 0.01, capi.d3(EDGES), 1e-12, None)
    with pytest.raises(capi.RmpcdError, match="outside"):
        capi.call("rmpcd_elastic_network", g.struct(), lk, dd, len(d), 3.0, capi.d3(EDGES), C.byref(e))
    assert same_bits(g.host("vel"), ps.vel)
    assert same_bits(g.host("force"), ps.force)


@pytest.mark.parametrize("centre", [(8.0, 6.0, 10.0), (0.2, 11.9, 10.0)])
def test_elastic_network(gpu, centre):
    ps, links, d, rng = sphere_body(40, 2.0, centre, 21)
    ps.pos[:] = fold(ps.pos + 0.05 * rng.normal(size=ps.pos.shape))
    g = to_device(ps)
    e = C.c_double(0)
    want = ps.force.copy()
    for rest in (d, d * 1.1):  # the second list differs by its distances only
        e_want = O.elastic_network(ps.pos, want, links, rest, 30.0, EDGES)
        capi.call("rmpcd_elastic_network", g.struct(), host(links), host(rest), len(d), 30.0, capi.d3(EDGES),
                  C.byref(e))
        assert e.value == e_want and e_want > 0
        assert same_bits(g.host("force"), want)


def test_elastic_network_known_answer(gpu):
    """test/test_elastic_network_0.f90: three links from particle 1 (one across the box),
    energy 8.75 and force (k, 0, 0) on particle 1"""
    ps = O.PState(4)
    ps.pos[:] = 1
    ps.pos[1, 0], ps.pos[2, 1], ps.pos[3, 0] = 3, 2, 4.5
    ps.force[:] = 0
    links = np.array([[0, 1], [0, 2], [0, 3]], np.int32)
    d = np.array([0.5, 1.0, 1.0])
    g = to_device(ps)
    e = C.c_double(0)
    capi.call("rmpcd_elastic_network", g.struct(), host(links), host(d), 3, 7.0, capi.d3([5.0] * 3), C.byref(e))
    assert abs(e.value - 8.75) < 1e-14
    f = g.host("force")
    np.testing.assert_allclose(f[0], [7.0, 0, 0], atol=1e-14)
    assert np.max(np.abs(f.sum(0))) < 1e-14


def body_fields(b):
    return {k: np.array(getattr(b, k)[:]) for k in ("pos", "vel", "force", "q", "L", "torque", "I_body",
                                                     "omega_body")} | {"mass": b.mass}


def test_rigid_body(gpu):
    ps, _, _, rng = sphere_body(30, 2.0, (15.0, 11.5, 1.0), 31)
    ps.image[:] = 0
    ps.image[ps.pos[:, 0] >= EDGES[0], 0] = 1
    ps.image[ps.pos[:, 1] >= EDGES[1], 1] = 1
    ps.pos[:] = fold(ps.pos)
    mass = np.array([1.0, 3.0])
    # two bodies: particles 0..9 and 10..29
    ranges = [(0, 9), (10, 29)]
    want_b = [O.rigid_body_init(ps.pos, ps.vel, ps.image, ps.species, mass, EDGES, a, z) for a, z in ranges]
    g = to_device(ps)
    m = (C.c_double * 2)(*mass)
    size = C.sizeof(capi.RigidBody)
    bodies = [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in ranges]
    rows = [torch.zeros(3 * (z - a + 1), dtype=torch.float64, device="cuda") for a, z in ranges]
    for b, r, (a, z) in zip(bodies, rows, ranges):
        capi.call("rmpcd_rigid_body_init", ptr(b), g.struct(), m, 2, capi.d3(EDGES), a, z, ptr(r))

    def check():
        torch.cuda.synchronize()
        for b, r, w in zip(bodies, rows, want_b):
            got = capi.RigidBody.from_buffer_copy(bytes(b.cpu().numpy()))
            gf, wf = body_fields(got), body_fields(w)
            for k in wf:
                assert same_bits(np.asarray(gf[k]), np.asarray(wf[k])), k
            assert same_bits(r.cpu().numpy().reshape(-1, 3), w.body_rows)
        assert same_bits(g.host("pos"), ps.pos)
        assert same_bits(g.host("vel"), ps.vel)

    check()
    for b, w in zip(bodies, want_b):  # spin both bodies up
        L = rng.normal(size=3) * 5
        w.L[:] = L
        hb = capi.RigidBody.from_buffer_copy(bytes(b.cpu().numpy()))
        hb.L[:] = L
        b.copy_(torch.frombuffer(bytearray(bytes(hb)), dtype=torch.uint8).cuda())
    it = C.c_int32(0)
    for step in range(8):
        ps.force[:] = rng.normal(size=ps.force.shape)
        g.load(force=ps.force)
        for b, w in zip(bodies, want_b):
            O.rigid_body_compute_force_torque(w, ps.force)
            capi.call("rmpcd_rigid_body_compute_force_torque", ptr(b), g.struct())
            n_want = O.rigid_body_vv1(w, ps.pos, 0.05, 1e-13, EDGES)
            capi.call("rmpcd_rigid_body_vv1", ptr(b), g.struct(), 0.05, 1e-13, capi.d3(EDGES),
                      C.byref(it) if step % 2 == 0 else None)
            if step % 2 == 0:
                assert it.value == n_want
            O.rigid_body_vv2(w, ps.vel, 0.05)
            capi.call("rmpcd_rigid_body_vv2", ptr(b), g.struct(), 0.05)
        check()
