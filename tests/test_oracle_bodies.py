"""The oracle's linked-body routines pinned by mathematics (md.f90 needs the absent
quaternion and threefry modules, so the reference cannot be built here).

- rattle_body_pos / rattle_body_vel (md.f90:321-480): a single link reproduces the
  closed-form dimer rattle; a network converges to its constraints, keeps the total
  momentum and the centre of mass, moves velocities by exactly the position correction
  over dt, is blind to where the periodic box cuts the body, and leaves a rigid motion
  untouched.
- elastic_network (md.f90:516-544): the closed form of one link, Newton's third law, the
  force as minus the gradient of the energy (central differences), minimum image.
- rigid_body_* (md.f90:547-700): centre of mass and principal moments of symmetric bodies,
  rotations compared with scipy's, torque as r x f, the closed-form quaternion step of a
  rotation about a principal axis, angular momentum from the particle velocities, and
  the conservation of L and of the kinetic energy of a free asymmetric top.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O

EDGES = np.array([12.0, 10.0, 14.0])


def tetra(rng, centre, scale=1.0):
    """4 particles at the corners of a slightly irregular tetrahedron, 6 links"""
    base = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0.9, 0], [0.5, 0.3, 0.8]], float) * scale
    pos = base + centre + 0.01 * rng.normal(size=(4, 3))
    links = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]], np.int32)
    d = np.linalg.norm(pos[links[:, 0]] - pos[links[:, 1]], axis=1)
    return pos, links, d


def moved(pos, rng, amount):
    return pos + amount * rng.normal(size=pos.shape)


def test_rattle_body_one_link_is_the_dimer_rattle():
    rng = np.random.default_rng(1)
    mass = np.array([1.0, 2.0])
    species = np.array([1, 2], np.int32)
    pr = np.array([[3.0, 4.0, 5.0], [4.1, 4.2, 5.3]])
    d = np.array([1.05])
    pos = moved(pr, rng, 0.05)
    vel = rng.normal(size=(2, 3))
    p1, v1, p2, v2 = pos.copy(), vel.copy(), pos.copy(), vel.copy()
    O.rattle_dimer_pos(p1, v1, pr.copy(), species, mass, d[0], 0.01, EDGES)
    n = O.rattle_body_pos(p2, v2, pr.copy(), species, mass, np.array([[0, 1]], np.int32), d, 0.01, EDGES, 1e-12)
    assert n <= 2
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(v2, v1, rtol=0, atol=1e-10)
    assert abs(np.linalg.norm(p2[0] - p2[1]) - d[0]) < 1e-12
    w1, w2 = vel.copy(), vel.copy()
    O.rattle_dimer_vel(p1, w1, species, mass, d[0], 0.01, EDGES)
    O.rattle_body_vel(p1.copy(), w2, species, mass, np.array([[0, 1]], np.int32), d, 0.01, EDGES, 1e-14)
    np.testing.assert_allclose(w2, w1, rtol=0, atol=1e-13)


@pytest.mark.parametrize("centre", [(5.0, 4.0, 6.0), (11.6, 9.7, 0.2)])  # inside / across 3 faces
def test_rattle_body_pos_network(centre):
    rng = np.random.default_rng(2)
    pr, links, d = tetra(rng, np.array(centre))
    mass = np.array([1.0, 3.0])
    species = np.array([1, 2, 1, 2], np.int32)
    m = mass[species - 1]
    dt = 0.02
    pos = moved(pr, rng, 0.03)
    vel = rng.normal(size=(4, 3))
    wrap = lambda x: x - np.floor(x / EDGES) * EDGES  # noqa: E731
    p, v = wrap(pos), vel.copy()
    n = O.rattle_body_pos(p, v, wrap(pr), species, mass, links, d, dt, EDGES, 1e-10)
    assert 1 < n <= 1000
    # unfolded around particle 0: the constraints, momentum, centre of mass, v = x / dt
    unf = p[0] + (p - p[0]) - np.round((p - p[0]) / EDGES) * EDGES
    assert np.max(np.abs(np.linalg.norm(unf[links[:, 0]] - unf[links[:, 1]], axis=1) - d)) < 1e-10
    np.testing.assert_allclose((m[:, None] * v).sum(0), (m[:, None] * vel).sum(0), atol=1e-12)
    dx = unf - (wrap(pos)[0] + (pos - pos[0]))
    np.testing.assert_allclose((m[:, None] * dx).sum(0), 0, atol=1e-12)
    np.testing.assert_allclose(v - vel, dx / dt, atol=1e-9)
    # the same body far from every face gives the same result, shifted
    shift = np.array([5.0, 4.0, 6.0]) - np.array(centre)
    q, w = pos + shift, vel.copy()
    O.rattle_body_pos(q, w, pr + shift, species, mass, links, d, dt, EDGES, 1e-10)
    np.testing.assert_allclose(wrap(q - shift), wrap(p), atol=1e-9)
    np.testing.assert_allclose(w, v, atol=1e-8)


def test_rattle_body_vel_network():
    rng = np.random.default_rng(3)
    pos, links, d = tetra(rng, np.array([4.0, 5.0, 6.0]))
    mass = np.array([1.0, 2.0])
    species = np.array([2, 1, 1, 2], np.int32)
    m = mass[species - 1]
    vel = rng.normal(size=(4, 3))
    v = vel.copy()
    n = O.rattle_body_vel(pos.copy(), v, species, mass, links, d, 0.01, EDGES, 1e-12)
    assert 1 < n <= 1000
    rel = np.einsum("ij,ij->i", v[links[:, 0]] - v[links[:, 1]], pos[links[:, 0]] - pos[links[:, 1]])
    assert np.max(np.abs(rel)) < 1e-10
    np.testing.assert_allclose((m[:, None] * v).sum(0), (m[:, None] * vel).sum(0), atol=1e-12)
    # a rigid motion has no component along the links: nothing to correct
    omega, V = np.array([0.3, -0.2, 0.5]), np.array([1.0, 2.0, -1.0])
    rigid = V + np.cross(omega, pos - pos.mean(0))
    w = rigid.copy()
    assert O.rattle_body_vel(pos.copy(), w, species, mass, links, d, 0.01, EDGES, 1e-12) == 1
    np.testing.assert_allclose(w, rigid, atol=1e-13)


def test_elastic_network():
    rng = np.random.default_rng(4)
    pos, links, d = tetra(rng, np.array([11.7, 0.1, 7.0]))
    pos = pos - np.floor(pos / EDGES) * EDGES  # the body crosses two faces
    d = d * (1 + 0.1 * rng.normal(size=d.shape))
    k = 7.0
    f = np.zeros_like(pos)
    e = O.elastic_network(pos, f, links, d, k, EDGES)

    def energy(x):
        unf = x - np.round((x - x[0]) / EDGES) * EDGES
        r = np.linalg.norm(unf[links[:, 0]] - unf[links[:, 1]], axis=1)
        return 0.5 * k * np.sum((r - d) ** 2)

    assert abs(e - energy(pos)) < 1e-12 * max(1, e)
    np.testing.assert_allclose(f.sum(0), 0, atol=1e-12)
    h = 1e-6
    grad = np.zeros_like(pos)
    for i in range(4):
        for c in range(3):
            a, b = pos.copy(), pos.copy()
            a[i, c] += h
            b[i, c] -= h
            grad[i, c] = (energy(a) - energy(b)) / (2 * h)
    np.testing.assert_allclose(f, -grad, atol=1e-6)
    # one stretched link: |f| = k (r - r0) along the link
    p2 = np.array([[1.0, 1.0, 1.0], [2.5, 1.0, 1.0]])
    f2 = np.zeros_like(p2)
    e2 = O.elastic_network(p2, f2, np.array([[0, 1]], np.int32), np.array([1.0]), 4.0, EDGES)
    assert e2 == 0.5 * 4.0 * 0.25
    np.testing.assert_array_equal(f2, [[2.0, 0, 0], [-2.0, 0, 0]])


def test_elastic_network_known_answer():
    """test/test_elastic_network_0.f90: box of side 5, three links from particle 1, the
    one to particle 4 across the box; energy (1.5^2 + 0 + 0.5^2) k / 2, force (k, 0, 0)
    on particle 1, no net force"""
    pos = np.ones((4, 3))
    pos[1, 0], pos[2, 1], pos[3, 0] = 3, 2, 4.5
    f = np.zeros((4, 3))
    e = O.elastic_network(pos, f, np.array([[0, 1], [0, 2], [0, 3]], np.int32), np.array([0.5, 1.0, 1.0]), 7.0,
                          [5.0] * 3)
    assert abs(e - (1.5 ** 2 + 0.5 ** 2) * 7 / 2) < 1e-14
    np.testing.assert_allclose(f[0], [7.0, 0, 0], atol=1e-14)
    assert np.max(np.abs(f.sum(0))) < 1e-14


def body(offset=(5.0, 5.0, 5.0)):
    """particles on +-x, +-y, +-z at distances 1, 1.5, 2 (principal axes = the frame),
    two species"""
    r = np.array([[1, 0, 0], [-1, 0, 0], [0, 1.5, 0], [0, -1.5, 0], [0, 0, 2], [0, 0, -2]], float)
    pos = r + np.array(offset)
    species = np.array([1, 1, 2, 2, 1, 1], np.int32)
    return pos, species


def test_rigid_body_init_closed_forms():
    rng = np.random.default_rng(5)
    pos, species = body(offset=(11.5, 5.0, 5.0))
    mass = np.array([2.0, 3.0])
    m = mass[species - 1]
    image = np.zeros((6, 3), np.int32)
    image[0, 0] = 1  # particle 0 unfolded one box to the right ...
    pos_w = pos.copy()
    pos_w[0, 0] -= EDGES[0]  # ... and stored folded
    vel = rng.normal(size=(6, 3))
    b = O.rigid_body_init(pos_w, vel, image, species, mass, EDGES, 0, 5)
    assert b.mass == m.sum()
    np.testing.assert_allclose(b.pos[:], [11.5, 5.0, 5.0], atol=1e-14)
    np.testing.assert_allclose(b.vel[:], (m[:, None] * vel).sum(0) / m.sum(), atol=1e-14)
    np.testing.assert_allclose(b.body_rows, pos - [11.5, 5.0, 5.0], atol=1e-14)
    Ix = 3 * 2 * 1.5 ** 2 + 2 * 2 * 2.0 ** 2
    Iy = 2 * 2 * 1.0 + 2 * 2 * 4.0
    Iz = 2 * 2 * 1.0 + 2 * 3 * 1.5 ** 2
    np.testing.assert_allclose(b.I_body[:], [Ix, Iy, Iz], rtol=1e-15)
    assert list(b.q) == [1, 0, 0, 0] and list(b.L) == [0, 0, 0]


def test_rigid_body_torque_rotated():
    rng = np.random.default_rng(6)
    pos, species = body()
    mass = np.array([1.0, 1.0])
    b = O.rigid_body_init(pos, np.zeros((6, 3)), np.zeros((6, 3), np.int32), species, mass, EDGES, 0, 5)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    b.q[:] = q
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]])  # scipy is scalar last
    f = rng.normal(size=(6, 3))
    O.rigid_body_compute_force_torque(b, f)
    np.testing.assert_allclose(b.force[:], f.sum(0), atol=1e-14)
    np.testing.assert_allclose(b.torque[:], np.cross(R.apply(b.body_rows), f).sum(0), atol=1e-13)


def test_rigid_body_principal_rotation_step():
    """L along the body z axis, no torque: the half-step iteration converges to
    q_n (cos a, 0, 0, sin a) with sin a = w dt / 4 (the fixed point of
    h = normalize(1 + dt/4 h (0, w z))), and each step is then a rotation about z by
    2 atan2(x cos a, 1 - x sin a) with x = w dt / 2"""
    rng = np.random.default_rng(7)
    pos, species = body()
    mass = np.array([1.0, 1.0])
    b = O.rigid_body_init(pos, np.zeros((6, 3)), np.zeros((6, 3), np.int32), species, mass, EDGES, 0, 5)
    Iz = b.I_body[2]
    b.L[:] = [0, 0, 0.9 * Iz]
    w, dt = 0.9, 0.1
    x, a = w * dt / 2, np.arcsin(w * dt / 4)
    beta = np.arctan2(x * np.cos(a), 1 - x * np.sin(a))  # half the rotation angle
    p = pos.copy()
    f = np.zeros((6, 3))
    for step in range(1, 6):
        O.rigid_body_compute_force_torque(b, f)
        it = O.rigid_body_vv1(b, p, dt, 1e-13, EDGES)
        assert it < 20
        np.testing.assert_allclose(b.q[:], [np.cos(step * beta), 0, 0, np.sin(step * beta)], atol=1e-13)
        R = Rotation.from_rotvec([0, 0, 2 * step * beta])
        np.testing.assert_allclose(p, b.pos[:] + R.apply(b.body_rows), atol=1e-12)
        v = np.zeros((6, 3))
        O.rigid_body_vv2(b, v, dt)
        np.testing.assert_allclose(b.omega_body[:], [0, 0, w], atol=1e-14)
        np.testing.assert_allclose(v, np.cross([0, 0, w], R.apply(b.body_rows)), atol=1e-12)


def test_rigid_body_free_asymmetric_top():
    rng = np.random.default_rng(8)
    pos, species = body()
    mass = np.array([1.0, 2.0])
    m = mass[species - 1]
    b = O.rigid_body_init(pos, np.zeros((6, 3)), np.zeros((6, 3), np.int32), species, mass, EDGES, 0, 5)
    b.L[:] = [0.4, 1.1, -0.7]
    b.vel[:] = [0.1, 0.0, -0.2]
    L0 = np.array(b.L[:])
    p, v, f = pos.copy(), np.zeros((6, 3)), np.zeros((6, 3))
    kin = []
    for step in range(200):
        O.rigid_body_compute_force_torque(b, f)
        O.rigid_body_vv1(b, p, 0.01, 1e-12, EDGES)
        O.rigid_body_vv2(b, v, 0.01)
        np.testing.assert_allclose(b.L[:], L0, atol=1e-15)
        assert abs(np.linalg.norm(b.q[:]) - 1) < 1e-14
        # the particles' angular momentum about the centre is L, their momentum M V
        unf = p - np.round((p - b.pos[:]) / EDGES) * EDGES
        r = unf - b.pos[:]
        np.testing.assert_allclose((m[:, None] * np.cross(r, v)).sum(0), L0, atol=1e-11)
        np.testing.assert_allclose((m[:, None] * v).sum(0), b.mass * np.array(b.vel[:]), atol=1e-13)
        # rigid: pairwise distances kept
        assert np.allclose(np.linalg.norm(r, axis=1), np.linalg.norm(b.body_rows, axis=1), atol=1e-12)
        kin.append(0.5 * np.sum(m[:, None] * v ** 2))
    assert np.ptp(kin) / np.mean(kin) < 1e-3
