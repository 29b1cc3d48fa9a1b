"""Pins the rigid-clump oracle (CPU only): Morton codes by hand-derived values,
rigid integration by closed forms (free spin about a principal axis, free
fall, torque-free top), the clump contact restatement against the already
pinned sphere law, and momentum conservation."""
import math

import numpy as np

import pytest

from stardust_amd import _lib as L
from stardust_amd import clumps as CL


def morton_of(oracle, centroids, dom):
    c = np.asarray(centroids, np.float32)
    lo = np.zeros((len(c), 4), np.float32)
    hi = np.zeros((len(c), 4), np.float32)
    lo[:, :3] = c
    hi[:, :3] = c
    return oracle.morton30(lo, hi, dom) >> np.uint64(32)


def test_morton_known_answers(oracle):
    dom = np.array([0, 0, 0, 1024, 1024, 1024], np.float32)  # cell = 1 unit
    got = morton_of(oracle, [[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.5, 1.5, 0.5], [0.5, 0.5, 1.5],
                             [1023.5, 1023.5, 1023.5], [3.5, 5.5, 6.5], [-5, 2000, 0.5]], dom)
    # x -> bit 0, y -> bit 1, z -> bit 2 of each triple; 3 = 0b11, 5 = 0b101, 6 = 0b110
    assert got[:6].tolist() == [0, 1, 2, 4, 0x3FFFFFFF, 0b110101011]
    # clamped: x -> 0, y -> 1023 (all y bits), z -> 0
    assert int(got[6]) == 0x12492492
    keys = oracle.morton30(np.zeros((3, 4), np.float32), np.zeros((3, 4), np.float32), dom)
    assert (keys & np.uint64(0xFFFFFFFF)).tolist() == [0, 1, 2]


def single_clump(offsets, radii, mass_each, q=(1.0, 0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0),
                 pos=(0.0, 0.0, 0.0)):
    offsets = np.asarray(offsets, np.float64)
    k = len(offsets)
    m = np.full(k, mass_each)
    I = np.zeros(6)
    for d, mm, r in zip(offsets, m, radii):
        own = 0.4 * mm * r * r
        dd = d @ d
        I[:3] += own + mm * (dd - d * d)
        I[3] += -mm * d[0] * d[1]
        I[4] += -mm * d[0] * d[2]
        I[5] += -mm * d[1] * d[2]
    s = {"pos4": np.array([[*pos, m.sum()]], np.float32), "vel4": np.array([[*v, 0]], np.float32),
         "quat4": np.array([q], np.float32), "angvel4": np.array([[*w, 0]], np.float32),
         "inertia6": np.array([I], np.float32), "first": np.array([0], np.int32),
         "count": np.array([k], np.int32),
         "local4": np.concatenate([offsets, np.asarray(radii)[:, None]], 1).astype(np.float32)}
    return s, I


def contact(g=(0.0, 0.0, 0.0), dt=1e-4):
    return L.ContactParams.make(1e5, 0.3, 0.5, dt, gravity=g, walls=None)


def test_free_spin_about_principal_axis(oracle):
    # symmetric dumbbell along x: principal axes are x, y, z. Spin about z.
    s, I = single_clump([[0.02, 0, 0], [-0.02, 0, 0]], [0.005, 0.005], 0.01, w=(0.0, 0.0, 50.0))
    st = oracle.ClumpState(s)
    dt, n = 1e-4, 400
    st.step(contact(dt=dt), nsteps=n)
    ang = 50.0 * dt * n
    q = st.Q[0].astype(np.float64)
    expect = np.array([math.cos(ang / 2), 0, 0, math.sin(ang / 2)])
    assert np.allclose(q, expect, atol=2e-5), (q, expect)
    assert np.allclose(st.L[0, :3], [0, 0, I[2] * 50.0], rtol=1e-6)


def test_free_fall_closed_form(oracle):
    s, _ = single_clump([[0.0, 0, 0]], [0.01], 0.02, pos=(0.1, 0.2, 3.0))
    st = oracle.ClumpState(s)
    dt, n = 1e-4, 300
    st.step(contact(g=(0.0, 0.0, -9.81), dt=dt), nsteps=n)
    assert abs(st.P[0, 2] - (3.0 - 9.81 * dt * dt * n * (n + 1) / 2)) < 1e-5
    assert abs(st.V[0, 2] + 9.81 * dt * n) < 1e-5


def test_torque_free_top_conserves_L_and_energy(oracle):
    # asymmetric body spinning about a generic axis: L is exactly constant, energy to O(dt)
    s, _ = single_clump([[0.02, 0, 0], [-0.02, 0, 0], [0, 0.01, 0.004], [0, -0.012, 0]], [0.004, 0.005, 0.006, 0.004],
                        0.01, q=(0.9, 0.1, 0.3, 0.3), w=(3.0, -7.0, 11.0))
    s["quat4"] /= np.linalg.norm(s["quat4"])
    st = oracle.ClumpState(s)
    L0 = st.L[0, :3].copy()

    def energy():
        R = oracle_rot(st.Q[0])
        i0, i1 = st.I0[0], st.I1[0]
        Iinv = np.array([[i0[0], i0[3], i1[0]], [i0[3], i0[1], i1[1]], [i1[0], i1[1], i0[2]]], np.float64)
        w = R @ Iinv @ R.T @ st.L[0, :3].astype(np.float64)
        return 0.5 * w @ st.L[0, :3]

    e0 = energy()
    st.step(contact(dt=1e-5), nsteps=2000)
    assert np.array_equal(st.L[0, :3], L0)
    assert abs(energy() - e0) / e0 < 1e-4
    assert abs(np.linalg.norm(st.Q[0]) - 1) < 1e-6


def oracle_rot(q):
    w, x, y, z = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_clump_contact_matches_sphere_law(oracle):
    # two single-sphere clumps: forces equal the (pinned) sphere contact law with zero history
    r = 0.01
    m = 2500 * 4 / 3 * math.pi * r ** 3
    s = {"pos4": np.array([[0.5, 0.5, 0.5, m], [0.5 + 1.9 * r, 0.5 + 0.2 * r, 0.5, m]], np.float32),
         "vel4": np.array([[0.3, 0.1, 0.0, 0], [-0.2, 0.0, 0.05, 0]], np.float32),
         "quat4": np.array([[1, 0, 0, 0], [1, 0, 0, 0]], np.float32),
         "angvel4": np.zeros((2, 4), np.float32), "inertia6": np.array([[0.4 * m * r * r] * 3 + [0, 0, 0]] * 2,
                                                                          np.float32),
         "first": np.array([0, 1], np.int32), "count": np.array([1, 1], np.int32),
         "local4": np.array([[0, 0, 0, r], [0, 0, 0, r]], np.float32)}
    st = oracle.ClumpState(s)
    c = contact(dt=1e-5)
    f, t = st.forces(c, np.array([[0, 1]], np.int32))
    sp = oracle.SphereState(np.array([0.5, 0.5 + 1.9 * r], np.float32), np.array([0.5, 0.5 + 0.2 * r], np.float32),
                            np.array([0.5, 0.5], np.float32), np.array([r, r], np.float32),
                            np.array([m, m], np.float32), vx=np.array([0.3, -0.2], np.float32),
                            vy=np.array([0.1, 0.0], np.float32), vz=np.array([0.0, 0.05], np.float32))
    g = L.Grid.make((0, 0, 0), 4 * r, (25, 25, 25))
    keys, vals = oracle.hash_keys(sp.pos4[:, 0], sp.pos4[:, 1], sp.pos4[:, 2], g)
    k, v = oracle.sort_pairs_u32(keys, vals, 0, 14)
    ranges = oracle.cell_ranges_u32(k, g.n_cells)
    out = oracle.contact_forces(sp.pos4[v], sp.vel4[v], sp.omg4[v], sp.ids[v], g, ranges, c, v.astype(np.int32),
                                sp.hcount, sp.hist)
    fs = np.empty_like(out["force"])
    fs[v] = out["force"]
    assert np.array_equal(f[:, :3], fs[:, :3])
    assert np.array_equal(f[0, :3], -f[1, :3])


def test_clump_cluster_conserves_momentum(oracle):
    s = CL.generate_clumps(77, 216, lattice=6, spacing=0.035, jitter=0.002, v_sigma=0.5)
    st = oracle.ClumpState(s)
    m = 1.0 / st.P[:, 3].astype(np.float64)
    p0 = (m[:, None] * st.V[:, :3]).sum(0)
    npairs = st.step(contact(dt=1e-5), nsteps=200)
    assert npairs > 50
    p1 = (m[:, None] * st.V[:, :3]).sum(0)
    scale = (m[:, None] * np.abs(st.V[:, :3])).sum()
    assert np.all(np.abs(p1 - p0) < 1e-6 * scale)


def test_body_reduce_against_float64_sums(oracle):
    """or_body_reduce: per-body sums of sphere forces and of (x - X) x f + tau,
    against float64 sums (FP32 rounding only)."""
    rng = np.random.default_rng(12)
    nb = 50
    count = rng.integers(1, 9, nb).astype(np.int32)
    first = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int32)
    ns = int(count.sum())
    bpos = np.zeros((nb, 4), np.float32)
    bpos[:, :3] = rng.uniform(-1, 1, (nb, 3))
    owner = np.repeat(np.arange(nb), count)
    spos = np.zeros((ns, 4), np.float32)
    spos[:, :3] = bpos[owner, :3] + rng.normal(0, 0.01, (ns, 3))
    sf = np.zeros((ns, 4), np.float32)
    st = np.zeros((ns, 4), np.float32)
    sf[:, :3] = rng.normal(0, 10, (ns, 3))
    st[:, :3] = rng.normal(0, 0.1, (ns, 3))
    bf, bt = oracle.body_reduce(bpos, first, count, spos, sf, st)
    a = spos[:, :3].astype(np.float64) - bpos[owner, :3]
    tq = np.cross(a, sf[:, :3].astype(np.float64)) + st[:, :3]
    F64 = np.zeros((nb, 3))
    T64 = np.zeros((nb, 3))
    np.add.at(F64, owner, sf[:, :3].astype(np.float64))
    np.add.at(T64, owner, tq)
    assert np.allclose(bf[:, :3], F64, rtol=1e-5, atol=1e-4)
    assert np.allclose(bt[:, :3], T64, rtol=1e-5, atol=1e-6)


def test_integrate_bodies_free_spin_and_fall(oracle):
    """or_integrate_bodies: a free spin about a principal axis follows
    q(t) = exp(w t / 2) q0 and the body falls as z0 - g dt^2 n (n + 1) / 2."""
    nb = 3
    P = np.array([[0, 0, 1, 1.0], [1, 2, 3, 0.5], [0, 0, 0, 0.0]], np.float32)  # last: static
    V = np.zeros((nb, 4), np.float32)
    Q = np.tile(np.array([1, 0, 0, 0], np.float32), (nb, 1))
    I = np.array([0.2, 0.3, 0.4], np.float64)  # principal moments (body frame = world at t = 0)
    inv6 = np.zeros((nb, 6), np.float32)
    inv6[:, :3] = 1.0 / I
    w = np.array([0.0, 0.0, 3.0])  # about z
    Lm = np.zeros((nb, 4), np.float32)
    Lm[:, :3] = (I * w).astype(np.float32)
    f = np.zeros((nb, 4), np.float32)
    c = L.ContactParams.make(1e5, 0.3, 0.5, 1e-3, gravity=(0.0, 0.0, -9.81), walls=None)
    n = 200
    for _ in range(n):
        P, V, Q, Lm = oracle.integrate_bodies(P, V, Q, Lm, inv6, f, f, c)
    t = n * 1e-3
    assert abs(Q[0, 0] - math.cos(1.5 * t)) < 2e-5 and abs(Q[0, 3] - math.sin(1.5 * t)) < 2e-5
    assert abs(P[0, 2] - (1.0 - 9.81 * 1e-3 ** 2 * n * (n + 1) / 2)) < 2e-5
    assert np.array_equal(P[2], np.array([0, 0, 0, 0], np.float32)) and np.array_equal(Q[2], [1, 0, 0, 0])


def test_clump_remove_and_add_far_bodies_change_nothing_else(oracle):
    """Bodies 10 m away never touch the others: carrying them and removing
    them at step k (or adding them at step k) leaves every other body
    bit-identical to a run that never had them, and the far bodies themselves
    move as when stepped alone."""
    from stardust_amd import clumps as CL
    c = CL.contact_c3()
    near = CL.generate_clumps(0xD1, 125, lattice=5, spacing=0.035, jitter=0.002, v_sigma=0.5)
    far = CL.generate_clumps(0xD2, 8, lattice=2, spacing=0.05, jitter=0.002, v_sigma=0.5)
    far["pos4"][:, :3] += np.float32(10.0)
    both = {k: np.concatenate([near[k], far[k] + (len(near["local4"]) if k == "first" else 0)])
            for k in near}
    a, b, f = oracle.ClumpState(both), oracle.ClumpState(near), oracle.ClumpState(far)
    for st in (a, b, f):
        st.step(c, 4)
    a.remove(np.arange(125, 133))
    for st in (a, b):
        st.step(c, 5)
    for k in ("P", "V", "Q", "L"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    a.add(far)
    f2 = oracle.ClumpState(far)
    for st in (a, b, f2):
        st.step(c, 6)
    for k in ("P", "V", "Q", "L"):
        assert np.array_equal(getattr(a, k)[:125], getattr(b, k)), k
        assert np.array_equal(getattr(a, k)[125:], getattr(f2, k)), k
    assert a.nb == 133 and a.ns == len(both["local4"])
