"""Pins the oracle's rigid convex-body step (or_hull_step) on the CPU.

Closed forms of the contact law and the integration: a head-on collision of
two sphere bodies has the restitution exp(-zeta pi / sqrt(1 - zeta^2)) of the
linear spring-dashpot law; a sphere body launched sliding on a floor of static
triangles reaches rolling at 5/7 of its speed (friction and torque); a cube
hull resting on a triangle floor sinks by m g / kn without tipping (the
contact sits at the face centre); and in a crowd of tumbling hulls with no
gravity and no static body, total linear and angular momentum stay constant
(pair forces are opposite and applied at one point). Scenes are independent:
stepped together they equal each stepped alone, bit for bit.
"""
import math

import numpy as np
import pytest

from stardust_amd import _lib as L
from stardust_amd import hullbody as HB
from stardust_amd import hulls as H


def law(kn, zeta, mu, dt, gravity=(0.0, 0.0, 0.0)):
    return L.ContactParams.make(kn, zeta, mu, dt, gravity=gravity, walls=None, max_contacts=16)


def sphere_bodies(centres, r, m, vel=None):
    n = len(centres)
    b = HB.triangle_bodies(np.zeros((n, 3, 3), np.float32))
    b["pos4"][:, :3] = np.asarray(centres, np.float32)
    b["pos4"][:, 3] = m
    b["nv"][:] = 0
    b["vstart"][:] = 0
    b["verts4"] = np.zeros((0, 4), np.float32)
    b["radius"][:] = r
    I = 0.4 * m * r * r
    b["inertia6"][:, :3] = I
    if vel is not None:
        b["vel4"][:, :3] = np.asarray(vel, np.float32)
    return b


def floor(size=1.0, z=0.0, centre=(0.5, 0.5)):
    cx, cy = centre
    lo, hi = -size, size
    c = np.array([[cx + lo, cy + lo, z], [cx + hi, cy + lo, z], [cx + hi, cy + hi, z], [cx + lo, cy + hi, z]],
                 np.float32)
    return HB.triangle_bodies(np.stack([c[[0, 1, 2]], c[[0, 2, 3]]]))


def cube(centre, half, m, vel=(0.0, 0.0, 0.0)):
    s = np.array([[x, y, z] for x in (-half, half) for y in (-half, half) for z in (-half, half)], np.float32)
    b = HB.triangle_bodies(np.zeros((1, 3, 3), np.float32))
    b["verts4"] = np.concatenate([s, np.zeros((8, 1), np.float32)], 1)
    b["nv"][:] = 8
    b["pos4"][0, :3] = centre
    b["pos4"][0, 3] = m
    b["vel4"][0, :3] = vel
    I = m * (2 * half) ** 2 / 6.0
    b["inertia6"][0, :3] = I
    return b


@pytest.mark.parametrize("zeta", [0.0, 0.3])
def test_sphere_bodies_head_on_restitution(oracle, zeta):
    r, v0, kn, dt = 0.01, 0.2, 1e5, 1e-6
    m = 2500 * 4 / 3 * math.pi * r ** 3
    b = sphere_bodies([[0.5 - r - 1e-5, 0.5, 0.5], [0.5 + r + 1e-5, 0.5, 0.5]], r, m, [[v0, 0, 0], [-v0, 0, 0]])
    st = oracle.HullBodyState(b)
    st.step(law(kn, zeta, 0.5, dt), nsteps=4000)
    va, vb = float(st.V[0, 0]), float(st.V[1, 0])
    e = (vb - va) / (2 * v0)
    assert abs(e - math.exp(-zeta * math.pi / math.sqrt(1 - zeta ** 2))) < 0.01, e
    assert abs(va + vb) < 1e-6


def test_sphere_body_sliding_on_triangles_reaches_rolling(oracle):
    r, v0, dt, kn = 0.01, 0.5, 1e-6, 1e5
    m = 2500 * 4 / 3 * math.pi * r ** 3
    settle = m * 9.81 / kn
    b, first = HB.concat([sphere_bodies([[0.3, 0.5, r - settle]], r, m, [[v0, 0, 0]]), floor()])
    st = oracle.HullBodyState(b)
    c = law(kn, 0.3, 0.5, dt, gravity=(0.0, 0.0, -9.81))
    st.step(c, nsteps=45000)  # the slip ends after v0 / (3.5 mu g) = 29 ms
    vx = float(st.V[0, 0])
    w = st.L[0, :3] / (0.4 * m * r * r)
    assert abs(vx - 5 / 7 * v0) / (5 / 7 * v0) < 0.01, vx
    assert abs(vx - w[1] * r) / vx < 0.01, (vx, w)
    assert np.array_equal(st.P[1:, :3], b["pos4"][1:, :3])  # the floor is static


def test_cube_rests_on_triangle_floor_at_mg_over_kn(oracle):
    half, kn, dt = 0.01, 1e5, 2e-5
    m = 2500 * (2 * half) ** 3
    # away from the floor's diagonal: the cube touches one of its two triangles
    b, _ = HB.concat([cube((0.7, 0.3, half), half, m), floor()])
    st = oracle.HullBodyState(b)
    st.step(law(kn, 0.7, 0.5, dt, gravity=(0.0, 0.0, -9.81)), nsteps=6000)
    sink = half - float(st.P[0, 2])
    assert abs(sink - m * 9.81 / kn) / (m * 9.81 / kn) < 2e-3, sink
    assert np.abs(st.Q[0] - np.array([1, 0, 0, 0], np.float32)).max() < 1e-5  # no tipping
    assert np.abs(st.P[0, :2] - np.float32([0.7, 0.3])).max() < 1e-6


def crowd(n=64, seed=3):
    h = H.generate_hulls(seed, n, lo=(0.0, 0.0, 0.0), hi=(0.06, 0.06, 0.06), v_sigma=0.3)
    return HB.make_bodies(h, seed)


def test_crowd_conserves_linear_and_angular_momentum(oracle):
    b = crowd()
    st = oracle.HullBodyState(b)
    c = law(2e4, 0.3, 0.5, 1e-5)

    def momenta():
        m = np.where(st.P[:, 3] > 0, 1.0 / st.P[:, 3].astype(np.float64), 0.0)
        X, V = st.P[:, :3].astype(np.float64), st.V[:, :3].astype(np.float64)
        p = (m[:, None] * V).sum(0)
        Lt = (np.cross(X, m[:, None] * V) + st.L[:, :3].astype(np.float64)).sum(0)
        scale_p = (m[:, None] * np.abs(V)).sum()
        scale_l = np.abs(np.cross(X, m[:, None] * V)).sum() + np.abs(st.L[:, :3]).sum()
        return p, Lt, scale_p, scale_l

    p0, l0, sp, sl = momenta()
    touched = 0
    for _ in range(40):
        touched += st.step(c, nsteps=5)
    p1, l1, _, _ = momenta()
    assert touched > 20  # the hulls were in contact
    assert np.all(np.abs(p1 - p0) < 1e-5 * sp), (p1 - p0, sp)
    assert np.all(np.abs(l1 - l0) < 1e-5 * sl), (l1 - l0, sl)


def test_scenes_are_independent(oracle):
    a, b2 = crowd(40, 5), crowd(50, 6)
    both, first = HB.concat([a, b2])
    c = law(2e4, 0.3, 0.5, 1e-5, gravity=(0.0, 0.0, -9.81))
    s_all = oracle.HullBodyState(both, first)
    s_a, s_b = oracle.HullBodyState(a), oracle.HullBodyState(b2)
    for st in (s_all, s_a, s_b):
        st.step(c, nsteps=20)
    for k in ("P", "V", "Q", "L"):
        assert np.array_equal(getattr(s_all, k)[:40], getattr(s_a, k)), k
        assert np.array_equal(getattr(s_all, k)[40:], getattr(s_b, k)), k


def test_grid_broadphase_equals_brute_force(oracle):
    """The body step's candidate pairs come from per-scene grids fixed at
    upload. After the bodies have moved, some of them far outside the initial
    grid (clamped to its border cells, several of them next to each other),
    the pairs are exactly the pairs of overlapping boxes (closed intervals)
    within a scene, without pairs of two static bodies, each once as (a, b)
    with a < b (brute force). The wall triangles, larger than any moving body,
    go to each scene's second grid; one scene lists them after its hulls, one
    before, and a fourth scene has a floor of 18 triangles spanning several
    cells of that grid."""
    h1 = H.generate_hulls(0x4E01, 300, lo=(0.0, 0.0, 0.0), hi=(0.12, 0.12, 0.12))
    s1, _ = HB.concat([HB.make_bodies(h1, 1), HB.box_walls((0.0, 0.0, 0.0), (0.12, 0.12, 0.12))])
    h2 = H.generate_hulls(0x4E02, 200, lo=(0.0, 0.0, 0.0), hi=(0.1, 0.1, 0.1))
    sp = sphere_bodies(np.random.default_rng(3).uniform(0.0, 0.1, (40, 3)), np.float32(0.006), np.float32(0.01))
    h3 = H.generate_hulls(0x4E03, 150, lo=(0.0, 0.0, 0.0), hi=(0.08, 0.08, 0.08))
    s3, _ = HB.concat([HB.box_walls((0.0, 0.0, 0.0), (0.08, 0.08, 0.08)), HB.make_bodies(h3, 3)])
    h4 = H.generate_hulls(0x4E04, 150, lo=(0.0, 0.0, 0.0), hi=(0.12, 0.12, 0.03))
    q = np.linspace(0.0, 0.12, 4, dtype=np.float32)
    tris = []
    for i in range(3):
        for j in range(3):
            a, b, c, d = ((q[i], q[j], 0.0), (q[i + 1], q[j], 0.0), (q[i + 1], q[j + 1], 0.0), (q[i], q[j + 1], 0.0))
            tris += [(a, b, c), (a, c, d)]
    s4, _ = HB.concat([HB.make_bodies(h4, 4), HB.triangle_bodies(np.array(tris, np.float32))])
    bodies, first = HB.concat([s1, HB.concat([HB.make_bodies(h2, 2), sp])[0], s3, s4])
    n = len(bodies["pos4"])
    st = oracle.HullBodyState(bodies, first)
    walls = np.r_[np.arange(300, 312), np.arange(first[2], first[2] + 12), np.arange(first[4] - 18, first[4])]
    assert np.array_equal(np.nonzero(st.large)[0], walls)
    assert st.grid16[3].view(np.int32)[12:14].min() >= 2  # the floor spans several cells of the second grid
    rng = np.random.default_rng(7)
    st.P[:, :3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    far = rng.choice(n, 40, replace=False)
    st.P[far, :3] += rng.uniform(-2.0, 2.0, (40, 3)).astype(np.float32)
    st.P[far[1::2], :3] = st.P[far[0::2], :3] + np.float32(0.003)  # far bodies in overlapping pairs
    spos = np.concatenate([st.P[:, :3], bodies["radius"][:, None]], 1).astype(np.float32)
    lo, hi = oracle.hull_aabbs({"pos4": spos, "quat4": st.Q, "vstart": bodies["vstart"], "nv": bodies["nv"],
                                "verts4": bodies["verts4"]})
    got = st.grid_pairs(lo, hi)
    assert np.all(got[:, 0] < got[:, 1])
    static = st.P[:, 3] == 0.0
    want = set()
    for s in range(len(first) - 1):
        ids = np.arange(first[s], first[s + 1])
        l, u = lo[ids, :3], hi[ids, :3]
        ov = np.all((l[:, None, :] <= u[None, :, :]) & (l[None, :, :] <= u[:, None, :]), axis=2)
        for i, j in zip(*np.nonzero(np.triu(ov, 1))):
            a, b = int(ids[i]), int(ids[j])
            if not (static[a] and static[b]):
                want.add((a, b))
    assert len(got) == len(set(map(tuple, got.tolist())))
    assert set(map(tuple, got.tolist())) == want
    assert len(want) > 300 and any(a in far or b in far for a, b in want)
    assert any(a in walls[12:24] for a, b in want) and any(b in walls[:12] for a, b in want)
    assert sum(b in walls[24:] for a, b in want) > 20  # hulls resting on the floor
