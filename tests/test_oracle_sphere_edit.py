"""Pins the oracle's runtime add / remove of spheres (CPU only).

The rule (drop the rows, renumber the rest in id order, keep the slot order
and every history entry whose partner stays) is checked against what it must
leave unchanged: spheres far from everything never touch the others, so a run
that carries them and removes them (or adds them) at step k is bit-identical,
for every other sphere, to a run that never had them; the added spheres fall
freely (the semi-implicit Euler closed form); and history entries of kept
pairs survive a removal unchanged while those of a removed partner go.
"""
import numpy as np

from stardust_amd import _lib as L
from stardust_amd import workloads as W

FIELDS = ("pos4", "vel4", "omg4")


def dense(cells=(4, 4, 4)):
    w = W.dem_c2_dense(cells)
    return w, w.generate()


def scene_with_far_spheres(w, s, k_far):
    """The lattice in the lower part of a taller walled box, k_far isolated
    spheres above it (no contact with anything within the run)."""
    hi = (w.hi[0], w.hi[1], w.hi[2] + 0.4)
    g = W.grid_for(w.lo, hi, w.cell)
    c = L.ContactParams.make(w.kn, 0.3, 0.5, w.dt, gravity=(0.0, 0.0, -9.81), walls=(w.lo, hi),
                             kt_ratio=W.KT_RATIO, max_contacts=w.max_contacts)
    # a vertical stack over the centre of the lattice, 5 cm apart, inside the walls
    fx = np.full(k_far, 0.5 * w.hi[0], np.float32)
    fy = np.full(k_far, 0.5 * w.hi[1], np.float32)
    fz = (w.hi[2] + 0.05 + 0.05 * np.arange(k_far)).astype(np.float32)
    fr = np.full(k_far, 0.004, np.float32)
    return g, c, (fx, fy, fz, fr, W.sphere_mass(fr))


def velocities(n, seed=3):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.3, (3, n)).astype(np.float32)


def test_remove_isolated_spheres_leaves_the_rest_unchanged(oracle):
    w, s = dense()
    k_far = 6
    g, c, far = scene_with_far_spheres(w, s, k_far)
    v = velocities(w.n)
    n = w.n
    # run A: lattice ids 0..n-1, far spheres n..n+k-1; removed at step 15
    cat = lambda a, b: np.concatenate([a, b]).astype(np.float32)
    z = np.zeros(k_far, np.float32)
    a = oracle.SphereState(cat(s["x"], far[0]), cat(s["y"], far[1]), cat(s["z"], far[2]), cat(s["r"], far[3]),
                           cat(s["m"], far[4]), vx=cat(v[0], z), vy=cat(v[1], z), vz=cat(v[2], z),
                           max_contacts=w.max_contacts)
    b = oracle.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2],
                           max_contacts=w.max_contacts)
    a.step(g, c, nsteps=15)
    b.step(g, c, nsteps=15)
    assert int((a.hcount > 0).sum()) > 0  # the lattice is in contact
    a.remove(np.arange(n, n + k_far))
    assert a.n == n
    a.step(g, c, nsteps=15)
    b.step(g, c, nsteps=15)
    ea, eb = a.by_id(), b.by_id()
    for f in FIELDS:
        assert np.array_equal(ea[f], eb[f]), f


def test_remove_renumbers_ids_and_keeps_other_history(oracle):
    w, s = dense()
    v = velocities(w.n, 5)
    g, c = w.grid(), w.contact()
    st = oracle.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2],
                            max_contacts=w.max_contacts)
    st.step(g, c, nsteps=10)
    hist = st.hist.reshape(st.maxc, st.n, 4).copy()
    ids_before, hc_before = st.ids.copy(), st.hcount.copy()
    gone = np.array([3, 17, 40, 41, 200])
    # entries (sphere, partner) before, by id
    def entries(ids, hc, h):
        out = {}
        hid = h[:, :, 3].view(np.int32)
        for i in range(len(ids)):
            out[int(ids[i])] = [(int(hid[k, i]), tuple(h[k, i, :3].tolist())) for k in range(hc[i])]
        return out
    before = entries(ids_before, hc_before, hist)
    assert any(p in gone for e in before.values() for p, _ in e)  # some removed sphere was a partner
    st.remove(gone)
    after = entries(st.ids, st.hcount, st.hist.reshape(st.maxc, st.n, 4))
    newid = {old: old - int((gone < old).sum()) for old in range(w.n) if old not in gone}
    assert sorted(after) == list(range(w.n - len(gone)))
    for old, new in newid.items():
        want = [((newid[p] if p >= 0 else p), t) for p, t in before[old] if p < 0 or p not in gone]
        assert after[new] == want
    # the slot order of the kept rows is unchanged
    kept = [i for i in ids_before if i not in gone]
    assert [newid[i] for i in kept] == st.ids.tolist()


def test_added_spheres_fall_freely_and_change_nothing_else(oracle):
    w, s = dense()
    k_far = 4
    g, c, far = scene_with_far_spheres(w, s, k_far)
    v = velocities(w.n, 9)
    a = oracle.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2],
                           max_contacts=w.max_contacts)
    b = oracle.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2],
                           max_contacts=w.max_contacts)
    a.step(g, c, nsteps=12)
    b.step(g, c, nsteps=12)
    a.add(*far[:4], far[4])
    assert a.n == w.n + k_far and a.ids[-k_far:].tolist() == list(range(w.n, w.n + k_far))
    steps = 20
    a.step(g, c, nsteps=steps)
    b.step(g, c, nsteps=steps)
    ea, eb = a.by_id(), b.by_id()
    for f in FIELDS:
        assert np.array_equal(ea[f][:w.n], eb[f]), f
    # free fall, semi-implicit Euler in float32: v += g dt, z += v dt
    vz = np.zeros(k_far, np.float32)
    z = far[2].copy()
    for _ in range(steps):
        vz = (vz + np.float32(-9.81) * np.float32(c.dt)).astype(np.float32)
        z = (z + vz * np.float32(c.dt)).astype(np.float32)
    assert np.array_equal(ea["pos4"][w.n:, 2], z)
    assert np.array_equal(ea["vel4"][w.n:, 2], vz)
    assert np.array_equal(ea["pos4"][w.n:, 0], far[0])
