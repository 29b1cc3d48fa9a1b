"""Batched multi-scene stepping on the GPU: every scene of a batch is
bit-identical to the same scene stepped alone (sd_ctx, packed u64 keys) and to
the oracle; the autotuner changes launch shapes but no result."""
import math
import sys

import numpy as np
import pytest

from stardust_amd import workloads as W
from stardust_amd.scene import BatchScene, SphereScene

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "wx", "wy", "wz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_batch(nscenes=12, seed=0x5EEDB001):
    # dense columns with random velocities so that contacts form within a few hundred steps
    return W.BatchWorkload("test_batch", nscenes, seed, n_min=200, n_max=3000, side=(0.04, 0.12),
                           r_med=(0.002, 0.005), kn=(5e4, 5e5), phi_seed=0.3)


def velocities(n, seed, sigma=1.5):
    rng = np.random.default_rng(seed)
    return [rng.normal(0.0, sigma, n).astype(np.float32) for _ in range(3)]


def batch_of(w, scenes, s, v, **kw):
    b = BatchScene(**kw)
    b.upload_scenes(scenes, s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2])
    return b


def solo(scene, s, v, lo, hi, graph=True):
    n, g, c = scene
    sc = SphereScene(g, c, key64=True, use_graph=graph)
    sc.upload(s["x"][lo:hi], s["y"][lo:hi], s["z"][lo:hi], s["r"][lo:hi], s["m"][lo:hi], vx=v[0][lo:hi],
              vy=v[1][lo:hi], vz=v[2][lo:hi])
    return sc


def assert_same(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), (what, k, np.max(np.abs(a[k] - b[k])))


def test_batch_equals_each_scene_alone():
    say("batch vs solo, 12 scenes")
    w = small_batch()
    s, scenes = w.generate()
    v = velocities(len(s["x"]), 1)
    b = batch_of(w, scenes, s, v)
    steps = 150
    b.step(steps)
    got = b.download()
    assert b.contact_count() > 0
    off = 0
    for k, sc_desc in enumerate(scenes):
        n = sc_desc[0]
        sc = solo(sc_desc, s, v, off, off + n)
        sc.step(steps)
        ref = sc.download()
        assert_same({f: got[f][off:off + n] for f in FIELDS}, ref, "scene %d" % k)
        sc.close()
        off += n
    b.close()


def test_batch_wave_boundaries():
    """Scene sizes that put some scene boundaries exactly on wave boundaries
    (multiples of 64 slots) and others inside a wave: single-scene waves take
    their scene's parameters as wave-uniform values, waves spanning two scenes
    run in a launch of their own; every scene is bit-identical to it alone."""
    sizes = [64, 128, 200, 56, 64, 192, 65, 63]  # boundaries 64 128+64 392 448 512 704 769 832
    say("batch vs solo, scene sizes", sizes)
    w = small_batch(nscenes=len(sizes), seed=0x5EEDB0A1)
    s_full, scenes_full = w.generate()
    keep, scenes, off = [], [], 0
    for (n, g, c), m in zip(scenes_full, sizes):
        assert n >= m
        keep.append(np.arange(off, off + m))
        scenes.append((m, g, c))
        off += n
    idx = np.concatenate(keep)
    s = {k: np.ascontiguousarray(s_full[k][idx]) for k in ("x", "y", "z", "r", "m")}
    v = velocities(len(idx), 3)
    b = batch_of(w, scenes, s, v)
    steps = 120
    b.step(steps)
    got = b.download()
    off = 0
    for k, sc_desc in enumerate(scenes):
        n = sc_desc[0]
        sc = solo(sc_desc, s, v, off, off + n)
        sc.step(steps)
        ref = sc.download()
        assert_same({f: got[f][off:off + n] for f in FIELDS}, ref, "scene %d" % k)
        sc.close()
        off += n
    b.close()


def test_batch_scene_equals_oracle(oracle):
    say("batch vs oracle, 3 scenes")
    w = small_batch(nscenes=3, seed=0x5EEDB002)
    s, scenes = w.generate()
    v = velocities(len(s["x"]), 2)
    b = batch_of(w, scenes, s, v, use_graph=False)
    b.step(40)
    got = b.download()
    off = 0
    for k, (n, g, c) in enumerate(scenes):
        sl = slice(off, off + n)
        st = oracle.SphereState(s["x"][sl], s["y"][sl], s["z"][sl], s["r"][sl], s["m"][sl], vx=v[0][sl],
                                vy=v[1][sl], vz=v[2][sl], max_contacts=c.max_contacts)
        st.step(g, c, nsteps=40, nthreads=8)
        e = st.by_id()
        for f, (arr, col) in {"x": ("pos4", 0), "y": ("pos4", 1), "z": ("pos4", 2), "vx": ("vel4", 0),
                              "vy": ("vel4", 1), "vz": ("vel4", 2), "wx": ("omg4", 0), "wy": ("omg4", 1),
                              "wz": ("omg4", 2)}.items():
            assert np.array_equal(got[f][sl], e[arr][:, col]), ("scene", k, f)
        off += n
    b.close()


def test_autotune_changes_no_result():
    say("autotune determinism")
    w = small_batch(nscenes=6, seed=0x5EEDB003)
    s, scenes = w.generate()
    v = velocities(len(s["x"]), 3)
    a = batch_of(w, scenes, s, v)
    b = batch_of(w, scenes, s, v)
    a.step(10)
    b.step(10)
    choice = b.autotune(trial_steps=3)
    assert choice["tile_items"] in (4, 8, 16) and choice["contact_block"] in (64, 128, 256)
    assert choice["ms_per_step"] > 0
    assert_same(a.download(), b.download(), "after autotune")  # the state was restored
    a.step(30)
    b.step(30)
    assert_same(a.download(), b.download(), "stepping with the tuned shapes")
    a.close()
    b.close()


def test_batch_rejects_mixed_dt():
    say("batch argument checks")
    from stardust_amd import _lib as L
    w = small_batch(nscenes=2, seed=0x5EEDB004)
    s, scenes = w.generate()
    n0, g0, c0 = scenes[0]
    c_bad = L.ContactParams.make(c0.kn, 0.3, 0.5, c0.dt * 2, walls=None, max_contacts=c0.max_contacts)
    b = BatchScene()
    with pytest.raises(L.SdError):
        b.upload_scenes([scenes[0], (scenes[1][0], scenes[1][1], c_bad)], s["x"], s["y"], s["z"], s["r"], s["m"])
    b.close()


def test_batch_c5_full_size_spot_checks():
    say("batch c5 full size")
    w = W.batch_c5()
    s, scenes = w.generate()
    zero = [np.zeros(len(s["x"]), np.float32)] * 3
    b = batch_of(w, scenes, s, zero)
    steps = 20
    b.step(steps)
    got = b.download()
    for f in FIELDS:
        assert np.all(np.isfinite(got[f])), f
    firsts = np.concatenate([[0], np.cumsum([d[0] for d in scenes])])
    ns = np.array([d[0] for d in scenes])
    picks = sorted({int(np.argmin(ns)), int(np.argmax(ns)), int(np.argmax(w.height / w.side)), 255})
    for k in picks:
        lo, hi = int(firsts[k]), int(firsts[k + 1])
        sc = solo(scenes[k], s, zero, lo, hi)
        sc.step(steps)
        assert_same({f: got[f][lo:hi] for f in FIELDS}, sc.download(), "c5 scene %d" % k)
        sc.close()
    b.close()


@pytest.mark.parametrize("scale,speed", [(1, 150.0), (4, 150.0), (1, 2.0)])
def test_batch_many_cell_changes(oracle, scale, speed):
    """Fast spheres in scenes of different origins and cell sizes: the re-sort
    steps place hundreds of changed spheres across scene cell blocks (scale 1),
    or give up and redo with the full radix sort (scale 4, a few thousand
    changes a step); every scene equals the oracle bit for bit."""
    from stardust_amd import _lib as L
    say("batch many cell changes", scale, speed)
    rng = np.random.default_rng(43)
    spec = [(1500, 0.2, 0.004, (0.0, 0.0, 0.0)), (250, 0.1, 0.003, (1.0, -2.0, 0.5)),
            (2000, 0.25, 0.005, (-3.0, 0.0, 0.0)), (750, 0.15, 0.0035, (0.0, 4.0, -1.0))]
    scenes, cols, states = [], {k: [] for k in ("x", "y", "z", "r", "m", "vx", "vy", "vz")}, []
    for n0, box, r, org in spec:
        n = n0 * scale
        box = box * scale ** (1.0 / 3.0)
        lo = np.array(org, np.float32)
        pos = (lo + rng.uniform(r, box - r, (n, 3))).astype(np.float32)
        rad = rng.uniform(0.5 * r, r, n).astype(np.float32)
        m = W.sphere_mass(rad)
        vel = rng.normal(0, speed, (n, 3)).astype(np.float32)
        nc = int(math.ceil(box / (2 * r)))
        g = L.Grid.make(tuple(float(v) for v in lo), 2 * r, (nc, nc, nc))
        c = L.ContactParams.make(1e5, 0.3, 0.5, 2e-6, gravity=(0.0, 0.0, -9.81),
                                 walls=(tuple(float(v) for v in lo), tuple(float(v) + box for v in lo)),
                                 max_contacts=32)
        scenes.append((n, g, c))
        for k, a in (("x", pos[:, 0]), ("y", pos[:, 1]), ("z", pos[:, 2]), ("r", rad), ("m", m),
                     ("vx", vel[:, 0]), ("vy", vel[:, 1]), ("vz", vel[:, 2])):
            cols[k].append(np.ascontiguousarray(a))
        st = oracle.SphereState(pos[:, 0], pos[:, 1], pos[:, 2], rad, m, vx=vel[:, 0], vy=vel[:, 1],
                                vz=vel[:, 2], max_contacts=32)
        st.step(g, c, nsteps=12, nthreads=8)
        states.append(st)
    s = {k: np.concatenate(v) for k, v in cols.items()}
    b = batch_of(None, scenes, s, [s["vx"], s["vy"], s["vz"]])
    for k in (5, 4, 3):
        b.step(k)
    got = b.download()
    changes, full_next = b.cell_changes()
    say("cell changes in the last step", changes, "full sort next:", full_next)
    off = 0
    for k, ((n, g, c), st) in enumerate(zip(scenes, states)):
        sl = slice(off, off + n)
        e = st.by_id()
        for f, (arr, col) in {"x": ("pos4", 0), "y": ("pos4", 1), "z": ("pos4", 2), "vx": ("vel4", 0),
                              "vy": ("vel4", 1), "vz": ("vel4", 2), "wx": ("omg4", 0), "wy": ("omg4", 1),
                              "wz": ("omg4", 2)}.items():
            assert np.array_equal(got[f][sl], e[arr][:, col]), ("scene", k, f)
        off += n
    if speed > 100.0:
        assert changes > 100 and full_next == (scale > 1)
    b.close()


def test_batch_tiny_scenes_equal_oracle(oracle):
    """Forty scenes of 6 to 40 spheres: every wave spans several scenes (with
    their own grids), and each scene equals the oracle bit for bit."""
    from stardust_amd import _lib as L
    say("batch of tiny scenes")
    rng = np.random.default_rng(59)
    scenes, cols, states = [], {k: [] for k in ("x", "y", "z", "r", "m", "vx", "vy", "vz")}, []
    for k in range(40):
        n = int(rng.integers(6, 41))
        r_max = float(rng.uniform(0.002, 0.004))
        box = 6 * r_max
        lo = np.array([k * 0.1, 0.0, 0.0], np.float32)
        pos = (lo + rng.uniform(r_max, box - r_max, (n, 3))).astype(np.float32)
        rad = rng.uniform(0.5 * r_max, r_max, n).astype(np.float32)
        m = W.sphere_mass(rad)
        vel = rng.normal(0, 0.05, (n, 3)).astype(np.float32)
        nc = 3
        g = L.Grid.make(tuple(float(v) for v in lo), 2 * r_max, (nc, nc, nc))
        c = L.ContactParams.make(1e4, 0.3, 0.5, 1e-5, gravity=(0.0, 0.0, -9.81),
                                 walls=(tuple(float(v) for v in lo), tuple(float(v) + box for v in lo)),
                                 max_contacts=32)
        scenes.append((n, g, c))
        for key, a in (("x", pos[:, 0]), ("y", pos[:, 1]), ("z", pos[:, 2]), ("r", rad), ("m", m),
                       ("vx", vel[:, 0]), ("vy", vel[:, 1]), ("vz", vel[:, 2])):
            cols[key].append(np.ascontiguousarray(a))
        st = oracle.SphereState(pos[:, 0], pos[:, 1], pos[:, 2], rad, m, vx=vel[:, 0], vy=vel[:, 1],
                                vz=vel[:, 2], max_contacts=32)
        st.step(g, c, nsteps=8, nthreads=1)
        states.append(st)
    s = {key: np.concatenate(v) for key, v in cols.items()}
    b = batch_of(None, scenes, s, [s["vx"], s["vy"], s["vz"]])
    b.step(8)
    got = b.download()
    assert b.contact_count() > 0
    off = 0
    for k, ((n, g, c), st) in enumerate(zip(scenes, states)):
        sl = slice(off, off + n)
        e = st.by_id()
        for f, (arr, col) in {"x": ("pos4", 0), "y": ("pos4", 1), "z": ("pos4", 2), "vx": ("vel4", 0),
                              "vy": ("vel4", 1), "vz": ("vel4", 2), "wx": ("omg4", 0), "wy": ("omg4", 1),
                              "wz": ("omg4", 2)}.items():
            assert np.array_equal(got[f][sl], e[arr][:, col]), ("scene", k, f)
        off += n
    b.close()


def test_batch_checkpoint_resume_and_scene_table_check():
    """sd_ctx_save / sd_ctx_load on a batch context: a resumed batch steps
    bit-identically to the uninterrupted one; a state saved from a batch with
    the same totals but another scene table (one scene's stiffness changed) is
    rejected."""
    say("batch checkpoint")
    w = small_batch(nscenes=5, seed=0x5EEDB004)
    s, scenes = w.generate()
    v = velocities(len(s["x"]), 4)
    a = batch_of(w, scenes, s, v)
    a.step(40)
    buf = a.save()
    a.step(30)
    ref = a.download()
    b = batch_of(w, scenes, s, v)
    b.load(buf)
    b.step(30)
    assert_same(ref, b.download(), "resumed batch")
    n0, g0, c0 = scenes[0]
    c1 = type(c0).from_buffer_copy(c0)
    c1.kn = 2.0 * c0.kn
    other = [(n0, g0, c1)] + list(scenes[1:])
    d = batch_of(w, other, s, v)
    with pytest.raises(Exception, match="scene table"):
        d.load(buf)
    a.close()
    b.close()
    d.close()


def test_batch_steps_attached_hull_scenes(warm_model):
    """sd_batch_attach_hulls: the rigid-body scenes step inside the batched
    step (graph-captured, forked stream); afterwards the spheres equal the
    batch stepped without them and the bodies equal their own context stepped
    alone, bit for bit; also after the attached context switches to the
    learned warm start (the batch captures its step again)."""
    say("batch + attached hull scenes")
    from stardust_amd import _lib as L
    from stardust_amd import hullbody as HB
    from stardust_amd import hulls as H
    w = small_batch(nscenes=6, seed=0x5EEDB005)
    s, scenes = w.generate()
    v = velocities(len(s["x"]), 5)
    sets = []
    for j, side in enumerate((0.1, 0.08)):
        h = H.generate_hulls(0x4C00 + j, 300, lo=(0.0, 0.0, 0.0), hi=(side, side, side), v_sigma=0.3)
        sets.append(HB.concat([HB.make_bodies(h, 0x4C10 + j), HB.box_walls((0, 0, 0), (side, side, side))])[0])
    bodies, first = HB.concat(sets)
    c = L.ContactParams.make(2e4, 0.3, 0.5, w.dt, gravity=(0.0, 0.0, -9.81), walls=None, max_contacts=16)
    a = batch_of(w, scenes, s, v)
    ha = HB.HullBodyScene(c)
    ha.upload(bodies, first)
    a.attach_hulls(ha)
    b = batch_of(w, scenes, s, v)
    hb = HB.HullBodyScene(c)
    hb.upload(bodies, first)
    for k in (1, 6, 13):
        a.step(k)
        b.step(k)
        hb.step(k)
    assert_same(a.download(), b.download(), "spheres with bodies attached")
    ga, gb = ha.download(), hb.download()
    for key in ("pos4", "vel4", "quat4", "angm4"):
        assert np.array_equal(ga[key], gb[key]), key
    assert len(ha.contacts()[0]) > 50
    ha.set_warm(warm_model)  # while attached
    hb.set_warm(warm_model)
    for k in (1, 5):
        a.step(k)
        b.step(k)
        hb.step(k)
    assert ha.narrowphase_stats()["warm_used"] > 0
    assert_same(a.download(), b.download(), "spheres with warm-started bodies attached")
    ga, gb = ha.download(), hb.download()
    for key in ("pos4", "vel4", "quat4", "angm4"):
        assert np.array_equal(ga[key], gb[key]), "warm " + key
    a.attach_hulls(None)
    a.step(3)
    b.step(3)
    assert_same(a.download(), b.download(), "after detaching")
    assert np.array_equal(ha.download()["pos4"], ga["pos4"])  # detached bodies no longer move
    for x in (a, b):
        x.close()
    ha.close()
    hb.close()


def test_persistent_step_kernel_changes_no_result():
    """sd_ctx_set_persistent: the persistent form of the fused step kernel
    (resident workgroups claiming 64-slot tiles per XCD) gives the default
    form's results bit for bit, batched and single-scene, graph and eager."""
    say("persistent step kernel")
    w = small_batch(nscenes=8, seed=0x5EEDB006)
    s, scenes = w.generate()
    v = velocities(len(s["x"]), 6)
    for graph in (True, False):
        a = batch_of(w, scenes, s, v, use_graph=graph)
        b = batch_of(w, scenes, s, v, use_graph=graph)
        b.set_persistent(True)
        for k in (1, 7, 30):
            a.step(k)
            b.step(k)
            assert_same(a.download(), b.download(), "batch, graph=%s, after %d" % (graph, k))
        b.set_persistent(False)
        a.step(5)
        b.step(5)
        assert_same(a.download(), b.download(), "back to the default form")
        a.close()
        b.close()
    n, g, c = scenes[0]
    x = solo(scenes[0], s, v, 0, n)
    y = solo(scenes[0], s, v, 0, n)
    y.set_persistent(True)
    x.step(40)
    y.step(40)
    assert_same(x.download(), y.download(), "single scene")
    x.close()
    y.close()
