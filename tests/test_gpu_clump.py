"""GPU parity of the LBVH broadphase and the rigid-clump step against the oracle.

Integer work (Morton keys, tree topology, pair sets) is checked exactly; the
clump state is checked bit-exact against the FP32 oracle, which uses the same
operation order (no contraction) and the same polynomial sin/cos below 0.25 rad
per half step.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from stardust_amd import _lib as L
from stardust_amd import clumps as CL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_boxes(rng, n, extent=1.0, size=0.02):
    c = rng.uniform(0, extent, (n, 3)).astype(np.float32)
    h = rng.uniform(0.2, 1.0, (n, 3)).astype(np.float32) * np.float32(size)
    lo = np.zeros((n, 4), np.float32)
    hi = np.zeros((n, 4), np.float32)
    lo[:, :3] = c - h
    hi[:, :3] = c + h
    return lo, hi


class Lbvh:
    """Runs domain, Morton, sort (oracle order: keys are unique), build and query through the C ABI."""

    def __init__(self, lo, hi):
        n = self.n = len(lo)
        self.dlo, self.dhi = L.DeviceBuffer.from_array(lo), L.DeviceBuffer.from_array(hi)
        self.dom = L.DeviceBuffer(32)
        self.scr = L.DeviceBuffer(256 * 6 * 4)
        self.keys = L.DeviceBuffer(8 * n)
        nn = 2 * n - 1
        self.buf = {k: L.DeviceBuffer(4 * max(nn, 1)) for k in ("left", "right", "parent", "first", "last", "rope")}
        self.buf["counters"] = L.DeviceBuffer(4 * (3 * n + 1))
        self.nlo, self.nhi = L.DeviceBuffer(16 * nn), L.DeviceBuffer(16 * nn)
        self.err = L.DeviceBuffer.from_array(np.zeros(4, np.int32))

    def domain(self):
        L.call("sd_global_domain", self.dlo.ptr, self.dhi.ptr, self.n, self.dom.ptr, self.scr.ptr, None)
        return self.dom.download(np.float32, 6)

    def morton(self):
        L.call("sd_morton30", self.dlo.ptr, self.dhi.ptr, self.n, self.dom.ptr, self.keys.ptr, None)
        return self.keys.download(np.uint64, self.n)

    def sort(self):
        keys = self.keys.download(np.uint64, self.n)
        alt = L.DeviceBuffer(8 * self.n)
        tb = C.c_size_t()
        L.call("sd_radix_sort_temp_bytes", self.n, C.byref(tb))
        tmp = L.DeviceBuffer(tb.value)
        L.call("sd_radix_sort_keys_u64", self.keys.ptr, alt.ptr, self.n, 32, 62, tmp.ptr, tb.value, None)
        L.synchronize()
        out = self.keys.download(np.uint64, self.n)
        assert np.array_equal(out, np.sort(keys))
        return out

    def build(self):
        b = self.buf
        L.call("sd_lbvh_build", self.keys.ptr, self.n, self.dlo.ptr, self.dhi.ptr, b["left"].ptr, b["right"].ptr,
               b["parent"].ptr, b["first"].ptr, b["last"].ptr, b["rope"].ptr, self.nlo.ptr, self.nhi.ptr,
               b["counters"].ptr, self.err.ptr, None)
        nn = 2 * self.n - 1
        t = {k: b[k].download(np.int32, nn) for k in ("left", "right", "parent", "first", "last", "rope")}
        t["left"], t["right"] = t["left"][:self.n - 1], t["right"][:self.n - 1]
        t["nlo"] = self.nlo.download(np.float32, 4 * nn).reshape(nn, 4)
        t["nhi"] = self.nhi.download(np.float32, 4 * nn).reshape(nn, 4)
        t["error"] = int(self.err.download(np.int32, 1)[0])
        return t

    def query(self, capacity=None):
        sb = C.c_size_t()
        L.call("sd_lbvh_query_scratch_bytes", self.n, C.byref(sb))
        scr = L.DeviceBuffer(sb.value)
        cap = 8 * self.n if capacity is None else capacity
        pairs = L.DeviceBuffer(8 * max(cap, 1))
        cnt = C.c_int()
        b = self.buf
        code = L.load().sd_lbvh_query(self.keys.ptr, self.n, b["left"].ptr, b["last"].ptr, b["rope"].ptr,
                                      self.nlo.ptr, self.nhi.ptr, pairs.ptr, cap, C.byref(cnt), self.err.ptr,
                                      scr.ptr, sb.value, None)
        if code != L.SD_OK:
            return code, cnt.value, None
        return code, cnt.value, pairs.download(np.int32, 2 * cnt.value).reshape(-1, 2)


def check_tree(t, keys, lo, hi):
    """Structural invariants of a Karras tree with ropes and refit boxes."""
    n = len(keys)
    assert t["error"] == 0
    nn = 2 * n - 1
    left, right, parent, first, last, rope = (t[k] for k in ("left", "right", "parent", "first", "last", "rope"))
    if n > 1:
        kids = np.concatenate([left, right])
        assert np.array_equal(np.sort(kids), np.arange(1, nn)), "every non-root node is a child exactly once"
        assert parent[0] == -1
        assert np.array_equal(parent[left], np.arange(n - 1)) and np.array_equal(parent[right], np.arange(n - 1))
        # leaf ranges are contiguous and split at the children
        assert np.array_equal(first[:n - 1], first[left]) and np.array_equal(last[:n - 1], last[right])
        assert np.array_equal(last[left] + 1, first[right])
        assert first[0] == 0 and last[0] == n - 1
        # Karras split: the highest bit in which the node's keys differ is the one that
        # changes between the last key of the left child and the first of the right
        kl = [int(x) for x in keys]
        for i in range(n - 1):
            g = int(last[left[i]])
            assert (kl[first[i]] ^ kl[last[i]]).bit_length() == (kl[g] ^ kl[g + 1]).bit_length()
    ids = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.array_equal(first[n - 1:], np.arange(n)) and np.array_equal(last[n - 1:], np.arange(n))
    assert np.array_equal(t["nlo"][n - 1:, :3], lo[ids, :3]) and np.array_equal(t["nhi"][n - 1:, :3], hi[ids, :3])
    if n > 1:
        assert np.array_equal(t["nlo"][:n - 1, :3], np.minimum(t["nlo"][left, :3], t["nlo"][right, :3]))
        assert np.array_equal(t["nhi"][:n - 1, :3], np.maximum(t["nhi"][left, :3], t["nhi"][right, :3]))
    # ropes: the node after each subtree in a depth-first (left first) walk
    order, nxt = [], np.full(nn, -1, np.int64)
    stack = [0]
    while stack:
        v = stack.pop()
        order.append(v)
        if v < n - 1:
            stack.append(int(right[v]))
            stack.append(int(left[v]))
    pos = {v: i for i, v in enumerate(order)}
    size = np.ones(nn, np.int64)
    for v in reversed(order):
        if v < n - 1:
            size[v] = 1 + size[left[v]] + size[right[v]]
    for v in range(nn):
        j = pos[v] + size[v]
        nxt[v] = order[j] if j < nn else -1
    assert np.array_equal(rope, nxt)


def pair_set(p):
    return set(map(tuple, np.asarray(p).tolist()))


@pytest.mark.parametrize("n,case", [(1, "single"), (2, "two"), (3, "three"), (1000, "random"), (65537, "random"),
                                    (512, "identical"), (4096, "lattice")])
def test_lbvh_tree_and_pairs(oracle, n, case):
    say("lbvh case", case, n)
    L.require_gpu()
    rng = np.random.default_rng(n + 17)
    if case == "identical":
        lo, hi = random_boxes(rng, n)
        lo[:, :3] = np.float32(0.4)
        hi[:, :3] = np.float32(0.6)
    elif case == "lattice":
        g = np.stack(np.meshgrid(*[np.arange(16)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
        c = g * np.float32(0.05)
        lo = np.zeros((n, 4), np.float32)
        hi = np.zeros((n, 4), np.float32)
        lo[:, :3] = c - np.float32(0.025)  # touching faces count as overlap
        hi[:, :3] = c + np.float32(0.025)
    else:
        lo, hi = random_boxes(rng, n, size=0.02 if n > 100 else 0.4)
    bv = Lbvh(lo, hi)
    dom = bv.domain()
    assert np.array_equal(dom, oracle.global_domain(lo, hi))
    keys = bv.morton()
    assert np.array_equal(keys, oracle.morton30(lo, hi, dom))
    skeys = bv.sort()
    t = bv.build()
    check_tree(t, skeys, lo, hi)
    want = oracle.brute_aabb_pairs(lo, hi) if n <= 5000 else oracle.sap_pairs(lo, hi)[0]
    code, cnt, pairs = bv.query(capacity=len(want) + 16)
    assert code == L.SD_OK
    assert cnt == len(want)
    assert np.all(pairs[:, 0] < pairs[:, 1])
    assert pair_set(pairs) == pair_set(want) and len(pair_set(pairs)) == cnt
    if case == "identical":
        assert cnt == n * (n - 1) // 2


def test_lbvh_many_crossing_nodes(oracle):
    """1,048,576 random boxes: each of the 4,095 refit chunk boundaries has its
    own crossing node (the split between its two leaves), more than the 2,048
    the refit finishes in LDS, so the top of the tree is finished by the
    global-memory walk; the smaller trees of the other tests finish it in LDS."""
    say("lbvh many crossing nodes")
    L.require_gpu()
    rng = np.random.default_rng(91)
    n = 1 << 20
    lo, hi = random_boxes(rng, n, size=0.001)
    bv = Lbvh(lo, hi)
    bv.domain()
    bv.morton()
    skeys = bv.sort()
    t = bv.build()
    check_tree(t, skeys, lo, hi)
    want = oracle.sap_pairs(lo, hi)[0]
    code, cnt, pairs = bv.query(capacity=len(want) + 16)
    assert code == L.SD_OK and cnt == len(want)
    assert pair_set(pairs) == pair_set(want)


def test_lbvh_query_capacity_error(oracle):
    say("lbvh capacity")
    L.require_gpu()
    rng = np.random.default_rng(5)
    lo, hi = random_boxes(rng, 2000, size=0.05)
    bv = Lbvh(lo, hi)
    bv.domain()
    bv.morton()
    bv.sort()
    bv.build()
    code, cnt, _ = bv.query(capacity=10)
    assert code == L.SD_ERR_CAPACITY and cnt == len(oracle.brute_aabb_pairs(lo, hi)) > 10
    code, cnt2, pairs = bv.query(capacity=cnt)
    assert code == L.SD_OK and cnt2 == cnt


def clump_set(nb, seed, static_every=0, v_sigma=0.5, spacing=0.035):
    lat = int(round(nb ** (1 / 3)))
    s = CL.generate_clumps(seed, nb, lattice=lat, spacing=spacing, jitter=0.002, v_sigma=v_sigma)
    if static_every:
        s["pos4"][::static_every, 3] = 0.0
    return s


def compare_state(g, o, what):
    for key, arr in (("pos4", o.P), ("vel4", o.V), ("quat4", o.Q), ("angm4", o.L)):
        a, b = g[key][:, :3 if key != "quat4" else 4], arr[:, :3 if key != "quat4" else 4]
        bad = np.nonzero(np.any(a != b, axis=1))[0]
        assert len(bad) == 0, "%s: %s differs at %d bodies, first %d: %s vs %s" % (
            what, key, len(bad), bad[0], a[bad[0]], b[bad[0]])


@pytest.mark.parametrize("nb,static_every", [(4096, 0), (1000, 7)])
def test_clump_steps_bit_exact(oracle, nb, static_every):
    say("clump steps", nb, static_every)
    s = clump_set(nb, 0xC1 + nb, static_every)
    c = CL.contact_c3()
    o = oracle.ClumpState(s)
    sx0, sv0, lo0, hi0 = o.pose()
    sc = CL.ClumpScene(c)
    sc.upload(s)
    sc.step(1)
    assert np.array_equal(sc.spheres(), sx0)
    assert pair_set(sc.pairs()) == pair_set(oracle.brute_aabb_pairs(lo0, hi0))
    o.step(c, 1)
    compare_state(sc.download(), o, "step 1")
    total = 1
    for k in (4, 15):
        sc.step(k)
        o.step(c, k)
        total += k
        compare_state(sc.download(), o, "step %d" % total)
    assert len(sc.pairs()) > nb // 4  # the bodies are in contact, not just free flight


def test_clump_c3_full_size(oracle):
    say("clump c3 full size")
    s = CL.clumps_c3()
    c = CL.contact_c3()
    o = oracle.ClumpState(s)
    sx0, sv0, lo0, hi0 = o.pose()
    sc = CL.ClumpScene(c)
    sc.upload(s)
    sc.step(1)
    gp = sc.pairs()
    want, _ = oracle.sap_pairs(lo0, hi0)
    assert len(gp) == len(want) and pair_set(gp) == pair_set(want)
    o.step(c, 1)
    compare_state(sc.download(), o, "c3 step 1")
    # momentum: contact forces cancel pairwise, so total momentum changes by M g dt per step
    g0 = sc.download()
    m = np.where(s["pos4"][:, 3] > 0, s["pos4"][:, 3], 0).astype(np.float64)
    p0 = (m[:, None] * g0["vel4"][:, :3]).sum(0)
    sc.step(20)
    g1 = sc.download()
    p1 = (m[:, None] * g1["vel4"][:, :3]).sum(0)
    dp = p1 - p0 - m.sum() * np.array([0.0, 0.0, -9.81]) * 1e-5 * 20
    scale = (m[:, None] * np.abs(g1["vel4"][:, :3])).sum()
    assert np.all(np.abs(dp) < 1e-5 * scale), dp


def test_clump_upload_rejects_bad_ranges():
    say("clump upload errors")
    s = clump_set(64, 3)
    sc = CL.ClumpScene(CL.contact_c3())
    bad = dict(s)
    bad["count"] = s["count"].copy()
    bad["count"][5] += 1
    with pytest.raises(L.SdError):
        sc.upload(bad)
    bad = dict(s)
    bad["inertia6"] = np.zeros_like(s["inertia6"])
    with pytest.raises(L.SdError):
        sc.upload(bad)
    sc.upload(s)
    sc.step(2)


def test_clump_steps_crowded_bit_exact(oracle):
    """Bodies 12 mm apart with AABBs about 6 cm wide: each has over a hundred
    candidate partners, so the step's partner search takes several passes of
    its 32-entry list; every step equals the oracle bit for bit."""
    say("clump steps crowded")
    s = clump_set(729, 0xC7, spacing=0.012, v_sigma=0.05)
    c = CL.contact_c3()
    o = oracle.ClumpState(s)
    sx0, sv0, lo0, hi0 = o.pose()
    per = np.bincount(oracle.brute_aabb_pairs(lo0, hi0).ravel(), minlength=729)
    assert per.max() > 3 * 32
    sc = CL.ClumpScene(c)
    sc.upload(s)
    total = 0
    for k in (1, 2, 3):
        sc.step(k)
        o.step(c, k)
        total += k
        compare_state(sc.download(), o, "step %d" % total)
    sc.close()


def test_clump_checkpoint_resume_and_body_check(oracle):
    """sd_clump_save / sd_clump_load: a context resumed from a checkpoint steps
    bit-identically to the uninterrupted one; a checkpoint of other bodies
    (same counts, one sphere offset changed) is rejected."""
    say("clump checkpoint")
    s = clump_set(1000, 0xC9, static_every=11)
    c = CL.contact_c3()
    a = CL.ClumpScene(c)
    a.upload(s)
    a.step(7)
    buf = a.save()
    a.step(9)
    ref = a.download()
    b = CL.ClumpScene(c)
    b.upload(s)
    b.load(buf)
    b.step(9)
    got = b.download()
    for k in ("pos4", "vel4", "quat4", "angm4"):
        assert np.array_equal(got[k], ref[k]), k
    other = dict(s)
    other["local4"] = s["local4"].copy()
    other["local4"][3, 0] += 1e-4
    d = CL.ClumpScene(c)
    d.upload(other)
    with pytest.raises(L.SdError, match="other bodies"):
        d.load(buf)
    for x in (a, b, d):
        x.close()


def test_clump_remove_then_add_matches_oracle(oracle):
    """Runtime removal and addition of bodies (sd_clump_remove_bodies,
    sd_clump_add_bodies) against the oracle's rule, bit for bit, in contact."""
    say("clump remove / add")
    s = clump_set(1000, 0xCA, static_every=13)
    c = CL.contact_c3()
    o = oracle.ClumpState(s)
    sc = CL.ClumpScene(c)
    sc.upload(s)
    sc.step(5)
    o.step(c, 5)
    compare_state(sc.download(), o, "before")
    rng = np.random.default_rng(5)
    gone = rng.choice(1000, 77, replace=False)
    sc.remove(gone)
    o.remove(gone)
    sc.step(6)
    o.step(c, 6)
    compare_state(sc.download(), o, "after remove")
    extra = CL.generate_clumps(0xCB, 27, lattice=3, spacing=0.035, jitter=0.002, v_sigma=0.5)
    extra["pos4"][:, :3] += np.float32(0.012)  # among the others: contacts with them
    ids = sc.add(extra)
    o.add(extra)
    assert ids.tolist() == list(range(923, 950))
    sc.step(7)
    o.step(c, 7)
    compare_state(sc.download(), o, "after add")
    assert len(sc.pairs()) > 200
    sc.close()
