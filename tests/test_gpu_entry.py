"""GPU parity of the per-kernel entry points that the fused steps do not
call: the segmented sort (against numpy's stable lexsort), the per-body force
reduction and the rigid-body integration (against the oracle, bit for bit),
and the FP32 MLP forward pass (against the float64 oracle)."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as M
from stardust_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    L.require_gpu()


def dbuf(a):
    return L.DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.mark.parametrize("n,nseg,bits", [(10, 3, 8), (100000, 37, 17), (1 << 20, 512, 24), (5000, 1, 32)])
def test_segmented_sort_pairs(n, nseg, bits):
    say("segmented sort", n, nseg, bits)
    rng = np.random.default_rng(n + nseg)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 5] = keys[0]  # equal keys: stability
    vals = rng.permutation(n).astype(np.uint32)
    cuts = np.sort(rng.integers(0, n + 1, nseg - 1))
    first = np.concatenate([[0], cuts, [n]]).astype(np.int32)  # empty segments allowed
    seg = np.repeat(np.arange(nseg), np.diff(first))
    order = np.lexsort((keys, seg))  # stable: by segment, then key, then input position
    k, v = dbuf(keys), dbuf(vals)
    ka, va = L.DeviceBuffer(4 * n), L.DeviceBuffer(4 * n)
    f = dbuf(first)
    nb = C.c_size_t()
    L.call("sd_segmented_sort_temp_bytes", n, C.byref(nb))
    tmp = L.DeviceBuffer(nb.value)
    L.call("sd_segmented_sort_pairs_u32", k.ptr, v.ptr, ka.ptr, va.ptr, n, f.ptr, nseg, bits, tmp.ptr, nb.value, None)
    L.synchronize()
    assert np.array_equal(k.download(np.uint32, n), keys[order])
    assert np.array_equal(v.download(np.uint32, n), vals[order])


def test_body_reduce_matches_oracle(oracle):
    say("body reduce")
    rng = np.random.default_rng(3)
    nb = 20000
    count = rng.integers(1, 17, nb).astype(np.int32)
    first = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int32)
    ns = int(count.sum())
    owner = np.repeat(np.arange(nb), count)
    bpos = np.zeros((nb, 4), np.float32)
    bpos[:, :3] = rng.uniform(0, 1, (nb, 3))
    spos = np.zeros((ns, 4), np.float32)
    spos[:, :3] = bpos[owner, :3] + rng.normal(0, 0.01, (ns, 3))
    sf = np.zeros((ns, 4), np.float32)
    st = np.zeros((ns, 4), np.float32)
    sf[:, :3] = rng.normal(0, 5, (ns, 3))
    st[:, :3] = rng.normal(0, 0.05, (ns, 3))
    ef, et = oracle.body_reduce(bpos, first, count, spos, sf, st)
    bf, bt = L.DeviceBuffer(16 * nb), L.DeviceBuffer(16 * nb)
    ins = [dbuf(a) for a in (bpos, first, count, spos, sf, st)]  # held until the kernel has run
    L.call("sd_body_reduce", ins[0].ptr, ins[1].ptr, ins[2].ptr, nb, ins[3].ptr, ins[4].ptr, ins[5].ptr, bf.ptr,
           bt.ptr, None)
    L.synchronize()
    assert np.array_equal(bf.download(np.float32, 4 * nb).reshape(-1, 4), ef)
    assert np.array_equal(bt.download(np.float32, 4 * nb).reshape(-1, 4), et)


def random_bodies(rng, nb, wscale):
    P = np.zeros((nb, 4), np.float32)
    P[:, :3] = rng.uniform(0, 1, (nb, 3))
    P[:, 3] = rng.uniform(0.5, 2.0, nb)
    P[::13, 3] = 0.0  # static
    V = np.zeros((nb, 4), np.float32)
    V[:, :3] = rng.normal(0, 1, (nb, 3))
    q = rng.normal(0, 1, (nb, 4))
    Q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    L_ = np.zeros((nb, 4), np.float32)
    L_[:, :3] = rng.normal(0, wscale, (nb, 3))
    inv6 = np.zeros((nb, 6), np.float32)
    inv6[:, :3] = rng.uniform(50, 200, (nb, 3))
    inv6[:, 3:] = rng.uniform(-10, 10, (nb, 3))
    f = np.zeros((nb, 4), np.float32)
    t = np.zeros((nb, 4), np.float32)
    f[:, :3] = rng.normal(0, 3, (nb, 3))
    t[:, :3] = rng.normal(0, 0.01, (nb, 3))
    return P, V, Q, L_, inv6, f, t


@pytest.mark.parametrize("wscale,exact", [(0.01, True), (2.0, False)])
def test_integrate_bodies_matches_oracle(oracle, wscale, exact):
    """Small rotation angles (below 0.25 rad per half step: the shared Taylor
    polynomials) give the oracle's bits; larger ones use the device's sinf/cosf
    against the host's (within 1e-6)."""
    say("integrate bodies", wscale)
    rng = np.random.default_rng(17)
    nb = 10000
    P, V, Q, Lm, inv6, f, t = random_bodies(rng, nb, wscale)
    c = L.ContactParams.make(1e5, 0.3, 0.5, 1e-3, gravity=(0.0, 0.0, -9.81), walls=None)
    eP, eV, eQ, eL = oracle.integrate_bodies(P, V, Q, Lm, inv6, f, t, c)
    d = [dbuf(a) for a in (P, V, Q, Lm)]
    ins = [dbuf(a) for a in (inv6, f, t)]  # held until the kernel has run
    L.call("sd_integrate_bodies", d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, ins[0].ptr, ins[1].ptr, ins[2].ptr, nb,
           C.byref(c), None)
    L.synchronize()
    got = [x.download(np.float32, 4 * nb).reshape(-1, 4) for x in d]
    for g, e, name in zip(got, (eP, eV, eQ, eL), "PVQL"):
        if exact:
            assert np.array_equal(g, e), name
        else:
            assert np.abs(g - e).max() <= 1e-6 * max(1.0, np.abs(e).max()), name


def test_mlp_forward_fp32_matches_float64():
    say("mlp forward fp32")
    rng = np.random.default_rng(5)
    w = rng.normal(0, 0.25, M.NPARAM).astype(np.float32)
    n = 5000
    X = rng.normal(0, 1, (n, M.NIN)).astype(np.float32)
    out = L.DeviceBuffer(12 * n)
    xd, wd = dbuf(X), dbuf(w)  # held until the kernel has run
    L.call("sd_mlp_forward", xd.ptr, n, wd.ptr, out.ptr, None)
    L.synchronize()
    got = out.download(np.float32, 3 * n).reshape(-1, 3)
    y = M.forward(w, X)[0]
    ref = y / np.linalg.norm(y, axis=1, keepdims=True)
    assert np.abs(got - ref).max() <= 1e-5
