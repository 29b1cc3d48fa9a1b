"""GPU checks of the learned warm start: features, gradients and Adam against
the float64 oracle, BF16 MFMA inference against a BF16 emulation, and the
effect of a trained network on GJK iteration counts (plain GJK as the judge)."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as M
from stardust_amd import _lib as L
from stardust_amd import hulls as H
from stardust_amd import warmstart as WS

pytestmark = pytest.mark.gpu


def rel_diff(a, b):
    """|a - b| / |b| per element; 0 where the two are equal (both zero included)."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.where(a == b, 0.0, np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    L.require_gpu()


def test_features_match_oracle():
    say("mlp features")
    h = H.generate_hulls(9, 3000)
    d = H.DeviceHulls(h)
    d.aabbs()
    pairs, npairs, _ = d.sap()
    X, _ = WS.pair_features(d, pairs, npairs)
    L.synchronize()
    got = X.download(np.float32, 22 * npairs).reshape(-1, 22)
    hp = pairs.download(np.int32, 2 * npairs).reshape(-1, 2)
    sel = np.arange(0, npairs, max(1, npairs // 300))
    exp = M.features(h, hp[sel])
    assert np.allclose(got[sel], exp, atol=2e-5, rtol=1e-4)


def rand_data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 22)).astype(np.float32)
    T = rng.normal(0, 1, (n, 3))
    T = (T / np.linalg.norm(T, axis=1, keepdims=True)).astype(np.float32)
    return X, T


def test_gradient_and_adam_match_oracle():
    say("mlp gradient")
    model = WS.WarmStartModel(seed=7)
    w0 = model.params()
    assert np.all(np.isfinite(w0)) and np.abs(w0).max() > 0
    for batch in (64, 1000):
        X, T = rand_data(batch, seed=batch)
        data = WS.PairData(batch)
        data.X.upload(X)
        data.T.upload(T)
        data.n = batch
        idx = L.DeviceBuffer.from_array(np.arange(batch, dtype=np.int32)[::-1].copy())
        g, loss = model.gradient(data, idx, batch)
        Lr, gr = M.loss_and_grad(w0, X[::-1], T[::-1])
        assert abs(loss - Lr) < 1e-5
        scale = np.abs(gr).max()
        assert np.abs(g - gr).max() < 2e-5 * scale + 1e-7, np.abs(g - gr).max() / scale
    # Adam step
    m0 = np.zeros(model.np)
    w1_ref, _, _ = M.adam(w0.astype(np.float64), g.astype(np.float64), m0, m0.copy(), 1, 1e-3)
    L.call("sd_adam_step", model.w.ptr, model.g.ptr, model.m.ptr, model.v.ptr, model.np, 1e-3, 0.9, 0.999, 1e-8, 1,
           None)
    w1 = model.params()
    assert np.allclose(w1, w1_ref, atol=1e-6)


def test_bf16_mfma_inference_matches_emulation():
    say("mlp bf16 mfma")
    model = WS.WarmStartModel(seed=3)
    rng = np.random.default_rng(1)
    w = (rng.normal(0, 0.25, model.np)).astype(np.float32)
    model.set_params(w)
    model.pack()
    for n in (1, 17, 1000):
        X, _ = rand_data(n, seed=n)
        Xd = L.DeviceBuffer.from_array(X)
        out = model.infer(Xd, n)
        L.synchronize()
        got = out.download(np.float32, 3 * n).reshape(-1, 3)
        exp = M.forward_bf16(w, X)
        cos = (got * exp).sum(1)
        assert cos.min() > 0.999, cos.min()
        full = M.forward(w, X)[0]
        full /= np.linalg.norm(full, axis=1, keepdims=True)
        assert (got * full).sum(1).mean() > 0.99


def test_trained_network_reduces_gjk_iterations(oracle):
    """The trained network, as the warm start of full-size configuration-4
    queries (131,072 hulls, 445,828 pairs): identical verdicts and depths within
    1e-6 relative of plain cold GJK on every pair, bit-identical to the
    oracle's warm query from the same directions, and fewer iterations than
    plain GJK. The network's own share of the gain is reported against the
    same seeded query started from the centre difference."""
    say("mlp train + warm gjk at full C4")
    data = WS.PairData(3_000_000)
    seed = 0x5EED1000
    while data.n < 2_500_000:
        data.append_scene(H.generate_hulls(seed, 131072))
        seed += 1
    model = WS.WarmStartModel(seed=11)
    hist = model.train(data, epochs=8, batch=8192, lr=3e-3, lr_final=3e-4, log=lambda *a: say(*a))
    assert hist[-1] < hist[0] and hist[-1] < 0.25, hist
    # held-out scene: the full configuration-4 set
    h = H.hulls_c4()
    d = H.DeviceHulls(h)
    d.aabbs()
    pairs, npairs, _ = d.sap()
    X, _ = WS.pair_features(d, pairs, npairs)
    warm = model.infer(X, npairs, pairs=pairs, quat=d.buf["quat4"])
    cold = d.download_gjk(d.gjk(pairs, npairs), npairs)
    ws = d.download_gjk(d.gjk(pairs, npairs, warm=warm, warm_budget=16), npairs)
    L.synchronize()
    wdir = warm.download(np.float32, 3 * npairs).reshape(-1, 3)
    hp = pairs.download(np.int32, 2 * npairs).reshape(-1, 2)
    # every verdict identical, every depth within 1e-6 relative (no floor)
    assert np.array_equal(cold["flags"] & 1, ws["flags"] & 1)
    diff = rel_diff(ws["sep"], cold["sep"])
    assert diff.max() <= 1e-6, diff.max()
    # the GPU's warm queries equal the oracle's from the same directions
    exp = oracle.gjk(h, hp, warm=wdir, warm_budget=16, nthreads=16)
    assert np.array_equal(ws["flags"], exp["flags"]) and np.array_equal(ws["iters"], exp["iters"])
    assert np.array_equal(ws["sep"].view(np.uint32), exp["sep"].view(np.uint32))
    sep = (cold["flags"] & 1) == 0
    used = float(((ws["flags"] & L.GJK_WARM_USED) != 0).mean())
    say("pairs %d: iterations cold %.3f warm %.3f (separated: %.3f -> %.3f), warm kept %.4f, bit-identical %.5f, "
        "max rel diff %.2e" % (npairs, cold["iters"].mean(), ws["iters"].mean(), cold["iters"][sep].mean(),
                               ws["iters"][sep].mean(), used, float((ws["sep"] == cold["sep"]).mean()), diff.max()))
    say("cold histogram", np.bincount(cold["iters"])[:30].tolist())
    say("warm histogram", np.bincount(ws["iters"])[:30].tolist())
    assert used > 0.95
    assert ws["iters"].mean() < 0.85 * cold["iters"].mean()
    assert ws["iters"][sep].mean() < 0.75 * cold["iters"][sep].mean()
    # what the network adds over its input pose: its directions are closer to the
    # converged normals than the centre difference is
    dc = (h["pos4"][hp[:, 1], :3] - h["pos4"][hp[:, 0], :3]).astype(np.float64)
    dc /= np.linalg.norm(dc, axis=1, keepdims=True)
    ang_net = np.degrees(np.arccos(np.clip((wdir * cold["normal"]).sum(1), -1, 1)))
    ang_dc = np.degrees(np.arccos(np.clip((dc * cold["normal"]).sum(1), -1, 1)))
    say("mean angle to the converged normal: network %.2f deg, centre difference %.2f deg" % (ang_net.mean(),
                                                                                                ang_dc.mean()))
    assert ang_net.mean() < ang_dc.mean()
    cdir = L.DeviceBuffer.from_array(np.ascontiguousarray(dc.astype(np.float32)))
    cs = d.download_gjk(d.gjk(pairs, npairs, warm=cdir, warm_budget=16), npairs)
    say("same seeded query from the centre difference: iterations %.3f (network %.3f)" % (cs["iters"].mean(),
                                                                                          ws["iters"].mean()))
