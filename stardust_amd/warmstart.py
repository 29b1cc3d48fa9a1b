"""Learned GJK warm start: data logging from plain GJK, training with the
library's forward/backward/Adam kernels, BF16 MFMA inference.

Training data are (pose, converged direction) pairs: every sweep-and-prune pair
of seeded hull scenes (the configuration-4 generator with consecutive seeds) is
solved by plain cold-start GJK/EPA and its converged normal, expressed in the
first hull's frame, is the target of the pair's 22 features.
"""
import ctypes as C
import time

import numpy as np

from . import _lib as L
from . import hulls as H

NIN = 22


def n_params():
    return L.load().sd_mlp_param_count()


class PairData:
    """Features X [n x 22] and targets T [n x 3] on the device."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.n = 0
        self.X = L.DeviceBuffer(4 * NIN * capacity)
        self.T = L.DeviceBuffer(4 * 3 * capacity)

    def append_scene(self, h, max_iter=H.MAX_ITER):
        """Solves every SAP pair of hull set h with plain GJK and appends the samples."""
        d = H.DeviceHulls(h)
        d.aabbs()
        pairs, npairs, _ = d.sap()
        out = d.gjk(pairs, npairs, max_iter=max_iter)
        take = min(npairs, self.capacity - self.n)
        if take <= 0:
            return 0
        desc = L.DeviceBuffer(4 * 7 * d.n)
        b = d.buf
        L.call("sd_hull_descriptors", b["pos4"].ptr, b["vstart"].ptr, b["nv"].ptr, b["verts4"].ptr, d.n, desc.ptr,
               None)
        xo = C.c_void_p(self.X.ptr.value + 4 * NIN * self.n)
        to = C.c_void_p(self.T.ptr.value + 4 * 3 * self.n)
        L.call("sd_mlp_features", pairs.ptr, take, b["pos4"].ptr, b["quat4"].ptr, desc.ptr, xo, None)
        L.call("sd_mlp_targets", pairs.ptr, take, b["quat4"].ptr, out["normal"].ptr, to, None)
        L.synchronize()
        self.n += take
        return take


class WarmStartModel:
    def __init__(self, seed=0x5EED00AA):
        L.require_gpu()
        self.np = n_params()
        self.w = L.DeviceBuffer(4 * self.np)
        self.g = L.DeviceBuffer(4 * self.np)
        self.m = L.DeviceBuffer(4 * self.np)
        self.v = L.DeviceBuffer(4 * self.np)
        L.call("sd_mlp_init", self.w.ptr, seed, None)
        L.call("sd_memset_u32", self.m.ptr, 0, self.np, None)
        L.call("sd_memset_u32", self.v.ptr, 0, self.np, None)
        self.step = 0
        self.packed = L.DeviceBuffer(L.load().sd_mlp_packed_bytes())
        self.loss = L.DeviceBuffer(4)

    def params(self):
        L.synchronize()
        return self.w.download(np.float32, self.np)

    def set_params(self, w):
        self.w.upload(np.ascontiguousarray(w, np.float32))

    def gradient(self, data, idx_dev, batch, scratch=None):
        nb = C.c_size_t()
        L.call("sd_mlp_train_scratch_bytes", batch, C.byref(nb))
        scratch = scratch or L.DeviceBuffer(nb.value)
        L.call("sd_mlp_forward_backward", data.X.ptr, data.T.ptr, idx_dev.ptr, batch, self.w.ptr, self.g.ptr,
               self.loss.ptr, scratch.ptr, nb.value, None)
        L.synchronize()
        return self.g.download(np.float32, self.np), float(self.loss.download(np.float32, 1)[0])

    def train(self, data, epochs=1, batch=8192, lr=3e-3, seed=1, log=None, lr_final=None):
        """Adam over shuffled mini-batches; returns the mean loss of each epoch.
        lr_final: the learning rate decays geometrically from lr to lr_final over
        the epochs (constant when None)."""
        nb = C.c_size_t()
        L.call("sd_mlp_train_scratch_bytes", batch, C.byref(nb))
        scratch = L.DeviceBuffer(nb.value)
        rng = np.random.default_rng(seed)
        idx = L.DeviceBuffer(4 * data.n)
        history = []
        for ep in range(epochs):
            t0 = time.perf_counter()
            lr_ep = lr if lr_final is None or epochs < 2 else lr * (lr_final / lr) ** (ep / (epochs - 1))
            idx.upload(rng.permutation(data.n).astype(np.int32))
            losses = []
            for b0 in range(0, data.n - batch + 1, batch):
                sub = C.c_void_p(idx.ptr.value + 4 * b0)
                L.call("sd_mlp_forward_backward", data.X.ptr, data.T.ptr, sub, batch, self.w.ptr, self.g.ptr,
                       self.loss.ptr, scratch.ptr, nb.value, None)
                self.step += 1
                L.call("sd_adam_step", self.w.ptr, self.g.ptr, self.m.ptr, self.v.ptr, self.np, lr_ep, 0.9, 0.999,
                       1e-8, self.step, None)
                if (b0 // batch) % 64 == 0:
                    L.synchronize()
                    losses.append(float(self.loss.download(np.float32, 1)[0]))
            L.synchronize()
            history.append(float(np.mean(losses)) if losses else float("nan"))
            if log:
                log("epoch %d: loss %.4f (%.1f s)" % (ep, history[-1], time.perf_counter() - t0))
        self.pack()
        return history

    def pack(self):
        L.call("sd_mlp_pack_bf16", self.w.ptr, self.packed.ptr, None)

    def infer(self, X, n, pairs=None, quat=None, out=None, stream=None):
        out = out or L.DeviceBuffer(12 * n)
        L.call("sd_mlp_infer_bf16", X.ptr, n, self.packed.ptr, pairs.ptr if pairs else None,
               quat.ptr if quat else None, out.ptr, stream)
        return out


def pair_features(d, pairs, npairs, stream=None):
    """Descriptors + features of the pairs of a DeviceHulls set (device buffer)."""
    desc = L.DeviceBuffer(4 * 7 * d.n)
    b = d.buf
    L.call("sd_hull_descriptors", b["pos4"].ptr, b["vstart"].ptr, b["nv"].ptr, b["verts4"].ptr, d.n, desc.ptr,
           stream)
    X = L.DeviceBuffer(4 * NIN * max(npairs, 1))
    L.call("sd_mlp_features", pairs.ptr, npairs, b["pos4"].ptr, b["quat4"].ptr, desc.ptr, X.ptr, stream)
    return X, desc
