"""Convex-hull sets on the GPU: AABBs, sweep-and-prune, GJK/EPA.

hulls_c4: 131,072 convex hulls, each the hull of 8-32 seeded points on an
ellipsoid with semi-axes U[5, 15] mm, random orientations, velocities
N(0, 0.5 m/s), centres uniform in a 1 m box (overlaps allowed). Seed 0x5EED0004.
"""
import ctypes as C

import numpy as np

from . import _lib as L

MAX_ITER = 64


def generate_hulls(seed, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), nv_min=8, nv_max=32, ax_min=0.005,
                   ax_max=0.015, v_sigma=0.5):
    cap = n * nv_max
    pos4 = np.zeros((n, 4), np.float32)
    quat4 = np.zeros((n, 4), np.float32)
    vel4 = np.zeros((n, 4), np.float32)
    vstart = np.zeros(n, np.int32)
    nv = np.zeros(n, np.int32)
    verts4 = np.zeros((cap, 4), np.float32)
    nverts = C.c_int()
    lo_a, hi_a = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    L.call("sd_gen_hulls", seed, n, L.ptr(lo_a), L.ptr(hi_a), nv_min, nv_max, ax_min, ax_max, v_sigma, L.ptr(pos4),
           L.ptr(quat4), L.ptr(vel4), L.ptr(vstart), L.ptr(nv), L.ptr(verts4), cap, C.byref(nverts))
    return {"pos4": pos4, "quat4": quat4, "vel4": vel4, "vstart": vstart, "nv": nv,
            "verts4": verts4[:nverts.value].copy()}


def hulls_c4(n=131072):
    return generate_hulls(0x5EED0004, n)


class DeviceHulls:
    """A hull set resident in device memory."""

    def __init__(self, h):
        L.require_gpu()
        self.n = len(h["pos4"])
        self.n_verts = len(h["verts4"])
        self.buf = {k: L.DeviceBuffer.from_array(np.ascontiguousarray(h[k])) for k in
                    ("pos4", "quat4", "vstart", "nv", "verts4")}
        self.lo4 = L.DeviceBuffer(16 * self.n)
        self.hi4 = L.DeviceBuffer(16 * self.n)
        nb = C.c_size_t()
        L.call("sd_sap_scratch_bytes", self.n, C.byref(nb))
        self.scratch_bytes = nb.value
        self.scratch = L.DeviceBuffer(nb.value)

    def _args(self):
        b = self.buf
        return b["pos4"].ptr, b["quat4"].ptr, b["vstart"].ptr, b["nv"].ptr, b["verts4"].ptr

    def aabbs(self, stream=None):
        L.call("sd_hull_aabbs", *self._args(), self.n, self.lo4.ptr, self.hi4.ptr, stream)

    def download_aabbs(self):
        L.synchronize()
        return (self.lo4.download(np.float32, 4 * self.n).reshape(-1, 4),
                self.hi4.download(np.float32, 4 * self.n).reshape(-1, 4))

    def sap(self, capacity=None, stream=None):
        """Overlapping AABB pairs (device buffer, count, axis). Grows once on overflow."""
        cap = capacity or max(16 * self.n, 1024)
        for _ in range(2):
            pairs = L.DeviceBuffer(8 * cap)
            npairs, axis = C.c_int(), C.c_int()
            st = L.load().sd_sap_pairs(self.lo4.ptr, self.hi4.ptr, self.n, pairs.ptr, cap, C.byref(npairs),
                                       C.byref(axis), self.scratch.ptr, self.scratch_bytes, stream)
            if st == L.SD_OK:
                return pairs, npairs.value, axis.value
            if st != L.SD_ERR_CAPACITY:
                L.check(st)
            cap = npairs.value
        raise L.SdError(L.SD_ERR_CAPACITY, "sap: pair count changed between attempts")

    def gjk(self, pairs, npairs, warm=None, warm_budget=16, max_iter=MAX_ITER, near_margin=1e-6, stream=None,
            out=None):
        """Runs GJK/EPA on device pair buffer `pairs`; returns device output buffers."""
        if out is None:
            out = {"sep": L.DeviceBuffer(4 * npairs), "normal": L.DeviceBuffer(12 * npairs),
                   "flags": L.DeviceBuffer(4 * npairs), "iters": L.DeviceBuffer(4 * npairs)}
        nb = C.c_size_t()
        L.call("sd_gjk_scratch_bytes", npairs, C.byref(nb))
        if getattr(self, "_gjk_scratch", None) is None or self._gjk_scratch.nbytes < nb.value:
            self._gjk_scratch = L.DeviceBuffer(nb.value)
        L.call("sd_gjk", pairs.ptr, npairs, *self._args(), self.n, self.n_verts, warm.ptr if warm else None,
               warm_budget, max_iter, near_margin, out["sep"].ptr, out["normal"].ptr, out["flags"].ptr,
               out["iters"].ptr, self._gjk_scratch.ptr, self._gjk_scratch.nbytes, stream)
        return out

    @staticmethod
    def download_gjk(out, npairs):
        L.synchronize()
        return {"sep": out["sep"].download(np.float32, npairs),
                "normal": out["normal"].download(np.float32, 3 * npairs).reshape(-1, 3),
                "flags": out["flags"].download(np.int32, npairs),
                "iters": out["iters"].download(np.int32, npairs)}


def concat_hulls(sets):
    """Concatenates hull sets (vertex ranges re-based); returns (hulls, scene_first)."""
    out = {k: [] for k in ("pos4", "quat4", "vel4", "vstart", "nv", "verts4")}
    first = [0]
    vbase = 0
    for h in sets:
        for k in ("pos4", "quat4", "vel4", "nv", "verts4"):
            out[k].append(h[k])
        out["vstart"].append(h["vstart"] + vbase)
        vbase += len(h["verts4"])
        first.append(first[-1] + len(h["pos4"]))
    cat = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in out.items()}
    cat["vstart"] = cat["vstart"].astype(np.int32)
    return cat, np.array(first, np.int32)


class BatchedHulls(DeviceHulls):
    """Hull sets of many scenes in one device set; sweep-and-prune per scene in one pass."""

    def __init__(self, sets):
        h, first = concat_hulls(sets)
        super().__init__(h)
        self.nscenes = len(sets)
        self.scene_first = first
        scene_of = np.repeat(np.arange(self.nscenes, dtype=np.int32), np.diff(first))
        self.d_first = L.DeviceBuffer.from_array(first)
        self.d_scene_of = L.DeviceBuffer.from_array(scene_of)
        self.d_axes = L.DeviceBuffer(4 * self.nscenes)
        nb = C.c_size_t()
        L.call("sd_sap_batched_scratch_bytes", self.n, self.nscenes, C.byref(nb))
        self.b_scratch_bytes = nb.value
        self.b_scratch = L.DeviceBuffer(nb.value)

    def sap_batched(self, capacity=None, stream=None):
        """Overlapping pairs of every scene (device buffer, count)."""
        cap = capacity or max(16 * self.n, 1024)
        for _ in range(2):
            pairs = L.DeviceBuffer(8 * cap)
            npairs = C.c_int()
            st = L.load().sd_sap_pairs_batched(self.lo4.ptr, self.hi4.ptr, self.n, self.d_scene_of.ptr,
                                               self.d_first.ptr, self.nscenes, pairs.ptr, cap, C.byref(npairs),
                                               self.d_axes.ptr, self.b_scratch.ptr, self.b_scratch_bytes, stream)
            if st == L.SD_OK:
                return pairs, npairs.value
            if st != L.SD_ERR_CAPACITY:
                L.check(st)
            cap = npairs.value
        raise L.SdError(L.SD_ERR_CAPACITY, "sap_batched: pair count changed between attempts")

    def axes(self):
        L.synchronize()
        return self.d_axes.download(np.int32, self.nscenes)
