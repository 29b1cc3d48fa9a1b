"""Host-side driver of one sphere scene on the GPU (wraps the sd_ctx_* C ABI).

Mirrors the reference's engine life cycle: construct, add spheres, upload
(prep + transfer), update(dt) steps, read back
(src/engine/engine.hpp:43-190, src/engine/cuda/simulation_kernel.cu:52).
"""
import ctypes as C

import numpy as np

from . import _lib as L


class SphereScene:
    def __init__(self, grid, contact, key64=False, use_graph=True, tile_items=0):
        L.require_gpu()
        self.params = L.SphereParams()
        self.params.grid = grid
        self.params.contact = contact
        self.params.key64 = 1 if key64 else 0
        self.params.use_graph = 1 if use_graph else 0
        self.params.tile_items = tile_items
        h = C.c_void_p()
        L.call("sd_ctx_create", C.byref(self.params), C.byref(h))
        self.h = h
        self.n = 0

    def upload(self, x, y, z, radius, mass, vx=None, vy=None, vz=None):
        n = len(x)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        zero = np.zeros(n, np.float32)
        arrs = [f(x), f(y), f(z), f(vx if vx is not None else zero), f(vy if vy is not None else zero),
                f(vz if vz is not None else zero), f(radius), f(mass)]
        L.call("sd_ctx_upload_spheres", self.h, n, *[L.ptr(a) for a in arrs])
        self.n = n

    def step(self, nsteps=1):
        L.call("sd_ctx_step", self.h, int(nsteps))

    def profile(self, nsteps):
        """Per-stage GPU time in ms summed over nsteps eager steps:
        sort, ranges (the cell starts), permute, contact (HIP events on the
        context stream). With u64 keys "sort" is the re-sort, positions permuted
        and cell starts updated in the same pass, unless the full radix sort ran
        (then "ranges" holds the cell starts and the permutation)."""
        out = np.zeros(4, np.float32)
        L.call("sd_ctx_profile", self.h, int(nsteps), L.ptr(out), 4)
        return {"sort": float(out[0]), "ranges": float(out[1]), "permute": float(out[2]),
                "contact": float(out[3])}

    def download(self):
        n = self.n
        out = {k: np.empty(n, np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "wx", "wy", "wz")}
        L.call("sd_ctx_download_spheres", self.h, *[L.ptr(out[k]) for k in
                                                    ("x", "y", "z", "vx", "vy", "vz", "wx", "wy", "wz")])
        return out

    def cell_changes(self):
        """(spheres that changed cell in the last step, whether the next step
        uses the full radix sort)."""
        m, full = C.c_int(), C.c_int()
        L.call("sd_ctx_cell_changes", self.h, C.byref(m), C.byref(full))
        return m.value, bool(full.value)

    def contact_count(self):
        c = C.c_longlong()
        L.call("sd_ctx_contact_count", self.h, C.byref(c))
        return c.value

    def diagnostics(self):
        out = np.zeros(8, np.float64)
        L.call("sd_ctx_diagnostics", self.h, L.ptr(out))
        return {"momentum": out[0:3].copy(), "angular_momentum": out[3:6].copy(), "kinetic": float(out[6]),
                "potential": float(out[7])}

    def autotune(self, trial_steps=4):
        """Sweeps sort tile size x contact block size; keeps the fastest (state unchanged)."""
        ti, cb, ms = C.c_int(), C.c_int(), C.c_float()
        L.call("sd_ctx_autotune", self.h, int(trial_steps), C.byref(ti), C.byref(cb), C.byref(ms))
        return {"tile_items": ti.value, "contact_block": cb.value, "ms_per_step": ms.value}

    def stream(self):
        s = C.c_void_p()
        L.call("sd_ctx_stream", self.h, C.byref(s))
        return s

    def save(self):
        nb = C.c_size_t()
        L.call("sd_ctx_state_bytes", self.h, C.byref(nb))
        buf = np.empty(nb.value, np.uint8)
        L.call("sd_ctx_save", self.h, L.ptr(buf), nb.value)
        return buf

    def load(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        L.call("sd_ctx_load", self.h, L.ptr(buf), buf.nbytes)

    def set_persistent(self, on=True):
        """Persistent form of the fused step kernel (packed 64-bit keys); same results."""
        L.call("sd_ctx_set_persistent", self.h, 1 if on else 0)

    def remove(self, ids):
        """Drops the spheres with these particle ids; the others are renumbered in id order."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        L.call("sd_ctx_remove_spheres", self.h, L.ptr(ids), len(ids))
        self.n -= len(ids)

    def add(self, x, y, z, radius, mass, vx=None, vy=None, vz=None):
        """Appends spheres with ids n, n + 1, ...; returns their ids."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        k = len(x)
        zero = np.zeros(k, np.float32)
        arrs = [f(x), f(y), f(z), f(vx if vx is not None else zero), f(vy if vy is not None else zero),
                f(vz if vz is not None else zero), f(radius), f(mass)]
        L.call("sd_ctx_add_spheres", self.h, k, *[L.ptr(a) for a in arrs])
        first = self.n
        self.n += k
        return np.arange(first, first + k)

    def close(self):
        if self.h is not None and self.h.value:
            L.load().sd_ctx_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchScene(SphereScene):
    """Many independent sphere scenes stepped as one (wraps sd_batch_create /
    sd_batch_upload; stepping, profiling, download and checkpoint are the
    sd_ctx_* calls of SphereScene). Arrays are concatenated in scene order."""

    def __init__(self, use_graph=True, tile_items=0, contact_block=0):
        L.require_gpu()
        self.params = L.BatchParams()
        self.params.use_graph = 1 if use_graph else 0
        self.params.tile_items = tile_items
        self.params.contact_block = contact_block
        h = C.c_void_p()
        L.call("sd_batch_create", C.byref(self.params), C.byref(h))
        self.h = h
        self.n = 0
        self.counts = []

    def upload_scenes(self, scenes, x, y, z, radius, mass, vx=None, vy=None, vz=None):
        """scenes: list of (n, Grid, ContactParams)."""
        desc = (L.SceneDesc * len(scenes))()
        for k, (n, g, c) in enumerate(scenes):
            desc[k].n = int(n)
            desc[k].grid = g
            desc[k].contact = c
        n = len(x)
        assert n == sum(int(sc[0]) for sc in scenes)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        zero = np.zeros(n, np.float32)
        arrs = [f(x), f(y), f(z), f(vx if vx is not None else zero), f(vy if vy is not None else zero),
                f(vz if vz is not None else zero), f(radius), f(mass)]
        L.call("sd_batch_upload", self.h, len(scenes), C.cast(desc, C.c_void_p), *[L.ptr(a) for a in arrs])
        self.n = n
        self.counts = [int(sc[0]) for sc in scenes]

    def upload(self, *a, **k):
        raise TypeError("BatchScene takes upload_scenes(scenes, x, y, z, radius, mass)")

    def attach_hulls(self, hulls):
        """Steps the bodies of a hullbody.HullBodyScene with these scenes (None detaches)."""
        L.call("sd_batch_attach_hulls", self.h, None if hulls is None else hulls.h)
        self.hulls = hulls
