"""Rigid convex bodies stepped on the GPU (wraps sd_hull_*): convex hulls,
spheres and static triangles in independent scenes; a uniform grid per scene,
GJK/EPA (cold or learned warm start) and the contact law of every touching
pair each step.

make_bodies turns a hull set of hulls.generate_hulls into bodies (mass from the
density and the ellipsoid of the vertex box, inertia of equal point masses at
the vertices, seeded velocities); box_walls adds the twelve triangles of a
box's faces as static bodies.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class HullParams(C.Structure):
    _fields_ = [("contact", L.ContactParams), ("max_iter", C.c_int), ("pairs_per_body", C.c_int),
                ("use_graph", C.c_int)]


def make_bodies(h, seed, density=2500.0, w_sigma=5.0):
    """Bodies from a hull set: pos4 = (centre, mass), vel4 from the set,
    angular velocities N(0, w_sigma) rad/s, inertia6 of equal point masses at
    the vertices about the centre (float64), radius 0."""
    n = len(h["pos4"])
    rng = np.random.Generator(np.random.PCG64(seed))
    pos4 = h["pos4"].copy()
    vs, nv = h["vstart"].astype(np.int64), h["nv"].astype(np.int64)
    owner = np.repeat(np.arange(n), nv)
    rows = np.concatenate([np.arange(a, a + k) for a, k in zip(vs, nv)]) if n else np.zeros(0, np.int64)
    v = h["verts4"][rows, :3].astype(np.float64)
    lo = np.full((n, 3), np.inf)
    hi = np.full((n, 3), -np.inf)
    np.minimum.at(lo, owner, v)
    np.maximum.at(hi, owner, v)
    half = 0.5 * (hi - lo)
    m = density * 4.0 / 3.0 * np.pi * half[:, 0] * half[:, 1] * half[:, 2]
    pm = (m / nv)[owner]
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    terms = np.stack([y * y + z * z, x * x + z * z, x * x + y * y, -x * y, -x * z, -y * z], 1) * pm[:, None]
    inertia = np.zeros((n, 6))
    np.add.at(inertia, owner, terms)
    inertia6 = inertia.astype(np.float32)
    pos4[:, 3] = m.astype(np.float32)
    angvel4 = np.zeros((n, 4), np.float32)
    angvel4[:, :3] = rng.normal(0.0, w_sigma, (n, 3)).astype(np.float32)
    return {"pos4": pos4, "vel4": h["vel4"].copy(), "quat4": h["quat4"].copy(), "angvel4": angvel4,
            "inertia6": inertia6, "radius": np.zeros(n, np.float32), "vstart": h["vstart"].copy(),
            "nv": h["nv"].copy(), "verts4": h["verts4"].copy()}


def triangle_bodies(tri):
    """Static bodies from world triangles tri[k] = (3, 3): centre = the
    vertices' mean, vertices relative to it, mass 0."""
    tri = np.asarray(tri, np.float32)
    n = len(tri)
    c = (tri.astype(np.float64).mean(1)).astype(np.float32)
    verts4 = np.zeros((3 * n, 4), np.float32)
    verts4[:, :3] = (tri - c[:, None, :]).reshape(-1, 3)
    pos4 = np.zeros((n, 4), np.float32)
    pos4[:, :3] = c
    quat4 = np.zeros((n, 4), np.float32)
    quat4[:, 0] = 1.0
    z4 = np.zeros((n, 4), np.float32)
    return {"pos4": pos4, "vel4": z4.copy(), "quat4": quat4, "angvel4": z4.copy(),
            "inertia6": np.zeros((n, 6), np.float32), "radius": np.zeros(n, np.float32),
            "vstart": (3 * np.arange(n)).astype(np.int32), "nv": np.full(n, 3, np.int32), "verts4": verts4}


def box_walls(lo, hi):
    """The twelve triangles of the faces of the box [lo, hi] (two per face)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0], [x0, y0, z1], [x1, y0, z1], [x1, y1, z1],
                  [x0, y1, z1]], np.float32)
    faces = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 5, 1), (0, 4, 5), (3, 2, 6), (3, 6, 7), (0, 3, 7),
             (0, 7, 4), (1, 5, 6), (1, 6, 2)]
    return triangle_bodies(np.stack([c[list(f)] for f in faces]))


def concat(sets):
    """Bodies of several sets in order (vertex ranges offset); returns the
    concatenation and the scene boundaries (one scene per set)."""
    out, first, voff = {}, [0], 0
    for s in sets:
        for k, v in s.items():
            v = v + voff if k == "vstart" else v
            out.setdefault(k, []).append(v)
        voff += len(s["verts4"])
        first.append(first[-1] + len(s["pos4"]))
    return {k: np.ascontiguousarray(np.concatenate(v)) for k, v in out.items()}, np.array(first, np.int32)


class HullBodyScene:
    def __init__(self, contact, max_iter=64, pairs_per_body=32, use_graph=True):
        L.require_gpu()
        self.params = HullParams()
        self.params.contact = contact
        self.params.max_iter = max_iter
        self.params.pairs_per_body = pairs_per_body
        self.params.use_graph = 1 if use_graph else 0
        h = C.c_void_p()
        L.call("sd_hull_create", C.byref(self.params), C.byref(h))
        self.h = h
        self.n = 0
        self.model = None  # kept alive while attached

    def upload(self, b, scene_first=None):
        self.n = len(b["pos4"])
        a = [np.ascontiguousarray(b[k], dtype=np.float32) for k in ("pos4", "vel4", "quat4", "angvel4", "inertia6",
                                                                    "radius")]
        vs = np.ascontiguousarray(b["vstart"], np.int32)
        nv = np.ascontiguousarray(b["nv"], np.int32)
        verts = np.ascontiguousarray(b["verts4"], np.float32)
        if scene_first is None:
            sf, ns = None, 1
        else:
            sf = np.ascontiguousarray(scene_first, np.int32)
            ns = len(sf) - 1
        L.call("sd_hull_upload", self.h, self.n, *[L.ptr(x) for x in a], L.ptr(vs), L.ptr(nv), len(verts),
               L.ptr(verts), None if sf is None else L.ptr(sf), ns)

    def set_warm(self, model, warm_budget=16):
        """Warm-starts the step's GJK queries from a trained
        warmstart.WarmStartModel (its packed BF16 network); None: cold."""
        self.model = model
        L.call("sd_hull_set_warm", self.h, model.packed.ptr if model is not None else None, int(warm_budget))

    def step(self, n=1):
        L.call("sd_hull_step", self.h, int(n))

    def download(self):
        out = {k: np.zeros((self.n, 4), np.float32) for k in ("pos4", "vel4", "quat4", "angm4")}
        L.call("sd_hull_download", self.h, L.ptr(out["pos4"]), L.ptr(out["vel4"]), L.ptr(out["quat4"]),
               L.ptr(out["angm4"]))
        return out

    def contacts(self):
        k = C.c_int()
        L.load().sd_hull_contacts(self.h, None, None, None, 0, C.byref(k))
        m = max(k.value, 1)
        pairs = np.zeros((m, 2), np.int32)
        f = np.zeros((m, 4), np.float32)
        p = np.zeros((m, 4), np.float32)
        L.call("sd_hull_contacts", self.h, L.ptr(pairs), L.ptr(f), L.ptr(p), m, C.byref(k))
        return pairs[:k.value], f[:k.value], p[:k.value]

    def narrowphase_stats(self):
        """Last step: candidate pairs, hits, warm-started results kept, support
        passes of GJK and EPA summed over the pairs."""
        a, b, w = C.c_int(), C.c_int(), C.c_int()
        it = C.c_longlong()
        L.call("sd_hull_narrowphase_stats", self.h, C.byref(a), C.byref(b), C.byref(w), C.byref(it))
        return {"pairs": a.value, "hits": b.value, "warm_used": w.value, "iters": it.value}

    def close(self):
        if self.h is not None and self.h.value:
            L.load().sd_hull_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
