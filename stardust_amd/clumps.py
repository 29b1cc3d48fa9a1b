"""Rigid clumps on the GPU (wraps the sd_clump_* C ABI).

clumps_c3: 262,144 rigid bodies on a 64^3 lattice of spacing 6 cm (jitter 5 mm),
each a clump of U{4..16} spheres with centres N(0, 1 cm) about the body centre
and radii U[3, 6] mm (density 2500 kg/m^3); inertia from the member spheres;
uniform orientations; velocities N(0, 0.5 m/s). LBVH over body AABBs with
30-bit Morton codes rebuilt every step; kn = 1e5 N/m, zeta = 0.3, mu = 0.5,
dt = 1e-5 s, 1,000 steps. Seed 0x5EED0003.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class ClumpParams(C.Structure):
    _fields_ = [("contact", L.ContactParams)]


def generate_clumps(seed, nb, lattice=None, spacing=0.06, jitter=0.005, c_min=4, c_max=16, sigma=0.01,
                    r_min=0.003, r_max=0.006, rho=2500.0, v_sigma=0.5):
    if lattice is None:
        lattice = int(np.ceil(nb ** (1.0 / 3.0) - 1e-9))
        while lattice ** 3 < nb:
            lattice += 1
    out = {k: np.zeros((nb, 4), np.float32) for k in ("pos4", "vel4", "quat4", "angvel4")}
    out["inertia6"] = np.zeros((nb, 6), np.float32)
    out["first"] = np.zeros(nb, np.int32)
    out["count"] = np.zeros(nb, np.int32)
    local = np.zeros((nb * c_max, 4), np.float32)
    ns = C.c_int()
    L.call("sd_gen_clumps", seed, nb, lattice, spacing, jitter, c_min, c_max, sigma, r_min, r_max, rho, v_sigma,
           L.ptr(out["pos4"]), L.ptr(out["vel4"]), L.ptr(out["quat4"]), L.ptr(out["angvel4"]),
           L.ptr(out["inertia6"]), L.ptr(out["first"]), L.ptr(out["count"]), L.ptr(local), C.byref(ns))
    out["local4"] = local[:ns.value].copy()
    return out


def clumps_c3(nb=262144):
    return generate_clumps(0x5EED0003, nb, lattice=64)


def contact_c3():
    return L.ContactParams.make(1e5, 0.3, 0.5, 1e-5, gravity=(0.0, 0.0, -9.81), walls=None, max_contacts=16)


class ClumpScene:
    # stage times of sd_clump_profile; [3] and [4] are empty (the partner search
    # runs inside the contact kernel)
    STAGES = ("pose", "morton_sort", "lbvh_build", None, None, "partners_contacts")

    def __init__(self, contact):
        L.require_gpu()
        self.params = ClumpParams()
        self.params.contact = contact
        h = C.c_void_p()
        L.call("sd_clump_create", C.byref(self.params), C.byref(h))
        self.h = h
        self.nb = 0

    def upload(self, s):
        self.nb, self.ns = len(s["pos4"]), len(s["local4"])
        self.counts = np.asarray(s["count"], np.int64).copy()
        a = [np.ascontiguousarray(s[k]) for k in ("pos4", "vel4", "quat4", "angvel4", "inertia6", "first", "count")]
        loc = np.ascontiguousarray(s["local4"], np.float32)
        L.call("sd_clump_upload", self.h, self.nb, *[L.ptr(x) for x in a], self.ns, L.ptr(loc))

    def step(self, n=1):
        L.call("sd_clump_step", self.h, int(n))

    def profile(self, n):
        out = np.zeros(6, np.float32)
        L.call("sd_clump_profile", self.h, int(n), L.ptr(out), 6)
        return {k: float(v) for k, v in zip(self.STAGES, out) if k}

    def download(self):
        out = {k: np.zeros((self.nb, 4), np.float32) for k in ("pos4", "vel4", "quat4", "angm4")}
        L.call("sd_clump_download", self.h, L.ptr(out["pos4"]), L.ptr(out["vel4"]), L.ptr(out["quat4"]),
               L.ptr(out["angm4"]))
        return out

    def pairs(self):
        n = C.c_int()
        L.load().sd_clump_pairs(self.h, None, 0, C.byref(n))
        out = np.zeros((max(n.value, 1), 2), np.int32)
        L.call("sd_clump_pairs", self.h, L.ptr(out), n.value, C.byref(n))
        return out[:n.value]

    def spheres(self):
        out = np.zeros((self.ns, 4), np.float32)
        L.call("sd_clump_spheres", self.h, L.ptr(out))
        return out

    def save(self):
        nb = C.c_size_t()
        L.call("sd_clump_state_bytes", self.h, C.byref(nb))
        buf = np.empty(nb.value, np.uint8)
        L.call("sd_clump_save", self.h, L.ptr(buf), nb.value)
        return buf

    def load(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        L.call("sd_clump_load", self.h, L.ptr(buf), buf.nbytes)

    def remove(self, ids):
        """Drops these bodies and their spheres; the rest are renumbered in order."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        L.call("sd_clump_remove_bodies", self.h, L.ptr(ids), len(ids))
        self.counts = np.delete(self.counts, ids)
        self.nb, self.ns = len(self.counts), int(self.counts.sum())

    def add(self, s):
        """Appends the bodies of a generate_clumps-style dict; returns their ids."""
        a = [np.ascontiguousarray(s[k]) for k in ("pos4", "vel4", "quat4", "angvel4", "inertia6", "first", "count")]
        loc = np.ascontiguousarray(s["local4"], np.float32)
        k = len(s["pos4"])
        L.call("sd_clump_add_bodies", self.h, k, *[L.ptr(x) for x in a], len(loc), L.ptr(loc))
        first = self.nb
        self.counts = np.concatenate([self.counts, np.asarray(s["count"], np.int64)])
        self.nb, self.ns = len(self.counts), int(self.counts.sum())
        return np.arange(first, first + k)

    def close(self):
        if self.h is not None and self.h.value:
            L.load().sd_clump_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
