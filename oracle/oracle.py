"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrGrid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("cell", C.c_float), ("n", C.c_int * 3)]


class OrContact(C.Structure):
    _fields_ = [("kn", C.c_float), ("damping_ratio", C.c_float), ("kt_ratio", C.c_float), ("mu", C.c_float),
                ("dt", C.c_float), ("gravity", C.c_float * 3), ("walls", C.c_int), ("wall_lo", C.c_float * 3),
                ("wall_hi", C.c_float * 3), ("max_contacts", C.c_int)]


_P = C.c_void_p
_I = C.c_int
_SIGS = {
    "or_hash_keys": ([_I, _P, _P, _P, _I, C.POINTER(OrGrid), _P, _P], None),
    "or_cell_id": ([C.POINTER(OrGrid), C.c_float, C.c_float, C.c_float], C.c_uint32),
    "or_sort_pairs_u32": ([_I, _P, _P, _I, _I], None),
    "or_sort_keys_u64": ([_I, _P, _I, _I], None),
    "or_cell_ranges_u32": ([_I, _P, _I, _P], None),
    "or_cell_ranges_u64": ([_I, _P, _I, _P], None),
    "or_neighbor_pairs": ([_I, _P, C.POINTER(OrGrid), _P, C.c_float, _P, C.c_long], C.c_long),
    "or_brute_pairs": ([_I, _P, C.c_float, _P, C.c_long], C.c_long),
    "or_contact_forces": ([_I, _P, _P, _P, _P, C.POINTER(OrGrid), _P, C.POINTER(OrContact), _P, _P, _P, _P, _P, _P,
                           _P, _I], _I),
    "or_contact_forces_f64": ([_I, _P, _P, _P, _P, C.POINTER(OrGrid), _P, C.POINTER(OrContact), _P, _P, _P, _P, _P,
                               _P, _P], _I),
    "or_integrate": ([_I, _P, _P, _P, _P, _P, C.POINTER(OrContact)], None),
    "or_sphere_step": ([_I, C.POINTER(OrGrid), C.POINTER(OrContact), _P, _P, _P, _P, _P, _P, _I], _I),
    "or_hull_aabbs": ([_I, _P, _P, _P, _P, _P, _P, _P], None),
    "or_sap_axis": ([_I, _P, _P], _I),
    "or_sap_pairs": ([_I, _P, _P, _I, _P, C.c_long], C.c_long),
    "or_brute_aabb_pairs": ([_I, _P, _P, _P, C.c_long], C.c_long),
    "or_gjk": ([_I, _P, _P, _P, _P, _P, _P, _P, _I, _I, C.c_float, _P, _P, _P, _P, _I], None),
    "or_gjk_ex": ([_I, _P, _P, _P, _P, _P, _P, _P, _I, _I, C.c_float, _P, _P, _P, _P, _P, _I], None),
    "or_clump_prepare": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], None),
    "or_clump_pose": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], None),
    "or_global_domain": ([_I, _P, _P, _P], None),
    "or_morton30": ([_I, _P, _P, _P, _P], None),
    "or_clump_contacts_integrate": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_long, _P, C.POINTER(OrContact),
                                     _P, _P, _I], _I),
    "or_compat_pair_forces": ([_I, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, _P, _P], None),
    "or_compat_inverse3x3": ([_I, _P, _P], None),
    "or_compat_angular_velocity": ([_I, _P, _P, _P, _P], None),
    "or_body_reduce": ([_I, _P, _P, _P, _P, _P, _P, _P, _P], None),
    "or_integrate_bodies": ([_I, _P, _P, _P, _P, _P, _P, _P, C.POINTER(OrContact)], None),
    "or_hull_grid": ([_I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P], C.c_long),
    "or_hull_grid_pairs": ([_I, _P, _P, _P, _P, _I, _P, _P, _P, C.c_long], C.c_long),
    "or_hull_step": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, C.POINTER(OrContact), _I, _I,
                      _P, _P, _P, C.c_long], C.c_long),
    "or_clump_step": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, C.POINTER(OrContact)], C.c_long),
}


class ClumpState:
    """Rigid-clump state as the GPU context prepares it (float4 rows)."""

    def __init__(self, s):
        nb = len(s["pos4"])
        self.nb, self.ns = nb, len(s["local4"])
        self.s = s
        self.P, self.V, self.Q, self.L, self.I0, self.I1 = (np.zeros((nb, 4), np.float32) for _ in range(6))
        a = [np.ascontiguousarray(s[k], np.float32) for k in ("pos4", "vel4", "quat4", "angvel4", "inertia6")]
        load().or_clump_prepare(nb, *[p(x) for x in a], p(self.P), p(self.V), p(self.Q), p(self.L), p(self.I0),
                                p(self.I1))
        self.first = np.ascontiguousarray(s["first"], np.int32)
        self.count = np.ascontiguousarray(s["count"], np.int32)
        self.local4 = np.ascontiguousarray(s["local4"], np.float32)

    def remove(self, ids):
        """Runtime removal (EntityHandler::removeEntity is a stub in the
        reference, src/engine/entities/entity_handler.cpp:108-119): drops the
        bodies and their sphere rows; the rest keep their order."""
        keep = np.ones(self.nb, bool)
        keep[np.asarray(ids, np.int64)] = False
        rows = np.concatenate([np.arange(f, f + c) for f, c in zip(self.first[keep], self.count[keep])])
        for k in ("P", "V", "Q", "L", "I0", "I1"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k)[keep]))
        self.count = np.ascontiguousarray(self.count[keep])
        self.first = np.ascontiguousarray(np.concatenate([[0], np.cumsum(self.count)[:-1]]).astype(np.int32))
        self.local4 = np.ascontiguousarray(self.local4[rows])
        self.nb, self.ns = int(keep.sum()), len(rows)

    def add(self, s):
        """Runtime addition: appends the bodies of s (prepared as at
        construction) after the current ones."""
        new = ClumpState(s)
        for k in ("P", "V", "Q", "L", "I0", "I1"):
            setattr(self, k, np.ascontiguousarray(np.concatenate([getattr(self, k), getattr(new, k)])))
        self.first = np.ascontiguousarray(np.concatenate([self.first, new.first + self.ns]).astype(np.int32))
        self.count = np.ascontiguousarray(np.concatenate([self.count, new.count]).astype(np.int32))
        self.local4 = np.ascontiguousarray(np.concatenate([self.local4, new.local4]))
        self.nb += new.nb
        self.ns += new.ns

    def pose(self):
        sx = np.zeros((self.ns, 4), np.float32)
        sv = np.zeros((self.ns, 4), np.float32)
        lo = np.zeros((self.nb, 4), np.float32)
        hi = np.zeros((self.nb, 4), np.float32)
        load().or_clump_pose(self.nb, p(self.P), p(self.V), p(self.Q), p(self.L), p(self.I0), p(self.I1),
                             p(self.first), p(self.count), p(self.local4), p(sx), p(sv), p(lo), p(hi))
        return sx, sv, lo, hi

    def step(self, c, nsteps=1):
        oc = contact(c)
        np_ = 0
        for _ in range(nsteps):
            np_ = load().or_clump_step(self.nb, p(self.P), p(self.V), p(self.Q), p(self.L), p(self.I0), p(self.I1),
                                       p(self.first), p(self.count), p(self.local4), self.ns, C.byref(oc))
        return np_

    def forces(self, c, pairs):
        """Force/torque of every body for the current state and a given pair set."""
        sx, sv, lo, hi = self.pose()
        f = np.zeros((self.nb, 4), np.float32)
        t = np.zeros((self.nb, 4), np.float32)
        oc = contact(c)
        pr = np.ascontiguousarray(pairs, np.int32)
        load().or_clump_contacts_integrate(self.nb, p(self.P), p(self.V), p(self.Q), p(self.L), p(self.I0),
                                           p(self.I1), p(self.first), p(self.count), p(sx), p(sv), len(pr), p(pr),
                                           C.byref(oc), p(f), p(t), 0)
        return f, t


class HullBodyState:
    """Rigid convex-body state as the GPU context prepares it (float4 rows);
    step() is one sd_hull_step."""

    def __init__(self, b, scene_first=None):
        n = len(b["pos4"])
        self.n = n
        self.P, self.V, self.Q, self.L, self.I0, self.I1 = (np.zeros((n, 4), np.float32) for _ in range(6))
        a = [np.ascontiguousarray(b[k], np.float32) for k in ("pos4", "vel4", "quat4", "angvel4", "inertia6")]
        load().or_clump_prepare(n, *[p(x) for x in a], p(self.P), p(self.V), p(self.Q), p(self.L), p(self.I0),
                                p(self.I1))
        self.radius = np.ascontiguousarray(b["radius"], np.float32)
        self.vstart = np.ascontiguousarray(b["vstart"], np.int32)
        self.nv = np.ascontiguousarray(b["nv"], np.int32)
        self.verts4 = np.ascontiguousarray(b["verts4"], np.float32)
        sf = [0, n] if scene_first is None else scene_first
        self.scene_first = np.ascontiguousarray(sf, np.int32)
        # the broadphase grids, fixed from the initial bodies as at upload
        self.grid16 = np.zeros((len(self.scene_first) - 1, 16), np.float32)
        self.large = np.zeros(n, np.int32)
        load().or_hull_grid(n, p(a[0]), p(self.radius), p(self.vstart), p(self.nv), p(self.verts4),
                            p(self.scene_first), len(self.scene_first) - 1, p(self.grid16), p(self.large), None)

    def grid_pairs(self, lo4, hi4):
        """The candidate pairs of the broadphase grids for the current positions
        and the given boxes."""
        lo4, hi4 = np.ascontiguousarray(lo4, np.float32), np.ascontiguousarray(hi4, np.float32)
        m = load().or_hull_grid_pairs(self.n, p(self.P), p(lo4), p(hi4), p(self.scene_first),
                                      len(self.scene_first) - 1, p(self.grid16), p(self.large), None, 0)
        out = np.zeros((max(m, 1), 2), np.int32)
        load().or_hull_grid_pairs(self.n, p(self.P), p(lo4), p(hi4), p(self.scene_first), len(self.scene_first) - 1,
                                  p(self.grid16), p(self.large), p(out), m)
        return out[:m]

    def step(self, c, nsteps=1, max_iter=64, nthreads=1):
        """Steps; returns the touching pairs of the last step (pairs, force on
        b, contact point)."""
        oc = contact(c)
        out = None
        for _ in range(nsteps):
            m = load().or_hull_step(self.n, p(self.P), p(self.V), p(self.Q), p(self.L), p(self.I0), p(self.I1),
                                    p(self.radius), p(self.vstart), p(self.nv), p(self.verts4), len(self.verts4),
                                    p(self.scene_first), len(self.scene_first) - 1, p(self.grid16), p(self.large), C.byref(oc),
                                    max_iter, nthreads, None, None, None, 0)
            out = m
        return out

    def contacts(self, c, max_iter=64, nthreads=1):
        """Touching pairs of one step from the current state, without
        advancing it (the step runs on a copy)."""
        oc = contact(c)
        P, V, Q, L = (x.copy() for x in (self.P, self.V, self.Q, self.L))
        m = load().or_hull_step(self.n, p(P), p(V), p(Q), p(L), p(self.I0), p(self.I1), p(self.radius),
                                p(self.vstart), p(self.nv), p(self.verts4), len(self.verts4), p(self.scene_first),
                                len(self.scene_first) - 1, p(self.grid16), p(self.large), C.byref(oc), max_iter, nthreads, None, None,
                                None, 0)
        pairs = np.zeros((max(m, 1), 2), np.int32)
        f = np.zeros((max(m, 1), 4), np.float32)
        q = np.zeros((max(m, 1), 4), np.float32)
        P, V, Q, L = (x.copy() for x in (self.P, self.V, self.Q, self.L))
        load().or_hull_step(self.n, p(P), p(V), p(Q), p(L), p(self.I0), p(self.I1), p(self.radius), p(self.vstart),
                            p(self.nv), p(self.verts4), len(self.verts4), p(self.scene_first),
                            len(self.scene_first) - 1, p(self.grid16), p(self.large), C.byref(oc), max_iter, nthreads, p(pairs), p(f),
                            p(q), m)
        return pairs[:m], f[:m], q[:m]


def global_domain(lo, hi):
    d = np.zeros(6, np.float32)
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    load().or_global_domain(len(lo), p(lo), p(hi), p(d))
    return d


def morton30(lo, hi, dom):
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    dom = np.ascontiguousarray(dom, np.float32)
    k = np.zeros(len(lo), np.uint64)
    load().or_morton30(len(lo), p(lo), p(hi), p(dom), p(k))
    return k


def hull_aabbs(h):
    n = len(h["pos4"])
    lo = np.zeros((n, 4), np.float32)
    hi = np.zeros((n, 4), np.float32)
    a = [np.ascontiguousarray(h[k]) for k in ("pos4", "quat4", "vstart", "nv", "verts4")]
    load().or_hull_aabbs(n, *[p(x) for x in a], p(lo), p(hi))
    return lo, hi


def sap_axis(lo, hi):
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    return load().or_sap_axis(len(lo), p(lo), p(hi))


def sap_pairs(lo, hi, axis=None):
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    if axis is None:
        axis = sap_axis(lo, hi)
    cnt = load().or_sap_pairs(len(lo), p(lo), p(hi), axis, None, 0)
    out = np.empty((max(cnt, 1), 2), np.int32)
    load().or_sap_pairs(len(lo), p(lo), p(hi), axis, p(out), cnt)
    return out[:cnt], axis


def brute_aabb_pairs(lo, hi):
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    cnt = load().or_brute_aabb_pairs(len(lo), p(lo), p(hi), None, 0)
    out = np.empty((max(cnt, 1), 2), np.int32)
    load().or_brute_aabb_pairs(len(lo), p(lo), p(hi), p(out), cnt)
    return out[:cnt]


def gjk(h, pairs, warm=None, warm_budget=16, max_iter=64, near_margin=1e-6, nthreads=1):
    pairs = np.ascontiguousarray(pairs, np.int32)
    n = len(pairs)
    sep = np.zeros(n, np.float32)
    nrm = np.zeros((n, 3), np.float32)
    flags = np.zeros(n, np.int32)
    iters = np.zeros(n, np.int32)
    a = [np.ascontiguousarray(h[k]) for k in ("pos4", "quat4", "vstart", "nv", "verts4")]
    w = None if warm is None else np.ascontiguousarray(warm, np.float32)
    small = np.zeros(n, np.int32)
    load().or_gjk_ex(n, p(pairs), *[p(x) for x in a], None if w is None else p(w), warm_budget, max_iter,
                     near_margin, p(sep), p(nrm), p(flags), p(iters), p(small), nthreads)
    return {"sep": sep, "normal": nrm, "flags": flags, "iters": iters, "small_iters": small}

_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def p(a):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def grid(g):
    """OrGrid from any object with lo, cell, n (e.g. stardust_amd.Grid)."""
    o = OrGrid()
    for a in range(3):
        o.lo[a] = g.lo[a]
        o.n[a] = g.n[a]
    o.cell = g.cell
    return o


def contact(c):
    o = OrContact()
    for f, _ in OrContact._fields_:
        v = getattr(c, f)
        if hasattr(v, "__len__"):
            arr = getattr(o, f)
            for a in range(3):
                arr[a] = v[a]
        else:
            setattr(o, f, v)
    return o


def hash_keys(x, y, z, g):
    n = len(x)
    keys = np.empty(n, np.uint32)
    vals = np.empty(n, np.uint32)
    og = grid(g)
    xs, ys, zs = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))  # kept alive across the call
    load().or_hash_keys(n, p(xs), p(ys), p(zs), 1, C.byref(og), p(keys), p(vals))
    return keys, vals


def sort_pairs_u32(keys, vals, begin_bit, end_bit):
    k = np.array(keys, np.uint32)
    v = np.array(vals, np.uint32)
    load().or_sort_pairs_u32(len(k), p(k), p(v), begin_bit, end_bit)
    return k, v


def sort_keys_u64(keys, begin_bit, end_bit):
    k = np.array(keys, np.uint64)
    load().or_sort_keys_u64(len(k), p(k), begin_bit, end_bit)
    return k


def cell_ranges_u32(keys, n_cells):
    r = np.empty(2 * n_cells, np.int32)
    k = np.ascontiguousarray(keys, np.uint32)
    load().or_cell_ranges_u32(len(k), p(k), n_cells, p(r))
    return r


def cell_ranges_u64(keys, n_cells):
    r = np.empty(2 * n_cells, np.int32)
    k = np.ascontiguousarray(keys, np.uint64)
    load().or_cell_ranges_u64(len(k), p(k), n_cells, p(r))
    return r


def neighbor_pairs(pos4, g, ranges, skin):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = len(pos4)
    og = grid(g)
    rng = np.ascontiguousarray(ranges, np.int32)
    cnt = load().or_neighbor_pairs(n, p(pos4), C.byref(og), p(rng), skin, None, 0)
    out = np.empty((max(cnt, 1), 2), np.int32)
    load().or_neighbor_pairs(n, p(pos4), C.byref(og), p(rng), skin, p(out), cnt)
    return out[:cnt]


def brute_pairs(pos4, skin):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = len(pos4)
    cnt = load().or_brute_pairs(n, p(pos4), skin, None, 0)
    out = np.empty((max(cnt, 1), 2), np.int32)
    load().or_brute_pairs(n, p(pos4), skin, p(out), cnt)
    return out[:cnt]


def contact_forces(pos4, vel4, omg4, ids, g, ranges, c, old_slot, hcount_old, hist4_old, nthreads=1):
    n = len(pos4)
    maxc = c.max_contacts
    og, oc = grid(g), contact(c)
    hcount_new = np.zeros(n, np.int32)
    hist_new = np.zeros((maxc * n, 4), np.float32)
    f = np.zeros((n, 4), np.float32)
    t = np.zeros((n, 4), np.float32)
    arrs = [np.ascontiguousarray(a) for a in (pos4, vel4, omg4, ids, ranges, old_slot, hcount_old, hist4_old)]
    mx = load().or_contact_forces(n, p(arrs[0]), p(arrs[1]), p(arrs[2]), p(arrs[3]), C.byref(og), p(arrs[4]),
                                  C.byref(oc), p(arrs[5]), p(arrs[6]), p(arrs[7]), p(hcount_new), p(hist_new), p(f),
                                  p(t), nthreads)
    return {"force": f, "torque": t, "hcount": hcount_new, "hist": hist_new, "max_contacts": mx}


def contact_forces_f64(pos4, vel4, omg4, ids, g, ranges, c, old_slot, hcount_old, hist4_old_f64):
    n = len(pos4)
    maxc = c.max_contacts
    og, oc = grid(g), contact(c)
    hcount_new = np.zeros(n, np.int32)
    hist_new = np.zeros((maxc * n, 4), np.float64)
    f = np.zeros((n, 4), np.float64)
    t = np.zeros((n, 4), np.float64)
    d = [np.ascontiguousarray(a, np.float64) for a in (pos4, vel4, omg4)]
    i = [np.ascontiguousarray(a, np.int32) for a in (ids, ranges, old_slot, hcount_old)]
    h = np.ascontiguousarray(hist4_old_f64, np.float64)
    load().or_contact_forces_f64(n, p(d[0]), p(d[1]), p(d[2]), p(i[0]), C.byref(og), p(i[1]), C.byref(oc), p(i[2]),
                                 p(i[3]), p(h), p(hcount_new), p(hist_new), p(f), p(t))
    return {"force": f, "torque": t, "hcount": hcount_new, "hist": hist_new}


def integrate(pos4, vel4, omg4, force4, torque4, c):
    pos4, vel4, omg4 = (np.array(a, np.float32) for a in (pos4, vel4, omg4))
    f, t = (np.ascontiguousarray(a, np.float32) for a in (force4, torque4))
    oc = contact(c)
    load().or_integrate(len(pos4), p(pos4), p(vel4), p(omg4), p(f), p(t), C.byref(oc))
    return pos4, vel4, omg4


class SphereState:
    """Sphere state in slot order, as the step keeps it (float4 rows)."""

    def __init__(self, x, y, z, r, m, vx=None, vy=None, vz=None, max_contacts=16):
        n = len(x)
        self.n = n
        self.maxc = max_contacts
        m = np.asarray(m, np.float32)
        r = np.asarray(r, np.float32)
        inv_m = np.where(m > 0, np.float32(1.0) / np.where(m > 0, m, 1), 0).astype(np.float32)
        I = (np.float32(0.4) * m * r * r).astype(np.float32)
        inv_I = np.where(m > 0, np.float32(1.0) / np.where(m > 0, I, 1), 0).astype(np.float32)
        z0 = np.zeros(n, np.float32)
        self.pos4 = np.stack([x, y, z, r], 1).astype(np.float32)
        vx = z0 if vx is None else np.where(m > 0, vx, 0)
        vy = z0 if vy is None else np.where(m > 0, vy, 0)
        vz = z0 if vz is None else np.where(m > 0, vz, 0)
        self.vel4 = np.stack([vx, vy, vz, inv_m], 1).astype(np.float32)
        self.omg4 = np.stack([z0, z0, z0, inv_I], 1).astype(np.float32)
        self.ids = np.arange(n, dtype=np.int32)
        self.hcount = np.zeros(n, np.int32)
        self.hist = np.zeros((max_contacts * n, 4), np.float32)

    def step(self, g, c, nsteps=1, nthreads=1):
        og, oc = grid(g), contact(c)
        mx = 0
        for _ in range(nsteps):
            mx = max(mx, load().or_sphere_step(self.n, C.byref(og), C.byref(oc), p(self.pos4), p(self.vel4),
                                               p(self.omg4), p(self.ids), p(self.hcount), p(self.hist), nthreads))
        return mx

    def reorder_by_id(self):
        """Permutes every row (and each sphere's contact history) into particle-id
        order. Stepping from id order sorts each cell's spheres by id, the order
        the reference's id-ordered arrays and a stable sort give."""
        inv = np.empty(self.n, np.int64)
        inv[self.ids] = np.arange(self.n)
        self.pos4, self.vel4, self.omg4 = self.pos4[inv], self.vel4[inv], self.omg4[inv]
        self.hcount = self.hcount[inv]
        self.hist = np.ascontiguousarray(self.hist.reshape(self.maxc, self.n, 4)[:, inv].reshape(-1, 4))
        self.ids = np.arange(self.n, dtype=np.int32)

    def remove(self, ids):
        """Runtime removal (the reference's EntityHandler::removeEntity is a
        stub, src/engine/entities/entity_handler.cpp:108-119): drops the spheres
        with these ids, renumbers the rest in id order, keeps the slot order of
        the remaining rows and every history entry whose partner stays (sphere
        partners renumbered, walls kept)."""
        ids = np.asarray(ids, np.int64)
        gone = np.zeros(self.n, bool)
        gone[ids] = True
        newid = np.cumsum(~gone) - 1
        newid[gone] = -1
        keep = ~gone[self.ids]
        m = self.n - len(ids)
        hist = self.hist.reshape(self.maxc, self.n, 4)
        hid = hist[:, :, 3].view(np.int32)
        nh = np.zeros((self.maxc, m, 4), np.float32)
        nc = np.zeros(m, np.int32)
        for j, i in enumerate(np.nonzero(keep)[0]):
            h = 0
            for k in range(min(int(self.hcount[i]), self.maxc)):
                pid = int(hid[k, i])
                e = hist[k, i].copy()
                if pid >= 0:
                    if newid[pid] < 0:
                        continue
                    e[3] = np.array([newid[pid]], np.int32).view(np.float32)[0]
                nh[h, j] = e
                h += 1
            nc[j] = h
        self.pos4, self.vel4, self.omg4 = self.pos4[keep], self.vel4[keep], self.omg4[keep]
        self.ids = newid[self.ids[keep]].astype(np.int32)
        self.hcount = nc
        self.hist = nh.reshape(-1, 4)
        self.n = m

    def add(self, x, y, z, r, m, vx=None, vy=None, vz=None):
        """Runtime addition: appends spheres with ids n, n + 1, ... and no
        contact history (1/m, 1/I as at construction)."""
        new = SphereState(x, y, z, r, m, vx, vy, vz, max_contacts=self.maxc)
        n, k = self.n, new.n
        hist = np.zeros((self.maxc, n + k, 4), np.float32)
        hist[:, :n] = self.hist.reshape(self.maxc, n, 4)
        self.pos4 = np.concatenate([self.pos4, new.pos4])
        self.vel4 = np.concatenate([self.vel4, new.vel4])
        self.omg4 = np.concatenate([self.omg4, new.omg4])
        self.ids = np.concatenate([self.ids, new.ids + n]).astype(np.int32)
        self.hcount = np.concatenate([self.hcount, new.hcount])
        self.hist = hist.reshape(-1, 4)
        self.n = n + k

    def by_id(self):
        """State rows reordered to particle-id order."""
        inv = np.empty(self.n, np.int64)
        inv[self.ids] = np.arange(self.n)
        return {"pos4": self.pos4[inv], "vel4": self.vel4[inv], "omg4": self.omg4[inv]}


def body_reduce(bpos4, first, count, spos4, sf4, st4):
    a = [np.ascontiguousarray(x, t) for x, t in ((bpos4, np.float32), (first, np.int32), (count, np.int32),
                                                  (spos4, np.float32), (sf4, np.float32), (st4, np.float32))]
    nb = len(a[0])
    bf, bt = np.zeros((nb, 4), np.float32), np.zeros((nb, 4), np.float32)
    load().or_body_reduce(nb, *[p(x) for x in a], p(bf), p(bt))
    return bf, bt


def integrate_bodies(P, V, Q, L, inv6, f4, t4, c):
    P, V, Q, L = (np.array(x, np.float32) for x in (P, V, Q, L))
    a = [np.ascontiguousarray(x, np.float32) for x in (inv6, f4, t4)]
    oc = contact(c)
    load().or_integrate_bodies(len(P), p(P), p(V), p(Q), p(L), p(a[0]), p(a[1]), p(a[2]), C.byref(oc))
    return P, V, Q, L


def compat_pair_forces(pairs, pos4, vel4, size, kn, kt, damping):
    pairs = np.ascontiguousarray(pairs, np.int32)
    a = [np.ascontiguousarray(x, np.float32) for x in (pos4, vel4, size)]
    n = len(pairs)
    fh, fp = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
    load().or_compat_pair_forces(n, p(pairs), p(a[0]), p(a[1]), p(a[2]), kn, kt, damping, p(fh), p(fp))
    return fh, fp


def compat_inverse3x3(m9):
    m = np.ascontiguousarray(m9, np.float32).reshape(-1, 9)
    o = np.zeros_like(m)
    load().or_compat_inverse3x3(len(m), p(m), p(o))
    return o


def compat_angular_velocity(quat4, inertia9, angmom4):
    a = [np.ascontiguousarray(x, np.float32) for x in (quat4, inertia9, angmom4)]
    o = np.zeros((len(a[0]), 4), np.float32)
    load().or_compat_angular_velocity(len(a[0]), p(a[0]), p(a[1]), p(a[2]), p(o))
    return o
