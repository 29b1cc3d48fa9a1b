"""Synthetic, seeded sphere workloads of the benchmark configurations.

dem_c1: plain DEM, CPU-runnable. 65,536 monodisperse spheres, r = 0.01 m,
        density 2500 kg/m^3, no initial overlap in a 1 m cube with 6 walls,
        v = 0, g = -9.81 z, kn = 1e5 N/m, zeta = 0.3, mu = 0.5, dt = 1e-5 s,
        1,000 steps, cell = 2r (50^3 cells), 32-bit cell keys.
dem_c2: plain DEM at scale. 4,194,304 polydisperse spheres, radius lognormal
        (median 5 mm, sigma 0.25, clipped to [2, 12] mm), kn = 2e5 N/m,
        dt = 5e-6 s, 2,000 steps, cell = largest diameter (24 mm), packed 64-bit
        (cell, index) keys sorted every step. The stated 0.5 x 0.5 x 4 m column
        cannot hold 4.19M such spheres without overlap (their solid volume is
        about 2.9 m^3), so the column keeps its 4 m height and its footprint is
        widened to 2.0 x 2.0 m (solid fraction 0.18 at seeding, where random
        sequential addition still finishes in about 20 s); the walled domain is
        the whole column.
dem_c2_dense: the dem_c2 law (kn = 2e5 N/m, dt = 5e-6 s, zeta = 0.3, mu = 0.5,
        gravity, six walls) on a contact-dense packing of the same 4,194,304
        spheres: an FCC lattice (128 x 128 x 64 cubic cells of 4 sites) with
        nearest-neighbour spacing 10 mm, 5.05 mm from the walls, radii lognormal
        (median 5 mm, sigma 0.25, numpy PCG64 seeded with 0x5EED0102) clipped
        to [2, 5.05] mm. Neighbours whose radii sum past 10 mm overlap by at
        most 0.1 mm, about 3 contacts per sphere from the first step; the
        walled box (1.81 x 1.81 x 0.91 m) holds the packing together.
batch_c5: 512 independent scenes in one batch. Per scene (numpy PCG64 seeded
        with 0x5EED0005): sphere count log-uniform in [1,024, 65,536]; box
        footprint side U[0.2, 1] m; radius lognormal with median U[2, 8] mm,
        sigma 0.25, clipped to [0.4, 2.4] x median; kn log-uniform in
        [5e4, 5e5] N/m; zeta = 0.3, mu = 0.5, dt = 1e-5 s, 1,000 steps; cell =
        largest diameter. A footprint side of 0.2 m cannot hold 65,536 spheres
        of 8 mm, so each box is a walled column whose height is the larger of
        the side and the height that seeds the spheres at solid fraction 0.15.
        Sphere positions: sd_gen_spheres with seed 0x5EED0005 + 1 + scene.
        A quarter of the scenes (chosen by the same generator) also carry
        U{1,000..8,000} convex hulls of the hulls_c4 kind, uniform in the scene's
        box, generated with seed 0x5EED0005 + 0x10000 + scene.
Seeds are 0x5EED0000 + k for configuration k (1-based).
"""
import math

import numpy as np

from . import _lib as L

DENSITY = 2500.0
KT_RATIO = 2.0 / 7.0


def _gen(seed, n, lo, hi, r_med, r_sigma, r_min, r_max):
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    z = np.empty(n, np.float32)
    r = np.empty(n, np.float32)
    lo_a = np.asarray(lo, np.float32)
    hi_a = np.asarray(hi, np.float32)
    L.call("sd_gen_spheres", seed, n, L.ptr(lo_a), L.ptr(hi_a), r_med, r_sigma, r_min, r_max, L.ptr(x), L.ptr(y),
           L.ptr(z), L.ptr(r))
    return x, y, z, r


def sphere_mass(r):
    r = np.asarray(r, np.float64)
    return (DENSITY * 4.0 / 3.0 * math.pi * r ** 3).astype(np.float32)


def grid_for(lo, hi, cell):
    n = [max(1, int(math.ceil((hi[a] - lo[a]) / cell))) for a in range(3)]
    return L.Grid.make(lo, cell, n)


class Workload:
    def __init__(self, name, n, lo, hi, r_med, r_sigma, r_min, r_max, kn, dt, steps, key64, seed, max_contacts):
        self.name, self.n, self.lo, self.hi = name, n, lo, hi
        self.r_med, self.r_sigma, self.r_min, self.r_max = r_med, r_sigma, r_min, r_max
        self.kn, self.dt, self.steps, self.key64, self.seed = kn, dt, steps, key64, seed
        self.cell = 2.0 * r_max
        self.max_contacts = max_contacts

    def grid(self):
        return grid_for(self.lo, self.hi, self.cell)

    def contact(self):
        return L.ContactParams.make(self.kn, 0.3, 0.5, self.dt, gravity=(0.0, 0.0, -9.81),
                                    walls=(self.lo, self.hi), kt_ratio=KT_RATIO, max_contacts=self.max_contacts)

    def generate(self, n=None):
        n = self.n if n is None else n
        x, y, z, r = _gen(self.seed, n, self.lo, self.hi, self.r_med, self.r_sigma, self.r_min, self.r_max)
        return {"x": x, "y": y, "z": z, "r": r, "m": sphere_mass(r)}

    def describe(self):
        return {"workload": self.name, "spheres": self.n, "steps": self.steps, "dt": self.dt, "kn": self.kn,
                "domain": [list(self.lo), list(self.hi)], "cell": self.cell,
                "keys": "u64 (cell<<32|index)" if self.key64 else "u32 cell + u32 index", "seed": self.seed}


def dem_c1(n=65536):
    return Workload("dem_c1_plain_65536", n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.01, 0.0, 0.01, 0.01, 1e5, 1e-5,
                    1000, False, 0x5EED0001, 16)


def dem_c2(n=4194304):
    return Workload("dem_c2_column_4194304", n, (0.0, 0.0, 0.0), (2.0, 2.0, 4.0), 0.005, 0.25, 0.002, 0.012, 2e5,
                    5e-6, 2000, True, 0x5EED0002, 24)


class LatticeWorkload(Workload):
    """Spheres on an FCC lattice (a contact-dense packing) under a sphere law."""

    def __init__(self, name, cells, spacing, r_med, r_sigma, r_min, r_max, kn, dt, steps, seed, max_contacts):
        self.cells3 = tuple(int(c) for c in cells)
        self.spacing = spacing
        self.a = spacing * math.sqrt(2.0)  # cubic cell edge of the FCC lattice
        self.margin = r_max
        n = 4 * self.cells3[0] * self.cells3[1] * self.cells3[2]
        hi = tuple(2.0 * self.margin + (c - 0.5) * self.a for c in self.cells3)
        super().__init__(name, n, (0.0, 0.0, 0.0), hi, r_med, r_sigma, r_min, r_max, kn, dt, steps, True, seed,
                         max_contacts)

    def generate(self, n=None):
        assert n is None or n == self.n, "the lattice fixes the sphere count"
        nx, ny, nz = self.cells3
        basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        cell = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float64)
        pts = (cell[:, None, :] + basis[None, :, :]).reshape(-1, 3) * self.a + self.margin
        rng = np.random.Generator(np.random.PCG64(self.seed))
        r = np.exp(math.log(self.r_med) + self.r_sigma * rng.standard_normal(self.n))
        r = np.clip(r, self.r_min, self.r_max).astype(np.float32)
        p = pts.astype(np.float32)
        return {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "r": r, "m": sphere_mass(r)}

    def describe(self):
        d = super().describe()
        d.update({"packing": "FCC lattice, nearest-neighbour spacing %g m" % self.spacing,
                  "lattice_cells": list(self.cells3)})
        return d


def dem_c2_dense(cells=(128, 128, 64)):
    return LatticeWorkload("dem_c2_dense_%d" % (4 * cells[0] * cells[1] * cells[2]), cells, 0.01, 0.005, 0.25,
                           0.002, 0.00505, 2e5, 5e-6, 2000, 0x5EED0102, 24)


class BatchWorkload:
    """A batch of independent sphere scenes (see batch_c5)."""

    def __init__(self, name, nscenes, seed, n_min=1024, n_max=65536, side=(0.2, 1.0), r_med=(0.002, 0.008),
                 kn=(5e4, 5e5), r_sigma=0.25, phi_seed=0.15, dt=1e-5, steps=1000, max_contacts=16):
        self.name, self.nscenes, self.seed, self.dt, self.steps = name, nscenes, seed, dt, steps
        self.max_contacts = max_contacts
        rng = np.random.Generator(np.random.PCG64(seed))
        self.scene_n = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), nscenes))).astype(np.int64)
        self.side = rng.uniform(side[0], side[1], nscenes)
        self.r_med = rng.uniform(r_med[0], r_med[1], nscenes)
        self.kn = np.exp(rng.uniform(math.log(kn[0]), math.log(kn[1]), nscenes))
        self.r_sigma, self.phi_seed = r_sigma, phi_seed
        solid = self.scene_n * (4.0 / 3.0) * math.pi * self.r_med ** 3 * math.exp(4.5 * r_sigma ** 2)
        self.height = np.maximum(self.side, solid / (phi_seed * self.side ** 2))
        self.n = int(self.scene_n.sum())
        self.hull_scenes = np.sort(rng.choice(nscenes, nscenes // 4, replace=False)) if nscenes >= 4 else \
            np.zeros(0, np.int64)
        self.hull_n = rng.integers(1000, 8001, len(self.hull_scenes))

    def generate_hull_scene(self, j):
        """Hulls of the j-th hull-carrying scene."""
        from . import hulls as H
        k = int(self.hull_scenes[j])
        return H.generate_hulls(self.seed + 0x10000 + k, int(self.hull_n[j]), lo=(0.0, 0.0, 0.0),
                                hi=(float(self.side[k]), float(self.side[k]), float(self.height[k])))

    def hull_bodies(self, j, spacing=0.020):
        """Rigid bodies of the j-th hull-carrying scene: its hulls (8-32
        vertices on ellipsoids with semi-axes U[5, 15] mm, seeded velocities
        and spins) on a cubic lattice of the given spacing over the scene's
        floor, and the twelve static triangles of the walls of a box around
        them (the scene's footprint, tall enough for the lattice). At the
        default 20 mm spacing (twice the mean semi-axis) neighbours overlap by
        up to a few mm: about 0.5 contacts per hull from the first step; 32 mm
        leaves every pair apart."""
        from . import hullbody as HB
        h = self.generate_hull_scene(j)
        k = int(self.hull_scenes[j])
        side = float(self.side[k])
        m = max(1, int(side / spacing))
        n = len(h["pos4"])
        layers = -(-n // (m * m))
        idx = np.arange(n)
        h["pos4"][:, 0] = ((idx % m) + 0.5) * spacing
        h["pos4"][:, 1] = (((idx // m) % m) + 0.5) * spacing
        h["pos4"][:, 2] = ((idx // (m * m)) + 0.5) * spacing
        b = HB.make_bodies(h, self.seed + 0x20000 + k)
        top = max(float(self.height[k]), (layers + 1) * spacing)
        walls = HB.box_walls((0.0, 0.0, 0.0), (m * spacing, m * spacing, top))
        out, _ = HB.concat([b, walls])
        return out

    def hull_set(self):
        """Bodies of every hull-carrying scene, one scene each, and the scene
        boundaries."""
        from . import hullbody as HB
        return HB.concat([self.hull_bodies(j) for j in range(len(self.hull_scenes))])

    def scene(self, k):
        lo = (0.0, 0.0, 0.0)
        hi = (float(self.side[k]), float(self.side[k]), float(self.height[k]))
        r_max = 2.4 * float(self.r_med[k])
        g = grid_for(lo, hi, 2.0 * r_max)
        c = L.ContactParams.make(float(self.kn[k]), 0.3, 0.5, self.dt, gravity=(0.0, 0.0, -9.81), walls=(lo, hi),
                                 kt_ratio=KT_RATIO, max_contacts=self.max_contacts)
        return int(self.scene_n[k]), g, c, lo, hi

    def cells(self):
        return sum(self.scene(k)[1].n_cells for k in range(self.nscenes))

    def generate_scene(self, k):
        n, g, c, lo, hi = self.scene(k)
        rm = float(self.r_med[k])
        x, y, z, r = _gen(self.seed + 1 + k, n, lo, hi, rm, self.r_sigma, 0.4 * rm, 2.4 * rm)
        return {"x": x, "y": y, "z": z, "r": r, "m": sphere_mass(r)}

    def generate(self, scenes=None):
        """Concatenated arrays of the given scenes (default: all) and their descriptors."""
        ks = range(self.nscenes) if scenes is None else scenes
        parts = [self.generate_scene(k) for k in ks]
        cat = {key: np.concatenate([p[key] for p in parts]) for key in ("x", "y", "z", "r", "m")}
        desc = [self.scene(k)[:3] for k in ks]
        return cat, desc


def batch_c5(nscenes=512):
    return BatchWorkload("batch_c5_512_scenes", nscenes, 0x5EED0005)
