"""Runtime add / remove of spheres on the GPU (sd_ctx_remove_spheres,
sd_ctx_add_spheres) against the oracle's rule, bit for bit: a contact-dense
lattice stepped, spheres removed (some of them in contact, so history entries
are dropped and partner ids renumbered), stepped, spheres added, stepped;
32-bit and packed 64-bit keys, graph-replayed steps (re-captured after each
edit)."""
import sys

import numpy as np
import pytest

from stardust_amd import _lib as L
from stardust_amd import workloads as W
from stardust_amd.scene import SphereScene

pytestmark = pytest.mark.gpu

QUANTITIES = {"x": ("pos4", 0), "y": ("pos4", 1), "z": ("pos4", 2), "vx": ("vel4", 0), "vy": ("vel4", 1),
              "vz": ("vel4", 2), "wx": ("omg4", 0), "wy": ("omg4", 1), "wz": ("omg4", 2)}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    L.require_gpu()


def assert_equal_state(sc, st, what):
    got, exp = sc.download(), st.by_id()
    assert len(got["x"]) == st.n
    for k, (arr, col) in QUANTITIES.items():
        assert np.array_equal(got[k], exp[arr][:, col]), (what, k)


@pytest.mark.parametrize("key64", [False, True])
def test_remove_then_add_matches_oracle(oracle, key64):
    w = W.dem_c2_dense((6, 6, 6))
    s = w.generate()
    g, c = w.grid(), w.contact()
    rng = np.random.default_rng(21)
    v = rng.normal(0.0, 0.3, (3, w.n)).astype(np.float32)
    say("edit", "key64" if key64 else "key32", w.n)
    sc = SphereScene(g, c, key64=key64, use_graph=True)
    sc.upload(s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2])
    st = oracle.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], vx=v[0], vy=v[1], vz=v[2],
                            max_contacts=w.max_contacts)
    sc.step(10)
    st.step(g, c, nsteps=10)
    assert_equal_state(sc, st, "before")
    gone = rng.choice(w.n, 60, replace=False).astype(np.int32)
    sc.remove(gone)
    st.remove(gone)
    sc.step(10)
    st.step(g, c, nsteps=10)
    assert_equal_state(sc, st, "after remove")
    k = 40
    ax = (w.margin + rng.random(k) * (w.hi[0] - 2 * w.margin)).astype(np.float32)
    ay = (w.margin + rng.random(k) * (w.hi[1] - 2 * w.margin)).astype(np.float32)
    az = (w.margin + rng.random(k) * (w.hi[2] - 2 * w.margin)).astype(np.float32)
    ar = np.full(k, 0.003, np.float32)
    am = W.sphere_mass(ar)
    am[::7] = 0.0  # some static
    av = rng.normal(0.0, 0.3, (3, k)).astype(np.float32)
    ids = sc.add(ax, ay, az, ar, am, vx=av[0], vy=av[1], vz=av[2])
    st.add(ax, ay, az, ar, am, vx=av[0], vy=av[1], vz=av[2])
    assert ids.tolist() == list(range(st.n - k, st.n))
    sc.step(11)
    st.step(g, c, nsteps=11)
    assert_equal_state(sc, st, "after add")
    assert int(st.hcount.sum()) > 0
    sc.close()


def test_edit_errors():
    w = W.dem_c2_dense((2, 2, 2))
    s = w.generate()
    sc = SphereScene(w.grid(), w.contact(), key64=True)
    sc.upload(s["x"], s["y"], s["z"], s["r"], s["m"])
    with pytest.raises(Exception, match="out of range"):
        sc.remove([w.n])
    with pytest.raises(Exception, match="twice"):
        sc.remove([1, 1])
    with pytest.raises(Exception, match="cell size"):
        sc.add([0.01], [0.01], [0.01], [w.cell], [1.0])
    sc.close()
