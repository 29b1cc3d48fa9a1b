"""The kept engine API (STARDUST::DEMEngine, DEMParticle, EntityHandler) through
the example driver stardust_amd/engine/bin/sd_dem_driver (built with the library).

CPU: the packed arrays of the reference driver's own entities
(sandbox/testDEM.cpp:76 and :95) and of setup/star_test.json equal the values
the reference's formulas give (entity.cpp:9-59, entity_handler.cpp:8-98).
GPU: DEMEngine::update steps equal the oracle's clump steps on the same bodies,
packed independently here (sphere radius = diameter / 2; inertia = own
2/5 m r^2 + parallel-axis term, in float64; mass >= 10 static).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from stardust_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "stardust_amd", "engine", "bin", "sd_dem_driver")
STAR = os.path.join(ROOT, "tests", "golden", "star_test.json")


def drive(*args, check=True):
    if not os.path.exists(DRIVER):
        pytest.fail("engine driver not built: run python -m stardust_amd.build")
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=600)
    if check:
        assert r.returncode == 0, r.stderr
        return json.loads(r.stdout)
    return r


def test_reference_driver_entities_kat():
    d = drive("pack")
    assert d["bodies"] == 2 and d["spheres"] == 9
    # DEMParticle(1, 2, 1.0, 0.5, 9, (0.5, 0, 3)): floor(1/0.5)^2 = 4 -> dx = 0.25, 2.25 kg per sphere,
    # lattice mirrored about its centroid -> (0.5 +- 0.125, +-0.125, 3 +- 0.125)
    pos = np.array(d["particle_position"])[:8, :3]
    want = {(0.5 + a, b, 3.0 + c) for a in (-0.125, 0.125) for b in (-0.125, 0.125) for c in (-0.125, 0.125)}
    assert {tuple(p) for p in pos.tolist()} == want
    assert d["particle_mass"][:8] == [2.25] * 8 and d["particle_size"] == [0.5] * 9
    assert np.array_equal(np.array(d["inertia"][0]), np.array([0.5625, 0, 0, 0, 0.5625, 0, 0, 0, 0.5625]))
    # DEMParticle(0, 1, .5, .5, 9, p): one sphere at p, full mass, zero point-mass inertia
    assert d["particle_mass"][8] == 9.0 and d["inertia"][1] == [0.0] * 9
    assert d["particle_relative_position"][8] == [0, 0, 0, 0]
    assert d["entity_start"] == [0, 8] and d["entity_length"] == [8, 1]
    assert d["particle_to_rigid"] == [0] * 8 + [1]


def test_star_test_json_kat():
    d = drive("pack-json", STAR)
    assert d["bodies"] == 2 and d["spheres"] == 2000
    # size = diameter = 0.15: floor(1)^2 = 1 -> dx = 0.15, each sphere carries the full 3 kg
    assert set(d["particle_mass"]) == {3.0}
    assert set(np.array(d["particle_size"], np.float32).tolist()) == {np.float32(0.15).item()}
    pos = np.array(d["particle_position"])[:, :3]
    for b, centre in enumerate(([0.5, 0.5, 0.5], [1.0, 1.0, 1.0])):
        p = pos[1000 * b:1000 * (b + 1)]
        span = p.max(0) - p.min(0)
        assert np.allclose(span, 1.35, atol=1e-5)  # a 10^3 lattice of spacing 0.15
        assert np.allclose(p.mean(0), centre, atol=1e-5)
        assert np.allclose(d["rigid_body_position"][b][:3], centre)
    # the two bodies overlap heavily at t = 0 (their lattices interpenetrate)
    assert pos[:1000].max(0)[0] > pos[1000:].min(0)[0]


def test_json_loader_rejects_unknown_entities(tmp_path):
    f = tmp_path / "bad.json"
    f.write_text(json.dumps({"scene": {"domain": 1, "entities": [{"type": "mesh"}]}}))
    r = drive("pack-json", f, check=False)
    assert r.returncode == 1 and "unsupported entity type" in r.stderr
    f.write_text("{\"scene\": {\"entities\": [}")
    r = drive("pack-json", f, check=False)
    assert r.returncode == 1 and "json" in r.stderr


def clump_input_from_pack(d, static_mass=10.0):
    """Independent restatement of the engine's upload (see module docstring)."""
    nb = d["bodies"]
    start, length = d["entity_start"], d["entity_length"]
    # body-frame offsets (the world ones of an unrotated lattice body)
    rel = np.array(d.get("particle_init_relative_position", d["particle_relative_position"]), np.float32)
    pm = np.array(d["particle_mass"], np.float32)
    rad = (np.float32(0.5) * np.array(d["particle_size"], np.float32)).astype(np.float32)
    s = {"pos4": np.zeros((nb, 4), np.float32), "vel4": np.zeros((nb, 4), np.float32),
         "quat4": np.tile(np.array([1, 0, 0, 0], np.float32), (nb, 1)), "angvel4": np.zeros((nb, 4), np.float32),
         "inertia6": np.zeros((nb, 6), np.float32), "first": np.array(start, np.int32),
         "count": np.array(length, np.int32),
         "local4": np.concatenate([rel[:, :3], rad[:, None]], 1).astype(np.float32)}
    for b in range(nb):
        m = np.float32(d["rigid_body_mass"][b])
        s["pos4"][b, :3] = d["rigid_body_position"][b][:3]
        s["pos4"][b, 3] = 0.0 if m >= static_mass else m
        s["vel4"][b, :3] = d["rigid_body_velocity"][b][:3]
        if "rigid_body_quaternion" in d:
            s["quat4"][b] = d["rigid_body_quaternion"][b]  # (w, x, y, z) as the engine's float4 holds it
            s["angvel4"][b, :3] = d["rigid_body_angular_velocity"][b][:3]
        I = np.zeros(6)
        for k in range(start[b], start[b] + length[b]):
            dx, dy, dz = (float(v) for v in rel[k, :3])
            mj, r = float(pm[k]), float(rad[k])
            own = 0.4 * mj * r * r
            I += [own + mj * (dy * dy + dz * dz), own + mj * (dx * dx + dz * dz), own + mj * (dx * dx + dy * dy),
                  -mj * dx * dy, -mj * dx * dz, -mj * dy * dz]
        s["inertia6"][b] = I.astype(np.float32)
    return s


def run_and_compare(oracle, steps, dt, *scene):
    d = drive("run", steps, dt, *scene)
    assert d["steps"] == steps
    s = clump_input_from_pack(d)
    st = oracle.ClumpState(s)
    c = L.ContactParams.make(2e6, 0.3, 0.5, np.float32(dt), gravity=(0.0, 0.0, -9.81), walls=None,
                             kt_ratio=2.0 / 7.0, max_contacts=16)
    st.step(c, steps)
    for key, arr, w in (("final_position", st.P, 3), ("final_velocity", st.V, 3), ("final_quaternion", st.Q, 4),
                        ("final_angular_momentum", st.L, 3)):
        got = np.array(d[key], np.float32)[:, :w]
        assert np.array_equal(got, arr[:, :w]), (key, got, arr[:, :w])
    return d, st


@pytest.mark.gpu
def test_engine_update_equals_oracle_two_bodies(oracle):
    print("[gpu-case] engine two bodies", flush=True)
    d, st = run_and_compare(oracle, 400, 1e-5)
    # the two dynamic bodies collided (they were closing at 2 m/s with a 2 mm gap)
    v = np.array(d["final_velocity"])
    assert v[0][0] < 0.99 and v[1][0] > -0.99
    # the static body (mass 50 >= 10) did not move
    assert np.array_equal(np.array(d["final_position"][2][:3], np.float32), np.float32([0.016, -0.02, 0.0]))


@pytest.mark.gpu
def test_engine_reference_other_names():
    """addParticle / prepArrays / setUpGrid / transferDataToDevice / step (the
    reference's other names, engine.hpp:172-185) give the run of add / prep /
    transfer / update bit for bit; cleanBuffers releases the device state."""
    print("[gpu-case] engine legacy names", flush=True)
    a = drive("run", 200, 1e-5)
    b = drive("run-legacy", 200, 1e-5)
    assert b["cleaned"] is True
    for key in ("final_position", "final_velocity", "final_quaternion", "final_angular_momentum"):
        assert a[key] == b[key], key


@pytest.mark.gpu
def test_engine_update_star_test_scene(oracle):
    print("[gpu-case] engine star_test.json", flush=True)
    d, st = run_and_compare(oracle, 3, 1e-5, STAR)
    assert np.all(np.isfinite(np.array(d["final_position"])))


def run_spheres_and_compare(oracle, steps, dt, n, domain, size, lo, hi, seed):
    """testDEM.cpp's pattern (single-sphere entities) through DEMEngine: the sphere
    path, bit-identical to the sphere oracle stepped from the same packed input
    (radius = diameter / 2, mass >= 10 static, the engine's default law)."""
    d = drive("run-spheres", steps, dt, n, domain, size, lo, hi, seed)
    assert d["sphere_mode"] is True and d["steps"] == steps
    gr = d["grid"]
    g = L.Grid.make(gr["lo"], np.float32(gr["cell"]), gr["n"])
    walls = (gr["wall_lo"], gr["wall_hi"]) if gr["walls"] else None
    c = L.ContactParams.make(2e6, 0.3, 0.5, np.float32(dt), gravity=(0.0, 0.0, -9.81), walls=walls,
                             kt_ratio=2.0 / 7.0, max_contacts=24)
    pos = np.array(d["rigid_body_position"], np.float32)
    vel = np.array(d["rigid_body_velocity"], np.float32)
    mass = np.array(d["rigid_body_mass"], np.float32)
    r = (np.float32(0.5) * np.array(d["particle_size"], np.float32)).astype(np.float32)
    m = np.where(mass >= 10.0, np.float32(0.0), mass).astype(np.float32)
    st = oracle.SphereState(pos[:, 0], pos[:, 1], pos[:, 2], r, m, vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2],
                            max_contacts=24)
    st.step(g, c, nsteps=steps, nthreads=8)
    exp = st.by_id()
    assert np.array_equal(np.array(d["final_position"], np.float32)[:, :3], exp["pos4"][:, :3])
    assert np.array_equal(np.array(d["final_velocity"], np.float32)[:, :3], exp["vel4"][:, :3])
    I = (np.float32(0.4) * mass * r * r).astype(np.float32)
    L_exp = (I[:, None] * exp["omg4"][:, :3]).astype(np.float32)
    assert np.array_equal(np.array(d["final_angular_momentum"], np.float32)[:, :3], L_exp)
    return d, st


@pytest.mark.gpu
def test_engine_single_spheres_walled_equals_sphere_oracle(oracle):
    print("[gpu-case] engine single spheres, walled domain", flush=True)
    d, st = run_spheres_and_compare(oracle, 40, 1e-4, 6000, 1.0, 0.05, 0.05, 0.95, 11)
    assert d["grid"]["walls"] is True
    assert int(st.hcount.sum()) > 0  # contacts formed


@pytest.mark.gpu
def test_engine_single_spheres_unbounded_equals_sphere_oracle(oracle):
    # the reference driver's own numbers: diameter 0.5, positions uniform in
    # [1, 40]^3, engine domain 0.005 (smaller than the scene: unbounded)
    print("[gpu-case] engine single spheres, unbounded (testDEM.cpp sizes)", flush=True)
    d, st = run_spheres_and_compare(oracle, 30, 5e-4, 100000, 0.005, 0.5, 1.0, 40.0, 5)
    assert d["grid"]["walls"] is False


def sphere_oracle_from_pack(oracle, d, dt):
    gr = d["grid"]
    g = L.Grid.make(gr["lo"], np.float32(gr["cell"]), gr["n"])
    walls = (gr["wall_lo"], gr["wall_hi"]) if gr["walls"] else None
    c = L.ContactParams.make(2e6, 0.3, 0.5, np.float32(dt), gravity=(0.0, 0.0, -9.81), walls=walls,
                             kt_ratio=2.0 / 7.0, max_contacts=24)
    pos = np.array(d["rigid_body_position"], np.float32)
    vel = np.array(d["rigid_body_velocity"], np.float32)
    mass = np.array(d["rigid_body_mass"], np.float32)
    r = (np.float32(0.5) * np.array(d["particle_size"], np.float32)).astype(np.float32)
    m = np.where(mass >= 10.0, np.float32(0.0), mass).astype(np.float32)
    st = oracle.SphereState(pos[:, 0], pos[:, 1], pos[:, 2], r, m, vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2],
                            max_contacts=24)
    return g, c, st, r


@pytest.mark.gpu
def test_engine_runtime_remove_and_add_spheres(oracle):
    """EntityHandler::removeEntityRuntime / addEntityRuntime on the sphere path:
    the engine's state equals the oracle's, which removes the same entities and
    appends the same spheres, bit for bit (contact history kept)."""
    print("[gpu-case] engine runtime edit, spheres", flush=True)
    steps, dt = 25, 1e-4
    d = drive("edit-spheres", steps, dt, 3000, 1.0, 0.05, 0.05, 0.95, 17, 9)
    g, c, st, r = sphere_oracle_from_pack(oracle, d, dt)
    st.step(g, c, nsteps=steps, nthreads=8)
    assert int(st.hcount.sum()) > 0
    st.remove(np.array(d["removed"]))
    st.step(g, c, nsteps=steps, nthreads=8)
    add = np.array(d["added"], np.float32)
    rr = np.full(len(add), np.float32(0.5) * np.float32(0.05), np.float32)
    st.add(add[:, 0], add[:, 1], add[:, 2], rr, add[:, 6], vx=add[:, 3], vy=add[:, 4], vz=add[:, 5])
    st.step(g, c, nsteps=steps, nthreads=8)
    exp = st.by_id()
    assert d["steps"] == 3 * steps
    assert np.array_equal(np.array(d["final_position"], np.float32)[:, :3], exp["pos4"][:, :3])
    assert np.array_equal(np.array(d["final_velocity"], np.float32)[:, :3], exp["vel4"][:, :3])


@pytest.mark.gpu
def test_engine_runtime_remove_and_add_clumps(oracle):
    print("[gpu-case] engine runtime edit, clumps", flush=True)
    steps, dt = 150, 1e-5
    d = drive("edit-clumps", steps, dt)
    st = oracle.ClumpState(clump_input_from_pack(d))
    c = L.ContactParams.make(2e6, 0.3, 0.5, np.float32(dt), gravity=(0.0, 0.0, -9.81), walls=None,
                             kt_ratio=2.0 / 7.0, max_contacts=16)
    st.step(c, steps)
    st.remove([2])
    st.step(c, steps)
    st.add(clump_input_from_pack(d["added"]))
    st.step(c, steps)
    assert d["bodies_after"] == 4
    for key, arr, w in (("final_position", st.P, 3), ("final_velocity", st.V, 3), ("final_quaternion", st.Q, 4),
                        ("final_angular_momentum", st.L, 3)):
        got = np.array(d[key], np.float32)[:, :w]
        assert np.array_equal(got, arr[:, :w]), (key, got, arr[:, :w])


def read_vtk(path):
    """Minimal reader of the legacy binary polydata DEMEngine::writeVTK writes."""
    raw = open(path, "rb").read()
    out, pos = {}, 0

    def line():
        nonlocal pos
        e = raw.index(b"\n", pos)
        t = raw[pos:e].decode()
        pos = e + 1
        return t

    assert line().startswith("# vtk DataFile")
    line()
    assert line() == "BINARY" and line() == "DATASET POLYDATA"
    _, n, _ = line().split()
    n = int(n)
    out["points"] = np.frombuffer(raw, ">f4", 3 * n, pos).reshape(n, 3)
    pos += 12 * n + 1
    assert line().startswith("VERTICES")
    out["verts"] = np.frombuffer(raw, ">i4", 2 * n, pos).reshape(n, 2)
    pos += 8 * n + 1
    assert line() == "POINT_DATA %d" % n and line() == "SCALARS radius float 1" and line() == "LOOKUP_TABLE default"
    out["radius"] = np.frombuffer(raw, ">f4", n, pos)
    pos += 4 * n + 1
    assert line() == "VECTORS body_velocity float"
    out["velocity"] = np.frombuffer(raw, ">f4", 3 * n, pos).reshape(n, 3)
    pos += 12 * n + 1
    assert line() == "SCALARS body int 1" and line() == "LOOKUP_TABLE default"
    out["body"] = np.frombuffer(raw, ">i4", n, pos)
    return out


@pytest.mark.gpu
def test_engine_checkpoint_resume_and_vtk(tmp_path):
    """saveState / loadState: a second engine resumed from the checkpoint ends
    bit-identical to the first; writeVTK holds the spheres' positions, radii,
    body velocities and owners."""
    print("[gpu-case] engine checkpoint", flush=True)
    ck, vtk = tmp_path / "state.bin", tmp_path / "state.vtk"
    d = drive("checkpoint", 120, 1e-5, ck, vtk)
    assert d["resumed_at"] == 120 and d["steps_a"] == d["steps_b"] == 240
    for k in ("position", "quaternion", "angular_momentum", "velocity"):
        assert np.array_equal(np.array(d["a_" + k], np.float32), np.array(d["b_" + k], np.float32)), k
    v = read_vtk(vtk)
    sp = np.array(d["a_sphere_position"], np.float32)
    assert np.array_equal(v["points"], sp[:, :3])
    assert np.array_equal(v["radius"], (np.float32(0.5) * np.array(d["a_sphere_size"], np.float32)))
    owner = np.array(d["a_particle_to_rigid"])
    assert np.array_equal(v["body"], owner)
    assert np.array_equal(v["verts"][:, 1], np.arange(len(sp)))
    assert np.array_equal(v["velocity"], np.array(d["a_velocity"], np.float32)[owner, :3])


def stl_triangles(path):
    """Corners and normals of a binary STL, read with numpy."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, "<u4", 1, 80)[0])
    rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    t = np.frombuffer(raw, rec, n, 84)
    return t["v"].copy(), t["n"].copy()


MESH = os.path.join(ROOT, "tests", "golden", "test_mesh.stl")


def test_mesh_loader_kat():
    """DEMMesh on the reference's test mesh: unique corners in lexicographic
    order, indices that give back every corner of the file, the file's
    normals; bounding spheres of the (right) triangles: the hypotenuse's
    midpoint and half its length."""
    d = drive("mesh-info", MESH)
    tri, nrm = stl_triangles(MESH)
    assert d["n_triangles"] == 12
    v = np.array(d["vertices"], np.float32)[:, :3]
    want = np.unique(tri.reshape(-1, 3), axis=0)  # lexicographic
    assert np.array_equal(v, want)
    idx = np.array(d["indices"]).reshape(-1, 3)
    assert np.array_equal(v[idx], tri)
    assert np.array_equal(np.array(d["normals"], np.float32)[:, :3], nrm)
    for t in range(12):
        p = tri[t].astype(np.float64)
        e = [np.linalg.norm(p[(k + 1) % 3] - p[(k + 2) % 3]) for k in range(3)]
        k = int(np.argmax(e))
        mid = 0.5 * (p[(k + 1) % 3] + p[(k + 2) % 3])
        assert np.allclose(np.array(d["centres"][t][:3]), mid, atol=1e-6)
        assert abs(d["radii"][t] - 0.5 * e[k]) < 1e-5


def test_mesh_loader_ascii_and_acute_circumsphere(tmp_path):
    """An ASCII STL of a tetrahedron with acute faces: the same vertices and
    indices as its binary twin; every bounding sphere passes through the
    three corners (the circumsphere)."""
    P = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.45, 0.9, 0.05], [0.5, 0.35, 0.8]], np.float32)
    faces = [(0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)]
    lines = ["solid tet"]
    for f in faces:
        a, b, c = P[list(f)]
        n = np.cross(b - a, c - a)
        n = n / np.linalg.norm(n)
        lines.append("facet normal %.9g %.9g %.9g" % tuple(n))
        lines.append("outer loop")
        for q in (a, b, c):
            lines.append("vertex %.9g %.9g %.9g" % tuple(q))
        lines += ["endloop", "endfacet"]
    lines.append("endsolid tet")
    asc = tmp_path / "tet_ascii.stl"
    asc.write_text("\n".join(lines) + "\n")
    rec = np.zeros(4, np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    for t, f in enumerate(faces):
        rec["v"][t] = P[list(f)]
    binf = tmp_path / "tet_binary.stl"
    binf.write_bytes(b"\0" * 80 + np.array([4], "<u4").tobytes() + rec.tobytes())
    da, db = drive("mesh-info", asc), drive("mesh-info", binf)
    assert da["n_triangles"] == db["n_triangles"] == 4
    assert da["vertices"] == db["vertices"] and da["indices"] == db["indices"]
    v = np.array(da["vertices"], np.float64)[:, :3]
    idx = np.array(da["indices"]).reshape(-1, 3)
    for t in range(4):
        c, r = np.array(da["centres"][t][:3]), da["radii"][t]
        dist = np.linalg.norm(v[idx[t]] - c, axis=1)
        assert np.allclose(dist, r, rtol=1e-6), (dist, r)


def mesh_drop_bodies(d):
    """The bodies DEMEngine::transfer builds for mesh-drop, restated: the
    particles (from the packed arrays), the cube hull, the mesh triangles."""
    from stardust_amd import hullbody as HB
    pos = np.array(d["rigid_body_position"], np.float32)
    vel = np.array(d["rigid_body_velocity"], np.float32)
    mass = np.array(d["rigid_body_mass"], np.float32)
    size = np.array(d["particle_size"], np.float32)
    n = len(pos)
    sp = HB.triangle_bodies(np.zeros((n, 3, 3), np.float32))
    sp["pos4"][:, :3] = pos[:, :3]
    sp["pos4"][:, 3] = np.where(mass >= 10.0, 0.0, mass)
    sp["vel4"][:, :3] = vel[:, :3]
    r = (np.float32(0.5) * size).astype(np.float32)
    sp["radius"][:] = r
    for b in range(n):
        Is = np.float32(0.4 * float(mass[b]) * float(r[b]) * float(r[b]))
        sp["inertia6"][b, :3] = Is
    sp["nv"][:] = 0
    sp["vstart"][:] = 0
    sp["verts4"] = np.zeros((0, 4), np.float32)
    cube = np.array([[0.3 if k & 1 else -0.3, 0.3 if k & 2 else -0.3, 0.3 if k & 4 else -0.3] for k in range(8)],
                    np.float32)
    hb = HB.triangle_bodies(np.zeros((1, 3, 3), np.float32))
    hb["verts4"] = np.concatenate([cube, np.zeros((8, 1), np.float32)], 1)
    hb["nv"][:] = 8
    hb["pos4"][0] = [-30.5, 18.5, 2.5, 2.0]
    hb["vel4"][0, :3] = [0.0, 0.0, -2.0]
    hb["quat4"][0] = [0.9659258, 0.2588190, 0.0, 0.0]
    pm = 2.0 / 8
    I = [0.0] * 6
    for x, y, z in cube.astype(np.float64):
        I[0] += pm * (y * y + z * z)
        I[1] += pm * (x * x + z * z)
        I[2] += pm * (x * x + y * y)
        I[3] -= pm * x * y
        I[4] -= pm * x * z
        I[5] -= pm * y * z
    hb["inertia6"][0] = np.array(I, np.float32)
    tri, _ = stl_triangles(MESH)
    c = (tri.astype(np.float64).sum(1) / 3.0).astype(np.float32)
    tb = HB.triangle_bodies(tri)
    tb["pos4"][:, :3] = c
    tb["verts4"][:, :3] = (tri - c[:, None, :]).reshape(-1, 3)
    bodies, _ = HB.concat([sp, hb, tb])
    return bodies


@pytest.mark.gpu
def test_engine_mesh_walls_and_hull_equal_oracle(oracle):
    """DEMMesh + DEMHull + single spheres through DEMEngine (the rigid-body
    path): spheres and a cube dropped onto the top face of the reference's
    test mesh equal the oracle's rigid-body steps bit for bit, and come to
    rest on it."""
    print("[gpu-case] engine mesh walls", flush=True)
    steps, dt = 3000, 2e-4
    d = drive("mesh-drop", MESH, steps, dt)
    assert d["hull_mode"] is True and d["steps"] == steps
    bodies = mesh_drop_bodies(d)
    st = oracle.HullBodyState(bodies)
    c = L.ContactParams.make(2e4, 0.3, 0.5, np.float32(dt), gravity=(0.0, 0.0, -9.81), walls=None, max_contacts=16)
    st.step(c, nsteps=steps, nthreads=8)
    n = len(d["final_position"])
    assert np.array_equal(np.array(d["final_position"], np.float32)[:, :3], st.P[:n, :3])
    assert np.array_equal(np.array(d["final_velocity"], np.float32)[:, :3], st.V[:n, :3])
    assert np.array_equal(np.array(d["hull_final"], np.float32)[0, :3], st.P[n, :3])
    assert np.array_equal(np.array(d["hull_final_quaternion"], np.float32)[0], st.Q[n])
    # the spheres came to rest vertically on the top face (z = 1; they may
    # still roll: the law has no rolling resistance)
    z = st.P[:n, 2]
    on_top = np.abs(z - (1.0 + 0.25)) < 0.01
    assert on_top.sum() >= 8, z
    assert np.abs(st.V[:n][on_top, 2]).max() < 0.05


def write_clumps(path, c, rho=2500.0):
    """C3-style clumps (the sd_gen_clumps generator) as the driver's clump
    file: per-sphere masses rho 4/3 pi r^3 (float32), diameters 2 r."""
    lines = [str(len(c["pos4"]))]
    for b in range(len(c["pos4"])):
        f, n = int(c["first"][b]), int(c["count"][b])
        p, v, q, w = c["pos4"][b], c["vel4"][b], c["quat4"][b], c["angvel4"][b]
        lines.append(" ".join(["%d" % n] + ["%.9g" % x for x in (*p[:3], *v[:3], *q, *w[:3])]))
        for k in range(f, f + n):
            r = np.float32(c["local4"][k, 3])
            m = np.float32(rho * 4.0 / 3.0 * np.pi * float(r) ** 3)
            lines.append("%.9g %.9g %.9g %.9g %.9g" % (*c["local4"][k, :3], np.float32(2.0) * r, m))
    path.write_text("\n".join(lines) + "\n")


def quat_rot64(q):
    w, x, y, z = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_engine_general_clumps_pack(tmp_path):
    """DEMParticle from arbitrary spheres (the configuration-3 bodies: 4-16
    spheres, Gaussian offsets, random orientation and spin): the handler holds
    the body-frame offsets, the orientation and the spin as given, the
    spheres' world positions are position + R(q) offset (float64, rounded
    once), and the body mass is the float sum of the sphere masses."""
    from stardust_amd import clumps as CL
    c = CL.generate_clumps(0x5EED0C01, 40, lattice=4)
    f = tmp_path / "clumps.txt"
    write_clumps(f, c)
    d = drive("pack-clumps", f)
    assert d["bodies"] == 40 and d["spheres"] == int(c["count"].sum())
    pos = np.array(d["particle_position"], np.float32)
    init = np.array(d["particle_init_relative_position"], np.float32)
    for b in range(40):
        fst, n = int(c["first"][b]), int(c["count"][b])
        assert d["entity_start"][b] == fst and d["entity_length"][b] == n
        assert np.array_equal(np.float32(d["rigid_body_quaternion"][b]), c["quat4"][b])
        assert np.array_equal(np.float32(d["rigid_body_angular_velocity"][b][:3]), c["angvel4"][b, :3])
        R = quat_rot64(c["quat4"][b])
        X = c["pos4"][b, :3].astype(np.float64)
        m = np.float32(0.0)
        for k in range(fst, fst + n):
            r = c["local4"][k, :3].astype(np.float64)
            assert np.array_equal(init[k, :3], c["local4"][k, :3])
            assert np.array_equal(pos[k, :3], (X + R @ r).astype(np.float32)), (b, k)
            m = np.float32(m + np.float32(d["particle_mass"][k]))
        assert np.float32(d["rigid_body_mass"][b]) == m


@pytest.mark.gpu
def test_engine_general_clumps_equal_oracle(oracle, tmp_path):
    """Configuration-3 style clumps (512 bodies of 4-16 spheres, random
    orientations, velocities N(0, 0.5 m/s), spins) built through the engine API
    and stepped by DEMEngine::update equal the oracle's clump steps on the
    packed bodies, bit for bit."""
    print("[gpu-case] engine general clumps", flush=True)
    from stardust_amd import clumps as CL
    c = CL.generate_clumps(0x5EED0C02, 512, lattice=8, spacing=0.035)
    f = tmp_path / "clumps.txt"
    write_clumps(f, c)
    steps, dt = 60, 1e-5
    d = drive("run-clumps", steps, dt, f)
    assert d["steps"] == steps
    s = clump_input_from_pack(d)
    st = oracle.ClumpState(s)
    cp = L.ContactParams.make(2e6, 0.3, 0.5, np.float32(dt), gravity=(0.0, 0.0, -9.81), walls=None,
                              kt_ratio=2.0 / 7.0, max_contacts=16)
    st.step(cp, steps)
    for key, arr, w in (("final_position", st.P, 3), ("final_velocity", st.V, 3), ("final_quaternion", st.Q, 4),
                        ("final_angular_momentum", st.L, 3)):
        got = np.array(d[key], np.float32)[:, :w]
        assert np.array_equal(got, arr[:, :w]), key
    # the bodies turned: the final orientations differ from the initial ones
    assert not np.array_equal(np.array(d["final_quaternion"], np.float32), c["quat4"])
