"""Pins the compat restatements of the reference's own arithmetic (CPU):
hand-derived known answers of computeParticleToParticleCUDA
(sap_reaction.cu:20-70), compute3x3Inverse (cuda_helpers.cuh:65-77) and the
angular velocity of advectParticlesCUDA (physics_update.cu:271-277), quirks
included."""
import math

import numpy as np


def test_pair_force_known_answers(oracle):
    kn, kt, c = 2e6, 20.0, 2000.0
    pos = np.array([[0.8, 0, 0, 0], [0, 0, 0, 0], [5, 5, 5, 0], [5, 5, 5, 0], [9, 0, 0, 0], [7, 0, 0, 0]],
                   np.float32)
    vel = np.array([[-1, 0.5, 0, 0], [0, 0, 0, 0], [0.3, -0.2, 0.1, 0], [0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]],
                   np.float32)
    size = np.full(6, 0.5, np.float32)  # diameters, summed as the separation (1.0)
    pairs = np.array([[0, 1], [2, 3], [4, 5]], np.int32)
    # static pair first: same positions, no velocity
    v0 = np.zeros_like(vel)
    fh, fp = oracle.compat_pair_forces(pairs[:1], pos, v0, size, kn, kt, c)
    # overlap 1.0 - 0.8 = 0.2 along n = (1, 0, 0): -kn * 0.2 * n; zero velocity has sign +1,
    # so the home sphere is pulled towards the other one (the reference's sign)
    assert np.allclose(fh[0], [-4e5, 0, 0, 0], rtol=1e-6) and np.allclose(fp[0], [4e5, 0, 0, 0], rtol=1e-6)
    fh, fp = oracle.compat_pair_forces(pairs, pos, vel, size, kn, kt, c)
    # v = (-1, 0.5, 0, 0): signs (-1, +1, +1, +1); vt = v - (v * n) * n = (0, 0.5, 0, 0);
    # total = (-4e5 - 2000, 1000 + 10, 0, 0); home force = signs * total
    assert np.allclose(fh[0], [402000.0, 1010.0, 0, 0], rtol=1e-6)
    assert np.allclose(fp[0], [-402000.0, -1010.0, 0, 0], rtol=1e-6)
    # coincident spheres: n = 0, so total = c v + kt v, times the signs of v
    v = np.array([0.3, -0.2, 0.1], np.float64)
    want = np.sign(np.where(v == 0, 1, v)) * (c * v + kt * v)
    assert np.allclose(fh[1, :3], want, rtol=1e-6)
    # separated (2 > 1): no force
    assert np.all(fh[2] == 0) and np.all(fp[2] == 0)


def test_inverse_known_answers(oracle):
    m = np.array([np.diag([2.0, 4.0, 8.0]).ravel(), np.diag([1e-3] * 3).ravel(), np.diag([-1.0, 2.0, 3.0]).ravel()],
                 np.float32)
    o = oracle.compat_inverse3x3(m)
    assert np.array_equal(o[0], np.diag([0.5, 0.25, 0.125]).ravel().astype(np.float32))
    # determinant 1e-9 < 1e-4: the input comes back (quirk Q10), and so for any negative determinant
    assert np.array_equal(o[1], m[1]) and np.array_equal(o[2], m[2])
    # the threshold is the double 1e-4: det = 1e-4f (below 1e-4 in double) returns the input,
    # the next float up is inverted
    a0 = np.float32(1e-4)
    a1 = np.nextafter(a0, np.float32(1))
    b = np.array([np.diag([a0, 1, 1]).ravel(), np.diag([a1, 1, 1]).ravel()], np.float32)
    o = oracle.compat_inverse3x3(b)
    assert np.array_equal(o[0], b[0]) and abs(o[1][0] * a1 - 1) < 1e-6
    # a general symmetric positive definite matrix: the adjugate formula is a true inverse
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3))
    S = (A @ A.T + 3 * np.eye(3)).astype(np.float32)
    o = oracle.compat_inverse3x3(S.ravel()[None])
    assert np.allclose(o[0].reshape(3, 3), np.linalg.inv(S.astype(np.float64)), rtol=1e-5, atol=1e-6)


def test_angular_velocity_uses_transposed_rotation(oracle):
    I = np.diag([2.0, 4.0, 8.0])
    L = np.array([1.0, 0.3, -0.2, 0.0], np.float32)
    th = math.radians(30.0)
    q = np.array([[1, 0, 0, 0], [math.cos(th / 2), 0, 0, math.sin(th / 2)]], np.float32)
    w = oracle.compat_angular_velocity(q, np.stack([I.ravel()] * 2).astype(np.float32), np.stack([L, L]))
    assert np.allclose(w[0, :3], np.linalg.inv(I) @ L[:3], rtol=1e-6)
    # the row-major rotation read by column-major helpers: w = R^T I^-1 R L (quirk Q11),
    # not the correct R I^-1 R^T L
    c, s = math.cos(th), math.sin(th)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    ref = R.T @ np.linalg.inv(I) @ R @ L[:3]
    correct = R @ np.linalg.inv(I) @ R.T @ L[:3]
    assert np.allclose(w[1, :3], ref, rtol=1e-6) and not np.allclose(w[1, :3], correct, rtol=1e-3)
    # a mm-scale body: det(I) < 1e-4, the inertia itself is applied (Q10 with Q11)
    Is = np.diag([1e-4, 2e-4, 3e-4])
    w = oracle.compat_angular_velocity(q[1:], Is.ravel()[None].astype(np.float32), L[None])
    assert np.allclose(w[0, :3], R.T @ Is @ R @ L[:3], rtol=1e-5)
