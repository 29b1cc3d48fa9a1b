"""GPU compat entry points (the reference's own arithmetic, quirks included)
against the oracle restatements, bit for bit: computeParticleToParticleCUDA
(sap_reaction.cu:20-70), compute3x3Inverse (cuda_helpers.cuh:65-77), the
angular velocity of advectParticlesCUDA (physics_update.cu:271-277)."""
import sys

import numpy as np
import pytest

from stardust_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    L.require_gpu()


def dbuf(a):
    return L.DeviceBuffer.from_array(np.ascontiguousarray(a))


def test_compat_pair_forces_match_reference_arithmetic(oracle):
    say("compat pair forces")
    rng = np.random.default_rng(8)
    n, npairs = 20000, 100000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, 2, (n, 3))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1, (n, 3))
    vel[::7, 1] = 0.0  # zero components: sign +1
    size = rng.uniform(0.05, 0.5, n).astype(np.float32)
    pairs = rng.integers(0, n, (npairs, 2)).astype(np.int32)
    pairs[::11, 1] = pairs[::11, 0]  # coincident: distance 0
    eh, ep = oracle.compat_pair_forces(pairs, pos, vel, size, 2e6, 20.0, 2000.0)
    fh, fp = L.DeviceBuffer(16 * npairs), L.DeviceBuffer(16 * npairs)
    ins = [dbuf(a) for a in (pos, vel, size, pairs)]
    L.call("sd_compat_pair_forces", ins[0].ptr, ins[1].ptr, ins[2].ptr, ins[3].ptr, npairs, 2e6, 20.0, 2000.0,
           fh.ptr, fp.ptr, None)
    L.synchronize()
    gh = fh.download(np.float32, 4 * npairs).reshape(-1, 4)
    gp = fp.download(np.float32, 4 * npairs).reshape(-1, 4)
    assert (np.abs(eh).sum(1) > 0).mean() > 0.05  # contacts among the pairs
    assert np.array_equal(gh.view(np.uint32), eh.view(np.uint32))
    assert np.array_equal(gp.view(np.uint32), ep.view(np.uint32))


def test_compat_inverse_matches_reference_arithmetic(oracle):
    say("compat inverse")
    rng = np.random.default_rng(9)
    n = 50000
    A = rng.normal(size=(n, 3, 3))
    S = np.einsum("nij,nkj->nik", A, A) * rng.choice([1e-9, 1e-3, 1.0, 100.0], n)[:, None, None]
    S[::5] *= -1.0  # negative determinants
    m = S.reshape(n, 9).astype(np.float32)
    e = oracle.compat_inverse3x3(m)
    o = L.DeviceBuffer(36 * n)
    md = dbuf(m)
    L.call("sd_compat_inverse3x3", md.ptr, n, o.ptr, None)
    L.synchronize()
    g = o.download(np.float32, 9 * n).reshape(-1, 9)
    assert np.array_equal(g.view(np.uint32), e.view(np.uint32))
    assert (g == m).all(1).mean() > 0.3  # the determinant rule returned inputs


def test_compat_angular_velocity_matches_reference_arithmetic(oracle):
    say("compat angular velocity")
    rng = np.random.default_rng(10)
    nb = 50000
    q = rng.normal(size=(nb, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    A = rng.normal(size=(nb, 3, 3))
    I = (np.einsum("nij,nkj->nik", A, A) + np.eye(3)) * rng.choice([1e-3, 1.0], nb)[:, None, None]
    I = I.reshape(nb, 9).astype(np.float32)
    Lm = np.zeros((nb, 4), np.float32)
    Lm[:, :3] = rng.normal(0, 1, (nb, 3))
    e = oracle.compat_angular_velocity(q, I, Lm)
    o = L.DeviceBuffer(16 * nb)
    ins = [dbuf(a) for a in (q, I, Lm)]
    L.call("sd_compat_angular_velocity", ins[0].ptr, ins[1].ptr, ins[2].ptr, nb, o.ptr, None)
    L.synchronize()
    g = o.download(np.float32, 4 * nb).reshape(-1, 4)
    assert np.array_equal(g.view(np.uint32), e.view(np.uint32))
