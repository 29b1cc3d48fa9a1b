/*
 * sd_oracle_compat.c — CPU restatement of the reference's OWN arithmetic for
 * the formulas the build replaces (TEST INFRASTRUCTURE ONLY; see sd_oracle.h).
 *
 * The product follows the configurations' contact law and the corrected rigid
 * body formulas. These functions restate, operation for operation, what the
 * reference computes instead, quirks included, so that tests can show the GPU
 * compat entry points (sd_compat_*) reproduce the reference's bits:
 *  - the pair force of computeParticleToParticleCUDA
 *    (src/engine/cuda/collisions/SAP/sap_reaction.cu:20-70): the diameter
 *    summed as the separation, a component-wise tangential projection, the
 *    force multiplied component-wise by the signs of the relative velocity;
 *  - compute3x3Inverse (src/engine/cuda/cuda_helpers.cuh:65-77): the input is
 *    returned when the determinant is below 1e-4;
 *  - the world angular velocity of advectParticlesCUDA
 *    (src/engine/cuda/physics_update.cu:271-277): quatToRotationCUDA fills a
 *    row-major matrix (cuda_helpers.cuh:114-128) that the column-major helpers
 *    (computeMatrixMultiplication, computeTranspose,
 *    computeMatrixVectorMultiplication, cuda_helpers.cuh:79-112) read as its
 *    transpose.
 * float4 arithmetic is component-wise (helper_math.hpp:863-881,1007-1010) and
 * length() includes the w lane (helper_math.hpp:1236-1238,1279-1282).
 */
#include <math.h>

#include "sd_oracle.h"

typedef struct {
    float x, y, z, w;
} F4;

static F4 f4(const float* p) {
    F4 r = {p[0], p[1], p[2], p[3]};
    return r;
}

/* getSignedFloat (cuda_helpers.cuh:43-53): +1 for positive and zero, -1 otherwise */
static float sgn(float v) { return (v > 0.0f) ? 1.0f : (v == 0.0f ? 1.0f : -1.0f); }

/* Per pair (home, phantom): the contributions the reference adds to
 * force[home] (sgn(v) * total) and force[phantom] (-sgn(v) * total). */
void or_compat_pair_forces(int npairs, const int* pairs, const float* pos4, const float* vel4, const float* size,
                           float kn, float kt, float damping, float* fh4, float* fp4) {
    for (int k = 0; k < npairs; ++k) {
        const int h = pairs[2 * (long)k], p = pairs[2 * (long)k + 1];
        const F4 ph = f4(pos4 + 4 * (long)h), pp = f4(pos4 + 4 * (long)p);
        const F4 vh = f4(vel4 + 4 * (long)h), vp = f4(vel4 + 4 * (long)p);
        const F4 rel = {ph.x - pp.x, ph.y - pp.y, ph.z - pp.z, ph.w - pp.w};
        const F4 rv = {vh.x - vp.x, vh.y - vp.y, vh.z - vp.z, vh.w - vp.w};
        const float distance = sqrtf(rel.x * rel.x + rel.y * rel.y + rel.z * rel.z + rel.w * rel.w);
        const float separation = size[h] + size[p];
        F4 nf = {0, 0, 0, 0}, df = {0, 0, 0, 0}, tf = {0, 0, 0, 0};
        if (distance <= separation) {
            F4 n = {0, 0, 0, 0};
            if (!(distance == 0.0f)) {
                n.x = rel.x / distance;
                n.y = rel.y / distance;
                n.z = rel.z / distance;
                n.w = rel.w / distance;
            }
            /* relative_velocity - (relative_velocity * n) * n, component-wise */
            const F4 tv = {rv.x - (rv.x * n.x) * n.x, rv.y - (rv.y * n.y) * n.y, rv.z - (rv.z * n.z) * n.z,
                           rv.w - (rv.w * n.w) * n.w};
            const float s = -kn * (separation - distance);
            nf.x = s * n.x;
            nf.y = s * n.y;
            nf.z = s * n.z;
            nf.w = s * n.w;
            df.x = damping * rv.x;
            df.y = damping * rv.y;
            df.z = damping * rv.z;
            df.w = damping * rv.w;
            tf.x = kt * tv.x;
            tf.y = kt * tv.y;
            tf.z = kt * tv.z;
            tf.w = kt * tv.w;
        }
        const F4 sv = {sgn(rv.x), sgn(rv.y), sgn(rv.z), sgn(rv.w)};
        const F4 tot = {(nf.x + df.x) + tf.x, (nf.y + df.y) + tf.y, (nf.z + df.z) + tf.z, (nf.w + df.w) + tf.w};
        float* a = fh4 + 4 * (long)k;
        float* b = fp4 + 4 * (long)k;
        a[0] = sv.x * tot.x;
        a[1] = sv.y * tot.y;
        a[2] = sv.z * tot.z;
        a[3] = sv.w * tot.w;
        b[0] = -sv.x * tot.x;
        b[1] = -sv.y * tot.y;
        b[2] = -sv.z * tot.z;
        b[3] = -sv.w * tot.w;
    }
}

/* compute3x3Inverse: `1e-4 > determinant` compares in double; for a float
 * determinant that is `determinant <= 1e-4f` (1e-4f is the largest float below
 * 1e-4), which is what is evaluated here. */
static void inverse9(const float* m, float* o) {
    const float c11 = m[4] * m[8] - m[7] * m[5], c12 = m[7] * m[2] - m[1] * m[8], c13 = m[1] * m[5] - m[4] * m[2];
    const float det = m[0] * c11 + m[3] * c12 + m[6] * c13;
    if (det <= 1e-4f) {
        for (int i = 0; i < 9; ++i) o[i] = m[i];
        return;
    }
    const float s = 1 / det;
    o[0] = s * c11;
    o[1] = s * c12;
    o[2] = s * c13;
    o[3] = s * m[6] * m[5] - s * m[3] * m[8];
    o[4] = s * m[0] * m[8] - s * m[6] * m[2];
    o[5] = s * m[3] * m[2] - s * m[0] * m[5];
    o[6] = s * m[3] * m[7] - s * m[6] * m[4];
    o[7] = s * m[6] * m[1] - s * m[0] * m[7];
    o[8] = s * m[0] * m[4] - s * m[3] * m[1];
}

void or_compat_inverse3x3(int n, const float* m9, float* o9) {
    for (int i = 0; i < n; ++i) inverse9(m9 + 9 * (long)i, o9 + 9 * (long)i);
}

/* computeMatrixMultiplication (cuda_helpers.cuh:89-103) */
static void mul9(const float* a, const float* b, float* o) {
    o[0] = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
    o[1] = a[1] * b[0] + a[4] * b[1] + a[7] * b[2];
    o[2] = a[2] * b[0] + a[5] * b[1] + a[8] * b[2];
    o[3] = a[0] * b[3] + a[3] * b[4] + a[6] * b[5];
    o[4] = a[1] * b[3] + a[4] * b[4] + a[7] * b[5];
    o[5] = a[2] * b[3] + a[5] * b[4] + a[8] * b[5];
    o[6] = a[0] * b[6] + a[3] * b[7] + a[6] * b[8];
    o[7] = a[1] * b[6] + a[4] * b[7] + a[7] * b[8];
    o[8] = a[2] * b[6] + a[5] * b[7] + a[8] * b[8];
}

/* quatToRotationCUDA, q = (s, vx, vy, vz) in (x, y, z, w) */
static void rot9(const float* q, float* r) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    r[0] = 1 - 2 * (z * z) - 2 * (w * w);
    r[1] = 2 * y * z - 2 * x * w;
    r[2] = 2 * y * w + 2 * x * z;
    r[3] = 2 * y * z + 2 * x * w;
    r[4] = 1 - 2 * (y * y) - 2 * (w * w);
    r[5] = 2 * z * w - 2 * x * y;
    r[6] = 2 * y * w - 2 * x * z;
    r[7] = 2 * z * w + 2 * x * y;
    r[8] = 1 - 2 * (y * y) - 2 * (z * z);
}

/* angular velocity of advectParticlesCUDA (physics_update.cu:271-277):
 * inverse(I) with the determinant rule, rotation_inverse = R * inv(I),
 * M = rotation_inverse * transpose(R), w = M L, with the column-major helpers */
void or_compat_angular_velocity(int nb, const float* quat4, const float* inertia9, const float* angmom4,
                                float* omega4) {
    for (int b = 0; b < nb; ++b) {
        float inv[9], R[9], RT[9], RI[9], M[9];
        inverse9(inertia9 + 9 * (long)b, inv);
        rot9(quat4 + 4 * (long)b, R);
        RT[0] = R[0];
        RT[1] = R[3];
        RT[2] = R[6];
        RT[3] = R[1];
        RT[4] = R[4];
        RT[5] = R[7];
        RT[6] = R[2];
        RT[7] = R[5];
        RT[8] = R[8];
        mul9(R, inv, RI);
        mul9(RI, RT, M);
        const float* L = angmom4 + 4 * (long)b;
        float* o = omega4 + 4 * (long)b;
        o[0] = M[0] * L[0] + M[3] * L[1] + M[6] * L[2];
        o[1] = M[1] * L[0] + M[4] * L[1] + M[7] * L[2];
        o[2] = M[2] * L[0] + M[5] * L[1] + M[8] * L[2];
        o[3] = 0.0f;
    }
}
