// Compat restatements of the reference's own arithmetic (no step path calls
// them). The library's steps use the configurations' contact law and the
// corrected rigid-body formulas; these entry points compute what the
// reference computes instead, quirks included, so that the parity tests can
// show its formulas are reproduced bit for bit:
//  - sd_compat_pair_forces: computeParticleToParticleCUDA
//    (src/engine/cuda/collisions/SAP/sap_reaction.cu:20-70): separation = sum
//    of the `size` fields (the diameters), component-wise tangential
//    projection, force multiplied component-wise by the signs of the relative
//    velocity (zero counts as positive); returned per pair instead of the
//    reference's unsynchronised scatter into both spheres;
//  - sd_compat_inverse3x3: compute3x3Inverse (src/engine/cuda/cuda_helpers.cuh:
//    65-77), which returns its input when the determinant is below 1e-4;
//  - sd_compat_angular_velocity: the world angular velocity of
//    advectParticlesCUDA (src/engine/cuda/physics_update.cu:271-277), whose
//    rotation matrix is filled row-major (quatToRotationCUDA,
//    cuda_helpers.cuh:114-128) and read column-major by the matrix helpers
//    (cuda_helpers.cuh:79-112).
// float4 arithmetic is component-wise and length() includes the w lane, as in
// the reference's helper_math.hpp.
#include "sd_internal.hpp"

namespace sd {

__device__ __forceinline__ float ref_sign(float v) { return (v > 0.0f) ? 1.0f : (v == 0.0f ? 1.0f : -1.0f); }

__global__ void k_compat_pair_forces(const int* __restrict__ pairs, int npairs, const float4* __restrict__ pos,
                                     const float4* __restrict__ vel, const float* __restrict__ size, float kn,
                                     float kt, float damping, float4* __restrict__ fh, float4* __restrict__ fp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int h = pairs[2 * (size_t)k], p = pairs[2 * (size_t)k + 1];
    const float4 ph = pos[h], pp = pos[p], vh = vel[h], vp = vel[p];
    const float4 rel = make_float4(ph.x - pp.x, ph.y - pp.y, ph.z - pp.z, ph.w - pp.w);
    const float4 rv = make_float4(vh.x - vp.x, vh.y - vp.y, vh.z - vp.z, vh.w - vp.w);
    const float distance = sqrtf(rel.x * rel.x + rel.y * rel.y + rel.z * rel.z + rel.w * rel.w);
    const float separation = size[h] + size[p];
    float4 nf = make_float4(0.0f, 0.0f, 0.0f, 0.0f), df = nf, tf = nf;
    if (distance <= separation) {
        float4 n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!(distance == 0.0f)) n = make_float4(rel.x / distance, rel.y / distance, rel.z / distance, rel.w / distance);
        const float4 tv = make_float4(rv.x - (rv.x * n.x) * n.x, rv.y - (rv.y * n.y) * n.y, rv.z - (rv.z * n.z) * n.z,
                                      rv.w - (rv.w * n.w) * n.w);
        const float s = -kn * (separation - distance);
        nf = make_float4(s * n.x, s * n.y, s * n.z, s * n.w);
        df = make_float4(damping * rv.x, damping * rv.y, damping * rv.z, damping * rv.w);
        tf = make_float4(kt * tv.x, kt * tv.y, kt * tv.z, kt * tv.w);
    }
    const float4 sv = make_float4(ref_sign(rv.x), ref_sign(rv.y), ref_sign(rv.z), ref_sign(rv.w));
    const float4 tot = make_float4((nf.x + df.x) + tf.x, (nf.y + df.y) + tf.y, (nf.z + df.z) + tf.z,
                                   (nf.w + df.w) + tf.w);
    fh[k] = make_float4(sv.x * tot.x, sv.y * tot.y, sv.z * tot.z, sv.w * tot.w);
    fp[k] = make_float4(-sv.x * tot.x, -sv.y * tot.y, -sv.z * tot.z, -sv.w * tot.w);
}

// `1e-4 > determinant` compares in double in the reference; for a float
// determinant that is `determinant <= 1e-4f` (the largest float below 1e-4).
__device__ __forceinline__ void ref_inverse9(const float* m, float* o) {
    const float c11 = m[4] * m[8] - m[7] * m[5], c12 = m[7] * m[2] - m[1] * m[8], c13 = m[1] * m[5] - m[4] * m[2];
    const float det = m[0] * c11 + m[3] * c12 + m[6] * c13;
    if (det <= 1e-4f) {
#pragma unroll
        for (int i = 0; i < 9; ++i) o[i] = m[i];
        return;
    }
    const float s = 1.0f / det;
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

__device__ __forceinline__ void ref_mul9(const float* a, const float* b, float* o) {
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

__global__ void k_compat_inverse(const float* __restrict__ m9, int n, float* __restrict__ o9) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float m[9], o[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = m9[9 * (size_t)i + k];
    ref_inverse9(m, o);
#pragma unroll
    for (int k = 0; k < 9; ++k) o9[9 * (size_t)i + k] = o[k];
}

__global__ void k_compat_angular_velocity(const float4* __restrict__ quat, const float* __restrict__ inertia9,
                                          const float4* __restrict__ angm, int nb, float4* __restrict__ omega) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    float I[9], inv[9], R[9], RT[9], RI[9], M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) I[k] = inertia9[9 * (size_t)b + k];
    ref_inverse9(I, inv);
    const float4 q = quat[b];
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    R[0] = 1.0f - 2.0f * (z * z) - 2.0f * (w * w);
    R[1] = 2.0f * y * z - 2.0f * x * w;
    R[2] = 2.0f * y * w + 2.0f * x * z;
    R[3] = 2.0f * y * z + 2.0f * x * w;
    R[4] = 1.0f - 2.0f * (y * y) - 2.0f * (w * w);
    R[5] = 2.0f * z * w - 2.0f * x * y;
    R[6] = 2.0f * y * w - 2.0f * x * z;
    R[7] = 2.0f * z * w + 2.0f * x * y;
    R[8] = 1.0f - 2.0f * (y * y) - 2.0f * (z * z);
    RT[0] = R[0];
    RT[1] = R[3];
    RT[2] = R[6];
    RT[3] = R[1];
    RT[4] = R[4];
    RT[5] = R[7];
    RT[6] = R[2];
    RT[7] = R[5];
    RT[8] = R[8];
    ref_mul9(R, inv, RI);
    ref_mul9(RI, RT, M);
    const float4 L = angm[b];
    omega[b] = make_float4(M[0] * L.x + M[3] * L.y + M[6] * L.z, M[1] * L.x + M[4] * L.y + M[7] * L.z,
                           M[2] * L.x + M[5] * L.y + M[8] * L.z, 0.0f);
}

}  // namespace sd

using namespace sd;

extern "C" sd_status sd_compat_pair_forces(const float* pos4, const float* vel4, const float* size, const int* pairs,
                                           int npairs, float kn, float kt, float damping, float* force_home4,
                                           float* force_phantom4, void* stream) {
    SD_REQUIRE(npairs >= 0, "sd_compat_pair_forces: npairs < 0");
    if (npairs == 0) return SD_OK;
    SD_REQUIRE(pos4 && vel4 && size && pairs && force_home4 && force_phantom4, "sd_compat_pair_forces: null pointer");
    hipLaunchKernelGGL(k_compat_pair_forces, dim3(div_up(npairs, 256)), dim3(256), 0, (hipStream_t)stream, pairs,
                       npairs, (const float4*)pos4, (const float4*)vel4, size, kn, kt, damping, (float4*)force_home4,
                       (float4*)force_phantom4);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_compat_inverse3x3(const float* m9, int n, float* out9, void* stream) {
    SD_REQUIRE(n >= 0, "sd_compat_inverse3x3: n < 0");
    if (n == 0) return SD_OK;
    SD_REQUIRE(m9 && out9, "sd_compat_inverse3x3: null pointer");
    hipLaunchKernelGGL(k_compat_inverse, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, m9, n, out9);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_compat_angular_velocity(const float* quat4, const float* inertia9, const float* angmom4, int nb,
                                                float* omega4, void* stream) {
    SD_REQUIRE(nb >= 0, "sd_compat_angular_velocity: nb < 0");
    if (nb == 0) return SD_OK;
    SD_REQUIRE(quat4 && inertia9 && angmom4 && omega4, "sd_compat_angular_velocity: null pointer");
    hipLaunchKernelGGL(k_compat_angular_velocity, dim3(div_up(nb, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float4*)quat4, inertia9, (const float4*)angmom4, nb, (float4*)omega4);
    SD_LAUNCH_CHECK();
    return SD_OK;
}
