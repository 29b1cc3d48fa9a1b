// Rigid-body integration shared by the clump and hull contexts: the
// semi-implicit step of advectParticlesCUDA (src/engine/cuda/physics_update.cu:
// 232-310) with the world inverse inertia R I^-1 R^T and a renormalised
// quaternion.
#ifndef SD_RIGID_DEV_HPP
#define SD_RIGID_DEV_HPP

#include "sd_gjk_dev.hpp"
#include "sd_internal.hpp"

namespace sd {

// sin(a), cos(a) for 0 <= a: Taylor polynomials in Horner form below 0.25.
__device__ __forceinline__ void sincos_small(float a, float& s, float& c) {
    if (a < 0.25f) {
        const float a2 = a * a;
        s = a * (1.0f - a2 / 6.0f * (1.0f - a2 / 20.0f * (1.0f - a2 / 42.0f * (1.0f - a2 / 72.0f))));
        c = 1.0f - a2 / 2.0f * (1.0f - a2 / 12.0f * (1.0f - a2 / 30.0f * (1.0f - a2 / 56.0f)));
    } else {
        s = sinf(a);
        c = cosf(a);
    }
}

// World inverse inertia R I^-1 R^T applied to L (I^-1 symmetric, body frame:
// i0 = (xx, yy, zz, xy), i1 = (xz, yz, -, -)).
__device__ __forceinline__ V3 inv_inertia_apply(const Rot& R, float4 i0, float4 i1, V3 L) {
    // u = R^T L, then I^-1 u, then R (.)
    const V3 u = rot_tmul(R, L);
    const V3 w = v3(i0.x * u.x + i0.w * u.y + i1.x * u.z, i0.w * u.x + i0.y * u.y + i1.y * u.z,
                    i1.x * u.x + i1.y * u.y + i0.z * u.z);
    return rot_mul(R, w);
}

// Semi-implicit step of one body with 1/m > 0 under force F and torque T
// (the integration above); a static body (1/m = 0) is left as it is.
__device__ __forceinline__ void integrate_body(float4& P, float4& V, float4& q, float4& L, float4 i0, float4 i1,
                                               float Fx, float Fy, float Fz, float Tx, float Ty, float Tz,
                                               const ContactDev& c) {
    if (!(P.w > 0.0f)) return;
    V.x = V.x + (Fx * P.w + c.gx) * c.dt;
    V.y = V.y + (Fy * P.w + c.gy) * c.dt;
    V.z = V.z + (Fz * P.w + c.gz) * c.dt;
    P.x = P.x + V.x * c.dt;
    P.y = P.y + V.y * c.dt;
    P.z = P.z + V.z * c.dt;
    L.x = L.x + Tx * c.dt;
    L.y = L.y + Ty * c.dt;
    L.z = L.z + Tz * c.dt;
    const Rot R = quat_rot(q);
    const V3 w = inv_inertia_apply(R, i0, i1, v3(L.x, L.y, L.z));
    const float wn = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    if (wn > 0.0f) {
        float sa, ca;
        sincos_small(wn * (0.5f * c.dt), sa, ca);
        const float k = sa / wn;
        const float dw = ca, dx = w.x * k, dy = w.y * k, dz = w.z * k;
        // dq * q
        const float nw = dw * q.x - (dx * q.y + dy * q.z + dz * q.w);
        const float nx = dw * q.y + q.x * dx + (dy * q.w - dz * q.z);
        const float ny = dw * q.z + q.x * dy + (dz * q.y - dx * q.w);
        const float nz = dw * q.w + q.x * dz + (dx * q.z - dy * q.y);
        const float inv = 1.0f / sqrtf(nw * nw + nx * nx + ny * ny + nz * nz);
        q = make_float4(nw * inv, nx * inv, ny * inv, nz * inv);
    }
}

}  // namespace sd

#endif
