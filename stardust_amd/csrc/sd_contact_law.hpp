// The per-contact force law shared by the sphere and clump kernels (see the
// operation-order specification in sd_contact.hip).
#pragma once

#include "sd_internal.hpp"

namespace sd {

struct PairForce {
    float fx, fy, fz, tx, ty, tz;
};

__device__ __forceinline__ PairForce contact_law(const ContactDev& c, float nx, float ny, float nz, float delta,
                                                 float ri, float rj, float inv_mi, float inv_mj, float vrx, float vry,
                                                 float vrz, float wix, float wiy, float wiz, float wjx, float wjy,
                                                 float wjz, float& hx, float& hy, float& hz) {
    const float m_eff = 1.0f / (inv_mi + inv_mj);
    const float cn = c.c2z * sqrtf(c.kn * m_eff);
    const float ct = c.c2z * sqrtf(c.kt * m_eff);
    const float li = ri - 0.5f * delta;
    const float lj = rj - 0.5f * delta;
    const float ax = li * wix + lj * wjx;
    const float ay = li * wiy + lj * wjy;
    const float az = li * wiz + lj * wjz;
    const float cx = ay * nz - az * ny;
    const float cy = az * nx - ax * nz;
    const float cz = ax * ny - ay * nx;
    const float vcx = vrx - cx;
    const float vcy = vry - cy;
    const float vcz = vrz - cz;
    const float vn = vcx * nx + vcy * ny + vcz * nz;
    const float fn = c.kn * delta - cn * vn;
    const float vtx = vcx - vn * nx;
    const float vty = vcy - vn * ny;
    const float vtz = vcz - vn * nz;
    const float hn = hx * nx + hy * ny + hz * nz;
    hx = hx - hn * nx;
    hy = hy - hn * ny;
    hz = hz - hn * nz;
    hx = hx + vtx * c.dt;
    hy = hy + vty * c.dt;
    hz = hz + vtz * c.dt;
    float ftx = -(c.kt * hx) - ct * vtx;
    float fty = -(c.kt * hy) - ct * vty;
    float ftz = -(c.kt * hz) - ct * vtz;
    const float ft2 = ftx * ftx + fty * fty + ftz * ftz;
    const float fmax = c.mu * fabsf(fn);
    if (ft2 > fmax * fmax) {
        const float sc = fmax / sqrtf(ft2);
        ftx = ftx * sc;
        fty = fty * sc;
        ftz = ftz * sc;
        hx = -(ftx + ct * vtx) / c.kt;
        hy = -(fty + ct * vty) / c.kt;
        hz = -(ftz + ct * vtz) / c.kt;
    }
    PairForce f;
    f.fx = fn * nx + ftx;
    f.fy = fn * ny + fty;
    f.fz = fn * nz + ftz;
    f.tx = -li * (ny * ftz - nz * fty);
    f.ty = -li * (nz * ftx - nx * ftz);
    f.tz = -li * (nx * fty - ny * ftx);
    return f;
}

}  // namespace sd
