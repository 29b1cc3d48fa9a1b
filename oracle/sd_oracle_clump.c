/*
 * sd_oracle_clump.c — CPU oracle for rigid clumps and the LBVH broadphase
 * (TEST INFRASTRUCTURE ONLY; see sd_oracle.h).
 *
 * Restates, as plain C with the documented FP32 operation order:
 *   - the clump pose, force/torque reduction and rigid integration of
 *     src/engine/cuda/physics_update.cu:26-348 with the corrected inertia
 *     (R I^-1 R^T, quirks Q10/Q11) and quaternion renormalisation (Q12);
 *   - the Morton keys of sandbox/testLBVH.cu:91-171 on a 1024^3 grid over the
 *     global domain (:61-88).
 * The broadphase pair set is computed by serial sweep-and-prune over the same
 * body boxes: the LBVH must return exactly this set (every overlapping pair once).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sd_oracle.h"

typedef struct {
    float x, y, z;
} V3;
static V3 mk(float x, float y, float z) {
    V3 r = {x, y, z};
    return r;
}

static void qmat(const float* q, float m[9]) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    m[0] = 1.0f - 2.0f * (yy + zz);
    m[1] = 2.0f * (xy - wz);
    m[2] = 2.0f * (xz + wy);
    m[3] = 2.0f * (xy + wz);
    m[4] = 1.0f - 2.0f * (xx + zz);
    m[5] = 2.0f * (yz - wx);
    m[6] = 2.0f * (xz - wy);
    m[7] = 2.0f * (yz + wx);
    m[8] = 1.0f - 2.0f * (xx + yy);
}
static V3 mv(const float m[9], V3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
static V3 mtv(const float m[9], V3 v) {
    return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

/* w = R I^-1 R^T L, I^-1 = [[a, d, e], [d, b, f], [e, f, c]] from i0 = (a, b, c, d), i1 = (e, f) */
static V3 inv_inertia(const float m[9], const float* i0, const float* i1, V3 L) {
    const V3 u = mtv(m, L);
    const V3 w = mk(i0[0] * u.x + i0[3] * u.y + i1[0] * u.z, i0[3] * u.x + i0[1] * u.y + i1[1] * u.z,
                    i1[0] * u.x + i1[1] * u.y + i0[2] * u.z);
    return mv(m, w);
}

static void sincos_small(float a, float* s, float* c) {
    if (a < 0.25f) {
        const float a2 = a * a;
        *s = a * (1.0f - a2 / 6.0f * (1.0f - a2 / 20.0f * (1.0f - a2 / 42.0f * (1.0f - a2 / 72.0f))));
        *c = 1.0f - a2 / 2.0f * (1.0f - a2 / 12.0f * (1.0f - a2 / 30.0f * (1.0f - a2 / 56.0f)));
    } else {
        *s = sinf(a);
        *c = cosf(a);
    }
}

/* Upload-time preparation, double precision on the host:
 * P = (x, y, z, 1/m), V, Q, L = R I R^T w, I^-1 body frame as i0/i1. */
void or_clump_prepare(int nb, const float* pos4, const float* vel4, const float* quat4, const float* angvel4,
                      const float* inertia6, float* P, float* V, float* Q, float* L, float* I0, float* I1) {
    for (int b = 0; b < nb; ++b) {
        const float m = pos4[4 * (long)b + 3];
        P[4 * (long)b] = pos4[4 * (long)b];
        P[4 * (long)b + 1] = pos4[4 * (long)b + 1];
        P[4 * (long)b + 2] = pos4[4 * (long)b + 2];
        P[4 * (long)b + 3] = m > 0.0f ? 1.0f / m : 0.0f;
        for (int a = 0; a < 3; ++a) V[4 * (long)b + a] = m > 0.0f ? vel4[4 * (long)b + a] : 0.0f;
        V[4 * (long)b + 3] = 0.0f;
        for (int a = 0; a < 4; ++a) Q[4 * (long)b + a] = quat4[4 * (long)b + a];
        const float* I = inertia6 + 6 * (long)b;
        const double xx = I[0], yy = I[1], zz = I[2], xy = I[3], xz = I[4], yz = I[5];
        const double c00 = yy * zz - yz * yz, c01 = xz * yz - xy * zz, c02 = xy * yz - xz * yy;
        const double det = xx * c00 + xy * c01 + xz * c02;
        double inv[6] = {0, 0, 0, 0, 0, 0};
        if (m > 0.0f && det > 0.0) {
            inv[0] = c00 / det;
            inv[1] = (xx * zz - xz * xz) / det;
            inv[2] = (xx * yy - xy * xy) / det;
            inv[3] = c01 / det;
            inv[4] = c02 / det;
            inv[5] = (xz * xy - xx * yz) / det;
        }
        I0[4 * (long)b] = (float)inv[0];
        I0[4 * (long)b + 1] = (float)inv[1];
        I0[4 * (long)b + 2] = (float)inv[2];
        I0[4 * (long)b + 3] = (float)inv[3];
        I1[4 * (long)b] = (float)inv[4];
        I1[4 * (long)b + 1] = (float)inv[5];
        I1[4 * (long)b + 2] = 0.0f;
        I1[4 * (long)b + 3] = 0.0f;
        const float* q = quat4 + 4 * (long)b;
        const float* w = angvel4 + 4 * (long)b;
        const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
        const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy),
                             2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                             2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)};
        const double Ib[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
        double u[3], t[3], Lw[3];
        for (int r = 0; r < 3; ++r) u[r] = R[0 + r] * w[0] + R[3 + r] * w[1] + R[6 + r] * w[2];
        for (int r = 0; r < 3; ++r) t[r] = Ib[3 * r] * u[0] + Ib[3 * r + 1] * u[1] + Ib[3 * r + 2] * u[2];
        for (int r = 0; r < 3; ++r) Lw[r] = R[3 * r] * t[0] + R[3 * r + 1] * t[1] + R[3 * r + 2] * t[2];
        for (int a = 0; a < 3; ++a) L[4 * (long)b + a] = m > 0.0f ? (float)Lw[a] : 0.0f;
        L[4 * (long)b + 3] = 0.0f;
    }
}

/* Sphere world positions / velocities and body AABBs of the current state. */
void or_clump_pose(int nb, const float* P, const float* V, const float* Q, const float* L, const float* I0,
                   const float* I1, const int* first, const int* count, const float* local4, float* sx, float* sv,
                   float* lo4, float* hi4) {
    for (int b = 0; b < nb; ++b) {
        float m[9];
        qmat(Q + 4 * (long)b, m);
        const V3 w = inv_inertia(m, I0 + 4 * (long)b, I1 + 4 * (long)b,
                                 mk(L[4 * (long)b], L[4 * (long)b + 1], L[4 * (long)b + 2]));
        float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
        float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
        for (int s = first[b]; s < first[b] + count[b]; ++s) {
            const float* l = local4 + 4 * (long)s;
            const V3 r = mv(m, mk(l[0], l[1], l[2]));
            float* x = sx + 4 * (long)s;
            float* v = sv + 4 * (long)s;
            x[0] = P[4 * (long)b] + r.x;
            x[1] = P[4 * (long)b + 1] + r.y;
            x[2] = P[4 * (long)b + 2] + r.z;
            x[3] = l[3];
            v[0] = V[4 * (long)b] + (w.y * r.z - w.z * r.y);
            v[1] = V[4 * (long)b + 1] + (w.z * r.x - w.x * r.z);
            v[2] = V[4 * (long)b + 2] + (w.x * r.y - w.y * r.x);
            v[3] = 0.0f;
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], x[a] - x[3]);
                hi[a] = fmaxf(hi[a], x[a] + x[3]);
            }
        }
        for (int a = 0; a < 3; ++a) {
            lo4[4 * (long)b + a] = lo[a];
            hi4[4 * (long)b + a] = hi[a];
        }
        lo4[4 * (long)b + 3] = 0.0f;
        hi4[4 * (long)b + 3] = 0.0f;
    }
}

void or_global_domain(int n, const float* lo4, const float* hi4, float* dom6) {
    for (int a = 0; a < 3; ++a) {
        dom6[a] = 3.402823466e38f;
        dom6[3 + a] = -3.402823466e38f;
    }
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            dom6[a] = fminf(dom6[a], lo4[4 * (long)i + a]);
            dom6[3 + a] = fmaxf(dom6[3 + a], hi4[4 * (long)i + a]);
        }
}

static uint32_t spread10(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0xFF0000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

static uint32_t maxis(float c, float lo, float hi) {
    const float ext = hi - lo;
    float f = ext > 0.0f ? floorf(((c - lo) / ext) * 1024.0f) : 0.0f;
    f = fmaxf(f, 0.0f);
    f = fminf(f, 1023.0f);
    return (uint32_t)f;
}

void or_morton30(int n, const float* lo4, const float* hi4, const float* dom6, uint64_t* keys) {
    for (int i = 0; i < n; ++i) {
        const float* l = lo4 + 4 * (long)i;
        const float* h = hi4 + 4 * (long)i;
        const float cx = 0.5f * (l[0] + h[0]), cy = 0.5f * (l[1] + h[1]), cz = 0.5f * (l[2] + h[2]);
        const uint32_t code = spread10(maxis(cx, dom6[0], dom6[3])) | (spread10(maxis(cy, dom6[1], dom6[4])) << 1) |
                              (spread10(maxis(cz, dom6[2], dom6[5])) << 2);
        keys[i] = ((uint64_t)code << 32) | (uint64_t)(uint32_t)i;
    }
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}

/* Contacts + integration of every body given the unique pair set (a, b).
 * Adjacency: (body, partner) ascending; per contact the sphere law with no
 * history; torque about the centre of mass. State arrays updated in place. */
/* Semi-implicit step of one body (P, V, q, L rows; P[3] = 1/m, 0 = static) under
 * force F and torque T: v += (F/m + g) dt, x += v dt, L += T dt, w = R I^-1 R^T L,
 * q = normalise(dq q) (physics_update.cu:232-310 with quirks Q10-Q12 corrected). */
static void integrate_body(float* Pn, float* Vn, float* q, float* Ln, const float* i0, const float* i1,
                           const float* F, const float* T, const or_contact* c) {
    if (!(Pn[3] > 0.0f)) return;
    Vn[0] = Vn[0] + (F[0] * Pn[3] + c->gravity[0]) * c->dt;
    Vn[1] = Vn[1] + (F[1] * Pn[3] + c->gravity[1]) * c->dt;
    Vn[2] = Vn[2] + (F[2] * Pn[3] + c->gravity[2]) * c->dt;
    Pn[0] = Pn[0] + Vn[0] * c->dt;
    Pn[1] = Pn[1] + Vn[1] * c->dt;
    Pn[2] = Pn[2] + Vn[2] * c->dt;
    Ln[0] = Ln[0] + T[0] * c->dt;
    Ln[1] = Ln[1] + T[1] * c->dt;
    Ln[2] = Ln[2] + T[2] * c->dt;
    float m[9];
    qmat(q, m);
    const V3 w = inv_inertia(m, i0, i1, mk(Ln[0], Ln[1], Ln[2]));
    const float wn = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    if (wn > 0.0f) {
        float sa, ca;
        sincos_small(wn * (0.5f * c->dt), &sa, &ca);
        const float k = sa / wn;
        const float dw = ca, dx = w.x * k, dy = w.y * k, dz = w.z * k;
        const float nw = dw * q[0] - (dx * q[1] + dy * q[2] + dz * q[3]);
        const float nx = dw * q[1] + q[0] * dx + (dy * q[3] - dz * q[2]);
        const float ny = dw * q[2] + q[0] * dy + (dz * q[1] - dx * q[3]);
        const float nz = dw * q[3] + q[0] * dz + (dx * q[2] - dy * q[1]);
        const float inv = 1.0f / sqrtf(nw * nw + nx * nx + ny * ny + nz * nz);
        q[0] = nw * inv;
        q[1] = nx * inv;
        q[2] = ny * inv;
        q[3] = nz * inv;
    }
}

/* Per-body sums of per-sphere forces (computeForcesAndTorquesCUDA,
 * physics_update.cu:144-195, without its gravity and static rule, which the
 * integration applies): F = sum f, T = sum ((x_s - X) x f + tau_s), spheres
 * of [first, first + count) in order. */
void or_body_reduce(int nb, const float* bpos4, const int* first, const int* count, const float* spos4,
                    const float* sf4, const float* st4, float* bf4, float* bt4) {
    for (int b = 0; b < nb; ++b) {
        const float* X = bpos4 + 4 * (long)b;
        float F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        for (int i = first[b]; i < first[b] + count[b]; ++i) {
            const float* x = spos4 + 4 * (long)i;
            const float* f = sf4 + 4 * (long)i;
            const float* t = st4 + 4 * (long)i;
            F[0] = F[0] + f[0];
            F[1] = F[1] + f[1];
            F[2] = F[2] + f[2];
            const float ax = x[0] - X[0], ay = x[1] - X[1], az = x[2] - X[2];
            T[0] = T[0] + ((ay * f[2] - az * f[1]) + t[0]);
            T[1] = T[1] + ((az * f[0] - ax * f[2]) + t[1]);
            T[2] = T[2] + ((ax * f[1] - ay * f[0]) + t[2]);
        }
        for (int a = 0; a < 3; ++a) {
            bf4[4 * (long)b + a] = F[a];
            bt4[4 * (long)b + a] = T[a];
        }
        bf4[4 * (long)b + 3] = 0.0f;
        bt4[4 * (long)b + 3] = 0.0f;
    }
}

/* One step of every body under given forces and torques; inverse inertia in
 * the body frame (xx, yy, zz, xy, xz, yz). */
void or_integrate_bodies(int nb, float* P, float* V, float* Q, float* L, const float* inv6, const float* f4,
                         const float* t4, const or_contact* c) {
    for (int b = 0; b < nb; ++b) {
        const float* I = inv6 + 6 * (long)b;
        const float i0[4] = {I[0], I[1], I[2], I[3]}, i1[4] = {I[4], I[5], 0.0f, 0.0f};
        integrate_body(P + 4 * (long)b, V + 4 * (long)b, Q + 4 * (long)b, L + 4 * (long)b, i0, i1,
                       f4 + 4 * (long)b, t4 + 4 * (long)b, c);
    }
}

int or_clump_contacts_integrate(int nb, float* P, float* V, float* Q, float* L, const float* I0, const float* I1,
                                const int* first, const int* count, const float* sx, const float* sv, long npairs,
                                const int* pairs, const or_contact* c, float* force4, float* torque4, int integrate) {
    const float kt = c->kt_ratio * c->kn, c2z = 2.0f * c->damping_ratio;
    uint64_t* adj = (uint64_t*)malloc((size_t)(2 * npairs + 1) * sizeof(uint64_t));
    for (long p = 0; p < npairs; ++p) {
        const uint64_t a = (uint64_t)pairs[2 * p], b = (uint64_t)pairs[2 * p + 1];
        adj[2 * p] = a * (uint64_t)nb + b;
        adj[2 * p + 1] = b * (uint64_t)nb + a;
    }
    qsort(adj, (size_t)(2 * npairs), sizeof(uint64_t), cmp_u64);
    long e = 0;
    float* np_ = (float*)malloc((size_t)nb * 16 * sizeof(float));
    /* omega of every body at the start of the step */
    float* om = (float*)malloc((size_t)nb * 3 * sizeof(float));
    for (int b = 0; b < nb; ++b) {
        float m[9];
        qmat(Q + 4 * (long)b, m);
        const V3 w = inv_inertia(m, I0 + 4 * (long)b, I1 + 4 * (long)b,
                                 mk(L[4 * (long)b], L[4 * (long)b + 1], L[4 * (long)b + 2]));
        om[3 * b] = w.x;
        om[3 * b + 1] = w.y;
        om[3 * b + 2] = w.z;
    }
    for (int b = 0; b < nb; ++b) {
        const float* X = P + 4 * (long)b;
        float F[3] = {0.0f, 0.0f, 0.0f}, T[3] = {0.0f, 0.0f, 0.0f};
        while (e < 2 * npairs && (int)(adj[e] / (uint64_t)nb) == b) {
            const int pb = (int)(adj[e] % (uint64_t)nb);
            ++e;
            const float inv_mp = P[4 * (long)pb + 3];
            if (X[3] + inv_mp == 0.0f) continue;
            for (int i = first[b]; i < first[b] + count[b]; ++i) {
                const float* xi = sx + 4 * (long)i;
                const float* vi = sv + 4 * (long)i;
                for (int j = first[pb]; j < first[pb] + count[pb]; ++j) {
                    const float* xj = sx + 4 * (long)j;
                    const float dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    const float rs = xi[3] + xj[3];
                    if (!(d2 < rs * rs) || d2 == 0.0f) continue;
                    const float* vj = sv + 4 * (long)j;
                    const float d = sqrtf(d2);
                    const float inv_d = 1.0f / d;
                    const float nx = dx * inv_d, ny = dy * inv_d, nz = dz * inv_d;
                    const float delta = rs - d;
                    /* contact law (sd_oracle.c order), zero tangential history */
                    const float m_eff = 1.0f / (X[3] + inv_mp);
                    const float cn = c2z * sqrtf(c->kn * m_eff);
                    const float ct = c2z * sqrtf(kt * m_eff);
                    const float li = xi[3] - 0.5f * delta, lj = xj[3] - 0.5f * delta;
                    const float ax = li * om[3 * b] + lj * om[3 * pb];
                    const float ay = li * om[3 * b + 1] + lj * om[3 * pb + 1];
                    const float az = li * om[3 * b + 2] + lj * om[3 * pb + 2];
                    const float vcx = (vi[0] - vj[0]) - (ay * nz - az * ny);
                    const float vcy = (vi[1] - vj[1]) - (az * nx - ax * nz);
                    const float vcz = (vi[2] - vj[2]) - (ax * ny - ay * nx);
                    const float vn = vcx * nx + vcy * ny + vcz * nz;
                    const float fn = c->kn * delta - cn * vn;
                    const float vtx = vcx - vn * nx, vty = vcy - vn * ny, vtz = vcz - vn * nz;
                    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
                    const float hn = hx * nx + hy * ny + hz * nz;
                    hx = hx - hn * nx;
                    hy = hy - hn * ny;
                    hz = hz - hn * nz;
                    hx = hx + vtx * c->dt;
                    hy = hy + vty * c->dt;
                    hz = hz + vtz * c->dt;
                    float ftx = -(kt * hx) - ct * vtx, fty = -(kt * hy) - ct * vty, ftz = -(kt * hz) - ct * vtz;
                    const float ft2 = ftx * ftx + fty * fty + ftz * ftz;
                    const float fmax = c->mu * fabsf(fn);
                    if (ft2 > fmax * fmax) {
                        const float sc = fmax / sqrtf(ft2);
                        ftx = ftx * sc;
                        fty = fty * sc;
                        ftz = ftz * sc;
                    }
                    const float fx = fn * nx + ftx, fy = fn * ny + fty, fz = fn * nz + ftz;
                    const float tx = -li * (ny * ftz - nz * fty);
                    const float ty = -li * (nz * ftx - nx * ftz);
                    const float tz = -li * (nx * fty - ny * ftx);
                    F[0] = F[0] + fx;
                    F[1] = F[1] + fy;
                    F[2] = F[2] + fz;
                    const float rx = xi[0] - X[0], ry = xi[1] - X[1], rz = xi[2] - X[2];
                    T[0] = T[0] + ((ry * fz - rz * fy) + tx);
                    T[1] = T[1] + ((rz * fx - rx * fz) + ty);
                    T[2] = T[2] + ((rx * fy - ry * fx) + tz);
                }
            }
        }
        if (force4) {
            for (int a = 0; a < 3; ++a) {
                force4[4 * (long)b + a] = F[a];
                torque4[4 * (long)b + a] = T[a];
            }
            force4[4 * (long)b + 3] = 0.0f;
            torque4[4 * (long)b + 3] = 0.0f;
        }
        float* o = np_ + 16 * (long)b;
        memcpy(o, P + 4 * (long)b, 16);
        memcpy(o + 4, V + 4 * (long)b, 16);
        memcpy(o + 8, Q + 4 * (long)b, 16);
        memcpy(o + 12, L + 4 * (long)b, 16);
        if (integrate) integrate_body(o, o + 4, o + 8, o + 12, I0 + 4 * (long)b, I1 + 4 * (long)b, F, T, c);
    }
    if (integrate)
        for (int b = 0; b < nb; ++b) {
            memcpy(P + 4 * (long)b, np_ + 16 * (long)b, 16);
            memcpy(V + 4 * (long)b, np_ + 16 * (long)b + 4, 16);
            memcpy(Q + 4 * (long)b, np_ + 16 * (long)b + 8, 16);
            memcpy(L + 4 * (long)b, np_ + 16 * (long)b + 12, 16);
        }
    free(adj);
    free(np_);
    free(om);
    return 0;
}

/* One full clump step (pose, AABBs, SAP pair set, contacts, integration).
 * Returns the number of unique body pairs. */
long or_clump_step(int nb, float* P, float* V, float* Q, float* L, const float* I0, const float* I1, const int* first,
                   const int* count, const float* local4, int ns, const or_contact* c) {
    float* sx = (float*)malloc((size_t)ns * 16);
    float* sv = (float*)malloc((size_t)ns * 16);
    float* lo = (float*)malloc((size_t)nb * 16);
    float* hi = (float*)malloc((size_t)nb * 16);
    or_clump_pose(nb, P, V, Q, L, I0, I1, first, count, local4, sx, sv, lo, hi);
    const int axis = or_sap_axis(nb, lo, hi);
    const long np = or_sap_pairs(nb, lo, hi, axis, NULL, 0);
    int* pairs = (int*)malloc((size_t)(2 * np + 2) * sizeof(int));
    or_sap_pairs(nb, lo, hi, axis, pairs, np);
    or_clump_contacts_integrate(nb, P, V, Q, L, I0, I1, first, count, sx, sv, np, pairs, c, NULL, NULL, 1);
    free(sx);
    free(sv);
    free(lo);
    free(hi);
    free(pairs);
    return np;
}

/* ------------------------------------------------------------------ rigid convex bodies
 * One step of the hull-body context (sd_hull_step), in its order: shape
 * centres and world angular velocities, hull boxes, sweep-and-prune of every
 * scene on its own axis (scene by scene, ids global), cold GJK/EPA, the
 * contact of every touching pair, the per-body sums in pair order and the
 * integration. The narrowphase replaces the boolean gjkCUDA of the reference
 * sandbox (sandbox/testGJK.cu:237-315); the contact force and integration
 * follow the clump path's law and advectParticlesCUDA
 * (src/engine/cuda/physics_update.cu:232-310). */
#define HB_FEATURE_REL 1e-3f

static V3 hb_add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 hb_sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 hb_scl(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static float hb_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 hb_cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

/* contact feature along u: the vertices within HB_FEATURE_REL times the
 * shape's radius (largest vertex distance from its centre) of the extreme one
 * along u, their mean (vertex order, world) and the
 * largest squared distance of one of them from it; a sphere's X + r u */
static V3 hb_feature(const float* verts4, int nv, float radius, const float m[9], V3 X, V3 u, float* spread2) {
    *spread2 = 0.0f;
    if (nv == 0) return hb_add(X, hb_scl(u, radius));
    const V3 l = mtv(m, u);
    float dmax = -3.402823466e38f, r2 = 0.0f;
    for (int k = 0; k < nv; ++k) {
        const float* v = verts4 + 4 * (long)k;
        const float d = v[0] * l.x + v[1] * l.y + v[2] * l.z;
        dmax = fmaxf(dmax, d);
        r2 = fmaxf(r2, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    const float cut = dmax - HB_FEATURE_REL * sqrtf(r2);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int cnt = 0;
    for (int k = 0; k < nv; ++k) {
        const float* v = verts4 + 4 * (long)k;
        const float d = v[0] * l.x + v[1] * l.y + v[2] * l.z;
        if (d >= cut) {
            sx = sx + v[0];
            sy = sy + v[1];
            sz = sz + v[2];
            ++cnt;
        }
    }
    const float inv = 1.0f / (float)cnt;
    const float mx = sx * inv, my = sy * inv, mz = sz * inv;
    for (int k = 0; k < nv; ++k) {
        const float* v = verts4 + 4 * (long)k;
        const float d = v[0] * l.x + v[1] * l.y + v[2] * l.z;
        if (d >= cut) {
            const float ex = v[0] - mx, ey = v[1] - my, ez = v[2] - mz;
            *spread2 = fmaxf(*spread2, ex * ex + ey * ey + ez * ez);
        }
    }
    return hb_add(mv(m, mk(mx, my, mz)), X);
}

/* the centroid of the feature of smaller spread; the midpoint on a tie */
static V3 hb_contact_point(V3 ca, float wa, V3 cb, float wb) {
    const float t = wa < wb ? 0.0f : (wb < wa ? 1.0f : 0.5f);
    return hb_add(ca, hb_scl(hb_sub(cb, ca), t));
}

typedef struct {
    int p;
    int e;
} HbEntry;
static int cmp_entry(const void* a, const void* b) {
    const int x = ((const HbEntry*)a)->e, y = ((const HbEntry*)b)->e;
    return (x > y) - (x < y);
}

/* ---------------------------------------------------------- body broadphase */

/* Bounding radius of a body about its centre: the largest vertex distance
 * (body frame), or the sphere radius of a body without vertices. */
static float hb_bound(const float* verts4, int nv, float radius) {
    if (nv == 0) return radius;
    float r = 0.0f;
    for (int k = 0; k < nv; ++k) {
        const float* v = verts4 + 4 * (long)k;
        r = fmaxf(r, sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    return r;
}

#define HB_GRID_AXIS_MAX 512
#define HB_GRID_CELLS_MAX 4194304L

/* Fits a grid of cell `cell` (inflated by 2^-10) over [lo, hi]: the cell
 * grows by 1.25 until no axis has more than 512 cells and the grid no more
 * than 2^22; writes 1 / cell and the cells per axis, returns the cell count. */
static long hb_grid_fit(float cell, const float* lo, const float* hi, float* inv_out, int* d) {
    cell = 2.0f * cell * 1.0009765625f;
    if (!(cell > 0.0f)) cell = 1e-3f;
    for (;;) {
        const float inv = 1.0f / cell;
        long cells = 1;
        int ok = 1;
        for (int a = 0; a < 3; ++a) {
            const float e = ceilf((hi[a] - lo[a]) * inv);
            d[a] = e >= 1.0f ? (e <= (float)HB_GRID_AXIS_MAX ? (int)e : HB_GRID_AXIS_MAX + 1) : 1;
            ok = ok && d[a] <= HB_GRID_AXIS_MAX;
            cells *= d[a];
        }
        if (ok && cells <= HB_GRID_CELLS_MAX) {
            *inv_out = inv;
            return cells;
        }
        cell = cell * 1.25f;
    }
}

/* Per-scene grids of the rigid-body broadphase, fixed from the initial bodies
 * (sd_hull_upload). rd = the largest bounding radius of the scene's moving
 * bodies (mass > 0; of all its bodies when none moves). A body whose bounding
 * radius exceeds rd (large[b] = 1: a static body such as a wall or mesh
 * triangle) goes to the scene's second grid, of the large bodies, and every
 * other body to the first. Each grid: cell = twice the largest bounding radius
 * of its bodies (inflated by 2^-10, so that a gridded body and a body of
 * either grid whose boxes overlap have centres at most one cell of that grid
 * apart per axis), over the box of its bodies' bounding spheres, the cell
 * grown by 1.25 until no axis has more than 512 cells and the grid no more
 * than 2^22. grid16[16 s ..] = (lo x, y, z, 1 / cell, nx, ny, nz, first cell
 * id) of the first grid, then the same of the second, the cell counts and
 * first ids as int bits. Returns the cell count of all scenes' first grids;
 * *large_cells (if given) receives that of the second grids. */
long or_hull_grid(int n, const float* pos4, const float* radius, const int* vstart, const int* nv,
                  const float* verts4, const int* scene_first, int nscenes, float* grid16, int* large,
                  long* large_cells) {
    (void)n;
    long cbase = 0, lbase = 0;
    for (int s = 0; s < nscenes; ++s) {
        float rd = 0.0f, ra = 0.0f;
        int moving = 0;
        for (int b = scene_first[s]; b < scene_first[s + 1]; ++b) {
            const float r = hb_bound(verts4 + 4 * (long)vstart[b], nv[b], radius[b]);
            ra = fmaxf(ra, r);
            if (pos4[4 * (long)b + 3] > 0.0f) {
                rd = fmaxf(rd, r);
                moving = 1;
            }
        }
        if (!moving) rd = ra;
        float lo[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}, hi[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        float rl = 0.0f;
        int first[2] = {1, 1};
        for (int b = scene_first[s]; b < scene_first[s + 1]; ++b) {
            const float r = hb_bound(verts4 + 4 * (long)vstart[b], nv[b], radius[b]);
            const int k = large[b] = r > rd;
            if (k) rl = fmaxf(rl, r);
            for (int a = 0; a < 3; ++a) {
                const float x = pos4[4 * (long)b + a];
                lo[k][a] = first[k] ? x - r : fminf(lo[k][a], x - r);
                hi[k][a] = first[k] ? x + r : fmaxf(hi[k][a], x + r);
            }
            first[k] = 0;
        }
        for (int k = 0; k < 2; ++k) {
            float inv;
            int d[3];
            const long cells = hb_grid_fit(k ? rl : rd, lo[k], hi[k], &inv, d);
            float* g = grid16 + 16 * (long)s + 8 * k;
            g[0] = lo[k][0];
            g[1] = lo[k][1];
            g[2] = lo[k][2];
            g[3] = inv;
            int* gi = (int*)(g + 4);
            gi[0] = d[0];
            gi[1] = d[1];
            gi[2] = d[2];
            gi[3] = (int)(k ? lbase : cbase);
            if (k) lbase += cells;
            else cbase += cells;
        }
    }
    if (large_cells) *large_cells = lbase;
    return cbase;
}

/* cell coordinate of x on one axis of a scene grid, clamped to [0, nd - 1] */
static int hb_cell_axis(float x, float lo, float inv, int nd) {
    const float t = fminf(fmaxf(floorf((x - lo) * inv), 0.0f), (float)(nd - 1));
    return (int)t;
}

static int hb_boxes_overlap(const float* l1, const float* h1, const float* l2, const float* h2) {
    return l2[0] <= h1[0] && l1[0] <= h2[0] && l2[1] <= h1[1] && l1[1] <= h2[1] && l2[2] <= h1[2] && l1[2] <= h2[2];
}

static int hb_pair_ok(const float* P, const float* lo, const float* hi, int a, int b) {
    if (P[4 * (long)a + 3] == 0.0f && P[4 * (long)b + 3] == 0.0f) return 0; /* two static bodies */
    return hb_boxes_overlap(lo + 4 * (long)a, hi + 4 * (long)a, lo + 4 * (long)b, hi + 4 * (long)b);
}

static int* hb_push(int* pairs, long* cap, long np, int a, int b) {
    if (np >= *cap) {
        *cap *= 2;
        pairs = (int*)realloc(pairs, 2 * (size_t)*cap * sizeof(int));
    }
    pairs[2 * np] = a;
    pairs[2 * np + 1] = b;
    return pairs;
}

/* The bodies of scene [f, f + k) with large[b] == kind bucketed by their
 * centre's cell in grid g (8 values): start[0 .. ncell] and order (index
 * order within a cell); cellof[b] = the body's local cell. */
static void hb_bucket(const float* P, const int* large, int kind, int f, int k, const float* g, int* cellof,
                      int** start_out, int** order_out) {
    const int* gi = (const int*)(g + 4);
    const int nx = gi[0], ny = gi[1], nz = gi[2];
    const long ncell = (long)nx * ny * nz;
    int* start = (int*)calloc((size_t)ncell + 1, sizeof(int));
    int* order = (int*)malloc((size_t)(k > 0 ? k : 1) * sizeof(int));
    for (int b = f; b < f + k; ++b) {
        if (large[b] != kind) continue;
        const int cx = hb_cell_axis(P[4 * (long)b], g[0], g[3], nx);
        const int cy = hb_cell_axis(P[4 * (long)b + 1], g[1], g[3], ny);
        const int cz = hb_cell_axis(P[4 * (long)b + 2], g[2], g[3], nz);
        cellof[b] = (cz * ny + cy) * nx + cx;
        ++start[cellof[b] + 1];
    }
    for (long c = 0; c < ncell; ++c) start[c + 1] += start[c];
    int* fill = (int*)calloc((size_t)ncell + 1, sizeof(int));
    for (int b = f; b < f + k; ++b)
        if (large[b] == kind) order[start[cellof[b]] + fill[cellof[b]]++] = b; /* stable: index order */
    free(fill);
    *start_out = start;
    *order_out = order;
}

/* Candidate pairs of the rigid-body step, body a by body a (index order); for
 * each gridded body a, the bodies of its scene whose boxes overlap a's
 * (closed intervals), except pairs of two static bodies (1/m = 0, P.w), each
 * pair written as (min, max): first the gridded bodies of a's own cell after
 * it (b > a) and of the 13 cells that follow its own in (dz, dy, dx) order
 * (dz = 1; dz = 0, dy = 1; dz = dy = 0, dx = 1), in that order and within a
 * cell in index order, then the large bodies in the 27 cells of the scene's
 * second grid around a's centre (dz, dy, dx ascending; within a cell in index
 * order). Two gridded bodies whose boxes overlap lie in the same cell or in
 * neighbouring ones (clamping to the border cells keeps that), so the pair is
 * found once, from the body whose cell comes first (the lower index in a
 * shared cell); a large body overlapping a gridded one lies within one cell
 * of the second grid. A large body is static (its bounding radius exceeds
 * every moving body's), so two large bodies never pair and a large body adds
 * no pairs of its own. Returns the pair count (pairs grown by realloc). */
static long hb_grid_pairs(int n, const float* P, const float* lo, const float* hi, const int* scene_first,
                          int nscenes, const float* grid16, const int* large, int** pairs_io, long* cap_io) {
    long np = 0;
    int* pairs = *pairs_io;
    long cap = *cap_io;
    int* cellof = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int s = 0; s < nscenes; ++s) {
        const int f = scene_first[s], k = scene_first[s + 1] - f;
        if (k < 2) continue;
        const float* g = grid16 + 16 * (long)s;
        const float* gl = g + 8;
        const int* gi = (const int*)(g + 4);
        const int* gli = (const int*)(gl + 4);
        const int nx = gi[0], ny = gi[1], nz = gi[2];
        const int lx = gli[0], ly = gli[1], lz = gli[2];
        int *start, *order, *lstart, *lorder;
        hb_bucket(P, large, 0, f, k, g, cellof, &start, &order);
        hb_bucket(P, large, 1, f, k, gl, cellof, &lstart, &lorder);
        for (int a = f; a < f + k; ++a) {
            if (large[a]) continue;
            const int ca = cellof[a];
            const int cx = ca % nx, cy = (ca / nx) % ny, cz = ca / (nx * ny);
            for (int dz = 0; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (dz == 0 && (dy < 0 || (dy == 0 && dx < 0))) continue; /* the other half */
                        const int x = cx + dx, y = cy + dy, z = cz + dz;
                        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) continue;
                        const int c = (z * ny + y) * nx + x;
                        const int own = dz == 0 && dy == 0 && dx == 0;
                        for (int t = start[c]; t < start[c + 1]; ++t) {
                            const int b = order[t];
                            if ((!own || b > a) && hb_pair_ok(P, lo, hi, a, b))
                                pairs = hb_push(pairs, &cap, np++, a < b ? a : b, a < b ? b : a);
                        }
                    }
            if (P[4 * (long)a + 3] == 0.0f) continue; /* large bodies are static */
            const int qx = hb_cell_axis(P[4 * (long)a], gl[0], gl[3], lx);
            const int qy = hb_cell_axis(P[4 * (long)a + 1], gl[1], gl[3], ly);
            const int qz = hb_cell_axis(P[4 * (long)a + 2], gl[2], gl[3], lz);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int x = qx + dx, y = qy + dy, z = qz + dz;
                        if (x < 0 || x >= lx || y < 0 || y >= ly || z < 0 || z >= lz) continue;
                        const int c = (z * ly + y) * lx + x;
                        for (int t = lstart[c]; t < lstart[c + 1]; ++t) {
                            const int b = lorder[t];
                            if (hb_pair_ok(P, lo, hi, a, b))
                                pairs = hb_push(pairs, &cap, np++, a < b ? a : b, a < b ? b : a);
                        }
                    }
        }
        free(start);
        free(order);
        free(lstart);
        free(lorder);
    }
    free(cellof);
    *pairs_io = pairs;
    *cap_io = cap;
    return np;
}

long or_hull_grid_pairs(int n, const float* P, const float* lo4, const float* hi4, const int* scene_first,
                        int nscenes, const float* grid16, const int* large, int* pairs, long capacity) {
    long cap = 1024;
    int* all = (int*)malloc(2 * (size_t)cap * sizeof(int));
    const long np = hb_grid_pairs(n, P, lo4, hi4, scene_first, nscenes, grid16, large, &all, &cap);
    for (long t = 0; t < np && t < capacity; ++t) {
        pairs[2 * t] = all[2 * t];
        pairs[2 * t + 1] = all[2 * t + 1];
    }
    free(all);
    return np;
}

long or_hull_step(int n, float* P, float* V, float* Q, float* L, const float* I0, const float* I1, const float* radius,
                  const int* vstart, const int* nv, const float* verts4, int n_verts, const int* scene_first,
                  int nscenes, const float* grid16, const int* large, const or_contact* c, int max_iter, int nthreads,
                  int* pairs_out, float* force_out, float* point_out, long out_capacity) {
    (void)n_verts;
    float* spos = (float*)calloc(4 * (size_t)(n > 0 ? n : 1), sizeof(float));
    float* om = (float*)malloc(3 * (size_t)n * sizeof(float));
    float* lo = (float*)malloc(4 * (size_t)n * sizeof(float));
    float* hi = (float*)malloc(4 * (size_t)n * sizeof(float));
    for (int b = 0; b < n; ++b) {
        spos[4 * b] = P[4 * b];
        spos[4 * b + 1] = P[4 * b + 1];
        spos[4 * b + 2] = P[4 * b + 2];
        spos[4 * b + 3] = radius[b];
        float m[9];
        qmat(Q + 4 * (long)b, m);
        const V3 w = inv_inertia(m, I0 + 4 * (long)b, I1 + 4 * (long)b, mk(L[4 * b], L[4 * b + 1], L[4 * b + 2]));
        om[3 * b] = w.x;
        om[3 * b + 1] = w.y;
        om[3 * b + 2] = w.z;
    }
    or_hull_aabbs(n, spos, Q, vstart, nv, verts4, lo, hi);
    /* candidate pairs: the per-scene grids */
    long cap = 1024;
    int* pairs = (int*)malloc(2 * (size_t)cap * sizeof(int));
    const long np = hb_grid_pairs(n, P, lo, hi, scene_first, nscenes, grid16, large, &pairs, &cap);
    float* sep = (float*)malloc((size_t)(np + 1) * sizeof(float));
    float* nrm = (float*)malloc(3 * (size_t)(np + 1) * sizeof(float));
    int* flags = (int*)malloc((size_t)(np + 1) * sizeof(int));
    int* iters = (int*)malloc((size_t)(np + 1) * sizeof(int));
    if (np) or_gjk((int)np, pairs, spos, Q, vstart, nv, verts4, NULL, 1, max_iter, 0.0f, sep, nrm, flags, iters,
                   nthreads);
    float* pf = (float*)calloc(4 * (size_t)(np + 1), sizeof(float));
    float* pp = (float*)calloc(4 * (size_t)(np + 1), sizeof(float));
    int* cnt = (int*)calloc((size_t)n, sizeof(int));
    const float c2z = 2.0f * c->damping_ratio;
    long ncontacts = 0;
    for (long t = 0; t < np; ++t) {
        const int a = pairs[2 * t], b = pairs[2 * t + 1];
        const float* Pa = P + 4 * (long)a;
        const float* Pb = P + 4 * (long)b;
        const float minv = Pa[3] + Pb[3];
        const int fl = flags[t];
        const float s = sep[t];
        if (fl < 0 || !(fl & 1) || !(s < 0.0f) || !(minv > 0.0f)) continue;
        const V3 nn = mk(nrm[3 * t], nrm[3 * t + 1], nrm[3 * t + 2]);
        const V3 Xa = mk(Pa[0], Pa[1], Pa[2]), Xb = mk(Pb[0], Pb[1], Pb[2]);
        float ma[9], mb[9];
        qmat(Q + 4 * (long)a, ma);
        qmat(Q + 4 * (long)b, mb);
        float sa, sb;
        const V3 ca = hb_feature(verts4 + 4 * (long)vstart[a], nv[a], radius[a], ma, Xa, nn, &sa);
        const V3 cb = hb_feature(verts4 + 4 * (long)vstart[b], nv[b], radius[b], mb, Xb, hb_scl(nn, -1.0f), &sb);
        const V3 pt = hb_contact_point(ca, sa, cb, sb);
        const V3 wa = mk(om[3 * a], om[3 * a + 1], om[3 * a + 2]), wb = mk(om[3 * b], om[3 * b + 1], om[3 * b + 2]);
        const V3 va = hb_add(mk(V[4 * a], V[4 * a + 1], V[4 * a + 2]), hb_cross(wa, hb_sub(pt, Xa)));
        const V3 vb = hb_add(mk(V[4 * b], V[4 * b + 1], V[4 * b + 2]), hb_cross(wb, hb_sub(pt, Xb)));
        const V3 vr = hb_sub(vb, va);
        const float vn = hb_dot(vr, nn);
        const float meff = 1.0f / minv;
        const float cn = c2z * sqrtf(c->kn * meff);
        const float fn = c->kn * (-s) - cn * vn;
        const V3 vt = hb_sub(vr, hb_scl(nn, vn));
        const float vtn = sqrtf(hb_dot(vt, vt));
        V3 f = hb_scl(nn, fn);
        if (vtn > 0.0f) {
            const float fmag = fminf(c->mu * fmaxf(fn, 0.0f), cn * vtn);
            f = hb_sub(f, hb_scl(vt, fmag / vtn));
        }
        pf[4 * t] = f.x;
        pf[4 * t + 1] = f.y;
        pf[4 * t + 2] = f.z;
        pf[4 * t + 3] = 1.0f;
        pp[4 * t] = pt.x;
        pp[4 * t + 1] = pt.y;
        pp[4 * t + 2] = pt.z;
        cnt[a]++;
        cnt[b]++;
        if (ncontacts < out_capacity) {
            if (pairs_out) {
                pairs_out[2 * ncontacts] = a;
                pairs_out[2 * ncontacts + 1] = b;
            }
            if (force_out) memcpy(force_out + 4 * ncontacts, pf + 4 * t, 4 * sizeof(float));
            if (point_out) memcpy(point_out + 4 * ncontacts, pp + 4 * t, 4 * sizeof(float));
        }
        ++ncontacts;
    }
    /* per body: its contacts in entry order (2 p for a, 2 p + 1 for b) */
    long* offs = (long*)malloc((size_t)(n + 1) * sizeof(long));
    offs[0] = 0;
    for (int b = 0; b < n; ++b) offs[b + 1] = offs[b] + cnt[b];
    HbEntry* list = (HbEntry*)malloc((size_t)(offs[n] + 1) * sizeof(HbEntry));
    long* fill = (long*)calloc((size_t)n, sizeof(long));
    for (long t = 0; t < np; ++t) {
        if (!(pf[4 * t + 3] > 0.0f)) continue;
        const int a = pairs[2 * t], b = pairs[2 * t + 1];
        list[offs[a] + fill[a]++] = (HbEntry){(int)t, (int)(2 * t)};
        list[offs[b] + fill[b]++] = (HbEntry){(int)t, (int)(2 * t + 1)};
    }
    for (int b = 0; b < n; ++b) {
        qsort(list + offs[b], (size_t)cnt[b], sizeof(HbEntry), cmp_entry);
        float* X = P + 4 * (long)b;
        float F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        for (long i = offs[b]; i < offs[b + 1]; ++i) {
            const int e = list[i].e, p = e >> 1;
            const float sg = (e & 1) ? 1.0f : -1.0f;
            const float fx = sg * pf[4 * (long)p], fy = sg * pf[4 * (long)p + 1], fz = sg * pf[4 * (long)p + 2];
            F[0] = F[0] + fx;
            F[1] = F[1] + fy;
            F[2] = F[2] + fz;
            const float rx = pp[4 * (long)p] - X[0], ry = pp[4 * (long)p + 1] - X[1], rz = pp[4 * (long)p + 2] - X[2];
            T[0] = T[0] + (ry * fz - rz * fy);
            T[1] = T[1] + (rz * fx - rx * fz);
            T[2] = T[2] + (rx * fy - ry * fx);
        }
        integrate_body(X, V + 4 * (long)b, Q + 4 * (long)b, L + 4 * (long)b, I0 + 4 * (long)b, I1 + 4 * (long)b, F, T,
                       c);
    }
    free(spos);
    free(om);
    free(lo);
    free(hi);
    free(pairs);
    free(sep);
    free(nrm);
    free(flags);
    free(iters);
    free(pf);
    free(pp);
    free(cnt);
    free(offs);
    free(list);
    free(fill);
    return ncontacts;
}
