/*
 * sd_oracle.c — CPU oracle for the DEM step path (TEST INFRASTRUCTURE ONLY;
 * see sd_oracle.h). Built with -ffp-contract=off -fno-fast-math on x86-64 SSE,
 * so FP32 expressions round exactly as written.
 */
#include "sd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ grid */

/* Per-axis cell: the reference truncates x / cell_dim (sh_collision_detection.cu:64);
 * here the domain origin is subtracted first and the index clamped to the grid. */
static int axis_cell(float p, float lo, float cell, int nc) {
    float f = floorf((p - lo) / cell);
    f = fmaxf(f, 0.0f);
    f = fminf(f, (float)(nc - 1));
    return (int)f;
}

static void cell_xyz(const or_grid* g, float x, float y, float z, int* cx, int* cy, int* cz) {
    *cx = axis_cell(x, g->lo[0], g->cell, g->n[0]);
    *cy = axis_cell(y, g->lo[1], g->cell, g->n[1]);
    *cz = axis_cell(z, g->lo[2], g->cell, g->n[2]);
}

uint32_t or_cell_id(const or_grid* g, float x, float y, float z) {
    int cx, cy, cz;
    cell_xyz(g, x, y, z, &cx, &cy, &cz);
    return (uint32_t)((cz * g->n[1] + cy) * g->n[0] + cx);
}

void or_hash_keys(int n, const float* x, const float* y, const float* z, int stride, const or_grid* g,
                  uint32_t* keys, uint32_t* vals) {
    for (int i = 0; i < n; ++i) {
        keys[i] = or_cell_id(g, x[(long)i * stride], y[(long)i * stride], z[(long)i * stride]);
        if (vals) vals[i] = (uint32_t)i;
    }
}

/* ------------------------------------------------------------------ sort */

/* One stable counting pass per 8-bit digit, least significant first:
 * tabulate (sh_collision_detection.cu:129-166), exclusive prefix (:168-214),
 * stable reorder (:216-283). */
void or_sort_pairs_u32(int n, uint32_t* keys, uint32_t* vals, int begin_bit, int end_bit) {
    if (n <= 1 || end_bit <= begin_bit) return;
    uint32_t* tk = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* tv = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    for (int bit = begin_bit; bit < end_bit; bit += 8) {
        const int w = end_bit - bit < 8 ? end_bit - bit : 8;
        const uint32_t mask = (1u << w) - 1u;
        size_t count[257];
        memset(count, 0, sizeof count);
        for (int i = 0; i < n; ++i) count[((keys[i] >> bit) & mask) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (int i = 0; i < n; ++i) {
            const uint32_t d = (keys[i] >> bit) & mask;
            const size_t o = count[d]++;
            tk[o] = keys[i];
            tv[o] = vals[i];
        }
        memcpy(keys, tk, (size_t)n * sizeof(uint32_t));
        memcpy(vals, tv, (size_t)n * sizeof(uint32_t));
    }
    free(tk);
    free(tv);
}

void or_sort_keys_u64(int n, uint64_t* keys, int begin_bit, int end_bit) {
    if (n <= 1 || end_bit <= begin_bit) return;
    uint64_t* tk = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
    for (int bit = begin_bit; bit < end_bit; bit += 8) {
        const int w = end_bit - bit < 8 ? end_bit - bit : 8;
        const uint64_t mask = (1ull << w) - 1ull;
        size_t count[257];
        memset(count, 0, sizeof count);
        for (int i = 0; i < n; ++i) count[((keys[i] >> bit) & mask) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (int i = 0; i < n; ++i) tk[count[(keys[i] >> bit) & mask]++] = keys[i];
        memcpy(keys, tk, (size_t)n * sizeof(uint64_t));
    }
    free(tk);
}

/* ------------------------------------------------------------------ ranges */

void or_cell_ranges_u32(int n, const uint32_t* keys, int n_cells, int* ranges) {
    memset(ranges, 0, 2 * (size_t)n_cells * sizeof(int));
    int i = 0;
    while (i < n) {
        int j = i;
        while (j < n && keys[j] == keys[i]) ++j;
        ranges[2 * (size_t)keys[i]] = i;
        ranges[2 * (size_t)keys[i] + 1] = j;
        i = j;
    }
}

void or_cell_ranges_u64(int n, const uint64_t* keys, int n_cells, int* ranges) {
    memset(ranges, 0, 2 * (size_t)n_cells * sizeof(int));
    int i = 0;
    while (i < n) {
        const uint32_t c = (uint32_t)(keys[i] >> 32);
        int j = i;
        while (j < n && (uint32_t)(keys[j] >> 32) == c) ++j;
        ranges[2 * (size_t)c] = i;
        ranges[2 * (size_t)c + 1] = j;
        i = j;
    }
}

/* ------------------------------------------------------------------ pairs */

/* Runs the body for every sorted position t in the 27-cell stencil of
 * cell (cx, cy, cz): z rows outer, y rows inner, cells x-1..x+1, positions
 * ascending inside each cell. */
#define FOR_STENCIL(g, ranges, cx, cy, cz, t, ...)                                  \
    for (int dz_ = -1; dz_ <= 1; ++dz_) {                                           \
        const int z_ = (cz) + dz_;                                                  \
        if (z_ < 0 || z_ >= (g)->n[2]) continue;                                    \
        for (int dy_ = -1; dy_ <= 1; ++dy_) {                                       \
            const int y_ = (cy) + dy_;                                              \
            if (y_ < 0 || y_ >= (g)->n[1]) continue;                                \
            for (int dx_ = -1; dx_ <= 1; ++dx_) {                                   \
                const int x_ = (cx) + dx_;                                          \
                if (x_ < 0 || x_ >= (g)->n[0]) continue;                            \
                const long c_ = ((long)z_ * (g)->n[1] + y_) * (g)->n[0] + x_;       \
                const int s0_ = (ranges)[2 * c_], s1_ = (ranges)[2 * c_ + 1];       \
                for (int t = s0_; t < s1_; ++t) { __VA_ARGS__ }                     \
            }                                                                       \
        }                                                                           \
    }

long or_neighbor_pairs(int n, const float* pos4, const or_grid* g, const int* ranges, float skin, int* pairs,
                       long capacity) {
    long np = 0;
    for (int a = 0; a < n; ++a) {
        const float* p = pos4 + 4 * (long)a;
        int cx, cy, cz;
        cell_xyz(g, p[0], p[1], p[2], &cx, &cy, &cz);
        FOR_STENCIL(g, ranges, cx, cy, cz, b, {
            if (b <= a) continue;
            const float* q = pos4 + 4 * (long)b;
            const float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float rs = p[3] + q[3] + skin;
            if (d2 < rs * rs) {
                if (np < capacity) {
                    pairs[2 * np] = a;
                    pairs[2 * np + 1] = b;
                }
                ++np;
            }
        })
    }
    return np;
}

long or_brute_pairs(int n, const float* pos4, float skin, int* pairs, long capacity) {
    long np = 0;
    for (int i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (long)i;
        for (int j = i + 1; j < n; ++j) {
            const float* q = pos4 + 4 * (long)j;
            const float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float rs = p[3] + q[3] + skin;
            if (d2 < rs * rs) {
                if (np < capacity) {
                    pairs[2 * np] = i;
                    pairs[2 * np + 1] = j;
                }
                ++np;
            }
        }
    }
    return np;
}

/* ------------------------------------------------------------------ contact law */

typedef struct {
    float x, y, z;
} v3f;

typedef struct {
    float kn, kt, c2z, mu, dt;
} law_f;

static law_f law_of(const or_contact* c) {
    law_f l;
    l.kn = c->kn;
    l.kt = c->kt_ratio * c->kn;
    l.c2z = 2.0f * c->damping_ratio;
    l.mu = c->mu;
    l.dt = c->dt;
    return l;
}

/* Linear spring-dashpot + tangential spring with Coulomb limit for one contact
 * seen from sphere i. The reference's law (sap_reaction.cu:38-69: -kn*(sep-d)*n,
 * damping*v_rel, kt*v_t, component-wise sign flips) is replaced by this one;
 * the operation order below is the specification the GPU kernel follows.
 * n points from j to i; h is the tangential spring history (updated). */
static void law_pair_f(const law_f* L, v3f n, float delta, float ri, float rj, float inv_mi, float inv_mj, v3f vr,
                       v3f wi, v3f wj, v3f* h, v3f* F, v3f* T) {
    const float m_eff = 1.0f / (inv_mi + inv_mj);
    const float cn = L->c2z * sqrtf(L->kn * m_eff);
    const float ct = L->c2z * sqrtf(L->kt * m_eff);
    /* lever arms to the mid-overlap point, shared by both spheres */
    const float li = ri - 0.5f * delta;
    const float lj = rj - 0.5f * delta;
    /* relative velocity of the contact points: v_r - (li wi + lj wj) x n */
    v3f a = {li * wi.x + lj * wj.x, li * wi.y + lj * wj.y, li * wi.z + lj * wj.z};
    v3f axn = {a.y * n.z - a.z * n.y, a.z * n.x - a.x * n.z, a.x * n.y - a.y * n.x};
    v3f vc = {vr.x - axn.x, vr.y - axn.y, vr.z - axn.z};
    const float vn = vc.x * n.x + vc.y * n.y + vc.z * n.z;
    const float fn = L->kn * delta - cn * vn;
    v3f vt = {vc.x - vn * n.x, vc.y - vn * n.y, vc.z - vn * n.z};
    /* history: drop its normal part, then integrate the tangential velocity */
    const float hn = h->x * n.x + h->y * n.y + h->z * n.z;
    h->x = h->x - hn * n.x;
    h->y = h->y - hn * n.y;
    h->z = h->z - hn * n.z;
    h->x = h->x + vt.x * L->dt;
    h->y = h->y + vt.y * L->dt;
    h->z = h->z + vt.z * L->dt;
    v3f ft = {-(L->kt * h->x) - ct * vt.x, -(L->kt * h->y) - ct * vt.y, -(L->kt * h->z) - ct * vt.z};
    const float ft2 = ft.x * ft.x + ft.y * ft.y + ft.z * ft.z;
    const float fmax = L->mu * fabsf(fn);
    if (ft2 > fmax * fmax) {
        const float sc = fmax / sqrtf(ft2);
        ft.x = ft.x * sc;
        ft.y = ft.y * sc;
        ft.z = ft.z * sc;
        h->x = -(ft.x + ct * vt.x) / L->kt;
        h->y = -(ft.y + ct * vt.y) / L->kt;
        h->z = -(ft.z + ct * vt.z) / L->kt;
    }
    F->x = fn * n.x + ft.x;
    F->y = fn * n.y + ft.y;
    F->z = fn * n.z + ft.z;
    T->x = -li * (n.y * ft.z - n.z * ft.y);
    T->y = -li * (n.z * ft.x - n.x * ft.z);
    T->z = -li * (n.x * ft.y - n.y * ft.x);
}

static void history_lookup_f(const float* hist4_old, int n, int os, int oc, int pid, v3f* h) {
    h->x = 0.0f;
    h->y = 0.0f;
    h->z = 0.0f;
    for (int k = 0; k < oc; ++k) {
        const float* e = hist4_old + 4 * ((long)k * n + os);
        int id;
        memcpy(&id, &e[3], sizeof id);
        if (id == pid) {
            h->x = e[0];
            h->y = e[1];
            h->z = e[2];
            return;
        }
    }
}

static void history_store_f(float* hist4_new, int n, int s, int k, int pid, v3f h) {
    float* e = hist4_new + 4 * ((long)k * n + s);
    e[0] = h.x;
    e[1] = h.y;
    e[2] = h.z;
    memcpy(&e[3], &pid, sizeof pid);
}

static int contacts_of(int s, int n, const float* pos4, const float* vel4, const float* omg4, const int* ids,
                       const or_grid* g, const int* ranges, const or_contact* c, const law_f* L, const int* old_slot,
                       const int* hcount_old, const float* hist4_old, int* hcount_new, float* hist4_new,
                       float* force4, float* torque4) {
    const float* p = pos4 + 4 * (long)s;
    const float* v = vel4 + 4 * (long)s;
    const float* w = omg4 + 4 * (long)s;
    const int os = old_slot[s];
    int oc = hcount_old[os];
    if (oc > c->max_contacts) oc = c->max_contacts;
    int cx, cy, cz;
    cell_xyz(g, p[0], p[1], p[2], &cx, &cy, &cz);
    v3f Fs = {0.0f, 0.0f, 0.0f}, Ts = {0.0f, 0.0f, 0.0f};
    int nc = 0;
    const v3f wi = {w[0], w[1], w[2]};
    FOR_STENCIL(g, ranges, cx, cy, cz, t, {
        if (t == s) continue;
        const float* q = pos4 + 4 * (long)t;
        const float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float rs = p[3] + q[3];
        if (!(d2 < rs * rs) || d2 == 0.0f) continue;
        const float* vq = vel4 + 4 * (long)t;
        if (v[3] + vq[3] == 0.0f) continue;
        const float* wq = omg4 + 4 * (long)t;
        const int pid = ids[t];
        const float d = sqrtf(d2);
        const float inv_d = 1.0f / d;
        const v3f nn = {dx * inv_d, dy * inv_d, dz * inv_d};
        const v3f vr = {v[0] - vq[0], v[1] - vq[1], v[2] - vq[2]};
        const v3f wj = {wq[0], wq[1], wq[2]};
        v3f h, F, T;
        history_lookup_f(hist4_old, n, os, oc, pid, &h);
        law_pair_f(L, nn, rs - d, p[3], q[3], v[3], vq[3], vr, wi, wj, &h, &F, &T);
        Fs.x = Fs.x + F.x;
        Fs.y = Fs.y + F.y;
        Fs.z = Fs.z + F.z;
        Ts.x = Ts.x + T.x;
        Ts.y = Ts.y + T.y;
        Ts.z = Ts.z + T.z;
        if (nc < c->max_contacts) history_store_f(hist4_new, n, s, nc, pid, h);
        ++nc;
    })
    if (c->walls && v[3] > 0.0f) {
        /* six planes: x_lo, x_hi, y_lo, y_hi, z_lo, z_hi; n points into the domain */
        for (int wl = 0; wl < 6; ++wl) {
            const int ax = wl / 2;
            const int upper = wl % 2;
            const float dist = upper ? c->wall_hi[ax] - p[ax] : p[ax] - c->wall_lo[ax];
            if (!(dist < p[3])) continue;
            const float sgn = upper ? -1.0f : 1.0f;
            const v3f nn = {ax == 0 ? sgn : 0.0f, ax == 1 ? sgn : 0.0f, ax == 2 ? sgn : 0.0f};
            const v3f vr = {v[0], v[1], v[2]};
            const v3f wj = {0.0f, 0.0f, 0.0f};
            const int pid = -(wl + 1);
            v3f h, F, T;
            history_lookup_f(hist4_old, n, os, oc, pid, &h);
            law_pair_f(L, nn, p[3] - dist, p[3], 0.0f, v[3], 0.0f, vr, wi, wj, &h, &F, &T);
            Fs.x = Fs.x + F.x;
            Fs.y = Fs.y + F.y;
            Fs.z = Fs.z + F.z;
            Ts.x = Ts.x + T.x;
            Ts.y = Ts.y + T.y;
            Ts.z = Ts.z + T.z;
            if (nc < c->max_contacts) history_store_f(hist4_new, n, s, nc, pid, h);
            ++nc;
        }
    }
    hcount_new[s] = nc < c->max_contacts ? nc : c->max_contacts;
    float* fo = force4 + 4 * (long)s;
    float* to = torque4 + 4 * (long)s;
    fo[0] = Fs.x;
    fo[1] = Fs.y;
    fo[2] = Fs.z;
    fo[3] = 0.0f;
    to[0] = Ts.x;
    to[1] = Ts.y;
    to[2] = Ts.z;
    to[3] = 0.0f;
    return nc;
}

int or_contact_forces(int n, const float* pos4, const float* vel4, const float* omg4, const int* ids,
                      const or_grid* g, const int* ranges, const or_contact* c, const int* old_slot,
                      const int* hcount_old, const float* hist4_old, int* hcount_new, float* hist4_new, float* force4,
                      float* torque4, int nthreads) {
    const law_f L = law_of(c);
    int maxc = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 512) reduction(max : maxc)
    for (int s = 0; s < n; ++s) {
        const int nc = contacts_of(s, n, pos4, vel4, omg4, ids, g, ranges, c, &L, old_slot, hcount_old, hist4_old,
                                   hcount_new, hist4_new, force4, torque4);
        if (nc > maxc) maxc = nc;
    }
    return maxc;
}

/* ------------------------------------------------------------------ FP64 law */

typedef struct {
    double x, y, z;
} v3d;

static void law_pair_d(const or_contact* c, v3d n, double delta, double ri, double rj, double inv_mi, double inv_mj,
                       v3d vr, v3d wi, v3d wj, v3d* h, v3d* F, v3d* T) {
    const double kn = c->kn, kt = (double)c->kt_ratio * c->kn, c2z = 2.0 * c->damping_ratio;
    const double m_eff = 1.0 / (inv_mi + inv_mj);
    const double cn = c2z * sqrt(kn * m_eff), ct = c2z * sqrt(kt * m_eff);
    const double li = ri - 0.5 * delta, lj = rj - 0.5 * delta;
    v3d a = {li * wi.x + lj * wj.x, li * wi.y + lj * wj.y, li * wi.z + lj * wj.z};
    v3d vc = {vr.x - (a.y * n.z - a.z * n.y), vr.y - (a.z * n.x - a.x * n.z), vr.z - (a.x * n.y - a.y * n.x)};
    const double vn = vc.x * n.x + vc.y * n.y + vc.z * n.z;
    const double fn = kn * delta - cn * vn;
    v3d vt = {vc.x - vn * n.x, vc.y - vn * n.y, vc.z - vn * n.z};
    const double hn = h->x * n.x + h->y * n.y + h->z * n.z;
    h->x = h->x - hn * n.x + vt.x * c->dt;
    h->y = h->y - hn * n.y + vt.y * c->dt;
    h->z = h->z - hn * n.z + vt.z * c->dt;
    v3d ft = {-kt * h->x - ct * vt.x, -kt * h->y - ct * vt.y, -kt * h->z - ct * vt.z};
    const double ft2 = ft.x * ft.x + ft.y * ft.y + ft.z * ft.z;
    const double fmax = c->mu * fabs(fn);
    if (ft2 > fmax * fmax) {
        const double sc = fmax / sqrt(ft2);
        ft.x *= sc;
        ft.y *= sc;
        ft.z *= sc;
        h->x = -(ft.x + ct * vt.x) / kt;
        h->y = -(ft.y + ct * vt.y) / kt;
        h->z = -(ft.z + ct * vt.z) / kt;
    }
    F->x = fn * n.x + ft.x;
    F->y = fn * n.y + ft.y;
    F->z = fn * n.z + ft.z;
    T->x = -li * (n.y * ft.z - n.z * ft.y);
    T->y = -li * (n.z * ft.x - n.x * ft.z);
    T->z = -li * (n.x * ft.y - n.y * ft.x);
}

int or_contact_forces_f64(int n, const double* pos4, const double* vel4, const double* omg4, const int* ids,
                          const or_grid* g, const int* ranges, const or_contact* c, const int* old_slot,
                          const int* hcount_old, const double* hist4_old, int* hcount_new, double* hist4_new,
                          double* force4, double* torque4) {
    int maxc = 0;
    for (int s = 0; s < n; ++s) {
        const double* p = pos4 + 4 * (long)s;
        const double* v = vel4 + 4 * (long)s;
        const double* w = omg4 + 4 * (long)s;
        const int os = old_slot[s];
        int oc = hcount_old[os] < c->max_contacts ? hcount_old[os] : c->max_contacts;
        /* cells from the FP32 rounding of the position so both oracles walk the same cells */
        int cx, cy, cz;
        cell_xyz(g, (float)p[0], (float)p[1], (float)p[2], &cx, &cy, &cz);
        v3d Fs = {0, 0, 0}, Ts = {0, 0, 0};
        const v3d wi = {w[0], w[1], w[2]};
        int nc = 0;
        FOR_STENCIL(g, ranges, cx, cy, cz, t, {
            if (t == s) continue;
            const double* q = pos4 + 4 * (long)t;
            const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const double d2 = dx * dx + dy * dy + dz * dz, rs = p[3] + q[3];
            if (!(d2 < rs * rs) || d2 == 0.0) continue;
            const double* vq = vel4 + 4 * (long)t;
            if (v[3] + vq[3] == 0.0) continue;
            const double* wq = omg4 + 4 * (long)t;
            const int pid = ids[t];
            const double d = sqrt(d2);
            const v3d nn = {dx / d, dy / d, dz / d};
            const v3d vr = {v[0] - vq[0], v[1] - vq[1], v[2] - vq[2]};
            const v3d wj = {wq[0], wq[1], wq[2]};
            v3d h = {0, 0, 0}, F, T;
            for (int k = 0; k < oc; ++k) {
                const double* e = hist4_old + 4 * ((long)k * n + os);
                if ((int)e[3] == pid) {
                    h.x = e[0];
                    h.y = e[1];
                    h.z = e[2];
                    break;
                }
            }
            law_pair_d(c, nn, rs - d, p[3], q[3], v[3], vq[3], vr, wi, wj, &h, &F, &T);
            Fs.x += F.x;
            Fs.y += F.y;
            Fs.z += F.z;
            Ts.x += T.x;
            Ts.y += T.y;
            Ts.z += T.z;
            if (nc < c->max_contacts) {
                double* e = hist4_new + 4 * ((long)nc * n + s);
                e[0] = h.x;
                e[1] = h.y;
                e[2] = h.z;
                e[3] = (double)pid;
            }
            ++nc;
        })
        if (c->walls && v[3] > 0.0) {
            for (int wl = 0; wl < 6; ++wl) {
                const int ax = wl / 2, upper = wl % 2;
                const double dist = upper ? (double)c->wall_hi[ax] - p[ax] : p[ax] - (double)c->wall_lo[ax];
                if (!(dist < p[3])) continue;
                const double sgn = upper ? -1.0 : 1.0;
                const v3d nn = {ax == 0 ? sgn : 0.0, ax == 1 ? sgn : 0.0, ax == 2 ? sgn : 0.0};
                const v3d vr = {v[0], v[1], v[2]}, wj = {0, 0, 0};
                const int pid = -(wl + 1);
                v3d h = {0, 0, 0}, F, T;
                for (int k = 0; k < oc; ++k) {
                    const double* e = hist4_old + 4 * ((long)k * n + os);
                    if ((int)e[3] == pid) {
                        h.x = e[0];
                        h.y = e[1];
                        h.z = e[2];
                        break;
                    }
                }
                law_pair_d(c, nn, p[3] - dist, p[3], 0.0, v[3], 0.0, vr, wi, wj, &h, &F, &T);
                Fs.x += F.x;
                Fs.y += F.y;
                Fs.z += F.z;
                Ts.x += T.x;
                Ts.y += T.y;
                Ts.z += T.z;
                if (nc < c->max_contacts) {
                    double* e = hist4_new + 4 * ((long)nc * n + s);
                    e[0] = h.x;
                    e[1] = h.y;
                    e[2] = h.z;
                    e[3] = (double)pid;
                }
                ++nc;
            }
        }
        hcount_new[s] = nc < c->max_contacts ? nc : c->max_contacts;
        double* fo = force4 + 4 * (long)s;
        double* to = torque4 + 4 * (long)s;
        fo[0] = Fs.x;
        fo[1] = Fs.y;
        fo[2] = Fs.z;
        fo[3] = 0;
        to[0] = Ts.x;
        to[1] = Ts.y;
        to[2] = Ts.z;
        to[3] = 0;
        if (nc > maxc) maxc = nc;
    }
    return maxc;
}

/* ------------------------------------------------------------------ integrate */

void or_integrate(int n, float* pos4, float* vel4, float* omg4, const float* force4, const float* torque4,
                  const or_contact* c) {
    const float dt = c->dt;
    for (int i = 0; i < n; ++i) {
        float* p = pos4 + 4 * (long)i;
        float* v = vel4 + 4 * (long)i;
        float* w = omg4 + 4 * (long)i;
        const float* f = force4 + 4 * (long)i;
        const float* t = torque4 + 4 * (long)i;
        if (!(v[3] > 0.0f)) continue; /* 1/m == 0: static */
        for (int a = 0; a < 3; ++a) v[a] = v[a] + (f[a] * v[3] + c->gravity[a]) * dt;
        for (int a = 0; a < 3; ++a) p[a] = p[a] + v[a] * dt;
        for (int a = 0; a < 3; ++a) w[a] = w[a] + (t[a] * w[3]) * dt;
    }
}

/* ------------------------------------------------------------------ step */

int or_sphere_step(int n, const or_grid* g, const or_contact* c, float* pos4, float* vel4, float* omg4, int* ids,
                   int* hcount, float* hist4, int nthreads) {
    const int n_cells = g->n[0] * g->n[1] * g->n[2];
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (unsigned long long)n_cells) ++bits;
    uint32_t* keys = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* perm = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    int* ranges = (int*)malloc(2 * (size_t)n_cells * sizeof(int));
    float* p2 = (float*)malloc(4 * (size_t)n * sizeof(float));
    float* v2 = (float*)malloc(4 * (size_t)n * sizeof(float));
    float* w2 = (float*)malloc(4 * (size_t)n * sizeof(float));
    int* id2 = (int*)malloc((size_t)n * sizeof(int));
    int* slot = (int*)malloc((size_t)n * sizeof(int));
    int* hc2 = (int*)malloc((size_t)n * sizeof(int));
    /* zeroed: history rows past a sphere's count stay 0 rather than stale */
    float* h2 = (float*)calloc(4 * (size_t)c->max_contacts * n, sizeof(float));
    float* f = (float*)malloc(4 * (size_t)n * sizeof(float));
    float* t = (float*)malloc(4 * (size_t)n * sizeof(float));

    or_hash_keys(n, pos4, pos4 + 1, pos4 + 2, 4, g, keys, perm);
    or_sort_pairs_u32(n, keys, perm, 0, bits);
    or_cell_ranges_u32(n, keys, n_cells, ranges);
    for (int s = 0; s < n; ++s) {
        const long o = perm[s];
        memcpy(p2 + 4 * (long)s, pos4 + 4 * o, 4 * sizeof(float));
        memcpy(v2 + 4 * (long)s, vel4 + 4 * o, 4 * sizeof(float));
        memcpy(w2 + 4 * (long)s, omg4 + 4 * o, 4 * sizeof(float));
        id2[s] = ids[o];
        slot[s] = (int)o;
    }
    const int maxc = or_contact_forces(n, p2, v2, w2, id2, g, ranges, c, slot, hcount, hist4, hc2, h2, f, t,
                                       nthreads);
    or_integrate(n, p2, v2, w2, f, t, c);
    memcpy(pos4, p2, 4 * (size_t)n * sizeof(float));
    memcpy(vel4, v2, 4 * (size_t)n * sizeof(float));
    memcpy(omg4, w2, 4 * (size_t)n * sizeof(float));
    memcpy(ids, id2, (size_t)n * sizeof(int));
    memcpy(hcount, hc2, (size_t)n * sizeof(int));
    memcpy(hist4, h2, 4 * (size_t)c->max_contacts * n * sizeof(float));
    free(keys);
    free(perm);
    free(ranges);
    free(p2);
    free(v2);
    free(w2);
    free(id2);
    free(slot);
    free(hc2);
    free(h2);
    free(f);
    free(t);
    return maxc;
}
