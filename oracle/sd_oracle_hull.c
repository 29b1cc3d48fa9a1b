/*
 * sd_oracle_hull.c — CPU oracle for the convex-hull path (TEST INFRASTRUCTURE
 * ONLY; see sd_oracle.h): hull AABBs, serial sweep-and-prune, GJK distance and
 * EPA penetration depth.
 *
 * Restates, as plain C, the algorithms the reference sketches in
 * src/engine/cuda/collisions/SAP/sap_detection.cu:135-706 (AABB, projection,
 * sort of lower extents, sweep) and sandbox/testGJK.cu:52-315 (support mapping
 * on point sets, Minkowski difference, simplex evolution), with the intended
 * outputs: complete overlap sets, terminating GJK with distances, EPA depths.
 * The operation order is the documented one of the GPU kernels so FP32
 * results can be compared bit for bit; the algorithms themselves are pinned in
 * tests/test_oracle_hull.py against scipy's Qhull and NNLS and closed forms.
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
static V3 vadd(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 vsub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 vneg(V3 a) { return mk(-a.x, -a.y, -a.z); }
static V3 vscl(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 vcross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

/* rotation matrix (row-major) of the unit quaternion q = (w, x, y, z) */
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
static V3 mvec(const float m[9], V3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
static V3 mtvec(const float m[9], V3 v) {
    return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

/* ------------------------------------------------------------------ AABBs */

void or_hull_aabbs(int n, const float* pos4, const float* quat4, const int* vstart, const int* nv,
                   const float* verts4, float* lo4, float* hi4) {
    for (int i = 0; i < n; ++i) {
        const float* c = pos4 + 4 * (long)i;
        float lo[3], hi[3];
        if (nv[i] == 0) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = c[a] - c[3];
                hi[a] = c[a] + c[3];
            }
        } else {
            float m[9];
            qmat(quat4 + 4 * (long)i, m);
            for (int a = 0; a < 3; ++a) {
                lo[a] = 3.402823466e38f;
                hi[a] = -3.402823466e38f;
            }
            for (int k = 0; k < nv[i]; ++k) {
                const float* p = verts4 + 4 * (long)(vstart[i] + k);
                const V3 w = mvec(m, mk(p[0], p[1], p[2]));
                const float ww[3] = {w.x, w.y, w.z};
                for (int a = 0; a < 3; ++a) {
                    lo[a] = fminf(lo[a], ww[a]);
                    hi[a] = fmaxf(hi[a], ww[a]);
                }
            }
            for (int a = 0; a < 3; ++a) {
                lo[a] = lo[a] + c[a];
                hi[a] = hi[a] + c[a];
            }
        }
        for (int a = 0; a < 3; ++a) {
            lo4[4 * (long)i + a] = lo[a];
            hi4[4 * (long)i + a] = hi[a];
        }
        lo4[4 * (long)i + 3] = 0.0f;
        hi4[4 * (long)i + 3] = 0.0f;
    }
}

/* ------------------------------------------------------------------ SAP */

static uint32_t fkey(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

int or_sap_axis(int n, const float* lo4, const float* hi4) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const double c = 0.5 * ((double)lo4[4 * (long)i + a] + (double)hi4[4 * (long)i + a]);
            s[a] += c;
            s[3 + a] += c * c;
        }
    int best = 0;
    double bv = -1.0;
    for (int a = 0; a < 3; ++a) {
        const double m = s[a] / n, var = s[3 + a] / n - m * m;
        if (var > bv) {
            bv = var;
            best = a;
        }
    }
    return best;
}

/* Serial sweep-and-prune on the given axis (the CPU baseline of the broadphase):
 * stable sort of the lower extents, then for every box walk forward while the
 * next lower extent is within its upper extent, testing the other two axes. */
long or_sap_pairs(int n, const float* lo4, const float* hi4, int axis, int* pairs, long capacity) {
    if (n <= 0) return 0;
    uint32_t* keys = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* order = (uint32_t*)malloc((size_t)n * 4);
    for (int i = 0; i < n; ++i) {
        keys[i] = fkey(lo4[4 * (long)i + axis]);
        order[i] = (uint32_t)i;
    }
    or_sort_pairs_u32(n, keys, order, 0, 32);
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    long np = 0;
    for (int t = 0; t < n; ++t) {
        const float* l = lo4 + 4 * (long)order[t];
        const float* h = hi4 + 4 * (long)order[t];
        for (int u = t + 1; u < n; ++u) {
            const float* lu = lo4 + 4 * (long)order[u];
            if (lu[axis] > h[axis]) break;
            const float* hu = hi4 + 4 * (long)order[u];
            if (lu[a1] <= h[a1] && l[a1] <= hu[a1] && lu[a2] <= h[a2] && l[a2] <= hu[a2]) {
                if (np < capacity) {
                    const int i = (int)order[t], j = (int)order[u];
                    pairs[2 * np] = i < j ? i : j;
                    pairs[2 * np + 1] = i < j ? j : i;
                }
                ++np;
            }
        }
    }
    free(keys);
    free(order);
    return np;
}

long or_brute_aabb_pairs(int n, const float* lo4, const float* hi4, int* pairs, long capacity) {
    long np = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            int ok = 1;
            for (int a = 0; a < 3; ++a)
                ok = ok && lo4[4 * (long)j + a] <= hi4[4 * (long)i + a] && lo4[4 * (long)i + a] <= hi4[4 * (long)j + a];
            if (ok) {
                if (np < capacity) {
                    pairs[2 * np] = i;
                    pairs[2 * np + 1] = j;
                }
                ++np;
            }
        }
    return np;
}

/* ------------------------------------------------------------------ GJK / EPA */

#define KEYMUL 1024
#define REL_EPS 1e-6      /* GJK (double) */
#define TOUCH_EPS2 1e-14
#define STALL_EPS 1e-4
#define EPA_REL_EPS 1e-6f /* EPA (float) */
#define MAXV 48
#define MAXF 96
#define MAXE 48

typedef struct {
    const float* verts; /* local (x, y, z, 0) rows */
    int nv;
    float radius;
    float m[9];
    V3 off;
} Shape;

typedef struct {
    V3 p;
    int key;
} MP;

static V3 support(const Shape* s, V3 u, int* idx) {
    if (s->nv == 0) {
        *idx = -1;
        const float uu = vdot(u, u);
        if (uu == 0.0f) return s->off;
        const float k = s->radius / sqrtf(uu);
        return vadd(s->off, vscl(u, k));
    }
    const V3 l = mtvec(s->m, u);
    int best = 0;
    float bs = -3.402823466e38f;
    for (int k = 0; k < s->nv; ++k) {
        const float* p = s->verts + 4 * (long)k;
        const float d = p[0] * l.x + p[1] * l.y + p[2] * l.z;
        if (d > bs) {
            bs = d;
            best = k;
        }
    }
    *idx = best;
    const float* p = s->verts + 4 * (long)best;
    return vadd(mvec(s->m, mk(p[0], p[1], p[2])), s->off);
}

static MP supM(const Shape* A, const Shape* B, V3 u) {
    int ia, ib;
    const V3 pa = support(A, u, &ia);
    const V3 pb = support(B, vneg(u), &ib);
    MP r;
    r.p = vsub(pa, pb);
    r.key = (ia < 0 || ib < 0) ? -1 : ia * KEYMUL + ib;
    return r;
}

/* The GJK simplex and everything derived from it are evaluated in double (the
 * support points themselves are FP32 values, stored exactly). */
typedef struct {
    double x, y, z;
} D3;
static D3 dk(double x, double y, double z) {
    D3 r = {x, y, z};
    return r;
}
static D3 d_of(V3 a) { return dk((double)a.x, (double)a.y, (double)a.z); }
static V3 f_of(D3 a) { return mk((float)a.x, (float)a.y, (float)a.z); }
static D3 dadd(D3 a, D3 b) { return dk(a.x + b.x, a.y + b.y, a.z + b.z); }
static D3 dsub(D3 a, D3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }
static D3 dneg(D3 a) { return dk(-a.x, -a.y, -a.z); }
static D3 dscl(D3 a, double s) { return dk(a.x * s, a.y * s, a.z * s); }
static double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static D3 dcross(D3 a, D3 b) { return dk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

typedef struct {
    D3 p;
    int key;
} SP;
static SP sp_of(MP m) {
    SP r;
    r.p = d_of(m.p);
    r.key = m.key;
    return r;
}

typedef struct {
    SP w[4];
    int n;
} Simplex;

static D3 cl_segment(Simplex* s) {
    const D3 a = s->w[0].p, b = s->w[1].p;
    const D3 ab = dsub(b, a);
    const double t = -ddot(a, ab);
    if (t <= 0.0) {
        s->n = 1;
        return a;
    }
    const double den = ddot(ab, ab);
    if (t >= den) {
        s->w[0] = s->w[1];
        s->n = 1;
        return b;
    }
    return dadd(a, dscl(ab, t / den));
}

/* closest point of a triangle to the origin by Voronoi regions (vertex, edge, face) */
static D3 cl_triangle(Simplex* s) {
    const SP A = s->w[0], B = s->w[1], C = s->w[2];
    const D3 a = A.p, b = B.p, c = C.p;
    const D3 ab = dsub(b, a), ac = dsub(c, a), ap = dneg(a);
    const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        s->n = 1;
        return a;
    }
    const D3 bp = dneg(b);
    const double d3 = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        s->w[0] = B;
        s->n = 1;
        return b;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double t = d1 / (d1 - d3);
        s->n = 2;
        return dadd(a, dscl(ab, t));
    }
    const D3 cp = dneg(c);
    const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        s->w[0] = C;
        s->n = 1;
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double t = d2 / (d2 - d6);
        s->w[1] = C;
        s->n = 2;
        return dadd(a, dscl(ac, t));
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        s->w[0] = B;
        s->w[1] = C;
        s->n = 2;
        return dadd(b, dscl(dsub(c, b), t));
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    return dadd(dadd(a, dscl(ab, v)), dscl(ac, w));
}

static int outside(D3 a, D3 b, D3 c, D3 d) {
    const D3 nrm = dcross(dsub(b, a), dsub(c, a));
    const double so = -ddot(a, nrm);
    const double sd = ddot(dsub(d, a), nrm);
    return so * sd < 0.0;
}

/* A near-flat tetrahedron (|6V| <= 1e-5 |ab||ac||ad|) never counts as enclosing
 * the origin; then every face is searched. */
static D3 cl_tetra(Simplex* s, int* inside) {
    const SP P[4] = {s->w[0], s->w[1], s->w[2], s->w[3]};
    static const int F[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
    const D3 ab = dsub(P[1].p, P[0].p), ac = dsub(P[2].p, P[0].p), ad = dsub(P[3].p, P[0].p);
    const double vol = ddot(ad, dcross(ab, ac));
    const double lim = 1e-5 * (sqrt(ddot(ab, ab)) * sqrt(ddot(ac, ac)) * sqrt(ddot(ad, ad)));
    const int flat = !(fabs(vol) > lim);
    double best = 1.79769313486231e308;
    D3 bestp = dk(0.0, 0.0, 0.0);
    Simplex bests = *s;
    int any = 0;
    for (int f = 0; f < 4; ++f) {
        if (!flat && !outside(P[F[f][0]].p, P[F[f][1]].p, P[F[f][2]].p, P[F[f][3]].p)) continue;
        any = 1;
        Simplex t = *s;
        t.w[0] = P[F[f][0]];
        t.w[1] = P[F[f][1]];
        t.w[2] = P[F[f][2]];
        t.n = 3;
        const D3 q = cl_triangle(&t);
        const double dq = ddot(q, q);
        if (dq < best) {
            best = dq;
            bestp = q;
            bests = t;
        }
    }
    *inside = !any;
    if (any) *s = bests;
    return bestp;
}

static D3 cl_simplex(Simplex* s, int* inside) {
    *inside = 0;
    if (s->n == 1) return s->w[0].p;
    if (s->n == 2) return cl_segment(s);
    if (s->n == 3) return cl_triangle(s);
    return cl_tetra(s, inside);
}

static void sort_by_key(SP* w, int n) {
    for (int i = 1; i < n; ++i) {
        SP x = w[i];
        int j = i - 1;
        while (j >= 0 && w[j].key > x.key) {
            w[j + 1] = w[j];
            --j;
        }
        w[j + 1] = x;
    }
}

typedef struct {
    int a, b, c;
    V3 n;
    float d;
} Face;

static int make_face(const MP* P, int a, int b, int c, V3 inner, Face* f) {
    V3 nrm = vcross(vsub(P[b].p, P[a].p), vsub(P[c].p, P[a].p));
    const float l2 = vdot(nrm, nrm);
    if (!(l2 > 0.0f)) return 0;
    if (vdot(nrm, vsub(P[a].p, inner)) < 0.0f) {
        const int t = b;
        b = c;
        c = t;
        nrm = vneg(nrm);
    }
    const float inv = 1.0f / sqrtf(l2);
    f->a = a;
    f->b = b;
    f->c = c;
    f->n = vscl(nrm, inv);
    f->d = vdot(f->n, P[a].p);
    return 1;
}

typedef struct {
    float sep;
    V3 n;
    int flags, iters, epa_iters;
} Out;

enum { HIT = 1, CONV = 2, BUDGET = 4, ELIMIT = 8, DEGEN = 16, WARM = 32 };

static void run_epa(const Shape* A, const Shape* B, const Simplex* s0, int max_iter, Out* o) {
    MP P[MAXV];
    int np = 0;
    for (int i = 0; i < s0->n; ++i) {
        P[np].p = f_of(s0->w[i].p); /* exact: FP32 support points */
        P[np].key = s0->w[i].key;
        ++np;
    }
    const V3 axes[3] = {mk(1.0f, 0.0f, 0.0f), mk(0.0f, 1.0f, 0.0f), mk(0.0f, 0.0f, 1.0f)};
    if (np == 1) {
        for (int k = 0; k < 6 && np == 1; ++k) {
            const V3 u = (k & 1) ? vneg(axes[k >> 1]) : axes[k >> 1];
            const MP m = supM(A, B, u);
            const V3 dd = vsub(m.p, P[0].p);
            if (vdot(dd, dd) > 1e-20f) P[np++] = m;
        }
    }
    if (np == 2) {
        const V3 ab = vsub(P[1].p, P[0].p);
        const float ax = fabsf(ab.x), ay = fabsf(ab.y), az = fabsf(ab.z);
        const V3 e = (ax <= ay && ax <= az) ? axes[0] : (ay <= az ? axes[1] : axes[2]);
        const V3 u = vcross(ab, e);
        for (int k = 0; k < 2 && np == 2; ++k) {
            const MP m = supM(A, B, k ? vneg(u) : u);
            const V3 c = vcross(ab, vsub(m.p, P[0].p));
            if (vdot(c, c) > 1e-24f) P[np++] = m;
        }
    }
    if (np == 3) {
        const V3 nrm = vcross(vsub(P[1].p, P[0].p), vsub(P[2].p, P[0].p));
        for (int k = 0; k < 2 && np == 3; ++k) {
            const MP m = supM(A, B, k ? vneg(nrm) : nrm);
            const float vol = vdot(nrm, vsub(m.p, P[0].p));
            if (fabsf(vol) > 1e-24f) P[np++] = m;
        }
    }
    if (np < 4) {
        o->flags |= HIT | DEGEN;
        o->sep = 0.0f;
        return;
    }
    const V3 inner = vscl(vadd(vadd(P[0].p, P[1].p), vadd(P[2].p, P[3].p)), 0.25f);
    Face F[MAXF];
    int nf = 0;
    static const int T[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    for (int k = 0; k < 4; ++k)
        if (make_face(P, T[k][0], T[k][1], T[k][2], inner, &F[nf])) ++nf;
    if (nf < 4) {
        o->flags |= HIT | DEGEN;
        o->sep = 0.0f;
        return;
    }
    int best = 0, it = 0, converged = 0;
    for (; it < max_iter; ++it) {
        best = 0;
        for (int f = 1; f < nf; ++f)
            if (F[f].d < F[best].d) best = f;
        const Face fb = F[best];
        const MP m = supM(A, B, fb.n);
        const float gap = vdot(m.p, fb.n) - fb.d;
        if (gap <= EPA_REL_EPS * fabsf(fb.d) + 1e-9f) {
            converged = 1;
            break;
        }
        int dup = 0;
        for (int i = 0; i < np; ++i) dup = dup || (m.key >= 0 && P[i].key == m.key);
        if (dup || np >= MAXV) {
            converged = dup;
            if (!dup) o->flags |= ELIMIT;
            break;
        }
        const int iv = np;
        P[np++] = m;
        /* faces that see the new point; the horizon is every edge (a, b), (b, c),
         * (c, a) of a visible face, in face order, whose reverse edge belongs to no
         * visible face; the other faces survive in order */
        int vis[MAXF];
        for (int f = 0; f < nf; ++f) vis[f] = vdot(F[f].n, vsub(m.p, P[F[f].a].p)) > 0.0f;
        int ea[MAXE], eb[MAXE], ne = 0, overflow = 0, w = 0;
        for (int f = 0; f < nf; ++f) {
            if (!vis[f]) continue;
            const int e[3][2] = {{F[f].a, F[f].b}, {F[f].b, F[f].c}, {F[f].c, F[f].a}};
            for (int k = 0; k < 3; ++k) {
                int twin = 0;
                for (int g = 0; g < nf && !twin; ++g) {
                    if (!vis[g]) continue;
                    twin = (F[g].a == e[k][1] && F[g].b == e[k][0]) || (F[g].b == e[k][1] && F[g].c == e[k][0]) ||
                           (F[g].c == e[k][1] && F[g].a == e[k][0]);
                }
                if (twin) continue;
                if (ne < MAXE) {
                    ea[ne] = e[k][0];
                    eb[ne] = e[k][1];
                    ++ne;
                } else {
                    overflow = 1;
                }
            }
        }
        for (int f = 0; f < nf; ++f)
            if (!vis[f]) F[w++] = F[f];
        nf = w;
        if (overflow || nf + ne > MAXF) {
            o->flags |= ELIMIT;
            break;
        }
        for (int q = 0; q < ne; ++q)
            if (make_face(P, ea[q], eb[q], iv, inner, &F[nf])) ++nf;
        if (nf == 0) {
            o->flags |= DEGEN;
            break;
        }
    }
    if (nf == 0) {
        o->flags |= HIT | DEGEN;
        o->sep = 0.0f;
        o->epa_iters = it;
        return;
    }
    best = 0;
    for (int f = 1; f < nf; ++f)
        if (F[f].d < F[best].d) best = f;
    /* canonical depth / normal: the face's vertices ordered by key */
    MP q[3] = {P[F[best].a], P[F[best].b], P[F[best].c]};
    for (int i = 1; i < 3; ++i) {
        MP x = q[i];
        int j = i - 1;
        while (j >= 0 && q[j].key > x.key) {
            q[j + 1] = q[j];
            --j;
        }
        q[j + 1] = x;
    }
    V3 nrm = vcross(vsub(q[1].p, q[0].p), vsub(q[2].p, q[0].p));
    const float l2 = vdot(nrm, nrm);
    if (vdot(nrm, vsub(q[0].p, inner)) < 0.0f) nrm = vneg(nrm);
    const V3 n = vscl(nrm, 1.0f / sqrtf(l2));
    const float d = vdot(n, q[0].p);
    o->flags |= HIT | (converged ? CONV : 0);
    if (!converged && it >= max_iter) o->flags |= BUDGET;
    if (d < 0.0f) o->flags |= DEGEN; /* polytope does not enclose the origin */
    o->sep = -d;
    o->n = n;
    o->epa_iters = it;
}

static Out query(const Shape* A, const Shape* B, V3 init, int max_iter) {
    Out o;
    o.flags = 0;
    o.iters = 0;
    o.epa_iters = 0;
    o.sep = 0.0f;
    o.n = mk(1.0f, 0.0f, 0.0f);
    if (vdot(init, init) == 0.0f) init = mk(1.0f, 0.0f, 0.0f);
    Simplex s;
    s.w[0] = sp_of(supM(A, B, init));
    s.n = 1;
    D3 v = s.w[0].p;
    int hit = 0, converged = 0, uncertain = 0, it = 0;
    while (it < max_iter) {
        const double vv = ddot(v, v);
        if (vv <= TOUCH_EPS2) {
            hit = 1;
            break;
        }
        const SP w = sp_of(supM(A, B, f_of(dneg(v))));
        ++it;
        const double vw = ddot(v, w.p);
        if (vv - vw <= REL_EPS * vv) {
            converged = 1;
            break;
        }
        /* a repeated support vertex or a stalled distance ends the query: a
         * separation only when the duality gap certifies it to 1e-4; without a
         * certificate and with v.w <= 0 the exit is uncertain and goes to EPA */
        int dup = 0;
        for (int i = 0; i < s.n; ++i) dup = dup || (w.key >= 0 && s.w[i].key == w.key);
        if (dup) {
            if (vv - vw <= STALL_EPS * vv) converged = 1;
            else if (vw <= 0.0) hit = uncertain = 1;
            break;
        }
        const Simplex sp = s;
        s.w[s.n++] = w;
        int inside = 0;
        const D3 vn = cl_simplex(&s, &inside);
        if (inside) {
            hit = 1;
            break;
        }
        /* the distance must strictly decrease; otherwise the previous feature is
         * final when the duality gap certifies it to 1e-4 */
        if (!(ddot(vn, vn) < vv)) {
            s = sp;
            if (vv - vw <= STALL_EPS * vv) converged = 1;
            else if (vw <= 0.0) hit = uncertain = 1;
            break;
        }
        v = vn;
    }
    o.iters = it;
    if (hit) {
        run_epa(A, B, &s, max_iter, &o);
        if (!(uncertain && (o.flags & DEGEN))) return o;
        /* no enclosing polytope after an uncertain exit: the uncertified separation */
        const int ei = o.epa_iters;
        o.flags = 0;
        o.epa_iters = ei;
        o.sep = 0.0f;
        o.n = mk(1.0f, 0.0f, 0.0f);
        int inside = 0;
        v = cl_simplex(&s, &inside);
        converged = 0;
    }
    o.flags |= converged ? CONV : BUDGET;
    int canon = 1;
    for (int i = 0; i < s.n; ++i) canon = canon && s.w[i].key >= 0;
    if (canon) {
        sort_by_key(s.w, s.n);
        int inside = 0;
        v = cl_simplex(&s, &inside);
    }
    const double dist = sqrt(ddot(v, v));
    o.sep = (float)dist;
    o.n = dist > 0.0 ? f_of(dscl(v, -1.0 / dist)) : mk(1.0f, 0.0f, 0.0f);
    return o;
}

/* Per pair (a, b): same outputs and warm/cold policy as the GPU entry point. */
void or_gjk(int npairs, const int* pairs, const float* pos4, const float* quat4, const int* vstart, const int* nv,
            const float* verts4, const float* warm, int warm_budget, int max_iter, float near_margin, float* sep,
            float* normal3, int* flags, int* iters, int nthreads) {
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (int p = 0; p < npairs; ++p) {
        const int ia = pairs[2 * (long)p], ib = pairs[2 * (long)p + 1];
        const float* ca = pos4 + 4 * (long)ia;
        const float* cb = pos4 + 4 * (long)ib;
        Shape A, B;
        A.verts = verts4 + 4 * (long)vstart[ia];
        A.nv = nv[ia];
        A.radius = ca[3];
        qmat(quat4 + 4 * (long)ia, A.m);
        A.off = mk(0.0f, 0.0f, 0.0f);
        B.verts = verts4 + 4 * (long)vstart[ib];
        B.nv = nv[ib];
        B.radius = cb[3];
        qmat(quat4 + 4 * (long)ib, B.m);
        const V3 dc = mk(cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]);
        B.off = dc;
        Out o;
        int spent = 0, done = 0;
        if (warm) {
            const V3 wd = mk(warm[3 * (long)p], warm[3 * (long)p + 1], warm[3 * (long)p + 2]);
            o = query(&A, &B, wd, warm_budget < max_iter ? warm_budget : max_iter);
            spent = o.iters + o.epa_iters;
            done = (o.flags & CONV) && !(o.flags & (DEGEN | ELIMIT)) && fabsf(o.sep) > near_margin;
            if (done) o.flags |= WARM;
        }
        if (!done) {
            o = query(&A, &B, dc, max_iter);
            spent += o.iters + o.epa_iters;
        }
        sep[p] = o.sep;
        normal3[3 * (long)p] = o.n.x;
        normal3[3 * (long)p + 1] = o.n.y;
        normal3[3 * (long)p + 2] = o.n.z;
        flags[p] = o.flags;
        iters[p] = spent;
    }
}
