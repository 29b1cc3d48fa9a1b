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

/* Axis of largest box-centre variance, in FP32 and in the order of the GPU's
 * reduction: centres relative to the first box's centre, per-thread sums over
 * a grid stride of nb = min(ceil(n / 256), 256) blocks of 256 threads, a tree
 * over each block's 256 sums (halving strides), block sums added in order,
 * var = S2 / n - (S1 / n)^2, the first axis of largest variance. */
int or_sap_axis(int n, const float* lo4, const float* hi4) {
    if (n <= 0) return 0;
    const int blk = 256;
    int nb = (n + blk - 1) / blk;
    if (nb > 256) nb = 256;
    float r[3];
    for (int a = 0; a < 3; ++a) r[a] = 0.5f * (lo4[a] + hi4[a]);
    float s[6] = {0, 0, 0, 0, 0, 0};
    float red[6][256];
    for (int b = 0; b < nb; ++b) {
        for (int t = 0; t < blk; ++t) {
            float acc[6] = {0, 0, 0, 0, 0, 0};
            for (long i = (long)b * blk + t; i < n; i += (long)nb * blk)
                for (int a = 0; a < 3; ++a) {
                    const float d = 0.5f * (lo4[4 * i + a] + hi4[4 * i + a]) - r[a];
                    acc[a] += d;
                    acc[3 + a] += d * d;
                }
            for (int k = 0; k < 6; ++k) red[k][t] = acc[k];
        }
        for (int h = blk / 2; h > 0; h >>= 1)
            for (int t = 0; t < h; ++t)
                for (int k = 0; k < 6; ++k) red[k][t] += red[k][t + h];
        for (int k = 0; k < 6; ++k) s[k] += red[k][0];
    }
    int best = 0;
    float bv = -1.0f;
    for (int a = 0; a < 3; ++a) {
        const float m = s[a] / (float)n;
        const float var = s[3 + a] / (float)n - m * m;
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

/* ------------------------------------------------------------------ GJK / EPA
 *
 * All arithmetic is FP32, written out in the order the GPU kernels use
 * (compiled without contraction). M = A - B in the query frame: A's centre at
 * the origin, B's centre at dc = c_B - c_A. Vertex keys are ia * 1024 + ib.
 *
 * GJK ends on an exact exit (the support point of M along -v is already a
 * simplex vertex), on a stall (adding it does not decrease |v|), inside (hit),
 * or on the iteration budget. The reported separation / depth and normal are
 * a function of the closest features alone: the sorted vertex sets IA, IB of
 * the final simplex (or EPA face) give vertex-vertex, vertex-edge, edge-vertex
 * distances, or the plane distance of a vertex-face, face-vertex or edge-edge
 * feature. Two queries that end on the same features report the same bits,
 * whatever their start.
 *
 * A warm-started query seeds its simplex from a predicted direction u: the
 * eight vertices of A furthest along u and of B furthest along -u form a small
 * problem (64 points of M) solved first, and GJK on the full hulls continues
 * from its simplex. A warm result is accepted only with a certificate that its
 * features are the unique closest ones: every vertex outside the features lies
 * a margin below the supporting plane along the normal, and the closest points
 * lie inside the features. Otherwise the pair is solved cold.
 */

#define KEYMUL 1024
#define TOUCH_EPS2 1e-14f   /* |v|^2 below this: origin on the boundary (hit) */
#define REL_EPS 1e-6f       /* sphere shapes: v.v - v.w <= eps v.v ends the query */
#define STALL_EPS 1e-4f     /* gap accepted when the distance stops decreasing */
#define EPA_REL_EPS 1e-6f
#define CERT_REL 1.52587890625e-05f /* 2^-16: support margin relative to the features' vertex radii */
#define CERT_TAU 9.765625e-04f      /* 2^-10: closest points this far inside their features */
#define WARM_K 8           /* warm seed: vertices of each hull furthest along the predicted direction */
#define SMALL_MAX_ITER 16
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

/* vertex k of s in the query frame */
static V3 vert_at(const Shape* s, int k) {
    const float* p = s->verts + 4 * (long)k;
    return vadd(mvec(s->m, mk(p[0], p[1], p[2])), s->off);
}

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
    return vert_at(s, best);
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

/* The K vertices of s furthest along u (larger projection first, lower index on
 * ties), in the query frame. Returns their number (min(K, nv)). */
static int top_k(const Shape* s, V3 u, int* idx, V3* pts) {
    const V3 l = mtvec(s->m, u);
    float val[WARM_K];
    int n = 0;
    for (int k = 0; k < s->nv; ++k) {
        const float* p = s->verts + 4 * (long)k;
        const float d = p[0] * l.x + p[1] * l.y + p[2] * l.z;
        int pos = n < WARM_K ? n : WARM_K;
        while (pos > 0 && d > val[pos - 1]) --pos;
        if (pos >= WARM_K) continue;
        const int last = n < WARM_K ? n : WARM_K - 1;
        for (int j = last; j > pos; --j) {
            val[j] = val[j - 1];
            idx[j] = idx[j - 1];
        }
        val[pos] = d;
        idx[pos] = k;
        if (n < WARM_K) ++n;
    }
    for (int j = 0; j < n; ++j) pts[j] = vert_at(s, idx[j]);
    return n;
}

/* Support over the full shapes, or over two small vertex sets (the warm seed). */
typedef struct {
    const Shape* A;
    const Shape* B;
    int small;
    int na, nb;
    int ia[WARM_K], ib[WARM_K];
    V3 pa[WARM_K], pb[WARM_K];
} Sup;

static MP sup_eval(const Sup* S, V3 u) {
    if (!S->small) return supM(S->A, S->B, u);
    /* larger projection, lower vertex index on ties (independent of the order of the sets) */
    int ka = 0, kb = 0;
    float ba = vdot(S->pa[0], u), bb = -vdot(S->pb[0], u);
    for (int j = 1; j < S->na; ++j) {
        const float d = vdot(S->pa[j], u);
        if (d > ba || (d == ba && S->ia[j] < S->ia[ka])) {
            ba = d;
            ka = j;
        }
    }
    for (int j = 1; j < S->nb; ++j) {
        const float d = -vdot(S->pb[j], u);
        if (d > bb || (d == bb && S->ib[j] < S->ib[kb])) {
            bb = d;
            kb = j;
        }
    }
    MP r;
    r.p = vsub(S->pa[ka], S->pb[kb]);
    r.key = S->ia[ka] * KEYMUL + S->ib[kb];
    return r;
}

typedef struct {
    MP w[4];
    int n;
} Simplex;

static V3 cl_segment(Simplex* s) {
    const V3 a = s->w[0].p, b = s->w[1].p;
    const V3 ab = vsub(b, a);
    const float t = -vdot(a, ab);
    if (t <= 0.0f) {
        s->n = 1;
        return a;
    }
    const float den = vdot(ab, ab);
    if (t >= den) {
        s->w[0] = s->w[1];
        s->n = 1;
        return b;
    }
    return vadd(a, vscl(ab, t / den));
}

/* closest point of a triangle to the origin by Voronoi regions (vertex, edge, face) */
static V3 cl_triangle(Simplex* s) {
    const MP A = s->w[0], B = s->w[1], C = s->w[2];
    const V3 a = A.p, b = B.p, c = C.p;
    const V3 ab = vsub(b, a), ac = vsub(c, a), ap = vneg(a);
    const float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        s->n = 1;
        return a;
    }
    const V3 bp = vneg(b);
    const float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        s->w[0] = B;
        s->n = 1;
        return b;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float t = d1 / (d1 - d3);
        s->n = 2;
        return vadd(a, vscl(ab, t));
    }
    const V3 cp = vneg(c);
    const float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        s->w[0] = C;
        s->n = 1;
        return c;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        s->w[1] = C;
        s->n = 2;
        return vadd(a, vscl(ac, t));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        s->w[0] = B;
        s->w[1] = C;
        s->n = 2;
        return vadd(b, vscl(vsub(c, b), t));
    }
    const float den = 1.0f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    return vadd(vadd(a, vscl(ab, v)), vscl(ac, w));
}

static int outside(V3 a, V3 b, V3 c, V3 d) {
    const V3 nrm = vcross(vsub(b, a), vsub(c, a));
    const float so = -vdot(a, nrm);
    const float sd = vdot(vsub(d, a), nrm);
    return so * sd < 0.0f;
}

/* A near-flat tetrahedron (|6V| <= 1e-5 |ab||ac||ad|) never counts as enclosing
 * the origin; then every face is searched. Faces in search order: (0,1,2),
 * (0,2,3), (0,3,1), (1,3,2); the first face of smallest distance wins. */
static V3 cl_tetra(Simplex* s, int* inside) {
    const MP P[4] = {s->w[0], s->w[1], s->w[2], s->w[3]};
    static const int F[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
    const V3 ab = vsub(P[1].p, P[0].p), ac = vsub(P[2].p, P[0].p), ad = vsub(P[3].p, P[0].p);
    const float vol = vdot(ad, vcross(ab, ac));
    const float lim = 1e-5f * (sqrtf(vdot(ab, ab)) * sqrtf(vdot(ac, ac)) * sqrtf(vdot(ad, ad)));
    const int flat = !(fabsf(vol) > lim);
    float best = 3.402823466e38f;
    int bi = -1;
    for (int f = 0; f < 4; ++f) {
        if (!flat && !outside(P[F[f][0]].p, P[F[f][1]].p, P[F[f][2]].p, P[F[f][3]].p)) continue;
        Simplex t;
        t.w[0] = P[F[f][0]];
        t.w[1] = P[F[f][1]];
        t.w[2] = P[F[f][2]];
        t.n = 3;
        const V3 q = cl_triangle(&t);
        const float dq = vdot(q, q);
        if (dq < best) {
            best = dq;
            bi = f;
        }
    }
    *inside = bi < 0;
    if (bi < 0) return mk(0.0f, 0.0f, 0.0f);
    s->w[0] = P[F[bi][0]];
    s->w[1] = P[F[bi][1]];
    s->w[2] = P[F[bi][2]];
    s->n = 3;
    return cl_triangle(s);
}

static V3 cl_simplex(Simplex* s, int* inside) {
    *inside = 0;
    if (s->n == 1) return s->w[0].p;
    if (s->n == 2) return cl_segment(s);
    if (s->n == 3) return cl_triangle(s);
    return cl_tetra(s, inside);
}

static void sort_by_key(MP* w, int n) {
    for (int i = 1; i < n; ++i) {
        MP x = w[i];
        int j = i - 1;
        while (j >= 0 && w[j].key > x.key) {
            w[j + 1] = w[j];
            --j;
        }
        w[j + 1] = x;
    }
}

typedef struct {
    int hit, converged, uncertain, it;
} GjkState;

/* GJK iterations from simplex s with closest point *vp, at most max_iter
 * support evaluations (counted in g->it). */
static void gjk_loop(const Sup* S, Simplex* s, V3* vp, int max_iter, GjkState* g) {
    V3 v = *vp;
    int it = 0;
    while (it < max_iter) {
        const float vv = vdot(v, v);
        if (vv <= TOUCH_EPS2) {
            g->hit = 1;
            break;
        }
        const MP w = sup_eval(S, vneg(v));
        ++it;
        const float vw = vdot(v, w.p);
        const float gap = vv - vw;
        int dup = 0;
        for (int i = 0; i < s->n; ++i) dup = dup || (w.key >= 0 && s->w[i].key == w.key);
        /* exact exit: the support point is already a simplex vertex, so in
         * exact arithmetic v.w >= v.v and any gap is rounding. v.w > 0 makes
         * v a separating axis (converged); v.w <= 0 leaves the verdict to the
         * rounding of v's direction: an uncertain hit, which finish() settles.
         * A sphere shape (keyless) ends on the relative gap instead. */
        if (dup || (w.key < 0 && gap <= REL_EPS * vv)) {
            if (vw > 0.0f) {
                g->converged = 1;
            } else {
                g->hit = 1;
                g->uncertain = 1;
            }
            break;
        }
        const Simplex sp = *s;
        s->w[s->n++] = w;
        int inside = 0;
        const V3 vn = cl_simplex(s, &inside);
        if (inside) {
            g->hit = 1;
            break;
        }
        if (!(vdot(vn, vn) < vv)) {
            /* stalled at FP32 resolution: the previous feature is final if the
             * duality gap is within 1e-4 */
            *s = sp;
            if (gap <= STALL_EPS * vv) {
                g->converged = 1;
            } else if (vw <= 0.0f) {
                g->hit = 1;
                g->uncertain = 1;
            }
            break;
        }
        v = vn;
    }
    g->it += it;
    *vp = v;
}

/* ---------------------------------------------------------- canonical features */

typedef struct {
    float sep;
    V3 n;
    int ok;   /* a recognised feature pair with its closest points inside it */
    int kind; /* 0 degenerate, else na * 4 + nb */
} Canon;

static V3 unit_or_x(V3 d, float l) { return l > 0.0f ? vscl(d, 1.0f / l) : mk(1.0f, 0.0f, 0.0f); }

/* barycentric (u, v, w) of p in triangle (a, b, c), all >= tau */
static int inside_tri(V3 p, V3 a, V3 b, V3 c) {
    const V3 e1 = vsub(b, a), e2 = vsub(c, a), r = vsub(p, a);
    const float d00 = vdot(e1, e1), d01 = vdot(e1, e2), d11 = vdot(e2, e2);
    const float d20 = vdot(r, e1), d21 = vdot(r, e2);
    const float den = d00 * d11 - d01 * d01;
    if (!(den > 0.0f)) return 0;
    const float v = (d11 * d20 - d01 * d21) / den;
    const float w = (d00 * d21 - d01 * d20) / den;
    const float u = 1.0f - v - w;
    return u >= CERT_TAU && v >= CERT_TAU && w >= CERT_TAU;
}

/* Distance of the features conv(pa[0..na)) of A and conv(pb[0..nb)) of B,
 * vertices ordered by index. hit: the shapes intersect (the plane distance is
 * then a depth, reported negative). The normal n points from A to B. */
static Canon canon_feature(const V3* pa, int na, const V3* pb, int nb, int hit) {
    Canon c;
    c.ok = 0;
    c.kind = 0;
    c.sep = 0.0f;
    c.n = mk(1.0f, 0.0f, 0.0f);
    if (!hit && na == 1 && nb == 1) {
        const V3 d = vsub(pb[0], pa[0]);
        const float l = sqrtf(vdot(d, d));
        c.sep = l;
        c.n = unit_or_x(d, l);
        c.ok = l > 0.0f;
        c.kind = 5;
        return c;
    }
    if (!hit && ((na == 1 && nb == 2) || (na == 2 && nb == 1))) {
        /* point p against segment (s0, s1); d points from A to B */
        const int pa_is_point = na == 1;
        const V3 p = pa_is_point ? pa[0] : pb[0];
        const V3 s0 = pa_is_point ? pb[0] : pa[0];
        const V3 s1 = pa_is_point ? pb[1] : pa[1];
        const V3 e = vsub(s1, s0);
        const float ee = vdot(e, e);
        if (!(ee > 0.0f)) return c;
        const float t = vdot(vsub(p, s0), e) / ee;
        const V3 q = vadd(s0, vscl(e, t));
        const V3 d = pa_is_point ? vsub(q, p) : vsub(p, q);
        const float l = sqrtf(vdot(d, d));
        c.sep = l;
        c.n = unit_or_x(d, l);
        c.ok = l > 0.0f && t >= CERT_TAU && t <= 1.0f - CERT_TAU;
        c.kind = na * 4 + nb;
        return c;
    }
    V3 e1, e2, ref_a = pa[0], ref_b = pb[0];
    if (na == 1 && nb == 3) {
        e1 = vsub(pb[1], pb[0]);
        e2 = vsub(pb[2], pb[0]);
    } else if (na == 3 && nb == 1) {
        e1 = vsub(pa[1], pa[0]);
        e2 = vsub(pa[2], pa[0]);
    } else if (na == 2 && nb == 2) {
        e1 = vsub(pa[1], pa[0]);
        e2 = vsub(pb[1], pb[0]);
    } else {
        return c;
    }
    const V3 cr = vcross(e1, e2);
    const float l = sqrtf(vdot(cr, cr));
    if (!(l > 0.0f)) return c;
    V3 m = vscl(cr, 1.0f / l);
    float h = vdot(vsub(ref_b, ref_a), m);
    if (hit ? h > 0.0f : h < 0.0f) {
        m = vneg(m);
        h = -h;
    }
    c.sep = h;
    c.n = m;
    c.kind = na * 4 + nb;
    if (na == 1) {
        /* a projected onto B's face plane */
        c.ok = inside_tri(vadd(pa[0], vscl(m, h)), pb[0], pb[1], pb[2]);
    } else if (nb == 1) {
        c.ok = inside_tri(vsub(pb[0], vscl(m, h)), pa[0], pa[1], pa[2]);
    } else {
        /* closest parameters of the two edge lines */
        const V3 r = vsub(pa[0], pb[0]);
        const float a = vdot(e1, e1), b = vdot(e1, e2), e = vdot(e2, e2);
        const float cc = vdot(e1, r), f = vdot(e2, r);
        const float den = a * e - b * b;
        if (den > 0.0f) {
            const float s = (b * f - cc * e) / den;
            const float t = (a * f - b * cc) / den;
            c.ok = s >= CERT_TAU && s <= 1.0f - CERT_TAU && t >= CERT_TAU && t <= 1.0f - CERT_TAU;
        }
    }
    return c;
}

/* Sorted distinct vertex indices of A (key / 1024) and B (key % 1024) of n
 * keyed points; returns 0 if a point has no key (sphere shape). */
static int feature_sets(const MP* w, int n, int* IA, int* na, int* IB, int* nb) {
    *na = 0;
    *nb = 0;
    for (int i = 0; i < n; ++i) {
        if (w[i].key < 0) return 0;
        const int a = w[i].key / KEYMUL, b = w[i].key % KEYMUL;
        int da = 0, db = 0;
        for (int j = 0; j < *na; ++j) da = da || IA[j] == a;
        for (int j = 0; j < *nb; ++j) db = db || IB[j] == b;
        if (!da) IA[(*na)++] = a;
        if (!db) IB[(*nb)++] = b;
    }
    for (int i = 1; i < *na; ++i)
        for (int j = i; j > 0 && IA[j - 1] > IA[j]; --j) {
            const int t = IA[j];
            IA[j] = IA[j - 1];
            IA[j - 1] = t;
        }
    for (int i = 1; i < *nb; ++i)
        for (int j = i; j > 0 && IB[j - 1] > IB[j]; --j) {
            const int t = IB[j];
            IB[j] = IB[j - 1];
            IB[j - 1] = t;
        }
    return 1;
}

/* Margin certificate: along n, every vertex of A outside IA lies more than
 * delta below the lowest vertex of IA, every vertex of B outside IB more than
 * delta above the highest of IB; delta = 2^-16 (|a0| + |b0|) of the features'
 * first vertices (local frames). */
static int certify(const Shape* A, const int* IA, int na, const Shape* B, const int* IB, int nb, V3 n) {
    const V3 la = mtvec(A->m, n), lb = mtvec(B->m, n);
    const float* a0 = A->verts + 4 * (long)IA[0];
    const float* b0 = B->verts + 4 * (long)IB[0];
    const float delta =
        CERT_REL * (sqrtf(a0[0] * a0[0] + a0[1] * a0[1] + a0[2] * a0[2]) + sqrtf(b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]));
    float in_a = 3.402823466e38f, out_a = -3.402823466e38f;
    for (int k = 0; k < A->nv; ++k) {
        const float* p = A->verts + 4 * (long)k;
        const float d = p[0] * la.x + p[1] * la.y + p[2] * la.z;
        int mem = 0;
        for (int j = 0; j < na; ++j) mem = mem || IA[j] == k;
        if (mem) in_a = fminf(in_a, d);
        else out_a = fmaxf(out_a, d);
    }
    float in_b = -3.402823466e38f, out_b = 3.402823466e38f;
    for (int k = 0; k < B->nv; ++k) {
        const float* p = B->verts + 4 * (long)k;
        const float d = p[0] * lb.x + p[1] * lb.y + p[2] * lb.z;
        int mem = 0;
        for (int j = 0; j < nb; ++j) mem = mem || IB[j] == k;
        if (mem) in_b = fmaxf(in_b, d);
        else out_b = fminf(out_b, d);
    }
    return out_a < in_a - delta && out_b > in_b + delta;
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
    int certified; /* closest features certified unique (computed on request) */
} Out;

enum { HIT = 1, CONV = 2, BUDGET = 4, ELIMIT = 8, DEGEN = 16, WARM = 32 };

/* Result of a feature (simplex points or an EPA face), canonical when every
 * point has a key; *ok receives whether the features were recognised with
 * interior closest points; cert: also run the margin certificate. */
static void feature_result(const Shape* A, const Shape* B, MP* w, int n, int hit, int cert, Out* o, V3 inner) {
    int IA[4], IB[4], na, nb;
    o->certified = 0;
    if (feature_sets(w, n, IA, &na, IB, &nb)) {
        V3 pa[3], pb[3];
        if (na <= 3 && nb <= 3) {
            for (int k = 0; k < na; ++k) pa[k] = vert_at(A, IA[k]);
            for (int k = 0; k < nb; ++k) pb[k] = vert_at(B, IB[k]);
            const Canon c = canon_feature(pa, na, pb, nb, hit);
            if (c.kind) {
                o->sep = c.sep;
                o->n = c.n;
                o->certified = c.ok && cert && certify(A, IA, na, B, IB, nb, c.n);
                return;
            }
        }
        /* degenerate features: the points ordered by key */
        sort_by_key(w, n);
    }
    if (!hit) {
        Simplex s;
        for (int i = 0; i < n; ++i) s.w[i] = w[i];
        s.n = n;
        int inside = 0;
        const V3 v = n == 1 ? w[0].p : cl_simplex(&s, &inside);
        const float l = sqrtf(vdot(v, v));
        o->sep = l;
        o->n = l > 0.0f ? vscl(v, -1.0f / l) : mk(1.0f, 0.0f, 0.0f);
        return;
    }
    V3 nrm = vcross(vsub(w[1].p, w[0].p), vsub(w[2].p, w[0].p));
    const float l2 = vdot(nrm, nrm);
    if (vdot(nrm, vsub(w[0].p, inner)) < 0.0f) nrm = vneg(nrm);
    const V3 nn = vscl(nrm, 1.0f / sqrtf(l2));
    o->sep = -vdot(nn, w[0].p);
    o->n = nn;
}

static void run_epa(const Shape* A, const Shape* B, const Simplex* s0, int max_iter, int cert, Out* o) {
    MP P[MAXV];
    int np = 0;
    for (int i = 0; i < s0->n; ++i) P[np++] = s0->w[i];
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
        o->certified = 0;
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
        o->certified = 0;
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
    o->epa_iters = it;
    if (nf == 0) {
        o->flags |= HIT | DEGEN;
        o->sep = 0.0f;
        o->certified = 0;
        return;
    }
    best = 0;
    for (int f = 1; f < nf; ++f)
        if (F[f].d < F[best].d) best = f;
    MP q[3] = {P[F[best].a], P[F[best].b], P[F[best].c]};
    o->flags |= HIT | (converged ? CONV : 0);
    if (!converged && it >= max_iter) o->flags |= BUDGET;
    if (F[best].d < 0.0f) o->flags |= DEGEN; /* polytope does not enclose the origin */
    feature_result(A, B, q, 3, 1, cert, o, inner);
}

/* Gap of A and B along the unit axis n (A -> B): the lowest point of B on n
 * less the highest point of A, from one support pass over both shapes;
 * positive proves the shapes disjoint. */
static float axis_gap(const Shape* A, const Shape* B, V3 n) {
    const MP m = supM(A, B, n);
    return -vdot(n, m.p);
}

/* Finishes a query from its GJK state: EPA for a hit, the feature result for a
 * separation. An uncertain hit is first tested for a separation along the
 * normal of its simplex's feature result: a positive gap along it is a
 * converged separation. An uncertain hit for which EPA finds no enclosing
 * polytope reports the uncertified separation. */
static void finish(const Shape* A, const Shape* B, Simplex* s, const GjkState* g, int max_iter, int cert, Out* o) {
    int converged = g->converged;
    if (g->hit && g->uncertain) {
        Out t = *o;
        MP w[4];
        for (int i = 0; i < s->n; ++i) w[i] = s->w[i];
        feature_result(A, B, w, s->n, 0, cert, &t, mk(0.0f, 0.0f, 0.0f));
        if (axis_gap(A, B, t.n) > 0.0f) {
            *o = t;
            o->flags |= CONV;
            return;
        }
    }
    if (g->hit) {
        run_epa(A, B, s, max_iter, cert, o);
        if (!(g->uncertain && (o->flags & DEGEN))) return;
        const int ei = o->epa_iters;
        o->flags = 0;
        o->epa_iters = ei;
        converged = 0;
    }
    o->flags |= converged ? CONV : BUDGET;
    feature_result(A, B, s->w, s->n, 0, cert, o, mk(0.0f, 0.0f, 0.0f));
}

static Out out_init(void) {
    Out o;
    o.flags = 0;
    o.iters = 0;
    o.epa_iters = 0;
    o.sep = 0.0f;
    o.n = mk(1.0f, 0.0f, 0.0f);
    o.certified = 0;
    return o;
}

/* Cold query: GJK from the support point along init (the centre difference). */
static Out query_cold(const Shape* A, const Shape* B, V3 init, int max_iter) {
    Out o = out_init();
    if (vdot(init, init) == 0.0f) init = mk(1.0f, 0.0f, 0.0f);
    Sup S;
    S.A = A;
    S.B = B;
    S.small = 0;
    Simplex s;
    s.w[0] = supM(A, B, init);
    s.n = 1;
    V3 v = s.w[0].p;
    GjkState g = {0, 0, 0, 0};
    gjk_loop(&S, &s, &v, max_iter, &g);
    o.iters = g.it;
    finish(A, B, &s, &g, max_iter, 0, &o);
    return o;
}

/* Warm query from the predicted direction u (A -> B): seed from the top-K
 * supports (one pass over both hulls, counted as one iteration), the small
 * problem from the seed support along u, then GJK on the full hulls within
 * budget iterations, then the certificate (one more pass, counted). */
static Out query_warm(const Shape* A, const Shape* B, V3 u, int budget, int* small_it) {
    Out o = out_init();
    *small_it = 0;
    if (vdot(u, u) == 0.0f) u = mk(1.0f, 0.0f, 0.0f);
    Sup S;
    S.A = A;
    S.B = B;
    S.small = 1;
    S.na = top_k(A, u, S.ia, S.pa);
    S.nb = top_k(B, vneg(u), S.ib, S.pb);
    Simplex s;
    s.w[0] = sup_eval(&S, u);
    s.n = 1;
    V3 v = s.w[0].p;
    GjkState g = {0, 0, 0, 0};
    gjk_loop(&S, &s, &v, SMALL_MAX_ITER, &g);
    *small_it = g.it;
    GjkState gf = {0, 0, 0, 0};
    if (g.hit && !g.uncertain) {
        gf.hit = 1;
    } else {
        S.small = 0;
        gjk_loop(&S, &s, &v, budget, &gf);
    }
    o.iters = 1 + gf.it + 1;
    finish(A, B, &s, &gf, budget, 1, &o);
    return o;
}

static void make_pair_shapes(const float* pos4, const float* quat4, const int* vstart, const int* nv,
                             const float* verts4, int ia, int ib, Shape* A, Shape* B, V3* dc) {
    const float* ca = pos4 + 4 * (long)ia;
    const float* cb = pos4 + 4 * (long)ib;
    A->verts = verts4 + 4 * (long)vstart[ia];
    A->nv = nv[ia];
    A->radius = ca[3];
    qmat(quat4 + 4 * (long)ia, A->m);
    A->off = mk(0.0f, 0.0f, 0.0f);
    B->verts = verts4 + 4 * (long)vstart[ib];
    B->nv = nv[ib];
    B->radius = cb[3];
    qmat(quat4 + 4 * (long)ib, B->m);
    *dc = mk(cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]);
    B->off = *dc;
}

/* Per pair (a, b): same outputs and warm/cold policy as the GPU entry point.
 * small_iters (nullable) receives the small-problem iterations of warm queries. */
void or_gjk_ex(int npairs, const int* pairs, const float* pos4, const float* quat4, const int* vstart, const int* nv,
               const float* verts4, const float* warm, int warm_budget, int max_iter, float near_margin, float* sep,
               float* normal3, int* flags, int* iters, int* small_iters, int nthreads) {
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (int p = 0; p < npairs; ++p) {
        const int ia = pairs[2 * (long)p], ib = pairs[2 * (long)p + 1];
        Shape A, B;
        V3 dc;
        make_pair_shapes(pos4, quat4, vstart, nv, verts4, ia, ib, &A, &B, &dc);
        Out o;
        int spent = 0, done = 0, sit = 0;
        if (warm && A.nv > 0 && B.nv > 0) {
            const V3 wd = mk(warm[3 * (long)p], warm[3 * (long)p + 1], warm[3 * (long)p + 2]);
            o = query_warm(&A, &B, wd, warm_budget < max_iter ? warm_budget : max_iter, &sit);
            spent = o.iters + o.epa_iters;
            done = (o.flags & CONV) && !(o.flags & (DEGEN | ELIMIT | BUDGET)) && o.certified &&
                   fabsf(o.sep) > near_margin;
            if (done) o.flags |= WARM;
        }
        if (!done) {
            o = query_cold(&A, &B, dc, max_iter);
            spent += o.iters + o.epa_iters;
        }
        sep[p] = o.sep;
        normal3[3 * (long)p] = o.n.x;
        normal3[3 * (long)p + 1] = o.n.y;
        normal3[3 * (long)p + 2] = o.n.z;
        flags[p] = o.flags;
        iters[p] = spent;
        if (small_iters) small_iters[p] = sit;
    }
}

void or_gjk(int npairs, const int* pairs, const float* pos4, const float* quat4, const int* vstart, const int* nv,
            const float* verts4, const float* warm, int warm_budget, int max_iter, float near_margin, float* sep,
            float* normal3, int* flags, int* iters, int nthreads) {
    or_gjk_ex(npairs, pairs, pos4, quat4, vstart, nv, verts4, warm, warm_budget, max_iter, near_margin, sep, normal3,
              flags, iters, NULL, nthreads);
}
