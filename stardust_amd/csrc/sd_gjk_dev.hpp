// Device-side convex-hull geometry: quaternion rotation, support mapping, the
// GJK distance phase with Voronoi-region simplex reduction, the warm seed, and
// the canonical result of a pair of closest features (EPA, run by 16 lanes per
// pair, is in sd_epa_wave.hpp).
//
// Replaces the boolean GJK of the reference sandbox (sandbox/testGJK.cu:52-315:
// supportPointSoupNaive, computeMinkowskiDifference, expandSimplex3/4, gjkCUDA),
// which never terminates on a hit and reports no distance (quirk Q16).
//
// Conventions: M = A - B is evaluated in A's frame (A's centre at the origin,
// B's centre at d = c_B - c_A). support_M(u) = s_A(u) - s_B(-u). The result
// normal n points from A to B. sep > 0 is the separation distance, sep < 0
// the penetration depth. All arithmetic is FP32; every expression is written
// out in an explicit order (the library is compiled without contraction) so a
// CPU restatement reproduces the bits.
//
// Results depend on the closest features only. GJK ends on an exact exit (the
// support point along -v is already a simplex vertex), a stall (adding it
// does not decrease |v| in FP32), a hit, or the budget. The vertex sets IA, IB
// of the final simplex (or EPA face), sorted, select the formula: point-point,
// point-segment, or the plane distance of a vertex-face, face-vertex or
// edge-edge pair. Two queries ending on the same features report the same
// bits whatever their start, which is what lets a warm start reproduce the
// cold result.
#pragma once

#include "sd_internal.hpp"

namespace sd {

constexpr int GJK_MAX_VERTS_PER_HULL = 1024;  // key = ia * 1024 + ib
constexpr float GJK_TOUCH_EPS2 = 1e-14f;      // |v|^2 below this: origin on the boundary
constexpr float GJK_REL_EPS = 1e-6f;          // sphere shapes: v.v - v.w <= eps v.v ends the query
constexpr float GJK_STALL_EPS = 1e-4f;         // gap accepted when the distance stops decreasing
constexpr float EPA_REL_EPS = 1e-6f;          // EPA: support gap <= eps * depth + 1e-9 -> converged
constexpr float CERT_REL = 1.52587890625e-05f;  // 2^-16: certificate margin relative to the vertex radii
constexpr float CERT_TAU = 9.765625e-04f;       // 2^-10: closest points this far inside their features
constexpr int WARM_K = 8;                     // warm seed: vertices per hull furthest along the prediction
constexpr int SMALL_MAX_ITER = 16;            // iterations of the seed problem
constexpr int EPA_MAX_V = 48;
constexpr int EPA_MAX_F = 96;
constexpr int EPA_MAX_E = 48;

enum {
    GJK_HIT = 1,          // shapes intersect
    GJK_CONVERGED = 2,    // result met the tolerance
    GJK_BUDGET = 4,       // iteration budget exhausted (or a stall left a gap)
    GJK_EPA_LIMIT = 8,    // EPA ran out of vertex/face capacity (best face returned)
    GJK_DEGENERATE = 16,  // EPA could not build a tetrahedron (depth 0 reported)
    GJK_WARM_USED = 32,   // result comes from the warm-started query
};

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 scl(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 sel3(bool c, V3 a, V3 b) { return v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }

// Row-major rotation of a unit quaternion stored as (w, x, y, z) in (.x, .y, .z, .w).
struct Rot {
    float m[9];
};
__device__ __forceinline__ Rot quat_rot(float4 q) {
    const float w = q.x, x = q.y, y = q.z, z = q.w;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    Rot r;
    r.m[0] = 1.0f - 2.0f * (yy + zz);
    r.m[1] = 2.0f * (xy - wz);
    r.m[2] = 2.0f * (xz + wy);
    r.m[3] = 2.0f * (xy + wz);
    r.m[4] = 1.0f - 2.0f * (xx + zz);
    r.m[5] = 2.0f * (yz - wx);
    r.m[6] = 2.0f * (xz - wy);
    r.m[7] = 2.0f * (yz + wx);
    r.m[8] = 1.0f - 2.0f * (xx + yy);
    return r;
}
__device__ __forceinline__ V3 rot_mul(const Rot& r, V3 v) {
    return v3(r.m[0] * v.x + r.m[1] * v.y + r.m[2] * v.z, r.m[3] * v.x + r.m[4] * v.y + r.m[5] * v.z,
              r.m[6] * v.x + r.m[7] * v.y + r.m[8] * v.z);
}
__device__ __forceinline__ V3 rot_tmul(const Rot& r, V3 v) {
    return v3(r.m[0] * v.x + r.m[3] * v.y + r.m[6] * v.z, r.m[1] * v.x + r.m[4] * v.y + r.m[7] * v.z,
              r.m[2] * v.x + r.m[5] * v.y + r.m[8] * v.z);
}

// One convex shape placed in the query frame: a vertex set (nv > 0) or a sphere
// (nv == 0, radius > 0).
struct Shape {
    const float4* verts;  // local-frame vertices
    int nv;
    float radius;
    Rot R;
    V3 off;  // centre in the query frame
};

// Vertex k of s in the query frame.
__device__ __forceinline__ V3 vert_at(const Shape& s, int k) {
    const float4 p = s.verts[k];
    return add(rot_mul(s.R, v3(p.x, p.y, p.z)), s.off);
}

// Index of the vertex with the largest projection on u (first one on ties),
// and that vertex. Vertices are read SUPPORT_BATCH at a time so that their
// loads are in flight together; indices past the last vertex read the last
// vertex again, which never wins (ties keep the first index).
constexpr int SUPPORT_BATCH = 16;
constexpr int SCAN_BATCH = 8;  // the seed selection and the certificate scan
__device__ __forceinline__ int support_index(const Shape& s, V3 u, float4& bp) {
    const V3 l = rot_tmul(s.R, u);
    int best = 0;
    float bs = -3.402823466e38f;
    bp = s.verts[0];
    const int last = s.nv - 1;
    for (int k = 0; k < s.nv; k += SUPPORT_BATCH) {
        float4 p[SUPPORT_BATCH];
#pragma unroll
        for (int j = 0; j < SUPPORT_BATCH; ++j) p[j] = s.verts[min(k + j, last)];
#pragma unroll
        for (int j = 0; j < SUPPORT_BATCH; ++j) {
            const float d = p[j].x * l.x + p[j].y * l.y + p[j].z * l.z;
            if (d > bs) {
                bs = d;
                best = k + j;
                bp = p[j];
            }
        }
    }
    return best;
}

__device__ __forceinline__ V3 support_point(const Shape& s, V3 u, int& idx) {
    if (s.nv == 0) {
        idx = -1;
        const float uu = dot(u, u);
        if (uu == 0.0f) return s.off;
        const float k = s.radius / sqrtf(uu);
        return add(s.off, scl(u, k));
    }
    float4 p;
    idx = support_index(s, u, p);
    return add(rot_mul(s.R, v3(p.x, p.y, p.z)), s.off);
}

struct MPoint {
    V3 p;
    int key;  // ia * 1024 + ib, or -1 when a sphere contributed
};

__device__ __forceinline__ MPoint support_M(const Shape& A, const Shape& B, V3 u) {
    int ia, ib;
    const V3 pa = support_point(A, u, ia);
    const V3 pb = support_point(B, neg(u), ib);
    MPoint m;
    m.p = sub(pa, pb);
    m.key = (ia < 0 || ib < 0) ? -1 : ia * GJK_MAX_VERTS_PER_HULL + ib;
    return m;
}

// ---------------------------------------------------------------- warm seed

// The WARM_K vertices of a hull furthest along u: the largest projections,
// lower index on ties (slots of a hull with fewer vertices repeat vertex
// idx[0]), with their query-frame points.
struct SeedSet {
    V3 p[WARM_K];
    int idx[WARM_K];
};

__device__ __forceinline__ void seed_select(const Shape& s, V3 u, SeedSet& out) {
    const V3 l = rot_tmul(s.R, u);
    float tv[WARM_K];
    int ti[WARM_K];
#pragma unroll
    for (int j = 0; j < WARM_K; ++j) {
        tv[j] = -3.402823466e38f;
        ti[j] = 0x7fffffff;
    }
    float mval = -3.402823466e38f;  // the smallest kept entry: (value, then larger index)
    int mslot = 0, filled = 0;
    const int last = s.nv - 1;
    for (int k = 0; k < s.nv; k += SCAN_BATCH) {
        float4 p[SCAN_BATCH];
#pragma unroll
        for (int j = 0; j < SCAN_BATCH; ++j) p[j] = s.verts[min(k + j, last)];
#pragma unroll
        for (int j = 0; j < SCAN_BATCH; ++j) {
            const float d = p[j].x * l.x + p[j].y * l.y + p[j].z * l.z;
            // repeated reads past the last vertex never enter (k + j > last)
            if (k + j <= last && (filled < WARM_K || d > mval)) {
                ++filled;
#pragma unroll
                for (int q = 0; q < WARM_K; ++q) {
                    const bool w = q == mslot;
                    tv[q] = w ? d : tv[q];
                    ti[q] = w ? k + j : ti[q];
                }
                mval = tv[0];
                int mi = ti[0];
                mslot = 0;
#pragma unroll
                for (int q = 1; q < WARM_K; ++q) {
                    const bool lower = (tv[q] < mval) | ((tv[q] == mval) & (ti[q] > mi));
                    mval = lower ? tv[q] : mval;
                    mi = lower ? ti[q] : mi;
                    mslot = lower ? q : mslot;
                }
            }
        }
    }
    // unfilled slots (fewer than WARM_K vertices) repeat slot 0
#pragma unroll
    for (int q = 0; q < WARM_K; ++q) out.idx[q] = ti[q] == 0x7fffffff ? ti[0] : ti[q];
#pragma unroll
    for (int q = 0; q < WARM_K; ++q) out.p[q] = vert_at(s, out.idx[q]);
}

// Support of M over the two seed sets: larger projection, lower vertex index
// on ties.
__device__ __forceinline__ MPoint seed_support(const SeedSet& a, const SeedSet& b, V3 u) {
    float ba = dot(a.p[0], u), bb = -dot(b.p[0], u);
    int ia = a.idx[0], ib = b.idx[0];
    V3 pa = a.p[0], pb = b.p[0];
#pragma unroll
    for (int j = 1; j < WARM_K; ++j) {
        const float d = dot(a.p[j], u);
        const bool t = (d > ba) | ((d == ba) & (a.idx[j] < ia));  // bitwise: no branches
        ba = t ? d : ba;
        ia = t ? a.idx[j] : ia;
        pa = sel3(t, a.p[j], pa);
        const float e = -dot(b.p[j], u);
        const bool r = (e > bb) | ((e == bb) & (b.idx[j] < ib));
        bb = r ? e : bb;
        ib = r ? b.idx[j] : ib;
        pb = sel3(r, b.p[j], pb);
    }
    MPoint m;
    m.p = sub(pa, pb);
    m.key = ia * GJK_MAX_VERTS_PER_HULL + ib;
    return m;
}

// ---------------------------------------------------------------- simplex

struct Simplex {
    MPoint w[4];
    int n;
};

// Closest point of segment (a, b) to the origin; keeps the feature's vertices.
__device__ __forceinline__ V3 closest_segment(Simplex& s) {
    const V3 a = s.w[0].p, b = s.w[1].p;
    const V3 ab = sub(b, a);
    const float t = -dot(a, ab);
    if (t <= 0.0f) {
        s.n = 1;
        return a;
    }
    const float den = dot(ab, ab);
    if (t >= den) {
        s.w[0] = s.w[1];
        s.n = 1;
        return b;
    }
    return add(a, scl(ab, t / den));
}

// Closest point of triangle (a, b, c) to the origin (Voronoi-region tests).
__device__ __forceinline__ V3 closest_triangle(Simplex& s) {
    const MPoint A = s.w[0], B = s.w[1], Cc = s.w[2];
    const V3 a = A.p, b = B.p, c = Cc.p;
    const V3 ab = sub(b, a), ac = sub(c, a);
    const V3 ap = neg(a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        s.n = 1;
        return a;
    }
    const V3 bp = neg(b);
    const float d3v = dot(ab, bp), d4 = dot(ac, bp);
    if (d3v >= 0.0f && d4 <= d3v) {
        s.w[0] = B;
        s.n = 1;
        return b;
    }
    const float vc = d1 * d4 - d3v * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3v <= 0.0f) {
        const float t = d1 / (d1 - d3v);
        s.n = 2;
        return add(a, scl(ab, t));
    }
    const V3 cp = neg(c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        s.w[0] = Cc;
        s.n = 1;
        return c;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        s.w[1] = Cc;
        s.n = 2;
        return add(a, scl(ac, t));
    }
    const float va = d3v * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3v) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3v) / ((d4 - d3v) + (d5 - d6));
        s.w[0] = B;
        s.w[1] = Cc;
        s.n = 2;
        return add(b, scl(sub(c, b), t));
    }
    const float den = 1.0f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    return add(add(a, scl(ab, v)), scl(ac, w));
}

// True if the origin and d lie strictly on opposite sides of plane (a, b, c).
__device__ __forceinline__ bool origin_outside(V3 a, V3 b, V3 c, V3 d) {
    const V3 nrm = cross(sub(b, a), sub(c, a));
    const float so = -dot(a, nrm);
    const float sd = dot(sub(d, a), nrm);
    return so * sd < 0.0f;
}

// Closest point of tetrahedron w[0..3] to the origin; *inside when it contains it.
// A near-flat tetrahedron (|6 V| <= 1e-5 |ab| |ac| |ad|) has unreliable side
// tests; it is never reported as containing the origin, and all four faces are
// searched instead. Faces in search order: (0,1,2) opposite 3, (0,2,3)
// opposite 1, (0,3,1) opposite 2, (1,3,2) opposite 0. Every index is a
// compile-time constant so the simplex stays in registers.
__device__ __forceinline__ float tetra_face(const MPoint& a, const MPoint& b, const MPoint& c, const MPoint& opp,
                                            bool flat) {
    if (!flat && !origin_outside(a.p, b.p, c.p, opp.p)) return -1.0f;
    Simplex t;
    t.w[0] = a;
    t.w[1] = b;
    t.w[2] = c;
    t.n = 3;
    const V3 q = closest_triangle(t);
    return dot(q, q);
}

// The best face is found first (squared distances only); its reduced feature
// and closest point are then computed once more (same arithmetic).
__device__ __forceinline__ V3 closest_tetra(Simplex& s, bool* inside) {
    const MPoint P0 = s.w[0], P1 = s.w[1], P2 = s.w[2], P3 = s.w[3];
    const V3 ab = sub(P1.p, P0.p), ac = sub(P2.p, P0.p), ad = sub(P3.p, P0.p);
    const float vol = dot(ad, cross(ab, ac));
    const float lim = 1e-5f * (sqrtf(dot(ab, ab)) * sqrtf(dot(ac, ac)) * sqrtf(dot(ad, ad)));
    const bool flat = !(fabsf(vol) > lim);
    float best = 3.402823466e38f;
    int bi = -1;
    float dq;
    dq = tetra_face(P0, P1, P2, P3, flat);
    if (dq >= 0.0f && dq < best) {
        best = dq;
        bi = 0;
    }
    dq = tetra_face(P0, P2, P3, P1, flat);
    if (dq >= 0.0f && dq < best) {
        best = dq;
        bi = 1;
    }
    dq = tetra_face(P0, P3, P1, P2, flat);
    if (dq >= 0.0f && dq < best) {
        best = dq;
        bi = 2;
    }
    dq = tetra_face(P1, P3, P2, P0, flat);
    if (dq >= 0.0f && dq < best) {
        best = dq;
        bi = 3;
    }
    *inside = bi < 0;
    if (bi < 0) return v3(0.0f, 0.0f, 0.0f);
    s.w[0] = bi == 3 ? P1 : P0;
    s.w[1] = bi == 0 ? P1 : (bi == 1 ? P2 : P3);
    s.w[2] = bi == 0 ? P2 : (bi == 1 ? P3 : (bi == 2 ? P1 : P2));
    s.n = 3;
    return closest_triangle(s);
}

// Reduces s to the feature closest to the origin and returns that point.
__device__ __forceinline__ V3 simplex_closest(Simplex& s, bool* inside) {
    *inside = false;
    if (s.n == 1) return s.w[0].p;
    if (s.n == 2) return closest_segment(s);
    if (s.n == 3) return closest_triangle(s);
    return closest_tetra(s, inside);
}

// The same, as a call: for the rare recomputations of kernels whose code
// should not hold a second inlined copy of the simplex code.
__device__ __noinline__ V3 simplex_closest_call(Simplex& s, bool* inside) { return simplex_closest(s, inside); }

// Every slot is assigned unconditionally (a select), so no store has a
// run-time address and the simplex stays in registers.
__device__ __forceinline__ MPoint sel(bool c, const MPoint& a, const MPoint& b) {
    MPoint r;
    r.p = sel3(c, a.p, b.p);
    r.key = c ? a.key : b.key;
    return r;
}
__device__ __forceinline__ void simplex_push(Simplex& s, const MPoint& w) {
    s.w[0] = sel(s.n == 0, w, s.w[0]);
    s.w[1] = sel(s.n == 1, w, s.w[1]);
    s.w[2] = sel(s.n == 2, w, s.w[2]);
    s.w[3] = sel(s.n >= 3, w, s.w[3]);
    ++s.n;
}

__device__ __forceinline__ bool simplex_has(const Simplex& s, int key) {
    bool dup = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) dup = dup | ((i < s.n) & (key >= 0) & (s.w[i].key == key));  // no branches
    return dup;
}

// ---------------------------------------------------------------- GJK loop

struct GjkState {
    bool hit, converged, uncertain;
    int it;
};

__device__ __forceinline__ GjkState gjk_state() { return GjkState{false, false, false, 0}; }

// GJK iterations from simplex s with closest point v, at most max_iter
// support evaluations (added to g.it). SEED: support over the seed sets
// instead of the full shapes. Exits:
//  - exact: the support point is already a simplex vertex; in exact
//    arithmetic v.w >= v.v then, so any gap is rounding. v.w > 0 makes v a
//    separating axis (converged); v.w <= 0 leaves the verdict to the rounding
//    of v's direction (an uncertain hit, settled by the caller with
//    axis_separates or EPA); a sphere shape (keyless) ends on the relative
//    gap instead;
//  - inside: the simplex encloses the origin (hit);
//  - stall: the new simplex is not strictly closer in FP32; the previous
//    feature is final if the duality gap is within 1e-4 v.v (converged), else
//    an uncertain hit when v.w <= 0;
//  - touching: |v|^2 <= 1e-14 (hit).
template <bool SEED>
__device__ __forceinline__ void gjk_loop(const Shape& A, const Shape& B, const SeedSet* sa, const SeedSet* sb,
                                         Simplex& s, V3& v, int max_iter, GjkState& g) {
    int it = 0;
    while (it < max_iter) {
        const float vv = dot(v, v);
        if (vv <= GJK_TOUCH_EPS2) {
            g.hit = true;
            break;
        }
        const MPoint w = SEED ? seed_support(*sa, *sb, neg(v)) : support_M(A, B, neg(v));
        ++it;
        const float vw = dot(v, w.p);
        const float gap = vv - vw;
        if (simplex_has(s, w.key) || (w.key < 0 && gap <= GJK_REL_EPS * vv)) {
            if (vw > 0.0f) {
                g.converged = true;
            } else {
                g.hit = true;
                g.uncertain = true;
            }
            break;
        }
        const Simplex sp = s;
        simplex_push(s, w);
        bool inside = false;
        const V3 vn = simplex_closest(s, &inside);
        if (inside) {
            g.hit = true;
            break;
        }
        if (!(dot(vn, vn) < vv)) {
            s = sp;
            if (gap <= GJK_STALL_EPS * vv) {
                g.converged = true;
            } else if (vw <= 0.0f) {
                g.hit = true;
                g.uncertain = true;
            }
            break;
        }
        v = vn;
    }
    g.it += it;
}

// ---------------------------------------------------------------- closest features

// Sorted distinct vertex indices of A (key / 1024) and B (key % 1024) of the
// first n points (n <= 4); false if a point has no key (sphere shape).
struct Features {
    int ia[4], ib[4];
    int na, nb;
};

__device__ __forceinline__ void sort4(int* x, int n) {
    // insertion order network over the used prefix; unused entries hold INT_MAX
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = i < n ? x[i] : 0x7fffffff;
#define SD_CX(i, j)                     \
    {                                   \
        const int lo = min(x[i], x[j]); \
        const int hi = max(x[i], x[j]); \
        x[i] = lo;                      \
        x[j] = hi;                      \
    }
    SD_CX(0, 1)
    SD_CX(2, 3)
    SD_CX(0, 2)
    SD_CX(1, 3)
    SD_CX(1, 2)
#undef SD_CX
}

__device__ __forceinline__ bool feature_sets(const MPoint* w, int n, Features& f) {
    f.na = 0;
    f.nb = 0;
    bool keyed = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f.ia[i] = 0x7fffffff;
        f.ib[i] = 0x7fffffff;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i < n) {
            keyed = keyed & (w[i].key >= 0);
            const int a = w[i].key >> 10, b = w[i].key & (GJK_MAX_VERTS_PER_HULL - 1);
            bool da = false, db = false;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                da = da | ((j < f.na) & (f.ia[j] == a));
                db = db | ((j < f.nb) & (f.ib[j] == b));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                f.ia[j] = (!da && j == f.na) ? a : f.ia[j];
                f.ib[j] = (!db && j == f.nb) ? b : f.ib[j];
            }
            f.na += da ? 0 : 1;
            f.nb += db ? 0 : 1;
        }
    }
    sort4(f.ia, f.na);
    sort4(f.ib, f.nb);
    return keyed;
}

struct Canon {
    float sep;
    V3 n;
    bool ok;   // a recognised feature pair with its closest points inside it
    int kind;  // 0: degenerate, else na * 4 + nb
};

// Barycentric coordinates of p in triangle (a, b, c) all >= CERT_TAU.
__device__ __forceinline__ bool inside_tri(V3 p, V3 a, V3 b, V3 c) {
    const V3 e1 = sub(b, a), e2 = sub(c, a), r = sub(p, a);
    const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
    const float d20 = dot(r, e1), d21 = dot(r, e2);
    const float den = d00 * d11 - d01 * d01;
    if (!(den > 0.0f)) return false;
    const float v = (d11 * d20 - d01 * d21) / den;
    const float w = (d00 * d21 - d01 * d20) / den;
    const float u = 1.0f - v - w;
    return u >= CERT_TAU && v >= CERT_TAU && w >= CERT_TAU;
}

// Distance between conv(pa[0..na)) on A and conv(pb[0..nb)) on B, vertices in
// index order; hit: the shapes intersect and a plane distance is a depth
// (reported negative). n points from A to B.
__device__ __forceinline__ Canon canon_feature(const V3* pa, int na, const V3* pb, int nb, bool hit) {
    Canon c;
    c.ok = false;
    c.kind = 0;
    c.sep = 0.0f;
    c.n = v3(1.0f, 0.0f, 0.0f);
    if (!hit && na == 1 && nb == 1) {
        const V3 d = sub(pb[0], pa[0]);
        const float l = sqrtf(dot(d, d));
        c.sep = l;
        c.n = l > 0.0f ? scl(d, 1.0f / l) : v3(1.0f, 0.0f, 0.0f);
        c.ok = l > 0.0f;
        c.kind = 5;
        return c;
    }
    if (!hit && ((na == 1 && nb == 2) || (na == 2 && nb == 1))) {
        const bool a_point = na == 1;
        const V3 p = a_point ? pa[0] : pb[0];
        const V3 s0 = a_point ? pb[0] : pa[0];
        const V3 s1 = a_point ? pb[1] : pa[1];
        const V3 e = sub(s1, s0);
        const float ee = dot(e, e);
        if (!(ee > 0.0f)) return c;
        const float t = dot(sub(p, s0), e) / ee;
        const V3 q = add(s0, scl(e, t));
        const V3 d = a_point ? sub(q, p) : sub(p, q);
        const float l = sqrtf(dot(d, d));
        c.sep = l;
        c.n = l > 0.0f ? scl(d, 1.0f / l) : v3(1.0f, 0.0f, 0.0f);
        c.ok = l > 0.0f && t >= CERT_TAU && t <= 1.0f - CERT_TAU;
        c.kind = na * 4 + nb;
        return c;
    }
    V3 e1, e2;
    if (na == 1 && nb == 3) {
        e1 = sub(pb[1], pb[0]);
        e2 = sub(pb[2], pb[0]);
    } else if (na == 3 && nb == 1) {
        e1 = sub(pa[1], pa[0]);
        e2 = sub(pa[2], pa[0]);
    } else if (na == 2 && nb == 2) {
        e1 = sub(pa[1], pa[0]);
        e2 = sub(pb[1], pb[0]);
    } else {
        return c;
    }
    const V3 cr = cross(e1, e2);
    const float l = sqrtf(dot(cr, cr));
    if (!(l > 0.0f)) return c;
    V3 m = scl(cr, 1.0f / l);
    float h = dot(sub(pb[0], pa[0]), m);
    if (hit ? h > 0.0f : h < 0.0f) {
        m = neg(m);
        h = -h;
    }
    c.sep = h;
    c.n = m;
    c.kind = na * 4 + nb;
    if (na == 1) {
        c.ok = inside_tri(add(pa[0], scl(m, h)), pb[0], pb[1], pb[2]);
    } else if (nb == 1) {
        c.ok = inside_tri(sub(pb[0], scl(m, h)), pa[0], pa[1], pa[2]);
    } else {
        const V3 r = sub(pa[0], pb[0]);
        const float a = dot(e1, e1), b = dot(e1, e2), e = dot(e2, e2);
        const float cc = dot(e1, r), f = dot(e2, r);
        const float den = a * e - b * b;
        if (den > 0.0f) {
            const float s = (b * f - cc * e) / den;
            const float t = (a * f - b * cc) / den;
            c.ok = s >= CERT_TAU && s <= 1.0f - CERT_TAU && t >= CERT_TAU && t <= 1.0f - CERT_TAU;
        }
    }
    return c;
}

// Margin of the features along n, per thread over every vertex: every vertex
// of A outside IA lies more than delta below the lowest vertex of IA, every
// vertex of B outside IB more than delta above the highest of IB; delta =
// 2^-16 (|a0| + |b0|) of the features' first vertices (local frames).
__device__ __forceinline__ float cert_delta(const Shape& A, const Shape& B, const Features& f) {
    const float4 a0 = A.verts[f.ia[0]], b0 = B.verts[f.ib[0]];
    return CERT_REL * (sqrtf(a0.x * a0.x + a0.y * a0.y + a0.z * a0.z) + sqrtf(b0.x * b0.x + b0.y * b0.y + b0.z * b0.z));
}

__device__ __forceinline__ bool in_set(int k, const int* set) {
    return (k == set[0]) | (k == set[1]) | (k == set[2]) | (k == set[3]);  // no branches
}

// Lowest projection of the set's vertices and highest of the others along l
// (local frame); sign -1 turns both into highest / lowest.
__device__ __forceinline__ void margin_scan(const Shape& s, V3 l, const int* set, float sg, float& in_v,
                                            float& out_v) {
    in_v = 3.402823466e38f;
    out_v = -3.402823466e38f;
    const int last = s.nv - 1;
    for (int k = 0; k < s.nv; k += SCAN_BATCH) {
        float4 p[SCAN_BATCH];
#pragma unroll
        for (int j = 0; j < SCAN_BATCH; ++j) p[j] = s.verts[min(k + j, last)];
#pragma unroll
        for (int j = 0; j < SCAN_BATCH; ++j) {
            if (k + j <= last) {
                const float d = sg * (p[j].x * l.x + p[j].y * l.y + p[j].z * l.z);
                if (in_set(k + j, set)) in_v = fminf(in_v, d);
                else out_v = fmaxf(out_v, d);
            }
        }
    }
}

__device__ __forceinline__ bool certify(const Shape& A, const Shape& B, const Features& f, V3 n) {
    const float delta = cert_delta(A, B, f);
    float in_a, out_a, in_b, out_b;
    margin_scan(A, rot_tmul(A.R, n), f.ia, 1.0f, in_a, out_a);
    margin_scan(B, rot_tmul(B.R, n), f.ib, -1.0f, in_b, out_b);
    // B's values are negated: out_b > in_b + delta  <=>  -out_b < -in_b - delta
    return out_a < in_a - delta && out_b < in_b - delta;
}

// ---------------------------------------------------------------- results

struct GjkOut {
    float sep;
    V3 n;
    int flags;
    int iters;
    int epa_iters;
    bool certified;
};

__device__ __forceinline__ GjkOut gjk_out_init() {
    GjkOut o;
    o.flags = 0;
    o.iters = 0;
    o.epa_iters = 0;
    o.sep = 0.0f;
    o.n = v3(1.0f, 0.0f, 0.0f);
    o.certified = false;
    return o;
}

// Stable sort of up to 4 points by key (insertion order on equal keys).
__device__ __forceinline__ void cmpx_key(MPoint& a, MPoint& b) {
    if (a.key > b.key) {
        const MPoint t = a;
        a = b;
        b = t;
    }
}

// Whether the unit axis n (A -> B) separates A and B: the lowest point of B
// on n above the highest point of A, from one support pass over both shapes.
__device__ __forceinline__ bool axis_separates(const Shape& A, const Shape& B, V3 n) {
    const MPoint m = support_M(A, B, n);
    return -dot(n, m.p) > 0.0f;
}

// Result of a separated query from its final simplex: the canonical feature
// distance when the features are recognised, else the closest point of the
// simplex with its points ordered by key (keyless: as they are). CERT: also
// evaluate the margin certificate (o.certified). CALL: the fallback's simplex
// code through simplex_closest_call instead of inline.
template <bool CERT, bool CALL = false>
__device__ __forceinline__ void separated_result(const Shape& A, const Shape& B, Simplex s, GjkOut& o) {
    Features f;
    o.certified = false;
    const bool keyed = feature_sets(s.w, s.n, f);
    if (keyed && f.na <= 3 && f.nb <= 3) {
        V3 pa[3], pb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pa[k] = k < f.na ? vert_at(A, f.ia[k]) : v3(0.0f, 0.0f, 0.0f);
            pb[k] = k < f.nb ? vert_at(B, f.ib[k]) : v3(0.0f, 0.0f, 0.0f);
        }
        const Canon c = canon_feature(pa, f.na, pb, f.nb, false);
        if (c.kind != 0) {
            o.sep = c.sep;
            o.n = c.n;
            if (CERT) o.certified = c.ok && certify(A, B, f, c.n);
            return;
        }
    }
    if (keyed) {
        // degenerate features: the points ordered by key (a stable network)
        if (s.n >= 2) cmpx_key(s.w[0], s.w[1]);
        if (s.n >= 3) {
            cmpx_key(s.w[1], s.w[2]);
            cmpx_key(s.w[0], s.w[1]);
        }
    }
    bool inside = false;
    const V3 v = s.n == 1 ? s.w[0].p : (CALL ? simplex_closest_call(s, &inside) : simplex_closest(s, &inside));
    const float l = sqrtf(dot(v, v));
    o.sep = l;
    o.n = l > 0.0f ? scl(v, -1.0f / l) : v3(1.0f, 0.0f, 0.0f);
}

}  // namespace sd
