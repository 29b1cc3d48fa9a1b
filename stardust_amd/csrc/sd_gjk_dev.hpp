// Device-side convex-hull geometry: quaternion rotation, support mapping and
// the GJK distance phase with Voronoi-region simplex reduction (EPA, run by a
// whole wavefront per pair, is in sd_epa_wave.hpp).
//
// Replaces the boolean GJK of the reference sandbox (sandbox/testGJK.cu:52-315:
// supportPointSoupNaive, computeMinkowskiDifference, expandSimplex3/4, gjkCUDA),
// which never terminates on a hit and reports no distance (quirk Q16).
//
// Conventions: M = A - B is evaluated in A's frame (A's centre at the origin,
// B's centre at d = c_B - c_A). support_M(u) = s_A(u) - s_B(-u). The result
// normal n points from A to B: for separated shapes n = -v/|v| (v the closest
// point of M), for intersecting shapes n is the normal of EPA's closest face.
// sep > 0 is the separation distance, sep < 0 the penetration depth.
// Every floating-point expression is written out with an explicit order
// (no contraction) so a CPU restatement can reproduce the bits.
#pragma once

#include "sd_internal.hpp"

namespace sd {

constexpr int GJK_MAX_VERTS_PER_HULL = 1024;  // key = ia * 1024 + ib
constexpr double GJK_REL_EPS = 1e-6;          // v.v - v.w <= eps * v.v  -> converged
constexpr double GJK_TOUCH_EPS2 = 1e-14;      // |v|^2 below this: origin on the boundary
constexpr double GJK_STALL_EPS = 1e-4;        // gap accepted when the distance stops decreasing
constexpr float EPA_REL_EPS = 1e-6f;          // EPA: support gap <= eps * depth + 1e-9 -> converged
constexpr int EPA_MAX_V = 48;
constexpr int EPA_MAX_F = 96;
constexpr int EPA_MAX_E = 48;

enum {
    GJK_HIT = 1,          // shapes intersect
    GJK_CONVERGED = 2,    // result met the tolerance
    GJK_BUDGET = 4,       // iteration budget exhausted
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

// Index of the vertex with the largest projection on u (first one on ties),
// and that vertex. Vertices are read SUPPORT_BATCH at a time so that their
// loads are in flight together; indices past the last vertex read the last
// vertex again, which never wins (ties keep the first index).
constexpr int SUPPORT_BATCH = 16;
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

// ---------------------------------------------------------------- simplex
//
// The simplex of the GJK phase and every quantity derived from it (closest
// point, v.v, v.w, duality gap) are evaluated in FP64; support points are the
// FP32 differences the support mapping returns, stored exactly. The closest
// point of a feature far from the origin relative to its vertices then keeps
// its accuracy, and the duality gap can certify the separation.

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 d3(V3 a) { return D3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ V3 to_v3(D3 a) { return v3((float)a.x, (float)a.y, (float)a.z); }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 neg(D3 a) { return d3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ D3 scl(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Simplex vertex: an FP32 support point (kept in FP32, widened where used).
struct SPoint {
    V3 p;
    int key;
};
__device__ __forceinline__ SPoint spoint(const MPoint& m) { return SPoint{m.p, m.key}; }

struct Simplex {
    SPoint w[4];
    int n;
};

// Closest point of segment (a, b) to the origin; keeps the feature's vertices.
__device__ __forceinline__ D3 closest_segment(Simplex& s) {
    const D3 a = d3(s.w[0].p), b = d3(s.w[1].p);
    const D3 ab = sub(b, a);
    const double t = -dot(a, ab);
    if (t <= 0.0) {
        s.n = 1;
        return a;
    }
    const double den = dot(ab, ab);
    if (t >= den) {
        s.w[0] = s.w[1];
        s.n = 1;
        return b;
    }
    return add(a, scl(ab, t / den));
}

// Closest point of triangle (a, b, c) to the origin (Voronoi-region tests).
__device__ __forceinline__ D3 closest_triangle(Simplex& s) {
    const SPoint A = s.w[0], B = s.w[1], Cc = s.w[2];
    const D3 a = d3(A.p), b = d3(B.p), c = d3(Cc.p);
    const D3 ab = sub(b, a), ac = sub(c, a);
    const D3 ap = neg(a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        s.n = 1;
        return a;
    }
    const D3 bp = neg(b);
    const double d3v = dot(ab, bp), d4 = dot(ac, bp);
    if (d3v >= 0.0 && d4 <= d3v) {
        s.w[0] = B;
        s.n = 1;
        return b;
    }
    const double vc = d1 * d4 - d3v * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0) {
        const double t = d1 / (d1 - d3v);
        s.n = 2;
        return add(a, scl(ab, t));
    }
    const D3 cp = neg(c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        s.w[0] = Cc;
        s.n = 1;
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double t = d2 / (d2 - d6);
        s.w[1] = Cc;
        s.n = 2;
        return add(a, scl(ac, t));
    }
    const double va = d3v * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3v) >= 0.0 && (d5 - d6) >= 0.0) {
        const double t = (d4 - d3v) / ((d4 - d3v) + (d5 - d6));
        s.w[0] = B;
        s.w[1] = Cc;
        s.n = 2;
        return add(b, scl(sub(c, b), t));
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    return add(add(a, scl(ab, v)), scl(ac, w));
}

// True if the origin and d lie strictly on opposite sides of plane (a, b, c).
__device__ __forceinline__ bool origin_outside(D3 a, D3 b, D3 c, D3 d) {
    const D3 nrm = cross(sub(b, a), sub(c, a));
    const double so = -dot(a, nrm);
    const double sd = dot(sub(d, a), nrm);
    return so * sd < 0.0;
}

// Closest point of tetrahedron w[0..3] to the origin; *inside when it contains it.
// A near-flat tetrahedron (|6 V| <= 1e-5 |ab| |ac| |ad|) has unreliable side
// tests; it is never reported as containing the origin, and all four faces are
// searched instead. Faces in search order: (0,1,2) opposite 3, (0,2,3)
// opposite 1, (0,3,1) opposite 2, (1,3,2) opposite 0. Every index is a
// compile-time constant so the simplex stays in registers.
__device__ __forceinline__ double tetra_face(const SPoint& a, const SPoint& b, const SPoint& c, const SPoint& opp,
                                             bool flat) {
    if (!flat && !origin_outside(d3(a.p), d3(b.p), d3(c.p), d3(opp.p))) return -1.0;
    Simplex t;
    t.w[0] = a;
    t.w[1] = b;
    t.w[2] = c;
    t.n = 3;
    const D3 q = closest_triangle(t);
    return dot(q, q);
}

// The best face is found first (squared distances only); its reduced feature
// and closest point are then computed once more (same arithmetic).
__device__ __forceinline__ D3 closest_tetra(Simplex& s, bool* inside) {
    const SPoint P0 = s.w[0], P1 = s.w[1], P2 = s.w[2], P3 = s.w[3];
    const D3 p0 = d3(P0.p);
    const D3 ab = sub(d3(P1.p), p0), ac = sub(d3(P2.p), p0), ad = sub(d3(P3.p), p0);
    const double vol = dot(ad, cross(ab, ac));
    const double lim = 1e-5 * (sqrt(dot(ab, ab)) * sqrt(dot(ac, ac)) * sqrt(dot(ad, ad)));
    const bool flat = !(fabs(vol) > lim);
    double best = 1.79769313486231e308;
    int bi = -1;
    double dq;
    dq = tetra_face(P0, P1, P2, P3, flat);
    if (dq >= 0.0 && dq < best) {
        best = dq;
        bi = 0;
    }
    dq = tetra_face(P0, P2, P3, P1, flat);
    if (dq >= 0.0 && dq < best) {
        best = dq;
        bi = 1;
    }
    dq = tetra_face(P0, P3, P1, P2, flat);
    if (dq >= 0.0 && dq < best) {
        best = dq;
        bi = 2;
    }
    dq = tetra_face(P1, P3, P2, P0, flat);
    if (dq >= 0.0 && dq < best) {
        best = dq;
        bi = 3;
    }
    *inside = bi < 0;
    if (bi < 0) return d3(0.0, 0.0, 0.0);
    s.w[0] = bi == 3 ? P1 : P0;
    s.w[1] = bi == 0 ? P1 : (bi == 1 ? P2 : P3);
    s.w[2] = bi == 0 ? P2 : (bi == 1 ? P3 : (bi == 2 ? P1 : P2));
    s.n = 3;
    return closest_triangle(s);
}

// Reduces s to the feature closest to the origin and returns that point.
__device__ __forceinline__ D3 simplex_closest(Simplex& s, bool* inside) {
    *inside = false;
    if (s.n == 1) return d3(s.w[0].p);
    if (s.n == 2) return closest_segment(s);
    if (s.n == 3) return closest_triangle(s);
    return closest_tetra(s, inside);
}

// The same, as a call: for the rare recomputations of kernels whose code
// should not hold a second inlined copy of the simplex code.
__device__ __noinline__ D3 simplex_closest_call(Simplex& s, bool* inside) { return simplex_closest(s, inside); }

// Sorts the simplex by vertex key so that the same feature always yields the
// same arithmetic regardless of the order its vertices were found in. Keys of
// one simplex are distinct; unused entries sort last. A fixed compare-exchange
// network keeps every index constant.
__device__ __forceinline__ void cmpx(SPoint& a, SPoint& b, int& ka, int& kb) {
    if (ka > kb) {
        const SPoint t = a;
        a = b;
        b = t;
        const int k = ka;
        ka = kb;
        kb = k;
    }
}
__device__ __forceinline__ void simplex_canonical(Simplex& s) {
    int k0 = s.w[0].key, k1 = s.n > 1 ? s.w[1].key : 0x7fffffff, k2 = s.n > 2 ? s.w[2].key : 0x7fffffff,
        k3 = s.n > 3 ? s.w[3].key : 0x7fffffff;
    cmpx(s.w[0], s.w[1], k0, k1);
    cmpx(s.w[2], s.w[3], k2, k3);
    cmpx(s.w[0], s.w[2], k0, k2);
    cmpx(s.w[1], s.w[3], k1, k3);
    cmpx(s.w[1], s.w[2], k1, k2);
}

// Every slot is assigned unconditionally (a select), so no store has a
// run-time address and the simplex stays in registers.
__device__ __forceinline__ SPoint sel(bool c, const SPoint& a, const SPoint& b) {
    SPoint r;
    r.p.x = c ? a.p.x : b.p.x;
    r.p.y = c ? a.p.y : b.p.y;
    r.p.z = c ? a.p.z : b.p.z;
    r.key = c ? a.key : b.key;
    return r;
}
__device__ __forceinline__ void simplex_push(Simplex& s, const SPoint& w) {
    s.w[0] = sel(s.n == 0, w, s.w[0]);
    s.w[1] = sel(s.n == 1, w, s.w[1]);
    s.w[2] = sel(s.n == 2, w, s.w[2]);
    s.w[3] = sel(s.n >= 3, w, s.w[3]);
    ++s.n;
}

__device__ __forceinline__ bool simplex_has(const Simplex& s, int key) {
    bool dup = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) dup = dup || (i < s.n && key >= 0 && s.w[i].key == key);
    return dup;
}

// ---------------------------------------------------------------- results

struct GjkOut {
    float sep;
    V3 n;
    int flags;
    int iters;
    int epa_iters;
};

// GJK phase of a distance query. init_dir: initial search direction (warm
// start or the centre difference); the first point is support_M(init_dir).
// On return s is the final feature, v its closest point, *hit whether the
// shapes intersect (then s encloses or touches the origin and EPA follows).
// A separation is only reported as converged when the duality gap certifies
// it: v.v - v.w <= GJK_REL_EPS v.v on the regular exit, <= GJK_STALL_EPS v.v
// when the support point repeats a simplex vertex or the distance stops
// decreasing. Such an exit without a certificate and with v.w <= 0 (a point of
// the Minkowski difference on the origin side) is *uncertain: it goes to EPA as
// a hit, and if EPA finds no enclosing polytope the uncertified separation is
// reported with GJK_BUDGET.
__device__ inline void gjk_phase(const Shape& A, const Shape& B, V3 init_dir, int max_iter, Simplex& s, D3& v,
                                 bool* hit, bool* converged, bool* uncertain, int* iters) {
    if (dot(init_dir, init_dir) == 0.0f) init_dir = v3(1.0f, 0.0f, 0.0f);
    s.w[0] = spoint(support_M(A, B, init_dir));
    s.n = 1;
    v = d3(s.w[0].p);
    *hit = false;
    *converged = false;
    *uncertain = false;
    int it = 0;
    while (it < max_iter) {
        const double vv = dot(v, v);
        if (vv <= GJK_TOUCH_EPS2) {
            *hit = true;
            break;
        }
        const SPoint w = spoint(support_M(A, B, to_v3(neg(v))));
        ++it;
        const double vw = dot(v, d3(w.p));
        if (vv - vw <= GJK_REL_EPS * vv) {
            *converged = true;
            break;
        }
        if (simplex_has(s, w.key)) {
            if (vv - vw <= GJK_STALL_EPS * vv) {
                *converged = true;
            } else if (vw <= 0.0) {
                *hit = true;
                *uncertain = true;
            }
            break;
        }
        const Simplex sp = s;
        simplex_push(s, w);
        bool inside = false;
        const D3 vn = simplex_closest(s, &inside);
        if (inside) {
            *hit = true;
            break;
        }
        if (!(dot(vn, vn) < vv)) {
            // no progress at FP64 resolution: the previous feature is final if the
            // duality gap certifies it to 1e-4
            s = sp;
            if (vv - vw <= GJK_STALL_EPS * vv) {
                *converged = true;
            } else if (vw <= 0.0) {
                *hit = true;
                *uncertain = true;
            }
            break;
        }
        v = vn;
    }
    *iters = it;
}

__device__ __forceinline__ GjkOut gjk_out_init() {
    GjkOut o;
    o.flags = 0;
    o.iters = 0;
    o.epa_iters = 0;
    o.sep = 0.0f;
    o.n = v3(1.0f, 0.0f, 0.0f);
    return o;
}

// Result of a separated query: distance and normal recomputed canonically from
// the converged feature (FP64, rounded to FP32 at the end). CALL: through
// simplex_closest_call instead of inline.
template <bool CALL = false>
__device__ inline void gjk_separated(Simplex& s, D3 v, bool converged, GjkOut& o) {
    if (!converged) o.flags |= GJK_BUDGET;
    else o.flags |= GJK_CONVERGED;
    bool canon = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) canon = canon && (i >= s.n || s.w[i].key >= 0);
    if (canon) {
        simplex_canonical(s);
        bool inside = false;
        v = CALL ? simplex_closest_call(s, &inside) : simplex_closest(s, &inside);
    }
    const double dist = sqrt(dot(v, v));
    o.sep = (float)dist;
    o.n = dist > 0.0 ? to_v3(scl(v, -1.0 / dist)) : v3(1.0f, 0.0f, 0.0f);
}

}  // namespace sd
