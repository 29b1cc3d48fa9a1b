// EPA penetration depth on a 16-lane group: four queries per wavefront.
//
// A group is one DPP row, so its reductions (the support argmax over hull
// vertices, the closest-face argmin) are four DPP steps over (value, index)
// pairs; the winning support vertex is then read from the lane that holds it
// (three ds_bpermute) instead of riding along in every step.
// The two hulls' vertices (local frame) are loaded into registers once per
// query, EPA_VSLOT per lane; larger hulls read the rest from memory. The
// polytope (vertices, faces with unit normal and plane distance, horizon
// edges) lives in LDS, one region per group. Faces are owned round-robin
// (face f by lane f mod 16); visibility, horizon, the compaction of the
// surviving faces and the new faces use ballots and prefix counts over the
// group. Values that steer control flow are equal on the group's lanes, so a
// group runs uniformly; the four groups of a wave diverge freely.
//
// (Eight-lane groups, eight queries per wave, were slower at C4: 210 VGPRs
// and the polytopes of 16 groups per block left 1.5 waves per SIMD.)
//
// The algorithm is the serial one of the CPU restatement, step for step:
//   grow the GJK simplex to a tetrahedron; faces (0,1,2) (0,3,1) (0,2,3) (1,3,2)
//   oriented away from the centroid; repeat: closest face (first on ties),
//   support point along its normal (first vertex on ties), stop when the gap is
//   <= 1e-6 |d| + 1e-9 or the point repeats a vertex; faces that see the point
//   are removed; the horizon is every edge (a, b), (b, c), (c, a) of a visible
//   face, in face order, whose reverse edge belongs to no visible face (the
//   visible faces' directed edges are marked in a bitmap first); one new
//   face per horizon edge (degenerate ones skipped), appended after the
//   surviving faces, which keep their order.
// Depth and normal come from the final closest face with its vertices ordered
// by key, so equal features give equal bits.
#pragma once

#include "sd_gjk_dev.hpp"

namespace sd {

constexpr int EPA_G = 16;                                // lanes per query (one DPP row)
constexpr int EPA_GPW = 64 / EPA_G;                      // queries per wavefront
constexpr int EPA_VSLOT = 2;                             // hull vertices per lane in registers
static_assert(EPA_G == 8 || EPA_G == 16, "groups are half or whole DPP rows");
constexpr int EPA_FJ = (EPA_MAX_F + EPA_G - 1) / EPA_G;  // faces per lane
constexpr int EPA_EK = (EPA_MAX_E + EPA_G - 1) / EPA_G;  // horizon edges per lane
// directed edges (a, b) of the visible faces, bit a * EPA_MAX_V + b
constexpr int EPA_EMAP_WORDS = (EPA_MAX_V * EPA_MAX_V + 31) / 32;

struct EpaShared {
    float4 P[EPA_MAX_V];   // (x, y, z, key bits)
    float4 fn[EPA_MAX_F];  // (nx, ny, nz, d)
    int fabc[EPA_MAX_F];   // a | b << 8 | c << 16
    unsigned char ea[EPA_MAX_E];
    unsigned char eb[EPA_MAX_E];
    uint32_t emap[EPA_EMAP_WORDS];  // all zero between iterations
};

__device__ __forceinline__ void emap_mark(EpaShared& S, int a, int b) {
    const int e = a * EPA_MAX_V + b;
    atomicOr(&S.emap[e >> 5], 1u << (e & 31));
}
__device__ __forceinline__ bool emap_has(const EpaShared& S, int a, int b) {
    const int e = a * EPA_MAX_V + b;
    return (S.emap[e >> 5] >> (e & 31)) & 1u;
}
__device__ __forceinline__ void emap_clear(EpaShared& S, int gl) {
    for (int w = gl; w < EPA_EMAP_WORDS; w += EPA_G) S.emap[w] = 0u;
}

// Orders this wave's LDS accesses (all lanes of a group are in one wave).
__device__ __forceinline__ void grp_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t grp_ballot(bool p, int base) {
    return (uint32_t)((__ballot(p) >> base) & ((1ull << EPA_G) - 1ull));
}
__device__ __forceinline__ int grp_below(uint32_t m, int gl) { return __popc(m & ((1u << gl) - 1u)); }

constexpr int DPP_XOR1 = 0xB1;          // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;          // quad_perm [2, 3, 0, 1]
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i <- 7 - i within 8 lanes
constexpr int DPP_MIRROR = 0x140;       // lane i <- 15 - i within 16 lanes

template <int C>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, C, 0xf, 0xf, false);
}
template <int C>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(dpp_i<C>(__float_as_int(v)));
}

template <int C>
__device__ __forceinline__ void amin_step(float& v, int& i) {
    const float ov = dpp_f<C>(v);
    const int oi = dpp_i<C>(i);
    const bool t = (ov < v) | ((ov == v) & (oi < i));  // bitwise: no branches
    v = t ? ov : v;
    i = t ? oi : i;
}

// A hull's first EPA_G * EPA_VSLOT local-frame vertices, vertex gl + EPA_G j
// in slot j of lane gl.
struct HullRegs {
    V3 v[EPA_VSLOT];
};

__device__ __forceinline__ HullRegs load_hull_regs(const Shape& s, int gl) {
    HullRegs r;
#pragma unroll
    for (int j = 0; j < EPA_VSLOT; ++j) {
        const int k = gl + EPA_G * j;
        r.v[j] = v3(0.0f, 0.0f, 0.0f);
        if (k < s.nv) {
            const float4 p = s.verts[k];
            r.v[j] = v3(p.x, p.y, p.z);
        }
    }
    return r;
}

// (value, index) with the larger value, the smaller index on ties.
template <int C>
__device__ __forceinline__ void amax_vi_step(float& v, int& i) {
    const float ov = dpp_f<C>(v);
    const int oi = dpp_i<C>(i);
    const bool t = (ov > v) | ((ov == v) & (oi < i));  // bitwise: no branches
    v = t ? ov : v;
    i = t ? oi : i;
}

// Support point of a shape along u across the group: the vertex with the
// largest local-frame projection, first index on ties (a serial scan's pick).
// The arg-max carries (value, index); the winning vertex then comes from the
// lane that holds it in registers (ds_bpermute), or from memory past them.
__device__ __forceinline__ V3 grp_support_point(const Shape& s, const HullRegs& hr, V3 u, int gl, int& idx) {
    if (s.nv == 0) {
        idx = -1;
        const float uu = dot(u, u);
        if (uu == 0.0f) return s.off;
        const float k = s.radius / sqrtf(uu);
        return add(s.off, scl(u, k));
    }
    const V3 l = rot_tmul(s.R, u);
    float bv = -3.402823466e38f;
    int bi = 0x7fffffff;
#pragma unroll
    for (int j = 0; j < EPA_VSLOT; ++j) {  // indices ascend within a lane: the first maximum is kept
        const int k = gl + EPA_G * j;
        if (k < s.nv) {
            const V3 p = hr.v[j];
            const float d = p.x * l.x + p.y * l.y + p.z * l.z;
            if (d > bv) {
                bv = d;
                bi = k;
            }
        }
    }
    for (int k = gl + EPA_G * EPA_VSLOT; k < s.nv; k += EPA_G) {
        const float4 q = s.verts[k];
        const float d = q.x * l.x + q.y * l.y + q.z * l.z;
        if (d > bv) {
            bv = d;
            bi = k;
        }
    }
    amax_vi_step<DPP_XOR1>(bv, bi);
    amax_vi_step<DPP_XOR2>(bv, bi);
    amax_vi_step<DPP_HALF_MIRROR>(bv, bi);
    if constexpr (EPA_G == 16) amax_vi_step<DPP_MIRROR>(bv, bi);
    if (bi == 0x7fffffff) bi = 0;  // no ordered projection: vertex 0, as a serial scan
    V3 bp;
    if (bi < EPA_G * EPA_VSLOT) {
        const int j = bi / EPA_G;  // equal on the group's lanes
        V3 mine = hr.v[0];
#pragma unroll
        for (int jj = 1; jj < EPA_VSLOT; ++jj)
            if (j == jj) mine = hr.v[jj];
        const int src = 4 * ((int)(__lane_id() & (64 - EPA_G)) + (bi & (EPA_G - 1)));
        bp = v3(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mine.x))),
                __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mine.y))),
                __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mine.z))));
    } else {
        const float4 q = s.verts[bi];
        bp = v3(q.x, q.y, q.z);
    }
    idx = bi;
    return add(rot_mul(s.R, bp), s.off);
}

__device__ __forceinline__ MPoint grp_support_M(const Shape& A, const HullRegs& hA, const Shape& B,
                                                const HullRegs& hB, V3 u, int gl) {
    int ia, ib;
    const V3 pa = grp_support_point(A, hA, u, gl, ia);
    const V3 pb = grp_support_point(B, hB, neg(u), gl, ib);
    MPoint m;
    m.p = sub(pa, pb);
    m.key = (ia < 0 || ib < 0) ? -1 : ia * GJK_MAX_VERTS_PER_HULL + ib;
    return m;
}

__device__ __forceinline__ V3 pv(const float4& q) { return v3(q.x, q.y, q.z); }

// Face (a, b, c) oriented away from `inner`; false if degenerate.
__device__ __forceinline__ bool face_of(V3 pa, V3 pb, V3 pc, V3 inner, int a, int b, int c, int& abc, float4& nd) {
    V3 nrm = cross(sub(pb, pa), sub(pc, pa));
    const float l2 = dot(nrm, nrm);
    if (!(l2 > 0.0f)) return false;
    if (dot(nrm, sub(pa, inner)) < 0.0f) {
        const int t = b;
        b = c;
        c = t;
        nrm = neg(nrm);
    }
    const float inv = 1.0f / sqrtf(l2);
    const V3 n = scl(nrm, inv);
    abc = a | (b << 8) | (c << 16);
    nd = make_float4(n.x, n.y, n.z, dot(n, pa));
    return true;
}

// Closest face (smallest d, first on ties) of the nf faces in S.
__device__ __forceinline__ int grp_closest_face(const EpaShared& S, int nf, int gl) {
    float bd = 3.402823466e38f;
    int bf = 0x7fffffff;
    for (int f = gl; f < nf; f += EPA_G) {
        const float d = S.fn[f].w;
        if (d < bd) {
            bd = d;
            bf = f;
        }
    }
    amin_step<DPP_XOR1>(bd, bf);
    amin_step<DPP_XOR2>(bd, bf);
    amin_step<DPP_HALF_MIRROR>(bd, bf);
    if constexpr (EPA_G == 16) amin_step<DPP_MIRROR>(bd, bf);
    return bf == 0x7fffffff ? 0 : bf;  // no ordered distance: the first face, as a serial scan
}

__device__ __forceinline__ void cmpx3(MPoint& a, MPoint& b) {
    if (a.key > b.key) {
        const MPoint t = a;
        a = b;
        b = t;
    }
}

// EPA's outcome: the closest face of the final polytope (its three points
// with their keys, in face order) and the interior point that orients it, or
// none (a degenerate start); flags and iterations.
struct EpaFace {
    MPoint q[3];
    V3 inner;
    float d;  // plane distance of the face in the polytope
    bool has_face;
    int flags;
    int iters;
};

__device__ inline void grp_epa(const Shape& A, const HullRegs& hA, const Shape& B, const HullRegs& hB,
                               const Simplex& s0, int max_iter, EpaFace& o, EpaShared& S, int gl, int base) {
    o.has_face = false;
    o.flags = 0;
    o.iters = 0;
    o.d = 0.0f;
    o.inner = v3(0.0f, 0.0f, 0.0f);
    // grow the simplex to a tetrahedron
    MPoint T0, T1, T2, T3;
    T0.p = s0.w[0].p;
    T0.key = s0.w[0].key;
    T1.p = s0.w[1].p;
    T1.key = s0.w[1].key;
    T2.p = s0.w[2].p;
    T2.key = s0.w[2].key;
    T3.p = s0.w[3].p;
    T3.key = s0.w[3].key;
    int np = s0.n;
    if (np == 1) {
        for (int k = 0; k < 6 && np == 1; ++k) {
            const int j = k >> 1;
            const float sg = (k & 1) ? -1.0f : 1.0f;
            const V3 u = v3(j == 0 ? sg : 0.0f, j == 1 ? sg : 0.0f, j == 2 ? sg : 0.0f);
            const MPoint m = grp_support_M(A, hA, B, hB, u, gl);
            const V3 dd = sub(m.p, T0.p);
            if (dot(dd, dd) > 1e-20f) {
                T1 = m;
                np = 2;
            }
        }
    }
    if (np == 2) {
        const V3 ab = sub(T1.p, T0.p);
        const float ax = fabsf(ab.x), ay = fabsf(ab.y), az = fabsf(ab.z);
        const int j = (ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2);
        const V3 e = v3(j == 0 ? 1.0f : 0.0f, j == 1 ? 1.0f : 0.0f, j == 2 ? 1.0f : 0.0f);
        const V3 u = cross(ab, e);
        for (int k = 0; k < 2 && np == 2; ++k) {
            const MPoint m = grp_support_M(A, hA, B, hB, k ? neg(u) : u, gl);
            const V3 c = cross(ab, sub(m.p, T0.p));
            if (dot(c, c) > 1e-24f) {
                T2 = m;
                np = 3;
            }
        }
    }
    if (np == 3) {
        const V3 nrm = cross(sub(T1.p, T0.p), sub(T2.p, T0.p));
        for (int k = 0; k < 2 && np == 3; ++k) {
            const MPoint m = grp_support_M(A, hA, B, hB, k ? neg(nrm) : nrm, gl);
            const float vol = dot(nrm, sub(m.p, T0.p));
            if (fabsf(vol) > 1e-24f) {
                T3 = m;
                np = 4;
            }
        }
    }
    if (np < 4) {
        o.flags |= GJK_HIT | GJK_DEGENERATE;
        return;
    }
    const V3 inner = scl(add(add(T0.p, T1.p), add(T2.p, T3.p)), 0.25f);
    int nf = 0;
    {
        const MPoint Tv[4] = {T0, T1, T2, T3};
        const int T[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
        int abc[4];
        float4 nd[4];
        bool ok[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            ok[k] = face_of(Tv[T[k][0]].p, Tv[T[k][1]].p, Tv[T[k][2]].p, inner, T[k][0], T[k][1], T[k][2], abc[k],
                            nd[k]);
        if (gl == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                S.P[k] = make_float4(Tv[k].p.x, Tv[k].p.y, Tv[k].p.z, __int_as_float(Tv[k].key));
                if (ok[k]) {
                    S.fabc[nf] = abc[k];
                    S.fn[nf] = nd[k];
                    ++nf;
                }
            }
        }
        nf = (ok[0] ? 1 : 0) + (ok[1] ? 1 : 0) + (ok[2] ? 1 : 0) + (ok[3] ? 1 : 0);
    }
    emap_clear(S, gl);
    grp_sync();
    if (nf < 4) {
        o.flags |= GJK_HIT | GJK_DEGENERATE;
        return;
    }
    int it = 0;
    bool converged = false;
    for (; it < max_iter; ++it) {
        const int best = grp_closest_face(S, nf, gl);
        const float4 fb = S.fn[best];
        const V3 fbn = v3(fb.x, fb.y, fb.z);
        const MPoint m = grp_support_M(A, hA, B, hB, fbn, gl);
        const float gap = dot(m.p, fbn) - fb.w;
        if (gap <= EPA_REL_EPS * fabsf(fb.w) + 1e-9f) {
            converged = true;
            break;
        }
        bool mine = false;
        for (int k = gl; k < np; k += EPA_G) mine = mine || (m.key >= 0 && __float_as_int(S.P[k].w) == m.key);
        const bool dup = grp_ballot(mine, base) != 0u;
        if (dup || np >= EPA_MAX_V) {
            converged = dup;
            if (!dup) o.flags |= GJK_EPA_LIMIT;
            break;
        }
        const int iv = np;
        if (gl == 0) S.P[iv] = make_float4(m.p.x, m.p.y, m.p.z, __int_as_float(m.key));
        ++np;
        grp_sync();
        // visibility of this lane's faces f = gl + EPA_G j
        const int nj = (nf + EPA_G - 1) / EPA_G;
        int abc[EPA_FJ];
        int hm[EPA_FJ];  // visible: bit k set while edge k may be on the horizon
        uint32_t vm[EPA_FJ];
#pragma unroll
        for (int j = 0; j < EPA_FJ; ++j) {
            const int f = gl + EPA_G * j;
            abc[j] = 0;
            hm[j] = 0;
            vm[j] = 0u;
            if (j < nj) {
                if (f < nf) {
                    abc[j] = S.fabc[f];
                    const float4 q = S.fn[f];
                    if (dot(v3(q.x, q.y, q.z), sub(m.p, pv(S.P[abc[j] & 0xff]))) > 0.0f) hm[j] = 7;
                }
                vm[j] = grp_ballot(hm[j] != 0, base);
            }
        }
        // horizon: an edge of a visible face survives unless a visible face
        // holds its reverse. The visible faces mark their directed edges in
        // the group's edge bitmap, then test the reverses of their own.
#pragma unroll
        for (int j = 0; j < EPA_FJ; ++j) {
            if (hm[j]) {
                const int a = abc[j] & 0xff, b = (abc[j] >> 8) & 0xff, c = (abc[j] >> 16) & 0xff;
                emap_mark(S, a, b);
                emap_mark(S, b, c);
                emap_mark(S, c, a);
            }
        }
        grp_sync();
#pragma unroll
        for (int j = 0; j < EPA_FJ; ++j) {
            if (hm[j]) {
                const int a = abc[j] & 0xff, b = (abc[j] >> 8) & 0xff, c = (abc[j] >> 16) & 0xff;
                if (emap_has(S, b, a)) hm[j] &= ~1;
                if (emap_has(S, c, b)) hm[j] &= ~2;
                if (emap_has(S, a, c)) hm[j] &= ~4;
            }
        }
        // horizon edge offsets (face order) and surviving face slots (face order)
        int ne = 0, nsurv = 0;
        int off[EPA_FJ], dst[EPA_FJ];
#pragma unroll
        for (int j = 0; j < EPA_FJ; ++j) {
            off[j] = 0;
            dst[j] = -1;
            if (j < nj) {
                const int h = __popc(hm[j]);
                const uint32_t b0 = grp_ballot(h & 1, base), b1 = grp_ballot(h & 2, base);
                off[j] = ne + grp_below(b0, gl) + 2 * grp_below(b1, gl);
                ne += __popc(b0) + 2 * __popc(b1);
                const bool sv = gl + EPA_G * j < nf && ((vm[j] >> gl) & 1u) == 0u;
                const uint32_t sm = grp_ballot(sv, base);
                if (sv) dst[j] = nsurv + grp_below(sm, gl);
                nsurv += __popc(sm);
            }
        }
        const bool overflow = ne > EPA_MAX_E;
        if (!overflow) {
#pragma unroll
            for (int j = 0; j < EPA_FJ; ++j) {
                if (hm[j]) {
                    const int a = abc[j] & 0xff, b = (abc[j] >> 8) & 0xff, c = (abc[j] >> 16) & 0xff;
                    int q = off[j];
                    if (hm[j] & 1) {
                        S.ea[q] = (unsigned char)a;
                        S.eb[q] = (unsigned char)b;
                        ++q;
                    }
                    if (hm[j] & 2) {
                        S.ea[q] = (unsigned char)b;
                        S.eb[q] = (unsigned char)c;
                        ++q;
                    }
                    if (hm[j] & 4) {
                        S.ea[q] = (unsigned char)c;
                        S.eb[q] = (unsigned char)a;
                    }
                }
            }
        }
        grp_sync();  // the horizon scan has read every face and the edge bitmap
        emap_clear(S, gl);
        // compaction: a surviving face moves to a slot <= its own; round j reads
        // faces [EPA_G j, EPA_G (j + 1)) and writes below that range's end
#pragma unroll
        for (int j = 0; j < EPA_FJ; ++j) {
            if (j < nj) {
                float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (dst[j] >= 0) q = S.fn[gl + EPA_G * j];
                grp_sync();
                if (dst[j] >= 0) {
                    S.fabc[dst[j]] = abc[j];
                    S.fn[dst[j]] = q;
                }
                grp_sync();
            }
        }
        nf = nsurv;
        if (overflow || nf + ne > EPA_MAX_F) {
            o.flags |= GJK_EPA_LIMIT;
            break;
        }
        // one new face per horizon edge, in edge order
        const int nk = (ne + EPA_G - 1) / EPA_G;
#pragma unroll
        for (int k = 0; k < EPA_EK; ++k) {
            if (k < nk) {
                const int q = gl + EPA_G * k;
                bool okf = false;
                int nabc = 0;
                float4 nnd = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (q < ne) {
                    const int a = S.ea[q], b = S.eb[q];
                    okf = face_of(pv(S.P[a]), pv(S.P[b]), pv(S.P[iv]), inner, a, b, iv, nabc, nnd);
                }
                const uint32_t om = grp_ballot(okf, base);
                if (okf) {
                    const int d = nf + grp_below(om, gl);
                    S.fabc[d] = nabc;
                    S.fn[d] = nnd;
                }
                nf += __popc(om);
            }
        }
        grp_sync();
        if (nf == 0) {
            o.flags |= GJK_DEGENERATE;
            break;
        }
    }
    o.iters = it;
    if (nf == 0) {
        o.flags |= GJK_HIT | GJK_DEGENERATE;
        return;
    }
    const int best = grp_closest_face(S, nf, gl);
    const int abc = S.fabc[best];
    const float4 pa = S.P[abc & 0xff], pb = S.P[(abc >> 8) & 0xff], pc = S.P[(abc >> 16) & 0xff];
    o.q[0].p = pv(pa);
    o.q[0].key = __float_as_int(pa.w);
    o.q[1].p = pv(pb);
    o.q[1].key = __float_as_int(pb.w);
    o.q[2].p = pv(pc);
    o.q[2].key = __float_as_int(pc.w);
    o.inner = inner;
    o.d = S.fn[best].w;
    o.has_face = true;
    o.flags |= GJK_HIT | (converged ? GJK_CONVERGED : 0);
    if (!converged && it >= max_iter) o.flags |= GJK_BUDGET;
    if (o.d < 0.0f) o.flags |= GJK_DEGENERATE;  // polytope does not enclose the origin
}

// Depth and normal of EPA's closest face: canonical from its closest features,
// else the face's plane with its points ordered by key. CERT: also evaluate
// the certificate (o.certified).
template <bool CERT>
__device__ __forceinline__ void hit_result(const Shape& A, const Shape& B, MPoint* q, V3 inner, GjkOut& o) {
    Features f;
    o.certified = false;
    if (feature_sets(q, 3, f) && f.na <= 3 && f.nb <= 3) {
        V3 fa[3], fb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            fa[k] = k < f.na ? vert_at(A, f.ia[k]) : v3(0.0f, 0.0f, 0.0f);
            fb[k] = k < f.nb ? vert_at(B, f.ib[k]) : v3(0.0f, 0.0f, 0.0f);
        }
        const Canon c = canon_feature(fa, f.na, fb, f.nb, true);
        if (c.kind != 0) {
            o.sep = c.sep;
            o.n = c.n;
            if (CERT) o.certified = c.ok && certify(A, B, f, c.n);
            return;
        }
    }
    cmpx_key(q[0], q[1]);  // stable for equal keys: the order of a serial insertion sort
    cmpx_key(q[1], q[2]);
    cmpx_key(q[0], q[1]);
    V3 nrm = cross(sub(q[1].p, q[0].p), sub(q[2].p, q[0].p));
    const float l2 = dot(nrm, nrm);
    if (dot(nrm, sub(q[0].p, inner)) < 0.0f) nrm = neg(nrm);
    const V3 n = scl(nrm, 1.0f / sqrtf(l2));
    o.sep = -dot(n, q[0].p);
    o.n = n;
}

}  // namespace sd
