// Rigid convex bodies stepped on the GPU (sd_hull_*): convex hulls given by
// their vertices, spheres (no vertices, a radius) and static triangles (three
// vertices, mass 0) in independent scenes, each step
//   pose      shape centre (x, y, z, radius) and world angular velocity
//   boxes     k_hull_aabb
//   pairs     a uniform grid per scene (fixed at upload): bodies sorted by
//             cell, each body's overlapping partners in the later part of its
//             own cell and the 13 cells after it, then the large static
//             bodies (walls, mesh triangles) of a second grid around it
//   narrow    GJK/EPA of every candidate pair: cold, or warm-started from
//             the directions of an attached learned model (sd_hull_set_warm),
//             the rejected warm queries redone cold
//   contacts  per touching pair (sep < 0): contact point, linear
//             spring-dashpot normal force, Coulomb-capped viscous friction
//   gather    per body, its contacts in pair order (no float atomics), then
//             the semi-implicit integration of the clump path
// with every launch sized to the pair capacity and every count read on the
// device, so the step is captured once in a hipGraph and replayed.
//
// Replaces the reference sandbox's hull pipeline (sandbox/testGJK.cu:237-315
// gjkCUDA on Hull pairs, primitives.hpp:37-57 Hull with SPHERE, TRIANGLE and
// POLYHEDRA types), which stops at a boolean overlap: here the narrowphase
// result drives the bodies, and triangles (the reference's DEMMesh,
// src/engine/entities/entity.hpp:64-93) are static bodies that hulls and
// spheres collide with.
#include <algorithm>
#include <cstring>
#include <vector>

#include "sd_gjk_dev.hpp"
#include "sd_internal.hpp"
#include "sd_rigid_dev.hpp"

namespace sd {

// Support-feature tolerance: vertices within this fraction of a hull's radius
// (its largest vertex distance from the centre) of its extreme one along the
// contact normal form the contact feature.
constexpr float HB_FEATURE_REL = 1e-3f;
// Vertices in flight together in the contact feature's passes.
constexpr int HB_FEATURE_BATCH = 8;
// A warm-started query is kept only beyond this distance from touching (m);
// nearer ones are redone cold.
constexpr float HB_WARM_NEAR_MARGIN = 1e-6f;

__global__ void k_hb_pose(const float4* __restrict__ P, const float4* __restrict__ Q, const float4* __restrict__ L,
                          const float4* __restrict__ I0, const float4* __restrict__ I1,
                          const float* __restrict__ radius, int n, float4* __restrict__ spos,
                          float4* __restrict__ omega) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const float4 p = P[b], l = L[b];
    spos[b] = make_float4(p.x, p.y, p.z, radius[b]);
    const V3 w = inv_inertia_apply(quat_rot(Q[b]), I0[b], I1[b], v3(l.x, l.y, l.z));
    omega[b] = make_float4(w.x, w.y, w.z, 0.0f);
}

// Support-feature tolerance of every body: HB_FEATURE_REL times its radius
// (largest vertex distance from its centre); 0 for a sphere. Fixed at upload.
__global__ void k_hb_feature_tol(const float4* __restrict__ verts, const int* __restrict__ vstart,
                                 const int* __restrict__ nv, int n, float* __restrict__ tol) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const float4* vb = verts + vstart[b];
    float r2 = 0.0f;
    for (int k = 0; k < nv[b]; ++k) {
        const float4 v = vb[k];
        r2 = fmaxf(r2, v.x * v.x + v.y * v.y + v.z * v.z);
    }
    tol[b] = HB_FEATURE_REL * sqrtf(r2);
}

// Contact feature of a shape along unit direction u: for a sphere the point
// c + r u (spread 0); for a vertex set the vertices whose projection on u is
// within tol (k_hb_feature_tol) of the largest, their mean in vertex order
// (world) and spread2 = the largest squared distance of one of them from that
// mean. The vertices of a hull of up to 32 are selected once, as a bit mask.
__device__ __forceinline__ V3 feature_centroid(const float4* __restrict__ verts, int nv, float radius, float tol,
                                               const Rot& R, V3 X, V3 u, float& spread2) {
    spread2 = 0.0f;
    if (nv == 0) return add(X, scl(u, radius));
    const V3 l = rot_tmul(R, u);
    // vertices read HB_FEATURE_BATCH at a time, their loads in flight
    // together; indices past the last vertex read it again (harmless for the
    // maximum, skipped by the sums)
    const int last = nv - 1;
    float dmax = -3.402823466e38f;
    for (int k = 0; k < nv; k += HB_FEATURE_BATCH) {
        float4 p[HB_FEATURE_BATCH];
#pragma unroll
        for (int j = 0; j < HB_FEATURE_BATCH; ++j) p[j] = verts[min(k + j, last)];
#pragma unroll
        for (int j = 0; j < HB_FEATURE_BATCH; ++j) dmax = fmaxf(dmax, p[j].x * l.x + p[j].y * l.y + p[j].z * l.z);
    }
    const float cut = dmax - tol;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int cnt = 0;
    uint32_t sel = 0u;
    for (int k = 0; k < nv; k += HB_FEATURE_BATCH) {
        float4 p[HB_FEATURE_BATCH];
#pragma unroll
        for (int j = 0; j < HB_FEATURE_BATCH; ++j) p[j] = verts[min(k + j, last)];
#pragma unroll
        for (int j = 0; j < HB_FEATURE_BATCH; ++j) {
            const float d = p[j].x * l.x + p[j].y * l.y + p[j].z * l.z;
            if (k + j < nv && d >= cut) {
                sx = sx + p[j].x;
                sy = sy + p[j].y;
                sz = sz + p[j].z;
                ++cnt;
                if (k + j < 32) sel |= 1u << (k + j);
            }
        }
    }
    const float inv = 1.0f / (float)cnt;
    const float mx = sx * inv, my = sy * inv, mz = sz * inv;
    if (nv <= 32) {
        for (uint32_t m = sel; m; m &= m - 1u) {
            const float4 v = verts[__builtin_ctz(m)];
            const float ex = v.x - mx, ey = v.y - my, ez = v.z - mz;
            spread2 = fmaxf(spread2, ex * ex + ey * ey + ez * ez);
        }
    } else {
        for (int k = 0; k < nv; ++k) {
            const float4 v = verts[k];
            const float d = v.x * l.x + v.y * l.y + v.z * l.z;
            if (d >= cut) {
                const float ex = v.x - mx, ey = v.y - my, ez = v.z - mz;
                spread2 = fmaxf(spread2, ex * ex + ey * ey + ez * ez);
            }
        }
    }
    return add(rot_mul(R, v3(mx, my, mz)), X);
}

// Contact point of two features: the centroid of the one of smaller spread
// (w = squared spread), the midpoint of the two when the spreads are equal
// (two points, two like faces), so that a point or a small face touching a
// large face acts where it touches.
__device__ __forceinline__ V3 contact_point(V3 ca, float wa, V3 cb, float wb) {
    const float t = wa < wb ? 0.0f : (wb < wa ? 1.0f : 0.5f);
    return add(ca, scl(sub(cb, ca), t));
}

// Broadphase grids of one scene, fixed at upload from the initial bodies:
// origin, 1 / cell, cells per axis and first global cell id of the grid of
// its bodies, then the same of the grid of its large bodies. The first cell
// is twice the largest bounding radius rd of the scene's moving bodies
// (inflated by 2^-10), so two gridded bodies whose boxes overlap have centres
// at most one cell apart on every axis; bodies outside a grid are clamped to
// its border cells, which only merges cells. A body whose bounding radius
// exceeds rd (a wall or mesh triangle) is static and goes to the second grid,
// whose cell is twice the largest such radius: a gridded body and a large
// body whose boxes overlap lie within one cell of it.
struct HbGrid {
    float lox, loy, loz, inv;
    int nx, ny, nz, cbase;
    float llox, lloy, lloz, linv;
    int lnx, lny, lnz, lcbase;
};

__device__ __forceinline__ int hb_cell_axis(float x, float lo, float inv, int nd) {
    return (int)fminf(fmaxf(floorf((x - lo) * inv), 0.0f), (float)(nd - 1));
}

// key = (global cell of the body's centre) << 32 | body index; a large body
// (bucketed in its scene's second grid at upload) takes cell ncells, past
// every scene's cells of the first grids
__global__ void k_hb_cell_keys(const float4* __restrict__ P, const int* __restrict__ scene_of,
                               const HbGrid* __restrict__ grid, const int* __restrict__ large, int n, int ncells,
                               uint64_t* __restrict__ keys) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    uint32_t cell = (uint32_t)ncells;
    if (!large[b]) {
        const HbGrid g = grid[scene_of[b]];
        const float4 p = P[b];
        const int cx = hb_cell_axis(p.x, g.lox, g.inv, g.nx);
        const int cy = hb_cell_axis(p.y, g.loy, g.inv, g.ny);
        const int cz = hb_cell_axis(p.z, g.loz, g.inv, g.nz);
        cell = (uint32_t)g.cbase + (uint32_t)((cz * g.ny + cy) * g.nx + cx);
    }
    keys[b] = ((uint64_t)cell << 32) | (uint64_t)(uint32_t)b;
}

// Pairs a walk of the grid keeps in a body's own slots (int2 each, slot k of
// body a at k n + a, so that neighbouring threads touch neighbouring slots); a
// body with more walks the grid again and writes them in place.
constexpr int HB_GRID_SLOTS = 32;

struct HbGridArgs {
    const float4* P;
    const float4* lo;
    const float4* hi;
    const int* scene_of;
    const HbGrid* grid;
    const int* large;
    const int* llist;        // large bodies sorted by cell of the second grid, then index
    const int* lstart;       // their cell starts
    const uint64_t* sorted;  // keys sorted by cell, then index
    const int* cell_start;
};

// Candidate pairs of a gridded body a whose boxes overlap a's on all three
// axes (closed intervals), except pairs of two static bodies, each as (min,
// max): the bodies after a in its own cell, then those of the 13 cells that
// follow its own in (dz, dy, dx) order, within a cell in index order, then the
// large bodies of the 27 cells of the second grid around a's centre (dz, dy,
// dx ascending; index order within a cell). Overlapping gridded bodies lie in
// the same or neighbouring cells, so each such pair is found once, from the
// body whose cell comes first. Pair k goes to emit(k, lo, hi); returns the
// count.
template <typename Emit>
__device__ __forceinline__ int hb_grid_walk(const HbGridArgs& G, int a, Emit emit) {
    const HbGrid g = G.grid[G.scene_of[a]];
    const float4 pa = G.P[a], la = G.lo[a], ha = G.hi[a];
    const bool sa = pa.w == 0.0f;  // static (1/m = 0)
    const int cx = hb_cell_axis(pa.x, g.lox, g.inv, g.nx);
    const int cy = hb_cell_axis(pa.y, g.loy, g.inv, g.ny);
    const int cz = hb_cell_axis(pa.z, g.loz, g.inv, g.nz);
    int k = 0;
    // the stencil's rows (dz, dy) in order; a row's cells are consecutive in
    // the sorted keys, so each row is one window of positions: from a's own
    // cell (the bodies after a) for the first row, x - 1 .. x + 1 (clamped)
    // for the others
    for (int dz = 0; dz <= 1; ++dz) {
        const int z = cz + dz;
        if (z >= g.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.ny || (dz == 0 && dy < 0)) continue;
            const bool own_row = dz == 0 && dy == 0;
            const int x0 = own_row ? cx : max(cx - 1, 0), x1 = min(cx + 1, g.nx - 1);
            const int row = g.cbase + (z * g.ny + y) * g.nx;
            const int t0 = G.cell_start[row + x0], t1 = G.cell_start[row + x1 + 1];
            const int own_end = own_row ? G.cell_start[row + cx + 1] : t0;  // a's own cell: [t0, own_end)
            for (int t = t0; t < t1; ++t) {
                const int b = (int)(uint32_t)G.sorted[t];
                if (t < own_end && b <= a) continue;
                if (sa && G.P[b].w == 0.0f) continue;
                const float4 lb = G.lo[b], hb = G.hi[b];
                if (!(lb.x <= ha.x && la.x <= hb.x && lb.y <= ha.y && la.y <= hb.y && lb.z <= ha.z && la.z <= hb.z))
                    continue;
                emit(k++, min(a, b), max(a, b));
            }
        }
    }
    if (sa) return k;  // large bodies are static
    const int qx = hb_cell_axis(pa.x, g.llox, g.linv, g.lnx);
    const int qy = hb_cell_axis(pa.y, g.lloy, g.linv, g.lny);
    const int qz = hb_cell_axis(pa.z, g.lloz, g.linv, g.lnz);
    const int qx0 = max(qx - 1, 0), qx1 = min(qx + 1, g.lnx - 1);
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = qz + dz;
        if (z < 0 || z >= g.lnz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = qy + dy;
            if (y < 0 || y >= g.lny) continue;
            const int row = g.lcbase + (z * g.lny + y) * g.lnx;
            const int t1 = G.lstart[row + qx1 + 1];
            for (int t = G.lstart[row + qx0]; t < t1; ++t) {
                const int b = G.llist[t];
                const float4 lb = G.lo[b], hb = G.hi[b];
                if (!(lb.x <= ha.x && la.x <= hb.x && lb.y <= ha.y && la.y <= hb.y && lb.z <= ha.z && la.z <= hb.z))
                    continue;
                emit(k++, min(a, b), max(a, b));
            }
        }
    }
    return k;
}

// One thread per body: its pair count (0 for a large body, which pairs only
// through the gridded bodies) and its first HB_GRID_SLOTS pairs in its slots.
__global__ void k_hb_grid_pairs(HbGridArgs G, int n, int* __restrict__ cnt, int2* __restrict__ slots) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    if (G.large[a]) {
        cnt[a] = 0;
        return;
    }
    cnt[a] = hb_grid_walk(G, a, [&](int k, int u, int v) {
        if (k < HB_GRID_SLOTS) slots[(size_t)k * n + a] = make_int2(u, v);
    });
}

// One thread per body: its pairs into the list from offs[a] (pairs past
// capacity are dropped: the total exceeds the capacity and the step reports
// it), copied from its slots, or walked again when they did not fit.
__global__ void k_hb_grid_place(HbGridArgs G, int n, const int* __restrict__ cnt, const int* __restrict__ offs,
                                const int2* __restrict__ slots, int* __restrict__ pairs, int capacity) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int m = cnt[a], base = offs[a];
    if (m == 0) return;
    auto put = [&](int k, int u, int v) {
        const int o = base + k;
        if (o < capacity) {
            pairs[2 * (size_t)o] = u;
            pairs[2 * (size_t)o + 1] = v;
        }
    };
    if (m <= HB_GRID_SLOTS) {
        for (int k = 0; k < m; ++k) {
            const int2 q = slots[(size_t)k * n + a];
            put(k, q.x, q.y);
        }
    } else {
        (void)hb_grid_walk(G, a, put);
    }
}

struct HbBodies {
    const float4* P;      // (x, y, z, 1/m)
    const float4* V;      // (v, 0)
    const float4* Q;      // (w, x, y, z)
    const float4* omega;  // world angular velocity
    const float* radius;
    const float* tol;  // support-feature tolerance
    const int* vstart;
    const int* nv;
    const float4* verts;
};

// One thread per candidate pair (a, b) of the first min(*total, capacity):
// a touching pair (GJK/EPA flags valid and HIT, sep < 0, not both static)
// gets pf = (force on b, 1) and pp = contact point, and counts one contact
// for each body; any other pair gets pf.w = 0.
__global__ void k_hb_pairs(const int* __restrict__ pairs, int capacity, const int* __restrict__ total, HbBodies B,
                           const float* __restrict__ sep, const float* __restrict__ normal,
                           const int* __restrict__ flags, ContactDev c, float4* __restrict__ pf,
                           float4* __restrict__ pp, int* __restrict__ cnt, int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = *total;
    if (t == 0 && m > capacity) atomicOr(err, 1);
    if (t >= min(m, capacity)) return;
    const int a = pairs[2 * (size_t)t], b = pairs[2 * (size_t)t + 1];
    const float4 Pa = B.P[a], Pb = B.P[b];
    const float minv = Pa.w + Pb.w;
    const int fl = flags[t];
    const float s = sep[t];
    if (fl < 0 || !(fl & 1) || !(s < 0.0f) || !(minv > 0.0f)) {
        pf[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const V3 n = v3(normal[3 * (size_t)t], normal[3 * (size_t)t + 1], normal[3 * (size_t)t + 2]);
    const V3 Xa = v3(Pa.x, Pa.y, Pa.z), Xb = v3(Pb.x, Pb.y, Pb.z);
    const Rot Ra = quat_rot(B.Q[a]), Rb = quat_rot(B.Q[b]);
    float wa, wb;
    const V3 ca = feature_centroid(B.verts + B.vstart[a], B.nv[a], B.radius[a], B.tol[a], Ra, Xa, n, wa);
    const V3 cb = feature_centroid(B.verts + B.vstart[b], B.nv[b], B.radius[b], B.tol[b], Rb, Xb, scl(n, -1.0f),
                                   wb);
    const V3 pt = contact_point(ca, wa, cb, wb);
    const float4 va4 = B.V[a], vb4 = B.V[b], wa4 = B.omega[a], wb4 = B.omega[b];
    const V3 va = add(v3(va4.x, va4.y, va4.z), cross(v3(wa4.x, wa4.y, wa4.z), sub(pt, Xa)));
    const V3 vb = add(v3(vb4.x, vb4.y, vb4.z), cross(v3(wb4.x, wb4.y, wb4.z), sub(pt, Xb)));
    const V3 vr = sub(vb, va);
    const float vn = dot(vr, n);
    const float meff = 1.0f / minv;
    const float cn = c.c2z * sqrtf(c.kn * meff);
    const float fn = c.kn * (-s) - cn * vn;
    const V3 vt = sub(vr, scl(n, vn));
    const float vtn = sqrtf(dot(vt, vt));
    V3 f = scl(n, fn);
    if (vtn > 0.0f) {
        const float fmag = fminf(c.mu * fmaxf(fn, 0.0f), cn * vtn);
        f = sub(f, scl(vt, fmag / vtn));
    }
    pf[t] = make_float4(f.x, f.y, f.z, 1.0f);
    pp[t] = make_float4(pt.x, pt.y, pt.z, 0.0f);
    // static bodies (1/m = 0) do not move: their contacts are not listed
    if (Pa.w > 0.0f) atomicAdd(&cnt[a], 1);
    if (Pb.w > 0.0f) atomicAdd(&cnt[b], 1);
}

// Lists every touching pair under each of its moving bodies: entry 2 p (body
// a) or 2 p + 1 (body b) at offs[body] + a slot; slot order is arbitrary (the
// gather orders the entries).
__global__ void k_hb_fill(const int* __restrict__ pairs, int capacity, const int* __restrict__ total,
                          const float4* __restrict__ P, const float4* __restrict__ pf, const int* __restrict__ offs,
                          int* __restrict__ fillc, int* __restrict__ list) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= min(*total, capacity) || !(pf[t].w > 0.0f)) return;
    const int a = pairs[2 * (size_t)t], b = pairs[2 * (size_t)t + 1];
    if (P[a].w > 0.0f) list[offs[a] + atomicAdd(&fillc[a], 1)] = 2 * t;
    if (P[b].w > 0.0f) list[offs[b] + atomicAdd(&fillc[b], 1)] = 2 * t + 1;
}

// Per moving body: F = sum f, T = sum (p - X) x f over its contacts in
// ascending entry order (f = +pf on b, -pf on a), then the integration step.
// A static body has no listed contacts (F = T = 0) and is left as it is: a
// wall triangle under a whole layer of hulls would otherwise order thousands
// of entries on one thread.
__global__ void k_hb_gather_integrate(int n, const int* __restrict__ offs, const int* __restrict__ cnt,
                                      const int* __restrict__ list, const float4* __restrict__ pf,
                                      const float4* __restrict__ pp, float4* __restrict__ P, float4* __restrict__ V,
                                      float4* __restrict__ Q, float4* __restrict__ L, const float4* __restrict__ I0,
                                      const float4* __restrict__ I1, ContactDev c, float4* __restrict__ fout,
                                      float4* __restrict__ tout) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    float4 X = P[b];
    float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f, Tx = 0.0f, Ty = 0.0f, Tz = 0.0f;
    const int o = offs[b], k = cnt[b];
    int last = -1;
    for (int i = 0; i < k; ++i) {
        int e = 0x7fffffff;
        for (int j = 0; j < k; ++j) {
            const int x = list[o + j];
            if (x > last && x < e) e = x;
        }
        last = e;
        const int p = e >> 1;
        const float4 f4 = pf[p], q4 = pp[p];
        const float sg = (e & 1) ? 1.0f : -1.0f;
        const float fx = sg * f4.x, fy = sg * f4.y, fz = sg * f4.z;
        Fx = Fx + fx;
        Fy = Fy + fy;
        Fz = Fz + fz;
        const float rx = q4.x - X.x, ry = q4.y - X.y, rz = q4.z - X.z;
        Tx = Tx + (ry * fz - rz * fy);
        Ty = Ty + (rz * fx - rx * fz);
        Tz = Tz + (rx * fy - ry * fx);
    }
    if (fout) {
        fout[b] = make_float4(Fx, Fy, Fz, 0.0f);
        tout[b] = make_float4(Tx, Ty, Tz, 0.0f);
    }
    float4 v = V[b], q = Q[b], l = L[b];
    integrate_body(X, v, q, l, I0[b], I1[b], Fx, Fy, Fz, Tx, Ty, Tz, c);
    P[b] = X;
    V[b] = v;
    Q[b] = q;
    L[b] = l;
}

}  // namespace sd

using namespace sd;

struct sd_hull_ctx {
    sd_hull_params p{};
    ContactDev c{};
    hipStream_t stream = nullptr;
    int n = 0, n_verts = 0, nscenes = 0, capacity = 0;
    float4 *P = nullptr, *V = nullptr, *Q = nullptr, *L = nullptr, *I0 = nullptr, *I1 = nullptr;
    float4 *spos = nullptr, *omega = nullptr, *lo = nullptr, *hi = nullptr, *verts = nullptr;
    float *radius = nullptr, *ftol = nullptr;
    int *vstart = nullptr, *nv = nullptr, *scene_of = nullptr, *scene_first = nullptr;
    int* pairs = nullptr;
    // broadphase: per-scene grids, the bodies kept out of them (flag per body,
    // per-scene lists), cell keys (two buffers), cell starts, per-body pair
    // counts and offsets (the total at offs[n]), sort scratch
    HbGrid* grid = nullptr;
    int *large = nullptr, *llist = nullptr, *lstart = nullptr;
    long long ncells = 0;
    uint64_t *gkeys = nullptr, *gkeys2 = nullptr;
    int *cell_start = nullptr, *gcnt = nullptr, *goff = nullptr;
    int2* gslots = nullptr;  // HB_GRID_SLOTS pairs per body
    void* sort_tmp = nullptr;
    int *total = nullptr, *sap_error = nullptr;  // total = goff + n; sap_error: a zero word
    float *sep = nullptr, *normal = nullptr;
    int *flags = nullptr, *iters = nullptr;
    void* gjk_scratch = nullptr;
    float4 *pf = nullptr, *pp = nullptr, *F = nullptr, *T = nullptr;
    int *cnt = nullptr, *fillc = nullptr, *offs = nullptr, *list = nullptr, *err = nullptr;
    void* scan_tmp = nullptr;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_obj = nullptr;
    long long steps_done = 0;
    // learned warm start (sd_hull_set_warm): the caller's packed network, the
    // hull descriptors, the pair features and the predicted directions
    const void* model = nullptr;
    int warm_budget = 16;
    float *desc = nullptr, *X = nullptr, *wdir = nullptr;
    // changes whenever a captured step would be stale (upload, parameters,
    // warm start): a batch that captured this step captures it again
    long long generation = 0;
};

namespace {

template <typename T>
sd_status halloc(T** p, size_t count) {
    SD_HIP(hipMalloc((void**)p, (count ? count : 1) * sizeof(T)));
    return SD_OK;
}

void hb_drop_graph(sd_hull_ctx* c) {
    ++c->generation;
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->graph_obj) (void)hipGraphDestroy(c->graph_obj);
    c->graph = nullptr;
    c->graph_obj = nullptr;
}

void hb_free_warm(sd_hull_ctx* c) {
    for (float** q : {&c->desc, &c->X, &c->wdir}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
}

// The warm-start buffers of the uploaded bodies and their descriptors (the
// shapes are fixed at upload).
sd_status hb_alloc_warm(sd_hull_ctx* c) {
    hb_free_warm(c);
    sd_status st = halloc(&c->desc, 7 * (size_t)c->n);
    if (st == SD_OK) st = halloc(&c->X, 22 * (size_t)c->capacity);
    if (st == SD_OK) st = halloc(&c->wdir, 3 * (size_t)c->capacity);
    if (st == SD_OK) st = launch_hull_descriptors(c->spos, c->vstart, c->nv, c->verts, c->n, c->desc, c->stream);
    if (st == SD_OK) SD_HIP(hipStreamSynchronize(c->stream));
    if (st != SD_OK) hb_free_warm(c);
    return st;
}

void hb_free(sd_hull_ctx* c) {
    hb_drop_graph(c);
    hb_free_warm(c);
    void** ptrs[] = {(void**)&c->P,          (void**)&c->V,        (void**)&c->Q,          (void**)&c->L,
                     (void**)&c->I0,         (void**)&c->I1,       (void**)&c->spos,       (void**)&c->omega,
                     (void**)&c->lo,         (void**)&c->hi,       (void**)&c->verts,      (void**)&c->radius,       (void**)&c->ftol,
                     (void**)&c->vstart,     (void**)&c->nv,       (void**)&c->scene_of,   (void**)&c->scene_first,
                     (void**)&c->pairs,      (void**)&c->grid,     (void**)&c->gkeys,      (void**)&c->gkeys2,
                     (void**)&c->large,      (void**)&c->llist,    (void**)&c->lstart,
                     (void**)&c->cell_start, (void**)&c->gcnt,     (void**)&c->gslots,     (void**)&c->goff,       (void**)&c->sort_tmp,
                     (void**)&c->sap_error,  (void**)&c->sep,      (void**)&c->normal,     (void**)&c->flags,
                     (void**)&c->iters,      (void**)&c->gjk_scratch, (void**)&c->pf,      (void**)&c->pp,
                     (void**)&c->F,          (void**)&c->T,        (void**)&c->cnt,        (void**)&c->fillc,
                     (void**)&c->offs,       (void**)&c->list,     (void**)&c->err,        (void**)&c->scan_tmp};
    for (void** q : ptrs) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    c->total = nullptr;
    c->ncells = 0;
    c->n = 0;
}

sd_status hb_check_params(const sd_hull_params* p) {
    SD_REQUIRE(p, "sd_hull: NULL parameters");
    const sd_contact_params& k = p->contact;
    SD_REQUIRE(k.kn > 0.0f && k.dt > 0.0f && k.damping_ratio >= 0.0f && k.mu >= 0.0f,
               "sd_hull: kn and dt must be positive, damping_ratio and mu non-negative");
    SD_REQUIRE(p->max_iter >= 0 && p->max_iter <= 1024, "sd_hull: max_iter must be in [0, 1024]");
    SD_REQUIRE(p->pairs_per_body >= 0 && p->pairs_per_body <= 1024, "sd_hull: pairs_per_body must be in [0, 1024]");
    return SD_OK;
}

}  // namespace

namespace sd {

// One step of the context, enqueued on s (no host synchronisation).
sd_status hull_enqueue_step(sd_hull_ctx* c, hipStream_t s) {
    const int n = c->n, cap = c->capacity;
    const int mi = c->p.max_iter ? c->p.max_iter : 64;
    hipLaunchKernelGGL(k_hb_pose, dim3(div_up(n, 256)), dim3(256), 0, s, (const float4*)c->P, (const float4*)c->Q,
                       (const float4*)c->L, (const float4*)c->I0, (const float4*)c->I1, (const float*)c->radius, n,
                       c->spos, c->omega);
    SD_LAUNCH_CHECK();
    sd_status st = launch_hull_aabbs(c->spos, c->Q, c->vstart, c->nv, c->verts, n, c->lo, c->hi, s);
    if (st != SD_OK) return st;
    // candidate pairs: bodies sorted by grid cell (index order within a
    // cell), cell starts, then per body its count, the offsets and the pairs
    hipLaunchKernelGGL(k_hb_cell_keys, dim3(div_up(n, 256)), dim3(256), 0, s, (const float4*)c->P,
                       (const int*)c->scene_of, (const HbGrid*)c->grid, (const int*)c->large, n, (int)c->ncells,
                       c->gkeys);
    SD_LAUNCH_CHECK();
    uint64_t *K = c->gkeys, *Ka = c->gkeys2;
    int cbits = 1;  // cell ids run to ncells (the bodies kept out of the grids)
    while (cbits < 32 && (1ll << cbits) <= c->ncells) ++cbits;
    st = radix_sort_keys_u64(&K, &Ka, n, 32, 32 + cbits, c->sort_tmp, radix_temp_bytes(n), s, 0);
    if (st == SD_OK)
        st = launch_step_starts_u64(K, n, (int)c->ncells, c->cell_start, nullptr, false, nullptr, nullptr, nullptr,
                                    s);
    if (st != SD_OK) return st;
    const HbGridArgs G{(const float4*)c->P,     (const float4*)c->lo,     (const float4*)c->hi,
                       (const int*)c->scene_of, (const HbGrid*)c->grid,   (const int*)c->large,
                       (const int*)c->llist,    (const int*)c->lstart,    (const uint64_t*)K,
                       (const int*)c->cell_start};
    hipLaunchKernelGGL(k_hb_grid_pairs, dim3(div_up(n, 256)), dim3(256), 0, s, G, n, c->gcnt, c->gslots);
    SD_LAUNCH_CHECK();
    st = exclusive_scan_i32(c->gcnt, c->goff, n, c->goff + n, c->scan_tmp, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_hb_grid_place, dim3(div_up(n, 256)), dim3(256), 0, s, G, n, (const int*)c->gcnt,
                       (const int*)c->goff, (const int2*)c->gslots, c->pairs, cap);
    SD_LAUNCH_CHECK();
    if (c->model) {
        st = enqueue_warm_dirs(c->pairs, cap, c->total, c->spos, c->Q, c->desc, c->X, c->model, c->wdir, s);
        if (st == SD_OK)
            st = enqueue_gjk_warm(c->pairs, cap, c->total, c->spos, c->Q, c->vstart, c->nv, c->verts, n, c->n_verts,
                                  c->wdir, std::min(c->warm_budget, mi), mi, HB_WARM_NEAR_MARGIN, c->sep, c->normal,
                                  c->flags, c->iters, c->gjk_scratch, s);
    } else {
        st = enqueue_gjk_cold(c->pairs, cap, c->total, c->spos, c->Q, c->vstart, c->nv, c->verts, n, c->n_verts, mi,
                              0.0f, c->sep, c->normal, c->flags, c->iters, c->gjk_scratch, s);
    }
    if (st != SD_OK) return st;
    st = launch_fill_u32((uint32_t*)c->cnt, 0u, n, s);
    if (st == SD_OK) st = launch_fill_u32((uint32_t*)c->fillc, 0u, n, s);
    if (st != SD_OK) return st;
    const HbBodies B{c->P, c->V, c->Q, c->omega, c->radius, c->ftol, c->vstart, c->nv, c->verts};
    hipLaunchKernelGGL(k_hb_pairs, dim3(div_up(cap, 256)), dim3(256), 0, s, (const int*)c->pairs, cap,
                       (const int*)c->total, B, (const float*)c->sep, (const float*)c->normal, (const int*)c->flags,
                       c->c, c->pf, c->pp, c->cnt, c->err);
    SD_LAUNCH_CHECK();
    st = exclusive_scan_i32(c->cnt, c->offs, n, nullptr, c->scan_tmp, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_hb_fill, dim3(div_up(cap, 256)), dim3(256), 0, s, (const int*)c->pairs, cap,
                       (const int*)c->total, (const float4*)c->P, (const float4*)c->pf, (const int*)c->offs, c->fillc,
                       c->list);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hb_gather_integrate, dim3(div_up(n, 256)), dim3(256), 0, s, n, (const int*)c->offs,
                       (const int*)c->cnt, (const int*)c->list, (const float4*)c->pf, (const float4*)c->pp, c->P,
                       c->V, c->Q, c->L, (const float4*)c->I0, (const float4*)c->I1, c->c, c->F, c->T);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// Errors recorded on the device by the steps since the last check: the pair
// capacity overflowed, or a scene range was malformed.
sd_status hull_check_errors(sd_hull_ctx* c) {
    int h[2] = {0, 0};
    SD_HIP(hipMemcpyAsync(&h[0], c->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipMemcpyAsync(&h[1], c->sap_error, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    if (h[1]) return set_error(SD_ERR_ARG, "sd_hull_step: bad scene ranges (bits %d)", h[1]);
    if (h[0])
        return set_error(SD_ERR_CAPACITY,
                         "sd_hull_step: more candidate pairs than the capacity (%d); the state is invalid, upload "
                         "again with a larger pairs_per_body",
                         c->capacity);
    return SD_OK;
}

int hull_body_count(const sd_hull_ctx* c) { return c->n; }
long long hull_generation(const sd_hull_ctx* c) { return c->generation; }
float hull_dt(const sd_hull_ctx* c) { return c->p.contact.dt; }
void hull_count_steps(sd_hull_ctx* c, int nsteps) { c->steps_done += nsteps; }

}  // namespace sd

namespace {

// Bounding radius of a body about its centre: the largest vertex distance
// (body frame), or the sphere radius of a body without vertices.
float hb_bound(const float* verts4, int nv, float radius) {
    if (nv == 0) return radius;
    float r = 0.0f;
    for (int k = 0; k < nv; ++k) {
        const float* v = verts4 + 4 * (size_t)k;
        r = fmaxf(r, sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    return r;
}

constexpr int HB_GRID_AXIS_MAX = 512;
constexpr long long HB_GRID_CELLS_MAX = 1ll << 22;

// Fits a grid of cell `cell` (twice it, inflated by 2^-10) over [lo, hi]: the
// cell grows by 1.25 until no axis has more than 512 cells and the grid no
// more than 2^22; writes 1 / cell and the cells per axis, returns the cells.
long long hb_grid_fit(float cell, const float* lo, const float* hi, float& inv_out, int* d) {
    cell = 2.0f * cell * 1.0009765625f;
    if (!(cell > 0.0f)) cell = 1e-3f;
    for (;;) {
        const float inv = 1.0f / cell;
        long long cells = 1;
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const float e = ceilf((hi[a] - lo[a]) * inv);
            d[a] = e >= 1.0f ? (e <= (float)HB_GRID_AXIS_MAX ? (int)e : HB_GRID_AXIS_MAX + 1) : 1;
            ok = ok && d[a] <= HB_GRID_AXIS_MAX;
            cells *= d[a];
        }
        if (ok && cells <= HB_GRID_CELLS_MAX) {
            inv_out = inv;
            return cells;
        }
        cell = cell * 1.25f;
    }
}

int hb_cell_axis_host(float x, float lo, float inv, int nd) {
    return (int)fminf(fmaxf(floorf((x - lo) * inv), 0.0f), (float)(nd - 1));
}

// The grids of every scene from the initial bodies: rd = the largest bounding
// radius of its moving bodies (of all its bodies when none moves); a body
// whose bounding radius exceeds rd is large (static) and goes to the second
// grid. Each grid: cell = twice the largest bounding radius of its bodies
// (times 1 + 2^-10) over the box of their bounding spheres, grown by 1.25
// until no axis has more than 512 cells and the grid no more than 2^22. The
// large bodies never move, so they are bucketed here: llist sorted by (second
// grid cell, index), lstart its cell starts. Returns the cells of the first
// grids; nlarge_cells receives those of the second.
long long hb_grids(const float* pos4, const float* radius, const int* vstart, const int* nv, const float* verts4,
                   const std::vector<int>& sfirst, std::vector<HbGrid>& grids, std::vector<int>& large,
                   std::vector<int>& llist, std::vector<int>& lstart, long long& nlarge_cells) {
    const int nscenes = (int)sfirst.size() - 1;
    grids.assign(nscenes, HbGrid{});
    large.assign(sfirst[nscenes], 0);
    std::vector<float> bound(sfirst[nscenes]);
    for (int b = 0; b < sfirst[nscenes]; ++b) bound[b] = hb_bound(verts4 + 4 * (size_t)vstart[b], nv[b], radius[b]);
    long long cbase = 0, lbase = 0;
    std::vector<std::pair<long long, int>> lkeys;  // (second grid cell, body)
    for (int s = 0; s < nscenes; ++s) {
        float rd = 0.0f, ra = 0.0f, rl = 0.0f;
        bool moving = false;
        for (int b = sfirst[s]; b < sfirst[s + 1]; ++b) {
            ra = fmaxf(ra, bound[b]);
            if (pos4[4 * (size_t)b + 3] > 0.0f) {
                rd = fmaxf(rd, bound[b]);
                moving = true;
            }
        }
        if (!moving) rd = ra;
        float lo[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}, hi[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        bool first[2] = {true, true};
        for (int b = sfirst[s]; b < sfirst[s + 1]; ++b) {
            const float r = bound[b];
            const int k = large[b] = r > rd ? 1 : 0;
            if (k) rl = fmaxf(rl, r);
            for (int a = 0; a < 3; ++a) {
                const float x = pos4[4 * (size_t)b + a];
                lo[k][a] = first[k] ? x - r : fminf(lo[k][a], x - r);
                hi[k][a] = first[k] ? x + r : fmaxf(hi[k][a], x + r);
            }
            first[k] = false;
        }
        HbGrid& g = grids[s];
        int d[3], e[3];
        float inv, linv;
        const long long cells = hb_grid_fit(rd, lo[0], hi[0], inv, d);
        const long long lcells = hb_grid_fit(rl, lo[1], hi[1], linv, e);
        g = HbGrid{lo[0][0], lo[0][1], lo[0][2], inv,  d[0], d[1], d[2], (int)cbase,
                   lo[1][0], lo[1][1], lo[1][2], linv, e[0], e[1], e[2], (int)lbase};
        for (int b = sfirst[s]; b < sfirst[s + 1]; ++b) {
            if (!large[b]) continue;
            const int cx = hb_cell_axis_host(pos4[4 * (size_t)b], g.llox, g.linv, g.lnx);
            const int cy = hb_cell_axis_host(pos4[4 * (size_t)b + 1], g.lloy, g.linv, g.lny);
            const int cz = hb_cell_axis_host(pos4[4 * (size_t)b + 2], g.lloz, g.linv, g.lnz);
            lkeys.emplace_back(lbase + ((long long)cz * g.lny + cy) * g.lnx + cx, b);
        }
        cbase += cells;
        lbase += lcells;
    }
    std::sort(lkeys.begin(), lkeys.end());
    llist.resize(lkeys.size());
    lstart.assign(lbase + 1, 0);
    for (size_t i = 0; i < lkeys.size(); ++i) {
        llist[i] = lkeys[i].second;
        ++lstart[lkeys[i].first + 1];
    }
    for (long long c = 0; c < lbase; ++c) lstart[c + 1] += lstart[c];
    nlarge_cells = lbase;
    return cbase;
}

}  // namespace

extern "C" sd_status sd_hull_create(const sd_hull_params* p, sd_hull_ctx** out) {
    SD_REQUIRE(out, "sd_hull_create: NULL out");
    *out = nullptr;
    sd_status st = hb_check_params(p);
    if (st != SD_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(SD_ERR_NO_DEVICE, "sd_hull_create: no HIP device");
    sd_hull_ctx* c = new sd_hull_ctx();
    c->p = *p;
    c->c = to_dev(p->contact);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_error(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    *out = c;
    return SD_OK;
}

extern "C" sd_status sd_hull_set_params(sd_hull_ctx* c, const sd_hull_params* p) {
    SD_REQUIRE(c, "sd_hull_set_params: NULL context");
    sd_status st = hb_check_params(p);
    if (st != SD_OK) return st;
    SD_REQUIRE(c->n == 0 || p->pairs_per_body == c->p.pairs_per_body,
               "sd_hull_set_params: pairs_per_body cannot change after upload");
    (void)hipStreamSynchronize(c->stream);
    hb_drop_graph(c);
    c->p = *p;
    c->c = to_dev(p->contact);
    return SD_OK;
}

extern "C" sd_status sd_hull_destroy(sd_hull_ctx* c) {
    if (!c) return SD_OK;
    (void)hipStreamSynchronize(c->stream);
    hb_free(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return SD_OK;
}

extern "C" sd_status sd_hull_upload(sd_hull_ctx* c, int n, const float* pos4, const float* vel4, const float* quat4,
                                    const float* angvel4, const float* inertia6, const float* radius,
                                    const int* vstart, const int* nv, int n_verts, const float* verts4,
                                    const int* scene_first, int nscenes) {
    SD_REQUIRE(c, "sd_hull_upload: NULL context");
    SD_REQUIRE(n >= 1 && n_verts >= 0 && nscenes >= 1, "sd_hull_upload: bad sizes");
    SD_REQUIRE(pos4 && vel4 && quat4 && angvel4 && inertia6 && radius && vstart && nv && (verts4 || n_verts == 0),
               "sd_hull_upload: NULL array");
    SD_REQUIRE(nscenes == 1 || scene_first, "sd_hull_upload: scene_first is required for more than one scene");
    std::vector<int> sfirst(nscenes + 1), sof(n);
    if (scene_first) sfirst.assign(scene_first, scene_first + nscenes + 1);
    else sfirst = {0, n};
    SD_REQUIRE(sfirst[0] == 0 && sfirst[nscenes] == n, "sd_hull_upload: scene_first must run from 0 to n");
    for (int k = 0; k < nscenes; ++k) {
        SD_REQUIRE(sfirst[k + 1] >= sfirst[k], "sd_hull_upload: scene_first must not decrease");
        SD_REQUIRE(sfirst[k + 1] - sfirst[k] <= 65536, "sd_hull_upload: scene %d has more than 65,536 bodies", k);
        for (int i = sfirst[k]; i < sfirst[k + 1]; ++i) sof[i] = k;
    }
    for (int i = 0; i < n; ++i) {
        SD_REQUIRE(vstart[i] >= 0 && nv[i] >= 0 && vstart[i] + nv[i] <= n_verts,
                   "sd_hull_upload: body %d has a bad vertex range", i);
        SD_REQUIRE(nv[i] > 0 || radius[i] > 0.0f, "sd_hull_upload: body %d has neither vertices nor a radius", i);
        SD_REQUIRE(nv[i] <= GJK_MAX_VERTS_PER_HULL, "sd_hull_upload: body %d has %d vertices (at most %d)", i, nv[i],
                   GJK_MAX_VERTS_PER_HULL);
        const float q0 = quat4[4 * (size_t)i], q1 = quat4[4 * (size_t)i + 1], q2 = quat4[4 * (size_t)i + 2],
                    q3 = quat4[4 * (size_t)i + 3];
        SD_REQUIRE(fabsf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3 - 1.0f) < 1e-4f,
                   "sd_hull_upload: body %d has a quaternion that is not of unit length", i);
    }
    std::vector<float4> P(n), V(n), Q(n), L(n), I0(n), I1(n);
    sd_status st = prepare_rigid("sd_hull_upload", n, pos4, vel4, quat4, angvel4, inertia6, P.data(), V.data(),
                                 Q.data(), L.data(), I0.data(), I1.data());
    if (st != SD_OK) return st;
    (void)hipStreamSynchronize(c->stream);
    hb_free(c);
    const int ppb = c->p.pairs_per_body ? c->p.pairs_per_body : 32;
    const long long capl = std::max(64ll, (long long)ppb * n);
    SD_REQUIRE(capl < (1ll << 30), "sd_hull_upload: pair capacity %lld too large", capl);
    const int cap = (int)capl;
    std::vector<HbGrid> grids;
    std::vector<int> large, llist, lstart;
    long long nlarge_cells = 0;
    const long long ncells = hb_grids(pos4, radius, vstart, nv, verts4, sfirst, grids, large, llist, lstart,
                                      nlarge_cells);
    SD_REQUIRE(ncells < (1ll << 31) - 1 && nlarge_cells < (1ll << 31) - 1,
               "sd_hull_upload: %lld + %lld broadphase cells", ncells, nlarge_cells);
    c->n = n;
    c->n_verts = n_verts;
    c->nscenes = nscenes;
    c->capacity = cap;
    c->ncells = ncells;
#define ALLOC(ptr, cnt) \
    if (st == SD_OK) st = halloc(&(ptr), (size_t)(cnt))
    ALLOC(c->P, n);
    ALLOC(c->V, n);
    ALLOC(c->Q, n);
    ALLOC(c->L, n);
    ALLOC(c->I0, n);
    ALLOC(c->I1, n);
    ALLOC(c->spos, n);
    ALLOC(c->omega, n);
    ALLOC(c->lo, n);
    ALLOC(c->hi, n);
    ALLOC(c->verts, n_verts);
    ALLOC(c->radius, n);
    ALLOC(c->ftol, n);
    ALLOC(c->vstart, n);
    ALLOC(c->nv, n);
    ALLOC(c->scene_of, n);
    ALLOC(c->scene_first, nscenes + 1);
    ALLOC(c->pairs, 2 * (size_t)cap);
    ALLOC(c->grid, nscenes);
    ALLOC(c->large, n);
    ALLOC(c->llist, llist.size());
    ALLOC(c->lstart, lstart.size());
    ALLOC(c->gkeys, n);
    ALLOC(c->gkeys2, n);
    ALLOC(c->cell_start, ncells + 1);
    ALLOC(c->gcnt, n);
    ALLOC(c->gslots, (size_t)n * HB_GRID_SLOTS);
    ALLOC(c->goff, n + 1);
    ALLOC(*(char**)&c->sort_tmp, radix_temp_bytes(n));
    ALLOC(c->sap_error, 1);
    ALLOC(c->sep, cap);
    ALLOC(c->normal, 3 * (size_t)cap);
    ALLOC(c->flags, cap);
    ALLOC(c->iters, cap);
    ALLOC(*(char**)&c->gjk_scratch, gjk_cold_scratch_bytes(cap));
    ALLOC(c->pf, cap);
    ALLOC(c->pp, cap);
    ALLOC(c->F, n);
    ALLOC(c->T, n);
    ALLOC(c->cnt, n);
    ALLOC(c->fillc, n);
    ALLOC(c->offs, n + 1);
    ALLOC(c->list, 2 * (size_t)cap);
    ALLOC(c->err, 1);
    ALLOC(*(char**)&c->scan_tmp, scan_temp_bytes(n));
#undef ALLOC
    if (st != SD_OK) {
        hb_free(c);
        return st;
    }
    hipStream_t s = c->stream;
    std::vector<float4> vv(std::max(n_verts, 1));
    for (int k = 0; k < n_verts; ++k)
        vv[k] = make_float4(verts4[4 * (size_t)k], verts4[4 * (size_t)k + 1], verts4[4 * (size_t)k + 2], 0.0f);
    SD_HIP(hipMemcpyAsync(c->P, P.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->V, V.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->Q, Q.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->L, L.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->I0, I0.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->I1, I1.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    if (n_verts) SD_HIP(hipMemcpyAsync(c->verts, vv.data(), n_verts * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->radius, radius, n * sizeof(float), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->vstart, vstart, n * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->nv, nv, n * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->scene_of, sof.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->scene_first, sfirst.data(), (nscenes + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->grid, grids.data(), nscenes * sizeof(HbGrid), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->large, large.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    if (!llist.empty())
        SD_HIP(hipMemcpyAsync(c->llist, llist.data(), llist.size() * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->lstart, lstart.data(), lstart.size() * sizeof(int), hipMemcpyHostToDevice, s));
    c->total = c->goff + n;
    hipLaunchKernelGGL(k_hb_feature_tol, dim3(div_up(n, 256)), dim3(256), 0, s, (const float4*)c->verts,
                       (const int*)c->vstart, (const int*)c->nv, n, c->ftol);
    SD_LAUNCH_CHECK();
    st = launch_fill_u32((uint32_t*)c->err, 0u, 1, s);
    if (st == SD_OK) st = launch_fill_u32((uint32_t*)c->sap_error, 0u, 1, s);
    if (st != SD_OK) return st;
    // spos (centre, radius) for the descriptors, as the step's pose stage writes it
    std::vector<float4> sp(n);
    for (int i = 0; i < n; ++i) sp[i] = make_float4(P[i].x, P[i].y, P[i].z, radius[i]);
    SD_HIP(hipMemcpyAsync(c->spos, sp.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipStreamSynchronize(s));
    c->steps_done = 0;
    if (c->model) {
        st = hb_alloc_warm(c);
        if (st != SD_OK) {
            hb_free(c);
            return st;
        }
    }
    return SD_OK;
}

extern "C" sd_status sd_hull_set_warm(sd_hull_ctx* c, const void* packed_model, int warm_budget) {
    SD_REQUIRE(c, "sd_hull_set_warm: NULL context");
    SD_REQUIRE(packed_model == nullptr || (warm_budget >= 1 && warm_budget <= 1024),
               "sd_hull_set_warm: warm_budget must be in [1, 1024]");
    (void)hipStreamSynchronize(c->stream);
    hb_drop_graph(c);
    c->model = packed_model;
    c->warm_budget = packed_model ? warm_budget : 16;
    if (!packed_model) {
        hb_free_warm(c);
        return SD_OK;
    }
    return c->n ? hb_alloc_warm(c) : SD_OK;
}

extern "C" sd_status sd_hull_step(sd_hull_ctx* c, int nsteps) {
    SD_REQUIRE(c && nsteps >= 0, "sd_hull_step: bad argument");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_hull_step: no bodies uploaded");
    for (int i = 0; i < nsteps; ++i) {
        if (!c->p.use_graph) {
            sd_status st = hull_enqueue_step(c, c->stream);
            if (st != SD_OK) return st;
            continue;
        }
        if (!c->graph) {
            SD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            sd_status st = hull_enqueue_step(c, c->stream);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (st != SD_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            SD_HIP(e);
            hipGraphExec_t ge = nullptr;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                SD_HIP(e);
            }
            c->graph = ge;
            c->graph_obj = g;
        }
        SD_HIP(hipGraphLaunch(c->graph, c->stream));
    }
    sd_status st = hull_check_errors(c);
    c->steps_done += nsteps;
    return st;
}

extern "C" sd_status sd_hull_download(sd_hull_ctx* c, float* pos4, float* vel4, float* quat4, float* angmom4) {
    SD_REQUIRE(c, "sd_hull_download: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_hull_download: no bodies uploaded");
    const size_t bytes = (size_t)c->n * sizeof(float4);
    hipStream_t s = c->stream;
    if (pos4) SD_HIP(hipMemcpyAsync(pos4, c->P, bytes, hipMemcpyDeviceToHost, s));
    if (vel4) SD_HIP(hipMemcpyAsync(vel4, c->V, bytes, hipMemcpyDeviceToHost, s));
    if (quat4) SD_HIP(hipMemcpyAsync(quat4, c->Q, bytes, hipMemcpyDeviceToHost, s));
    if (angmom4) SD_HIP(hipMemcpyAsync(angmom4, c->L, bytes, hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    if (pos4)
        for (int b = 0; b < c->n; ++b) pos4[4 * (size_t)b + 3] = 0.0f;
    return SD_OK;
}

extern "C" sd_status sd_hull_contacts(sd_hull_ctx* c, int* pairs, float* force4, float* point4, int capacity,
                                      int* n_contacts) {
    SD_REQUIRE(c && n_contacts, "sd_hull_contacts: NULL");
    if (c->n == 0 || c->steps_done == 0) return set_error(SD_ERR_STATE, "sd_hull_contacts: no step has run");
    int m = 0;
    SD_HIP(hipMemcpyAsync(&m, c->total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    m = std::min(m, c->capacity);
    std::vector<int> pr(2 * (size_t)std::max(m, 1));
    std::vector<float4> f(std::max(m, 1)), q(std::max(m, 1));
    if (m) {
        SD_HIP(hipMemcpyAsync(pr.data(), c->pairs, 2 * (size_t)m * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SD_HIP(hipMemcpyAsync(f.data(), c->pf, (size_t)m * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        SD_HIP(hipMemcpyAsync(q.data(), c->pp, (size_t)m * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        SD_HIP(hipStreamSynchronize(c->stream));
    }
    int k = 0;
    for (int t = 0; t < m; ++t) {
        if (!(f[t].w > 0.0f)) continue;
        if (k < capacity) {
            if (pairs) {
                pairs[2 * (size_t)k] = pr[2 * (size_t)t];
                pairs[2 * (size_t)k + 1] = pr[2 * (size_t)t + 1];
            }
            if (force4) memcpy(force4 + 4 * (size_t)k, &f[t], sizeof(float4));
            if (point4) memcpy(point4 + 4 * (size_t)k, &q[t], sizeof(float4));
        }
        ++k;
    }
    *n_contacts = k;
    if (k > capacity) return set_error(SD_ERR_CAPACITY, "sd_hull_contacts: %d contacts", k);
    return SD_OK;
}

extern "C" sd_status sd_hull_narrowphase_stats(sd_hull_ctx* c, int* n_pairs, int* n_hits, int* n_warm_used,
                                               long long* iters_total) {
    SD_REQUIRE(c, "sd_hull_narrowphase_stats: NULL");
    if (c->n == 0 || c->steps_done == 0) return set_error(SD_ERR_STATE, "sd_hull_narrowphase_stats: no step has run");
    int m = 0;
    SD_HIP(hipMemcpyAsync(&m, c->total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    m = std::min(m, c->capacity);
    std::vector<int> fl(std::max(m, 1)), it(std::max(m, 1));
    if (m) {
        SD_HIP(hipMemcpyAsync(fl.data(), c->flags, (size_t)m * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SD_HIP(hipMemcpyAsync(it.data(), c->iters, (size_t)m * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SD_HIP(hipStreamSynchronize(c->stream));
    }
    int hits = 0, warm = 0;
    long long total = 0;
    for (int t = 0; t < m; ++t) {
        hits += (fl[t] & GJK_HIT) ? 1 : 0;
        warm += (fl[t] & GJK_WARM_USED) ? 1 : 0;
        total += it[t];
    }
    if (n_pairs) *n_pairs = m;
    if (n_hits) *n_hits = hits;
    if (n_warm_used) *n_warm_used = warm;
    if (iters_total) *iters_total = total;
    return SD_OK;
}
