// Convex-hull broadphase and narrowphase.
//
// Sweep-and-prune replaces SAPCollision::processCollisions and its kernels
// (src/engine/cuda/collisions/SAP/sap_detection.cu:135-706): AABB, projection,
// sort of the lower extents, windowed sweep. Differences by design: the sweep
// axis is the one of largest centre variance (the reference's TODO at
// sap_detection.cu:589), float keys are mapped order-preservingly to u32
// (quirk Q3), the sweep stops at the first lower extent beyond the upper one
// and tests all three axes (quirk Q5), and the pair list is exact and complete
// (count, scan, write) instead of 16 fixed slots per object.
//
// GJK/EPA replaces gjkCUDA (sandbox/testGJK.cu:237-315); see sd_gjk_dev.hpp.
#include "sd_epa_wave.hpp"
#include "sd_gjk_dev.hpp"
#include "sd_internal.hpp"

namespace sd {

// ------------------------------------------------------------------ AABBs

__global__ void k_hull_aabb(const float4* __restrict__ pos, const float4* __restrict__ quat,
                            const int* __restrict__ vstart, const int* __restrict__ nv,
                            const float4* __restrict__ verts, int n, float4* __restrict__ lo4,
                            float4* __restrict__ hi4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = pos[i];
    const Rot R = quat_rot(quat[i]);
    const int k0 = vstart[i], kn = nv[i];
    float lx, ly, lz, hx, hy, hz;
    if (kn == 0) {  // sphere of radius c.w
        lx = c.x - c.w;
        ly = c.y - c.w;
        lz = c.z - c.w;
        hx = c.x + c.w;
        hy = c.y + c.w;
        hz = c.z + c.w;
    } else {
        lx = ly = lz = 3.402823466e38f;
        hx = hy = hz = -3.402823466e38f;
        // 8 vertices per step with their loads in flight; indices past the
        // last vertex repeat it (min and max do not change)
        for (int k = 0; k < kn; k += 8) {
            float4 p[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) p[j] = verts[k0 + min(k + j, kn - 1)];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const V3 w = rot_mul(R, v3(p[j].x, p[j].y, p[j].z));
                lx = fminf(lx, w.x);
                ly = fminf(ly, w.y);
                lz = fminf(lz, w.z);
                hx = fmaxf(hx, w.x);
                hy = fmaxf(hy, w.y);
                hz = fmaxf(hz, w.z);
            }
        }
        lx = lx + c.x;
        ly = ly + c.y;
        lz = lz + c.z;
        hx = hx + c.x;
        hy = hy + c.y;
        hz = hz + c.z;
    }
    lo4[i] = make_float4(lx, ly, lz, 0.0f);
    hi4[i] = make_float4(hx, hy, hz, 0.0f);
}

// ------------------------------------------------------------------ axis

constexpr int AX_BLOCK = 256;
constexpr int AX_BLOCKS = 256;

// The sweep axis is the one of largest variance of the box centres, computed
// in FP32 in a fixed order: centres relative to the first box's centre (which
// keeps the sums well conditioned), per-thread sums over a grid stride of nb
// blocks of 256 threads, a tree over each block's 256 sums (halving strides),
// the block sums added in block order, var = S2 / n - (S1 / n)^2 per axis,
// the first axis of largest variance. The order is fixed for a given n, so
// the axis does not depend on the launch.
__device__ __forceinline__ void box_centre(const float4& l, const float4& h, float& cx, float& cy, float& cz) {
    cx = 0.5f * (l.x + h.x);
    cy = 0.5f * (l.y + h.y);
    cz = 0.5f * (l.z + h.z);
}

// Tree sum of red[k][0..AX_BLOCK) into red[k][0] (halving strides).
__device__ __forceinline__ void axis_tree(float (*red)[AX_BLOCK]) {
    __syncthreads();
    for (int h = AX_BLOCK / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
}

__device__ __forceinline__ int axis_of_sums(const float* s, int n) {
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

__global__ __launch_bounds__(AX_BLOCK) void k_axis_partial(const float4* __restrict__ lo4,
                                                           const float4* __restrict__ hi4, int n,
                                                           float* __restrict__ part) {
    __shared__ float red[6][AX_BLOCK];
    float rx, ry, rz;
    box_centre(lo4[0], hi4[0], rx, ry, rz);
    float acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * AX_BLOCK + threadIdx.x; i < n; i += gridDim.x * AX_BLOCK) {
        float cx, cy, cz;
        box_centre(lo4[i], hi4[i], cx, cy, cz);
        const float dx = cx - rx, dy = cy - ry, dz = cz - rz;
        acc[0] += dx;
        acc[1] += dy;
        acc[2] += dz;
        acc[3] += dx * dx;
        acc[4] += dy * dy;
        acc[5] += dz * dz;
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = acc[k];
    axis_tree(red);
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Adds the partials in block order and picks the axis. The partials are
// staged in LDS by the whole block first (one coalesced read), so the six
// serial sums read LDS, not a chain of dependent global loads.
__global__ __launch_bounds__(AX_BLOCK) void k_axis_pick(const float* __restrict__ part, int nparts, int n,
                                                        int* __restrict__ axis) {
    __shared__ float sp[AX_BLOCKS * 6];
    __shared__ float sum[6];
    for (int i = threadIdx.x; i < nparts * 6; i += AX_BLOCK) sp[i] = part[i];
    __syncthreads();
    if (threadIdx.x < 6) {
        float acc = 0.0f;
        for (int b = 0; b < nparts; ++b) acc += sp[b * 6 + threadIdx.x];
        sum[threadIdx.x] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) *axis = axis_of_sums(sum, n);
}

// Order-preserving map of a float to u32 (negative values flip all bits).
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float axis_of(const float4& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

__global__ void k_sap_keys(const float4* __restrict__ lo4, int n, const int* __restrict__ axis,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = float_key(axis_of(lo4[i], *axis));
    vals[i] = (uint32_t)i;
}

__global__ void k_sap_gather(const uint32_t* __restrict__ order, int n, const float4* __restrict__ lo4,
                             const float4* __restrict__ hi4, const int* __restrict__ axis,
                             float4* __restrict__ slo, float4* __restrict__ shi) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = order[t];
    const int a = *axis;
    float4 l = lo4[i], h = hi4[i];
    // store the sweep axis first: (sweep, other1, other2, id)
    const float l0 = axis_of(l, a), h0 = axis_of(h, a);
    const float l1 = axis_of(l, (a + 1) % 3), h1 = axis_of(h, (a + 1) % 3);
    const float l2 = axis_of(l, (a + 2) % 3), h2 = axis_of(h, (a + 2) % 3);
    slo[t] = make_float4(l0, l1, l2, __int_as_float((int)i));
    shi[t] = make_float4(h0, h1, h2, 0.0f);
}

// Sweep over the objects sorted by lower extent on the sweep axis (slo/shi:
// sweep axis first, then the other two): the number of overlapping partners
// after t, and the sorted positions of the first SAP_SLOTS of them. A partner
// search stops at the first lower extent beyond t's upper one (and, with
// SCENES, at the first object of another scene: shi[].w holds the scene). A
// block stages consecutive SAP_TILE-object slices of the sorted extents in LDS
// and its threads scan them from there, until no thread of the block is still
// searching; a thread tests SAP_CHUNK candidates per step with all their loads
// in flight at once.
constexpr int SAP_TILE = 256;
constexpr int SAP_CHUNK = 8;
constexpr int SAP_SLOTS = 16;

__device__ __forceinline__ bool sap_stop(const float4& lu, const float4& hu, const float4& h, int sc, bool scenes) {
    return (scenes & (__float_as_int(hu.w) != sc)) | (lu.x > h.x);
}
__device__ __forceinline__ bool sap_overlap(const float4& lu, const float4& hu, const float4& l, const float4& h) {
    return (lu.y <= h.y) & (l.y <= hu.y) & (lu.z <= h.z) & (l.z <= hu.z);
}

template <bool SCENES>
__global__ __launch_bounds__(SAP_TILE) void k_sap_sweep(const float4* __restrict__ slo,
                                                        const float4* __restrict__ shi, int n,
                                                        int* __restrict__ counts, int* __restrict__ slots,
                                                        int* __restrict__ ovf, int* __restrict__ ovf_count) {
    __shared__ float4 tl[SAP_TILE], th[SAP_TILE];
    const int t0 = blockIdx.x * SAP_TILE;
    const int t = t0 + (int)threadIdx.x;
    float4 l = make_float4(0.0f, 0.0f, 0.0f, 0.0f), h = l;
    bool active = t < n;
    int cnt = 0;
    if (active) {
        l = slo[t];
        h = shi[t];
    }
    const int sc = __float_as_int(h.w);
    int* my_slots = slots + (size_t)(active ? t : 0) * SAP_SLOTS;
    for (int base = t0 + 1; base < n; base += SAP_TILE) {
        if (!__syncthreads_or(active)) break;  // also: every thread is done with the previous slice
        const int u = base + (int)threadIdx.x;
        if (u < n) {
            tl[threadIdx.x] = slo[u];
            th[threadIdx.x] = shi[u];
        }
        __syncthreads();
        if (active) {
            const int kend = min(SAP_TILE, n - base);
            for (int k = max(0, t + 1 - base); k < kend; k += SAP_CHUNK) {
                float4 lu[SAP_CHUNK], hu[SAP_CHUNK];
#pragma unroll
                for (int i = 0; i < SAP_CHUNK; ++i) {
                    const int kk = min(k + i, SAP_TILE - 1);
                    lu[i] = tl[kk];
                    hu[i] = th[kk];
                }
                uint32_t stop = 0u, hit = 0u;
#pragma unroll
                for (int i = 0; i < SAP_CHUNK; ++i) {  // bitwise: no branch per candidate
                    const bool in = k + i < kend;
                    stop |= (uint32_t)(in & sap_stop(lu[i], hu[i], h, sc, SCENES)) << i;
                    hit |= (uint32_t)(in & sap_overlap(lu[i], hu[i], l, h)) << i;
                }
                if (stop) hit &= (1u << __builtin_ctz(stop)) - 1u;  // the first stop ends the search
                int c = cnt;
                for (uint32_t m = hit; m && c < SAP_SLOTS; m &= m - 1u) my_slots[c++] = base + k + __builtin_ctz(m);
                cnt += __popc(hit);
                if (stop) {
                    active = false;
                    break;
                }
            }
        }
    }
    if (t < n) {
        counts[t] = cnt;
        if (cnt > SAP_SLOTS) ovf[atomicAdd(ovf_count, 1)] = t;  // finished by k_sap_overflow
    }
}

// The first min(count, SAP_SLOTS) pairs of object t at offsets[t], partners in
// sorted order.
__global__ void k_sap_emit(const float4* __restrict__ slo, int n, const int* __restrict__ counts,
                           const int* __restrict__ offsets, const int* __restrict__ slots, int* __restrict__ pairs,
                           int capacity) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = min(counts[t], SAP_SLOTS);
    if (c == 0) return;
    const int id = __float_as_int(slo[t].w);
    const int* my_slots = slots + (size_t)t * SAP_SLOTS;
    const int o = offsets[t];
    for (int i = 0; i < c; ++i) {
        if (o + i < capacity) {
            const int iu = __float_as_int(slo[my_slots[i]].w);
            pairs[2 * (size_t)(o + i)] = id < iu ? id : iu;
            pairs[2 * (size_t)(o + i) + 1] = id < iu ? iu : id;
        }
    }
}

// Objects with more than SAP_SLOTS partners: one wavefront per listed object
// resumes the search after the last recorded partner, SAP_OVF_STEPS * 64
// candidates per step (one per lane and sub-step, loads in flight together),
// and writes the remaining pairs in sorted order (ballot prefix counts).
constexpr int SAP_OVF_BLOCKS = 64;
constexpr int SAP_OVF_STEPS = 4;

template <bool SCENES>
__global__ __launch_bounds__(256) void k_sap_overflow(const float4* __restrict__ slo, const float4* __restrict__ shi,
                                                      int n, const int* __restrict__ counts,
                                                      const int* __restrict__ offsets, const int* __restrict__ slots,
                                                      const int* __restrict__ ovf, const int* __restrict__ ovf_count,
                                                      int* __restrict__ pairs, int capacity) {
    const int lane = threadIdx.x & 63;
    const int nov = *ovf_count;
    for (int w = blockIdx.x * 4 + (int)threadIdx.x / 64; w < nov; w += gridDim.x * 4) {
        const int t = ovf[w];
        const float4 l = slo[t], h = shi[t];
        const int id = __float_as_int(l.w);
        const int sc = __float_as_int(h.w);
        int o = offsets[t] + SAP_SLOTS;
        int left = counts[t] - SAP_SLOTS;
        for (int u0 = slots[(size_t)t * SAP_SLOTS + SAP_SLOTS - 1] + 1; left > 0 && u0 < n;
             u0 += 64 * SAP_OVF_STEPS) {
            float4 lu[SAP_OVF_STEPS], hu[SAP_OVF_STEPS];
#pragma unroll
            for (int j = 0; j < SAP_OVF_STEPS; ++j) {
                const int u = min(u0 + 64 * j + lane, n - 1);
                lu[j] = slo[u];
                hu[j] = shi[u];
            }
            bool done = false;
#pragma unroll
            for (int j = 0; j < SAP_OVF_STEPS; ++j) {
                if (done) break;
                const int u = u0 + 64 * j + lane;
                const bool in = u < n;
                const uint64_t stop = __ballot(in && sap_stop(lu[j], hu[j], h, sc, SCENES));
                uint64_t hit = __ballot(in && sap_overlap(lu[j], hu[j], l, h));
                if (stop) hit &= (1ull << __builtin_ctzll(stop)) - 1ull;
                if ((hit >> lane) & 1ull) {
                    const int q = o + __popcll(hit & ((1ull << lane) - 1ull));
                    if (q < capacity) {
                        const int iu = __float_as_int(lu[j].w);
                        pairs[2 * (size_t)q] = id < iu ? id : iu;
                        pairs[2 * (size_t)q + 1] = id < iu ? iu : id;
                    }
                }
                o += __popcll(hit);
                left -= __popcll(hit);
                done = stop != 0ull || !(u0 + 64 * (j + 1) < n);
            }
            if (done) break;
        }
    }
}

// ------------------------------------------------------------------ batched SAP
//
// Many hull scenes in one pass. Hulls are concatenated scene by scene
// (scene_first[s] .. scene_first[s+1]); every scene keeps its own sweep axis and
// no pair crosses a scene. Keys are (scene, order-preserving float key of the
// lower extent on the scene's axis, hull index) packed into 64 bits, so one
// radix sort orders every scene exactly as sd_sap_pairs orders it alone, and
// the sweep stops at the end of the scene. The per-scene axis is reduced in the
// same FP32 order as k_axis_partial + k_axis_pick use for a scene of up to
// AX_BLOCK * AX_BLOCKS hulls (centres relative to the scene's first box,
// 256-element chunk sums by a tree, chunk sums in order), so it is the axis
// sd_sap_pairs picks for that scene.

constexpr int SAPB_MAX_SCENE = AX_BLOCK * AX_BLOCKS;

__global__ __launch_bounds__(AX_BLOCK) void k_axis_scenes(const float4* __restrict__ lo4,
                                                          const float4* __restrict__ hi4, int n,
                                                          const int* __restrict__ scene_first, int nscenes,
                                                          int* __restrict__ axis, int* __restrict__ error) {
    __shared__ float red[6][AX_BLOCK];
    const int sc = blockIdx.x;
    const int b0 = scene_first[sc], b1 = scene_first[sc + 1];
    const bool bad = b0 < 0 || b1 > n || b1 < b0 || b1 - b0 > SAPB_MAX_SCENE;
    if (bad) {
        if (threadIdx.x == 0) {
            axis[sc] = 0;
            atomicOr(error, 1);
        }
        return;
    }
    if (b1 == b0) {
        if (threadIdx.x == 0) axis[sc] = 0;
        return;
    }
    float rx, ry, rz;
    box_centre(lo4[b0], hi4[b0], rx, ry, rz);
    float s[6] = {0, 0, 0, 0, 0, 0};  // meaningful in thread 0
    for (int c0 = b0; c0 < b1; c0 += AX_BLOCK) {
        float acc[6] = {0, 0, 0, 0, 0, 0};
        const int i = c0 + (int)threadIdx.x;
        if (i < b1) {
            float cx, cy, cz;
            box_centre(lo4[i], hi4[i], cx, cy, cz);
            const float dx = cx - rx, dy = cy - ry, dz = cz - rz;
            acc[0] += dx;
            acc[1] += dy;
            acc[2] += dz;
            acc[3] += dx * dx;
            acc[4] += dy * dy;
            acc[5] += dz * dz;
        }
        for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = acc[k];
        axis_tree(red);
        if (threadIdx.x == 0)
            for (int k = 0; k < 6; ++k) s[k] += red[k][0];
        __syncthreads();
    }
    if (threadIdx.x == 0) axis[sc] = axis_of_sums(s, b1 - b0);
}

__global__ void k_sapb_keys(const float4* __restrict__ lo4, int n, const int* __restrict__ scene_of, int nscenes,
                            const int* __restrict__ axis, int ibits, uint64_t* __restrict__ keys,
                            int* __restrict__ error) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int sc = scene_of[i];
    if (sc < 0 || sc >= nscenes) {
        atomicOr(error, 2);
        sc = 0;
    }
    const uint64_t fk = float_key(axis_of(lo4[i], axis[sc]));
    keys[i] = ((uint64_t)sc << (32 + ibits)) | (fk << ibits) | (uint64_t)i;
}

__global__ void k_sapb_gather(const uint64_t* __restrict__ keys, int n, int ibits, const float4* __restrict__ lo4,
                              const float4* __restrict__ hi4, const int* __restrict__ scene_of,
                              const int* __restrict__ axis, int nscenes, float4* __restrict__ slo,
                              float4* __restrict__ shi) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(keys[t] & ((1ull << ibits) - 1ull));
    int sc = scene_of[i];
    sc = sc < 0 || sc >= nscenes ? 0 : sc;
    const int a = axis[sc];
    const float4 l = lo4[i], h = hi4[i];
    slo[t] = make_float4(axis_of(l, a), axis_of(l, (a + 1) % 3), axis_of(l, (a + 2) % 3), __int_as_float(i));
    shi[t] = make_float4(axis_of(h, a), axis_of(h, (a + 1) % 3), axis_of(h, (a + 2) % 3), __int_as_float(sc));
}

// ------------------------------------------------------------------ GJK

struct HullSet {
    const float4* pos;
    const float4* quat;
    const int* vstart;
    const int* nv;
    const float4* verts;
};

__device__ __forceinline__ void make_shapes(const HullSet& h, int ia, int ib, Shape& A, Shape& B, V3& dc) {
    const float4 ca = h.pos[ia], cb = h.pos[ib];
    A.verts = h.verts + h.vstart[ia];
    A.nv = h.nv[ia];
    A.radius = ca.w;
    A.R = quat_rot(h.quat[ia]);
    A.off = v3(0.0f, 0.0f, 0.0f);
    B.verts = h.verts + h.vstart[ib];
    B.nv = h.nv[ib];
    B.radius = cb.w;
    B.R = quat_rot(h.quat[ib]);
    dc = v3(cb.x - ca.x, cb.y - ca.y, cb.z - ca.z);
    B.off = dc;
}

// GJK/EPA runs as two kernels so that the GJK kernel keeps its simplex in
// registers at full occupancy and the polytope of EPA is paid only by
// intersecting pairs:
//   k_gjk_phase  one thread per pair: the GJK phase; separated pairs are
//                finished, intersecting pairs append their final simplex to a
//                list (order of the list is irrelevant: every entry is
//                finished independently and written by pair index)
//   k_epa_wave   one 16-lane group per listed pair: EPA from that simplex
//                with the polytope in LDS; a warm-started entry whose result
//                is rejected goes to a retry list, whose pairs get the cold
//                query (k_gjk_phase<true, false>, then k_epa_wave again).
// warm (nullable): per-pair predicted direction (A -> B). A warm-started query
// selects the WARM_K vertices of each hull furthest along it (one pass over
// both hulls), solves GJK over those seed sets, continues on the full hulls
// within warm_budget iterations, and is accepted only when it converged, its
// closest features carry the margin certificate (one more pass), and it lies
// beyond near_margin of contact; otherwise the pair is redone cold with
// max_iter iterations. Pairs involving a sphere shape are always cold.
struct EpaEntry {
    int4 meta;     // pair, EPA iteration limit, iterations so far,
                   // simplex size | mode << 8 (0 cold, 1 warm) | uncertain << 16
    float4 w[4];   // simplex points (x, y, z, key bits)
    // written by k_epa_wave, read by k_epa_finish:
    float4 f[3];   // closest face of the final polytope (x, y, z, key bits)
    float4 inner;  // interior point of the polytope (x, y, z), face plane distance
    int4 res;      // flags, EPA iterations, has_face, 0
};

// Ids outside the set, vertex ranges outside the vertex array, or a hull of
// more vertices than a feature key holds.
__device__ __forceinline__ bool pair_is_bad(const HullSet& h, int ia, int ib, int n_hulls, int n_verts) {
    if (ia < 0 || ia >= n_hulls || ib < 0 || ib >= n_hulls) return true;
    if (h.nv[ia] > GJK_MAX_VERTS_PER_HULL || h.nv[ib] > GJK_MAX_VERTS_PER_HULL) return true;
    return h.vstart[ia] + h.nv[ia] > n_verts || h.vstart[ib] + h.nv[ib] > n_verts;
}

__device__ __forceinline__ void write_out(int p, const GjkOut& o, int flags, int spent, float* __restrict__ sep,
                                          float* __restrict__ normal, int* __restrict__ oflags,
                                          int* __restrict__ iters) {
    sep[p] = o.sep;
    normal[3 * (size_t)p] = o.n.x;
    normal[3 * (size_t)p + 1] = o.n.y;
    normal[3 * (size_t)p + 2] = o.n.z;
    oflags[p] = flags;
    iters[p] = spent;
}

// The EPA of a query runs with that query's iteration limit (the warm budget
// for a warm-started query, max_iter for a cold one).
__device__ __forceinline__ void push_epa(EpaEntry* __restrict__ list, int* __restrict__ count, int p, int mode,
                                         bool uncertain, int limit, int spent, const Simplex& s) {
    const int k = atomicAdd(count, 1);
    EpaEntry e;
    e.meta = make_int4(p, limit, spent, s.n | (mode << 8) | ((uncertain ? 1 : 0) << 16));
#pragma unroll
    for (int i = 0; i < 4; ++i)
        e.w[i] = make_float4(s.w[i].p.x, s.w[i].p.y, s.w[i].p.z, __int_as_float(i < s.n ? s.w[i].key : -1));
    list[k] = e;
}

// A warm result is kept when it converged cleanly, its features are certified
// unique, and it is not within near_margin of contact.
__device__ __forceinline__ bool warm_accepted(const GjkOut& o, float near_margin) {
    return (o.flags & GJK_CONVERGED) && !(o.flags & (GJK_DEGENERATE | GJK_EPA_LIMIT | GJK_BUDGET)) && o.certified &&
           fabsf(o.sep) > near_margin;
}

// The seed problem of a warm-started pair, solved by k_gjk_seed and continued
// on the full hulls by k_gjk_phase<false, true>: the simplex (a seed hit may
// hold four points), its closest point, its size and whether it is a hit.
struct SeedState {
    float4 w[4];  // simplex points (x, y, z, key bits)
    float4 v;     // closest point (x, y, z); w: simplex size | hit << 8 (int bits)
};

// Seed problem of every warm-started pair (one thread per pair): the WARM_K
// vertices of each hull furthest along the prediction u (a pass over both
// hulls), then GJK over those seed sets from their support along u. A hit of
// the seed problem is a hit of the hulls (the seed sets are subsets). Pairs
// with a sphere shape or invalid ids are not seeded (the continuation kernel
// does not read their state).
__global__ __launch_bounds__(64) void k_gjk_seed(const int* __restrict__ pairs, int npairs, HullSet h, int n_hulls,
                                                 int n_verts, const float* __restrict__ warm,
                                                 SeedState* __restrict__ state, const int* __restrict__ dcount) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (dcount ? min(*dcount, npairs) : npairs)) return;
    const int ia = pairs[2 * (size_t)p], ib = pairs[2 * (size_t)p + 1];
    if (pair_is_bad(h, ia, ib, n_hulls, n_verts)) return;
    Shape A, B;
    V3 dc;
    make_shapes(h, ia, ib, A, B, dc);
    if (A.nv == 0 || B.nv == 0) return;
    V3 u = v3(warm[3 * (size_t)p], warm[3 * (size_t)p + 1], warm[3 * (size_t)p + 2]);
    if (dot(u, u) == 0.0f) u = v3(1.0f, 0.0f, 0.0f);
    SeedSet sa, sb;
    seed_select(A, u, sa);
    seed_select(B, neg(u), sb);
    Simplex s;
    s.w[0] = seed_support(sa, sb, u);
    s.n = 1;
    V3 v = s.w[0].p;
    GjkState gs = gjk_state();
    gjk_loop<true>(A, B, &sa, &sb, s, v, SMALL_MAX_ITER, gs);
    const bool hit = gs.hit && !gs.uncertain;
    SeedState st;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        st.w[i] = make_float4(s.w[i].p.x, s.w[i].p.y, s.w[i].p.z, __int_as_float(i < s.n ? s.w[i].key : -1));
    st.v = make_float4(v.x, v.y, v.z, __int_as_float(s.n | (hit ? 1 << 8 : 0)));
    state[p] = st;
}

// GJK phase of every pair (RETRY = false, one thread per pair: warm-started
// with WARM = true, else cold), or the cold query of the
// warm-started pairs whose result was rejected (RETRY = true, threads striding
// over the retry list of (pair, iterations so far)). Hits go to the EPA list;
// a rejected warm separation goes to the retry list. Iterations count support
// passes over the full hulls: a warm query adds one for the seed selection and
// one for the certificate.
template <bool RETRY, bool WARM>
__global__ __launch_bounds__(64) void k_gjk_phase(const int* __restrict__ pairs, int npairs, HullSet h, int n_hulls,
                                                  int n_verts, const float* __restrict__ warm, int warm_budget,
                                                  int max_iter, float near_margin, float* __restrict__ sep,
                                                  float* __restrict__ normal, int* __restrict__ flags,
                                                  int* __restrict__ iters, EpaEntry* __restrict__ list,
                                                  int* __restrict__ count, int2* __restrict__ retry,
                                                  int* __restrict__ retry_count, const SeedState* __restrict__ seed,
                                                  const int* __restrict__ dcount) {
    // dcount (nullable): pairs held on the device (the first min(*dcount,
    // npairs) of the list), for launches sized to a capacity
    const int n = RETRY ? *retry_count : (dcount ? min(*dcount, npairs) : npairs);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += RETRY ? gridDim.x * blockDim.x : n) {
        const int p = RETRY ? retry[t].x : t;
        const int ia = pairs[2 * (size_t)p], ib = pairs[2 * (size_t)p + 1];
        if (!RETRY && pair_is_bad(h, ia, ib, n_hulls, n_verts)) {  // reported, never dereferenced
            sep[p] = 0.0f;
            flags[p] = -1;
            iters[p] = 0;
            continue;
        }
        Shape A, B;
        V3 dc;
        make_shapes(h, ia, ib, A, B, dc);
        const bool warm_try = WARM && !RETRY && A.nv > 0 && B.nv > 0;
        const int spent = RETRY ? retry[t].y : 0;
        Simplex s;
        V3 v;
        GjkState g = gjk_state();
        int limit = max_iter, extra = 0;
        bool full = true;
        if (WARM && warm_try) {
            // continue the seed problem of k_gjk_seed on the full hulls; a seed
            // hit goes to EPA as it is
            limit = warm_budget < max_iter ? warm_budget : max_iter;
            extra = 2;
            const SeedState st = seed[p];
            const int meta = __float_as_int(st.v.w);
            s.n = meta & 0xff;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                s.w[i].p = v3(st.w[i].x, st.w[i].y, st.w[i].z);
                s.w[i].key = __float_as_int(st.w[i].w);
            }
            v = v3(st.v.x, st.v.y, st.v.z);
            if (meta >> 8) {
                g.hit = true;
                full = false;
            }
        } else {
            V3 dir = dc;
            if (dot(dir, dir) == 0.0f) dir = v3(1.0f, 0.0f, 0.0f);
            s.w[0] = support_M(A, B, dir);
            s.n = 1;
            v = s.w[0].p;
        }
        if (full) gjk_loop<false>(A, B, nullptr, nullptr, s, v, limit, g);
        const int it = g.it + extra;
        GjkOut o = gjk_out_init();
        if (!g.hit || g.uncertain) {
            // the separated result; an uncertain hit whose feature normal
            // separates the shapes is a converged separation, else EPA decides
            if (WARM && warm_try) separated_result<true>(A, B, s, o);
            else separated_result<false>(A, B, s, o);
            if (g.hit && axis_separates(A, B, o.n)) {
                g.hit = false;
                g.converged = true;
            }
        }
        if (g.hit) {
            push_epa(list, count, p, warm_try ? 1 : 0, g.uncertain, limit, spent + it, s);
            continue;
        }
        o.flags = g.converged ? GJK_CONVERGED : GJK_BUDGET;
        if (WARM && warm_try) {
            if (warm_accepted(o, near_margin)) write_out(p, o, o.flags | GJK_WARM_USED, it, sep, normal, flags, iters);
            else retry[atomicAdd(retry_count, 1)] = make_int2(p, it);
        } else {
            write_out(p, o, o.flags, spent + it, sep, normal, flags, iters);
        }
    }
}

constexpr int EPA_WAVES_PER_BLOCK = 2;
constexpr int EPA_GROUPS_PER_BLOCK = EPA_WAVES_PER_BLOCK * EPA_GPW;
constexpr int GJK_RETRY_MAX_BLOCKS = 1024;

// One 16-lane group per listed pair, persistent over the list (its length is
// only known on the device); the grid is the number of blocks the device holds
// at once. The polytope is in the group's LDS region. The group stores EPA's
// final face in the entry; k_epa_finish turns it into the result.
__global__ __launch_bounds__(64 * EPA_WAVES_PER_BLOCK) void k_epa_wave(const int* __restrict__ pairs, HullSet h,
                                                                       EpaEntry* __restrict__ list,
                                                                       const int* __restrict__ count) {
    __shared__ EpaShared sh[EPA_GROUPS_PER_BLOCK];
    const int gl = threadIdx.x & (EPA_G - 1);
    const int base = threadIdx.x & (64 - EPA_G);  // first lane of the group within its wave
    const int grp = threadIdx.x / EPA_G;
    EpaShared& S = sh[grp];
    const int n = *count;
    for (int k = blockIdx.x * EPA_GROUPS_PER_BLOCK + grp; k < n; k += gridDim.x * EPA_GROUPS_PER_BLOCK) {
        const int4 meta = list[k].meta;
        const int p = meta.x;
        const int ia = pairs[2 * (size_t)p], ib = pairs[2 * (size_t)p + 1];
        Shape A, B;
        V3 dc;
        make_shapes(h, ia, ib, A, B, dc);
        const HullRegs hA = load_hull_regs(A, gl), hB = load_hull_regs(B, gl);
        Simplex s;
        s.n = meta.w & 0xff;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 w = list[k].w[i];
            s.w[i].p = v3(w.x, w.y, w.z);
            s.w[i].key = __float_as_int(w.w);
        }
        EpaFace r;
        grp_epa(A, hA, B, hB, s, meta.y, r, S, gl, base);
        if (gl == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                list[k].f[i] = make_float4(r.q[i].p.x, r.q[i].p.y, r.q[i].p.z, __int_as_float(r.q[i].key));
            list[k].inner = make_float4(r.inner.x, r.inner.y, r.inner.z, r.d);
            list[k].res = make_int4(r.flags, r.iters, r.has_face ? 1 : 0, 0);
        }
        grp_sync();  // the next entry reuses this group's LDS region
    }
}

// Result of every listed pair from EPA's face (one thread per entry): depth
// and normal of the closest features, the certificate of a warm-started entry;
// an uncertain GJK exit whose EPA found no enclosing polytope reports the
// uncertified separation of its GJK simplex (GJK_BUDGET). A warm-started
// entry whose result is rejected goes to the retry list with the iterations
// spent so far.
__global__ __launch_bounds__(64) void k_epa_finish(const int* __restrict__ pairs, HullSet h, float near_margin,
                                                   float* __restrict__ sep, float* __restrict__ normal,
                                                   int* __restrict__ flags, int* __restrict__ iters,
                                                   const EpaEntry* __restrict__ list, const int* __restrict__ count,
                                                   int2* __restrict__ retry, int* __restrict__ retry_count) {
    const int n = *count;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const EpaEntry& e = list[k];
        const int4 meta = e.meta, res = e.res;
        const int p = meta.x;
        const int ia = pairs[2 * (size_t)p], ib = pairs[2 * (size_t)p + 1];
        Shape A, B;
        V3 dc;
        make_shapes(h, ia, ib, A, B, dc);
        const bool warm_started = ((meta.w >> 8) & 0xff) == 1;
        const bool uncertain = ((meta.w >> 16) & 1) != 0;
        GjkOut o = gjk_out_init();
        o.flags = res.x;
        o.epa_iters = res.y;
        if (res.z) {
            MPoint q[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float4 f = e.f[i];
                q[i].p = v3(f.x, f.y, f.z);
                q[i].key = __float_as_int(f.w);
            }
            const float4 in = e.inner;
            if (warm_started) hit_result<true>(A, B, q, v3(in.x, in.y, in.z), o);
            else hit_result<false>(A, B, q, v3(in.x, in.y, in.z), o);
        }
        if (uncertain && (o.flags & GJK_DEGENERATE)) {
            Simplex s;
            s.n = meta.w & 0xff;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 w = e.w[i];
                s.w[i].p = v3(w.x, w.y, w.z);
                s.w[i].key = __float_as_int(w.w);
            }
            const int ei = o.epa_iters;
            o = gjk_out_init();
            o.epa_iters = ei;
            o.flags = GJK_BUDGET;  // never accepted for a warm query: no certificate needed
            separated_result<false, true>(A, B, s, o);
        }
        const int spent = meta.z + o.epa_iters;
        if (!warm_started || warm_accepted(o, near_margin))
            write_out(p, o, o.flags | (warm_started ? GJK_WARM_USED : 0), spent, sep, normal, flags, iters);
        else
            retry[atomicAdd(retry_count, 1)] = make_int2(p, spent);
    }
}

// Blocks of k_epa_wave resident on the current device at once.
static int epa_resident_blocks() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    if (cached[dev] == 0) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_epa_wave, 64 * EPA_WAVES_PER_BLOCK, 0) !=
                hipSuccess ||
            cus < 1 || per < 1)
            return 1024;
        cached[dev] = cus * per;
    }
    return cached[dev];
}

size_t gjk_scratch_bytes(int npairs) {
    return 256 + (size_t)(npairs > 0 ? npairs : 1) * (sizeof(EpaEntry) + sizeof(int2) + sizeof(SeedState));
}

sd_status check_hulls(const float* pos4, const float* quat4, const int* vstart, const int* nv, const float* verts4) {
    SD_REQUIRE(pos4 && quat4 && vstart && nv && verts4, "hulls: null pointer");
    return SD_OK;
}

}  // namespace sd

using namespace sd;

extern "C" sd_status sd_hull_aabbs(const float* pos4, const float* quat4, const int* vstart, const int* nv,
                                   const float* verts4, int n, float* lo4, float* hi4, void* stream) {
    SD_REQUIRE(n >= 0, "sd_hull_aabbs: n < 0");
    if (n == 0) return SD_OK;
    sd_status st = check_hulls(pos4, quat4, vstart, nv, verts4);
    if (st != SD_OK) return st;
    SD_REQUIRE(lo4 && hi4, "sd_hull_aabbs: null output");
    hipLaunchKernelGGL(k_hull_aabb, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, (const float4*)pos4,
                       (const float4*)quat4, vstart, nv, (const float4*)verts4, n, (float4*)lo4, (float4*)hi4);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// Scratch regions are carved at 256-byte boundaries; sizes below add each
// region rounded up the same way.
static size_t pad256(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

extern "C" sd_status sd_sap_scratch_bytes(int n, size_t* bytes) {
    SD_REQUIRE(bytes && n >= 0, "sd_sap_scratch_bytes: bad argument");
    const size_t nn = (size_t)(n > 0 ? n : 1);
    *bytes = 4 * pad256(nn * sizeof(uint32_t))           // keys, vals, alt keys, alt vals
             + 2 * pad256(nn * sizeof(float4))           // sorted lower / upper boxes
             + 2 * pad256((nn + 2) * sizeof(int))        // counts, offsets (+ total)
             + pad256(nn * SAP_SLOTS * sizeof(int))      // first partners of each object
             + pad256((nn + 1) * sizeof(int))            // objects with more partners (+ count)
             + pad256((size_t)AX_BLOCKS * 6 * sizeof(float)) + pad256(64)  // axis partials, axis
             + pad256(radix_temp_bytes(n)) + pad256(scan_temp_bytes(n));
    return SD_OK;
}

static char* carve(char*& p, size_t bytes) {
    char* r = p;
    p += (bytes + 255) & ~(size_t)255;
    return r;
}

extern "C" sd_status sd_sap_pairs(const float* lo4, const float* hi4, int n, int* pairs, int capacity, int* n_pairs,
                                  int* axis_out, void* scratch, size_t scratch_bytes, void* stream) {
    SD_REQUIRE(n >= 0 && capacity >= 0 && n_pairs, "sd_sap_pairs: bad argument");
    if (n == 0) {
        *n_pairs = 0;
        return SD_OK;
    }
    SD_REQUIRE(lo4 && hi4 && scratch && (pairs || capacity == 0), "sd_sap_pairs: null pointer");
    size_t need;
    sd_sap_scratch_bytes(n, &need);
    SD_REQUIRE(scratch_bytes >= need, "sd_sap_pairs: scratch too small (%zu < %zu)", scratch_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)scratch;
    uint32_t* keys = (uint32_t*)carve(p, (size_t)n * 4);
    uint32_t* vals = (uint32_t*)carve(p, (size_t)n * 4);
    uint32_t* keys2 = (uint32_t*)carve(p, (size_t)n * 4);
    uint32_t* vals2 = (uint32_t*)carve(p, (size_t)n * 4);
    float4* slo = (float4*)carve(p, (size_t)n * sizeof(float4));
    float4* shi = (float4*)carve(p, (size_t)n * sizeof(float4));
    int* counts = (int*)carve(p, (size_t)(n + 2) * sizeof(int));
    int* offsets = (int*)carve(p, (size_t)(n + 2) * sizeof(int));
    int* slots = (int*)carve(p, (size_t)n * SAP_SLOTS * sizeof(int));
    int* ovf = (int*)carve(p, (size_t)(n + 1) * sizeof(int));
    int* ovf_count = ovf + n;
    float* part = (float*)carve(p, (size_t)AX_BLOCKS * 6 * sizeof(float));
    int* axis = (int*)carve(p, 64);
    void* sort_tmp = (void*)carve(p, radix_temp_bytes(n));
    void* scan_tmp = (void*)carve(p, scan_temp_bytes(n));
    int* total = offsets + n;

    const float4* L = (const float4*)lo4;
    const float4* H = (const float4*)hi4;
    const int nb = div_up(n, AX_BLOCK) < AX_BLOCKS ? div_up(n, AX_BLOCK) : AX_BLOCKS;
    hipLaunchKernelGGL(k_axis_partial, dim3(nb), dim3(AX_BLOCK), 0, s, L, H, n, part);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_axis_pick, dim3(1), dim3(AX_BLOCK), 0, s, part, nb, n, axis);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sap_keys, dim3(div_up(n, 256)), dim3(256), 0, s, L, n, axis, keys, vals);
    SD_LAUNCH_CHECK();
    sd_status st = radix_sort_pairs_u32(&keys, &vals, &keys2, &vals2, n, 0, 32, sort_tmp, radix_temp_bytes(n), s, 0);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_sap_gather, dim3(div_up(n, 256)), dim3(256), 0, s, vals, n, L, H, axis, slo, shi);
    SD_LAUNCH_CHECK();
    st = launch_fill_u32((uint32_t*)ovf_count, 0u, 1, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_sap_sweep<false>, dim3(div_up(n, SAP_TILE)), dim3(SAP_TILE), 0, s, slo, shi, n, counts,
                       slots, ovf, ovf_count);
    SD_LAUNCH_CHECK();
    st = exclusive_scan_i32(counts, offsets, n, total, scan_tmp, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_sap_emit, dim3(div_up(n, 256)), dim3(256), 0, s, slo, n, counts, offsets, slots, pairs,
                       capacity);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sap_overflow<false>, dim3(SAP_OVF_BLOCKS), dim3(256), 0, s, slo, shi, n, counts, offsets,
                       slots, ovf, ovf_count, pairs, capacity);
    SD_LAUNCH_CHECK();
    int h_total = 0, h_axis = 0;
    SD_HIP(hipMemcpyAsync(&h_total, total, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(&h_axis, axis, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    *n_pairs = h_total;
    if (axis_out) *axis_out = h_axis;
    if (h_total > capacity)
        return set_error(SD_ERR_CAPACITY, "sd_sap_pairs: %d pairs exceed capacity %d", h_total, capacity);
    return SD_OK;
}

static int bits_of(uint64_t v) {  // bits needed for values 0 .. v - 1
    int b = 1;
    while (b < 63 && (1ull << b) < v) ++b;
    return b;
}

extern "C" sd_status sd_sap_batched_scratch_bytes(int n, int nscenes, size_t* bytes) {
    SD_REQUIRE(bytes && n >= 0 && nscenes >= 1, "sd_sap_batched_scratch_bytes: bad argument");
    const size_t nn = (size_t)(n > 0 ? n : 1);
    *bytes = 2 * pad256(nn * sizeof(uint64_t))         // keys, alt keys
             + 2 * pad256(nn * sizeof(float4))         // sorted lower / upper boxes
             + 2 * pad256((nn + 2) * sizeof(int))      // counts, offsets (+ total)
             + pad256(nn * SAP_SLOTS * sizeof(int))    // first partners of each object
             + pad256((nn + 1) * sizeof(int))          // objects with more partners (+ count)
             + pad256((size_t)nscenes * sizeof(int)) + pad256(64)  // axes, error flag
             + pad256(radix_temp_bytes(n)) + pad256(scan_temp_bytes(n));
    return SD_OK;
}

extern "C" sd_status sd_sap_pairs_batched(const float* lo4, const float* hi4, int n, const int* scene_of,
                                          const int* scene_first, int nscenes, int* pairs, int capacity,
                                          int* n_pairs, int* axes, void* scratch, size_t scratch_bytes,
                                          void* stream) {
    SD_REQUIRE(n >= 0 && nscenes >= 1 && capacity >= 0 && n_pairs, "sd_sap_pairs_batched: bad argument");
    if (n == 0) {
        *n_pairs = 0;
        return SD_OK;
    }
    SD_REQUIRE(lo4 && hi4 && scene_of && scene_first && scratch && (pairs || capacity == 0),
               "sd_sap_pairs_batched: null pointer");
    hipStream_t s = (hipStream_t)stream;
    int *total = nullptr, *error = nullptr;
    sd_status st = enqueue_sap_batched((const float4*)lo4, (const float4*)hi4, n, scene_of, scene_first, nscenes,
                                       pairs, capacity, &total, &error, scratch, scratch_bytes, s);
    if (st != SD_OK) return st;
    if (axes) {
        // the axes follow the keys in the scratch layout: carve again to find them
        char* p = (char*)scratch;
        carve(p, (size_t)n * 8);
        carve(p, (size_t)n * 8);
        carve(p, (size_t)n * sizeof(float4));
        carve(p, (size_t)n * sizeof(float4));
        carve(p, (size_t)(n + 2) * sizeof(int));
        carve(p, (size_t)(n + 2) * sizeof(int));
        carve(p, (size_t)n * SAP_SLOTS * sizeof(int));
        carve(p, (size_t)(n + 1) * sizeof(int));
        const int* axis = (const int*)p;
        SD_HIP(hipMemcpyAsync(axes, axis, (size_t)nscenes * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    int h[2] = {0, 0};
    SD_HIP(hipMemcpyAsync(&h[0], total, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(&h[1], error, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    *n_pairs = h[0];
    if (h[1]) return set_error(SD_ERR_ARG, "sd_sap_pairs_batched: bad scene ranges or scene ids (bits %d)", h[1]);
    if (h[0] > capacity)
        return set_error(SD_ERR_CAPACITY, "sd_sap_pairs_batched: %d pairs exceed capacity %d", h[0], capacity);
    return SD_OK;
}

namespace {

// GJK/EPA of the first min(*dcount, capacity) pairs (all capacity pairs when
// dcount is null), warm-started from warm (world directions, A -> B) when it
// is given; every launch is sized to capacity. scratch: counters (EPA list,
// retry list, second EPA list), the EPA list, the retry list, the seed states
// (gjk_scratch_bytes(capacity)); the second EPA list reuses the first one's
// memory.
sd_status enqueue_gjk(const int* pairs, int capacity, const int* dcount, const HullSet& h, int n_hulls,
                      int n_verts, const float* warm, int warm_budget, int max_iter, float near_margin, float* sep,
                      float* normal3, int* flags, int* iters, void* scratch, hipStream_t s) {
    if (capacity == 0) return SD_OK;
    int* count = (int*)scratch;
    EpaEntry* list = (EpaEntry*)((char*)scratch + 256);
    int2* retry = (int2*)((char*)scratch + 256 + (size_t)capacity * sizeof(EpaEntry));
    SeedState* seed = (SeedState*)((char*)retry + (size_t)capacity * sizeof(int2));
    sd_status st = launch_fill_u32((uint32_t*)count, 0u, 3, s);
    if (st != SD_OK) return st;
    const int epa_blocks = std::min(div_up(capacity, EPA_GROUPS_PER_BLOCK), epa_resident_blocks());
    const dim3 epa_block(64 * EPA_WAVES_PER_BLOCK);
    const dim3 grid(div_up(capacity, 64));
    if (warm != nullptr) {
        hipLaunchKernelGGL(k_gjk_seed, grid, dim3(64), 0, s, pairs, capacity, h, n_hulls, n_verts, warm, seed, dcount);
        SD_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_gjk_phase<false, true>), grid, dim3(64), 0, s, pairs, capacity, h, n_hulls, n_verts,
                           warm, warm_budget, max_iter, near_margin, sep, normal3, flags, iters, list, count, retry,
                           count + 1, (const SeedState*)seed, dcount);
    } else {
        hipLaunchKernelGGL((k_gjk_phase<false, false>), grid, dim3(64), 0, s, pairs, capacity, h, n_hulls, n_verts,
                           warm, warm_budget, max_iter, near_margin, sep, normal3, flags, iters, list, count, retry,
                           count + 1, (const SeedState*)nullptr, dcount);
    }
    SD_LAUNCH_CHECK();
    const int fin_blocks = std::min(div_up(capacity, 64), GJK_RETRY_MAX_BLOCKS);
    hipLaunchKernelGGL(k_epa_wave, dim3(epa_blocks), epa_block, 0, s, pairs, h, list, (const int*)count);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_epa_finish, dim3(fin_blocks), dim3(64), 0, s, pairs, h, near_margin, sep, normal3, flags,
                       iters, (const EpaEntry*)list, (const int*)count, retry, count + 1);
    SD_LAUNCH_CHECK();
    if (warm == nullptr) return SD_OK;  // cold entries are never rejected
    // the rejected warm queries, cold
    hipLaunchKernelGGL((k_gjk_phase<true, false>), dim3(fin_blocks), dim3(64), 0, s, pairs, capacity, h, n_hulls,
                       n_verts, nullptr, 1, max_iter, near_margin, sep, normal3, flags, iters, list, count + 2, retry,
                       count + 1, (const SeedState*)nullptr, (const int*)nullptr);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_epa_wave, dim3(epa_blocks), epa_block, 0, s, pairs, h, list, (const int*)(count + 2));
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_epa_finish, dim3(fin_blocks), dim3(64), 0, s, pairs, h, near_margin, sep, normal3, flags,
                       iters, (const EpaEntry*)list, (const int*)(count + 2), retry, count + 1);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

}  // namespace

extern "C" sd_status sd_gjk_scratch_bytes(int n_pairs, size_t* bytes) {
    SD_REQUIRE(bytes && n_pairs >= 0, "sd_gjk_scratch_bytes: bad argument");
    *bytes = gjk_scratch_bytes(n_pairs);
    return SD_OK;
}

extern "C" sd_status sd_gjk(const int* pairs, int n_pairs, const float* pos4, const float* quat4, const int* vstart,
                            const int* nv, const float* verts4, int n_hulls, int n_verts, const float* warm_dir,
                            int warm_budget, int max_iter, float near_margin, float* sep, float* normal3, int* flags,
                            int* iters, void* scratch, size_t scratch_bytes, void* stream) {
    SD_REQUIRE(n_pairs >= 0 && n_hulls >= 1 && n_verts >= 0, "sd_gjk: bad sizes");
    if (n_pairs == 0) return SD_OK;
    sd_status st = check_hulls(pos4, quat4, vstart, nv, verts4);
    if (st != SD_OK) return st;
    SD_REQUIRE(pairs && sep && normal3 && flags && iters && scratch, "sd_gjk: null pointer");
    SD_REQUIRE(scratch_bytes >= gjk_scratch_bytes(n_pairs), "sd_gjk: scratch too small (%zu < %zu)", scratch_bytes,
               gjk_scratch_bytes(n_pairs));
    SD_REQUIRE(max_iter >= 1 && max_iter <= 1024, "sd_gjk: max_iter must be in [1, 1024]");
    SD_REQUIRE(warm_dir == nullptr || warm_budget >= 1, "sd_gjk: warm_budget must be >= 1");
    return enqueue_gjk(pairs, n_pairs, nullptr,
                       HullSet{(const float4*)pos4, (const float4*)quat4, vstart, nv, (const float4*)verts4}, n_hulls,
                       n_verts, warm_dir, warm_budget, max_iter, near_margin, sep, normal3, flags, iters, scratch,
                       (hipStream_t)stream);
}

namespace sd {

sd_status launch_hull_aabbs(const float4* pos4, const float4* quat4, const int* vstart, const int* nv,
                            const float4* verts4, int n, float4* lo4, float4* hi4, hipStream_t s) {
    if (n == 0) return SD_OK;
    hipLaunchKernelGGL(k_hull_aabb, dim3(div_up(n, 256)), dim3(256), 0, s, pos4, quat4, vstart, nv, verts4, n, lo4,
                       hi4);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status enqueue_sap_batched(const float4* L, const float4* H, int n, const int* scene_of, const int* scene_first,
                              int nscenes, int* pairs, int capacity, int** total_out, int** error_out, void* scratch,
                              size_t scratch_bytes, hipStream_t s) {
    const int ibits = bits_of((uint64_t)n);
    const int sbits = bits_of((uint64_t)nscenes);
    SD_REQUIRE(sbits + 32 + ibits <= 64, "sweep-and-prune: %d hulls in %d scenes do not fit 64-bit keys", n,
               nscenes);
    size_t need;
    sd_sap_batched_scratch_bytes(n, nscenes, &need);
    SD_REQUIRE(scratch_bytes >= need, "sweep-and-prune: scratch too small (%zu < %zu)", scratch_bytes, need);
    char* p = (char*)scratch;
    uint64_t* keys = (uint64_t*)carve(p, (size_t)n * 8);
    uint64_t* keys2 = (uint64_t*)carve(p, (size_t)n * 8);
    float4* slo = (float4*)carve(p, (size_t)n * sizeof(float4));
    float4* shi = (float4*)carve(p, (size_t)n * sizeof(float4));
    int* counts = (int*)carve(p, (size_t)(n + 2) * sizeof(int));
    int* offsets = (int*)carve(p, (size_t)(n + 2) * sizeof(int));
    int* slots = (int*)carve(p, (size_t)n * SAP_SLOTS * sizeof(int));
    int* ovf = (int*)carve(p, (size_t)(n + 1) * sizeof(int));
    int* ovf_count = ovf + n;
    int* axis = (int*)carve(p, (size_t)nscenes * sizeof(int));
    int* error = (int*)carve(p, 64);
    void* sort_tmp = (void*)carve(p, radix_temp_bytes(n));
    void* scan_tmp = (void*)carve(p, scan_temp_bytes(n));
    *total_out = offsets + n;
    *error_out = error;
    sd_status st = launch_fill_u32((uint32_t*)error, 0u, 1, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_axis_scenes, dim3(nscenes), dim3(AX_BLOCK), 0, s, L, H, n, scene_first, nscenes, axis,
                       error);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sapb_keys, dim3(div_up(n, 256)), dim3(256), 0, s, L, n, scene_of, nscenes,
                       (const int*)axis, ibits, keys, error);
    SD_LAUNCH_CHECK();
    uint64_t* K = keys;
    uint64_t* Ka = keys2;
    st = radix_sort_keys_u64(&K, &Ka, n, 0, sbits + 32 + ibits, sort_tmp, radix_temp_bytes(n), s, 0);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_sapb_gather, dim3(div_up(n, 256)), dim3(256), 0, s, (const uint64_t*)K, n, ibits, L, H,
                       scene_of, (const int*)axis, nscenes, slo, shi);
    SD_LAUNCH_CHECK();
    st = launch_fill_u32((uint32_t*)ovf_count, 0u, 1, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_sap_sweep<true>, dim3(div_up(n, SAP_TILE)), dim3(SAP_TILE), 0, s, slo, shi, n, counts,
                       slots, ovf, ovf_count);
    SD_LAUNCH_CHECK();
    st = exclusive_scan_i32(counts, offsets, n, offsets + n, scan_tmp, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_sap_emit, dim3(div_up(n, 256)), dim3(256), 0, s, slo, n, counts, offsets, slots, pairs,
                       capacity);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sap_overflow<true>, dim3(SAP_OVF_BLOCKS), dim3(256), 0, s, slo, shi, n, counts, offsets,
                       slots, ovf, ovf_count, pairs, capacity);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status enqueue_gjk_cold(const int* pairs, int capacity, const int* dcount, const float4* pos4,
                           const float4* quat4, const int* vstart, const int* nv, const float4* verts4, int n_hulls,
                           int n_verts, int max_iter, float near_margin, float* sep, float* normal3, int* flags,
                           int* iters, void* scratch, hipStream_t s) {
    return enqueue_gjk(pairs, capacity, dcount, HullSet{pos4, quat4, vstart, nv, verts4}, n_hulls, n_verts, nullptr,
                       1, max_iter, near_margin, sep, normal3, flags, iters, scratch, s);
}

sd_status enqueue_gjk_warm(const int* pairs, int capacity, const int* dcount, const float4* pos4,
                           const float4* quat4, const int* vstart, const int* nv, const float4* verts4, int n_hulls,
                           int n_verts, const float* warm_dir3, int warm_budget, int max_iter, float near_margin,
                           float* sep, float* normal3, int* flags, int* iters, void* scratch, hipStream_t s) {
    return enqueue_gjk(pairs, capacity, dcount, HullSet{pos4, quat4, vstart, nv, verts4}, n_hulls, n_verts,
                       warm_dir3, warm_budget, max_iter, near_margin, sep, normal3, flags, iters, scratch, s);
}

size_t gjk_cold_scratch_bytes(int capacity) { return gjk_scratch_bytes(capacity); }

}  // namespace sd
