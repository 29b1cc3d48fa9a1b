// Linear BVH broadphase: global domain, 30-bit Morton codes, Karras (2012)
// radix-tree build, bottom-up refit, ropes and a stackless overlap query.
//
// Replaces the sandbox LBVH pipeline (sandbox/testLBVH.cu:42-954):
// constructGlobalDomain (:61-88, a Thrust reduce with a host copy) becomes a
// device reduction; computeSpaceFillingCurveCUDA (:114-151) keeps its 10-bit
// per-axis interleave; generatePrefixes / buildBinaryRadix* (:213-535) become
// one Karras build on 64-bit (code << 32 | body) keys with a 64-bit count of
// leading zeros (quirk Q15: the reference applied 32-bit __clz to 64-bit keys);
// the depth probe + level-by-level refit (:546-699, depth+1 launches and a host
// copy) becomes a bottom-up refit in three launches (chunk-local arrival counters
// in LDS, then the chunk-crossing nodes in one workgroup); the 128-entry stack
// traversal (:769-864, atomicAdd pair append with a read-after-add race) becomes
// a stackless rope traversal with count / scan / write output, visiting only
// leaves to the right of the query so every pair is found once.
//
// Node numbering: internal nodes 0 .. n-2 (root 0), leaf k is node n-1+k.
#include "sd_internal.hpp"

namespace sd {

constexpr int LB_SENTINEL = -1;

__global__ void k_domain_partial(const float4* __restrict__ lo4, const float4* __restrict__ hi4, int n,
                                 float* __restrict__ part) {
    __shared__ float red[6][256];
    float a[6] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f, -3.402823466e38f, -3.402823466e38f,
                  -3.402823466e38f};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 l = lo4[i], h = hi4[i];
        a[0] = fminf(a[0], l.x);
        a[1] = fminf(a[1], l.y);
        a[2] = fminf(a[2], l.z);
        a[3] = fmaxf(a[3], h.x);
        a[4] = fmaxf(a[4], h.y);
        a[5] = fmaxf(a[5], h.z);
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
            for (int k = 3; k < 6; ++k) red[k][threadIdx.x] = fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// One thread per partial (nb <= DOMAIN_PARTS), then a tree reduction in LDS;
// min and max are exact, so the order does not matter (threads past nb repeat
// partial 0).
constexpr int DOMAIN_PARTS = 256;
__global__ __launch_bounds__(DOMAIN_PARTS) void k_domain_final(const float* __restrict__ part, int nb,
                                                               float* __restrict__ dom) {
    __shared__ float s[6][DOMAIN_PARTS];
    const int t = threadIdx.x;
    const int b = t < nb ? t : 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k][t] = part[b * 6 + k];
    __syncthreads();
    for (int h = DOMAIN_PARTS / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int k = 0; k < 6; ++k)
                s[k][t] = k < 3 ? fminf(s[k][t], s[k][t + h]) : fmaxf(s[k][t], s[k][t + h]);
        }
        __syncthreads();
    }
    if (t < 6) dom[t] = s[t][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0xFF0000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Cell of a centroid coordinate on a 1024 grid over [lo, hi].
__device__ __forceinline__ uint32_t morton_axis(float c, float lo, float hi) {
    const float ext = hi - lo;
    float f = ext > 0.0f ? floorf(((c - lo) / ext) * 1024.0f) : 0.0f;
    f = fmaxf(f, 0.0f);
    f = fminf(f, 1023.0f);
    return (uint32_t)f;
}

// keys[i] = morton30(centroid of box i) << 32 | i, with the code's drop_bits
// lowest bits cleared (0 for the public Morton codes)
__global__ void k_morton(const float4* __restrict__ lo4, const float4* __restrict__ hi4, int n,
                         const float* __restrict__ dom, uint64_t* __restrict__ keys, int drop_bits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 l = lo4[i], h = hi4[i];
    const float cx = 0.5f * (l.x + h.x), cy = 0.5f * (l.y + h.y), cz = 0.5f * (l.z + h.z);
    const uint32_t mx = morton_axis(cx, dom[0], dom[3]);
    const uint32_t my = morton_axis(cy, dom[1], dom[4]);
    const uint32_t mz = morton_axis(cz, dom[2], dom[5]);
    const uint32_t code = (spread10(mx) | (spread10(my) << 1) | (spread10(mz) << 2)) >> drop_bits << drop_bits;
    keys[i] = ((uint64_t)code << 32) | (uint64_t)(uint32_t)i;
}

// Length of the common prefix of sorted keys i and j (-1 outside [0, n)).
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(k[i] ^ k[j]);  // keys are distinct (index in the low half)
}

// Karras 2012, one thread per internal node.
__global__ void k_lbvh_build(const uint64_t* __restrict__ keys, int n, int* __restrict__ left,
                             int* __restrict__ right, int* __restrict__ parent, int* __restrict__ first,
                             int* __restrict__ last) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (i == 0) parent[0] = LB_SENTINEL;  // the root is nobody's child
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int lc = (lo == gamma) ? (n - 1 + gamma) : gamma;
    const int rc = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
    first[i] = lo;
    last[i] = hi;
}

// Leaf boxes (sorted order) + ranges; the refit's list count is cleared.
__global__ void k_lbvh_leaves(const uint64_t* __restrict__ keys, int n, const float4* __restrict__ lo4,
                              const float4* __restrict__ hi4, float4* __restrict__ nlo, float4* __restrict__ nhi,
                              int* __restrict__ first, int* __restrict__ last, int* __restrict__ parent,
                              int* __restrict__ list_count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *list_count = 0;
    if (k >= n) return;
    const int b = (int)(uint32_t)keys[k];
    const int node = n - 1 + k;
    nlo[node] = lo4[b];
    nhi[node] = hi4[b];
    first[node] = k;
    last[node] = k;
    if (n == 1) parent[node] = LB_SENTINEL;
}

// Bottom-up refit without device-wide fences. A node's box is the union of its
// children's boxes (exact, so the order of evaluation does not matter).
//  k_refit_chunks    one workgroup per chunk of RF_CHUNK consecutive leaves: every
//                    internal node whose leaf range lies inside the chunk is
//                    refit there (an internal node i is an end of its own range,
//                    so its arrival counter lives in LDS at i - chunk start); the
//                    second child to arrive computes the box
//  k_refit_mark      nodes whose range crosses a chunk boundary are listed with
//                    the number of their children already final (and their
//                    place in the list, both in ready[])
//  k_refit_boundary  one workgroup finishes the listed nodes bottom-up with
//                    arrival counters (workgroup-scope ordering only); up to
//                    RF_LDS_CAP listed nodes are finished in LDS (boxes,
//                    counters and parent/child places copied in first)
// Every walk is bounded; a malformed tree sets *error.
constexpr int RF_CHUNK = 1024;
constexpr int RF_FINISH_THREADS = 1024;
constexpr int RF_LDS_CAP = 2048;  // listed nodes finished in LDS (48 B each)

__device__ __forceinline__ bool rf_crosses(const int* __restrict__ first, const int* __restrict__ last, int node) {
    return first[node] / RF_CHUNK != last[node] / RF_CHUNK;
}

__device__ __forceinline__ void rf_union(int node, const int* __restrict__ left, const int* __restrict__ right,
                                         float4* nlo, float4* nhi) {
    const int a = left[node], b = right[node];
    const float4 la = nlo[a], lb = nlo[b], ha = nhi[a], hb = nhi[b];
    nlo[node] = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.0f);
    nhi[node] = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.0f);
}

// (node boxes are written by other threads during the kernel: no __restrict__)
__global__ __launch_bounds__(RF_CHUNK) void k_refit_chunks(int n, const int* __restrict__ left,
                                                           const int* __restrict__ right,
                                                           const int* __restrict__ parent,
                                                           const int* __restrict__ first,
                                                           const int* __restrict__ last, float4* nlo, float4* nhi,
                                                           int* error) {
    __shared__ int cnt[RF_CHUNK];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int c0 = blockIdx.x * RF_CHUNK;
    const int k = c0 + (int)threadIdx.x;
    if (k >= n || n < 2) return;
    int node = parent[n - 1 + k];
    for (int guard = 0; node >= 0; ++guard) {
        if (guard > RF_CHUNK || node >= n - 1) {
            atomicOr(error, 1);
            return;
        }
        if (rf_crosses(first, last, node)) return;  // finished by k_refit_boundary
        const int slot = node - c0;
        if (slot < 0 || slot >= RF_CHUNK) {
            atomicOr(error, 1);
            return;
        }
        __threadfence_block();
        if (atomicAdd(&cnt[slot], 1) == 0) return;  // first arrival: the sibling finishes the node
        __threadfence_block();
        rf_union(node, left, right, nlo, nhi);
        node = parent[node];
    }
}

// scratch: cnt[n], list[n], ready[n], count[1] (ints). ready[i] of a listed
// node: (its place in list) << 2 | (children already final).
__global__ void k_refit_mark(int n, const int* __restrict__ left, const int* __restrict__ right,
                             const int* __restrict__ first, const int* __restrict__ last, int* __restrict__ cnt,
                             int* __restrict__ list, int* __restrict__ ready, int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1 || !rf_crosses(first, last, i)) return;
    const int a = left[i], b = right[i];
    const int done = (a >= n - 1 || !rf_crosses(first, last, a) ? 1 : 0) +
                     (b >= n - 1 || !rf_crosses(first, last, b) ? 1 : 0);
    cnt[i] = done;
    const int e = atomicAdd(count, 1);
    ready[i] = (e << 2) | done;
    list[e] = i;
}

// Finishes the m <= RF_LDS_CAP listed nodes in LDS. A listed node's entry
// holds the union of its final (non-listed) children's boxes, its counter
// (children final), the entry of its parent and those of its listed
// children; the second child to arrive completes the node.
__device__ void refit_listed_lds(int n, int m, const int* __restrict__ left, const int* __restrict__ right,
                                 const int* __restrict__ parent, const int* __restrict__ first,
                                 const int* __restrict__ last, float4* nlo, float4* nhi,
                                 const int* __restrict__ list, const int* __restrict__ ready, int* error) {
    __shared__ float4 blo[RF_LDS_CAP], bhi[RF_LDS_CAP];
    __shared__ int lcnt[RF_LDS_CAP], lpar[RF_LDS_CAP];
    __shared__ int2 lch[RF_LDS_CAP];
    const float inf = __int_as_float(0x7f800000);
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
        const int node = list[e];
        if (node < 0 || node >= n - 1) {
            atomicOr(error, 1);
            lcnt[e] = 0;
            lpar[e] = -1;
            continue;
        }
        float4 lo = make_float4(inf, inf, inf, 0.0f), hi = make_float4(-inf, -inf, -inf, 0.0f);
        int ch[2] = {-1, -1};
        const int kids[2] = {left[node], right[node]};
        for (int j = 0; j < 2; ++j) {
            const int k = kids[j];
            if (k < n - 1 && rf_crosses(first, last, k)) {
                const int ek = ready[k] >> 2;
                if (ek < 0 || ek >= m) atomicOr(error, 1);
                else ch[j] = ek;
            } else {
                const float4 a = nlo[k], b = nhi[k];
                lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.0f);
                hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.0f);
            }
        }
        blo[e] = lo;
        bhi[e] = hi;
        lch[e] = make_int2(ch[0], ch[1]);
        lcnt[e] = ready[node] & 3;
        int ep = -1;
        const int p = node == 0 ? -1 : parent[node];
        if (p >= 0) {
            if (p >= n - 1 || !rf_crosses(first, last, p)) atomicOr(error, 1);  // a crossing node's parent crosses
            else ep = ready[p] >> 2;
            if (ep >= m) {
                atomicOr(error, 1);
                ep = -1;
            }
        }
        lpar[e] = ep;
    }
    __syncthreads();
    for (int e0 = threadIdx.x; e0 < m; e0 += blockDim.x) {
        if ((ready[list[e0]] & 3) != 2) continue;  // waits for a listed child (read where nothing changes it)
        int e = e0;
        for (int guard = 0;; ++guard) {
            if (guard > m) {
                atomicOr(error, 1);
                break;
            }
            float4 lo = blo[e], hi = bhi[e];
            const int2 ch = lch[e];
            if (ch.x >= 0) {
                const float4 a = blo[ch.x], b = bhi[ch.x];
                lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.0f);
                hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.0f);
            }
            if (ch.y >= 0) {
                const float4 a = blo[ch.y], b = bhi[ch.y];
                lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.0f);
                hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.0f);
            }
            blo[e] = lo;
            bhi[e] = hi;
            const int ep = lpar[e];
            if (ep < 0) break;
            __threadfence_block();  // (only LDS accesses are outstanding here)
            if (atomicAdd(&lcnt[ep], 1) + 1 < 2) break;  // the other child is not final yet
            __threadfence_block();
            e = ep;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
        const int node = list[e];
        if (node >= 0 && node < n - 1) {
            nlo[node] = blo[e];
            nhi[node] = bhi[e];
        }
    }
}

__global__ __launch_bounds__(RF_FINISH_THREADS) void k_refit_boundary(int n, const int* __restrict__ left,
                                                                       const int* __restrict__ right,
                                                                       const int* __restrict__ parent,
                                                                       const int* __restrict__ first,
                                                                       const int* __restrict__ last, float4* nlo,
                                                                       float4* nhi, int* cnt,
                                                                       const int* __restrict__ list,
                                                                       const int* __restrict__ ready,
                                                                       const int* __restrict__ count, int* error) {
    const int m = *count;
    if (m <= RF_LDS_CAP) {
        refit_listed_lds(n, m, left, right, parent, first, last, nlo, nhi, list, ready, error);
        return;
    }
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
        int node = list[e];
        if ((ready[node] & 3) != 2) continue;  // waits for a crossing child
        for (int guard = 0;; ++guard) {
            if (guard > n || node < 0 || node >= n - 1) {
                atomicOr(error, 1);
                break;
            }
            rf_union(node, left, right, nlo, nhi);
            const int p = node == 0 ? -1 : parent[node];
            if (p < 0) break;
            if (p >= n - 1 || !rf_crosses(first, last, p)) {
                atomicOr(error, 1);  // a crossing node's parent crosses too
                break;
            }
            __threadfence_block();
            if (atomicAdd(&cnt[p], 1) + 1 < 2) break;  // the other child is not final yet
            __threadfence_block();
            node = p;
        }
    }
}

// Rope of a node: the right sibling of its nearest ancestor-or-self that is a
// left child (the next node in depth-first order after its subtree).
// With pnode, also the tree in the form a stackless walk reads in one load
// step per node: pnode[2i] = (box low corner, left child or a leaf's index
// from its sorted key, as int bits), pnode[2i+1] = (box high corner, rope).
__global__ void k_lbvh_ropes(int n, const int* __restrict__ left, const int* __restrict__ right,
                             const int* __restrict__ parent, int* __restrict__ rope, int* __restrict__ error,
                             const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                             const uint64_t* __restrict__ sorted_keys, float4* __restrict__ pnode) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= 2 * n - 1) return;
    int c = node;
    int r = LB_SENTINEL;
    for (int guard = 0; c != 0; ++guard) {  // node 0 is the root (internal, or the only leaf)
        const int p = parent[c];
        if (guard > n || p < 0 || p >= n - 1) {
            atomicOr(error, 2);
            break;
        }
        if (left[p] == c) {
            r = right[p];
            break;
        }
        c = p;
    }
    rope[node] = r;
    if (pnode) {
        const int down = node < n - 1 ? left[node] : (int)(uint32_t)sorted_keys[node - (n - 1)];
        const float4 l = nlo[node], h = nhi[node];
        pnode[2 * (size_t)node] = make_float4(l.x, l.y, l.z, __int_as_float(down));
        pnode[2 * (size_t)node + 1] = make_float4(h.x, h.y, h.z, __int_as_float(r));
    }
}

__device__ __forceinline__ bool box_overlap(float4 al, float4 ah, float4 bl, float4 bh) {
    return al.x <= bh.x && bl.x <= ah.x && al.y <= bh.y && bl.y <= ah.y && al.z <= bh.z && bl.z <= ah.z;
}

// Stackless query of leaf k against leaves k+1 .. n-1. WRITE = false counts.
// A depth-first walk visits each of the 2n-1 nodes at most once; the bound
// turns a malformed tree into an error instead of an endless loop.
template <bool WRITE>
__global__ void k_lbvh_query(const uint64_t* __restrict__ keys, int n, const int* __restrict__ left,
                             const int* __restrict__ last, const int* __restrict__ rope,
                             const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                             int* __restrict__ counts, const int* __restrict__ offsets, int* __restrict__ pairs,
                             int capacity, int* __restrict__ error) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 ql = nlo[n - 1 + k], qh = nhi[n - 1 + k];
    const int qb = (int)(uint32_t)keys[k];
    int o = WRITE ? offsets[k] : 0;
    int cnt = 0;
    int node = n == 1 ? LB_SENTINEL : 0;
    for (int guard = 0; node != LB_SENTINEL; ++guard) {
        if (guard > 2 * n + 2 || node < 0 || node >= 2 * n - 1) {
            atomicOr(error, 4);
            break;
        }
        const bool leaf = node >= n - 1;
        // only leaves right of k: skip subtrees that end at or before k
        if (last[node] <= k || !box_overlap(ql, qh, nlo[node], nhi[node])) {
            node = rope[node];
            continue;
        }
        if (leaf) {
            if (WRITE && o < capacity) {
                const int b = (int)(uint32_t)keys[node - (n - 1)];
                pairs[2 * (size_t)o] = qb < b ? qb : b;
                pairs[2 * (size_t)o + 1] = qb < b ? b : qb;
            }
            ++o;
            ++cnt;
            node = rope[node];
        } else {
            node = left[node];
        }
    }
    if (!WRITE) counts[k] = cnt;
}

size_t lbvh_scratch_bytes(int n) {
    const size_t nn = (size_t)(n > 0 ? n : 1);
    return (2 * nn + 8) * sizeof(int) * 2 + 256 * 6 * sizeof(float) + 64 + scan_temp_bytes(n) + 1024;
}

sd_status launch_morton(const float* lo4, const float* hi4, int n, const float* domain6, uint64_t* keys,
                        int drop_bits, hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_morton, dim3(div_up(n, 256)), dim3(256), 0, s, (const float4*)lo4, (const float4*)hi4, n,
                       domain6, keys, drop_bits);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

}  // namespace sd

using namespace sd;

extern "C" sd_status sd_global_domain(const float* lo4, const float* hi4, int n, float* domain6, void* scratch,
                                      void* stream) {
    SD_REQUIRE(n >= 1 && lo4 && hi4 && domain6 && scratch, "sd_global_domain: bad argument");
    const int nb = div_up(n, 256) < DOMAIN_PARTS ? div_up(n, 256) : DOMAIN_PARTS;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_domain_partial, dim3(nb), dim3(256), 0, s, (const float4*)lo4, (const float4*)hi4, n,
                       (float*)scratch);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_domain_final, dim3(1), dim3(DOMAIN_PARTS), 0, s, (const float*)scratch, nb, domain6);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_morton30(const float* lo4, const float* hi4, int n, const float* domain6, uint64_t* keys,
                                 void* stream) {
    SD_REQUIRE(n >= 0 && (n == 0 || (lo4 && hi4 && domain6 && keys)), "sd_morton30: bad argument");
    if (n == 0) return SD_OK;
    return sd::launch_morton(lo4, hi4, n, domain6, keys, 0, (hipStream_t)stream);
}

sd_status sd::lbvh_build(const uint64_t* sorted_keys, int n, const float* lo4, const float* hi4, int* left, int* right,
                         int* parent, int* first, int* last, int* rope, float* node_lo4, float* node_hi4, int* counters,
                         int* error, float4* pnode, hipStream_t s) {
    SD_REQUIRE(n >= 1, "sd_lbvh_build: n must be >= 1");
    SD_REQUIRE(sorted_keys && lo4 && hi4 && left && right && parent && first && last && rope && node_lo4 && node_hi4 &&
                   counters && error,
               "sd_lbvh_build: null pointer");
    if (n > 1) {
        hipLaunchKernelGGL(k_lbvh_build, dim3(div_up(n - 1, 256)), dim3(256), 0, s, sorted_keys, n, left, right,
                           parent, first, last);
        SD_LAUNCH_CHECK();
    }
    int* rf_cnt = counters;
    int* rf_list = counters + n;
    int* rf_ready = counters + 2 * (size_t)n;
    int* rf_count = counters + 3 * (size_t)n;
    hipLaunchKernelGGL(k_lbvh_leaves, dim3(div_up(n, 256)), dim3(256), 0, s, sorted_keys, n, (const float4*)lo4,
                       (const float4*)hi4, (float4*)node_lo4, (float4*)node_hi4, first, last, parent, rf_count);
    SD_LAUNCH_CHECK();
    if (n > 1) {
        hipLaunchKernelGGL(k_refit_chunks, dim3(div_up(n, RF_CHUNK)), dim3(RF_CHUNK), 0, s, n, left, right, parent,
                           first, last, (float4*)node_lo4, (float4*)node_hi4, error);
        SD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_refit_mark, dim3(div_up(n - 1, 256)), dim3(256), 0, s, n, left, right, first, last,
                           rf_cnt, rf_list, rf_ready, rf_count);
        SD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_refit_boundary, dim3(1), dim3(RF_FINISH_THREADS), 0, s, n, left, right, parent, first,
                           last, (float4*)node_lo4, (float4*)node_hi4, rf_cnt, (const int*)rf_list,
                           (const int*)rf_ready, (const int*)rf_count, error);
        SD_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_lbvh_ropes, dim3(div_up(2 * n - 1, 256)), dim3(256), 0, s, n, left, right, parent, rope,
                       error, (const float4*)node_lo4, (const float4*)node_hi4, sorted_keys, pnode);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_lbvh_build(const uint64_t* sorted_keys, int n, const float* lo4, const float* hi4, int* left,
                                   int* right, int* parent, int* first, int* last, int* rope, float* node_lo4,
                                   float* node_hi4, int* counters, int* error, void* stream) {
    return lbvh_build(sorted_keys, n, lo4, hi4, left, right, parent, first, last, rope, node_lo4, node_hi4, counters,
                      error, nullptr, (hipStream_t)stream);
}

extern "C" sd_status sd_lbvh_query_scratch_bytes(int n, size_t* bytes) {
    SD_REQUIRE(bytes && n >= 0, "sd_lbvh_query_scratch_bytes: bad argument");
    *bytes = lbvh_scratch_bytes(n);
    return SD_OK;
}

extern "C" sd_status sd_lbvh_query(const uint64_t* sorted_keys, int n, const int* left, const int* last,
                                   const int* rope, const float* node_lo4, const float* node_hi4, int* pairs,
                                   int capacity, int* n_pairs, int* error, void* scratch, size_t scratch_bytes,
                                   void* stream) {
    SD_REQUIRE(n >= 1 && capacity >= 0 && n_pairs && error, "sd_lbvh_query: bad argument");
    SD_REQUIRE(sorted_keys && left && last && rope && node_lo4 && node_hi4 && scratch && (pairs || capacity == 0),
               "sd_lbvh_query: null pointer");
    SD_REQUIRE(scratch_bytes >= lbvh_scratch_bytes(n), "sd_lbvh_query: scratch too small");
    hipStream_t s = (hipStream_t)stream;
    int* counts = (int*)scratch;
    int* offsets = counts + (n + 2);
    int* total = offsets + n;
    void* scan_tmp = (void*)(offsets + (n + 2));
    hipLaunchKernelGGL(k_lbvh_query<false>, dim3(div_up(n, 128)), dim3(128), 0, s, sorted_keys, n, left, last, rope,
                       (const float4*)node_lo4, (const float4*)node_hi4, counts, nullptr, nullptr, 0, error);
    SD_LAUNCH_CHECK();
    sd_status st = exclusive_scan_i32(counts, offsets, n, total, scan_tmp, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_lbvh_query<true>, dim3(div_up(n, 128)), dim3(128), 0, s, sorted_keys, n, left, last, rope,
                       (const float4*)node_lo4, (const float4*)node_hi4, counts, offsets, pairs, capacity, error);
    SD_LAUNCH_CHECK();
    int h[2] = {0, 0};
    SD_HIP(hipMemcpyAsync(&h[0], total, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(&h[1], error, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    *n_pairs = h[0];
    if (h[1]) return set_error(SD_ERR_STATE, "sd_lbvh_query: malformed tree (error bits %d)", h[1]);
    if (h[0] > capacity) return set_error(SD_ERR_CAPACITY, "sd_lbvh_query: %d pairs exceed capacity %d", h[0], capacity);
    return SD_OK;
}
