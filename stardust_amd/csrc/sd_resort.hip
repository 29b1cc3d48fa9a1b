// Sorted order of a step's cell keys from the previous step's order.
//
// The keys a step sorts are key(s) = cell(s) << 32 | s, s being the sphere's
// position in the previous step's sorted order, in which cells were
// non-decreasing. A sphere whose cell did not change (a stayer) keeps the key
// old(s) = old_cell(s) << 32 | s, and old() increases with s: the stayers are
// already in order. The contact kernel lists the spheres whose cell changed
// (movers) with their new and old keys. Every key's place in the sorted order
// is then its rank among all keys:
//   stayer s:  (stayers before s) + (movers whose new key is smaller)
//              = s - (movers before s) + (movers whose new key is smaller)
//   mover x:   (stayers with a smaller key) + (movers whose new key is smaller)
//              stayers with a smaller key = (positions whose old key is smaller)
//                                           - (movers whose old key is smaller),
//              the first read off the previous step's cell ranges (a binary
//              search over the previous sorted positions when x's cell was empty).
// The result is the same array a stable radix sort of the keys produces, and
// the positions are permuted into that order in the same pass. With more than
// RESORT_CAP changes the step is given up: step_ctl[0] is set, every later
// kernel of the launch sequence returns at once (the state is untouched), and
// the host redoes that step with the full radix sort.
//
// Replaces nothing in the reference (which re-sorts from scratch every step,
// src/engine/cuda/sh_collision_detection.cu:310-360); the full radix sort runs
// instead whenever more than RESORT_CAP spheres changed cell, and after every
// upload, load or restore.
#include "sd_grid_dev.hpp"
#include "sd_internal.hpp"

namespace sd {

constexpr int RESORT_MERGE_BLOCK = 512;

// Old key of position s: its cell in the previous step's sorted positions.
__device__ __forceinline__ uint64_t old_key_at(const float4* __restrict__ posB, int s, const GridDev& g1,
                                               const SceneDev* __restrict__ scenes,
                                               const int* __restrict__ scene_of) {
    GridDev g = g1;
    uint32_t cbase = 0u;
    if (scenes) {
        const SceneDev& S = scenes[scene_of[s]];
        g = S.g;
        cbase = S.cbase;
    }
    const float4 p = posB[s];
    return ((uint64_t)(cbase + cell_id(g, p.x, p.y, p.z)) << 32) | (uint64_t)(uint32_t)s;
}

// One workgroup: sorts the movers' new and old keys (rank by counting, keys are
// distinct) and places every mover.
__global__ __launch_bounds__(RESORT_CAP) void k_resort_movers(
    const int* __restrict__ mcount, const uint64_t* __restrict__ mnew, const uint64_t* __restrict__ mold,
    uint64_t* __restrict__ msn, uint64_t* __restrict__ mso, int n, int* ranges, GridDev g,
    const SceneDev* __restrict__ scenes, const int* __restrict__ scene_of, const float4* __restrict__ pos_state,
    float4* __restrict__ posB, uint64_t* __restrict__ keys_out, int* __restrict__ step_ctl) {
    if (step_ctl[0]) return;  // an earlier step of this launch sequence was given up
    const int M = *mcount;
    if (M > RESORT_CAP) {  // too many changes: this step is given up (the host redoes it with the full sort)
        if (threadIdx.x == 0) step_ctl[0] = 1;
        return;
    }
    __shared__ uint64_t sn[RESORT_CAP], so[RESORT_CAP];
    const int i = threadIdx.x;
    if (i < M) {
        sn[i] = mnew[i];
        so[i] = mold[i];
    }
    __syncthreads();
    int dest = 0;
    uint64_t x = 0;
    if (i < M) {
        x = sn[i];
        const uint64_t o = so[i];
        int rn = 0, ro = 0, old_less = 0;
        for (int j = 0; j < M; ++j) {
            const uint64_t a = sn[j], b = so[j];
            rn += a < x ? 1 : 0;
            ro += b < o ? 1 : 0;
            old_less += b < x ? 1 : 0;
        }
        msn[rn] = x;
        mso[ro] = o;
        const uint32_t cell = (uint32_t)(x >> 32);  // < the cell count: written from a clamped cell
        const int sx = (int)(uint32_t)x;
        const int2 r = *reinterpret_cast<const int2*>(ranges + 2 * (size_t)cell);  // the previous step's
        int all_less;
        if (r.y > r.x) {
            all_less = r.x + (sx < r.x ? 0 : (sx > r.y ? r.y - r.x : sx - r.x));
        } else {  // empty cell before: binary search over the previous order
            int lo = 0, hi = n;
            while (lo < hi) {
                const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
                if (old_key_at(posB, mid, g, scenes, scene_of) < x) lo = mid + 1;
                else hi = mid;
            }
            all_less = lo;
        }
        dest = all_less - old_less + rn;
    }
    __syncthreads();  // every search is done with posB and the ranges
    if (i < M && dest >= 0 && dest < n && (uint32_t)x < (uint32_t)n) {  // (always, for consistent keys)
        keys_out[dest] = x;
        posB[dest] = pos_state[(uint32_t)x];
        // a cell left by a mover may be empty now; the step's range marking
        // rewrites every cell that is not, so only these need clearing
        const uint32_t cell_old = (uint32_t)(so[i] >> 32);
        ranges[2 * (size_t)cell_old] = 0;
        ranges[2 * (size_t)cell_old + 1] = 0;
    }
}

// Every stayer: its place, key and position.
__global__ __launch_bounds__(RESORT_MERGE_BLOCK) void k_resort_merge(
    const int* __restrict__ mcount, const uint64_t* __restrict__ msn, const uint64_t* __restrict__ mso,
    const uint64_t* __restrict__ keys, int n, const float4* __restrict__ pos_state, float4* __restrict__ posB,
    uint64_t* __restrict__ keys_out, const int* __restrict__ step_ctl) {
    if (step_ctl[0]) return;
    const int M = *mcount;
    if (M > RESORT_CAP) return;
    __shared__ uint64_t sn[RESORT_CAP], so[RESORT_CAP];
    for (int j = threadIdx.x; j < M; j += RESORT_MERGE_BLOCK) {
        sn[j] = msn[j];
        so[j] = mso[j];
    }
    __syncthreads();
    const int s = blockIdx.x * RESORT_MERGE_BLOCK + (int)threadIdx.x;
    if (s >= n) return;
    // movers before s: old keys ascend with their positions (the low halves)
    int lo = 0, hi = M;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)so[mid] < (uint32_t)s) lo = mid + 1;
        else hi = mid;
    }
    if (lo < M && (uint32_t)so[lo] == (uint32_t)s) return;  // a mover, placed by k_resort_movers
    const uint64_t y = keys[s];
    int a = 0, b = M;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (sn[mid] < y) a = mid + 1;
        else b = mid;
    }
    const int d = s - lo + a;
    if (d < 0 || d >= n) return;  // (never, for consistent keys)
    keys_out[d] = y;
    posB[d] = pos_state[s];
}

sd_status launch_resort_u64(const int* mcount, const uint64_t* mnew, const uint64_t* mold, uint64_t* msorted_new,
                            uint64_t* msorted_old, const uint64_t* keys, int n, int* ranges,
                            const GridDev& g, const SceneDev* scenes, const int* scene_of, const float4* pos_state,
                            float4* posB, uint64_t* keys_out, int* step_ctl, hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_resort_movers, dim3(1), dim3(RESORT_CAP), 0, s, mcount, mnew, mold, msorted_new,
                       msorted_old, n, ranges, g, scenes, scene_of, pos_state, posB, keys_out, step_ctl);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_resort_merge, dim3(div_up(n, RESORT_MERGE_BLOCK)), dim3(RESORT_MERGE_BLOCK), 0, s, mcount,
                       (const uint64_t*)msorted_new, (const uint64_t*)msorted_old, keys, n, pos_state, posB, keys_out,
                       (const int*)step_ctl);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

}  // namespace sd
