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
//              the first read off the previous step's cell starts.
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

// One workgroup: sorts the movers' new and old keys (rank by counting, keys are
// distinct) and places every mover.
__global__ __launch_bounds__(RESORT_CAP) void k_resort_movers(
    const int* __restrict__ mcount, const uint64_t* __restrict__ mnew, const uint64_t* __restrict__ mold,
    uint64_t* __restrict__ msn, uint64_t* __restrict__ mso, int n, const int* __restrict__ cell_start,
    const float4* __restrict__ pos_state, float4* __restrict__ posB, uint64_t* __restrict__ keys_out,
    int* __restrict__ step_ctl) {
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
    if (i >= M) return;
    const uint64_t x = sn[i];
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
    // previous positions with a smaller key: those of the cells before x's
    // cell, and those of x's cell below x's position
    const uint32_t cell = (uint32_t)(x >> 32);  // < the cell count: written from a clamped cell
    const int sx = (int)(uint32_t)x;
    const int c0 = cell_start[cell], c1 = cell_start[cell + 1];
    const int all_less = c0 + (sx < c0 ? 0 : (sx > c1 ? c1 - c0 : sx - c0));
    const int dest = all_less - old_less + rn;
    if (dest >= 0 && dest < n && (uint32_t)x < (uint32_t)n) {  // (always, for consistent keys)
        keys_out[dest] = x;
        posB[dest] = pos_state[(uint32_t)x];
    }
}

__device__ __forceinline__ int lower_bound_u64(const uint64_t* a, int lo, int hi, uint64_t key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Every stayer: its place, key and position. Also the new order's cell starts,
// in place: start[c] counts the keys below c << 32, so it changes by the
// movers' new keys below c << 32 minus their old keys below c << 32. Each
// block updates cells_per_block consecutive cells, RESORT_MERGE_BLOCK at a
// time; a run of cells that holds no mover key has one change for all its
// cells (most runs: none). Sets *next_count = 0 for the next step's movers.
__global__ __launch_bounds__(RESORT_MERGE_BLOCK) void k_resort_merge(
    const int* __restrict__ mcount, const uint64_t* __restrict__ msn, const uint64_t* __restrict__ mso,
    const uint64_t* __restrict__ keys, int n, const float4* __restrict__ pos_state, float4* __restrict__ posB,
    uint64_t* __restrict__ keys_out, int* __restrict__ cell_start, int n_cells, int cells_per_block,
    int* __restrict__ next_count, const int* __restrict__ step_ctl) {
    if (step_ctl[0]) return;
    const int M = *mcount;
    if (M > RESORT_CAP) return;
    // this thread's key and position, loaded first: their latency overlaps the
    // movers' staging and the searches below
    const int s = blockIdx.x * RESORT_MERGE_BLOCK + (int)threadIdx.x;
    uint64_t y = 0u;
    float4 ps = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (s < n) {
        y = keys[s];
        ps = pos_state[s];
    }
    __shared__ uint64_t sn[RESORT_CAP], so[RESORT_CAP];
    for (int j = threadIdx.x; j < M; j += RESORT_MERGE_BLOCK) {
        sn[j] = msn[j];
        so[j] = mso[j];
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    const int c0 = blockIdx.x * cells_per_block;
    const int c1 = min(c0 + cells_per_block, n_cells);
    if (M > 0 && c0 < c1) {
        // mover keys below the run start (new: alo, old: blo); the runs are
        // consecutive, so each next bound is a short scan from the last one
        int alo = lower_bound_u64(sn, 0, M, (uint64_t)c0 << 32);
        int blo = lower_bound_u64(so, 0, M, (uint64_t)c0 << 32);
        for (int cs = c0; cs < c1; cs += RESORT_MERGE_BLOCK) {
            const int ce = min(cs + RESORT_MERGE_BLOCK, c1);
            const uint64_t khi = (uint64_t)ce << 32;
            int ahi = alo, bhi = blo;
            while (ahi < M && sn[ahi] < khi) ++ahi;
            while (bhi < M && so[bhi] < khi) ++bhi;
            const int c = cs + (int)threadIdx.x;
            if (c < ce) {
                int d = alo - blo;
                if (alo != ahi || blo != bhi) {
                    const uint64_t k = (uint64_t)c << 32;
                    d = lower_bound_u64(sn, alo, ahi, k) - lower_bound_u64(so, blo, bhi, k);
                }
                if (d != 0) cell_start[c] += d;
            }
            alo = ahi;
            blo = bhi;
        }
    }
    if (s >= n) return;
    // movers before s: old keys ascend with their positions (the low halves)
    int lo = 0, hi = M;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)so[mid] < (uint32_t)s) lo = mid + 1;
        else hi = mid;
    }
    if (lo < M && (uint32_t)so[lo] == (uint32_t)s) return;  // a mover, placed by k_resort_movers
    const int a = lower_bound_u64(sn, 0, M, y);
    const int d = s - lo + a;
    if (d < 0 || d >= n) return;  // (never, for consistent keys)
    keys_out[d] = y;
    posB[d] = ps;
}

sd_status launch_resort_u64(const int* mcount, const uint64_t* mnew, const uint64_t* mold, uint64_t* msorted_new,
                            uint64_t* msorted_old, const uint64_t* keys, int n, int* cell_start, int n_cells,
                            const float4* pos_state, float4* posB, uint64_t* keys_out, int* next_count,
                            int* step_ctl, hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_resort_movers, dim3(1), dim3(RESORT_CAP), 0, s, mcount, mnew, mold, msorted_new,
                       msorted_old, n, (const int*)cell_start, pos_state, posB, keys_out, step_ctl);
    SD_LAUNCH_CHECK();
    const int nblocks = div_up(n, RESORT_MERGE_BLOCK);
    const int cells_per_block = div_up(div_up(n_cells, nblocks), RESORT_MERGE_BLOCK) * RESORT_MERGE_BLOCK;
    hipLaunchKernelGGL(k_resort_merge, dim3(nblocks), dim3(RESORT_MERGE_BLOCK), 0, s, mcount,
                       (const uint64_t*)msorted_new, (const uint64_t*)msorted_old, keys, n, pos_state, posB, keys_out,
                       cell_start, n_cells, cells_per_block, next_count, (const int*)step_ctl);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

}  // namespace sd
