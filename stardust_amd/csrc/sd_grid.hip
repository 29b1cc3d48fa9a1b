// Uniform-grid kernels: home-cell hash keys, cell start/end ranges (the public
// form) and cell starts (every cell's first sorted position, the form the
// stepping contexts keep), state permutation and neighbour-pair generation,
// plus a device-wide int scan.
//
// Replaces constructCollisionListCUDA (src/engine/cuda/sh_collision_detection.cu:27-127),
// the run detection of tranverseCollisionListCUDA (src/engine/cuda/collision_resolution.cu:100-199)
// and the pair loop of resolveCollisionsCUDA (collision_resolution.cu:56-97).
// The reference emits up to 9 (cell, sphere) entries per sphere (home + phantom
// cells); here every sphere emits one home-cell key and the neighbour search
// walks the 27-cell stencil of the sorted order instead.
#include "sd_internal.hpp"
#include "sd_grid_dev.hpp"

namespace sd {

__global__ void k_fill_u32(uint32_t* __restrict__ dst, uint32_t value, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

__global__ void k_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

sd_status launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t count, hipStream_t s) {
    if (count == 0) return SD_OK;
    const int block = 256;
    long blocks = (long)((count + block - 1) / block);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_copy_u32, dim3((int)blocks), dim3(block), 0, s, dst, src, count);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_fill_u32(uint32_t* dst, uint32_t value, size_t count, hipStream_t s) {
    if (count == 0) return SD_OK;
    const int block = 256;
    long blocks = (long)((count + block - 1) / block);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3((int)blocks), dim3(block), 0, s, dst, value, count);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// ------------------------------------------------------------------ keys

__global__ void k_hash_xyz_u32(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                               int n, GridDev g, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = cell_id(g, x[i], y[i], z[i]);
    vals[i] = (uint32_t)i;
}

__global__ void k_hash_xyz_u64(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                               int n, GridDev g, uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((uint64_t)cell_id(g, x[i], y[i], z[i]) << 32) | (uint64_t)(uint32_t)i;
}

__global__ void k_hash_pos4_u32(const float4* __restrict__ p, int n, GridDev g, uint32_t* __restrict__ keys,
                                uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = p[i];
    keys[i] = cell_id(g, q.x, q.y, q.z);
    vals[i] = (uint32_t)i;
}

__global__ void k_hash_pos4_u64(const float4* __restrict__ p, int n, GridDev g, uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = p[i];
    keys[i] = ((uint64_t)cell_id(g, q.x, q.y, q.z) << 32) | (uint64_t)(uint32_t)i;
}

sd_status launch_hash_keys_u32(const float4* pos4, int n, const GridDev& g, uint32_t* keys, uint32_t* vals,
                               hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_hash_pos4_u32, dim3(div_up(n, 256)), dim3(256), 0, s, pos4, n, g, keys, vals);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_hash_keys_u64(const float4* pos4, int n, const GridDev& g, uint64_t* keys, hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_hash_pos4_u64, dim3(div_up(n, 256)), dim3(256), 0, s, pos4, n, g, keys);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

__global__ void k_hash_pos4_batch(const float4* __restrict__ p, int n, const SceneDev* __restrict__ scenes,
                                  const int* __restrict__ scene_of, uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = p[i];
    const SceneDev& S = scenes[scene_of[i]];
    keys[i] = ((uint64_t)(S.cbase + cell_id(S.g, q.x, q.y, q.z)) << 32) | (uint64_t)(uint32_t)i;
}

sd_status launch_hash_keys_batch(const float4* pos4, int n, const SceneDev* scenes, const int* scene_of,
                                 uint64_t* keys, hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_hash_pos4_batch, dim3(div_up(n, 256)), dim3(256), 0, s, pos4, n, scenes, scene_of, keys);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

int grid_bits(const sd_grid& g) {
    const uint64_t cells = (uint64_t)g.n[0] * (uint64_t)g.n[1] * (uint64_t)g.n[2];
    int b = 0;
    while (b < 32 && (1ull << b) < cells) ++b;
    return b == 0 ? 1 : b;
}

sd_status check_grid(const sd_grid* g) {
    SD_REQUIRE(g != nullptr, "grid is NULL");
    SD_REQUIRE(g->n[0] > 0 && g->n[1] > 0 && g->n[2] > 0, "grid: cell counts must be positive");
    SD_REQUIRE((uint64_t)g->n[0] * g->n[1] * g->n[2] < (1ull << 31), "grid: more than 2^31 cells");
    SD_REQUIRE(g->cell > 0.0f, "grid: cell size must be positive");
    return SD_OK;
}

// ------------------------------------------------------------------ ranges

template <typename K, int SHIFT>
__global__ void k_cell_ranges(const K* __restrict__ keys, int n, int n_cells, int* __restrict__ ranges) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = (uint32_t)(keys[i] >> SHIFT);
    if (c >= (uint32_t)n_cells) return;  // not a cell of this grid
    if (i == 0 || (uint32_t)(keys[i - 1] >> SHIFT) != c) ranges[2 * (size_t)c] = i;
    if (i == n - 1 || (uint32_t)(keys[i + 1] >> SHIFT) != c) ranges[2 * (size_t)c + 1] = i + 1;
}

sd_status launch_cell_ranges_u32(const uint32_t* keys, int n, int n_cells, int* ranges, hipStream_t s) {
    sd_status st = launch_fill_u32((uint32_t*)ranges, 0u, 2 * (size_t)n_cells, s);
    if (st != SD_OK || n <= 0) return st;
    hipLaunchKernelGGL((k_cell_ranges<uint32_t, 0>), dim3(div_up(n, 256)), dim3(256), 0, s, keys, n, n_cells, ranges);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_cell_ranges_u64(const uint64_t* keys, int n, int n_cells, int* ranges, hipStream_t s) {
    sd_status st = launch_fill_u32((uint32_t*)ranges, 0u, 2 * (size_t)n_cells, s);
    if (st != SD_OK || n <= 0) return st;
    hipLaunchKernelGGL((k_cell_ranges<uint64_t, 32>), dim3(div_up(n, 256)), dim3(256), 0, s, keys, n, n_cells,
                       ranges);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// First sorted position of every cell: start[c] = the number of keys whose
// cell is below c, for c = 0 .. n_cells (start[n_cells] = n), so that cells
// a .. b hold the sorted positions [start[a], start[b + 1]) whether or not
// they are empty. Key i writes the entries of the cells after key i-1's cell
// up to its own; the last key also those after its cell. Every entry is
// written once per step. Up to CELL_GAP_INLINE entries are written by the
// key's lane, longer gaps (empty stretches of a sparse grid) by its whole wave.
constexpr int CELL_GAP_INLINE = 8;

__device__ __forceinline__ void fill_cell_starts(int* __restrict__ start, int lo, int hi, int val) {
    const int e = min(hi, lo + CELL_GAP_INLINE - 1);
    for (int c = lo; c <= e; ++c) start[c] = val;
    uint64_t m = __ballot(hi > e);  // lanes with a longer gap (all lanes take part)
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1u;
        const int a = __builtin_amdgcn_readlane(e + 1, l);
        const int b = __builtin_amdgcn_readlane(hi, l);
        const int v = __builtin_amdgcn_readlane(val, l);
        for (int c = a + lane_id(); c <= b; c += WAVE) start[c] = v;
    }
}

// The cell starts of a step's sorted keys, fused with what else reads them:
// with permute, the positions gathered into sorted order (the re-sort does it
// otherwise); next_count (if given) set to 0. Does nothing while *skip (the
// step was given up). Cell ids at or past n_cells (never written by this
// library) count as n_cells.
template <typename K, int SHIFT>
__global__ void k_cell_starts(const K* __restrict__ keys, int n, int n_cells, int* __restrict__ start,
                              const int* __restrict__ skip, int permute, const float4* __restrict__ pos,
                              float4* __restrict__ pos_o, int* __restrict__ next_count) {
    if (skip && *skip) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && next_count) *next_count = 0;
    int lo = 0, hi = -1, tail_lo = 0, tail_hi = -1;  // every lane stays for the wave-wide fills
    if (i < n) {
        const K k = keys[i];
        if (permute) {
            const uint32_t src = (uint32_t)k;
            if (src < (uint32_t)n) pos_o[i] = pos[src];
        }
        const uint32_t c = min((uint32_t)(k >> SHIFT), (uint32_t)n_cells);
        const uint32_t prev = i == 0 ? 0u : min((uint32_t)(keys[i - 1] >> SHIFT), (uint32_t)n_cells) + 1u;
        lo = (int)prev;
        hi = (int)c;
        if (i == n - 1) {
            tail_lo = (int)c + 1;
            tail_hi = n_cells;
        }
    }
    fill_cell_starts(start, lo, hi, i);
    fill_cell_starts(start, tail_lo, tail_hi, n);
}

template <typename K, int SHIFT>
static sd_status launch_cell_starts(const K* keys, int n, int n_cells, int* start, const int* skip, bool permute,
                                    const float4* pos, float4* pos_o, int* next_count, hipStream_t s) {
    if (n <= 0) return launch_fill_u32((uint32_t*)start, 0u, (size_t)n_cells + 1, s);
    hipLaunchKernelGGL((k_cell_starts<K, SHIFT>), dim3(div_up(n, 256)), dim3(256), 0, s, keys, n, n_cells, start,
                       skip, permute ? 1 : 0, pos, pos_o, next_count);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_cell_starts_u32(const uint32_t* keys, int n, int n_cells, int* start, hipStream_t s) {
    return launch_cell_starts<uint32_t, 0>(keys, n, n_cells, start, nullptr, false, nullptr, nullptr, nullptr, s);
}

sd_status launch_step_starts_u64(const uint64_t* keys, int n, int n_cells, int* start, const int* skip, bool permute,
                                 const float4* pos, float4* pos_o, int* next_count, hipStream_t s) {
    return launch_cell_starts<uint64_t, 32>(keys, n, n_cells, start, skip, permute, pos, pos_o, next_count, s);
}

// ------------------------------------------------------------------ permute

__global__ void k_permute(const uint32_t* __restrict__ perm32, const uint64_t* __restrict__ perm64, int n,
                          const float4* __restrict__ pos, const float4* __restrict__ vel,
                          const float4* __restrict__ omg, const int* __restrict__ ids, float4* __restrict__ pos_o,
                          float4* __restrict__ vel_o, float4* __restrict__ omg_o, int* __restrict__ ids_o) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t src = perm32 ? perm32[t] : (uint32_t)perm64[t];
    if (src >= (uint32_t)n) return;  // not a permutation of 0..n-1: leave the row untouched
    pos_o[t] = pos[src];
    vel_o[t] = vel[src];
    omg_o[t] = omg[src];
    ids_o[t] = ids[src];
}

sd_status launch_permute(const uint32_t* perm32, const uint64_t* perm64, int n, const float4* pos, const float4* vel,
                         const float4* omg, const int* ids, float4* pos_o, float4* vel_o, float4* omg_o, int* ids_o,
                         hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_permute, dim3(div_up(n, 256)), dim3(256), 0, s, perm32, perm64, n, pos, vel, omg, ids, pos_o,
                       vel_o, omg_o, ids_o);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

__global__ void k_permute_pos(const uint32_t* __restrict__ perm32, const uint64_t* __restrict__ perm64, int n,
                              const float4* __restrict__ pos, float4* __restrict__ pos_o,
                              const int* __restrict__ gate, int gate_cap) {
    if (gate && *gate <= gate_cap) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t src = perm32 ? perm32[t] : (uint32_t)perm64[t];
    if (src >= (uint32_t)n) return;
    pos_o[t] = pos[src];
}

sd_status launch_permute_pos(const uint32_t* perm32, const uint64_t* perm64, int n, const float4* pos, float4* pos_o,
                             hipStream_t s, const int* gate, int gate_cap) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_permute_pos, dim3(div_up(n, 256)), dim3(256), 0, s, perm32, perm64, n, pos, pos_o, gate,
                       gate_cap);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// ------------------------------------------------------------------ scan

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Phase 1: per-tile sums.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const int* __restrict__ in, int n, int* __restrict__ sums) {
    __shared__ uint32_t wsum[SCAN_BLOCK / WAVE];
    const long b = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
        if (b + j < n) acc += (uint32_t)in[b + j];
    uint32_t tot;
    block_exclusive_scan(acc, wsum, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = (int)tot;
}

// Phase 2: single block scans the tile sums (exclusive) and writes the total.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int* __restrict__ sums, int nt, int* __restrict__ total) {
    __shared__ uint32_t wsum[SCAN_BLOCK / WAVE];
    uint32_t carry = 0;
    for (int b = 0; b < nt; b += SCAN_BLOCK) {
        const int i = b + threadIdx.x;
        const uint32_t v = i < nt ? (uint32_t)sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, wsum, &tot);
        if (i < nt) sums[i] = (int)(carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = (int)carry;
}

// Phase 3: each tile scans its values and adds the tile offset.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const int* __restrict__ in, int* __restrict__ out, int n,
                                                           const int* __restrict__ sums) {
    __shared__ uint32_t wsum[SCAN_BLOCK / WAVE];
    const long b = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = (b + j < n) ? (uint32_t)in[b + j] : 0u;
        acc += v[j];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(acc, wsum, &tot) + (uint32_t)sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        if (b + j < n) out[b + j] = (int)run;
        run += v[j];
    }
}

size_t scan_temp_bytes(int n) { return ((size_t)div_up(n > 0 ? n : 1, SCAN_TILE) + 1) * sizeof(int); }

sd_status exclusive_scan_i32(const int* in, int* out, int n, int* total, void* scratch, hipStream_t s) {
    if (n <= 0) {
        if (total) SD_HIP(hipMemsetAsync(total, 0, sizeof(int), s));
        return SD_OK;
    }
    const int nt = div_up(n, SCAN_TILE);
    int* sums = (int*)scratch;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SCAN_BLOCK), 0, s, in, n, sums);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nt, total);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(SCAN_BLOCK), 0, s, in, out, n, sums);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// ------------------------------------------------------------------ pairs

// WRITE = false: count pairs per sorted position; WRITE = true: emit them.
template <bool WRITE>
__global__ void k_neighbor_pairs(const float4* __restrict__ pos, int n, GridDev g, const int* __restrict__ ranges,
                                 float skin, int* __restrict__ counts, const int* __restrict__ offsets,
                                 int* __restrict__ pairs, int capacity) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const float4 p = pos[a];
    int cx, cy, cz;
    cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
    int cnt = 0;
    int o = WRITE ? offsets[a] : 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= g.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.ny) continue;
            int lo, hi;
            row_range(g, ranges, cx, y, z, lo, hi);
            for (int b = lo > a + 1 ? lo : a + 1; b < hi; ++b) {
                const float4 q = pos[b];
                const float dx = p.x - q.x, dy2 = p.y - q.y, dz2 = p.z - q.z;
                const float d2 = dx * dx + dy2 * dy2 + dz2 * dz2;
                const float rs = p.w + q.w + skin;
                if (d2 < rs * rs) {
                    if (WRITE && o < capacity) {
                        pairs[2 * (size_t)o] = a;
                        pairs[2 * (size_t)o + 1] = b;
                    }
                    ++o;
                    ++cnt;
                }
            }
        }
    }
    if (!WRITE) counts[a] = cnt;
}

}  // namespace sd

// ------------------------------------------------------------------ C ABI

using namespace sd;

extern "C" sd_status sd_hash_keys_u32(const float* x, const float* y, const float* z, int n, const sd_grid* grid,
                                      uint32_t* keys, uint32_t* vals, void* stream) {
    sd_status st = check_grid(grid);
    if (st != SD_OK) return st;
    SD_REQUIRE(n >= 0, "sd_hash_keys_u32: n < 0");
    if (n == 0) return SD_OK;
    SD_REQUIRE(x && y && z && keys && vals, "sd_hash_keys_u32: null pointer");
    hipLaunchKernelGGL(k_hash_xyz_u32, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, x, y, z, n,
                       to_dev(*grid), keys, vals);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_hash_keys_u64(const float* x, const float* y, const float* z, int n, const sd_grid* grid,
                                      uint64_t* keys, void* stream) {
    sd_status st = check_grid(grid);
    if (st != SD_OK) return st;
    SD_REQUIRE(n >= 0, "sd_hash_keys_u64: n < 0");
    if (n == 0) return SD_OK;
    SD_REQUIRE(x && y && z && keys, "sd_hash_keys_u64: null pointer");
    hipLaunchKernelGGL(k_hash_xyz_u64, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, x, y, z, n,
                       to_dev(*grid), keys);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_cell_ranges_u32(const uint32_t* sorted_keys, int n, int n_cells, int* ranges, void* stream) {
    SD_REQUIRE(n >= 0 && n_cells > 0 && ranges && (n == 0 || sorted_keys), "sd_cell_ranges_u32: bad argument");
    return launch_cell_ranges_u32(sorted_keys, n, n_cells, ranges, (hipStream_t)stream);
}

extern "C" sd_status sd_cell_ranges_u64(const uint64_t* sorted_keys, int n, int n_cells, int* ranges, void* stream) {
    SD_REQUIRE(n >= 0 && n_cells > 0 && ranges && (n == 0 || sorted_keys), "sd_cell_ranges_u64: bad argument");
    return launch_cell_ranges_u64(sorted_keys, n, n_cells, ranges, (hipStream_t)stream);
}

extern "C" sd_status sd_permute_spheres(const int* perm, int n, const float* pos4, const float* vel4,
                                        const float* omg4, const int* ids, float* pos4_out, float* vel4_out,
                                        float* omg4_out, int* ids_out, void* stream) {
    SD_REQUIRE(n >= 0, "sd_permute_spheres: n < 0");
    if (n == 0) return SD_OK;
    SD_REQUIRE(perm && pos4 && vel4 && omg4 && ids && pos4_out && vel4_out && omg4_out && ids_out,
               "sd_permute_spheres: null pointer");
    return launch_permute((const uint32_t*)perm, nullptr, n, (const float4*)pos4, (const float4*)vel4,
                          (const float4*)omg4, ids, (float4*)pos4_out, (float4*)vel4_out, (float4*)omg4_out, ids_out,
                          (hipStream_t)stream);
}

extern "C" sd_status sd_neighbor_pairs_scratch_bytes(int n, size_t* bytes) {
    SD_REQUIRE(bytes && n >= 0, "sd_neighbor_pairs_scratch_bytes: bad argument");
    *bytes = (size_t)(n + 2) * sizeof(int) * 2 + scan_temp_bytes(n) + 256;
    return SD_OK;
}

extern "C" sd_status sd_neighbor_pairs(const float* pos4, int n, const sd_grid* grid, const int* ranges, float skin,
                                       int* pairs, int capacity, int* n_pairs, void* scratch, size_t scratch_bytes,
                                       void* stream) {
    sd_status st = check_grid(grid);
    if (st != SD_OK) return st;
    SD_REQUIRE(n >= 0 && capacity >= 0 && n_pairs, "sd_neighbor_pairs: bad argument");
    if (n == 0) {
        *n_pairs = 0;
        return SD_OK;
    }
    SD_REQUIRE(pos4 && ranges && scratch && (pairs || capacity == 0), "sd_neighbor_pairs: null pointer");
    size_t need;
    sd_neighbor_pairs_scratch_bytes(n, &need);
    SD_REQUIRE(scratch_bytes >= need, "sd_neighbor_pairs: scratch too small (%zu < %zu)", scratch_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    int* counts = (int*)scratch;
    int* offsets = counts + (n + 2);
    int* total = offsets + n;
    void* scan_tmp = (void*)(counts + 2 * (size_t)(n + 2));
    const GridDev g = to_dev(*grid);
    hipLaunchKernelGGL(k_neighbor_pairs<false>, dim3(div_up(n, 128)), dim3(128), 0, s, (const float4*)pos4, n, g,
                       ranges, skin, counts, nullptr, nullptr, 0);
    SD_LAUNCH_CHECK();
    st = exclusive_scan_i32(counts, offsets, n, total, scan_tmp, s);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_neighbor_pairs<true>, dim3(div_up(n, 128)), dim3(128), 0, s, (const float4*)pos4, n, g,
                       ranges, skin, counts, offsets, pairs, capacity);
    SD_LAUNCH_CHECK();
    int h_total = 0;
    SD_HIP(hipMemcpyAsync(&h_total, total, sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    *n_pairs = h_total;
    if (h_total > capacity)
        return set_error(SD_ERR_CAPACITY, "sd_neighbor_pairs: %d pairs exceed capacity %d", h_total, capacity);
    return SD_OK;
}
