// Stable LSD radix sort of (key, value) pairs for gfx950.
//
// Replaces the reference's 16-block, 12-group serialised-LDS sort
// (src/engine/cuda/sh_collision_detection.cu:129-283, driven by :310-360).
// Design (reduce-then-scan, three kernels per pass, no inter-block waiting):
//   1. k_rs_hist    one tile per block; wave64 ballot "match" aggregates equal
//                   digits so a single lane per digit touches LDS; per-tile
//                   digit counts are written digit-major: hist[d * ntiles + t].
//   2. k_rs_scan    one block per digit: exclusive scan of that digit's row over
//                   tiles, row total -> totals[d].
//   3. k_rs_scatter re-reads the tile, ranks every element among equal digits
//                   (stable: waves own contiguous element ranges, items are
//                   ranked in element order), places the tile digit-sorted in
//                   LDS, then writes it out so consecutive threads write
//                   consecutive addresses of each digit bucket.
// Digit width per pass is ceil(bits / passes) with passes = ceil(bits / 8),
// so a 17-bit key costs 3 passes of 6/6/5 bits.
#include <mutex>

#include "sd_internal.hpp"

namespace sd {

constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / WAVE;
constexpr int RS_RADIX = 256;  // largest digit: 8 bits

template <typename K>
__device__ __forceinline__ uint32_t rs_digit(K k, int bit, uint32_t mask) {
    return (uint32_t)(k >> bit) & mask;
}

// Lanes of this wave that hold a valid element with the same digit as this lane.
__device__ __forceinline__ uint64_t rs_match(uint32_t d, bool valid, int nbits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
        const bool set = (d >> b) & 1u;
        const uint64_t bb = __ballot(set);
        m &= set ? bb : ~bb;
    }
    return m;
}

template <typename K, int ITEMS>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const K* __restrict__ keys, int n, int bit, int nbits,
                                                      uint32_t* __restrict__ hist, int ntiles,
                                                      const int* __restrict__ gate, int gate_cap) {
    if (gate && *gate <= gate_cap) return;
    __shared__ uint32_t cnt[RS_RADIX];
    const int radix = 1 << nbits;
    const uint32_t mask = (uint32_t)radix - 1u;
    for (int i = threadIdx.x; i < radix; i += RS_BLOCK) cnt[i] = 0;
    __syncthreads();

    const int wave = threadIdx.x >> 6;
    const long base = (long)blockIdx.x * (RS_BLOCK * ITEMS) + (long)wave * (WAVE * ITEMS) + lane_id();
    K k[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const long i = base + (long)j * WAVE;
        k[j] = (i < n) ? keys[i] : K(0);
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const bool valid = base + (long)j * WAVE < n;
        const uint32_t d = rs_digit(k[j], bit, mask);
        const uint64_t m = rs_match(d, valid, nbits);
        if (valid && (m & lt) == 0) atomicAdd(&cnt[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < radix; i += RS_BLOCK) hist[(long)i * ntiles + blockIdx.x] = cnt[i];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scan(uint32_t* __restrict__ hist, int ntiles,
                                                      uint32_t* __restrict__ totals, const int* __restrict__ gate,
                                                      int gate_cap) {
    if (gate && *gate <= gate_cap) return;
    __shared__ uint32_t wsum[RS_WAVES];
    uint32_t* row = hist + (long)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int b = 0; b < ntiles; b += RS_BLOCK) {
        const int i = b + threadIdx.x;
        const uint32_t v = (i < ntiles) ? row[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, wsum, &tot);
        if (i < ntiles) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

template <typename K, bool HASV, int ITEMS>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                         K* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                                                         int bit, int nbits, const uint32_t* __restrict__ hist,
                                                         const uint32_t* __restrict__ totals, int ntiles,
                                                         const int* __restrict__ gate, int gate_cap) {
    if (gate && *gate <= gate_cap) return;
    constexpr int TILE = RS_BLOCK * ITEMS;
    __shared__ uint32_t wcnt[RS_WAVES][RS_RADIX];
    __shared__ uint32_t tpre[RS_RADIX];
    __shared__ uint32_t gofs[RS_RADIX];
    __shared__ uint32_t wsum[RS_WAVES];
    __shared__ K skey[TILE];
    __shared__ uint32_t sval[HASV ? TILE : 1];

    const int radix = 1 << nbits;
    const uint32_t mask = (uint32_t)radix - 1u;
    const int wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < RS_WAVES * RS_RADIX; i += RS_BLOCK) (&wcnt[0][0])[i] = 0;

    const long tile0 = (long)blockIdx.x * TILE;
    const long base = tile0 + (long)wave * (WAVE * ITEMS) + lane_id();
    K k[ITEMS];
    uint32_t v[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const long i = base + (long)j * WAVE;
        k[j] = (i < n) ? kin[i] : K(0);
        if (HASV) v[j] = (i < n) ? vin[i] : 0u;
    }
    __syncthreads();

    // Rank each item among the wave's earlier items with the same digit.
    uint32_t rank[ITEMS];
    uint32_t dig[ITEMS];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const bool valid = base + (long)j * WAVE < n;
        const uint32_t d = rs_digit(k[j], bit, mask);
        dig[j] = d;
        const uint64_t m = rs_match(d, valid, nbits);
        uint32_t c = 0;
        if (valid) c = wcnt[wave][d];
        if (valid && (m & lt) == 0) wcnt[wave][d] = c + (uint32_t)__popcll(m);
        rank[j] = c + (uint32_t)__popcll(m & lt);
    }
    __syncthreads();

    // Per digit: exclusive prefix over waves (in place) and the tile count.
    for (int d = threadIdx.x; d < radix; d += RS_BLOCK) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = s;
            s += c;
        }
        tpre[d] = s;
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // radix <= RS_BLOCK
        const uint32_t tc = (d < radix) ? tpre[d] : 0u;
        const uint32_t gt = (d < radix) ? totals[d] : 0u;
        uint32_t t1, t2;
        const uint32_t ex_tc = block_exclusive_scan(tc, wsum, &t1);
        const uint32_t ex_gt = block_exclusive_scan(gt, wsum, &t2);
        if (d < radix) {
            tpre[d] = ex_tc;
            gofs[d] = ex_gt + hist[(long)d * ntiles + blockIdx.x];
        }
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (base + (long)j * WAVE < n) {
            const uint32_t local = tpre[dig[j]] + wcnt[wave][dig[j]] + rank[j];
            skey[local] = k[j];
            if (HASV) sval[local] = v[j];
        }
    }
    __syncthreads();

    const long rem = (long)n - tile0;
    const int tile_n = rem < TILE ? (int)rem : TILE;
    for (int i = threadIdx.x; i < tile_n; i += RS_BLOCK) {
        const K kk = skey[i];
        const uint32_t d = rs_digit(kk, bit, mask);
        const uint32_t o = gofs[d] + ((uint32_t)i - tpre[d]);
        kout[o] = kk;
        if (HASV) vout[o] = sval[i];
    }
}

// ------------------------------------------------------------------ host

int radix_passes(int begin_bit, int end_bit) {
    const int bits = end_bit - begin_bit;
    return bits <= 0 ? 0 : (bits + 7) / 8;
}

int radix_default_items(int n) {
    if (n <= (1 << 18)) return 4;
    if (n <= (1 << 21)) return 8;
    return 16;
}

size_t radix_temp_bytes(int n) {
    const long ntiles = n <= 0 ? 1 : div_up(n, RS_BLOCK * 4);
    return ((size_t)RS_RADIX * (size_t)ntiles + RS_RADIX) * sizeof(uint32_t) + 256;
}

template <typename K, bool HASV, int ITEMS>
static sd_status rs_pass(const K* kin, const uint32_t* vin, K* kout, uint32_t* vout, int n, int bit, int nbits,
                         uint32_t* hist, uint32_t* totals, hipStream_t s, const int* gate, int gate_cap) {
    const int ntiles = div_up(n, RS_BLOCK * ITEMS);
    hipLaunchKernelGGL((k_rs_hist<K, ITEMS>), dim3(ntiles), dim3(RS_BLOCK), 0, s, kin, n, bit, nbits, hist, ntiles,
                       gate, gate_cap);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rs_scan, dim3(1 << nbits), dim3(RS_BLOCK), 0, s, hist, ntiles, totals, gate, gate_cap);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_rs_scatter<K, HASV, ITEMS>), dim3(ntiles), dim3(RS_BLOCK), 0, s, kin, vin, kout, vout, n,
                       bit, nbits, hist, totals, ntiles, gate, gate_cap);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

// gate (nullable): every kernel returns at once while *gate <= gate_cap, so a
// captured sort can be switched off on the device. odd_passes: at least one and
// an odd number of passes, so the result always ends in the alternate buffer.
template <typename K, bool HASV>
static sd_status rs_sort(K** keys, uint32_t** vals, K** keys_alt, uint32_t** vals_alt, int n, int begin_bit,
                         int end_bit, void* temp, size_t temp_bytes, hipStream_t s, int items,
                         const int* gate = nullptr, int gate_cap = 0, bool odd_passes = false) {
    SD_REQUIRE(n >= 0, "radix sort: n < 0");
    SD_REQUIRE(begin_bit >= 0 && end_bit <= (int)(8 * sizeof(K)) && begin_bit <= end_bit,
               "radix sort: bad bit range [%d, %d)", begin_bit, end_bit);
    if (n <= 1) return SD_OK;
    SD_REQUIRE(keys && *keys && keys_alt && *keys_alt, "radix sort: null key buffer");
    if (HASV) SD_REQUIRE(vals && *vals && vals_alt && *vals_alt, "radix sort: null value buffer");
    SD_REQUIRE(temp != nullptr && temp_bytes >= radix_temp_bytes(n), "radix sort: temp too small (%zu < %zu)",
               temp_bytes, radix_temp_bytes(n));
    if (items == 0) items = radix_default_items(n);
    SD_REQUIRE(items == 4 || items == 8 || items == 16, "radix sort: items must be 4, 8 or 16");

    const int bits = end_bit - begin_bit;
    int passes = radix_passes(begin_bit, end_bit);
    if (odd_passes && passes % 2 == 0) ++passes;  // (a zero-bit pass is a stable copy)
    uint32_t* hist = (uint32_t*)temp;
    const long ntiles_max = div_up(n, RS_BLOCK * 4);
    uint32_t* totals = hist + RS_RADIX * ntiles_max;

    int bit = begin_bit;
    for (int p = 0; p < passes; ++p) {
        const int nbits = bits / passes + (p < bits % passes ? 1 : 0);
        K* kin = *keys;
        K* kout = *keys_alt;
        uint32_t* vin = HASV ? *vals : nullptr;
        uint32_t* vout = HASV ? *vals_alt : nullptr;
        sd_status st;
        if (items == 4)
            st = rs_pass<K, HASV, 4>(kin, vin, kout, vout, n, bit, nbits, hist, totals, s, gate, gate_cap);
        else if (items == 8)
            st = rs_pass<K, HASV, 8>(kin, vin, kout, vout, n, bit, nbits, hist, totals, s, gate, gate_cap);
        else
            st = rs_pass<K, HASV, 16>(kin, vin, kout, vout, n, bit, nbits, hist, totals, s, gate, gate_cap);
        if (st != SD_OK) return st;
        *keys = kout;
        *keys_alt = kin;
        if (HASV) {
            *vals = vout;
            *vals_alt = vin;
        }
        bit += nbits;
    }
    return SD_OK;
}

sd_status radix_sort_pairs_u32(uint32_t** keys, uint32_t** vals, uint32_t** keys_alt, uint32_t** vals_alt, int n,
                               int begin_bit, int end_bit, void* temp, size_t temp_bytes, hipStream_t s, int items) {
    return rs_sort<uint32_t, true>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, temp, temp_bytes, s,
                                   items);
}

sd_status radix_sort_keys_u64(uint64_t** keys, uint64_t** keys_alt, int n, int begin_bit, int end_bit, void* temp,
                              size_t temp_bytes, hipStream_t s, int items) {
    return rs_sort<uint64_t, false>(keys, nullptr, keys_alt, nullptr, n, begin_bit, end_bit, temp, temp_bytes, s,
                                    items);
}

sd_status radix_sort_keys_u64_gated(uint64_t** keys, uint64_t** keys_alt, int n, int begin_bit, int end_bit,
                                    void* temp, size_t temp_bytes, hipStream_t s, int items, const int* gate,
                                    int gate_cap) {
    return rs_sort<uint64_t, false>(keys, nullptr, keys_alt, nullptr, n, begin_bit, end_bit, temp, temp_bytes, s,
                                    items, gate, gate_cap, true);
}

}  // namespace sd

// ------------------------------------------------------------------ C ABI

extern "C" sd_status sd_radix_sort_temp_bytes(int n, size_t* bytes) {
    SD_REQUIRE(bytes != nullptr && n >= 0, "sd_radix_sort_temp_bytes: bad argument");
    *bytes = sd::radix_temp_bytes(n);
    return SD_OK;
}

extern "C" sd_status sd_radix_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                                             int n, int begin_bit, int end_bit, void* temp, size_t temp_bytes,
                                             void* stream) {
    uint32_t *k = keys, *v = vals, *ka = keys_alt, *va = vals_alt;
    hipStream_t s = (hipStream_t)stream;
    sd_status st = sd::radix_sort_pairs_u32(&k, &v, &ka, &va, n, begin_bit, end_bit, temp, temp_bytes, s, 0);
    if (st != SD_OK) return st;
    if (k != keys) {
        SD_HIP(hipMemcpyAsync(keys, k, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        SD_HIP(hipMemcpyAsync(vals, v, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    return SD_OK;
}

extern "C" sd_status sd_radix_sort_keys_u64(uint64_t* keys, uint64_t* keys_alt, int n, int begin_bit, int end_bit,
                                            void* temp, size_t temp_bytes, void* stream) {
    uint64_t *k = keys, *ka = keys_alt;
    hipStream_t s = (hipStream_t)stream;
    sd_status st = sd::radix_sort_keys_u64(&k, &ka, n, begin_bit, end_bit, temp, temp_bytes, s, 0);
    if (st != SD_OK) return st;
    if (k != keys) SD_HIP(hipMemcpyAsync(keys, k, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    return SD_OK;
}

// ------------------------------------------------------------------ segmented sort
//
// Stable sort of (key, value) pairs within each segment of a CSR partition
// (seg_first[0] = 0 <= ... <= seg_first[nseg] = n): a stable LSD sort by key
// over the whole array, then a stable sort of that order by segment, then one
// gather. The batched step does not call it: there the segment (scene) rides
// in the high bits of one packed u64 key, which sorts every scene in one pass.
namespace sd {

__global__ void k_seg_init(const uint32_t* __restrict__ keys, int n, uint32_t* __restrict__ k,
                           uint32_t* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    k[i] = keys[i];
    v[i] = (uint32_t)i;
}

// Segment of element index e: the last s with seg_first[s] <= e.
__global__ void k_seg_keys(const uint32_t* __restrict__ perm, int n, const int* __restrict__ seg_first, int nseg,
                           uint32_t* __restrict__ sk, uint32_t* __restrict__ sv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int e = (int)perm[j];
    int lo = 0, hi = nseg;  // seg_first[lo] <= e < seg_first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg_first[mid] <= e) lo = mid;
        else hi = mid;
    }
    sk[j] = (uint32_t)lo;
    sv[j] = (uint32_t)e;
}

__global__ void k_seg_gather(const uint32_t* __restrict__ perm, int n, const uint32_t* __restrict__ keys,
                             const uint32_t* __restrict__ vals, uint32_t* __restrict__ ko,
                             uint32_t* __restrict__ vo) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t e = perm[t];
    ko[t] = keys[e];
    vo[t] = vals[e];
}

size_t segmented_temp_bytes(int n) {
    return radix_temp_bytes(n) + 4 * ((size_t)(n > 0 ? n : 1) * sizeof(uint32_t) + 256);
}

}  // namespace sd

extern "C" sd_status sd_segmented_sort_temp_bytes(int n, size_t* bytes) {
    SD_REQUIRE(bytes != nullptr && n >= 0, "sd_segmented_sort_temp_bytes: bad argument");
    *bytes = sd::segmented_temp_bytes(n);
    return SD_OK;
}

extern "C" sd_status sd_segmented_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                                                 uint32_t* vals_alt, int n, const int* seg_first, int nseg,
                                                 int key_bits, void* temp, size_t temp_bytes, void* stream) {
    using namespace sd;
    SD_REQUIRE(n >= 0 && nseg >= 1, "sd_segmented_sort_pairs_u32: bad sizes");
    SD_REQUIRE(key_bits >= 1 && key_bits <= 32, "sd_segmented_sort_pairs_u32: key_bits must be in [1, 32]");
    if (n <= 1) return SD_OK;
    SD_REQUIRE(keys && vals && keys_alt && vals_alt && seg_first && temp, "sd_segmented_sort_pairs_u32: null pointer");
    SD_REQUIRE(temp_bytes >= segmented_temp_bytes(n), "sd_segmented_sort_pairs_u32: temp too small (%zu < %zu)",
               temp_bytes, segmented_temp_bytes(n));
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)temp;
    const size_t rt = radix_temp_bytes(n), nb = ((size_t)n * sizeof(uint32_t) + 255) & ~(size_t)255;
    void* rtemp = p;
    uint32_t* b0 = (uint32_t*)(p + rt);
    uint32_t* b1 = (uint32_t*)(p + rt + nb);
    uint32_t* b2 = (uint32_t*)(p + rt + 2 * nb);
    uint32_t* b3 = (uint32_t*)(p + rt + 3 * nb);
    const dim3 grid(div_up(n, 256)), block(256);
    hipLaunchKernelGGL(k_seg_init, grid, block, 0, s, keys, n, b0, b1);
    SD_LAUNCH_CHECK();
    uint32_t *k = b0, *v = b1, *ka = b2, *va = b3;
    sd_status st = radix_sort_pairs_u32(&k, &v, &ka, &va, n, 0, key_bits, rtemp, rt, s, 0);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_seg_keys, grid, block, 0, s, v, n, seg_first, nseg, ka, va);
    SD_LAUNCH_CHECK();
    int seg_bits = 1;
    while (seg_bits < 31 && (1 << seg_bits) < nseg) ++seg_bits;
    uint32_t *k2 = ka, *v2 = va, *ka2 = k, *va2 = v;
    st = radix_sort_pairs_u32(&k2, &v2, &ka2, &va2, n, 0, seg_bits, rtemp, rt, s, 0);
    if (st != SD_OK) return st;
    hipLaunchKernelGGL(k_seg_gather, grid, block, 0, s, v2, n, keys, vals, keys_alt, vals_alt);
    SD_LAUNCH_CHECK();
    SD_HIP(hipMemcpyAsync(keys, keys_alt, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    SD_HIP(hipMemcpyAsync(vals, vals_alt, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    return SD_OK;
}

// ------------------------------------------------------------------ no-allocation form
//
// The histogram workspace of sd_radix_sort_pairs_u32_noalloc is a file-scope
// device array sized for SD_SORT_NOALLOC_MAX_N elements. Every GPU the library
// runs on holds its own copy until the process ends. Host code never passes
// the array's name to a kernel (in host code it is only a stand-in): its
// device address is looked up once per device with hipGetSymbolAddress and
// kept here.
namespace sd {
constexpr size_t SORT_NOALLOC_BYTES =
    ((size_t)RS_RADIX * (size_t)((SD_SORT_NOALLOC_MAX_N + RS_BLOCK * 4 - 1) / (RS_BLOCK * 4)) + RS_RADIX) *
        sizeof(uint32_t) +
    256;
static __device__ __align__(256) unsigned char scratch[SORT_NOALLOC_BYTES];

constexpr int SORT_NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SORT_NOALLOC_MAX_DEVICES];

// Device address of this device's workspace (looked up on the first call per
// device, with that device current).
static sd_status noalloc_scratch(void** out) {
    int dev = 0;
    SD_HIP(hipGetDevice(&dev));
    SD_REQUIRE(dev >= 0 && dev < SORT_NOALLOC_MAX_DEVICES, "sort workspace: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        SD_REQUIRE(p != nullptr, "sort workspace: no device address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SD_OK;
}
}  // namespace sd

extern "C" sd_status sd_radix_sort_pairs_u32_noalloc(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                                                     uint32_t* vals_alt, int n, int begin_bit, int end_bit,
                                                     void* stream) {
    SD_REQUIRE(n >= 0, "sd_radix_sort_pairs_u32_noalloc: n < 0");
    SD_REQUIRE(n <= SD_SORT_NOALLOC_MAX_N,
               "sd_radix_sort_pairs_u32_noalloc: n = %d exceeds the static workspace limit %d "
               "(use sd_radix_sort_pairs_u32 with a caller-provided temp buffer)",
               n, SD_SORT_NOALLOC_MAX_N);
    static_assert(sd::SORT_NOALLOC_BYTES >= 256, "workspace size");
    if (n <= 1) return SD_OK;
    void* temp = nullptr;
    const sd_status st = sd::noalloc_scratch(&temp);
    if (st != SD_OK) return st;
    // radix_temp_bytes(n) <= SORT_NOALLOC_BYTES for every n <= SD_SORT_NOALLOC_MAX_N
    return sd_radix_sort_pairs_u32(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, temp,
                                   sd::radix_temp_bytes(n), stream);
}
