// Learned GJK warm start: pair features, MLP training (forward, backward,
// Adam) in FP32 and inference as BF16 MFMA GEMMs with FP32 accumulation.
//
// No counterpart in the reference (its GJK starts from the centre difference,
// sandbox/testGJK.cu:266). The network maps a pair's relative pose in A's frame
// plus a shape descriptor of both hulls to the unit direction from A to B that
// plain GJK converges to; that direction seeds a warm-started query (sd_gjk).
//
// Network: 22 -> 64 -> 64 -> 64 -> 3, SiLU hidden activations, output
// normalised to a unit vector. Parameters (FP32 master copy, 9,987 floats):
//   W1[22][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][3] b4[3]
// (W[i][j]: input i -> output j). Loss per sample: 1 - u . t (u = y/|y|).
#include "sd_gjk_dev.hpp"
#include "sd_internal.hpp"

namespace sd {

constexpr int NIN = 22, NH = 64, NOUT = 3;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH * NOUT,
              NPARAM = OFF_B4 + NOUT;
static_assert(NPARAM == 9987, "parameter count");
constexpr int DESC = 7;

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// ------------------------------------------------------------------ features

// Per hull: local half extents |v_x|max, |v_y|max, |v_z|max, mean / max / min
// vertex radius, vertex count / 32 (a sphere: radius in every slot, count 0).
__global__ void k_hull_desc(const float4* __restrict__ pos, const int* __restrict__ vstart,
                            const int* __restrict__ nv, const float4* __restrict__ verts, int n,
                            float* __restrict__ desc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* d = desc + (size_t)i * DESC;
    const int k0 = vstart[i], kn = nv[i];
    if (kn == 0) {
        const float r = pos[i].w;
        d[0] = r;
        d[1] = r;
        d[2] = r;
        d[3] = r;
        d[4] = r;
        d[5] = r;
        d[6] = 0.0f;
        return;
    }
    float ex = 0.0f, ey = 0.0f, ez = 0.0f, rs = 0.0f, rmax = 0.0f, rmin = 3.402823466e38f;
    for (int k = 0; k < kn; ++k) {
        const float4 v = verts[k0 + k];
        ex = fmaxf(ex, fabsf(v.x));
        ey = fmaxf(ey, fabsf(v.y));
        ez = fmaxf(ez, fabsf(v.z));
        const float r = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        rs = rs + r;
        rmax = fmaxf(rmax, r);
        rmin = fminf(rmin, r);
    }
    d[0] = ex;
    d[1] = ey;
    d[2] = ez;
    d[3] = rs / (float)kn;
    d[4] = rmax;
    d[5] = rmin;
    d[6] = (float)kn * (1.0f / 32.0f);
}

// Features of pair (a, b), all lengths divided by s = rmax_a + rmax_b:
//   [0..2]  R_a^T (c_b - c_a) / s
//   [3..6]  q_a^-1 q_b (w, x, y, z), sign chosen so that w >= 0
//   [7..13] descriptor of a (lengths / s), [14..20] descriptor of b,
//   [21]    |c_b - c_a| / s
// dcount (nullable): the pairs held on the device (the first min(*dcount,
// npairs)), for launches sized to a capacity.
__global__ void k_mlp_features(const int* __restrict__ pairs, int npairs, const float4* __restrict__ pos,
                               const float4* __restrict__ quat, const float* __restrict__ desc,
                               float* __restrict__ X, const int* __restrict__ dcount) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (dcount ? min(*dcount, npairs) : npairs)) return;
    const int a = pairs[2 * (size_t)p], b = pairs[2 * (size_t)p + 1];
    const float4 ca = pos[a], cb = pos[b];
    const float4 qa = quat[a], qb = quat[b];
    const float* da = desc + (size_t)a * DESC;
    const float* db = desc + (size_t)b * DESC;
    const float s = da[4] + db[4];
    const float is = s > 0.0f ? 1.0f / s : 1.0f;
    const V3 dc = v3(cb.x - ca.x, cb.y - ca.y, cb.z - ca.z);
    const V3 pl = rot_tmul(quat_rot(qa), dc);
    // conj(qa) * qb
    const float aw = qa.x, ax = -qa.y, ay = -qa.z, az = -qa.w;
    const float bw = qb.x, bx = qb.y, by = qb.z, bz = qb.w;
    float rw = aw * bw - ax * bx - ay * by - az * bz;
    float rx = aw * bx + ax * bw + ay * bz - az * by;
    float ry = aw * by - ax * bz + ay * bw + az * bx;
    float rz = aw * bz + ax * by - ay * bx + az * bw;
    if (rw < 0.0f) {
        rw = -rw;
        rx = -rx;
        ry = -ry;
        rz = -rz;
    }
    float* x = X + (size_t)p * NIN;
    x[0] = pl.x * is;
    x[1] = pl.y * is;
    x[2] = pl.z * is;
    x[3] = rw;
    x[4] = rx;
    x[5] = ry;
    x[6] = rz;
    for (int k = 0; k < 6; ++k) {
        x[7 + k] = da[k] * is;
        x[14 + k] = db[k] * is;
    }
    x[13] = da[6];
    x[20] = db[6];
    x[21] = sqrtf(dot(dc, dc)) * is;
}

// World normal (a -> b) of each pair expressed in a's frame.
__global__ void k_mlp_targets(const int* __restrict__ pairs, int npairs, const float4* __restrict__ quat,
                              const float* __restrict__ normal3, float* __restrict__ T) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const int a = pairs[2 * (size_t)p];
    const V3 n = v3(normal3[3 * (size_t)p], normal3[3 * (size_t)p + 1], normal3[3 * (size_t)p + 2]);
    const V3 l = rot_tmul(quat_rot(quat[a]), n);
    T[3 * (size_t)p] = l.x;
    T[3 * (size_t)p + 1] = l.y;
    T[3 * (size_t)p + 2] = l.z;
}

// ------------------------------------------------------------------ init

// Uniform(-sqrt(6/fan_in)/2, sqrt(6/fan_in)/2) weights from SplitMix64(seed + index), zero biases.
__global__ void k_mlp_init(float* __restrict__ w, uint64_t seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    int fan_in;
    bool bias;
    if (i < OFF_B1) {
        fan_in = NIN;
        bias = false;
    } else if (i < OFF_W2) {
        fan_in = NIN;
        bias = true;
    } else if (i < OFF_B2) {
        fan_in = NH;
        bias = false;
    } else if (i < OFF_W3) {
        fan_in = NH;
        bias = true;
    } else if (i < OFF_B3) {
        fan_in = NH;
        bias = false;
    } else if (i < OFF_W4) {
        fan_in = NH;
        bias = true;
    } else if (i < OFF_B4) {
        fan_in = NH;
        bias = false;
    } else {
        fan_in = NH;
        bias = true;
    }
    if (bias) {
        w[i] = 0.0f;
        return;
    }
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
    const float lim = sqrtf(6.0f / (float)fan_in) * 0.5f;
    w[i] = (2.0f * u - 1.0f) * lim;
}

// ------------------------------------------------------------------ training

constexpr int TB_S = 64;       // samples per block
constexpr int TB_T = 256;      // threads per block
constexpr int XS = NIN + 2;    // padded row of the input tile

// Forward + backward of one block of samples; writes this block's gradient
// (summed over its samples, already divided by the batch size) and loss sum.
// Sample s of the block is dataset row idx[blk * 64 + s] (rows past `count`
// contribute nothing).
__global__ __launch_bounds__(TB_T) void k_mlp_fwd_bwd(const float* __restrict__ X, const float* __restrict__ T,
                                                     const int* __restrict__ idx, int count, float inv_batch,
                                                     const float* __restrict__ w, float* __restrict__ gpart,
                                                     float* __restrict__ lpart) {
    __shared__ float sx[TB_S][XS];
    __shared__ float z1[TB_S][NH], h1[TB_S][NH], z2[TB_S][NH], h2[TB_S][NH], z3[TB_S][NH], h3[TB_S][NH];
    __shared__ float g[TB_S][NH];
    __shared__ float sy[TB_S][4], st[TB_S][4], sdy[TB_S][4], sl[TB_S];
    __shared__ int valid[TB_S];
    const int tid = threadIdx.x;
    const int base = blockIdx.x * TB_S;
    for (int e = tid; e < TB_S * NIN; e += TB_T) {
        const int s = e / NIN, i = e % NIN;
        const int r = base + s < count ? idx[base + s] : -1;
        sx[s][i] = r >= 0 ? X[(size_t)r * NIN + i] : 0.0f;
    }
    for (int e = tid; e < TB_S * NOUT; e += TB_T) {
        const int s = e / NOUT, k = e % NOUT;
        const int r = base + s < count ? idx[base + s] : -1;
        st[s][k] = r >= 0 ? T[(size_t)r * NOUT + k] : 0.0f;
    }
    if (tid < TB_S) valid[tid] = base + tid < count ? 1 : 0;
    __syncthreads();

    // thread -> (output column j, 16 samples)
    const int j = tid & (NH - 1);
    const int s0 = (tid >> 6) * 16;
    float acc[16];
    // layer 1
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = w[OFF_B1 + j];
    for (int i = 0; i < NIN; ++i) {
        const float wij = w[OFF_W1 + i * NH + j];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = acc[q] + sx[s0 + q][i] * wij;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        z1[s0 + q][j] = acc[q];
        h1[s0 + q][j] = acc[q] * sigm(acc[q]);
    }
    __syncthreads();
    // layer 2
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = w[OFF_B2 + j];
    for (int i = 0; i < NH; ++i) {
        const float wij = w[OFF_W2 + i * NH + j];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = acc[q] + h1[s0 + q][i] * wij;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        z2[s0 + q][j] = acc[q];
        h2[s0 + q][j] = acc[q] * sigm(acc[q]);
    }
    __syncthreads();
    // layer 3
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = w[OFF_B3 + j];
    for (int i = 0; i < NH; ++i) {
        const float wij = w[OFF_W3 + i * NH + j];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = acc[q] + h2[s0 + q][i] * wij;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        z3[s0 + q][j] = acc[q];
        h3[s0 + q][j] = acc[q] * sigm(acc[q]);
    }
    __syncthreads();
    // output layer + loss + dL/dy
    if (tid < TB_S * NOUT) {
        const int s = tid / NOUT, k = tid % NOUT;
        float y = w[OFF_B4 + k];
        for (int i = 0; i < NH; ++i) y = y + h3[s][i] * w[OFF_W4 + i * NOUT + k];
        sy[s][k] = y;
    }
    __syncthreads();
    if (tid < TB_S) {
        const int s = tid;
        const float y0 = sy[s][0], y1 = sy[s][1], y2 = sy[s][2];
        const float ny = sqrtf(y0 * y0 + y1 * y1 + y2 * y2) + 1e-12f;
        const float u0 = y0 / ny, u1 = y1 / ny, u2 = y2 / ny;
        const float ut = u0 * st[s][0] + u1 * st[s][1] + u2 * st[s][2];
        const float sc = valid[s] ? inv_batch : 0.0f;
        sl[s] = valid[s] ? 1.0f - ut : 0.0f;
        // d(1 - u.t)/dy = -(t - (u.t) u) / |y|
        sdy[s][0] = -(st[s][0] - ut * u0) / ny * sc;
        sdy[s][1] = -(st[s][1] - ut * u1) / ny * sc;
        sdy[s][2] = -(st[s][2] - ut * u2) / ny * sc;
    }
    __syncthreads();
    float* gp = gpart + (size_t)blockIdx.x * NPARAM;
    // dW4, db4
    if (tid < NH * NOUT) {
        const int i = tid / NOUT, k = tid % NOUT;
        float a = 0.0f;
        for (int s = 0; s < TB_S; ++s) a = a + h3[s][i] * sdy[s][k];
        gp[OFF_W4 + i * NOUT + k] = a;
    }
    if (tid < NOUT) {
        float a = 0.0f;
        for (int s = 0; s < TB_S; ++s) a = a + sdy[s][tid];
        gp[OFF_B4 + tid] = a;
    }
    // dz3 = (dy W4^T) * silu'(z3) -> g
    for (int e = tid; e < TB_S * NH; e += TB_T) {
        const int s = e / NH, i = e % NH;
        const float dh = sdy[s][0] * w[OFF_W4 + i * NOUT] + sdy[s][1] * w[OFF_W4 + i * NOUT + 1] +
                         sdy[s][2] * w[OFF_W4 + i * NOUT + 2];
        const float z = z3[s][i], sg = sigm(z);
        g[s][i] = dh * (sg * (1.0f + z * (1.0f - sg)));
    }
    __syncthreads();
    // dW3 = h2^T g, db3 ; then dz2 = (g W3^T) * silu'(z2) -> h3 (free now)
    {
        const int i0 = (tid >> 6) * 16;  // 16 input rows per thread, column j
        for (int q = 0; q < 16; ++q) {
            float a = 0.0f;
            for (int s = 0; s < TB_S; ++s) a = a + h2[s][i0 + q] * g[s][j];
            gp[OFF_W3 + (i0 + q) * NH + j] = a;
        }
        if (tid < NH) {
            float a = 0.0f;
            for (int s = 0; s < TB_S; ++s) a = a + g[s][tid];
            gp[OFF_B3 + tid] = a;
        }
    }
    for (int e = tid; e < TB_S * NH; e += TB_T) {
        const int s = e / NH, i = e % NH;
        float dh = 0.0f;
        for (int k = 0; k < NH; ++k) dh = dh + g[s][k] * w[OFF_W3 + i * NH + k];
        const float z = z2[s][i], sg = sigm(z);
        h3[s][i] = dh * (sg * (1.0f + z * (1.0f - sg)));
    }
    __syncthreads();
    // dW2 = h1^T dz2, db2 ; dz1 = (dz2 W2^T) * silu'(z1) -> g
    {
        const int i0 = (tid >> 6) * 16;
        for (int q = 0; q < 16; ++q) {
            float a = 0.0f;
            for (int s = 0; s < TB_S; ++s) a = a + h1[s][i0 + q] * h3[s][j];
            gp[OFF_W2 + (i0 + q) * NH + j] = a;
        }
        if (tid < NH) {
            float a = 0.0f;
            for (int s = 0; s < TB_S; ++s) a = a + h3[s][tid];
            gp[OFF_B2 + tid] = a;
        }
    }
    __syncthreads();  // g is rewritten below; everyone finished reading it above
    for (int e = tid; e < TB_S * NH; e += TB_T) {
        const int s = e / NH, i = e % NH;
        float dh = 0.0f;
        for (int k = 0; k < NH; ++k) dh = dh + h3[s][k] * w[OFF_W2 + i * NH + k];
        const float z = z1[s][i], sg = sigm(z);
        g[s][i] = dh * (sg * (1.0f + z * (1.0f - sg)));
    }
    __syncthreads();
    // dW1 = x^T dz1, db1
    for (int e = tid; e < NIN * NH; e += TB_T) {
        const int i = e / NH, jj = e % NH;
        float a = 0.0f;
        for (int s = 0; s < TB_S; ++s) a = a + sx[s][i] * g[s][jj];
        gp[OFF_W1 + i * NH + jj] = a;
    }
    if (tid < NH) {
        float a = 0.0f;
        for (int s = 0; s < TB_S; ++s) a = a + g[s][tid];
        gp[OFF_B1 + tid] = a;
    }
    if (tid == 0) {
        float a = 0.0f;
        for (int s = 0; s < TB_S; ++s) a = a + sl[s];
        lpart[blockIdx.x] = a;
    }
}

// grad[p] = sum over blocks (fixed order); loss = sum of block losses * inv_batch.
__global__ void k_mlp_reduce(const float* __restrict__ gpart, const float* __restrict__ lpart, int nblocks,
                             float inv_batch, float* __restrict__ grad, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) {
        float a = 0.0f;
        for (int b = 0; b < nblocks; ++b) a = a + gpart[(size_t)b * NPARAM + p];
        grad[p] = a;
    }
    if (p == NPARAM && loss) {
        float a = 0.0f;
        for (int b = 0; b < nblocks; ++b) a = a + lpart[b];
        *loss = a * inv_batch;
    }
}

// Adam with bias correction: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// w -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------ BF16 MFMA inference

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> BF16 bits.
__device__ __forceinline__ unsigned short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short b) { return __uint_as_float((uint32_t)b << 16); }

// Packed inference weights: BF16 matrices with K padded (layer 1: 22 -> 32,
// output: N 3 -> 16), row-major [K][N], then FP32 biases.
constexpr int K1 = 32, N4 = 16;
constexpr int PK_W1 = 0, PK_W2 = PK_W1 + K1 * NH, PK_W3 = PK_W2 + NH * NH, PK_W4 = PK_W3 + NH * NH,
              PK_END = PK_W4 + NH * N4;  // bf16 elements
constexpr int PK_BYTES = PK_END * 2 + (3 * NH + N4) * 4;

__global__ void k_mlp_pack(const float* __restrict__ w, unsigned short* __restrict__ wb, float* __restrict__ bias) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < K1 * NH) {
        const int k = e / NH, n = e % NH;
        wb[PK_W1 + e] = to_bf16(k < NIN ? w[OFF_W1 + k * NH + n] : 0.0f);
    }
    if (e < NH * NH) {
        wb[PK_W2 + e] = to_bf16(w[OFF_W2 + e]);
        wb[PK_W3 + e] = to_bf16(w[OFF_W3 + e]);
    }
    if (e < NH * N4) {
        const int k = e / N4, n = e % N4;
        wb[PK_W4 + e] = to_bf16(n < NOUT ? w[OFF_W4 + k * NOUT + n] : 0.0f);
    }
    if (e < NH) {
        bias[e] = w[OFF_B1 + e];
        bias[NH + e] = w[OFF_B2 + e];
        bias[2 * NH + e] = w[OFF_B3 + e];
    }
    if (e < N4) bias[3 * NH + e] = e < NOUT ? w[OFF_B4 + e] : 0.0f;
}

constexpr int INF_WAVES = 4;
constexpr int INF_ROWS = 16;  // samples per wave per tile

// v_mfma_f32_16x16x16_bf16 operand layout (wave64): lane l holds
//   A: row l % 16, k = 4 (l / 16) .. + 3      B: k = 4 (l / 16) .. + 3, column l % 16
//   C/D: rows 4 (l / 16) .. + 3, column l % 16
// One wave pushes 16 samples through the whole network; activations go through
// LDS (as BF16) to turn the D layout into the next layer's A layout. The block
// keeps the weights in LDS transposed, [n][k] with rows padded by 4, so that a
// lane's four k values of A and of B are each one 8-byte LDS read.
constexpr int WT_K1 = K1 + 4, WT_NH = NH + 4;  // transposed row lengths (bf16)
constexpr int WT_1 = 0, WT_2 = WT_1 + NH * WT_K1, WT_3 = WT_2 + NH * WT_NH, WT_4 = WT_3 + NH * WT_NH,
              WT_END = WT_4 + N4 * WT_NH;
static_assert(PK_END % 8 == 0 && (3 * NH + N4) % 4 == 0, "weights are staged in 16-byte pieces");

// SiLU of the inference path: hardware exp and reciprocal (a few ulp of FP32,
// far below the BF16 rounding of the activations it feeds)
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// dcount (nullable): rows held on the device (the first min(*dcount, n_rows)).
__global__ __launch_bounds__(INF_WAVES * 64) void k_mlp_infer(const float* __restrict__ X, int n_rows,
                                                              const unsigned short* __restrict__ wb_g,
                                                              const float* __restrict__ bias_g,
                                                              const int* __restrict__ pairs,
                                                              const float4* __restrict__ quat,
                                                              float* __restrict__ out_dir,
                                                              const int* __restrict__ dcount) {
    const int n = dcount ? min(*dcount, n_rows) : n_rows;
    __shared__ __attribute__((aligned(16))) unsigned short wt[WT_END];
    __shared__ __attribute__((aligned(16))) float bias[3 * NH + N4];
    __shared__ __attribute__((aligned(16))) unsigned short act[INF_WAVES][INF_ROWS][NH + 8];
    __shared__ float yout[INF_WAVES][INF_ROWS][N4];
    // packed [K][N] rows of 8 bf16 (one 16-byte load) -> transposed rows
    for (int c = threadIdx.x; c < PK_END / 8; c += blockDim.x) {
        const uint4 v = reinterpret_cast<const uint4*>(wb_g)[c];
        const int e = 8 * c;
        int base, row, k, n0;
        if (e < PK_W2) {
            base = WT_1, row = WT_K1, k = e / NH, n0 = e % NH;
        } else if (e < PK_W3) {
            base = WT_2, row = WT_NH, k = (e - PK_W2) / NH, n0 = (e - PK_W2) % NH;
        } else if (e < PK_W4) {
            base = WT_3, row = WT_NH, k = (e - PK_W3) / NH, n0 = (e - PK_W3) % NH;
        } else {
            base = WT_4, row = WT_NH, k = (e - PK_W4) / N4, n0 = (e - PK_W4) % N4;
        }
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 8; ++j)
            wt[base + (n0 + j) * row + k] = (unsigned short)(w4[j >> 1] >> (16 * (j & 1)));
    }
    for (int c = threadIdx.x; c < (3 * NH + N4) / 4; c += blockDim.x)
        reinterpret_cast<float4*>(bias)[c] = reinterpret_cast<const float4*>(bias_g)[c];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, q4 = lane >> 4;
    const int ntiles = (n + INF_ROWS - 1) / INF_ROWS;
    for (int tile = blockIdx.x * INF_WAVES + wave; tile < ntiles; tile += gridDim.x * INF_WAVES) {
        const int row0 = tile * INF_ROWS;
        // input tile -> act (bf16, K padded to 32)
        for (int e = lane; e < INF_ROWS * K1; e += 64) {
            const int r = e / K1, k = e % K1;
            const int gr = row0 + r;
            act[wave][r][k] = to_bf16((gr < n && k < NIN) ? X[(size_t)gr * NIN + k] : 0.0f);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // three hidden layers
        for (int layer = 0; layer < 3; ++layer) {
            const int K = layer == 0 ? K1 : NH;
            const int row = layer == 0 ? WT_K1 : WT_NH;
            const unsigned short* W = wt + (layer == 0 ? WT_1 : (layer == 1 ? WT_2 : WT_3));
            f32x4 accs[NH / 16];
            for (int nt = 0; nt < NH / 16; ++nt) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int k0 = 0; k0 < K; k0 += 16) {
                    const bf16x4 a = *reinterpret_cast<const bf16x4*>(&act[wave][r16][k0 + 4 * q4]);
                    const bf16x4 b = *reinterpret_cast<const bf16x4*>(W + (nt * 16 + r16) * row + k0 + 4 * q4);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
                }
                accs[nt] = acc;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            for (int nt = 0; nt < NH / 16; ++nt) {
                for (int t = 0; t < 4; ++t) {
                    const int r = 4 * q4 + t, c = nt * 16 + r16;
                    act[wave][r][c] = to_bf16(silu_fast(accs[nt][t] + bias[layer * NH + c]));
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
        // output layer: N padded to 16
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = 0; k0 < NH; k0 += 16) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(&act[wave][r16][k0 + 4 * q4]);
            const bf16x4 b = *reinterpret_cast<const bf16x4*>(wt + WT_4 + r16 * WT_NH + k0 + 4 * q4);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        for (int t = 0; t < 4; ++t) yout[wave][4 * q4 + t][r16] = acc[t] + bias[3 * NH + r16];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (lane < INF_ROWS && row0 + lane < n) {
            const int gr = row0 + lane;
            V3 y = v3(yout[wave][lane][0], yout[wave][lane][1], yout[wave][lane][2]);
            if (quat) y = rot_mul(quat_rot(quat[pairs[2 * (size_t)gr]]), y);  // a's frame -> world
            const float ny = sqrtf(dot(y, y));
            const float inv = ny > 0.0f ? 1.0f / ny : 0.0f;
            out_dir[3 * (size_t)gr] = y.x * inv;
            out_dir[3 * (size_t)gr + 1] = y.y * inv;
            out_dir[3 * (size_t)gr + 2] = y.z * inv;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
}

}  // namespace sd

using namespace sd;

// FP32 forward pass of n feature rows (one thread per row, parameters in LDS,
// hidden activations in registers): the unit direction u = y / |y|. The
// reference-precision form of k_mlp_infer (BF16 MFMA); the training kernel
// runs the same arithmetic with the backward pass.
constexpr int FW_BLOCK = 128;
__global__ __launch_bounds__(FW_BLOCK) void k_mlp_forward(const float* __restrict__ X, int n,
                                                          const float* __restrict__ w, float* __restrict__ out) {
    __shared__ float sw[NPARAM];
    for (int i = threadIdx.x; i < NPARAM; i += FW_BLOCK) sw[i] = w[i];
    __syncthreads();
    const int r = blockIdx.x * FW_BLOCK + threadIdx.x;
    if (r >= n) return;
    float h0[NH], h1[NH];
    const float* x = X + (size_t)r * NIN;
#pragma unroll
    for (int j = 0; j < NH; ++j) h0[j] = sw[OFF_B1 + j];
    for (int i = 0; i < NIN; ++i) {
        const float xi = x[i];
#pragma unroll
        for (int j = 0; j < NH; ++j) h0[j] = h0[j] + xi * sw[OFF_W1 + i * NH + j];
    }
#pragma unroll
    for (int j = 0; j < NH; ++j) h0[j] = h0[j] * sigm(h0[j]);
#pragma unroll
    for (int j = 0; j < NH; ++j) h1[j] = sw[OFF_B2 + j];
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        const float a = h0[i];
#pragma unroll
        for (int j = 0; j < NH; ++j) h1[j] = h1[j] + a * sw[OFF_W2 + i * NH + j];
    }
#pragma unroll
    for (int j = 0; j < NH; ++j) h1[j] = h1[j] * sigm(h1[j]);
#pragma unroll
    for (int j = 0; j < NH; ++j) h0[j] = sw[OFF_B3 + j];
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        const float a = h1[i];
#pragma unroll
        for (int j = 0; j < NH; ++j) h0[j] = h0[j] + a * sw[OFF_W3 + i * NH + j];
    }
    float y[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) y[k] = sw[OFF_B4 + k];
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        const float a = h0[i] * sigm(h0[i]);
#pragma unroll
        for (int k = 0; k < NOUT; ++k) y[k] = y[k] + a * sw[OFF_W4 + i * NOUT + k];
    }
    const float l = sqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    const float il = l > 0.0f ? 1.0f / l : 0.0f;
    out[3 * (size_t)r] = y[0] * il;
    out[3 * (size_t)r + 1] = y[1] * il;
    out[3 * (size_t)r + 2] = y[2] * il;
}

extern "C" int sd_mlp_param_count(void) { return NPARAM; }
extern "C" int sd_mlp_packed_bytes(void) { return PK_BYTES; }

extern "C" sd_status sd_hull_descriptors(const float* pos4, const int* vstart, const int* nv, const float* verts4,
                                         int n, float* desc7, void* stream) {
    SD_REQUIRE(n >= 0 && (n == 0 || (pos4 && vstart && nv && verts4 && desc7)), "sd_hull_descriptors: bad argument");
    if (n == 0) return SD_OK;
    hipLaunchKernelGGL(k_hull_desc, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, (const float4*)pos4,
                       vstart, nv, (const float4*)verts4, n, desc7);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_mlp_features(const int* pairs, int npairs, const float* pos4, const float* quat4,
                                     const float* desc7, float* X, void* stream) {
    SD_REQUIRE(npairs >= 0 && (npairs == 0 || (pairs && pos4 && quat4 && desc7 && X)), "sd_mlp_features: bad argument");
    if (npairs == 0) return SD_OK;
    hipLaunchKernelGGL(k_mlp_features, dim3(div_up(npairs, 256)), dim3(256), 0, (hipStream_t)stream, pairs, npairs,
                       (const float4*)pos4, (const float4*)quat4, desc7, X, (const int*)nullptr);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_mlp_targets(const int* pairs, int npairs, const float* quat4, const float* normal3, float* T,
                                    void* stream) {
    SD_REQUIRE(npairs >= 0 && (npairs == 0 || (pairs && quat4 && normal3 && T)), "sd_mlp_targets: bad argument");
    if (npairs == 0) return SD_OK;
    hipLaunchKernelGGL(k_mlp_targets, dim3(div_up(npairs, 256)), dim3(256), 0, (hipStream_t)stream, pairs, npairs,
                       (const float4*)quat4, normal3, T);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_mlp_init(float* params, uint64_t seed, void* stream) {
    SD_REQUIRE(params, "sd_mlp_init: NULL");
    hipLaunchKernelGGL(k_mlp_init, dim3(div_up(NPARAM, 256)), dim3(256), 0, (hipStream_t)stream, params, seed);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_mlp_train_scratch_bytes(int batch, size_t* bytes) {
    SD_REQUIRE(bytes && batch > 0, "sd_mlp_train_scratch_bytes: bad argument");
    const size_t nb = (size_t)div_up(batch, TB_S);
    *bytes = nb * NPARAM * sizeof(float) + nb * sizeof(float) + 256;
    return SD_OK;
}

extern "C" sd_status sd_mlp_forward_backward(const float* X, const float* T, const int* idx, int batch,
                                             const float* params, float* grad, float* loss, void* scratch,
                                             size_t scratch_bytes, void* stream) {
    SD_REQUIRE(X && T && idx && params && grad && scratch && batch > 0, "sd_mlp_forward_backward: bad argument");
    size_t need;
    sd_mlp_train_scratch_bytes(batch, &need);
    SD_REQUIRE(scratch_bytes >= need, "sd_mlp_forward_backward: scratch too small");
    const int nb = div_up(batch, TB_S);
    float* gpart = (float*)scratch;
    float* lpart = gpart + (size_t)nb * NPARAM;
    const float inv_batch = 1.0f / (float)batch;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(TB_T), 0, s, X, T, idx, batch, inv_batch, params, gpart, lpart);
    SD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_reduce, dim3(div_up(NPARAM + 1, 256)), dim3(256), 0, s, gpart, lpart, nb, inv_batch,
                       grad, loss);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_adam_step(float* params, const float* grad, float* m, float* v, int n, float lr, float beta1,
                                  float beta2, float eps, int step, void* stream) {
    SD_REQUIRE(params && grad && m && v && n > 0 && step >= 1, "sd_adam_step: bad argument");
    const float c1 = 1.0f - powf(beta1, (float)step);
    const float c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, params, grad, m, v, n, lr,
                       beta1, beta2, eps, c1, c2);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_mlp_pack_bf16(const float* params, void* packed, void* stream) {
    SD_REQUIRE(params && packed, "sd_mlp_pack_bf16: NULL");
    unsigned short* wb = (unsigned short*)packed;
    float* bias = (float*)((char*)packed + PK_END * 2);
    hipLaunchKernelGGL(k_mlp_pack, dim3(div_up(NH * NH, 256)), dim3(256), 0, (hipStream_t)stream, params, wb, bias);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

extern "C" sd_status sd_mlp_infer_bf16(const float* X, int n, const void* packed, const int* pairs,
                                       const float* quat4, float* out_dir3, void* stream) {
    SD_REQUIRE(n >= 0 && (n == 0 || (X && packed && out_dir3)), "sd_mlp_infer_bf16: bad argument");
    SD_REQUIRE((pairs == nullptr) == (quat4 == nullptr), "sd_mlp_infer_bf16: pairs and quat4 go together");
    if (n == 0) return SD_OK;
    const int ntiles = div_up(n, INF_ROWS);
    int blocks = div_up(ntiles, INF_WAVES);
    if (blocks > 2048) blocks = 2048;
    const unsigned short* wb = (const unsigned short*)packed;
    const float* bias = (const float*)((const char*)packed + PK_END * 2);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(INF_WAVES * 64), 0, (hipStream_t)stream, X, n, wb, bias, pairs,
                       (const float4*)quat4, out_dir3, (const int*)nullptr);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

namespace sd {

sd_status launch_hull_descriptors(const float4* pos4, const int* vstart, const int* nv, const float4* verts4, int n,
                                  float* desc7, hipStream_t s) {
    if (n == 0) return SD_OK;
    hipLaunchKernelGGL(k_hull_desc, dim3(div_up(n, 256)), dim3(256), 0, s, pos4, vstart, nv, verts4, n, desc7);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status enqueue_warm_dirs(const int* pairs, int capacity, const int* dcount, const float4* pos4,
                            const float4* quat4, const float* desc7, float* X, const void* packed, float* out_dir3,
                            hipStream_t s) {
    if (capacity == 0) return SD_OK;
    hipLaunchKernelGGL(k_mlp_features, dim3(div_up(capacity, 256)), dim3(256), 0, s, pairs, capacity, pos4, quat4,
                       desc7, X, dcount);
    SD_LAUNCH_CHECK();
    const int blocks = std::min(div_up(div_up(capacity, INF_ROWS), INF_WAVES), 2048);
    const unsigned short* wb = (const unsigned short*)packed;
    const float* bias = (const float*)((const char*)packed + PK_END * 2);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(INF_WAVES * 64), 0, s, (const float*)X, capacity, wb, bias,
                       pairs, quat4, out_dir3, dcount);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

}  // namespace sd

extern "C" sd_status sd_mlp_forward(const float* X, int n, const float* params, float* out_dir3, void* stream) {
    SD_REQUIRE(n >= 0 && (n == 0 || (X && params && out_dir3)), "sd_mlp_forward: bad argument");
    if (n == 0) return SD_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3(div_up(n, FW_BLOCK)), dim3(FW_BLOCK), 0, (hipStream_t)stream, X, n, params,
                       out_dir3);
    SD_LAUNCH_CHECK();
    return SD_OK;
}
