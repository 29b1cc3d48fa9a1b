// Deterministic per-sphere contact gather and semi-implicit integration.
//
// Replaces computeParticleToParticleCUDA / reactCollisionsCUDA
// (src/engine/cuda/collisions/SAP/sap_reaction.cu:20-160) and, for single-sphere
// bodies, advectParticlesCUDA (src/engine/cuda/physics_update.cu:232-310).
// The reference scatters pair forces into both spheres without atomics; here
// each sphere gathers the forces of all its contacts in a fixed traversal order
// (stencil rows (dz, dy) ascending, sorted positions ascending within a row,
// then walls 0..5), so results do not depend on the launch configuration.
//
// Contact law (per contact, FP32, no fused multiply-add so that the pair force
// seen by each side is the exact negation of the other):
//   n      = (x_i - x_j) / |x_i - x_j|,  delta = r_i + r_j - |x_i - x_j| > 0
//   l_i    = r_i - delta/2 (lever arm to the mid-overlap contact point; the same
//            point for both spheres, so pair torques conserve angular momentum)
//   m_eff  = 1 / (1/m_i + 1/m_j),  c_n = 2 zeta sqrt(k_n m_eff),  c_t = 2 zeta sqrt(k_t m_eff)
//   v_c    = (v_i - v_j) - (l_i w_i + l_j w_j) x n,  v_n = v_c . n,  v_t = v_c - v_n n
//   f_n    = k_n delta - c_n v_n
//   xi_t   = (xi_t - (xi_t . n) n) + v_t dt                (tangential spring history)
//   F_t    = -k_t xi_t - c_t v_t, limited to mu |f_n|; on slip xi_t = -(F_t + c_t v_t) / k_t
//   F_i    = f_n n + F_t,  T_i = -l_i (n x F_t)
// Walls are partners with r = 0, 1/m = 0, v = w = 0 and ids -1..-6.
#include "sd_contact_law.hpp"
#include "sd_grid_dev.hpp"
#include <algorithm>

#include "sd_internal.hpp"

namespace sd {

ContactDev to_dev(const sd_contact_params& c) {
    ContactDev d;
    d.kn = c.kn;
    d.kt = c.kt_ratio * c.kn;
    d.c2z = 2.0f * c.damping_ratio;
    d.mu = c.mu;
    d.dt = c.dt;
    d.gx = c.gravity[0];
    d.gy = c.gravity[1];
    d.gz = c.gravity[2];
    d.walls = c.walls;
    for (int a = 0; a < 3; ++a) {
        d.wlo[a] = c.wall_lo[a];
        d.whi[a] = c.wall_hi[a];
    }
    d.maxc = c.max_contacts;
    return d;
}

sd_status check_contact(const sd_contact_params* c) {
    SD_REQUIRE(c != nullptr, "contact params are NULL");
    SD_REQUIRE(c->kn > 0.0f && c->kt_ratio > 0.0f, "contact: kn and kt_ratio must be positive");
    SD_REQUIRE(c->damping_ratio >= 0.0f && c->mu >= 0.0f, "contact: damping_ratio and mu must be >= 0");
    SD_REQUIRE(c->dt > 0.0f, "contact: dt must be positive");
    SD_REQUIRE(c->max_contacts > 0 && c->max_contacts <= 256, "contact: max_contacts must be in [1, 256]");
    return SD_OK;
}

__device__ __forceinline__ void integrate_sphere(const ContactDev& c, float4& p, float4& v, float4& w, float fx,
                                                 float fy, float fz, float tx, float ty, float tz) {
    if (v.w > 0.0f) {
        v.x = v.x + (fx * v.w + c.gx) * c.dt;
        v.y = v.y + (fy * v.w + c.gy) * c.dt;
        v.z = v.z + (fz * v.w + c.gz) * c.dt;
        p.x = p.x + v.x * c.dt;
        p.y = p.y + v.y * c.dt;
        p.z = p.z + v.z * c.dt;
        w.x = w.x + (tx * w.w) * c.dt;
        w.y = w.y + (ty * w.w) * c.dt;
        w.z = w.z + (tz * w.w) * c.dt;
    }
}

__device__ __forceinline__ int old_slot_of(const ContactArgs& a, int s) {
    if (a.old_slot_i32) return a.old_slot_i32[s];
    if (a.old_slot_u32) return (int)a.old_slot_u32[s];
    return (int)(uint32_t)a.old_slot_u64[s];
}

// Looks up the tangential history stored for partner pid in the old list
// (partner ids in a list are distinct, so the entry found does not depend on
// the search order). The search starts at entry `start`, the position the
// contact has in the new list: contacts are found in the same order step after
// step, so a lasting contact is usually at the same position and the first
// probe finds it.
__device__ __forceinline__ void find_history(const ContactArgs& a, int os, int oc, int pid, int start, float& hx,
                                             float& hy, float& hz) {
    hx = 0.0f;
    hy = 0.0f;
    hz = 0.0f;
    int k = start < oc ? start : 0;
    for (int j = 0; j < oc; ++j) {
        const float4 h = a.hist_old[(size_t)k * a.n + os];
        if (__float_as_int(h.w) == pid) {
            hx = h.x;
            hy = h.y;
            hz = h.z;
            break;
        }
        k = k + 1 < oc ? k + 1 : 0;
    }
}

constexpr int XCDS = 8;       // gfx950: 8 accelerator complex dies, one L2 each
constexpr int CAND_UNROLL = 4; // candidate positions loaded per batch (loads in flight)
static_assert(CAND_UNROLL - 1 <= CONTACT_POS_PAD, "sorted positions must be padded for a whole batch");
constexpr int CONTACT_STAGE = 256;  // staged candidate positions per wave (4 KiB of LDS)

// Orders one wave's LDS writes and reads (the lanes of a wave only).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup i is dispatched to XCD i % 8. Remap so that every XCD works on a
// contiguous range of logical blocks: blocks resident on one XCD then touch
// neighbouring cells and share that XCD's L2. Returns -1 for padding blocks.
__device__ __forceinline__ int xcd_logical_block(int nblocks) {
    const int per = (nblocks + XCDS - 1) / XCDS;
    const int lb = (int)(blockIdx.x % XCDS) * per + (int)(blockIdx.x / XCDS);
    return lb < nblocks ? lb : -1;
}

// MODE: CONTACT_FORCES writes force/torque; CONTACT_FUSED_* integrates and writes
// the next state and keys. INDIRECT: vel/omg/ids are in the previous step's order
// and are read through old_slot (positions are always in sorted order).
// CONTACT_FUSED_BATCH: the grid, contact law and cell starts come from the
// sphere's scene; the arithmetic is the same code, so a scene stepped in a
// batch is bit-identical to the same scene stepped alone.
// a wave-uniform struct in global memory, word by word into scalar registers
// (the compiler does not use scalar loads for memory the kernel also writes)
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& src) {
    static_assert(sizeof(T) % 4 == 0, "uniform_copy: whole 32-bit words");
    T dst;
    const int* const w = reinterpret_cast<const int*>(&src);
    int* const d = reinterpret_cast<int*>(&dst);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(w[i]);
    return dst;
}

// STRADDLE (batch mode only): one 64-thread block per listed wave whose slots
// span more than one scene; every lane then takes its own scene's parameters.
// The work of one sphere slot s of a wave (its lanes are consecutive slots
// of one wave, or the lanes of a straddling wave): everything the kernel does
// once the slot is known.
template <int MODE, bool INDIRECT, bool STRADDLE>
__device__ __forceinline__ void contact_wave(const ContactArgs& a, int s) {
    constexpr bool BATCH = MODE == CONTACT_FUSED_BATCH;
    constexpr bool PADDED = MODE != CONTACT_FORCES;  // fused modes read a context's padded buffer
    if (s >= a.n) return;
    GridDev g;
    ContactDev c;
    const int* ranges;  // CONTACT_FORCES: (start, end) per cell; fused modes: cell starts
    uint32_t cbase = 0u;
    if constexpr (STRADDLE) {
        const SceneDev& S = a.scenes[a.scene_of[s]];
        g = S.g;
        c = S.c;
        cbase = S.cbase;
        ranges = a.cell_start + cbase;
    } else if constexpr (BATCH) {
        // one scene in the wave (the waves spanning two are in a.straddle and
        // run by the straddle launch): its parameters are wave-uniform
        const int my_scene = a.scene_of[s];
        const int sid = __builtin_amdgcn_readfirstlane(my_scene);
        if (__ballot(my_scene != sid) != 0u) return;
        const SceneDev S = uniform_copy(a.scenes[sid]);
        g = S.g;
        c = S.c;
        cbase = S.cbase;
        ranges = a.cell_start + cbase;
    } else {
        g = a.g;
        c = a.c;
        ranges = MODE == CONTACT_FORCES ? a.ranges : a.cell_start;
    }
    float4 p = a.pos[s];
    const int os = old_slot_of(a, s);
    const int self = INDIRECT ? os : s;
    float4 v = a.vel[self];
    float4 w = a.omg[self];
    const int myid = a.ids[self];
    int oc = a.hcount_old[os];
    oc = oc < c.maxc ? oc : c.maxc;

    int cx, cy, cz;
    cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f, tx = 0.0f, ty = 0.0f, tz = 0.0f;
    int nc = 0;

    // one sphere-sphere contact with the sphere at sorted position t; the
    // sphere itself and coincident spheres (d2 == 0) are not contacts
    auto visit = [&](int t, const float4* sl) {
        const float4 q = sl[t];  // the staged copy (LDS) of position t
        const float dx = p.x - q.x;
        const float dyy = p.y - q.y;
        const float dzz = p.z - q.z;
        const float d2 = dx * dx + dyy * dyy + dzz * dzz;
        const float rs = p.w + q.w;
        if (!(d2 < rs * rs) || d2 == 0.0f) return;
        const int ot = INDIRECT ? old_slot_of(a, t) : t;
        const float4 vq = a.vel[ot];
        if (v.w + vq.w == 0.0f) return;  // two static spheres
        const float4 wq = a.omg[ot];
        const int pid = a.ids[ot];
        const float d = sqrtf(d2);
        const float inv_d = 1.0f / d;
        const float nx = dx * inv_d, ny = dyy * inv_d, nz = dzz * inv_d;
        float hx, hy, hz;
        find_history(a, os, oc, pid, nc, hx, hy, hz);
        const PairForce f = contact_law(c, nx, ny, nz, rs - d, p.w, q.w, v.w, vq.w, v.x - vq.x, v.y - vq.y,
                                        v.z - vq.z, w.x, w.y, w.z, wq.x, wq.y, wq.z, hx, hy, hz);
        fx = fx + f.fx;
        fy = fy + f.fy;
        fz = fz + f.fz;
        tx = tx + f.tx;
        ty = ty + f.ty;
        tz = tz + f.tz;
        if (nc < c.maxc) a.hist_new[(size_t)nc * a.n + s] = make_float4(hx, hy, hz, __int_as_float(pid));
        ++nc;
    };

    // Distance test of the candidates at sorted positions t0 .. t0+3 (q), then
    // the contacts among them in ascending position order. Only the distance
    // is tested per candidate; a batch with a possible contact in any lane
    // (rare, and always the batch holding the sphere itself) then drops
    // positions past hi and the sphere itself (position s, without a memory
    // access), and visit() drops coincident spheres (d2 == 0).
    auto test_batch = [&](const float4 (&q)[CAND_UNROLL], int t0, int hi, const float4* sl) {
        bool h[CAND_UNROLL];
        uint64_t any = 0u;  // lanes with a possible contact (scalar masks)
#pragma unroll
        for (int k = 0; k < CAND_UNROLL; ++k) {
            const float dx = p.x - q[k].x;
            const float dyy = p.y - q[k].y;
            const float dzz = p.z - q[k].z;
            const float d2 = dx * dx + dyy * dyy + dzz * dzz;
            const float rs = p.w + q[k].w;
            h[k] = d2 < rs * rs;
            any |= __builtin_amdgcn_ballot_w64(h[k]);
        }
        if (any) {
            uint32_t hit = 0u;
#pragma unroll
            for (int k = 0; k < CAND_UNROLL; ++k) hit |= h[k] ? (1u << k) : 0u;
            while (hit) {
                const int k = __builtin_ctz(hit);
                hit &= hit - 1u;
                if (t0 + k < hi && t0 + k != s) visit(t0 + k, sl);
            }
        }
    };

    // Candidate staging. The lanes of a wave are consecutive sorted spheres, so
    // for one stencil row offset (dz, dy) their windows ascend with the lane
    // (rows and cells are in sorted order). The lanes in the first active
    // lane's grid row (segment A) and the others (segment B, normally the next
    // row) each have a compact union of windows. The two unions, each followed
    // by the 3 positions a batch may read past a window, form one staged
    // sequence: sorted position t of a lane is entry t + off of it. The wave
    // copies the sequence into its LDS region with coalesced loads (in chunks
    // of CONTACT_STAGE entries; one chunk unless the wave spans several sparse
    // rows) and every lane tests its candidates in the chunk from LDS.
    extern __shared__ float4 cstage[];
    float4* const stg = cstage + (threadIdx.x >> 6) * CONTACT_STAGE;
    const uint64_t amask = __ballot(1);  // active lanes: a prefix of the wave
    const int arank = __popcll(amask & lanemask_lt());
    const int nact = __popcll(amask);
    const uint32_t rowid = cbase + (uint32_t)((cz * g.ny + cy) * g.nx);
    const bool segB = rowid != (uint32_t)__builtin_amdgcn_readfirstlane((int)rowid);
    const uint64_t bmask = __ballot(segB);
    constexpr int PAD = CAND_UNROLL - 1;        // a batch reads up to 3 positions past its window
    constexpr int CHUNK = CONTACT_STAGE - PAD;  // sequence entries tested per chunk
    const char* const pos_bytes = reinterpret_cast<const char*>(a.pos);

    // window of stencil row r (rows (dz, dy) in ascending order); the next
    // row's window is loaded while a row is copied and tested
    auto window = [&](int r, int& lo, int& hi) {
        const int z = cz + r / 3 - 1, y = cy + r % 3 - 1;
        lo = 0;
        hi = 0;
        if (z >= 0 && z < g.nz && y >= 0 && y < g.ny) {
            if constexpr (MODE == CONTACT_FORCES) row_range(g, ranges, cx, y, z, lo, hi);
            else row_window(g, ranges, cx, y, z, lo, hi);
        }
    };
    int next_lo, next_hi;
    window(0, next_lo, next_hi);
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        {
            const int lo = next_lo, hi = next_hi;
            if (r + 1 < 9) window(r + 1, next_lo, next_hi);
            const uint64_t ne = __ballot(hi > lo);
            if (ne == 0u) continue;
            const uint64_t ma = ne & ~bmask, mb = ne & bmask;
            int la = 0, na = 0, lb = 0, nb = 0;
            if (ma) {
                la = __builtin_amdgcn_readlane(lo, __builtin_ctzll(ma));
                na = __builtin_amdgcn_readlane(hi, 63 - __builtin_clzll(ma)) - la;
            }
            if (mb) {
                lb = __builtin_amdgcn_readlane(lo, __builtin_ctzll(mb));
                nb = __builtin_amdgcn_readlane(hi, 63 - __builtin_clzll(mb)) - lb;
            }
            const int offb = na + PAD;                 // sequence entry of position lb
            const int seq = offb + nb + PAD;           // sequence length
            const int off = segB ? offb - lb : -la;    // sequence entry of sorted position t: t + off
            const int lb_b = lb - offb;                // source of entry e >= offb: lb_b + e
            for (int e0 = 0; e0 < seq - PAD; e0 += CHUNK) {
                const int m = min(CONTACT_STAGE, seq - e0);
                for (int i = arank; i < m; i += 2 * nact) {
                    const int e1 = e0 + i, e2 = e1 + nact;
                    // sources reach at most n + 2 (3 padding entries past a
                    // window that ends at n): inside a context's padded buffer,
                    // clamped for a caller's unpadded one
                    const int s1 = PADDED ? (e1 < offb ? la : lb_b) + e1 : min((e1 < offb ? la : lb_b) + e1, a.n - 1);
                    const int s2 = PADDED ? (e2 < offb ? la : lb_b) + e2 : min((e2 < offb ? la : lb_b) + e2, a.n - 1);
                    const float4 v1 = *reinterpret_cast<const float4*>(pos_bytes + (size_t)((uint32_t)s1 * 16u));
                    if (i + nact < m) {
                        const float4 v2 = *reinterpret_cast<const float4*>(pos_bytes + (size_t)((uint32_t)s2 * 16u));
                        stg[i + nact] = v2;
                    }
                    stg[i] = v1;
                }
                wave_lds_sync();
                // this lane's positions in the chunk; a batch may read 3 entries
                // past them (still inside the chunk's copy)
                const int tlo = max(lo, e0 - off), thi = min(hi, e0 + CHUNK - off);
                const float4* const sl = stg + (off - e0);
                for (int t0 = tlo; t0 < thi; t0 += CAND_UNROLL) {
                    float4 q[CAND_UNROLL];
#pragma unroll
                    for (int k = 0; k < CAND_UNROLL; ++k) q[k] = sl[t0 + k];
                    test_batch(q, t0, thi, sl);
                }
                wave_lds_sync();  // every lane's reads precede the next copy
            }
        }
    }

    // walls: the touched ones first, so that the contact-law set-up of a wall
    // contact runs only where some lane of the wave touches a wall
    uint32_t touch = 0u;
    if (c.walls && v.w > 0.0f) {
#pragma unroll
        for (int wl = 0; wl < 6; ++wl) {
            const int ax = wl >> 1;
            const bool upper = wl & 1;
            const float pc = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
            const float dist = upper ? c.whi[ax] - pc : pc - c.wlo[ax];
            touch |= dist < p.w ? (1u << wl) : 0u;
        }
    }
    if (touch) {
#pragma unroll
        for (int wl = 0; wl < 6; ++wl) {
            if (!((touch >> wl) & 1u)) continue;
            const int ax = wl >> 1;
            const bool upper = wl & 1;
            const float pc = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
            const float dist = upper ? c.whi[ax] - pc : pc - c.wlo[ax];
            const float sgn = upper ? -1.0f : 1.0f;
            const float nx = ax == 0 ? sgn : 0.0f;
            const float ny = ax == 1 ? sgn : 0.0f;
            const float nz = ax == 2 ? sgn : 0.0f;
            const int pid = -(wl + 1);
            float hx, hy, hz;
            find_history(a, os, oc, pid, nc, hx, hy, hz);
            const PairForce f = contact_law(c, nx, ny, nz, p.w - dist, p.w, 0.0f, v.w, 0.0f, v.x, v.y, v.z, w.x,
                                            w.y, w.z, 0.0f, 0.0f, 0.0f, hx, hy, hz);
            fx = fx + f.fx;
            fy = fy + f.fy;
            fz = fz + f.fz;
            tx = tx + f.tx;
            ty = ty + f.ty;
            tz = tz + f.tz;
            if (nc < c.maxc) a.hist_new[(size_t)nc * a.n + s] = make_float4(hx, hy, hz, __int_as_float(pid));
            ++nc;
        }
    }

    a.hcount_new[s] = nc < c.maxc ? nc : c.maxc;
    if (nc > c.maxc) atomicMax(a.overflow, nc);

    if (MODE == CONTACT_FORCES) {
        a.force[s] = make_float4(fx, fy, fz, 0.0f);
        a.torque[s] = make_float4(tx, ty, tz, 0.0f);
    } else {
        integrate_sphere(c, p, v, w, fx, fy, fz, tx, ty, tz);
        a.pos_o[s] = p;
        a.vel_o[s] = v;
        a.omg_o[s] = w;
        a.ids_o[s] = myid;
        if (MODE == CONTACT_FUSED_U32) {
            a.keys_o32[s] = cell_id(g, p.x, p.y, p.z);
            a.vals_o32[s] = (uint32_t)s;
        } else {
            const uint32_t cell_new = cbase + cell_id(g, p.x, p.y, p.z);
            const uint64_t key_new = ((uint64_t)cell_new << 32) | (uint64_t)(uint32_t)s;
            a.keys_o64[s] = key_new;
            // a sphere that left the cell it was sorted in: listed for the next
            // step's re-sort (one atomic per wave)
            const uint32_t cell_old = cbase + (uint32_t)((cz * g.ny + cy) * g.nx + cx);
            const bool moved = a.mcount != nullptr && cell_new != cell_old;
            const uint64_t m = __ballot(moved);
            if (moved) {
                const int leader = __ffsll((long long)m) - 1;
                int first = 0;
                if (lane_id() == leader) first = atomicAdd(a.mcount, __popcll(m));
                first = __shfl(first, leader);
                const int i = first + __popcll(m & lanemask_lt());
                if (i < a.mcap) {
                    a.mnew[i] = key_new;
                    a.mold[i] = ((uint64_t)cell_old << 32) | (uint64_t)(uint32_t)s;
                }
            }
        }
    }
}


template <int MODE, bool INDIRECT, bool STRADDLE = false>
__global__ __launch_bounds__(256) void k_contacts(ContactArgs a) {
    static_assert(!STRADDLE || MODE == CONTACT_FUSED_BATCH, "k_contacts: straddling waves exist in batch mode only");
    if (a.step_ctl) {
        if (a.step_ctl[0]) return;  // the step was given up
        if (!STRADDLE && blockIdx.x == 0 && threadIdx.x == 0) a.step_ctl[1] += 1;
    }
    int s;
    if constexpr (STRADDLE) {
        s = a.straddle[blockIdx.x] * WAVE + (int)threadIdx.x;
    } else {
        const int lb = xcd_logical_block((a.n + (int)blockDim.x - 1) / (int)blockDim.x);
        if (lb < 0) return;
        s = lb * (int)blockDim.x + (int)threadIdx.x;
    }
    contact_wave<MODE, INDIRECT, STRADDLE>(a, s);
}

constexpr int PERSIST_CLAIM = 8;    // tiles a wave claims at a time
constexpr int PERSIST_STRIDE = 64;  // ints between the XCDs' counters

// Persistent form of the fused step kernel: a grid of resident blocks whose
// waves claim runs of PERSIST_CLAIM 64-slot tiles from a counter per XCD (tiles of XCD x are the
// x-th eighth of the slots, so an XCD's L2 still holds neighbouring cells)
// until its range is exhausted; the results are those of k_contacts (every
// tile computes the same slots the same way). counters[0..7] start at 0.
template <int MODE, bool INDIRECT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_contacts_persistent(
    ContactArgs a, int* __restrict__ counters) {
    if (a.step_ctl) {
        if (a.step_ctl[0]) return;  // the step was given up
        if (blockIdx.x == 0 && threadIdx.x == 0) a.step_ctl[1] += 1;
    }
    const int ntiles = (a.n + WAVE - 1) / WAVE;
    const int per = (ntiles + XCDS - 1) / XCDS;
    const int x = (int)(blockIdx.x % XCDS);
    const int t0 = x * per, t1 = min(ntiles, t0 + per);
    int* const ctr = counters + PERSIST_STRIDE * x;  // one counter per XCD, each on its own 256-byte line
    for (;;) {
        int t = 0;
        if (lane_id() == 0) t = atomicAdd(ctr, PERSIST_CLAIM);
        t = t0 + __shfl(t, 0);
        if (t >= t1) break;
        const int te = min(t + PERSIST_CLAIM, t1);
        for (; t < te; ++t) contact_wave<MODE, INDIRECT, false>(a, t * WAVE + lane_id());
    }
}

__global__ void k_integrate(float4* __restrict__ pos, float4* __restrict__ vel, float4* __restrict__ omg,
                            const float4* __restrict__ f, const float4* __restrict__ t, int n, ContactDev c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i], v = vel[i], w = omg[i];
    const float4 F = f[i], T = t[i];
    integrate_sphere(c, p, v, w, F.x, F.y, F.z, T.x, T.y, T.z);
    pos[i] = p;
    vel[i] = v;
    omg[i] = w;
}

sd_status launch_contacts(int mode, const ContactArgs& a, hipStream_t s, int block) {
    if (a.n <= 0) return SD_OK;
    if (block <= 0) block = 128;
    SD_REQUIRE(block % WAVE == 0 && block <= 256, "contact kernel: block must be a multiple of 64, <= 256");
    const int nblocks = div_up(a.n, block);
    const dim3 grid(div_up(nblocks, XCDS) * XCDS);  // padded for the XCD remap
    const size_t lds = (size_t)(block / WAVE) * CONTACT_STAGE * sizeof(float4);
    if (mode == CONTACT_FORCES)
        hipLaunchKernelGGL((k_contacts<CONTACT_FORCES, false>), grid, dim3(block), lds, s, a);
    else if (mode == CONTACT_FUSED_U32)
        hipLaunchKernelGGL((k_contacts<CONTACT_FUSED_U32, true>), grid, dim3(block), lds, s, a);
    else if (mode == CONTACT_FUSED_U64)
        hipLaunchKernelGGL((k_contacts<CONTACT_FUSED_U64, true>), grid, dim3(block), lds, s, a);
    else
        hipLaunchKernelGGL((k_contacts<CONTACT_FUSED_BATCH, true>), grid, dim3(block), lds, s, a);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_contacts_persistent(int mode, const ContactArgs& a, hipStream_t s, int block, int* counters) {
    if (a.n <= 0) return SD_OK;
    if (block <= 0) block = 128;
    SD_REQUIRE(block % WAVE == 0 && block <= 256, "contact kernel: block must be a multiple of 64, <= 256");
    SD_REQUIRE(mode == CONTACT_FUSED_U64 || mode == CONTACT_FUSED_BATCH,
               "contact kernel: the persistent form is for packed 64-bit keys");
    const size_t lds = (size_t)(block / WAVE) * CONTACT_STAGE * sizeof(float4);
    int dev = 0, cus = 0, per_cu = 0;
    SD_HIP(hipGetDevice(&dev));
    SD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (mode == CONTACT_FUSED_U64)
        SD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_contacts_persistent<CONTACT_FUSED_U64, true>), block, lds));
    else
        SD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_contacts_persistent<CONTACT_FUSED_BATCH, true>), block, lds));
    // every resident block, at most one per tile, a multiple of the XCD count
    const int tiles = div_up(a.n, WAVE);
    int nb = std::max(1, per_cu) * std::max(1, cus);
    nb = std::min(nb, div_up(tiles, block / WAVE));
    nb = std::max(XCDS, nb / XCDS * XCDS);
    sd_status st = launch_fill_u32((uint32_t*)counters, 0u, (size_t)PERSIST_STRIDE * XCDS, s);
    if (st != SD_OK) return st;
    if (mode == CONTACT_FUSED_U64)
        hipLaunchKernelGGL((k_contacts_persistent<CONTACT_FUSED_U64, true>), dim3(nb), dim3(block), lds, s, a,
                           counters);
    else
        hipLaunchKernelGGL((k_contacts_persistent<CONTACT_FUSED_BATCH, true>), dim3(nb), dim3(block), lds, s, a,
                           counters);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_contacts_straddle(const ContactArgs& a, hipStream_t s) {
    if (a.n <= 0 || a.n_straddle <= 0) return SD_OK;
    SD_REQUIRE(a.straddle && a.scenes && a.scene_of, "contact kernel: straddle launch without its lists");
    const size_t lds = (size_t)CONTACT_STAGE * sizeof(float4);
    hipLaunchKernelGGL((k_contacts<CONTACT_FUSED_BATCH, true, true>), dim3(a.n_straddle), dim3(WAVE), lds, s, a);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

sd_status launch_integrate(float4* pos, float4* vel, float4* omg, const float4* f, const float4* t, int n,
                           const ContactDev& c, hipStream_t s) {
    if (n <= 0) return SD_OK;
    hipLaunchKernelGGL(k_integrate, dim3(div_up(n, 256)), dim3(256), 0, s, pos, vel, omg, f, t, n, c);
    SD_LAUNCH_CHECK();
    return SD_OK;
}

}  // namespace sd

using namespace sd;

extern "C" sd_status sd_contact_forces(const float* pos4, const float* vel4, const float* omg4, const int* ids, int n,
                                       const sd_grid* grid, const int* ranges, const sd_contact_params* cp,
                                       const int* old_slot, const int* hist_count_old, const float* hist4_old,
                                       int* hist_count_new, float* hist4_new, float* force4, float* torque4,
                                       int* overflow, void* stream) {
    sd_status st = check_grid(grid);
    if (st != SD_OK) return st;
    st = check_contact(cp);
    if (st != SD_OK) return st;
    SD_REQUIRE(n >= 0, "sd_contact_forces: n < 0");
    if (n == 0) return SD_OK;
    SD_REQUIRE(pos4 && vel4 && omg4 && ids && ranges && old_slot && hist_count_old && hist4_old && hist_count_new &&
                   hist4_new && force4 && torque4 && overflow,
               "sd_contact_forces: null pointer");
    ContactArgs a = {};
    a.pos = (const float4*)pos4;
    a.vel = (const float4*)vel4;
    a.omg = (const float4*)omg4;
    a.ids = ids;
    a.n = n;
    a.g = to_dev(*grid);
    a.ranges = ranges;
    a.c = to_dev(*cp);
    a.old_slot_i32 = old_slot;
    a.hcount_old = hist_count_old;
    a.hist_old = (const float4*)hist4_old;
    a.hcount_new = hist_count_new;
    a.hist_new = (float4*)hist4_new;
    a.overflow = overflow;
    a.force = (float4*)force4;
    a.torque = (float4*)torque4;
    return launch_contacts(CONTACT_FORCES, a, (hipStream_t)stream, 128);
}

extern "C" sd_status sd_integrate_spheres(float* pos4, float* vel4, float* omg4, const float* force4,
                                          const float* torque4, int n, const sd_contact_params* cp, void* stream) {
    sd_status st = check_contact(cp);
    if (st != SD_OK) return st;
    SD_REQUIRE(n >= 0, "sd_integrate_spheres: n < 0");
    if (n == 0) return SD_OK;
    SD_REQUIRE(pos4 && vel4 && omg4 && force4 && torque4, "sd_integrate_spheres: null pointer");
    return launch_integrate((float4*)pos4, (float4*)vel4, (float4*)omg4, (const float4*)force4,
                            (const float4*)torque4, n, to_dev(*cp), (hipStream_t)stream);
}
