// Sphere-scene context: owns the device state and runs whole timesteps.
//
// Replaces DEMEngine::update and cleanBuffers (src/engine/cuda/simulation_kernel.cu:30-153)
// together with EntityHandler::allocateCUDA / transferDataToDevice
// (src/engine/entities/entity_handler.cpp:196-472) for sphere scenes.
//
// One step, all on the context stream, no host synchronisation:
//   sort(keys)            keys were written by the previous step (or upload):
//                         the radix sort, or with u64 keys the re-sort from
//                         the previous order (which also permutes positions)
//   cell starts           first sorted position of every cell
//   permute positions     sorted positions gathered from the current state
//   fused contact kernel  reads sorted positions, the current state (through the
//                         sort permutation) and the old history; writes the
//                         integrated state into the other state buffer in the
//                         new order, the new history and the next step's keys
// Buffer roles repeat with period two, so two steps are captured once into a
// hipGraph per starting parity and replayed.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "sd_internal.hpp"

namespace sd {

constexpr int DIAG_BLOCK = 256;
constexpr int DIAG_VALUES = 9;  // 8 diagnostics + contact count

__global__ __launch_bounds__(DIAG_BLOCK) void k_diag_partial(const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel,
                                                             const float4* __restrict__ omg,
                                                             const int* __restrict__ hcount, int n, float gx, float gy,
                                                             float gz, double* __restrict__ partial) {
    __shared__ double red[DIAG_VALUES][DIAG_BLOCK];
    double acc[DIAG_VALUES] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * DIAG_BLOCK + threadIdx.x; i < n; i += gridDim.x * DIAG_BLOCK) {
        const float4 p = pos[i], v = vel[i], w = omg[i];
        acc[8] += (double)hcount[i];
        if (v.w <= 0.0f) continue;  // static spheres carry no momentum
        const double m = 1.0 / (double)v.w;
        const double I = w.w > 0.0f ? 1.0 / (double)w.w : 0.0;
        const double px = m * v.x, py = m * v.y, pz = m * v.z;
        acc[0] += px;
        acc[1] += py;
        acc[2] += pz;
        acc[3] += (double)p.y * pz - (double)p.z * py + I * w.x;
        acc[4] += (double)p.z * px - (double)p.x * pz + I * w.y;
        acc[5] += (double)p.x * py - (double)p.y * px + I * w.z;
        acc[6] += 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) +
                  0.5 * I * ((double)w.x * w.x + (double)w.y * w.y + (double)w.z * w.z);
        acc[7] += -m * ((double)gx * p.x + (double)gy * p.y + (double)gz * p.z);
    }
    for (int k = 0; k < DIAG_VALUES; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = DIAG_BLOCK / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < DIAG_VALUES; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < DIAG_VALUES) partial[(size_t)blockIdx.x * DIAG_VALUES + threadIdx.x] = red[threadIdx.x][0];
}

struct GraphKey {
    int keys, hist, state;
    bool full;  // two steps with the full radix sort (dense runs)
    bool operator<(const GraphKey& o) const {
        if (keys != o.keys) return keys < o.keys;
        if (hist != o.hist) return hist < o.hist;
        if (state != o.state) return state < o.state;
        return full < o.full;
    }
};

// Full-sort steps run between two reads of the changed-cell count while a
// scene is dense (more changed cells per step than the re-sort takes).
constexpr int DENSE_CHUNK = 16;

}  // namespace sd

using namespace sd;

struct sd_ctx {
    sd_sphere_params p{};
    hipStream_t stream = nullptr;
    int n = 0;
    int n_cells = 0;
    int key_bits = 0;
    int items = 0;
    int contact_block = 128;
    // state in the current order (the keys refer to it), double-buffered:
    // [cur_state] is read, [1 - cur_state] receives the integrated state;
    // posB holds the positions gathered into sorted order
    float4* pos[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    float4* omg[2] = {nullptr, nullptr};
    int* ids[2] = {nullptr, nullptr};
    int cur_state = 0;
    float4* posB = nullptr;
    uint32_t* k32[2] = {nullptr, nullptr};
    uint32_t* v32[2] = {nullptr, nullptr};
    uint64_t* k64[2] = {nullptr, nullptr};
    int cur_keys = 0;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int* cell_start = nullptr;  // n_cells + 1 entries
    int* hcount[2] = {nullptr, nullptr};
    float4* hist[2] = {nullptr, nullptr};
    int cur_hist = 0;
    int* overflow = nullptr;
    // spheres that changed cell in the last step (u64 keys): count (set above
    // RESORT_CAP to force the full sort), new and old keys, and the same sorted
    int* mcount = nullptr;
    uint64_t* mnew = nullptr;
    uint64_t* mold = nullptr;
    uint64_t* msn = nullptr;
    uint64_t* mso = nullptr;
    // [0] != 0: a re-sort step was given up; [1] steps run since the last reset
    int* step_ctl = nullptr;
    bool need_full = true;  // the next step's keys are not from a step of this order
    bool dense = false;     // the last full-sort step had more changed cells than the re-sort takes
    double* diag = nullptr;
    int diag_blocks = 0;
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> graph_objs;
    long long steps_done = 0;
    // batch of independent scenes (sd_batch_upload): scene table and the scene
    // of every slot; scenes occupy consecutive slot and cell-id ranges
    bool batch = false;
    int nscenes = 0;
    std::vector<SceneDev> scenes_h;
    SceneDev* scenes_d = nullptr;
    int* scene_of = nullptr;
    // waves (64 consecutive slots from a multiple of 64) spanning two or more
    // scenes, run by a launch of their own
    int* straddle = nullptr;
    int n_straddle = 0;
    // the straddling waves' launch runs on a side stream, forked from and
    // joined back into the context's stream around the main contact launch
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join = nullptr;
    // rigid-body scenes stepped with the batch (sd_batch_attach_hulls), on a
    // stream of their own forked from and joined to the step
    sd_hull_ctx* hulls = nullptr;
    long long hull_gen = 0;  // hull_generation(hulls) when the step graphs were captured
    // persistent form of the fused step kernel (sd_ctx_set_persistent)
    bool persistent = false;
    int* pcounters = nullptr;
    hipStream_t hside = nullptr;
    hipEvent_t hev_fork = nullptr, hev_join = nullptr;
};

namespace {

void drop_graphs(sd_ctx* c) {
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto g : c->graph_objs) (void)hipGraphDestroy(g);
    c->graphs.clear();
    c->graph_objs.clear();
}

void release(sd_ctx* c) {
    drop_graphs(c);
    void* bufs[] = {c->pos[0], c->pos[1], c->vel[0], c->vel[1], c->omg[0], c->omg[1], c->ids[0], c->ids[1],
                    c->posB, c->k32[0], c->k32[1], c->v32[0], c->v32[1], c->k64[0], c->k64[1], c->sort_tmp,
                    c->cell_start, c->hcount[0], c->hcount[1], c->hist[0], c->hist[1], c->overflow, c->diag,
                    c->scenes_d, c->scene_of, c->straddle, c->mcount, c->mnew, c->mold, c->msn, c->mso,
                    c->step_ctl};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int k = 0; k < 2; ++k) {
        c->pos[k] = c->vel[k] = c->omg[k] = nullptr;
        c->ids[k] = nullptr;
    }
    c->posB = nullptr;
    c->k32[0] = c->k32[1] = c->v32[0] = c->v32[1] = nullptr;
    c->k64[0] = c->k64[1] = nullptr;
    c->sort_tmp = nullptr;
    c->cell_start = nullptr;
    c->hcount[0] = c->hcount[1] = nullptr;
    c->hist[0] = c->hist[1] = nullptr;
    c->overflow = nullptr;
    c->mcount = nullptr;
    c->mnew = c->mold = c->msn = c->mso = nullptr;
    c->step_ctl = nullptr;
    c->diag = nullptr;
    c->scenes_d = nullptr;
    c->scene_of = nullptr;
    c->straddle = nullptr;
    c->n_straddle = 0;
    c->scenes_h.clear();
    c->nscenes = 0;
    c->n = 0;
}

template <typename T>
sd_status dalloc(T** p, size_t count) {
    SD_HIP(hipMalloc((void**)p, (count ? count : 1) * sizeof(T)));
    return SD_OK;
}

// Enqueues one step on the context stream. If ev is given, ev[0..4] are
// recorded before the sort and after the sort, cell-start, permute and contact stages.
// full (u64 keys): the radix sort instead of the re-sort from the previous
// order. A re-sort step that finds more than RESORT_CAP changed cells is given
// up on the device (step_ctl[0]), together with every later step of the same
// launch sequence; the host then redoes it as a full step (run_steps).
sd_status enqueue_step(sd_ctx* c, hipEvent_t* ev = nullptr, bool full = false) {
    hipStream_t s = c->stream;
#define MARK(i) \
    if (ev) SD_HIP(hipEventRecord(ev[i], s))
    const int n = c->n;
    const int ho = c->cur_hist, hn = 1 - c->cur_hist;
    const int so = c->cur_state, sn = 1 - c->cur_state;
    ContactArgs a = {};
    a.pos = c->posB;
    a.vel = c->vel[so];
    a.omg = c->omg[so];
    a.ids = c->ids[so];
    a.n = n;
    a.g = to_dev(c->p.grid);
    a.cell_start = c->cell_start;
    a.c = to_dev(c->p.contact);
    a.hcount_old = c->hcount[ho];
    a.hist_old = c->hist[ho];
    a.hcount_new = c->hcount[hn];
    a.hist_new = c->hist[hn];
    a.overflow = c->overflow;
    a.pos_o = c->pos[sn];
    a.vel_o = c->vel[sn];
    a.omg_o = c->omg[sn];
    a.ids_o = c->ids[sn];
    sd_status st;
    MARK(0);
    if (c->hulls) {  // the rigid-body scenes' step runs beside the sphere step
        SD_HIP(hipEventRecord(c->hev_fork, s));
        SD_HIP(hipStreamWaitEvent(c->hside, c->hev_fork, 0));
        st = hull_enqueue_step(c->hulls, c->hside);
        if (st != SD_OK) return st;
        SD_HIP(hipEventRecord(c->hev_join, c->hside));
    }
    if (c->p.key64) {
        // sorted keys: the radix sort (full), or the re-sort from the previous
        // order and the spheres that changed cell, with the positions permuted
        // in the same pass. Either leaves them in the other key buffer.
        uint64_t* K = c->k64[c->cur_keys];
        uint64_t* Ka = c->k64[1 - c->cur_keys];
        // changed-cell counters alternate with the history buffers: this step
        // reads [ho] (written by the previous step) and appends to [hn]
        int* mc_read = c->mcount + ho;
        int* mc_write = c->mcount + hn;
        if (full) {
            st = radix_sort_keys_u64_gated(&K, &Ka, n, 32, 32 + c->key_bits, c->sort_tmp, c->sort_tmp_bytes, s,
                                           c->items, nullptr, 0);
        } else {
            // also updates the cell starts and sets mc_write = 0
            st = launch_resort_u64(mc_read, c->mnew, c->mold, c->msn, c->mso, K, n, c->cell_start, c->n_cells,
                                   c->pos[so], c->posB, Ka, mc_write, c->step_ctl, s);
            uint64_t* t = K;
            K = Ka;
            Ka = t;
        }
        if (st != SD_OK) return st;
        MARK(1);
        // after the full sort: cell starts, the permutation, mc_write = 0
        if (full) st = launch_step_starts_u64(K, n, c->n_cells, c->cell_start, c->step_ctl, true, c->pos[so],
                                              c->posB, mc_write, s);
        if (st != SD_OK) return st;
        MARK(2);
        MARK(3);
        a.old_slot_u64 = K;
        a.keys_o64 = Ka;
        a.mcount = mc_write;
        a.mnew = c->mnew;
        a.mold = c->mold;
        a.mcap = RESORT_CAP;
        a.step_ctl = c->step_ctl;
        a.scenes = c->scenes_d;
        a.scene_of = c->scene_of;
        a.straddle = c->straddle;
        a.n_straddle = c->n_straddle;
        // waves spanning two scenes: their own launch, beside the main one
        // (disjoint spheres; shared counters are atomics)
        const bool fork = c->batch && c->n_straddle > 0;
        if (fork) {
            SD_HIP(hipEventRecord(c->ev_fork, s));
            SD_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
            st = launch_contacts_straddle(a, c->side);
            if (st != SD_OK) return st;
            SD_HIP(hipEventRecord(c->ev_join, c->side));
        }
        if (c->persistent)
            st = launch_contacts_persistent(c->batch ? CONTACT_FUSED_BATCH : CONTACT_FUSED_U64, a, s,
                                            c->contact_block, c->pcounters);
        else
            st = launch_contacts(c->batch ? CONTACT_FUSED_BATCH : CONTACT_FUSED_U64, a, s, c->contact_block);
        if (st != SD_OK) return st;
        if (fork) SD_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
        MARK(4);
        c->cur_keys = (Ka == c->k64[0]) ? 0 : 1;
    } else {
        uint32_t* K = c->k32[c->cur_keys];
        uint32_t* V = c->v32[c->cur_keys];
        uint32_t* Ka = c->k32[1 - c->cur_keys];
        uint32_t* Va = c->v32[1 - c->cur_keys];
        st = radix_sort_pairs_u32(&K, &V, &Ka, &Va, n, 0, c->key_bits, c->sort_tmp, c->sort_tmp_bytes, s, c->items);
        if (st != SD_OK) return st;
        MARK(1);
        st = launch_cell_starts_u32(K, n, c->n_cells, c->cell_start, s);
        if (st != SD_OK) return st;
        MARK(2);
        st = launch_permute_pos(V, nullptr, n, c->pos[so], c->posB, s);
        if (st != SD_OK) return st;
        MARK(3);
        a.old_slot_u32 = V;
        a.keys_o32 = Ka;
        a.vals_o32 = Va;
        st = launch_contacts(CONTACT_FUSED_U32, a, s, c->contact_block);
        if (st != SD_OK) return st;
        MARK(4);
        c->cur_keys = (Ka == c->k32[0]) ? 0 : 1;
    }
#undef MARK
    if (c->hulls) SD_HIP(hipStreamWaitEvent(s, c->hev_join, 0));
    c->cur_hist = hn;
    c->cur_state = sn;
    return SD_OK;
}

sd_status run_graph_pair(sd_ctx* c, bool full = false) {
    const GraphKey key{c->cur_keys, c->cur_hist, c->cur_state, full};
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        const int k0 = c->cur_keys, h0 = c->cur_hist, s0 = c->cur_state;
        SD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        sd_status st = enqueue_step(c, nullptr, full);
        if (st == SD_OK) st = enqueue_step(c, nullptr, full);
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
        c->graph_objs.push_back(g);
        it = c->graphs.emplace(key, ge).first;
        // capture advanced the host-side roles; rewind them for the replay
        c->cur_keys = k0;
        c->cur_hist = h0;
        c->cur_state = s0;
    }
    SD_HIP(hipGraphLaunch(it->second, c->stream));
    // two steps advance the roles by an even amount: keys and history return to
    // their starting buffers (period two)
    return SD_OK;
}

// Runs nsteps steps. u64 keys: re-sort steps (graph-replayed in pairs when
// use_graph), and a full step first whenever the keys are not from a step of
// this order, or a re-sort step was given up. The device counts the steps that
// ran, which sets the buffer roles; one synchronisation per launch sequence.
sd_status run_steps(sd_ctx* c, int nsteps) {
    sd_status st;
    if (c->hulls && hull_generation(c->hulls) != c->hull_gen) {
        drop_graphs(c);  // captured with the attached bodies' previous step
        c->hull_gen = hull_generation(c->hulls);
    }
    if (!c->p.key64) {
        int left = nsteps;
        if (c->p.use_graph) {
            while (left >= 2) {
                st = run_graph_pair(c);
                if (st != SD_OK) return st;
                left -= 2;
            }
        }
        for (; left > 0; --left) {
            st = enqueue_step(c);
            if (st != SD_OK) return st;
        }
        return SD_OK;
    }
    int left = nsteps;
    while (left > 0) {
        st = launch_fill_u32((uint32_t*)c->step_ctl, 0u, 2, c->stream);
        if (st != SD_OK) return st;
        const int h0 = c->cur_hist, s0 = c->cur_state;
        const bool full = c->need_full || c->dense;
        if (full) {
            // one full-sort step after a given-up step or an upload; while the
            // scene is dense, DENSE_CHUNK of them (graph pairs) before the
            // changed-cell count is read again
            int l = c->dense ? std::min(left, DENSE_CHUNK) : 1;
            if (c->p.use_graph) {
                while (l >= 2) {
                    st = run_graph_pair(c, true);
                    if (st != SD_OK) return st;
                    l -= 2;
                }
            }
            for (; l > 0; --l) {
                st = enqueue_step(c, nullptr, true);
                if (st != SD_OK) return st;
            }
            c->need_full = false;
        } else {
            int l = left;
            if (c->p.use_graph) {
                while (l >= 2) {
                    st = run_graph_pair(c);
                    if (st != SD_OK) return st;
                    l -= 2;
                }
            }
            for (; l > 0; --l) {
                st = enqueue_step(c);
                if (st != SD_OK) return st;
            }
        }
        int ctl[2] = {0, 0};
        SD_HIP(hipMemcpyAsync(ctl, c->step_ctl, sizeof ctl, hipMemcpyDeviceToHost, c->stream));
        SD_HIP(hipStreamSynchronize(c->stream));
        const int ran = ctl[1];
        if (ran < 0 || ran > left || (ran == 0 && ctl[0] == 0))
            return set_error(SD_ERR_STATE, "step: %d of %d steps ran", ran, left);
        c->cur_hist = h0 ^ (ran & 1);  // the roles of the steps that ran
        c->cur_state = s0 ^ (ran & 1);
        left -= ran;
        if (ctl[0]) c->need_full = true;  // redo the given-up step with the full sort
        if (full) {  // still dense: the last full-sort step's changed cells exceed the re-sort's cap
            int m = 0;
            SD_HIP(hipMemcpyAsync(&m, c->mcount + c->cur_hist, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            SD_HIP(hipStreamSynchronize(c->stream));
            c->dense = m > RESORT_CAP;
        }
    }
    return SD_OK;
}

}  // namespace

extern "C" sd_status sd_ctx_create(const sd_sphere_params* p, sd_ctx** out) {
    SD_REQUIRE(p && out, "sd_ctx_create: NULL argument");
    *out = nullptr;
    sd_status st = check_grid(&p->grid);
    if (st != SD_OK) return st;
    st = check_contact(&p->contact);
    if (st != SD_OK) return st;
    SD_REQUIRE(p->tile_items == 0 || p->tile_items == 4 || p->tile_items == 8 || p->tile_items == 16,
               "sd_ctx_create: tile_items must be 0, 4, 8 or 16");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(SD_ERR_NO_DEVICE, "sd_ctx_create: no HIP device");
    sd_ctx* c = new sd_ctx();
    c->p = *p;
    c->n_cells = p->grid.n[0] * p->grid.n[1] * p->grid.n[2];
    c->key_bits = grid_bits(p->grid);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_error(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    *out = c;
    return SD_OK;
}

extern "C" sd_status sd_ctx_set_contact(sd_ctx* c, const sd_contact_params* cp) {
    SD_REQUIRE(c && cp, "sd_ctx_set_contact: NULL argument");
    if (c->batch) return set_error(SD_ERR_STATE, "sd_ctx_set_contact: a batch context keeps its scenes' laws");
    sd_status st = check_contact(cp);
    if (st != SD_OK) return st;
    SD_REQUIRE(cp->max_contacts == c->p.contact.max_contacts,
               "sd_ctx_set_contact: max_contacts cannot change (%d != %d)", cp->max_contacts,
               c->p.contact.max_contacts);
    if (c->stream && hipStreamSynchronize(c->stream) != hipSuccess)
        return set_error(SD_ERR_HIP, "sd_ctx_set_contact: stream synchronisation failed");
    c->p.contact = *cp;
    drop_graphs(c);  // captured with the old law
    return SD_OK;
}

extern "C" sd_status sd_ctx_set_persistent(sd_ctx* c, int on) {
    SD_REQUIRE(c, "sd_ctx_set_persistent: NULL context");
    SD_REQUIRE(c->p.key64, "sd_ctx_set_persistent: the persistent step kernel needs packed 64-bit keys");
    if (on && !c->pcounters) {
        sd_status st = dalloc(&c->pcounters, 64 * 8);  // one 256-byte line per XCD
        if (st != SD_OK) return st;
    }
    (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);  // captured with the other kernel form
    c->persistent = on != 0;
    return SD_OK;
}

extern "C" sd_status sd_batch_attach_hulls(sd_ctx* c, sd_hull_ctx* h) {
    SD_REQUIRE(c, "sd_batch_attach_hulls: NULL context");
    SD_REQUIRE(c->batch, "sd_batch_attach_hulls: not a batch context");
    if (h) {
        if (c->n == 0 || hull_body_count(h) == 0)
            return set_error(SD_ERR_STATE, "sd_batch_attach_hulls: upload the scenes and the bodies first");
        SD_REQUIRE(hull_dt(h) == c->p.contact.dt, "sd_batch_attach_hulls: dt differs (%g vs %g)", hull_dt(h),
                   c->p.contact.dt);
        if (!c->hside) {
            SD_HIP(hipStreamCreateWithFlags(&c->hside, hipStreamNonBlocking));
            SD_HIP(hipEventCreateWithFlags(&c->hev_fork, hipEventDisableTiming));
            SD_HIP(hipEventCreateWithFlags(&c->hev_join, hipEventDisableTiming));
        }
    }
    (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);  // the step graphs are captured again with (or without) the bodies
    c->hulls = h;
    if (h) c->hull_gen = hull_generation(h);
    return SD_OK;
}

extern "C" sd_status sd_ctx_destroy(sd_ctx* c) {
    if (!c) return SD_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->hside) (void)hipStreamSynchronize(c->hside);
    if (c->hev_fork) (void)hipEventDestroy(c->hev_fork);
    if (c->hev_join) (void)hipEventDestroy(c->hev_join);
    if (c->hside) (void)hipStreamDestroy(c->hside);
    if (c->pcounters) (void)hipFree(c->pcounters);
    release(c);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SD_OK;
}

extern "C" sd_status sd_ctx_stream(sd_ctx* c, void** stream) {
    SD_REQUIRE(c && stream, "sd_ctx_stream: NULL");
    *stream = (void*)c->stream;
    return SD_OK;
}

namespace {

sd_status check_spheres(const char* who, int n, const float* x, const float* y, const float* z, const float* radius,
                        const float* mass, float cell) {
    float rmax = 0.0f;
    for (int i = 0; i < n; ++i) {
        if (!(radius[i] > 0.0f) || !(mass[i] >= 0.0f) || !isfinite(x[i]) || !isfinite(y[i]) || !isfinite(z[i]))
            return set_error(SD_ERR_ARG, "%s: sphere %d has a bad radius, mass or position", who, i);
        rmax = radius[i] > rmax ? radius[i] : rmax;
    }
    if (2.0f * rmax > cell)
        return set_error(SD_ERR_ARG, "%s: largest diameter %g exceeds the cell size %g", who, 2.0f * rmax, cell);
    return SD_OK;
}

// The next step's keys do not come from a step of this order (upload,
// restore): the next sort is the full one.
sd_status force_full_sort(sd_ctx* c) {
    c->need_full = true;
    c->dense = false;
    sd_status st = launch_fill_u32((uint32_t*)c->step_ctl, 0u, 2, c->stream);
    if (st != SD_OK) return st;
    return launch_fill_u32((uint32_t*)(c->mcount + c->cur_hist), 0x7fffffffu, 1, c->stream);
}

// Allocates the device state for n spheres and uploads it (slot = particle id);
// keys are computed by the caller.
sd_status upload_state(sd_ctx* c, int n, const float* x, const float* y, const float* z, const float* vx,
                       const float* vy, const float* vz, const float* radius, const float* mass) {
    c->n = n;
    const int maxc = c->p.contact.max_contacts;
    sd_status st;
#define ALLOC(ptr, count)                 \
    st = dalloc(&(ptr), (size_t)(count)); \
    if (st != SD_OK) {                    \
        release(c);                       \
        return st;                        \
    }
    for (int k = 0; k < 2; ++k) {
        ALLOC(c->pos[k], n);
        ALLOC(c->vel[k], n);
        ALLOC(c->omg[k], n);
        ALLOC(c->ids[k], n);
    }
    ALLOC(c->posB, (size_t)n + CONTACT_POS_PAD);
    if (c->p.key64) {
        ALLOC(c->k64[0], n);
        ALLOC(c->k64[1], n);
    } else {
        ALLOC(c->k32[0], n);
        ALLOC(c->k32[1], n);
        ALLOC(c->v32[0], n);
        ALLOC(c->v32[1], n);
    }
    c->sort_tmp_bytes = radix_temp_bytes(n);
    ALLOC(*(char**)&c->sort_tmp, c->sort_tmp_bytes);
    ALLOC(c->cell_start, (size_t)c->n_cells + 1);
    ALLOC(c->hcount[0], n);
    ALLOC(c->hcount[1], n);
    ALLOC(c->hist[0], (size_t)maxc * n);
    ALLOC(c->hist[1], (size_t)maxc * n);
    ALLOC(c->overflow, 1);
    ALLOC(c->mcount, 2);
    ALLOC(c->mnew, RESORT_CAP);
    ALLOC(c->mold, RESORT_CAP);
    ALLOC(c->msn, RESORT_CAP);
    ALLOC(c->mso, RESORT_CAP);
    ALLOC(c->step_ctl, 2);
    c->diag_blocks = div_up(n, DIAG_BLOCK) < 1024 ? div_up(n, DIAG_BLOCK) : 1024;
    ALLOC(c->diag, (size_t)c->diag_blocks * DIAG_VALUES);
#undef ALLOC
    c->items = c->p.tile_items ? c->p.tile_items : radix_default_items(n);

    std::vector<float4> pos(n), vel(n), omg(n);
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) {
        const float inv_m = mass[i] > 0.0f ? 1.0f / mass[i] : 0.0f;
        const float I = 0.4f * mass[i] * radius[i] * radius[i];
        const float inv_I = mass[i] > 0.0f ? 1.0f / I : 0.0f;
        pos[i] = make_float4(x[i], y[i], z[i], radius[i]);
        vel[i] = make_float4(mass[i] > 0.0f ? vx[i] : 0.0f, mass[i] > 0.0f ? vy[i] : 0.0f,
                             mass[i] > 0.0f ? vz[i] : 0.0f, inv_m);
        omg[i] = make_float4(0.0f, 0.0f, 0.0f, inv_I);
        ids[i] = i;
    }
    SD_HIP(hipMemcpyAsync(c->pos[0], pos.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    SD_HIP(hipMemcpyAsync(c->vel[0], vel.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    SD_HIP(hipMemcpyAsync(c->omg[0], omg.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    SD_HIP(hipMemcpyAsync(c->ids[0], ids.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    st = launch_fill_u32((uint32_t*)c->hcount[0], 0u, n, c->stream);
    if (st == SD_OK) st = launch_fill_u32((uint32_t*)c->hcount[1], 0u, n, c->stream);
    if (st == SD_OK) st = launch_fill_u32((uint32_t*)(c->posB + n), 0u, 4 * (size_t)CONTACT_POS_PAD, c->stream);
    if (st == SD_OK) st = launch_fill_u32((uint32_t*)c->overflow, 0u, 1, c->stream);
    if (st != SD_OK) return st;
    c->cur_keys = 0;
    c->cur_hist = 0;
    c->cur_state = 0;
    c->steps_done = 0;
    return force_full_sort(c);  // (reads cur_hist)
}

}  // namespace

extern "C" sd_status sd_ctx_upload_spheres(sd_ctx* c, int n, const float* x, const float* y, const float* z,
                                           const float* vx, const float* vy, const float* vz, const float* radius,
                                           const float* mass) {
    SD_REQUIRE(c, "sd_ctx_upload_spheres: NULL context");
    SD_REQUIRE(!c->batch || c->n == 0, "sd_ctx_upload_spheres: this is a batch context; use sd_batch_upload");
    SD_REQUIRE(n > 0, "sd_ctx_upload_spheres: n must be positive");
    SD_REQUIRE(x && y && z && vx && vy && vz && radius && mass, "sd_ctx_upload_spheres: NULL array");
    sd_status st = check_spheres("sd_ctx_upload_spheres", n, x, y, z, radius, mass, c->p.grid.cell);
    if (st != SD_OK) return st;
    (void)hipStreamSynchronize(c->stream);
    release(c);
    st = upload_state(c, n, x, y, z, vx, vy, vz, radius, mass);
    if (st != SD_OK) {
        release(c);
        return st;
    }
    const GridDev g = to_dev(c->p.grid);
    st = c->p.key64 ? launch_hash_keys_u64(c->pos[0], n, g, c->k64[0], c->stream)
                    : launch_hash_keys_u32(c->pos[0], n, g, c->k32[0], c->v32[0], c->stream);
    if (st != SD_OK) return st;
    SD_HIP(hipStreamSynchronize(c->stream));
    return SD_OK;
}

extern "C" sd_status sd_batch_create(const sd_batch_params* p, sd_ctx** out) {
    SD_REQUIRE(p && out, "sd_batch_create: NULL argument");
    *out = nullptr;
    SD_REQUIRE(p->tile_items == 0 || p->tile_items == 4 || p->tile_items == 8 || p->tile_items == 16,
               "sd_batch_create: tile_items must be 0, 4, 8 or 16");
    SD_REQUIRE(p->contact_block == 0 || p->contact_block == 64 || p->contact_block == 128 ||
                   p->contact_block == 256,
               "sd_batch_create: contact_block must be 0, 64, 128 or 256");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(SD_ERR_NO_DEVICE, "sd_batch_create: no HIP device");
    sd_ctx* c = new sd_ctx();
    c->p.key64 = 1;
    c->p.use_graph = p->use_graph;
    c->p.tile_items = p->tile_items;
    c->contact_block = p->contact_block ? p->contact_block : 128;
    c->batch = true;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_error(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    *out = c;
    return SD_OK;
}

extern "C" sd_status sd_batch_upload(sd_ctx* c, int nscenes, const sd_scene_desc* scenes, const float* x,
                                     const float* y, const float* z, const float* vx, const float* vy,
                                     const float* vz, const float* radius, const float* mass) {
    SD_REQUIRE(c && scenes, "sd_batch_upload: NULL argument");
    SD_REQUIRE(c->batch, "sd_batch_upload: not a batch context (create it with sd_batch_create)");
    SD_REQUIRE(nscenes >= 1, "sd_batch_upload: nscenes must be >= 1");
    SD_REQUIRE(x && y && z && vx && vy && vz && radius && mass, "sd_batch_upload: NULL array");
    std::vector<SceneDev> table(nscenes);
    long long n = 0;
    uint64_t cells = 0;
    for (int k = 0; k < nscenes; ++k) {
        const sd_scene_desc& d = scenes[k];
        SD_REQUIRE(d.n >= 1, "sd_batch_upload: scene %d has no spheres", k);
        sd_status st = check_grid(&d.grid);
        if (st == SD_OK) st = check_contact(&d.contact);
        if (st != SD_OK) return set_error(st, "sd_batch_upload: scene %d: %s", k, sd_last_error());
        const sd_contact_params& c0 = scenes[0].contact;
        SD_REQUIRE(d.contact.dt == c0.dt && d.contact.gravity[0] == c0.gravity[0] &&
                       d.contact.gravity[1] == c0.gravity[1] && d.contact.gravity[2] == c0.gravity[2] &&
                       d.contact.max_contacts == c0.max_contacts,
                   "sd_batch_upload: scene %d: dt, gravity and max_contacts must be the same in every scene", k);
        st = check_spheres("sd_batch_upload", d.n, x + n, y + n, z + n, radius + n, mass + n, d.grid.cell);
        if (st != SD_OK) return set_error(st, "sd_batch_upload: scene %d: %s", k, sd_last_error());
        table[k].g = to_dev(d.grid);
        table[k].c = to_dev(d.contact);
        table[k].cbase = (uint32_t)cells;
        table[k].first = (int)n;
        table[k].n = d.n;
        table[k].pad = 0;
        cells += (uint64_t)d.grid.n[0] * (uint64_t)d.grid.n[1] * (uint64_t)d.grid.n[2];
        n += d.n;
        SD_REQUIRE(n < (1ll << 31) && cells < (1ull << 31), "sd_batch_upload: more than 2^31 spheres or cells");
    }
    (void)hipStreamSynchronize(c->stream);
    release(c);
    c->p.grid = scenes[0].grid;
    c->p.contact = scenes[0].contact;
    c->n_cells = (int)cells;
    int bits = 1;
    while (bits < 32 && (1ull << bits) < cells) ++bits;
    c->key_bits = bits;
    sd_status st = upload_state(c, (int)n, x, y, z, vx, vy, vz, radius, mass);
    if (st == SD_OK) st = dalloc(&c->scenes_d, (size_t)nscenes);
    if (st == SD_OK) st = dalloc(&c->scene_of, (size_t)n);
    if (st != SD_OK) {
        release(c);
        return st;
    }
    c->nscenes = nscenes;
    c->scenes_h = table;
    std::vector<int> sof((size_t)n);
    for (int k = 0; k < nscenes; ++k)
        for (int i = 0; i < table[k].n; ++i) sof[(size_t)table[k].first + i] = k;
    // scenes are consecutive, so a wave spans two scenes exactly when its first
    // and last slot differ in scene
    std::vector<int> straddle;
    for (long long w = 0; w * WAVE < n; ++w) {
        const long long last = std::min(w * WAVE + WAVE - 1, n - 1);
        if (sof[(size_t)(w * WAVE)] != sof[(size_t)last]) straddle.push_back((int)w);
    }
    if (!straddle.empty()) {
        if (!c->side) {
            SD_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
            SD_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
            SD_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        }
        st = dalloc(&c->straddle, straddle.size());
        if (st != SD_OK) {
            release(c);
            return st;
        }
        SD_HIP(hipMemcpyAsync(c->straddle, straddle.data(), straddle.size() * sizeof(int), hipMemcpyHostToDevice,
                              c->stream));
    }
    c->n_straddle = (int)straddle.size();
    SD_HIP(hipMemcpyAsync(c->scenes_d, table.data(), nscenes * sizeof(SceneDev), hipMemcpyHostToDevice, c->stream));
    SD_HIP(hipMemcpyAsync(c->scene_of, sof.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    st = launch_hash_keys_batch(c->pos[0], (int)n, c->scenes_d, c->scene_of, c->k64[0], c->stream);
    if (st != SD_OK) return st;
    SD_HIP(hipStreamSynchronize(c->stream));
    return SD_OK;
}

extern "C" sd_status sd_ctx_step(sd_ctx* c, int nsteps) {
    SD_REQUIRE(c, "sd_ctx_step: NULL context");
    SD_REQUIRE(nsteps >= 0, "sd_ctx_step: nsteps < 0");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_step: no spheres uploaded");
    sd_status st = launch_fill_u32((uint32_t*)c->overflow, 0u, 1, c->stream);
    if (st != SD_OK) return st;
    st = run_steps(c, nsteps);
    if (st != SD_OK) return st;
    int ovf = 0;
    SD_HIP(hipMemcpyAsync(&ovf, c->overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    c->steps_done += nsteps;
    if (c->hulls) {
        hull_count_steps(c->hulls, nsteps);
        st = hull_check_errors(c->hulls);
        if (st != SD_OK) return st;
    }
    if (ovf > 0)
        return set_error(SD_ERR_CAPACITY, "sd_ctx_step: a sphere had %d contacts, capacity is %d", ovf,
                         c->p.contact.max_contacts);
    return SD_OK;
}

extern "C" sd_status sd_ctx_profile(sd_ctx* c, int nsteps, float* stage_ms, int nstages) {
    SD_REQUIRE(c && stage_ms && nstages >= 1 && nsteps >= 1, "sd_ctx_profile: bad argument");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_profile: no spheres uploaded");
    const int ns = nstages < 4 ? nstages : 4;
    for (int k = 0; k < nstages; ++k) stage_ms[k] = 0.0f;
    hipEvent_t ev[5];
    for (int k = 0; k < 5; ++k) SD_HIP(hipEventCreate(&ev[k]));
    sd_status st = SD_OK;
    for (int i = 0; i < nsteps && st == SD_OK; ++i) {
        bool full = c->need_full;
        if (c->p.key64 && !full) {  // full sort when the re-sort would give the step up
            int m = 0;
            if (hipMemcpyAsync(&m, c->mcount + c->cur_hist, sizeof(int), hipMemcpyDeviceToHost, c->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                st = set_error(SD_ERR_HIP, "sd_ctx_profile: counter read failed");
                break;
            }
            full = m > RESORT_CAP;
        }
        c->need_full = false;
        st = enqueue_step(c, ev, full);
        if (st != SD_OK) break;
        hipError_t e = hipEventSynchronize(ev[4]);
        if (e != hipSuccess) {
            st = hip_error(e, "hipEventSynchronize", __FILE__, __LINE__);
            break;
        }
        for (int k = 0; k < ns; ++k) {
            float ms = 0.0f;
            e = hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            if (e != hipSuccess) {
                st = hip_error(e, "hipEventElapsedTime", __FILE__, __LINE__);
                break;
            }
            stage_ms[k] += ms;
        }
    }
    for (int k = 0; k < 5; ++k) (void)hipEventDestroy(ev[k]);
    if (st == SD_OK) c->steps_done += nsteps;
    if (st == SD_OK && c->hulls) {
        hull_count_steps(c->hulls, nsteps);
        st = hull_check_errors(c->hulls);
    }
    return st;
}

// ------------------------------------------------------------------ autotuner
//
// Sweeps the launch shapes of the step (radix-sort items per thread x contact
// kernel block size), timing trial_steps eager steps of each with HIP events on
// the context stream, and keeps the fastest. Every output of the step is
// independent of these shapes (per-sphere gathers, stable sort), so the choice
// changes time only; the state is snapshotted on the device before the trials
// and restored after them, so the simulation does not advance.

namespace {

struct Snapshot {
    char* buf = nullptr;
    size_t off[8] = {};
    size_t sz[8] = {};
    int roles[3] = {0, 0, 0};
};

// Role buffers of the current state: pos, vel, omg, ids, keys (k64 or k32), v32, hcount, hist.
void role_buffers(sd_ctx* c, void* b[8], size_t sz[8]) {
    const size_t n = (size_t)c->n;
    const int cs = c->cur_state, ck = c->cur_keys, ch = c->cur_hist;
    b[0] = c->pos[cs];
    sz[0] = n * sizeof(float4);
    b[1] = c->vel[cs];
    sz[1] = n * sizeof(float4);
    b[2] = c->omg[cs];
    sz[2] = n * sizeof(float4);
    b[3] = c->ids[cs];
    sz[3] = n * sizeof(int);
    b[4] = c->p.key64 ? (void*)c->k64[ck] : (void*)c->k32[ck];
    sz[4] = n * (c->p.key64 ? sizeof(uint64_t) : sizeof(uint32_t));
    b[5] = c->p.key64 ? nullptr : (void*)c->v32[ck];
    sz[5] = c->p.key64 ? 0 : n * sizeof(uint32_t);
    b[6] = c->hcount[ch];
    sz[6] = n * sizeof(int);
    b[7] = c->hist[ch];
    sz[7] = (size_t)c->p.contact.max_contacts * n * sizeof(float4);
}

sd_status snapshot_take(sd_ctx* c, Snapshot& sn) {
    void* b[8];
    size_t total = 0;
    role_buffers(c, b, sn.sz);
    for (int k = 0; k < 8; ++k) {
        sn.off[k] = total;
        total += (sn.sz[k] + 255) & ~(size_t)255;
    }
    SD_HIP(hipMalloc((void**)&sn.buf, total));
    for (int k = 0; k < 8; ++k)
        if (sn.sz[k]) SD_HIP(hipMemcpyAsync(sn.buf + sn.off[k], b[k], sn.sz[k], hipMemcpyDeviceToDevice, c->stream));
    sn.roles[0] = c->cur_keys;
    sn.roles[1] = c->cur_hist;
    sn.roles[2] = c->cur_state;
    return SD_OK;
}

sd_status snapshot_restore(sd_ctx* c, const Snapshot& sn) {
    c->cur_keys = sn.roles[0];
    c->cur_hist = sn.roles[1];
    c->cur_state = sn.roles[2];
    void* b[8];
    size_t sz[8];
    role_buffers(c, b, sz);
    for (int k = 0; k < 8; ++k)
        if (sz[k]) SD_HIP(hipMemcpyAsync(b[k], sn.buf + sn.off[k], sz[k], hipMemcpyDeviceToDevice, c->stream));
    sd_status st = force_full_sort(c);
    if (st != SD_OK) return st;
    return launch_fill_u32((uint32_t*)c->overflow, 0u, 1, c->stream);
}

}  // namespace

extern "C" sd_status sd_ctx_autotune(sd_ctx* c, int trial_steps, int* tile_items, int* contact_block,
                                     float* best_ms) {
    SD_REQUIRE(c && trial_steps >= 1, "sd_ctx_autotune: bad argument");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_autotune: no spheres uploaded");
    const int items_c[3] = {4, 8, 16};
    const int block_c[3] = {64, 128, 256};
    Snapshot sn;
    sd_status st = snapshot_take(c, sn);
    if (st != SD_OK) {
        if (sn.buf) (void)hipFree(sn.buf);
        return st;
    }
    const long long steps0 = c->steps_done;
    // trials step the spheres only: attached bodies are not advanced (nor restored)
    sd_hull_ctx* const hulls = c->hulls;
    c->hulls = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 3.4e38f;
    int bi = c->items, bb = c->contact_block;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = SD_ERR_HIP;
    // each trial runs the steps as a run does (run_steps: graph pairs, a
    // given-up re-sort step redone with the full sort, dense chunks of
    // full-sort pairs) from the snapshot, after a warm-up of three steps that
    // makes the full sort the restore asks for and captures the trial's graphs
    for (int i = 0; i < 3 && st == SD_OK; ++i) {
        for (int j = 0; j < 3 && st == SD_OK; ++j) {
            c->items = items_c[i];
            c->contact_block = block_c[j];
            drop_graphs(c);  // captured with other launch shapes
            st = snapshot_restore(c, sn);
            if (st == SD_OK) st = run_steps(c, 3);
            if (st != SD_OK) break;
            if (hipEventRecord(e0, c->stream) != hipSuccess) {
                st = set_error(SD_ERR_HIP, "sd_ctx_autotune: event record failed");
                break;
            }
            st = run_steps(c, trial_steps);
            if (st != SD_OK) break;
            float ms = 0.0f;
            if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                st = set_error(SD_ERR_HIP, "sd_ctx_autotune: event timing failed");
                break;
            }
            if (ms < best) {
                best = ms;
                bi = items_c[i];
                bb = block_c[j];
            }
        }
    }
    c->items = bi;
    c->contact_block = bb;
    sd_status st2 = snapshot_restore(c, sn);
    if (hipStreamSynchronize(c->stream) != hipSuccess && st2 == SD_OK)
        st2 = set_error(SD_ERR_HIP, "sd_ctx_autotune: stream synchronisation failed");
    (void)hipFree(sn.buf);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    drop_graphs(c);  // captured with other launch shapes
    c->steps_done = steps0;
    c->hulls = hulls;
    if (st != SD_OK) return st;
    if (st2 != SD_OK) return st2;
    if (tile_items) *tile_items = bi;
    if (contact_block) *contact_block = bb;
    if (best_ms) *best_ms = best / (float)trial_steps;
    return SD_OK;
}

extern "C" sd_status sd_ctx_download_spheres(sd_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                                             float* wx, float* wy, float* wz) {
    SD_REQUIRE(c, "sd_ctx_download_spheres: NULL context");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_download_spheres: no spheres uploaded");
    const int n = c->n;
    std::vector<float4> pos(n), vel(n), omg(n);
    std::vector<int> ids(n);
    const int cs = c->cur_state;
    SD_HIP(hipMemcpyAsync(pos.data(), c->pos[cs], n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipMemcpyAsync(vel.data(), c->vel[cs], n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipMemcpyAsync(omg.data(), c->omg[cs], n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipMemcpyAsync(ids.data(), c->ids[cs], n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < n; ++s) {
        const int i = ids[s];
        if (i < 0 || i >= n) return set_error(SD_ERR_STATE, "sd_ctx_download_spheres: corrupt id %d", i);
        if (x) x[i] = pos[s].x;
        if (y) y[i] = pos[s].y;
        if (z) z[i] = pos[s].z;
        if (vx) vx[i] = vel[s].x;
        if (vy) vy[i] = vel[s].y;
        if (vz) vz[i] = vel[s].z;
        if (wx) wx[i] = omg[s].x;
        if (wy) wy[i] = omg[s].y;
        if (wz) wz[i] = omg[s].z;
    }
    return SD_OK;
}

static sd_status diag_values(sd_ctx* c, double out[DIAG_VALUES]) {
    const ContactDev cd = to_dev(c->p.contact);
    const int cs = c->cur_state;
    hipLaunchKernelGGL(k_diag_partial, dim3(c->diag_blocks), dim3(DIAG_BLOCK), 0, c->stream, c->pos[cs], c->vel[cs],
                       c->omg[cs], c->hcount[c->cur_hist], c->n, cd.gx, cd.gy, cd.gz, c->diag);
    SD_LAUNCH_CHECK();
    std::vector<double> part((size_t)c->diag_blocks * DIAG_VALUES);
    SD_HIP(hipMemcpyAsync(part.data(), c->diag, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < DIAG_VALUES; ++k) out[k] = 0.0;
    for (int b = 0; b < c->diag_blocks; ++b)
        for (int k = 0; k < DIAG_VALUES; ++k) out[k] += part[(size_t)b * DIAG_VALUES + k];
    return SD_OK;
}

extern "C" sd_status sd_ctx_diagnostics(sd_ctx* c, double out[8]) {
    SD_REQUIRE(c && out, "sd_ctx_diagnostics: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_diagnostics: no spheres uploaded");
    double v[DIAG_VALUES];
    sd_status st = diag_values(c, v);
    if (st != SD_OK) return st;
    for (int k = 0; k < 8; ++k) out[k] = v[k];
    return SD_OK;
}

extern "C" sd_status sd_ctx_contact_count(sd_ctx* c, long long* count) {
    SD_REQUIRE(c && count, "sd_ctx_contact_count: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_contact_count: no spheres uploaded");
    double v[DIAG_VALUES];
    sd_status st = diag_values(c, v);
    if (st != SD_OK) return st;
    *count = (long long)llround(v[8]);
    return SD_OK;
}

extern "C" sd_status sd_ctx_cell_changes(sd_ctx* c, int* count, int* full_next) {
    SD_REQUIRE(c && count && full_next, "sd_ctx_cell_changes: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_cell_changes: no spheres uploaded");
    *count = 0;
    *full_next = c->p.key64 ? 0 : 1;
    if (!c->p.key64) return SD_OK;
    int m = 0;  // written by the last step, read by the next one
    SD_HIP(hipMemcpyAsync(&m, c->mcount + c->cur_hist, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SD_HIP(hipStreamSynchronize(c->stream));
    *full_next = m > RESORT_CAP ? 1 : 0;
    *count = m == 0x7fffffff ? 0 : m;
    return SD_OK;
}

// ------------------------------------------------------------------ checkpoint

namespace {
struct StateHeader {
    uint32_t magic;
    uint32_t version;
    int32_t n;
    int32_t maxc;
    int32_t key64;
    int32_t n_cells;
    int64_t steps_done;
    int32_t nscenes;  // 0 for a single-scene context
    int32_t reserved;
    uint64_t layout_hash;  // scene table (batch) or grid (single scene)
};
constexpr uint32_t STATE_MAGIC = 0x53445354u;  // "SDST"
constexpr uint32_t STATE_VERSION = 2u;

// FNV-1a 64 over the bytes of the scene table of a batch context (grids,
// contact laws, cell-id bases, slot ranges: 4-byte fields, no padding) or of
// the grid of a single-scene context, so that a state saved from another
// batch with the same totals is rejected.
uint64_t layout_hash(const sd_ctx* c) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](const void* p, size_t sz) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < sz; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
    };
    if (c->nscenes > 0) {
        mix(c->scenes_h.data(), c->scenes_h.size() * sizeof(SceneDev));
    } else {
        const GridDev g = to_dev(c->p.grid);
        mix(&g, sizeof g);
    }
    return h;
}

size_t state_bytes(const sd_ctx* c) {
    const size_t n = (size_t)c->n;
    return sizeof(StateHeader) + n * (3 * sizeof(float4) + sizeof(int)) +
           n * (c->p.key64 ? sizeof(uint64_t) : 2 * sizeof(uint32_t)) + n * sizeof(int) +
           (size_t)c->p.contact.max_contacts * n * sizeof(float4);
}
}  // namespace

// ------------------------------------------------------------------ runtime add / remove

namespace {

// Host copy of the current state rows in slot order; hist holds entry k of
// slot i at k * n + i, as on the device.
struct HostRows {
    int n = 0;
    std::vector<float4> pos, vel, omg, hist;
    std::vector<int> ids, hcount;
};

sd_status read_rows(sd_ctx* c, HostRows& r) {
    const size_t n = (size_t)c->n, maxc = (size_t)c->p.contact.max_contacts;
    r.n = c->n;
    r.pos.resize(n);
    r.vel.resize(n);
    r.omg.resize(n);
    r.ids.resize(n);
    r.hcount.resize(n);
    r.hist.resize(maxc * n);
    const int cs = c->cur_state, ch = c->cur_hist;
    hipStream_t s = c->stream;
    SD_HIP(hipMemcpyAsync(r.pos.data(), c->pos[cs], n * sizeof(float4), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(r.vel.data(), c->vel[cs], n * sizeof(float4), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(r.omg.data(), c->omg[cs], n * sizeof(float4), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(r.ids.data(), c->ids[cs], n * sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(r.hcount.data(), c->hcount[ch], n * sizeof(int), hipMemcpyDeviceToHost, s));
    SD_HIP(hipMemcpyAsync(r.hist.data(), c->hist[ch], maxc * n * sizeof(float4), hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    return SD_OK;
}

// Reallocates the context for r.n spheres and uploads the rows; keys are
// hashed from the positions (index = slot) and the next sort is the full one,
// as after an upload. The step count is kept.
sd_status write_rows(sd_ctx* c, const HostRows& r) {
    const int n = r.n;
    const size_t maxc = (size_t)c->p.contact.max_contacts;
    const long long steps = c->steps_done;
    std::vector<float> x(n), y(n), z(n), rad(n), one(n, 1.0f), zero(n, 0.0f);
    for (int i = 0; i < n; ++i) {
        x[i] = r.pos[i].x;
        y[i] = r.pos[i].y;
        z[i] = r.pos[i].z;
        rad[i] = r.pos[i].w;
    }
    release(c);
    sd_status st = upload_state(c, n, x.data(), y.data(), z.data(), zero.data(), zero.data(), zero.data(),
                                rad.data(), one.data());
    if (st != SD_OK) {
        release(c);
        return st;
    }
    hipStream_t s = c->stream;
    SD_HIP(hipMemcpyAsync(c->pos[0], r.pos.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->vel[0], r.vel.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->omg[0], r.omg.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->ids[0], r.ids.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->hcount[0], r.hcount.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    SD_HIP(hipMemcpyAsync(c->hist[0], r.hist.data(), maxc * n * sizeof(float4), hipMemcpyHostToDevice, s));
    const GridDev g = to_dev(c->p.grid);
    st = c->p.key64 ? launch_hash_keys_u64(c->pos[0], n, g, c->k64[0], s)
                    : launch_hash_keys_u32(c->pos[0], n, g, c->k32[0], c->v32[0], s);
    if (st != SD_OK) return st;
    SD_HIP(hipStreamSynchronize(s));
    c->steps_done = steps;
    return SD_OK;
}

}  // namespace

extern "C" sd_status sd_ctx_remove_spheres(sd_ctx* c, const int* ids, int count) {
    SD_REQUIRE(c, "sd_ctx_remove_spheres: NULL context");
    SD_REQUIRE(!c->batch, "sd_ctx_remove_spheres: not available on a batch context");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_remove_spheres: no spheres uploaded");
    SD_REQUIRE(count >= 0 && (count == 0 || ids), "sd_ctx_remove_spheres: bad id list");
    if (count == 0) return SD_OK;
    const int n = c->n;
    std::vector<int> newid(n, 0);  // 1 = removed, then the new id of every kept one
    for (int k = 0; k < count; ++k) {
        SD_REQUIRE(ids[k] >= 0 && ids[k] < n, "sd_ctx_remove_spheres: id %d out of range [0, %d)", ids[k], n);
        SD_REQUIRE(newid[ids[k]] == 0, "sd_ctx_remove_spheres: id %d listed twice", ids[k]);
        newid[ids[k]] = 1;
    }
    SD_REQUIRE(count < n, "sd_ctx_remove_spheres: removing every sphere (destroy the context instead)");
    int next = 0;
    for (int i = 0; i < n; ++i) newid[i] = newid[i] ? -1 : next++;
    (void)hipStreamSynchronize(c->stream);
    HostRows r, o;
    sd_status st = read_rows(c, r);
    if (st != SD_OK) return st;
    const size_t maxc = (size_t)c->p.contact.max_contacts;
    const int m = n - count;
    o.n = m;
    o.pos.resize(m);
    o.vel.resize(m);
    o.omg.resize(m);
    o.ids.resize(m);
    o.hcount.assign(m, 0);
    o.hist.assign(maxc * m, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    int j = 0;
    for (int i = 0; i < n; ++i) {
        const int nid = newid[r.ids[i]];
        if (nid < 0) continue;
        o.pos[j] = r.pos[i];
        o.vel[j] = r.vel[i];
        o.omg[j] = r.omg[i];
        o.ids[j] = nid;
        int h = 0;
        for (int k = 0; k < r.hcount[i] && k < (int)maxc; ++k) {
            float4 e = r.hist[(size_t)k * n + i];
            int pid;
            memcpy(&pid, &e.w, sizeof pid);  // partner id stored as int bits; walls are -1..-6
            if (pid >= 0) {                   // a sphere partner: renumbered, or dropped with it
                if (pid >= n || newid[pid] < 0) continue;
                memcpy(&e.w, &newid[pid], sizeof pid);
            }
            o.hist[(size_t)h * m + j] = e;
            ++h;
        }
        o.hcount[j] = h;
        ++j;
    }
    return write_rows(c, o);
}

extern "C" sd_status sd_ctx_add_spheres(sd_ctx* c, int count, const float* x, const float* y, const float* z,
                                        const float* vx, const float* vy, const float* vz, const float* radius,
                                        const float* mass) {
    SD_REQUIRE(c, "sd_ctx_add_spheres: NULL context");
    SD_REQUIRE(!c->batch, "sd_ctx_add_spheres: not available on a batch context");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_add_spheres: no spheres uploaded (use sd_ctx_upload_spheres)");
    SD_REQUIRE(count >= 0, "sd_ctx_add_spheres: count < 0");
    if (count == 0) return SD_OK;
    SD_REQUIRE(x && y && z && vx && vy && vz && radius && mass, "sd_ctx_add_spheres: NULL array");
    SD_REQUIRE((long long)c->n + count < (1ll << 31), "sd_ctx_add_spheres: more than 2^31 spheres");
    sd_status st = check_spheres("sd_ctx_add_spheres", count, x, y, z, radius, mass, c->p.grid.cell);
    if (st != SD_OK) return st;
    (void)hipStreamSynchronize(c->stream);
    HostRows r, o;
    st = read_rows(c, r);
    if (st != SD_OK) return st;
    const int n = c->n, m = n + count;
    const size_t maxc = (size_t)c->p.contact.max_contacts;
    o.n = m;
    o.pos = r.pos;
    o.vel = r.vel;
    o.omg = r.omg;
    o.ids = r.ids;
    o.hcount = r.hcount;
    o.pos.resize(m);
    o.vel.resize(m);
    o.omg.resize(m);
    o.ids.resize(m);
    o.hcount.resize(m, 0);
    o.hist.assign(maxc * m, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (size_t k = 0; k < maxc; ++k)
        for (int i = 0; i < n; ++i) o.hist[k * m + i] = r.hist[k * n + i];
    for (int k = 0; k < count; ++k) {
        const float ms = mass[k];
        const float inv_m = ms > 0.0f ? 1.0f / ms : 0.0f;
        const float I = 0.4f * ms * radius[k] * radius[k];
        o.pos[n + k] = make_float4(x[k], y[k], z[k], radius[k]);
        o.vel[n + k] = make_float4(ms > 0.0f ? vx[k] : 0.0f, ms > 0.0f ? vy[k] : 0.0f, ms > 0.0f ? vz[k] : 0.0f,
                                   inv_m);
        o.omg[n + k] = make_float4(0.0f, 0.0f, 0.0f, ms > 0.0f ? 1.0f / I : 0.0f);
        o.ids[n + k] = n + k;
    }
    return write_rows(c, o);
}

extern "C" sd_status sd_ctx_state_bytes(sd_ctx* c, size_t* bytes) {
    SD_REQUIRE(c && bytes, "sd_ctx_state_bytes: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_state_bytes: no spheres uploaded");
    *bytes = state_bytes(c);
    return SD_OK;
}

extern "C" sd_status sd_ctx_save(sd_ctx* c, void* buffer, size_t bytes) {
    SD_REQUIRE(c && buffer, "sd_ctx_save: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_save: no spheres uploaded");
    SD_REQUIRE(bytes >= state_bytes(c), "sd_ctx_save: buffer too small (%zu < %zu)", bytes, state_bytes(c));
    const size_t n = (size_t)c->n;
    char* b = (char*)buffer;
    StateHeader h{STATE_MAGIC, STATE_VERSION, c->n, c->p.contact.max_contacts, c->p.key64, c->n_cells,
                  c->steps_done,  c->nscenes,    0,    layout_hash(c)};
    memcpy(b, &h, sizeof h);
    b += sizeof h;
    auto d2h = [&](const void* src, size_t sz) -> sd_status {
        SD_HIP(hipMemcpyAsync(b, src, sz, hipMemcpyDeviceToHost, c->stream));
        b += sz;
        return SD_OK;
    };
    sd_status st = SD_OK;
    const int cs = c->cur_state;
    if (st == SD_OK) st = d2h(c->pos[cs], n * sizeof(float4));
    if (st == SD_OK) st = d2h(c->vel[cs], n * sizeof(float4));
    if (st == SD_OK) st = d2h(c->omg[cs], n * sizeof(float4));
    if (st == SD_OK) st = d2h(c->ids[cs], n * sizeof(int));
    if (c->p.key64) {
        if (st == SD_OK) st = d2h(c->k64[c->cur_keys], n * sizeof(uint64_t));
    } else {
        if (st == SD_OK) st = d2h(c->k32[c->cur_keys], n * sizeof(uint32_t));
        if (st == SD_OK) st = d2h(c->v32[c->cur_keys], n * sizeof(uint32_t));
    }
    if (st == SD_OK) st = d2h(c->hcount[c->cur_hist], n * sizeof(int));
    if (st == SD_OK) st = d2h(c->hist[c->cur_hist], (size_t)c->p.contact.max_contacts * n * sizeof(float4));
    if (st != SD_OK) return st;
    SD_HIP(hipStreamSynchronize(c->stream));
    return SD_OK;
}

extern "C" sd_status sd_ctx_load(sd_ctx* c, const void* buffer, size_t bytes) {
    SD_REQUIRE(c && buffer, "sd_ctx_load: NULL");
    if (c->n == 0) return set_error(SD_ERR_STATE, "sd_ctx_load: upload a scene of the same size first");
    SD_REQUIRE(bytes >= state_bytes(c), "sd_ctx_load: buffer too small");
    StateHeader h;
    memcpy(&h, buffer, sizeof h);
    SD_REQUIRE(h.magic == STATE_MAGIC, "sd_ctx_load: not a state buffer");
    SD_REQUIRE(h.version == STATE_VERSION, "sd_ctx_load: state buffer version %u (this library reads %u)", h.version,
               STATE_VERSION);
    SD_REQUIRE(h.n == c->n && h.maxc == c->p.contact.max_contacts && h.key64 == c->p.key64 &&
                   h.n_cells == c->n_cells,
               "sd_ctx_load: state was saved from a context with different sizes");
    SD_REQUIRE(h.nscenes == c->nscenes && h.layout_hash == layout_hash(c),
               "sd_ctx_load: state was saved from a context with a different scene table or grid");
    const size_t n = (size_t)c->n;
    const char* b = (const char*)buffer + sizeof h;
    c->cur_keys = 0;
    c->cur_hist = 0;
    c->cur_state = 0;
    auto h2d = [&](void* dst, size_t sz) -> sd_status {
        SD_HIP(hipMemcpyAsync(dst, b, sz, hipMemcpyHostToDevice, c->stream));
        b += sz;
        return SD_OK;
    };
    sd_status st = SD_OK;
    if (st == SD_OK) st = h2d(c->pos[0], n * sizeof(float4));
    if (st == SD_OK) st = h2d(c->vel[0], n * sizeof(float4));
    if (st == SD_OK) st = h2d(c->omg[0], n * sizeof(float4));
    if (st == SD_OK) st = h2d(c->ids[0], n * sizeof(int));
    if (c->p.key64) {
        if (st == SD_OK) st = h2d(c->k64[0], n * sizeof(uint64_t));
    } else {
        if (st == SD_OK) st = h2d(c->k32[0], n * sizeof(uint32_t));
        if (st == SD_OK) st = h2d(c->v32[0], n * sizeof(uint32_t));
    }
    if (st == SD_OK) st = h2d(c->hcount[0], n * sizeof(int));
    if (st == SD_OK) st = h2d(c->hist[0], (size_t)c->p.contact.max_contacts * n * sizeof(float4));
    if (st == SD_OK) st = force_full_sort(c);
    if (st != SD_OK) return st;
    SD_HIP(hipStreamSynchronize(c->stream));
    c->steps_done = h.steps_done;
    return SD_OK;
}
