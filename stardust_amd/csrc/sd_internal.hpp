// Internal C++ interfaces shared between the translation units of
// libstardust_hip.so (not part of the C ABI).
#pragma once

#include "sd_common.hpp"

namespace sd {

// ------------------------------------------------------------------ sort
// Bytes of scratch the radix sort needs for n elements (any tile size).
size_t radix_temp_bytes(int n);
// Items per thread the sort uses for n elements when the caller passes 0.
int radix_default_items(int n);

// Stable LSD sort over key bits [begin_bit, end_bit). On return *keys (and
// *vals) point at the buffer holding the sorted data; the other buffer of each
// pair is left in *alt. items = 4, 8 or 16 (0 = default for n).
sd_status radix_sort_pairs_u32(uint32_t** keys, uint32_t** vals, uint32_t** keys_alt, uint32_t** vals_alt,
                               int n, int begin_bit, int end_bit, void* temp, size_t temp_bytes,
                               hipStream_t s, int items);
sd_status radix_sort_keys_u64(uint64_t** keys, uint64_t** keys_alt, int n, int begin_bit, int end_bit,
                              void* temp, size_t temp_bytes, hipStream_t s, int items);
// As radix_sort_keys_u64, with an odd number of passes (the result always ends
// in *keys_alt's original buffer) and every kernel skipped on the device while
// *gate <= gate_cap.
sd_status radix_sort_keys_u64_gated(uint64_t** keys, uint64_t** keys_alt, int n, int begin_bit, int end_bit,
                                    void* temp, size_t temp_bytes, hipStream_t s, int items, const int* gate,
                                    int gate_cap);
// Number of LSD passes the sort makes over a bit range.
int radix_passes(int begin_bit, int end_bit);

// ------------------------------------------------------------------ scan
// Exclusive prefix sum of n int32 values (in may equal out). scratch must
// hold scan_temp_bytes(n) bytes. If total != nullptr the grand total is
// written there (device pointer).
size_t scan_temp_bytes(int n);
sd_status exclusive_scan_i32(const int* in, int* out, int n, int* total, void* scratch, hipStream_t s);

// ------------------------------------------------------------------ grid
struct GridDev {
    float lox, loy, loz, cell;
    int nx, ny, nz;
};
inline GridDev to_dev(const sd_grid& g) {
    return GridDev{g.lo[0], g.lo[1], g.lo[2], g.cell, g.n[0], g.n[1], g.n[2]};
}
int grid_bits(const sd_grid& g);  // bits needed for the largest cell id
sd_status check_grid(const sd_grid* g);

sd_status launch_fill_u32(uint32_t* dst, uint32_t value, size_t count, hipStream_t s);
sd_status launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t count, hipStream_t s);
sd_status launch_hash_keys_u32(const float4* pos4, int n, const GridDev& g, uint32_t* keys, uint32_t* vals,
                               hipStream_t s);
sd_status launch_hash_keys_u64(const float4* pos4, int n, const GridDev& g, uint64_t* keys, hipStream_t s);
struct SceneDev;
// keys[i] = (scene cbase + cell of pos4[i] in the scene grid) << 32 | i
sd_status launch_hash_keys_batch(const float4* pos4, int n, const SceneDev* scenes, const int* scene_of,
                                 uint64_t* keys, hipStream_t s);
sd_status launch_cell_ranges_u32(const uint32_t* keys, int n, int n_cells, int* ranges, hipStream_t s);
sd_status launch_cell_ranges_u64(const uint64_t* keys, int n, int n_cells, int* ranges, hipStream_t s);
// Cell starts of sorted keys: start[c] = number of keys whose cell is below c,
// c = 0 .. n_cells (n_cells + 1 entries), every entry written. The u64 form
// also gathers pos into sorted order (pos_o) with permute and sets
// *next_count = 0; it does nothing while *skip.
sd_status launch_cell_starts_u32(const uint32_t* keys, int n, int n_cells, int* start, hipStream_t s);
sd_status launch_step_starts_u64(const uint64_t* keys, int n, int n_cells, int* start, const int* skip, bool permute,
                                 const float4* pos, float4* pos_o, int* next_count, hipStream_t s);
// perm is either u32 values (perm32) or the low halves of u64 keys (perm64).
sd_status launch_permute(const uint32_t* perm32, const uint64_t* perm64, int n, const float4* pos, const float4* vel,
                         const float4* omg, const int* ids, float4* pos_o, float4* vel_o, float4* omg_o, int* ids_o,
                         hipStream_t s);

// gate (nullable): skipped on the device while *gate <= gate_cap.
sd_status launch_permute_pos(const uint32_t* perm32, const uint64_t* perm64, int n, const float4* pos, float4* pos_o,
                             hipStream_t s, const int* gate = nullptr, int gate_cap = 0);

// ------------------------------------------------------------------ contacts
struct ContactDev {
    float kn, kt, c2z, mu, dt;
    float gx, gy, gz;
    int walls;
    float wlo[3], whi[3];
    int maxc;
};
ContactDev to_dev(const sd_contact_params& c);
sd_status check_contact(const sd_contact_params* c);

// Entries past n that the sorted-position buffer of a context carries, so the
// contact kernel's batched candidate loads need no clamping.
constexpr int CONTACT_POS_PAD = 4;

// Fused step kernel modes. CONTACT_FUSED_BATCH: packed u64 keys, every sphere
// takes its grid, contact law and cell-id base from its scene's SceneDev.
enum { CONTACT_FORCES = 0, CONTACT_FUSED_U32 = 1, CONTACT_FUSED_U64 = 2, CONTACT_FUSED_BATCH = 3 };

// One scene of a batch: its grid, its contact law, the first global cell id of
// its cells and its slot range [first, first + n) (sorted positions of a batch
// keep scenes contiguous because cell ids are scene-major).
struct SceneDev {
    GridDev g;
    ContactDev c;
    uint32_t cbase;
    int first;
    int n;
    int pad;
};

struct ContactArgs {
    const float4* pos;  // sorted state
    const float4* vel;
    const float4* omg;
    const int* ids;
    int n;
    GridDev g;
    const int* ranges;      // CONTACT_FORCES: (start, end) per cell, empty cells (0, 0)
    const int* cell_start;  // fused modes: first sorted position of every cell (n_cells + 1 entries)
    ContactDev c;
    // old order: either an int array, u32 sort values or packed u64 keys
    const int* old_slot_i32;
    const uint32_t* old_slot_u32;
    const uint64_t* old_slot_u64;
    const int* hcount_old;
    const float4* hist_old;
    int* hcount_new;
    float4* hist_new;
    int* overflow;       // receives max contacts seen when > capacity
    // CONTACT_FORCES outputs
    float4* force;
    float4* torque;
    // fused outputs: integrated state + next keys
    float4* pos_o;
    float4* vel_o;
    float4* omg_o;
    int* ids_o;
    uint32_t* keys_o32;
    uint32_t* vals_o32;
    uint64_t* keys_o64;
    // fused u64 modes: spheres whose cell changed this step, appended as
    // (new key, old key) while *mcount < mcap (nullable: not recorded)
    int* mcount;
    uint64_t* mnew;
    uint64_t* mold;
    int mcap;
    // nullable: [0] != 0 -> the step was given up, return at once; [1] counts
    // the steps that ran
    int* step_ctl;
    // CONTACT_FUSED_BATCH: scene table and scene of every slot
    const SceneDev* scenes;
    const int* scene_of;
    // CONTACT_FUSED_BATCH: the waves (64 consecutive slots from a multiple of
    // 64) whose slots span more than one scene, n_straddle of them. The main
    // launch leaves them out (its waves take their scene's parameters in scalar
    // registers); launch_contacts_straddle runs them with per-lane parameters.
    const int* straddle;
    int n_straddle;
};
sd_status launch_contacts(int mode, const ContactArgs& a, hipStream_t s, int block);
// Persistent form of the fused u64 / batch step kernel (same results): a
// grid of resident blocks claiming runs of 64-slot tiles from one counter per
// XCD (counters: 512 ints, cleared here by a fill kernel) until every tile is
// done.
sd_status launch_contacts_persistent(int mode, const ContactArgs& a, hipStream_t s, int block, int* counters);
sd_status launch_contacts_straddle(const ContactArgs& a, hipStream_t s);

// Sorted order of the next step's u64 keys from the previous sorted order and
// the spheres that changed cell (see sd_resort.hip): writes the sorted keys to
// keys_out and the positions in that order to posB, turns cell_start (the
// previous order's cell starts, n_cells + 1 entries) into the new order's and
// sets *next_count = 0. With *mcount > RESORT_CAP it gives the step up instead
// (step_ctl[0] = 1).
constexpr int RESORT_CAP = 1024;
// sd_lbvh_build, and with pnode (2 per node) also the packed tree of a
// stackless walk: (box low corner, left child or leaf index), (box high corner, rope)
sd_status lbvh_build(const uint64_t* sorted_keys, int n, const float* lo4, const float* hi4, int* left, int* right,
                     int* parent, int* first, int* last, int* rope, float* node_lo4, float* node_hi4, int* counters,
                     int* error, float4* pnode, hipStream_t s);
// keys[i] = (morton30 of box i's centroid with its drop_bits lowest bits cleared) << 32 | i
sd_status launch_morton(const float* lo4, const float* hi4, int n, const float* domain6, uint64_t* keys,
                        int drop_bits, hipStream_t s);
sd_status launch_resort_u64(const int* mcount, const uint64_t* mnew, const uint64_t* mold, uint64_t* msorted_new,
                            uint64_t* msorted_old, const uint64_t* keys, int n, int* cell_start, int n_cells,
                            const float4* pos_state, float4* posB, uint64_t* keys_out, int* next_count,
                            int* step_ctl, hipStream_t s);
sd_status launch_integrate(float4* pos, float4* vel, float4* omg, const float4* f, const float4* t, int n,
                           const ContactDev& c, hipStream_t s);

// Rigid-body rows from upload arguments (pos4 = (x, y, z, mass), mass 0 =
// static; quat4 = (w, x, y, z); angvel4 world; inertia6 = (xx, yy, zz, xy, xz,
// yz) body frame): P = (x, y, z, 1/m), V, Q, L = world angular momentum,
// I0/I1 = body-frame inverse inertia (xx, yy, zz, xy) / (xz, yz, 0, 0).
sd_status prepare_rigid(const char* who, int nb, const float* pos4, const float* vel4, const float* quat4,
                        const float* angvel4, const float* inertia6, float4* P, float4* V, float4* Q, float4* L,
                        float4* I0, float4* I1);

// Rigid-body context (sd_hull_ctx, sd_hullbody.hip): one step enqueued on s
// with no host synchronisation; device errors of the steps since the last
// check; accessors used by a batch context that steps the bodies with its
// scenes.
sd_status hull_enqueue_step(sd_hull_ctx* c, hipStream_t s);
sd_status hull_check_errors(sd_hull_ctx* c);
int hull_body_count(const sd_hull_ctx* c);
// Changes whenever the context's step changes (upload, parameters, warm start).
long long hull_generation(const sd_hull_ctx* c);
float hull_dt(const sd_hull_ctx* c);
void hull_count_steps(sd_hull_ctx* c, int nsteps);

// ------------------------------------------------------------------ hull narrowphase, enqueued only
// Hull boxes (pos4 = centre and sphere radius, quat4 = (w, x, y, z)).
sd_status launch_hull_aabbs(const float4* pos4, const float4* quat4, const int* vstart, const int* nv,
                            const float4* verts4, int n, float4* lo4, float4* hi4, hipStream_t s);
// sd_sap_pairs_batched without its host read-back: *total_out points at the
// device pair count (may exceed capacity; pairs past capacity are not
// written), *error_out at the device error bits of the scene ranges.
sd_status enqueue_sap_batched(const float4* lo4, const float4* hi4, int n, const int* scene_of,
                              const int* scene_first, int nscenes, int* pairs, int capacity, int** total_out,
                              int** error_out, void* scratch, size_t scratch_bytes, hipStream_t s);
// Cold GJK/EPA (as sd_gjk with no warm start) of the first min(*dcount,
// capacity) pairs; launches sized to capacity, scratch of
// gjk_cold_scratch_bytes(capacity) bytes.
size_t gjk_cold_scratch_bytes(int capacity);
sd_status enqueue_gjk_cold(const int* pairs, int capacity, const int* dcount, const float4* pos4,
                           const float4* quat4, const int* vstart, const int* nv, const float4* verts4, int n_hulls,
                           int n_verts, int max_iter, float near_margin, float* sep, float* normal3, int* flags,
                           int* iters, void* scratch, hipStream_t s);
// Warm-started GJK/EPA (as sd_gjk with warm_dir3) of the first min(*dcount,
// capacity) pairs, the rejected ones redone cold; same scratch.
sd_status enqueue_gjk_warm(const int* pairs, int capacity, const int* dcount, const float4* pos4,
                           const float4* quat4, const int* vstart, const int* nv, const float4* verts4, int n_hulls,
                           int n_verts, const float* warm_dir3, int warm_budget, int max_iter, float near_margin,
                           float* sep, float* normal3, int* flags, int* iters, void* scratch, hipStream_t s);

// ------------------------------------------------------------------ learned warm start, enqueued only
// sd_hull_descriptors on a stream (desc7: 7 floats per hull).
sd_status launch_hull_descriptors(const float4* pos4, const int* vstart, const int* nv, const float4* verts4, int n,
                                  float* desc7, hipStream_t s);
// Features (X: 22 floats per pair) and BF16 inference of the network packed
// by sd_mlp_pack_bf16 for the first min(*dcount, capacity) pairs: world
// directions (A -> B) in out_dir3.
sd_status enqueue_warm_dirs(const int* pairs, int capacity, const int* dcount, const float4* pos4,
                            const float4* quat4, const float* desc7, float* X, const void* packed, float* out_dir3,
                            hipStream_t s);

}  // namespace sd
