/*
 * sd_api.h — C ABI of libstardust_hip.so, the MI355X (gfx950) implementation of
 * stardust's per-timestep DEM / rigid-body step path.
 *
 * Conventions
 *  - Every call returns sd_status (SD_OK == 0). Nothing calls exit().
 *  - "stream" arguments are hipStream_t passed as void*; NULL selects the
 *    null stream. Per-kernel entry points take DEVICE pointers; the context
 *    API (sd_ctx_*) takes HOST pointers for upload/download and owns every
 *    device buffer it uses.
 *  - Results are deterministic and independent of the launch configuration
 *    (block size, tile size, items per thread).
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the spacesuit825/stardust source tree).
 */
#ifndef SD_API_H
#define SD_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef int sd_status;
#define SD_OK               0
#define SD_ERR_ARG          1   /* bad argument (null pointer, negative size, ...) */
#define SD_ERR_HIP          2   /* HIP runtime error; see sd_last_error() */
#define SD_ERR_CAPACITY     3   /* a fixed-capacity buffer overflowed; required size reported */
#define SD_ERR_STATE        4   /* call made in the wrong state (e.g. step before upload) */
#define SD_ERR_NO_DEVICE    5   /* no HIP device visible */
#define SD_ERR_NONFINITE    6   /* NaN/Inf detected in particle state */

/* Text of the last error recorded on the calling thread (never NULL). */
const char* sd_last_error(void);
/* Library version string. */
const char* sd_version(void);

/* ------------------------------------------------- device memory helpers
 * Thin wrappers over the HIP runtime so that callers without a HIP toolchain
 * (ctypes, cgo, JNI) can stage buffers for the per-kernel entry points. */
sd_status sd_device_count(int* count);
sd_status sd_malloc(void** dptr, size_t bytes);
sd_status sd_free(void* dptr);
sd_status sd_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
sd_status sd_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
sd_status sd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
sd_status sd_memset_u32(void* dst, uint32_t value, size_t count, void* stream); /* fill kernel, graph-safe */
sd_status sd_stream_create(void** stream);
sd_status sd_stream_destroy(void* stream);
sd_status sd_stream_synchronize(void* stream);
sd_status sd_event_create(void** ev);
sd_status sd_event_destroy(void* ev);
sd_status sd_event_record(void* ev, void* stream);
sd_status sd_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms);

/* ------------------------------------------------------------ grid / hash
 * Uniform hash grid over an axis-aligned domain. Cell coordinates are
 *   c_a = clamp(floor((p_a - lo_a) / cell), 0, n_a - 1)
 * and the linear cell id is (c_z * n_y + c_y) * n_x + c_x (x fastest).
 * Replaces the 3x8-bit truncated cell hash of constructCollisionListCUDA
 * (src/engine/cuda/sh_collision_detection.cu:27-127), one entry per particle
 * (its home cell) instead of 9. */
typedef struct sd_grid {
    float lo[3];        /* domain lower corner */
    float cell;         /* cell edge length; must be >= the largest sphere diameter */
    int   n[3];         /* cells per axis */
} sd_grid;

/* keys[i] = cell id of particle i, vals[i] = i.
 * Replaces constructCollisionListCUDA (sh_collision_detection.cu:27). */
sd_status sd_hash_keys_u32(const float* x, const float* y, const float* z, int n,
                           const sd_grid* grid, uint32_t* keys, uint32_t* vals, void* stream);
/* keys[i] = (uint64)cell << 32 | i  — the packed 64-bit (cellId, index) key. */
sd_status sd_hash_keys_u64(const float* x, const float* y, const float* z, int n,
                           const sd_grid* grid, uint64_t* keys, void* stream);

/* ------------------------------------------------------------ radix sort
 * Stable LSD radix sort over key bits [begin_bit, end_bit).
 * Replaces RadixTabulateCUDA / RadixSumCUDA / RadixReorderCUDA and
 * SpatialPartition::sortCollisionList (sh_collision_detection.cu:129-283,310-360).
 * The sorted result is left in keys/vals; keys_alt/vals_alt are ping-pong
 * buffers of the same size. temp must hold sd_radix_sort_temp_bytes(n) bytes. */
sd_status sd_radix_sort_temp_bytes(int n, size_t* bytes);
sd_status sd_radix_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                                  int n, int begin_bit, int end_bit, void* temp, size_t temp_bytes,
                                  void* stream);
sd_status sd_radix_sort_keys_u64(uint64_t* keys, uint64_t* keys_alt, int n, int begin_bit, int end_bit,
                                 void* temp, size_t temp_bytes, void* stream);
/* Stable sort of (key, value) pairs within each segment of a CSR partition
 * (device seg_first[0] = 0 <= ... <= seg_first[nseg] = n): every segment ends
 * sorted by the key bits [0, key_bits), equal keys in input order. temp must
 * hold sd_segmented_sort_temp_bytes(n) bytes. Replaces a sortCollisionList
 * call per scene (sh_collision_detection.cu:310-360) for callers that hold
 * segments; the batched step sorts its scenes in one sd_radix_sort_keys_u64
 * pass over scene-major packed keys instead. */
sd_status sd_segmented_sort_temp_bytes(int n, size_t* bytes);
sd_status sd_segmented_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                                      int n, const int* seg_first, int nseg, int key_bits, void* temp,
                                      size_t temp_bytes, void* stream);
/* As sd_radix_sort_pairs_u32, with no temp argument and no allocation: the
 * histogram workspace is a static device array of each GPU (about 8.4 MB,
 * held until the process ends), so n may not exceed SD_SORT_NOALLOC_MAX_N
 * (SD_ERR_ARG otherwise). No hipMalloc, hipFree or allocating library call is
 * made. Calls that overlap on one GPU (on different streams or host threads)
 * share that workspace and race; serialise them. The workspace needs no
 * clearing (every entry is written before it is read).
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
#define SD_SORT_NOALLOC_MAX_N 8388608
sd_status sd_radix_sort_pairs_u32_noalloc(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                                          int n, int begin_bit, int end_bit, void* stream);

/* --------------------------------------------------- cell ranges / pairs
 * ranges[2*c+0] = first sorted position whose cell is c, ranges[2*c+1] = one
 * past the last; an empty cell has start == end == 0. Every cell is written.
 * key_shift is 0 for u32 cell keys and 32 for packed u64 keys.
 * Replaces the run detection of tranverseCollisionListCUDA
 * (src/engine/cuda/collision_resolution.cu:100-199). */
sd_status sd_cell_ranges_u32(const uint32_t* sorted_keys, int n, int n_cells, int* ranges, void* stream);
sd_status sd_cell_ranges_u64(const uint64_t* sorted_keys, int n, int n_cells, int* ranges, void* stream);

/* Neighbour pairs of a sorted sphere set: every unordered pair (a < b, in sorted
 * positions) whose cells are 27-neighbours and whose spheres satisfy
 *   |p_a - p_b|^2 < (r_a + r_b + skin)^2.
 * pos4 is (x, y, z, r) per sorted position. Output pairs are written as
 * (a, b) int pairs, grouped by a ascending and, within a, in stencil order.
 * If the pair count exceeds capacity, *n_pairs receives the required count and
 * SD_ERR_CAPACITY is returned. Replaces resolveCollisionsCUDA's pair loop
 * (collision_resolution.cu:56-97) and the SAP pair rows (sap_detection.cu:293-324). */
sd_status sd_neighbor_pairs(const float* pos4, int n, const sd_grid* grid, const int* ranges, float skin,
                            int* pairs, int capacity, int* n_pairs, void* scratch, size_t scratch_bytes,
                            void* stream);
sd_status sd_neighbor_pairs_scratch_bytes(int n, size_t* bytes);

/* ------------------------------------------------------- contact law
 * Linear spring-dashpot normal force, tangential spring with history and
 * Coulomb limit, six optional planar walls. */
typedef struct sd_contact_params {
    float kn;            /* normal stiffness [N/m] */
    float damping_ratio; /* zeta; c = 2 zeta sqrt(k m_eff) for normal and tangential */
    float kt_ratio;      /* kt = kt_ratio * kn */
    float mu;            /* Coulomb coefficient */
    float dt;            /* time step [s] */
    float gravity[3];    /* [m/s^2] */
    int   walls;         /* 1: six planar walls at the domain boundary */
    float wall_lo[3];    /* wall planes (usually the grid domain) */
    float wall_hi[3];
    int   max_contacts;  /* per-particle contact-history capacity */
} sd_contact_params;

/* Sphere state in sorted order (struct of float4 arrays):
 *   pos4 = (x, y, z, radius), vel4 = (vx, vy, vz, 1/m), omg4 = (wx, wy, wz, 1/I).
 * Contact history: hist_count[s], hist4[k * n + s] = (dtx, dty, dtz, partner id bits).
 * Partner ids are stable particle ids (ids[s]); walls are ids -1..-6.
 * old_slot[s] is the position, in the previous step's order, of the particle now
 * at s (identity if the order did not change); old history is read there.
 *
 * sd_contact_forces: force4/torque4 of every sorted sphere, new history.
 * Replaces computeParticleToParticleCUDA / reactCollisionsCUDA
 * (src/engine/cuda/collisions/SAP/sap_reaction.cu:20-160); gather, not scatter. */
sd_status sd_contact_forces(const float* pos4, const float* vel4, const float* omg4, const int* ids, int n,
                            const sd_grid* grid, const int* ranges, const sd_contact_params* cp,
                            const int* old_slot, const int* hist_count_old, const float* hist4_old,
                            int* hist_count_new, float* hist4_new,
                            float* force4, float* torque4, int* overflow, void* stream);

/* Per-body force and torque from per-sphere forces: F_b = sum f_s and
 * T_b = sum ((x_s - X_b) x f_s + tau_s) over the body's spheres
 * [first[b], first[b] + count[b]) in sphere order (body_pos4 (X, *), sphere_pos4
 * (x, *), force4 / torque4 rows (., 0)). Replaces computeForcesAndTorquesCUDA
 * (src/engine/cuda/physics_update.cu:144-195, declared physics_update.cuh:19-59);
 * gravity and the static rule are applied by sd_integrate_bodies. The clump
 * step gathers each body's contacts directly and does not go through it. */
sd_status sd_body_reduce(const float* body_pos4, const int* first, const int* count, int nb,
                         const float* sphere_pos4, const float* sphere_force4, const float* sphere_torque4,
                         float* body_force4, float* body_torque4, void* stream);
/* One semi-implicit rigid-body step per body, in place: v += (F/m + g) dt,
 * x += v dt, L += T dt, w = R I^-1 R^T L, q = normalise(dq(w dt) q) with
 * pos4 = (x, y, z, 1/m) (1/m = 0: static), quat4 = (w, x, y, z), angmom4 =
 * world L, inv_inertia6 = body-frame inverse inertia (xx, yy, zz, xy, xz, yz);
 * dt and gravity from cp. The same arithmetic as the clump step's
 * integration. Replaces advectParticlesCUDA (physics_update.cu:232-310). */
sd_status sd_integrate_bodies(float* pos4, float* vel4, float* quat4, float* angmom4, const float* inv_inertia6,
                              const float* force4, const float* torque4, int nb, const sd_contact_params* cp,
                              void* stream);

/* Semi-implicit Euler for free spheres: v += (F/m + g) dt; x += v dt; w += T/I dt.
 * Spheres with 1/m == 0 are static. Replaces advectParticlesCUDA
 * (src/engine/cuda/physics_update.cu:232-310) for single-sphere bodies. */
sd_status sd_integrate_spheres(float* pos4, float* vel4, float* omg4, const float* force4, const float* torque4,
                               int n, const sd_contact_params* cp, void* stream);

/* Gather a sphere state into a new order: dst[t] = src[perm[t]] for pos4, vel4,
 * omg4 and ids. */
sd_status sd_permute_spheres(const int* perm, int n,
                             const float* pos4, const float* vel4, const float* omg4, const int* ids,
                             float* pos4_out, float* vel4_out, float* omg4_out, int* ids_out, void* stream);

/* ------------------------------------------------------- scene contexts
 * A context owns the device state of one sphere scene and runs whole steps:
 *   keys -> sort -> cell starts -> permute -> fused contact + integrate + next keys,
 * replayed from a hipGraph. Replaces DEMEngine::update
 * (src/engine/cuda/simulation_kernel.cu:52-153) and cleanBuffers (:30-50). */
typedef struct sd_sphere_params {
    sd_grid grid;
    sd_contact_params contact;
    int key64;           /* 0: u32 cell keys + u32 index values; 1: packed u64 (cell<<32 | index) */
    int use_graph;       /* 1: replay steps from a captured hipGraph */
    int tile_items;      /* sort items per thread (0 = autotuned) */
} sd_sphere_params;

typedef struct sd_ctx sd_ctx;

sd_status sd_ctx_create(const sd_sphere_params* p, sd_ctx** ctx);
sd_status sd_ctx_destroy(sd_ctx* ctx);
/* New contact law, dt, gravity or walls for the following steps (state and
 * contact history kept; max_contacts must stay the same). Lets
 * DEMEngine::update(float dt) (simulation_kernel.cu:52) take dt per call.
 * Not for batch contexts. */
sd_status sd_ctx_set_contact(sd_ctx* ctx, const sd_contact_params* cp);
/* Host arrays in particle-id order. mass may not be NULL; mass[i] == 0 marks a static sphere. */
sd_status sd_ctx_upload_spheres(sd_ctx* ctx, int n, const float* x, const float* y, const float* z,
                                const float* vx, const float* vy, const float* vz,
                                const float* radius, const float* mass);
/* Advance nsteps steps. Returns SD_ERR_CAPACITY if any step overflowed the
 * contact-history capacity (the run still completes; forces of the overflowed
 * contacts are included, their history is not kept). */
sd_status sd_ctx_step(sd_ctx* ctx, int nsteps);
/* Runs nsteps steps without graph replay, recording HIP events on the context
 * stream around each stage, and adds the time of each stage summed over the
 * steps to stage_ms: [0] sort, [1] cell starts (first sorted position of
 * every cell), [2] permute, [3] fused contact + integrate + next keys. With u64
 * keys, [0] is the re-sort from the previous order (positions permuted and
 * cell starts updated in the same pass; [1] is empty) when at most 1024
 * spheres changed cell in the previous step, else the radix sort (all passes),
 * and the positions are then permuted within [1]; [2] is empty. */
sd_status sd_ctx_profile(sd_ctx* ctx, int nsteps, float* stage_ms, int nstages);
/* Spheres that changed cell in the last step (u64 keys; 0 with u32 keys):
 * *full_next = 1 if the next step uses the full radix sort (more than 1024
 * changes, or after upload / load), else 0. Diagnostic with no counterpart in
 * the reference, whose sort does not depend on the previous order. */
sd_status sd_ctx_cell_changes(sd_ctx* ctx, int* count, int* full_next);
/* Host arrays in particle-id order (any pointer may be NULL). */
sd_status sd_ctx_download_spheres(sd_ctx* ctx, float* x, float* y, float* z,
                                  float* vx, float* vy, float* vz, float* wx, float* wy, float* wz);
/* Number of stored contacts (directed, sphere-sphere + sphere-wall) after the last step. */
sd_status sd_ctx_contact_count(sd_ctx* ctx, long long* count);
/* FP64 diagnostics: out[0..2] linear momentum, out[3..5] angular momentum about the
 * origin (orbital + spin), out[6] kinetic energy, out[7] gravitational potential energy. */
sd_status sd_ctx_diagnostics(sd_ctx* ctx, double out[8]);
/* Stream the context launches on (hipStream_t as void*). */
sd_status sd_ctx_stream(sd_ctx* ctx, void** stream);
/* Save / restore the complete device state (positions, velocities, contact history,
 * order) to a host byte buffer. sd_ctx_state_bytes gives the size. The buffer
 * (format version 2) records the sizes, the number of scenes and a hash of the
 * scene table (batch context: every scene's grid, contact law, cell and slot
 * range) or of the grid (single scene); sd_ctx_load rejects a buffer whose
 * sizes, scene count or hash differ from the context's with SD_ERR_ARG. */
sd_status sd_ctx_state_bytes(sd_ctx* ctx, size_t* bytes);
sd_status sd_ctx_save(sd_ctx* ctx, void* buffer, size_t bytes);
sd_status sd_ctx_load(sd_ctx* ctx, const void* buffer, size_t bytes);
/* Runtime add / remove of spheres (single-scene contexts; the reference's
 * EntityHandler::addEntity / removeEntity are stubs, entity_handler.cpp:108-119).
 * sd_ctx_remove_spheres drops the spheres with the given particle ids and
 * renumbers the rest in id order (new id = old id minus the removed ids below
 * it); the state rows keep their slot order, the contact history keeps every
 * entry whose partner stays (partner ids renumbered). sd_ctx_add_spheres
 * appends count spheres with ids n .. n + count - 1 (no history; 1/m and 1/I
 * from mass as at upload). Both reallocate the context, hash the keys from the
 * positions, make the next sort the full one and drop the captured step
 * graphs (the next sd_ctx_step captures new ones). */
sd_status sd_ctx_remove_spheres(sd_ctx* ctx, const int* ids, int count);
sd_status sd_ctx_add_spheres(sd_ctx* ctx, int count, const float* x, const float* y, const float* z, const float* vx,
                             const float* vy, const float* vz, const float* radius, const float* mass);
/* Persistent form of the fused step kernel (contexts with packed 64-bit keys,
 * single-scene or batch): instead of one 64-slot wave per launched wave, a
 * grid of resident workgroups (the occupancy limit times the CUs) whose waves
 * claim runs of eight 64-slot tiles from one counter per XCD (the XCD's eighth
 * of the slots) until they run out. Results are bit-identical to the default form; the
 * captured step graphs are dropped and captured again. on = 0 restores the
 * default form. */
sd_status sd_ctx_set_persistent(sd_ctx* ctx, int on);
/* Autotuner: times trial_steps eager steps for every radix-sort tile size
 * (4, 8, 16 items per thread) and contact block size (64, 128, 256), keeps the
 * fastest pair and returns it with its time per step. Results of a step do not
 * depend on these shapes; the state is restored after the trials. Generalises
 * the block-size sweeps of the reference's profiling runs (profiling/profile_1.py). */
sd_status sd_ctx_autotune(sd_ctx* ctx, int trial_steps, int* tile_items, int* contact_block, float* ms_per_step);

/* ------------------------------------------------------- batched scenes
 * Many independent sphere scenes stepped as one: one scene-segmented sort
 * (keys = (scene cell base + cell) << 32 | slot, so every scene keeps its own
 * sorted order), one cell-range pass, one contact launch in which every sphere
 * takes its scene's grid, contact law and walls, one graph per step pair.
 * A scene stepped in a batch is bit-identical to the same scene stepped alone
 * in an sd_ctx. dt, gravity and max_contacts must be equal in all scenes.
 * sd_ctx_step / _profile / _autotune / _download_spheres / _diagnostics /
 * _save / _load work on a batch context (arrays concatenated in scene order).
 * Replaces running DEMEngine::update (simulation_kernel.cu:52) once per scene. */
typedef struct sd_batch_params {
    int use_graph;       /* 1: replay steps from a captured hipGraph */
    int tile_items;      /* sort items per thread (0 = default for the size) */
    int contact_block;   /* contact kernel block size (0 = 128) */
} sd_batch_params;
typedef struct sd_scene_desc {
    int n;                      /* spheres of the scene */
    sd_grid grid;               /* the scene's own grid (cell >= largest diameter) */
    sd_contact_params contact;  /* the scene's law and walls */
} sd_scene_desc;
sd_status sd_batch_create(const sd_batch_params* p, sd_ctx** ctx);
sd_status sd_batch_upload(sd_ctx* ctx, int nscenes, const sd_scene_desc* scenes, const float* x, const float* y,
                          const float* z, const float* vx, const float* vy, const float* vz, const float* radius,
                          const float* mass);

/* ------------------------------------------------------- scene generators
 * Seeded SplitMix64 generators (host only). uniform = (u >> 40) * 2^-24,
 * normal = Box-Muller, lognormal = exp(ln(median) + sigma * normal).
 * Random sequential addition on a hash grid guarantees no initial overlap;
 * returns SD_ERR_CAPACITY if the requested count cannot be placed. */
sd_status sd_gen_spheres(uint64_t seed, int n, const float lo[3], const float hi[3],
                         float r_median, float r_sigma, float r_min, float r_max,
                         float* x, float* y, float* z, float* radius);

/* Convex hulls: points on ellipsoids (all of them hull vertices). Arrays:
 * pos4/vel4 (x, y, z, 0), quat4 (w, x, y, z) per hull, vstart/nv per hull,
 * verts4 (x, y, z, 0) local-frame vertices, capacity verts_capacity. */
sd_status sd_gen_hulls(uint64_t seed, int n, const float lo[3], const float hi[3], int nv_min, int nv_max,
                       float ax_min, float ax_max, float v_sigma, float* pos4, float* quat4, float* vel4,
                       int* vstart, int* nv, float* verts4, int verts_capacity, int* n_verts);

/* ------------------------------------------------------- convex hulls
 * Hull set (device arrays): pos4 = (x, y, z, r) centre (r: radius of a sphere
 * shape, nv == 0), quat4 = unit quaternion (w, x, y, z), vstart/nv = vertex
 * range of each hull in verts4 (local-frame (x, y, z, 0)).
 *
 * World AABB of every hull: lo4/hi4 = (x, y, z, 0).
 * Replaces computeAABBCUDA (sap_detection.cu:178-217) for rotated vertex sets. */
sd_status sd_hull_aabbs(const float* pos4, const float* quat4, const int* vstart, const int* nv, const float* verts4,
                        int n, float* lo4, float* hi4, void* stream);

/* Sweep-and-prune over AABBs: every unordered overlapping pair (closed
 * intervals on all three axes), written as (min id, max id), grouped by the
 * sweep position of the first box. The sweep axis is the one with the largest
 * variance of box centres (written to *axis_out if not NULL). If the pair count
 * exceeds capacity, *n_pairs receives the required count and SD_ERR_CAPACITY is
 * returned. Replaces SAPCollision::processCollisions (sap_detection.cu:668-706),
 * initIdx / projectAABB / sortLowestExtents / sweepAndPruneCUDA (:135-488). */
sd_status sd_sap_scratch_bytes(int n, size_t* bytes);
sd_status sd_sap_pairs(const float* lo4, const float* hi4, int n, int* pairs, int capacity, int* n_pairs,
                       int* axis_out, void* scratch, size_t scratch_bytes, void* stream);
/* Sweep-and-prune over many hull scenes at once: hulls concatenated scene by
 * scene (scene_first[0] = 0 <= ... <= scene_first[nscenes] = n, scene_of[i] the
 * scene of hull i; device arrays), at most 65,536 hulls per scene. Each scene is
 * swept on its own axis and gives exactly the pairs, in the same order, that
 * sd_sap_pairs gives for it alone (hull ids are global); no pair crosses a
 * scene. axes (device, nullable) receives the axis of every scene. Replaces
 * running processCollisions (sap_detection.cu:668-706) once per scene. */
sd_status sd_sap_batched_scratch_bytes(int n, int nscenes, size_t* bytes);
sd_status sd_sap_pairs_batched(const float* lo4, const float* hi4, int n, const int* scene_of,
                               const int* scene_first, int nscenes, int* pairs, int capacity, int* n_pairs,
                               int* axes, void* scratch, size_t scratch_bytes, void* stream);

/* GJK distance / EPA penetration for each pair (a, b) of hull ids. Outputs per
 * pair: sep (> 0 separation distance, < 0 penetration depth), normal3 (unit,
 * from a to b), flags (SD_GJK_* bits; -1 for ids or vertex ranges outside the
 * set, or a hull of more than 1,024 vertices), iters (support iterations spent,
 * GJK + EPA, warm + cold).
 * warm_dir (nullable, 3 floats per pair): initial direction of a warm-started
 * query limited to warm_budget iterations; a warm query that does not converge
 * within its budget, or ends within near_margin of contact, is redone from the
 * cold start (centre difference) with max_iter iterations.
 * Replaces gjkCUDA (sandbox/testGJK.cu:237-315). */
#define SD_GJK_HIT        1
#define SD_GJK_CONVERGED  2
#define SD_GJK_BUDGET     4
#define SD_GJK_EPA_LIMIT  8
#define SD_GJK_DEGENERATE 16
#define SD_GJK_WARM_USED  32
/* scratch: sd_gjk_scratch_bytes(n_pairs) bytes of device memory (the list of
 * intersecting pairs handed from the GJK kernel to the EPA kernel). */
sd_status sd_gjk_scratch_bytes(int n_pairs, size_t* bytes);
sd_status sd_gjk(const int* pairs, int n_pairs, const float* pos4, const float* quat4, const int* vstart,
                 const int* nv, const float* verts4, int n_hulls, int n_verts, const float* warm_dir, int warm_budget,
                 int max_iter, float near_margin, float* sep, float* normal3, int* flags, int* iters, void* scratch,
                 size_t scratch_bytes, void* stream);

/* ------------------------------------------------------- LBVH broadphase
 * Replaces the sandbox LBVH pipeline (sandbox/testLBVH.cu:42-954).
 * Global domain: domain6 = (min x, y, z of lo4, max x, y, z of hi4), computed
 * on the device (scratch: 256 * 6 floats). Replaces constructGlobalDomain (:61-88). */
sd_status sd_global_domain(const float* lo4, const float* hi4, int n, float* domain6, void* scratch, void* stream);
/* keys[i] = morton30(box centroid on a 1024^3 grid over the domain) << 32 | i.
 * Replaces computeSpaceFillingCurveCUDA (:114-151). */
sd_status sd_morton30(const float* lo4, const float* hi4, int n, const float* domain6, uint64_t* keys, void* stream);
/* Karras radix tree over sorted keys (internal nodes 0..n-2, leaf k = node n-1+k),
 * node boxes by a bottom-up refit, and ropes (next node in depth-first order
 * after a subtree, -1 at the end). Arrays of 2n-1 entries except left/right
 * (n-1) and counters (3n + 1 ints of scratch). *error (device int, cleared by the caller) gets
 * nonzero bits for a malformed tree. Replaces generatePrefixesCUDA ...
 * findLeafIndexRangeCUDA (:213-751). */
sd_status sd_lbvh_build(const uint64_t* sorted_keys, int n, const float* lo4, const float* hi4, int* left,
                        int* right, int* parent, int* first, int* last, int* rope, float* node_lo4, float* node_hi4,
                        int* counters, int* error, void* stream);
/* Every unordered pair of overlapping leaf boxes once, as (min id, max id),
 * grouped by the sorted position of the first leaf. Replaces
 * findPotentialCollisionsCUDA (:769-864). */
sd_status sd_lbvh_query_scratch_bytes(int n, size_t* bytes);
sd_status sd_lbvh_query(const uint64_t* sorted_keys, int n, const int* left, const int* last, const int* rope,
                        const float* node_lo4, const float* node_hi4, int* pairs, int capacity, int* n_pairs,
                        int* error, void* scratch, size_t scratch_bytes, void* stream);

/* ------------------------------------------------------- rigid clumps
 * A clump scene context: bodies made of spheres, LBVH broadphase over body
 * AABBs, sphere-sphere contacts (the sphere contact law without tangential
 * history), forces and torques summed per body, quaternion integration.
 * Walls are not applied. Replaces DEMEngine::update for multi-sphere entities
 * (simulation_kernel.cu:52-153 with physics_update.cu:26-348). */
/* ------------------------------------------------------- rigid convex bodies
 * Convex hulls (vertices in the body frame), spheres (no vertices, a radius)
 * and static triangles (three vertices, mass 0: the faces of a DEMMesh,
 * src/engine/entities/entity.hpp:64-93) stepped on the GPU in independent
 * scenes. Each step: a uniform grid per scene over the body boxes (fixed at
 * upload: the cell twice the largest moving body's bounding radius; larger
 * static bodies such as wall or mesh triangles in a second grid of their own,
 * bucketed at upload), GJK/EPA of every candidate pair (the sd_gjk results: TRIANGLE
 * and SPHERE shapes of sandbox/primitives.hpp:11-14 included; cold, or
 * warm-started by a learned network, sd_hull_set_warm), and
 * for each touching pair (sep < 0) a contact between the two contact
 * features (a hull's feature: its vertices within 1e-3 of its radius of the
 * extreme one along the normal, averaged; a sphere's surface point), placed at
 * the centroid of the feature of smaller spread (the midpoint of the two on a
 * tie: two points, two like faces), with
 *   f_n = kn * depth - c_n * (v_rel . n),  c_n = 2 zeta sqrt(kn m_eff),
 *   f_t = -min(mu * max(f_n, 0), c_n |v_t|) v_t / |v_t|   (no history),
 * applied to both bodies (no float atomics: each body sums its contacts in
 * pair order), then the semi-implicit integration of the clump path (1/m = 0:
 * static). The step is captured in a hipGraph and replayed. Replaces the
 * hull pipeline of the reference sandbox (testGJK.cu:237-315 on Hull pairs,
 * testGridSAP.cu), which reports overlap only. */
typedef struct sd_hull_params {
    sd_contact_params contact; /* kn, damping_ratio, mu, dt, gravity (kt_ratio, walls, max_contacts unused) */
    int max_iter;              /* GJK/EPA iteration limit; 0 = 64 */
    int pairs_per_body;        /* candidate-pair capacity per body; 0 = 32 */
    int use_graph;             /* capture the step in a hipGraph */
} sd_hull_params;
typedef struct sd_hull_ctx sd_hull_ctx;
sd_status sd_hull_create(const sd_hull_params* p, sd_hull_ctx** ctx);
sd_status sd_hull_destroy(sd_hull_ctx* ctx);
/* New contact law / dt / gravity / iteration limit for the following steps
 * (state kept; pairs_per_body cannot change after an upload). A context
 * attached to a batch must keep the batch's dt. */
sd_status sd_hull_set_params(sd_hull_ctx* ctx, const sd_hull_params* p);
/* n bodies: pos4 (x, y, z, mass; mass 0 = static), vel4, quat4 (w, x, y, z,
 * unit), angvel4 (world), inertia6 (body frame about the centre: xx, yy, zz,
 * xy, xz, yz), radius (sphere radius where nv == 0), vstart/nv into verts4
 * (body-frame vertices, n_verts in all; at most 1,024 per body); scenes are consecutive body ranges
 * (scene_first[0] = 0 ... scene_first[nscenes] = n; NULL for one scene), at
 * most 65,536 bodies each, with no contacts across scenes. */
sd_status sd_hull_upload(sd_hull_ctx* ctx, int n, const float* pos4, const float* vel4, const float* quat4,
                         const float* angvel4, const float* inertia6, const float* radius, const int* vstart,
                         const int* nv, int n_verts, const float* verts4, const int* scene_first, int nscenes);
/* nsteps steps. SD_ERR_CAPACITY when a step had more candidate pairs than the
 * capacity (its result is then invalid: upload again with a larger
 * pairs_per_body). */
sd_status sd_hull_step(sd_hull_ctx* ctx, int nsteps);
/* Body state: pos4 (x, y, z, 0), vel4, quat4, angmom4 (world L); any may be NULL. */
sd_status sd_hull_download(sd_hull_ctx* ctx, float* pos4, float* vel4, float* quat4, float* angmom4);
/* Learned warm start of the step's GJK queries: packed_model is a device
 * buffer of sd_mlp_packed_bytes() bytes written by sd_mlp_pack_bf16, kept
 * alive by the caller while attached; NULL returns to cold queries. Each step
 * then computes the pair features, infers the directions (BF16 MFMA) and runs
 * the warm-started queries of sd_gjk (warm_budget iterations, near margin
 * 1e-6 m), redoing cold every query without a converged, certified result.
 * No counterpart in the reference (its GJK starts from the centre difference,
 * sandbox/testGJK.cu:266). */
sd_status sd_hull_set_warm(sd_hull_ctx* ctx, const void* packed_model, int warm_budget);
/* Narrowphase of the last step (diagnostics, copies to the host): candidate
 * pairs, intersecting pairs, pairs whose result came from the warm-started
 * query, and the support passes of GJK and EPA summed over the pairs (as
 * sd_gjk's iters); any may be NULL. */
sd_status sd_hull_narrowphase_stats(sd_hull_ctx* ctx, int* n_pairs, int* n_hits, int* n_warm_used,
                                    long long* iters_total);
/* Touching pairs of the last step in pair order: (a, b), the force on b
 * (x, y, z, 1; -force on a) and the contact point (x, y, z, 0). */
sd_status sd_hull_contacts(sd_hull_ctx* ctx, int* pairs, float* force4, float* point4, int capacity,
                           int* n_contacts);
/* Steps the bodies of hulls with the scenes of a batch context: every
 * sd_ctx_step / sd_ctx_profile of batch then advances both, the body step on
 * a stream forked from and joined to the sphere step inside the same captured
 * graph (the bodies do not touch the spheres: they are separate scene sets
 * that share the timestep). Both must have the same dt. NULL detaches; detach
 * before destroying hulls. The autotuner's trials step the spheres only. */
sd_status sd_batch_attach_hulls(sd_ctx* batch, sd_hull_ctx* hulls);

typedef struct sd_clump_params {
    sd_contact_params contact;
} sd_clump_params;
typedef struct sd_clump_ctx sd_clump_ctx;
sd_status sd_clump_create(const sd_clump_params* p, sd_clump_ctx** ctx);
sd_status sd_clump_destroy(sd_clump_ctx* ctx);
/* New contact law / dt / gravity for the following steps (state kept). Lets
 * DEMEngine::update(float dt) (simulation_kernel.cu:52) take dt per call. */
sd_status sd_clump_set_params(sd_clump_ctx* ctx, const sd_clump_params* p);
/* Bodies: pos4 (x, y, z, mass; mass 0 = static), vel4 (v, 0), quat4 (w, x, y, z),
 * angvel4 (world angular velocity, 0), inertia6 (body frame: xx, yy, zz, xy, xz, yz
 * about the centre of mass), first/count sphere ranges. Spheres: local4
 * (body-frame offset from the centre of mass, radius). */
sd_status sd_clump_upload(sd_clump_ctx* ctx, int nb, const float* pos4, const float* vel4, const float* quat4,
                          const float* angvel4, const float* inertia6, const int* first, const int* count, int ns,
                          const float* local4);
sd_status sd_clump_step(sd_clump_ctx* ctx, int nsteps);
/* Stage times (ms, summed over nsteps eager steps): [0] prep + pose + body AABBs,
 * [1] domain + Morton + sort, [2] LBVH build, [3] and [4] empty (kept for the
 * layout), [5] partner search in the tree + contacts + integration. */
sd_status sd_clump_profile(sd_clump_ctx* ctx, int nsteps, float* stage_ms, int nstages);
sd_status sd_clump_download(sd_clump_ctx* ctx, float* pos4, float* vel4, float* quat4, float* angmom4);
/* Unique candidate body pairs (a < b, AABBs overlap) of the last step's tree,
 * built on this call (the step itself finds partners inside its contact
 * kernel); SD_ERR_CAPACITY with *n_pairs set when capacity is too small. */
sd_status sd_clump_pairs(sd_clump_ctx* ctx, int* pairs, int capacity, int* n_pairs);
sd_status sd_clump_spheres(sd_clump_ctx* ctx, float* pos4);
/* Checkpoint of a clump context: body positions, velocities, orientations and
 * angular momenta plus the step count (sd_clump_state_bytes gives the size).
 * sd_clump_load restores it into a context holding the same bodies (same
 * counts and the same inertia, sphere ranges and offsets, checked by a hash;
 * SD_ERR_ARG otherwise); stepping on from it is bit-identical to stepping on
 * from the saved context. Replaces the position read-back of the reference's
 * GL interop path (src/engine/engine.cpp:36-62) for resume and file output. */
sd_status sd_clump_state_bytes(sd_clump_ctx* ctx, size_t* bytes);
sd_status sd_clump_save(sd_clump_ctx* ctx, void* buffer, size_t bytes);
sd_status sd_clump_load(sd_clump_ctx* ctx, const void* buffer, size_t bytes);
/* Runtime add / remove of bodies (the reference's EntityHandler::addEntity /
 * removeEntity are stubs, entity_handler.cpp:108-119). sd_clump_remove_bodies
 * drops the listed bodies and their spheres; the rest keep their order and are
 * renumbered (new id = old id minus the removed ids below it).
 * sd_clump_add_bodies appends bodies given as in sd_clump_upload (first[] is
 * relative to the added local4 rows); they get ids nb, nb + 1, .... Both keep
 * the state of the other bodies, reallocate and drop the captured graphs. */
sd_status sd_clump_remove_bodies(sd_clump_ctx* ctx, const int* ids, int count);
sd_status sd_clump_add_bodies(sd_clump_ctx* ctx, int nb, const float* pos4, const float* vel4, const float* quat4,
                              const float* angvel4, const float* inertia6, const int* first, const int* count,
                              int ns, const float* local4);
/* Clumps of count U{c_min..c_max} spheres with centres N(0, sigma) about the
 * body centre, radii U[r_min, r_max], density rho; bodies on a jittered cubic
 * lattice (spacing, jitter) with uniform orientations and velocities
 * N(0, v_sigma). Outputs as sd_clump_upload takes them (centre of mass and
 * inertia from the member spheres). local4 capacity: nb * c_max. */
sd_status sd_gen_clumps(uint64_t seed, int nb, int lattice, float spacing, float jitter, int c_min, int c_max,
                        float sigma, float r_min, float r_max, float rho, float v_sigma, float* pos4, float* vel4,
                        float* quat4, float* angvel4, float* inertia6, int* first, int* count, float* local4,
                        int* ns);

/* ------------------------------------------------------- learned warm start
 * No counterpart in the reference (gjkCUDA starts from the centre difference,
 * sandbox/testGJK.cu:266). MLP 22 -> 64 -> 64 -> 64 -> 3 (SiLU), FP32 master
 * parameters (sd_mlp_param_count() floats, layout W1 b1 W2 b2 W3 b3 W4 b4 with
 * W[in][out]); loss 1 - u . t with u = y / |y|.
 *
 * Per hull: local half extents (3), mean / max / min vertex radius, count / 32. */
int sd_mlp_param_count(void);
int sd_mlp_packed_bytes(void);
sd_status sd_hull_descriptors(const float* pos4, const int* vstart, const int* nv, const float* verts4, int n,
                              float* desc7, void* stream);
/* 22 features per pair (relative pose in a's frame, both descriptors). */
sd_status sd_mlp_features(const int* pairs, int npairs, const float* pos4, const float* quat4, const float* desc7,
                          float* X, void* stream);
/* GJK normals (world, a -> b) expressed in a's frame: the training targets. */
sd_status sd_mlp_targets(const int* pairs, int npairs, const float* quat4, const float* normal3, float* T,
                         void* stream);
sd_status sd_mlp_init(float* params, uint64_t seed, void* stream);
/* Gradient of the mean loss over rows idx[0..batch) of X/T (deterministic,
 * fixed reduction order) and the loss itself (loss may be NULL). */
sd_status sd_mlp_train_scratch_bytes(int batch, size_t* bytes);
sd_status sd_mlp_forward_backward(const float* X, const float* T, const int* idx, int batch, const float* params,
                                  float* grad, float* loss, void* scratch, size_t scratch_bytes, void* stream);
/* Adam with bias correction; step counts from 1. */
sd_status sd_adam_step(float* params, const float* grad, float* m, float* v, int n, float lr, float beta1,
                       float beta2, float eps, int step, void* stream);
/* FP32 forward pass: unit directions of n feature rows from the FP32 master
 * parameters (the reference-precision form of sd_mlp_infer_bf16). */
sd_status sd_mlp_forward(const float* X, int n, const float* params, float* out_dir3, void* stream);
/* BF16 copy of the weights in the MFMA inference layout (sd_mlp_packed_bytes() bytes). */
sd_status sd_mlp_pack_bf16(const float* params, void* packed, void* stream);
/* Unit directions for n feature rows (BF16 MFMA GEMMs, FP32 accumulation).
 * With pairs/quat4 given the direction is rotated from a's frame to the world. */
sd_status sd_mlp_infer_bf16(const float* X, int n, const void* packed, const int* pairs, const float* quat4,
                            float* out_dir3, void* stream);

/* ------------------------------------------------------- compat restatements
 * The reference's own arithmetic, quirks included, for the formulas the
 * library replaces (no step calls these): parity tests use them to show the
 * reference's formulas are reproduced bit for bit.
 *
 * Pair force of computeParticleToParticleCUDA
 * (src/engine/cuda/collisions/SAP/sap_reaction.cu:20-70) for each pair
 * (home, phantom): separation = size[home] + size[phantom] (the reference
 * stores diameters there), relative position / velocity as float4 (w lanes
 * included), tangential velocity v - (v * n) * n component-wise, total =
 * -kn (sep - d) n + damping v + kt vt, multiplied component-wise by the signs
 * of v (zero counts as +1): force_home4[k] = sgn * total, force_phantom4[k] =
 * -sgn * total (the reference adds them unsynchronised into both spheres). */
sd_status sd_compat_pair_forces(const float* pos4, const float* vel4, const float* size, const int* pairs,
                                int npairs, float kn, float kt, float damping, float* force_home4,
                                float* force_phantom4, void* stream);
/* compute3x3Inverse (src/engine/cuda/cuda_helpers.cuh:65-77) of n row-ordered
 * float9 matrices: the input itself when the determinant is below 1e-4. */
sd_status sd_compat_inverse3x3(const float* m9, int n, float* out9, void* stream);
/* World angular velocity of advectParticlesCUDA (physics_update.cu:271-277):
 * w = M L with M = (R inv(I)) R^T evaluated by the column-major helpers
 * (cuda_helpers.cuh:79-112) on the row-major quatToRotationCUDA matrix
 * (:114-128), inv() the compat inverse; quat4 = (s, vx, vy, vz), inertia9 the
 * body's float9 inertia tensor, angmom4 world L. */
sd_status sd_compat_angular_velocity(const float* quat4, const float* inertia9, const float* angmom4, int nb,
                                     float* omega4, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SD_API_H */
