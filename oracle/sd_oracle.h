/*
 * sd_oracle.h — CPU oracle for the DEM step path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the parity tests, smoke() and
 * the bench's cpu_baseline leg compare the HIP library against. Nothing in
 * the product (libstardust_hip.so, the stardust_amd package) links, loads or
 * calls it.
 *
 * It restates, in plain C (FP32, no contraction, same operation order as the
 * documented contact law), the reference algorithms of spacesuit825/stardust
 * with the intended well-defined outputs (stable sort order, complete pair
 * sets); every function cites the reference file:line it follows. Each function
 * is pinned by the tests in tests/test_oracle.py against closed forms, brute
 * force or textbook routines.
 */
#ifndef SD_ORACLE_H
#define SD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same memory layout as sd_grid / sd_contact_params of the C ABI. */
typedef struct or_grid {
    float lo[3];
    float cell;
    int n[3];
} or_grid;

typedef struct or_contact {
    float kn;
    float damping_ratio;
    float kt_ratio;
    float mu;
    float dt;
    float gravity[3];
    int walls;
    float wall_lo[3];
    float wall_hi[3];
    int max_contacts;
} or_contact;

/* Home-cell key of each particle (sh_collision_detection.cu:60-66). */
void or_hash_keys(int n, const float* x, const float* y, const float* z, int stride, const or_grid* g,
                  uint32_t* keys, uint32_t* vals);
uint32_t or_cell_id(const or_grid* g, float x, float y, float z);

/* Stable LSD radix sort, 8-bit digits, over key bits [begin, end)
 * (sh_collision_detection.cu:129-283,310-360: tabulate, prefix, stable reorder). */
void or_sort_pairs_u32(int n, uint32_t* keys, uint32_t* vals, int begin_bit, int end_bit);
void or_sort_keys_u64(int n, uint64_t* keys, int begin_bit, int end_bit);

/* Cell start/end of a sorted key array (collision_resolution.cu:145-178 run detection);
 * empty cells are (0, 0). key_shift selects the cell bits of a packed key. */
void or_cell_ranges_u32(int n, const uint32_t* keys, int n_cells, int* ranges);
void or_cell_ranges_u64(int n, const uint64_t* keys, int n_cells, int* ranges);

/* Neighbour pairs (a < b in sorted order, 27-cell stencil, |d|^2 < (ra+rb+skin)^2),
 * grouped by a, stencil order within a (collision_resolution.cu:56-97).
 * Returns the pair count; writes at most capacity pairs. */
long or_neighbor_pairs(int n, const float* pos4, const or_grid* g, const int* ranges, float skin, int* pairs,
                       long capacity);
/* Brute force: all unordered pairs with |d|^2 < (ra+rb+skin)^2 (i < j). */
long or_brute_pairs(int n, const float* pos4, float skin, int* pairs, long capacity);

/* Contact forces of every sphere of a sorted state (sap_reaction.cu:20-70,72-134,
 * restated as a deterministic gather with the documented contact law).
 * Returns the largest contact count seen. */
int or_contact_forces(int n, const float* pos4, const float* vel4, const float* omg4, const int* ids,
                      const or_grid* g, const int* ranges, const or_contact* c, const int* old_slot,
                      const int* hcount_old, const float* hist4_old, int* hcount_new, float* hist4_new, float* force4,
                      float* torque4, int nthreads);
/* The same in FP64 (reference for error reports; history in double). */
int or_contact_forces_f64(int n, const double* pos4, const double* vel4, const double* omg4, const int* ids,
                          const or_grid* g, const int* ranges, const or_contact* c, const int* old_slot,
                          const int* hcount_old, const double* hist4_old, int* hcount_new, double* hist4_new,
                          double* force4, double* torque4);

/* Semi-implicit Euler (physics_update.cu:264-269 for the linear part). */
void or_integrate(int n, float* pos4, float* vel4, float* omg4, const float* force4, const float* torque4,
                  const or_contact* c);

/* One full step on a state in slot order: keys, stable sort, ranges, permute,
 * forces, integrate (simulation_kernel.cu:52-153 step order). All arrays are
 * updated in place to the new order. Returns the largest contact count. */
int or_sphere_step(int n, const or_grid* g, const or_contact* c, float* pos4, float* vel4, float* omg4, int* ids,
                   int* hcount, float* hist4, int nthreads);

/* ------------------------------------------------------------------ hulls
 * (sd_oracle_hull.c) hull set: pos4 (x, y, z, sphere radius), quat4 (w, x, y, z),
 * vstart/nv vertex ranges into verts4 (local (x, y, z, 0)). */
void or_hull_aabbs(int n, const float* pos4, const float* quat4, const int* vstart, const int* nv,
                   const float* verts4, float* lo4, float* hi4);
/* Axis of largest box-centre variance (sap_detection.cu:589 TODO). */
int or_sap_axis(int n, const float* lo4, const float* hi4);
/* Serial sweep-and-prune (sap_detection.cu:668-706 pipeline, exact pair set). */
long or_sap_pairs(int n, const float* lo4, const float* hi4, int axis, int* pairs, long capacity);
long or_brute_aabb_pairs(int n, const float* lo4, const float* hi4, int* pairs, long capacity);
/* GJK distance + EPA depth per pair (testGJK.cu:52-315 restated, terminating). */
void or_gjk(int npairs, const int* pairs, const float* pos4, const float* quat4, const int* vstart, const int* nv,
            const float* verts4, const float* warm, int warm_budget, int max_iter, float near_margin, float* sep,
            float* normal3, int* flags, int* iters, int nthreads);
/* As or_gjk; small_iters (nullable) receives the iterations each warm query
 * spent on its seed problem (the top-3 supports of both hulls). */
void or_gjk_ex(int npairs, const int* pairs, const float* pos4, const float* quat4, const int* vstart, const int* nv,
               const float* verts4, const float* warm, int warm_budget, int max_iter, float near_margin, float* sep,
               float* normal3, int* flags, int* iters, int* small_iters, int nthreads);

/* ------------------------------------------------------------------ clumps
 * (sd_oracle_clump.c) body state P (x, y, z, 1/m), V, Q (w, x, y, z), L (world
 * angular momentum), inverse body inertia I0 (xx, yy, zz, xy) / I1 (xz, yz). */
void or_clump_prepare(int nb, const float* pos4, const float* vel4, const float* quat4, const float* angvel4,
                      const float* inertia6, float* P, float* V, float* Q, float* L, float* I0, float* I1);
void or_clump_pose(int nb, const float* P, const float* V, const float* Q, const float* L, const float* I0,
                   const float* I1, const int* first, const int* count, const float* local4, float* sx, float* sv,
                   float* lo4, float* hi4);
void or_global_domain(int n, const float* lo4, const float* hi4, float* dom6);
void or_morton30(int n, const float* lo4, const float* hi4, const float* dom6, uint64_t* keys);
int or_clump_contacts_integrate(int nb, float* P, float* V, float* Q, float* L, const float* I0, const float* I1,
                                const int* first, const int* count, const float* sx, const float* sv, long npairs,
                                const int* pairs, const or_contact* c, float* force4, float* torque4, int integrate);
void or_body_reduce(int nb, const float* bpos4, const int* first, const int* count, const float* spos4,
                    const float* sf4, const float* st4, float* bf4, float* bt4);
void or_integrate_bodies(int nb, float* P, float* V, float* Q, float* L, const float* inv6, const float* f4,
                         const float* t4, const or_contact* c);
/* Per-scene broadphase grids of the rigid-body context from its initial
 * bodies (16 values per scene: the grid of the bodies, then that of the
 * large static bodies, large[b] = 1); returns the total cell count of the
 * first grids, *large_cells (nullable) that of the second. */
long or_hull_grid(int n, const float* pos4, const float* radius, const int* vstart, const int* nv,
                  const float* verts4, const int* scene_first, int nscenes, float* grid16, int* large,
                  long* large_cells);
long or_hull_grid_pairs(int n, const float* P, const float* lo4, const float* hi4, const int* scene_first,
                        int nscenes, const float* grid16, const int* large, int* pairs, long capacity);
/* One step of the rigid convex-body context (sd_hull_step) over the grids of
 * or_hull_grid: returns the number of touching pairs; the first out_capacity
 * are written (pairs, force on b, contact point; all nullable). */
long or_hull_step(int n, float* P, float* V, float* Q, float* L, const float* I0, const float* I1, const float* radius,
                  const int* vstart, const int* nv, const float* verts4, int n_verts, const int* scene_first,
                  int nscenes, const float* grid16, const int* large, const or_contact* c, int max_iter, int nthreads,
                  int* pairs_out, float* force_out, float* point_out, long out_capacity);
long or_clump_step(int nb, float* P, float* V, float* Q, float* L, const float* I0, const float* I1, const int* first,
                   const int* count, const float* local4, int ns, const or_contact* c);

/* ------------------------------------------------------------------ compat
 * (sd_oracle_compat.c) the reference's own arithmetic, quirks included:
 * sap_reaction.cu:20-70 pair force, cuda_helpers.cuh:65-77 inverse,
 * physics_update.cu:271-277 angular velocity with cuda_helpers.cuh:79-128. */
void or_compat_pair_forces(int npairs, const int* pairs, const float* pos4, const float* vel4, const float* size,
                           float kn, float kt, float damping, float* fh4, float* fp4);
void or_compat_inverse3x3(int n, const float* m9, float* o9);
void or_compat_angular_velocity(int nb, const float* quat4, const float* inertia9, const float* angmom4,
                                float* omega4);

#ifdef __cplusplus
}
#endif
#endif
