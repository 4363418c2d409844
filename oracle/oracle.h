/*
 * oracle.h -- CPU restatement of krABMaga's flockers agent-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against, and the single-threaded CPU timing
 * baseline.  It is never linked into, loaded by, or called from the product
 * library (krabmaga_amd/libkrabgpu.so).
 *
 * It follows the reference's sequential Field2D (Vec<Vec<O>> bags with a
 * read and a write buffer) and the in-tree flockers model line by line, in
 * f32 with IEEE division and square root and no FMA contraction (build with
 * -ffp-contract=off).  Reference files:
 *   src/engine/fields/field_2d.rs   (Field2D, sequential variant, :267-1014)
 *   tests/model/flockers/bird.rs    (Bird::step, :38-154)
 *   tests/model/flockers/state.rs   (Flocker::init/update, :41-60)
 *   src/engine/schedule.rs          (Schedule::step ordering, :347-413)
 *
 * Two choices the reference leaves open are fixed here:
 *   - randomness: the reference draws from rand::rng() (OS-seeded, not
 *     reproducible).  Here every draw comes from Philox4x32-10 keyed by a run
 *     seed, so runs are reproducible (see or_step_uniforms).
 *   - agent order inside one schedule step: the reference pops equal
 *     priorities in binary-heap order.  Here agents step in ascending id
 *     order unless a permutation is supplied (or_flocker_set_order), so each
 *     bag is filled in ascending id order.
 */
#ifndef KRAB_ORACLE_H
#define KRAB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 (Salmon et al., SC'11) ------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* u32 -> f32 in [0,1): top 24 bits times 2^-24 (rand's Standard f32 map). */
float or_u01(uint32_t u);
/* The two uniforms Bird::step draws for agent `id` at schedule step `step`. */
void or_step_uniforms(uint64_t seed, uint32_t id, uint64_t step, float *r1, float *r2);

/* ---- field_2d.rs free helpers (:926-1014) ------------------------------ */
int32_t or_t_transform(int32_t n, int32_t size);
float   or_toroidal_transform(float val, float dim);
float   or_toroidal_distance(float val1, float val2, float dim);
float   or_distance(float x1, float y1, float x2, float y2, float dim1, float dim2, int tor);
int     or_check_circle(int32_t bx, int32_t by, float discretization, float width, float height,
                        float lx, float ly, float dis, int tor);
int32_t or_f32_to_i32(float f); /* Rust `as i32`: saturating, NaN -> 0 */

/* ---- Field2D ----------------------------------------------------------- */
typedef struct {
    uint32_t id;
    float px, py; /* pos    */
    float hx, hy; /* last_d */
} or_bird;

typedef struct {
    or_bird *v;
    size_t n, cap;
} or_bag;

typedef struct {
    or_bag *bags[2];
    int read, write;
    size_t nbags;
    size_t nagents;
    float width, height, discretization;
    int toroidal;
    int32_t dh, dw;
    size_t density_estimation;
    int density_estimation_check;
} or_field2d;

or_field2d *or_field_new(float w, float h, float d, int toroidal);
void or_field_free(or_field2d *f);
void or_discretize(const or_field2d *f, float x, float y, int32_t *cx, int32_t *cy);
/* Returns 0, or -1 when the discretized index falls outside the bag array
 * (where the reference panics on the out-of-bounds Vec index). */
int  or_set_object_location(or_field2d *f, or_bird b, float x, float y);
void or_remove_object_location(or_field2d *f, uint32_t id, float x, float y);
void or_lazy_update(or_field2d *f);
/* Neighbour queries write up to `cap` birds into `out` and return the total
 * count (which may exceed cap).  Order is the reference's bag scan order. */
size_t or_get_neighbors_within_distance(const or_field2d *f, float x, float y, float dist,
                                        or_bird *out, size_t cap);
size_t or_get_neighbors_within_relax_distance(const or_field2d *f, float x, float y, float dist,
                                              or_bird *out, size_t cap);
size_t or_get_objects(const or_field2d *f, float x, float y, or_bird *out, size_t cap);
size_t or_get_objects_unbuffered(const or_field2d *f, float x, float y, or_bird *out, size_t cap);
size_t or_num_objects_at_location(const or_field2d *f, float x, float y);
size_t or_num_empty_bags(const or_field2d *f);
/* get_empty_bags as not_discretize positions in bag index order; writes up
 * to cap, returns the count */
size_t or_get_empty_bags(const or_field2d *f, float *x, float *y, size_t cap);
/* Bulk helpers for large test inputs (the same calls in a loop):
 * set_object_location of n birds in array order (returns the index of the
 * first out-of-range bird, or -1), and get_neighbors_within_distance at n
 * query points as CSR: row_ptr[n+1], ids (up to cap written); returns the
 * total id count. */
int64_t or_set_object_locations(or_field2d *f, const uint32_t *ids, const float *x, const float *y,
                                const float *hx, const float *hy, size_t n);
size_t or_neighbor_sets(const or_field2d *f, const float *qx, const float *qy, size_t n, float dist,
                        uint64_t *row_ptr, uint32_t *ids, size_t cap);
size_t or_nagents(const or_field2d *f);

/* ---- Flockers model ---------------------------------------------------- */
typedef struct {
    float cohesion, avoidance, consistency, randomness, momentum, jump, radius;
    uint64_t seed;
} or_flock_params;

typedef struct {
    or_field2d *field;
    or_bird *agents;   /* the schedule's copies, indexed by id */
    uint32_t *order;   /* stepping order within a schedule step */
    uint32_t n;
    float width, height;
    uint64_t step;
    or_flock_params p;
    or_bird *scratch;  /* neighbour vector (the reference allocates per call) */
    size_t scratch_cap;
} or_flocker;

/* Builds a flocker whose agents have ids 0..n-1 and the given state, and
 * puts them into the field's write buffer in id order (state.rs:41-56). */
or_flocker *or_flocker_new(uint32_t n, float w, float h, float d, int toroidal,
                           const float *x, const float *y, const float *hx, const float *hy,
                           const or_flock_params *p);
void or_flocker_free(or_flocker *fl);
/* Replace the stepping order (a permutation of 0..n-1). */
void or_flocker_set_order(or_flocker *fl, const uint32_t *order);
/* Schedule::step (schedule.rs:347-413) for the flocker state. */
void or_flocker_step(or_flocker *fl);
/* Bird::step (bird.rs:38-154) for agent `id`; exposed for unit tests. */
void or_bird_step(or_flocker *fl, uint32_t id);
void or_flocker_get(const or_flocker *fl, float *x, float *y, float *hx, float *hy);

/* ---- Initial state generators shared with the device library ---------- */
/* pos = dim * U[0,1) per axis (state.rs:42-46 with Philox in place of
 * rand::rng()).  heading_mode 0: (0,0) as state.rs:47; 1: unit vector at a
 * uniform angle by rejection from the square [-1,1)^2. */
void or_init_uniform(uint32_t n, float w, float h, uint64_t seed, int heading_mode,
                     float *x, float *y, float *hx, float *hy);

#ifdef __cplusplus
}
#endif
#endif
