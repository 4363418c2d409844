/*
 * krabgpu.h -- C ABI of the MI355X (gfx950) agent-step engine for krABMaga's
 * continuous 2-D field and flockers model.
 *
 * Every entry point returns KG_OK (0) or a negative KG_E_* status, never
 * unwinds across the boundary, and records a message retrievable with
 * kg_last_error() on the calling thread.  Host pointers are borrowed for the
 * duration of a call; device memory is owned by the library objects.
 *
 * A kg_field is one Field2D<Bird> (src/engine/fields/field_2d.rs:267-923)
 * living in HBM: a read buffer and a write buffer of agents, each agent an
 * id (u32), a position (Real2D, f32) and a heading (the flockers model's
 * last_d, f32).  One host thread uses a kg_field at a time; distinct fields
 * may be used concurrently.  Each kg_field owns one HIP stream and all its
 * work is ordered on it.
 *
 * Ordering contract.  Inside a bag the reference keeps insertion order,
 * which for a schedule step is the scheduler's pop order.  This library keeps
 * every bag sorted by agent id, i.e. the order a scheduler stepping agents in
 * ascending id produces; neighbour queries return agents in bag-scan order
 * (cell x ascending, then cell y ascending, then id ascending), as
 * field_2d.rs:418-435 does for such a run.
 */
#ifndef KRABGPU_H
#define KRABGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KG_OK 0
#define KG_E_INVALID (-1)  /* bad argument */
#define KG_E_CAPACITY (-2) /* caller buffer or field capacity too small; the
                              required size is reported where the call says */
#define KG_E_RANGE (-3)    /* a location discretizes outside the bag grid
                              (the reference indexes out of bounds and panics) */
#define KG_E_HIP (-4)      /* HIP runtime error */
#define KG_E_NOMEM (-5)    /* device or host allocation failed */
#define KG_E_STATE (-6)    /* call not valid in the field's current state */
#define KG_E_ADMIT (-7)    /* a run does not fit the engine's memory budget */

typedef struct kg_field kg_field;

typedef struct kg_field_desc {
    float width;          /* Field2D::new(w, ...) */
    float height;         /* Field2D::new(_, h, ...) */
    float discretization; /* Field2D::new(_, _, d, _) */
    int32_t toroidal;     /* Field2D::new(_, _, _, t) */
    uint32_t capacity;    /* maximum agents per buffer */
    uint32_t bin_rows;    /* device bin height in cells; 0 = automatic */
    uint32_t tile_cols;   /* device tile width in cells (power of two); 0 = automatic */
    uint32_t tile_rows;   /* device tile height in cells (power of two); 0 = automatic */
} kg_field_desc;

/* Weights of the flockers step (tests/model/flockers/bird.rs:12-17) and the
 * neighbour radius (bird.rs:41).  `seed` keys the Philox4x32-10 stream that
 * replaces rand::rng(): agent `id` at schedule step `s` draws the first two
 * words of Philox(ctr = {id, s_lo, s_hi, 0x53544550}, key = seed) mapped to
 * [0,1) by (u >> 8) * 2^-24. */
typedef struct kg_flock_params {
    float cohesion;
    float avoidance;
    float consistency;
    float randomness;
    float momentum;
    float jump;
    float radius;
    uint32_t reserved; /* must be 0 */
    uint64_t seed;
} kg_flock_params;

/* Field2D::new (field_2d.rs:304-322). */
int kg_field_create(const kg_field_desc *desc, kg_field **out);
void kg_field_destroy(kg_field *f);

/* Field2D::set_object_location (field_2d.rs:838-846), in bulk: appends n
 * agents to the write buffer.  Returns KG_E_CAPACITY if the write buffer
 * would exceed the capacity, KG_E_RANGE if a location falls outside the bag
 * grid (nothing is appended then). */
int kg_set_object_locations(kg_field *f, const uint32_t *ids, const float *x, const float *y,
                            const float *hx, const float *hy, size_t n);

/* Flocker::init (tests/model/flockers/state.rs:41-56) on the device: appends
 * agents 0..n-1 with pos = (w, h) * U[0,1)^2 drawn from Philox(ctr = {id, k,
 * 0, 0x494E4954}, key = seed) and heading (0,0) (heading_mode 0) or a unit
 * vector at a uniform angle (heading_mode 1). */
int kg_init_uniform(kg_field *f, uint32_t n, uint64_t seed, int heading_mode);

/* Field2D::remove_object_location (field_2d.rs:885-898): removes from the
 * write buffer every agent with this id in the bag of (x, y). */
int kg_remove_object_location(kg_field *f, uint32_t id, float x, float y);

/* Field::lazy_update (field_2d.rs:905-921): the write buffer becomes the read
 * buffer (binned on the device) and the write buffer is emptied. */
int kg_lazy_update(kg_field *f);

/* Field2D::get_neighbors_within_distance (field_2d.rs:386-440) and
 * get_neighbors_within_relax_distance (field_2d.rs:472-516) on the read
 * buffer.  Writes up to cap ids; *n_out receives the full count and
 * KG_E_CAPACITY is returned if it exceeds cap. */
int kg_get_neighbors_within_distance(kg_field *f, float x, float y, float dist, uint32_t *ids,
                                     size_t cap, size_t *n_out);
int kg_get_neighbors_within_relax_distance(kg_field *f, float x, float y, float dist,
                                           uint32_t *ids, size_t cap, size_t *n_out);

/* Field2D::get_objects / get_objects_unbuffered (field_2d.rs:544-582): the
 * agents in the bag of (x, y) in the read (unbuffered = 0) or write
 * (unbuffered = 1) buffer.  Same capacity contract as above. */
int kg_get_objects(kg_field *f, float x, float y, int unbuffered, uint32_t *ids, size_t cap,
                   size_t *n_out);

/* Field2D::num_objects_at_location (field_2d.rs:803-808). */
int kg_num_objects_at_location(kg_field *f, float x, float y, size_t *n);

/* Field2D::get_empty_bags (field_2d.rs:727-739), count only. */
int kg_num_empty_bags(kg_field *f, size_t *n);

/* Field2D::get_empty_bags (field_2d.rs:727-762): the position
 * not_discretize(bag) = (cx * d, cy * d) (field_2d.rs:345-353) of every bag
 * of the read buffer that holds no agent, in bag index order (cx ascending,
 * then cy).  Writes up to cap positions to x, y; *n_out receives the full
 * count and KG_E_CAPACITY is returned if it exceeds cap. */
int kg_get_empty_bags(kg_field *f, float *x, float *y, size_t cap, size_t *n_out);

/* Field2D::get_random_empty_bag (field_2d.rs:764-800) with a seeded draw in
 * place of rand::rng(): the empty bag number floor(u * count / 2^64) in
 * kg_get_empty_bags order, u = (w1 << 32) | w0 from the words w0, w1 of
 * Philox4x32-10(ctr = {0, 0, 0, 0x42414753}, key = seed).  *found is 0 and
 * x, y are untouched when no bag is empty (the reference returns None). */
int kg_get_random_empty_bag(kg_field *f, uint64_t seed, float *x, float *y, int *found);

/* Field2D::iter_objects / iter_objects_unbuffered (field_2d.rs:619-726):
 * calls fn once per agent of the read (unbuffered = 0) or write
 * (unbuffered = 1) buffer, bag by bag in bag index order, with the bag's
 * not_discretize position and the agent.  Inside a bag the agents come in
 * ascending id (the read buffer and a flock step's output) or in insertion
 * order (set_object_location entries not yet swapped in). */
typedef void (*kg_object_fn)(void *user, float bag_x, float bag_y, uint32_t id, float x, float y,
                             float hx, float hy);
int kg_iter_objects(kg_field *f, int unbuffered, kg_object_fn fn, void *user);

/* Field2D::nagents (field_2d.rs:277): set_object_location calls counted
 * until the first lazy_update. */
int kg_nagents(kg_field *f, size_t *n);

/* Number of agents in the read (buffer 0) or write (buffer 1) buffer. */
int kg_count(kg_field *f, int buffer, size_t *n);

/* get_neighbors_within_distance for every agent of the read buffer at once,
 * as CSR over the read buffer's storage order (see kg_read_agents):
 * row_ptr has count+1 entries, ids receives the neighbour ids of each row in
 * bag-scan order.  *total receives the number of ids; KG_E_CAPACITY if it
 * exceeds cap (row_ptr is still filled). */
int kg_neighbor_sets(kg_field *f, float dist, uint32_t *row_ptr, uint32_t *ids, size_t cap,
                     size_t *total);

/* Bird::step (bird.rs:38-154) for every agent of the read buffer, i.e. the
 * agent loop of Schedule::step (src/engine/schedule.rs:395-408) at schedule
 * step `step`: reads the read buffer, writes the write buffer.  The write
 * buffer must be empty. */
int kg_flock_step(kg_field *f, const kg_flock_params *p, uint64_t step);

/* The simulate! loop (src/lib.rs:1163-1173) for a flockers state:
 * `nsteps` calls of Schedule::step starting at schedule step `first_step`,
 * each = [lazy_update if the step is 0] + kg_flock_step + lazy_update
 * (schedule.rs:349-351, 395-412).  Asynchronous: call kg_field_sync to wait.
 *
 * Two device paths give identical results: the three-kernel path (flock +
 * scan + scatter and bin sort) and, for toroidal fields whose window
 * (floor(radius/d) cells) fits one tile of halo, a tile-resident layout that
 * runs each whole step as ONE kernel launch.  By default fields of fewer
 * than 32,768 agents take the tile-resident step and larger ones the
 * three-kernel path (the faster of the two at each size on MI355X).
 * `flags`: KG_RUN_GRAPH replays launches from HIP graphs; KG_RUN_BINNED
 * forces the three-kernel path; KG_RUN_TILE_RESIDENT the tile-resident step
 * where the field suits it (the two exclude each other). */
#define KG_RUN_GRAPH 1
#define KG_RUN_BINNED 2
#define KG_RUN_TILE_RESIDENT 4
int kg_schedule_run(kg_field *f, const kg_flock_params *p, uint64_t first_step, uint32_t nsteps,
                    int flags);

/* One Schedule::step of a flockers field whose agents live in caller-owned
 * device memory, allocating nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * ids, x, y, hx, hy are device pointers to n agents.  They go into a field
 * described by `desc` (the geometry kg_field_create makes; n <= capacity)
 * by set_object_location, the field is lazily updated, every agent takes
 * Bird::step at schedule step `step`, and the field is lazily updated again.
 * out_ids .. out_hy (device pointers to n entries each, not overlapping the
 * inputs) receive the new read buffer in storage order: the agents
 * kg_read_agents(f, 0) returns after kg_set_object_locations,
 * kg_lazy_update, kg_flock_step(f, p, step) and kg_lazy_update on a
 * kg_field of the same description.  `stream` is a hipStream_t (NULL: the
 * null stream); the call synchronizes it.  KG_E_RANGE if an input location
 * falls outside the bag grid.
 *
 * Every buffer beyond the inputs and outputs is carved from one static
 * device array of 64 MiB per GPU, held until the process ends.  A field that
 * needs more (40 B per agent plus 16 B per bin: about 1.3 M agents on a
 * sparse field) is rejected with KG_E_CAPACITY before any work.  Calls on
 * one GPU share that array: calls that overlap on one GPU race. */
int kg_schedule_step_static(const kg_field_desc *desc, const kg_flock_params *p, uint64_t step,
                            const uint32_t *ids, const float *x, const float *y, const float *hx,
                            const float *hy, uint32_t n, uint32_t *out_ids, float *out_x,
                            float *out_y, float *out_hx, float *out_hy, void *stream);

/* Copies a buffer (0 = read, 1 = write) to the host in storage order.  Any
 * output pointer may be NULL.  *n receives the agent count; KG_E_CAPACITY
 * if it exceeds cap. */
int kg_read_agents(kg_field *f, int buffer, uint32_t *ids, float *x, float *y, float *hx,
                   float *hy, size_t cap, size_t *n);

/* Waits for all work queued on the field's stream. */
int kg_field_sync(kg_field *f);

/* The field's hipStream_t, for callers that time or order work around it. */
void *kg_field_stream(kg_field *f);

/* Instrumentation (no counterpart in the reference): when enabled, every
 * kernel launch of the field is bracketed by HIP events on the field's
 * stream and its duration accumulated per kernel; graph replay is disabled
 * meanwhile.  kg_field_profile(f, on) also resets the totals. */
#define KG_KERNEL_FLOCK 0          /* fused query + Bird::step */
#define KG_KERNEL_SCAN 1           /* bin offsets */
#define KG_KERNEL_SCATTER 2        /* rebinning of the flock output */
#define KG_KERNEL_COUNT_APPEND 3   /* histogram of set_object_location agents */
#define KG_KERNEL_SCATTER_APPEND 4 /* binning of set_object_location agents */
#define KG_KERNEL_RT_STEP 5        /* tile-resident schedule step */
#define KG_KERNEL_RT_LOAD 6        /* entry into the tile-resident layout */
#define KG_KERNEL_SORT_BINS 7      /* (cell y, id) order inside every bin */
#define KG_KERNEL_COUNT 8
int kg_field_profile(kg_field *f, int enable);
int kg_field_profile_read(kg_field *f, int kernel, double *total_ms, uint64_t *launches);

/* Layout statistics (no counterpart in the reference), 16 values:
 * [0] tile-resident layout active, [1] last plan did not fit the LDS
 * budget, [2] plans made, [3] OR of the tile-resident flags seen
 * (1 spill, 2 global path, 4 out of grid, 8 stay overflow), [4] tiles,
 * [5] tile columns, [6] tile rows, [7] bin rows, [8] LDS staged-agent
 * capacity, [9] LDS own-agent capacity, [10] LDS extras capacity, [11] LDS
 * bytes per work-group ([8]-[11] of the normal tile class), [12] LDS limit
 * per work-group, [13] largest tile capacity, [14] tiles in the dense class,
 * [15] schedule steps run tile-resident. */
int kg_field_stats(kg_field *f, uint64_t out[16]);

/* ---- ensembles of runs (explore_parallel!, src/explore/model_exploration.rs:
 * 354-432; the serving loop) --------------------------------------------
 *
 * A kg_ensemble advances many independent flockers runs -- each its own
 * toroidal Field2D, parameters, Philox key and schedule step -- in one
 * device layout, the same launches per schedule step for all of them: a
 * tile-resident layout (one launch per tile class), or the bin-grouped read
 * buffer stepped by the three-kernel path over every run at once.  A run's
 * results are bit-identical to the same run on its own kg_field
 * (kg_schedule_run), whatever else shares the ensemble.  Runs can join and
 * leave between kg_ensemble_run calls; the layout is rebuilt then. */
typedef struct kg_ensemble kg_ensemble;

/* One run: Field2D::new(width, height, discretization, toroidal = 1) with
 * `agents` birds placed by Flocker::init (state.rs:41-56) from
 * Philox(init_seed) (see kg_init_uniform), stepped with `params`. */
typedef struct kg_run_desc {
    float width;
    float height;
    float discretization;
    int32_t toroidal;     /* must be 1 */
    uint32_t agents;      /* >= 1 */
    int32_t heading_mode; /* as kg_init_uniform */
    uint64_t init_seed;
    kg_flock_params params;
} kg_run_desc;

/* Layout choices an autotuner may set; 0 = automatic. */
typedef struct kg_tuning {
    uint32_t tile_cols;  /* tile width in cells (power of two) */
    uint32_t tile_rows;  /* tile height in cells (power of two) */
    uint32_t bin_rows;   /* bin height in cells (power of two) */
    uint32_t lds_normal; /* LDS bytes per class-0 work-group; 0: what keeps a CU full */
} kg_tuning;

/* Device bytes a run of this description occupies in an ensemble. */
int kg_run_footprint(const kg_run_desc *run, const kg_tuning *tuning, uint64_t *bytes);

int kg_ensemble_create(const kg_tuning *tuning, kg_ensemble **out);
void kg_ensemble_destroy(kg_ensemble *e);

/* Adds runs (their schedule step 0 comes with the next kg_ensemble_run);
 * ids[i] receives run i's id.  KG_E_INVALID if a run cannot use the
 * tile-resident step (not toroidal, window wider than a tile). */
int kg_ensemble_admit(kg_ensemble *e, const kg_run_desc *runs, uint32_t n, uint32_t *ids);

/* Removes a run (its state is gone; read it first). */
int kg_ensemble_retire(kg_ensemble *e, uint32_t id);

/* Every run advances `nsteps` schedule steps (simulate!'s loop,
 * src/lib.rs:1163-1173).  flags: KG_RUN_GRAPH; KG_RUN_BINNED or
 * KG_RUN_TILE_RESIDENT choose the layout's form (the layout is rebuilt when
 * the form changes); without either, ensembles whose live runs average
 * 32,768 agents or more take the bin-grouped form.  Asynchronous. */
int kg_ensemble_run(kg_ensemble *e, uint32_t nsteps, int flags);

/* Schedule steps run so far by run `id`. */
int kg_ensemble_steps(kg_ensemble *e, uint32_t id, uint64_t *steps);

/* Copies run `id`'s birds to the host (any order); same contract as
 * kg_read_agents.  Synchronises. */
int kg_ensemble_read(kg_ensemble *e, uint32_t id, uint32_t *ids, float *x, float *y, float *hx,
                     float *hy, size_t cap, size_t *n);

int kg_ensemble_sync(kg_ensemble *e);
void *kg_ensemble_stream(kg_ensemble *e);

/* Statistics, 8 values: [0] runs, [1] agents, [2] tiles, [3] dense tiles,
 * [4] layouts built, [5] OR of the tile-resident flags seen, [6] schedule
 * steps run, [7] device bytes held. */
int kg_ensemble_stats(kg_ensemble *e, uint64_t out[8]);

/* Message of the last failing call on this thread ("" if none). */
const char *kg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
