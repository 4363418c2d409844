/*
 * oracle_par.c -- run-level parallel CPU path (TEST INFRASTRUCTURE ONLY).
 *
 * Restates explore_parallel! (src/explore/model_exploration.rs:387-421): a
 * pool of worker threads, each taking whole independent runs (State::new +
 * simulate_explore!, model_exploration.rs:159-190) from a shared counter.
 * Used as the CPU baseline for the ensemble and serving workloads.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <time.h>

typedef struct {
    uint32_t n;
    float side;
    float discretization;
    uint32_t steps;
    or_flock_params p;
    int heading_mode;
} or_run_spec;

typedef struct {
    const or_run_spec *specs;
    size_t nruns;
    atomic_size_t next;
    double *seconds;
    float *final_xyhh; /* optional: 4 floats per agent, concatenated */
    const size_t *final_off;
} pool_t;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void run_one(const or_run_spec *s, double *secs, float *out)
{
    uint32_t n = s->n;
    float *x = (float *)malloc(4 * (size_t)(n ? n : 1) * sizeof(float));
    float *y = x + n, *hx = y + n, *hy = hx + n;
    or_init_uniform(n, s->side, s->side, s->p.seed, s->heading_mode, x, y, hx, hy);
    double t0 = now_s();
    or_flocker *fl = or_flocker_new(n, s->side, s->side, s->discretization, 1, x, y, hx, hy, &s->p);
    for (uint32_t k = 0; k < s->steps; ++k) or_flocker_step(fl);
    *secs = now_s() - t0;
    if (out) {
        or_flocker_get(fl, x, y, hx, hy);
        for (uint32_t i = 0; i < n; ++i) {
            out[4 * (size_t)i + 0] = x[i];
            out[4 * (size_t)i + 1] = y[i];
            out[4 * (size_t)i + 2] = hx[i];
            out[4 * (size_t)i + 3] = hy[i];
        }
    }
    or_flocker_free(fl);
    free(x);
}

static void *worker(void *arg)
{
    pool_t *pl = (pool_t *)arg;
    for (;;) {
        size_t r = atomic_fetch_add(&pl->next, 1);
        if (r >= pl->nruns) break;
        float *out = pl->final_xyhh ? pl->final_xyhh + 4 * pl->final_off[r] : NULL;
        run_one(&pl->specs[r], &pl->seconds[r], out);
    }
    return NULL;
}

/* Runs `nruns` independent flocker runs on `nthreads` threads.  seconds[r]
 * receives each run's wall time; final_xyhh (optional, with offsets
 * final_off[r] in agents) receives each run's final x, y, hx, hy
 * interleaved.  Returns the wall time of the whole ensemble in seconds. */
double or_run_ensemble(const or_run_spec *specs, size_t nruns, int nthreads, double *seconds,
                       float *final_xyhh, const size_t *final_off)
{
    pool_t pl;
    pl.specs = specs;
    pl.nruns = nruns;
    atomic_init(&pl.next, 0);
    pl.seconds = seconds;
    pl.final_xyhh = final_xyhh;
    pl.final_off = final_off;
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    double t0 = now_s();
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &pl);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    double wall = now_s() - t0;
    free(th);
    return wall;
}

/* Times `steps` schedule steps of one run already initialised from arrays
 * (the plain CPU path).  Returns seconds; final state is written back. */
double or_time_steps(uint32_t n, float w, float h, float d, int toroidal, float *x, float *y,
                     float *hx, float *hy, const or_flock_params *p, uint32_t steps)
{
    or_flocker *fl = or_flocker_new(n, w, h, d, toroidal, x, y, hx, hy, p);
    double t0 = now_s();
    for (uint32_t k = 0; k < steps; ++k) or_flocker_step(fl);
    double secs = now_s() - t0;
    or_flocker_get(fl, x, y, hx, hy);
    or_flocker_free(fl);
    return secs;
}
