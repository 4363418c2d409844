/*
 * oracle.c -- CPU restatement of krABMaga's flockers agent-step path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the
 * single-threaded CPU baseline.  Never linked into the product library.
 * Build with -ffp-contract=off so every f32 operation rounds exactly once,
 * as Rust's f32 arithmetic does.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================= */
/* Philox4x32-10, from Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11 (round function, multipliers and Weyl key
 * increments as published there). */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float or_u01(uint32_t u)
{
    return (float)(u >> 8) * (1.0f / 16777216.0f);
}

#define TAG_STEP 0x53544550u /* "STEP" */
#define TAG_INIT 0x494E4954u /* "INIT" */

void or_step_uniforms(uint64_t seed, uint32_t id, uint64_t step, float *r1, float *r2)
{
    uint32_t ctr[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), TAG_STEP};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    *r1 = or_u01(o[0]);
    *r2 = or_u01(o[1]);
}

void or_init_uniform(uint32_t n, float w, float h, uint64_t seed, int heading_mode,
                     float *x, float *y, float *hx, float *hy)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t id = 0; id < n; ++id) {
        uint32_t ctr[4] = {id, 0u, 0u, TAG_INIT};
        uint32_t o[4];
        or_philox4x32_10(ctr, key, o);
        /* state.rs:42-46: pos = dim * r */
        x[id] = w * or_u01(o[0]);
        y[id] = h * or_u01(o[1]);
        if (heading_mode == 0) {
            hx[id] = 0.0f; /* state.rs:47 */
            hy[id] = 0.0f;
            continue;
        }
        float u = 2.0f * or_u01(o[2]) - 1.0f;
        float v = 2.0f * or_u01(o[3]) - 1.0f;
        float s = u * u + v * v;
        uint32_t k = 0;
        while (!(s > 0.0f && s <= 1.0f)) {
            ctr[1] = ++k;
            or_philox4x32_10(ctr, key, o);
            u = 2.0f * or_u01(o[2]) - 1.0f;
            v = 2.0f * or_u01(o[3]) - 1.0f;
            s = u * u + v * v;
        }
        float r = sqrtf(s);
        hx[id] = u / r;
        hy[id] = v / r;
    }
}

/* ======================================================================= */
/* field_2d.rs free helpers */

int32_t or_f32_to_i32(float f)
{
    /* Rust `f as i32` (saturating since Rust 1.45; NaN -> 0). */
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

/* field_2d.rs:926-932 (Rust `%` on i32 truncates like C). */
int32_t or_t_transform(int32_t n, int32_t size)
{
    if (n >= 0) return n % size;
    return (n % size) + size;
}

/* field_2d.rs:1004-1014 (Rust `%` on f32 is fmod). */
float or_toroidal_transform(float val, float dim)
{
    if (val >= 0.0f && val < dim) return val;
    float v = fmodf(val, dim);
    if (v < 0.0f) v += dim;
    return v;
}

/* field_2d.rs:988-1002 */
float or_toroidal_distance(float val1, float val2, float dim)
{
    if (fabsf(val1 - val2) <= dim / 2.0f) return val1 - val2;
    float d = or_toroidal_transform(val1, dim) - or_toroidal_transform(val2, dim);
    if (d * 2.0f > dim) return d - dim;
    if (d * 2.0f < -dim) return d + dim;
    return d;
}

/* field_2d.rs:974-986 */
float or_distance(float x1, float y1, float x2, float y2, float dim1, float dim2, int tor)
{
    float dx, dy;
    if (tor) {
        dx = or_toroidal_distance(x1, x2, dim1);
        dy = or_toroidal_distance(y1, y2, dim2);
    } else {
        dx = x1 - x2;
        dy = y1 - y2;
    }
    return sqrtf(dx * dx + dy * dy);
}

/* field_2d.rs:934-972 */
int or_check_circle(int32_t bx, int32_t by, float discretization, float width, float height,
                    float lx, float ly, float dis, int tor)
{
    float nw_x = (float)bx * discretization;
    float nw_y = (float)by * discretization;
    float ne_x = nw_x;
    float ne_y = fminf(nw_y + discretization, height);
    float sw_x = fminf(nw_x + discretization, width);
    float sw_y = nw_y;
    float se_x = sw_x;
    float se_y = ne_y;
    float d_nw = or_distance(nw_x, nw_y, lx, ly, width, height, tor);
    float d_ne = or_distance(ne_x, ne_y, lx, ly, width, height, tor);
    float d_sw = or_distance(sw_x, sw_y, lx, ly, width, height, tor);
    float d_se = or_distance(se_x, se_y, lx, ly, width, height, tor);
    if (d_nw <= dis && d_ne <= dis && d_sw <= dis && d_se <= dis) return 1;
    if (d_nw > dis && d_ne > dis && d_sw > dis && d_se > dis) return -1;
    return 0;
}

/* ======================================================================= */
/* Field2D (field_2d.rs:267-923, sequential variant) */

static void bag_push(or_bag *b, or_bird v)
{
    if (b->n == b->cap) {
        size_t nc = b->cap ? b->cap * 2 : 4;
        b->v = (or_bird *)realloc(b->v, nc * sizeof(or_bird));
        b->cap = nc;
    }
    b->v[b->n++] = v;
}

static or_bag *bags_new(size_t n, size_t cap)
{
    or_bag *b = (or_bag *)calloc(n ? n : 1, sizeof(or_bag));
    if (cap) {
        for (size_t i = 0; i < n; ++i) {
            b[i].v = (or_bird *)malloc(cap * sizeof(or_bird));
            b[i].cap = cap;
        }
    }
    return b;
}

static void bags_free(or_bag *b, size_t n)
{
    if (!b) return;
    for (size_t i = 0; i < n; ++i) free(b[i].v);
    free(b);
}

/* field_2d.rs:304-322 */
or_field2d *or_field_new(float w, float h, float d, int toroidal)
{
    or_field2d *f = (or_field2d *)calloc(1, sizeof(or_field2d));
    float nbags_f = (ceilf(w / d) + 1.0f) * (ceilf(h / d) + 1.0f);
    int32_t nb = or_f32_to_i32(nbags_f);
    f->nbags = nb > 0 ? (size_t)nb : 0; /* `as usize` saturates at 0 */
    f->bags[0] = bags_new(f->nbags, 0);
    f->bags[1] = bags_new(f->nbags, 0);
    f->read = 0;
    f->write = 1;
    f->nagents = 0;
    f->width = w;
    f->height = h;
    f->discretization = d;
    f->toroidal = toroidal;
    f->dh = or_f32_to_i32(ceilf(h / d)) + 1;
    f->dw = or_f32_to_i32(ceilf(w / d)) + 1;
    f->density_estimation = 0;
    f->density_estimation_check = 0;
    return f;
}

void or_field_free(or_field2d *f)
{
    if (!f) return;
    bags_free(f->bags[0], f->nbags);
    bags_free(f->bags[1], f->nbags);
    free(f);
}

/* field_2d.rs:328-339 */
void or_discretize(const or_field2d *f, float x, float y, int32_t *cx, int32_t *cy)
{
    *cx = or_f32_to_i32(floorf(x / f->discretization));
    *cy = or_f32_to_i32(floorf(y / f->discretization));
}

static int64_t bag_index(const or_field2d *f, float x, float y)
{
    int32_t cx, cy;
    or_discretize(f, x, y, &cx, &cy);
    /* field_2d.rs:840: ((bag.x * self.dh) + bag.y) as usize */
    int64_t idx = (int64_t)cx * f->dh + cy;
    if (idx < 0 || (uint64_t)idx >= f->nbags) return -1;
    return idx;
}

/* field_2d.rs:838-846 */
int or_set_object_location(or_field2d *f, or_bird b, float x, float y)
{
    int64_t idx = bag_index(f, x, y);
    if (idx < 0) return -1;
    bag_push(&f->bags[f->write][idx], b);
    if (!f->density_estimation_check) f->nagents += 1;
    return 0;
}

/* field_2d.rs:885-898 */
void or_remove_object_location(or_field2d *f, uint32_t id, float x, float y)
{
    int64_t idx = bag_index(f, x, y);
    if (idx < 0) return;
    or_bag *b = &f->bags[f->write][idx];
    if (b->n == 0) return;
    size_t before = b->n, k = 0;
    for (size_t i = 0; i < b->n; ++i)
        if (b->v[i].id != id) b->v[k++] = b->v[i];
    b->n = k;
    if (!f->density_estimation_check) f->nagents -= before - k;
}

/* field_2d.rs:905-921 */
void or_lazy_update(or_field2d *f)
{
    int t = f->read;
    f->read = f->write;
    f->write = t;
    if (!f->density_estimation_check) {
        f->density_estimation = f->nagents / (size_t)((int64_t)f->dw * f->dh);
        f->density_estimation_check = 1;
        bags_free(f->bags[f->write], f->nbags);
        f->bags[f->write] = bags_new(f->nbags, f->density_estimation);
    } else {
        or_bag *b = f->bags[f->write];
        for (size_t i = 0; i < f->nbags; ++i) b[i].n = 0;
    }
}

static size_t emit(or_bird *out, size_t cap, size_t n, const or_bird *b)
{
    if (n < cap) out[n] = *b;
    return n + 1;
}

/* field_2d.rs:386-440 */
size_t or_get_neighbors_within_distance(const or_field2d *f, float x, float y, float dist,
                                        or_bird *out, size_t cap)
{
    size_t n = 0;
    if (dist <= 0.0f) return 0;
    int32_t disc_dist = or_f32_to_i32(floorf(dist / f->discretization));
    int32_t cx, cy;
    or_discretize(f, x, y, &cx, &cy);
    int32_t max_x = or_f32_to_i32(ceilf(f->width / f->discretization));
    int32_t max_y = or_f32_to_i32(ceilf(f->height / f->discretization));
    int32_t min_i = cx - disc_dist, max_i = cx + disc_dist;
    int32_t min_j = cy - disc_dist, max_j = cy + disc_dist;
    if (f->toroidal) {
        if (min_i < 0) min_i = 0;
        if (max_i > max_x - 1) max_i = max_x - 1;
        if (min_j < 0) min_j = 0;
        if (max_j > max_y - 1) max_j = max_y - 1;
    }
    const or_bag *bags = f->bags[f->read];
    for (int32_t i = min_i; i < max_i + 1; ++i) {
        for (int32_t j = min_j; j < max_j + 1; ++j) {
            int32_t bx = or_t_transform(i, max_x);
            int32_t by = or_t_transform(j, max_y);
            int check = or_check_circle(bx, by, f->discretization, f->width, f->height, x, y, dist,
                                        f->toroidal);
            size_t index = (size_t)((int64_t)bx * f->dh + by);
            const or_bag *b = &bags[index];
            for (size_t e = 0; e < b->n; ++e) {
                const or_bird *el = &b->v[e];
                if ((check == 0 && or_distance(x, y, el->px, el->py, f->width, f->height,
                                               f->toroidal) <= dist) ||
                    check == 1)
                    n = emit(out, cap, n, el);
            }
        }
    }
    return n;
}

int64_t or_set_object_locations(or_field2d *f, const uint32_t *ids, const float *x, const float *y,
                                const float *hx, const float *hy, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        or_bird b = {ids[i], x[i], y[i], hx ? hx[i] : 0.0f, hy ? hy[i] : 0.0f};
        if (or_set_object_location(f, b, x[i], y[i]) != 0) return (int64_t)i;
    }
    return -1;
}

size_t or_neighbor_sets(const or_field2d *f, const float *qx, const float *qy, size_t n, float dist,
                        uint64_t *row_ptr, uint32_t *ids, size_t cap)
{
    size_t vcap = 256, total = 0;
    or_bird *v = (or_bird *)malloc(vcap * sizeof(or_bird));
    for (size_t k = 0; k < n; ++k) {
        size_t m = or_get_neighbors_within_distance(f, qx[k], qy[k], dist, v, vcap);
        if (m > vcap) {
            vcap = m;
            v = (or_bird *)realloc(v, vcap * sizeof(or_bird));
            or_get_neighbors_within_distance(f, qx[k], qy[k], dist, v, vcap);
        }
        row_ptr[k] = total;
        for (size_t e = 0; e < m; ++e, ++total)
            if (total < cap) ids[total] = v[e].id;
    }
    row_ptr[n] = total;
    free(v);
    return total;
}

/* field_2d.rs:472-516 */
size_t or_get_neighbors_within_relax_distance(const or_field2d *f, float x, float y, float dist,
                                              or_bird *out, size_t cap)
{
    size_t n = 0;
    if (dist <= 0.0f) return 0;
    int32_t disc_dist = or_f32_to_i32(floorf(dist / f->discretization));
    int32_t cx, cy;
    or_discretize(f, x, y, &cx, &cy);
    int32_t max_x = or_f32_to_i32(ceilf(f->width / f->discretization));
    int32_t max_y = or_f32_to_i32(ceilf(f->height / f->discretization));
    int32_t min_i = cx - disc_dist, max_i = cx + disc_dist;
    int32_t min_j = cy - disc_dist, max_j = cy + disc_dist;
    if (f->toroidal) {
        if (min_i < 0) min_i = 0;
        if (max_i > max_x - 1) max_i = max_x - 1;
        if (min_j < 0) min_j = 0;
        if (max_j > max_y - 1) max_j = max_y - 1;
    }
    const or_bag *bags = f->bags[f->read];
    for (int32_t i = min_i; i < max_i + 1; ++i) {
        for (int32_t j = min_j; j < max_j + 1; ++j) {
            int32_t bx = or_t_transform(i, max_x);
            int32_t by = or_t_transform(j, max_y);
            size_t index = (size_t)((int64_t)bx * f->dh + by);
            const or_bag *b = &bags[index];
            for (size_t e = 0; e < b->n; ++e) n = emit(out, cap, n, &b->v[e]);
        }
    }
    return n;
}

static size_t copy_bag(const or_bag *b, or_bird *out, size_t cap)
{
    for (size_t e = 0; e < b->n && e < cap; ++e) out[e] = b->v[e];
    return b->n;
}

/* field_2d.rs:544-549 */
size_t or_get_objects(const or_field2d *f, float x, float y, or_bird *out, size_t cap)
{
    int64_t idx = bag_index(f, x, y);
    if (idx < 0) return 0;
    return copy_bag(&f->bags[f->read][idx], out, cap);
}

/* field_2d.rs:577-582 */
size_t or_get_objects_unbuffered(const or_field2d *f, float x, float y, or_bird *out, size_t cap)
{
    int64_t idx = bag_index(f, x, y);
    if (idx < 0) return 0;
    return copy_bag(&f->bags[f->write][idx], out, cap);
}

/* field_2d.rs:803-808 */
size_t or_num_objects_at_location(const or_field2d *f, float x, float y)
{
    int64_t idx = bag_index(f, x, y);
    if (idx < 0) return 0;
    return f->bags[f->read][idx].n;
}

/* field_2d.rs:727-739 (count only) */
size_t or_num_empty_bags(const or_field2d *f)
{
    size_t n = 0;
    for (int32_t i = 0; i < f->dw; ++i)
        for (int32_t j = 0; j < f->dh; ++j)
            if (f->bags[f->read][(size_t)((int64_t)i * f->dh + j)].n == 0) ++n;
    return n;
}

/* field_2d.rs:727-762 with not_discretize (:345-353): writes up to cap
 * positions, returns the count */
size_t or_get_empty_bags(const or_field2d *f, float *x, float *y, size_t cap)
{
    size_t n = 0;
    for (int32_t i = 0; i < f->dw; ++i)
        for (int32_t j = 0; j < f->dh; ++j)
            if (f->bags[f->read][(size_t)((int64_t)i * f->dh + j)].n == 0) {
                if (n < cap) {
                    x[n] = (float)i * f->discretization;
                    y[n] = (float)j * f->discretization;
                }
                ++n;
            }
    return n;
}

size_t or_nagents(const or_field2d *f) { return f->nagents; }

/* ======================================================================= */
/* Flockers (tests/model/flockers) */

or_flocker *or_flocker_new(uint32_t n, float w, float h, float d, int toroidal,
                           const float *x, const float *y, const float *hx, const float *hy,
                           const or_flock_params *p)
{
    or_flocker *fl = (or_flocker *)calloc(1, sizeof(or_flocker));
    fl->field = or_field_new(w, h, d, toroidal);
    fl->agents = (or_bird *)malloc((n ? n : 1) * sizeof(or_bird));
    fl->order = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    fl->n = n;
    fl->width = w;
    fl->height = h;
    fl->step = 0;
    fl->p = *p;
    for (uint32_t i = 0; i < n; ++i) {
        or_bird b = {i, x[i], y[i], hx[i], hy[i]};
        fl->agents[i] = b;
        fl->order[i] = i;
        /* state.rs:50-51: field1.set_object_location(bird, pos) */
        or_set_object_location(fl->field, b, b.px, b.py);
    }
    return fl;
}

void or_flocker_free(or_flocker *fl)
{
    if (!fl) return;
    or_field_free(fl->field);
    free(fl->agents);
    free(fl->order);
    free(fl->scratch);
    free(fl);
}

void or_flocker_set_order(or_flocker *fl, const uint32_t *order)
{
    memcpy(fl->order, order, fl->n * sizeof(uint32_t));
}

/* bird.rs:38-154 */
void or_bird_step(or_flocker *fl, uint32_t id)
{
    or_bird *self = &fl->agents[id];
    const or_field2d *f = fl->field;
    const or_flock_params *p = &fl->p;
    /* bird.rs:41, with the per-call Vec of field_2d.rs:387-391 */
    size_t cap = f->density_estimation_check ? f->density_estimation * 2 : 0;
    if (cap < 16) cap = 16;
    or_bird *vec = (or_bird *)malloc(cap * sizeof(or_bird));
    size_t nv = or_get_neighbors_within_distance(f, self->px, self->py, p->radius, vec, cap);
    if (nv > cap) {
        vec = (or_bird *)realloc(vec, nv * sizeof(or_bird));
        or_get_neighbors_within_distance(f, self->px, self->py, p->radius, vec, nv);
    }

    float width = fl->width, height = fl->height;
    float avoid_x = 0.0f, avoid_y = 0.0f;
    float cohe_x = 0.0f, cohe_y = 0.0f;
    float rand_x = 0.0f, rand_y = 0.0f;
    float cons_x = 0.0f, cons_y = 0.0f;

    if (nv != 0) {
        float x_avoid = 0.0f, y_avoid = 0.0f;
        float x_cohe = 0.0f, y_cohe = 0.0f;
        float x_cons = 0.0f, y_cons = 0.0f;
        int32_t count = 0;
        for (size_t e = 0; e < nv; ++e) {
            const or_bird *el = &vec[e];
            if (self->id != el->id) {
                float dx = or_toroidal_distance(self->px, el->px, width);
                float dy = or_toroidal_distance(self->py, el->py, height);
                count += 1;
                float square = dx * dx + dy * dy;
                x_avoid += dx / (square * square + 1.0f);
                y_avoid += dy / (square * square + 1.0f);
                x_cohe += dx;
                y_cohe += dy;
                x_cons += el->hx;
                y_cons += el->hy;
            }
        }
        if (count > 0) {
            float c = (float)count;
            x_avoid /= c;
            y_avoid /= c;
            x_cohe /= c;
            y_cohe /= c;
            x_cons /= c;
            y_cons /= c;
            cons_x = x_cons / c;
            cons_y = y_cons / c;
        } else {
            cons_x = x_cons;
            cons_y = y_cons;
        }
        avoid_x = 400.0f * x_avoid;
        avoid_y = 400.0f * y_avoid;
        cohe_x = -x_cohe / 10.0f;
        cohe_y = -y_cohe / 10.0f;

        float r1, r2;
        or_step_uniforms(p->seed, self->id, fl->step, &r1, &r2);
        float x_rand = r1 * 2.0f - 1.0f;
        float y_rand = r2 * 2.0f - 1.0f;
        float square = sqrtf(x_rand * x_rand + y_rand * y_rand);
        rand_x = 0.05f * x_rand / square;
        rand_y = 0.05f * y_rand / square;
    }
    free(vec);

    float mom_x = self->hx, mom_y = self->hy;
    float dx = p->cohesion * cohe_x + p->avoidance * avoid_x + p->consistency * cons_x +
               p->randomness * rand_x + p->momentum * mom_x;
    float dy = p->cohesion * cohe_y + p->avoidance * avoid_y + p->consistency * cons_y +
               p->randomness * rand_y + p->momentum * mom_y;
    float dis = sqrtf(dx * dx + dy * dy);
    if (dis > 0.0f) {
        dx = dx / dis * p->jump;
        dy = dy / dis * p->jump;
    }
    self->hx = dx;
    self->hy = dy;
    float loc_x = or_toroidal_transform(self->px + dx, width);
    float loc_y = or_toroidal_transform(self->py + dy, width); /* bird.rs:147 wraps y by width */
    self->px = loc_x;
    self->py = loc_y;
    or_set_object_location(fl->field, *self, loc_x, loc_y);
}

/* schedule.rs:347-413 with Flocker::update = lazy_update (state.rs:58-60) */
void or_flocker_step(or_flocker *fl)
{
    if (fl->step == 0) or_lazy_update(fl->field);
    for (uint32_t k = 0; k < fl->n; ++k) or_bird_step(fl, fl->order[k]);
    fl->step += 1;
    or_lazy_update(fl->field);
}

void or_flocker_get(const or_flocker *fl, float *x, float *y, float *hx, float *hy)
{
    for (uint32_t i = 0; i < fl->n; ++i) {
        x[i] = fl->agents[i].px;
        y[i] = fl->agents[i].py;
        hx[i] = fl->agents[i].hx;
        hy[i] = fl->agents[i].hy;
    }
}
