/*
 * CPU restatement of the reference's compact local-inertial step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the parity oracle and the CPU baseline. It is compiled
 * into oracle/_build/liboracle_f32.so (REAL = float, the reference's arithmetic) and
 * liboracle_f64.so (REAL = double, the FP64 comparison run), both exporting the entry points of
 * include/nhflood.h over HOST arrays. Only tests, the smoke check and the benchmark's CPU
 * baseline load it; the product library never links or calls it.
 *
 * It follows the reference pass by pass, in the reference's order and operation order:
 *   prepare_step        solver_compact.py:364-382, StepClock.choose clock.py:94-167
 *   update_pressure     solver_compact.py:346-349, pressure.py:66-116
 *   tick                solver_compact.py:422-613 (face averages 430-436, flux 439-486,
 *                       donor 491-514, scale 516-524, sources 525-566, depth 567-610)
 *   compute_cfl         solver_compact.py:617-641
 *   finish_step         solver_compact.py:384-391, StepClock.commit clock.py:173-181
 *   face_discharge      solver_compact.py:118-132
 *   boundary_neighbor_sum solver_compact.py:135-149
 * Deliberate difference: budget diagnostics are added per cell into error-free double-double
 * accumulators instead of F32 block sums (device_budget.py keeps exact integers of F32 block
 * sums whose F32 rounding depends on atomic order); the oracle therefore reports the exact sum
 * of the same per-cell F32 terms.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nhflood.h"

#ifndef NHF_REAL
#define NHF_REAL float
#endif
typedef NHF_REAL real;

/* x^y and e^x for the F32 step: evaluated in FP64 and rounded once to F32, by range reduction
 * plus series (log: m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) as an odd series;
 * 2^f: |f| <= 1/2, Taylor series of e^(f ln 2); see J.-M. Muller, Elementary Functions, ch. 5).
 * Every FP64 operation is an IEEE add, multiply, divide or explicit fma, so any IEEE machine
 * (this oracle, the GPU kernels) produces the same bits. */
static double det_log2(double x) {
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 13.0;  // |s| < 0.1716: the first omitted term is below 2^-37 of ln m
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    p = fma(p, s2, 1.0);
    double ln_m = 2.0 * s * p;
    return fma(ln_m, 1.44269504088896340736, (double)e);
}

static double det_exp2(double t) {
    if (t < -1100.0) return 0.0;
    if (t > 1100.0) return INFINITY;
    double n = rint(t);
    double u = (t - n) * 0.69314718055994530942;
    double p = 1.0 / 3628800.0;  // 1/10!; |u| <= ln(2)/2: the first omitted term is below 2^-36
    p = fma(p, u, 1.0 / 362880.0);
    p = fma(p, u, 1.0 / 40320.0);
    p = fma(p, u, 1.0 / 5040.0);
    p = fma(p, u, 1.0 / 720.0);
    p = fma(p, u, 1.0 / 120.0);
    p = fma(p, u, 1.0 / 24.0);
    p = fma(p, u, 1.0 / 6.0);
    p = fma(p, u, 0.5);
    p = fma(p, u, 1.0);
    p = fma(p, u, 1.0);
    return ldexp(p, (int)n);
}

static float det_powf(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x == 1.0f) return 1.0f;
    return (float)det_exp2((double)y * det_log2((double)x));
}

static float det_expf(float x) { return (float)det_exp2((double)x * 1.44269504088896340736); }

#if NHF_REAL_IS_DOUBLE
#define RPOW pow
#define REXP exp
#define RSQRT sqrt
#define RABS fabs
#else
#define RPOW det_powf
#define REXP det_expf
#define RSQRT sqrtf
#define RABS fabsf
#endif

#define TICKS_PER_SECOND 1099511627776.0 /* 2^40 */
#define OMP_MIN_ITEMS 32768                /* loops shorter than this run on one thread */
#define SNAP_TICKS 1099512ull            /* ceil(1e-6 * 2^40), clock.py:11 */

static __thread char g_error[256];

static int fail(const char *msg) {
    snprintf(g_error, sizeof g_error, "%s", msg);
    return -1;
}

static inline real rmax(real a, real b) { return a > b ? a : b; }
static inline real rmin(real a, real b) { return a < b ? a : b; }

typedef struct {
    double hi, lo;
} ddsum;

static inline void dd_add(ddsum *acc, double x) {
    /* Knuth two-sum: hi + err == acc->hi + x exactly */
    double s = acc->hi + x;
    double bp = s - acc->hi;
    double err = (acc->hi - (s - bp)) + (x - bp);
    acc->hi = s;
    acc->lo += err;
}

struct nhf_ctx {
    nhf_mesh m;
    nhf_params p;
    int terms;
    /* clock (core/clock.py) */
    uint64_t now, next, origin, step_end;
    real dt, local, end_local;
    int active;
    uint32_t trace[NHF_MAX_BATCH * NHF_TRACE_COLS];
    int slot;
    int pending_steps; /* steps of the batch run by nhf_launch_batch, until nhf_wait_batch */
    /* CFL and limiter state */
    real ndt;
    int32_t cfl_arg;
    uint32_t limited_cells, limited_faces;
    real min_scale;
    float limit_value[5];
    int32_t limit_ids[3];
    /* forcing */
    float *coefficients;
    real *values, *level_values;
    float *rain_rates;
    int raining;
    /* working arrays */
    real *depth, *bed, *rain_time, *capacity, *max_depth, *max_speed, *q;
    real *donor, *face_qn, *gradients, *cfl, *flow_depth;
    real *eq, *ssq, *increment, *rain_depth, *diag;
    ddsum *budget;
    /* gate control (core/gate.py), when bound */
    int gates_on;
    nhf_gates g;
    ddsum *gate_budget;  /* [10*count + 8] */
    real *gate_volume;   /* [s_num] post-limiter q*length*dt of the step */
    real *gate_sources;  /* [e_num*6] record_cell + collect values of the step */
};

int32_t nhf_abi_version(void) { return NHF_ABI_VERSION; }
const char *nhf_last_error(void) { return g_error; }
size_t nhf_workspace_bytes(const nhf_mesh *mesh, const nhf_params *params) {
    (void)mesh;
    (void)params;
    return 0;
}

static int cell_boundary(const nhf_ctx *c, int ei) { return (int)((c->m.cell_meta[ei] >> 8) & 255u) - 1; }
static int cell_class(const nhf_ctx *c, int ei) { return (int)(c->m.cell_meta[ei] & 255u); }
static int cell_pixel(const nhf_ctx *c, int ei) { return c->m.n_pixels > 1 ? c->m.rain_index[ei] : 0; }

static int face_count(const nhf_ctx *c, int ei, int d) {
    return (int)((c->m.csr_count[ei] >> (8 * d)) & 255u);
}

static int face_start(const nhf_ctx *c, int ei, int d) {
    int start = c->m.csr_start[ei];
    for (int k = 0; k < d; ++k) start += face_count(c, ei, k);
    return start;
}

static void *alloc(size_t n, size_t size) {
    void *ptr = calloc(n ? n : 1, size);
    return ptr;
}

/* Load the F32 state arrays into the working arrays (a no-op copy for REAL = float). */
static void load_state(nhf_ctx *c) {
    const nhf_mesh *m = &c->m;
    for (int i = 0; i < m->e_num; ++i) {
        c->depth[i] = m->depth[i];
        c->bed[i] = m->bed[i];
        c->rain_time[i] = m->rain_time[i];
        c->capacity[i] = m->capacity[i];
        c->max_depth[i] = m->max_depth[i];
        c->max_speed[i] = m->max_speed[i];
    }
    for (int i = 0; i < m->s_num; ++i) c->q[i] = m->q[i];
}

/* Take in what the host wrote since the last batch (extrema resets, restored state): an element
 * whose stored F32 value no longer equals the F32 rounding of the working value was written by
 * the host. Unchanged elements keep their working value, so the FP64 build stays FP64. */
static void sync_host_writes(nhf_ctx *c) {
    const nhf_mesh *m = &c->m;
#define NHF_SYNC(arr, n)                                          \
    for (int i = 0; i < (n); ++i)                                 \
        if ((float)c->arr[i] != m->arr[i]) c->arr[i] = m->arr[i];
    NHF_SYNC(depth, m->e_num)
    NHF_SYNC(bed, m->e_num)
    NHF_SYNC(rain_time, m->e_num)
    NHF_SYNC(capacity, m->e_num)
    NHF_SYNC(max_depth, m->e_num)
    NHF_SYNC(max_speed, m->e_num)
    NHF_SYNC(q, m->s_num)
#undef NHF_SYNC
}

static void store_state(nhf_ctx *c) {
    const nhf_mesh *m = &c->m;
    for (int i = 0; i < m->e_num; ++i) {
        m->depth[i] = (float)c->depth[i];
        m->bed[i] = (float)c->bed[i];
        m->rain_time[i] = (float)c->rain_time[i];
        m->capacity[i] = (float)c->capacity[i];
        m->max_depth[i] = (float)c->max_depth[i];
        m->max_speed[i] = (float)c->max_speed[i];
    }
    for (int i = 0; i < m->s_num; ++i) m->q[i] = (float)c->q[i];
}

int32_t nhf_create(nhf_ctx **out, const nhf_mesh *mesh, const nhf_params *params,
                   void *workspace, size_t workspace_bytes, void *stream) {
    (void)workspace;
    (void)workspace_bytes;
    (void)stream;
    if (!out || !mesh || !params) return fail("null argument");
    if (params->batch_capacity < 1 || params->batch_capacity > NHF_MAX_BATCH)
        return fail("batch_capacity out of range");
    nhf_ctx *c = (nhf_ctx *)calloc(1, sizeof(nhf_ctx));
    c->m = *mesh;
    c->p = *params;
    int e = mesh->e_num, s = mesh->s_num, b = mesh->b_num;
    c->terms = 4 + 6 * b;
    c->coefficients = (float *)alloc((size_t)2 * (b > 0 ? b : 1), sizeof(float));
    c->values = (real *)alloc(b > 0 ? b : 1, sizeof(real));
    c->level_values = (real *)alloc(b > 0 ? b : 1, sizeof(real));
    c->rain_rates = (float *)alloc(mesh->n_pixels, sizeof(float));
    c->depth = (real *)alloc(e, sizeof(real));
    c->bed = (real *)alloc(e, sizeof(real));
    c->rain_time = (real *)alloc(e, sizeof(real));
    c->capacity = (real *)alloc(e, sizeof(real));
    c->max_depth = (real *)alloc(e, sizeof(real));
    c->max_speed = (real *)alloc(e, sizeof(real));
    c->q = (real *)alloc(s, sizeof(real));
    c->donor = (real *)alloc(e, sizeof(real));
    c->face_qn = (real *)alloc((size_t)e * 4, sizeof(real));
    c->gradients = (real *)alloc((size_t)(mesh->n_grad > 0 ? mesh->n_grad : 1) * 2, sizeof(real));
    c->cfl = (real *)alloc(s, sizeof(real));
    c->flow_depth = (real *)alloc(s, sizeof(real));
    c->eq = (real *)alloc((size_t)e * 4, sizeof(real));
    c->ssq = (real *)alloc(e, sizeof(real));
    c->increment = (real *)alloc(e, sizeof(real));
    c->rain_depth = (real *)alloc(e, sizeof(real));
    c->diag = (real *)alloc((size_t)e * 8, sizeof(real));
    c->budget = (ddsum *)alloc(c->terms, sizeof(ddsum));
    c->limit_value[0] = INFINITY;
    *out = c;
    return 0;
}

int32_t nhf_destroy(nhf_ctx *c) {
    if (!c) return 0;
    void *ptrs[] = {c->coefficients, c->values, c->level_values, c->rain_rates, c->depth, c->bed,
                    c->rain_time, c->capacity, c->max_depth, c->max_speed, c->q, c->donor,
                    c->face_qn, c->gradients, c->cfl, c->flow_depth, c->eq, c->ssq, c->increment,
                    c->rain_depth, c->diag, c->budget, c->gate_budget, c->gate_volume, c->gate_sources};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(c);
    return 0;
}

int32_t nhf_reset(nhf_ctx *c, uint64_t now_ticks) {
    c->now = now_ticks;
    c->next = now_ticks;
    c->origin = now_ticks;
    c->active = 0;
    c->limit_value[0] = INFINITY;
    for (int t = 0; t < c->terms; ++t) c->budget[t].hi = c->budget[t].lo = 0.0;
    for (int i = 0; i < c->m.e_num; ++i) c->donor[i] = 0;
    load_state(c);
    return 0;
}

int32_t nhf_set_forcing(nhf_ctx *c, const float *coefficients, const float *rain_rates,
                        int32_t raining) {
    int b = c->m.b_num;
    if (b > 0 && coefficients) memcpy(c->coefficients, coefficients, sizeof(float) * 2 * b);
    if (rain_rates) memcpy(c->rain_rates, rain_rates, sizeof(float) * c->m.n_pixels);
    c->raining = raining;
    return 0;
}

int32_t nhf_begin(nhf_ctx *c, uint64_t target_ticks, uint64_t origin_ticks) {
    if (target_ticks < c->now) return fail("target precedes the current clock");
    c->next = target_ticks;
    c->origin = origin_ticks;
    memset(c->trace, 0, sizeof c->trace);
    c->slot = 0;
    return 0;
}

/* ---------------------------------------------------------------- clock (clock.py:94-181) */

static real ticks_to_local(uint64_t ticks) {
    uint32_t hi = (uint32_t)(ticks >> 32), lo = (uint32_t)ticks;
    return (real)hi * (real)(1.0 / 256.0) + (real)lo * (real)(1.0 / TICKS_PER_SECOND);
}

static uint32_t f32_bits(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    return u;
}

/* F32 (or F64) dt -> ticks with round-half-even below one tick; returns 0 if unrepresentable. */
static int dt_to_ticks(real chosen, uint64_t *ticks) {
#if NHF_REAL_IS_DOUBLE
    if (!(chosen > 0.0)) return 0;
    int e;
    double m = frexp(chosen, &e); /* chosen = m * 2^e, m in [0.5, 1) */
    if (e - 1 < -40 || e - 1 > 0) return 0;
    uint64_t mant = (uint64_t)ldexp(m, 53); /* 53-bit integer */
    int shift = e - 53 + 40;
    uint64_t t;
    if (shift >= 0) {
        t = mant << shift;
    } else {
        int right = -shift;
        t = right >= 64 ? 0 : mant >> right;
        uint64_t dropped = right >= 64 ? mant : mant & ((1ull << right) - 1);
        uint64_t half = 1ull << (right - 1);
        if (dropped > half || (dropped == half && (t & 1))) t += 1;
    }
    *ticks = t;
    return t != 0;
#else
    uint32_t bits = f32_bits(chosen);
    int exponent = (int)((bits >> 23) & 255u) - 127;
    uint32_t mantissa = (bits & 0x7fffffu) | 0x800000u;
    if (!(chosen > 0.0f) || exponent < -40 || exponent > 0) return 0;
    uint32_t dt_hi = 0, dt_lo = 0;
    int shift = exponent + 17;
    if (shift >= 0) {
        dt_lo = mantissa << shift;
        if (shift >= 9) dt_hi = mantissa >> (32 - shift);
    } else {
        int right = -shift;
        dt_lo = mantissa >> right;
        uint32_t dropped = mantissa & ((1u << right) - 1u);
        uint32_t half = 1u << (right - 1);
        if (dropped > half || (dropped == half && (dt_lo & 1u) != 0)) dt_lo += 1;
    }
    *ticks = ((uint64_t)dt_hi << 32) | dt_lo;
    return *ticks != 0;
#endif
}

static void clock_choose(nhf_ctx *c, real candidate, int slot) {
    uint32_t *row = &c->trace[slot * NHF_TRACE_COLS];
    c->active = 0;
    c->dt = 0;
    c->local = 0;
    c->end_local = 0;
    if (!(c->now < c->next)) return;
    if (!(candidate >= (real)c->p.min_dt)) {
        row[3] = NHF_BAD_CFL;
        return;
    }
    uint64_t rem = c->next - c->now;
    real remaining = ticks_to_local(rem);
    real capped = rmin(candidate, (real)1.0);
    real chosen = rmin(capped, remaining);
    row[4] = f32_bits((float)candidate);
    row[5] = chosen < capped ? 1u : 0u;
    uint64_t dt_ticks;
    if (!dt_to_ticks(chosen, &dt_ticks)) {
        row[3] = NHF_BAD_DT;
        return;
    }
    int exceeds = dt_ticks > rem;
    uint64_t diff = exceeds ? dt_ticks - rem : rem - dt_ticks;
    uint64_t end;
    if (diff < SNAP_TICKS) {
        end = c->next;
    } else if (exceeds) {
        row[3] = NHF_OVERSHOOT;
        return;
    } else {
        end = c->now + dt_ticks;
    }
    c->local = ticks_to_local(c->now - c->origin);
    c->end_local = ticks_to_local(end - c->origin);
    c->dt = chosen;
    c->step_end = end;
    c->active = 1;
}

static void clock_commit(nhf_ctx *c, int slot) {
    if (!c->active) return;
    uint32_t *row = &c->trace[slot * NHF_TRACE_COLS];
    c->now = c->step_end;
    row[0] = f32_bits((float)c->dt);
    row[1] = (uint32_t)(c->step_end >> 32);
    row[2] = (uint32_t)c->step_end;
    row[3] = NHF_COMPLETED;
}

int32_t nhf_clock_probe(const uint64_t *now, const uint64_t *target, const uint64_t *origin,
                        const float *candidate, float min_dt, int32_t n, uint32_t *out_words,
                        uint64_t *out_ticks, float *out_times) {
    nhf_ctx c;
    memset(&c, 0, sizeof c);
    c.p.min_dt = min_dt;
    for (int i = 0; i < n; ++i) {
        c.now = now[i];
        c.next = target[i];
        c.origin = origin[i];
        memset(c.trace, 0, sizeof(uint32_t) * NHF_TRACE_COLS);
        c.step_end = 0;
        clock_choose(&c, (real)candidate[i], 0);
        out_words[i * 4 + 0] = c.active ? NHF_COMPLETED : c.trace[3];
        out_words[i * 4 + 1] = f32_bits((float)c.dt);
        out_words[i * 4 + 2] = c.trace[4];
        out_words[i * 4 + 3] = c.trace[5];
        uint64_t dt_ticks = 0;
        if (c.active) dt_to_ticks(c.dt, &dt_ticks);
        out_ticks[i * 2 + 0] = c.active ? c.step_end : 0;
        out_ticks[i * 2 + 1] = dt_ticks;
        out_times[i * 2 + 0] = (float)c.local;
        out_times[i * 2 + 1] = (float)c.end_local;
    }
    return 0;
}

/* ------------------------------------------------------------ step passes (solver_compact.py) */

static void add_diag(nhf_ctx *c, int term, real value) {
    if (value != 0) dd_add(&c->budget[term], (double)value);
}

static void prepare_step(nhf_ctx *c) {
    clock_choose(c, c->ndt, c->slot);
    if (!c->active) return;
    c->limited_cells = c->limited_faces = 0;
    c->min_scale = 1;
    for (int bi = 0; bi < c->m.b_num; ++bi) {
        real c0 = c->coefficients[2 * bi], c1 = c->coefficients[2 * bi + 1];
        c->values[bi] = c0 + c1 * (c->local + c->dt / (real)2.0);
        c->level_values[bi] = c0 + c1 * c->end_local;
    }
    if ((float)c->ndt < c->limit_value[0]) {
        int si = c->cfl_arg;
        int el = c->m.side_cells[2 * si], eh = c->m.side_cells[2 * si + 1];
        real h = c->flow_depth[si];
        c->limit_value[0] = (float)c->ndt;
        c->limit_value[1] = (float)c->depth[el];
        c->limit_value[2] = (float)c->depth[eh];
        c->limit_value[3] = (float)h;
        c->limit_value[4] = (float)(c->q[si] / rmax(h, (real)1e-30));
        c->limit_ids[0] = si;
        c->limit_ids[1] = el;
        c->limit_ids[2] = eh;
    }
}

/* PressureReconstruction.face_difference (pressure.py:66-90) */
static void face_difference(const nhf_ctx *c, int cell, int direction, int axis, real *value,
                            real *distance, int *usable) {
    const nhf_mesh *m = &c->m;
    real v = 0, dist = 0, total = 0;
    int ok = 1;
    real eta = c->bed[cell] + c->depth[cell];
    int start = face_start(c, cell, direction);
    int count = face_count(c, cell, direction);
    for (int j = 0; j < count; ++j) {
        int side = m->isl[start + j];
        int other = m->side_cells[2 * side] + m->side_cells[2 * side + 1] - cell;
        if (other <= 0 || c->depth[other] <= (real)c->p.min_h) {
            ok = 0;
        } else {
            real weight = m->length[side];
            v += weight * ((c->bed[other] + c->depth[other]) - eta);
            real delta = (real)m->x[other] - (real)m->x[cell];
            if (axis == 1) delta = (real)m->y[other] - (real)m->y[cell];
            dist += weight * delta;
            total += weight;
        }
    }
    if (total > 0) {
        v /= total;
        dist /= total;
    } else {
        ok = 0;
    }
    *value = v;
    *distance = dist;
    *usable = ok;
}

/* PressureReconstruction.gradient (pressure.py:92-106) */
static real gradient(const nhf_ctx *c, int cell, int axis) {
    int minus = axis == 1 ? 2 : 0, plus = axis == 1 ? 3 : 1;
    real dm, xm, dp, xp;
    int okm, okp;
    face_difference(c, cell, minus, axis, &dm, &xm, &okm);
    face_difference(c, cell, plus, axis, &dp, &xp, &okp);
    if (okm && okp) return (dp - dm) / (xp - xm);
    if (okm) return dm / xm;
    if (okp) return dp / xp;
    return 0;
}

/* update_pressure / update_cell (solver_compact.py:346-349, pressure.py:108-116) */
static void update_pressure(nhf_ctx *c) {
    const nhf_mesh *m = &c->m;
#pragma omp parallel for schedule(static) if (m->n_grad > OMP_MIN_ITEMS)
    for (int slot = 0; slot < m->n_grad; ++slot) {
        int cell = m->grad_cell[slot];
        for (int axis = 0; axis < 2; ++axis) {
            if (m->grad_axes[slot] & (1 << axis)) {
                real value = 0;
                if (c->depth[cell] > (real)c->p.min_h) value = gradient(c, cell, axis);
                c->gradients[slot * 2 + axis] = value;
            }
        }
    }
}

/* PressureReconstruction.difference (pressure.py:118-129) */
static real pressure_difference(const nhf_ctx *c, int side, int lower, int higher, real raw) {
    const nhf_mesh *m = &c->m;
    real result = raw;
    int index = m->edge_index[side];
    if (index >= 0) {
        if (c->depth[lower] > (real)c->p.min_h && c->depth[higher] > (real)c->p.min_h) {
            int axis = (m->side_meta[side] & NHF_SIDE_ORIENT_MASK) == 2 ? 1 : 0;
            real o0 = m->hang_offset[2 * index], o1 = m->hang_offset[2 * index + 1];
            if (o0 != 0) result -= c->gradients[m->hang_slot[2 * index] * 2 + axis] * o0;
            if (o1 != 0) result += c->gradients[m->hang_slot[2 * index + 1] * 2 + axis] * o1;
        }
    }
    return result;
}

/* face_discharge (solver_compact.py:118-132) */
static real face_discharge(const nhf_ctx *c, int start, int count) {
    const nhf_mesh *m = &c->m;
    real result = 0;
    if (count == 1) {
        result = c->q[m->isl[start]];
    } else if (count > 1) {
        real total = 0, width = 0;
        for (int k = 0; k < count; ++k) {
            int side = m->isl[start + k];
            total += c->q[side] * (real)m->length[side];
            width += (real)m->length[side];
        }
        result = total / width;
    }
    return result;
}

/* boundary_neighbor_sum (solver_compact.py:135-149) */
static real boundary_neighbor_sum(real q, real left, real right, real left_ext, real right_ext) {
    real result = left + right;
    if (left_ext > 0 || right_ext > 0) {
        if (left_ext == 1 || right_ext == 1) {
            result = (real)2.0 * q;
        } else {
            real product = left_ext * right_ext;
            result = ((1 - right_ext) * left + (1 - left_ext) * right +
                      (real)2.0 * q * (left_ext + right_ext - (real)2.0 * product)) /
                     (1 - product);
        }
    }
    return result;
}

/* ---------------------------------------------------------------- gate control (core/gate.py) */

static int gate_source_offset(const nhf_ctx *c, int cell) {
    const int count = c->g.count;
    if (c->g.pond[cell]) return 4 * count + 2;
    if (c->g.cell_gate[cell] >= 0) return 4 * count + 8 + 6 * c->g.cell_gate[cell];
    return -1;
}

/* GateController.move (core/gate.py:72-86) at `tick`. */
static void gate_move(nhf_ctx *c, uint64_t tick) {
    const nhf_gates *g = &c->g;
    const uint32_t hi = (uint32_t)(tick >> 32), lo = (uint32_t)tick;
    for (int i = 0; i < g->n_cells; ++i) {
        int cell = g->cells[i], k = g->cell_gate[cell];
        if (g->flags[3 * k] != 1) continue;
        uint32_t chi = (uint32_t)(g->ticks[3 * k] >> 32), clo = (uint32_t)g->ticks[3 * k];
        uint32_t borrow = lo < clo ? 1u : 0u;
        uint32_t high = hi - chi - borrow, low = lo - clo;
        float since = (float)high / 256.0f + (float)low * 0x1p-40f;
        float fraction = since / g->opening_seconds;
        if (!(fraction < 1.0f)) fraction = 1.0f;
        if (fraction >= 1.0f) {
            c->bed[cell] = (real)g->open_elevation;
            if (!g->flags[3 * k + 2]) {
                g->flags[3 * k + 2] = 1;
                g->ticks[3 * k + 2] = tick;
            }
        } else {
            float closed = g->closed[cell];
            c->bed[cell] = (real)(closed + (g->open_elevation - closed) * fraction);
        }
    }
}

/* GateController.check (core/gate.py:88-98): the trigger test at the step start, then move. */
static void gate_check(nhf_ctx *c) {
    const nhf_gates *g = &c->g;
    for (int k = 0; k < g->count; ++k) {
        int cell = g->control[k];
        if (c->now >= g->formal_tick && g->flags[3 * k] == 0 && c->depth[cell] >= (real)g->min_h &&
            c->bed[cell] + c->depth[cell] > (real)g->trigger_level) {
            g->flags[3 * k] = 1;
            g->ticks[3 * k] = c->now;
        }
    }
    gate_move(c, c->now);
}

static void gate_directed(nhf_ctx *c, int term, real into) {
    dd_add(&c->gate_budget[term], (double)rmax(into, 0));
    dd_add(&c->gate_budget[term + 1], (double)rmax(-into, 0));
}

/* record_flux (core/gate.py:115-139) for every side, in side order, after the limiter. */
static void gate_record_sides(nhf_ctx *c) {
    const nhf_gates *g = &c->g;
    const int offset = 4 * g->count;
    for (int si = 1; si < c->m.s_num; ++si) {
        int el = c->m.side_cells[2 * si], eh = c->m.side_cells[2 * si + 1];
        int gl = g->cell_gate[el], gr = g->cell_gate[eh], pl = g->pond[el], pr = g->pond[eh];
        real volume = c->gate_volume[si];
        int flow_gate = -1;
        if (gl >= 0 && gr < 0) {
            if (pr) {
                gate_directed(c, 4 * gl, volume);
                flow_gate = gl;
            } else {
                gate_directed(c, 4 * gl + 2, -volume);
            }
        } else if (gr >= 0 && gl < 0) {
            if (pl) {
                gate_directed(c, 4 * gr, -volume);
                flow_gate = gr;
            } else {
                gate_directed(c, 4 * gr + 2, volume);
            }
        } else if (pl != pr) {
            gate_directed(c, offset, pr ? volume : -volume);
        }
        if (flow_gate >= 0 && volume != 0 && !g->flags[3 * flow_gate + 1]) {
            g->flags[3 * flow_gate + 1] = 1;
            g->ticks[3 * flow_gate + 1] = c->step_end;
        }
    }
}

/* record_cell + measure_inventory + collect (core/gate.py:149-188), in cell order. */
static void gate_record_cells(nhf_ctx *c) {
    for (int ei = 1; ei < c->m.e_num; ++ei) {
        int offset = gate_source_offset(c, ei);
        if (offset < 0) continue;
        for (int k = 0; k < 6; ++k) dd_add(&c->gate_budget[offset + k], (double)c->gate_sources[ei * 6 + k]);
    }
}

static void tick(nhf_ctx *c) {
    const nhf_mesh *m = &c->m;
    const real dt = c->dt;
    const real g = (real)c->p.g;
    const real theta = (real)c->p.theta, theta_half = (real)c->p.theta_half;
    const real min_h = (real)c->p.min_h;
    const int theta_one = c->p.theta == 1.0f;
    const int hanging = m->n_hang > 0;
    const int e_num = m->e_num, s_num = m->s_num, b_num = m->b_num;

    /* face averages of the previous discharge (:430-436) */
    if (!theta_one) {
#pragma omp parallel for schedule(static) if (e_num > OMP_MIN_ITEMS)
        for (int ei = 0; ei < e_num; ++ei)
            for (int d = 0; d < 4; ++d)
                c->face_qn[ei * 4 + d] = face_discharge(c, face_start(c, ei, d), face_count(c, ei, d));
    }

    /* momentum on every side (:439-486) */
#pragma omp parallel for schedule(static) if (s_num > OMP_MIN_ITEMS)
    for (int si = 1; si < s_num; ++si) {
        real bf = (m->side_meta[si] & NHF_SIDE_EXTERIOR) ? (real)0.0 : (real)1.0;
        int orient = m->side_meta[si] & NHF_SIDE_ORIENT_MASK;
        int el = m->side_cells[2 * si], eh = m->side_cells[2 * si + 1];
        int dl = orient == 2 ? 0 : 2, dh_face = orient == 2 ? 1 : 3;
        real hl = c->bed[el] + c->depth[el], hr = c->bed[eh] + c->depth[eh];
        real dx = (real)m->sdc[si];
        real xq = c->q[si];
        real crest = rmax(rmax(c->bed[el], c->bed[eh]), (real)-999.0);
        real xwh = rmax(hl, hr) - crest;
        real difference = hr - hl;
        if (hanging) difference = pressure_difference(c, si, el, eh, difference);
        real xdq = ((-g * difference) / dx) * rmax(xwh, (real)0.0) * dt;
        real n = (real)m->manning[si];
        real xf = (real)1.0 + g * dt * (n * n) *
                                  RABS(xq / RPOW(rmax(xwh, (real)0.00001), (real)(7.0 / 3.0)));
        real neighbours = c->face_qn[el * 4 + dl] + c->face_qn[eh * 4 + dh_face];
        if (!theta_one)
            neighbours = boundary_neighbor_sum(xq, c->face_qn[el * 4 + dl],
                                               c->face_qn[eh * 4 + dh_face],
                                               (real)m->open_face[el * 4 + dl],
                                               (real)m->open_face[eh * 4 + dh_face]);
        real new_q = (theta * xq + theta_half * neighbours + xdq) / xf;
        new_q = xwh < min_h ? (real)0.0 : new_q;
        new_q *= bf;
        c->q[si] = new_q;
    }

    /* donor scale per cell (:491-514) */
    uint32_t limited_cells = 0;
    real min_scale = c->min_scale;
#pragma omp parallel for schedule(static) reduction(+ : limited_cells) reduction(min : min_scale) if (e_num > OMP_MIN_ITEMS)
    for (int ei = 1; ei < e_num; ++ei) {
        int bi = cell_boundary(c, ei);
        real discharge = 0;
        if (bi >= 0 && m->kind[bi] == 1) discharge = c->values[bi] * (real)m->bweight[ei];
        real rainq = 0;
        if (c->raining) rainq = (real)c->rain_rates[cell_pixel(c, ei)];
        real pq = m->pipe_q ? (real)m->pipe_q[ei] : (real)0.0;
        real positive = (rmax(pq, 0) + rmax(discharge, 0)) * dt;
        real available = c->depth[ei] * (real)m->area[ei] + positive + rainq * dt * (real)m->area[ei];
        real outgoing = (rmax(-pq, 0) + rmax(-discharge, 0)) * dt;
        for (int d = 0; d < 4; ++d) {
            int start = face_start(c, ei, d), count = face_count(c, ei, d);
            for (int j = 0; j < count; ++j) {
                int side = m->isl[start + j];
                real signed_q = c->q[side] * (real)m->length[side];
                if (d == 0 || d == 2) signed_q = -signed_q;
                outgoing += rmax(signed_q, 0) * dt;
            }
        }
        c->donor[ei] = 1;
        if (outgoing > available && !(bi >= 0 && m->kind[bi] == 0)) {
            c->donor[ei] = available / outgoing;
            limited_cells += 1;
            min_scale = rmin(min_scale, c->donor[ei]);
        }
    }
    c->limited_cells += limited_cells;
    c->min_scale = min_scale;

    /* shared-subedge scaling (:516-524) */
    uint32_t limited_faces = 0;
#pragma omp parallel for schedule(static) reduction(+ : limited_faces) if (s_num > OMP_MIN_ITEMS)
    for (int si = 1; si < s_num; ++si) {
        int el = m->side_cells[2 * si], eh = m->side_cells[2 * si + 1];
        int donor = c->q[si] >= 0 ? el : eh;
        if (c->q[si] != 0 && c->donor[donor] < 1) limited_faces += 1;
        c->q[si] *= c->donor[donor];
        if (c->gates_on) c->gate_volume[si] = c->q[si] * (real)m->length[si] * dt;
    }
    c->limited_faces += limited_faces;
    if (c->gates_on) gate_record_sides(c);

    /* sources, Horton capacity and face totals (:525-566) */
#pragma omp parallel for schedule(static) if (e_num > OMP_MIN_ITEMS)
    for (int ei = 1; ei < e_num; ++ei) {
        int bi = cell_boundary(c, ei);
        real ds = c->donor[ei];
        real pq = m->pipe_q ? (real)m->pipe_q[ei] : (real)0.0;
        real pipe = rmax(pq, 0) + rmin(pq, 0) * ds;
        real discharge = 0;
        if (bi >= 0 && m->kind[bi] == 1) {
            real requested = c->values[bi] * (real)m->bweight[ei];
            discharge = rmax(requested, 0) + rmin(requested, 0) * ds;
        }
        c->ssq[ei] = pipe + discharge;
        if (m->pipe_volume) { /* compensated window volume (:535-539), F32 as the reference's fields */
            float increment = (float)(pipe * dt) - m->pipe_correction[ei];
            float before = m->pipe_volume[ei];
            float total = before + increment;
            m->pipe_correction[ei] = (total - before) - increment;
            m->pipe_volume[ei] = total;
        }
        real rainq = 0;
        if (c->raining) rainq = (real)c->rain_rates[cell_pixel(c, ei)];
        if (rainq > 0) {
            c->rain_time[ei] += dt;
            const float *hp = &m->class_horton[3 * cell_class(c, ei)];
            real f0 = hp[0], fc = hp[1], hk = hp[2];
            c->capacity[ei] = fc + (f0 - fc) * REXP(-hk * c->rain_time[ei] / (real)3600.0);
        }
        real f[4];
        for (int d = 0; d < 4; ++d) {
            real face_flux = 0;
            int start = face_start(c, ei, d), count = face_count(c, ei, d);
            for (int j = 0; j < count; ++j) {
                int side = m->isl[start + j];
                face_flux += c->q[side] * (real)m->length[side];
            }
            f[d] = face_flux;
            c->eq[ei * 4 + d] = face_flux;
        }
        real tq = f[0] - f[1] + f[2] - f[3] + c->ssq[ei];
        real rain_depth = rainq * dt;
        c->increment[ei] = tq * dt / (real)m->area[ei] + rain_depth;
        c->rain_depth[ei] = rain_depth;
    }
    /* discharge diagnostic in cell order (:533) */
    for (int ei = 1; ei < e_num; ++ei) {
        int bi = cell_boundary(c, ei);
        if (bi >= 0 && m->kind[bi] == 1) {
            real requested = c->values[bi] * (real)m->bweight[ei];
            real ds = c->donor[ei];
            real discharge = rmax(requested, 0) + rmin(requested, 0) * ds;
            add_diag(c, 4 + 5 * b_num + bi, discharge * dt);
        }
    }

    /* depth, positivity, level cells and extrema (:567-610) */
    const uint64_t step_end = c->step_end;
#pragma omp parallel for schedule(static) if (e_num > OMP_MIN_ITEMS)
    for (int ei = 1; ei < e_num; ++ei) {
        int bi = cell_boundary(c, ei);
        real ea = (real)m->area[ei], previous = c->depth[ei];
        real pq = m->pipe_q ? (real)m->pipe_q[ei] : (real)0.0;
        real pipe = rmax(pq, 0) + rmin(pq, 0) * c->donor[ei];
        real ql = c->eq[ei * 4 + 0], qr = c->eq[ei * 4 + 1], qb = c->eq[ei * 4 + 2], qt = c->eq[ei * 4 + 3];
        real rain_depth = c->rain_depth[ei];
        real raw = previous + c->increment[ei];
        real removed = rmin(c->capacity[ei] * dt, rmax(raw, 0));
        raw -= removed;
        real floor_ = rmax(-raw, 0);
        real *dg = &c->diag[ei * 8];
        dg[0] = removed * ea;
        dg[2] = pipe * dt;
        if (bi >= 0 && m->kind[bi] == 0) {
            dg[1] = 0;
            dg[7] = floor_ * ea;
            real target = rmax(c->level_values[bi] - c->bed[ei], 0);
            dg[3] = (target - previous) * ea;
            dg[4] = -(ql - qr + qb - qt) * dt;
            dg[5] = -pipe * dt;
            dg[6] = -rain_depth * ea;
            c->depth[ei] = target;
        } else {
            dg[1] = floor_ * ea;
            c->depth[ei] = rmax(raw, 0);
        }
        if (c->gates_on && gate_source_offset(c, ei) >= 0) {
            /* record_cell (solver_compact.py:595-603, core/gate.py:149-163) and collect (:179-188) */
            int level = bi >= 0 && m->kind[bi] == 0;
            real applied_floor = level ? (real)0.0 : floor_;
            real level_exchange = 0;
            if (level)
                level_exchange = (c->depth[ei] - previous) * ea - (ql - qr + qb - qt) * dt - pipe * dt -
                                 rain_depth * ea + removed * ea;
            real delta = c->depth[ei] - previous;
            real rounding = delta - c->increment[ei] + removed - applied_floor;
            real *gs = &c->gate_sources[ei * 6];
            gs[0] = rain_depth * ea;
            gs[1] = removed * ea;
            gs[2] = applied_floor * ea;
            gs[3] = pipe * dt;
            gs[4] = (c->ssq[ei] - pipe) * dt + level_exchange;
            gs[5] = rounding * ea - level_exchange;
        }
        real h = rmax(c->depth[ei], (real)0.01);
        real u = 0, v = 0;
        if (c->depth[ei] >= min_h) {
            u = (ql + qr) / (real)m->size[ei] / h / (real)2.0;
            v = (qb + qt) / (real)m->size[ei] / h / (real)2.0;
        }
        c->max_speed[ei] = rmax(c->max_speed[ei], RSQRT(u * u + v * v));
        if (c->depth[ei] > c->max_depth[ei]) {
            c->max_depth[ei] = c->depth[ei];
            m->peak_tick[ei] = step_end;
        }
    }
    if (c->gates_on) gate_record_cells(c);
    /* budget terms, added per cell in cell order (:578-590) */
    for (int ei = 1; ei < e_num; ++ei) {
        const real *dg = &c->diag[ei * 8];
        int bi = cell_boundary(c, ei);
        add_diag(c, 0, dg[0]);
        add_diag(c, 3, dg[2]);
        if (bi >= 0 && m->kind[bi] == 0) {
            int offset = 4 + 5 * bi;
            add_diag(c, 2, dg[7]);
            add_diag(c, offset, dg[3]);
            add_diag(c, offset + 1, dg[4]);
            add_diag(c, offset + 2, dg[5]);
            add_diag(c, offset + 3, dg[6]);
            add_diag(c, offset + 4, dg[0]);
        } else {
            add_diag(c, 1, dg[1]);
        }
    }
}

static void compute_cfl(nhf_ctx *c) {
    const nhf_mesh *m = &c->m;
    const real g = (real)c->p.g, afa = (real)c->p.afa;
    real ndt = INFINITY;
#pragma omp parallel for schedule(static) reduction(min : ndt) if (m->s_num > OMP_MIN_ITEMS)
    for (int si = 1; si < m->s_num; ++si) {
        int el = m->side_cells[2 * si], eh = m->side_cells[2 * si + 1];
        c->cfl[si] = INFINITY;
        if (el != 0 && eh != 0) {
            real head = rmax(c->bed[el] + c->depth[el], c->bed[eh] + c->depth[eh]);
            real h = rmax(head - rmax(rmax(c->bed[el], c->bed[eh]), (real)-999.0), (real)0.01);
            c->flow_depth[si] = h;
            c->cfl[si] = afa * (real)m->length[si] / (RSQRT(g * h) + RABS(c->q[si]) / h);
            ndt = rmin(ndt, c->cfl[si]);
        }
    }
    c->ndt = ndt;
    int32_t arg = 0;
    uint32_t best = UINT32_MAX;
    for (int si = 1; si < m->s_num; ++si) {
        int el = m->side_cells[2 * si], eh = m->side_cells[2 * si + 1];
        if (c->cfl[si] == ndt && el != 0 && eh != 0 && m->side_key[si] < best) {
            best = m->side_key[si];
            arg = si;
        }
    }
    c->cfl_arg = arg;
}

static void finish_step(nhf_ctx *c) {
    if (c->active) {
        uint32_t *row = &c->trace[c->slot * NHF_TRACE_COLS];
        row[6] = c->limited_cells;
        row[7] = c->limited_faces;
        row[8] = f32_bits((float)c->min_scale);
    }
    clock_commit(c, c->slot);
    c->slot += 1;
}

/* Kernel timing has no meaning on the CPU: reports zeros. */
int32_t nhf_time_kernels(nhf_ctx *c, int32_t repeats, double *us) {
    for (int i = 0; i < 9; ++i) us[i] = 0.0;
    return 0;
}

/* Launch tuning has no meaning on the CPU: accepted and ignored. */
int32_t nhf_set_tuning(nhf_ctx *c, int32_t cfl_blocks, int32_t cell_block, int32_t cells_per_thread) { return 0; }

int32_t nhf_tuning_range(nhf_ctx *c, int32_t *out4) {
    out4[0] = 8;
    out4[1] = 8;
    out4[2] = 256;
    out4[3] = 1;
    return 0;
}

int32_t nhf_launch_batch(nhf_ctx *c, int32_t steps) {
    if (c->pending_steps) return fail("the previous batch was not waited for");
    int32_t rc = nhf_run_batch(c, steps, NULL);
    if (rc == 0) c->pending_steps = steps;
    return rc;
}

int32_t nhf_wait_batch(nhf_ctx *c, uint32_t *trace) {
    if (!c->pending_steps) return fail("no batch in flight");
    if (trace) memcpy(trace, c->trace, sizeof(uint32_t) * NHF_TRACE_COLS * c->pending_steps);
    c->pending_steps = 0;
    return 0;
}

int32_t nhf_run_batch(nhf_ctx *c, int32_t steps, uint32_t *trace) {
    if (steps < 1 || steps > c->p.batch_capacity) return fail("steps out of range");
    sync_host_writes(c);
    c->slot = 0;
    memset(c->trace, 0, sizeof c->trace);
    for (int s = 0; s < steps; ++s) {
        prepare_step(c);
        if (c->active) {
            if (c->gates_on) gate_check(c);
            if (c->m.n_grad > 0) update_pressure(c);
            tick(c);
            if (c->gates_on) gate_move(c, c->step_end); /* finish_gate_bed (core/gate.py:100-101) */
            compute_cfl(c);
        }
        finish_step(c);
    }
    if (trace) memcpy(trace, c->trace, sizeof(uint32_t) * NHF_TRACE_COLS * steps);
    store_state(c);
    return 0;
}

int32_t nhf_compute_cfl(nhf_ctx *c) {
    sync_host_writes(c);
    compute_cfl(c);
    return 0;
}

int32_t nhf_flush_budget(nhf_ctx *c, double *out) {
    for (int t = 0; t < c->terms; ++t)
        if (!isfinite(c->budget[t].hi) || !isfinite(c->budget[t].lo)) return 1;
    for (int t = 0; t < c->terms; ++t) {
        out[2 * t] = c->budget[t].hi;
        out[2 * t + 1] = c->budget[t].lo;
        c->budget[t].hi = c->budget[t].lo = 0.0;
    }
    return 0;
}

int32_t nhf_limit_capture(nhf_ctx *c, float *values, int32_t *ids, int32_t reset) {
    if (values) memcpy(values, c->limit_value, sizeof c->limit_value);
    if (ids) memcpy(ids, c->limit_ids, sizeof c->limit_ids);
    if (reset) c->limit_value[0] = INFINITY;
    return 0;
}

int32_t nhf_cfl_state(nhf_ctx *c, float *ndt, int32_t *side) {
    if (ndt) *ndt = (float)c->ndt;
    if (side) *side = c->cfl_arg;
    return 0;
}

int32_t nhf_sync(nhf_ctx *c) {
    (void)c;
    return 0;
}

int32_t nhf_set_gates(nhf_ctx *c, const nhf_gates *gates) {
    free(c->gate_budget);
    free(c->gate_volume);
    free(c->gate_sources);
    c->gate_budget = NULL;
    c->gate_volume = c->gate_sources = NULL;
    c->gates_on = 0;
    if (!gates) return 0;
    if (gates->count < 1 || !(gates->opening_seconds > 0.0f)) return fail("nhf_set_gates: invalid counts or opening time");
    c->g = *gates;
    c->gate_budget = (ddsum *)alloc((size_t)(10 * gates->count + 8), sizeof(ddsum));
    c->gate_volume = (real *)alloc(c->m.s_num, sizeof(real));
    c->gate_sources = (real *)alloc((size_t)c->m.e_num * 6, sizeof(real));
    c->gates_on = 1;
    return 0;
}

int32_t nhf_gate_check(nhf_ctx *c) {
    if (!c->gates_on) return fail("nhf_gate_check: no gates bound");
    sync_host_writes(c);
    gate_check(c);
    store_state(c);
    return 0;
}

int32_t nhf_flush_gate_budget(nhf_ctx *c, double *out) {
    if (!c->gates_on) return fail("nhf_flush_gate_budget: no gates bound");
    for (int t = 0; t < 10 * c->g.count + 8; ++t) {
        out[2 * t] = c->gate_budget[t].hi;
        out[2 * t + 1] = c->gate_budget[t].lo;
        c->gate_budget[t].hi = c->gate_budget[t].lo = 0.0;
    }
    return 0;
}

/* The oracle writes max_depth and peak_tick on every rise, so there is nothing to settle. */
int32_t nhf_settle_extrema(nhf_ctx *c) {
    (void)c;
    return 0;
}
/* The oracle always keeps the per-cell rain time and capacity arrays. */
int32_t nhf_set_rain_state(nhf_ctx *c, int32_t uniform, float rain_time, const float *class_capacity) {
    (void)c;
    (void)uniform;
    (void)rain_time;
    (void)class_capacity;
    return 0;
}
int32_t nhf_set_graph(nhf_ctx *c, int32_t enable) {
    (void)c;
    (void)enable;
    return 0;
}
int32_t nhf_set_timing(nhf_ctx *c, int32_t enable) {
    (void)c;
    (void)enable;
    return 0;
}
int32_t nhf_kernel_times(nhf_ctx *c, double *ms, int64_t *count, int32_t n, int32_t reset) {
    (void)c;
    (void)reset;
    for (int i = 0; i < n; ++i) {
        ms[i] = 0;
        count[i] = 0;
    }
    return 0;
}

/* ------------------------------------------------------------------ oracle-only test hooks
 * Not part of include/nhflood.h: they expose single reference functions so the tests can pin
 * them against the reference's own known answers. */

int32_t nhf_oracle_face_discharge(const float *q, const float *lengths, const int32_t *ids,
                                  int32_t start, int32_t count, float *out) {
    nhf_ctx c;
    memset(&c, 0, sizeof c);
    int32_t n_sides = 0;
    for (int k = 0; k < count; ++k)
        if (ids[start + k] + 1 > n_sides) n_sides = ids[start + k] + 1;
    real *qq = (real *)alloc(n_sides, sizeof(real));
    for (int k = 0; k < count; ++k) qq[ids[start + k]] = q[ids[start + k]];
    c.q = qq;
    c.m.isl = ids;
    c.m.length = lengths;
    *out = (float)face_discharge(&c, start, count);
    free(qq);
    return 0;
}

int32_t nhf_oracle_neighbor_sum(const float *cases, int32_t n, float *out) {
    for (int i = 0; i < n; ++i) {
        const float *v = &cases[5 * i];
        out[i] = (float)boundary_neighbor_sum(v[0], v[1], v[2], v[3], v[4]);
    }
    return 0;
}

int32_t nhf_oracle_theta_update(const float *cases, int32_t n, float theta, float theta_half,
                                float *out) {
    /* (theta q + (1 - theta)/2 N + dq) / f with N from boundary_neighbor_sum, as in tick */
    for (int i = 0; i < n; ++i) {
        const float *v = &cases[7 * i];
        real nb = boundary_neighbor_sum(v[0], v[1], v[2], v[3], v[4]);
        out[i] = (float)(((real)theta * v[0] + (real)theta_half * nb + (real)v[5]) / (real)v[6]);
    }
    return 0;
}

/* Pressure reconstruction on a bound mesh: update every gradient cell from the mesh's depth
 * and bed, then correct `raw` on the listed sides. */
int32_t nhf_oracle_pressure(const nhf_mesh *mesh, float min_h, int32_t n_probe,
                            const int32_t *sides, const float *raw, float *gradients_out,
                            float *difference_out) {
    nhf_ctx *c = NULL;
    nhf_params p;
    memset(&p, 0, sizeof p);
    p.min_h = min_h;
    p.batch_capacity = 1;
    if (nhf_create(&c, mesh, &p, NULL, 0, NULL)) return -1;
    for (int i = 0; i < mesh->e_num; ++i) {
        c->depth[i] = mesh->depth[i];
        c->bed[i] = mesh->bed[i];
    }
    update_pressure(c);
    for (int i = 0; i < 2 * mesh->n_grad; ++i) gradients_out[i] = (float)c->gradients[i];
    for (int i = 0; i < n_probe; ++i) {
        int s = sides[i];
        difference_out[i] = (float)pressure_difference(c, s, mesh->side_cells[2 * s],
                                                       mesh->side_cells[2 * s + 1], raw[i]);
    }
    nhf_destroy(c);
    return 0;
}

/* Error-free accumulation of F32 values into one double-double, as the budget terms are added. */
int32_t nhf_oracle_dd_sum(const float *values, int64_t n, double *out) {
    ddsum acc = {0.0, 0.0};
    for (int64_t i = 0; i < n; ++i)
        if (values[i] != 0.0f) dd_add(&acc, (double)values[i]);
    out[0] = acc.hi;
    out[1] = acc.lo;
    return 0;
}

int32_t nhf_oracle_math(const float *x, const float *y, int32_t n, float *pow_out, float *exp_out) {
    for (int i = 0; i < n; ++i) {
        pow_out[i] = det_powf(x[i], y[i]);
        exp_out[i] = det_expf(y[i]);
    }
    return 0;
}

/* ------------------------------------------------------------------------ flood-map raster */

/* rasterize_tile_kernel (output/flood_map.py:180-222) over host arrays, one cell after another in
 * cell order: the F32 bounding-box test against the tile, then the pixel range
 *   columns ceil((min_x - left) / hr - 0.5) .. ceil((max_x - left) / hr - 0.5) - 1,
 *   rows    ceil((top - max_y) / vr - 0.5) .. ceil((top - min_y) / vr - 0.5) - 1,
 * each operation rounded to F32 as written, clamped to the tile. Always F32, whatever REAL is:
 * the maps are F32 in the reference. */
int32_t nhf_rasterize(const float *x, const float *y, const float *half_w, const float *half_h, const float *value,
                      int32_t n, float tile_min_x, float tile_max_y, float hr, float vr, int32_t rows, int32_t cols,
                      float *tile, int64_t ld, void *stream) {
    (void)stream;
    if (n < 0 || rows < 0 || cols < 0 || ld < cols) return fail("nhf_rasterize: negative size or ld < cols");
    if (!(hr > 0.0f) || !(vr > 0.0f)) return fail("nhf_rasterize: pixel sizes must be positive");
    if (n == 0 || rows == 0 || cols == 0) return 0;
    const float tile_max_x = tile_min_x + (float)cols * hr;
    const float tile_min_y = tile_max_y - (float)rows * vr;
    for (int32_t i = 0; i < n; ++i) {
        const float e_min_x = x[i] - half_w[i], e_max_x = x[i] + half_w[i];
        const float e_min_y = y[i] - half_h[i], e_max_y = y[i] + half_h[i];
        if (!(e_max_x >= tile_min_x && e_min_x < tile_max_x && e_max_y >= tile_min_y && e_min_y < tile_max_y))
            continue;
        const float cl = (e_min_x - tile_min_x) / hr, ch = (e_max_x - tile_min_x) / hr;
        const float rl = (tile_max_y - e_max_y) / vr, rh = (tile_max_y - e_min_y) / vr;
        int32_t c0 = (int32_t)ceilf(cl - 0.5f), c1 = (int32_t)(ceilf(ch - 0.5f) - 1.0f);
        int32_t r0 = (int32_t)ceilf(rl - 0.5f), r1 = (int32_t)(ceilf(rh - 0.5f) - 1.0f);
        if (c0 < 0) c0 = 0;
        if (c1 > cols - 1) c1 = cols - 1;
        if (r0 < 0) r0 = 0;
        if (r1 > rows - 1) r1 = rows - 1;
        for (int32_t r = r0; r <= r1; ++r)
            for (int32_t c = c0; c <= c1; ++c) tile[(int64_t)r * ld + c] = value[i];
    }
    return 0;
}
