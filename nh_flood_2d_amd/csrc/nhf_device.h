// Device code shared by the single-mesh step (nhflood.hip) and the scenario-batched step
// (nhbatch.hip): constants, the clock, deterministic x^y / e^x, wave reductions and the
// per-context control block. Everything here is static or inline: each translation unit gets
// its own copy of the device functions.
#ifndef NHF_DEVICE_H
#define NHF_DEVICE_H

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "nhflood.h"

#define NHF_BLOCK 256
#define NHF_WAVES (NHF_BLOCK / 64)
#define TICKS_PER_SECOND_F 1099511627776.0f  // 2^40 exactly representable in F32
#define SNAP_TICKS 1099512ull                 // ceil(1e-6 s * 2^40)

// Thread-local text of the last error, defined once in nhflood.hip.
extern thread_local char nhf_error_text[512];

static inline int32_t set_error(const char *what, hipError_t err) {
    snprintf(nhf_error_text, sizeof nhf_error_text, "%s: %s", what, hipGetErrorString(err));
    return -1;
}

static inline int32_t set_error_text(const char *what) {
    snprintf(nhf_error_text, sizeof nhf_error_text, "%s", what);
    return -2;
}

#define HIP_TRY(expr)                                      \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return set_error(#expr, e_); \
    } while (0)

// Name of the first required mesh array that is NULL, or NULL when all are present. `state`
// also checks the per-cell and per-side state arrays (the single-mesh context reads them from
// the mesh; a batch has its own).
static inline const char *missing_mesh_array(const nhf_mesh *m, bool state, bool manning) {
    struct Field {
        const void *p;
        const char *name;
    } fields[] = {{m->bed, "bed"}, {m->area, "area"}, {m->size, "size"}, {m->cell_meta, "cell_meta"},
                  {m->class_horton, "class_horton"}, {m->bweight, "bweight"}, {m->csr_start, "csr_start"},
                  {m->csr_count, "csr_count"}, {m->isl, "isl"}, {m->isl_cell, "isl_cell"},
                  {m->side_cells, "side_cells"}, {m->side_meta, "side_meta"}, {m->length, "length"},
                  {m->sdc, "sdc"}, {m->side_key, "side_key"}, {m->kind, "kind"},
                  {m->cell_quad, "cell_quad"}};
    for (const Field &f : fields)
        if (!f.p) return f.name;
    if (m->n_pixels > 1 && !m->rain_index) return "rain_index";
    if (manning && !m->manning) return "manning";
    if (state) {
        const Field st[] = {{m->depth, "depth"}, {m->rain_time, "rain_time"}, {m->capacity, "capacity"},
                            {m->max_depth, "max_depth"}, {m->max_speed, "max_speed"}, {m->peak_tick, "peak_tick"},
                            {m->q, "q"}};
        for (const Field &f : st)
            if (!f.p) return f.name;
    }
    return nullptr;
}

// --------------------------------------------------------------------------- side lists

// On a 2:1 quadtree a cell lists at most 8 sides (two per face). The cell passes load such a list
// with all of its loads issued together (ids, then every side's values) instead of one dependent
// load chain per entry; longer lists take a plain loop.
#define NHF_LIST_MAX 8

// Side count of face d (0 left, 1 right, 2 bottom, 3 top) from a cell's packed face counts.
static __device__ __forceinline__ int face_n(uint32_t counts, int d) { return (int)((counts >> (8 * d)) & 255u); }

// Every entry is loaded, past the end of the list from its first entry (a real cell lists at least
// one side), so the loads carry no branch and are issued together.
static __device__ __forceinline__ void load_list(const int32_t *isl, int k0, int total, int ids[NHF_LIST_MAX]) {
#pragma unroll
    for (int j = 0; j < NHF_LIST_MAX; ++j) ids[j] = isl[k0 + (j < total ? j : 0)];
}

// A cell's side list as the cell passes read it: one 16-byte record per cell gives the four side
// ids of a regular cell (one side per face, every side as long as the cell is wide, area equal to
// size*size in F32, all checked when the record is built), which needs no further index, length
// or area load; any other cell's record points into the packed lists and holds the list length.
// Both forms use all four words, so the record is read with one 16-byte load. `total` entries (at most
// NHF_LIST_MAX on the short path) and the running face counts come back; `short_list` is false for
// a list that needs the plain loop.
struct CellList {
    int ids[NHF_LIST_MAX];
    int total, n0, n01, n012, k0;
    uint32_t cnt;
    bool short_list, regular;
};

static __device__ __forceinline__ int4 cell_record(const nhf_mesh &m, int cell) {
    return reinterpret_cast<const int4 *>(m.cell_quad)[cell];
}

// REGULAR: the caller knows the record is a regular cell's, so no list code is generated.
template <bool REGULAR = false>
static __device__ __forceinline__ void cell_list(const nhf_mesh &m, const int4 rec, CellList &l) {
    if (REGULAR || rec.x >= 0) {
        l.ids[0] = rec.x;
        l.ids[1] = rec.y;
        l.ids[2] = rec.z;
        l.ids[3] = rec.w;
#pragma unroll
        for (int j = 4; j < NHF_LIST_MAX; ++j) l.ids[j] = rec.x;
        l.total = 4;
        l.n0 = 1;
        l.n01 = 2;
        l.n012 = 3;
        l.k0 = -1;
        l.cnt = 0x01010101u;
        l.short_list = true;
        l.regular = true;
        return;
    }
    l.k0 = rec.y;
    l.cnt = (uint32_t)rec.z;
    l.n0 = face_n(l.cnt, 0);
    l.n01 = l.n0 + face_n(l.cnt, 1);
    l.n012 = l.n01 + face_n(l.cnt, 2);
    l.total = rec.w;
    l.short_list = l.total <= NHF_LIST_MAX;
    l.regular = false;
    if (l.short_list) load_list(m.isl, l.k0, l.total, l.ids);
}

// One value per entry of a short list, load(side) for each. The first four loads carry no
// condition (entries past the end name the list's first side), the other four are made only when
// the list is longer than four; entries past the end are not to be used.
template <class Load>
static __device__ __forceinline__ void gather_list(const CellList &l, Load load, float out[NHF_LIST_MAX]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = load(l.ids[j]);
    if (l.total > 4) {
#pragma unroll
        for (int j = 4; j < NHF_LIST_MAX; ++j) out[j] = load(l.ids[j]);
    } else {
#pragma unroll
        for (int j = 4; j < NHF_LIST_MAX; ++j) out[j] = 0.0f;
    }
}

// Side lengths of a short list: the cell's size for a regular cell, else the stored lengths.
static __device__ __forceinline__ void list_lengths(const nhf_mesh &m, const CellList &l, float size,
                                                    float len[NHF_LIST_MAX]) {
    if (l.regular) {
#pragma unroll
        for (int j = 0; j < NHF_LIST_MAX; ++j) len[j] = j < 4 ? size : 0.0f;
    } else {
        gather_list(l, [&](int side) { return m.length[side]; }, len);
    }
}

// Face (0 left, 1 right, 2 bottom, 3 top) of list entry j, from the running face counts.
static __device__ __forceinline__ int entry_face(int j, int n0, int n01, int n012) {
    return j < n0 ? 0 : (j < n01 ? 1 : (j < n012 ? 2 : 3));
}

// --------------------------------------------------------------------------- control block

// One scenario's clock, step scalars, limiter counters, CFL-limit capture and batch trace.
struct Ctl {
    uint64_t now, next, origin, step_end;
    float dt, local, end_local, ndt;
    int32_t active, cfl_arg, raining, slot;
    uint32_t limited_cells, limited_faces, min_scale_bits, pad0;
    float limit_value[5];
    int32_t limit_ids[3];
    uint32_t trace[NHF_MAX_BATCH * NHF_TRACE_COLS];
};

// Comparisons written as selects so +0/-0 and NaN behave the same in every pass.
static __device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
static __device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }

// Workgroups are dealt to the 8 XCDs round-robin. For a grid that is a multiple of 8, this maps
// the hardware order so each XCD runs contiguous runs of logical blocks: Morton-near cells then
// share one XCD's L2 instead of being spread over all eight. The runs are chunks of NHF_XCD_CHUNK
// consecutive blocks per XCD in turn (0: one run of an eighth of the grid per XCD). Measured, 32-mesh union: 29.84 G cell-updates/s at 256 against 29.15 G with eighths
// (16 and 64: 29.80 and 29.81 G); single mesh and 64-slot batch the same within noise.
#ifndef NHF_XCD_CHUNK
#define NHF_XCD_CHUNK 256
#endif
static __device__ __forceinline__ int xcd_block() {
    const int per = (int)(gridDim.x >> 3);
#if NHF_XCD_CHUNK > 0
    // XCD x takes chunks of NHF_XCD_CHUNK consecutive blocks in turn with the other XCDs; the
    // tail that does not fill a chunk on every XCD is split into eight runs
    const int x = (int)(blockIdx.x & 7u), j = (int)(blockIdx.x >> 3);
    const int full = per / NHF_XCD_CHUNK * NHF_XCD_CHUNK;
    if (j < full) return ((j / NHF_XCD_CHUNK) * 8 + x) * NHF_XCD_CHUNK + j % NHF_XCD_CHUNK;
    return full * 8 + x * (per - full) + (j - full);
#else
    return (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
#endif
}

static __device__ __forceinline__ int meta_boundary(uint32_t meta) { return (int)((meta >> 8) & 255u) - 1; }
static __device__ __forceinline__ int meta_class(uint32_t meta) { return (int)(meta & 255u); }

// A cell's size: size_base * 2^level from its meta word when the mesh is of that form (exact: a
// power-of-two scaling), else the stored value.
static __device__ __forceinline__ float cell_size(const nhf_mesh &m, uint32_t meta, int cell) {
    return m.size_base > 0.0f ? ldexpf(m.size_base, (int)((meta >> 16) & 15u)) : m.size[cell];
}

// --------------------------------------------------------------------------- x^y and e^x
// Evaluated in FP64 and rounded once to F32, so the result is the same on any IEEE machine:
// log2 by range reduction to m in [sqrt(1/2), sqrt(2)) and the odd series of
// ln m = 2 atanh((m-1)/(m+1)); 2^t by t = n + f, |f| <= 1/2, and the Taylor series of
// e^(f ln 2) (J.-M. Muller, Elementary Functions, ch. 5). Only IEEE add, multiply, divide and
// explicit fma are used.
static __device__ __forceinline__ double det_log2(double x) {
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

static __device__ __forceinline__ double det_exp2(double t) {
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

static __device__ __forceinline__ float det_powf(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x == 1.0f) return 1.0f;
    return (float)det_exp2((double)y * det_log2((double)x));
}

static __device__ __forceinline__ float det_expf(float x) { return (float)det_exp2((double)x * 1.44269504088896340736); }

// --------------------------------------------------------------------------- block reductions

static __device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

static __device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

static __device__ __forceinline__ uint32_t wave_min_u(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// (key, index) minimum: the key orders by CFL bits, then by the reference side id.
static __device__ __forceinline__ void wave_min_key(uint64_t &key, int32_t &idx) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t k2 = __shfl_xor((unsigned long long)key, o, 64);
        int32_t i2 = __shfl_xor(idx, o, 64);
        if (k2 < key) {
            key = k2;
            idx = i2;
        }
    }
}

// --------------------------------------------------------------------------- clock

static __device__ __forceinline__ float ticks_local(uint64_t ticks) {
    uint32_t hi = (uint32_t)(ticks >> 32), lo = (uint32_t)ticks;
    return (float)hi * (1.0f / 256.0f) + (float)lo * (1.0f / TICKS_PER_SECOND_F);
}

// Exact F32 seconds -> 2^-40 s ticks, half-even below one tick; 0 when not representable.
static __device__ __forceinline__ uint64_t f32_ticks(float v) {
    uint32_t bits = __float_as_uint(v);
    int e = (int)((bits >> 23) & 255u) - 127;
    if (!(v > 0.0f) || e < -40 || e > 0) return 0;
    uint64_t mant = (uint64_t)((bits & 0x7fffffu) | 0x800000u);
    int shift = e + 17;
    if (shift >= 0) return mant << shift;
    int right = -shift;
    uint64_t t = mant >> right;
    uint64_t dropped = mant & ((1ull << right) - 1ull);
    uint64_t half = 1ull << (right - 1);
    if (dropped > half || (dropped == half && (t & 1ull))) t += 1;
    return t;
}

// StepClock.choose (core/clock.py:94-167) with u64 tick arithmetic. Writes the trace row's
// status / candidate / event-short columns and the step's dt, local times and end tick.
static __device__ void clock_choose(Ctl *c, float candidate, float min_dt, int slot) {
    uint32_t *row = &c->trace[slot * NHF_TRACE_COLS];
    c->active = 0;
    c->dt = 0.0f;
    c->local = 0.0f;
    c->end_local = 0.0f;
    uint64_t now = c->now, target = c->next;
    if (!(now < target)) return;
    if (!(candidate >= min_dt)) {
        row[3] = NHF_BAD_CFL;
        return;
    }
    uint64_t rem = target - now;
    float remaining = ticks_local(rem);
    float capped = fmn(candidate, 1.0f);
    float chosen = fmn(capped, remaining);
    row[4] = __float_as_uint(candidate);
    row[5] = chosen < capped ? 1u : 0u;
    uint64_t dt_ticks = f32_ticks(chosen);
    if (dt_ticks == 0) {
        row[3] = NHF_BAD_DT;
        return;
    }
    bool exceeds = dt_ticks > rem;
    uint64_t diff = exceeds ? dt_ticks - rem : rem - dt_ticks;
    uint64_t end;
    if (diff < SNAP_TICKS) {
        end = target;
    } else if (exceeds) {
        row[3] = NHF_OVERSHOOT;
        return;
    } else {
        end = now + dt_ticks;
    }
    c->local = ticks_local(now - c->origin);
    c->end_local = ticks_local(end - c->origin);
    c->dt = chosen;
    c->step_end = end;
    c->active = 1;
}

// finish_step + StepClock.commit (solver_compact.py:384-391, core/clock.py:173-181).
static __device__ void commit_slot(Ctl *c, int slot) {
    if (!c->active) return;
    uint32_t *row = &c->trace[slot * NHF_TRACE_COLS];
    row[6] = c->limited_cells;
    row[7] = c->limited_faces;
    row[8] = c->min_scale_bits;
    c->now = c->step_end;
    row[0] = __float_as_uint(c->dt);
    row[1] = (uint32_t)(c->step_end >> 32);
    row[2] = (uint32_t)c->step_end;
    row[3] = NHF_COMPLETED;
}

// Linear continuation of missing theta neighbours at prescribed faces
// (boundary_neighbor_sum, solver_compact.py:135-149).
static __device__ __forceinline__ float neighbour_sum(float q, float left, float right, float fl, float fr) {
    float result = left + right;
    if (fl > 0.0f || fr > 0.0f) {
        if (fl == 1.0f || fr == 1.0f) {
            result = 2.0f * q;
        } else {
            float product = fl * fr;
            result = ((1.0f - fr) * left + (1.0f - fl) * right + 2.0f * q * (fl + fr - 2.0f * product)) /
                     (1.0f - product);
        }
    }
    return result;
}

// The donor certificate (k_donor's test, solver_compact.py:491-514). A cell's donor scale is
// below 1 only when the volume it sends out exceeds what it holds. With no boundary or pipe term
// and at most NHF_LIST_MAX sides, each no longer than the cell is wide (a 2:1 quadtree of
// squares), that cannot happen while every side the cell donates to carries
// |q| dt <= NHF_DONOR_SHARE h len, h the cell's depth: the outgoing sum is then at most
// 8 x 0.12 h area, with F32 rounding, below the available volume of at least h area. The face
// pass flags the donor of any side that breaks that bound, or whose bound is too small to hold
// through rounding (below 2^-100); the donor pass looks at flagged cells only, and every other
// cell keeps scale 1 exactly. The donor of a side is its lower cell when q >= 0.
#define NHF_DONOR_SHARE 0.12f

static __device__ __forceinline__ bool donor_share_exceeded(float nq, float dt, float h_donor, float len) {
    if (nq == 0.0f) return false;  // a dry face, or no flow: the donor sends nothing
    const float bound = NHF_DONOR_SHARE * h_donor * len;
    return !(fabsf(nq) * dt <= bound && bound >= 0x1p-100f);
}

// Neumaier step on a (sum, compensation) pair.
static __device__ __forceinline__ void acc_add(double2 &a, double v) {
    double s = a.x + v;
    if (fabs(a.x) >= fabs(v))
        a.y += (a.x - s) + v;
    else
        a.y += (v - s) + a.x;
    a.x = s;
}

#endif  // NHF_DEVICE_H
