// libnhflood.so: the compact local-inertial step as hand-written HIP kernels for gfx950.
//
// One time step is five launches on the context's stream (plus two optional ones):
//   k_gradients   hanging-edge tangential water-level gradients   (core/pressure.py:108-116)
//   k_face_avg    theta < 1 face averages of the previous q         (solver_compact.py:430-436)
//   k_flux        face flux: theta-weighted discharge, semi-implicit Manning friction, min_h
//                 gating, exterior zeroing, hanging pressure correction (solver_compact.py:439-486)
//   k_donor       positivity limiter: outgoing vs available volume, then the shared-subedge
//                 scaling of every side this cell is the donor of (solver_compact.py:491-524)
//   k_depth       sources, Horton capacity, face totals, continuity, infiltration removal,
//                 positivity floor, level cells, extrema and the per-wavefront FP64 water
//                 budget                                               (solver_compact.py:525-613)
//   k_cfl         CFL bound per interior side, wave-shuffle + LDS min/argmin per workgroup
//                                                                      (solver_compact.py:617-641)
//   k_finish      one workgroup (per part) reduces the partials, commits the clock and prepares
//                 the next slot                (solver_compact.py:364-391; core/clock.py:94-181)
// k_finish stays a launch of its own: a last-workgroup-done reduction inside k_cfl needs an
// agent-scope release per workgroup, an L2 write-back on each XCD, and made k_cfl 2.4x slower
// (also with the acquire, an L2 invalidate, left to the last workgroup alone: 14.0 -> 36.8 us).
// The clock, dt and every reduction stay on the device; the host reads one trace per batch.
//
// Layout: structure of arrays, cells and sides renumbered along a Morton curve by the host so a
// wavefront's neighbour gathers stay within a few cache lines. Cell 0 / side 0 are sentinels.
// The path is HBM-bandwidth bound (about 0.2 flop/byte); no MFMA.
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "nhf_device.h"

// --------------------------------------------------------------------------- errors

thread_local char nhf_error_text[512];

// --------------------------------------------------------------------------- device state

// Parts: several meshes concatenated into one (nhu_create), each with its own clock, forcing and
// budget. Every part's cells start at 1 + 1024 k and its sides likewise, so a workgroup of a cell
// or side pass lies in one part and finds it with one load.
#define NHF_PART_ALIGN 1024
struct PartDev {
    int32_t count;                // 0: one mesh (nhf_create)
    int32_t pad;
    const int32_t *cell_chunk;    // part of cells 1 + 1024 k .. 1024 (k + 1)
    const int32_t *side_chunk;    // part of sides 1 + 1024 k .. 1024 (k + 1)
    const int4 *range;            // per part: cell begin, cell end, side begin, side end
    const int4 *extra;            // per part: first boundary, boundaries, first CFL workgroup, CFL workgroups
    const int32_t *cfl_part;      // part of each CFL workgroup
};

struct Dev {
    nhf_mesh m;
    nhf_params p;
    Ctl *ctl;  // one per part
    float *coefficients, *values, *level_values, *rain_rates;
    float *donor, *face_qn, *grad;
    uint64_t *cfl_key;
    int32_t *cfl_idx;
    double2 *acc;  // per depth-block, per term: (sum, compensation)
    int32_t terms, cfl_blocks, depth_blocks, theta_one, cfl_max;
    int32_t cell_block;  // workgroup size of the two cell passes (128, 256 or 512)
    int32_t cells_per_thread;  // cells per thread of the two cell passes (1 or 2)
    int32_t acc_owners;  // budget accumulators allocated (wavefronts of the largest grid)
    int32_t gates_on;    // gate control bound (nhf_set_gates)
    // Per cell: nonzero when the donor pass must look at the cell (see donor_check): bit 0 set by
    // the flux pass for a cell one of whose outflows exceeds its certified share, bit 1 always set
    // for cells the certificate does not cover (boundary cells, long side lists, pipe exchange).
    uint8_t *donor_flag;
    int32_t donor_keep;  // 1: k_donor leaves the flags as they are (the per-kernel timing pass)
    // Uniform rain state of a part with one rain pixel (rain_uniform[part] = 1): every cell's rain
    // time is rt_u[part] and its capacity cap_u[class], so the depth pass neither reads nor
    // writes the per-cell arrays (k_settle_extrema writes them out); the flux pass advances
    // rt_u and the capacity of the part's classes (class_part) once per step (advance_rain).
    int32_t *rain_uniform, *rain_pixel_u;  // [parts]
    float *rt_u;                           // [parts]
    float *cap_u, *cap_stage;              // [256 classes]
    int32_t *class_part;                   // [256]: the part whose cells have the class, -1 none
    nhf_gates g;
    PartDev parts;
};

// The cells the donor certificate (donor_check) does not cover: the donor pass looks at them
// every step.
static __device__ __forceinline__ uint8_t donor_always(const nhf_mesh &m, uint32_t meta, int4 rec) {
    return (meta_boundary(meta) >= 0 || m.pipe_q || (rec.x < 0 && rec.w > NHF_LIST_MAX)) ? 2 : 0;
}

// The part of a workgroup whose first cell (side) is `first`; 0 for a single mesh.
template <bool PARTS>
static __device__ __forceinline__ int cell_part(const Dev &d, int first) {
    return PARTS ? d.parts.cell_chunk[(first - 1) >> 10] : 0;
}
template <bool PARTS>
static __device__ __forceinline__ int side_part(const Dev &d, int first) {
    return PARTS ? d.parts.side_chunk[(first - 1) >> 10] : 0;
}

// Boundaries [base, base + count) of a part: all of them for a single mesh.
static __device__ __forceinline__ int2 part_boundaries(const Dev &d, int part) {
    if (d.parts.count == 0) return make_int2(0, d.m.b_num);
    const int4 e = d.parts.extra[part];
    return make_int2(e.x, e.y);
}

// prepare_step (solver_compact.py:364-382) of one part, run by one thread.
__device__ void prepare_slot(const Dev &d, int part, int slot) {
    Ctl *c = d.ctl + part;
    clock_choose(c, c->ndt, d.p.min_dt, slot);
    if (!c->active) return;
    c->limited_cells = 0;
    c->limited_faces = 0;
    c->min_scale_bits = __float_as_uint(1.0f);
    const float dt = c->dt, local = c->local, end_local = c->end_local;
    const int2 bounds = part_boundaries(d, part);
    for (int bi = bounds.x; bi < bounds.x + bounds.y; ++bi) {
        float c0 = d.coefficients[2 * bi], c1 = d.coefficients[2 * bi + 1];
        d.values[bi] = c0 + c1 * (local + dt / 2.0f);
        d.level_values[bi] = c0 + c1 * end_local;
    }
    if (c->ndt < c->limit_value[0]) {
        int si = c->cfl_arg;
        int el = d.m.side_cells[2 * si], eh = d.m.side_cells[2 * si + 1];
        float zl = d.m.bed[el], zh = d.m.bed[eh], hl = d.m.depth[el], hh = d.m.depth[eh];
        float head = fmx(zl + hl, zh + hh);
        float h = fmx(head - fmx(fmx(zl, zh), -999.0f), 0.01f);
        c->limit_value[0] = c->ndt;
        c->limit_value[1] = hl;
        c->limit_value[2] = hh;
        c->limit_value[3] = h;
        c->limit_value[4] = d.m.q[si] / fmx(h, 1e-30f);
        c->limit_ids[0] = si;
        c->limit_ids[1] = el;
        c->limit_ids[2] = eh;
    }
}

// The rain state a step of a uniform part advances to, by one wavefront (lane l takes classes
// l, l + 64, ...): where the part's one pixel rains, every cell's rain time becomes rt_u + dt and
// its capacity its class's Horton value there, the depth pass's arithmetic
// (solver_compact.py:543-548) once per class instead of per cell. The first wavefront of the
// part's flux pass runs it, off the path between steps; the depth pass of the same step reads it.
// Lane 0 stores the new rain time with a store that depends on the wavefront's one load of it.
__device__ void advance_rain(const Dev &d, int part, const Ctl *c, int lane) {
    if (!d.rain_uniform[part] || !c->raining || !(d.rain_rates[d.rain_pixel_u[part]] > 0.0f)) return;
    const float rt = d.rt_u[part] + c->dt;
    for (int k = lane; k < 256; k += 64) {
        if (d.class_part[k] != part) continue;
        const float *hp = d.m.class_horton + 3 * k;
        const float f0 = hp[0], fc = hp[1], hk = hp[2];
        d.cap_u[k] = fc + (f0 - fc) * det_expf(-hk * rt / 3600.0f);
    }
    if (lane == 0) d.rt_u[part] = rt;
}

// --------------------------------------------------------------------------- small kernels

// A part's cells [begin, padded end) and budget accumulators [first, last): the whole mesh for a
// single mesh. Accumulator k belongs to the wavefront whose first cell is 1 + 64 k.
static __device__ __forceinline__ int4 part_span(const Dev &d, int part) {
    if (d.parts.count == 0) return make_int4(0, d.m.e_num, 0, d.acc_owners);
    const int4 r = d.parts.range[part];
    const int end = r.x + ((r.y - r.x + NHF_PART_ALIGN - 1) / NHF_PART_ALIGN) * NHF_PART_ALIGN;
    return make_int4(r.x, end, (r.x - 1) >> 6, (end - 1) >> 6);
}

__global__ void k_reset(Dev d, int part, uint64_t now) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int stride = gridDim.x * blockDim.x;
    if (tid == 0) {
        Ctl *c = d.ctl + part;
        c->now = c->next = c->origin = c->step_end = now;
        c->dt = c->local = c->end_local = 0.0f;
        c->active = 0;
        c->slot = 0;
        c->limited_cells = c->limited_faces = 0;
        c->min_scale_bits = __float_as_uint(1.0f);
        c->limit_value[0] = INFINITY;
    }
    const int4 span = part_span(d, part);
    for (int i = span.z * d.terms + tid; i < span.w * d.terms; i += stride) d.acc[i] = make_double2(0.0, 0.0);
    for (int i = span.x + tid; i < span.y; i += stride) d.donor[i] = 0.0f;
    // the donor flags of the part's real cells: their standing values (padding cells stay 0)
    const int cell_end = d.parts.count ? d.parts.range[part].y : d.m.e_num;
    for (int i = (span.x > 0 ? span.x : 1) + tid; i < span.y; i += stride)
        d.donor_flag[i] = i < cell_end ? donor_always(d.m, d.m.cell_meta[i], cell_record(d.m, i)) : 0;
    if (d.face_qn)
        for (int i = 4 * span.x + tid; i < 4 * span.y; i += stride) d.face_qn[i] = 0.0f;
}

// Per context, before k_init_flags: no class has a part yet, and every part starts in the
// per-cell rain state.
__global__ void k_init_classes(Dev d, int parts) {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) d.class_part[k] = -1;
    for (int p = threadIdx.x; p < parts; p += blockDim.x) {
        d.rain_uniform[p] = 0;
        d.rt_u[p] = 0.0f;
        d.rain_pixel_u[p] = 0;
    }
}

// The standing donor flags of every real cell, and the part of each land class, once per context
// (k_reset sets a part's flags again).
__global__ __launch_bounds__(NHF_BLOCK) void k_init_flags(Dev d) {
    for (int cell = blockIdx.x * NHF_BLOCK + threadIdx.x + 1; cell < d.m.e_num; cell += gridDim.x * NHF_BLOCK) {
        const int part = d.parts.count ? d.parts.cell_chunk[(cell - 1) >> 10] : 0;
        const bool real = !d.parts.count || cell < d.parts.range[part].y;
        d.donor_flag[cell] = real ? donor_always(d.m, d.m.cell_meta[cell], cell_record(d.m, cell)) : 0;
        if (real) d.class_part[meta_class(d.m.cell_meta[cell])] = part;  // a class lies in one part
    }
}

// Enter (uniform != 0) or leave the uniform rain state of one part: `rt` the rain time of every
// cell, cap_stage the capacity of each class code (one workgroup).
__global__ void k_set_rain_state(Dev d, int part, int uniform, float rt, int pixel) {
    if (uniform)
        for (int k = threadIdx.x; k < 256; k += blockDim.x)
            if (d.class_part[k] == part) d.cap_u[k] = d.cap_stage[k];
    if (threadIdx.x == 0) {
        d.rain_uniform[part] = uniform ? 1 : 0;
        d.rt_u[part] = rt;
        d.rain_pixel_u[part] = pixel;
    }
}

__global__ void k_begin(Dev d, int part, uint64_t target, uint64_t origin) {
    Ctl *c = d.ctl + part;
    for (int i = threadIdx.x; i < NHF_MAX_BATCH * NHF_TRACE_COLS; i += blockDim.x) c->trace[i] = 0u;
    if (threadIdx.x == 0) {
        c->next = target;
        c->origin = origin;
        c->slot = 0;
    }
}

__global__ void k_set_raining(Dev d, int part, int raining) {
    if (threadIdx.x == 0) d.ctl[part].raining = raining;
}

// First slot of a batch: clear the trace rows, then prepare slot 0 (workgroup p: part p).
__global__ void k_prepare(Dev d, int slot) {
    const int part = blockIdx.x;
    for (int i = threadIdx.x; i < NHF_MAX_BATCH * NHF_TRACE_COLS; i += blockDim.x) d.ctl[part].trace[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) prepare_slot(d, part, slot);
}

// The interval budget of nhf_flush_budget reduced on the device (nhf_run_static, which has no
// host buffer to take the accumulators to): workgroup t sums term t's per-wavefront (sum,
// compensation) pairs in wavefront order with the host's two-sum, staged through LDS a workgroup
// at a time, so the result has the host flush's bits. A non-finite entry leaves a non-finite
// total (a non-finite partial sum stays non-finite under these additions).
__global__ __launch_bounds__(NHF_BLOCK) void k_flush_terms(Dev d, double *out) {
    const int t = blockIdx.x;
    __shared__ double2 s_pair[NHF_BLOCK];
    double hi = 0.0, lo = 0.0;
    for (int base = 0; base < d.acc_owners; base += NHF_BLOCK) {
        const int b = base + (int)threadIdx.x;
        s_pair[threadIdx.x] = b < d.acc_owners ? d.acc[(size_t)b * d.terms + t] : make_double2(0.0, 0.0);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = d.acc_owners - base < NHF_BLOCK ? d.acc_owners - base : NHF_BLOCK;
            for (int j = 0; j < n; ++j) {
                const double pair[2] = {s_pair[j].x, s_pair[j].y};
                for (double x : pair) {
                    const double s = hi + x;
                    const double bp = s - hi;
                    lo += (hi - (s - bp)) + (x - bp);
                    hi = s;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * t] = hi;
        out[2 * t + 1] = lo;
    }
}

__global__ void k_clear_acc(Dev d, int part) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int4 span = part_span(d, part);
    for (int i = span.z * d.terms + tid; i < span.w * d.terms; i += gridDim.x * blockDim.x)
        d.acc[i] = make_double2(0.0, 0.0);
}

// Writes out the deferred extrema: a cell marked as rising (max_depth < 0) gets its depth as
// maximum and its clock's current tick (the end of its last step) as peak tick.
// In the uniform rain state it also writes out every cell's rain time and capacity.
__global__ __launch_bounds__(NHF_BLOCK) void k_settle_extrema(Dev d) {
    const nhf_mesh &m = d.m;
    for (int cell = blockIdx.x * NHF_BLOCK + threadIdx.x + 1; cell < m.e_num; cell += gridDim.x * NHF_BLOCK) {
        const int part = d.parts.count ? d.parts.cell_chunk[(cell - 1) >> 10] : 0;
        if (d.parts.count && cell >= d.parts.range[part].y) continue;  // padding
        if (m.max_depth[cell] < 0.0f) {
            m.max_depth[cell] = m.depth[cell];
            m.peak_tick[cell] = d.ctl[part].now;
        }
        if (d.rain_uniform[part]) {
            m.rain_time[cell] = d.rt_u[part];
            m.capacity[cell] = d.cap_u[meta_class(m.cell_meta[cell])];
        }
    }
}

__global__ void k_reset_limit(Dev d, int part) {
    if (threadIdx.x == 0) d.ctl[part].limit_value[0] = INFINITY;
}

__global__ void k_clock_probe(const uint64_t *now, const uint64_t *target, const uint64_t *origin,
                              const float *candidate, float min_dt, int n, uint32_t *words,
                              uint64_t *ticks, float *times, Ctl *scratch) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ctl *c = &scratch[i];
    for (int k = 0; k < NHF_TRACE_COLS; ++k) c->trace[k] = 0u;
    c->now = now[i];
    c->next = target[i];
    c->origin = origin[i];
    c->step_end = 0;
    clock_choose(c, candidate[i], min_dt, 0);
    words[4 * i + 0] = c->active ? NHF_COMPLETED : c->trace[3];
    words[4 * i + 1] = __float_as_uint(c->dt);
    words[4 * i + 2] = c->trace[4];
    words[4 * i + 3] = c->trace[5];
    ticks[2 * i + 0] = c->active ? c->step_end : 0ull;
    ticks[2 * i + 1] = c->active ? f32_ticks(c->dt) : 0ull;
    times[2 * i + 0] = c->local;
    times[2 * i + 1] = c->end_local;
}

// --------------------------------------------------------------------------- hanging gradients

// Length-weighted mean water-level difference and centre distance over one face's wet
// neighbours (PressureReconstruction.face_difference, core/pressure.py:66-90).
__device__ void face_difference(const Dev &d, int cell, int dir, int axis, float &value,
                                float &distance, int &usable) {
    const nhf_mesh &m = d.m;
    float v = 0.0f, dist = 0.0f, total = 0.0f;
    int ok = 1;
    float eta = m.bed[cell] + m.depth[cell];
    uint32_t cnt = m.csr_count[cell];
    int start = m.csr_start[cell];
    for (int k = 0; k < dir; ++k) start += face_n(cnt, k);
    int n = face_n(cnt, dir);
    for (int j = 0; j < n; ++j) {
        int side = m.isl[start + j];
        int other = m.isl_cell[start + j];
        if (other <= 0 || m.depth[other] <= d.p.min_h) {
            ok = 0;
        } else {
            float w = m.length[side];
            v += w * ((m.bed[other] + m.depth[other]) - eta);
            float delta = axis == 1 ? m.y[other] - m.y[cell] : m.x[other] - m.x[cell];
            dist += w * delta;
            total += w;
        }
    }
    if (total > 0.0f) {
        v /= total;
        dist /= total;
    } else {
        ok = 0;
    }
    value = v;
    distance = dist;
    usable = ok;
}

template <bool PARTS>
static __device__ __forceinline__ void gradient_slot(const Dev &d, int slot) {
    if (slot >= d.m.n_grad) return;
    int cell = d.m.grad_cell[slot];
    if (PARTS && !d.ctl[cell_part<true>(d, cell)].active) return;
    int axes = d.m.grad_axes[slot];
    float h = d.m.depth[cell];
    for (int axis = 0; axis < 2; ++axis) {
        if (!(axes & (1 << axis))) continue;
        float value = 0.0f;
        if (h > d.p.min_h) {
            float dm, xm, dp, xp;
            int okm, okp;
            face_difference(d, cell, axis == 1 ? 2 : 0, axis, dm, xm, okm);
            face_difference(d, cell, axis == 1 ? 3 : 1, axis, dp, xp, okp);
            if (okm && okp)
                value = (dp - dm) / (xp - xm);
            else if (okm)
                value = dm / xm;
            else if (okp)
                value = dp / xp;
        }
        d.grad[2 * slot + axis] = value;
    }
}

template <bool PARTS>
__global__ __launch_bounds__(NHF_BLOCK) void k_gradients(Dev d) {
    if (!PARTS && !d.ctl->active) return;
    gradient_slot<PARTS>(d, blockIdx.x * NHF_BLOCK + threadIdx.x);
}

// --------------------------------------------------------------------------- theta < 1

template <bool PARTS>
__global__ __launch_bounds__(NHF_BLOCK) void k_face_avg(Dev d) {
    if (!PARTS && !d.ctl->active) return;
    int cell = blockIdx.x * NHF_BLOCK + threadIdx.x;
    if (cell >= d.m.e_num) return;
    if (PARTS && cell > 0 && !d.ctl[cell_part<true>(d, cell)].active) return;
    const nhf_mesh &m = d.m;
    uint32_t cnt = m.csr_count[cell];
    int start = m.csr_start[cell];
    for (int dir = 0; dir < 4; ++dir) {
        int n = face_n(cnt, dir);
        float result = 0.0f;
        if (n == 1) {
            result = m.q[m.isl[start]];
        } else if (n > 1) {
            float total = 0.0f, width = 0.0f;
            for (int j = 0; j < n; ++j) {
                int side = m.isl[start + j];
                total += m.q[side] * m.length[side];
                width += m.length[side];
            }
            result = total / width;
        }
        d.face_qn[4 * cell + dir] = result;
        start += n;
    }
}

// --------------------------------------------------------------------------- (a) face flux

// Sides per thread of the flux pass: each thread issues the loads of both its sides before it
// uses either (the pass is bound by memory latency). Measured on a 3.2 M-side mesh: 1 side
// 11.8 us, 2 sides 256 apart 11.4 us, 4 sides 12.3 us, 2 adjacent sides (one 16-byte load of
// their cells) 12.0 us; on the 3.17 M-side benchmark state in launch order 19.1 / 18.5 / 19.8 us
// for 1 / 2 / 4 sides.
#ifndef NHF_FLUX_SIDES
#define NHF_FLUX_SIDES 2
#endif

struct FluxSide {
    int2 lh;
    uint32_t meta;
    float q, zl, zh, dl, dh;
};

// The new discharge of side si from its loaded values (solver_compact.py:439-486). `len` gets
// the side's length on a wet face (loaded with its coefficients), 0 on a dry one.
static __device__ __forceinline__ float flux_value(const Dev &d, int si, float dt, const FluxSide &f, float &len) {
    const nhf_mesh &m = d.m;
    const float g = d.p.g;
    const float bf = (f.meta & NHF_SIDE_EXTERIOR) ? 0.0f : 1.0f;
    const float hl = f.zl + f.dl, hr = f.zh + f.dh;
    const float q = f.q;
    const float wh = fmx(hl, hr) - fmx(fmx(f.zl, f.zh), -999.0f);
    // a face below min_h carries nothing (solver_compact.py:462,484): its arithmetic is skipped
    float nq = 0.0f;
    len = 0.0f;
    if (!(wh < d.p.min_h)) {
        len = m.length[si];
        float difference = hr - hl;
        if (m.edge_index) {
            int idx = m.edge_index[si];
            if (idx >= 0 && f.dl > d.p.min_h && f.dh > d.p.min_h) {
                int axis = (f.meta & NHF_SIDE_ORIENT_MASK) == 2 ? 1 : 0;
                float o0 = m.hang_offset[2 * idx], o1 = m.hang_offset[2 * idx + 1];
                if (o0 != 0.0f) difference -= d.grad[2 * m.hang_slot[2 * idx] + axis] * o0;
                if (o1 != 0.0f) difference += d.grad[2 * m.hang_slot[2 * idx + 1] + axis] * o1;
            }
        }
        float dq = ((-g * difference) / m.sdc[si]) * fmx(wh, 0.0f) * dt;
        float n = m.manning[si];
        // q == 0 gives f == 1 exactly for any finite positive power, so the power is skipped there.
        float fr = 1.0f;
        if (q != 0.0f) fr = 1.0f + g * dt * (n * n) * fabsf(q / det_powf(fmx(wh, 0.00001f), 7.0f / 3.0f));
        float neighbours = 0.0f;
        if (!d.theta_one) {
            int vertical = (f.meta & NHF_SIDE_ORIENT_MASK) == 2;
            int fa = vertical ? 0 : 2, fb = vertical ? 1 : 3;
            neighbours = neighbour_sum(q, d.face_qn[4 * f.lh.x + fa], d.face_qn[4 * f.lh.y + fb],
                                       m.open_face[4 * f.lh.x + fa], m.open_face[4 * f.lh.y + fb]);
        }
        nq = (d.p.theta * q + d.p.theta_half * neighbours + dq) / fr;
    }
    return nq * bf;
}

// The donor certificate (donor_share_exceeded): the flux pass flags the donor of a side that
// carries more than its share of the donor's volume.
static __device__ __forceinline__ void donor_check(const Dev &d, const FluxSide &f, float nq, float len, float dt) {
    const bool lower = nq >= 0.0f;  // the donor test of k_donor
    if (donor_share_exceeded(nq, dt, lower ? f.dl : f.dh, len)) d.donor_flag[lower ? f.lh.x : f.lh.y] = 1;
}

// A workgroup takes NHF_BLOCK * SIDES consecutive sides; thread t the sides t, t + NHF_BLOCK, ...
// Workgroups take the side ranges in launch order, not the XCD-contiguous order of the cell
// passes: the flux pass streams its sides, and one front over the array across all XCDs ran
// faster (17.9 against 18.7 us on the benchmark mesh; the 64-slot batch's flux 402 against 446 us).
template <int SIDES, bool PARTS>
__global__ __launch_bounds__(NHF_BLOCK) void k_flux(Dev d) {
    const nhf_mesh &m = d.m;
    const int block_first = (int)blockIdx.x * NHF_BLOCK * SIDES + 1;
    const Ctl *c = d.ctl;
    int end = m.s_num, part = 0, begin = 1;
    if (PARTS) {
        if (block_first >= m.s_num) return;
        part = side_part<true>(d, block_first);
        c = d.ctl + part;
        begin = d.parts.range[part].z;
        end = d.parts.range[part].w;
    }
    if (!c->active) return;
    if (block_first == begin && threadIdx.x < 64) advance_rain(d, part, c, threadIdx.x);
    const int first = block_first + threadIdx.x;
    if (first >= end) return;
    const float dt = c->dt;
    FluxSide f[SIDES];
#pragma unroll
    for (int u = 0; u < SIDES; ++u) {
        const int si = first + u * NHF_BLOCK;
        if (si < end) {
            f[u].lh = reinterpret_cast<const int2 *>(m.side_cells)[si];
            f[u].meta = m.side_meta[si];
            f[u].q = m.q[si];
        }
    }
#pragma unroll
    for (int u = 0; u < SIDES; ++u) {
        if (first + u * NHF_BLOCK < end) {
            f[u].zl = m.bed[f[u].lh.x];
            f[u].zh = m.bed[f[u].lh.y];
            f[u].dl = m.depth[f[u].lh.x];
            f[u].dh = m.depth[f[u].lh.y];
        }
    }
    float nq[SIDES], len[SIDES];
#pragma unroll
    for (int u = 0; u < SIDES; ++u)
        if (first + u * NHF_BLOCK < end) nq[u] = flux_value(d, first + u * NHF_BLOCK, dt, f[u], len[u]);
    // In place, as the reference's sq_t (solver_compact.py:486): no other thread of this pass
    // reads these sides' discharges. k_donor scales them afterwards where their donor is limited.
    // A discharge that keeps its bits (a dry face stays at zero) is not stored.
#pragma unroll
    for (int u = 0; u < SIDES; ++u)
        if (first + u * NHF_BLOCK < end && __float_as_uint(nq[u]) != __float_as_uint(f[u].q))
            m.q[first + u * NHF_BLOCK] = nq[u];
#pragma unroll
    for (int u = 0; u < SIDES; ++u)
        if (first + u * NHF_BLOCK < end) donor_check(d, f[u], nq[u], len[u], dt);
}

// --------------------------------------------------------------------------- donor scale

// The two cell passes are bound by memory latency rather than bandwidth, so their loads come in
// rounds: first every load that needs the cell index only, the side record included (issued
// together, before any of their values is used), then the side gathers and the boundary and rain
// lookups, which need the first round, and every store last (a store ahead of a load would make
// the load wait for it, since the arrays may alias as far as the compiler can tell). A wavefront
// whose cells are all regular (nearly all of them) runs a body with no list code, so nothing in
// it waits on a load it does not need.

struct DonorCell {
    uint32_t meta, pix;
    float h, size, pq;
    int4 rec;
};

template <bool REGULAR>
static __device__ __forceinline__ void donor_cell(const Dev &d, int cell, float dt, bool raining, const DonorCell &v,
                                                  uint32_t &limited, uint32_t &faces, uint32_t &scale_bits) {
    const nhf_mesh &m = d.m;
    const float size = m.size_base > 0.0f ? cell_size(m, v.meta, cell) : v.size;
    CellList l;
    cell_list<REGULAR>(m, v.rec, l);
    // second round: the sides' discharges, the boundary and the rain rate
    const int n0 = l.n0, n01 = l.n01, n012 = l.n012, total = l.total, k0 = l.k0;
    const uint32_t cnt = l.cnt;
    const bool short_list = l.short_list;
    const int *ids = l.ids;
    float qf[NHF_LIST_MAX], len[NHF_LIST_MAX];
    if (short_list) {
        gather_list(l, [&](int side) { return m.q[side]; }, qf);
        list_lengths(m, l, size, len);
    }
    const int bi = meta_boundary(v.meta);
    int kind = -1;
    float bvalue = 0.0f, bw = 0.0f;
    if (bi >= 0) {
        kind = m.kind[bi];
        bvalue = d.values[bi];
        bw = m.bweight[cell];
    }
    const float rainq = raining ? d.rain_rates[v.pix] : 0.0f;
    const float area = l.regular ? size * size : m.area[cell];
    const float discharge = kind == 1 ? bvalue * bw : 0.0f;
    const bool level = kind == 0;
    const float pq = v.pq;
    float positive = (fmx(pq, 0.0f) + fmx(discharge, 0.0f)) * dt;
    float available = v.h * area + positive + rainq * dt * area;
    float outgoing = (fmx(-pq, 0.0f) + fmx(-discharge, 0.0f)) * dt;
    if (short_list) {
#pragma unroll
        for (int j = 0; j < NHF_LIST_MAX; ++j) {
            if (j < total) {
                float sq = qf[j] * len[j];
                if (!(entry_face(j, n0, n01, n012) & 1)) sq = -sq;  // left and bottom faces
                outgoing += fmx(sq, 0.0f) * dt;
            }
        }
    } else {
        int k = k0;
        for (int dir = 0; dir < 4; ++dir) {
            int n = face_n(cnt, dir);
            for (int j = 0; j < n; ++j, ++k) {
                int side = m.isl[k];
                float sq = m.q[side] * m.length[side];
                if (dir == 0 || dir == 2) sq = -sq;
                outgoing += fmx(sq, 0.0f) * dt;
            }
        }
    }
    float ds = 1.0f;
    if (outgoing > available && !level) {
        ds = available / outgoing;
        limited += 1;
        scale_bits = min(scale_bits, __float_as_uint(ds));  // non-negative floats order as their bits
    }
    // Shared-subedge scaling (solver_compact.py:516-524): every side takes the scale of its
    // donor, the lower cell when q >= 0 and the higher cell otherwise. This cell is the lower
    // cell of its right/top sides, so it is the donor of a side exactly when
    // (q >= 0) == (face is right/top), and it scales that side's discharge in place: one writer
    // per side, and only where the scale is not 1 (q * 1 is q). A neighbour may read a side
    // while its donor scales it: the scale is in [0, 1), so the sign stays, or a zero of the
    // same sign results. For the neighbour the side is inflow either way (no outgoing volume),
    // and if it reads that zero it takes itself for the donor and stores the same zero.
    if (ds != 1.0f) {
        if (short_list) {
#pragma unroll
            for (int j = 0; j < NHF_LIST_MAX; ++j) {
                if (j < total && (qf[j] >= 0.0f) == (bool)(entry_face(j, n0, n01, n012) & 1)) {
                    m.q[ids[j]] = qf[j] * ds;
                    if (qf[j] != 0.0f && ds < 1.0f) faces += 1;
                }
            }
        } else {
            int k = k0;
            for (int dir = 0; dir < 4; ++dir) {
                int n = face_n(cnt, dir);
                const bool upper = dir & 1;
                for (int j = 0; j < n; ++j, ++k) {
                    int side = m.isl[k];
                    float q1 = m.q[side];
                    if ((q1 >= 0.0f) == upper) {
                        m.q[side] = q1 * ds;
                        if (q1 != 0.0f && ds < 1.0f) faces += 1;
                    }
                }
            }
        }
    }
    // the depth pass needs the scale itself only for boundary and pipe exchange
    if (bi >= 0 || m.pipe_q) d.donor[cell] = ds;
}

// First round of the donor pass: every load that needs only the cell index.
static __device__ __forceinline__ void load_donor_cell(const nhf_mesh &m, int cell, bool raining, DonorCell &v) {
    v.meta = m.cell_meta[cell];
    v.h = m.depth[cell];
    v.size = m.size_base > 0.0f ? 0.0f : m.size[cell];  // else computed from meta
    v.pix = (raining && m.n_pixels > 1) ? (uint32_t)m.rain_index[cell] : 0u;
    v.pq = m.pipe_q ? m.pipe_q[cell] : 0.0f;
    v.rec = cell_record(m, cell);
}

// CPT cells per thread (1 or 2): thread t of a workgroup takes cells t, t + BLOCK, ... of the
// workgroup's BLOCK*CPT cells, and issues the first-round loads of all of them before using any.
template <int BLOCK, int CPT, bool PARTS>
__global__ __launch_bounds__(BLOCK) void k_donor(Dev d) {
    const nhf_mesh &m = d.m;
    const int block_first = xcd_block() * BLOCK * CPT + 1;
    Ctl *c = d.ctl;
    int end = m.e_num;
    if (PARTS) {
        if (block_first >= m.e_num) return;
        const int part = cell_part<true>(d, block_first);
        c = d.ctl + part;
        end = d.parts.range[part].y;
    }
    if (!c->active) return;
    const float dt = c->dt;
    const bool raining = c->raining;
    const int first = block_first + threadIdx.x;
    uint32_t limited = 0, faces = 0, scale_bits = __float_as_uint(1.0f);
    // Only cells the flux pass flagged, or that the certificate does not cover, are looked at
    // (donor_check): every other cell's scale is 1 and it changes nothing. A flagged cell's flag
    // goes back to its standing value once it has been looked at.
    auto unflag = [&](int cell, uint8_t flag, const DonorCell &v) {
        const uint8_t standing = donor_always(m, v.meta, v.rec);
        if (!d.donor_keep && flag != standing) d.donor_flag[cell] = standing;
    };
    // first round: the cells' own values and their side records
    if constexpr (CPT == 1) {   // the plain form keeps every value in registers
        const uint8_t flag = first < end ? d.donor_flag[first] : 0;
        const bool live = flag != 0;
        if (!__any(live)) return;
        DonorCell v = {};
        if (live) load_donor_cell(m, first, raining, v);
        if (__all(!live || v.rec.x >= 0)) {
            if (live) donor_cell<true>(d, first, dt, raining, v, limited, faces, scale_bits);
        } else if (live) {
            donor_cell<false>(d, first, dt, raining, v, limited, faces, scale_bits);
        }
        if (live) unflag(first, flag, v);
    } else {   // two cells: both first rounds, then both bodies (named values, no arrays)
        static_assert(CPT == 2, "one or two cells per thread");
        const int second = first + BLOCK;
        const uint8_t flag0 = first < end ? d.donor_flag[first] : 0, flag1 = second < end ? d.donor_flag[second] : 0;
        const bool live0 = flag0 != 0, live1 = flag1 != 0;
        if (!__any(live0 || live1)) return;
        DonorCell v0 = {}, v1 = {};
        if (live0) load_donor_cell(m, first, raining, v0);
        if (live1) load_donor_cell(m, second, raining, v1);
        if (__all(!live0 || v0.rec.x >= 0)) {
            if (live0) donor_cell<true>(d, first, dt, raining, v0, limited, faces, scale_bits);
        } else if (live0) {
            donor_cell<false>(d, first, dt, raining, v0, limited, faces, scale_bits);
        }
        if (__all(!live1 || v1.rec.x >= 0)) {
            if (live1) donor_cell<true>(d, second, dt, raining, v1, limited, faces, scale_bits);
        } else if (live1) {
            donor_cell<false>(d, second, dt, raining, v1, limited, faces, scale_bits);
        }
        if (live0) unflag(first, flag0, v0);
        if (live1) unflag(second, flag1, v1);
    }
    // limiter counters: one integer atomic per wavefront that limited a cell or face (rare)
    if (__any(limited | faces)) {
        uint32_t wc = wave_sum_u(limited), wf = wave_sum_u(faces), wm = wave_min_u(scale_bits);
        if ((threadIdx.x & 63) == 0) {
            if (wc) atomicAdd(&c->limited_cells, wc);
            if (wf) atomicAdd(&c->limited_faces, wf);
            if (wc) atomicMin(&c->min_scale_bits, wm);
        }
    }
}

// --------------------------------------------------------------------------- (b) depth update

struct DepthCell {
    uint32_t meta, pix;
    float h, size, max_d, max_s, rain_time, pq, pv_before, pv_correction;
    int4 rec;
};

// Per-cell budget contributions (solver_compact.py:533, 578-590).
struct DepthTerms {
    float infil, floor, level_floor, pipe, disc;
    float lv[5];
    int bi, kind;
};

// A gate-accounted cell's six source terms (record_cell and collect, core/gate.py:149-188): rain,
// infiltration, positivity floor, pipe, boundary exchange, inventory rounding (all m^3, F32).
struct GateSources {
    float v[6];
};

// RECORD: compute the cell's update without storing anything and return its gate source terms
// (the gate accounting runs this between the donor and depth passes, on the same inputs).
template <bool REGULAR, bool RECORD = false>
static __device__ __forceinline__ void depth_cell(const Dev &d, int cell, float dt, bool raining, bool uniform,
                                                  uint64_t step_start, const DepthCell &v, DepthTerms &t,
                                                  GateSources *gs = nullptr) {
    const nhf_mesh &m = d.m;
    const float size = m.size_base > 0.0f ? cell_size(m, v.meta, cell) : v.size;
    CellList l;
    cell_list<REGULAR>(m, v.rec, l);
    // second round: the donor-scaled side discharges (stored in q by k_donor and k_flux), the
    // boundary, the cell's own scale, the rain rate and the Horton parameters
    const int n0 = l.n0, n01 = l.n01, n012 = l.n012, total = l.total, k0 = l.k0;
    const uint32_t cnt = l.cnt;
    float qv[NHF_LIST_MAX], len[NHF_LIST_MAX];
    if (l.short_list) {
        gather_list(l, [&](int side) { return m.q[side]; }, qv);
        list_lengths(m, l, size, len);
    }
    const int bi = meta_boundary(v.meta);
    int kind = -1;
    float bvalue = 0.0f, bw = 0.0f, lvalue = 0.0f, bed = 0.0f;
    if (bi >= 0) {
        kind = m.kind[bi];
        bvalue = d.values[bi];
        bw = m.bweight[cell];
        lvalue = d.level_values[bi];
        bed = m.bed[cell];
    }
    // the cell's own scale, stored by k_donor for boundary and pipe cells only
    const float ds = (bi >= 0 || m.pipe_q) ? d.donor[cell] : 1.0f;
    const float rainq = raining ? d.rain_rates[v.pix] : 0.0f;
    float f0 = 0.0f, fc = 0.0f, hk = 0.0f;
    if (raining) {
        const float *hp = m.class_horton + 3 * meta_class(v.meta);
        f0 = hp[0];
        fc = hp[1];
        hk = hp[2];
    }
    const float area = l.regular ? size * size : m.area[cell];
    t.bi = bi;
    t.kind = kind;

    const float pq = v.pq;
    float pipe = fmx(pq, 0.0f) + fmn(pq, 0.0f) * ds;
    float discharge = 0.0f;
    if (kind == 1) {
        float requested = bvalue * bw;
        discharge = fmx(requested, 0.0f) + fmn(requested, 0.0f) * ds;
        t.disc = discharge * dt;
    }
    float ssq = pipe + discharge;
    // compensated window volume (solver_compact.py:535-539)
    const float pv_increment = pipe * dt - v.pv_correction;
    const float pv_total = v.pv_before + pv_increment;
    // a cell with rain recomputes its capacity from its rain time; only a cell without reads it;
    // in the uniform rain state the class's capacity is already this step's (advance_rain)
    const float cap = uniform ? d.cap_u[meta_class(v.meta)]
                              : (rainq > 0.0f ? fc + (f0 - fc) * det_expf(-hk * (v.rain_time + dt) / 3600.0f)
                                              : m.capacity[cell]);
    float ql = 0.0f, qr = 0.0f, qb = 0.0f, qt = 0.0f;
    if (l.short_list) {
#pragma unroll
        for (int j = 0; j < NHF_LIST_MAX; ++j) {
            if (j < total) {
                const float p = qv[j] * len[j];
                const int dir = entry_face(j, n0, n01, n012);
                if (dir == 0)
                    ql += p;
                else if (dir == 1)
                    qr += p;
                else if (dir == 2)
                    qb += p;
                else
                    qt += p;
            }
        }
    } else {
        float fsum[4];
        int k = k0;
        for (int dir = 0; dir < 4; ++dir) {
            int n = face_n(cnt, dir);
            float acc = 0.0f;
            for (int j = 0; j < n; ++j, ++k) {
                int side = m.isl[k];
                acc += m.q[side] * m.length[side];
            }
            fsum[dir] = acc;
        }
        ql = fsum[0];
        qr = fsum[1];
        qb = fsum[2];
        qt = fsum[3];
    }
    const float previous = v.h;
    float tq = ql - qr + qb - qt + ssq;
    float rain_depth = rainq * dt;
    float increment = tq * dt / area + rain_depth;
    float raw = previous + increment;
    float removed = fmn(cap * dt, fmx(raw, 0.0f));
    raw -= removed;
    float floor_v = fmx(-raw, 0.0f);
    t.infil = removed * area;
    t.pipe = pipe * dt;
    float h_new;
    if (kind == 0) {
        t.level_floor = floor_v * area;
        float target = fmx(lvalue - bed, 0.0f);
        t.lv[0] = (target - previous) * area;
        t.lv[1] = -(ql - qr + qb - qt) * dt;
        t.lv[2] = -pipe * dt;
        t.lv[3] = -rain_depth * area;
        t.lv[4] = removed * area;
        h_new = target;
    } else {
        t.floor = floor_v * area;
        h_new = fmx(raw, 0.0f);
    }
    float hv = fmx(h_new, 0.01f);
    float u = 0.0f, w = 0.0f;
    if (h_new >= d.p.min_h) {
        u = (ql + qr) / size / hv / 2.0f;
        w = (qb + qt) / size / hv / 2.0f;
    }
    float speed = sqrtf(u * u + w * w);
    if constexpr (RECORD) {
        // record_cell / collect (core/gate.py:149-188; solver_compact.py:595-603), in F32
        const float applied_floor = kind == 0 ? 0.0f : floor_v;
        float level_exchange = 0.0f;
        if (kind == 0)
            level_exchange = (h_new - previous) * area - (ql - qr + qb - qt) * dt - pipe * dt - rain_depth * area +
                             removed * area;
        const float rounding = (h_new - previous) - increment + removed - applied_floor;
        gs->v[0] = rain_depth * area;
        gs->v[1] = removed * area;
        gs->v[2] = applied_floor * area;
        gs->v[3] = pipe * dt;
        gs->v[4] = (ssq - pipe) * dt + level_exchange;
        gs->v[5] = rounding * area - level_exchange;
        return;
    }
    // stores last
    if (m.pipe_volume) {
        m.pipe_correction[cell] = (pv_total - v.pv_before) - pv_increment;
        m.pipe_volume[cell] = pv_total;
    }
    if (rainq > 0.0f && !uniform) {
        m.rain_time[cell] = v.rain_time + dt;
        m.capacity[cell] = cap;
    }
    if (__float_as_uint(h_new) != __float_as_uint(previous)) m.depth[cell] = h_new;
    if (speed > v.max_s) m.max_speed[cell] = speed;
    // Extrema (solver_compact.py:600-606): the maximum depth and the end tick of the step that
    // first reached it. A negative max_depth marks a cell whose depth rose in its last step: its
    // maximum is its depth and its tick the end of that step, written out only when the rise
    // stops (or by k_settle_extrema), so a filling cell writes nothing step after step.
    if (v.max_d < 0.0f) {
        if (!(h_new > previous)) {
            m.max_depth[cell] = previous;
            m.peak_tick[cell] = step_start;
        }
    } else if (h_new > v.max_d) {
        m.max_depth[cell] = -1.0f;
    }
}

// First round of the depth pass: every load that needs only the cell index.
static __device__ __forceinline__ void load_depth_cell(const nhf_mesh &m, int cell, bool raining, bool uniform,
                                                       DepthCell &v) {
    v.meta = m.cell_meta[cell];
    v.h = m.depth[cell];
    v.size = m.size_base > 0.0f ? 0.0f : m.size[cell];  // else computed from meta
    v.max_d = m.max_depth[cell];
    v.max_s = m.max_speed[cell];
    v.rain_time = (raining && !uniform) ? m.rain_time[cell] : 0.0f;
    v.pix = (raining && m.n_pixels > 1) ? (uint32_t)m.rain_index[cell] : 0u;
    v.pq = m.pipe_q ? m.pipe_q[cell] : 0.0f;
    if (m.pipe_volume) {
        v.pv_correction = m.pipe_correction[cell];
        v.pv_before = m.pipe_volume[cell];
    }
    v.rec = cell_record(m, cell);
}

template <int BLOCK, int CPT, bool PARTS>
__global__ __launch_bounds__(BLOCK) void k_depth(Dev d) {
    const nhf_mesh &m = d.m;
    const int block_first = xcd_block() * BLOCK * CPT + 1;
    Ctl *c = d.ctl;
    int end = m.e_num;
    int2 bounds = make_int2(0, m.b_num);  // the part's boundaries: first, count
    int part = 0;
    if (PARTS) {
        if (block_first >= m.e_num) return;
        part = cell_part<true>(d, block_first);
        c = d.ctl + part;
        end = d.parts.range[part].y;
        const int4 e = d.parts.extra[part];
        bounds = make_int2(e.x, e.y);
    }
    if (!c->active) return;
    const float dt = c->dt;
    const bool raining = c->raining;
    const bool uniform = d.rain_uniform[part] != 0;
    const uint64_t step_start = c->now;  // the clock commits this step's end in k_finish
    const int b_num = bounds.y;
    const int first = block_first + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // accumulator k belongs to the wavefront whose first cell is 1 + 64 k (any workgroup shape)
    double2 *acc = d.acc + ((size_t)((block_first - 1) >> 6) + wave) * d.terms;
    // first round: the cells' own values and their side records
    DepthTerms t[CPT];
    if constexpr (CPT == 1) {   // the plain form keeps every value in registers
        const bool live = first < end;
        DepthCell v = {};
        if (live) load_depth_cell(m, first, raining, uniform, v);
        t[0] = DepthTerms{};
        t[0].bi = -1;
        t[0].kind = -1;
        if (__all(!live || v.rec.x >= 0)) {
            if (live) depth_cell<true>(d, first, dt, raining, uniform, step_start, v, t[0]);
        } else if (live) {
            depth_cell<false>(d, first, dt, raining, uniform, step_start, v, t[0]);
        }
    } else {   // two cells: both first rounds, then both bodies (named values, no arrays)
        static_assert(CPT == 2, "one or two cells per thread");
        const int second = first + BLOCK;
        const bool live0 = first < end, live1 = second < end;
        DepthCell v0 = {}, v1 = {};
        if (live0) load_depth_cell(m, first, raining, uniform, v0);
        if (live1) load_depth_cell(m, second, raining, uniform, v1);
        t[0] = DepthTerms{};
        t[0].bi = t[0].kind = -1;
        t[1] = DepthTerms{};
        t[1].bi = t[1].kind = -1;
        if (__all(!live0 || v0.rec.x >= 0)) {
            if (live0) depth_cell<true>(d, first, dt, raining, uniform, step_start, v0, t[0]);
        } else if (live0) {
            depth_cell<false>(d, first, dt, raining, uniform, step_start, v0, t[0]);
        }
        if (__all(!live1 || v1.rec.x >= 0)) {
            if (live1) depth_cell<true>(d, second, dt, raining, uniform, step_start, v1, t[1]);
        } else if (live1) {
            depth_cell<false>(d, second, dt, raining, uniform, step_start, v1, t[1]);
        }
    }
    // FP64 budget: each wavefront owns one compensated accumulator per term and adds the sum of
    // its cells' F32 terms; wavefronts whose cells are all zero for a term skip it. With at most
    // 64 terms, lane t holds the accumulator of term t, loaded once while the sums are formed.
    // Terms 0-3 (infiltration, floors, pipe) come first; the level components (4 + 5 b + c) and
    // the applied discharges (4 + 5 B + b) only in a wavefront that holds a boundary cell.
    const bool lane_terms = d.terms <= 64;
    double2 held = make_double2(0.0, 0.0);
    if (lane_terms && lane < d.terms) held = acc[lane];
    bool touched = false;
    auto add_term = [&](int k, double x) {
        if (!__any(x != 0.0)) return;
        const double w = wave_sum(x);
        if (lane_terms) {
            if (lane == k) {
                acc_add(held, w);
                touched = true;
            }
        } else if (lane == 0) {
            double2 a = acc[k];
            acc_add(a, w);
            acc[k] = a;
        }
    };
    double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
    bool boundary = false;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        x0 += (double)t[j].infil;
        x1 += (double)t[j].floor;
        x2 += (double)t[j].level_floor;
        x3 += (double)t[j].pipe;
        boundary = boundary || t[j].kind >= 0;
    }
    add_term(0, x0);
    add_term(1, x1);
    add_term(2, x2);
    add_term(3, x3);
    if (__any(boundary)) {
        for (int tb = 0; tb < b_num; ++tb) {
            bool level = false, inflow = false;
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                level = level || (t[j].kind == 0 && t[j].bi == bounds.x + tb);
                inflow = inflow || (t[j].kind == 1 && t[j].bi == bounds.x + tb);
            }
            if (__any(level)) {
#pragma unroll
                for (int comp = 0; comp < 5; ++comp) {
                    double x = 0.0;
#pragma unroll
                    for (int j = 0; j < CPT; ++j)
                        x += (double)((t[j].kind == 0 && t[j].bi == bounds.x + tb) ? t[j].lv[comp] : 0.0f);
                    add_term(4 + 5 * tb + comp, x);
                }
            }
            if (__any(inflow)) {
                double x = 0.0;
#pragma unroll
                for (int j = 0; j < CPT; ++j)
                    x += (double)((t[j].kind == 1 && t[j].bi == bounds.x + tb) ? t[j].disc : 0.0f);
                add_term(4 + 5 * b_num + tb, x);
            }
        }
    }
    if (touched) acc[lane] = held;
}

// --------------------------------------------------------------------------- gate control

// check_gates / finish_gate_bed (core/gate.py:77-101), one workgroup. At a step start (at_end 0)
// a gate not yet commanded is commanded when the formal start is reached and its control cell is
// wet (depth >= min_h) with a level above the trigger; then every commanded gate's cells take the
// bed of the opening fraction at the step's start or end tick. `standalone` runs it outside a
// batch (the clock need not be active).
__global__ __launch_bounds__(1024) void k_gate_move(Dev d, int at_end, int standalone) {
    const Ctl *c = d.ctl;
    if (!standalone && !c->active) return;
    const nhf_gates &g = d.g;
    const uint64_t tick = at_end ? c->step_end : c->now;
    if (!at_end) {
        for (int k = threadIdx.x; k < g.count; k += blockDim.x) {
            const int cell = g.control[k];
            const float h = d.m.depth[cell];
            if (tick >= g.formal_tick && g.flags[3 * k] == 0 && h >= g.min_h && d.m.bed[cell] + h > g.trigger_level) {
                g.flags[3 * k] = 1;
                g.ticks[3 * k] = tick;
            }
        }
        __syncthreads();
    }
    const uint32_t hi = (uint32_t)(tick >> 32), lo = (uint32_t)tick;
    for (int i = threadIdx.x; i < g.n_cells; i += blockDim.x) {
        const int cell = g.cells[i];
        const int k = g.cell_gate[cell];
        if (g.flags[3 * k] != 1) continue;
        // since_command (core/gate.py:66-70) on the tick's 32-bit words
        const uint64_t command = g.ticks[3 * k];
        const uint32_t chi = (uint32_t)(command >> 32), clo = (uint32_t)command;
        const uint32_t borrow = lo < clo ? 1u : 0u;
        const uint32_t high = hi - chi - borrow, low = lo - clo;
        const float since = (float)high / 256.0f + (float)low * 0x1p-40f;
        const float fraction = fmn(since / g.opening_seconds, 1.0f);
        if (fraction >= 1.0f) {
            d.m.bed[cell] = g.open_elevation;
            if (atomicOr(&g.flags[3 * k + 2], 1) == 0) g.ticks[3 * k + 2] = tick;
        } else {
            const float closed = g.closed[cell];
            d.m.bed[cell] = closed + (g.open_elevation - closed) * fraction;
        }
    }
}

// Gate accounting of one step, between the donor and depth passes: one workgroup per chunk of
// one term group. A side item adds its post-limiter volume into/out of its two terms (and marks
// the gate's first pond flow); a cell item computes the cell's update as k_depth will and adds
// its six source terms. Sums are FP64 in a fixed order per chunk, then one compensated
// accumulator per chunk and term.
__global__ __launch_bounds__(NHF_BLOCK) void k_gate_account(Dev d) {
    const Ctl *c = d.ctl;
    if (!c->active) return;
    const nhf_gates &g = d.g;
    const nhf_mesh &m = d.m;
    const int4 ch = reinterpret_cast<const int4 *>(g.chunks)[blockIdx.x];
    const int kind = ch.z & 255, flow_gate = (ch.z >> 8) - 1, width = kind == 0 ? 2 : 6;
    const float dt = c->dt;
    float v[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if ((int)threadIdx.x < ch.y) {
        const uint32_t code = (uint32_t)g.items[ch.x + threadIdx.x];
        if (kind == 0) {
            const int side = (int)(code & 0x7fffffffu);
            const float volume = m.q[side] * m.length[side] * dt;
            const float into = (code >> 31) ? -volume : volume;
            v[0] = fmx(into, 0.0f);
            v[1] = fmx(-into, 0.0f);
            if (flow_gate >= 0 && volume != 0.0f && atomicOr(&g.flags[3 * flow_gate + 1], 1) == 0)
                g.ticks[3 * flow_gate + 1] = c->step_end;
        } else {
            const int cell = (int)code;
            const bool raining = c->raining;
            DepthCell cv = {};
            const bool uniform = d.rain_uniform[0] != 0;
            load_depth_cell(m, cell, raining, uniform, cv);
            DepthTerms t = {};
            GateSources gs;
            if (cv.rec.x >= 0)
                depth_cell<true, true>(d, cell, dt, raining, uniform, c->now, cv, t, &gs);
            else
                depth_cell<false, true>(d, cell, dt, raining, uniform, c->now, cv, t, &gs);
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = gs.v[k];
        }
    }
    __shared__ double s_part[NHF_WAVES][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < width; ++k) {
        const double w = wave_sum((double)v[k]);
        if (lane == 0) s_part[wave][k] = w;
    }
    __syncthreads();
    if ((int)threadIdx.x < width) {
        double sum = 0.0;
        for (int w = 0; w < NHF_WAVES; ++w) sum += s_part[w][threadIdx.x];
        double2 *acc = reinterpret_cast<double2 *>(g.acc) + (size_t)blockIdx.x * 6 + threadIdx.x;
        double2 a = *acc;
        acc_add(a, sum);
        *acc = a;
    }
}

// --------------------------------------------------------------------------- (c) CFL dt

// Per-workgroup CFL minimum over a grid-stride range of sides. The key orders by the F32 bits of
// the bound (non-negative, so bit order is numeric order), then by the reference side id, which
// makes the argmin the smallest reference id among ties (solver_compact.py:633-641).
// The first `grad_blocks` workgroups (a multiple of 8, so each CFL workgroup keeps its XCD)
// compute the next step's hanging-edge gradients instead: they need only the depth and bed this
// step ends with, which are final once the depth pass is done.
// With parts, each part has its own range of CFL workgroups (a multiple of 8, starting at a
// multiple of 8) striding over its own sides.
template <bool PARTS>
__global__ __launch_bounds__(NHF_BLOCK) void k_cfl(Dev d, int mode, int grad_blocks) {
    if (!PARTS && mode == 1 && !d.ctl->active) return;
    if ((int)blockIdx.x < grad_blocks) {
        gradient_slot<PARTS>(d, blockIdx.x * NHF_BLOCK + threadIdx.x);
        return;
    }
    const int block = (int)blockIdx.x - grad_blocks;
    int begin = 1, end = d.m.s_num, count = d.cfl_blocks, lb = block;
    if (PARTS) {
        const int part = d.parts.cfl_part[block];
        if (mode == 1 && !d.ctl[part].active) return;
        const int4 r = d.parts.range[part], e = d.parts.extra[part];
        begin = r.z;
        end = r.w;
        count = e.w;
        lb = block - e.z;
    }
    const int xb = (lb & 7) * (count >> 3) + (lb >> 3);
    __shared__ uint64_t s_key[NHF_WAVES];
    __shared__ int32_t s_idx[NHF_WAVES];
    const nhf_mesh &m = d.m;
    const float afa = d.p.afa, g = d.p.g;
    // The thread's minimum of (bound bits, reference id): the reference id of a side is read only
    // where the bounds tie, and for the thread's winner at the end.
    uint32_t best_bits = ~0u;
    int32_t idx = -1;
    for (int si = begin + xb * NHF_BLOCK + threadIdx.x; si < end; si += count * NHF_BLOCK) {
        int2 lh = reinterpret_cast<const int2 *>(m.side_cells)[si];
        if (lh.x != 0 && lh.y != 0) {
            float zl = m.bed[lh.x], zh = m.bed[lh.y];
            float head = fmx(zl + m.depth[lh.x], zh + m.depth[lh.y]);
            float h = fmx(head - fmx(fmx(zl, zh), -999.0f), 0.01f);
            float cfl = afa * m.length[si] / (sqrtf(g * h) + fabsf(m.q[si]) / h);
            const uint32_t bits = __float_as_uint(cfl);
            if (bits < best_bits || (bits == best_bits && idx >= 0 && m.side_key[si] < m.side_key[idx])) {
                best_bits = bits;
                idx = si;
            }
        }
    }
    uint64_t key = idx >= 0 ? ((uint64_t)best_bits << 32) | m.side_key[idx] : ~0ull;
    if (idx < 0) idx = 0;
    wave_min_key(key, idx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_key[wave] = key;
        s_idx[wave] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NHF_WAVES; ++w)
            if (s_key[w] < key) {
                key = s_key[w];
                idx = s_idx[w];
            }
        d.cfl_key[block] = key;
        d.cfl_idx[block] = idx;
    }
}

// One workgroup: reduce the CFL partials to ndt and its side; in a batch (mode 1) also commit
// `slot` (finish_step) and prepare `slot + 1` (prepare_step) when it is below `steps`.
template <bool PARTS>
__global__ __launch_bounds__(1024) void k_finish(Dev d, int mode, int slot, int steps) {
    const int part = PARTS ? (int)blockIdx.x : 0;  // one workgroup per part
    Ctl *c = d.ctl + part;
    if (mode == 1 && !c->active) return;
    __shared__ uint64_t s_key[16];
    __shared__ int32_t s_idx[16];
    uint64_t key = ~0ull;
    int32_t idx = 0;
    int b0 = 0, b1 = d.cfl_blocks;
    if (PARTS) {
        const int4 e = d.parts.extra[part];
        b0 = e.z;
        b1 = e.z + e.w;
    }
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        uint64_t k2 = d.cfl_key[b];
        if (k2 < key) {
            key = k2;
            idx = d.cfl_idx[b];
        }
    }
    wave_min_key(key, idx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_key[wave] = key;
        s_idx[wave] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (s_key[w] < key) {
                key = s_key[w];
                idx = s_idx[w];
            }
        if (key == ~0ull) {
            c->ndt = INFINITY;
            c->cfl_arg = 0;
        } else {
            c->ndt = __uint_as_float((uint32_t)(key >> 32));
            c->cfl_arg = idx;
        }
        if (mode == 1) {
            commit_slot(c, slot);
            if (slot + 1 < steps) prepare_slot(d, part, slot + 1);
        }
    }
}

// --------------------------------------------------------------------------- host side

enum { K_PREPARE = 0, K_GRAD, K_FACE, K_FLUX, K_DONOR, K_DEPTH, K_CFL, K_FINISH, K_GATE, K_COUNT };

// The cell passes' instantiation for the context's workgroup size and cells per thread.
#define NHF_CELL_PASS(KERNEL)                                                                       \
    template <bool PARTS>                                                                          \
    static void launch_##KERNEL##_as(const Dev &d, hipStream_t s) {                                \
        const dim3 grid(d.depth_blocks);                                                           \
        switch (d.cell_block * 10 + d.cells_per_thread) {                                          \
        case 1281: hipLaunchKernelGGL((KERNEL<128, 1, PARTS>), grid, dim3(128), 0, s, d); break;    \
        case 1282: hipLaunchKernelGGL((KERNEL<128, 2, PARTS>), grid, dim3(128), 0, s, d); break;    \
        case 2562: hipLaunchKernelGGL((KERNEL<256, 2, PARTS>), grid, dim3(256), 0, s, d); break;    \
        case 5121: hipLaunchKernelGGL((KERNEL<512, 1, PARTS>), grid, dim3(512), 0, s, d); break;    \
        case 5122: hipLaunchKernelGGL((KERNEL<512, 2, PARTS>), grid, dim3(512), 0, s, d); break;    \
        default: hipLaunchKernelGGL((KERNEL<256, 1, PARTS>), grid, dim3(256), 0, s, d); break;      \
        }                                                                                          \
    }                                                                                              \
    static void launch_##KERNEL(const Dev &d, hipStream_t s) {                                     \
        if (d.parts.count)                                                                         \
            launch_##KERNEL##_as<true>(d, s);                                                      \
        else                                                                                       \
            launch_##KERNEL##_as<false>(d, s);                                                     \
    }
NHF_CELL_PASS(k_donor)
NHF_CELL_PASS(k_depth)
#undef NHF_CELL_PASS

struct nhf_ctx {
    Dev d;
    hipStream_t stream;
    int graph_enabled, timing;
    hipGraphExec_t exec[NHF_MAX_BATCH + 1];
    std::vector<hipEvent_t> events;
    std::vector<int> event_kind;
    double kernel_ms[K_COUNT];
    int64_t kernel_count[K_COUNT];
    uint32_t *host_trace;
    double2 *host_acc;
    Ctl *probe_scratch;
    std::vector<int32_t> gate_chunks;  // host copy of the gate accounting chunks
    hipEvent_t done;    // recorded after a launched batch's trace copy
    int pending_steps;  // steps of the batch in flight, 0 when none
    int parts;          // 0: one mesh (nhf_create); else the parts of nhu_create
    std::vector<int4> part_range, part_extra;
    std::vector<int2> part_pixels;
    Ctl *host_ctl;      // pinned copy of every part's control block (parts contexts)
};

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

struct Carve {
    size_t ctl, cell_chunk, side_chunk, range, cfl_part, coeff, values, level, rain, donor, face_qn, grad, key, idx,
        acc, donor_flag, rain_uniform, rain_pixel_u, rt_u, cap_u, cap_stage, class_part, total;
};

static int blocks_for(int n) { return (n + NHF_BLOCK - 1) / NHF_BLOCK; }

// Grids of the XCD-mapped kernels are rounded up to a multiple of the 8 XCDs.
static int blocks_xcd(int n) { return (blocks_for(n) + 7) & ~7; }

// Grid of the flux pass: NHF_FLUX_SIDES sides per thread, a multiple of the 8 XCDs.
static int flux_blocks(int sides) {
    const int per = NHF_BLOCK * NHF_FLUX_SIDES;
    return ((sides + per - 1) / per + 7) & ~7;
}
static void launch_k_flux(const Dev &d, int sides, hipStream_t s) {
    if (d.parts.count)
        hipLaunchKernelGGL((k_flux<NHF_FLUX_SIDES, true>), dim3(flux_blocks(sides)), dim3(NHF_BLOCK), 0, s, d);
    else
        hipLaunchKernelGGL((k_flux<NHF_FLUX_SIDES, false>), dim3(flux_blocks(sides)), dim3(NHF_BLOCK), 0, s, d);
}

// CFL partials: the workspace holds up to NHF_CFL_MAX_BLOCKS (no more than the sides need); the
// default grid of 2048 workgroups fills 256 CUs and stays small for the finishing workgroup.
// nhf_set_tuning picks another grid within that range.
#define NHF_CFL_MAX_BLOCKS 4096

static int cfl_max_blocks(int s_num) {
    int n = blocks_xcd(s_num - 1 > 0 ? s_num - 1 : 1);
    return n < NHF_CFL_MAX_BLOCKS ? n : NHF_CFL_MAX_BLOCKS;
}

static int cfl_grid(int s_num) {
    int n = cfl_max_blocks(s_num);
    return n < 2048 ? n : 2048;
}

// Grid of a cell pass with workgroups of `block` threads, a multiple of the 8 XCDs.
static int cell_grid(int e_num, int block) {
    int n = e_num - 1 > 0 ? e_num - 1 : 1;
    return ((n + block - 1) / block + 7) & ~7;
}

// Budget accumulators: accumulator k belongs to the wavefront whose first cell is 1 + 64 k, so
// one per 64 cells of the largest grid of any workgroup shape; a change of shape keeps every
// accumulated term.
static int acc_owner_count(int e_num) {
    int most = 0;
    for (int b = 128; b <= 512; b *= 2)
        for (int cpt = 1; cpt <= 2; ++cpt) {
            int owners = cell_grid(e_num, b * cpt) * (b * cpt / 64);
            most = owners > most ? owners : most;
        }
    return most;
}

// Parts of a concatenated mesh (nhu_create): ranges checked, and the device tables.
struct PartPlan {
    std::vector<int4> range, extra;  // as PartDev
    std::vector<int32_t> cell_chunk, side_chunk, cfl_part;
    std::vector<int2> pixels;        // per part: first rain pixel, pixels
    int cfl_total = 0, max_b = 0;
};

static int padded(int n) { return (n + NHF_PART_ALIGN - 1) / NHF_PART_ALIGN * NHF_PART_ALIGN; }

// CFL workgroups of a part with `sides` sides: about six sides per thread, a multiple of 8 in
// [8, 2048] (the single-mesh default grid).
static int part_cfl_blocks(int sides) {
    int n = (sides + 6 * NHF_BLOCK - 1) / (6 * NHF_BLOCK);
    n = (n + 7) & ~7;
    return n < 8 ? 8 : n > 2048 ? 2048 : n;
}

static const char *plan_parts(const nhf_mesh *m, const nhf_part *parts, int count, PartPlan &plan) {
    if (count < 1 || count > NHF_MAX_PARTS) return "part count out of range";
    int cell = 1, side = 1, bnd = 0, pix = 0;
    for (int p = 0; p < count; ++p) {
        const nhf_part &q = parts[p];
        if (q.cell_begin != cell || q.cell_end <= q.cell_begin || q.side_begin != side || q.side_end < q.side_begin)
            return "parts must tile the cells and sides in order, each starting at 1 + 1024 k";
        if (q.b_base != bnd || q.b_count < 0 || q.pix_base != pix || q.pix_count < 1)
            return "parts must tile the boundaries and rain pixels in order";
        cell = q.cell_begin + padded(q.cell_end - q.cell_begin);
        side = q.side_begin + padded(q.side_end - q.side_begin > 0 ? q.side_end - q.side_begin : 1);
        bnd += q.b_count;
        pix += q.pix_count;
        const int blocks = part_cfl_blocks(q.side_end - q.side_begin);
        plan.range.push_back(make_int4(q.cell_begin, q.cell_end, q.side_begin, q.side_end));
        plan.extra.push_back(make_int4(q.b_base, q.b_count, plan.cfl_total, blocks));
        plan.pixels.push_back(make_int2(q.pix_base, q.pix_count));
        for (int k = 0; k < blocks; ++k) plan.cfl_part.push_back(p);
        plan.cfl_total += blocks;
        if (q.b_count > plan.max_b) plan.max_b = q.b_count;
        for (int k = 0; k < padded(q.cell_end - q.cell_begin) / NHF_PART_ALIGN; ++k) plan.cell_chunk.push_back(p);
        for (int k = 0; k < padded(q.side_end - q.side_begin > 0 ? q.side_end - q.side_begin : 1) / NHF_PART_ALIGN; ++k)
            plan.side_chunk.push_back(p);
    }
    if (cell != m->e_num || side != m->s_num) return "the parts' padded ranges must end at e_num and s_num";
    if (bnd != m->b_num || pix != m->n_pixels) return "the parts' boundaries and pixels must sum to b_num and n_pixels";
    return nullptr;
}

static Carve carve(const nhf_mesh *m, const nhf_params *p, const PartPlan *plan = nullptr) {
    Carve c;
    size_t off = 0;
    int b = m->b_num > 0 ? m->b_num : 1;
    int terms = 4 + 6 * (plan ? plan->max_b : m->b_num);
    int cfl_blocks = plan ? plan->cfl_total : cfl_max_blocks(m->s_num);
    bool theta_one = p->theta == 1.0f;
    const int parts = plan ? (int)plan->range.size() : 1;
    c.ctl = off; off = align_up(off + sizeof(Ctl) * parts);
    c.cell_chunk = off; off = align_up(off + (plan ? sizeof(int32_t) * plan->cell_chunk.size() : 0));
    c.side_chunk = off; off = align_up(off + (plan ? sizeof(int32_t) * plan->side_chunk.size() : 0));
    c.range = off; off = align_up(off + (plan ? sizeof(int4) * 2 * parts : 0));
    c.cfl_part = off; off = align_up(off + (plan ? sizeof(int32_t) * plan->cfl_part.size() : 0));
    c.coeff = off; off = align_up(off + sizeof(float) * 2 * b);
    c.values = off; off = align_up(off + sizeof(float) * b);
    c.level = off; off = align_up(off + sizeof(float) * b);
    c.rain = off; off = align_up(off + sizeof(float) * (m->n_pixels > 0 ? m->n_pixels : 1));
    c.donor = off; off = align_up(off + sizeof(float) * m->e_num);
    c.face_qn = off; off = align_up(off + (theta_one ? 0 : sizeof(float) * 4 * (size_t)m->e_num));
    c.grad = off; off = align_up(off + sizeof(float) * 2 * (m->n_grad > 0 ? m->n_grad : 1));
    c.key = off; off = align_up(off + sizeof(uint64_t) * cfl_blocks);
    c.idx = off; off = align_up(off + sizeof(int32_t) * cfl_blocks);
    c.acc = off; off = align_up(off + sizeof(double2) * (size_t)acc_owner_count(m->e_num) * terms);
    c.donor_flag = off; off = align_up(off + (size_t)m->e_num);
    c.rain_uniform = off; off = align_up(off + sizeof(int32_t) * parts);
    c.rain_pixel_u = off; off = align_up(off + sizeof(int32_t) * parts);
    c.rt_u = off; off = align_up(off + sizeof(float) * parts);
    c.cap_u = off; off = align_up(off + sizeof(float) * 256);
    c.cap_stage = off; off = align_up(off + sizeof(float) * 256);
    c.class_part = off; off = align_up(off + sizeof(int32_t) * 256);
    c.total = off;
    return c;
}

extern "C" {

int32_t nhf_abi_version(void) { return NHF_ABI_VERSION; }

const char *nhf_last_error(void) { return nhf_error_text; }

size_t nhf_workspace_bytes(const nhf_mesh *mesh, const nhf_params *params) {
    if (!mesh || !params) return 0;
    return carve(mesh, params).total;
}

// Argument checks shared by nhf_create, nhu_create and nhf_run_static.
static int32_t check_mesh(const char *who, const nhf_mesh *mesh, const nhf_params *params, bool parts) {
    if (params->batch_capacity < 1 || params->batch_capacity > NHF_MAX_BATCH) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "%s: batch_capacity out of range", who);
        return -2;
    }
    const char *why = nullptr;
    if (mesh->e_num < 2 || mesh->s_num < 1)
        why = "empty mesh";
    else if ((params->theta != 1.0f) && !mesh->open_face)
        why = "open_face is required when sita != 1";
    else if (mesh->n_hang > 0 && (!mesh->edge_index || !mesh->x || !mesh->y || !mesh->grad_cell))
        why = "hanging reconstruction arrays missing";
    else if ((mesh->pipe_volume != nullptr) != (mesh->pipe_correction != nullptr) ||
             (mesh->pipe_volume && !mesh->pipe_q))
        why = "pipe volume and correction go together, with pipe_q";
    else if (parts && mesh->pipe_q)
        why = "pipe exchange is not supported with parts";
    if (why) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "%s: %s", who, why);
        return -2;
    }
    if (const char *name = missing_mesh_array(mesh, true, true)) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "%s: mesh array %s is NULL", who, name);
        return -2;
    }
    return 0;
}

// The device view of one mesh whose scratch is carved from `ws` (default launch shapes).
static void bind_dev(Dev &d, const nhf_mesh *mesh, const nhf_params *params, const PartPlan *plan, char *ws,
                     const Carve &cv) {
    d = Dev{};
    d.m = *mesh;
    if (mesh->n_hang == 0) d.m.edge_index = nullptr;
    d.p = *params;
    d.ctl = (Ctl *)(ws + cv.ctl);
    d.coefficients = (float *)(ws + cv.coeff);
    d.values = (float *)(ws + cv.values);
    d.level_values = (float *)(ws + cv.level);
    d.rain_rates = (float *)(ws + cv.rain);
    d.donor = (float *)(ws + cv.donor);
    d.face_qn = params->theta == 1.0f ? nullptr : (float *)(ws + cv.face_qn);
    d.grad = (float *)(ws + cv.grad);
    d.cfl_key = (uint64_t *)(ws + cv.key);
    d.cfl_idx = (int32_t *)(ws + cv.idx);
    d.acc = (double2 *)(ws + cv.acc);
    d.donor_flag = (uint8_t *)(ws + cv.donor_flag);
    d.donor_keep = 0;
    d.rain_uniform = (int32_t *)(ws + cv.rain_uniform);
    d.rain_pixel_u = (int32_t *)(ws + cv.rain_pixel_u);
    d.rt_u = (float *)(ws + cv.rt_u);
    d.cap_u = (float *)(ws + cv.cap_u);
    d.cap_stage = (float *)(ws + cv.cap_stage);
    d.class_part = (int32_t *)(ws + cv.class_part);
    d.terms = 4 + 6 * (plan ? plan->max_b : mesh->b_num);
    d.cfl_blocks = plan ? plan->cfl_total : cfl_grid(mesh->s_num);
    d.cfl_max = plan ? plan->cfl_total : cfl_max_blocks(mesh->s_num);
    d.cell_block = NHF_BLOCK;
    d.cells_per_thread = 1;
    d.depth_blocks = cell_grid(mesh->e_num, d.cell_block);
    d.acc_owners = acc_owner_count(mesh->e_num);
    d.theta_one = params->theta == 1.0f;
    d.parts = PartDev{};
    if (plan) {
        const int n = (int)plan->range.size();
        d.parts.count = n;
        d.parts.cell_chunk = (const int32_t *)(ws + cv.cell_chunk);
        d.parts.side_chunk = (const int32_t *)(ws + cv.side_chunk);
        d.parts.range = (const int4 *)(ws + cv.range);
        d.parts.extra = d.parts.range + n;
        d.parts.cfl_part = (const int32_t *)(ws + cv.cfl_part);
    }
}

static int32_t create_ctx(nhf_ctx **out, const nhf_mesh *mesh, const nhf_params *params, const PartPlan *plan,
                          void *workspace, size_t workspace_bytes, void *stream) {
    if (!out || !mesh || !params || !workspace) return set_error_text("nhf_create: null argument");
    if (int32_t rc = check_mesh(plan ? "nhu_create" : "nhf_create", mesh, params, plan != nullptr)) return rc;
    Carve cv = carve(mesh, params, plan);
    if (workspace_bytes < cv.total) return set_error_text("nhf_create: workspace too small");
    char *ws = (char *)workspace;
    if (plan) {
        // the part tables, copied once into the workspace
        HIP_TRY(hipMemcpy(ws + cv.cell_chunk, plan->cell_chunk.data(), sizeof(int32_t) * plan->cell_chunk.size(),
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ws + cv.side_chunk, plan->side_chunk.data(), sizeof(int32_t) * plan->side_chunk.size(),
                          hipMemcpyHostToDevice));
        std::vector<int4> table(plan->range);
        table.insert(table.end(), plan->extra.begin(), plan->extra.end());
        HIP_TRY(hipMemcpy(ws + cv.range, table.data(), sizeof(int4) * table.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ws + cv.cfl_part, plan->cfl_part.data(), sizeof(int32_t) * plan->cfl_part.size(),
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(ws + cv.ctl, 0, sizeof(Ctl) * plan->range.size()));
    }
    nhf_ctx *ctx = new nhf_ctx();
    bind_dev(ctx->d, mesh, params, plan, ws, cv);
    {
        const int blocks = blocks_for(mesh->e_num);
        hipLaunchKernelGGL(k_init_classes, dim3(1), dim3(NHF_BLOCK), 0, (hipStream_t)stream, ctx->d,
                           plan ? (int)plan->range.size() : 1);
        hipLaunchKernelGGL(k_init_flags, dim3(blocks < 2048 ? blocks : 2048), dim3(NHF_BLOCK), 0, (hipStream_t)stream,
                           ctx->d);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) {
            delete ctx;
            return set_error("nhf_create: donor flags", err);
        }
    }
    ctx->parts = 0;
    ctx->host_ctl = nullptr;
    if (plan) {
        ctx->parts = (int)plan->range.size();
        ctx->part_range = plan->range;
        ctx->part_extra = plan->extra;
        ctx->part_pixels = plan->pixels;
    }
    ctx->stream = (hipStream_t)stream;
    ctx->graph_enabled = 0;
    ctx->timing = 0;
    for (int i = 0; i <= NHF_MAX_BATCH; ++i) ctx->exec[i] = nullptr;
    ctx->host_acc = nullptr;
    ctx->probe_scratch = nullptr;
    ctx->pending_steps = 0;
    hipError_t err = hipHostMalloc((void **)&ctx->host_trace, sizeof(uint32_t) * NHF_MAX_BATCH * NHF_TRACE_COLS,
                                   hipHostMallocDefault);
    if (err == hipSuccess && plan)
        err = hipHostMalloc((void **)&ctx->host_ctl, sizeof(Ctl) * plan->range.size(), hipHostMallocDefault);
    if (err != hipSuccess) {
        if (ctx->host_trace) (void)hipHostFree(ctx->host_trace);
        delete ctx;
        return set_error("hipHostMalloc trace", err);
    }
    err = hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming);
    if (err != hipSuccess) {
        (void)hipHostFree(ctx->host_trace);
        if (ctx->host_ctl) (void)hipHostFree(ctx->host_ctl);
        delete ctx;
        return set_error("hipEventCreate", err);
    }
    memset(ctx->kernel_ms, 0, sizeof ctx->kernel_ms);
    memset(ctx->kernel_count, 0, sizeof ctx->kernel_count);
    *out = ctx;
    return 0;
}

int32_t nhf_create(nhf_ctx **out, const nhf_mesh *mesh, const nhf_params *params, void *workspace,
                   size_t workspace_bytes, void *stream) {
    return create_ctx(out, mesh, params, nullptr, workspace, workspace_bytes, stream);
}

int32_t nhf_destroy(nhf_ctx *ctx) {
    if (!ctx) return 0;
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i <= NHF_MAX_BATCH; ++i)
        if (ctx->exec[i]) (void)hipGraphExecDestroy(ctx->exec[i]);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->done);
    if (ctx->host_trace) (void)hipHostFree(ctx->host_trace);
    if (ctx->host_ctl) (void)hipHostFree(ctx->host_ctl);
    if (ctx->host_acc) (void)hipHostFree(ctx->host_acc);
    delete ctx;
    return 0;
}

// Entry points of one mesh refuse a parts context and the nhu_* forms a single mesh.
static int32_t single_only(const nhf_ctx *ctx, const char *name) {
    if (!ctx) return set_error_text("null context");
    if (ctx->parts) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "%s: a parts context takes the nhu_* form", name);
        return -2;
    }
    return 0;
}

static int32_t part_only(const nhf_ctx *ctx, int part, const char *name) {
    if (!ctx) return set_error_text("null context");
    if (!ctx->parts || part < 0 || part >= ctx->parts) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "%s: not a parts context or part out of range", name);
        return -2;
    }
    return 0;
}

static int32_t reset_part(nhf_ctx *ctx, int part, uint64_t now_ticks) {
    hipLaunchKernelGGL(k_reset, dim3(256), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, part, now_ticks);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int32_t forcing_part(nhf_ctx *ctx, int part, const float *coefficients, const float *rain_rates,
                            int32_t raining) {
    const Dev &d = ctx->d;
    int b0 = 0, nb = d.m.b_num, p0 = 0, np = d.m.n_pixels;
    if (ctx->parts) {
        b0 = ctx->part_extra[part].x;
        nb = ctx->part_extra[part].y;
        p0 = ctx->part_pixels[part].x;
        np = ctx->part_pixels[part].y;
    }
    if (nb > 0 && coefficients)
        HIP_TRY(hipMemcpyAsync(d.coefficients + 2 * b0, coefficients, sizeof(float) * 2 * nb, hipMemcpyHostToDevice,
                               ctx->stream));
    if (rain_rates && np > 0)
        HIP_TRY(hipMemcpyAsync(d.rain_rates + p0, rain_rates, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_set_raining, dim3(1), dim3(64), 0, ctx->stream, ctx->d, part, (int)raining);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int32_t begin_part(nhf_ctx *ctx, int part, uint64_t target_ticks, uint64_t origin_ticks) {
    hipLaunchKernelGGL(k_begin, dim3(1), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, part, target_ticks, origin_ticks);
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhf_reset(nhf_ctx *ctx, uint64_t now_ticks) {
    if (int32_t rc = single_only(ctx, "nhf_reset")) return rc;
    return reset_part(ctx, 0, now_ticks);
}

int32_t nhf_set_forcing(nhf_ctx *ctx, const float *coefficients, const float *rain_rates, int32_t raining) {
    if (int32_t rc = single_only(ctx, "nhf_set_forcing")) return rc;
    return forcing_part(ctx, 0, coefficients, rain_rates, raining);
}

int32_t nhf_begin(nhf_ctx *ctx, uint64_t target_ticks, uint64_t origin_ticks) {
    if (int32_t rc = single_only(ctx, "nhf_begin")) return rc;
    return begin_part(ctx, 0, target_ticks, origin_ticks);
}

static hipEvent_t timing_event(nhf_ctx *ctx, size_t i) {
    while (ctx->events.size() <= i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->events.push_back(e);
    }
    return ctx->events[i];
}

// Launch the kernels of `steps` slots. With timing on, each launch is bracketed by events.
static int32_t launch_batch(nhf_ctx *ctx, int steps, bool timed) {
    const Dev &d = ctx->d;
    hipStream_t s = ctx->stream;
    size_t ev = 0;
    ctx->event_kind.clear();
#define TIMED(kind, stmt)                                                   \
    do {                                                                    \
        if (timed) {                                                        \
            hipEvent_t e0 = timing_event(ctx, ev++);                        \
            if (!e0) return set_error_text("hipEventCreate failed");        \
            HIP_TRY(hipEventRecord(e0, s));                                 \
        }                                                                   \
        stmt;                                                               \
        if (timed) {                                                        \
            hipEvent_t e1 = timing_event(ctx, ev++);                        \
            if (!e1) return set_error_text("hipEventCreate failed");        \
            HIP_TRY(hipEventRecord(e1, s));                                 \
            ctx->event_kind.push_back(kind);                                \
        }                                                                   \
    } while (0)
#define LAUNCH(kind, ...) TIMED(kind, hipLaunchKernelGGL(__VA_ARGS__))
    const int sides = d.m.s_num - 1;
    const bool parts = ctx->parts > 0;
    const int groups = parts ? ctx->parts : 1;  // workgroups of the prepare and finish launches
    const bool fuse_grad = d.m.n_grad > 0 && !d.gates_on;
    const int grad_blocks = blocks_xcd(d.m.n_grad > 0 ? d.m.n_grad : 1);
    LAUNCH(K_PREPARE, k_prepare, dim3(groups), dim3(64), 0, s, d, 0);
    for (int slot = 0; slot < steps; ++slot) {
        if (d.gates_on) LAUNCH(K_GATE, k_gate_move, dim3(1), dim3(1024), 0, s, d, 0, 0);
        // gradients: a launch of their own at the batch's first step (the host may have written
        // the state) and where gates move the bed at the step start; otherwise the previous
        // step's CFL launch computed them
        if (d.m.n_grad > 0 && (slot == 0 || !fuse_grad)) {
            if (parts)
                LAUNCH(K_GRAD, k_gradients<true>, dim3(blocks_for(d.m.n_grad)), dim3(NHF_BLOCK), 0, s, d);
            else
                LAUNCH(K_GRAD, k_gradients<false>, dim3(blocks_for(d.m.n_grad)), dim3(NHF_BLOCK), 0, s, d);
        }
        if (!d.theta_one) {
            if (parts)
                LAUNCH(K_FACE, k_face_avg<true>, dim3(blocks_for(d.m.e_num)), dim3(NHF_BLOCK), 0, s, d);
            else
                LAUNCH(K_FACE, k_face_avg<false>, dim3(blocks_for(d.m.e_num)), dim3(NHF_BLOCK), 0, s, d);
        }
        if (sides > 0) TIMED(K_FLUX, launch_k_flux(d, sides, s));
        TIMED(K_DONOR, launch_k_donor(d, s));
        if (d.gates_on && d.g.n_chunks > 0)
            LAUNCH(K_GATE, k_gate_account, dim3(d.g.n_chunks), dim3(NHF_BLOCK), 0, s, d);
        TIMED(K_DEPTH, launch_k_depth(d, s));
        if (d.gates_on) LAUNCH(K_GATE, k_gate_move, dim3(1), dim3(1024), 0, s, d, 1, 0);
        const int gb = fuse_grad && slot + 1 < steps ? grad_blocks : 0;
        if (parts) {
            LAUNCH(K_CFL, k_cfl<true>, dim3(gb + d.cfl_blocks), dim3(NHF_BLOCK), 0, s, d, 1, gb);
            LAUNCH(K_FINISH, k_finish<true>, dim3(groups), dim3(1024), 0, s, d, 1, slot, steps);
        } else {
            LAUNCH(K_CFL, k_cfl<false>, dim3(gb + d.cfl_blocks), dim3(NHF_BLOCK), 0, s, d, 1, gb);
            LAUNCH(K_FINISH, k_finish<false>, dim3(1), dim3(1024), 0, s, d, 1, slot, steps);
        }
    }
#undef LAUNCH
#undef TIMED
    HIP_TRY(hipGetLastError());
    return 0;
}

static int32_t launch_any(nhf_ctx *ctx, int32_t steps) {
    if (steps < 1 || steps > ctx->d.p.batch_capacity) return set_error_text("nhf_launch_batch: steps out of range");
    if (ctx->pending_steps) return set_error_text("nhf_launch_batch: the previous batch was not waited for");
    hipStream_t s = ctx->stream;
    if (ctx->graph_enabled && !ctx->timing && s != nullptr) {
        if (!ctx->exec[steps]) {
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int32_t rc = launch_batch(ctx, steps, false);
            hipError_t end = hipStreamEndCapture(s, &graph);
            if (rc) return rc;
            if (end != hipSuccess) return set_error("hipStreamEndCapture", end);
            hipError_t inst = hipGraphInstantiate(&ctx->exec[steps], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (inst != hipSuccess) return set_error("hipGraphInstantiate", inst);
        }
        HIP_TRY(hipGraphLaunch(ctx->exec[steps], s));
    } else {
        int32_t rc = launch_batch(ctx, steps, ctx->timing != 0);
        if (rc) return rc;
    }
    if (ctx->parts)  // every part's control block, trace rows included, in one copy
        HIP_TRY(hipMemcpyAsync(ctx->host_ctl, ctx->d.ctl, sizeof(Ctl) * ctx->parts, hipMemcpyDeviceToHost, s));
    else
        HIP_TRY(hipMemcpyAsync(ctx->host_trace, ctx->d.ctl->trace, sizeof(uint32_t) * NHF_TRACE_COLS * steps,
                               hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(ctx->done, s));
    ctx->pending_steps = steps;
    return 0;
}

static int32_t wait_any(nhf_ctx *ctx, uint32_t *trace) {
    const int steps = ctx->pending_steps;
    if (!steps) return set_error_text("nhf_wait_batch: no batch in flight");
    ctx->pending_steps = 0;
    HIP_TRY(hipEventSynchronize(ctx->done));
    if (ctx->timing) {
        for (size_t i = 0; i < ctx->event_kind.size(); ++i) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, ctx->events[2 * i], ctx->events[2 * i + 1]));
            ctx->kernel_ms[ctx->event_kind[i]] += ms;
            ctx->kernel_count[ctx->event_kind[i]] += 1;
        }
    }
    if (!trace) return 0;
    const size_t words = sizeof(uint32_t) * NHF_TRACE_COLS * steps;
    if (ctx->parts)
        for (int p = 0; p < ctx->parts; ++p) memcpy(trace + (size_t)p * NHF_TRACE_COLS * steps, ctx->host_ctl[p].trace, words);
    else
        memcpy(trace, ctx->host_trace, words);
    return 0;
}

int32_t nhf_launch_batch(nhf_ctx *ctx, int32_t steps) {
    if (int32_t rc = single_only(ctx, "nhf_launch_batch")) return rc;
    return launch_any(ctx, steps);
}

int32_t nhf_wait_batch(nhf_ctx *ctx, uint32_t *trace) {
    if (int32_t rc = single_only(ctx, "nhf_wait_batch")) return rc;
    return wait_any(ctx, trace);
}

int32_t nhf_run_batch(nhf_ctx *ctx, int32_t steps, uint32_t *trace) {
    if (!ctx) return set_error_text("nhf_run_batch: null context");
    int32_t rc = nhf_launch_batch(ctx, steps);
    if (rc) return rc;
    return nhf_wait_batch(ctx, trace);
}

int32_t nhf_time_kernels(nhf_ctx *ctx, int32_t repeats, double *us) {
    if (int32_t rc = single_only(ctx, "nhf_time_kernels")) return rc;
    if (repeats < 1) return set_error_text("nhf_time_kernels: repeats must be positive");
    if (ctx->pending_steps) return set_error_text("nhf_time_kernels: a batch is in flight");
    const Dev &d = ctx->d;
    hipStream_t s = ctx->stream;
    hipEvent_t e0 = timing_event(ctx, 0), e1 = timing_event(ctx, 1);
    if (!e0 || !e1) return set_error_text("hipEventCreate failed");
    for (int i = 0; i < K_COUNT; ++i) us[i] = 0.0;
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(64), 0, s, d, 0);
    const int sides = d.m.s_num - 1;
    // the donor pass's repeats look at the cells the flux pass flagged, as the step's one launch
    Dev keep = d;
    keep.donor_keep = 1;
    for (int k = K_GRAD; k < K_COUNT; ++k) {
        if ((k == K_GRAD && d.m.n_grad == 0) || (k == K_FACE && d.theta_one) || (k == K_FLUX && sides <= 0) ||
            k == K_GATE)
            continue;
        HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < repeats; ++r) {
            switch (k) {
            case K_GRAD: hipLaunchKernelGGL(k_gradients<false>, dim3(blocks_for(d.m.n_grad)), dim3(NHF_BLOCK), 0, s, d); break;
            case K_FACE: hipLaunchKernelGGL(k_face_avg<false>, dim3(blocks_for(d.m.e_num)), dim3(NHF_BLOCK), 0, s, d); break;
            case K_FLUX: launch_k_flux(d, sides, s); break;
            case K_DONOR: launch_k_donor(keep, s); break;
            case K_DEPTH: launch_k_depth(d, s); break;
            case K_CFL: hipLaunchKernelGGL(k_cfl<false>, dim3(d.cfl_blocks), dim3(NHF_BLOCK), 0, s, d, 1, 0); break;
            case K_FINISH: hipLaunchKernelGGL(k_finish<false>, dim3(1), dim3(1024), 0, s, d, 0, 0, 0); break;
            }
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        us[k] = 1e3 * (double)ms / repeats;
    }
    return 0;
}

static int32_t compute_cfl_any(nhf_ctx *ctx) {
    const Dev &d = ctx->d;
    if (ctx->parts) {
        hipLaunchKernelGGL(k_cfl<true>, dim3(d.cfl_blocks), dim3(NHF_BLOCK), 0, ctx->stream, d, 0, 0);
        hipLaunchKernelGGL(k_finish<true>, dim3(ctx->parts), dim3(1024), 0, ctx->stream, d, 0, 0, 0);
    } else {
        hipLaunchKernelGGL(k_cfl<false>, dim3(d.cfl_blocks), dim3(NHF_BLOCK), 0, ctx->stream, d, 0, 0);
        hipLaunchKernelGGL(k_finish<false>, dim3(1), dim3(1024), 0, ctx->stream, d, 0, 0, 0);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhf_compute_cfl(nhf_ctx *ctx) {
    if (int32_t rc = single_only(ctx, "nhf_compute_cfl")) return rc;
    return compute_cfl_any(ctx);
}

// Budget of one part (the whole mesh for a single mesh): its accumulators summed in a fixed
// order in double-double, then cleared.
static int32_t flush_part(nhf_ctx *ctx, int part, double *out) {
    const Dev &d = ctx->d;
    int o0 = 0, o1 = d.acc_owners, terms = d.terms;
    if (ctx->parts) {
        const int4 r = ctx->part_range[part];
        o0 = (r.x - 1) >> 6;
        o1 = (r.x + padded(r.y - r.x) - 1) >> 6;
        terms = 4 + 6 * ctx->part_extra[part].y;
    }
    const size_t n = (size_t)(o1 - o0) * d.terms;
    if (!ctx->host_acc)
        HIP_TRY(hipHostMalloc((void **)&ctx->host_acc, sizeof(double2) * (size_t)d.acc_owners * d.terms,
                              hipHostMallocDefault));
    HIP_TRY(hipMemcpyAsync(ctx->host_acc, d.acc + (size_t)o0 * d.terms, sizeof(double2) * n, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(ctx->host_acc[i].x) || !isfinite(ctx->host_acc[i].y)) return 1;
    for (int t = 0; t < terms; ++t) {
        // two-sum accumulation of every wavefront's (sum, compensation) pair
        double hi = 0.0, lo = 0.0;
        for (int b = 0; b < o1 - o0; ++b) {
            const double2 a = ctx->host_acc[(size_t)b * d.terms + t];
            const double pair[2] = {a.x, a.y};
            for (double x : pair) {
                double s = hi + x;
                double bp = s - hi;
                lo += (hi - (s - bp)) + (x - bp);
                hi = s;
            }
        }
        out[2 * t] = hi;
        out[2 * t + 1] = lo;
    }
    hipLaunchKernelGGL(k_clear_acc, dim3(64), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, part);
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhf_flush_budget(nhf_ctx *ctx, double *out) {
    if (int32_t rc = single_only(ctx, "nhf_flush_budget")) return rc;
    return flush_part(ctx, 0, out);
}

static int32_t limit_part(nhf_ctx *ctx, int part, float *values, int32_t *ids, int32_t reset) {
    Ctl *ctl = ctx->d.ctl + part;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (values) HIP_TRY(hipMemcpy(values, ctl->limit_value, sizeof(float) * 5, hipMemcpyDeviceToHost));
    if (ids) HIP_TRY(hipMemcpy(ids, ctl->limit_ids, sizeof(int32_t) * 3, hipMemcpyDeviceToHost));
    if (reset) {
        hipLaunchKernelGGL(k_reset_limit, dim3(1), dim3(64), 0, ctx->stream, ctx->d, part);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

int32_t nhf_limit_capture(nhf_ctx *ctx, float *values, int32_t *ids, int32_t reset) {
    if (int32_t rc = single_only(ctx, "nhf_limit_capture")) return rc;
    return limit_part(ctx, 0, values, ids, reset);
}

static int32_t cfl_state_part(nhf_ctx *ctx, int part, float *ndt, int32_t *side) {
    Ctl *ctl = ctx->d.ctl + part;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ndt) HIP_TRY(hipMemcpy(ndt, &ctl->ndt, sizeof(float), hipMemcpyDeviceToHost));
    if (side) HIP_TRY(hipMemcpy(side, &ctl->cfl_arg, sizeof(int32_t), hipMemcpyDeviceToHost));
    return 0;
}

int32_t nhf_cfl_state(nhf_ctx *ctx, float *ndt, int32_t *side) {
    if (int32_t rc = single_only(ctx, "nhf_cfl_state")) return rc;
    return cfl_state_part(ctx, 0, ndt, side);
}

int32_t nhf_sync(nhf_ctx *ctx) {
    if (!ctx) return set_error_text("nhf_sync: null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int32_t nhf_set_gates(nhf_ctx *ctx, const nhf_gates *gates) {
    if (int32_t rc = single_only(ctx, "nhf_set_gates")) return rc;
    if (ctx->pending_steps) return set_error_text("nhf_set_gates: a batch is in flight");
    Dev &d = ctx->d;
    if (gates) {
        const nhf_gates &g = *gates;
        if (g.count < 1 || g.n_cells < 0 || g.n_items < 0 || g.n_chunks < 0 || !(g.opening_seconds > 0.0f))
            return set_error_text("nhf_set_gates: invalid counts or opening time");
        if (!g.control || !g.cell_gate || !g.pond || !g.closed || !g.flags || !g.ticks ||
            (g.n_cells && !g.cells) || (g.n_chunks && (!g.items || !g.chunks || !g.acc)))
            return set_error_text("nhf_set_gates: an array is NULL");
        std::vector<int32_t> chunks((size_t)g.n_chunks * 4);
        if (g.n_chunks)
            HIP_TRY(hipMemcpy(chunks.data(), g.chunks, sizeof(int32_t) * chunks.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < g.n_chunks; ++c) {
            const int32_t *ch = &chunks[4 * (size_t)c];
            const int kind = ch[2] & 255, width = kind == 0 ? 2 : 6;
            if (ch[0] < 0 || ch[1] < 1 || ch[1] > NHF_BLOCK || ch[0] + ch[1] > g.n_items || kind > 1 ||
                (ch[2] >> 8) - 1 >= g.count || ch[3] < 0 || ch[3] + width > 10 * g.count + 8)
                return set_error_text("nhf_set_gates: accounting chunk out of range");
        }
        // every index the kernels follow is checked once here, on host copies
        const int e_num = d.m.e_num, s_num = d.m.s_num;
        std::vector<int32_t> control(g.count), cells(g.n_cells), cell_gate(e_num), items(g.n_items);
        HIP_TRY(hipMemcpy(control.data(), g.control, sizeof(int32_t) * g.count, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(cell_gate.data(), g.cell_gate, sizeof(int32_t) * e_num, hipMemcpyDeviceToHost));
        if (g.n_cells) HIP_TRY(hipMemcpy(cells.data(), g.cells, sizeof(int32_t) * g.n_cells, hipMemcpyDeviceToHost));
        if (g.n_items) HIP_TRY(hipMemcpy(items.data(), g.items, sizeof(int32_t) * g.n_items, hipMemcpyDeviceToHost));
        for (int k = 0; k < g.count; ++k)
            if (control[k] < 1 || control[k] >= e_num) return set_error_text("nhf_set_gates: control cell out of range");
        for (int i = 0; i < g.n_cells; ++i)
            if (cells[i] < 1 || cells[i] >= e_num || cell_gate[cells[i]] < 0 || cell_gate[cells[i]] >= g.count)
                return set_error_text("nhf_set_gates: gate cell out of range or without a gate");
        for (int c = 0; c < g.n_chunks; ++c) {
            const int32_t *ch = &chunks[4 * (size_t)c];
            for (int j = ch[0]; j < ch[0] + ch[1]; ++j) {
                const int id = (ch[2] & 255) == 0 ? (int)((uint32_t)items[j] & 0x7fffffffu) : items[j];
                if (id < 1 || id >= ((ch[2] & 255) == 0 ? s_num : e_num))
                    return set_error_text("nhf_set_gates: accounting item out of range");
            }
        }
        ctx->gate_chunks.swap(chunks);
        d.g = g;
    } else {
        ctx->gate_chunks.clear();
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i <= NHF_MAX_BATCH; ++i)
        if (ctx->exec[i]) {
            (void)hipGraphExecDestroy(ctx->exec[i]);
            ctx->exec[i] = nullptr;
        }
    d.gates_on = gates != nullptr;
    return 0;
}

int32_t nhf_gate_check(nhf_ctx *ctx) {
    if (!ctx->d.gates_on) return set_error_text("nhf_gate_check: no gates bound");
    if (ctx->pending_steps) return set_error_text("nhf_gate_check: a batch is in flight");
    hipLaunchKernelGGL(k_gate_move, dim3(1), dim3(1024), 0, ctx->stream, ctx->d, 0, 1);
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhf_flush_gate_budget(nhf_ctx *ctx, double *out) {
    if (!ctx) return set_error_text("nhf_flush_gate_budget: null context");
    const Dev &d = ctx->d;
    if (!d.gates_on) return set_error_text("nhf_flush_gate_budget: no gates bound");
    if (ctx->pending_steps) return set_error_text("nhf_flush_gate_budget: a batch is in flight");
    const int terms = 10 * d.g.count + 8, n = d.g.n_chunks;
    std::vector<double> acc((size_t)n * 12);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) {
        HIP_TRY(hipMemcpy(acc.data(), d.g.acc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemsetAsync(d.g.acc, 0, sizeof(double) * acc.size(), ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    // per term, the chunks' compensated pairs added in chunk order (double-double)
    std::vector<double> hi(terms, 0.0), lo(terms, 0.0);
    for (int c = 0; c < n; ++c) {
        const int32_t *ch = &ctx->gate_chunks[4 * (size_t)c];
        const int width = (ch[2] & 255) == 0 ? 2 : 6;
        for (int k = 0; k < width; ++k) {
            const int t = ch[3] + k;
            for (int part = 0; part < 2; ++part) {
                const double x = acc[(size_t)c * 12 + 2 * k + part];
                const double s = hi[t] + x, bp = s - hi[t];
                lo[t] += (hi[t] - (s - bp)) + (x - bp);
                hi[t] = s;
            }
        }
    }
    for (int t = 0; t < terms; ++t) {
        out[2 * t] = hi[t];
        out[2 * t + 1] = lo[t];
    }
    return 0;
}

int32_t nhf_settle_extrema(nhf_ctx *ctx) {
    if (!ctx) return set_error_text("nhf_settle_extrema: null context");
    if (ctx->pending_steps) return set_error_text("nhf_settle_extrema: a batch is in flight");
    const Dev &d = ctx->d;
    hipLaunchKernelGGL(k_settle_extrema, dim3(blocks_for(d.m.e_num) < 2048 ? blocks_for(d.m.e_num) : 2048),
                       dim3(NHF_BLOCK), 0, ctx->stream, d);
    HIP_TRY(hipGetLastError());
    return 0;
}

// The rain state of one part (part 0 of a single mesh): the part's one rain pixel when it has
// at most one, the capacities copied through cap_stage.
static int32_t rain_state_part(nhf_ctx *ctx, int part, int32_t uniform, float rain_time, const float *class_capacity,
                               const char *name) {
    if (ctx->pending_steps) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "%s: a batch is in flight", name);
        return -2;
    }
    const int pixels = ctx->parts ? ctx->part_pixels[part].y : ctx->d.m.n_pixels;
    const int pixel = ctx->parts ? ctx->part_pixels[part].x : 0;
    // the flux pass advances the uniform state, so a mesh without sides keeps the per-cell one
    if (uniform && (pixels > 1 || !class_capacity || ctx->d.m.s_num <= 1)) {
        snprintf(nhf_error_text, sizeof nhf_error_text,
                 "%s: the uniform state needs one rain pixel, sides and the class capacities", name);
        return -2;
    }
    if (uniform)
        HIP_TRY(hipMemcpyAsync(ctx->d.cap_stage, class_capacity, sizeof(float) * 256, hipMemcpyHostToDevice,
                               ctx->stream));
    hipLaunchKernelGGL(k_set_rain_state, dim3(1), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, part, (int)(uniform != 0),
                       rain_time, pixel);
    HIP_TRY(hipGetLastError());
    // the host's capacities are read by the copy before this returns
    if (uniform) HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int32_t nhf_set_rain_state(nhf_ctx *ctx, int32_t uniform, float rain_time, const float *class_capacity) {
    if (int32_t rc = single_only(ctx, "nhf_set_rain_state")) return rc;
    return rain_state_part(ctx, 0, uniform, rain_time, class_capacity, "nhf_set_rain_state");
}

int32_t nhf_set_graph(nhf_ctx *ctx, int32_t enable) {
    if (!ctx) return set_error_text("nhf_set_graph: null context");
    ctx->graph_enabled = enable != 0;
    return 0;
}

int32_t nhf_set_tuning(nhf_ctx *ctx, int32_t cfl_blocks, int32_t cell_block, int32_t cells_per_thread) {
    if (!ctx) return set_error_text("nhf_set_tuning: null context");
    if (ctx->parts) cfl_blocks = ctx->d.cfl_blocks;  // each part's CFL grid is fixed at creation
    if (cfl_blocks < 8 || cfl_blocks > ctx->d.cfl_max || (cfl_blocks & 7))
        return set_error_text("nhf_set_tuning: CFL grid out of range or not a multiple of 8");
    if (cell_block != 128 && cell_block != 256 && cell_block != 512)
        return set_error_text("nhf_set_tuning: cell workgroup size must be 128, 256 or 512");
    if (cells_per_thread != 1 && cells_per_thread != 2)
        return set_error_text("nhf_set_tuning: cells per thread must be 1 or 2");
    if (ctx->pending_steps) return set_error_text("nhf_set_tuning: a batch is in flight");
    if (cfl_blocks != ctx->d.cfl_blocks || cell_block != ctx->d.cell_block ||
        cells_per_thread != ctx->d.cells_per_thread) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i <= NHF_MAX_BATCH; ++i)
            if (ctx->exec[i]) {
                (void)hipGraphExecDestroy(ctx->exec[i]);
                ctx->exec[i] = nullptr;
            }
        ctx->d.cfl_blocks = cfl_blocks;
        ctx->d.cell_block = cell_block;
        ctx->d.cells_per_thread = cells_per_thread;
        ctx->d.depth_blocks = cell_grid(ctx->d.m.e_num, cell_block * cells_per_thread);
    }
    return 0;
}

int32_t nhf_tuning_range(nhf_ctx *ctx, int32_t *out4) {
    if (!ctx) return set_error_text("nhf_tuning_range: null context");
    out4[0] = ctx->d.cfl_blocks;
    out4[1] = ctx->d.cfl_max;
    out4[2] = ctx->d.cell_block;
    out4[3] = ctx->d.cells_per_thread;
    return 0;
}

int32_t nhf_set_timing(nhf_ctx *ctx, int32_t enable) {
    if (!ctx) return set_error_text("nhf_set_timing: null context");
    ctx->timing = enable != 0;
    return 0;
}

int32_t nhf_kernel_times(nhf_ctx *ctx, double *ms, int64_t *count, int32_t n, int32_t reset) {
    if (!ctx) return set_error_text("nhf_kernel_times: null context");
    for (int i = 0; i < n && i < K_COUNT; ++i) {
        ms[i] = ctx->kernel_ms[i];
        count[i] = ctx->kernel_count[i];
    }
    if (reset) {
        memset(ctx->kernel_ms, 0, sizeof ctx->kernel_ms);
        memset(ctx->kernel_count, 0, sizeof ctx->kernel_count);
    }
    return 0;
}

int32_t nhf_clock_probe(const uint64_t *now, const uint64_t *target, const uint64_t *origin, const float *candidate,
                        float min_dt, int32_t n, uint32_t *out_words, uint64_t *out_ticks, float *out_times) {
    if (n <= 0) return 0;
    char *buf = nullptr;
    size_t in_bytes = n * (3 * sizeof(uint64_t) + sizeof(float));
    size_t out_bytes = n * (4 * sizeof(uint32_t) + 2 * sizeof(uint64_t) + 2 * sizeof(float));
    size_t scratch = n * sizeof(Ctl);
    HIP_TRY(hipMalloc((void **)&buf, in_bytes + out_bytes + scratch));
    uint64_t *d_now = (uint64_t *)buf, *d_target = d_now + n, *d_origin = d_target + n;
    float *d_cand = (float *)(d_origin + n);
    uint32_t *d_words = (uint32_t *)(d_cand + n);
    uint64_t *d_ticks = (uint64_t *)(d_words + 4 * n);
    float *d_times = (float *)(d_ticks + 2 * n);
    Ctl *d_scratch = (Ctl *)(buf + in_bytes + out_bytes);
    int32_t rc = 0;
    hipError_t e = hipSuccess;
    if ((e = hipMemcpy(d_now, now, n * sizeof(uint64_t), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_target, target, n * sizeof(uint64_t), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_origin, origin, n * sizeof(uint64_t), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_cand, candidate, n * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
        rc = set_error("nhf_clock_probe upload", e);
    } else {
        hipLaunchKernelGGL(k_clock_probe, dim3(blocks_for(n)), dim3(NHF_BLOCK), 0, 0, d_now, d_target, d_origin,
                           d_cand, min_dt, n, d_words, d_ticks, d_times, d_scratch);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess ||
            (e = hipMemcpy(out_words, d_words, 4 * n * sizeof(uint32_t), hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(out_ticks, d_ticks, 2 * n * sizeof(uint64_t), hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(out_times, d_times, 2 * n * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
            rc = set_error("nhf_clock_probe", e);
    }
    (void)hipFree(buf);
    return rc;
}

// --------------------------------------------------------------------------- no allocation

}  // extern "C"

// Scratch of nhf_run_static: a file-scope device array (zero-filled when the code object loads),
// so that entry point needs no workspace and allocates nothing. Each GPU the library runs on holds
// its own copy, at its own address, until the process ends. In host code `scratch` names a host
// stand-in, not device memory: only the address hipGetSymbolAddress returns for the current
// device may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[NHF_STATIC_SCRATCH_BYTES];

#define NHF_STATIC_DEVICES 64
static std::mutex g_static_lock;
static char *g_static_ptr[NHF_STATIC_DEVICES];

// The current device's address of `scratch`, looked up once per device.
static int32_t static_scratch(char **out) {
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (device < 0 || device >= NHF_STATIC_DEVICES) return set_error_text("nhf_run_static: device index out of range");
    std::lock_guard<std::mutex> hold(g_static_lock);
    if (!g_static_ptr[device]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return set_error_text("nhf_run_static: hipGetSymbolAddress returned NULL");
        g_static_ptr[device] = (char *)ptr;
    }
    *out = g_static_ptr[device];
    return 0;
}

// Scratch bytes nhf_run_static needs for a mesh: nhf_create's workspace plus the budget totals.
static size_t static_bytes(const nhf_mesh *mesh, const nhf_params *params) {
    return align_up(carve(mesh, params).total) + sizeof(double) * 2 * (size_t)(4 + 6 * mesh->b_num);
}

extern "C" {

int32_t nhf_run_static(const nhf_mesh *mesh, const nhf_params *params, uint64_t now_ticks, uint64_t target_ticks,
                       uint64_t origin_ticks, const float *coefficients, const float *rain_rates, int32_t raining,
                       int32_t steps, uint32_t *trace, double *budget, void *stream) {
    if (!mesh || !params || !trace) return set_error_text("nhf_run_static: null argument");
    if (int32_t rc = check_mesh("nhf_run_static", mesh, params, false)) return rc;
    if (steps < 1 || steps > params->batch_capacity) return set_error_text("nhf_run_static: steps out of range");
    // the scratch keeps the previous call's contents (maybe another mesh's): no forcing carries over
    if ((mesh->b_num > 0 && !coefficients) || (raining && mesh->n_pixels > 0 && !rain_rates))
        return set_error_text("nhf_run_static: coefficients and, when raining, rain_rates are required");
    const size_t need = static_bytes(mesh, params);
    if (need > NHF_STATIC_SCRATCH_BYTES) {
        snprintf(nhf_error_text, sizeof nhf_error_text,
                 "nhf_run_static: this mesh needs %zu bytes of scratch, more than the static maximum of %zu "
                 "(use nhf_create with a workspace)", need, (size_t)NHF_STATIC_SCRATCH_BYTES);
        return -2;
    }
    char *ws = nullptr;
    if (int32_t rc = static_scratch(&ws)) return rc;
    const Carve cv = carve(mesh, params);
    // A context on the stack over the static scratch: the launches are nhf_create's, the vectors
    // stay empty (eager launches without timing touch none of them).
    nhf_ctx ctx{};
    bind_dev(ctx.d, mesh, params, nullptr, ws, cv);
    ctx.stream = (hipStream_t)stream;
    double *totals = (double *)(ws + align_up(cv.total));
    const int terms = ctx.d.terms;
    int32_t rc = 0;
    // the per-cell rain state: the scratch holds whatever the last call left
    hipLaunchKernelGGL(k_set_rain_state, dim3(1), dim3(NHF_BLOCK), 0, ctx.stream, ctx.d, 0, 0, 0.0f, 0);
    if ((rc = reset_part(&ctx, 0, now_ticks)) || (rc = forcing_part(&ctx, 0, coefficients, rain_rates, raining)) ||
        (rc = compute_cfl_any(&ctx)) || (rc = begin_part(&ctx, 0, target_ticks, origin_ticks)) ||
        (rc = launch_batch(&ctx, steps, false)))
        return rc;
    const Dev &d = ctx.d;
    hipStream_t s = ctx.stream;
    hipLaunchKernelGGL(k_settle_extrema, dim3(blocks_for(d.m.e_num) < 2048 ? blocks_for(d.m.e_num) : 2048),
                       dim3(NHF_BLOCK), 0, s, d);
    if (budget) hipLaunchKernelGGL(k_flush_terms, dim3(terms), dim3(NHF_BLOCK), 0, s, d, totals);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(trace, d.ctl->trace, sizeof(uint32_t) * NHF_TRACE_COLS * steps, hipMemcpyDeviceToHost, s));
    if (budget) HIP_TRY(hipMemcpyAsync(budget, totals, sizeof(double) * 2 * terms, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (budget)
        for (int i = 0; i < 2 * terms; ++i)
            if (!isfinite(budget[i])) return 1;
    return 0;
}

// --------------------------------------------------------------------------- parts (nhu_*)

size_t nhu_workspace_bytes(const nhf_mesh *mesh, const nhf_params *params, const nhf_part *parts, int32_t count) {
    if (!mesh || !params || !parts) return 0;
    PartPlan plan;
    if (plan_parts(mesh, parts, count, plan)) return 0;
    return carve(mesh, params, &plan).total;
}

int32_t nhu_create(nhf_ctx **out, const nhf_mesh *mesh, const nhf_params *params, const nhf_part *parts,
                   int32_t count, void *workspace, size_t workspace_bytes, void *stream) {
    if (!mesh || !parts) return set_error_text("nhu_create: null argument");
    PartPlan plan;
    if (const char *why = plan_parts(mesh, parts, count, plan)) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "nhu_create: %s", why);
        return -2;
    }
    return create_ctx(out, mesh, params, &plan, workspace, workspace_bytes, stream);
}

int32_t nhu_reset(nhf_ctx *ctx, int32_t part, uint64_t now_ticks) {
    if (int32_t rc = part_only(ctx, part, "nhu_reset")) return rc;
    return reset_part(ctx, part, now_ticks);
}

int32_t nhu_set_forcing(nhf_ctx *ctx, int32_t part, const float *coefficients, const float *rain_rates,
                        int32_t raining) {
    if (int32_t rc = part_only(ctx, part, "nhu_set_forcing")) return rc;
    return forcing_part(ctx, part, coefficients, rain_rates, raining);
}

int32_t nhu_begin(nhf_ctx *ctx, int32_t part, uint64_t target_ticks, uint64_t origin_ticks) {
    if (int32_t rc = part_only(ctx, part, "nhu_begin")) return rc;
    return begin_part(ctx, part, target_ticks, origin_ticks);
}

int32_t nhu_launch_batch(nhf_ctx *ctx, int32_t steps) {
    if (int32_t rc = part_only(ctx, 0, "nhu_launch_batch")) return rc;
    return launch_any(ctx, steps);
}

int32_t nhu_wait_batch(nhf_ctx *ctx, uint32_t *trace) {
    if (int32_t rc = part_only(ctx, 0, "nhu_wait_batch")) return rc;
    return wait_any(ctx, trace);
}

int32_t nhu_run_batch(nhf_ctx *ctx, int32_t steps, uint32_t *trace) {
    if (!ctx) return set_error_text("nhu_run_batch: null context");
    int32_t rc = nhu_launch_batch(ctx, steps);
    if (rc) return rc;
    return nhu_wait_batch(ctx, trace);
}

int32_t nhu_compute_cfl(nhf_ctx *ctx) {
    if (int32_t rc = part_only(ctx, 0, "nhu_compute_cfl")) return rc;
    return compute_cfl_any(ctx);
}

int32_t nhu_flush_budget(nhf_ctx *ctx, int32_t part, double *out) {
    if (int32_t rc = part_only(ctx, part, "nhu_flush_budget")) return rc;
    return flush_part(ctx, part, out);
}

int32_t nhu_limit_capture(nhf_ctx *ctx, int32_t part, float *values, int32_t *ids, int32_t reset) {
    if (int32_t rc = part_only(ctx, part, "nhu_limit_capture")) return rc;
    return limit_part(ctx, part, values, ids, reset);
}

int32_t nhu_cfl_state(nhf_ctx *ctx, int32_t part, float *ndt, int32_t *side) {
    if (int32_t rc = part_only(ctx, part, "nhu_cfl_state")) return rc;
    return cfl_state_part(ctx, part, ndt, side);
}

int32_t nhu_set_rain_state(nhf_ctx *ctx, int32_t part, int32_t uniform, float rain_time, const float *class_capacity) {
    if (int32_t rc = part_only(ctx, part, "nhu_set_rain_state")) return rc;
    return rain_state_part(ctx, part, uniform, rain_time, class_capacity, "nhu_set_rain_state");
}

}  // extern "C"
