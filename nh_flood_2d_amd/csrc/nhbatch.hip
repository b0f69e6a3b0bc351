// libnhflood.so, scenario-batched step: up to 64 scenarios ("slots") on one shared mesh, each
// with its own state, forcing, clock, dt, CFL, limiter counters, trace and water budget, advanced
// together by the same launches.
//
// Layout: every per-scenario array is [entity][slot] with the slot innermost, so the S slots of
// one cell or side are S consecutive words: a wavefront (64 lanes = 64 / S entities x S slots)
// reads state in whole cache lines, and the static mesh arrays (side lists, lengths, ids, bed,
// area, ...) are fetched once per entity for all of its slots.
//
// Each slot runs exactly the single-mesh step (nhflood.hip) with the same arithmetic in the same
// order, so a scenario gives the same bits alone and in a batch; only the FP64 budget is summed
// in a different grouping. Slot s is idle in a step when dt[s] == 0 (its clock is at its target,
// failed, or the slot is unused): its lanes load and store nothing.
//
// Kernels per step (reference lines as in nhflood.hip):
//   kb_gradients, kb_face_avg (2:1 meshes, sita < 1), kb_flux, kb_donor, kb_depth (grid-stride,
//   per-thread FP64 budget), kb_cfl (grid-stride, per-slot partials), kb_finish (one workgroup
//   per slot: CFL reduction, boundary budget terms, commit, prepare).
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "nhf_device.h"

// kb_cfl loads the next round's side records while the current round's rows are in flight.
#ifndef NHB_CFL_PREFETCH
#define NHB_CFL_PREFETCH 1
#endif

struct BDev {
    nhf_mesh m;  // static mesh arrays (the per-scenario state pointers of nhf_mesh are unused)
    nhf_params p;
    int32_t shift, terms, cfl_blocks, depth_blocks, n_bcells, theta_one, b_num, n_pixels;
    int32_t cfl_max, depth_max;  // allocated partial and accumulator blocks
    // per-scenario state, [entity][slot]
    float *depth, *rain_time, *capacity, *max_depth, *max_speed, *q;
    uint64_t *peak_tick;
    const float *manning;  // [s_num][slot] or NULL: mesh manning shared by every slot
    // [256 classes][slot] or NULL: Manning n per land class, side n rebuilt from side_class
    const float *class_manning;
    uint32_t *side_class;  // [s_num] class of the lower | higher cell << 16, NHB_NO_CLASS outside
    // boundary-cell list (static)
    const int32_t *bslot, *bcell_start;
    // Uniform rain state of a slot with one rain pixel (rain_uniform[slot] = 1): every cell's rain
    // time is rt_u[slot] and its capacity cap_u[class][slot], so the depth pass neither reads nor
    // writes the per-cell arrays (kb_settle_state writes them out); the flux pass advances
    // rt_u and the capacity of each used class (class_used) once per step (advance_rain_b).
    int32_t *rain_uniform;        // [slot]
    float *rt_u;                  // [slot]
    float *cap_u;                 // [256 classes][slot]
    int32_t *class_used;          // [256]: 1 for a class some real cell has
    float *cap_stage;             // [256]: nhb_set_rain_state's capacities on their way in
    // Per cell: nonzero when the donor pass must look at the cell: bit 0 set by the flux pass
    // when one of the cell's outflows exceeds its certified share in any slot
    // (donor_share_exceeded), bit 1 standing for boundary cells and long side lists.
    uint8_t *donor_flag;
    // scratch
    Ctl *ctl;                     // [slot]
    float *dt;                    // [slot]; 0 = the slot does not step
    int32_t *raining;             // [slot]
    float *coefficients;          // [slot][b][2]
    float *values, *level_values; // [slot][b]
    float *rain_rates;            // [slot][n_pixels]
    float *bdonor;                // [boundary cell][slot]
    float *face_qn;               // [e_num][4][slot]
    float *grad;                  // [n_grad][2][slot]
    uint64_t *cfl_key;            // [cfl_blocks][slot]
    int32_t *cfl_idx;
    double2 *acc;                 // [depth_blocks][2][slot]: infiltration, non-level floor
    float *bvals;                 // [boundary cell][6][slot]
    double2 *bacc;                // [slot][terms]: boundary-cell terms
    double2 *flush;               // [terms]
};

// Host -> device staging of one batch: begin words and forcing per slot, uploaded in one copy.
struct Stage {
    uint64_t *target, *origin;  // [slot]
    int32_t *flags;             // [slot]: see STAGE_*
    float *coeff;               // [slot][b][2]
    float *rain;                // [slot][n_pixels]
};
enum { STAGE_BEGIN = 1, STAGE_COEFF = 2, STAGE_RAIN = 4, STAGE_RAINING = 8, STAGE_RAINING_ON = 16 };

// Class half-word of the exterior cell in side_class: its n is zero.
#define NHB_NO_CLASS 0xffffu

// Index of entity e, slot s in a per-scenario array; `sh` is log2 of the slot count in scope.
#define IX(e) ((((size_t)(e)) << sh) | (size_t)s)

// Entity of flat thread index t. With 64 slots a wavefront is exactly one entity; readfirstlane
// makes that explicit, so the entity index and everything derived from it (static mesh loads,
// loop bounds over side lists, array bases) are wave-uniform and live in scalar registers.
template <bool W>
__device__ __forceinline__ int entity_of(size_t t, int sh) {
    return W ? __builtin_amdgcn_readfirstlane((int)(t >> 6)) : (int)(t >> sh);
}

// --------------------------------------------------------------------------- clock per slot

// prepare_step (solver_compact.py:364-382) for slot s.
__device__ void prepare_slot_b(const BDev &d, int s, int step) {
    const int sh = d.shift;
    Ctl *c = &d.ctl[s];
    clock_choose(c, c->ndt, d.p.min_dt, step);
    d.dt[s] = c->active ? c->dt : 0.0f;
    if (!c->active) return;
    c->limited_cells = 0;
    c->limited_faces = 0;
    c->min_scale_bits = __float_as_uint(1.0f);
    const float dt = c->dt, local = c->local, end_local = c->end_local;
    for (int bi = 0; bi < d.b_num; ++bi) {
        float c0 = d.coefficients[(s * d.b_num + bi) * 2], c1 = d.coefficients[(s * d.b_num + bi) * 2 + 1];
        d.values[s * d.b_num + bi] = c0 + c1 * (local + dt / 2.0f);
        d.level_values[s * d.b_num + bi] = c0 + c1 * end_local;
    }
    if (c->ndt < c->limit_value[0]) {
        int si = c->cfl_arg;
        int el = d.m.side_cells[2 * si], eh = d.m.side_cells[2 * si + 1];
        float zl = d.m.bed[el], zh = d.m.bed[eh], hl = d.depth[IX(el)], hh = d.depth[IX(eh)];
        float head = fmx(zl + hl, zh + hh);
        float h = fmx(head - fmx(fmx(zl, zh), -999.0f), 0.01f);
        c->limit_value[0] = c->ndt;
        c->limit_value[1] = hl;
        c->limit_value[2] = hh;
        c->limit_value[3] = h;
        c->limit_value[4] = d.q[IX(si)] / fmx(h, 1e-30f);
        c->limit_ids[0] = si;
        c->limit_ids[1] = el;
        c->limit_ids[2] = eh;
    }
}

// --------------------------------------------------------------------------- small kernels

__global__ void kb_reset(BDev d, int s, uint64_t now) {
    const int S = 1 << d.shift;
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int stride = gridDim.x * blockDim.x;
    if (tid == 0) {
        Ctl *c = &d.ctl[s];
        c->now = c->next = c->origin = c->step_end = now;
        c->dt = c->local = c->end_local = 0.0f;
        c->active = 0;
        c->slot = 0;
        c->limited_cells = c->limited_faces = 0;
        c->min_scale_bits = __float_as_uint(1.0f);
        c->limit_value[0] = INFINITY;
        d.dt[s] = 0.0f;
    }
    for (int i = tid; i < d.depth_max * 2; i += stride) d.acc[(size_t)i * S + s] = make_double2(0.0, 0.0);
    for (int i = tid; i < d.terms; i += stride) d.bacc[(size_t)s * d.terms + i] = make_double2(0.0, 0.0);
}

// Apply the staged begin words and forcing: one workgroup per slot.
__global__ void kb_apply(BDev d, Stage st) {
    const int s = blockIdx.x;
    const int flags = st.flags[s];
    Ctl *c = &d.ctl[s];
    if (threadIdx.x == 0) {
        if (flags & STAGE_BEGIN) {
            c->next = st.target[s];
            c->origin = st.origin[s];
            c->slot = 0;
        }
        if (flags & STAGE_RAINING) d.raining[s] = (flags & STAGE_RAINING_ON) ? 1 : 0;
    }
    if (flags & STAGE_COEFF)
        for (int i = threadIdx.x; i < 2 * d.b_num; i += blockDim.x)
            d.coefficients[s * 2 * d.b_num + i] = st.coeff[s * 2 * d.b_num + i];
    if (flags & STAGE_RAIN)
        for (int i = threadIdx.x; i < d.n_pixels; i += blockDim.x)
            d.rain_rates[(size_t)s * d.n_pixels + i] = st.rain[(size_t)s * d.n_pixels + i];
}

// The rain state a step of a uniform slot advances to, by the slot's thread of the first side
// entity in the flux pass: where the slot's one pixel rains, every cell's rain time becomes
// rt_u + dt and its capacity its class's Horton value there, the arithmetic of the depth pass
// (solver_compact.py:543-548) once per class instead of per cell. The depth pass of the same
// step reads it.
__device__ void advance_rain_b(const BDev &d, int s, float dt) {
    const int sh = d.shift;
    if (!d.rain_uniform[s] || !d.raining[s] || !(d.rain_rates[s] > 0.0f)) return;
    const float rt = d.rt_u[s] + dt;
    for (int k = 0; k < 256; ++k) {
        if (!d.class_used[k]) continue;
        const float *hp = d.m.class_horton + 3 * k;
        const float f0 = hp[0], fc = hp[1], hk = hp[2];
        d.cap_u[((size_t)k << sh) | (size_t)s] = fc + (f0 - fc) * det_expf(-hk * rt / 3600.0f);
    }
    d.rt_u[s] = rt;
}

// First step of a batch: clear the trace rows of every slot (one workgroup per slot), then
// prepare step 0.
__global__ void kb_prepare(BDev d, int steps) {
    const int s = blockIdx.x;
    for (int i = threadIdx.x; i < steps * NHF_TRACE_COLS; i += blockDim.x) d.ctl[s].trace[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) prepare_slot_b(d, s, 0);
}

// The classes real cells have (for the uniform rain state), once per context.
__global__ void kb_class_used(BDev d) {
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x + 1; cell < d.m.e_num; cell += gridDim.x * blockDim.x)
        d.class_used[meta_class(d.m.cell_meta[cell])] = 1;
}

// Make a slot's state arrays current (slot < 0: every slot): the deferred extrema of rising cells
// (a negative max_depth: the maximum is the depth, reached at the end of the slot's last step)
// and, in the uniform rain state, every cell's rain time and capacity.
__global__ __launch_bounds__(NHF_BLOCK) void kb_settle_state(BDev d, int slot) {
    const int sh = d.shift;
    const int S = 1 << sh;
    const int s0 = slot < 0 ? 0 : slot, s1 = slot < 0 ? S : slot + 1;
    for (int cell = blockIdx.x * NHF_BLOCK + threadIdx.x + 1; cell < d.m.e_num; cell += gridDim.x * NHF_BLOCK) {
        const int cls = meta_class(d.m.cell_meta[cell]);
        for (int s = s0; s < s1; ++s) {
            if (d.max_depth[IX(cell)] < 0.0f) {
                d.max_depth[IX(cell)] = d.depth[IX(cell)];
                d.peak_tick[IX(cell)] = d.ctl[s].now;
            }
            if (d.rain_uniform[s]) {
                d.rain_time[IX(cell)] = d.rt_u[s];
                d.capacity[IX(cell)] = d.cap_u[((size_t)cls << sh) | (size_t)s];
            }
        }
    }
}

// Enter (uniform != 0) or leave the uniform rain state of one slot; `cap` = [256] capacity by
// class, `rt` the rain time of every cell (one workgroup).
__global__ void kb_set_rain_state(BDev d, int s, int uniform, float rt, const float *cap) {
    const int sh = d.shift;
    if (uniform)
        for (int k = threadIdx.x; k < 256; k += blockDim.x) d.cap_u[((size_t)k << sh) | (size_t)s] = cap[k];
    if (threadIdx.x == 0) {
        d.rain_uniform[s] = uniform ? 1 : 0;
        d.rt_u[s] = rt;
    }
}

__global__ void kb_reset_limit(BDev d, int s) {
    if (threadIdx.x == 0) d.ctl[s].limit_value[0] = INFINITY;
}

// --------------------------------------------------------------------------- hanging gradients

__device__ __forceinline__ void face_difference_b(const BDev &d, int sh, int s, int cell, int dir, int axis,
                                                  float &value, float &distance, int &usable) {
    const nhf_mesh &m = d.m;
    float v = 0.0f, dist = 0.0f, total = 0.0f;
    int ok = 1;
    float eta = m.bed[cell] + d.depth[IX(cell)];
    uint32_t cnt = m.csr_count[cell];
    int start = m.csr_start[cell];
    for (int k = 0; k < dir; ++k) start += face_n(cnt, k);
    int n = face_n(cnt, dir);
    for (int j = 0; j < n; ++j) {
        int side = m.isl[start + j];
        int other = m.isl_cell[start + j];
        if (other <= 0 || d.depth[IX(other)] <= d.p.min_h) {
            ok = 0;
        } else {
            float w = m.length[side];
            v += w * ((m.bed[other] + d.depth[IX(other)]) - eta);
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

// The hanging-edge gradients of flat thread index t (one gradient cell and slot), for the step
// about to run: from the depth and bed the previous step ended with.
template <bool W>
__device__ __forceinline__ void gradient_b(const BDev &d, size_t t) {
    const int sh = W ? 6 : d.shift;
    int g = entity_of<W>(t, sh);
    int s = (int)(t & ((1u << sh) - 1));
    if (g >= d.m.n_grad || d.dt[s] == 0.0f) return;
    int cell = d.m.grad_cell[g];
    int axes = d.m.grad_axes[g];
    float h = d.depth[IX(cell)];
    for (int axis = 0; axis < 2; ++axis) {
        if (!(axes & (1 << axis))) continue;
        float value = 0.0f;
        if (h > d.p.min_h) {
            float dm, xm, dp, xp;
            int okm, okp;
            face_difference_b(d, sh, s, cell, axis == 1 ? 2 : 0, axis, dm, xm, okm);
            face_difference_b(d, sh, s, cell, axis == 1 ? 3 : 1, axis, dp, xp, okp);
            if (okm && okp)
                value = (dp - dm) / (xp - xm);
            else if (okm)
                value = dm / xm;
            else if (okp)
                value = dp / xp;
        }
        d.grad[IX(2 * g + axis)] = value;
    }
}

template <bool W>
__global__ __launch_bounds__(NHF_BLOCK) void kb_gradients(BDev d) {
    gradient_b<W>(d, (size_t)blockIdx.x * NHF_BLOCK + threadIdx.x);
}

// --------------------------------------------------------------------------- theta < 1

template <bool W>
__global__ __launch_bounds__(NHF_BLOCK) void kb_face_avg(BDev d) {
    const int sh = W ? 6 : d.shift;
    size_t t = (size_t)blockIdx.x * NHF_BLOCK + threadIdx.x;
    int cell = entity_of<W>(t, sh);
    int s = (int)(t & ((1u << sh) - 1));
    if (cell >= d.m.e_num || d.dt[s] == 0.0f) return;
    const nhf_mesh &m = d.m;
    uint32_t cnt = m.csr_count[cell];
    int start = m.csr_start[cell];
    for (int dir = 0; dir < 4; ++dir) {
        int n = face_n(cnt, dir);
        float result = 0.0f;
        if (n == 1) {
            result = d.q[IX(m.isl[start])];
        } else if (n > 1) {
            float total = 0.0f, width = 0.0f;
            for (int j = 0; j < n; ++j) {
                int side = m.isl[start + j];
                total += d.q[IX(side)] * m.length[side];
                width += m.length[side];
            }
            result = total / width;
        }
        d.face_qn[IX(4 * cell + dir)] = result;
        start += n;
    }
}

// --------------------------------------------------------------------------- (a) face flux

// A side's loaded values for one slot.
struct BFlux {
    int2 lh;
    uint32_t meta, cls;
    float q, zl, zh, dl, dh, sdc, len, n, nl, nh;
};

// The cells the donor certificate does not cover: the donor pass looks at them every step.
static __device__ __forceinline__ uint8_t bdonor_always(const nhf_mesh &m, int cell) {
    const int4 rec = cell_record(m, cell);
    return (meta_boundary(m.cell_meta[cell]) >= 0 || (rec.x < 0 && rec.w > NHF_LIST_MAX)) ? 2 : 0;
}

__global__ __launch_bounds__(NHF_BLOCK) void kb_init_flags(BDev d) {
    for (int cell = blockIdx.x * NHF_BLOCK + threadIdx.x + 1; cell < d.m.e_num; cell += gridDim.x * NHF_BLOCK)
        d.donor_flag[cell] = bdonor_always(d.m, cell);
}

// The side classes of the per-scenario Manning table: one word per side, built once.
__global__ void kb_side_classes(BDev d) {
    const nhf_mesh &m = d.m;
    for (int si = blockIdx.x * blockDim.x + threadIdx.x; si < m.s_num; si += gridDim.x * blockDim.x) {
        const int l = m.side_cells[2 * si], h = m.side_cells[2 * si + 1];
        const uint32_t cl = l ? (uint32_t)meta_class(m.cell_meta[l]) : NHB_NO_CLASS;
        const uint32_t ch = h ? (uint32_t)meta_class(m.cell_meta[h]) : NHB_NO_CLASS;
        d.side_class[si] = cl | (ch << 16);
    }
}

// Side n of one slot from its cells' class n (map_land_parameters, preprocess/hydraulics.py:63-93,
// with the exterior's n zero): the F32 mean, or n + 0 on an exterior side.
static __device__ __forceinline__ float class_side_n(uint32_t cls, float nl, float nh) {
    const float both = nl + nh;
    const bool one_sided = (cls & 0xffffu) == NHB_NO_CLASS || (cls >> 16) == NHB_NO_CLASS;
    return one_sided ? both : both * 0.5f;
}

// The new discharge of side si for slot s (k_flux's arithmetic, solver_compact.py:439-486).
static __device__ __forceinline__ float bflux_value(const BDev &d, int s, int sh, int si, float dt, const BFlux &f) {
    const nhf_mesh &m = d.m;
    const float g = d.p.g;
    const float bf = (f.meta & NHF_SIDE_EXTERIOR) ? 0.0f : 1.0f;
    const float hl = f.zl + f.dl, hr = f.zh + f.dh;
    const float q = f.q;
    const float wh = fmx(hl, hr) - fmx(fmx(f.zl, f.zh), -999.0f);
    // a face below min_h carries nothing (solver_compact.py:462,484): its arithmetic is skipped
    float nq = 0.0f;
    if (!(wh < d.p.min_h)) {
        float difference = hr - hl;
        if (m.edge_index) {
            int idx = m.edge_index[si];
            if (idx >= 0 && f.dl > d.p.min_h && f.dh > d.p.min_h) {
                int axis = (f.meta & NHF_SIDE_ORIENT_MASK) == 2 ? 1 : 0;
                float o0 = m.hang_offset[2 * idx], o1 = m.hang_offset[2 * idx + 1];
                if (o0 != 0.0f) difference -= d.grad[IX(2 * m.hang_slot[2 * idx] + axis)] * o0;
                if (o1 != 0.0f) difference += d.grad[IX(2 * m.hang_slot[2 * idx + 1] + axis)] * o1;
            }
        }
        float dq = ((-g * difference) / f.sdc) * fmx(wh, 0.0f) * dt;
        float n = f.n;
        float fr = 1.0f;
        if (q != 0.0f) fr = 1.0f + g * dt * (n * n) * fabsf(q / det_powf(fmx(wh, 0.00001f), 7.0f / 3.0f));
        float neighbours = 0.0f;
        if (!d.theta_one) {
            int vertical = (f.meta & NHF_SIDE_ORIENT_MASK) == 2;
            int fa = vertical ? 0 : 2, fb = vertical ? 1 : 3;
            neighbours = neighbour_sum(q, d.face_qn[IX(4 * f.lh.x + fa)], d.face_qn[IX(4 * f.lh.y + fb)],
                                       m.open_face[4 * f.lh.x + fa], m.open_face[4 * f.lh.y + fb]);
        }
        nq = (d.p.theta * q + d.p.theta_half * neighbours + dq) / fr;
    }
    return nq * bf;
}

// K consecutive sides per thread (per wavefront with 64 slots): every side's loads are issued
// before any arithmetic, since a lone side is one short chain of dependent loads. The centre
// distance and Manning n (or the cells' class n) are loaded for every side (shared by the 64
// slots, they cost little), so a wet face waits for no third round (501 against 527 us). The
// hanging-edge index stays in the wet branch: loaded with the records it measured slower (482
// against 455 us on the 64-slot benchmark batch).
// Workgroups take the sides in launch order, as k_flux does (402 against 446 us per launch on
// the 64-slot benchmark batch with the XCD-contiguous order).
template <bool W, int K>
__global__ __launch_bounds__(NHF_BLOCK) void kb_flux(BDev d) {
    const int sh = W ? 6 : d.shift;
    size_t t = (size_t)blockIdx.x * NHF_BLOCK + threadIdx.x;
    const int first = entity_of<W>(t, sh) * K + 1;
    const int s = (int)(t & ((1u << sh) - 1));
    if (first >= d.m.s_num) return;
    const float dt = d.dt[s];
    if (dt == 0.0f) return;
    // the first entity's threads, one per slot, advance the slots' uniform rain state
    if (first == 1) advance_rain_b(d, s, dt);
    const nhf_mesh &m = d.m;
    BFlux f[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int si = first + k;
        if (si < m.s_num) {
            f[k].lh = reinterpret_cast<const int2 *>(m.side_cells)[si];
            f[k].meta = m.side_meta[si];
            f[k].q = d.q[IX(si)];
            if (d.class_manning) f[k].cls = d.side_class[si];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (first + k < m.s_num) {
            f[k].zl = m.bed[f[k].lh.x];
            f[k].zh = m.bed[f[k].lh.y];
            f[k].dl = d.depth[IX(f[k].lh.x)];
            f[k].dh = d.depth[IX(f[k].lh.y)];
            f[k].sdc = m.sdc[first + k];
            f[k].len = m.length[first + k];
            if (d.class_manning) {
                const uint32_t cl = f[k].cls & 0xffffu, ch = f[k].cls >> 16;
                f[k].nl = cl != NHB_NO_CLASS ? d.class_manning[IX(cl)] : 0.0f;
                f[k].nh = ch != NHB_NO_CLASS ? d.class_manning[IX(ch)] : 0.0f;
            } else {
                f[k].n = d.manning ? d.manning[IX(first + k)] : m.manning[first + k];
            }
        }
    }
    if (d.class_manning)
#pragma unroll
        for (int k = 0; k < K; ++k) f[k].n = class_side_n(f[k].cls, f[k].nl, f[k].nh);
    float nq[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (first + k < m.s_num) nq[k] = bflux_value(d, s, sh, first + k, dt, f[k]);
    // in place, as in k_flux; kb_donor scales them where their donor is limited; a discharge
    // that keeps its bits (a dry face stays at zero) is not stored
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (first + k < m.s_num && __float_as_uint(nq[k]) != __float_as_uint(f[k].q)) d.q[IX(first + k)] = nq[k];
    // the donor certificate: the donor of a side that carries more than its share in any slot
    // is flagged for the donor pass (the slots store the same byte: measured faster than one
    // store per wavefront after a ballot, 446 against 471 us)
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (first + k < m.s_num) {
            const bool lower = nq[k] >= 0.0f;
            if (donor_share_exceeded(nq[k], dt, lower ? f[k].dl : f[k].dh, f[k].len))
                d.donor_flag[lower ? f[k].lh.x : f[k].lh.y] = 1;
        }
    }
}

// --------------------------------------------------------------------------- donor scale

// One cell's values for kb_donor, gathered in two rounds before any arithmetic.
struct BDonor {
    uint32_t meta;
    int4 rec;
    float h;
    int pix;
    CellList l;
    float qf[NHF_LIST_MAX];
    float discharge, rainq;
    int bi;
    bool level;
};

// First round: what needs the cell index only.
static __device__ __forceinline__ void bdonor_first(const BDev &d, int s, int sh, int cell, BDonor &v) {
    const nhf_mesh &m = d.m;
    v.meta = m.cell_meta[cell];
    v.rec = cell_record(m, cell);
    v.h = d.depth[IX(cell)];
    v.pix = d.n_pixels > 1 ? m.rain_index[cell] : 0;
}

// Second round: the side discharges of a short list, the boundary values and the rain rate.
static __device__ __forceinline__ void bdonor_second(const BDev &d, int s, int sh, int cell, BDonor &v) {
    const nhf_mesh &m = d.m;
    cell_list(m, v.rec, v.l);
    if (v.l.short_list) gather_list(v.l, [&](int side) { return d.q[IX(side)]; }, v.qf);
    v.bi = meta_boundary(v.meta);
    v.discharge = 0.0f;
    v.level = false;
    if (v.bi >= 0) {
        int kind = m.kind[v.bi];
        if (kind == 1) v.discharge = d.values[s * d.b_num + v.bi] * m.bweight[cell];
        v.level = kind == 0;
    }
    v.rainq = d.raining[s] ? d.rain_rates[(size_t)s * d.n_pixels + v.pix] : 0.0f;
}

// The donor scale of the cell (solver_compact.py:491-514) and the in-place scaling of the sides
// it is the donor of (:516-524), only where the scale is not 1.
static __device__ __forceinline__ void bdonor_finish(const BDev &d, int s, int sh, int cell, float dt, Ctl *c,
                                                     BDonor &v) {
    const nhf_mesh &m = d.m;
    const CellList &l = v.l;
    const float size = l.regular ? cell_size(m, v.meta, cell) : 0.0f;
    const float area = l.regular ? size * size : m.area[cell];
    float positive = (fmx(0.0f, 0.0f) + fmx(v.discharge, 0.0f)) * dt;
    float available = v.h * area + positive + v.rainq * dt * area;
    float outgoing = (fmx(-0.0f, 0.0f) + fmx(-v.discharge, 0.0f)) * dt;
    const int n0 = l.n0, n01 = l.n01, n012 = l.n012, total = l.total, k0 = l.k0;
    const uint32_t cnt = l.cnt;
    if (l.short_list) {
        float len[NHF_LIST_MAX];
        list_lengths(m, l, size, len);
#pragma unroll
        for (int j = 0; j < NHF_LIST_MAX; ++j) {
            if (j < total) {
                float sq = v.qf[j] * len[j];
                if (!(entry_face(j, n0, n01, n012) & 1)) sq = -sq;
                outgoing += fmx(sq, 0.0f) * dt;
            }
        }
    } else {
        int k = k0;
        for (int dir = 0; dir < 4; ++dir) {
            int n = face_n(cnt, dir);
            for (int j = 0; j < n; ++j, ++k) {
                int side = m.isl[k];
                float sq = d.q[IX(side)] * m.length[side];
                if (dir == 0 || dir == 2) sq = -sq;
                outgoing += fmx(sq, 0.0f) * dt;
            }
        }
    }
    float ds = 1.0f;
    bool limited = false;
    if (outgoing > available && !v.level) {
        ds = available / outgoing;
        limited = true;
    }
    // in-place scaling of the sides this cell is the donor of, only where the scale is not 1
    // (the argument of k_donor covers the neighbour that reads a side while it is scaled)
    uint32_t faces = 0;
    if (ds != 1.0f) {
        if (l.short_list) {
#pragma unroll
            for (int j = 0; j < NHF_LIST_MAX; ++j) {
                if (j < total && (v.qf[j] >= 0.0f) == (bool)(entry_face(j, n0, n01, n012) & 1)) {
                    d.q[IX(l.ids[j])] = v.qf[j] * ds;
                    if (v.qf[j] != 0.0f && ds < 1.0f) faces += 1;
                }
            }
        } else {
            int k = k0;
            for (int dir = 0; dir < 4; ++dir) {
                int n = face_n(cnt, dir);
                const bool upper = dir & 1;
                for (int j = 0; j < n; ++j, ++k) {
                    int side = m.isl[k];
                    float q1 = d.q[IX(side)];
                    if ((q1 >= 0.0f) == upper) {
                        d.q[IX(side)] = q1 * ds;
                        if (q1 != 0.0f && ds < 1.0f) faces += 1;
                    }
                }
            }
        }
    }
    if (v.bi >= 0) d.bdonor[IX(d.bslot[cell])] = ds;
    if (limited) {
        atomicAdd(&c->limited_cells, 1u);
        atomicMin(&c->min_scale_bits, __float_as_uint(ds));
    }
    if (faces) atomicAdd(&c->limited_faces, faces);
}

// The donor pass of a batch with fewer than 64 slots (kb_donor<false, 1>: one (cell, slot) per
// thread; 64-slot batches take kb_donor_wide below). With K consecutive cells per thread both
// load rounds of every cell are issued before any arithmetic or store: a later cell may read a
// side before an earlier cell of the same thread scales it, which is one of the orders k_donor's
// argument already covers.
template <bool W, int K>
__global__ __launch_bounds__(NHF_BLOCK) void kb_donor(BDev d) {
    const int sh = W ? 6 : d.shift;
    size_t t = (size_t)xcd_block() * NHF_BLOCK + threadIdx.x;
    const int first = entity_of<W>(t, sh) * K + 1;
    const int s = (int)(t & ((1u << sh) - 1));
    if (first >= d.m.e_num) return;
    const float dt = d.dt[s];
    if (dt == 0.0f) return;
    Ctl *c = &d.ctl[s];
    const int e_num = d.m.e_num;
    // only cells the flux pass flagged, or that the certificate does not cover, are looked at:
    // every other cell's scale is 1 in every slot and changes nothing
    uint8_t flag[K];
#pragma unroll
    for (int k = 0; k < K; ++k) flag[k] = first + k < e_num ? d.donor_flag[first + k] : 0;
    BDonor v[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (flag[k]) bdonor_first(d, s, sh, first + k, v[k]);
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (flag[k]) bdonor_second(d, s, sh, first + k, v[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (flag[k]) {
            bdonor_finish(d, s, sh, first + k, dt, c, v[k]);
            const uint8_t standing = (meta_boundary(v[k].meta) >= 0 || (v[k].rec.x < 0 && v[k].rec.w > NHF_LIST_MAX)) ? 2 : 0;
            if (flag[k] != standing) d.donor_flag[first + k] = standing;
        }
    }
}

// The 64-slot donor pass over a block of 64 consecutive cells per wavefront: lane i loads the
// flag of cell `base + i` (one 64-byte load for the block), a ballot gives the flagged cells as a
// wave-uniform mask, and the wavefront then takes them K at a time, each cell with its 64 slots
// on the 64 lanes, both load rounds of the K cells before their arithmetic (as kb_donor). Few
// cells are flagged, so most wavefronts end after the one load: the pass launches a 64th of the
// wavefronts of one cell per wavefront. Every lane stays live to the end (the flags are read
// back across lanes); lanes of a slot that does not step load and store nothing.
template <int K, int B>
__global__ __launch_bounds__(NHF_BLOCK) void kb_donor_wide(BDev d) {
    const int sh = 6;
    const int s = (int)(threadIdx.x & 63u);
    // one eighth of the grid per XCD (the cell passes' chunked order measured 22.9 against 18.7 us)
    const int xb = (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    const int base = __builtin_amdgcn_readfirstlane(xb * NHF_WAVES + (int)(threadIdx.x >> 6)) * B + 1;
    const int e_num = d.m.e_num;
    if (base >= e_num) return;
    const int own = base + s;
    const int fl = (s < B && own < e_num) ? (int)d.donor_flag[own] : 0;
    uint64_t mask = __ballot(fl != 0);
    const float dt = d.dt[s];
    const bool on = dt != 0.0f;
    Ctl *c = &d.ctl[s];
    while (mask) {
        int cell[K], flag[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cell[k] = 0;
            flag[k] = 0;
            if (mask) {
                const int b = __builtin_ctzll(mask);
                mask &= mask - 1;
                cell[k] = base + b;
                flag[k] = __builtin_amdgcn_readlane(fl, b);
            }
        }
        BDonor v[K];
        if (on) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (cell[k]) bdonor_first(d, s, sh, cell[k], v[k]);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (cell[k]) bdonor_second(d, s, sh, cell[k], v[k]);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (cell[k]) {
                    bdonor_finish(d, s, sh, cell[k], dt, c, v[k]);
                    const uint8_t standing =
                        (meta_boundary(v[k].meta) >= 0 || (v[k].rec.x < 0 && v[k].rec.w > NHF_LIST_MAX)) ? 2 : 0;
                    if (flag[k] != standing) d.donor_flag[cell[k]] = standing;
                }
            }
        }
    }
}

// --------------------------------------------------------------------------- (b) depth update

// Grid-stride over (cell, slot): the stride is a multiple of the slot count, so each thread
// keeps one slot and sums its cells' infiltration and floor terms in FP64 registers; the block
// then adds them to its own compensated accumulators. Boundary cells leave their terms in bvals
// for kb_finish.
// One (cell, slot) of kb_depth: its values gathered in two rounds before any arithmetic.
struct BDepth {
    uint32_t meta;
    int4 rec;
    int pix, j;
    float h, cap, rain_time, max_s, max_d;
    CellList l;
    float qv[NHF_LIST_MAX];
    int kind, bi;
    float ds, bvalue, bweight, level_value, bed, rainq, f0, fc, hk, area_stored;
};

// First round: what needs the cell index only. A cell with rain recomputes its capacity from its
// rain time, so the stored capacity is read only where the cell may have none: `rained` is true
// when the slot's one rain pixel has a positive rate (every cell of the slot has rain).
// In the uniform rain state (`uniform`) the cell's rain time and capacity are the slot's: neither
// is loaded (the capacity comes from the class table in the second round).
static __device__ __forceinline__ void bdepth_first(const BDev &d, int s, int sh, int cell, bool raining,
                                                    bool rained, bool uniform, BDepth &v) {
    const nhf_mesh &m = d.m;
    v.meta = m.cell_meta[cell];
    v.rec = cell_record(m, cell);
    v.pix = d.n_pixels > 1 ? m.rain_index[cell] : 0;
    v.j = d.bslot[cell];
    v.h = d.depth[IX(cell)];
    v.cap = (rained || uniform) ? 0.0f : d.capacity[IX(cell)];
    v.rain_time = (raining && !uniform) ? d.rain_time[IX(cell)] : 0.0f;
    v.max_s = d.max_speed[IX(cell)];
    v.max_d = d.max_depth[IX(cell)];
}

// Second round: side discharges of a short list, boundary values, rain rate, Horton parameters.
static __device__ __forceinline__ void bdepth_second(const BDev &d, int s, int sh, int cell, bool raining,
                                                     bool uniform, BDepth &v) {
    const nhf_mesh &m = d.m;
    if (uniform) v.cap = d.cap_u[((size_t)meta_class(v.meta) << sh) | (size_t)s];
    cell_list(m, v.rec, v.l);
    if (v.l.short_list) gather_list(v.l, [&](int side) { return d.q[IX(side)]; }, v.qv);
    v.bi = meta_boundary(v.meta);
    v.kind = -1;
    v.ds = 1.0f;
    if (v.bi >= 0) {
        v.kind = m.kind[v.bi];
        v.ds = d.bdonor[IX(v.j)];
        v.bvalue = d.values[s * d.b_num + v.bi];
        v.bweight = m.bweight[cell];
        v.level_value = d.level_values[s * d.b_num + v.bi];
        v.bed = m.bed[cell];
    }
    v.rainq = raining ? d.rain_rates[(size_t)s * d.n_pixels + v.pix] : 0.0f;
    if (raining) {
        const float *hp = m.class_horton + 3 * meta_class(v.meta);
        v.f0 = hp[0];
        v.fc = hp[1];
        v.hk = hp[2];
    }
    v.area_stored = v.l.regular ? 0.0f : m.area[cell];
}

// The update of one (cell, slot): sources, Horton, face totals, continuity, infiltration,
// floor, level cells and extrema (k_depth's arithmetic, solver_compact.py:525-613), stores last.
static __device__ __forceinline__ void bdepth_finish(const BDev &d, int s, int sh, int cell, float dt,
                                                     uint64_t step_start, bool uniform, BDepth &v, double &a_infil,
                                                     double &a_floor) {
    const nhf_mesh &m = d.m;
    const int bi = v.bi, kind = v.kind, j = v.j;
    const float ds = v.ds;
    float pipe = fmx(0.0f, 0.0f) + fmn(0.0f, 0.0f) * ds;
    float discharge = 0.0f;
    if (kind == 1) {
        float requested = v.bvalue * v.bweight;
        discharge = fmx(requested, 0.0f) + fmn(requested, 0.0f) * ds;
    }
    float ssq = pipe + discharge;
    const float rainq = v.rainq;
    // the uniform state's capacity is already this step's (advance_rain_b)
    float cap = v.cap, rt = 0.0f;
    if (rainq > 0.0f && !uniform) {
        rt = v.rain_time + dt;
        cap = v.fc + (v.f0 - v.fc) * det_expf(-v.hk * rt / 3600.0f);
    }
    float ql = 0.0f, qr = 0.0f, qb = 0.0f, qt = 0.0f;
    const CellList &l = v.l;
    const float reg_size = l.regular ? cell_size(m, v.meta, cell) : 0.0f;
    const int n0 = l.n0, n01 = l.n01, n012 = l.n012, total = l.total, k0 = l.k0;
    const uint32_t cnt = l.cnt;
    if (l.short_list) {
        float len[NHF_LIST_MAX];
        list_lengths(m, l, reg_size, len);
#pragma unroll
        for (int jj = 0; jj < NHF_LIST_MAX; ++jj) {
            if (jj < total) {
                const float p = v.qv[jj] * len[jj];
                const int dir = entry_face(jj, n0, n01, n012);
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
            for (int jj = 0; jj < n; ++jj, ++k) {
                int side = m.isl[k];
                acc += d.q[IX(side)] * m.length[side];
            }
            fsum[dir] = acc;
        }
        ql = fsum[0];
        qr = fsum[1];
        qb = fsum[2];
        qt = fsum[3];
    }
    float tq = ql - qr + qb - qt + ssq;
    const float area = l.regular ? reg_size * reg_size : v.area_stored;
    float rain_depth = rainq * dt;
    float increment = tq * dt / area + rain_depth;
    float previous = v.h;
    float raw = previous + increment;
    float removed = fmn(cap * dt, fmx(raw, 0.0f));
    raw -= removed;
    float floor_v = fmx(-raw, 0.0f);
    a_infil += (double)(removed * area);
    float h_new;
    float hv_;
    if (kind == 0) {
        float target = fmx(v.level_value - v.bed, 0.0f);
        float *bv = d.bvals;
        bv[IX(6 * j + 0)] = floor_v * area;
        bv[IX(6 * j + 1)] = (target - previous) * area;
        bv[IX(6 * j + 2)] = -(ql - qr + qb - qt) * dt;
        bv[IX(6 * j + 3)] = -pipe * dt;
        bv[IX(6 * j + 4)] = -rain_depth * area;
        bv[IX(6 * j + 5)] = removed * area;
        h_new = target;
    } else {
        a_floor += (double)(floor_v * area);
        if (kind == 1) d.bvals[IX(6 * j + 0)] = discharge * dt;
        h_new = fmx(raw, 0.0f);
    }
    if (rainq > 0.0f && !uniform) {
        d.rain_time[IX(cell)] = rt;
        d.capacity[IX(cell)] = cap;
    }
    d.depth[IX(cell)] = h_new;
    hv_ = fmx(h_new, 0.01f);
    float u = 0.0f, w = 0.0f;
    if (h_new >= d.p.min_h) {
        const float size = l.regular ? reg_size : cell_size(m, v.meta, cell);
        u = (ql + qr) / size / hv_ / 2.0f;
        w = (qb + qt) / size / hv_ / 2.0f;
    }
    float speed = sqrtf(u * u + w * w);
    if (speed > v.max_s) d.max_speed[IX(cell)] = speed;
    // Extrema as the single-mesh step defers them (solver_compact.py:600-606): a cell whose depth
    // rises marks max_depth negative once (its maximum is its depth, reached at the end of its
    // last step); the pair is written out when the rise stops, or by kb_settle_state.
    if (v.max_d < 0.0f) {
        if (!(h_new > v.h)) {
            d.max_depth[IX(cell)] = v.h;
            d.peak_tick[IX(cell)] = step_start;
        }
    } else if (h_new > v.max_d) {
        d.max_depth[IX(cell)] = -1.0f;
    }
}

// Grid-stride over (cell, slot): the stride is a multiple of the slot count, so each thread
// keeps one slot and sums its cells' infiltration and floor terms in FP64 registers, in the
// order of its cells; the block then adds them to its own compensated accumulators. U strides
// at a time: both load rounds of all U cells come before their arithmetic and stores. Boundary
// cells leave their terms in bvals for kb_finish.
// With 64 slots the pass is held to 64 registers (8 wavefronts per SIMD): the compiler's own
// choice of 78 leaves 6, and the pass is bound by how many loads are in flight. It then keeps a
// few loop-invariant 64-bit values in scratch (52 bytes per lane, reloaded from cache), and still
// runs faster: 432 against 467 us on the 1 M-cell batch mesh.
template <bool W, int U>
__global__ __launch_bounds__(NHF_BLOCK) __attribute__((amdgpu_waves_per_eu(W ? 8 : 1))) void kb_depth(BDev d) {
    const int sh = W ? 6 : d.shift;
    const int S = 1 << sh;
    const int s = threadIdx.x & (S - 1);
    const float dt = d.dt[s];
    const nhf_mesh &m = d.m;
    const Ctl *c = &d.ctl[s];
    double a_infil = 0.0, a_floor = 0.0;
    if (dt != 0.0f) {
        const size_t total = (size_t)(m.e_num - 1) << sh;
        const size_t stride = (size_t)gridDim.x * NHF_BLOCK;
        const bool raining = d.raining[s] != 0;
        const bool rained = raining && d.n_pixels == 1 && d.rain_rates[s] > 0.0f;
        const bool uniform = d.rain_uniform[s] != 0;
        const uint64_t step_start = c->now;  // kb_finish commits this step's end
        for (size_t t = (size_t)xcd_block() * NHF_BLOCK + threadIdx.x; t < total; t += U * stride) {
            BDepth v[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (t + u * stride < total)
                    bdepth_first(d, s, sh, entity_of<W>(t + u * stride, sh) + 1, raining, rained, uniform, v[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (t + u * stride < total)
                    bdepth_second(d, s, sh, entity_of<W>(t + u * stride, sh) + 1, raining, uniform, v[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (t + u * stride < total)
                    bdepth_finish(d, s, sh, entity_of<W>(t + u * stride, sh) + 1, dt, step_start, uniform, v[u], a_infil,
                                  a_floor);
        }
    }
    // per-slot block totals: lanes of one slot within a wavefront, then the wavefronts
    for (int o = 32; o >= S; o >>= 1) {
        a_infil += __shfl_xor(a_infil, o, 64);
        a_floor += __shfl_xor(a_floor, o, 64);
    }
    __shared__ double s_part[NHF_WAVES][2][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < S) {
        s_part[wave][0][lane] = a_infil;
        s_part[wave][1][lane] = a_floor;
    }
    __syncthreads();
    if (threadIdx.x < S && dt != 0.0f) {
        double t0 = 0.0, t1 = 0.0;
        for (int w = 0; w < NHF_WAVES; ++w) {
            t0 += s_part[w][0][threadIdx.x];
            t1 += s_part[w][1][threadIdx.x];
        }
        double2 *a = d.acc + ((size_t)blockIdx.x * 2) * S + s;
        double2 v0 = a[0], v1 = a[S];
        acc_add(v0, t0);
        acc_add(v1, t1);
        a[0] = v0;
        a[S] = v1;
    }
}

// --------------------------------------------------------------------------- (c) CFL dt

// U strides per round: the side records of all U sides first, then their bed, depth, length,
// discharge and id loads, then the bounds (the minimum of the total order does not depend on the
// order the sides are taken in).
// The first `grad_blocks` workgroups (a multiple of 8, so each CFL workgroup keeps its XCD)
// compute the next step's hanging-edge gradients instead: they need only the depth and bed this
// step ends with, final once the depth pass is done. They are computed for the slots that stepped
// in this step; a slot that stops stepping does not step again before the next batch, whose first
// step has a gradient launch of its own.
template <bool W, int U>
__global__ __launch_bounds__(NHF_BLOCK) void kb_cfl(BDev d, int mode, int grad_blocks) {
    if ((int)blockIdx.x < grad_blocks) {
        gradient_b<W>(d, (size_t)blockIdx.x * NHF_BLOCK + threadIdx.x);
        return;
    }
    const int block = (int)blockIdx.x - grad_blocks, blocks = (int)gridDim.x - grad_blocks;
    const int xb = (block & 7) * (blocks >> 3) + (block >> 3);  // the XCD-contiguous order
    const int sh = W ? 6 : d.shift;
    const int S = 1 << sh;
    const int s = threadIdx.x & (S - 1);
    const bool on = mode == 0 || d.dt[s] != 0.0f;
    const nhf_mesh &m = d.m;
    uint64_t key = ~0ull;
    int32_t idx = 0;
    if (on) {
        const float afa = d.p.afa, g = d.p.g;
        const size_t total = (size_t)(m.s_num - 1) << sh;
        const size_t stride = (size_t)blocks * NHF_BLOCK;
        // the cell pairs of a round are loaded during the round before (NHB_CFL_PREFETCH), so
        // each round waits for one load latency instead of two
        int si[U];
        int2 lh[U];
        auto records = [&](size_t t0, int *ids, int2 *cells) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool live = t0 + u * stride < total;
                ids[u] = live ? entity_of<W>(t0 + u * stride, sh) + 1 : 0;
                cells[u] = live ? reinterpret_cast<const int2 *>(m.side_cells)[ids[u]] : make_int2(0, 0);
            }
        };
        size_t t = (size_t)xb * NHF_BLOCK + threadIdx.x;
        if (NHB_CFL_PREFETCH) records(t, si, lh);
        for (; t < total; t += U * stride) {
            if (!NHB_CFL_PREFETCH) records(t, si, lh);
            float zl[U], zh[U], dl[U], dh[U], len[U], qs[U];
            uint32_t ref[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (lh[u].x != 0 && lh[u].y != 0) {
                    zl[u] = m.bed[lh[u].x];
                    zh[u] = m.bed[lh[u].y];
                    dl[u] = d.depth[IX(lh[u].x)];
                    dh[u] = d.depth[IX(lh[u].y)];
                    len[u] = m.length[si[u]];
                    qs[u] = d.q[IX(si[u])];
                    ref[u] = m.side_key[si[u]];
                }
            }
            int si_next[U];
            int2 lh_next[U];
            if (NHB_CFL_PREFETCH) records(t + U * stride, si_next, lh_next);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (lh[u].x != 0 && lh[u].y != 0) {
                    float head = fmx(zl[u] + dl[u], zh[u] + dh[u]);
                    float h = fmx(head - fmx(fmx(zl[u], zh[u]), -999.0f), 0.01f);
                    float cfl = afa * len[u] / (sqrtf(g * h) + fabsf(qs[u]) / h);
                    uint64_t k2 = ((uint64_t)__float_as_uint(cfl) << 32) | ref[u];
                    if (k2 < key) {
                        key = k2;
                        idx = si[u];
                    }
                }
            }
            if (NHB_CFL_PREFETCH)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    si[u] = si_next[u];
                    lh[u] = lh_next[u];
                }
        }
    }
    for (int o = 32; o >= S; o >>= 1) {
        uint64_t k2 = __shfl_xor((unsigned long long)key, o, 64);
        int32_t i2 = __shfl_xor(idx, o, 64);
        if (k2 < key) {
            key = k2;
            idx = i2;
        }
    }
    __shared__ uint64_t s_key[NHF_WAVES][64];
    __shared__ int32_t s_idx[NHF_WAVES][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < S) {
        s_key[wave][lane] = key;
        s_idx[wave][lane] = idx;
    }
    __syncthreads();
    if (threadIdx.x < S) {
        for (int w = 1; w < NHF_WAVES; ++w)
            if (s_key[w][threadIdx.x] < key) {
                key = s_key[w][threadIdx.x];
                idx = s_idx[w][threadIdx.x];
            }
        d.cfl_key[(size_t)block * S + s] = key;
        d.cfl_idx[(size_t)block * S + s] = idx;
    }
}

// One workgroup per slot: reduce its CFL partials to ndt and side; in a batch (mode 1) also add
// the boundary cells' budget terms, commit `step` and prepare `step + 1`.
__global__ __launch_bounds__(NHF_BLOCK) void kb_finish(BDev d, int mode, int step, int steps) {
    const int sh = d.shift;
    const int S = 1 << sh;
    const int s = blockIdx.x;
    Ctl *c = &d.ctl[s];
    if (mode == 1 && d.dt[s] == 0.0f) return;
    __shared__ uint64_t s_key[NHF_WAVES];
    __shared__ int32_t s_idx[NHF_WAVES];
    // Keys only, four partials per round (every key is unique: it holds the side's reference id),
    // remembering which partial holds the minimum; its side index is read once at the end.
    uint64_t key = ~0ull;
    int32_t idx = -1;  // the partial holding `key`
    for (int b0 = threadIdx.x; b0 < d.cfl_blocks; b0 += 4 * NHF_BLOCK) {
        uint64_t k[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int b = b0 + u * NHF_BLOCK;
            k[u] = b < d.cfl_blocks ? d.cfl_key[(size_t)b * S + s] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k[u] < key) {
                key = k[u];
                idx = b0 + u * NHF_BLOCK;
            }
        }
    }
    wave_min_key(key, idx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_key[wave] = key;
        s_idx[wave] = idx;
    }
    if (mode == 1 && d.n_bcells > 0) {
        // boundary-cell terms: one job per term, one wavefront per job, lanes over the cells
        const int b_num = d.b_num;
        const int jobs = 1 + 6 * b_num;
        for (int job = wave; job < jobs; job += NHF_WAVES) {
            int term, comp, lo_b, hi_b, want;
            if (job == 0) {  // floor at level cells, every level boundary
                term = 2; comp = 0; lo_b = 0; hi_b = b_num; want = 0;
            } else if (job <= 5 * b_num) {
                int tb = (job - 1) / 5;
                term = 4 + (job - 1); comp = 1 + (job - 1) % 5; lo_b = tb; hi_b = tb + 1; want = 0;
            } else {
                int tb = job - 1 - 5 * b_num;
                term = 4 + 5 * b_num + tb; comp = 0; lo_b = tb; hi_b = tb + 1; want = 1;
            }
            double sum = 0.0;
            for (int bb = lo_b; bb < hi_b; ++bb) {
                if (d.m.kind[bb] != want) continue;
                for (int j = d.bcell_start[bb] + lane; j < d.bcell_start[bb + 1]; j += 64)
                    sum += (double)d.bvals[IX(6 * j + comp)];
            }
            sum = wave_sum(sum);
            if (lane == 0 && sum != 0.0) {
                double2 a = d.bacc[(size_t)s * d.terms + term];
                acc_add(a, sum);
                d.bacc[(size_t)s * d.terms + term] = a;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NHF_WAVES; ++w)
            if (s_key[w] < key) {
                key = s_key[w];
                idx = s_idx[w];
            }
        if (key == ~0ull) {
            c->ndt = INFINITY;
            c->cfl_arg = 0;
        } else {
            c->ndt = __uint_as_float((uint32_t)(key >> 32));
            c->cfl_arg = d.cfl_idx[(size_t)idx * S + s];
        }
        if (mode == 1) {
            commit_slot(c, step);
            d.dt[s] = 0.0f;
            if (step + 1 < steps) prepare_slot_b(d, s, step + 1);
        }
    }
}

// Interval budget of slot s: the depth blocks' (sum, compensation) pairs of terms 0 and 1 are
// combined in double-double in a fixed order; the boundary terms are copied; all are cleared.
__global__ __launch_bounds__(NHF_BLOCK) void kb_flush(BDev d, int s) {
    const int S = 1 << d.shift;
    __shared__ double s_hi[NHF_BLOCK], s_lo[NHF_BLOCK];
    for (int term = 0; term < 2; ++term) {
        double hi = 0.0, lo = 0.0;
        for (int b = threadIdx.x; b < d.depth_max; b += NHF_BLOCK) {
            double2 *p = d.acc + ((size_t)b * 2 + term) * S + s;
            const double2 a = *p;
            *p = make_double2(0.0, 0.0);
            const double parts[2] = {a.x, a.y};
            for (double x : parts) {
                double sum = hi + x;
                double bp = sum - hi;
                lo += (hi - (sum - bp)) + (x - bp);
                hi = sum;
            }
        }
        s_hi[threadIdx.x] = hi;
        s_lo[threadIdx.x] = lo;
        __syncthreads();
        for (int o = NHF_BLOCK / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                double a = s_hi[threadIdx.x], b = s_hi[threadIdx.x + o];
                double sum = a + b;
                double bp = sum - a;
                double err = (a - (sum - bp)) + (b - bp);
                s_hi[threadIdx.x] = sum;
                s_lo[threadIdx.x] += s_lo[threadIdx.x + o] + err;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) d.flush[term] = make_double2(s_hi[0], s_lo[0]);
        __syncthreads();
    }
    for (int t = 2 + threadIdx.x; t < d.terms; t += NHF_BLOCK) {
        d.flush[t] = d.bacc[(size_t)s * d.terms + t];
        d.bacc[(size_t)s * d.terms + t] = make_double2(0.0, 0.0);
    }
}

// --------------------------------------------------------------------------- host side

enum { KB_PREPARE = 0, KB_GRAD, KB_FACE, KB_FLUX, KB_DONOR, KB_DEPTH, KB_CFL, KB_FINISH, KB_COUNT };

struct nhb_ctx {
    BDev d;
    int slots;
    hipStream_t stream;
    int graph_enabled, timing;
    hipGraphExec_t exec[NHF_MAX_BATCH + 1];
    std::vector<hipEvent_t> events;
    std::vector<int> event_kind;
    double kernel_ms[KB_COUNT];
    int64_t kernel_count[KB_COUNT];
    // staging: host (pinned) and device copies of the same layout
    char *stage_host, *stage_dev;
    size_t stage_bytes;
    Stage st_host, st_dev;
    Ctl *host_ctl;
    double2 *host_flush;
    hipEvent_t done, stage_done;  // batch trace copied; staged words copied
    int pending_steps, stage_inflight;
};

static size_t align_b(size_t v) { return (v + 255) & ~(size_t)255; }

static int b_blocks(size_t n) { return (int)((n + NHF_BLOCK - 1) / NHF_BLOCK); }

// Sides per wavefront of the 64-slot flux pass (measured on the 1 M-cell batch mesh in launch
// order: 434 / 401 / 472 us for 2 / 4 / 8 sides) and flagged cells per load round of its donor
// pass.
#ifndef NHB_FLUX_SIDES
#define NHB_FLUX_SIDES 4
#endif
#ifndef NHB_DONOR_CELLS
#define NHB_DONOR_CELLS 2
#endif
// Cells whose flags one wavefront of the 64-slot donor pass scans (at most 64).
#ifndef NHB_DONOR_BLOCK
#define NHB_DONOR_BLOCK 16
#endif
// Grid strides per load round of the 64-slot depth pass (1 / 2 / 3: 559 / 745 / 966 us; two or
// more lower the occupancy) and of its CFL pass.
#ifndef NHB_DEPTH_ROUND
#define NHB_DEPTH_ROUND 1
#endif
#ifndef NHB_CFL_ROUND
#define NHB_CFL_ROUND 2
#endif

static int b_blocks_xcd(size_t n) { return (b_blocks(n) + 7) & ~7; }

// Grids of the grid-stride kernels: by default about 2048 workgroups (8 per CU), fewer on small
// meshes; nhb_set_tuning picks others up to the allocated maximum (cells or sides per thread).
#define NHB_MAX_STRIDE_BLOCKS 8192

static int b_stride_max(size_t n) {
    int g = b_blocks_xcd(n);
    return g < NHB_MAX_STRIDE_BLOCKS ? g : NHB_MAX_STRIDE_BLOCKS;
}

static int b_stride_grid(size_t n) {
    int g = b_stride_max(n);
    return g < 2048 ? g : 2048;
}

static int slot_shift(int slots) {
    int sh = 0;
    while ((1 << sh) < slots) ++sh;
    return sh;
}

struct BCarve {
    size_t ctl, dt, raining, coeff, values, level, rain, bdonor, face_qn, grad, key, idx, acc, bvals, bacc,
        flush, side_class, rain_uniform, rt_u, cap_u, class_used, cap_stage, donor_flag, total;
};

static BCarve bcarve(const nhf_mesh *m, const nhb_state *st, const nhf_params *p, int slots) {
    BCarve c;
    size_t off = 0;
    const size_t S = slots;
    const int b = m->b_num > 0 ? m->b_num : 1;
    const int terms = 4 + 6 * m->b_num;
    const int nb = st->n_bcells > 0 ? st->n_bcells : 1;
    const int cfl_blocks = b_stride_max(((size_t)(m->s_num > 1 ? m->s_num - 1 : 1)) * S);
    const int depth_blocks = b_stride_max(((size_t)(m->e_num - 1)) * S);
    c.ctl = off; off = align_b(off + sizeof(Ctl) * S);
    c.dt = off; off = align_b(off + sizeof(float) * S);
    c.raining = off; off = align_b(off + sizeof(int32_t) * S);
    c.coeff = off; off = align_b(off + sizeof(float) * 2 * b * S);
    c.values = off; off = align_b(off + sizeof(float) * b * S);
    c.level = off; off = align_b(off + sizeof(float) * b * S);
    c.rain = off; off = align_b(off + sizeof(float) * (m->n_pixels > 0 ? m->n_pixels : 1) * S);
    c.bdonor = off; off = align_b(off + sizeof(float) * nb * S);
    c.face_qn = off; off = align_b(off + (p->theta == 1.0f ? 0 : sizeof(float) * 4 * (size_t)m->e_num * S));
    c.grad = off; off = align_b(off + sizeof(float) * 2 * (size_t)(m->n_grad > 0 ? m->n_grad : 1) * S);
    c.key = off; off = align_b(off + sizeof(uint64_t) * cfl_blocks * S);
    c.idx = off; off = align_b(off + sizeof(int32_t) * cfl_blocks * S);
    c.acc = off; off = align_b(off + sizeof(double2) * 2 * (size_t)depth_blocks * S);
    c.bvals = off; off = align_b(off + sizeof(float) * 6 * (size_t)nb * S);
    c.bacc = off; off = align_b(off + sizeof(double2) * terms * S);
    c.flush = off; off = align_b(off + sizeof(double2) * terms);
    c.side_class = off; off = align_b(off + (st->class_manning ? sizeof(uint32_t) * (size_t)m->s_num : 0));
    c.rain_uniform = off; off = align_b(off + sizeof(int32_t) * S);
    c.rt_u = off; off = align_b(off + sizeof(float) * S);
    c.cap_u = off; off = align_b(off + sizeof(float) * 256 * S);
    c.class_used = off; off = align_b(off + sizeof(int32_t) * 256);
    c.cap_stage = off; off = align_b(off + sizeof(float) * 256);
    c.donor_flag = off; off = align_b(off + (size_t)m->e_num);
    c.total = off;
    return c;
}

static size_t stage_layout(const BDev &d, int slots, Stage &st, char *base) {
    size_t off = 0;
    st.target = (uint64_t *)(base + off); off = align_b(off + sizeof(uint64_t) * slots);
    st.origin = (uint64_t *)(base + off); off = align_b(off + sizeof(uint64_t) * slots);
    st.flags = (int32_t *)(base + off); off = align_b(off + sizeof(int32_t) * slots);
    st.coeff = (float *)(base + off); off = align_b(off + sizeof(float) * 2 * (d.b_num > 0 ? d.b_num : 1) * slots);
    st.rain = (float *)(base + off); off = align_b(off + sizeof(float) * (d.n_pixels > 0 ? d.n_pixels : 1) * slots);
    return off;
}

static hipEvent_t b_timing_event(nhb_ctx *ctx, size_t i) {
    while (ctx->events.size() <= i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->events.push_back(e);
    }
    return ctx->events[i];
}

static int32_t b_launch_batch(nhb_ctx *ctx, int steps, bool timed) {
    const BDev &d = ctx->d;
    hipStream_t s = ctx->stream;
    const size_t S = ctx->slots;
    size_t ev = 0;
    ctx->event_kind.clear();
#define BLAUNCH(kind, ...)                                                  \
    do {                                                                    \
        if (timed) {                                                        \
            hipEvent_t e0 = b_timing_event(ctx, ev++);                      \
            if (!e0) return set_error_text("hipEventCreate failed");        \
            HIP_TRY(hipEventRecord(e0, s));                                 \
        }                                                                   \
        hipLaunchKernelGGL(__VA_ARGS__);                                    \
        if (timed) {                                                        \
            hipEvent_t e1 = b_timing_event(ctx, ev++);                      \
            if (!e1) return set_error_text("hipEventCreate failed");        \
            HIP_TRY(hipEventRecord(e1, s));                                 \
            ctx->event_kind.push_back(kind);                                \
        }                                                                   \
    } while (0)
    const size_t cells = (size_t)(d.m.e_num - 1), sides = (size_t)(d.m.s_num - 1);
    BLAUNCH(KB_PREPARE, kb_prepare, dim3((unsigned)S), dim3(64), 0, s, d, steps);
    const bool w = S == 64;
    const int grad_blocks = d.m.n_grad > 0 ? b_blocks_xcd((size_t)d.m.n_grad * S) : 0;
    for (int step = 0; step < steps; ++step) {
        // gradients: a launch of their own at the batch's first step (the host may have written
        // the state); the previous step's CFL launch computes them for every later step
        if (d.m.n_grad > 0 && step == 0)
            BLAUNCH(KB_GRAD, w ? kb_gradients<true> : kb_gradients<false>, dim3(b_blocks((size_t)d.m.n_grad * S)),
                    dim3(NHF_BLOCK), 0, s, d);
        if (!d.theta_one)
            BLAUNCH(KB_FACE, w ? kb_face_avg<true> : kb_face_avg<false>, dim3(b_blocks((size_t)d.m.e_num * S)),
                    dim3(NHF_BLOCK), 0, s, d);
        if (sides > 0)
            BLAUNCH(KB_FLUX, (w ? kb_flux<true, NHB_FLUX_SIDES> : kb_flux<false, 1>),
                    dim3(b_blocks_xcd((w ? (sides + NHB_FLUX_SIDES - 1) / NHB_FLUX_SIDES : sides) * S)), dim3(NHF_BLOCK), 0, s,
                    d);
        if (w)
            BLAUNCH(KB_DONOR, (kb_donor_wide<NHB_DONOR_CELLS, NHB_DONOR_BLOCK>),
                    dim3(b_blocks_xcd((cells + NHB_DONOR_BLOCK - 1) / NHB_DONOR_BLOCK * 64)), dim3(NHF_BLOCK), 0, s, d);
        else
            BLAUNCH(KB_DONOR, (kb_donor<false, 1>), dim3(b_blocks_xcd(cells * S)), dim3(NHF_BLOCK), 0, s, d);
        BLAUNCH(KB_DEPTH, (w ? kb_depth<true, NHB_DEPTH_ROUND> : kb_depth<false, 1>), dim3(d.depth_blocks), dim3(NHF_BLOCK), 0, s, d);
        const int gb = step + 1 < steps ? grad_blocks : 0;
        BLAUNCH(KB_CFL, (w ? kb_cfl<true, NHB_CFL_ROUND> : kb_cfl<false, 1>), dim3(gb + d.cfl_blocks), dim3(NHF_BLOCK), 0,
                s, d, 1, gb);
        BLAUNCH(KB_FINISH, kb_finish, dim3((unsigned)S), dim3(NHF_BLOCK), 0, s, d, 1, step, steps);
    }
#undef BLAUNCH
    HIP_TRY(hipGetLastError());
    return 0;
}

static int32_t settle_stage(nhb_ctx *ctx);

extern "C" {

size_t nhb_workspace_bytes(const nhf_mesh *mesh, const nhb_state *state, const nhf_params *params,
                           int32_t slots) {
    if (!mesh || !state || !params || slots < 1 || slots > NHB_MAX_SLOTS) return 0;
    return bcarve(mesh, state, params, 1 << slot_shift(slots)).total;
}

int32_t nhb_create(nhb_ctx **out, const nhf_mesh *mesh, const nhb_state *state, const nhf_params *params,
                   int32_t slots, void *workspace, size_t workspace_bytes, void *stream) {
    if (!out || !mesh || !state || !params || !workspace) return set_error_text("nhb_create: null argument");
    if (slots < 1 || slots > NHB_MAX_SLOTS || (slots & (slots - 1)))
        return set_error_text("nhb_create: slots must be a power of two in 1..64");
    if (params->batch_capacity < 1 || params->batch_capacity > NHF_MAX_BATCH)
        return set_error_text("nhb_create: batch_capacity out of range");
    if (mesh->e_num < 2 || mesh->s_num < 1) return set_error_text("nhb_create: empty mesh");
    if (mesh->pipe_q) return set_error_text("nhb_create: pipe exchange is not supported in a batch");
    if (!state->depth || !state->rain_time || !state->capacity || !state->max_depth || !state->max_speed ||
        !state->peak_tick || !state->q)
        return set_error_text("nhb_create: state array missing");
    if (mesh->b_num > 0 && (!state->bslot || !state->bcell_start))
        return set_error_text("nhb_create: boundary-cell list missing");
    if ((params->theta != 1.0f) && !mesh->open_face)
        return set_error_text("nhb_create: open_face is required when sita != 1");
    if (mesh->n_hang > 0 && (!mesh->edge_index || !mesh->x || !mesh->y || !mesh->grad_cell))
        return set_error_text("nhb_create: hanging reconstruction arrays missing");
    if (const char *name = missing_mesh_array(mesh, false, !state->manning && !state->class_manning)) {
        snprintf(nhf_error_text, sizeof nhf_error_text, "nhb_create: mesh array %s is NULL", name);
        return -2;
    }
    BCarve cv = bcarve(mesh, state, params, slots);
    if (workspace_bytes < cv.total) return set_error_text("nhb_create: workspace too small");
    nhb_ctx *ctx = new nhb_ctx();
    char *ws = (char *)workspace;
    BDev &d = ctx->d;
    d.m = *mesh;
    if (mesh->n_hang == 0) d.m.edge_index = nullptr;
    d.p = *params;
    d.shift = slot_shift(slots);
    d.terms = 4 + 6 * mesh->b_num;
    d.cfl_blocks = b_stride_grid(((size_t)(mesh->s_num > 1 ? mesh->s_num - 1 : 1)) * slots);
    d.depth_blocks = b_stride_grid(((size_t)(mesh->e_num - 1)) * slots);
    d.cfl_max = b_stride_max(((size_t)(mesh->s_num > 1 ? mesh->s_num - 1 : 1)) * slots);
    d.depth_max = b_stride_max(((size_t)(mesh->e_num - 1)) * slots);
    d.n_bcells = state->n_bcells;
    d.theta_one = params->theta == 1.0f;
    d.b_num = mesh->b_num;
    d.n_pixels = mesh->n_pixels > 0 ? mesh->n_pixels : 1;
    d.depth = state->depth;
    d.rain_time = state->rain_time;
    d.capacity = state->capacity;
    d.max_depth = state->max_depth;
    d.max_speed = state->max_speed;
    d.peak_tick = state->peak_tick;
    d.q = state->q;
    d.manning = state->manning;
    d.class_manning = state->class_manning;
    d.bslot = state->bslot;
    d.bcell_start = state->bcell_start;
    d.ctl = (Ctl *)(ws + cv.ctl);
    d.dt = (float *)(ws + cv.dt);
    d.raining = (int32_t *)(ws + cv.raining);
    d.coefficients = (float *)(ws + cv.coeff);
    d.values = (float *)(ws + cv.values);
    d.level_values = (float *)(ws + cv.level);
    d.rain_rates = (float *)(ws + cv.rain);
    d.bdonor = (float *)(ws + cv.bdonor);
    d.face_qn = d.theta_one ? nullptr : (float *)(ws + cv.face_qn);
    d.grad = (float *)(ws + cv.grad);
    d.cfl_key = (uint64_t *)(ws + cv.key);
    d.cfl_idx = (int32_t *)(ws + cv.idx);
    d.acc = (double2 *)(ws + cv.acc);
    d.bvals = (float *)(ws + cv.bvals);
    d.bacc = (double2 *)(ws + cv.bacc);
    d.flush = (double2 *)(ws + cv.flush);
    d.side_class = d.class_manning ? (uint32_t *)(ws + cv.side_class) : nullptr;
    d.rain_uniform = (int32_t *)(ws + cv.rain_uniform);
    d.rt_u = (float *)(ws + cv.rt_u);
    d.cap_u = (float *)(ws + cv.cap_u);
    d.class_used = (int32_t *)(ws + cv.class_used);
    d.cap_stage = (float *)(ws + cv.cap_stage);
    d.donor_flag = (uint8_t *)(ws + cv.donor_flag);
    ctx->slots = slots;
    ctx->stream = (hipStream_t)stream;
    ctx->graph_enabled = 0;
    ctx->timing = 0;
    for (int i = 0; i <= NHF_MAX_BATCH; ++i) ctx->exec[i] = nullptr;
    memset(ctx->kernel_ms, 0, sizeof ctx->kernel_ms);
    memset(ctx->kernel_count, 0, sizeof ctx->kernel_count);
    Stage probe;
    ctx->stage_bytes = stage_layout(d, slots, probe, nullptr);
    ctx->stage_host = ctx->stage_dev = nullptr;
    ctx->host_ctl = nullptr;
    ctx->host_flush = nullptr;
    hipError_t err;
    if ((err = hipHostMalloc((void **)&ctx->stage_host, ctx->stage_bytes, hipHostMallocDefault)) != hipSuccess ||
        (err = hipMalloc((void **)&ctx->stage_dev, ctx->stage_bytes)) != hipSuccess ||
        (err = hipHostMalloc((void **)&ctx->host_ctl, sizeof(Ctl) * slots, hipHostMallocDefault)) != hipSuccess ||
        (err = hipHostMalloc((void **)&ctx->host_flush, sizeof(double2) * d.terms, hipHostMallocDefault)) !=
            hipSuccess) {
        if (ctx->stage_host) (void)hipHostFree(ctx->stage_host);
        if (ctx->stage_dev) (void)hipFree(ctx->stage_dev);
        if (ctx->host_ctl) (void)hipHostFree(ctx->host_ctl);
        delete ctx;
        return set_error("nhb_create: staging allocation", err);
    }
    memset(ctx->stage_host, 0, ctx->stage_bytes);
    ctx->pending_steps = ctx->stage_inflight = 0;
    if ((err = hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming)) != hipSuccess ||
        (err = hipEventCreateWithFlags(&ctx->stage_done, hipEventDisableTiming)) != hipSuccess) {
        (void)hipHostFree(ctx->stage_host);
        (void)hipFree(ctx->stage_dev);
        (void)hipHostFree(ctx->host_ctl);
        (void)hipHostFree(ctx->host_flush);
        delete ctx;
        return set_error("nhb_create: events", err);
    }
    stage_layout(d, slots, ctx->st_host, ctx->stage_host);
    stage_layout(d, slots, ctx->st_dev, ctx->stage_dev);
    // every slot starts idle with a clean clock, in the per-cell rain state
    for (int s = 0; s < slots; ++s) {
        hipLaunchKernelGGL(kb_reset, dim3(64), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, s, (uint64_t)0);
        hipLaunchKernelGGL(kb_set_rain_state, dim3(1), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, s, 0, 0.0f,
                           (const float *)nullptr);
    }
    (void)hipMemsetAsync(d.class_used, 0, sizeof(int32_t) * 256, ctx->stream);
    hipLaunchKernelGGL(kb_init_flags, dim3(b_blocks((size_t)mesh->e_num) < 1024 ? b_blocks((size_t)mesh->e_num) : 1024),
                       dim3(NHF_BLOCK), 0, ctx->stream, ctx->d);
    hipLaunchKernelGGL(kb_class_used, dim3(b_blocks((size_t)mesh->e_num) < 1024 ? b_blocks((size_t)mesh->e_num) : 1024),
                       dim3(NHF_BLOCK), 0, ctx->stream, ctx->d);
    if (d.class_manning)
        hipLaunchKernelGGL(kb_side_classes, dim3(b_blocks((size_t)mesh->s_num)), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d);
    if ((err = hipGetLastError()) != hipSuccess) {
        (void)hipHostFree(ctx->stage_host);
        (void)hipFree(ctx->stage_dev);
        (void)hipHostFree(ctx->host_ctl);
        (void)hipHostFree(ctx->host_flush);
        delete ctx;
        return set_error("nhb_create: reset", err);
    }
    *out = ctx;
    return 0;
}

int32_t nhb_destroy(nhb_ctx *ctx) {
    if (!ctx) return 0;
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i <= NHF_MAX_BATCH; ++i)
        if (ctx->exec[i]) (void)hipGraphExecDestroy(ctx->exec[i]);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->done);
    (void)hipEventDestroy(ctx->stage_done);
    (void)hipHostFree(ctx->stage_host);
    (void)hipFree(ctx->stage_dev);
    (void)hipHostFree(ctx->host_ctl);
    (void)hipHostFree(ctx->host_flush);
    delete ctx;
    return 0;
}

static int32_t check_slot(const nhb_ctx *ctx, int32_t slot) {
    if (!ctx) return set_error_text("nhb: null context");
    if (slot < 0 || slot >= ctx->slots) return set_error_text("nhb: slot out of range");
    return 0;
}

int32_t nhb_reset(nhb_ctx *ctx, int32_t slot, uint64_t now_ticks) {
    if (check_slot(ctx, slot)) return -2;
    if (int32_t rc = settle_stage(ctx)) return rc;
    ctx->st_host.flags[slot] = 0;
    hipLaunchKernelGGL(kb_reset, dim3(64), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, (int)slot, now_ticks);
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhb_set_forcing(nhb_ctx *ctx, int32_t slot, const float *coefficients, const float *rain_rates,
                        int32_t raining) {
    if (check_slot(ctx, slot)) return -2;
    if (int32_t rc = settle_stage(ctx)) return rc;
    const BDev &d = ctx->d;
    Stage &st = ctx->st_host;
    int32_t flags = st.flags[slot] & ~(STAGE_RAINING | STAGE_RAINING_ON);
    if (d.b_num > 0 && coefficients) {
        memcpy(st.coeff + (size_t)slot * 2 * d.b_num, coefficients, sizeof(float) * 2 * d.b_num);
        flags |= STAGE_COEFF;
    }
    if (rain_rates) {
        memcpy(st.rain + (size_t)slot * d.n_pixels, rain_rates, sizeof(float) * d.n_pixels);
        flags |= STAGE_RAIN;
    }
    flags |= STAGE_RAINING | (raining ? STAGE_RAINING_ON : 0);
    st.flags[slot] = flags;
    return 0;
}

int32_t nhb_begin(nhb_ctx *ctx, int32_t slot, uint64_t target_ticks, uint64_t origin_ticks) {
    if (check_slot(ctx, slot)) return -2;
    if (int32_t rc = settle_stage(ctx)) return rc;
    ctx->st_host.target[slot] = target_ticks;
    ctx->st_host.origin[slot] = origin_ticks;
    ctx->st_host.flags[slot] |= STAGE_BEGIN;
    return 0;
}

// The staged words go to the device with an asynchronous copy from pinned memory; the host
// copy is cleared (and may be written again) only once that copy has completed.
static int32_t settle_stage(nhb_ctx *ctx) {
    if (!ctx->stage_inflight) return 0;
    HIP_TRY(hipEventSynchronize(ctx->stage_done));
    for (int s = 0; s < ctx->slots; ++s) ctx->st_host.flags[s] = 0;
    ctx->stage_inflight = 0;
    return 0;
}

// Upload and apply the staged words of every slot.
static int32_t apply_stage(nhb_ctx *ctx) {
    int32_t rc = settle_stage(ctx);
    if (rc) return rc;
    bool any = false;
    for (int s = 0; s < ctx->slots; ++s) any |= ctx->st_host.flags[s] != 0;
    if (!any) return 0;
    HIP_TRY(hipMemcpyAsync(ctx->stage_dev, ctx->stage_host, ctx->stage_bytes, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(kb_apply, dim3(ctx->slots), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, ctx->st_dev);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->stage_done, ctx->stream));
    ctx->stage_inflight = 1;
    return 0;
}

int32_t nhb_launch_batch(nhb_ctx *ctx, int32_t steps) {
    if (!ctx) return set_error_text("nhb_launch_batch: null context");
    if (steps < 1 || steps > ctx->d.p.batch_capacity) return set_error_text("nhb_launch_batch: steps out of range");
    if (ctx->pending_steps) return set_error_text("nhb_launch_batch: the previous batch was not waited for");
    int32_t rc = apply_stage(ctx);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    if (ctx->graph_enabled && !ctx->timing && s != nullptr) {
        if (!ctx->exec[steps]) {
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            rc = b_launch_batch(ctx, steps, false);
            hipError_t end = hipStreamEndCapture(s, &graph);
            if (rc) return rc;
            if (end != hipSuccess) return set_error("hipStreamEndCapture", end);
            hipError_t inst = hipGraphInstantiate(&ctx->exec[steps], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (inst != hipSuccess) return set_error("hipGraphInstantiate", inst);
        }
        HIP_TRY(hipGraphLaunch(ctx->exec[steps], s));
    } else {
        rc = b_launch_batch(ctx, steps, ctx->timing != 0);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(ctx->host_ctl, ctx->d.ctl, sizeof(Ctl) * ctx->slots, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(ctx->done, s));
    ctx->pending_steps = steps;
    return 0;
}

int32_t nhb_wait_batch(nhb_ctx *ctx, uint32_t *trace) {
    if (!ctx) return set_error_text("nhb_wait_batch: null context");
    const int steps = ctx->pending_steps;
    if (!steps) return set_error_text("nhb_wait_batch: no batch in flight");
    ctx->pending_steps = 0;
    HIP_TRY(hipEventSynchronize(ctx->done));
    int32_t rc = settle_stage(ctx);
    if (rc) return rc;
    if (ctx->timing) {
        for (size_t i = 0; i < ctx->event_kind.size(); ++i) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, ctx->events[2 * i], ctx->events[2 * i + 1]));
            ctx->kernel_ms[ctx->event_kind[i]] += ms;
            ctx->kernel_count[ctx->event_kind[i]] += 1;
        }
    }
    if (trace)
        for (int sl = 0; sl < ctx->slots; ++sl)
            memcpy(trace + (size_t)sl * steps * NHF_TRACE_COLS, ctx->host_ctl[sl].trace,
                   sizeof(uint32_t) * NHF_TRACE_COLS * steps);
    return 0;
}

int32_t nhb_run_batch(nhb_ctx *ctx, int32_t steps, uint32_t *trace) {
    if (!ctx) return set_error_text("nhb_run_batch: null context");
    int32_t rc = nhb_launch_batch(ctx, steps);
    if (rc) return rc;
    return nhb_wait_batch(ctx, trace);
}

int32_t nhb_compute_cfl(nhb_ctx *ctx) {
    if (!ctx) return set_error_text("nhb_compute_cfl: null context");
    const BDev &d = ctx->d;
    int32_t rc = apply_stage(ctx);
    if (rc) return rc;
    if (ctx->slots == 64)
        hipLaunchKernelGGL((kb_cfl<true, NHB_CFL_ROUND>), dim3(d.cfl_blocks), dim3(NHF_BLOCK), 0, ctx->stream, d, 0, 0);
    else
        hipLaunchKernelGGL((kb_cfl<false, 1>), dim3(d.cfl_blocks), dim3(NHF_BLOCK), 0, ctx->stream, d, 0, 0);
    hipLaunchKernelGGL(kb_finish, dim3(ctx->slots), dim3(NHF_BLOCK), 0, ctx->stream, d, 0, 0, 0);
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhb_flush_budget(nhb_ctx *ctx, int32_t slot, double *out) {
    if (check_slot(ctx, slot)) return -2;
    const BDev &d = ctx->d;
    hipLaunchKernelGGL(kb_flush, dim3(1), dim3(NHF_BLOCK), 0, ctx->stream, d, (int)slot);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->host_flush, d.flush, sizeof(double2) * d.terms, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int nonfinite = 0;
    for (int t = 0; t < d.terms; ++t) {
        double hi = t == 3 ? 0.0 : ctx->host_flush[t].x, lo = t == 3 ? 0.0 : ctx->host_flush[t].y;
        if (!isfinite(hi) || !isfinite(lo)) nonfinite = 1;
        out[2 * t] = hi;
        out[2 * t + 1] = lo;
    }
    return nonfinite;
}

int32_t nhb_limit_capture(nhb_ctx *ctx, int32_t slot, float *values, int32_t *ids, int32_t reset) {
    if (check_slot(ctx, slot)) return -2;
    Ctl *ctl = ctx->d.ctl + slot;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (values) HIP_TRY(hipMemcpy(values, ctl->limit_value, sizeof(float) * 5, hipMemcpyDeviceToHost));
    if (ids) HIP_TRY(hipMemcpy(ids, ctl->limit_ids, sizeof(int32_t) * 3, hipMemcpyDeviceToHost));
    if (reset) {
        hipLaunchKernelGGL(kb_reset_limit, dim3(1), dim3(64), 0, ctx->stream, ctx->d, (int)slot);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

int32_t nhb_cfl_state(nhb_ctx *ctx, int32_t slot, float *ndt, int32_t *side) {
    if (check_slot(ctx, slot)) return -2;
    Ctl *ctl = ctx->d.ctl + slot;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ndt) HIP_TRY(hipMemcpy(ndt, &ctl->ndt, sizeof(float), hipMemcpyDeviceToHost));
    if (side) HIP_TRY(hipMemcpy(side, &ctl->cfl_arg, sizeof(int32_t), hipMemcpyDeviceToHost));
    return 0;
}

int32_t nhb_settle_state(nhb_ctx *ctx, int32_t slot) {
    if (!ctx) return set_error_text("nhb_settle_state: null context");
    if (slot < -1 || slot >= ctx->slots) return set_error_text("nhb_settle_state: slot out of range");
    if (ctx->pending_steps) return set_error_text("nhb_settle_state: a batch is in flight");
    const int blocks = b_blocks((size_t)ctx->d.m.e_num);
    hipLaunchKernelGGL(kb_settle_state, dim3(blocks < 2048 ? blocks : 2048), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d,
                       (int)slot);
    HIP_TRY(hipGetLastError());
    return 0;
}

int32_t nhb_set_rain_state(nhb_ctx *ctx, int32_t slot, int32_t uniform, float rain_time, const float *class_capacity) {
    if (check_slot(ctx, slot)) return -2;
    if (ctx->pending_steps) return set_error_text("nhb_set_rain_state: a batch is in flight");
    // the flux pass advances the uniform state, so a mesh without sides keeps the per-cell one
    if (uniform && (ctx->d.m.n_pixels > 1 || !class_capacity || ctx->d.m.s_num <= 1))
        return set_error_text("nhb_set_rain_state: the uniform state needs one rain pixel, sides and the class "
                              "capacities");
    if (uniform)
        HIP_TRY(hipMemcpyAsync(ctx->d.cap_stage, class_capacity, sizeof(float) * 256, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(kb_set_rain_state, dim3(1), dim3(NHF_BLOCK), 0, ctx->stream, ctx->d, (int)slot, (int)(uniform != 0),
                       rain_time, (const float *)ctx->d.cap_stage);
    HIP_TRY(hipGetLastError());
    // the host's capacities are read by the copy before this returns
    if (uniform) HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int32_t nhb_sync(nhb_ctx *ctx) {
    if (!ctx) return set_error_text("nhb_sync: null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int32_t nhb_set_graph(nhb_ctx *ctx, int32_t enable) {
    if (!ctx) return set_error_text("nhb_set_graph: null context");
    ctx->graph_enabled = enable != 0;
    return 0;
}

int32_t nhb_set_tuning(nhb_ctx *ctx, int32_t depth_blocks, int32_t cfl_blocks) {
    if (!ctx) return set_error_text("nhb_set_tuning: null context");
    BDev &d = ctx->d;
    if (depth_blocks < 8 || depth_blocks > d.depth_max || (depth_blocks & 7) || cfl_blocks < 8 ||
        cfl_blocks > d.cfl_max || (cfl_blocks & 7))
        return set_error_text("nhb_set_tuning: grid out of range or not a multiple of 8");
    if (ctx->pending_steps) return set_error_text("nhb_set_tuning: a batch is in flight");
    if (depth_blocks != d.depth_blocks || cfl_blocks != d.cfl_blocks) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i <= NHF_MAX_BATCH; ++i)
            if (ctx->exec[i]) {
                (void)hipGraphExecDestroy(ctx->exec[i]);
                ctx->exec[i] = nullptr;
            }
        d.depth_blocks = depth_blocks;
        d.cfl_blocks = cfl_blocks;
    }
    return 0;
}

int32_t nhb_tuning_range(nhb_ctx *ctx, int32_t *out4) {
    if (!ctx) return set_error_text("nhb_tuning_range: null context");
    out4[0] = ctx->d.depth_blocks;
    out4[1] = ctx->d.depth_max;
    out4[2] = ctx->d.cfl_blocks;
    out4[3] = ctx->d.cfl_max;
    return 0;
}

int32_t nhb_set_timing(nhb_ctx *ctx, int32_t enable) {
    if (!ctx) return set_error_text("nhb_set_timing: null context");
    ctx->timing = enable != 0;
    return 0;
}

int32_t nhb_kernel_times(nhb_ctx *ctx, double *ms, int64_t *count, int32_t n, int32_t reset) {
    if (!ctx) return set_error_text("nhb_kernel_times: null context");
    for (int i = 0; i < n && i < KB_COUNT; ++i) {
        ms[i] = ctx->kernel_ms[i];
        count[i] = ctx->kernel_count[i];
    }
    if (reset) {
        memset(ctx->kernel_ms, 0, sizeof ctx->kernel_ms);
        memset(ctx->kernel_count, 0, sizeof ctx->kernel_count);
    }
    return 0;
}

}  // extern "C"
