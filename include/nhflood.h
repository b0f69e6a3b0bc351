/*
 * nhflood.h - C ABI of the compact local-inertial shallow-water step.
 *
 * One library, libnhflood.so, implements this ABI with hand-written HIP kernels for gfx950.
 * Every array in nhf_mesh is a DEVICE pointer owned by the caller (the Python host allocates
 * them as torch tensors); the library never frees them and, apart from the nhf_clock_probe test
 * hook, allocates no device memory of its own: its scratch lives in the caller's workspace
 * (nhf_workspace_bytes). The CPU oracle under oracle/ (test infrastructure) exports the same
 * entry points over host arrays so one host driver can run either.
 *
 * The reference has no FFI: its device step is a set of Taichi kernels closed over one Python
 * function (src/nh_flood_2d/core/solver_compact.py:179-1008). Each entry point below replaces
 * one piece of that Python/Taichi contract; the citation on each says which.
 *
 * Conventions kept from the reference:
 *   - index 0 of every cell and side array is the exterior sentinel (README.md:407);
 *   - a side's lower/higher cells are bottom/top for horizontal sides (orient 1) and left/right
 *     for vertical sides (orient 2) (preprocess/domain.py:456-460);
 *   - a cell's side list holds its left, right, bottom, top faces in that order
 *     (preprocess/domain.py:83-174);
 *   - the clock counts 2^-40 s ticks (core/clock.py:1-30), here as one uint64_t;
 *   - the per-step trace row has 9 uint32 columns (core/clock.py:35-39): dt bits, end tick hi,
 *     end tick lo, status, CFL candidate bits, event-short flag, limited cells, limited faces,
 *     minimum donor scale bits.
 * All functions return 0 on success, a negative value on a HIP or argument error
 * (nhf_last_error() then describes it).
 */
#ifndef NHFLOOD_H
#define NHFLOOD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NHF_ABI_VERSION 12
#define NHF_TRACE_COLS 9
#define NHF_MAX_BATCH 64

/* Clock status codes (core/clock.py:42-46). */
#define NHF_DONE 0
#define NHF_COMPLETED 1
#define NHF_BAD_CFL 2
#define NHF_BAD_DT 3
#define NHF_OVERSHOOT 4

/* side_meta bits */
#define NHF_SIDE_ORIENT_MASK 3u /* 1 = horizontal (bottom/top), 2 = vertical (left/right) */
#define NHF_SIDE_EXTERIOR 4u    /* sbf == 1: exactly one neighbour is the sentinel */

/* Physics constants, all already rounded to F32 the way the reference's Taichi kernels see
 * them (solver_compact.py:208 g = 9.81; input/domain.py:34-36 afa, sita, min_h). */
typedef struct nhf_params {
    float g;          /* 9.81f */
    float afa;        /* CFL number */
    float theta;      /* sita */
    float theta_half; /* F32 of the F64 value (1 - sita) / 2 (solver_compact.py:461) */
    float min_h;      /* wet/dry threshold */
    float min_dt;     /* F32 of 1e-4, the smallest CFL candidate (clock.py:49-57) */
    int32_t batch_capacity; /* steps per batch the trace can hold, 1..NHF_MAX_BATCH */
    int32_t reserved;
} nhf_params;

/* Device arrays of one mesh in the library's layout. Sizes: e_num cells (with sentinel),
 * s_num sides (with sentinel), r_len side-list entries, b_num boundaries, n_pixels rain pixels,
 * n_grad gradient cells and n_hang hanging edges (both 0 on aligned meshes). */
typedef struct nhf_mesh {
    int32_t e_num, s_num, r_len, b_num;
    int32_t n_pixels, n_grad, n_hang, flags; /* flags: bit0 pipe_q present */

    /* per-cell state (read-write) */
    float *depth;      /* water depth h (m) */
    float *bed;        /* bed elevation z (m) */
    float *rain_time;  /* Horton cumulative rain time (s) */
    float *capacity;   /* Horton infiltration capacity (m/s) */
    float *max_depth;  /* extrema since reset */
    float *max_speed;
    uint64_t *peak_tick; /* first tick of max_depth */
    /* per-cell static */
    const float *area;                /* F32 of the F64 width x height (MeshGeometry.area) */
    const float *size;                /* edge length; cells are square (force.py:52-53) */
    const uint32_t *cell_meta;        /* land class | (boundary index + 1) << 8 (0 = none)
                                         | size level << 16 when size_base > 0 */
    const float *class_horton;        /* [256*3] Horton f0, fc (m/s), k (1/h) per land class */
    const int32_t *rain_index;        /* cell -> rain pixel, read only when n_pixels > 1 */
    const float *bweight;             /* discharge share of the boundary */
    const int32_t *csr_start;         /* first entry of the cell's l, r, b, t side list */
    const uint32_t *csr_count;        /* side counts: l | r << 8 | b << 16 | t << 24 */
    const float *pipe_q;              /* optional pipe exchange (m^3/s), NULL when absent */
    const float *open_face;           /* [e_num*4] exterior open fraction, NULL when sita == 1 */
    const float *x, *y;               /* cell centres, NULL on aligned meshes */
    const int32_t *isl;               /* [r_len] side ids */
    const int32_t *isl_cell;          /* [r_len] the cell across each listed side, 0 = exterior */
    /* per-side */
    float *q;                         /* unit discharge (m^2/s), state */
    const int32_t *side_cells;        /* [s_num*2] lower, higher */
    const uint8_t *side_meta;         /* orient | exterior flag */
    const float *length, *sdc, *manning;
    const uint32_t *side_key;         /* reference side id, smallest wins CFL ties */
    /* hanging-edge pressure reconstruction (core/pressure.py:19-45), NULL on aligned meshes */
    const int32_t *edge_index;        /* [s_num] -1 or hanging edge */
    const float *hang_offset;         /* [n_hang*2] tangential offsets of lower, higher centre */
    const int32_t *hang_slot;         /* [n_hang*2] gradient slot of lower, higher or -1 */
    const int32_t *grad_cell;         /* [n_grad] */
    const int32_t *grad_axes;         /* [n_grad] bit0 x-gradient needed, bit1 y-gradient */
    /* per-boundary */
    const int32_t *kind;              /* [b_num] 0 level, 1 discharge */
    /* pipe coupling (solver_compact.py:535-539), NULL unless pipe_q is set: the volume each cell
       exchanged with the pipe network in the current coupling window, summed in F32 with a
       compensation term; the host reads and clears both at each window end */
    float *pipe_volume, *pipe_correction;
    /* [e_num*4] the cell passes' view of the side lists: a cell with one side per face, each of
       length size[cell], and area size*size (in F32) holds its left, right, bottom and top side
       ids; any other cell holds (-1, csr_start, csr_count, list length) */
    const int32_t *cell_quad;
    /* > 0: every cell's size is size_base * 2^level exactly (level = cell_meta bits 16..19), and
       the kernels compute it rather than read `size`; 0: they read `size` */
    float size_base;
    int32_t reserved_tail;
} nhf_mesh;

/* Gate control (core/gate.py:35-268; layout preprocess/gate.py:8-96). Arrays are in the library's
 * numbering and owned by the caller, like nhf_mesh's: device arrays for the HIP library, host
 * arrays for the CPU oracle. flags, ticks and acc are state the library updates in place. */
typedef struct nhf_gates {
    int32_t count;            /* gates */
    int32_t n_cells;          /* gate cells listed in `cells` */
    int32_t n_items;          /* accounting items */
    int32_t n_chunks;         /* accounting groups of at most 256 items */
    float trigger_level, opening_seconds, open_elevation, min_h; /* F32, as the reference's kernels */
    uint64_t formal_tick;     /* a gate can be commanded from this clock tick on (the formal start) */
    const int32_t *control;   /* [count] control cell of each gate */
    const int32_t *cells;     /* [n_cells] every gate cell */
    const int32_t *cell_gate; /* [e_num] gate of each cell, -1 elsewhere */
    const uint8_t *pond;      /* [e_num] 1 for a pond cell */
    const float *closed;      /* [e_num] the closed bed (the mesh bed) */
    /* accounting: side ids (bit 31 set: the side's volume counts negated) and cell ids; each
       chunk [4] = first item, items, kind (0 sides, 1 cells) | (first-flow gate + 1) << 8, first
       term. Sides add max(v, 0), max(-v, 0) of v = +-q*length*dt to two terms, cells their six
       source terms (record_flux, record_cell, collect: core/gate.py:119-188). */
    const int32_t *items;
    const int32_t *chunks;
    int32_t *flags;           /* [count*3] commanded, first flow seen, fully open seen */
    uint64_t *ticks;          /* [count*3] their clock ticks */
    double *acc;              /* [n_chunks*6*2] per chunk and term: sum, compensation (zeroed) */
} nhf_gates;

typedef struct nhf_ctx nhf_ctx;

/* Version of this ABI; the loader refuses a library that reports another. */
int32_t nhf_abi_version(void);

/* Thread-local text of the last error. */
const char *nhf_last_error(void);

/* Bytes of scratch the library needs for this mesh (control block, budget accumulators, CFL
 * partials, face averages, gradients, donor scales, unscaled discharge). */
size_t nhf_workspace_bytes(const nhf_mesh *mesh, const nhf_params *params);

/* Bind a mesh, parameters, a caller-owned device workspace and a HIP stream (hipStream_t, may be
 * NULL for the null stream). Replaces the field allocation and kernel definitions of _evolve
 * (solver_compact.py:204-687). */
int32_t nhf_create(nhf_ctx **out, const nhf_mesh *mesh, const nhf_params *params,
                   void *workspace, size_t workspace_bytes, void *stream);
int32_t nhf_destroy(nhf_ctx *ctx);

/* Set the device clock (StepClock.__init__ now words, clock.py:58-69). Also zeroes the budget
 * accumulators, donor scales and the CFL limit capture. */
int32_t nhf_reset(nhf_ctx *ctx, uint64_t now_ticks);

/* Upload the forcing of a batch (upload_forcing, solver_compact.py:769-790): coefficients
 * [b_num*2] (value at batch start, slope per second) and, when rain_rates != NULL, the rain rate
 * of every pixel (m/s). Host pointers. */
int32_t nhf_set_forcing(nhf_ctx *ctx, const float *coefficients, const float *rain_rates,
                        int32_t raining);

/* Start a batch (StepClock.begin, clock.py:75-92): target and origin ticks, trace cleared. */
int32_t nhf_begin(nhf_ctx *ctx, uint64_t target_ticks, uint64_t origin_ticks);

/* Run `steps` slots of [prepare_step, (update_pressure), tick, compute_cfl, finish_step]
 * (the graph of solver_compact.py:894-909, run by graph.run at :941) and copy the
 * steps x 9 trace to `trace` (host). Slots after the target, or after a failure status, do no
 * physics (clock.py:94-167). */
int32_t nhf_run_batch(nhf_ctx *ctx, int32_t steps, uint32_t *trace);

/* nhf_run_batch in two halves, so one host thread can keep several contexts (each on its own
 * stream) busy at once: nhf_launch_batch enqueues the batch and its trace copy and returns;
 * nhf_wait_batch waits for that batch and copies its trace. One batch in flight per context. */
int32_t nhf_launch_batch(nhf_ctx *ctx, int32_t steps);
int32_t nhf_wait_batch(nhf_ctx *ctx, uint32_t *trace);

/* Standalone compute_cfl on the current state (solver_compact.py:617-641, called once at
 * :889 before the first batch). */
int32_t nhf_compute_cfl(nhf_ctx *ctx);

/* Interval water-budget totals (DeviceBudget.flush, device_budget.py:68-84) as double-double
 * pairs: out[2*t] + out[2*t+1] for each of the 4 + 6*b_num terms (solver_compact.py:393-396,
 * 533, 578-590); the accumulators are cleared. Returns 1 (not an error code) if a term is
 * non-finite, after clearing nothing. */
int32_t nhf_flush_budget(nhf_ctx *ctx, double *out);

/* CFL-limit capture of prepare_step (solver_compact.py:374-382): values = candidate, lower depth,
 * higher depth, face depth, face velocity; ids = side, lower cell, higher cell. `reset` puts the
 * captured candidate back to +inf (solver_compact.py:854). */
int32_t nhf_limit_capture(nhf_ctx *ctx, float *values, int32_t *ids, int32_t reset);

/* Last CFL candidate and its side (ndt_t, cfl_arg_t; used by supported_dt, :644-655). */
int32_t nhf_cfl_state(nhf_ctx *ctx, float *ndt, int32_t *side);

/* Wait for all work of the context's stream. */
int32_t nhf_sync(nhf_ctx *ctx);

/* Bind gate control to the step (NULL: none). Each step then runs, as the reference's graph
 * (solver_compact.py:894-909): the trigger test and gate-bed motion at its start (check_gates),
 * the exchange and source accounting after the limiter (record_flux, record_cell, collect), and
 * the gate-bed motion to its end (finish_gate_bed). The arrays must outlive the binding. */
int32_t nhf_set_gates(nhf_ctx *ctx, const nhf_gates *gates);

/* The trigger test and bed motion at the current clock, once, outside a batch (the check_gates
 * call before the first CFL, solver_compact.py:885-889). */
int32_t nhf_gate_check(nhf_ctx *ctx);

/* Gate budget since the last flush (GateController.budget.flush, core/gate.py:222) as
 * double-double pairs: out[2*t] + out[2*t+1] for each of the 10*count + 8 terms; cleared. */
int32_t nhf_flush_gate_budget(nhf_ctx *ctx, double *out);

/* Make max_depth and peak_tick final (the extrema of solver_compact.py:600-606 as
 * reset_extrema/write_extrema see them, :674-679, output.py:130-150). The step leaves a cell
 * whose depth rose in its last step marked by a negative max_depth (its maximum is its depth,
 * reached at the end of that step) and writes the pair out when the rise stops; this call writes
 * out every marked cell, and in the uniform rain state (nhf_set_rain_state) every cell's
 * rain_time and capacity. Call it before reading max_depth, peak_tick, rain_time or capacity,
 * and before writing any of them or depth alone (every part of a parts context).
 * No reference counterpart (the reference writes both arrays on every rise). */
int32_t nhf_settle_extrema(nhf_ctx *ctx);

/* The rain state (no reference counterpart). With one rain pixel every cell rains alike (the
 * depth update of solver_compact.py:543-548), so once every cell has the same rain time and each
 * land class one capacity, that stays so: in the uniform state (uniform != 0; rain_time of every
 * cell, class_capacity [256] the capacity of each class code, a host array) the context keeps one
 * rain time and a capacity per class, advanced once per step, and the depth pass neither reads
 * nor writes the per-cell arrays; nhf_settle_extrema writes them out. uniform == 0 returns to the
 * per-cell arrays, which must be current (settled) then. The caller sets the state after writing
 * rain_time or capacity; a context starts per-cell. Results are the same bits in both states.
 * Refused with more than one rain pixel. */
int32_t nhf_set_rain_state(nhf_ctx *ctx, int32_t uniform, float rain_time, const float *class_capacity);

/* Launch mode: 1 replays each batch as one captured hipGraph, 0 launches kernels one by one. */
int32_t nhf_set_graph(nhf_ctx *ctx, int32_t enable);

/* Launch tuning (no reference counterpart; used by the autotuner): the number of CFL partial
 * workgroups (a multiple of 8 up to the maximum the workspace holds), the workgroup size of the
 * two cell passes (128, 256 or 512 threads) and their cells per thread (1 or 2). Every choice
 * computes the same state bits; budget accumulators are kept across a change (the sums are
 * grouped differently, FP64 rounding only). out4 of nhf_tuning_range = CFL grid, CFL maximum,
 * cell workgroup size, cells per thread. Captured graphs are rebuilt after a change. */
int32_t nhf_set_tuning(nhf_ctx *ctx, int32_t cfl_blocks, int32_t cell_block, int32_t cells_per_thread);
int32_t nhf_tuning_range(nhf_ctx *ctx, int32_t *out4);

/* Per-kernel timing with HIP events on the context's stream (eager mode only). Kernel ids:
 * 0 prepare, 1 gradients, 2 face averages, 3 flux, 4 donor, 5 depth, 6 cfl, 7 finish, 8 gate
 * control. `ms` and `count` receive totals since the last call with reset != 0. */
int32_t nhf_set_timing(nhf_ctx *ctx, int32_t enable);
int32_t nhf_kernel_times(nhf_ctx *ctx, double *ms, int64_t *count, int32_t n, int32_t reset);

/* Benchmark hook with no reference entry point: prepare one step towards the begun target, then
 * time each step kernel over `repeats` back-to-back launches with HIP events on the context's
 * stream; us[id] (9 entries) receives the mean per launch (kernel ids as for nhf_kernel_times, 0
 * for kernels that do not run on this mesh, and for gate control). The launches repeat each pass
 * on the same inputs except the depth pass, which updates the state in place, so the context's
 * state is not meaningful afterwards. */
int32_t nhf_time_kernels(nhf_ctx *ctx, int32_t repeats, double *us);

/* Flood-map rasteriser (replaces rasterize_tile_kernel, output/flood_map.py:180-222, and
 * rasterize_wse, :688-721, whose masking the caller applies to `value` first). Cell i is the
 * rectangle x[i] +- half_w[i], y[i] +- half_h[i]; it writes value[i] to every pixel of the tile
 * whose centre it holds ([min, max) along x, (min, max] along y), where the tile is rows x cols
 * pixels of hr x vr with its top-left corner at (tile_min_x, tile_max_y) and row r starts at
 * tile + r * ld. F32 arithmetic in the reference's order; pixels no cell holds are left as they
 * are (the caller fills the tile with its nodata value first). All arrays are DEVICE pointers;
 * the launch is asynchronous on `stream`. */
int32_t nhf_rasterize(const float *x, const float *y, const float *half_w, const float *half_h, const float *value,
                      int32_t n, float tile_min_x, float tile_max_y, float hr, float vr, int32_t rows, int32_t cols,
                      float *tile, int64_t ld, void *stream);

/* Test hook with no reference entry point: evaluate StepClock.choose (clock.py:94-167) for n
 * independent cases on the device. Inputs and outputs are host arrays. out_words[i*4 + 0..3]
 * = status, dt bits, CFL bits, event-short flag; out_ticks[i*2 + 0..1] = step end tick and the
 * tick distance of the F32 dt; out_times[i*2 + 0..1] = local, end_local. */
int32_t nhf_clock_probe(const uint64_t *now, const uint64_t *target, const uint64_t *origin,
                        const float *candidate, float min_dt, int32_t n, uint32_t *out_words,
                        uint64_t *out_ticks, float *out_times);

/* ------------------------------------------------------------------------------------------
 * One batch with no context and no allocation.
 *
 * No counterpart in the reference. It is for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. nhf_run_static(...) computes
 * what this sequence on a fresh context computes, bit for bit:
 *   nhf_reset(now_ticks); nhf_set_forcing(coefficients, rain_rates, raining); nhf_compute_cfl;
 *   nhf_begin(target_ticks, origin_ticks); nhf_run_batch(steps, trace);
 *   nhf_flush_budget(budget) when budget != NULL; nhf_settle_extrema.
 * The mesh and its state arrays are the caller's device arrays, as for nhf_create, and are
 * updated in place; coefficients, rain_rates, trace and budget are host arrays. The library's
 * scratch (control block, CFL partials, donor scales, face averages, gradients, budget
 * accumulators) is one file-scope device array of NHF_STATIC_SCRATCH_BYTES per GPU, held from the
 * library's load until the process ends; a mesh whose scratch would not fit is refused with an
 * error (about 10.5 B per cell at sita 1 with one boundary, 16 B more per cell at sita < 1).
 * The call makes no hipMalloc, hipFree, hipHostMalloc or event, stream or graph creation; it
 * launches on `stream` (NULL: the null stream) and returns after synchronising it.
 * Every call on one GPU uses the same scratch and leaves its contents behind: each call clears
 * what must start at zero, and calls that overlap on one GPU (several host threads, or streams
 * running at once) race and give wrong results. Gate control and parts are not supported.
 * Returns 1 (not an error code) when a budget total is non-finite, as nhf_flush_budget.
 * ------------------------------------------------------------------------------------------ */
#define NHF_STATIC_SCRATCH_BYTES (64u << 20)

int32_t nhf_run_static(const nhf_mesh *mesh, const nhf_params *params, uint64_t now_ticks, uint64_t target_ticks,
                       uint64_t origin_ticks, const float *coefficients, const float *rain_rates, int32_t raining,
                       int32_t steps, uint32_t *trace, double *budget, void *stream);

/* ------------------------------------------------------------------------------------------
 * Scenario batch: up to NHB_MAX_SLOTS scenarios ("slots") on one shared mesh, advanced by the
 * same launches, each with its own state, forcing, clock, dt, CFL, trace and budget.
 *
 * The reference runs one scenario per process and has no batched entry point; each nhb_*
 * function is the per-slot form of the nhf_* function of the same name above (same semantics,
 * same reference lines), so a slot behaves exactly like a context of its own. Per-scenario
 * arrays are laid out [entity][slot] (the slot index innermost, `slots` words per cell or
 * side); the mesh's static arrays are shared. The state pointers inside nhf_mesh are ignored.
 * ------------------------------------------------------------------------------------------ */
#define NHB_MAX_SLOTS 64

typedef struct nhb_state {
    /* per-scenario state, [e_num * slots] or [s_num * slots] (q), device pointers */
    float *depth, *rain_time, *capacity, *max_depth, *max_speed;
    uint64_t *peak_tick;
    float *q;
    const float *manning;        /* [s_num * slots] per-scenario side Manning n, or NULL to share
                                    mesh->manning */
    const float *class_manning;  /* [256 * slots] per-scenario Manning n by land class code, slot
                                    innermost, or NULL. When set, `manning` is not read: a side's n
                                    is rebuilt per scenario from its two cells' classes (the class
                                    bits of mesh->cell_meta) as map_land_parameters forms it
                                    (preprocess/hydraulics.py:63-93): the F32 mean of the two cells'
                                    n, or the interior cell's n on an exterior side. */
    /* boundary cells (static): bslot[cell] = index into the list or -1; the list is sorted by
       boundary, and bcell_start[b] .. bcell_start[b + 1] are boundary b's cells */
    const int32_t *bslot;        /* [e_num] */
    const int32_t *bcell_start;  /* [b_num + 1] */
    int32_t n_bcells, reserved;
} nhb_state;

typedef struct nhb_ctx nhb_ctx;

/* Workspace bytes for `slots` scenarios (a power of two, 1..NHB_MAX_SLOTS). */
size_t nhb_workspace_bytes(const nhf_mesh *mesh, const nhb_state *state, const nhf_params *params,
                           int32_t slots);
/* nhf_create for a batch; every slot starts idle at tick 0. pipe_q must be NULL. */
int32_t nhb_create(nhb_ctx **out, const nhf_mesh *mesh, const nhb_state *state, const nhf_params *params,
                   int32_t slots, void *workspace, size_t workspace_bytes, void *stream);
int32_t nhb_destroy(nhb_ctx *ctx);
/* nhf_reset of one slot (clock, budget, limiter capture). */
int32_t nhb_reset(nhb_ctx *ctx, int32_t slot, uint64_t now_ticks);
/* nhf_set_forcing / nhf_begin of one slot: staged on the host, applied by the next
 * nhb_run_batch or nhb_compute_cfl in one upload for all slots. */
int32_t nhb_set_forcing(nhb_ctx *ctx, int32_t slot, const float *coefficients, const float *rain_rates,
                        int32_t raining);
int32_t nhb_begin(nhb_ctx *ctx, int32_t slot, uint64_t target_ticks, uint64_t origin_ticks);
/* nhf_run_batch for every slot at once: each slot takes up to `steps` steps towards its own
 * target with its own dt; trace receives slots x steps x 9 words. */
int32_t nhb_run_batch(nhb_ctx *ctx, int32_t steps, uint32_t *trace);
/* nhb_run_batch in two halves, as nhf_launch_batch / nhf_wait_batch. */
int32_t nhb_launch_batch(nhb_ctx *ctx, int32_t steps);
int32_t nhb_wait_batch(nhb_ctx *ctx, uint32_t *trace);
/* nhf_compute_cfl for every slot. */
int32_t nhb_compute_cfl(nhb_ctx *ctx);
/* nhf_flush_budget / nhf_limit_capture / nhf_cfl_state of one slot. */
int32_t nhb_flush_budget(nhb_ctx *ctx, int32_t slot, double *out);
int32_t nhb_limit_capture(nhb_ctx *ctx, int32_t slot, float *values, int32_t *ids, int32_t reset);
int32_t nhb_cfl_state(nhb_ctx *ctx, int32_t slot, float *ndt, int32_t *side);
int32_t nhb_sync(nhb_ctx *ctx);
int32_t nhb_set_graph(nhb_ctx *ctx, int32_t enable);
/* Launch tuning of a batch: grid of the grid-stride depth pass (cells x slots per thread follow
 * from it) and of the CFL pass, multiples of 8 up to the allocated maxima. out4 of
 * nhb_tuning_range = depth grid, depth maximum, CFL grid, CFL maximum. */
int32_t nhb_set_tuning(nhb_ctx *ctx, int32_t depth_blocks, int32_t cfl_blocks);
int32_t nhb_tuning_range(nhb_ctx *ctx, int32_t *out4);
int32_t nhb_set_timing(nhb_ctx *ctx, int32_t enable);
int32_t nhb_kernel_times(nhb_ctx *ctx, double *ms, int64_t *count, int32_t n, int32_t reset);
/* nhf_settle_extrema of one slot (slot -1: every slot), which also writes out the rain time and
 * capacity of a slot in the uniform rain state (below). Call it before reading a slot's
 * max_depth, peak_tick, rain_time or capacity, and before writing its depth, max_depth or
 * peak_tick. No reference counterpart. */
int32_t nhb_settle_state(nhb_ctx *ctx, int32_t slot);
/* The rain state of one slot (no reference counterpart). With one rain pixel every cell rains
 * alike (the depth update of solver_compact.py:543-548), so once every cell has the same rain time
 * and each land class one capacity, that stays so: in the uniform state (uniform != 0; rain_time
 * of every cell, class_capacity [256] the capacity of each class code, a host array) the slot
 * keeps one rain time and a capacity per class, advanced once per step, and the depth pass does
 * not read or write the per-cell arrays; nhb_settle_state writes them out. uniform == 0 returns
 * the slot to its per-cell arrays, which must be current (settled) then. The caller sets the
 * state after writing a slot's rain_time and capacity; every slot starts per-cell. Results are
 * the same bits in both states. Refused with more than one rain pixel. */
int32_t nhb_set_rain_state(nhb_ctx *ctx, int32_t slot, int32_t uniform, float rain_time, const float *class_capacity);

/* ------------------------------------------------------------------------------------------
 * Meshes of different sizes advanced by the same launches (nhu_*).
 *
 * No reference counterpart either: the reference runs one mesh per process. Several meshes
 * ("parts") are concatenated into one nhf_mesh: part p's cells occupy [cell_begin, cell_end)
 * with cell_begin = 1 + 1024 k and the range padded to a multiple of 1024 (padding cells have
 * no sides and are never stepped); its sides likewise; every index inside a part's arrays
 * (side cells, side lists, cell records, boundary ids in cell_meta, rain pixels, hanging-edge
 * data) is global; cell 0 and side 0 are the one shared sentinel. Each part keeps its own clock,
 * dt, forcing, CFL reduction, limiter capture and budget, so a part steps exactly as it would in
 * a context of its own (nhf_create) on the same arrays; one nhu_run_batch advances every part
 * with a begun target. Gate control and pipe exchange are not supported here. The nhf_* entry
 * points that take no part (destroy, sync, graph, timing, tuning of the cell passes, settle
 * extrema) apply to a parts context too; the others are replaced by the nhu_* forms below.
 * ------------------------------------------------------------------------------------------ */
#define NHF_MAX_PARTS 256

typedef struct nhf_part {
    int32_t cell_begin, cell_end;  /* real cells of the part */
    int32_t side_begin, side_end;  /* real sides of the part (begin = 1 + 1024 k) */
    int32_t b_base, b_count;       /* its prescribed boundaries (ids in cell_meta are global) */
    int32_t pix_base, pix_count;   /* its rain pixels */
} nhf_part;

/* Workspace bytes of a parts context; parts is a host array of `count` entries that tile the
 * mesh's cells, sides, boundaries and rain pixels in order. */
size_t nhu_workspace_bytes(const nhf_mesh *mesh, const nhf_params *params, const nhf_part *parts,
                           int32_t count);
/* nhf_create for parts; every part starts idle at tick 0. pipe_q must be NULL. */
int32_t nhu_create(nhf_ctx **out, const nhf_mesh *mesh, const nhf_params *params, const nhf_part *parts,
                   int32_t count, void *workspace, size_t workspace_bytes, void *stream);
/* nhf_reset / nhf_set_forcing / nhf_begin of one part: coefficients [b_count*2] and rain_rates
 * [pix_count] of that part only. */
int32_t nhu_reset(nhf_ctx *ctx, int32_t part, uint64_t now_ticks);
int32_t nhu_set_forcing(nhf_ctx *ctx, int32_t part, const float *coefficients, const float *rain_rates,
                        int32_t raining);
int32_t nhu_begin(nhf_ctx *ctx, int32_t part, uint64_t target_ticks, uint64_t origin_ticks);
/* nhf_run_batch for every part at once; trace receives count x steps x 9 words. */
int32_t nhu_run_batch(nhf_ctx *ctx, int32_t steps, uint32_t *trace);
int32_t nhu_launch_batch(nhf_ctx *ctx, int32_t steps);
int32_t nhu_wait_batch(nhf_ctx *ctx, uint32_t *trace);
/* nhf_compute_cfl for every part. */
int32_t nhu_compute_cfl(nhf_ctx *ctx);
/* nhf_flush_budget of one part: out receives 2 * (4 + 6 * b_count) doubles. */
int32_t nhu_flush_budget(nhf_ctx *ctx, int32_t part, double *out);
/* nhf_limit_capture / nhf_cfl_state of one part (side and cell ids are global). */
int32_t nhu_limit_capture(nhf_ctx *ctx, int32_t part, float *values, int32_t *ids, int32_t reset);
int32_t nhu_cfl_state(nhf_ctx *ctx, int32_t part, float *ndt, int32_t *side);
/* nhf_set_rain_state of one part: the part's own rain pixel, class_capacity indexed by the
 * context's class codes (those in the part's cell_meta). */
int32_t nhu_set_rain_state(nhf_ctx *ctx, int32_t part, int32_t uniform, float rain_time, const float *class_capacity);

#ifdef __cplusplus
}
#endif

#endif /* NHFLOOD_H */
