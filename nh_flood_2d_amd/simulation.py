"""Host driver of one flood run: the reference's ``_evolve`` loop around the HIP step.

Restates ``src/nh_flood_2d/core/solver_compact.py:179-1008`` on the host side:
set-up (land mapping, boundaries, still water, initial levels, initial CFL and extrema), the
batch loop (target selection :930-937, forcing upload :769-790, one trace readback per batch
:941-970), the F64 water ledger (``WaterBudget``, core/run_record.py:107-161) and the output
instants (:815-855). The device step itself is the HIP library behind :class:`Engine`.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field
from fractions import Fraction
from time import perf_counter

import numpy as np

from . import abi
from .clock import TICKS_PER_SECOND, f32_from_bits, row_end_ticks, seconds_to_ticks
from .engine import Engine, make_params
from .forcing import Forcing, cell_forcing_arrays
from .hydraulics import connected_still_water
from .layout import build_layout, initial_state, relayout
from .mesh import Mesh

MIN_DT = 1e-4
CHECKPOINT_INTERVAL_S = 30 * 60
FRAME_INTERVAL_S = 900


class WaterBudget:
    """The host's F64 water ledger of one run, closed against the stored volume at each output
    (the ``mass_balance.csv`` rows of core/run_record.py:107-161).

    The host integrates the nominal rain and each discharge boundary's requested inflow from the
    forcing; the device supplies its cumulative source terms, in the step's order: 0 infiltration,
    1 the positivity floor at non-level cells, 2 the floor at level cells (diagnostic: the level
    overwrites it), 3 the pipe exchange, then five components per boundary (its level cells'
    storage change, face flux, pipe, rain and infiltration) and last the discharge each boundary
    applied after the limiter.
    """

    TAIL = ('bed_clamp_nonlevel_m3', 'bed_clamp_level_m3', 'gate_bed_m3', 'pipe_exchange_m3',
            'initial_storage_m3', 'storage_m3', 'residual_m3', 'relative_residual')

    def __init__(self, boundaries, initial_storage: float):
        self.boundaries = boundaries
        self.initial = float(initial_storage)
        self.rain = np.float64(0)
        self.discharge = np.zeros(len(boundaries), dtype=np.float64)
        self.device_terms = np.zeros(4 + 6 * len(boundaries), dtype=np.float64)
        self.last_row = None
        self.rows = []

    def fields(self):
        inflows = [b for b in self.boundaries if b.kind == 'discharge']
        return (['time', 'elapsed_seconds', 'phase', 'rainfall_m3', 'infiltration_m3'] +
                [f'{b.kind}:{b.name}_m3' for b in self.boundaries] +
                [f'requested_discharge:{b.name}_m3' for b in inflows] + list(self.TAIL))

    def add_forcing(self, dt, rain_rate_volume, boundary_values):
        step = np.float64(dt)
        self.rain += step * rain_rate_volume
        for i, b in enumerate(self.boundaries):
            if b.kind == 'discharge':
                self.discharge[i] += step * boundary_values[i]

    def boundary_exchange(self):
        """Volume each boundary put into the domain so far (negative: taken out): the net of its
        level cells' five components, and the discharge it applied (0 for a level boundary)."""
        nb = len(self.boundaries)
        t = self.device_terms
        level_net = t[4:4 + 5 * nb].reshape(nb, 5).sum(axis=1, dtype=np.float64)
        return level_net, t[4 + 5 * nb:4 + 6 * nb]

    def output(self, time_text, seconds, storage, phase):
        t = self.device_terms
        infiltration, floor, level_floor, pipe = (float(v) for v in t[:4])
        level_net, applied = self.boundary_exchange()
        boundaries_in = applied.sum(dtype=np.float64) + level_net.sum(dtype=np.float64)
        net_input = self.rain - infiltration + floor + pipe + boundaries_in
        residual = float(storage - self.initial - net_input)
        gross_in = (self.rain + np.maximum(applied, 0.).sum() + np.maximum(level_net, 0.).sum() +
                    max(pipe, 0.))
        turnover = (np.abs(applied).sum() + abs(infiltration) + abs(floor) + abs(pipe) +
                    np.abs(level_net).sum())
        scale = max(self.initial, float(gross_in), float(turnover))
        row = dict(time=time_text, elapsed_seconds=seconds, phase=phase, rainfall_m3=float(self.rain),
                   infiltration_m3=infiltration)
        for i, b in enumerate(self.boundaries):
            row[f'{b.kind}:{b.name}_m3'] = float(applied[i] if b.kind == 'discharge' else level_net[i])
        for i, b in enumerate(self.boundaries):
            if b.kind == 'discharge':
                row[f'requested_discharge:{b.name}_m3'] = float(self.discharge[i])
        row.update(bed_clamp_nonlevel_m3=floor, bed_clamp_level_m3=level_floor, gate_bed_m3=0.,
                   pipe_exchange_m3=pipe, initial_storage_m3=self.initial, storage_m3=float(storage),
                   residual_m3=residual, relative_residual=residual / scale if scale else None)
        self.last_row = row
        self.rows.append(row)
        return row


class ExactTerms:
    """Cumulative budget terms held as exact rationals of the device's double-double pairs."""

    def __init__(self, terms: int):
        self.total = [Fraction(0)] * terms

    def add_pairs(self, pairs: np.ndarray) -> None:
        for t, (hi, lo) in enumerate(pairs):
            self.total[t] += Fraction(float(hi)) + Fraction(float(lo))

    def add_values(self, term: int, values: np.ndarray) -> None:
        for v in values[values != 0]:
            self.total[term] += Fraction(float(v))

    def as_float(self) -> np.ndarray:
        return np.array([float(t) for t in self.total], dtype=np.float64)

    def state(self):
        return [[t.numerator, t.denominator] for t in self.total]

    def restore(self, saved):
        self.total = [Fraction(int(n), int(d)) for n, d in saved]


def face_totals(mesh: Mesh, q: np.ndarray) -> np.ndarray:
    """[e_num, 4] F32 sums of q * length over each face, side order as stored (solver_compact.py:548-557)."""
    lengths = mesh.length
    out = np.zeros((mesh.e_num, 4), dtype=np.float32)
    flux = (q * lengths).astype(np.float32)
    for d in range(4):
        count = mesh.face_count[:, d]
        ptr = mesh.face_ptr[:, d]
        acc = np.zeros(mesh.e_num, dtype=np.float32)
        for j in range(int(count.max()) if len(count) else 0):
            has = count > j
            idx = np.where(has, ptr + j, 0)
            add = np.where(has, flux[mesh.isl[idx]] if len(mesh.isl) else 0, np.float32(0))
            acc = np.where(has, (acc + add).astype(np.float32), acc)
        out[:, d] = acc
    return out


def velocities(mesh: Mesh, depth: np.ndarray, q: np.ndarray, min_h: float):
    """u, v of ``velocity`` (solver_compact.py:355-362) in F32."""
    eq = face_totals(mesh, q)
    _, widths, heights, _ = mesh.geometry()
    w = widths.astype(np.float32)
    h_ = heights.astype(np.float32)
    h = np.maximum(depth, np.float32(0.01))
    wet = depth >= np.float32(min_h)
    with np.errstate(divide='ignore', invalid='ignore'):
        u = ((eq[:, 0] + eq[:, 1]) / h_ / h / np.float32(2.)).astype(np.float32)
        v = ((eq[:, 2] + eq[:, 3]) / w / h / np.float32(2.)).astype(np.float32)
    u = np.where(wet, u, np.float32(0.))
    v = np.where(wet, v, np.float32(0.))
    u[0] = v[0] = 0.
    return u, v


@dataclass
class RunSettings:
    afa: float = 0.5
    sita: float = 1.0
    min_h: float = 0.02
    end_time: float = 3600.
    warmup_duration: float = 0.
    yield_step: float = 300.
    batch_size: int = 32
    order: str = 'morton'
    start_text: str = ''          # ISO start time, used to label output rows
    stations: dict = field(default_factory=dict)   # name -> cell id (original numbering)


class Simulation:
    """One scenario advanced by the device step with the reference's host protocol."""

    def __init__(self, mesh: Mesh, land_parameters: dict, forcing: Forcing, settings: RunSettings,
                 checkpoint=None, engine_factory=None, outputs=None, record: dict | None = None,
                 initial_depth: np.ndarray | None = None, base_layout=None, coupling=None, gates=None):
        self.mesh = mesh
        self.forcing = forcing
        self.settings = s = settings
        self.outputs = outputs
        self.record = record if record is not None else {}
        self.land_parameters = land_parameters
        self.boundaries = forcing.boundaries
        self.b_num = len(self.boundaries)
        end_time = float(s.end_time)
        self.end_time = end_time
        meta = checkpoint.meta if checkpoint is not None else {}
        self.continuation = bool(meta.get('run_state', False))
        self.coupling = coupling
        if self.continuation and coupling is not None:
            raise ValueError('A surface checkpoint cannot restore the pipe model state')
        self.origin_before_start = (meta['clock_origin_before_start_s'] if self.continuation
                                    else float(s.warmup_duration))
        initial_clock = (((checkpoint.clock_hi << 32) | checkpoint.clock_lo) / TICKS_PER_SECOND
                         if checkpoint is not None else 0.)
        if checkpoint is not None:
            end_tick = seconds_to_ticks(end_time + self.origin_before_start)
            state_tick = (checkpoint.clock_hi << 32) | checkpoint.clock_lo
            if state_tick > end_tick:
                raise ValueError('Checkpoint time is later than the requested simulation end')
            if self.continuation and state_tick == end_tick:
                self.record['completion_reason'] = 'checkpoint_at_end'
        self.host_now_ticks = seconds_to_ticks(initial_clock)

        cell_forcing = cell_forcing_arrays(mesh, forcing)
        if base_layout is not None:
            # another scenario on the same mesh: reuse its ordering and lists
            self.layout = relayout(base_layout, mesh, land_parameters, cell_forcing)
        else:
            self.layout = build_layout(mesh, land_parameters, cell_forcing, s.sita, s.order)
        self.params = make_params(s.afa, s.sita, s.min_h, s.batch_size)
        _, _, _, area = mesh.geometry()
        self.area = area
        state = initial_state(mesh, self.layout)
        if checkpoint is not None:
            if len(checkpoint.depth) != mesh.e_num or len(checkpoint.q) != mesh.s_num:
                raise ValueError('Checkpoint size differs from mesh')
            for name in ('depth', 'bed', 'q', 'rain_time', 'max_depth', 'max_speed'):
                state[name] = np.asarray(getattr(checkpoint, name), dtype=np.float32).copy()
            state['capacity'] = np.asarray(checkpoint.infiltration, dtype=np.float32).copy()
            state['peak_tick'] = ((np.asarray(checkpoint.first_time_hi, dtype=np.uint64) << np.uint64(32)) |
                                  np.asarray(checkpoint.first_time_lo, dtype=np.uint64))
            self.record['initialization'] = dict(method='checkpoint')
        elif s.warmup_duration > 0:
            depth, info = connected_still_water(state['bed'], np.full(mesh.s_num, -999., dtype=np.float32),
                                                mesh.sbf, mesh.topo, self.boundaries, area)
            state['depth'] = depth
            self.record['initialization'] = dict(method='connected_still_water', **info)
        elif initial_depth is not None:
            depth = np.asarray(initial_depth, dtype=np.float32).copy()
            if depth.shape != (mesh.e_num,) or not np.all(depth >= 0):
                raise ValueError('initial depth must be a non-negative array over the cells')
            depth[0] = 0.
            state['depth'] = depth
            self.record['initialization'] = dict(method='given_depth')
        else:
            self.record['initialization'] = dict(method='dry')
        self.initial_bed = state['bed'].copy()
        initial_storage = float(np.sum(state['depth'].astype(float) * area, dtype=np.float64))
        self.budget = WaterBudget(self.boundaries, initial_storage)
        self.device_units = ExactTerms(4 + 6 * self.b_num)
        if self.continuation:
            saved = meta['budget']
            self.budget.initial = saved['initial_storage_m3']
            self.budget.rain = np.float64(saved['rainfall_m3'])
            self.budget.discharge[:] = saved['requested_discharge_m3']
            self.device_units.restore(saved['device_terms_exact'])
        else:
            # apply_initial_levels (solver_compact.py:412-420, 876-883)
            for bi, b in enumerate(self.boundaries):
                if b.kind == 'level':
                    cells = np.asarray(b.cells)
                    target = np.maximum(b.series.at(0.) - state['bed'][cells].astype(float), 0.).astype(np.float32)
                    previous = state['depth'][cells].copy()
                    state['depth'][cells] = target
                    self.device_units.add_values(4 + 5 * bi, ((target - previous) * area.astype(np.float32)[cells]).astype(np.float32))
        # initial_diagnostics / reset_extrema (solver_compact.py:657-680, 889-892)
        now = self.host_now_ticks
        u, v = velocities(mesh, state['depth'], state['q'], s.min_h)
        speed = np.sqrt(u * u + v * v).astype(np.float32)
        state['max_speed'] = np.maximum(state['max_speed'], speed)
        higher = state['depth'] > state['max_depth']
        state['max_depth'] = np.where(higher, state['depth'], state['max_depth']).astype(np.float32)
        state['peak_tick'] = np.where(higher, np.uint64(now), state['peak_tick']).astype(np.uint64)
        self.elapsed = self.host_now_ticks / TICKS_PER_SECOND - self.origin_before_start
        if self.elapsed == 0. and end_time > 0.:
            state['max_depth'] = state['depth'].copy()
            state['max_speed'] = speed
            state['peak_tick'] = np.full(mesh.e_num, now, dtype=np.uint64)
        state['max_depth'][0] = state['max_speed'][0] = 0.
        if coupling is not None:
            self.engine = (engine_factory or Engine)(self.layout, self.params, state,
                                                     pipe_q=np.zeros(mesh.e_num, np.float32))
        else:
            self.engine = (engine_factory or Engine)(self.layout, self.params, state)
        self.engine.reset(self.host_now_ticks)
        self.gates = gates
        if gates is not None:
            if not hasattr(self.engine, 'set_gates'):
                raise ValueError('gate control needs a single-mesh engine (not a batch slot)')
            # the controller's device lists and state, then check_gates before the first CFL
            # (solver_compact.py:885-889); a continuation restores the controller first (:723-724)
            gates.bind(self.engine, self.layout, mesh.z, s.min_h)
            if self.continuation:
                gates.restore(meta['gate_controller'])
            if gates.initial_storage is None:
                gates.initialize_inventory(state['depth'], area)
            self.engine.gate_check()
        self.engine.compute_cfl()
        self.steps = self.warmup_steps = 0
        self.previous_steps = meta.get('total_steps', 0) if self.continuation else 0
        self.previous_warmup_steps = meta.get('total_warmup_steps', 0) if self.continuation else 0
        self.dt_hasher = hashlib.sha256()
        self.dt_stats = [math.inf, 0., 0.]
        self.interval_dts, self.interval_event_dts = [], []
        self.interval_activity = [0, 0, 1.]
        self.total_activity = [0, 0, 1.]
        self.stats_rows = []
        self.discharge_peaks = np.zeros(self.b_num, dtype=np.float64)
        self.q_exact = np.zeros(self.b_num, dtype=np.float64)
        self.coefficients = np.zeros((max(1, self.b_num), 2), dtype=np.float32)
        self.previous_coefficients = np.full_like(self.coefficients, np.nan)
        self.rain_index = -1
        self.rain_volume_rate = np.float64(0)
        ob = self.origin_before_start
        self.event_samples = {seconds_to_ticks(float(t) + ob): float(t)
                              for t in forcing.events if -ob <= t <= end_time}
        self.event_ticks = np.array(sorted(self.event_samples), dtype=np.uint64)
        el = self.elapsed
        self.next_output = min((math.floor(el / s.yield_step) + 1) * s.yield_step, 0. if el < 0. else end_time)
        self.next_frame = min((math.floor(el / FRAME_INTERVAL_S) + 1) * FRAME_INTERVAL_S, 0. if el < 0. else end_time)
        self.batches = 0
        self.device_seconds = 0.
        self.pipe_end, self.pipe_pending = self.elapsed, False

    # ------------------------------------------------------------------ helpers
    @property
    def host_now_seconds(self) -> float:
        return self.host_now_ticks / TICKS_PER_SECOND

    def upload_forcing(self, seconds: float, warming: bool) -> None:
        """upload_forcing (solver_compact.py:769-790)."""
        sample = 0. if warming else self.event_samples.get(seconds_to_ticks(self.host_now_seconds), seconds)
        for bi, b in enumerate(self.boundaries):
            t = b.series.times
            i = min(int(np.searchsorted(t, sample, side='right')) - 1, len(t) - 2)
            self.coefficients[bi, 0] = b.series.at(sample)
            self.coefficients[bi, 1] = 0. if warming else ((b.series.values[i + 1] - b.series.values[i]) /
                                                           (t[i + 1] - t[i]))
        coeff_changed = not np.array_equal(self.previous_coefficients, self.coefficients)
        if coeff_changed:
            self.previous_coefficients[:] = self.coefficients
        rates = None
        if not warming:
            index = self.forcing.rain.index_at(sample)
            if index != self.rain_index:
                r = self.forcing.rain.rates(index)
                rates = r.astype(np.float32)
                self.rain_volume_rate = np.sum(r[self.forcing.rain.pixel_index[1:]] * self.area[1:], dtype=np.float64)
                self.rain_index = index
        self.engine.set_forcing(self.coefficients[:max(1, self.b_num)].ravel() if self.b_num else np.zeros(0, np.float32),
                                rates, not warming)

    def flush_budget(self) -> np.ndarray:
        self.device_units.add_pairs(self.engine.flush_budget())
        return self.device_units.as_float()

    # ------------------------------------------------------------------ outputs
    def write_output(self, seconds: float, phase: str, frame=False, final=False):
        self.budget.device_terms = self.flush_budget()
        depth = self.engine.read_state('depth')
        storage = float(np.sum(depth.astype(float) * self.area, dtype=np.float64))
        row = self.budget.output(self.time_text(seconds), seconds, storage, phase)
        if self.gates is not None:
            self.record['gate_control'] = self.gates.report(depth, self.area, seconds, self.mesh.z)
        limits, ids = self.engine.limit_capture(reset=True)
        regular, short = np.array(self.interval_dts), np.array(self.interval_event_dts)
        stats = dict(time=self.time_text(seconds), elapsed_seconds=seconds,
                     steps=len(regular) + len(short), standard_steps=len(regular),
                     standard_dt_min_s=float(regular.min()) if len(regular) else '',
                     standard_dt_max_s=float(regular.max()) if len(regular) else '',
                     standard_dt_mean_s=float(regular.mean()) if len(regular) else '',
                     event_short_steps=len(short),
                     event_dt_min_s=float(short.min()) if len(short) else '',
                     event_dt_sum_s=float(short.sum()) if len(short) else 0.,
                     cfl_min_s=float(limits[0]),
                     limited_cell_steps=self.interval_activity[0], limited_face_steps=self.interval_activity[1],
                     minimum_donor_scale=self.interval_activity[2],
                     limiting_edge=int(ids[0]), lower_cell=int(ids[1]), higher_cell=int(ids[2]),
                     lower_depth_m=float(limits[1]), higher_depth_m=float(limits[2]),
                     face_depth_m=float(limits[3]), face_velocity_m_s=float(limits[4]))
        self.stats_rows.append(stats)
        if self.outputs is not None:
            self.outputs.emit(self, seconds, phase, frame, final, depth, row, stats)
        self.interval_dts.clear()
        self.interval_event_dts.clear()
        self.interval_activity[:] = [0, 0, 1.]
        return row

    def time_text(self, seconds: float) -> str:
        if not self.settings.start_text:
            return ''
        from datetime import datetime, timedelta
        return (datetime.fromisoformat(self.settings.start_text) + timedelta(seconds=seconds)).isoformat()

    def reset_extrema(self) -> None:
        """reset_extrema (solver_compact.py:674-679) from the current state."""
        depth = self.engine.read_state('depth')
        q = self.engine.read_state('q')
        u, v = velocities(self.mesh, depth, q, self.settings.min_h)
        speed = np.sqrt(u * u + v * v).astype(np.float32)
        speed[0] = 0.
        md = depth.copy()
        md[0] = 0.
        self.engine.write_state('max_depth', md)
        self.engine.write_state('max_speed', speed)
        peak = np.full(self.mesh.e_num, self.host_now_ticks, dtype=np.uint64)
        self.engine.write_state('peak_tick', peak)

    def supported_dt(self):
        ndt, si = self.engine.cfl_state()
        if ndt < MIN_DT:
            cells = self.mesh.topo[si, 1:]
            depth = self.engine.read_state('depth')
            bed = self.engine.read_state('bed')
            q = float(self.engine.read_state('q')[si])
            head = max(bed[cells[0]] + depth[cells[0]], bed[cells[1]] + depth[cells[1]])
            h = float(max(np.float32(head) - max(bed[cells[0]], bed[cells[1]], np.float32(-999.)), np.float32(0.01)))
            raise RuntimeError(f'CFL {ndt:.12g}s below supported minimum {MIN_DT}s; '
                               f'edge={si}, cells={cells.tolist()}, depth_m={depth[cells].tolist()}, '
                               f'face_depth_m={h}, unit_discharge_m2_s={q}, velocity_m_s={q / h if h else 0.}')
        return ndt

    # ------------------------------------------------------------------ main loop
    # The host loop of solver_compact.py:930-985, split at the device call so that a batch runner
    # can advance many scenarios with one launch sequence: start(); while plan(): consume(rows);
    # finish(). run() is that loop for a scenario on its own.
    def run(self):
        self.start()
        while self.plan():
            t0 = perf_counter()
            rows = self.engine.run_batch(self.settings.batch_size)
            self.device_seconds += perf_counter() - t0
            self.consume(rows)
        return self.finish()

    def start(self) -> None:
        """Write the initial output row and frame."""
        elapsed, end_time = self.elapsed, self.end_time
        self.write_output(elapsed, 'warmup' if elapsed < 0 or end_time == 0 else 'simulation',
                          frame=True, final=elapsed == end_time)

    def plan(self) -> bool:
        """Choose the next target, upload forcing and begin a batch; False when the run is done."""
        s = self.settings
        elapsed, end_time, ob = self.elapsed, self.end_time, self.origin_before_start
        if not elapsed < end_time:
            return False
        warming = elapsed < 0.
        if self.coupling is not None and elapsed == self.pipe_end:
            # coupling window boundary (solver_compact.py:921-929)
            if self.pipe_pending:
                self.coupling.end_window(self.engine.take_pipe_volume().astype(np.float64))
            window_end = min(elapsed + self.coupling.interval, 0. if warming else end_time)
            source = self.coupling.begin_window(elapsed, window_end - elapsed, self.engine.read_state('depth'),
                                                self.engine.read_state('bed'), self.area, self.mesh.land)
            self.engine.set_pipe_q(source)
            self.pipe_pending, self.pipe_end = True, window_end
        target = min(self.next_output, self.next_frame, 0. if warming else end_time)
        if not warming:
            now_tick = seconds_to_ticks(self.host_now_seconds)
            i = np.searchsorted(self.event_ticks, np.uint64(now_tick), side='right')
            if i < len(self.event_ticks):
                target = min(target, int(self.event_ticks[i]) / TICKS_PER_SECOND - ob)
        if self.coupling is not None:
            target = min(target, self.pipe_end)
        self.upload_forcing(elapsed, warming)
        target_ticks = seconds_to_ticks(target + ob)
        if target_ticks < self.host_now_ticks:
            raise ValueError('Target precedes the current clock')
        self.engine.begin(target_ticks, self.host_now_ticks)
        self._warming = warming
        return True

    def consume(self, rows: np.ndarray) -> None:
        """Read one batch's trace rows (clock, forcing integration, statistics) and write outputs.

        The rows are handled as arrays, with every F64 running total accumulated in step order
        (np.add.accumulate is sequential), so the totals equal those of a step-by-step loop."""
        s = self.settings
        end_time, ob = self.end_time, self.origin_before_start
        warming = self._warming
        elapsed = self.elapsed
        self.batches += 1
        status = rows[:, 3].astype(np.int64)
        stop = np.flatnonzero(status != abi.COMPLETED)
        n = int(stop[0]) if len(stop) else len(rows)
        if n < len(rows) and int(status[n]) != abi.DONE:
            # the steps before the failure still count, as in a step-by-step loop
            self.elapsed = self._consume_rows(rows[:n], warming, elapsed)
            if int(status[n]) == abi.BAD_CFL:
                self.supported_dt()
            raise RuntimeError(f'Device clock status {int(status[n])} at {self.elapsed}s')
        elapsed = self._consume_rows(rows[:n], warming, elapsed)
        self.elapsed = elapsed
        is_output = abs(elapsed - self.next_output) < 1e-9 or elapsed == end_time
        is_frame = abs(elapsed - self.next_frame) < 1e-9 or elapsed == end_time
        if is_output or is_frame:
            if warming and elapsed == 0. and end_time > 0.:
                self.reset_extrema()
            self.write_output(elapsed, 'warmup' if elapsed < 0 or end_time == 0 else 'simulation',
                              frame=is_frame, final=elapsed == end_time)
            if is_output:
                self.next_output = min(self.next_output + s.yield_step, 0. if elapsed < 0 else end_time)
            if is_frame:
                self.next_frame = min(self.next_frame + FRAME_INTERVAL_S, 0. if elapsed < 0 else end_time)

    def _consume_rows(self, r: np.ndarray, warming: bool, elapsed: float) -> float:
        """Account completed trace rows `r`; returns the elapsed seconds after the last one."""
        n = len(r)
        if n == 0:
            return elapsed
        ob = self.origin_before_start
        words = np.ascontiguousarray(r[:, 0], dtype=np.uint32)
        dt = words.view(np.float32).astype(np.float64)
        ends = (r[:, 1].astype(np.uint64) << np.uint64(32)) | r[:, 2].astype(np.uint64)
        self.host_now_ticks = int(ends[-1])
        scales = np.ascontiguousarray(r[:, 8], dtype=np.uint32).view(np.float32)
        for activity in (self.interval_activity, self.total_activity):
            activity[0] += int(r[:, 6].astype(np.int64).sum())
            activity[1] += int(r[:, 7].astype(np.int64).sum())
            activity[2] = min(activity[2], float(scales.min()))
        after = ends.astype(np.float64) / TICKS_PER_SECOND - ob
        before = np.concatenate([[elapsed], after[:-1]])
        mid = np.zeros(n) if warming else before + dt / 2.
        rain_rate = 0. if warming else self.rain_volume_rate
        b = self.budget
        b.rain = np.add.accumulate(np.concatenate([[b.rain], dt * rain_rate]))[-1]
        for bi, bd in enumerate(self.boundaries):
            values = np.interp(mid, bd.series.times, bd.series.values)
            if bd.kind == 'discharge':
                b.discharge[bi] = np.add.accumulate(np.concatenate([[b.discharge[bi]], dt * values]))[-1]
            self.discharge_peaks[bi] = max(self.discharge_peaks[bi], float(values.max()))
            self.q_exact[bi] = values[-1]
        self.steps += n
        self.warmup_steps += n if warming else 0
        self.dt_stats = [min(self.dt_stats[0], float(dt.min())), max(self.dt_stats[1], float(dt.max())),
                         float(np.add.accumulate(np.concatenate([[self.dt_stats[2]], dt]))[-1])]
        short = r[:, 5] != 0
        self.interval_event_dts.extend(dt[short].tolist())
        self.interval_dts.extend(dt[~short].tolist())
        self.dt_hasher.update(words.tobytes())
        return float(after[-1])

    def finish(self):
        """Record the run's totals and close the outputs; returns the last mass-balance row."""
        if self.coupling is not None:
            if self.pipe_pending:   # solver_compact.py:986-988
                self.coupling.end_window(self.engine.take_pipe_volume().astype(np.float64))
                self.pipe_pending = False
            self.record['coupling'] = self.coupling.metadata()
        self.record.update(
            steps=self.steps, warmup_steps=self.warmup_steps, mass_balance_final=self.budget.last_row,
            dt_seconds={'min': self.dt_stats[0] if self.steps else 0., 'max': self.dt_stats[1],
                        'sum': self.dt_stats[2]},
            dt_sequence_sha256=self.dt_hasher.hexdigest(),
            limiter_activity=dict(limited_cell_steps=self.total_activity[0],
                                  limited_face_steps=self.total_activity[1],
                                  minimum_donor_scale=self.total_activity[2]))
        if self.outputs is not None:
            self.outputs.finish(self)
        return self.budget.last_row

    # ------------------------------------------------------------------ state access
    def state(self) -> dict:
        names = ('depth', 'bed', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick', 'q')
        return {n: self.engine.read_state(n) for n in names}
