"""Scenario batches on one shared mesh: many flood solves advanced by the same launches.

:class:`BatchEngine` binds ``nhb_*`` (include/nhflood.h): up to 64 slots on one mesh, each with its
own state, forcing, clock and dt, stored [entity][slot] on the device. :meth:`BatchEngine.view`
hands a :class:`Simulation` an engine for one slot with the interface of
:class:`~nh_flood_2d_amd.engine.Engine`, so every scenario runs the unchanged host protocol of
``solver_compact.py:930-985``; :class:`BatchRunner` interleaves those protocols so that one
``nhb_run_batch`` advances every scenario that has a target, and admits queued scenarios into
slots as earlier ones finish (continuous batching).

The reference runs one scenario per process; batching has no counterpart there. A scenario gives
the same depth, discharge, dt sequence and outputs alone and in a batch (the device arithmetic per
slot is the single-scenario step's); only the FP64 budget sums group terms differently.
"""
from __future__ import annotations

import ctypes
from collections import deque
from time import perf_counter

import numpy as np

from . import abi
from .engine import SETTLED, STATE_CELLS, STATE_SIDES, TorchDevice, check_extrema, uniform_rain
from .layout import Layout, meta_boundary

# per-slot cell state of a batch; the bed is shared
SLOT_CELLS = tuple(n for n in STATE_CELLS if n != 'bed')
# arrays the device may hold deferred (extrema of rising cells, the uniform rain state)


def lanes_for(count: int) -> int:
    """Slots of a batch holding `count` scenarios: the next power of two, at most 64."""
    if not 1 <= count <= abi.MAX_SLOTS:
        raise ValueError(f'a batch holds 1..{abi.MAX_SLOTS} scenarios, not {count}')
    n = 1
    while n < count:
        n *= 2
    return n


def boundary_cells(layout: Layout):
    """Boundary-cell list of a layout (device order): per-cell index or -1, and per-boundary ranges."""
    bid = meta_boundary(layout.arrays['cell_meta'])
    e_num = layout.e_num
    cells = np.nonzero(bid[1:] >= 0)[0] + 1
    order = np.lexsort((cells, bid[cells]))
    cells = cells[order]
    bslot = np.full(e_num, -1, dtype=np.int32)
    bslot[cells] = np.arange(len(cells), dtype=np.int32)
    start = np.searchsorted(bid[cells], np.arange(layout.b_num + 1)).astype(np.int32)
    return bslot, start, len(cells)


class BatchEngine:
    """Device state of up to 64 scenarios sharing one mesh, and the nhb_* context over it."""

    def __init__(self, slots: int, device=None, shared_manning: bool = False):
        self.slots = lanes_for(slots)
        self.device = device if device is not None else TorchDevice()
        self.lib = self.device.lib
        if not hasattr(self.lib, 'nhb_create'):
            raise RuntimeError('the loaded step library has no batch entry points')
        self.shared_manning = shared_manning
        self.layout = None
        self.params = None
        self.ctx = None
        self.owner = [None] * self.slots
        self.batches = 0

    # ------------------------------------------------------------------ setup
    def _create(self, layout: Layout, params: abi.Params, bed: np.ndarray) -> None:
        torch = self.device.torch
        dev, S = self.device, self.slots
        self.layout, self.params = layout, params
        e_num, s_num = layout.e_num, layout.s_num
        self.static = {}
        for name, values in layout.arrays.items():
            if not name.startswith('_'):
                self.static[name] = dev.array(values)
        # the bed is part of the shared mesh in a batch (one array for every slot)
        self.bed = np.asarray(bed, dtype=np.float32).copy()
        self.static['bed'] = dev.array(layout.cells_to_device(self.bed))
        self.state = {}
        with torch.cuda.stream(dev.stream):
            for name in SLOT_CELLS:
                dtype = torch.int64 if name == 'peak_tick' else torch.float32
                self.state[name] = torch.zeros(e_num * S, dtype=dtype, device=dev.device)
            self.state['q'] = torch.zeros(s_num * S, dtype=torch.float32, device=dev.device)
            # per scenario: Manning n by land class, from which the kernels rebuild each side's n
            self.class_manning = None if self.shared_manning else torch.zeros(256 * S, dtype=torch.float32,
                                                                              device=dev.device)
        bslot, bstart, n_bcells = boundary_cells(layout)
        self.static['_bslot'] = dev.array(bslot)
        self.static['_bcell_start'] = dev.array(bstart)
        m = abi.MeshStruct()
        m.e_num, m.s_num = e_num, s_num
        m.r_len = int(layout.arrays['_r_len'][0])
        m.b_num, m.n_pixels = layout.b_num, layout.n_pixels
        m.n_grad, m.n_hang, m.flags = layout.n_grad, layout.n_hang, 0
        m.size_base = layout.size_base
        for name, kind in abi.MESH_FIELDS:
            if kind == 'p':
                buf = self.static.get(name)
                setattr(m, name, dev.ptr(buf) if buf is not None else None)
        st = abi.BatchState()
        for name in ('depth', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick', 'q'):
            setattr(st, name, dev.ptr(self.state[name]))
        st.manning = None
        st.class_manning = dev.ptr(self.class_manning) if self.class_manning is not None else None
        st.bslot = dev.ptr(self.static['_bslot'])
        st.bcell_start = dev.ptr(self.static['_bcell_start'])
        st.n_bcells = n_bcells
        self.mesh_struct, self.state_struct = m, st
        nbytes = self.lib.nhb_workspace_bytes(ctypes.byref(m), ctypes.byref(st), ctypes.byref(params), S)
        if not nbytes:
            raise RuntimeError('nhb_workspace_bytes rejected the batch')
        self.workspace = dev.empty(nbytes)
        handle = ctypes.c_void_p()
        rc = self.lib.nhb_create(ctypes.byref(handle), ctypes.byref(m), ctypes.byref(st), ctypes.byref(params), S,
                                 dev.ptr(self.workspace), nbytes, dev.stream_handle())
        abi.check(self.lib, rc, 'nhb_create')
        self.ctx = handle
        self.terms = 4 + 6 * layout.b_num
        self.cell_class = (layout.arrays['cell_meta'] & 255).astype(np.intp)
        self._pending = 0

    def _check_compatible(self, layout: Layout, params: abi.Params) -> None:
        if (layout.e_num, layout.s_num) != (self.layout.e_num, self.layout.s_num):
            raise ValueError('scenario mesh differs from the batch mesh')
        for name, values in self.layout.arrays.items():
            if name in ('manning', '_class_manning') and not self.shared_manning:
                continue
            other = layout.arrays.get(name)
            if other is None or not np.array_equal(other, values):
                raise ValueError(f'scenario differs from the batch mesh in {name!r}')
        for f, _ in abi.Params._fields_:
            if getattr(params, f) != getattr(self.params, f) and f != 'reserved':
                raise ValueError(f'scenario parameter {f} differs from the batch')

    def free_slots(self):
        return [s for s in range(self.slots) if self.owner[s] is None]

    def view(self, slot: int | None = None):
        """An engine factory for Simulation: binds the next free slot (or `slot`)."""
        def factory(layout, params, state, pipe_q=None):
            if pipe_q is not None:
                raise NotImplementedError('pipe exchange is not supported in a scenario batch')
            if self.ctx is None:
                self._create(layout, params, state['bed'])
            else:
                self._check_compatible(layout, params)
                if not np.array_equal(np.asarray(state['bed'], dtype=np.float32), self.bed):
                    raise ValueError('scenario bed differs from the batch mesh')
            s = slot if slot is not None else (self.free_slots() or [None])[0]
            if s is None:
                raise RuntimeError('no free slot in the batch')
            if self.owner[s] is not None:
                raise RuntimeError(f'slot {s} is in use')
            v = ScenarioView(self, s)
            self.owner[s] = v
            for name in SLOT_CELLS + STATE_SIDES:
                v.write_state(name, state[name])
            if self.class_manning is not None:
                v._write_column(self.class_manning,
                                np.ascontiguousarray(layout.arrays['_class_manning'], dtype=np.float32))
            return v
        return factory

    def release(self, slot: int) -> None:
        self.owner[slot] = None

    # ------------------------------------------------------------------ batch protocol
    def run_batch(self, steps: int) -> np.ndarray:
        """Advance every slot with a begun target; returns the [slots, steps, 9] trace."""
        buf = np.zeros((self.slots, steps, abi.TRACE_COLS), dtype=np.uint32)
        rc = self.lib.nhb_run_batch(self.ctx, int(steps), buf.ctypes.data)
        abi.check(self.lib, rc, 'nhb_run_batch')
        self.batches += 1
        return buf

    def launch_batch(self, steps: int) -> None:
        """Enqueue one batch for every slot with a begun target and return at once."""
        abi.check(self.lib, self.lib.nhb_launch_batch(self.ctx, int(steps)), 'nhb_launch_batch')
        self._pending = int(steps)
        marker = getattr(self.device, 'marker', None)
        self._done = marker() if marker is not None else None

    def ready(self) -> bool:
        """True when the launched batch and its trace copy have finished."""
        done = getattr(self, '_done', None)
        return done is None or done.query()

    def wait_batch(self) -> np.ndarray:
        """Wait for the launched batch; returns the [slots, steps, 9] trace."""
        buf = np.zeros((self.slots, self._pending, abi.TRACE_COLS), dtype=np.uint32)
        abi.check(self.lib, self.lib.nhb_wait_batch(self.ctx, buf.ctypes.data), 'nhb_wait_batch')
        self.batches += 1
        return buf

    def compute_cfl(self) -> None:
        abi.check(self.lib, self.lib.nhb_compute_cfl(self.ctx), 'nhb_compute_cfl')

    def set_graph(self, enable: bool) -> None:
        self.lib.nhb_set_graph(self.ctx, int(enable))

    def set_timing(self, enable: bool) -> None:
        self.lib.nhb_set_timing(self.ctx, int(enable))

    def set_tuning(self, depth_blocks: int, cfl_blocks: int) -> None:
        rc = self.lib.nhb_set_tuning(self.ctx, int(depth_blocks), int(cfl_blocks))
        abi.check(self.lib, rc, 'nhb_set_tuning')

    def tuning_range(self) -> dict:
        out = np.zeros(4, dtype=np.int32)
        self.lib.nhb_tuning_range(self.ctx, out.ctypes.data)
        return dict(depth_blocks=int(out[0]), depth_max=int(out[1]), cfl_blocks=int(out[2]), cfl_max=int(out[3]))

    def kernel_times(self, reset: bool = True):
        ms = np.zeros(8, dtype=np.float64)
        count = np.zeros(8, dtype=np.int64)
        self.lib.nhb_kernel_times(self.ctx, ms.ctypes.data, count.ctypes.data, 8, int(reset))
        return ms, count

    def synchronize(self) -> None:
        abi.check(self.lib, self.lib.nhb_sync(self.ctx), 'nhb_sync')

    def settle(self, slot: int = -1) -> None:
        """Write out a slot's deferred extrema and uniform rain state (every slot for -1)."""
        abi.check(self.lib, self.lib.nhb_settle_state(self.ctx, int(slot)), 'nhb_settle_state')

    def close(self) -> None:
        if self.ctx:
            self.lib.nhb_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScenarioView:
    """One slot of a BatchEngine with the Engine interface used by Simulation."""

    def __init__(self, batch: BatchEngine, slot: int):
        self.batch = batch
        self.slot = slot
        self.layout = batch.layout
        self.lib = batch.lib
        self.terms = batch.terms

    # ------------------------------------------------------------------ state in original order
    def _column(self, buf):
        n = buf.numel() // self.batch.slots
        return buf.view(n, self.batch.slots)[:, self.slot]

    def _write_column(self, buf, values: np.ndarray) -> None:
        dev = self.batch.device
        torch = dev.torch
        host = values.view(np.int64) if values.dtype == np.uint64 else values
        with torch.cuda.stream(dev.stream):
            self._column(buf).copy_(torch.from_numpy(np.ascontiguousarray(host)).to(dev.device))
        dev.stream.synchronize()

    def read_state(self, name: str) -> np.ndarray:
        if name == 'bed':
            return self.batch.bed.copy()
        if name in SETTLED:
            self.batch.settle(self.slot)
        dev = self.batch.device
        dev.stream.synchronize()
        with dev.torch.cuda.stream(dev.stream):
            values = self._column(self.batch.state[name]).contiguous().cpu().numpy()
        if name == 'peak_tick':
            values = values.view(np.uint64)
        if name in STATE_SIDES:
            return self.layout.sides_from_device(values)
        return self.layout.cells_from_device(values)

    def write_state(self, name: str, values: np.ndarray) -> None:
        dtype = np.uint64 if name == 'peak_tick' else np.float32
        values = np.asarray(values, dtype=dtype)
        if name == 'bed':
            if not np.array_equal(values, self.batch.bed):
                raise ValueError('the bed is shared by every scenario of a batch')
            return
        if name in SETTLED or name == 'depth':
            self.batch.settle(self.slot)   # a marked cell's maximum is its depth: keep the pair final
        check_extrema(name, values)
        values = self.layout.sides_to_device(values) if name in STATE_SIDES else self.layout.cells_to_device(values)
        self._write_column(self.batch.state[name], values)
        if name in ('rain_time', 'capacity'):
            self._rain_state()

    def _rain_state(self) -> None:
        """The slot's rain state after its rain time or capacity was written: uniform (one rain
        time, one capacity per land class, one rain pixel) when the arrays allow it, else per cell
        (nhb_set_rain_state). Values are compared bit for bit."""
        b = self.batch
        rt = self._column(b.state['rain_time']).contiguous().cpu().numpy()[1:]
        cap = self._column(b.state['capacity']).contiguous().cpu().numpy()[1:]
        uniform, rt0, table = uniform_rain(rt, cap, b.cell_class[1:], b.layout.n_pixels)
        uniform = uniform and b.layout.s_num > 1   # the flux pass advances the uniform state
        rc = self.lib.nhb_set_rain_state(b.ctx, self.slot, int(uniform), ctypes.c_float(rt0),
                                         table.ctypes.data if uniform else None)
        abi.check(self.lib, rc, 'nhb_set_rain_state')

    # ------------------------------------------------------------------ per-slot protocol
    def reset(self, now_ticks: int) -> None:
        abi.check(self.lib, self.lib.nhb_reset(self.batch.ctx, self.slot, int(now_ticks)), 'nhb_reset')

    def set_forcing(self, coefficients: np.ndarray, rain_rates: np.ndarray | None, raining: bool):
        coeff = np.ascontiguousarray(coefficients, dtype=np.float32)
        rates = None if rain_rates is None else np.ascontiguousarray(rain_rates, dtype=np.float32)
        rc = self.lib.nhb_set_forcing(self.batch.ctx, self.slot, coeff.ctypes.data if coeff.size else None,
                                      rates.ctypes.data if rates is not None else None, int(raining))
        abi.check(self.lib, rc, 'nhb_set_forcing')

    def begin(self, target_ticks: int, origin_ticks: int) -> None:
        rc = self.lib.nhb_begin(self.batch.ctx, self.slot, int(target_ticks), int(origin_ticks))
        abi.check(self.lib, rc, 'nhb_begin')

    def run_batch(self, steps: int) -> np.ndarray:
        """Run one batch of the whole engine (every slot with a target) and return this slot's rows."""
        return self.batch.run_batch(steps)[self.slot].copy()

    def compute_cfl(self) -> None:
        self.batch.compute_cfl()

    def cfl_state(self):
        ndt = ctypes.c_float()
        side = ctypes.c_int32()
        rc = self.lib.nhb_cfl_state(self.batch.ctx, self.slot, ctypes.byref(ndt), ctypes.byref(side))
        abi.check(self.lib, rc, 'nhb_cfl_state')
        return float(ndt.value), int(self.layout.side_perm[side.value])

    def flush_budget(self) -> np.ndarray:
        out = np.zeros(2 * self.terms, dtype=np.float64)
        rc = self.lib.nhb_flush_budget(self.batch.ctx, self.slot, out.ctypes.data)
        if rc == 1:
            raise FloatingPointError('Nonfinite F32 device budget diagnostic')
        abi.check(self.lib, rc, 'nhb_flush_budget')
        return out.reshape(self.terms, 2)

    def limit_capture(self, reset: bool = False):
        values = np.zeros(5, dtype=np.float32)
        ids = np.zeros(3, dtype=np.int32)
        rc = self.lib.nhb_limit_capture(self.batch.ctx, self.slot, values.ctypes.data, ids.ctypes.data, int(reset))
        abi.check(self.lib, rc, 'nhb_limit_capture')
        return values, np.array([self.layout.side_perm[ids[0]], self.layout.cell_perm[ids[1]],
                                 self.layout.cell_perm[ids[2]]], dtype=np.int64)

    def set_graph(self, enable: bool) -> None:
        self.batch.set_graph(enable)

    def set_timing(self, enable: bool) -> None:
        self.batch.set_timing(enable)

    def synchronize(self) -> None:
        self.batch.synchronize()

    def close(self) -> None:
        self.batch.release(self.slot)


class BatchRunner:
    """Continuous batching of Simulations that share one BatchEngine.

    `make` is a callable (engine_factory) -> Simulation that builds one scenario; scenarios are
    admitted in submission order while slots are free, and a slot is reused as soon as its
    scenario has written its last output. Each loop iteration plans every live scenario, runs one
    batch for all of them and hands each its own trace rows.
    """

    def __init__(self, batch: BatchEngine, batch_size: int = 32):
        self.batch = batch
        self.batch_size = batch_size
        self.queue = deque()
        self.live = {}
        self.results = {}
        self.latency = {}
        self.admitted = {}

    def submit(self, key, make, arrival: float | None = None) -> None:
        self.queue.append((key, make, perf_counter() if arrival is None else arrival))

    def _admit(self) -> None:
        while self.queue and self.batch.free_slots():
            key, make, arrival = self.queue.popleft()
            sim = make(self.batch.view())
            if sim.settings.batch_size != self.batch_size:
                raise ValueError('every scenario of a batch uses the same batch size')
            sim.start()
            self.live[key] = (sim, arrival)
            self.admitted[key] = perf_counter()

    def step(self) -> bool:
        """One batch for every live scenario; False when nothing is left to run."""
        self._admit()
        planned = []
        for key, (sim, arrival) in list(self.live.items()):
            if sim.plan():
                planned.append(sim)
            else:
                self.results[key] = sim.finish()
                self.latency[key] = perf_counter() - arrival
                sim.engine.close()
                del self.live[key]
        if not planned:
            return bool(self.queue)
        t0 = perf_counter()
        trace = self.batch.run_batch(self.batch_size)
        dt = perf_counter() - t0
        for sim in planned:
            sim.device_seconds += dt
            sim.consume(trace[sim.engine.slot])
        return True

    def run(self) -> dict:
        while self.step():
            pass
        return self.results
