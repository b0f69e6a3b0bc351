"""One mesh bound to a library implementing include/nhflood.h.

:class:`Engine` owns the device arrays of one scenario (torch tensors on the GPU), passes their
addresses to ``nhf_create`` and exposes the batch protocol of the reference's host loop
(``solver_compact.py:930-970``): set forcing, begin a batch, run it, read the trace. All state
exchanged with the caller is in the reference's numbering; the Morton permutation stays inside.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .layout import Layout

STATE_CELLS = ('depth', 'bed', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick')
STATE_SIDES = ('q',)
# arrays the step may hold back on the device until nhf_settle_extrema writes them out
SETTLED = ('max_depth', 'peak_tick', 'rain_time', 'capacity')


def uniform_rain(rain_time: np.ndarray, capacity: np.ndarray, classes: np.ndarray, n_pixels: int):
    """Whether cells with these rain times, capacities and land class codes (the real cells of one
    mesh, any order) can keep the uniform rain state: at most one rain pixel, one rain time, one
    capacity per class, compared bit for bit. Returns (uniform, rain time, capacity table [256])."""
    rt = np.asarray(rain_time, np.float32)
    cap = np.asarray(capacity, np.float32)
    cls = np.asarray(classes, np.int64)
    table = np.zeros(256, dtype=np.float32)
    table[cls] = cap
    uniform = (n_pixels <= 1 and rt.size > 0
               and bool(np.all(rt.view(np.uint32) == rt.view(np.uint32)[0]))
               and np.array_equal(table[cls].view(np.uint32), cap.view(np.uint32)))
    return uniform, (float(rt[0]) if uniform else 0.0), table


def check_extrema(name: str, values) -> None:
    """The step keeps max_depth >= 0 and reads a negative value as the mark of a cell whose depth
    is rising (its maximum is its depth), so a negative max_depth from the host is rejected."""
    if name == 'max_depth' and np.any(np.asarray(values, dtype=np.float32) < 0):
        raise ValueError(f'{name}: maxima must not be negative')


class TorchDevice:
    """Device memory and stream for the HIP library (torch is plumbing only)."""

    def __init__(self, device: str = 'cuda:0'):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('no GPU visible: the HIP step needs a ROCm device')
        self.torch = torch
        self.device = torch.device(device)
        self.stream = torch.cuda.Stream(device=self.device)
        self.lib = abi.load_hip_library()
        self._dtypes = {np.dtype(np.float32): torch.float32, np.dtype(np.int32): torch.int32,
                        np.dtype(np.uint32): torch.int32, np.dtype(np.uint8): torch.uint8,
                        np.dtype(np.uint64): torch.int64, np.dtype(np.int64): torch.int64,
                        np.dtype(np.float64): torch.float64}

    def array(self, values: np.ndarray):
        values = np.ascontiguousarray(values)
        host = values.view(np.int32) if values.dtype == np.uint32 else (
            values.view(np.int64) if values.dtype == np.uint64 else values)
        with self.torch.cuda.stream(self.stream):
            t = self.torch.from_numpy(host.copy()).to(self.device, non_blocking=False)
        return t

    def empty(self, nbytes: int):
        with self.torch.cuda.stream(self.stream):
            return self.torch.zeros(max(nbytes, 1), dtype=self.torch.uint8, device=self.device)

    def read(self, buf, dtype) -> np.ndarray:
        self.stream.synchronize()
        out = buf.cpu().numpy()
        return out.view(dtype)

    def write(self, buf, values: np.ndarray) -> None:
        values = np.ascontiguousarray(values)
        host = values.view(np.int32) if values.dtype == np.uint32 else (
            values.view(np.int64) if values.dtype == np.uint64 else values)
        with self.torch.cuda.stream(self.stream):
            buf.copy_(self.torch.from_numpy(host.copy()).to(self.device))
        self.stream.synchronize()

    def _pinned(self, buf):
        pin = self._pins.get(id(buf)) if hasattr(self, '_pins') else None
        if pin is None or pin.numel() != buf.numel() or pin.dtype != buf.dtype:
            self._pins = getattr(self, '_pins', {})
            pin = self.torch.empty(buf.numel(), dtype=buf.dtype, pin_memory=True)
            self._pins[id(buf)] = pin
        return pin

    def upload(self, buf, values: np.ndarray) -> None:
        """Host -> device through a pinned staging buffer kept per device array."""
        pin = self._pinned(buf)
        pin.numpy()[:] = np.ascontiguousarray(values).view(pin.numpy().dtype)
        with self.torch.cuda.stream(self.stream):
            buf.copy_(pin, non_blocking=True)
        self.stream.synchronize()   # the staging buffer is reused at the next upload

    def download(self, buf, dtype) -> np.ndarray:
        """Device -> host through the pinned staging buffer."""
        pin = self._pinned(buf)
        with self.torch.cuda.stream(self.stream):
            pin.copy_(buf, non_blocking=True)
        self.stream.synchronize()
        return pin.numpy().copy().view(dtype)

    def zero(self, buf) -> None:
        with self.torch.cuda.stream(self.stream):
            buf.zero_()

    @staticmethod
    def ptr(buf) -> int:
        return buf.data_ptr()

    def stream_handle(self) -> int:
        return self.stream.cuda_stream

    def synchronize(self) -> None:
        self.stream.synchronize()

    def marker(self):
        """An event recorded now on this device's stream: its query() turns true once everything
        enqueued before it has run."""
        event = self.torch.cuda.Event()
        event.record(self.stream)
        return event


class Engine:
    """A mesh, its state and the step library (``nhf_*`` entry points)."""

    def __init__(self, layout: Layout, params: abi.Params, state: dict, device=None,
                 pipe_q: np.ndarray | None = None):
        self.layout = layout
        self.device = device if device is not None else TorchDevice()
        self.lib = self.device.lib
        self.params = params
        self.buffers = {}
        dev = self.device
        arr = layout.arrays
        e_num, s_num = layout.e_num, layout.s_num
        # state (device order)
        for name in STATE_CELLS:
            check_extrema(name, state[name])
            self.buffers[name] = dev.array(layout.cells_to_device(state[name]))
        self.buffers['q'] = dev.array(layout.sides_to_device(state['q']))
        for name, values in arr.items():
            if not name.startswith('_'):
                self.buffers[name] = dev.array(values)
        if pipe_q is not None:
            self.buffers['pipe_q'] = dev.array(layout.cells_to_device(pipe_q.astype(np.float32)))
            self.buffers['pipe_volume'] = dev.array(np.zeros(e_num, np.float32))
            self.buffers['pipe_correction'] = dev.array(np.zeros(e_num, np.float32))
        m = abi.MeshStruct()
        m.e_num, m.s_num = e_num, s_num
        m.r_len = int(arr['_r_len'][0])
        m.b_num = layout.b_num
        m.n_pixels = layout.n_pixels
        m.n_grad, m.n_hang = layout.n_grad, layout.n_hang
        m.flags = 1 if pipe_q is not None else 0
        m.size_base = layout.size_base
        for name, kind in abi.MESH_FIELDS:
            if kind == 'p':
                buf = self.buffers.get(name)
                setattr(m, name, dev.ptr(buf) if buf is not None else None)
        self.mesh_struct = m
        nbytes = self.lib.nhf_workspace_bytes(ctypes.byref(m), ctypes.byref(params))
        self.workspace = dev.empty(nbytes) if nbytes else None
        handle = ctypes.c_void_p()
        rc = self.lib.nhf_create(ctypes.byref(handle), ctypes.byref(m), ctypes.byref(params),
                                 dev.ptr(self.workspace) if self.workspace is not None else None,
                                 nbytes, dev.stream_handle())
        abi.check(self.lib, rc, 'nhf_create')
        self.ctx = handle
        self.terms = 4 + 6 * layout.b_num
        self._trace = np.zeros((abi.MAX_BATCH, abi.TRACE_COLS), dtype=np.uint32)
        self._rain_state()

    def close(self):
        if getattr(self, 'ctx', None):
            self.lib.nhf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------------- state in original order
    def settle_extrema(self) -> None:
        """Write out the extrema the step defers for rising cells (nhf_settle_extrema)."""
        abi.check(self.lib, self.lib.nhf_settle_extrema(self.ctx), 'nhf_settle_extrema')

    def read_state(self, name: str) -> np.ndarray:
        if name in SETTLED:
            self.settle_extrema()
        dtype = np.uint64 if name == 'peak_tick' else np.float32
        values = self.device.read(self.buffers[name], dtype)
        if name in STATE_SIDES:
            return self.layout.sides_from_device(values)
        return self.layout.cells_from_device(values)

    def write_state(self, name: str, values: np.ndarray) -> None:
        if name in SETTLED or name == 'depth':
            self.settle_extrema()     # a marked cell's maximum is its depth: keep the pair final
        dtype = np.uint64 if name == 'peak_tick' else np.float32
        values = np.asarray(values, dtype=dtype)
        check_extrema(name, values)
        if name in STATE_SIDES:
            values = self.layout.sides_to_device(values)
        else:
            values = self.layout.cells_to_device(values)
        self.device.write(self.buffers[name], values)
        if name in ('rain_time', 'capacity'):
            self._rain_state()

    def _rain_state(self) -> None:
        """The rain state after the rain time or capacity was written (nhf_set_rain_state):
        uniform when uniform_rain allows it, else per cell."""
        rt = self.device.read(self.buffers['rain_time'], np.float32)[1:]
        cap = self.device.read(self.buffers['capacity'], np.float32)[1:]
        cls = self.layout.arrays['cell_meta'][1:].astype(np.int64) & 255
        uniform, rt0, table = uniform_rain(rt, cap, cls, self.layout.n_pixels)
        uniform = uniform and self.layout.s_num > 1   # the flux pass advances the uniform state
        rc = self.lib.nhf_set_rain_state(self.ctx, int(uniform), ctypes.c_float(rt0),
                                         table.ctypes.data if uniform else None)
        abi.check(self.lib, rc, 'nhf_set_rain_state')
        self.rain_uniform = uniform

    # ----------------------------------------------------------------- gate control
    def set_gates(self, spec: dict) -> None:
        """Bind gate control (nhf_set_gates): lists in device order from gate.GateController, fresh
        flags, ticks and budget accumulators."""
        dev = self.device
        count = int(spec['count'])
        chunks = np.ascontiguousarray(spec['chunks'], np.int32).reshape(-1, 4)
        bufs = {name: dev.array(np.ascontiguousarray(spec[name], dtype))
                for name, dtype in (('control', np.int32), ('cells', np.int32), ('cell_gate', np.int32),
                                    ('pond', np.uint8), ('closed', np.float32), ('items', np.int32))}
        bufs['chunks'] = dev.array(chunks.ravel())
        bufs['flags'] = dev.array(np.zeros(3 * count, np.int32))
        bufs['ticks'] = dev.array(np.zeros(3 * count, np.uint64))
        bufs['acc'] = dev.array(np.zeros(max(1, len(chunks)) * 12, np.float64))
        g = abi.GatesStruct(count=count, n_cells=len(spec['cells']), n_items=len(spec['items']),
                            n_chunks=len(chunks), trigger_level=np.float32(spec['trigger_level']),
                            opening_seconds=np.float32(spec['opening_seconds']),
                            open_elevation=np.float32(spec['open_elevation']), min_h=np.float32(spec['min_h']),
                            formal_tick=int(spec['formal_tick']))
        for name, buf in bufs.items():
            setattr(g, name, dev.ptr(buf))
        self.gate_buffers, self.gate_struct, self.gate_count = bufs, g, count
        abi.check(self.lib, self.lib.nhf_set_gates(self.ctx, ctypes.byref(g)), 'nhf_set_gates')

    def gate_check(self) -> None:
        abi.check(self.lib, self.lib.nhf_gate_check(self.ctx), 'nhf_gate_check')

    def gate_state(self):
        """(flags [count, 3] commanded/flow seen/opened, ticks [count, 3] their clock ticks)."""
        flags = self.device.read(self.gate_buffers['flags'], np.int32).reshape(-1, 3).copy()
        ticks = self.device.read(self.gate_buffers['ticks'], np.uint64).reshape(-1, 3).copy()
        return flags, ticks

    def set_gate_state(self, flags: np.ndarray, ticks: np.ndarray) -> None:
        self.device.write(self.gate_buffers['flags'], np.asarray(flags, np.int32).ravel())
        self.device.write(self.gate_buffers['ticks'], np.asarray(ticks, np.uint64).ravel())

    def flush_gate_budget(self) -> np.ndarray:
        out = np.zeros(2 * (10 * self.gate_count + 8), np.float64)
        abi.check(self.lib, self.lib.nhf_flush_gate_budget(self.ctx, out.ctypes.data), 'nhf_flush_gate_budget')
        return out.reshape(-1, 2)

    # ----------------------------------------------------------------- pipe coupling
    def set_pipe_q(self, rates: np.ndarray) -> None:
        """Per-cell pipe source rates (m^3/s, reference numbering) for the next coupling window."""
        self.device.upload(self.buffers['pipe_q'], self.layout.cells_to_device(np.asarray(rates, np.float32)))

    def take_pipe_volume(self) -> np.ndarray:
        """The window's applied pipe volume per cell (m^3, reference numbering); clears it and its
        compensation term for the next window."""
        volume = self.device.download(self.buffers['pipe_volume'], np.float32)
        self.device.zero(self.buffers['pipe_volume'])
        self.device.zero(self.buffers['pipe_correction'])
        return self.layout.cells_from_device(volume)

    # ----------------------------------------------------------------- batch protocol
    def reset(self, now_ticks: int) -> None:
        abi.check(self.lib, self.lib.nhf_reset(self.ctx, int(now_ticks)), 'nhf_reset')

    def set_forcing(self, coefficients: np.ndarray, rain_rates: np.ndarray | None, raining: bool):
        coeff = np.ascontiguousarray(coefficients, dtype=np.float32)
        rates = None if rain_rates is None else np.ascontiguousarray(rain_rates, dtype=np.float32)
        rc = self.lib.nhf_set_forcing(self.ctx, coeff.ctypes.data if coeff.size else None,
                                      rates.ctypes.data if rates is not None else None, int(raining))
        abi.check(self.lib, rc, 'nhf_set_forcing')

    def begin(self, target_ticks: int, origin_ticks: int) -> None:
        abi.check(self.lib, self.lib.nhf_begin(self.ctx, int(target_ticks), int(origin_ticks)), 'nhf_begin')

    def run_batch(self, steps: int) -> np.ndarray:
        rc = self.lib.nhf_run_batch(self.ctx, int(steps), self._trace.ctypes.data)
        abi.check(self.lib, rc, 'nhf_run_batch')
        return self._trace[:steps].copy()

    def launch_batch(self, steps: int) -> None:
        """Enqueue a batch on this engine's stream and return at once (see wait_batch)."""
        abi.check(self.lib, self.lib.nhf_launch_batch(self.ctx, int(steps)), 'nhf_launch_batch')
        self._pending = int(steps)
        marker = getattr(self.device, 'marker', None)
        self._done = marker() if marker is not None else None

    def ready(self) -> bool:
        """True when the launched batch and its trace copy have finished (wait_batch will not block)."""
        done = getattr(self, '_done', None)
        return done is None or done.query()

    def wait_batch(self) -> np.ndarray:
        """Wait for the launched batch and return its trace rows."""
        rc = self.lib.nhf_wait_batch(self.ctx, self._trace.ctypes.data)
        abi.check(self.lib, rc, 'nhf_wait_batch')
        return self._trace[:self._pending].copy()

    def compute_cfl(self) -> None:
        abi.check(self.lib, self.lib.nhf_compute_cfl(self.ctx), 'nhf_compute_cfl')

    def cfl_state(self):
        ndt = ctypes.c_float()
        side = ctypes.c_int32()
        abi.check(self.lib, self.lib.nhf_cfl_state(self.ctx, ctypes.byref(ndt), ctypes.byref(side)),
                  'nhf_cfl_state')
        return float(ndt.value), int(self.layout.side_perm[side.value])

    def flush_budget(self) -> np.ndarray:
        """Interval budget totals as [terms, 2] double-double pairs."""
        out = np.zeros(2 * self.terms, dtype=np.float64)
        rc = self.lib.nhf_flush_budget(self.ctx, out.ctypes.data)
        if rc == 1:
            raise FloatingPointError('Nonfinite F32 device budget diagnostic')
        abi.check(self.lib, rc, 'nhf_flush_budget')
        return out.reshape(self.terms, 2)

    def limit_capture(self, reset: bool = False):
        values = np.zeros(5, dtype=np.float32)
        ids = np.zeros(3, dtype=np.int32)
        rc = self.lib.nhf_limit_capture(self.ctx, values.ctypes.data, ids.ctypes.data, int(reset))
        abi.check(self.lib, rc, 'nhf_limit_capture')
        return values, np.array([self.layout.side_perm[ids[0]], self.layout.cell_perm[ids[1]],
                                 self.layout.cell_perm[ids[2]]], dtype=np.int64)

    def set_graph(self, enable: bool) -> None:
        self.lib.nhf_set_graph(self.ctx, int(enable))

    def set_timing(self, enable: bool) -> None:
        self.lib.nhf_set_timing(self.ctx, int(enable))

    def time_kernels(self, repeats: int = 20) -> np.ndarray:
        """Mean microseconds per launch of each step kernel (ids as kernel_times) over `repeats`
        back-to-back launches; leaves the state meaningless (benchmark use)."""
        us = np.zeros(9, dtype=np.float64)
        abi.check(self.lib, self.lib.nhf_time_kernels(self.ctx, int(repeats), us.ctypes.data), 'nhf_time_kernels')
        return us

    def set_tuning(self, cfl_blocks: int, cell_block: int = 256, cells_per_thread: int = 1) -> None:
        rc = self.lib.nhf_set_tuning(self.ctx, int(cfl_blocks), int(cell_block), int(cells_per_thread))
        abi.check(self.lib, rc, 'nhf_set_tuning')

    def tuning_range(self) -> dict:
        out = np.zeros(4, dtype=np.int32)
        self.lib.nhf_tuning_range(self.ctx, out.ctypes.data)
        return dict(cfl_blocks=int(out[0]), cfl_max=int(out[1]), cell_block=int(out[2]),
                    cells_per_thread=int(out[3]))

    def kernel_times(self, reset: bool = True):
        ms = np.zeros(8, dtype=np.float64)
        count = np.zeros(8, dtype=np.int64)
        self.lib.nhf_kernel_times(self.ctx, ms.ctypes.data, count.ctypes.data, 8, int(reset))
        return ms, count

    def synchronize(self) -> None:
        abi.check(self.lib, self.lib.nhf_sync(self.ctx), 'nhf_sync')


def make_params(afa: float, sita: float, min_h: float, batch: int = 32) -> abi.Params:
    """Constants as F32, the way the reference's kernels see its Python floats."""
    return abi.Params(g=np.float32(9.81), afa=np.float32(afa), theta=np.float32(sita),
                      theta_half=np.float32((1.0 - sita) / 2.0), min_h=np.float32(min_h),
                      min_dt=np.float32(1e-4), batch_capacity=batch, reserved=0)
