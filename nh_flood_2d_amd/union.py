"""Meshes of different sizes advanced by the same launches (the nhu_* entry points).

:class:`UnionEngine` concatenates several meshes ("parts") into one device mesh. Part p's real cells
occupy [cell_begin, cell_end) with cell_begin = 1 + 1024 k, padded to a multiple of 1024 cells,
and its sides likewise; every index inside its arrays is offset into the concatenation, and cell 0
and side 0 are the one shared sentinel. On the device each part keeps its own clock, dt, forcing,
CFL reduction, limiter capture and budget, so it takes exactly the steps it would take alone, and
one nhu_run_batch advances every part that has a target: small meshes no longer pay a launch
sequence each. :meth:`UnionEngine.view` hands a Simulation an engine for one part with the
interface of :class:`~nh_flood_2d_amd.engine.Engine`; :class:`UnionRunner` runs such Simulations
in lockstep, one batch of launches for all of them.

The reference runs one mesh per process; this has no counterpart there. The sentinel is shared,
so every part's sentinel bed and depth must be 0 (the synthetic and NE meshes give 0); gate
control and pipe exchange stay with single-mesh engines.
"""
from __future__ import annotations

import ctypes
from time import perf_counter

import numpy as np

from . import abi
from .engine import SETTLED, STATE_CELLS, STATE_SIDES, TorchDevice, check_extrema, uniform_rain
from .layout import META_LEVEL_SHIFT, Layout

ALIGN = abi.PART_ALIGN


def _padded(n: int) -> int:
    return max(1, -(-n // ALIGN)) * ALIGN


def _union_size_base(layouts):
    """A common size base for the level bits, with each part's level offset; (0.0, None) when the
    parts' bases are not power-of-two multiples of the smallest (the kernels then read sizes)."""
    bases = [lay.size_base for lay in layouts]
    if min(bases) <= 0:
        return 0.0, None
    base = np.float32(min(bases))
    shifts = []
    for lay in layouts:
        shift = int(np.rint(np.log2(lay.size_base / float(base))))
        levels = (lay.arrays['cell_meta'][1:] >> META_LEVEL_SHIFT) & 15
        top = int(levels.max()) if len(levels) else 0
        if np.float32(np.ldexp(base, shift)) != np.float32(lay.size_base) or top + shift > 15:
            return 0.0, None
        shifts.append(shift)
    return float(base), shifts


class UnionPlan:
    """Ranges of every part in the concatenation and the concatenated static arrays."""

    def __init__(self, layouts: list[Layout]):
        if not 1 <= len(layouts) <= abi.MAX_PARTS:
            raise ValueError(f'a union holds 1..{abi.MAX_PARTS} meshes, not {len(layouts)}')
        theta = {lay.theta for lay in layouts}
        if len(theta) != 1:
            raise ValueError('every part of a union uses the same sita')
        self.layouts = layouts
        n = len(layouts)
        self.cell_begin = np.zeros(n, np.int64)
        self.side_begin = np.zeros(n, np.int64)
        self.isl_base = np.zeros(n, np.int64)
        self.b_base = np.zeros(n, np.int64)
        self.pix_base = np.zeros(n, np.int64)
        self.hang_base = np.zeros(n, np.int64)
        self.grad_base = np.zeros(n, np.int64)
        self.class_base = np.zeros(n, np.int64)
        cell, side, isl, bnd, pix, hang, grad, cls = 1, 1, 0, 0, 0, 0, 0, 0
        for p, lay in enumerate(layouts):
            self.cell_begin[p], self.side_begin[p], self.isl_base[p] = cell, side, isl
            self.b_base[p], self.pix_base[p], self.class_base[p] = bnd, pix, cls
            self.hang_base[p], self.grad_base[p] = hang, grad
            cell += _padded(lay.e_num - 1)
            side += _padded(lay.s_num - 1)
            isl += int(lay.arrays['_r_len'][0])
            bnd += lay.b_num
            pix += max(1, lay.n_pixels)
            hang += lay.n_hang
            grad += lay.n_grad
            classes = lay.arrays['cell_meta'][1:] & 255
            cls += int(classes.max()) + 1 if len(classes) else 1
        if cls > 256:
            raise ValueError('the parts of a union use more than 256 land classes together')
        if bnd > 254:
            raise ValueError('the parts of a union have more than 254 prescribed boundaries together')
        self.e_num, self.s_num, self.r_len = cell, side, isl
        self.b_num, self.n_pixels, self.n_hang, self.n_grad = bnd, pix, hang, grad
        self.theta = theta.pop()
        self.size_base, shifts = _union_size_base(layouts)
        self.arrays = self._arrays(shifts)

    # global ids of a part's local ids (0 stays the shared sentinel)
    def cells(self, p: int, ids: np.ndarray) -> np.ndarray:
        ids = np.asarray(ids, np.int64)
        return np.where(ids == 0, 0, ids + self.cell_begin[p] - 1)

    def sides(self, p: int, ids: np.ndarray) -> np.ndarray:
        ids = np.asarray(ids, np.int64)
        return np.where(ids == 0, 0, ids + self.side_begin[p] - 1)

    def part_struct(self) -> ctypes.Array:
        parts = (abi.PartStruct * len(self.layouts))()
        for p, lay in enumerate(self.layouts):
            cb, sb = int(self.cell_begin[p]), int(self.side_begin[p])
            parts[p] = abi.PartStruct(cb, cb + lay.e_num - 1, sb, sb + lay.s_num - 1, int(self.b_base[p]),
                                      lay.b_num, int(self.pix_base[p]), max(1, lay.n_pixels))
        return parts

    def _arrays(self, shifts) -> dict:
        E, S, R = self.e_num, self.s_num, max(1, self.r_len)
        hanging = self.n_hang > 0
        out = dict(
            area=np.zeros(E, np.float32), size=np.zeros(E, np.float32), cell_meta=np.zeros(E, np.uint32),
            class_horton=np.zeros(256 * 3, np.float32), rain_index=np.zeros(E, np.int32),
            bweight=np.zeros(E, np.float32), csr_start=np.zeros(E, np.int32), csr_count=np.zeros(E, np.uint32),
            isl=np.zeros(R, np.int32), isl_cell=np.zeros(R, np.int32),
            side_cells=np.zeros(2 * S, np.int32), side_meta=np.zeros(S, np.uint8),
            length=np.zeros(S, np.float32), sdc=np.zeros(S, np.float32), manning=np.zeros(S, np.float32),
            side_key=np.zeros(S, np.uint32), kind=np.zeros(max(1, self.b_num), np.int32),
            cell_quad=np.tile(np.array([-1, 0, 0, 0], np.int32), E),
        )
        if self.theta != 1.0:
            out['open_face'] = np.zeros(4 * E, np.float32)
        if hanging:
            out.update(x=np.zeros(E, np.float32), y=np.zeros(E, np.float32), edge_index=np.full(S, -1, np.int32),
                       hang_offset=np.zeros(2 * self.n_hang, np.float32),
                       hang_slot=np.zeros(2 * self.n_hang, np.int32), grad_cell=np.zeros(max(1, self.n_grad), np.int32),
                       grad_axes=np.zeros(max(1, self.n_grad), np.int32))
        # padding cells are never stepped; unit area and size keep every value finite where a
        # whole concatenation is stepped as one mesh (the oracle check of the concatenation)
        out['area'][:] = 1.0
        out['size'][:] = 1.0
        out['area'][0] = out['size'][0] = 0.0
        horton = out['class_horton'].reshape(256, 3)
        for p, lay in enumerate(self.layouts):
            a = lay.arrays
            nc, ns, r = lay.e_num - 1, lay.s_num - 1, int(a['_r_len'][0])
            c0, s0, i0 = int(self.cell_begin[p]), int(self.side_begin[p]), int(self.isl_base[p])
            cs, ss = slice(c0, c0 + nc), slice(s0, s0 + ns)
            for name in ('area', 'size', 'bweight'):
                out[name][cs] = a[name][1:]
            meta = a['cell_meta'][1:].astype(np.int64)
            cls = (meta & 255) + int(self.class_base[p])
            bid = ((meta >> 8) & 255) - 1
            bid = np.where(bid >= 0, bid + int(self.b_base[p]) + 1, 0)
            level = (meta >> META_LEVEL_SHIFT) & 15
            level = level + shifts[p] if shifts is not None else np.zeros_like(level)
            out['cell_meta'][cs] = (cls | (bid << 8) | (level << META_LEVEL_SHIFT)).astype(np.uint32)
            k = (np.max(a['cell_meta'][1:] & 255) + 1) if nc else 1
            horton[int(self.class_base[p]):int(self.class_base[p]) + k] = a['class_horton'].reshape(256, 3)[:k]
            out['rain_index'][cs] = a['rain_index'][1:] + int(self.pix_base[p])
            out['csr_start'][cs] = a['csr_start'][1:] + i0
            out['csr_count'][cs] = a['csr_count'][1:]
            quad = a['cell_quad'].reshape(-1, 4)[1:].astype(np.int64).copy()
            regular = quad[:, 0] >= 0
            quad[regular] = self.sides(p, quad[regular])
            quad[~regular, 1] += i0
            out['cell_quad'].reshape(-1, 4)[cs] = quad.astype(np.int32)
            if r:
                out['isl'][i0:i0 + r] = self.sides(p, a['isl'][:r])
                out['isl_cell'][i0:i0 + r] = self.cells(p, a['isl_cell'][:r])
            sc = a['side_cells'].reshape(-1, 2)[1:]
            out['side_cells'].reshape(-1, 2)[ss] = self.cells(p, sc)
            for name in ('side_meta', 'length', 'sdc', 'manning', 'side_key'):
                out[name][ss] = a[name][1:]
            if lay.b_num:
                out['kind'][int(self.b_base[p]):int(self.b_base[p]) + lay.b_num] = a['kind'][:lay.b_num]
            if self.theta != 1.0:
                out['open_face'][4 * c0:4 * (c0 + nc)] = a['open_face'][4:4 * (nc + 1)]
            if hanging and lay.n_hang:
                h0, g0 = int(self.hang_base[p]), int(self.grad_base[p])
                out['x'][cs], out['y'][cs] = a['x'][1:], a['y'][1:]
                e = a['edge_index'][1:].astype(np.int64)
                out['edge_index'][ss] = np.where(e >= 0, e + h0, -1)
                out['hang_offset'][2 * h0:2 * (h0 + lay.n_hang)] = a['hang_offset']
                out['hang_slot'][2 * h0:2 * (h0 + lay.n_hang)] = a['hang_slot'] + g0
                out['grad_cell'][g0:g0 + lay.n_grad] = self.cells(p, a['grad_cell'])
                out['grad_axes'][g0:g0 + lay.n_grad] = a['grad_axes']
        return out


class UnionEngine:
    """Device state of several meshes concatenated into one, and the nhu_* context over it."""

    def __init__(self, layouts: list[Layout], params: abi.Params, device=None):
        self.plan = plan = UnionPlan(layouts)
        self.layouts = layouts
        self.params = params
        self.device = device if device is not None else TorchDevice()
        self.lib = self.device.lib
        if not hasattr(self.lib, 'nhu_create'):
            raise RuntimeError('this library has no nhu_* entry points')
        dev = self.device
        self.state = {name: dev.array(np.zeros(plan.e_num, np.uint64 if name == 'peak_tick' else np.float32))
                      for name in STATE_CELLS}
        self.state['q'] = dev.array(np.zeros(plan.s_num, np.float32))
        self.static = {name: dev.array(values) for name, values in plan.arrays.items()}
        m = abi.MeshStruct()
        m.e_num, m.s_num, m.r_len = plan.e_num, plan.s_num, plan.r_len
        m.b_num, m.n_pixels, m.n_grad, m.n_hang = plan.b_num, plan.n_pixels, plan.n_grad, plan.n_hang
        m.flags = 0
        m.size_base = plan.size_base
        for name, kind in abi.MESH_FIELDS:
            if kind == 'p':
                buf = self.state.get(name, self.static.get(name))
                setattr(m, name, dev.ptr(buf) if buf is not None else None)
        self.mesh_struct = m
        self.parts = plan.part_struct()
        count = len(layouts)
        nbytes = self.lib.nhu_workspace_bytes(ctypes.byref(m), ctypes.byref(params), self.parts, count)
        if not nbytes:
            raise ValueError('nhu_workspace_bytes refused the parts')
        self.workspace = dev.empty(nbytes)
        handle = ctypes.c_void_p()
        rc = self.lib.nhu_create(ctypes.byref(handle), ctypes.byref(m), ctypes.byref(params), self.parts, count,
                                 dev.ptr(self.workspace), nbytes, dev.stream_handle())
        abi.check(self.lib, rc, 'nhu_create')
        self.ctx = handle
        self.owner = [None] * count
        self.batches = 0
        self._pending = 0

    # ------------------------------------------------------------------ parts
    def view(self, part: int):
        """An engine factory for Simulation bound to `part` (the layout must be that part's)."""
        def factory(layout, params, state, pipe_q=None):
            if pipe_q is not None:
                raise NotImplementedError('pipe exchange is not supported in a union')
            ref = self.layouts[part]
            if layout is not ref and not (
                    np.array_equal(layout.cell_perm, ref.cell_perm) and np.array_equal(layout.side_perm, ref.side_perm)
                    and layout.arrays.keys() == ref.arrays.keys()
                    and all(np.array_equal(layout.arrays[k], ref.arrays[k]) for k in ref.arrays)):
                raise ValueError(f'the layout differs from the one of part {part}')
            for f, _ in abi.Params._fields_:
                if f not in ('reserved', 'batch_capacity') and getattr(params, f) != getattr(self.params, f):
                    raise ValueError(f'parameter {f} differs from the union')
            if self.owner[part] is not None:
                raise RuntimeError(f'part {part} is in use')
            v = PartView(self, part)
            for name in STATE_CELLS + STATE_SIDES:
                v.write_state(name, state[name])
            self.owner[part] = v
            return v
        return factory

    def release(self, part: int) -> None:
        self.owner[part] = None

    # ------------------------------------------------------------------ batch protocol
    def run_batch(self, steps: int) -> np.ndarray:
        """Advance every part with a begun target; returns the [parts, steps, 9] trace."""
        buf = np.zeros((len(self.layouts), steps, abi.TRACE_COLS), dtype=np.uint32)
        abi.check(self.lib, self.lib.nhu_run_batch(self.ctx, int(steps), buf.ctypes.data), 'nhu_run_batch')
        self.batches += 1
        return buf

    def launch_batch(self, steps: int) -> None:
        abi.check(self.lib, self.lib.nhu_launch_batch(self.ctx, int(steps)), 'nhu_launch_batch')
        self._pending = int(steps)
        marker = getattr(self.device, 'marker', None)
        self._done = marker() if marker is not None else None

    def ready(self) -> bool:
        """True when the launched batch and its trace copy have finished."""
        done = getattr(self, '_done', None)
        return done is None or done.query()

    def wait_batch(self) -> np.ndarray:
        buf = np.zeros((len(self.layouts), self._pending, abi.TRACE_COLS), dtype=np.uint32)
        abi.check(self.lib, self.lib.nhu_wait_batch(self.ctx, buf.ctypes.data), 'nhu_wait_batch')
        self.batches += 1
        return buf

    def compute_cfl(self) -> None:
        abi.check(self.lib, self.lib.nhu_compute_cfl(self.ctx), 'nhu_compute_cfl')

    def settle_extrema(self) -> None:
        abi.check(self.lib, self.lib.nhf_settle_extrema(self.ctx), 'nhf_settle_extrema')

    def set_graph(self, enable: bool) -> None:
        self.lib.nhf_set_graph(self.ctx, int(enable))

    def set_timing(self, enable: bool) -> None:
        self.lib.nhf_set_timing(self.ctx, int(enable))

    def set_tuning(self, cell_block: int = 256, cells_per_thread: int = 1) -> None:
        """Workgroup shape of the cell passes (each part's CFL grid is fixed at creation)."""
        rc = self.lib.nhf_set_tuning(self.ctx, 8, int(cell_block), int(cells_per_thread))
        abi.check(self.lib, rc, 'nhf_set_tuning')

    def synchronize(self) -> None:
        abi.check(self.lib, self.lib.nhf_sync(self.ctx), 'nhf_sync')

    def close(self) -> None:
        if getattr(self, 'ctx', None):
            self.lib.nhf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PartView:
    """One part of a UnionEngine with the Engine interface used by Simulation."""

    def __init__(self, union: UnionEngine, part: int):
        self.union = union
        self.part = part
        self.layout = union.layouts[part]
        self.lib = union.lib
        self.terms = 4 + 6 * self.layout.b_num
        plan = union.plan
        self.c0, self.nc = int(plan.cell_begin[part]), self.layout.e_num - 1
        self.s0, self.ns = int(plan.side_begin[part]), self.layout.s_num - 1
        self.sentinel = {}   # the part's own values at cell 0 / side 0 (kept on the host)

    @property
    def slot(self) -> int:
        """The part's row in the union's trace (as a batch slot's in a batch trace)."""
        return self.part

    def _slice(self, name):
        buf = self.union.state[name]
        return buf[self.s0:self.s0 + self.ns] if name in STATE_SIDES else buf[self.c0:self.c0 + self.nc]

    def read_state(self, name: str) -> np.ndarray:
        if name in SETTLED:
            self.union.settle_extrema()
        dev = self.union.device
        dev.stream.synchronize()
        with dev.torch.cuda.stream(dev.stream):
            body = self._slice(name).cpu().numpy()
        dtype = np.uint64 if name == 'peak_tick' else np.float32
        body = body.view(dtype)
        local = np.concatenate([[self.sentinel.get(name, dtype(0))], body]).astype(dtype)
        if name in STATE_SIDES:
            return self.layout.sides_from_device(local)
        return self.layout.cells_from_device(local)

    def write_state(self, name: str, values: np.ndarray) -> None:
        if name in SETTLED or name == 'depth':
            self.union.settle_extrema()
        dtype = np.uint64 if name == 'peak_tick' else np.float32
        values = np.asarray(values, dtype=dtype)
        check_extrema(name, values)
        local = self.layout.sides_to_device(values) if name in STATE_SIDES else self.layout.cells_to_device(values)
        if name in ('bed', 'depth') and local[0] != 0:
            raise ValueError(f'a union shares one sentinel cell: the part\'s {name} at cell 0 must be 0')
        self.sentinel[name] = local[0]
        dev = self.union.device
        host = local[1:].view(np.int64) if dtype == np.uint64 else local[1:]
        with dev.torch.cuda.stream(dev.stream):
            self._slice(name).copy_(dev.torch.from_numpy(np.ascontiguousarray(host)).to(dev.device))
        dev.stream.synchronize()
        if name in ('rain_time', 'capacity'):
            self._rain_state()

    def _rain_state(self) -> None:
        """The part's rain state after its rain time or capacity was written (nhu_set_rain_state),
        its class codes those of the union's cell_meta."""
        dev = self.union.device
        with dev.torch.cuda.stream(dev.stream):
            rt = self._slice('rain_time').cpu().numpy()
            cap = self._slice('capacity').cpu().numpy()
        meta = self.union.plan.arrays['cell_meta'][self.c0:self.c0 + self.nc]
        uniform, rt0, table = uniform_rain(rt, cap, meta.astype(np.int64) & 255, self.layout.n_pixels)
        uniform = uniform and self.layout.s_num > 1   # the flux pass advances the uniform state
        rc = self.lib.nhu_set_rain_state(self.union.ctx, self.part, int(uniform), ctypes.c_float(rt0),
                                         table.ctypes.data if uniform else None)
        abi.check(self.lib, rc, 'nhu_set_rain_state')

    # ------------------------------------------------------------------ per-part protocol
    def reset(self, now_ticks: int) -> None:
        abi.check(self.lib, self.lib.nhu_reset(self.union.ctx, self.part, int(now_ticks)), 'nhu_reset')

    def set_forcing(self, coefficients: np.ndarray, rain_rates: np.ndarray | None, raining: bool):
        coeff = np.ascontiguousarray(coefficients, dtype=np.float32)
        rates = None if rain_rates is None else np.ascontiguousarray(rain_rates, dtype=np.float32)
        if self.layout.b_num and coeff.size < 2 * self.layout.b_num:
            raise ValueError('coefficients must hold two values per boundary')
        rc = self.lib.nhu_set_forcing(self.union.ctx, self.part, coeff.ctypes.data if coeff.size else None,
                                      rates.ctypes.data if rates is not None else None, int(raining))
        abi.check(self.lib, rc, 'nhu_set_forcing')

    def begin(self, target_ticks: int, origin_ticks: int) -> None:
        rc = self.lib.nhu_begin(self.union.ctx, self.part, int(target_ticks), int(origin_ticks))
        abi.check(self.lib, rc, 'nhu_begin')

    def run_batch(self, steps: int) -> np.ndarray:
        """Run one batch of the whole union (every part with a target) and return this part's rows."""
        return self.union.run_batch(steps)[self.part].copy()

    def compute_cfl(self) -> None:
        self.union.compute_cfl()

    def _local_side(self, side: int) -> int:
        return 0 if side == 0 else side - self.s0 + 1

    def _local_cell(self, cell: int) -> int:
        return 0 if cell == 0 else cell - self.c0 + 1

    def cfl_state(self):
        ndt = ctypes.c_float()
        side = ctypes.c_int32()
        rc = self.lib.nhu_cfl_state(self.union.ctx, self.part, ctypes.byref(ndt), ctypes.byref(side))
        abi.check(self.lib, rc, 'nhu_cfl_state')
        return float(ndt.value), int(self.layout.side_perm[self._local_side(side.value)])

    def flush_budget(self) -> np.ndarray:
        out = np.zeros(2 * self.terms, dtype=np.float64)
        rc = self.lib.nhu_flush_budget(self.union.ctx, self.part, out.ctypes.data)
        if rc == 1:
            raise FloatingPointError('Nonfinite F32 device budget diagnostic')
        abi.check(self.lib, rc, 'nhu_flush_budget')
        return out.reshape(self.terms, 2)

    def limit_capture(self, reset: bool = False):
        values = np.zeros(5, dtype=np.float32)
        ids = np.zeros(3, dtype=np.int32)
        rc = self.lib.nhu_limit_capture(self.union.ctx, self.part, values.ctypes.data, ids.ctypes.data, int(reset))
        abi.check(self.lib, rc, 'nhu_limit_capture')
        return values, np.array([self.layout.side_perm[self._local_side(ids[0])],
                                 self.layout.cell_perm[self._local_cell(ids[1])],
                                 self.layout.cell_perm[self._local_cell(ids[2])]], dtype=np.int64)

    def set_graph(self, enable: bool) -> None:
        self.union.set_graph(enable)

    def set_timing(self, enable: bool) -> None:
        self.union.set_timing(enable)

    def synchronize(self) -> None:
        self.union.synchronize()

    def close(self) -> None:
        self.union.release(self.part)


class UnionRunner:
    """Simulations bound to the parts of one UnionEngine, advanced in lockstep: each iteration
    plans every live Simulation, runs one union batch and hands each its own trace rows."""

    def __init__(self, union: UnionEngine, sims: dict, batch_size: int = 32):
        for sim in sims.values():
            if not isinstance(sim.engine, PartView) or sim.engine.union is not union:
                raise ValueError('every Simulation must run on a part of this union')
            if sim.settings.batch_size != batch_size:
                raise ValueError('every Simulation of a union uses the same batch size')
        self.union = union
        self.sims = dict(sims)
        self.batch_size = batch_size
        self.results = {}
        self.seconds = {}

    def run(self) -> dict:
        t0 = perf_counter()
        live = dict(self.sims)
        for sim in live.values():
            sim.start()
        while live:
            planned = []
            for key, sim in list(live.items()):
                if sim.plan():
                    planned.append(sim)
                else:
                    self.results[key] = sim.finish()
                    self.seconds[key] = perf_counter() - t0
                    del live[key]
            if not planned:
                break
            t1 = perf_counter()
            trace = self.union.run_batch(self.batch_size)
            dt = perf_counter() - t1
            for sim in planned:
                sim.device_seconds += dt
                sim.consume(trace[sim.engine.part])
        return self.results
