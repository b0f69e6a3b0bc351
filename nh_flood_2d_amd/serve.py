"""Serving: a request loop that runs flood solves of uneven size and duration on one GPU.

Requests arrive over time. Each is admitted when its device footprint fits the HBM budget; it then
keeps its mesh and state resident and runs the unchanged host protocol of
``solver_compact.py:930-985`` (:class:`~nh_flood_2d_amd.simulation.Simulation`). Every loop
iteration plans all live solves, enqueues one batch for each on its own HIP stream (so the
device overlaps small meshes instead of running them one after another) and then consumes the
traces in turn. A solve that finishes frees its memory and its slot for the queue at once
(continuous batching). Requests that share a mesh can share a :class:`BatchEngine`, one slot each.

Depth frames leave the loop through :class:`FrameStream`, whose file writes run on a background
thread. The reference serves one solve per process; this loop has no counterpart there, and each
solve gives the same results alone and in the mix (same kernels, same arithmetic per solve).
"""
from __future__ import annotations

import queue
import threading
from dataclasses import dataclass, field
from pathlib import Path
from time import perf_counter, sleep
from typing import Callable

import numpy as np

from .engine import Engine, TorchDevice


def footprint_bytes(layout, slots: int = 1) -> int:
    """Device bytes of one solve (or a batch of `slots`): static arrays, state, scratch."""
    static = sum(int(np.asarray(v).nbytes) for k, v in layout.arrays.items() if not k.startswith('_'))
    e, s = layout.e_num, layout.s_num
    state = slots * (e * (5 * 4 + 8) + s * 4) + e * 4
    scratch = slots * (e * 4 + s * 4 + (16 * e if layout.theta != 1. else 0)) + (1 << 20)
    return static + state + scratch


@dataclass
class Request:
    """One flood solve: `build(engine_factory)` returns its Simulation; `arrival` is in seconds
    after the server starts; `cells` counts cell-updates; `footprint` is its device bytes."""
    key: object
    build: Callable
    arrival: float
    cells: int
    footprint: int
    group: object = None       # requests with the same group share a BatchEngine (same mesh)
    layout: object = None      # the solve's layout and step parameters (afa, sita, min_h): with
    params: tuple | None = None  # both, the request may be served as a part of a union


@dataclass
class ServeReport:
    latency: dict = field(default_factory=dict)      # key -> arrival to final output, s
    queued: dict = field(default_factory=dict)       # key -> arrival to admission, s
    results: dict = field(default_factory=dict)      # key -> final mass-balance row
    steps: dict = field(default_factory=dict)
    cell_updates: int = 0
    wall: float = 0.
    peak_live: int = 0
    peak_bytes: int = 0
    batches: int = 0
    unions: int = 0            # union engines formed (each holding two or more solves)

    def percentile(self, p: float) -> float:
        values = np.array(sorted(self.latency.values()))
        return float(np.percentile(values, p)) if len(values) else float('nan')

    def summary(self) -> dict:
        return dict(requests=len(self.latency), p50_latency_s=self.percentile(50),
                    p99_latency_s=self.percentile(99), max_latency_s=max(self.latency.values(), default=0.),
                    mean_queue_s=float(np.mean(list(self.queued.values()))) if self.queued else 0.,
                    cell_updates=self.cell_updates, wall_s=self.wall,
                    cell_updates_per_s=self.cell_updates / self.wall if self.wall else 0.,
                    peak_live=self.peak_live, peak_device_bytes=self.peak_bytes, loop_iterations=self.batches,
                    unions=self.unions)


class FrameStream:
    """Output sink for served solves: mass-balance rows are kept, depth frames are written as FDB
    frame tables by a background thread (or only counted when `root` is None)."""

    def __init__(self, root: Path | None = None, every_output: bool = False):
        self.root = Path(root) if root is not None else None
        self.every_output = every_output
        self.rows = {}
        self.frames = {}
        self._queue = queue.Queue()
        self._thread = threading.Thread(target=self._writer, daemon=True)
        self._thread.start()

    def _writer(self):
        from .output import write_depth_frame
        while True:
            item = self._queue.get()
            if item is None:
                break
            path, depth = item
            write_depth_frame(path, depth)

    def sink(self, key):
        stream = self

        class _Sink:
            def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
                stream.rows.setdefault(key, []).append(row)
                if frame or stream.every_output:
                    stream.frames[key] = stream.frames.get(key, 0) + 1
                    if stream.root is not None:
                        stream._queue.put((stream.root / str(key) / f'depth_{int(seconds)}.fdb', depth.copy()))

            def finish(self, sim):
                pass
        return _Sink()

    def close(self):
        self._queue.put(None)
        self._thread.join()


class FloodServer:
    """Continuous batching of flood solves on one GPU, admitted by device footprint.

    With `union_window` > 0, requests that carry their layout and step parameters are admitted
    in groups: arrivals are held until the oldest has waited `union_window` seconds, `union_parts`
    are waiting, or no more arrivals are due, and the group (equal parameters, within the budget)
    becomes one UnionEngine on its own stream: one launch sequence steps every solve of the group,
    each with its own clock (nhu_*), instead of one per solve. A solve computes the same bits
    either way. The union's memory is released when its last solve finishes."""

    def __init__(self, budget_bytes: int | None = None, batch_size: int = 32, device: str = 'cuda:0',
                 max_live: int = 64, group_slots: int = 8, device_factory: Callable | None = None,
                 graph: bool = True, poll: bool = True, union_window: float = 0.0, union_parts: int = 8):
        if budget_bytes is None:
            import torch
            total = torch.cuda.get_device_properties(device).total_memory
            budget_bytes = int(0.85 * total)
        self.budget = budget_bytes
        self.batch_size = batch_size
        self.device_factory = device_factory or (lambda: TorchDevice(device))
        self.max_live = max_live
        self.group_slots = group_slots
        self.graph = graph
        self.poll = poll   # consume batches in the order they finish (else oldest first)
        self.union_window = union_window
        self.union_parts = union_parts

    def _engine_factory(self, stream_device):
        def make(layout, params, state, pipe_q=None):
            return Engine(layout, params, state, device=stream_device, pipe_q=pipe_q)
        return make

    def _unionable(self, r: Request) -> bool:
        return self.union_window > 0 and r.group is None and r.layout is not None and r.params is not None

    def _union_due(self, waiting, pending, now: float, realtime: bool) -> list[Request]:
        """The requests to admit as one union now (empty: keep holding)."""
        held = [r for r in waiting if self._unionable(r)]
        if not held:
            return []
        key = held[0].params
        group = [r for r in held if r.params == key][:self.union_parts]
        due = (len(group) >= self.union_parts or not pending or not realtime
               or now - min(r.arrival for r in group) >= self.union_window)
        return group if due else []

    def serve(self, requests: list[Request], realtime: bool = True) -> ServeReport:
        """Run every request; arrivals are honoured on the wall clock when `realtime`."""
        from .batch import BatchEngine
        from .engine import make_params
        from .union import UnionEngine
        report = ServeReport()
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        live = {}          # key -> (sim, request, owner): owner is the Engine or the BatchEngine
        groups = {}        # group -> BatchEngine
        group_use = {}     # id(shared engine: BatchEngine or UnionEngine) -> live members
        used = 0
        t0 = perf_counter()
        inflight = []      # (owner, members) in launch order
        busy = set()       # ids of owners with a batch in flight
        while pending or waiting or live:
            now = perf_counter() - t0
            while pending and (pending[0].arrival <= now or not realtime):
                waiting.append(pending.pop(0))
            # unions: a due group of held requests, within the budget, in one UnionEngine
            group = self._union_due(waiting, pending, now, realtime)
            while group and len(live) + len(group) > self.max_live:
                group = group[:-1]
            while group and used + sum(r.footprint for r in group) > self.budget and live:
                group = group[:-1]
            if len(group) >= 2:
                sb = group[0].params
                union = UnionEngine([r.layout for r in group], make_params(sb[0], sb[1], sb[2], self.batch_size),
                                    device=self.device_factory())
                if self.graph:
                    union.set_graph(True)
                group_use[id(union)] = len(group)
                report.unions += 1
                for p, r in enumerate(group):
                    waiting.remove(r)
                    sim = r.build(union.view(p))
                    if sim.settings.batch_size != self.batch_size:
                        raise ValueError('every served solve uses the server batch size')
                    sim.start()
                    live[r.key] = (sim, r, union)
                    used += r.footprint
                    report.queued[r.key] = perf_counter() - t0 - r.arrival
            # admission, first come first served, by footprint (a held union request is admitted
            # alone once it is due and no other solve of its parameters waits)
            while waiting and len(live) < self.max_live:
                r = waiting[0]
                if self._unionable(r) and r not in group:
                    break
                if used + r.footprint > self.budget and live:
                    break
                if r.group is not None and groups.get(r.group) is not None and id(groups[r.group]) in busy \
                        and groups[r.group].free_slots():
                    break   # binding a slot waits until the group's batch in flight is consumed
                waiting.pop(0)
                if r.group is not None:
                    batch = groups.get(r.group)
                    if batch is None or not batch.free_slots():
                        batch = BatchEngine(self.group_slots, device=self.device_factory())
                        groups[r.group] = batch
                        group_use[id(batch)] = 0
                    sim = r.build(batch.view())
                    owner = batch
                    group_use[id(batch)] += 1
                else:
                    sim = r.build(self._engine_factory(self.device_factory()))
                    owner = sim.engine
                if sim.settings.batch_size != self.batch_size:
                    raise ValueError('every served solve uses the server batch size')
                if self.graph:
                    sim.engine.set_graph(True)   # one captured graph launch per batch
                sim.start()
                live[r.key] = (sim, r, owner)
                used += r.footprint
                report.queued[r.key] = perf_counter() - t0 - r.arrival
            report.peak_live = max(report.peak_live, len(live))
            report.peak_bytes = max(report.peak_bytes, used)
            # plan every solve whose engine is idle and launch one batch per such engine; finish the
            # solves that are done
            planned = {}
            for key, (sim, r, owner) in list(live.items()):
                if id(owner) in busy:
                    continue
                if sim.plan():
                    planned.setdefault(id(owner), (owner, []))[1].append((key, sim))
                    continue
                report.results[key] = sim.finish()
                report.latency[key] = perf_counter() - t0 - r.arrival
                report.steps[key] = sim.steps
                report.cell_updates += r.cells * sim.steps
                sim.engine.close()
                if owner is not sim.engine:
                    group_use[id(owner)] -= 1
                    if group_use[id(owner)] == 0:
                        for g, b in list(groups.items()):
                            if b is owner:
                                del groups[g]
                        owner.close()
                del live[key]
                used -= r.footprint
            for oid, (owner, members) in planned.items():
                owner.launch_batch(self.batch_size)
                inflight.append((owner, members))
                busy.add(oid)
            if not inflight:
                if not live and not waiting and pending and realtime:
                    sleep(max(0., min(0.05, pending[0].arrival - (perf_counter() - t0))))
                continue
            # consume a finished batch while the others run, the first to finish rather than the
            # oldest (a small mesh's batch must not wait behind a large one's); its engine is planned
            # and relaunched at the top of the next pass. Until one finishes, keep polling, and go
            # back to admission as soon as a request arrives.
            ready = 0
            if self.poll:
                ready = None
                while ready is None:
                    ready = next((i for i, (o, _) in enumerate(inflight) if o.ready()), None)
                    if ready is None and pending and realtime and pending[0].arrival <= perf_counter() - t0:
                        break       # a request has arrived: admit it first
                if ready is None:
                    continue
            owner, members = inflight.pop(ready)
            trace = owner.wait_batch()
            busy.discard(id(owner))
            for key, sim in members:
                sim.consume(trace[sim.engine.slot] if owner is not sim.engine else trace)
            report.batches += 1
        report.wall = perf_counter() - t0
        return report
