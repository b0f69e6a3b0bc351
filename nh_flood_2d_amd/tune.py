"""Autotuner for the launch parameters of the step.

What it picks, by timing candidates on a scratch engine built from the solve's own layout and
initial state (the real solve's state, clock and budget are never touched):

* single mesh: the workgroup size and cells per thread of the two cell passes, the CFL partial
  grid (workgroups of the CFL pass, hence sides per thread) and the steps per captured graph;
* scenario batch: the grid of the grid-stride depth pass (cells x slots per thread), the CFL
  grid and the steps per graph;
* meshes of different sizes in one union: the workgroup size and cells per thread of the cell
  passes and the steps per graph (each part's CFL grid follows its side count).

The choice changes launch shapes only: every candidate computes the same bits. Results are kept
per size class (log2 of the cell count, slots, theta, 2:1 or aligned) in a :class:`TuningCache`,
optionally persisted as JSON. The reference has no counterpart (Taichi picks its own launches).
"""
from __future__ import annotations

import json
import math
from pathlib import Path
from time import perf_counter

import numpy as np

CFL_GRIDS = (256, 512, 1024, 2048, 4096)
DEPTH_GRIDS = (512, 1024, 2048, 4096, 8192)
STEPS_PER_GRAPH = (16, 32, 64)
# (workgroup size, cells per thread) of the two cell passes
CELL_SHAPES = ((128, 1), (256, 1), (512, 1), (128, 2), (256, 2))


def size_class(layout, slots: int = 1) -> str:
    cells = max(layout.e_num - 1, 1)
    return (f'cells2^{round(math.log2(cells))}-slots{slots}-theta{"1" if layout.theta == 1. else "x"}'
            f'-{"hanging" if layout.n_hang else "aligned"}')


class TuningCache:
    """Chosen parameters per size class; `path` (JSON) makes them persist across processes."""

    def __init__(self, path: str | Path | None = None):
        self.path = Path(path) if path else None
        self.entries = {}
        if self.path and self.path.exists():
            self.entries = json.loads(self.path.read_text())

    def get(self, key: str):
        return self.entries.get(key)

    def put(self, key: str, value: dict) -> None:
        self.entries[key] = value
        if self.path:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            self.path.write_text(json.dumps(self.entries, indent=1, sort_keys=True))


def _advance(begin, run, steps: int, per_graph: int, horizon: int) -> None:
    """Run `steps` steps towards a far target (the forcing origin stays at tick 0: the local
    times only feed boundary values, which do not change the work)."""
    done = 0
    while done < steps:
        n = min(per_graph, steps - done)
        begin(horizon, 0)
        trace = run(n)
        rows = trace.reshape(-1, n, trace.shape[-1])
        if not (rows[:, :, 3] == 1).any():
            raise RuntimeError('tuning run stopped: no slot completed a step')
        done += n


def _time(engine_sync, begin, run, steps, per_graph, horizon) -> float:
    _advance(begin, run, per_graph, per_graph, horizon)        # capture the graph first
    engine_sync()
    t0 = perf_counter()
    _advance(begin, run, steps, per_graph, horizon)
    engine_sync()
    return (perf_counter() - t0) / steps


def pick(times: dict) -> tuple:
    """The candidate with the smallest time (ties: the earlier candidate)."""
    return min(times, key=lambda k: (times[k], list(times).index(k)))


def tune_engine(engine, steps: int = 128, rounds: int = 2, cfl_grids=CFL_GRIDS, cell_shapes=CELL_SHAPES,
                steps_per_graph=STEPS_PER_GRAPH) -> dict:
    """Time a single-mesh Engine (already reset, forcing set): the cell-pass shape (workgroup size
    and cells per thread) first (at the default CFL grid, 32 steps per graph), then the CFL grid,
    then the steps per graph."""
    rng = engine.tuning_range()
    grids = [g for g in cfl_grids if g <= rng['cfl_max']] or [rng['cfl_max'] - rng['cfl_max'] % 8 or 8]
    default_grid = rng['cfl_blocks']
    horizon = 1 << 62
    begin = lambda target, origin: engine.begin(target, origin)
    engine.set_graph(True)
    times = {}

    def measure(g, shape, pg):
        engine.set_tuning(g, shape[0], shape[1])
        t = _time(engine.synchronize, begin, engine.run_batch, steps, pg, horizon)
        times[(g, shape, pg)] = min(times.get((g, shape, pg), math.inf), t)

    for _ in range(rounds):                       # interleaved rounds, minimum kept
        for shape in cell_shapes:
            measure(default_grid, shape, 32)
    best_shape = pick({k: v for k, v in times.items() if k[0] == default_grid and k[2] == 32})[1]
    for _ in range(rounds):
        for g in grids:
            measure(g, best_shape, 32)
    best_g = pick({k: v for k, v in times.items() if k[1] == best_shape and k[2] == 32})[0]
    for _ in range(rounds):
        for pg in steps_per_graph:
            measure(best_g, best_shape, pg)
    best = pick({k: v for k, v in times.items() if k[0] == best_g and k[1] == best_shape})
    engine.set_tuning(best[0], best[1][0], best[1][1])
    return dict(cfl_blocks=best[0], cell_block=best[1][0], cells_per_thread=best[1][1], steps_per_graph=best[2],
                us_per_step=times[best] * 1e6,
                table={f'{a}/{b[0]}x{b[1]}/{c}': v * 1e6 for (a, b, c), v in times.items()})


def tune_batch(batch, views, steps: int = 64, rounds: int = 2, depth_grids=DEPTH_GRIDS, cfl_grids=CFL_GRIDS,
               steps_per_graph=STEPS_PER_GRAPH) -> dict:
    """Time a BatchEngine whose slots are bound (views reset, forcing set): the depth grid first,
    then the CFL grid at the best depth grid, then the steps per graph."""
    rng = batch.tuning_range()
    depths = [g for g in depth_grids if g <= rng['depth_max']] or [rng['depth_max']]
    cfls = [g for g in cfl_grids if g <= rng['cfl_max']] or [rng['cfl_max']]
    horizon = 1 << 62

    def begin(target, origin):
        for v in views:
            v.begin(target, origin)
    batch.set_graph(True)
    cur_cfl = rng['cfl_blocks']
    times = {}

    def measure(dg, cg, pg):
        batch.set_tuning(dg, cg)
        t = _time(batch.synchronize, begin, batch.run_batch, steps, pg, horizon)
        times[(dg, cg, pg)] = min(times.get((dg, cg, pg), math.inf), t)
        return times[(dg, cg, pg)]

    for _ in range(rounds):
        for dg in depths:
            measure(dg, cur_cfl if cur_cfl <= rng['cfl_max'] else cfls[-1], 32)
    best_d = pick({k: v for k, v in times.items() if k[2] == 32})[0]
    for _ in range(rounds):
        for cg in cfls:
            measure(best_d, cg, 32)
    best_c = pick({k: v for k, v in times.items() if k[0] == best_d and k[2] == 32})[1]
    for _ in range(rounds):
        for pg in steps_per_graph:
            measure(best_d, best_c, pg)
    best = pick({k: v for k, v in times.items() if k[0] == best_d and k[1] == best_c})
    batch.set_tuning(best[0], best[1])
    return dict(depth_blocks=best[0], cfl_blocks=best[1], steps_per_graph=best[2], us_per_step=times[best] * 1e6,
                table={f'{a}/{b}/{c}': v * 1e6 for (a, b, c), v in times.items()})


def tuned_engine_parameters(layout, params, state, device, forcing=None, cache: TuningCache | None = None,
                            steps: int = 128) -> dict:
    """Tuning of a single-mesh solve from a scratch copy of `state` (cached per size class)."""
    from .engine import Engine
    key = size_class(layout)
    if cache is not None and cache.get(key):
        return cache.get(key)
    scratch = Engine(layout, params, {k: np.array(v, copy=True) for k, v in state.items()}, device=device)
    try:
        if forcing is not None:
            scratch.set_forcing(*forcing)
        else:
            scratch.set_forcing(np.zeros(2 * layout.b_num, np.float32), None, False)
        scratch.reset(0)
        scratch.compute_cfl()
        result = tune_engine(scratch, steps=steps)
    finally:
        scratch.close()
    if cache is not None:
        cache.put(key, result)
    return result


def tune_union(union, views, steps: int = 64, rounds: int = 2, cell_shapes=CELL_SHAPES,
               steps_per_graph=STEPS_PER_GRAPH) -> dict:
    """Time a UnionEngine whose parts are bound (views reset, forcing set): the cell-pass shape at
    32 steps per graph, then the steps per graph. The union's own state advances while it is timed
    (every candidate computes the same bits, so only the timing depends on the candidate)."""
    horizon = 1 << 62

    def begin(target, origin):
        for v in views:
            v.begin(target, origin)
    union.set_graph(True)
    times = {}

    def measure(shape, pg):
        union.set_tuning(shape[0], shape[1])
        t = _time(union.synchronize, begin, union.run_batch, steps, pg, horizon)
        times[(shape, pg)] = min(times.get((shape, pg), math.inf), t)

    for _ in range(rounds):
        for shape in cell_shapes:
            measure(shape, 32)
    best_shape = pick({k: v for k, v in times.items() if k[1] == 32})[0]
    for _ in range(rounds):
        for pg in steps_per_graph:
            measure(best_shape, pg)
    best = pick({k: v for k, v in times.items() if k[0] == best_shape})
    union.set_tuning(best[0][0], best[0][1])
    return dict(cell_block=best[0][0], cells_per_thread=best[0][1], steps_per_graph=best[1],
                us_per_step=times[best] * 1e6,
                table={f'{a[0]}x{a[1]}/{b}': v * 1e6 for (a, b), v in times.items()})
