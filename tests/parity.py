"""Helpers that drive the HIP engine and the CPU oracle through the same host protocol."""
from __future__ import annotations

import numpy as np

from nh_flood_2d_amd import abi
from nh_flood_2d_amd.clock import seconds_to_ticks
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, initial_state
from nh_flood_2d_amd.synth import scenario_forcing

import oracle


def coefficients_at(forcing, t: float) -> np.ndarray:
    """Boundary value at t and slope of its segment (upload_forcing, solver_compact.py:774-779)."""
    c = np.zeros((max(1, len(forcing.boundaries)), 2), dtype=np.float32)
    for bi, b in enumerate(forcing.boundaries):
        times = b.series.times
        i = min(int(np.searchsorted(times, t, side='right')) - 1, len(times) - 2)
        c[bi, 0] = b.series.at(t)
        c[bi, 1] = (b.series.values[i + 1] - b.series.values[i]) / (times[i + 1] - times[i])
    return c


class Pair:
    """The same scenario bound to a HIP engine and an oracle engine."""

    def __init__(self, sc, gpu_device, order='morton', theta=None, precision='f32', stored_sizes=False):
        self.sc = sc
        self.forcing = scenario_forcing(sc)
        theta = sc.sita if theta is None else theta
        self.layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, self.forcing),
                                   theta, order)
        if stored_sizes:        # the kernels read the size array instead of computing sizes
            self.layout.size_base = 0.0
        self.params = make_params(sc.afa, theta, sc.min_h, 32)
        state = initial_state(sc.mesh, self.layout)
        if sc.initial_depth is not None:
            state['depth'] = sc.initial_depth.astype(np.float32).copy()
        self.gpu = Engine(self.layout, self.params, state, device=gpu_device)
        self.cpu = Engine(self.layout, self.params, state, device=oracle.HostDevice(precision))
        self.engines = (self.gpu, self.cpu)
        self.now = 0
        for e in self.engines:
            e.reset(0)
            e.compute_cfl()

    def forcing_at(self, t: float):
        rain = self.forcing.rain
        rates = rain.rates(rain.index_at(t)).astype(np.float32)
        coeff = coefficients_at(self.forcing, t)
        for e in self.engines:
            e.set_forcing(coeff.ravel()[:2 * max(1, len(self.forcing.boundaries))] if self.forcing.boundaries
                          else np.zeros(0, np.float32), rates, True)

    def fixed_step(self, dt: float):
        """One step of the same dt in both engines; returns both traces.

        The step is dt, or half the smaller CFL candidate when that is shorter, and is forced by
        placing the batch target exactly dt ahead (remaining time then decides the step).
        """
        t0 = self.now / 2 ** 40
        self.forcing_at(t0)
        limit = min(e.cfl_state()[0] for e in self.engines)
        step = min(dt, 0.5 * limit)
        target = seconds_to_ticks(t0 + step)
        rows = []
        for e in self.engines:
            e.begin(target, self.now)
            rows.append(e.run_batch(1))
        if rows[0][0, 3] == abi.COMPLETED:
            self.now = (int(rows[0][0, 1]) << 32) | int(rows[0][0, 2])
        return rows

    def adaptive_batch(self, target_seconds: float, steps: int = 32, engines=None):
        """One batch towards a target; the clock follows the first engine's trace."""
        engines = self.engines if engines is None else engines
        self.forcing_at(self.now / 2 ** 40)
        target = seconds_to_ticks(target_seconds)
        rows = []
        for e in engines:
            e.begin(target, self.now)
            rows.append(e.run_batch(steps))
        done = rows[0][rows[0][:, 3] == abi.COMPLETED]
        if len(done):
            self.now = (int(done[-1, 1]) << 32) | int(done[-1, 2])
        return rows

    def diff(self, name):
        a, b = self.gpu.read_state(name), self.cpu.read_state(name)
        return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) if len(a) else 0.
