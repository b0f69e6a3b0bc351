"""nhf_run_static, the entry point that allocates nothing, against the context entry points it
stands for (needs an MI355X).

Each call of nhf_run_static on a mesh must give what nhf_reset, nhf_set_forcing, nhf_compute_cfl,
nhf_begin, nhf_run_batch, nhf_flush_budget and nhf_settle_extrema give on a context with its own
workspace: the same trace rows, the same state bits and the same budget pairs. The calls run in a
row through the one static scratch: a 2:1 mesh at sita 0.7 (face averages, hanging gradients, a
tide and a river boundary), the same mesh again from where the first call ended, then a smaller
sita 1 mesh, then the first mesh once more, so every call starts from another call's leftovers.
"""
import ctypes

import numpy as np
import pytest

from nh_flood_2d_amd import abi, synth
from nh_flood_2d_amd.clock import seconds_to_ticks
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, initial_state

from parity import coefficients_at

pytestmark = pytest.mark.gpu

STATE = ('depth', 'q', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick')


class Twin:
    """One scenario bound twice: `ctx` steps through a context, `bare` only lends its arrays to
    nhf_run_static."""

    def __init__(self, sc, device):
        self.sc = sc
        self.forcing = synth.scenario_forcing(sc)
        self.layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, self.forcing), sc.sita)
        self.params = make_params(sc.afa, sc.sita, sc.min_h, 32)
        state = initial_state(sc.mesh, self.layout)
        if sc.initial_depth is not None:
            state['depth'] = sc.initial_depth.astype(np.float32).copy()
        self.device = device
        self.ctx = Engine(self.layout, self.params, state, device=device)
        self.bare = Engine(self.layout, self.params, state, device=device)
        # nhf_run_static keeps the rain time and capacity per cell in the bare arrays: the bare
        # context must not hold a uniform rain state over them (its settle would write it out)
        abi.check(device.lib, device.lib.nhf_set_rain_state(self.bare.ctx, 0, ctypes.c_float(0.), None),
                  'nhf_set_rain_state')
        self.now = 0

    def forcing_at(self, t):
        rain = self.forcing.rain
        rates = rain.rates(rain.index_at(t)).astype(np.float32)
        b = len(self.forcing.boundaries)
        coeff = coefficients_at(self.forcing, t).ravel()[:2 * b].copy() if b else np.zeros(0, np.float32)
        return np.ascontiguousarray(coeff, np.float32), np.ascontiguousarray(rates, np.float32)

    def call(self, seconds, steps):
        t0 = self.now / 2 ** 40
        coeff, rates = self.forcing_at(t0)
        target = seconds_to_ticks(t0 + seconds)
        e = self.ctx
        e.reset(self.now)
        e.set_forcing(coeff, rates, True)
        e.compute_cfl()
        e.begin(target, 0)
        rows_ctx = e.run_batch(steps)
        budget_ctx = e.flush_budget()
        trace = np.zeros((steps, abi.TRACE_COLS), np.uint32)
        budget = np.zeros(2 * (4 + 6 * self.layout.b_num), np.float64)
        lib = self.device.lib
        rc = lib.nhf_run_static(ctypes.byref(self.bare.mesh_struct), ctypes.byref(self.params), self.now, target, 0,
                                coeff.ctypes.data if coeff.size else None, rates.ctypes.data, 1, steps,
                                trace.ctypes.data, budget.ctypes.data, self.device.stream_handle())
        abi.check(lib, rc, 'nhf_run_static')
        assert rc == 0
        np.testing.assert_array_equal(trace, rows_ctx)
        np.testing.assert_array_equal(budget.view(np.uint64), budget_ctx.ravel().view(np.uint64))
        for name in STATE:
            a, b = self.ctx.read_state(name), self.bare.read_state(name)
            np.testing.assert_array_equal(a.view(np.uint64 if name == 'peak_tick' else np.uint32),
                                          b.view(np.uint64 if name == 'peak_tick' else np.uint32), err_msg=name)
        done = rows_ctx[rows_ctx[:, 3] == abi.COMPLETED]
        assert len(done) > 0
        self.now = (int(done[-1, 1]) << 32) | int(done[-1, 2])
        return len(done)


def test_static_entry_point_matches_the_context_entry_points(gpu_device):
    big = Twin(synth.small_multires(2, pond=0.3, discharge=2.0, theta=0.7), gpu_device)
    small = Twin(synth.small_multires(1, blocks=12, pond=0.3, discharge=1.0), gpu_device)
    assert small.sc.mesh.e_num < big.sc.mesh.e_num
    for twin, seconds, steps in ((big, 3., 24), (big, 3., 32), (small, 2., 16), (big, 2., 20)):
        print(f'[no-alloc] {twin.sc.mesh.e_num - 1} cells, sita {twin.sc.sita}, {steps} steps', flush=True)
        assert twin.call(seconds, steps) > 0


def test_static_entry_point_refuses_bad_arguments(gpu_device):
    twin = Twin(synth.small_multires(1, blocks=12, pond=0.3, discharge=1.0), gpu_device)
    lib = gpu_device.lib
    trace = np.zeros((4, abi.TRACE_COLS), np.uint32)
    coeff, rates = twin.forcing_at(0.)
    args = lambda steps, c: (ctypes.byref(twin.bare.mesh_struct), ctypes.byref(twin.params), 0, 1 << 40, 0,
                             c, rates.ctypes.data, 1, steps, trace.ctypes.data, None, gpu_device.stream_handle())
    assert lib.nhf_run_static(*args(0, coeff.ctypes.data)) < 0            # no steps
    assert lib.nhf_run_static(*args(33, coeff.ctypes.data)) < 0           # beyond the batch capacity
    assert lib.nhf_run_static(*args(4, None)) < 0                          # boundaries without coefficients
    assert b'coefficients' in lib.nhf_last_error()
