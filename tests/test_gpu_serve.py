"""Serving loop on the GPU: uneven solves overlapped on per-solve streams, or grouped into
unions, each the same as alone.

Plain-engine solves must match their run alone bit for bit, mass balance included; solves that
share a mesh group run in one BatchEngine, and with a union window the others run as parts of one
UnionEngine: both match in state and dt sequence, with budget rows equal to 1e-9 relative.
"""
import math

import numpy as np
import pytest

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.engine import Engine
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout
from nh_flood_2d_amd.serve import FloodServer, FrameStream, Request, footprint_bytes

pytestmark = pytest.mark.gpu


def _say(text):
    print(text, flush=True)


@pytest.mark.parametrize('union_window', [0.0, 10.0])
def test_served_mix_matches_each_solve_alone(gpu_device, union_window):
    scenarios, groups = [], []
    for i, (blocks, duration) in enumerate([(40, 600.), (16, 300.), (64, 420.), (24, 540.)]):
        base = synth.two_level_fractal(40 + i, blocks=blocks, duration=duration)
        scenarios.append(synth.scenario_variants(base, 1, seed=50 + i)[0])
        groups.append(None)
    shared = synth.small_multires(7, duration=360., discharge=1.0)
    for v in synth.scenario_variants(shared, 3, seed=60):
        scenarios.append(v)
        groups.append('shared')
    _say(f'[serve] {len(scenarios)} solves, cells {[s.mesh.e_num - 1 for s in scenarios]}: alone runs')

    def factory(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
    alone = []
    for sc in scenarios:
        sim = synth.simulation(sc, engine_factory=factory, yield_step=60.)
        row = sim.run()
        alone.append((row, sim.state(), sim.record['dt_sequence_sha256']))
        sim.engine.close()
    _say('[serve] served run')
    stream = FrameStream(None)
    finals = {}
    requests = []
    for i, (sc, g) in enumerate(zip(scenarios, groups)):
        layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)),
                              sc.sita)

        class Capture:
            def __init__(self, key):
                self.key, self.inner = key, stream.sink(key)

            def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
                self.inner.emit(sim, seconds, phase, frame, final, depth, row, stats)
                if final:
                    finals[self.key] = (sim.state(), sim.record)

            def finish(self, sim):
                finals[self.key] = (finals[self.key][0], dict(sim.record))

        def build(engine_factory, sc=sc, layout=layout, i=i):
            return synth.simulation(sc, engine_factory=engine_factory, yield_step=60., base_layout=layout,
                                    outputs=Capture(i))
        requests.append(Request(i, build, 0.02 * i, sc.mesh.e_num - 1, footprint_bytes(layout), group=g,
                                layout=layout, params=(sc.afa, sc.sita, sc.min_h)))
    server = FloodServer(batch_size=32, group_slots=4, union_window=union_window)
    report = server.serve(requests, realtime=True)
    stream.close()
    _say(f'[serve] {report.summary()}')
    assert report.peak_live >= 2
    assert report.unions == (1 if union_window else 0)
    for i, (row, state, sha) in enumerate(alone):
        got_state, record = finals[i]
        assert record['dt_sequence_sha256'] == sha, f'solve {i}'
        for name, values in state.items():
            np.testing.assert_array_equal(got_state[name], values, err_msg=f'solve {i} {name}')
        if groups[i] is None and not union_window:
            assert report.results[i] == row
        else:
            for k, x in row.items():
                y = report.results[i][k]
                if isinstance(x, float) and (math.isfinite(x) or math.isfinite(y)):
                    assert abs(x - y) <= 1e-9 + 1e-9 * max(abs(x), abs(y)), (i, k, x, y)
                else:
                    assert x == y
