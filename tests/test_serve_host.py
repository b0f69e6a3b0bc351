"""Serving loop, host side (no GPU): admission by footprint, interleaving, frames off the loop.

The devices are oracle contexts, so every served solve must match its run alone bit for bit.
"""
import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import synth
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout
from nh_flood_2d_amd.output import read_depth_frame
from nh_flood_2d_amd.serve import FloodServer, FrameStream, Request, footprint_bytes


def _scenarios():
    out = []
    for i, (blocks, duration) in enumerate([(12, 360.), (20, 240.), (16, 480.), (10, 300.)]):
        base = synth.two_level_fractal(20 + i, blocks=blocks, duration=duration)
        out.append(synth.scenario_variants(base, 1, seed=30 + i)[0])
    return out


def _request(i, sc, arrival, sink=None):
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)), sc.sita)

    def build(factory, sc=sc, layout=layout):
        return synth.simulation(sc, engine_factory=factory, yield_step=60., base_layout=layout,
                                outputs=sink.sink(i) if sink is not None else None)
    return Request(i, build, arrival, sc.mesh.e_num - 1, footprint_bytes(layout))


def test_served_solves_match_alone_and_respect_the_budget(tmp_path):
    scenarios = _scenarios()
    alone = []
    for sc in scenarios:
        sim = synth.simulation(sc, engine_factory=oracle.engine_factory('f32'), yield_step=60.)
        alone.append((sim.run(), sim.state(), sim.steps))
    stream = FrameStream(tmp_path / 'frames')
    requests = [_request(i, sc, 0.001 * i, stream) for i, sc in enumerate(scenarios)]
    budget = sorted(r.footprint for r in requests)[-1] + sorted(r.footprint for r in requests)[-2]
    server = FloodServer(budget_bytes=budget, batch_size=32, device_factory=lambda: oracle.HostDevice('f32'))
    states = {}
    report = server.serve(requests, realtime=False)
    stream.close()
    assert sorted(report.results) == list(range(len(scenarios)))
    assert report.peak_bytes <= budget and report.peak_live <= 3
    for i, (row, state, steps) in enumerate(alone):
        assert report.results[i] == row
        assert report.steps[i] == steps
    summary = report.summary()
    assert summary['requests'] == len(scenarios) and summary['p99_latency_s'] >= summary['p50_latency_s'] > 0
    assert summary['cell_updates'] == sum(r.cells * report.steps[r.key] for r in requests)
    # frames written off the loop: every 900 s plus start and end; the final frame equals the state
    for i, sc in enumerate(scenarios):
        frames = sorted((tmp_path / 'frames' / str(i)).glob('depth_*.fdb'))
        assert len(frames) == stream.frames[i] >= 2
        last = read_depth_frame(tmp_path / 'frames' / str(i) / f'depth_{int(sc.duration)}.fdb')
        np.testing.assert_array_equal(last, alone[i][1]['depth'])


def test_admission_waits_for_memory():
    scenarios = _scenarios()[:3]
    requests = [_request(i, sc, 0.) for i, sc in enumerate(scenarios)]
    budget = max(r.footprint for r in requests)   # one solve at a time
    server = FloodServer(budget_bytes=budget, batch_size=32, device_factory=lambda: oracle.HostDevice('f32'))
    report = server.serve(requests, realtime=False)
    assert report.peak_live == 1 and len(report.results) == 3
    # each later request queued until the earlier ones had finished
    assert report.queued[2] >= report.queued[1] >= report.queued[0]


def test_footprint_counts_state_and_static_arrays():
    sc = _scenarios()[0]
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)), 1.0)
    one = footprint_bytes(layout)
    static = sum(v.nbytes for k, v in layout.arrays.items() if not k.startswith('_'))
    assert one > static + layout.e_num * 28
    assert footprint_bytes(layout, 4) - one == 3 * (footprint_bytes(layout, 2) - one)


@pytest.mark.parametrize('p', [50, 99])
def test_percentiles(p):
    from nh_flood_2d_amd.serve import ServeReport
    r = ServeReport(latency={i: float(i) for i in range(1, 101)})
    assert r.percentile(p) == pytest.approx(np.percentile(np.arange(1, 101), p))


def test_union_admission_rule():
    """Held requests (layout and parameters given, no mesh group) are admitted together once the
    oldest has waited the window, the group is full, or no more arrivals are due; only requests of
    equal parameters share a union."""
    from nh_flood_2d_amd.serve import FloodServer, Request

    def req(key, arrival, params=(.5, 1., .02), group=None, layout=object()):
        return Request(key, None, arrival, 10, 1, group=group, layout=layout, params=params)
    server = FloodServer(budget_bytes=1 << 30, union_window=2.0, union_parts=3)
    waiting = [req(0, 0.), req(1, .5), req(2, .6, params=(.5, .7, .02)), req(3, .7, group='g')]
    later = [req(9, 5.)]
    assert server._union_due(waiting, later, 1.0, True) == []                 # holding
    assert [r.key for r in server._union_due(waiting, later, 2.0, True)] == [0, 1]
    assert [r.key for r in server._union_due(waiting, [], 1.0, True)] == [0, 1]  # nothing more to come
    full = [req(k, .1 * k) for k in range(5)]
    assert [r.key for r in server._union_due(full, later, .5, True)] == [0, 1, 2]
    assert FloodServer(budget_bytes=1 << 30)._union_due(waiting, later, 9., True) == []   # no window: off
    assert server._union_due([req(5, 0., layout=None)], later, 9., True) == []
