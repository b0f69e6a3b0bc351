"""The uneven-size batch and the serving load at their configured sizes, each result against its
run alone (needs an MI355X).

* Uneven-size batch: the benchmark's 32 meshes (cell counts log-uniform in 50 k-4 M, 1-3 levels,
  2-4 Horton land classes, 34.7 M cells in all) advanced in one union for 120 simulated seconds.
  Every part takes the dt sequence it takes on a single-mesh engine and ends with the same state
  bits.
* Serving load: the first five solves of the benchmark's serving load (Poisson arrivals at
  2 per second, 100 k-8 M-cell two-level meshes, 1-6 h simulated, rain peaks U[20, 150] mm/h,
  tide on some), served with continuous admission and overlapped on the device, each with its
  outputs every 300 s. Every solve gives the same trace, state and mass-balance rows as alone.
"""
import numpy as np
import pytest

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout
from nh_flood_2d_amd.serve import FloodServer, FrameStream, Request, footprint_bytes
from nh_flood_2d_amd.union import UnionEngine, UnionRunner

pytestmark = pytest.mark.gpu


def _say(text):
    print(text, flush=True)


def test_uneven_32_mesh_union_at_size(gpu_device):
    items = synth.uneven_batch(21, 32, 50_000, 4_000_000, workers=16)
    cells = [sc.mesh.e_num - 1 for sc, _ in items]
    _say(f'[union of 32 meshes] {len(items)} meshes, {sum(cells)} cells ({min(cells)} to {max(cells)}); alone runs')

    def plain(layout, params, state, pipe_q=None):
        e = Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
        e.set_graph(True)
        return e
    alone = []
    for sc, layout in items:
        sim = synth.simulation(sc, engine_factory=plain, end_time=120., yield_step=60., base_layout=layout)
        sim.run()
        alone.append((sim.state(), sim.record['dt_sequence_sha256'], sim.steps))
        sim.engine.close()
    _say(f'[union of 32 meshes] alone: steps {[a[2] for a in alone]}; the union')
    sc0 = items[0][0]
    union = UnionEngine([layout for _, layout in items], make_params(sc0.afa, sc0.sita, sc0.min_h, 32),
                        device=gpu_device)
    union.set_tuning(128, 2)        # the shape the uneven benchmark's tuner picks
    union.set_graph(True)
    sims = {p: synth.simulation(sc, engine_factory=union.view(p), end_time=120., yield_step=60., base_layout=layout)
            for p, (sc, layout) in enumerate(items)}
    UnionRunner(union, sims).run()
    for p, (state, sha, steps) in enumerate(alone):
        assert sims[p].steps == steps, f'part {p}'
        assert sims[p].record['dt_sequence_sha256'] == sha, f'part {p}'
        got = sims[p].state()
        for name, values in state.items():
            assert got[name].tobytes() == values.tobytes(), f'part {p} {name}'
    union.close()
    _say(f'[union of 32 meshes] {len(items)} parts match their runs alone bit for bit')


def test_serving_load_at_size(gpu_device):
    load = synth.serving_load(0, 5)
    _say(f'[serving load] {len(load)} solves, cells {[sc.mesh.e_num - 1 for _, sc, _ in load]}, '
         f'hours {[round(sc.duration / 3600, 2) for _, sc, _ in load]}; alone runs')
    assert max(sc.mesh.e_num - 1 for _, sc, _ in load) > 1_000_000
    assert min(sc.duration for _, sc, _ in load) >= 3600.

    def plain(layout, params, state, pipe_q=None):
        e = Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
        e.set_graph(True)
        return e
    layouts, alone = {}, {}
    for i, sc, _ in load:
        layouts[i] = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)),
                                  sc.sita)
        rows = []

        class Rows:
            def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
                rows.append(dict(row))

            def finish(self, sim):
                pass
        sim = synth.simulation(sc, engine_factory=plain, yield_step=300., base_layout=layouts[i], batch_size=64,
                               outputs=Rows())
        sim.run()
        alone[i] = (rows, sim.state(), sim.record['dt_sequence_sha256'], sim.steps)
        sim.engine.close()
    _say(f'[serving load] alone: steps {[alone[i][3] for i in alone]}; served run')
    stream = FrameStream(None, every_output=True)
    served_rows, finals = {}, {}
    requests = []
    for i, sc, arrival in load:
        class Capture:
            def __init__(self, key):
                self.key, self.inner = key, stream.sink(key)
                served_rows[key] = []

            def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
                self.inner.emit(sim, seconds, phase, frame, final, depth, row, stats)
                served_rows[self.key].append(dict(row))
                if final:
                    finals[self.key] = (sim.state(), None, sim.steps)

            def finish(self, sim):
                self.inner.finish(sim)
                state, _, steps = finals[self.key]
                finals[self.key] = (state, sim.record['dt_sequence_sha256'], steps)

        def build(factory, sc=sc, i=i):
            return synth.simulation(sc, engine_factory=factory, yield_step=300., base_layout=layouts[i],
                                    batch_size=64, outputs=Capture(i))
        requests.append(Request(i, build, arrival, sc.mesh.e_num - 1, footprint_bytes(layouts[i])))
    report = FloodServer(batch_size=64).serve(requests, realtime=True)
    stream.close()
    summary = report.summary()
    _say(f"[serving load] served: p50 {summary['p50_latency_s']:.2f} s, p99 {summary['p99_latency_s']:.2f} s, "
         f"{summary['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s, peak live {summary['peak_live']}")
    assert report.peak_live >= 2      # solves overlapped on the device
    for i, (rows, state, sha, steps) in alone.items():
        got_state, got_sha, got_steps = finals[i]
        assert got_steps == steps and got_sha == sha, f'solve {i}'
        for name, values in state.items():
            assert got_state[name].tobytes() == values.tobytes(), f'solve {i} {name}'
        assert served_rows[i] == rows, f'solve {i} mass balance'
