"""The batched and union forms at the sizes the benchmark times, against each run alone (needs an
MI355X).

The small-mesh batch and union tests pin every kernel path; these two repeat the comparison on
the benchmark's meshes for two minutes of simulated time, so the large grids, the 64-slot wide
kernels over a 1 M-cell mesh and parts of up to 3.5 M cells are covered too. Each scenario or
mesh must take the dt sequence it takes alone and end in the same state bit for bit.
"""
import numpy as np
import pytest

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.batch import BatchEngine, BatchRunner
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout
from nh_flood_2d_amd.union import UnionEngine, UnionRunner
from test_gpu_batch import _compare, _run_alone, _say

pytestmark = pytest.mark.gpu


def test_batch_on_the_benchmark_mesh(gpu_device):
    # the batched benchmark's mesh: 790 x 790 blocks of 10 m, two levels, 998,560 cells
    base = synth.two_level_fractal(0, blocks=790, duration=120.)
    variants = synth.scenario_variants(base, 6, seed=1)
    _say(f'[fullsize batch] {base.mesh.e_num - 1} cells, {len(variants)} scenarios in 64 slots; alone runs')
    alone = [_run_alone(v, gpu_device, 60.) for v in variants]
    _say('[fullsize batch] batched run')
    batch = BatchEngine(64, device=gpu_device)
    sims, states = {}, {}
    runner = BatchRunner(batch, batch_size=32)
    for i, v in enumerate(variants):
        def make(factory, i=i, v=v):
            sims[i] = synth.simulation(v, engine_factory=factory, yield_step=60.)
            return sims[i]
        runner.submit(i, make)
    while runner.step():
        pass
    for i in sims:
        states[i] = sims[i].state()
    _compare(variants, alone, runner.results, sims)
    for i, (_, _, state, _) in enumerate(alone):
        for name, values in state.items():
            np.testing.assert_array_equal(states[i][name], values, err_msg=f'scenario {i} {name}')
    batch.close()


def test_union_of_benchmark_sized_meshes(gpu_device):
    # meshes of the uneven-size benchmark's range: 1, 2 and 3 levels, 60 k to 3.5 M cells
    scenarios = [synth.uneven_scenario(seed, cells, levels, classes, duration=120.)
                 for seed, cells, levels, classes in ((21, 60_000, 1, 2), (22, 700_000, 2, 3), (23, 3_500_000, 3, 4))]
    _say(f'[fullsize union] cells {[sc.mesh.e_num - 1 for sc in scenarios]}; alone runs')

    def plain(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
    alone = []
    for sc in scenarios:
        sim = synth.simulation(sc, engine_factory=plain, yield_step=60.)
        sim.run()
        alone.append((sim.state(), sim.record['dt_sequence_sha256']))
        sim.engine.close()
    _say('[fullsize union] lockstep run')
    layouts = [build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)),
                            sc.sita) for sc in scenarios]
    sc0 = scenarios[0]
    union = UnionEngine(layouts, make_params(sc0.afa, sc0.sita, sc0.min_h, 32), device=gpu_device)
    union.set_tuning(128, 2)        # the shape the uneven benchmark's tuner picks
    union.set_graph(True)
    sims = {p: synth.simulation(sc, engine_factory=union.view(p), yield_step=60., base_layout=layouts[p])
            for p, sc in enumerate(scenarios)}
    UnionRunner(union, sims).run()
    for p, (state, sha) in enumerate(alone):
        assert sims[p].record['dt_sequence_sha256'] == sha, f'part {p}'
        got = sims[p].state()
        for name, values in state.items():
            assert got[name].tobytes() == values.tobytes(), f'part {p} {name}'
    union.close()
