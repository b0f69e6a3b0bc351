"""Scenario batches on the GPU: each scenario alone and in the mix (needs an MI355X).

Every scenario runs once on the single-scenario HIP engine and once inside a batch on the same
mesh (more scenarios than slots, so a slot is reused). The dt sequence, final state and extrema
must be identical bit for bit; the mass-balance rows may differ only by the grouping of the FP64
budget sums (relative 1e-9, far inside the 1e-7 volume-budget tolerance).
"""
import math

import numpy as np
import pytest

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.batch import BatchEngine, BatchRunner

pytestmark = pytest.mark.gpu


def _say(text):
    print(text, flush=True)


def _close_rows(a: dict, b: dict, rel=1e-9, atol=1e-9):
    assert a.keys() == b.keys()
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, float) and isinstance(y, float):
            if math.isfinite(x) or math.isfinite(y):
                assert abs(x - y) <= atol + rel * max(abs(x), abs(y)), (k, x, y)
        else:
            assert x == y, (k, x, y)


def _run_alone(v, device, yield_step):
    from nh_flood_2d_amd.engine import Engine

    def factory(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=device, pipe_q=pipe_q)
    sim = synth.simulation(v, engine_factory=factory, yield_step=yield_step)
    row = sim.run()
    out = row, sim.record['dt_sequence_sha256'], sim.state(), sim.steps
    sim.engine.close()
    return out


def _compare(variants, alone, results, sims):
    for i, (row, sha, state, steps) in enumerate(alone):
        assert sims[i].record['dt_sequence_sha256'] == sha, f'scenario {i}: dt sequence differs'
        assert sims[i].steps == steps
        _close_rows(results[i], row)


# 64 slots take the wide kernels (several sides or cells per wavefront, one wavefront holding
# the 64 slots of each); fewer slots take the narrow ones. The wide cases bind more scenarios than
# slots (every lane busy, slots reused) or few scenarios with boundaries, hanging edges and sita < 1.
@pytest.mark.parametrize('case', ['two_level', 'multires_tide_river', 'multires_theta', 'two_level_wide',
                                  'multires_tide_river_wide', 'multires_theta_wide'])
def test_batch_matches_each_scenario_alone(gpu_device, case):
    if case == 'two_level':
        base = synth.two_level_fractal(5, blocks=48, duration=600.)
        variants, slots = synth.scenario_variants(base, 6, seed=7), 4
    elif case == 'multires_tide_river':
        base = synth.small_multires(3, duration=600., discharge=1.5)
        variants, slots = synth.scenario_variants(base, 3, seed=8), 2
    elif case == 'multires_theta':
        base = synth.small_multires(4, duration=480., theta=0.7, pond=0.2)
        variants, slots = synth.scenario_variants(base, 3, seed=9), 4
    elif case == 'two_level_wide':
        base = synth.two_level_fractal(12, blocks=20, duration=300.)
        variants, slots = synth.scenario_variants(base, 70, seed=13), 64
    elif case == 'multires_tide_river_wide':
        base = synth.small_multires(14, duration=480., discharge=1.5)
        variants, slots = synth.scenario_variants(base, 5, seed=15), 64
    else:
        base = synth.small_multires(16, duration=420., theta=0.7, pond=0.2)
        variants, slots = synth.scenario_variants(base, 5, seed=17), 64
    _say(f'[batch] {case}: {base.mesh.e_num - 1} cells, {len(variants)} scenarios, {slots} slots; alone runs')
    alone = [_run_alone(v, gpu_device, 120.) for v in variants]
    _say(f'[batch] {case}: batched run')
    states = {}
    batch = BatchEngine(slots, device=gpu_device)
    sims = {}
    runner = BatchRunner(batch, batch_size=32)
    for i, v in enumerate(variants):
        def make(factory, i=i, v=v):
            sims[i] = synth.simulation(v, engine_factory=factory, yield_step=120.)
            return sims[i]
        runner.submit(i, make)
    # the wide cases change the grids after the first batch: the smallest (every workgroup strides
    # over many cells and sides) or the largest the context allows; every grid gives the same bits
    regrid = case.endswith('_wide')
    # read each scenario's final state before its slot is reused: run until it leaves
    while runner.step():
        if regrid:
            rng = batch.tuning_range()
            if case == 'two_level_wide':
                batch.set_tuning(8, 8)
            else:
                batch.set_tuning(rng['depth_max'] // 8 * 8, rng['cfl_max'] // 8 * 8)
            regrid = False
        for i, sim in sims.items():
            if i not in states and sim.elapsed >= sim.end_time:
                states[i] = sim.state()
    for i in sims:
        if i not in states:
            states[i] = sims[i].state()
    _compare(variants, alone, runner.results, sims)
    for i, (_, _, state, _) in enumerate(alone):
        for name, values in state.items():
            np.testing.assert_array_equal(states[i][name], values, err_msg=f'{case} scenario {i} {name}')
    batch.close()


def test_batch_graph_replays_match_eager(gpu_device):
    base = synth.small_multires(6, duration=300., discharge=1.0)
    variants = synth.scenario_variants(base, 2, seed=10)
    _say('[batch] graph vs eager, small_multires seed 6')
    out = []
    for graph in (False, True):
        batch = BatchEngine(2, device=gpu_device)
        sims = {}
        runner = BatchRunner(batch, batch_size=16)
        for i, v in enumerate(variants):
            def make(factory, i=i, v=v):
                sims[i] = synth.simulation(v, engine_factory=factory, yield_step=60., batch_size=16)
                sims[i].engine.set_graph(graph)
                return sims[i]
            runner.submit(i, make)
        finals = {}
        while runner.step():
            for i, sim in sims.items():
                if i not in finals and sim.elapsed >= sim.end_time:
                    finals[i] = sim.state()
        out.append((finals, {i: s.record['dt_sequence_sha256'] for i, s in sims.items()}))
        batch.close()
    (fa, sa), (fb, sb) = out
    assert sa == sb
    for i in fa:
        for name in fa[i]:
            np.testing.assert_array_equal(fa[i][name], fb[i][name], err_msg=f'{i} {name}')
