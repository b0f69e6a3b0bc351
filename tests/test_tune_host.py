"""Autotuner, host side (no GPU): candidate timing loop, choice, cache and size classes.

On the oracle every launch grid is accepted and ignored, so only the search and the bookkeeping
are exercised here; the GPU benchmark runs the same tuner on the HIP engines.
"""
import numpy as np

import oracle
from nh_flood_2d_amd import synth
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, initial_state
from nh_flood_2d_amd.tune import CELL_SHAPES, TuningCache, pick, size_class, tune_engine, tuned_engine_parameters


def _setup():
    sc = synth.small_multires(2, pond=0.3)
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)), 1.0)
    state = initial_state(sc.mesh, layout)
    state['depth'] = sc.initial_depth.copy()
    return sc, layout, state


def test_tuner_times_every_candidate_and_keeps_the_state_untouched(tmp_path):
    sc, layout, state = _setup()
    before = {k: v.copy() for k, v in state.items()}
    cache = TuningCache(tmp_path / 'tuning.json')
    params = make_params(sc.afa, sc.sita, sc.min_h, 64)
    got = tuned_engine_parameters(layout, params, state, oracle.HostDevice('f32'), cache=cache, steps=16)
    assert got['cfl_blocks'] == 8 and got['steps_per_graph'] in (16, 32, 64)
    assert len(got['table']) == len(CELL_SHAPES) + 2 and got['us_per_step'] > 0
    assert (got['cell_block'], got['cells_per_thread']) in CELL_SHAPES
    for k in before:
        np.testing.assert_array_equal(state[k], before[k])
    # cached per size class, also across processes through the JSON file
    assert TuningCache(tmp_path / 'tuning.json').get(size_class(layout)) == got
    assert tuned_engine_parameters(layout, params, state, None, cache=cache) == got


def test_tune_engine_on_a_live_engine_chooses_the_fastest():
    sc, layout, state = _setup()
    eng = Engine(layout, make_params(sc.afa, sc.sita, sc.min_h, 64), state, device=oracle.HostDevice('f32'))
    eng.set_forcing(np.zeros(2 * layout.b_num, np.float32), None, False)
    eng.reset(0)
    eng.compute_cfl()
    result = tune_engine(eng, steps=8, rounds=1, steps_per_graph=(8, 16))
    assert len(result['table']) == len(CELL_SHAPES) + 2
    assert result['us_per_step'] == min(result['table'].values())


def test_pick_and_size_class():
    assert pick({('a', 1): 2.0, ('b', 2): 1.0, ('c', 3): 1.0}) == ('b', 2)
    sc, layout, _ = _setup()
    assert size_class(layout, 4).endswith('-slots4-theta1-hanging')
