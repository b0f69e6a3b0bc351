"""Meshes of different sizes advanced by the same launches (nhu_*), on the GPU.

Each Simulation bound to a part of a UnionEngine, run in lockstep with the others, must take the
dt sequence it takes alone on a single-mesh engine and end in the same state bit for bit; its
mass-balance row agrees to 1e-9 relative (the budget's FP64 sums are grouped the same way, but the
test does not rely on it).
"""
import math

import numpy as np
import pytest

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout
from nh_flood_2d_amd.union import UnionEngine, UnionRunner

pytestmark = pytest.mark.gpu


def _say(text):
    print(text, flush=True)


def _scenarios(theta: float = 1.0):
    if theta != 1.0:      # the face-average pass and open-boundary fractions
        return [synth.small_multires(1, blocks=20, pond=0.3, duration=240., theta=theta),
                synth.small_multires(2, blocks=12, pond=0.3, discharge=0.5, duration=300., theta=theta)]
    out = [synth.small_multires(1, blocks=20, pond=0.3, duration=300.),
           synth.scenario_variants(synth.two_level_fractal(41, blocks=16, duration=240.), 1, seed=5)[0],
           synth.small_multires(2, blocks=12, pond=0.3, discharge=0.5, duration=360.)]
    uniform = synth.plain_uniform(3, n=48, dx=10.)
    uniform.duration = 200.
    out.insert(1, uniform)
    return out


@pytest.mark.parametrize('cell_shape,theta', [((256, 1), 1.0), ((128, 2), 1.0), ((256, 1), 0.7)])
def test_union_parts_match_each_mesh_alone(gpu_device, cell_shape, theta):
    scenarios = _scenarios(theta)
    _say(f'[union] {len(scenarios)} meshes, cells {[s.mesh.e_num - 1 for s in scenarios]}, shape {cell_shape}, '
         f'sita {theta}')

    def plain(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
    alone = []
    for sc in scenarios:
        sim = synth.simulation(sc, engine_factory=plain, yield_step=60.)
        row = sim.run()
        alone.append((row, sim.state(), sim.record['dt_sequence_sha256']))
        sim.engine.close()
    _say('[union] lockstep run')
    layouts = [build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)),
                            sc.sita) for sc in scenarios]
    sc0 = scenarios[0]
    union = UnionEngine(layouts, make_params(sc0.afa, sc0.sita, sc0.min_h, 32), device=gpu_device)
    union.set_tuning(*cell_shape)
    union.set_graph(True)
    sims = {p: synth.simulation(sc, engine_factory=union.view(p), yield_step=60., base_layout=layouts[p])
            for p, sc in enumerate(scenarios)}
    results = UnionRunner(union, sims).run()
    for p, (row, state, sha) in enumerate(alone):
        assert sims[p].record['dt_sequence_sha256'] == sha, f'part {p}'
        got = sims[p].state()
        for name, values in state.items():
            assert got[name].tobytes() == values.tobytes(), f'part {p} {name}'
        for key, value in row.items():
            other = results[p][key]
            if isinstance(value, float) and math.isfinite(value):
                assert abs(other - value) <= 1e-9 * max(1.0, abs(value)), (p, key, value, other)
            else:
                assert other == value, (p, key)
    _say(f'[union] {union.batches} union batches for {len(scenarios)} meshes')
    union.close()
