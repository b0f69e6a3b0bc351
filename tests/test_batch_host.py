"""Scenario batches, host side (no GPU): the runner's interleaving and continuous admission.

The batch here is made of independent oracle contexts, so every scenario must reproduce its own
run alone bit for bit: dt sequence, mass-balance rows, final state and extrema.
"""
import numpy as np
import pytest

import oracle
from batch_oracle import OracleBatch
from nh_flood_2d_amd import synth
from nh_flood_2d_amd.batch import BatchRunner, boundary_cells, lanes_for
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, meta_boundary, relayout


@pytest.fixture(scope='module')
def variants():
    base = synth.two_level_fractal(3, blocks=20, duration=480.)
    return synth.scenario_variants(base, 5, seed=4)


def _alone(v):
    sim = synth.simulation(v, engine_factory=oracle.engine_factory('f32'), yield_step=120.)
    row = sim.run()
    return row, sim.record['dt_sequence_sha256'], sim.state(), sim.steps


def test_runner_reproduces_each_scenario_alone_with_slot_reuse(variants):
    alone = [_alone(v) for v in variants]
    assert len({a[1] for a in alone}) == len(variants), 'variants should step differently'
    sims = {}
    runner = BatchRunner(OracleBatch(4), batch_size=32)   # 5 scenarios, 4 slots: one waits
    for i, v in enumerate(variants):
        def make(factory, i=i, v=v):
            sims[i] = synth.simulation(v, engine_factory=factory, yield_step=120.)
            return sims[i]
        runner.submit(i, make)
    results = runner.run()
    assert sorted(results) == list(range(len(variants)))
    for i, (row, sha, state, steps) in enumerate(alone):
        assert results[i] == row
        assert sims[i].record['dt_sequence_sha256'] == sha
        assert sims[i].steps == steps
        got = sims[i].state()
        for name, values in state.items():
            np.testing.assert_array_equal(got[name], values, err_msg=f'scenario {i} {name}')


def test_relayout_changes_only_land_arrays(variants):
    a, b = variants[0], variants[1]
    fa = cell_forcing_arrays(a.mesh, synth.scenario_forcing(a))
    fb = cell_forcing_arrays(b.mesh, synth.scenario_forcing(b))
    la = build_layout(a.mesh, a.land_parameters, fa, 1.0)
    fresh = build_layout(b.mesh, b.land_parameters, fb, 1.0)
    lb = relayout(la, b.mesh, b.land_parameters, fb)
    assert set(lb.arrays) == set(fresh.arrays)
    for name in fresh.arrays:
        np.testing.assert_array_equal(lb.arrays[name], fresh.arrays[name], err_msg=name)
    np.testing.assert_array_equal(lb.cell_f0, fresh.cell_f0)
    assert not np.array_equal(la.arrays['manning'], lb.arrays['manning'])


def test_boundary_cell_list_is_sorted_by_boundary():
    sc = synth.small_multires(2, discharge=1.0)
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)), 1.0)
    bslot, start, n = boundary_cells(layout)
    bid = meta_boundary(layout.arrays['cell_meta'])
    cells = np.argsort(np.where(bslot >= 0, bslot, 1 << 30))[:n]
    assert n == int(np.sum(bid[1:] >= 0)) and n > 0
    assert np.all(np.diff(bid[cells]) >= 0)
    for b in range(layout.b_num):
        assert np.all(bid[cells[start[b]:start[b + 1]]] == b)
    assert start[-1] == n


@pytest.mark.parametrize('count,lanes', [(1, 1), (2, 2), (3, 4), (33, 64), (64, 64)])
def test_lane_count(count, lanes):
    assert lanes_for(count) == lanes


def test_lane_count_limit():
    with pytest.raises(ValueError):
        lanes_for(65)


def test_class_manning_rebuilds_every_side_n(variants):
    """A batch keeps one Manning n per land class and scenario; its flux kernel rebuilds each
    side's n from the classes of the side's two cells (the class bits of cell_meta), with the
    exterior's n zero. That must be map_land_parameters' side n in layout order, bit for bit, on
    every side: here for scenario land tables on a one-class mesh and on a four-class 2:1 mesh
    whose class codes include 0 (so the exterior is told apart by its cell, not its class)."""
    from nh_flood_2d_amd.hydraulics import side_manning
    multi = synth.small_multires(2, discharge=1.0)
    cases = list(variants[:2]) + synth.scenario_variants(multi, 3, seed=6)
    for sc in cases:
        layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)),
                              sc.sita)
        table = layout.arrays['_class_manning']
        assert table.dtype == np.float32 and table.shape == (256,)
        cell_n = table[layout.arrays['cell_meta'] & 255]
        cell_n[0] = 0.
        ends = layout.arrays['side_cells'].reshape(-1, 2)
        assert np.any(ends[1:] == 0, axis=1).any()
        assert side_manning(cell_n, ends).tobytes() == layout.arrays['manning'].tobytes(), sc.name
    classes = np.unique(multi.mesh.land[1:])
    assert 0 in classes and len(classes) == 4
    # scenarios on one mesh differ only in the table
    layouts = [build_layout(v.mesh, v.land_parameters, cell_forcing_arrays(v.mesh, synth.scenario_forcing(v)), v.sita)
               for v in cases[2:]]
    assert not np.array_equal(layouts[0].arrays['_class_manning'], layouts[1].arrays['_class_manning'])
    assert np.array_equal(layouts[0].arrays['cell_meta'], layouts[1].arrays['cell_meta'])
