"""Gate control through the file-based solver, the CPU oracle as the device step (no GPU).

Restates the mechanism tests of the reference's ``tests/test_gate_solver.py`` (setups of
:22-58, known answers of :80-221) and layout checks of ``tests/test_gate_config.py``: the
trigger is strict and needs a wet control cell, commands latch and never reclose, control waits
for the formal start, the gate bed moves without moving water, exchanges close the pond and
channel budgets, and a run split at a checkpoint continues bit for bit. The gate-bed timeline
replay (``read_bed_at``; here ``output.BedTimeline``) is tested with the flood maps
(``tests/test_maps_host.py``).
"""
import csv
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.config import DomainConfig, ForceConfig, GateConfig
from nh_flood_2d_amd.gate import load_gate_layout
from nh_flood_2d_amd.output import read_checkpoint, write_checkpoint
import cases


@pytest.fixture(autouse=True)
def oracle_engine(monkeypatch):
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))


def setup(path, *, cells=None, types=None, bed=None, depth=None, gate=None, warmup=0, yield_step=1,
          boundaries=None):
    """Three 10 m cells in a row by default: control (type 1), gate (11), pond (10)."""
    path.mkdir(parents=True, exist_ok=True)
    cells = cells or [(i * 10, (i + 1) * 10, 0, 10) for i in range(3)]
    types = types or [1, 11, 10]
    bed = np.r_[0, bed if bed is not None else [0] * len(cells)].astype(np.float32)
    depth = np.r_[0, depth if depth is not None else [2] * len(cells)].astype(np.float32)
    mesh = cases.boxes_mesh(cells, bed=list(bed[1:]), land=np.array(types, np.uint8))
    cases.write_mesh(path, mesh)
    centre = lambda i: [(cells[i][0] + cells[i][1]) / 2, (cells[i][2] + cells[i][3]) / 2]
    config = dict(land_type=11, pond_land_type=10, enabled=True, trigger_level=1, opening_seconds=1,
                  open_elevation=-.1, gates=[dict(name='A', inside_point=centre(1), control_point=centre(0))])
    config.update(gate or {})
    x1, y1 = max(c[1] for c in cells), max(c[3] for c in cells)
    rain = cases.write_rain(path, np.zeros((1, 1, 1), np.float32), 0., y1, max(x1, y1), [(0, 3600)])
    force = ForceConfig(boundaries=boundaries or [], gate=config, rainfall=rain)
    zeros, words = np.zeros(len(depth), np.float32), np.zeros(len(depth), np.uint32)
    write_checkpoint(path / 'seed.fdb', depth=depth, bed=bed, q=np.zeros(mesh.s_num, np.float32),
                     rain_time=zeros, infiltration=zeros, max_depth=depth, max_speed=zeros, first_time_hi=words,
                     first_time_lo=words, clock_hi=0, clock_lo=0, gate_open=False,
                     metadata_json=json.dumps(dict(run_state=False, ne_sha256=solver.sha256(path / 'ne.txt'),
                                                   ns_sha256=solver.sha256(path / 'ns.txt'),
                                                   gate_config=force.gate.model_dump_json())))
    domain = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=cases.EPSG,
                          domain_dir=str(path / 'out'), start=cases.START, duration=3600, warmup_duration=warmup,
                          yield_step=yield_step, restart_state=str(path / 'seed.fdb'),
                          land_parameters={t: dict(manning_n=0, infiltration='none') for t in set(types)})
    return domain, force, mesh, depth, bed


def run(domain, force, seconds, *, output=None, restart=None):
    seconds = round(seconds * 2 ** 40) / 2 ** 40
    changes = {}
    if output is not None:
        changes['domain_dir'] = str(output)
    if restart is not None:
        changes['restart_state'] = str(restart)
    domain = domain.model_copy(update=changes)
    Path(domain.domain_dir).mkdir(parents=True, exist_ok=True)
    with solver.RunRecord(domain, force) as record:
        record.hash_inputs(domain, force)
        solver.preprocess(domain, force)
        solver._solve(domain, force, record, end_time=seconds)
    root = Path(domain.domain_dir)
    return read_checkpoint(root / 'output/state_final.fdb'), json.loads((root / 'manifest.json').read_text())


def level_boundary(path, rows):
    return [cases.write_boundary(path, 'river', 'level', [[0, 0], [0, 10]], rows)]


def test_local_bed_motion_preserves_water_and_drives_expected_exchange(tmp_path):
    domain, force, mesh, *_ = setup(tmp_path, gate=dict(opening_seconds=.1))
    # the first step moves the bed after its hydraulic substep; the next uses it
    run(domain, force, .1)
    state, manifest = run(domain, force, .2, output=tmp_path / 'next', restart=tmp_path / 'out/output/state_final.fdb')
    internal = np.flatnonzero((mesh.topo[:, 1] > 0) & (mesh.topo[:, 2] > 0))
    assert state.q[internal[0]] * 10 == pytest.approx(.1962, rel=2e-5)
    assert np.sum(state.depth) * 100 == pytest.approx(600, abs=3e-5)
    assert state.depth[2] > state.depth[1]
    assert manifest['gate_control']['gates'][0]['command_seconds'] == 0


@pytest.mark.parametrize('control_depth,threshold,expected', [(2, 1, 0), (2, 2, None), (0, -1, None),
                                                              (.001, -1, None), (2, 3, None)])
def test_strict_threshold_and_dry_control(tmp_path, control_depth, threshold, expected):
    domain, force, *_ = setup(tmp_path, depth=[control_depth, 0, 0], bed=[0, 10, 0], gate=dict(trigger_level=threshold))
    _, manifest = run(domain, force, .1)
    assert manifest['gate_control']['gates'][0]['command_seconds'] == expected


def test_end_bed_and_pond_exchange_accounting(tmp_path):
    domain, force, *_ = setup(tmp_path, depth=[2, 2, 1])
    state, manifest = run(domain, force, 3)
    gate = manifest['gate_control']['gates'][0]
    pond = manifest['gate_control']['pond_budget']
    assert state.bed[2] == pytest.approx(-.1)
    assert gate['pond_in_m3'] > 0
    assert gate['first_effective_flow_seconds'] is not None
    assert gate['fully_open_seconds'] == 1
    assert abs(pond['residual_m3']) < 3e-5
    assert pond['storage_m3'] - pond['initial_storage_m3'] == pytest.approx(gate['pond_net_m3'] + pond['rounding_m3'],
                                                                             abs=3e-5)
    assert abs(gate['channel_budget']['residual_m3']) < 3e-5
    rows = list(csv.DictReader(open(tmp_path / 'out/gate_balance.csv')))
    assert [r['region'] for r in rows[-2:]] == ['A', 'pond_control_volume']


@pytest.mark.parametrize('yield_step', [1, 2])
def test_short_pulse_latches_each_step_and_never_recloses(tmp_path, yield_step):
    boundaries = level_boundary(tmp_path, [(0, 1), (.125, 2), (.25, 1), (3600, 1)])
    domain, force, *_ = setup(tmp_path, bed=[0, 10, 0], depth=[1, 0, 0],
                              gate=dict(trigger_level=1.5, opening_seconds=1.875, open_elevation=0),
                              boundaries=boundaries, yield_step=yield_step)
    state, manifest = run(domain, force, 2)
    gate = manifest['gate_control']['gates'][0]
    assert gate['command_seconds'] == .125
    assert gate['fully_open_seconds'] == 2
    assert gate['first_fully_open_sample_seconds'] == 2
    assert state.bed[2] == 0
    assert abs(manifest['gate_control']['pond_budget']['residual_m3']) < 1e-5


def test_gates_trigger_independently(tmp_path):
    domain, force, *_ = setup(tmp_path, cells=[(i * 10, (i + 1) * 10, 0, 10) for i in range(6)],
                              types=[1, 11, 10, 1, 11, 10], depth=[2, 0, 0, 0, 0, 0], bed=[0, 10, 0, 0, 10, 0],
                              gate=dict(gates=[dict(name='A', inside_point=(15, 5), control_point=(5, 5)),
                                               dict(name='B', inside_point=(45, 5), control_point=(35, 5))]))
    state, manifest = run(domain, force, 1)
    assert [g['command_seconds'] for g in manifest['gate_control']['gates']] == [0, None]
    assert state.bed[2] < 10 and state.bed[5] == 10


def test_warmup_disables_control_until_formal_zero(tmp_path):
    domain, force, *_ = setup(tmp_path, warmup=2, depth=[2, 0, 0], bed=[0, 10, 0])
    _, warming = run(domain, force, 0)
    assert warming['gate_control']['gates'][0]['command_seconds'] is None
    state, formal = run(domain, force, 1, output=tmp_path / 'formal', restart=tmp_path / 'out/output/state_final.fdb')
    assert formal['gate_control']['gates'][0]['command_seconds'] == 0
    assert state.bed[2] == np.float32(-.1)


@pytest.mark.parametrize('moving', [False, True])
def test_fixed_still_water_and_uniform_motion_have_zero_flux(tmp_path, moving):
    domain, force, _, depth, _ = setup(tmp_path, cells=[(100, 110, 0, 10), (10, 20, 0, 10), (20, 30, 0, 10),
                                                        (30, 40, 0, 10), (0, 10, 0, 10)],
                                       types=[1, 11, 11, 10, 1], bed=[0, 0, 0, 10, 10], depth=[2, 2, 2, 0, 0],
                                       gate=dict(trigger_level=1 if moving else 3))
    state, _ = run(domain, force, 2)
    np.testing.assert_array_equal(state.depth, depth)
    assert np.all(state.q == 0)
    np.testing.assert_array_equal(state.bed[2:4], np.full(2, -.1 if moving else 0, dtype=np.float32))


@pytest.mark.parametrize('split_at', [1, 3, 5])
def test_restart_before_during_after_opening_is_bitwise(tmp_path, split_at):
    boundaries = level_boundary(tmp_path, [(0, 1), (1, 1), (2, 2), (3600, 2)])
    domain, force, *_ = setup(tmp_path, boundaries=boundaries, depth=[1, 0, 0], bed=[0, 3, 0],
                              gate=dict(trigger_level=1.5, opening_seconds=2))
    full, full_manifest = run(domain, force, 6)
    run(domain, force, split_at, output=tmp_path / 'part')
    resumed, resumed_manifest = run(domain, force, 6, output=tmp_path / 'resumed',
                                    restart=tmp_path / 'part/output/state_final.fdb')
    for name in ('depth', 'bed', 'q', 'rain_time', 'infiltration', 'max_depth', 'max_speed', 'first_time_hi',
                 'first_time_lo'):
        assert getattr(full, name).tobytes() == getattr(resumed, name).tobytes(), name
    assert full.metadata_json == resumed.metadata_json
    assert full_manifest['gate_control'] == resumed_manifest['gate_control']

    def rows(path):
        with path.open() as stream:
            return [r for r in csv.DictReader(stream) if float(r['elapsed_seconds']) > split_at]
    assert rows(tmp_path / 'out/gate_balance.csv') == rows(tmp_path / 'resumed/gate_balance.csv')


@pytest.mark.parametrize('pond_depth,gate_depth', [(0, 2), (2, .021), (1, 1.00001), (2, 1)])
def test_dry_near_dry_small_and_reverse_heads(tmp_path, pond_depth, gate_depth):
    domain, force, *_ = setup(tmp_path, depth=[gate_depth, gate_depth, pond_depth],
                              gate=dict(trigger_level=-1, open_elevation=0))
    state, manifest = run(domain, force, 2)
    gate = manifest['gate_control']['gates'][0]
    assert np.all(state.depth >= 0)
    assert abs(manifest['gate_control']['pond_budget']['residual_m3']) < 4e-5
    if gate_depth > pond_depth:
        assert gate['pond_in_m3'] > 0
    else:
        assert gate['pond_out_m3'] > 0


# ------------------------------------------------------------------------------ layout checks
# (tests/test_gate_config.py)

def _config(**changes):
    base = dict(land_type=11, pond_land_type=10, enabled=True, trigger_level=1, opening_seconds=1, open_elevation=-.1,
                gates=[dict(name='A', inside_point=(15, 5), control_point=(5, 5))])
    base.update(changes)
    return GateConfig(**base)


def _mesh(types, bed=None):
    cells = [(i * 10, (i + 1) * 10, 0, 10) for i in range(len(types))]
    return cases.boxes_mesh(cells, bed=bed or [0.] * len(types), land=np.array(types, np.uint8))


@pytest.mark.parametrize('changes', [dict(opening_seconds=0), dict(pond_land_type=11), dict(gates=[]),
                                     dict(gates=[dict(name='A', inside_point=(1, 1), control_point=(2, 2))] * 2)])
def test_invalid_config(changes):
    with pytest.raises(ValueError):
        _config(**changes)


def test_component_partition_and_controls():
    mesh = _mesh([1, 11, 11, 10, 1, 11, 10])
    config = _config(gates=[dict(name='A', inside_point=(25, 5), control_point=(5, 5)),
                            dict(name='B', inside_point=(55, 5), control_point=(45, 5))])
    layout = load_gate_layout(config, mesh, [], mesh.z)
    assert [c.tolist() for c in layout.cells] == [[2, 3], [6]]
    assert layout.controls.tolist() == [1, 5]
    assert layout.cell_gate.tolist() == [-1, -1, 0, 0, -1, -1, 1, -1]


@pytest.mark.parametrize('types,config,error', [
    ([1, 11, 1], {}, 'selects no pond'),
    ([1, 11, 10], dict(gates=[dict(name='A', inside_point=(5, 5), control_point=(25, 5))]), 'not on a gate cell'),
    ([1, 11, 10], dict(gates=[dict(name='A', inside_point=(15, 5), control_point=(15, 5))]), 'control_point is on'),
    ([1, 11, 10], dict(gates=[dict(name='A', inside_point=(15, 5), control_point=(500, 5))]), 'outside the mesh'),
    ([1, 11, 10, 11, 1], {}, 'exactly one gate'),
    ([10, 11, 10], dict(gates=[dict(name='A', inside_point=(15, 5), control_point=(5, 5))]), 'non-pond non-gate'),
])
def test_layout_rejections(types, config, error):
    mesh = _mesh(types)
    with pytest.raises(ValueError, match=error):
        load_gate_layout(_config(**config), mesh, [], mesh.z)


def test_equal_f32_closed_target_is_allowed_but_next_higher_target_is_not():
    bed = np.float32(.1)
    mesh = _mesh([1, 11, 10], bed=[0., float(bed), 0.])
    load_gate_layout(_config(open_elevation=float(bed)), mesh, [], mesh.z)
    with pytest.raises(ValueError, match='exceeds a closed gate bed'):
        load_gate_layout(_config(open_elevation=float(np.nextafter(bed, np.float32(1)))), mesh, [], mesh.z)


# ------------------------------------------------------------------------------ a larger case

def channel_and_pond(path, nx=40, ny=20, dx=5.):
    """River cells (type 1) with a rising tide on the west edge, a one-cell-wide gate column
    (type 11, closed at 1 m, opening to 0 m over 30 s once the control level passes 1.5 m), and
    a pond (type 10, Horton infiltration, 540 cells) east of it; rain on everything."""
    from nh_flood_2d_amd import synth
    path.mkdir(parents=True, exist_ok=True)
    col = np.tile(np.arange(nx), ny)
    land = np.where(col < 12, 1, np.where(col == 12, 11, 10)).astype(np.uint8)
    bed = np.where(land == 11, 1., np.where(land == 10, .5, 0.)).astype(np.float32)
    mesh = synth.uniform_grid(nx, ny, dx, bed, land)
    cases.write_mesh(path, mesh)
    depth = np.r_[0., np.where(land == 1, 1.2, np.where(land == 10, .2, 0.))].astype(np.float32)
    rain = cases.write_rain(path, np.full((1, 1, 1), 30., np.float32), 0., ny * dx, max(nx, ny) * dx, [(0, 3600)])
    tide = cases.write_boundary(path, 'sea', 'level', [[0, 0], [0, ny * dx]], [(0, 1.2), (100, 2.0), (3600, 2.0)])
    gate = dict(land_type=11, pond_land_type=10, enabled=True, trigger_level=1.5, opening_seconds=30,
                open_elevation=0., gates=[dict(name='G', inside_point=(12.5 * dx, 10.5 * dx), control_point=(6 * dx, 10.5 * dx))])
    force = ForceConfig(boundaries=[tide], gate=gate, rainfall=rain)
    zeros, words = np.zeros(mesh.e_num, np.float32), np.zeros(mesh.e_num, np.uint32)
    write_checkpoint(path / 'seed.fdb', depth=depth, bed=np.r_[0., bed].astype(np.float32),
                     q=np.zeros(mesh.s_num, np.float32), rain_time=zeros, infiltration=zeros, max_depth=depth,
                     max_speed=zeros, first_time_hi=words, first_time_lo=words, clock_hi=0, clock_lo=0,
                     gate_open=False,
                     metadata_json=json.dumps(dict(run_state=False, ne_sha256=solver.sha256(path / 'ne.txt'),
                                                   ns_sha256=solver.sha256(path / 'ns.txt'),
                                                   gate_config=force.gate.model_dump_json())))
    land_parameters = {1: dict(manning_n=.03, infiltration='none'), 11: dict(manning_n=.02, infiltration='none'),
                       10: dict(manning_n=.05, infiltration=dict(f0_mm_h=60, fc_mm_h=10, k_per_h=2))}
    domain = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=cases.EPSG,
                          domain_dir=str(path / 'out'), start=cases.START, duration=3600, warmup_duration=0,
                          yield_step=40, restart_state=str(path / 'seed.fdb'), land_parameters=land_parameters)
    return domain, force


def test_channel_and_pond_budgets_close(tmp_path):
    domain, force = channel_and_pond(tmp_path)
    state, manifest = run(domain, force, 160)
    report = manifest['gate_control']
    gate, pond = report['gates'][0], report['pond_budget']
    assert 20 < gate['command_seconds'] < 100 and gate['fully_open_seconds'] == gate['command_seconds'] + 30
    assert gate['river_in_m3'] > 0 and gate['pond_in_m3'] > 0
    assert pond['rain_m3'] > 0 and pond['infiltration_m3'] > 0
    assert abs(pond['residual_m3']) < 1e-6 * max(1., pond['storage_m3'])
    assert abs(gate['channel_budget']['residual_m3']) < 1e-6 * max(1., gate['channel_storage_m3'])
