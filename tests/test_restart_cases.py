"""Checkpoint / resume: identity checks and bit-for-bit continuation.

Restates ``tests/test_restart.py`` of the reference against this package's solver (CPU oracle
as the device step): a run split at a checkpoint and resumed must reproduce the uninterrupted
run's trace, state, budget rows and metadata exactly; checkpoints from other meshes, physics or
boundary configurations are refused.
"""
import csv
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import simulation
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.engine import Engine
from nh_flood_2d_amd.output import read_checkpoint
import cases


@pytest.fixture(autouse=True)
def oracle_engine(monkeypatch):
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))


def inputs(path: Path, band_edges=((0, 5), (5, 10)), bands=((100, 0), (0, 100))):
    """Two connected 10 m cells with Horton land and a 2-pixel rain raster."""
    path.mkdir(parents=True, exist_ok=True)
    mesh = cases.two_cells(True)
    rain = cases.write_rain(path, np.asarray(bands, np.float32)[:, None, :], 0., 10., 10., list(band_edges))
    return mesh, rain


LAND = {10: dict(manning_n=.03, infiltration=dict(f0_mm_h=36, fc_mm_h=18, k_per_h=2))}


def run(path, mesh, rain, out, duration, restart='', boundaries=(), afa=.5, sita=1., warmup=0, yield_step=1,
        stations=None):
    from nh_flood_2d_amd.config import DomainConfig, ForceConfig
    cases.write_mesh(path, mesh)
    dc = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=2326, domain_dir=str(out),
                      start=cases.START, duration=duration, warmup_duration=warmup, yield_step=yield_step,
                      afa=afa, sita=sita, land_parameters=LAND, restart_state=str(restart) if restart else '',
                      stations=stations or {})
    fc = ForceConfig(boundaries=list(boundaries), rainfall=rain)
    solver.preprocess(dc, fc)
    budget = solver.solver(dc, fc)
    state = read_checkpoint(Path(dc.output_dir) / 'state_final.fdb')
    manifest = json.loads((Path(out) / 'manifest.json').read_text())
    return budget, state, manifest


def capture(monkeypatch):
    traces = []
    original = Engine.run_batch

    def wrapped(self, steps):
        rows = original(self, steps)
        traces.extend(rows[rows[:, 3] == 1].copy())
        return rows

    monkeypatch.setattr(Engine, 'run_batch', wrapped)
    return traces


def test_restart_allows_changed_afa_but_rejects_changed_sita(tmp_path):
    mesh, rain = inputs(tmp_path / 'in')
    first_budget, first, _ = run(tmp_path / 'in', mesh, rain, tmp_path / 'first', 5, afa=.25)
    checkpoint = tmp_path / 'first/output/state_final.fdb'
    physics = json.loads(first.metadata_json)['physics_config']
    assert physics['afa'] == .25
    budget, final, manifest = run(tmp_path / 'in', mesh, rain, tmp_path / 'resumed', 10, restart=checkpoint, afa=.0625)
    meta = json.loads(final.metadata_json)
    assert manifest['steps'] > 0 and meta['elapsed_seconds'] == 10
    assert budget['rainfall_m3'] > first_budget['rainfall_m3']
    assert meta['physics_config'] == {**physics, 'afa': .0625}
    with pytest.raises(ValueError, match='Checkpoint clock or physical parameters'):
        run(tmp_path / 'in', mesh, rain, tmp_path / 'changed', 10, restart=checkpoint, afa=.0625, sita=.7)


def test_restart_after_requested_end_rejects_before_results(tmp_path):
    mesh, rain = inputs(tmp_path / 'in')
    run(tmp_path / 'in', mesh, rain, tmp_path / 'first', 10)
    target = tmp_path / 'invalid'
    with pytest.raises(ValueError, match='Checkpoint time.*simulation end'):
        run(tmp_path / 'in', mesh, rain, target, 5, restart=tmp_path / 'first/output/state_final.fdb')
    assert not (target / 'output').exists()
    assert not (target / 'mass_balance.csv').exists()
    assert not (target / 'step_statistics.csv').exists()


def test_restart_at_requested_end_completes_without_steps(tmp_path):
    mesh, rain = inputs(tmp_path / 'in')
    expected, first, _ = run(tmp_path / 'in', mesh, rain, tmp_path / 'first', 10)
    actual, final, manifest = run(tmp_path / 'in', mesh, rain, tmp_path / 'same', 10,
                                  restart=tmp_path / 'first/output/state_final.fdb')
    assert manifest['steps'] == 0
    assert manifest['completion_reason'] == 'checkpoint_at_end'
    assert actual == expected
    np.testing.assert_array_equal(final.depth, first.depth)


@pytest.mark.parametrize('change', ['order', 'name', 'kind', 'edges'])
def test_restart_rejects_changed_boundary_identity(tmp_path, change):
    mesh, rain = inputs(tmp_path / 'in')
    path = tmp_path / 'in'
    boundaries = [cases.write_boundary(path, 'A', 'level', [[0, 0], [0, 10]], [(0, .2), (10, .2)]),
                  cases.write_boundary(path, 'B', 'discharge', [[20, 0], [20, 10]], [(0, .1), (10, .1)])]
    run(path, mesh, rain, tmp_path / 'first', 5, boundaries=boundaries)
    if change == 'order':
        boundaries.reverse()
    elif change in ('name', 'kind'):
        boundaries[0][change] = 'renamed' if change == 'name' else 'discharge'
    else:
        data = json.loads((path / 'A.geojson').read_text())
        data['coordinates'] = [[0, 0], [10, 0]]
        (path / 'A.geojson').write_text(json.dumps(data))
    with pytest.raises(ValueError, match='Checkpoint boundary identity'):
        run(path, mesh, rain, tmp_path / 'changed', 10, restart=tmp_path / 'first/output/state_final.fdb',
            boundaries=boundaries)


def test_restart_matches_uninterrupted_rain_and_horton_evolution(tmp_path, monkeypatch):
    traces = capture(monkeypatch)
    mesh, rain = inputs(tmp_path / 'in')
    full_budget, full, full_manifest = run(tmp_path / 'in', mesh, rain, tmp_path / 'full', 10)
    full_trace = np.array(traces, dtype=np.uint32)
    traces.clear()
    first_budget, split, first_manifest = run(tmp_path / 'in', mesh, rain, tmp_path / 'first', 5)
    traces.clear()
    resumed_budget, resumed, resumed_manifest = run(tmp_path / 'in', mesh, rain, tmp_path / 'resumed', 10,
                                                    restart=tmp_path / 'first/output/state_final.fdb')
    resumed_trace = np.array(traces, dtype=np.uint32)
    after = ((full_trace[:, 1].astype(np.uint64) << np.uint64(32)) | full_trace[:, 2].astype(np.uint64)) > 5 * 2 ** 40
    np.testing.assert_array_equal(full_trace[after], resumed_trace)
    assert first_budget['rainfall_m3'] > 0 and first_budget['infiltration_m3'] > 0
    assert np.any(split.rain_time[1:] > 0) and np.any(split.q[1:] != 0)
    for name in ('depth', 'bed', 'q', 'rain_time', 'infiltration', 'max_depth', 'max_speed',
                 'first_time_hi', 'first_time_lo'):
        np.testing.assert_array_equal(getattr(full, name), getattr(resumed, name), err_msg=name)
    assert (full.clock_hi, full.clock_lo) == (resumed.clock_hi, resumed.clock_lo)
    assert full_budget == resumed_budget
    assert full_manifest['steps'] == first_manifest['steps'] + resumed_manifest['steps']
    assert json.loads(resumed.metadata_json)['total_steps'] == full_manifest['steps']
    assert json.loads(split.metadata_json)['elapsed_seconds'] == 5


@pytest.mark.parametrize('interrupt_at', [-900, 1800])
def test_periodic_checkpoint_resumes_bitwise(tmp_path, monkeypatch, interrupt_at):
    from nh_flood_2d_amd import output

    class Interrupted(RuntimeError):
        pass

    traces = capture(monkeypatch)
    mesh, rain = inputs(tmp_path / 'in', band_edges=((0, 900), (900, 2100)))
    checkpoints = []
    wall = [0.]
    replace = output.os.replace
    original_run_batch = Engine.run_batch

    def clocked(self, steps):
        rows = original_run_batch(self, steps)
        done = rows[rows[:, 3] == 1]
        if len(done):
            wall[0] = ((int(done[-1, 1]) << 32) | int(done[-1, 2])) / 2 ** 40
        return rows

    def interrupt_after_replace(source, target):
        if Path(target).name != 'state_latest.fdb':
            return replace(source, target)
        pending = read_checkpoint(source)
        seconds = json.loads(pending.metadata_json)['elapsed_seconds']
        assert seconds % 900 == 0 and seconds != 0
        replace(source, target)
        checkpoints.append(seconds)
        if seconds == interrupt_at:
            raise Interrupted('process stopped after atomic checkpoint replacement')

    def go(name, restart=''):
        wall[0] = 0.
        return run(tmp_path / 'in', mesh, rain, tmp_path / name, 2100, restart=restart, warmup=1800, yield_step=300,
                   stations={'left': (5, 5), 'right': (15, 5)})

    monkeypatch.setattr(Engine, 'run_batch', clocked)
    traces_clocked = []
    monkeypatch.setattr(solver, 'perf_counter', lambda: wall[0])
    monkeypatch.setattr(solver, 'CHECKPOINT_INTERVAL_S', float('inf'))
    full_budget, full, _ = go('full')
    del traces_clocked
    assert not (tmp_path / 'full/output/state_latest.fdb').exists()
    monkeypatch.setattr(solver, 'CHECKPOINT_INTERVAL_S', 500)
    monkeypatch.setattr(output.os, 'replace', interrupt_after_replace)
    with pytest.raises(Interrupted):
        go('interrupted')
    assert checkpoints == ([-900] if interrupt_at < 0 else [-900, 900, 1800])
    interrupted = tmp_path / 'interrupted/output'
    assert not (interrupted / 'state_final.fdb').exists()
    assert not list(interrupted.glob('*.tmp')) and not list(interrupted.glob('.*.tmp'))
    latest = interrupted / 'state_latest.fdb'
    saved = read_checkpoint(latest)
    assert json.loads(saved.metadata_json)['elapsed_seconds'] == interrupt_at
    monkeypatch.setattr(output.os, 'replace', replace)
    resumed_budget, resumed, _ = go('resumed', latest)
    for name in ('depth', 'bed', 'q', 'rain_time', 'infiltration', 'max_depth', 'max_speed',
                 'first_time_hi', 'first_time_lo'):
        assert getattr(full, name).tobytes() == getattr(resumed, name).tobytes(), name
    assert (full.clock_hi, full.clock_lo) == (resumed.clock_hi, resumed.clock_lo)
    assert full.metadata_json == resumed.metadata_json
    assert full_budget == resumed_budget
    for filename in ('output/stations.csv', 'mass_balance.csv', 'step_statistics.csv'):
        with (tmp_path / 'full' / filename).open() as stream:
            expected = [r for r in csv.DictReader(stream) if float(r['elapsed_seconds']) > interrupt_at]
        with (tmp_path / 'resumed' / filename).open() as stream:
            actual = [r for r in csv.DictReader(stream) if float(r['elapsed_seconds']) > interrupt_at]
        assert actual == expected, filename


def test_formal_restart_discards_warmup_extrema(tmp_path):
    from nh_flood_2d_amd.config import DomainConfig, ForceConfig
    path = tmp_path / 'in'
    mesh, rain = inputs(path, bands=((0, 0), (0, 0)))
    depth = np.array([0, .5, 0], np.float32)
    seed = cases.seed(path, mesh, depth, infiltration=np.array([0, .1, .1], np.float32))
    land = {10: dict(manning_n=.03, infiltration=dict(f0_mm_h=360000, fc_mm_h=360000, k_per_h=0))}
    dc = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=2326, domain_dir=str(tmp_path / 'warm'),
                      start=cases.START, duration=2, warmup_duration=2, yield_step=1, restart_state=str(seed),
                      land_parameters=land)
    fc = ForceConfig(boundaries=[], rainfall=rain)
    solver.preprocess(dc, fc)
    solver.warmup_solver(dc, fc)
    warm = read_checkpoint(Path(dc.output_dir) / 'state_final.fdb')
    assert warm.max_depth[1] == .5 and warm.depth[1] < .5
    resumed = dc.model_copy(update={'domain_dir': str(tmp_path / 'formal'), 'warmup_duration': 0,
                                    'restart_state': str(Path(dc.output_dir) / 'state_final.fdb')})
    solver.preprocess(resumed, fc)
    solver.solver(resumed, fc)
    final = read_checkpoint(Path(resumed.output_dir) / 'state_final.fdb')
    assert final.max_depth[1] == warm.depth[1]
    assert int(final.first_time_hi[1]) * 2 ** 32 + int(final.first_time_lo[1]) == 2 * 2 ** 40
