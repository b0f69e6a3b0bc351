"""Whole-solver physics and protocol cases on tiny meshes, driven through the file entry points.

Restates the reference's integration tests (``tests/test_solver_forcing.py``,
``tests/test_inventory_rounding.py``, ``tests/test_still_water.py:171-210``) against this
package's host driver with the CPU oracle as the device step; the HIP step is tied to the
oracle bit for bit by the GPU parity tests. Closed forms and balances are the assertions.
"""
import csv
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.engine import Engine
from nh_flood_2d_amd.output import list_depth_frames, read_checkpoint, read_depth_frame, read_extrema
import cases


@pytest.fixture(autouse=True)
def oracle_engine(monkeypatch):
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))


def experiment(path: Path, bands, duration, land_type=10, q=0., warmup=0, level=None, bed=0.,
               connected=False, level_rows=None, q_rows=None, yield_step=None, warmup_only=False,
               infiltration='none', stations=None):
    """Two 10 x 10 m cells under a 2-pixel rain raster, optional river (x = 0) and sea (x = 20)."""
    path.mkdir(parents=True, exist_ok=True)
    mesh = cases.two_cells(connected, bed, land_type)
    values = np.asarray(bands, dtype=np.float32).reshape(-1, 2)
    n = len(values)
    rain = cases.write_rain(path, values[:, None, :], 0., 10., 10.,
                            [(i * duration / n, (i + 1) * duration / n) for i in range(n)])
    boundaries = []
    if q != 0:
        boundaries.append(cases.write_boundary(path, 'river', 'discharge', [[0, 1], [0, 9]],
                                               q_rows or [(0, q), (duration, q)]))
    if level is not None:
        boundaries.append(cases.write_boundary(path, 'sea', 'level', [[20, 1], [20, 9]],
                                               level_rows or [(0, level), (duration, level)]))
    land = {land_type: dict(manning_n=.03, infiltration=infiltration)}
    row, manifest, dc = cases.run(path, mesh, duration=duration, rain=rain, boundaries=boundaries, warmup=warmup,
                                  land=land, yield_step=yield_step, warmup_only=warmup_only, stations=stations)
    state = read_checkpoint(Path(dc.output_dir) / 'state_final.fdb')
    return row, state.bed + state.depth, manifest


def test_closed_basin_rain_and_discharge_balance(tmp_path):
    result, h, manifest = experiment(tmp_path, [[6, 6]], 60, q=.1, connected=True)
    assert result['rainfall_m3'] == pytest.approx(1.2, abs=1e-8)
    assert result['discharge:river_m3'] == pytest.approx(6., abs=1e-7)
    assert abs(result['residual_m3']) / 7.2 < 1e-4
    assert result['storage_m3'] == pytest.approx(7.2, rel=1e-4)
    assert np.all(np.isfinite(h))
    assert manifest['status'] == 'completed' and manifest['steps'] >= 60
    assert len(manifest['input_sha256']) == 6


def test_horton_clocks_are_local(tmp_path):
    horton = dict(f0_mm_h=76.2, fc_mm_h=2.54, k_per_h=2)
    _, h_late, _ = experiment(tmp_path / 'late', [[100, 0], [0, 100]], 120, land_type=1, infiltration=horton)
    _, h_ref, _ = experiment(tmp_path / 'ref', [[0, 100]], 60, land_type=1, infiltration=horton)
    assert h_late[2] == pytest.approx(h_ref[2], abs=1e-6)


def test_actual_infiltration_is_bounded_by_water(tmp_path):
    result, h, _ = experiment(tmp_path, [[.001, .001]], 60, land_type=6,
                              infiltration=dict(f0_mm_h=101.6, fc_mm_h=7.62, k_per_h=2))
    assert result['infiltration_m3'] == pytest.approx(result['rainfall_m3'], rel=2e-6)
    assert result['bed_clamp_nonlevel_m3'] == 0 and result['bed_clamp_level_m3'] == 0
    assert np.max(h) == 0


def test_constant_horton_capacity_sets_actual_infiltration(tmp_path):
    high, _, _ = experiment(tmp_path / 'high', [[100, 100]], 60, infiltration=dict(f0_mm_h=36, fc_mm_h=36, k_per_h=0))
    low, _, _ = experiment(tmp_path / 'low', [[100, 100]], 60, infiltration=dict(f0_mm_h=18, fc_mm_h=18, k_per_h=0))
    none, _, _ = experiment(tmp_path / 'none', [[100, 100]], 60)
    assert high['infiltration_m3'] == pytest.approx(.12, rel=1e-4)
    assert low['infiltration_m3'] == pytest.approx(.06, rel=1e-4)
    assert none['infiltration_m3'] == 0.


def test_warmup_holds_boundaries_and_has_no_rain(tmp_path):
    result, _, _ = experiment(tmp_path, [[6, 6]], 60, q=.01, warmup=20, level=.1,
                              level_rows=[(0, .1), (60, .2)], q_rows=[(0, .01), (60, .03)], yield_step=10)
    rows = list(csv.DictReader((tmp_path / 'out/mass_balance.csv').open()))
    at_start = next(r for r in rows if float(r['elapsed_seconds']) == 0)
    assert float(at_start['rainfall_m3']) == 0
    assert float(at_start['discharge:river_m3']) == pytest.approx(.2)
    assert result['rainfall_m3'] == pytest.approx(1.2)
    assert result['discharge:river_m3'] == pytest.approx(1.4)
    assert [float(r['elapsed_seconds']) for r in rows if r['phase'] == 'warmup'] == [-20., -10.]
    state = read_checkpoint(tmp_path / 'out/output/state_start.fdb')
    assert state.depth[2] == pytest.approx(.1)
    assert abs(result['relative_residual']) < 1e-4


@pytest.mark.parametrize('bed,land_type', [(-5., 10), (1., 8)])
def test_cold_start_is_dry_without_boundaries(tmp_path, bed, land_type):
    result, h, _ = experiment(tmp_path, [[0, 0]], 2, bed=bed, land_type=land_type)
    assert result['initial_storage_m3'] == 0. and result['storage_m3'] == 0.
    np.testing.assert_array_equal(h[1:], [bed, bed])


def test_warmup_only_stops_at_start_and_outputs_each_interval(tmp_path):
    result, _, manifest = experiment(tmp_path, [[6, 6]], 60, q=.01, warmup=20, level=.1, yield_step=10,
                                     warmup_only=True)
    rows = list(csv.DictReader((tmp_path / 'out/mass_balance.csv').open()))
    assert [float(r['elapsed_seconds']) for r in rows] == [-20., -10., 0.]
    assert all(r['phase'] == 'warmup' for r in rows)
    assert result['rainfall_m3'] == 0.
    assert result['discharge:river_m3'] == pytest.approx(.2)
    assert manifest['run_kind'] == 'warmup'
    assert manifest['dt_seconds']['sum'] == pytest.approx(20.)
    assert sorted(p.stem for p in (tmp_path / 'out/output/frames').glob('*.fdb')) == ['depth_-20', 'depth_0']
    assert (tmp_path / 'out/output/state_final.fdb').exists()


def test_small_rain_is_retained_above_large_bed(tmp_path):
    result, _, _ = experiment(tmp_path, [[.01, .01]], 60, bed=100.)
    assert result['storage_m3'] == pytest.approx(.002, rel=2e-6)
    assert abs(result['residual_m3']) < 4e-9


def test_existing_output_is_not_overwritten(tmp_path):
    from nh_flood_2d_amd.config import DomainConfig, ForceConfig
    experiment(tmp_path, [[0, 0]], 1)
    before = ((tmp_path / 'out/manifest.json').read_bytes(), (tmp_path / 'out/mass_balance.csv').read_bytes())
    dc = DomainConfig(ne=str(tmp_path / 'ne.txt'), ns=str(tmp_path / 'ns.txt'), epsg_code=2326,
                      domain_dir=str(tmp_path / 'out'), start=cases.START, duration=1, warmup_duration=0,
                      land_parameters={10: dict(manning_n=.03, infiltration='none')})
    fc = ForceConfig(boundaries=[], rainfall=dict(raster='missing', intervals='missing'))
    with pytest.raises(FileExistsError):
        solver.RunRecord(dc, fc)
    assert before == ((tmp_path / 'out/manifest.json').read_bytes(), (tmp_path / 'out/mass_balance.csv').read_bytes())


def test_input_failure_is_recorded(tmp_path):
    from nh_flood_2d_amd.config import DomainConfig, ForceConfig
    (tmp_path / 'ne').write_text('mesh')
    (tmp_path / 'ns').write_text('mesh')
    dc = DomainConfig(ne=str(tmp_path / 'ne'), ns=str(tmp_path / 'ns'), epsg_code=2326,
                      domain_dir=str(tmp_path / 'out'), start=cases.START, duration=1, warmup_duration=0,
                      land_parameters={10: dict(manning_n=.03, infiltration='none')})
    fc = ForceConfig(boundaries=[], rainfall=dict(raster='missing', intervals='missing'))
    with pytest.raises(FileNotFoundError):
        solver.solver(dc, fc)
    assert json.loads((tmp_path / 'out/manifest.json').read_text())['status'] == 'failed'


def test_level_accounts_for_small_rain_at_high_bed(tmp_path):
    result, _, _ = experiment(tmp_path, [[0, .01]], 60, bed=99., level=100.)
    assert abs(result['residual_m3']) < 1e-8
    assert result['level:sea_m3'] == pytest.approx(100. - .001, abs=1e-8)


def test_level_below_bed_is_zero_depth_and_closes_budget(tmp_path):
    result, _, _ = experiment(tmp_path, [[0, .01]], 60, bed=100., level=50.)
    assert result['storage_m3'] == 0.
    assert result['bed_clamp_nonlevel_m3'] == 0. and result['bed_clamp_level_m3'] == 0.
    assert result['rainfall_m3'] == pytest.approx(.001, abs=1e-10)
    assert abs(result['residual_m3']) < 1e-7


def test_negative_discharge_shares_available_water_without_creating_water(tmp_path):
    result, _, _ = experiment(tmp_path, [[6, 6]], 60, q=-10.)
    assert result['discharge:river_m3'] == pytest.approx(-.6, abs=2e-6)
    assert result['requested_discharge:river_m3'] == pytest.approx(-600.)
    assert result['storage_m3'] == pytest.approx(.6, abs=2e-6)
    assert result['bed_clamp_nonlevel_m3'] < 1e-6
    assert abs(result['residual_m3']) < 2e-6


def test_horton_uses_f0_on_wet_cell_before_any_rain(tmp_path):
    result, _, _ = experiment(tmp_path, [[0, 0]], 60, level=.1,
                              infiltration=dict(f0_mm_h=76.2, fc_mm_h=2.54, k_per_h=2))
    assert result['infiltration_m3'] == pytest.approx(76.2 / 1000 / 3600 * 100 * 60, rel=2e-6)


def test_true_small_cfl_stops_with_limiting_edge_diagnostics(tmp_path):
    with pytest.raises(RuntimeError, match=r'CFL.*edge=\d+, cells=\[1, 2\].*depth_m=.*velocity_m_s='):
        experiment(tmp_path, [[0, 0]], 1, level=1e9, connected=True)


def test_subsecond_rain_knot_selects_the_next_band(tmp_path):
    path = tmp_path / 'knot'
    experiment(path, [[1, 0], [0, 2]], 2)
    from nh_flood_2d_amd.config import DomainConfig, ForceConfig
    with open(path / 'intervals.csv', 'w') as f:
        f.write('interval_start,interval_end\n')
        f.write(f'{cases.iso(0)},{cases.iso(0.000002)}\n{cases.iso(0.000002)},{cases.iso(1)}\n')
    dc = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=2326,
                      domain_dir=str(path / 'knot_out'), start=cases.START, duration=1, warmup_duration=0,
                      land_parameters={10: dict(manning_n=.03, infiltration='none')}, yield_step=1)
    fc = ForceConfig(boundaries=[], rainfall=dict(raster=str(path / 'rain.tif'), intervals=str(path / 'intervals.csv')))
    solver.preprocess(dc, fc)
    result = solver.solver(dc, fc)
    state = read_checkpoint(Path(dc.output_dir) / 'state_final.fdb')
    np.testing.assert_allclose(state.depth[1:], [.001, .002], rtol=2e-6, atol=0)
    assert result['storage_m3'] == pytest.approx(.3, rel=2e-6)


def test_subsecond_level_knot_and_between_frame_peak(tmp_path):
    _, h, _ = experiment(tmp_path, [[0, 0]], 1, level=0., level_rows=[(0, 0.), (.000002, 1.), (1, 0.)])
    assert abs(h[2]) < 2e-6
    root = tmp_path / 'out/output'
    extrema = read_extrema(root / 'extrema.fdb')
    assert extrema.max_depth[2] == pytest.approx(1., abs=2e-6)
    assert all(read_depth_frame(p)[2] < 2e-6 for _, p in list_depth_frames(root))
    peak = (int(extrema.first_time_hi[2]) * 2 ** 32 + int(extrema.first_time_lo[2])) / 2 ** 40
    assert peak == pytest.approx(.000002, abs=5e-13)


def _capture_traces(monkeypatch):
    traces = []
    original = Engine.run_batch

    def capture(self, steps):
        rows = original(self, steps)
        traces.extend(rows[rows[:, 3] == 1].copy())
        return rows

    monkeypatch.setattr(Engine, 'run_batch', capture)
    return traces


def test_next_cfl_uses_depth_after_the_level_update(tmp_path, monkeypatch):
    traces = _capture_traces(monkeypatch)
    experiment(tmp_path, [[0, 0]], 1, connected=True, level=0, level_rows=[(0, 0), (.000002, 100), (1, 100)])
    candidates = np.array(traces, dtype=np.uint32)[:, 4].copy().view(np.float32)
    # the first event step raises cell 2 from dry to 100 m while q is still zero
    assert candidates[1] == pytest.approx(.5 * 10 / np.sqrt(9.81 * 100), rel=2e-7)
    assert candidates[1] < 1
    last = list(csv.DictReader((tmp_path / 'out/step_statistics.csv').open()))[-1]
    assert float(last['face_depth_m']) == max(float(last['lower_depth_m']), float(last['higher_depth_m']))
    expected = .5 * 10 / (np.sqrt(9.81 * float(last['face_depth_m'])) + abs(float(last['face_velocity_m_s'])))
    assert float(last['cfl_min_s']) == pytest.approx(expected, rel=2e-7)


def test_all_event_steps_have_explicit_standard_statistics(tmp_path):
    experiment(tmp_path, [[0, 0]], 1, q=.1, q_rows=[(0, .1), (.000002, .1), (1, .1)])
    last = list(csv.DictReader((tmp_path / 'out/step_statistics.csv').open()))[-1]
    assert last['steps'] == last['event_short_steps'] == '2'
    assert last['standard_steps'] == '0'
    for name in ('standard_dt_min_s', 'standard_dt_max_s', 'standard_dt_mean_s'):
        assert last[name] == ''
    assert float(last['event_dt_min_s']) == pytest.approx(.000002)
    assert float(last['event_dt_sum_s']) == pytest.approx(1., abs=1e-7)


def test_transport_and_small_rain_have_one_inventory_rounding(tmp_path):
    """inventory_rounding.py: the F32 depth takes one rounding of (flux + rain) increments."""
    mesh = cases.two_cells(True)
    mm = np.float32(1.4911491234670393e-5)
    rain = cases.write_rain(tmp_path, np.array([[[mm, 0]]], np.float32), 0., 10., 10., [(0, 1)])
    depth = np.array([0, .5, .4], np.float32)
    seed = cases.seed(tmp_path, mesh, depth)
    row, manifest, dc = cases.run(tmp_path, mesh, duration=1, rain=rain, restart=seed)
    state = read_checkpoint(Path(dc.output_dir) / 'state_final.fdb')
    assert manifest['steps'] == 1
    shared = cases.side_at(mesh, 10, 5)
    rain_rate = float(np.float32(float(mm) * .001))
    ideal = float(depth[1]) - float(state.q[shared]) * 10 / 100 + rain_rate
    assert abs(float(state.depth[1]) - ideal) < .55 * float(np.spacing(np.float32(ideal)))


def test_warmup_fills_connected_cells_into_initial_inventory(tmp_path):
    result, _, meta = experiment(tmp_path, [[0, 0]], 1, connected=True, level=1, warmup=1, warmup_only=True)
    assert meta['initialization']['method'] == 'connected_still_water'
    assert result['initial_storage_m3'] == 200
    assert result['level:sea_m3'] == 0 and result['residual_m3'] == 0
    state = read_checkpoint(tmp_path / 'out/output/state_final.fdb')
    np.testing.assert_array_equal(state.depth, [0, 1, 1])
    assert not np.any(state.q)


def test_no_warmup_preserves_dry_initialization(tmp_path):
    result, _, meta = experiment(tmp_path, [[0, 0]], 1, connected=True, level=1)
    assert meta['initialization']['method'] == 'dry'
    assert result['initial_storage_m3'] == 0
