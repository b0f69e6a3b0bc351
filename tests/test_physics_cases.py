"""Positivity limiter, hanging-interface pressure reconstruction and steady channel flow.

Restates the reference's ``tests/test_positivity.py``, ``tests/test_pressure_gradient.py:98-151``
and ``tests/test_face_flux.py:113-271`` through this package's solver with the CPU oracle as the
device step. Assertions are the physical invariants and closed forms of those tests.
"""
import csv
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.config import DomainConfig, ForceConfig
from nh_flood_2d_amd.hydraulics import hanging_geometry
from nh_flood_2d_amd.output import list_depth_frames, read_checkpoint, read_depth_frame
import cases


@pytest.fixture(autouse=True)
def oracle_engine(monkeypatch):
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))


# ------------------------------------------------------------------------------ positivity

def _grid_run(path: Path, wet_cell, face_flows, extraction=False):
    mesh = cases.grid(3, 3, 10.)
    depth = np.zeros(mesh.e_num, np.float32)
    depth[wet_cell] = np.float32(.1)
    q = np.zeros(mesh.s_num, np.float32)
    for side, value in face_flows.items():
        q[side] = value
    seed = cases.seed(path, mesh, depth, q=q)
    rain = cases.write_rain(path, np.zeros((1, 1, 1), np.float32), 0., 30., 30., [(0, 1)])
    boundaries = []
    if extraction:
        boundaries.append(cases.write_boundary(path, 'river', 'discharge', [[0, 0], [0, 10]], [(0, -100), (1, -100)]))
    row, manifest, dc = cases.run(path, mesh, duration=1, rain=rain, boundaries=boundaries, restart=seed,
                                  land={10: dict(manning_n=1e-6, infiltration='none')}, yield_step=1)
    return row, read_checkpoint(Path(dc.output_dir) / 'state_final.fdb'), manifest, mesh


def _cell(row, col):
    return row * 3 + col + 1


def test_four_outgoing_faces_share_one_cell_inventory(tmp_path):
    mesh = cases.grid(3, 3, 10.)
    flows = {cases.side_at(mesh, 10, 15): -1, cases.side_at(mesh, 20, 15): 1,
             cases.side_at(mesh, 15, 10): -1, cases.side_at(mesh, 15, 20): 1}
    first_cfl = .5 * 10 / (np.sqrt(9.81 * .1) + 1 / .1)
    assert 4 * 10 * first_cfl > 10
    row, state, manifest, _ = _grid_run(tmp_path / 'center', _cell(1, 1), flows)
    neighbours = state.depth[[_cell(1, 0), _cell(1, 2), _cell(0, 1), _cell(2, 1)]]
    assert np.all(state.depth >= 0) and np.all(neighbours > 0)
    np.testing.assert_allclose(neighbours, neighbours[0], rtol=0, atol=2e-7)
    assert row['initial_storage_m3'] == pytest.approx(10)
    assert row['storage_m3'] == pytest.approx(10, abs=1e-5)
    assert row['rainfall_m3'] == row['infiltration_m3'] == 0
    assert row['bed_clamp_nonlevel_m3'] <= 1e-5 and abs(row['residual_m3']) <= 1e-5
    assert manifest['steps'] >= 2
    activity = manifest['limiter_activity']
    assert activity['limited_cell_steps'] > 0 and activity['limited_face_steps'] >= 4
    assert 0 <= activity['minimum_donor_scale'] < 1
    rows = list(csv.DictReader((tmp_path / 'center/out/step_statistics.csv').open()))
    assert sum(int(r['limited_cell_steps']) for r in rows) == activity['limited_cell_steps']
    assert sum(int(r['limited_face_steps']) for r in rows) == activity['limited_face_steps']
    assert min(float(r['minimum_donor_scale']) for r in rows) == activity['minimum_donor_scale']


def test_negative_discharge_and_faces_share_corner_inventory(tmp_path):
    mesh = cases.grid(3, 3, 10.)
    flows = {cases.side_at(mesh, 10, 5): 1, cases.side_at(mesh, 5, 10): 1}
    row, state, _, _ = _grid_run(tmp_path / 'corner', _cell(0, 0), flows, extraction=True)
    assert np.all(state.depth >= 0)
    assert state.depth[_cell(0, 1)] > 0 and state.depth[_cell(1, 0)] > 0
    assert row['requested_discharge:river_m3'] == pytest.approx(-100)
    assert -10 < row['discharge:river_m3'] < 0
    assert row['storage_m3'] > 0
    assert row['bed_clamp_nonlevel_m3'] <= 1e-5 and abs(row['residual_m3']) <= 1e-5


# ------------------------------------------------------------------------------ pressure

def strip_cells(rotated=False):
    cells = [(0, 8, y, y + 8) for y in range(0, 24, 8)] + [(8, 12, y, y + 4) for y in range(0, 24, 4)]
    return [(c, d, a, b) for a, b, c, d in cells] if rotated else cells


def evolve(path: Path, cells, bed_at, depth_at, seconds=.125):
    mesh = cases.boxes_mesh(cells, bed=lambda x, y, i: bed_at(x, y, i))
    depth = np.r_[0, [depth_at(mesh.x[i], mesh.y[i], mesh.z[i], i) for i in range(1, mesh.e_num)]].astype(np.float32)
    seed = cases.seed(path, mesh, depth)
    x1, y1 = max(c[1] for c in cells), max(c[3] for c in cells)
    rain = cases.write_rain(path, np.zeros((1, 1, 1), np.float32), 0., y1, float(max(x1, y1)), [(0, 3600)])
    dc = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=2326, domain_dir=str(path / 'out'),
                      start=cases.START, duration=max(1, int(seconds)), warmup_duration=0,
                      land_parameters={10: dict(manning_n=.03, infiltration='none')}, restart_state=str(seed))
    fc = ForceConfig(boundaries=[], rainfall=rain)
    with solver.RunRecord(dc, fc) as record:
        record.hash_inputs(dc, fc)
        solver.preprocess(dc, fc)
        solver._solve(dc, fc, record, end_time=seconds)
    return read_checkpoint(path / 'out/output/state_final.fdb'), depth, mesh


@pytest.mark.parametrize('rotated', [False, True])
def test_linear_tangential_surface_has_only_roundoff_crossflow(tmp_path, rotated):
    axis = 0 if rotated else 1
    result, _, mesh = evolve(tmp_path, strip_cells(rotated), lambda x, y, i: .001 * (x if rotated else y),
                             lambda x, y, z, i: 1.)
    topo = mesh.topo
    interface = (topo[:, 0] == (1 if rotated else 2)) & (topo[:, 1] != 0) & (topo[:, 2] != 0)
    assert np.max(np.abs(result.q[interface])) < 3e-7
    assert .001 / 3 * 9.81 * .125 > 1000 * np.max(np.abs(result.q[interface]))
    layout = hanging_geometry(mesh.x, mesh.y, mesh.sx, mesh.sy, topo)
    np.testing.assert_array_equal(layout.cells, [1, 2, 3])
    assert len(layout.offsets) == 6
    np.testing.assert_array_equal(layout.axes[:, axis], 1)
    np.testing.assert_array_equal(layout.axes[:, 1 - axis], 0)


@pytest.mark.parametrize('rotated', [False, True])
def test_variable_bed_still_water_is_bitwise_unchanged(tmp_path, rotated):
    result, initial, _ = evolve(tmp_path, strip_cells(rotated), lambda x, y, i: ((i % 7) - 3) * .125,
                                lambda x, y, z, i: 2. - z, seconds=60.)
    np.testing.assert_array_equal(result.depth.view(np.uint32), initial.view(np.uint32))
    assert np.all(result.q == 0.)


@pytest.mark.parametrize('dry_cells', [(1,), (3,), (3, .02), (1, 3)])
def test_dry_tangent_neighbours_use_one_side_or_zero(tmp_path, dry_cells):
    threshold = dry_cells == (3, .02)
    unavailable = (3,) if threshold else dry_cells
    result, _, mesh = evolve(tmp_path, strip_cells(), lambda x, y, i: .001 * y,
                             lambda x, y, z, i: (.02 if threshold else 0.) if i in unavailable else 1.)
    topo = mesh.topo
    middle = (topo[:, 0] == 2) & (topo[:, 1] == 2)
    if len(unavailable) == 1:
        assert np.max(np.abs(result.q[middle])) < 3e-7
    else:
        np.testing.assert_allclose(np.sort(result.q[middle]), [-.00040875, .00040875], atol=3e-7, rtol=0)


def test_one_cell_caches_both_axes_and_reconstructs_linear_plane(tmp_path):
    cells = [(0, 8, 0, 8), (8, 12, 0, 4), (8, 12, 4, 8), (0, 4, 8, 12), (4, 8, 8, 12), (8, 12, 8, 12)]
    a, b = .0013, -.0007
    result, _, mesh = evolve(tmp_path, cells, lambda x, y, i: a * x + b * y, lambda x, y, z, i: 1.)
    layout = hanging_geometry(mesh.x, mesh.y, mesh.sx, mesh.sy, mesh.topo)
    np.testing.assert_array_equal(layout.cells, [1])
    np.testing.assert_array_equal(layout.axes, [[1, 1]])
    edges = layout.edge_index >= 0
    expected = -9.81 * .125 * np.where(mesh.topo[edges, 0] == 2, a, b)
    np.testing.assert_allclose(result.q[edges], expected, atol=3e-7, rtol=0)


def test_aligned_mesh_allocates_no_reconstruction():
    mesh = cases.boxes_mesh([(x, x + 4, y, y + 4) for x in (0, 4) for y in (0, 4, 8)])
    assert hanging_geometry(mesh.x, mesh.y, mesh.sx, mesh.sy, mesh.topo) is None


# ------------------------------------------------------------------------------ channel flow

def channel_cells(length=1024):
    half = length // 2
    cells = [(x, x + 2, 0, 2) for x in range(0, half, 2)]
    cells += [(x, x + 1, y, y + 1) for x in range(half, length) for y in (0, 1)]
    return cells


@pytest.mark.parametrize('theta', [0.7, 1.0])
def test_coarse_fine_channel_reaches_continuous_steady_flow(tmp_path, theta):
    path = tmp_path / f'theta_{theta}'
    path.mkdir()
    slope, duration = 0.002, 4000
    cells = channel_cells()
    mesh = cases.boxes_mesh(cells, bed=lambda x, y, i: 10 - slope * x)
    expected_q = np.sqrt(slope) / 0.03
    rain = cases.write_rain(path, np.zeros((1, 1, 1), np.float32), 0., 2., 1024., [(0, duration)])
    bnd = [cases.write_boundary(path, 'upstream', 'discharge', [[0, .1], [0, 1.9]], [(0, 2 * expected_q), (duration, 2 * expected_q)]),
           cases.write_boundary(path, 'downstream', 'level', [[1024, .1], [1024, 1.9]],
                                [(0, float(mesh.z[-1] + 1)), (duration, float(mesh.z[-1] + 1))])]
    depth = np.ones(mesh.e_num, np.float32)
    depth[0] = 0
    seed = cases.seed(path, mesh, depth)
    coarse = next(i for i, c in enumerate(cells, 1) if c == (510, 512, 0, 2))
    fine = [i for i, c in enumerate(cells, 1) if c[0] == 512 and c[1] == 513]
    stations = {str(i): (float(mesh.x[i]), float(mesh.y[i])) for i in [coarse, *fine]}
    row, _, dc = cases.run(path, mesh, duration=duration, rain=rain, boundaries=bnd, restart=seed, sita=theta,
                           yield_step=300, stations=stations)
    samples = list(csv.DictReader((Path(dc.output_dir) / 'stations.csv').open()))
    frames = list_depth_frames(dc.output_dir)[-4:-1]
    history = []
    for seconds, frame in frames:
        water = read_depth_frame(frame)
        u = {int(s['cell_id']): float(s['u']) for s in samples if float(s['elapsed_seconds']) == seconds}
        history.append((float(water[coarse] * u[coarse]), float(sum(water[i] * u[i] for i in fine) / 2)))
    q_coarse, q_fine = history[-1]
    water = read_depth_frame(frames[-1][1])
    interface_depth = float((water[coarse] + sum(water[i] for i in fine) / 2) / 2)
    assert q_coarse == pytest.approx(expected_q, rel=2e-3)
    assert q_fine == pytest.approx(expected_q, rel=2e-3)
    assert q_coarse == pytest.approx(q_fine, rel=1e-4)
    assert max(np.ptp([q[j] for q in history]) for j in (0, 1)) < 1e-4
    assert interface_depth == pytest.approx(1, abs=.02)
    assert abs(row['residual_m3']) / (row['initial_storage_m3'] + row['discharge:upstream_m3']) < 1e-4


def test_open_channel_steady_profile_is_stable_across_theta_and_time_step(tmp_path):
    nx, ny, dx = 40, 2, 10
    profiles = []
    for theta in (1., .7):
        for afa in (.4, .1):
            path = tmp_path / f'channel_theta{theta}_afa{afa}'
            path.mkdir()
            mesh = cases.grid(nx, ny, dx, bed=0.)
            duration = 6000
            rain = cases.write_rain(path, np.zeros((1, 1, 1), np.float32), 0., ny * dx, float(nx * dx), [(0, duration)])
            bnd = [cases.write_boundary(path, 'upstream', 'discharge', [[0, .1], [0, ny * dx - .1]], [(0, 10.), (duration, 10.)]),
                   cases.write_boundary(path, 'downstream', 'level', [[nx * dx, .1], [nx * dx, ny * dx - .1]], [(0, 1.), (duration, 1.)])]
            depth = np.r_[0., np.full(nx * ny, 1.)].astype(np.float32)
            seed = cases.seed(path, mesh, depth)
            row, _, dc = cases.run(path, mesh, duration=duration, rain=rain, boundaries=bnd, restart=seed,
                                   sita=theta, afa=afa, yield_step=500, preprocess=True)
            frames = list_depth_frames(dc.output_dir)
            eta = [read_depth_frame(f)[1:].reshape(ny, nx) for _, f in frames[-3:]]
            profile = eta[-1].mean(axis=0)
            profiles.append(profile)
            assert abs(profile[0] - profile[1]) < .005
            assert abs(profile[-2] - profile[-1]) < .005
            assert np.max(np.abs(eta[-1] - eta[-2])) < 2e-4
            rows = list(csv.DictReader((path / 'out' / 'mass_balance.csv').open()))
            outflow = [-(float(b['level:downstream_m3']) - float(a['level:downstream_m3'])) /
                       (float(b['elapsed_seconds']) - float(a['elapsed_seconds'])) for a, b in zip(rows[-3:-1], rows[-2:])]
            assert all(abs(q - 10.) < .01 for q in outflow), (theta, afa, outflow)
            assert abs(row['residual_m3']) / (row['initial_storage_m3'] + row['discharge:upstream_m3']) < 1e-4
    assert np.ptp(np.stack(profiles), axis=0).max() < 2e-4
