"""The HIP flood-map rasteriser against the C oracle's restatement, bit for bit (needs an MI355X).

* The benchmark mesh (1.58 M cells, 2/4/8 m) on its own grid, in one tile and in 1000-pixel tiles.
* The same mesh moved to projected coordinates (x ~ 8.1e5, y ~ 2.5e6): F32 holds quarter metres
  there, so every corner and pixel range comes from non-integral F32 values, tile by tile.
* A shared grid of two meshes (corners in F64, rounded when passed), as the water-level MAE uses.
* The maps of a finished run written through the GPU and through the oracle: the same files.
"""
import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import maps, synth
from nh_flood_2d_amd.maps import Rasterizer, area_meta

pytestmark = pytest.mark.gpu


def _say(text):
    print(text, flush=True)


def _both(meta, values, grid, gpu_device, tile=maps.TILE, fill=maps.NODATA):
    gpu = Rasterizer(meta, gpu_device).raster(values, grid, fill=fill, tile=tile)
    ref = Rasterizer(meta, oracle.HostDevice('f32')).raster(values, grid, fill=fill, tile=tile)
    return gpu, ref


@pytest.fixture(scope='module')
def bench_mesh():
    return synth.multires_channel(0).mesh


@pytest.mark.parametrize('tile', [maps.TILE, 1000])
def test_benchmark_mesh_raster_matches_the_oracle(gpu_device, bench_mesh, tile):
    meta = area_meta(bench_mesh)
    values = np.random.default_rng(1).normal(0., 2., len(meta.x)).astype(np.float32)
    _say(f'[maps] {len(meta.x)} cells on {meta.width} x {meta.height} px of {meta.hr} m, tile {tile}')
    gpu, ref = _both(meta, values, meta.grid(), gpu_device, tile)
    assert gpu.tobytes() == ref.tobytes()
    assert np.count_nonzero(gpu != maps.NODATA) * float(meta.hr) * float(meta.vr) == bench_mesh.geometry()[3].sum()


def test_projected_coordinates_match_the_oracle(gpu_device, bench_mesh):
    """Centres and side midpoints moved by (812345.3, 2512345.7) m and rounded to F32. The mesh's
    integral coordinates all move by the same 0.25 and 0.75 m there (F32 holds quarter metres),
    so the cells still tile without overlap (one writer per pixel); the corners, half sizes and
    pixel ranges are F32 arithmetic on non-integral values on both sides, in 777-pixel tiles."""
    mesh = synth.multires_channel(0).mesh
    for name, shift in (('x', 812345.3), ('y', 2512345.7)):
        setattr(mesh, name, (getattr(mesh, name).astype(np.float64) + shift).astype(np.float32))
        side = 's' + name
        setattr(mesh, side, (getattr(mesh, side).astype(np.float64) + shift).astype(np.float32))
    meta = area_meta(mesh)
    values = np.arange(1, len(meta.x) + 1, dtype=np.float32)
    gpu, ref = _both(meta, values, meta.grid(), gpu_device, 777)
    painted = np.count_nonzero(gpu > 0)
    _say(f'[maps] projected: {meta.width} x {meta.height} px from ({meta.min_x}, {meta.max_y}), {painted} painted')
    assert gpu.tobytes() == ref.tobytes()
    assert painted * float(meta.hr) * float(meta.vr) == bench_mesh.geometry()[3].sum()


def test_shared_grid_matches_the_oracle(gpu_device, bench_mesh):
    """The benchmark mesh on the grid it shares with a coarser mesh over part of its box (the
    water-level MAE's grid: intersection of the boxes, finer pixel, F64 corners)."""
    meta = area_meta(bench_mesh)
    other = area_meta(synth.small_multires(2, discharge=1.0).mesh)
    grid = maps.shared_grid(meta, other)
    levels = np.random.default_rng(2).uniform(-5., 50., len(meta.x)).astype(np.float32)
    gpu, ref = _both(meta, levels, grid, gpu_device, 500)
    _say(f'[maps] shared grid {grid.width} x {grid.height}')
    assert gpu.tobytes() == ref.tobytes()


def test_flood_maps_of_a_run_match_the_oracle(tmp_path, gpu_device, monkeypatch):
    """A storm on a 2:1 mesh run by the solver; its flood maps and binary extent written through
    the GPU and through the oracle are the same bytes."""
    import cases
    from nh_flood_2d_amd import solver_compact as solver
    from nh_flood_2d_amd.geotiff import read_geotiff
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))
    sc = synth.small_multires(2, discharge=1.0)
    mesh = sc.mesh
    bounds = mesh.geometry()[0][1:]
    west, south, east, north = bounds[:, 0].min(), bounds[:, 1].min(), bounds[:, 2].max(), bounds[:, 3].max()
    tmp_path.mkdir(exist_ok=True)
    rain = cases.write_rain(tmp_path, np.full((1, 1, 1), 40., np.float32), west, north,
                            max(east - west, north - south), [(0, 1800)])
    _, _, dc = cases.run(tmp_path, mesh, duration=1800, rain=rain, land=sc.land_parameters)
    images = {}
    for name, device in (('gpu', gpu_device), ('oracle', oracle.HostDevice('f32'))):
        paths = maps.generate_flood_map(dc, device=device)
        extent = maps.generate_max_inundation_extent_map(dc, min_depth=.02, is_absolute=True, device=device)
        images[name] = [read_geotiff(p).bands.tobytes() for p in paths] + [read_geotiff(extent).bands.tobytes()]
    assert len(images['gpu']) == 4 and images['gpu'] == images['oracle']
