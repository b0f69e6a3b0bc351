"""Flood maps, extents, F1 scores, water-level MAE and hydrographs on the CPU, with the C oracle's
rasteriser behind the same C ABI as the HIP kernel (tests/test_gpu_maps.py ties the two bit for
bit). The assertions are closed forms: a pixel belongs to the cell whose box holds its centre
(x0 <= x < x1, y0 < y <= y1, the reference's half-open rule, flood_map.py:211-212), so a map is a
point-location problem solved here independently of the kernel's ceil arithmetic.
"""
import json
from datetime import timedelta

import numpy as np
import pytest

import cases
import oracle
from nh_flood_2d_amd import maps
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.geotiff import read_geotiff, write_geotiff
from nh_flood_2d_amd.maps import Grid, Rasterizer, area_meta
from nh_flood_2d_amd.output import BedTimeline, read_depth_frame, read_extrema, write_bed

X0, Y0 = 1000., 5000.
# a 2:1 mesh of 40 x 40 m: three 20 m cells and four 10 m cells in the lower right quarter
BOXES = [(0, 20, 0, 20), (0, 20, 20, 40), (20, 40, 20, 40),
         (20, 30, 0, 10), (30, 40, 0, 10), (20, 30, 10, 20), (30, 40, 10, 20)]


def _mesh(boxes=BOXES, x0=X0, y0=Y0):
    return cases.boxes_mesh([(a + x0, b + x0, c + y0, d + y0) for a, b, c, d in boxes])


def _boxes(mesh):
    bounds = mesh.geometry()[0][1:]
    return [tuple(b) for b in bounds]          # (left, bottom, right, top) per real cell


def _located(mesh, grid, values, nodata=maps.NODATA):
    """Point location of every pixel centre (F64, exact for these coordinates)."""
    out = np.full((grid.height, grid.width), nodata, dtype=np.float32)
    for i, (l, b, r, t) in enumerate(_boxes(mesh)):
        for row in range(grid.height):
            y = grid.top - (row + .5) * grid.vr
            if not b < y <= t:
                continue
            for col in range(grid.width):
                x = grid.left + (col + .5) * grid.hr
                if l <= x < r:
                    assert out[row, col] == nodata, 'two cells hold one pixel centre'
                    out[row, col] = values[i]
    return out


def test_area_meta_of_a_two_level_mesh():
    meta = area_meta(_mesh())
    assert (meta.min_x, meta.min_y, meta.max_x, meta.max_y) == (X0, Y0, X0 + 40, Y0 + 40)
    assert (meta.hr, meta.vr, meta.width, meta.height) == (10, 10, 4, 4)
    np.testing.assert_array_equal(meta.half_w, [10, 10, 10, 5, 5, 5, 5])
    assert meta.half_w.dtype == np.float32 and meta.x.dtype == np.float32


def test_raster_is_the_point_location_of_pixel_centres():
    mesh = _mesh()
    meta = area_meta(mesh)
    values = np.arange(1, 8, dtype=np.float32)
    image = Rasterizer(meta, oracle.HostDevice('f32')).raster(values, meta.grid())
    np.testing.assert_array_equal(image, [[2, 2, 3, 3], [2, 2, 3, 3], [1, 1, 6, 7], [1, 1, 4, 5]])
    np.testing.assert_array_equal(image, _located(mesh, meta.grid(), values))


def test_pixel_centres_on_cell_edges_follow_the_half_open_rule():
    """A grid half a pixel off the mesh puts every pixel centre on cell edges: a centre on a
    vertical edge goes to the cell on its right, one on a horizontal edge to the cell below, and
    the last column and row (on the mesh's right and bottom edges) stay nodata."""
    mesh = _mesh()
    meta = area_meta(mesh)
    grid = Grid(X0 - 5, Y0 + 45, 10., 10., 5, 5, True)
    values = np.arange(1, 8, dtype=np.float32)
    image = Rasterizer(meta, oracle.HostDevice('f32')).raster(values, grid)
    np.testing.assert_array_equal(image, _located(mesh, grid, values))
    assert np.all(image[:, 4] == maps.NODATA) and np.all(image[4] == maps.NODATA)
    assert image[0, 0] == 2 and image[1, 0] == 2 and image[2, 2] == 6


@pytest.mark.parametrize('tile', [4096, 3, 1])
def test_tiles_give_the_same_raster(tile):
    """The tile walk (each tile's corner in F32, every tile a separate launch) leaves the map as
    one pass over the whole grid draws it, down to one-pixel tiles."""
    mesh = _mesh()
    meta = area_meta(mesh)
    values = np.arange(1, 8, dtype=np.float32)
    image = Rasterizer(meta, oracle.HostDevice('f32')).raster(values, meta.grid(), tile=tile)
    np.testing.assert_array_equal(image, _located(mesh, meta.grid(), values))


def test_every_real_cell_paints_its_own_area_once():
    """On a three-level quadtree every cell paints (2 hw / hr) x (2 hh / vr) pixels, no more (no
    pixel claimed twice) and no fewer; only pixels outside the domain stay nodata."""
    from nh_flood_2d_amd import synth
    mesh = synth.small_multires(2, discharge=1.0).mesh
    meta = area_meta(mesh)
    ids = np.arange(1, mesh.e_num, dtype=np.float32)
    image = Rasterizer(meta, oracle.HostDevice('f32')).raster(ids, meta.grid(), tile=64)
    painted = np.bincount(image[image > 0].astype(np.int64), minlength=mesh.e_num)[1:]
    expected = (2 * meta.half_w / meta.hr) * (2 * meta.half_h / meta.vr)
    np.testing.assert_array_equal(painted, expected)
    area = mesh.geometry()[3][1:].sum()
    assert painted.sum() * float(meta.hr) * float(meta.vr) == area
    assert len(np.unique(meta.half_w)) >= 2


def test_masked_depth_and_extent_rasters():
    mesh = _mesh()
    meta = area_meta(mesh)
    depth = np.array([0., .01, .05, .2, .049, 1., .0], dtype=np.float32)
    masked = maps.masked_depth(depth, .02, maps.NODATA)
    np.testing.assert_array_equal(masked, np.array([-9999, -9999, .05, .2, .049, 1., -9999], np.float32))
    dev = oracle.HostDevice('f32')
    extent = maps.max_inundation_raster(meta, depth, .05, device=dev)
    np.testing.assert_array_equal(extent, _located(mesh, meta.grid(), np.where(depth >= np.float32(.05), depth,
                                                                                np.float32(-9999))))
    binary = maps.max_inundation_raster(meta, depth, .05, is_absolute=True, device=dev)
    assert binary.dtype == np.int8
    np.testing.assert_array_equal(binary, _located(mesh, meta.grid(), (depth >= np.float32(.05)).astype(np.float32)))


def test_f1_score_map_counts_and_nearest_resampling(tmp_path):
    """Truth 4 x 4 at 10 m; prediction 8 x 8 at 5 m over the same box, so the truth pixel (r, c)
    samples the prediction pixel (2r + 1, 2c + 1) (its centre's pixel); -128 marks no domain."""
    truth = np.array([[1, 1, 0, 0], [1, 0, 0, -128], [0, 0, 1, 1], [1, 0, 1, 0]], np.int8)
    wanted = np.array([[1, 0, 0, 1], [1, 1, 0, 0], [0, 1, 1, 0], [-128, 0, 1, 1]], np.int8)
    pred = np.full((8, 8), 7, np.int8)           # unsampled pixels hold a value no class uses
    pred[1::2, 1::2] = wanted
    write_geotiff(tmp_path / 't.tif', truth, X0, Y0 + 40, 10., 10., 2326, nodata=-128, dtype='<i1')
    write_geotiff(tmp_path / 'p.tif', pred, X0, Y0 + 40, 5., 5., 2326, nodata=-128, dtype='<i1')
    scores = maps.generate_f1_score_map(tmp_path / 't.tif', tmp_path / 'p.tif', tmp_path / 'f1' / 'f1.tif')
    assert (scores['tp'], scores['fp'], scores['fn'], scores['tn']) == (4, 4, 2, 4)
    assert scores['precision'] == pytest.approx(1 / 2) and scores['recall'] == pytest.approx(2 / 3)
    assert scores['f1'] == pytest.approx(4 / 7)
    out = read_geotiff(tmp_path / 'f1' / 'f1.tif')
    expected = np.array([[1, 3, 0, 2], [1, 2, 0, -128], [0, 2, 1, 3], [-128, 0, 1, 2]], np.int8)
    np.testing.assert_array_equal(out.bands[0], expected)
    assert out.nodata == -128 and (out.west, out.north, out.xres) == (X0, Y0 + 40, 10.)
    write_geotiff(tmp_path / 'q.tif', pred, X0, Y0 + 40, 5., 5., 4326, nodata=-128, dtype='<i1')
    with pytest.raises(ValueError, match='coordinate system'):
        maps.generate_f1_score_map(tmp_path / 't.tif', tmp_path / 'q.tif', tmp_path / 'x.tif')


def test_bed_timeline_replays_bed_events_and_gate_opening(tmp_path):
    """Base bed, then bed events in time order, then each commanded gate moved from its closed
    bed toward the open elevation by the fraction of its opening time since its command."""
    root = tmp_path / 'output'
    write_bed(root / 'bed.fdb', np.array([0, 1, 2, 3, 4], np.float32))
    command = 100.
    origin = 60.
    tick = round((command + origin) * 2 ** 40)
    manifest = dict(bed_events=[dict(elapsed_seconds=50., cell_ids=[1], new_bed=[9.]),
                                dict(elapsed_seconds=10., cell_ids=[2], new_bed=[7.])],
                    gate_control=dict(opening_seconds=40., open_elevation=-2., clock_origin_before_start_s=origin,
                                      gates=[dict(cell_ids=[3, 4], closed_elevation=[3., 4.],
                                                  command_seconds=command,
                                                  command_clock_ticks=[tick >> 32, tick & 0xffffffff])]))
    (tmp_path / 'manifest.json').write_text(json.dumps(manifest))
    line = BedTimeline(root, expected_count=5)
    np.testing.assert_array_equal(line.at(0.), [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(line.at(20.), [0, 1, 7, 3, 4])
    np.testing.assert_array_equal(line.at(110.), [0, 9, 7, 3 - 5 * .25, 4 - 6 * .25])
    np.testing.assert_array_equal(line.at(500.), [0, 9, 7, -2, -2])
    with pytest.raises(ValueError, match='nondecreasing'):
        line.at(100.)


@pytest.fixture()
def oracle_solver(monkeypatch):
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))


def _run(path, mm, mesh=None, stations=None):
    """1800 s of uniform rain (`mm` in all) from one raster pixel over the whole mesh."""
    mesh = mesh if mesh is not None else _mesh()
    path.mkdir(parents=True, exist_ok=True)
    bounds = mesh.geometry()[0][1:]
    west, south, east, north = bounds[:, 0].min(), bounds[:, 1].min(), bounds[:, 2].max(), bounds[:, 3].max()
    rain = cases.write_rain(path, np.full((1, 1, 1), mm, np.float32), west, north, max(east - west, north - south),
                            [(0, 1800)])
    return cases.run(path, mesh, duration=1800, rain=rain, stations=stations or {})


def test_flood_maps_of_a_run(tmp_path, oracle_solver):
    """A 1800 s storm on the 2:1 mesh: one GeoTIFF per frame (0, 900 and 1800 s), each the
    frame's depth at the pixels of the cells that reach min_h, on the mesh's grid."""
    _, _, dc = _run(tmp_path, 60.)
    dev = oracle.HostDevice('f32')
    paths = maps.generate_flood_map(dc, device=dev)
    assert [p.name for p in paths] == [f'flood_map_{i}_{(cases.START + timedelta(seconds=s)).strftime("%Y%m%d-%H%M%S")}.tif'
                                       for i, s in enumerate((0, 900, 1800))]
    mesh = solver.load_preprocessed(dc)
    meta = area_meta(mesh)
    for path, seconds in zip(paths, (0, 900, 1800)):
        image = read_geotiff(path)
        assert (image.west, image.north, image.xres, image.yres, image.epsg) == (X0, Y0 + 40, 10., 10., cases.EPSG)
        assert float(image.nodata) == maps.NODATA
        depth = read_depth_frame(f'{dc.domain_dir}/output/frames/depth_{seconds}.fdb')[1:]
        expected = _located(mesh, meta.grid(), maps.masked_depth(depth, dc.min_h, maps.NODATA))
        np.testing.assert_array_equal(image.bands[0], expected)
    assert np.all(read_geotiff(paths[-1]).bands[0] > 0)       # 30 mm on every cell by the end
    frames = list(maps.flood_video_frames(dc))
    assert [label for label, _ in frames] == [(cases.START + timedelta(seconds=s)).strftime('%Y-%m-%d  %H:%M:%S')
                                             for s in (0, 900, 1800)]
    assert all(frame.shape == (4, 4, 3) and frame.dtype == np.uint8 for _, frame in frames)
    depths = [read_geotiff(p).bands[0] for p in paths]
    wet = np.concatenate([d[d != maps.NODATA] for d in depths])
    low, high = float(wet.min()), float(wet.max())
    colours = [maps.colour_depths(d, np.float32(maps.NODATA), low, high) for d in depths]
    assert colours[-1][depths[-1] == wet.max()].tolist()[0] == [0, 0, 139]        # the deepest: dark blue
    assert colours[1][depths[1] == wet.min()].tolist()[0] == [173, 216, 230]      # the shallowest: light blue
    assert not colours[0].any()                                                   # all dry at 0 s: black
    middle = (low + high) / 2
    assert maps.colour_depths(np.array([[middle]], np.float32), None, low, high)[0, 0].tolist() == [86, 108, 184]
    from PIL import Image
    gif = maps.generate_flood_video(dc, tmp_path / 'video' / 'flood.gif')
    assert Image.open(gif).n_frames == 3
    import importlib.util
    if importlib.util.find_spec('imageio') is None:
        with pytest.raises(RuntimeError, match='imageio'):
            maps.generate_flood_video(dc, tmp_path / 'video' / 'flood.mp4')
    extent = maps.generate_max_inundation_extent_map(dc, min_depth=.02, is_absolute=True, device=dev)
    binary = read_geotiff(extent)
    maximum = read_extrema(f'{dc.domain_dir}/output/extrema.fdb').max_depth[1:]
    np.testing.assert_array_equal(binary.bands[0], _located(mesh, meta.grid(),
                                                            (maximum >= np.float32(.02)).astype(np.float32)))


def test_spatial_mae_between_two_runs(tmp_path, oracle_solver):
    """Two storms on one uniform 10 m grid (one pixel per cell): the MAE at each shared frame is
    the mean |level difference| over cells wet in both; a run against itself scores 0."""
    mesh = cases.grid(4, 3, 10.)
    _, _, a = _run(tmp_path / 'a', 60., mesh)
    _, _, b = _run(tmp_path / 'b', 30., mesh)
    dev = oracle.HostDevice('f32')
    times, errors, counts = maps.spatial_mae(a, a, .02, device=dev)
    assert errors[1:] == [0., 0.] and counts[1:] == [12, 12] and np.isnan(errors[0])
    times, errors, counts = maps.spatial_mae(a, b, .01, device=dev)
    assert len(times) == 3 and counts[1:] == [12, 12]
    for seconds, error in zip((900, 1800), errors[1:]):
        ha = read_depth_frame(f'{a.domain_dir}/output/frames/depth_{seconds}.fdb')[1:]
        hb = read_depth_frame(f'{b.domain_dir}/output/frames/depth_{seconds}.fdb')[1:]
        wet = (ha >= np.float32(.01)) & (hb >= np.float32(.01))
        assert wet.all() and error == pytest.approx(float(np.abs(ha - hb)[wet].mean()), rel=1e-6)
        assert error == pytest.approx(seconds / 1800 * .03, rel=1e-3)     # 60 against 30 mm of rain
    values = maps.plot_spatial_mae_curve(a, b, min_depth=.01, output_path=tmp_path / 'mae.png', show=False,
                                         device=dev)
    assert values[1:] == errors[1:] and (tmp_path / 'mae.png').stat().st_size > 0


def _stations_csv(path, rows):
    path.parent.mkdir(parents=True, exist_ok=True)
    lines = ['elapsed_seconds,time,station,cell_id,depth,level,u,v']
    for seconds, name, depth in rows:
        t = (cases.START + timedelta(seconds=seconds)).isoformat()
        lines.append(f'{seconds},{t},{name},1,{depth},{depth + 1},0,0')
    path.write_text('\n'.join(lines) + '\n')


def test_hydrograph_scores_and_plots(tmp_path):
    """Observed depth every 10 min; run A matches it, run B is 0.1 m higher, run C samples every
    20 min and is shifted by 600 s, so after the shift it lands on the observations."""
    from types import SimpleNamespace as NS
    from nh_flood_2d_amd import hydrograph
    obs_dir = tmp_path / 'obs'
    obs_dir.mkdir()
    truth = [0., .2, .5, .3, .1]
    lines = ['Date,Time,Water Depth (m),Water Level (m)']
    for k, value in enumerate(truth):
        t = cases.START.replace(tzinfo=None) + timedelta(minutes=10 * k)
        lines.append(f'{t:%d/%m/%Y},{t:%H:%M:%S},{value},{value + 1}')
    (obs_dir / 'S1.csv').write_text('\n'.join(lines) + '\n')
    cfgs = []
    for name, shift, step, offset in (('a', 0, 1, 0.), ('b', 0, 1, .1), ('c', 600, 2, 0.)):
        rows = [(600 * k - shift, 'S1', truth[k] + offset) for k in range(0, 5, step)]
        rows += [(600 * k - shift, 'S2', 9.) for k in range(0, 5, step)]
        _stations_csv(tmp_path / name / 'output' / 'stations.csv', rows)
        cfgs.append(NS(domain_dir=str(tmp_path / name), observation_dir=str(obs_dir)))
    observed = hydrograph.observation_series(obs_dir / 'S1.csv', 'depth')
    np.testing.assert_array_equal(observed['depth'], truth)
    rmse, nse = hydrograph.compare_hydrograph(cfgs[:2], 'S1', 'depth', show=False)
    assert rmse == pytest.approx([0., .1]) and nse[0] == 1.
    spread = np.sum((np.array(truth) - np.mean(truth)) ** 2)
    assert nse[1] == pytest.approx(1 - 5 * .01 / spread)
    rmse, _ = hydrograph.compare_hydrograph(cfgs[:2], 'S1', 'depth', baseline=cfgs[0], show=False)
    assert rmse == pytest.approx([.1])
    _, runs = hydrograph.comparison_data([cfgs[2]], 'S1', 'depth', translation_second=600)
    # the sample before the run start is not read; the others land on the 08:20 and 08:40 observations
    np.testing.assert_array_equal(runs[0][1]['depth'], [truth[2], truth[4]])
    rmse, nse = hydrograph.compare_hydrograph([cfgs[2]], 'S1', 'depth', translation_second=600, show=False)
    assert rmse == pytest.approx([0.]) and nse == [1.]
    rmse, _ = hydrograph.compare_hydrograph(cfgs[:1], 'S1', 'level', show=False,
                                            output_path=tmp_path / 'cmp.png')
    assert rmse == pytest.approx([0.]) and (tmp_path / 'cmp.png').stat().st_size > 0
    hydrograph.draw_hydrograph(cfgs[0], 'S1', 'depth', show=False, output_path=tmp_path / 'one.png')
    fig = hydrograph.compare_hydrograph_panels(cfgs[:2], ['S1', 'S2'], 'depth', show=False,
                                               output_path=tmp_path / 'panels.png', dpi=50)
    assert len(fig.axes) == 2 and (tmp_path / 'panels.png').stat().st_size > 0
    with pytest.raises(ValueError, match='No simulated samples'):
        hydrograph.station_series(cfgs[0], 'S3', 'depth')
