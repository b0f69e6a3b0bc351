"""Flood maps from a run's outputs, rasterised on the GPU (the reference's output/flood_map.py).

A map is a north-up grid over the mesh's bounding box at its finest cell size. Every cell writes
its value to the pixels whose centres it holds (``nhf_rasterize``, one thread per cell). The
grid and its tiles are formed in F32 as the reference forms them, so a map has the reference's
pixels and georeferencing.

* :func:`area_meta` is the grid of a mesh (``get_area_meta``, flood_map.py:57-115). A cell's half
  sizes are ``floor(centre - first left/bottom side midpoint + 0.5)``. The pixel size is twice the
  smallest half size. The bounding box starts from the reference's initial values
  (99999999 for the minima, 0 for the maxima).
* :class:`Rasterizer` keeps the cell geometry and one raster resident on the device and fills the
  raster tile by tile (4096 x 4096, the reference's TILE_SIZE), with each tile's corner computed as
  the reference computes it.
* :func:`generate_flood_map` writes one GeoTIFF per depth frame (flood_map.py:117-290).
* :func:`generate_max_inundation_extent_map` writes the extent from the extrema (:293-462).
* :func:`generate_f1_score_map` scores two binary extents (:464-578).
* :func:`spatial_mae` and :func:`plot_spatial_mae_curve` give the water-level error between two
  runs on their shared grid (:580-832).
* :func:`flood_video_frames` and :func:`generate_flood_video` animate the maps (:834-955).

GeoTIFFs go through this package's own writer (uncompressed strips): the pixel values and
georeferencing are the reference's, the LZW tiling and the F1 map's embedded colour table are not.
"""
from __future__ import annotations

import shutil
from dataclasses import dataclass
from datetime import timedelta
from pathlib import Path

import numpy as np

from . import abi
from .geotiff import read_geotiff, write_geotiff
from .output import BedTimeline, list_depth_frames, read_depth_frame, read_extrema

TILE = 4096
NODATA = -9999.0
LEFT, BOTTOM = 0, 2            # face columns of Mesh.face_ptr (left, right, bottom, top)
_F32 = np.float32


@dataclass
class Grid:
    """A north-up pixel grid: top-left corner, pixel sizes and shape. `f32_corners`: tile corners
    are formed in F32 from F32 corner and sizes (a mesh's own grid); otherwise in F64 and rounded
    to F32 when passed (the shared grid of two meshes)."""
    left: float
    top: float
    hr: float
    vr: float
    width: int
    height: int
    f32_corners: bool = True

    def tiles(self, size: int = TILE):
        """(row offset, column offset, rows, columns, tile left, tile top) of every tile."""
        for row_off in range(0, self.height, size):
            for col_off in range(0, self.width, size):
                rows, cols = min(size, self.height - row_off), min(size, self.width - col_off)
                if self.f32_corners:
                    top = _F32(self.top) - _F32(row_off) * _F32(self.vr)
                    left = _F32(self.left) + _F32(col_off) * _F32(self.hr)
                else:
                    top = _F32(self.top - row_off * self.vr)
                    left = _F32(self.left + col_off * self.hr)
                yield row_off, col_off, rows, cols, left, top


@dataclass
class AreaMeta:
    """Bounding box and pixel sizes (F32) of a mesh's maps and its real cells' rectangles."""
    min_x: np.float32
    min_y: np.float32
    max_x: np.float32
    max_y: np.float32
    hr: np.float32
    vr: np.float32
    x: np.ndarray
    y: np.ndarray
    half_w: np.ndarray
    half_h: np.ndarray

    @property
    def width(self) -> int:
        return int(np.ceil((self.max_x - self.min_x) / self.hr))

    @property
    def height(self) -> int:
        return int(np.ceil((self.max_y - self.min_y) / self.vr))

    def grid(self) -> Grid:
        return Grid(self.min_x, self.max_y, self.hr, self.vr, self.width, self.height, True)


def area_meta(mesh) -> AreaMeta:
    """The map grid of a mesh in F32: half sizes from each cell's first left and bottom sides,
    then the smallest pixel and the bounding box, over the real cells."""
    x, y = mesh.x[1:].astype(_F32), mesh.y[1:].astype(_F32)
    left = mesh.isl[mesh.face_ptr[1:, LEFT]]
    bottom = mesh.isl[mesh.face_ptr[1:, BOTTOM]]
    half_w = np.floor(x - mesh.sx[left].astype(_F32) + _F32(0.5))
    half_h = np.floor(y - mesh.sy[bottom].astype(_F32) + _F32(0.5))
    big, zero = _F32(99999999.0), _F32(0.)
    return AreaMeta(min(big, (x - half_w).min()), min(big, (y - half_h).min()),
                    max(zero, (x + half_w).max()), max(zero, (y + half_h).max()),
                    min(big, (half_w * _F32(2.)).min()), min(big, (half_h * _F32(2.)).min()),
                    x, y, half_w, half_h)


def masked_depth(depth: np.ndarray, min_depth: float, invalid_data: float) -> np.ndarray:
    """The display threshold on stored depth (flood_map.py:53-55)."""
    depth = np.asarray(depth, dtype=_F32)
    return np.where(depth >= _F32(min_depth), depth, _F32(invalid_data)).astype(_F32)


def _default_device():
    from .engine import TorchDevice
    return TorchDevice()


class Rasterizer:
    """Cell rectangles resident on a device (TorchDevice: the HIP kernel; the oracle's HostDevice
    runs its C restatement behind the same C ABI) and one raster buffer reused across calls."""

    def __init__(self, meta: AreaMeta, device=None):
        self.device = device if device is not None else _default_device()
        self.lib = self.device.lib
        self.n = len(meta.x)
        dev = self.device
        self.cells = [dev.array(np.ascontiguousarray(a, dtype=_F32)) for a in (meta.x, meta.y, meta.half_w,
                                                                              meta.half_h)]
        self.values = dev.array(np.zeros(max(self.n, 1), dtype=_F32))
        self.buffer, self.size = None, 0

    def _fill(self, value: float) -> None:
        dev = self.device
        if hasattr(dev, 'torch'):
            with dev.torch.cuda.stream(dev.stream):
                self.buffer.fill_(value)
        else:
            self.buffer[...] = value

    def raster(self, values: np.ndarray, grid: Grid, fill: float = NODATA, tile: int = TILE) -> np.ndarray:
        """[height, width] F32: `values` (one per real cell) on `grid`, `fill` where no cell is."""
        values = np.ascontiguousarray(values, dtype=_F32)
        if values.shape != (self.n,):
            raise ValueError(f'raster needs {self.n} cell values, got {values.shape}')
        dev = self.device
        size = grid.width * grid.height
        if size == 0:
            return np.zeros((grid.height, grid.width), dtype=_F32)
        if size > self.size:
            self.buffer, self.size = dev.array(np.zeros(size, dtype=_F32)), size
        dev.write(self.values, values)
        self._fill(fill)
        self.enqueue(grid, tile)
        return dev.read(self.buffer, _F32)[:size].reshape(grid.height, grid.width).copy()

    def enqueue(self, grid: Grid, tile: int = TILE) -> None:
        """Launch every tile of `grid` on the device's stream over the values already uploaded,
        into the raster buffer (its size checked by raster()); returns without waiting."""
        dev = self.device
        if grid.width * grid.height > self.size:
            raise ValueError('enqueue needs a raster buffer of the grid size: call raster() first')
        base = dev.ptr(self.buffer)
        ptrs = [dev.ptr(a) for a in self.cells]
        for row_off, col_off, rows, cols, left, top in grid.tiles(tile):
            rc = self.lib.nhf_rasterize(*ptrs, dev.ptr(self.values), self.n, float(left), float(top), float(grid.hr),
                                        float(grid.vr), rows, cols, base + 4 * (row_off * grid.width + col_off),
                                        grid.width, dev.stream_handle())
            abi.check(self.lib, rc, 'nhf_rasterize')


def _mesh_and_root(cfg):
    from .solver_compact import load_preprocessed
    ne, ns, output = Path(cfg.ne_fdb), Path(cfg.ns_fdb), Path(cfg.domain_dir) / 'output'
    if not ne.exists() or not ns.exists():
        raise FileNotFoundError(f'Preprocessed mesh not found: {ne}, {ns}')
    return load_preprocessed(cfg), output


def generate_flood_map(cfg, device=None) -> list[Path]:
    """One F32 GeoTIFF per depth frame, depth where it reaches min_h and -9999 elsewhere, named
    flood_map_<index>_<wall time YYYYmmdd-HHMMSS>.tif in the domain's flood_maps directory (which
    is emptied first). Returns the paths in frame order."""
    mesh, output = _mesh_and_root(cfg)
    if not (output / 'frames').exists():
        raise FileNotFoundError(f'Depth frames not found: {output / "frames"}')
    out = Path(cfg.domain_dir) / 'flood_maps'
    if out.exists():
        shutil.rmtree(out)
    out.mkdir(parents=True)
    meta = area_meta(mesh)
    grid = meta.grid()
    raster = Rasterizer(meta, device)
    paths = []
    for index, (seconds, frame) in enumerate(list_depth_frames(output)):
        depth = read_depth_frame(frame, expected_count=mesh.e_num)
        image = raster.raster(masked_depth(depth[1:], cfg.min_h, NODATA), grid)
        stamp = (cfg.start + timedelta(seconds=seconds)).strftime('%Y%m%d-%H%M%S')
        path = out / f'flood_map_{index}_{stamp}.tif'
        write_geotiff(path, image, meta.min_x, meta.max_y, meta.hr, meta.vr, cfg.epsg_code, nodata=NODATA)
        paths.append(path)
    return paths


def max_inundation_raster(meta: AreaMeta, max_depth: np.ndarray, min_depth: float = 0.05,
                          invalid_data: float = NODATA, is_absolute: bool = False, device=None) -> np.ndarray:
    """The extent raster of per-cell maximum depths (real cells): the maximum where it reaches
    `min_depth` and `invalid_data` elsewhere, or with `is_absolute` int8 1 (reached) / 0 (not) /
    -128 (no cell)."""
    maximum = np.asarray(max_depth, dtype=_F32)
    reached = maximum >= _F32(min_depth)
    raster = Rasterizer(meta, device)
    if is_absolute:
        image = raster.raster(np.where(reached, _F32(1.), _F32(0.)), meta.grid(), fill=-128.)
        return image.astype(np.int8)
    image = raster.raster(np.where(reached, maximum, _F32(invalid_data)), meta.grid(), fill=invalid_data)
    return np.where(image >= _F32(min_depth), image, _F32(invalid_data)).astype(_F32)


def generate_max_inundation_extent_map(cfg, min_depth: float = 0.05, invalid_data: float = NODATA,
                                       is_absolute: bool = False, device=None) -> Path:
    """max_inundation/max_inundation.tif from output/extrema.fdb (see max_inundation_raster).
    The binary map's nodata is -128 whatever `invalid_data` is: the reference casts its
    `invalid_data` fill to int8 there, which only -128 survives."""
    mesh, output = _mesh_and_root(cfg)
    extrema_path = output / 'extrema.fdb'
    if not extrema_path.exists():
        raise FileNotFoundError(f'Extrema not found: {extrema_path}')
    meta = area_meta(mesh)
    maximum = read_extrema(extrema_path, expected_count=mesh.e_num).max_depth[1:]
    image = max_inundation_raster(meta, maximum, min_depth, invalid_data, is_absolute, device)
    path = Path(cfg.max_inundation_dir) / 'max_inundation.tif'
    write_geotiff(path, image, meta.min_x, meta.max_y, meta.hr, meta.vr, cfg.epsg_code,
                  nodata=-128 if is_absolute else invalid_data, dtype='<i1' if is_absolute else '<f4')
    return path


def nearest_on(source, target, invalid_data: int) -> np.ndarray:
    """`source`'s first band resampled to `target`'s grid by nearest neighbour: each target pixel
    takes the source pixel holding its centre; outside the source, or at its nodata, invalid."""
    height, width = target.bands.shape[1:]
    xs = target.west + (np.arange(width) + 0.5) * target.xres
    ys = target.north - (np.arange(height) + 0.5) * target.yres
    cols = np.floor((xs - source.west) / source.xres + 1e-10).astype(np.int64)
    rows = np.floor((source.north - ys) / source.yres + 1e-10).astype(np.int64)
    band = source.bands[0].astype(np.int16)
    sh, sw = band.shape
    inside = (rows[:, None] >= 0) & (rows[:, None] < sh) & (cols[None, :] >= 0) & (cols[None, :] < sw)
    out = np.full((height, width), invalid_data, dtype=np.int16)
    rr, cc = np.broadcast_to(rows[:, None], inside.shape), np.broadcast_to(cols[None, :], inside.shape)
    out[inside] = band[rr[inside], cc[inside]]
    if source.nodata is not None:
        out[inside & (out == int(source.nodata))] = invalid_data
    return out


def generate_f1_score_map(truth_path, pred_path, output_path, invalid_data: int = -128) -> dict:
    """Classify the prediction extent against the truth extent (binary int8 maps from
    generate_max_inundation_extent_map(is_absolute=True)) on the truth grid: 1 TP, 2 FP, 3 FN,
    0 TN, `invalid_data` where either is outside its domain. Writes the int8 map and returns f1,
    precision, recall and the four counts. Both maps must be in one coordinate system."""
    truth, pred = read_geotiff(truth_path), read_geotiff(pred_path)
    if truth.epsg != pred.epsg:
        raise ValueError('generate_f1_score_map needs both maps in one coordinate system')
    t = truth.bands[0].astype(np.int16)
    p = nearest_on(pred, truth, invalid_data)
    invalid = (t == invalid_data) | (p == invalid_data)
    tp, fp = (t == 1) & (p == 1) & ~invalid, (t == 0) & (p == 1) & ~invalid
    fn, tn = (t == 1) & (p == 0) & ~invalid, (t == 0) & (p == 0) & ~invalid
    result = np.zeros(t.shape, dtype=np.int8)
    result[tp], result[fp], result[fn] = 1, 2, 3
    result[invalid] = np.int8(invalid_data)
    counts = {k: int(v.sum()) for k, v in (('tp', tp), ('fp', fp), ('fn', fn), ('tn', tn))}
    precision = counts['tp'] / (counts['tp'] + counts['fp']) if counts['tp'] + counts['fp'] else 0.
    recall = counts['tp'] / (counts['tp'] + counts['fn']) if counts['tp'] + counts['fn'] else 0.
    f1 = 2 * precision * recall / (precision + recall) if precision + recall else 0.
    out = Path(output_path)
    out.parent.mkdir(parents=True, exist_ok=True)
    write_geotiff(out, result, truth.west, truth.north, truth.xres, truth.yres, truth.epsg,
                  nodata=invalid_data, dtype='<i1')
    return dict(f1=f1, precision=precision, recall=recall, **counts)


def shared_grid(a: AreaMeta, b: AreaMeta) -> Grid:
    """The grid two meshes are compared on: their bounding boxes' intersection at the finer
    pixel size, corners and sizes as F64 (flood_map.py:643-654)."""
    left, bottom = float(max(a.min_x, b.min_x)), float(max(a.min_y, b.min_y))
    right, top = float(min(a.max_x, b.max_x)), float(min(a.max_y, b.max_y))
    if left >= right or bottom >= top:
        raise ValueError('The bounding boxes of the two runs do not overlap.')
    hr, vr = float(min(a.hr, b.hr)), float(min(a.vr, b.vr))
    return Grid(left, top, hr, vr, int(np.ceil((right - left) / hr)), int(np.ceil((top - bottom) / vr)), False)


def water_levels(depth: np.ndarray, bed: np.ndarray, min_depth: float) -> np.ndarray:
    """Real cells' water level bed + depth (F32), -9999 where depth is below `min_depth`."""
    level = (np.asarray(bed)[1:] + np.asarray(depth)[1:]).astype(_F32)
    return np.where(np.asarray(depth)[1:] < _F32(min_depth), _F32(NODATA), level).astype(_F32)


def mae_of(ref: np.ndarray, cmp: np.ndarray, tile: int = TILE) -> tuple[float, int]:
    """Mean |ref - cmp| over pixels wet in both (> -9000), summed tile by tile in F32 (NumPy's
    pairwise sum) and across tiles in F64, as the reference reduces it."""
    total, count = 0.0, 0
    for r0 in range(0, ref.shape[0], tile):
        for c0 in range(0, ref.shape[1], tile):
            a, b = ref[r0:r0 + tile, c0:c0 + tile], cmp[r0:r0 + tile, c0:c0 + tile]
            mask = (a > -9000.0) & (b > -9000.0)
            if mask.any():
                total += float(np.sum(np.abs(a[mask] - b[mask])))
                count += int(mask.sum())
    return (total / count if count else float('nan')), count


def spatial_mae(cfg_ref, cfg_cmp, min_depth: float = 0.05, device=None):
    """Per shared frame time: the mean absolute water-level difference over pixels wet in both
    runs, on their shared grid; returns (wall times, MAE in m, wet pixel counts)."""
    mesh_r, out_r = _mesh_and_root(cfg_ref)
    mesh_c, out_c = _mesh_and_root(cfg_cmp)
    meta_r, meta_c = area_meta(mesh_r), area_meta(mesh_c)
    grid = shared_grid(meta_r, meta_c)
    frames_r = {cfg_ref.start + timedelta(seconds=s): (s, p) for s, p in list_depth_frames(out_r)}
    frames_c = {cfg_cmp.start + timedelta(seconds=s): (s, p) for s, p in list_depth_frames(out_c)}
    common = sorted(set(frames_r) & set(frames_c))
    if not common:
        raise ValueError('No common depth-frame times found between the two runs.')
    raster_r = Rasterizer(meta_r, device)
    raster_c = Rasterizer(meta_c, raster_r.device)
    bed_r = BedTimeline(out_r, expected_count=mesh_r.e_num)
    bed_c = BedTimeline(out_c, expected_count=mesh_c.e_num)
    times, errors, counts = [], [], []
    for instant in common:
        (s_r, p_r), (s_c, p_c) = frames_r[instant], frames_c[instant]
        depth_r = read_depth_frame(p_r, expected_count=mesh_r.e_num)
        depth_c = read_depth_frame(p_c, expected_count=mesh_c.e_num)
        image_r = raster_r.raster(water_levels(depth_r, bed_r.at(s_r), min_depth), grid)
        image_c = raster_c.raster(water_levels(depth_c, bed_c.at(s_c), min_depth), grid)
        mae, count = mae_of(image_r, image_c)
        times.append(instant.replace(tzinfo=None))
        errors.append(mae)
        counts.append(count)
    return times, errors, counts


def rainfall_plot_series(domain_cfg, force_cfg):
    """Interval starts (naive wall times), mesh-mean rain in mm/h and interval widths in s, clipped
    to the run (flood_map.py:23-50)."""
    from .forcing import RainField
    mesh, _ = _mesh_and_root(domain_cfg)
    rain = RainField.load(force_cfg.rainfall, mesh, domain_cfg.epsg_code, domain_cfg.start, 0, domain_cfg.duration)
    area = mesh.geometry()[3][1:]
    pixel_area = np.bincount(rain.pixel_index[1:], weights=area, minlength=rain.millimetres.shape[1])
    total = pixel_area.sum(dtype=np.float64)
    if not np.isfinite(total) or total <= 0:
        raise ValueError('Rainfall plot requires positive finite mesh area')
    mean_mm = rain.millimetres @ pixel_area / total
    starts, ends = np.maximum(rain.starts, 0.), np.minimum(rain.ends, float(domain_cfg.duration))
    keep = ends > starts
    rates = mean_mm[keep] * (3600. / (rain.ends[keep] - rain.starts[keep]))
    dates = [(domain_cfg.start + timedelta(seconds=float(t))).replace(tzinfo=None) for t in starts[keep]]
    return dates, rates, ends[keep] - starts[keep]


def plot_spatial_mae_curve(cfg_ref, cfg_cmp, force_cfg=None, min_depth: float = 0.05, output_path=None,
                           show: bool = True, device=None) -> list[float]:
    """The spatial MAE of water level against time, with the mesh-mean hyetograph on an inverted
    second axis when `force_cfg` is given; saved to `output_path` and/or shown. Returns the MAE
    list in time order."""
    import matplotlib
    if not show:
        matplotlib.use('Agg')
    import matplotlib.dates as mdates
    import matplotlib.pyplot as plt
    times, errors, _ = spatial_mae(cfg_ref, cfg_cmp, min_depth, device)
    fig, ax1 = plt.subplots(figsize=(14, 5))
    ax1.plot(times, errors, color='steelblue', linewidth=2, label='Spatial MAE (m)')
    ax1.set_xlabel('Time', fontweight='bold')
    ax1.set_ylabel('Spatial MAE of WSE (m)', color='steelblue', fontweight='bold')
    ax1.xaxis.set_major_formatter(mdates.DateFormatter('%m/%d %H:%M'))
    fig.autofmt_xdate()
    if force_cfg is not None:
        dates, rates, widths = rainfall_plot_series(cfg_ref, force_cfg)
        if dates:
            ax2 = ax1.twinx()
            ax2.bar(dates, rates, width=widths / 86400., align='edge', alpha=0.35, color='dimgray',
                    label='Rainfall (mm/hr)')
            ax2.set_ylabel('Rainfall (mm/hr)', color='dimgray', fontweight='bold')
            ax2.invert_yaxis()
            ax2.legend(loc='upper right')
    ref, cmp = Path(cfg_ref.domain_dir).name, Path(cfg_cmp.domain_dir).name
    ax1.set_title(f'Spatial MAE of WSE - {ref} vs {cmp} (min_depth = {min_depth} m)', fontweight='bold')
    ax1.legend(loc='upper left')
    fig.tight_layout()
    if output_path:
        fig.savefig(output_path, dpi=150, bbox_inches='tight')
    if show:
        plt.show()
    plt.close(fig)
    return errors


SHALLOW = np.array([173, 216, 230], dtype=_F32)     # light blue
DEEP = np.array([0, 0, 139], dtype=_F32)            # dark blue


def colour_depths(data: np.ndarray, nodata, low: float, high: float) -> np.ndarray:
    """[h, w, 3] uint8: finite pixels other than `nodata` on the light-to-dark blue ramp over
    [low, high] (F32, truncated to 8 bits), the rest black."""
    frame = np.zeros(data.shape + (3,), dtype=np.uint8)
    mask = ((data != nodata) & np.isfinite(data)) if nodata is not None else np.isfinite(data)
    if mask.any():
        t = np.clip((data[mask] - low) / (high - low), 0.0, 1.0)
        frame[mask] = ((1.0 - t[:, None]) * SHALLOW + t[:, None] * DEEP).astype(np.uint8)
    return frame


def flood_video_frames(cfg):
    """(label, RGB frame) of every flood map in index order (generate_flood_video,
    flood_map.py:834-955): one depth scale over all maps (their minimum to maximum, 0-1 m when
    flat), each wet pixel coloured linearly from light blue (shallow) to dark blue (deep) and
    truncated to 8 bits, dry and nodata pixels black, the map's wall time drawn top right."""
    from datetime import datetime as wall
    from PIL import Image, ImageDraw, ImageFont
    paths = sorted(Path(cfg.flood_map_dir).glob('flood_map_*.tif'), key=lambda p: int(p.stem.split('_')[2]))
    if not paths:
        return
    first = read_geotiff(paths[0])
    nodata = first.nodata
    height, width = first.bands.shape[1:]

    def valid(data):
        return (data != nodata) & np.isfinite(data) if nodata is not None else np.isfinite(data)
    low, high = np.inf, -np.inf
    for path in paths:
        data = read_geotiff(path).bands[0]
        mask = valid(data)
        if mask.any():
            low, high = min(low, float(data[mask].min())), max(high, float(data[mask].max()))
    if low >= high:
        low, high = 0.0, 1.0
    size = max(14, int(width * 0.015))
    try:
        font = ImageFont.truetype('arial.ttf', size)
    except OSError:
        font = ImageFont.load_default()
    pad = max(6, size // 3)
    for path in paths:
        frame = colour_depths(read_geotiff(path).bands[0], nodata, low, high)
        label = wall.strptime(path.stem.split('_')[-1], '%Y%m%d-%H%M%S').strftime('%Y-%m-%d  %H:%M:%S')
        image = Image.fromarray(frame)
        draw = ImageDraw.Draw(image)
        box = draw.textbbox((0, 0), label, font=font)
        x, y = width - (box[2] - box[0]) - 2 * pad, pad
        draw.rectangle([x - pad, y - pad, x + (box[2] - box[0]) + pad, y + (box[3] - box[1]) + pad], fill=(0, 0, 0))
        draw.text((x, y), label, font=font, fill=(255, 255, 255))
        yield label, np.array(image)


def generate_flood_video(cfg, output_path, interval_second: float = 0.2) -> Path | None:
    """The flood maps as an animation, one map per `interval_second`: an animated GIF (Pillow)
    for a .gif path, an H.264 MP4 for .mp4 when imageio with ffmpeg is installed. Returns the
    path, or None when there are no maps."""
    from PIL import Image
    frames = flood_video_frames(cfg)
    first = next(frames, None)
    if first is None:
        return None
    out = Path(output_path)
    out.parent.mkdir(parents=True, exist_ok=True)
    if out.suffix.lower() == '.mp4':
        try:
            import imageio
        except ImportError as exc:
            raise RuntimeError('an MP4 needs imageio with ffmpeg; give a .gif path instead') from exc
        with imageio.get_writer(str(out), fps=1.0 / interval_second, codec='libx264', quality=8,
                                macro_block_size=1) as writer:
            writer.append_data(first[1])
            for _, frame in frames:
                writer.append_data(frame)
        return out
    images = [Image.fromarray(first[1])] + [Image.fromarray(frame) for _, frame in frames]
    images[0].save(out, save_all=True, append_images=images[1:], duration=int(round(interval_second * 1000)), loop=0)
    return out
