"""Absolute-time forcing: boundary series, banded rainfall and forcing events.

Semantics follow ``src/nh_flood_2d/preprocess/force.py``:
  TimeSeries.at          linear interpolation of the series (force.py:100-101)
  RainField.index_at     band containing t (right-closed search, force.py:198-199)
  RainField.rates        band millimetres -> m/s (force.py:201-202)
  Forcing.events         rain band ends, 0, end and every boundary knot (force.py:217)
  load_boundaries        exterior edges whose segment overlaps the boundary line with positive
                         length, length-weighted cell shares, no shared cells (force.py:136-160)
File loaders read the reference's inputs (CSV series, CSV intervals, GeoJSON lines, GeoTIFF
rain) with numpy and the standard library only.
"""
from __future__ import annotations

import csv
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .config import absolute_time
from .mesh import Mesh


def read_series_csv(path, origin) -> tuple[np.ndarray, np.ndarray]:
    """(seconds from `origin`, values) of a ``time,value`` CSV, in file order."""
    with open(path, newline='') as f:
        pairs = [((absolute_time(r['time']) - origin).total_seconds(), float(r['value'])) for r in csv.DictReader(f)]
    table = np.array(pairs, dtype=np.float64).reshape(-1, 2)
    return table[:, 0].copy(), table[:, 1].copy()


@dataclass
class TimeSeries:
    """A boundary's prescribed level or discharge: knots in seconds from the simulation start
    (F64), linear between knots (preprocess/force.py:84-101)."""
    times: np.ndarray
    values: np.ndarray

    @classmethod
    def load(cls, path, origin, begin, end):
        """Read and check a series: at least two knots, finite values, strictly increasing
        times, and knots at or beyond both ends of [begin, end]."""
        times, values = read_series_csv(path, origin)
        usable = len(times) >= 2 and bool(np.isfinite(values).all()) and bool((np.diff(times) > 0).all())
        if not usable:
            raise ValueError(f'{path}: need finite values and strictly increasing times (at least two)')
        if not (times[0] <= begin and times[-1] >= end):
            raise ValueError(f'{path}: series does not cover [{begin}, {end}] seconds from start')
        return cls(times, values)

    def at(self, seconds: float) -> float:
        return float(np.interp(seconds, self.times, self.values))


@dataclass
class RainField:
    starts: np.ndarray
    ends: np.ndarray
    millimetres: np.ndarray   # [bands, pixels]
    pixel_index: np.ndarray   # [e_num] cell -> pixel (cell 0 unused)

    def index_at(self, seconds: float) -> int:
        return int(np.searchsorted(self.ends, seconds, side='right'))

    def rates(self, index: int) -> np.ndarray:
        return self.millimetres[index] * (0.001 / (self.ends[index] - self.starts[index]))

    @classmethod
    def load(cls, config, mesh: Mesh, epsg: int, origin, begin, end):
        from .geotiff import read_geotiff
        with open(config.intervals, newline='') as f:
            rows = list(csv.DictReader(f))
        starts = np.array([(absolute_time(r['interval_start']) - origin).total_seconds() for r in rows])
        ends = np.array([(absolute_time(r['interval_end']) - origin).total_seconds() for r in rows])
        if not len(rows) or np.any(ends <= starts) or np.any(starts[1:] != ends[:-1]):
            raise ValueError('Rainfall intervals must be continuous, positive and nonoverlapping')
        if starts[0] > begin or ends[-1] < end:
            raise ValueError('Rainfall intervals do not cover the simulation window')
        raster = read_geotiff(config.raster)
        if raster.epsg != epsg:
            raise ValueError('Rainfall CRS differs from mesh CRS')
        if raster.bands.shape[0] != len(rows):
            raise ValueError('Rainfall band count differs from interval count')
        rr, cc = raster.rowcol(mesh.x[1:].astype(float), mesh.y[1:].astype(float))
        height, width = raster.bands.shape[1:]
        if np.any((rr < 0) | (rr >= height) | (cc < 0) | (cc >= width)):
            raise ValueError('Cell centre is outside the rainfall raster')
        pixels, inverse = np.unique(rr * width + cc, return_inverse=True)
        samples = raster.bands[:, pixels // width, pixels % width]
        rain = samples.astype(float)
        if np.any(raster.is_nodata(samples) | ~np.isfinite(rain) | (rain < 0)):
            raise ValueError('Rainfall at a cell centre is NoData, nonfinite or negative')
        mapping = np.zeros(mesh.e_num, dtype=np.int32)
        mapping[1:] = inverse
        return cls(starts, ends, rain, mapping)


@dataclass
class Boundary:
    """A prescribed boundary line: the interior cells of the exterior sides it covers, each with
    its share of the selected length (F64, summing to 1), and its series."""
    name: str
    kind: str            # 'level' | 'discharge'
    cells: np.ndarray
    weights: np.ndarray
    side_ids: np.ndarray
    series: TimeSeries

    def distribute(self, q):
        """A total discharge split over the boundary's cells by length share."""
        return q * self.weights


@dataclass
class Forcing:
    boundaries: list
    rain: RainField
    events: np.ndarray
    end: float

    @classmethod
    def build(cls, boundaries, rain, end):
        """Events are rain band ends, 0, the end and every boundary knot (force.py:217)."""
        events = np.unique(np.concatenate([rain.ends, [0., float(end)]] +
                                          [b.series.times for b in boundaries]))
        return cls(boundaries, rain, events, float(end))


# ----------------------------------------------------------------------------- boundary lines

def _line_epsg(data) -> int | None:
    try:
        name = data['crs']['properties']['name']
    except (KeyError, TypeError):
        return None
    digits = ''.join(ch if ch.isdigit() else ' ' for ch in str(name)).split()
    return int(digits[-1]) if digits else None


def load_line(path, epsg: int) -> list[np.ndarray]:
    """GeoJSON LineString / MultiLineString (or features of them) as vertex arrays."""
    with open(path) as f:
        data = json.load(f)
    if 'crs' not in data:
        raise ValueError(f'{path}: GeoJSON must declare its CRS')
    if _line_epsg(data) != epsg:
        raise ValueError(f'{path}: boundary CRS differs from mesh EPSG:{epsg}')
    if data['type'] == 'FeatureCollection':
        geometries = [f['geometry'] for f in data['features']]
    elif data['type'] == 'Feature':
        geometries = [data['geometry']]
    else:
        geometries = [data]
    lines = []
    for g in geometries:
        if g['type'] == 'LineString':
            parts = [g['coordinates']]
        elif g['type'] == 'MultiLineString':
            parts = g['coordinates']
        else:
            raise ValueError(f'{path}: expected nonempty LineString or MultiLineString')
        for p in parts:
            arr = np.asarray(p, dtype=float)[:, :2]
            if len(arr) < 2:
                raise ValueError(f'{path}: expected nonempty LineString or MultiLineString')
            lines.append(arr)
    if not lines:
        raise ValueError(f'{path}: expected nonempty LineString or MultiLineString')
    return lines


def _overlap_length(segments: np.ndarray, line: np.ndarray) -> np.ndarray:
    """Length of the collinear overlap of axis-parallel edge segments with a polyline."""
    total = np.zeros(len(segments))
    a, b = segments[:, 0, :], segments[:, 1, :]
    vertical = a[:, 0] == b[:, 0]
    for p, r in zip(line[:-1], line[1:]):
        d = r - p
        if d[0] == 0 and d[1] == 0:
            continue
        if d[0] == 0:   # vertical polyline piece
            sel = vertical & (a[:, 0] == p[0])
            lo, hi = min(p[1], r[1]), max(p[1], r[1])
            ov = np.minimum(np.maximum(a[:, 1], b[:, 1]), hi) - np.maximum(np.minimum(a[:, 1], b[:, 1]), lo)
            total += np.where(sel, np.maximum(ov, 0), 0)
        elif d[1] == 0:  # horizontal polyline piece
            sel = ~vertical & (a[:, 1] == p[1])
            lo, hi = min(p[0], r[0]), max(p[0], r[0])
            ov = np.minimum(np.maximum(a[:, 0], b[:, 0]), hi) - np.maximum(np.minimum(a[:, 0], b[:, 0]), lo)
            total += np.where(sel, np.maximum(ov, 0), 0)
        # an oblique piece meets axis-parallel edges in points only: no positive length
    return total


def select_boundary_edges(mesh: Mesh, lines: list[np.ndarray]):
    """Exterior sides overlapping the lines with positive length, in exterior-edge order."""
    side_ids, cells, _, lengths, segments = mesh.exterior_edges()
    overlap = np.zeros(len(side_ids))
    for line in lines:
        overlap += _overlap_length(segments, line)
    selected = np.flatnonzero(overlap > 0)
    return side_ids, cells, lengths, selected


def make_boundary(mesh: Mesh, name: str, kind: str, side_ids: np.ndarray, series: TimeSeries,
                  claimed: np.ndarray | None = None) -> Boundary:
    """Boundary over given exterior sides: unique cells with length-weighted shares."""
    all_sides, cells_of, lengths_of, _ = select_boundary_edges(mesh, [])
    pos = np.searchsorted(all_sides, side_ids)
    if np.any(pos >= len(all_sides)) or np.any(all_sides[np.minimum(pos, len(all_sides) - 1)] != side_ids):
        raise ValueError(f'{name}: boundary sides must be exterior sides')
    cells, inverse = np.unique(cells_of[pos], return_inverse=True)
    if claimed is not None:
        if np.any(claimed[cells]):
            raise ValueError(f'{name}: boundary cells overlap another boundary')
        claimed[cells] = True
    lengths = np.bincount(inverse, weights=lengths_of[pos])
    weights = lengths / lengths.sum(dtype=np.float64)
    return Boundary(name, kind, cells, weights, np.asarray(side_ids), series)


def load_boundaries(configs, mesh: Mesh, epsg: int, origin, begin, end) -> list[Boundary]:
    if not configs:
        return []
    claimed = np.zeros(mesh.e_num, dtype=bool)
    out = []
    for b in configs:
        lines = load_line(b.line, epsg)
        side_ids, _, _, selected = select_boundary_edges(mesh, lines)
        chosen = np.sort(side_ids[selected])
        if len(chosen) == 0:
            raise ValueError(f'{b.name}: boundary selects no exterior edges; '
                             'snap the line to mesh exterior edges in GIS')
        out.append(make_boundary(mesh, b.name, b.kind, chosen,
                                 TimeSeries.load(b.series, origin, begin, end), claimed))
    return out


def load_forcing(domain, force, mesh: Mesh) -> Forcing:
    """Forcing.load (force.py:211-218) for file inputs."""
    begin, end = 0., float(domain.duration)
    boundaries = load_boundaries(force.boundaries, mesh, domain.epsg_code, domain.start, begin, end)
    rain = RainField.load(force.rainfall, mesh, domain.epsg_code, domain.start, begin, end)
    return Forcing.build(boundaries, rain, end)


def cell_forcing_arrays(mesh: Mesh, forcing: Forcing, n_pixels: int | None = None):
    """Per-cell boundary ids, discharge shares and the selected-side mask (solver_compact.py:244-255)."""
    from .layout import Forcing as LayoutForcing
    e_num = mesh.e_num
    boundary_id = np.full(e_num, -1, dtype=np.int32)
    weights = np.zeros(e_num, dtype=np.float32)
    kinds = np.zeros(len(forcing.boundaries), dtype=np.int32)
    selected = np.zeros(mesh.s_num, dtype=bool)
    for i, b in enumerate(forcing.boundaries):
        boundary_id[b.cells] = i
        weights[b.cells] = b.weights
        kinds[i] = int(b.kind == 'discharge')
        selected[b.side_ids] = True
    pixels = forcing.rain.millimetres.shape[1] if n_pixels is None else n_pixels
    return LayoutForcing(forcing.rain.pixel_index.astype(np.int32), int(pixels), boundary_id,
                         weights, kinds, selected)
