"""Frames, extrema, bed and checkpoints in the reference's logical FDB layout.

The tables, column names and dtypes are the reference's (``core/output.py`` and
``schema/feature.py``): ``F32Value 'depth'`` frames of e_num cells, ``F32Value 'bed'``,
``F32Value max_depth/max_speed`` + ``IndexLike first_time_hi/lo`` extrema, and checkpoints with
F32 tables depth/bed/q/rain_time/infiltration/max_depth/max_speed, one-row ``IndexLike``
clock_hi/clock_lo/gate_open, per-cell first_time_hi/lo and the ``U8Value 'state_meta'`` JSON.
Validation and the atomic sibling-temporary replace follow ``core/output.py:53-243``.

The byte container is this package's own: fastdb4py 0.2.4's on-disk format is not available
here, so files carry a small self-describing header (magic, JSON table directory) followed by the
little-endian column data. Reading and writing go through :func:`save_tables` / :func:`load_tables`
only, so a byte-compatible fastdb4py backend can replace them without touching callers.
"""
from __future__ import annotations

import csv
import re
import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

MAGIC = b'NHFDB\x01\x00\x00'
FEATURES = {'F32Value': ('value', '<f4'), 'IndexLike': ('index', '<u4'), 'U8Value': ('value', 'u1')}
STATION_FIELDS = ('elapsed_seconds', 'time', 'station', 'cell_id', 'depth', 'level', 'u', 'v')
_CHECKPOINT_FLOATS = ('depth', 'bed', 'q', 'rain_time', 'infiltration', 'max_depth', 'max_speed')
_CHECKPOINT_INTS = ('clock_hi', 'clock_lo', 'gate_open')
_CHECKPOINT_TIME = ('first_time_hi', 'first_time_lo')


def save_tables(path, tables: list[tuple[str, str, np.ndarray]]) -> None:
    """Write (feature, name, column) tables atomically via a sibling temporary file."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    directory, blobs, offset = [], [], 0
    for feature, name, values in tables:
        column, dtype = FEATURES[feature]
        data = np.ascontiguousarray(values, dtype=dtype).tobytes()
        directory.append(dict(feature=feature, name=name, column=column, dtype=dtype,
                              rows=len(values), offset=offset, nbytes=len(data)))
        blobs.append(data)
        offset += len(data)
    header = json.dumps(directory).encode('utf-8')
    temporary = path.with_name(f'.{path.name}.{os.getpid()}.tmp')
    with open(temporary, 'wb') as f:
        f.write(MAGIC)
        f.write(len(header).to_bytes(8, 'little'))
        f.write(header)
        for blob in blobs:
            f.write(blob)
    os.replace(temporary, path)


def load_tables(path) -> dict[str, np.ndarray]:
    with open(path, 'rb') as f:
        if f.read(8) != MAGIC:
            raise ValueError(f'{path}: not a frame/checkpoint table file')
        n = int.from_bytes(f.read(8), 'little')
        directory = json.loads(f.read(n).decode('utf-8'))
        base = f.tell()
        out = {}
        for entry in directory:
            f.seek(base + entry['offset'])
            out[entry['name']] = np.frombuffer(f.read(entry['nbytes']), dtype=entry['dtype']).copy()
    return out


def _float_array(value, name):
    array = np.asarray(value, dtype=np.float32)
    if array.ndim != 1 or not np.all(np.isfinite(array)):
        raise ValueError(f'{name} must be a finite one-dimensional F32 array')
    return array


def _uint_array(value, name):
    array = np.asarray(value)
    if (array.ndim != 1 or not np.issubdtype(array.dtype, np.number) or
            not np.all(np.isfinite(array)) or np.any(array != np.floor(array)) or
            np.any(array < 0) or np.any(array > np.iinfo(np.uint32).max)):
        raise ValueError(f'{name} must be a one-dimensional U32 array')
    return array.astype(np.uint32)


def _read_named(path, name, expected_count=None):
    values = load_tables(path)[name]
    if expected_count is not None and len(values) != expected_count:
        raise ValueError(f'{path}: {name} has {len(values)} rows, expected {expected_count}')
    return values


def write_bed(path, bed):
    save_tables(path, [('F32Value', 'bed', _float_array(bed, 'bed'))])


def read_bed(path, expected_count=None):
    return _read_named(path, 'bed', expected_count)


def write_depth_frame(path, depth):
    values = _float_array(depth, 'depth')
    if np.any(values < 0):
        raise ValueError('depth must be nonnegative')
    save_tables(path, [('F32Value', 'depth', values)])


def read_depth_frame(path, expected_count=None):
    values = _read_named(path, 'depth', expected_count)
    if not np.all(np.isfinite(values)) or np.any(values < 0):
        raise ValueError(f'{path}: invalid depth frame')
    return values


FRAME_NAME = re.compile(r'depth_(.*)')


def list_depth_frames(output_root):
    """(elapsed seconds, path) of every ``frames/depth_<s>.fdb`` under output_root, by time; a
    frame whose name is not a whole number of seconds is an error (core/output.py:118-127)."""
    def seconds_of(path):
        try:
            return int(FRAME_NAME.fullmatch(path.stem).group(1))
        except ValueError as exc:
            raise ValueError(f'Invalid depth frame name: {path}') from exc
    paths = list((Path(output_root) / 'frames').glob('depth_*.fdb'))
    keyed = [(seconds_of(p), p) for p in paths]
    return sorted(keyed)


@dataclass
class Extrema:
    max_depth: np.ndarray
    first_time_hi: np.ndarray
    first_time_lo: np.ndarray
    max_speed: np.ndarray


def write_extrema(path, max_depth, first_time_hi, first_time_lo, max_speed):
    arrays = {'max_depth': _float_array(max_depth, 'max_depth'),
              'max_speed': _float_array(max_speed, 'max_speed'),
              'first_time_hi': _uint_array(first_time_hi, 'first_time_hi'),
              'first_time_lo': _uint_array(first_time_lo, 'first_time_lo')}
    if len({len(a) for a in arrays.values()}) != 1:
        raise ValueError('extrema arrays must have the same number of cells')
    save_tables(path, [('F32Value', 'max_depth', arrays['max_depth']),
                       ('F32Value', 'max_speed', arrays['max_speed']),
                       ('IndexLike', 'first_time_hi', arrays['first_time_hi']),
                       ('IndexLike', 'first_time_lo', arrays['first_time_lo'])])


def read_extrema(path, expected_count=None):
    t = load_tables(path)
    result = Extrema(t['max_depth'], t['first_time_hi'], t['first_time_lo'], t['max_speed'])
    if len({len(result.max_depth), len(result.first_time_hi), len(result.first_time_lo),
            len(result.max_speed)}) != 1 or (expected_count is not None and len(result.max_depth) != expected_count):
        raise ValueError(f'{path}: extrema cell count differs from mesh')
    return result


@dataclass
class Checkpoint:
    depth: np.ndarray
    bed: np.ndarray
    q: np.ndarray
    rain_time: np.ndarray
    infiltration: np.ndarray
    max_depth: np.ndarray
    first_time_hi: np.ndarray
    first_time_lo: np.ndarray
    max_speed: np.ndarray
    clock_hi: int
    clock_lo: int
    gate_open: bool
    metadata_json: str

    @property
    def meta(self) -> dict:
        return json.loads(self.metadata_json)


def write_checkpoint(path, *, depth, bed, q, rain_time, infiltration, max_depth, first_time_hi,
                     first_time_lo, max_speed, clock_hi, clock_lo, gate_open, metadata_json):
    arrays = {name: _float_array(value, name) for name, value in
              [('depth', depth), ('bed', bed), ('q', q), ('rain_time', rain_time),
               ('infiltration', infiltration), ('max_depth', max_depth), ('max_speed', max_speed)]}
    if len({len(arrays[n]) for n in ('depth', 'bed', 'rain_time', 'infiltration', 'max_depth', 'max_speed')}) != 1:
        raise ValueError('checkpoint cell arrays must have the same length')
    times = {name: _uint_array(value, name) for name, value in
             [('first_time_hi', first_time_hi), ('first_time_lo', first_time_lo)]}
    if any(len(v) != len(arrays['depth']) for v in times.values()):
        raise ValueError('checkpoint first-time arrays must match cell count')
    if np.any(arrays['depth'] < 0):
        raise ValueError('checkpoint depth must be nonnegative')
    if not isinstance(metadata_json, str):
        raise TypeError('metadata_json must be a JSON string')
    json.loads(metadata_json)
    if not isinstance(gate_open, (bool, np.bool_)):
        raise TypeError('gate_open must be boolean')
    ints = [_uint_array([v], n) for n, v in [('clock_hi', clock_hi), ('clock_lo', clock_lo),
                                             ('gate_open', int(gate_open))]]
    encoded = np.frombuffer(metadata_json.encode('utf-8'), dtype=np.uint8)
    tables = [('F32Value', n, a) for n, a in arrays.items()]
    tables += [('IndexLike', n, a) for n, a in zip(_CHECKPOINT_INTS, ints)]
    tables += [('IndexLike', n, a) for n, a in times.items()]
    tables.append(('U8Value', 'state_meta', encoded))
    save_tables(path, tables)


def read_checkpoint(path, expected_cells=None, expected_sides=None) -> Checkpoint:
    t = load_tables(path)
    floats = {n: t[n] for n in _CHECKPOINT_FLOATS}
    ints = {n: t[n] for n in _CHECKPOINT_INTS}
    times = {n: t[n] for n in _CHECKPOINT_TIME}
    metadata_json = t['state_meta'].tobytes().decode('utf-8')
    if len({len(floats[n]) for n in ('depth', 'bed', 'rain_time', 'infiltration', 'max_depth', 'max_speed')}) != 1:
        raise ValueError(f'{path}: inconsistent checkpoint cell counts')
    if any(len(v) != len(floats['depth']) for v in times.values()):
        raise ValueError(f'{path}: inconsistent checkpoint first-time counts')
    if expected_cells is not None and len(floats['depth']) != expected_cells:
        raise ValueError(f'{path}: checkpoint cell count differs from mesh')
    if expected_sides is not None and len(floats['q']) != expected_sides:
        raise ValueError(f'{path}: checkpoint side count differs from mesh')
    if any(len(v) != 1 for v in ints.values()):
        raise ValueError(f'{path}: checkpoint scalar count must be one')
    if any(not np.all(np.isfinite(v)) for v in floats.values()) or np.any(floats['depth'] < 0):
        raise ValueError(f'{path}: invalid checkpoint F32 state')
    if ints['gate_open'][0] not in (0, 1):
        raise ValueError(f'{path}: invalid checkpoint gate state')
    json.loads(metadata_json)
    return Checkpoint(floats['depth'], floats['bed'], floats['q'], floats['rain_time'],
                      floats['infiltration'], floats['max_depth'], times['first_time_hi'],
                      times['first_time_lo'], floats['max_speed'], int(ints['clock_hi'][0]),
                      int(ints['clock_lo'][0]), bool(ints['gate_open'][0]), metadata_json)


def seed_checkpoint(depth, bed, q, *, meta: dict | None = None, rain_time=None, infiltration=None):
    """In-memory initial condition with ``run_state=False`` (as the reference's tests seed runs)."""
    depth = np.asarray(depth, dtype=np.float32)
    zeros = np.zeros_like(depth)
    words = np.zeros(len(depth), dtype=np.uint32)
    return Checkpoint(depth, np.asarray(bed, np.float32), np.asarray(q, np.float32),
                      zeros if rain_time is None else rain_time,
                      zeros if infiltration is None else infiltration, depth.copy(), words, words.copy(),
                      zeros.copy(), 0, 0, False, json.dumps(dict(run_state=False, **(meta or {}))))


def append_station_samples(path, *, elapsed_seconds, time, samples):
    """Append one row per station, stations in name order, to ``stations.csv`` (header on first
    write; core/output.py:271-284). `samples` maps station name -> cell_id, depth, level, u, v."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    rows = [dict(elapsed_seconds=elapsed_seconds, time=time, station=name,
                 **{key: samples[name][key] for key in STATION_FIELDS[3:]}) for name in sorted(samples)]
    fresh = not path.exists()
    with path.open('a', newline='') as stream:
        out = csv.DictWriter(stream, fieldnames=STATION_FIELDS)
        if fresh:
            out.writeheader()
        out.writerows(rows)


def resolve_station_cells(points: dict, mesh) -> dict:
    """Station -> smallest real cell whose closed rectangle holds the point (output.py:246-268)."""
    bounds = mesh.geometry()[0]
    left, bottom, right, top = (bounds[1:, i] for i in range(4))
    resolved = {}
    for name, (x, y) in points.items():
        if not np.isfinite(x) or not np.isfinite(y):
            raise ValueError(f'Station {name}: coordinates must be finite')
        hit = np.flatnonzero((left <= x) & (x <= right) & (bottom <= y) & (y <= top))
        if not len(hit):
            raise ValueError(f'Station {name}: ({x}, {y}) lies outside the mesh')
        resolved[name] = int(hit[0] + 1)
    return resolved


def write_manifest(path, manifest):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    temporary = path.with_name(f'.{path.name}.{os.getpid()}.tmp')
    temporary.write_text(json.dumps(manifest, ensure_ascii=False, indent=2) + '\n')
    os.replace(temporary, path)


def read_manifest(path) -> dict:
    return json.loads(Path(path).read_text())


class BedTimeline:
    """The bed a run had at each output time, for water levels from depth frames (the reference's
    BedTimeline, core/output.py:299-343): ``bed.fdb`` under the output root, then the manifest's
    ``bed_events`` ({elapsed_seconds, cell_ids, new_bed}) applied in time order up to the asked
    time, then every gate of its ``gate_control`` report set between its closed bed and the open
    elevation by the fraction of its opening time elapsed since its command (F32, as the device
    moves it). Times must not decrease from one call to the next."""

    def __init__(self, output_root, manifest: dict | None = None, expected_count=None):
        root = Path(output_root)
        if manifest is None:
            manifest = read_manifest(root.parent / 'manifest.json')
        self.bed = read_bed(root / 'bed.fdb', expected_count).copy()
        self.events = sorted(manifest.get('bed_events', []), key=lambda e: e['elapsed_seconds'])
        self.gates = manifest.get('gate_control') or {}
        self.applied = 0
        self.last = float('-inf')

    def _gate_fraction(self, gate: dict, seconds: float) -> np.float32:
        command = gate.get('command_seconds')
        if command is None or seconds < command:
            return np.float32(0.)
        hi, lo = gate['command_clock_ticks']
        tick = round((seconds + self.gates['clock_origin_before_start_s']) * 2 ** 40)
        age = max(tick - ((int(hi) << 32) | int(lo)), 0)
        age_s = np.float32(age >> 32) / np.float32(256.) + np.float32(age & 0xffffffff) * np.float32(2. ** -40)
        return np.minimum(age_s / np.float32(self.gates['opening_seconds']), np.float32(1.))

    def at(self, seconds: float) -> np.ndarray:
        if seconds < self.last:
            raise ValueError('BedTimeline needs frame times in nondecreasing order')
        self.last = seconds
        while self.applied < len(self.events) and self.events[self.applied]['elapsed_seconds'] <= seconds:
            event = self.events[self.applied]
            cells = _uint_array(event['cell_ids'], 'bed event cell_ids')
            values = _float_array(event['new_bed'], 'bed event new_bed')
            if len(cells) != len(values) or np.any(cells >= len(self.bed)):
                raise ValueError('bed event has invalid cell ids or values')
            self.bed[cells] = values
            self.applied += 1
        open_bed = np.float32(self.gates.get('open_elevation', 0.))
        for gate in self.gates.get('gates', []):
            closed = np.asarray(gate['closed_elevation'], dtype=np.float32)
            fraction = self._gate_fraction(gate, seconds)
            cells = np.asarray(gate['cell_ids'], dtype=np.int64)
            self.bed[cells] = open_bed if fraction >= 1. else closed + (open_bed - closed) * fraction
        return self.bed
