"""Small input builders for solver tests: NE/NS text meshes, rain GeoTIFFs, series and lines.

The meshes are the shapes the reference's own tests use (two 10 m cells, a 3 x 3 grid, a
coarse/fine channel, 2:1 strips); they are generated here from cell boxes, and sides are looked
up by position, so no test depends on a particular side numbering.
"""
from __future__ import annotations

import csv
import json
from datetime import datetime, timedelta, timezone
from pathlib import Path

import numpy as np

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.geotiff import write_geotiff
from nh_flood_2d_amd.mesh import Mesh, write_text_mesh

START = datetime(2023, 9, 7, 8, tzinfo=timezone.utc)
EPSG = 2326


def iso(seconds: float) -> str:
    return (START + timedelta(seconds=seconds)).isoformat()


def boxes_mesh(cells, bed=None, land=10) -> Mesh:
    """Mesh from square cell boxes (x0, x1, y0, y1) in metres; faces split at finer neighbours."""
    cells = [tuple(float(v) for v in c) for c in cells]
    breaks = [{}, {}]
    for x0, x1, y0, y1 in cells:
        for x in (x0, x1):
            breaks[0].setdefault(x, set()).update((y0, y1))
        for y in (y0, y1):
            breaks[1].setdefault(y, set()).update((x0, x1))
    owners, faces = {}, []
    for ci, (x0, x1, y0, y1) in enumerate(cells, 1):
        groups = []
        for orient, coord, lo, hi, role in ((2, x0, y0, y1, 'higher'), (2, x1, y0, y1, 'lower'),
                                            (1, y0, x0, x1, 'higher'), (1, y1, x0, x1, 'lower')):
            pts = sorted(p for p in breaks[2 - orient][coord] if lo <= p <= hi)
            group = []
            for a, b in zip(pts, pts[1:]):
                key = (orient, coord, a, b)
                owners.setdefault(key, {})[role] = ci
                group.append(key)
            groups.append(group)
        faces.append(groups)
    ids = {key: i for i, key in enumerate(sorted(owners), 1)}
    e_num, s_num = len(cells) + 1, len(ids) + 1
    counts = np.zeros((e_num, 4), np.int32)
    isl = []
    for ci, groups in enumerate(faces, 1):
        counts[ci] = [len(g) for g in groups]
        isl += [ids[k] for g in groups for k in g]
    from nh_flood_2d_amd.mesh import csr_pointers
    topo = np.zeros((s_num, 3), np.int32)
    length = np.zeros(s_num, np.float32)
    sx = np.zeros(s_num, np.float32)
    sy = np.zeros(s_num, np.float32)
    for (orient, coord, a, b), i in ids.items():
        own = owners[(orient, coord, a, b)]
        topo[i] = (orient, own.get('lower', 0), own.get('higher', 0))
        length[i] = b - a
        sx[i], sy[i] = (coord, (a + b) / 2) if orient == 2 else ((a + b) / 2, coord)
    x = np.zeros(e_num, np.float32)
    y = np.zeros(e_num, np.float32)
    for ci, (x0, x1, y0, y1) in enumerate(cells, 1):
        x[ci], y[ci] = (x0 + x1) / 2, (y0 + y1) / 2
    z = np.zeros(e_num, np.float32)
    if bed is not None:
        z[1:] = [bed(xx, yy, i) for i, (xx, yy) in enumerate(zip(x[1:], y[1:]), 1)] if callable(bed) else bed
    lt = np.zeros(e_num, np.uint8)
    lt[1:] = land
    return Mesh(x, y, z, lt, counts, csr_pointers(counts), np.asarray(isl, np.int32), topo, length, sx, sy,
                np.zeros(s_num, np.float32), np.zeros(s_num, np.uint8))


def side_at(mesh: Mesh, x: float, y: float) -> int:
    hit = np.flatnonzero((mesh.sx == np.float32(x)) & (mesh.sy == np.float32(y)))
    assert len(hit) == 1, (x, y, hit)
    return int(hit[0])


def cell_at(mesh: Mesh, x: float, y: float) -> int:
    hit = np.flatnonzero((mesh.x == np.float32(x)) & (mesh.y == np.float32(y)))
    assert len(hit) == 1, (x, y, hit)
    return int(hit[0])


def write_mesh(path: Path, mesh: Mesh) -> None:
    path.mkdir(parents=True, exist_ok=True)
    write_text_mesh(mesh, path / 'ne.txt', path / 'ns.txt')


def write_rain(path: Path, bands_mm, x0, y_top, pixel, intervals):
    """Rain raster of [bands, rows, cols] mm with its interval CSV (start, end seconds)."""
    bands = np.asarray(bands_mm, dtype=np.float32)
    if bands.ndim == 2:
        bands = bands[:, None, :]
    write_geotiff(path / 'rain.tif', bands, x0, y_top, pixel, pixel, EPSG)
    with open(path / 'intervals.csv', 'w', newline='') as f:
        w = csv.writer(f)
        w.writerow(['interval_start', 'interval_end'])
        for a, b in intervals:
            w.writerow([iso(a), iso(b)])
    return dict(raster=str(path / 'rain.tif'), intervals=str(path / 'intervals.csv'))


def write_boundary(path: Path, name: str, kind: str, coords, rows):
    (path / f'{name}.geojson').write_text(json.dumps({
        'type': 'LineString', 'coordinates': coords,
        'crs': {'type': 'name', 'properties': {'name': f'EPSG:{EPSG}'}}}))
    with open(path / f'{name}.csv', 'w', newline='') as f:
        w = csv.writer(f)
        w.writerow(['time', 'value'])
        for t, v in rows:
            w.writerow([iso(t), v])
    return dict(name=name, kind=kind, line=str(path / f'{name}.geojson'), series=str(path / f'{name}.csv'))


def two_cells(connected=False, bed=0., land=10) -> Mesh:
    """Two 10 x 10 m cells side by side; separate (closed) faces unless connected."""
    if connected:
        return boxes_mesh([(0, 10, 0, 10), (10, 20, 0, 10)], bed=[bed, bed], land=land)
    a = boxes_mesh([(0, 10, 0, 10)], bed=[bed], land=land)
    b = boxes_mesh([(10, 20, 0, 10)], bed=[bed], land=land)
    # disjoint union: renumber b's cells and sides after a's
    e_a, s_a = a.e_num - 1, a.s_num - 1
    topo_b = b.topo[1:].copy()
    topo_b[:, 1:] = np.where(topo_b[:, 1:] > 0, topo_b[:, 1:] + e_a, 0)
    counts = np.vstack([a.face_count, b.face_count[1:]])
    from nh_flood_2d_amd.mesh import csr_pointers
    return Mesh(np.r_[a.x, b.x[1:]], np.r_[a.y, b.y[1:]], np.r_[a.z, b.z[1:]], np.r_[a.land, b.land[1:]],
                counts, csr_pointers(counts), np.r_[a.isl, b.isl + s_a].astype(np.int32),
                np.vstack([a.topo, topo_b]), np.r_[a.length, b.length[1:]], np.r_[a.sx, b.sx[1:]],
                np.r_[a.sy, b.sy[1:]], np.zeros(a.s_num + b.s_num - 1, np.float32),
                np.zeros(a.s_num + b.s_num - 1, np.uint8))


def grid(nx: int, ny: int, dx: float, bed=0., land=10) -> Mesh:
    m = synth.uniform_grid(nx, ny, dx, np.full(nx * ny, bed, np.float32), np.full(nx * ny, land, np.uint8))
    return m


def run(path: Path, mesh: Mesh, *, duration, rain, boundaries=(), warmup=0, land=None, yield_step=None,
        sita=1., afa=.5, restart='', stations=None, warmup_only=False, preprocess=True, pipe=None):
    """Write the mesh and run the file-based solver; returns (row, manifest, domain config)."""
    from nh_flood_2d_amd.config import DomainConfig, ForceConfig
    from nh_flood_2d_amd import solver_compact as solver
    write_mesh(path, mesh)
    land = land or {10: dict(manning_n=.03, infiltration='none')}
    dc = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=EPSG,
                      domain_dir=str(path / 'out'), start=START, duration=duration, warmup_duration=warmup,
                      land_parameters=land, yield_step=yield_step or duration, sita=sita, afa=afa,
                      restart_state=str(restart) if restart else '', stations=stations or {})
    fc = ForceConfig(boundaries=list(boundaries), rainfall=rain)
    if preprocess:
        solver.preprocess(dc, fc)
    if pipe is not None:
        from nh_flood_2d_amd.coupling import prepare_pipe
        prepare_pipe(pipe, solver.load_preprocessed(dc))
    fn = solver.warmup_solver if warmup_only else solver.solver
    row = fn(dc, fc, pipe)
    manifest = json.loads((path / 'out' / 'manifest.json').read_text())
    return row, manifest, dc


def seed(path: Path, mesh: Mesh, depth, q=None, bed=None, rain_time=None, infiltration=None):
    """A run_state=False checkpoint carrying an initial state for the mesh written at path."""
    from nh_flood_2d_amd.output import write_checkpoint
    from nh_flood_2d_amd.solver_compact import sha256
    write_mesh(path, mesh)
    depth = np.asarray(depth, np.float32)
    zeros = np.zeros_like(depth)
    words = np.zeros(len(depth), np.uint32)
    write_checkpoint(path / 'seed.fdb', depth=depth, bed=mesh.z if bed is None else bed,
                     q=np.zeros(mesh.s_num, np.float32) if q is None else q,
                     rain_time=zeros if rain_time is None else rain_time,
                     infiltration=zeros if infiltration is None else infiltration,
                     max_depth=depth, first_time_hi=words, first_time_lo=words, max_speed=zeros,
                     clock_hi=0, clock_lo=0, gate_open=False,
                     metadata_json=json.dumps(dict(run_state=False, ne_sha256=sha256(path / 'ne.txt'),
                                                   ns_sha256=sha256(path / 'ns.txt'), gate_config='null')))
    return path / 'seed.fdb'
