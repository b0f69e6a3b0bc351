"""Seeded synthetic meshes and forcing for the benchmark workloads.

Meshes are built directly as :class:`~nh_flood_2d_amd.mesh.Mesh` arrays in the reference's
conventions (sentinel 0, compact l/r/b/t side lists ordered along each face, sides of a coarse
face split at the finer neighbours, as the reference's own test generators build them:
``tests/test_face_flux.py:113-173`` and ``tests/test_pressure_gradient.py:23-61``).

Two families are provided:

* ``uniform_grid``: an nx x ny grid of square cells (the 256 x 256, 10 m plain workload);
* ``quadtree_mesh``: a 2:1-balanced multi-resolution mesh over a grid of base blocks, each block
  refined uniformly to its own level (the 1.5 M-cell, 2/4/8 m workload and the batch meshes).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import Mesh, csr_pointers


# ----------------------------------------------------------------------------- mesh builders

def _assemble(cells_xy0: np.ndarray, cells_size: np.ndarray, sides: dict, unit: float,
              bed: np.ndarray, land: np.ndarray) -> Mesh:
    """Turn fine-unit cell boxes and side records into a compact Mesh.

    ``sides`` holds arrays ``orient, lower, higher, coord, start, length`` in fine units: a vertical
    side (orient 2) lies at x = coord and spans y in [start, start + length); a horizontal side
    (orient 1) lies at y = coord and spans x in [start, start + length).
    """
    n_cells = len(cells_size)
    orient, lower, higher = sides['orient'], sides['lower'], sides['higher']
    coord, start, seglen = sides['coord'], sides['start'], sides['length']
    order = np.lexsort((start, coord, orient))          # reference generators sort by key
    orient, lower, higher = orient[order], lower[order], higher[order]
    coord, start, seglen = coord[order], start[order], seglen[order]
    n_sides = len(orient)
    side_id = np.arange(1, n_sides + 1, dtype=np.int64)
    # Each side sits on the higher-direction face of `lower` and the lower face of `higher`.
    face_of_lower = np.where(orient == 2, 1, 3)
    face_of_higher = np.where(orient == 2, 0, 2)
    ref_cell = np.concatenate([lower, higher])
    ref_face = np.concatenate([face_of_lower, face_of_higher])
    ref_pos = np.concatenate([start, start])
    ref_side = np.concatenate([side_id, side_id])
    keep = ref_cell != 0
    ref_cell, ref_face, ref_pos, ref_side = ref_cell[keep], ref_face[keep], ref_pos[keep], ref_side[keep]
    o = np.lexsort((ref_pos, ref_face, ref_cell))
    ref_cell, ref_face, ref_side = ref_cell[o], ref_face[o], ref_side[o]
    counts = np.zeros((n_cells + 1, 4), dtype=np.int32)
    np.add.at(counts, (ref_cell, ref_face), 1)
    isl = ref_side.astype(np.int32)
    ptr = csr_pointers(counts)
    e_num, s_num = n_cells + 1, n_sides + 1
    x = np.zeros(e_num, dtype=np.float32)
    y = np.zeros(e_num, dtype=np.float32)
    x[1:] = ((cells_xy0[:, 0] + cells_size / 2) * unit).astype(np.float32)
    y[1:] = ((cells_xy0[:, 1] + cells_size / 2) * unit).astype(np.float32)
    topo = np.zeros((s_num, 3), dtype=np.int32)
    topo[1:, 0], topo[1:, 1], topo[1:, 2] = orient, lower, higher
    length = np.zeros(s_num, dtype=np.float32)
    length[1:] = seglen * unit
    sx = np.zeros(s_num, dtype=np.float32)
    sy = np.zeros(s_num, dtype=np.float32)
    mid = (start + seglen / 2) * unit
    sx[1:] = np.where(orient == 2, coord * unit, mid)
    sy[1:] = np.where(orient == 2, mid, coord * unit)
    z = np.zeros(e_num, dtype=np.float32)
    z[1:] = bed
    lt = np.zeros(e_num, dtype=np.uint8)
    lt[1:] = land
    return Mesh(x, y, z, lt, counts, ptr, isl, topo, length, sx, sy,
                np.zeros(s_num, dtype=np.float32), np.zeros(s_num, dtype=np.uint8))


def quadtree_cells(block_level: np.ndarray, max_level: int):
    """Cells of a block grid whose blocks are refined uniformly to ``block_level``.

    Returns fine-unit lower-left corners, sizes, and the owner grid (cell id per fine unit).
    ``block_level`` is indexed [by, bx]; neighbouring blocks may differ by at most one level.
    """
    nby, nbx = block_level.shape
    if np.any(np.abs(np.diff(block_level, axis=0)) > 1) or np.any(np.abs(np.diff(block_level, axis=1)) > 1):
        raise ValueError('block levels must be 2:1 balanced across block faces')
    fine = 1 << max_level
    xy0, sizes = [], []
    for level in range(max_level + 1):
        by, bx = np.nonzero(block_level == level)
        if not len(by):
            continue
        n = 1 << level
        s = fine >> level
        jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
        cx = (bx[:, None] * fine + ii.ravel()[None, :] * s).ravel()
        cy = (by[:, None] * fine + jj.ravel()[None, :] * s).ravel()
        xy0.append(np.column_stack([cx, cy]))
        sizes.append(np.full(len(cx), s, dtype=np.int64))
    xy0 = np.concatenate(xy0).astype(np.int64)
    sizes = np.concatenate(sizes)
    # A GIS-like natural order: by row of the lower-left corner, then x.
    order = np.lexsort((xy0[:, 0], xy0[:, 1]))
    xy0, sizes = xy0[order], sizes[order]
    owner = np.zeros((nby * fine, nbx * fine), dtype=np.int32)
    for level in np.unique(sizes):
        sel = np.flatnonzero(sizes == level)
        s = int(level)
        ids = (sel + 1).astype(np.int32)
        for dy in range(s):
            for dx in range(s):
                owner[xy0[sel, 1] + dy, xy0[sel, 0] + dx] = ids
    return xy0, sizes, owner


def quadtree_sides(xy0: np.ndarray, sizes: np.ndarray, owner: np.ndarray) -> dict:
    """Side records of a 2:1 balanced quadtree (each side created once by its lower cell)."""
    height, width = owner.shape
    n = len(sizes)
    cid = np.arange(1, n + 1, dtype=np.int64)
    out = {k: [] for k in ('orient', 'lower', 'higher', 'coord', 'start', 'length')}

    def add(orient, lower, higher, coord, start, length):
        out['orient'].append(np.full(len(lower), orient, dtype=np.int64))
        out['lower'].append(lower.astype(np.int64))
        out['higher'].append(higher.astype(np.int64))
        out['coord'].append(coord.astype(np.int64))
        out['start'].append(start.astype(np.int64))
        out['length'].append(length.astype(np.int64))

    x0, y0, s = xy0[:, 0], xy0[:, 1], sizes
    for axis in (0, 1):   # 0: vertical sides on right faces, 1: horizontal sides on top faces
        orient = 2 if axis == 0 else 1
        along0 = y0 if axis == 0 else x0
        across1 = (x0 + s) if axis == 0 else (y0 + s)
        limit = width if axis == 0 else height
        # exterior faces on the low edge
        low = (x0 if axis == 0 else y0) == 0
        add(orient, np.zeros(low.sum()), cid[low], np.zeros(low.sum()), along0[low], s[low])
        # exterior faces on the high edge
        high = across1 == limit
        add(orient, cid[high], np.zeros(high.sum()), across1[high], along0[high], s[high])
        inner = ~high
        c, a0, a1, sz = cid[inner], along0[inner], across1[inner], s[inner]

        def owner_at(pos_along):
            return owner[pos_along, a1] if axis == 0 else owner[a1, pos_along]

        along_limit = height if axis == 0 else width
        first = owner_at(a0)
        half = np.maximum(sz // 2, 1)
        second = owner_at(np.minimum(a0 + half, along_limit - 1))
        split = (sz > 1) & (first != second)
        whole = ~split
        add(orient, c[whole], first[whole], a1[whole], a0[whole], sz[whole])
        add(orient, c[split], first[split], a1[split], a0[split], half[split])
        add(orient, c[split], second[split], a1[split], a0[split] + half[split], half[split])
    return {k: np.concatenate(v) for k, v in out.items()}


def uniform_grid(nx: int, ny: int, dx: float, bed: np.ndarray | None = None,
                 land: np.ndarray | None = None) -> Mesh:
    """nx x ny square cells of size dx, numbered row by row from the lower-left corner."""
    levels = np.zeros((ny, nx), dtype=np.int64)
    xy0, sizes, owner = quadtree_cells(levels, 0)
    sides = quadtree_sides(xy0, sizes, owner)
    n = nx * ny
    bed = np.zeros(n, dtype=np.float32) if bed is None else np.asarray(bed, dtype=np.float32).ravel()
    land = np.zeros(n, dtype=np.uint8) if land is None else np.asarray(land, dtype=np.uint8).ravel()
    return _assemble(xy0, sizes, sides, float(dx), bed, land)


def quadtree_mesh(block_level: np.ndarray, block_size: float, bed_fn, land_fn) -> Mesh:
    """2:1 balanced multi-resolution mesh; ``bed_fn``/``land_fn`` map centre x, y to values."""
    max_level = int(block_level.max())
    xy0, sizes, owner = quadtree_cells(block_level, max_level)
    sides = quadtree_sides(xy0, sizes, owner)
    unit = block_size / (1 << max_level)
    cx = (xy0[:, 0] + sizes / 2) * unit
    cy = (xy0[:, 1] + sizes / 2) * unit
    return _assemble(xy0, sizes, sides, unit, bed_fn(cx, cy), land_fn(cx, cy))


def balance_levels(levels: np.ndarray) -> np.ndarray:
    """Raise block levels until face neighbours differ by at most one level."""
    levels = levels.copy()
    while True:
        need = levels.copy()
        need[1:, :] = np.maximum(need[1:, :], levels[:-1, :] - 1)
        need[:-1, :] = np.maximum(need[:-1, :], levels[1:, :] - 1)
        need[:, 1:] = np.maximum(need[:, 1:], levels[:, :-1] - 1)
        need[:, :-1] = np.maximum(need[:, :-1], levels[:, 1:] - 1)
        if np.array_equal(need, levels):
            return levels
        levels = need


# ----------------------------------------------------------------------------- terrain

def diamond_square(power: int, roughness: float, rng: np.random.Generator) -> np.ndarray:
    """(2**power + 1)^2 fractal surface in [0, 1]; displacement scale shrinks by ``roughness``."""
    n = (1 << power) + 1
    h = np.zeros((n, n))
    h[::n - 1, ::n - 1] = rng.uniform(-1, 1, size=(2, 2))
    step, scale = n - 1, 1.0
    while step > 1:
        half = step // 2
        # diamond: square centres from four corners
        c = (h[0:-1:step, 0:-1:step] + h[0:-1:step, step::step] +
             h[step::step, 0:-1:step] + h[step::step, step::step]) / 4
        h[half::step, half::step] = c + rng.uniform(-scale, scale, size=c.shape)
        # square: edge midpoints from their (up to four) diamond neighbours
        for r0, c0 in ((0, half), (half, 0)):
            rows = np.arange(r0, n, step)
            cols = np.arange(c0, n, step)
            rr, cc = np.meshgrid(rows, cols, indexing='ij')
            total = np.zeros(rr.shape)
            count = np.zeros(rr.shape)
            for dr, dc in ((-half, 0), (half, 0), (0, -half), (0, half)):
                r2, c2 = rr + dr, cc + dc
                ok = (r2 >= 0) & (r2 < n) & (c2 >= 0) & (c2 < n)
                total[ok] += h[r2[ok], c2[ok]]
                count[ok] += 1
            h[rr, cc] = total / count + rng.uniform(-scale, scale, size=rr.shape)
        step = half
        scale *= roughness
    h -= h.min()
    return h / max(h.max(), 1e-12)


def sample_grid(grid: np.ndarray, extent: float, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Bilinear sample of a square grid spanning [0, extent]^2 (grid[row=y, col=x])."""
    n = grid.shape[0] - 1
    fx = np.clip(x / extent * n, 0, n - 1e-9)
    fy = np.clip(y / extent * n, 0, n - 1e-9)
    i0, j0 = np.floor(fx).astype(int), np.floor(fy).astype(int)
    tx, ty = fx - i0, fy - j0
    return ((1 - tx) * (1 - ty) * grid[j0, i0] + tx * (1 - ty) * grid[j0, i0 + 1] +
            (1 - tx) * ty * grid[j0 + 1, i0] + tx * ty * grid[j0 + 1, i0 + 1])


def meander(rng: np.random.Generator, x0: float, x1: float, y_mid: float, amplitude: float,
            wavelength: float, samples: int = 4000) -> np.ndarray:
    """A seeded meandering polyline from x0 to x1 around y_mid."""
    xs = np.linspace(x0, x1, samples)
    phase = rng.uniform(0, 2 * np.pi)
    wobble = rng.uniform(0.15, 0.3)
    ys = y_mid + amplitude * np.sin(2 * np.pi * xs / wavelength + phase) \
        + amplitude * wobble * np.sin(2 * np.pi * xs / (wavelength / 2.7) + 2 * phase)
    return np.column_stack([xs, ys])


def distance_to_polylines(points_x, points_y, lines: list[np.ndarray]) -> np.ndarray:
    """Approximate distance from points to densely sampled polylines (via a KD-tree)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.concatenate(lines))
    d, _ = tree.query(np.column_stack([points_x, points_y]))
    return d


# ----------------------------------------------------------------------------- forcing

@dataclass
class SeriesSpec:
    times: np.ndarray     # seconds from start (F64)
    values: np.ndarray


@dataclass
class BoundarySpec:
    name: str
    kind: str             # 'level' or 'discharge'
    side_ids: np.ndarray  # exterior side ids (original numbering)
    series: SeriesSpec


@dataclass
class Scenario:
    """A complete synthetic run: mesh, land table, rain bands, boundaries and physics."""
    name: str
    mesh: Mesh
    land_parameters: dict
    rain_starts: np.ndarray
    rain_ends: np.ndarray
    rain_mm: np.ndarray           # [bands, pixels]
    rain_pixel: np.ndarray        # [e_num] cell -> pixel
    boundaries: list = field(default_factory=list)
    duration: float = 3600.
    afa: float = 0.5
    sita: float = 1.0
    min_h: float = 0.02
    initial_depth: np.ndarray | None = None


def triangular_hyetograph(peak_mm_h: float, peak_s: float, total_s: float, band_s: float):
    """Band starts/ends and mm per band of a triangular rain-rate hyetograph."""
    edges = np.append(np.arange(0., total_s, band_s), total_s)  # the last band may be shorter
    starts, ends = edges[:-1], edges[1:]
    # exact band average of the piecewise-linear rate (mm/h), times band hours
    def cumulative(t):
        t = np.clip(t, 0, total_s)
        rise = np.minimum(t, peak_s)
        fall = np.clip(t - peak_s, 0, total_s - peak_s)
        up = peak_mm_h * rise ** 2 / (2 * peak_s)
        down = peak_mm_h * (fall - fall ** 2 / (2 * (total_s - peak_s)))
        return (up + down) / 3600.
    mm = cumulative(ends) - cumulative(starts)
    return starts, ends, mm


def plain_uniform(seed: int = 0, n: int = 256, dx: float = 10.) -> Scenario:
    """256 x 256 cells of 10 m: sloping plane + 3 Gaussian hills + 1.5 m meandering channel.

    One land class (n = 0.03, no infiltration), 50 mm/h for 1800 s then dry, closed boundaries,
    afa 0.5, theta 1, min_h 0.02, 3600 s from a dry start."""
    rng = np.random.default_rng(seed)
    extent = n * dx
    cx = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(cx, cx)
    bed = 0.001 * X
    for _ in range(3):
        hx, hy = rng.uniform(0.15, 0.85, size=2) * extent
        height, sigma = rng.uniform(2, 8), rng.uniform(150, 400)
        bed = bed + height * np.exp(-((X - hx) ** 2 + (Y - hy) ** 2) / (2 * sigma ** 2))
    line = meander(rng, 0, extent, extent / 2, extent * 0.12, extent / 1.7)
    d = distance_to_polylines(X.ravel(), Y.ravel(), [line]).reshape(X.shape)
    bed = bed - 1.5 * np.clip(1 - d / (1.5 * dx), 0, 1)
    mesh = uniform_grid(n, n, dx, bed.astype(np.float32), np.ones(n * n, dtype=np.uint8))
    land = {1: dict(manning_n=0.03, infiltration='none')}
    starts = np.array([0., 1800.])
    ends = np.array([1800., 3600.])
    mm = np.array([[25.], [0.]])
    return Scenario('plain_256x256_10m', mesh, land, starts, ends, mm,
                    np.zeros(mesh.e_num, dtype=np.int32), [], 3600., 0.5, 1.0, 0.02)


def multires_channel(seed: int = 0, extent_blocks: int = 600, block_size: float = 8.,
                     band_m: float = 300., channels: int = 3, duration: float = 10800.,
                     amplitude: float = 40., roughness: float = 0.6) -> Scenario:
    """2/4/8 m quadtree around a seeded channel network (the ~1.5 M-cell workload).

    2 m cells lie within ``band_m`` of the channels, 8 m cells elsewhere, 4 m cells keep the 2:1
    balance. Diamond-square DEM (amplitude 40 m, roughness 0.6) on a seaward slope, channels
    carved to the sea at x = 0, four land classes (n 0.015/0.03/0.05/0.1; Horton for 1-3,
    class 0 impervious), triangular hyetograph peaking at 80 mm/h at 1 h over 3 h, and a level
    boundary on the x = 0 edge with tide 1.0 m * sin(2 pi t / 44640 s)."""
    rng = np.random.default_rng(seed)
    extent = extent_blocks * block_size
    lines = [meander(rng, 0., extent, extent * rng.uniform(0.35, 0.65),
                     extent * 0.08, extent / rng.uniform(1.5, 2.5))]
    for _ in range(channels - 1):
        xa = rng.uniform(0.2, 0.7) * extent
        ya = np.interp(xa, lines[0][:, 0], lines[0][:, 1])
        ytop = extent * (0.9 if rng.uniform() < 0.5 else 0.1)
        t = np.linspace(0, 1, 3000)[:, None]
        xs = xa + (rng.uniform(0.1, 0.25) * extent) * t[:, 0] + 30 * np.sin(7 * np.pi * t[:, 0])
        ys = ya + (ytop - ya) * t[:, 0]
        lines.append(np.column_stack([xs, ys]))
    bc = (np.arange(extent_blocks) + 0.5) * block_size
    BX, BY = np.meshgrid(bc, bc)
    dist = distance_to_polylines(BX.ravel(), BY.ravel(), lines).reshape(BX.shape)
    half_diag = block_size / np.sqrt(2)
    levels = np.where(dist <= band_m + half_diag, 2, 0).astype(np.int64)
    levels = balance_levels(levels)
    fractal = diamond_square(10, roughness, rng)

    def bed_fn(x, y):
        base = 0.5 + 0.003 * x + amplitude * sample_grid(fractal, extent, x, y)
        d = distance_to_polylines(x, y, lines)
        channel = -1.5 + 0.0015 * x
        weight = np.clip(1 - d / 12., 0, 1)
        return (base * (1 - weight) + np.minimum(channel, base) * weight).astype(np.float32)

    classes = diamond_square(7, 0.7, rng)

    def land_fn(x, y):
        v = sample_grid(classes, extent, x, y)
        return np.searchsorted(np.quantile(v, [0.15, 0.5, 0.8]), v).astype(np.uint8)

    mesh = quadtree_mesh(levels, block_size, bed_fn, land_fn)
    land = {0: dict(manning_n=0.015, infiltration='none'),
            1: dict(manning_n=0.03, infiltration=dict(f0_mm_h=60, fc_mm_h=10, k_per_h=2)),
            2: dict(manning_n=0.05, infiltration=dict(f0_mm_h=90, fc_mm_h=15, k_per_h=3)),
            3: dict(manning_n=0.1, infiltration=dict(f0_mm_h=120, fc_mm_h=20, k_per_h=4))}
    starts, ends, mm = triangular_hyetograph(80., 3600., duration, 300.)
    side_ids, cells, faces, _, _ = mesh.exterior_edges()
    sea = side_ids[faces == 0]
    knots = np.arange(0., duration + 300., 300.)
    tide = SeriesSpec(knots, 1.0 * np.sin(2 * np.pi * knots / 44640.))
    boundaries = [BoundarySpec('sea', 'level', np.sort(sea), tide)]
    return Scenario(f'multires_{mesh.e_num - 1}_cells', mesh, land, starts, ends, mm[:, None],
                    np.zeros(mesh.e_num, dtype=np.int32), boundaries, duration, 0.5, 1.0, 0.02)


def scenario_forcing(sc: Scenario):
    """Forcing object (boundaries with cell shares, rain field, events) of a synthetic scenario."""
    from .forcing import Forcing, RainField, TimeSeries, make_boundary
    claimed = np.zeros(sc.mesh.e_num, dtype=bool)
    boundaries = [make_boundary(sc.mesh, b.name, b.kind, np.asarray(b.side_ids),
                                TimeSeries(np.asarray(b.series.times, float), np.asarray(b.series.values, float)),
                                claimed) for b in sc.boundaries]
    rain = RainField(np.asarray(sc.rain_starts, float), np.asarray(sc.rain_ends, float),
                     np.asarray(sc.rain_mm, float), np.asarray(sc.rain_pixel, np.int32))
    return Forcing.build(boundaries, rain, sc.duration)


def small_multires(seed: int = 1, blocks: int = 24, block_size: float = 8., band_m: float = 30.,
                   duration: float = 1800., theta: float = 1.0, pond: float = 0.0,
                   discharge: float = 0.0) -> Scenario:
    """A small 2/4/8 m quadtree with hanging faces, Horton classes, a tide and a river inflow."""
    rng = np.random.default_rng(seed)
    extent = blocks * block_size
    line = meander(rng, 0., extent, extent * 0.5, extent * 0.1, extent / 1.3, samples=600)
    bc = (np.arange(blocks) + 0.5) * block_size
    BX, BY = np.meshgrid(bc, bc)
    dist = distance_to_polylines(BX.ravel(), BY.ravel(), [line]).reshape(BX.shape)
    levels = balance_levels(np.where(dist <= band_m, 2, 0).astype(np.int64))
    fractal = diamond_square(6, 0.6, rng)

    def bed_fn(x, y):
        base = 0.2 + 0.004 * x + 3.0 * sample_grid(fractal, extent, x, y)
        d = distance_to_polylines(x, y, [line])
        w = np.clip(1 - d / 6., 0, 1)
        return (base * (1 - w) + np.minimum(-0.8 + 0.002 * x, base) * w).astype(np.float32)

    classes = diamond_square(5, 0.7, rng)

    def land_fn(x, y):
        v = sample_grid(classes, extent, x, y)
        return np.searchsorted(np.quantile(v, [0.25, 0.5, 0.75]), v).astype(np.uint8)

    mesh = quadtree_mesh(levels, block_size, bed_fn, land_fn)
    land = {0: dict(manning_n=0.015, infiltration='none'),
            1: dict(manning_n=0.03, infiltration=dict(f0_mm_h=60, fc_mm_h=10, k_per_h=2)),
            2: dict(manning_n=0.05, infiltration=dict(f0_mm_h=90, fc_mm_h=15, k_per_h=3)),
            3: dict(manning_n=0.1, infiltration=dict(f0_mm_h=120, fc_mm_h=20, k_per_h=4))}
    starts, ends, mm = triangular_hyetograph(80., duration / 3, duration, 60.)
    side_ids, cells, faces, _, _ = mesh.exterior_edges()
    knots = np.arange(0., duration + 60., 60.)
    boundaries = [BoundarySpec('sea', 'level', np.sort(side_ids[faces == 0]),
                               SeriesSpec(knots, 0.3 * np.sin(2 * np.pi * knots / 600.)))]
    if discharge:
        right = side_ids[faces == 1]
        ys = mesh.sy[right]
        pick = right[np.abs(ys - np.interp(extent, line[:, 0], line[:, 1])) < 12]
        if len(pick) == 0:
            pick = right[:2]
        boundaries.append(BoundarySpec('river', 'discharge', np.sort(pick),
                                       SeriesSpec(np.array([0., duration]), np.array([discharge, discharge * 0.5]))))
    depth = None
    if pond:
        depth = np.zeros(mesh.e_num, dtype=np.float32)
        _, widths, _, _ = mesh.geometry()
        depth[1:] = np.where(mesh.z[1:] < 1.0, pond, 0.).astype(np.float32)
    return Scenario(f'small_multires_{mesh.e_num - 1}', mesh, land, starts, ends, mm[:, None],
                    np.zeros(mesh.e_num, dtype=np.int32), boundaries, duration, 0.5, theta, 0.02, depth)


def two_level_fractal(seed: int = 0, blocks: int = 790, block_size: float = 10., refine: float = 0.2,
                      duration: float = 7200., tide: float = 0.) -> Scenario:
    """One shared mesh for scenario batches: a 2-level (block_size / block_size/2) quadtree over a
    seeded diamond-square DEM (amplitude 20 m), with the lowest `refine` fraction of the blocks
    refined; about blocks^2 x (1 + 3 refine) cells (one million at the defaults). One land class
    with Manning n 0.03 and Horton (60, 10, 2) mm/h, 1/h; closed boundaries, or with `tide` > 0 a
    level boundary on the left edge at tide x sin(2 pi t / 44640 s)."""
    rng = np.random.default_rng(seed)
    extent = blocks * block_size
    power = int(np.ceil(np.log2(blocks))) + 1
    fractal = diamond_square(min(power, 11), 0.6, rng)
    bc = (np.arange(blocks) + 0.5) * block_size
    BX, BY = np.meshgrid(bc, bc)
    zb = sample_grid(fractal, extent, BX.ravel(), BY.ravel()).reshape(BX.shape)
    levels = np.where(zb <= np.quantile(zb, refine), 1, 0).astype(np.int64)
    levels = balance_levels(levels)

    def bed_fn(x, y):
        return (20.0 * sample_grid(fractal, extent, x, y) + 0.0005 * x).astype(np.float32)

    def land_fn(x, y):
        return np.ones(len(x), dtype=np.uint8)

    mesh = quadtree_mesh(levels, block_size, bed_fn, land_fn)
    land = {1: dict(manning_n=0.03, infiltration=dict(f0_mm_h=60, fc_mm_h=10, k_per_h=2))}
    starts, ends, mm = triangular_hyetograph(80., duration / 3, duration, 300.)
    boundaries = []
    if tide > 0:
        side_ids, _, faces, _, _ = mesh.exterior_edges()
        knots = np.arange(0., duration + 300., 300.)
        boundaries.append(BoundarySpec('tide', 'level', np.sort(side_ids[faces == 0]),
                                       SeriesSpec(knots, tide * np.sin(2 * np.pi * knots / 44640.))))
    return Scenario(f'two_level_{mesh.e_num - 1}_cells', mesh, land, starts, ends, mm[:, None],
                    np.zeros(mesh.e_num, dtype=np.int32), boundaries, duration, 0.5, 1.0, 0.02)


def uneven_scenario(seed: int, cells: float, levels: int, classes: int = 4, finest: float = 5.,
                    duration: float = 7200.) -> Scenario:
    """One mesh of an uneven-size batch: about `cells` cells over `levels` (1-3) refinement levels
    whose finest cells are `finest` m (so meshes of different levels share one size base), the
    lowest 25 % of a seeded diamond-square DEM (amplitude 20 m) refined one level and, with three
    levels, its lowest 10 % two; `classes` (2-4) land classes of the Horton table (class 0
    impervious); a triangular hyetograph with peak U[20, 150] mm/h; a tide boundary on the left
    edge for half of the seeds; a dry start or a 0.3 m pond on the lowest quarter."""
    if not 1 <= levels <= 3 or not 2 <= classes <= 4:
        raise ValueError('1-3 levels and 2-4 land classes')
    rng = np.random.default_rng(seed)
    per_block = {1: 1.0, 2: 0.75 + 0.25 * 4, 3: 0.65 + 0.25 * 4 + 0.10 * 16}[levels]
    blocks = max(8, int(round(np.sqrt(cells / per_block))))
    block_size = finest * (1 << (levels - 1))
    extent = blocks * block_size
    power = int(np.ceil(np.log2(blocks))) + 1
    fractal = diamond_square(min(power, 11), 0.6, rng)
    bc = (np.arange(blocks) + 0.5) * block_size
    BX, BY = np.meshgrid(bc, bc)
    zb = sample_grid(fractal, extent, BX.ravel(), BY.ravel()).reshape(BX.shape)
    lv = np.zeros_like(zb, dtype=np.int64)
    if levels >= 2:
        lv[zb <= np.quantile(zb, 0.25)] = 1
    if levels == 3:
        lv[zb <= np.quantile(zb, 0.10)] = 2
    lv = balance_levels(lv)
    class_field = diamond_square(min(power, 9), 0.7, rng)
    cuts = np.linspace(0, 1, classes + 1)[1:-1]

    def bed_fn(x, y):
        return (20.0 * sample_grid(fractal, extent, x, y) + 0.0005 * x).astype(np.float32)

    def land_fn(x, y):
        v = sample_grid(class_field, extent, x, y)
        return np.searchsorted(np.quantile(v, cuts), v).astype(np.uint8)

    mesh = quadtree_mesh(lv, block_size, bed_fn, land_fn)
    table = {0: dict(manning_n=0.015, infiltration='none'),
             1: dict(manning_n=0.03, infiltration=dict(f0_mm_h=60, fc_mm_h=10, k_per_h=2)),
             2: dict(manning_n=0.05, infiltration=dict(f0_mm_h=90, fc_mm_h=15, k_per_h=3)),
             3: dict(manning_n=0.1, infiltration=dict(f0_mm_h=120, fc_mm_h=20, k_per_h=4))}
    land = {k: table[k] for k in range(classes)}
    starts, ends, mm = triangular_hyetograph(float(rng.uniform(20., 150.)), duration / 3, duration, 300.)
    boundaries = []
    if rng.uniform() < 0.5:
        side_ids, _, faces, _, _ = mesh.exterior_edges()
        knots = np.arange(0., duration + 300., 300.)
        boundaries.append(BoundarySpec('tide', 'level', np.sort(side_ids[faces == 0]),
                                       SeriesSpec(knots, 1.0 * np.sin(2 * np.pi * knots / 44640.))))
    depth = None
    if rng.uniform() < 0.5:
        low = mesh.z <= np.quantile(mesh.z[1:], 0.25)
        depth = np.where(low, np.float32(0.3), np.float32(0.)).astype(np.float32)
        depth[0] = 0.
    return Scenario(f'uneven_{levels}l_{mesh.e_num - 1}_cells', mesh, land, starts, ends, mm[:, None],
                    np.zeros(mesh.e_num, dtype=np.int32), boundaries, duration, 0.5, 1.0, 0.02, depth)


def _uneven_item(args):
    seed, cells, levels, classes = args
    from .forcing import cell_forcing_arrays
    from .layout import build_layout
    sc = uneven_scenario(seed, cells, levels, classes)
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, scenario_forcing(sc)), sc.sita)
    return sc, layout


def uneven_batch(seed: int, count: int = 32, min_cells: float = 50_000, max_cells: float = 4_000_000,
                 workers: int = 16) -> list:
    """(scenario, layout) of `count` meshes with cell counts log-uniform in [min_cells, max_cells],
    1-3 levels and 2-4 land classes drawn per mesh, built in `workers` processes (spawned, so a
    parent that already holds a GPU context is not forked)."""
    rng = np.random.default_rng(seed)
    jobs = [(seed * 1000 + i, float(np.exp(rng.uniform(np.log(min_cells), np.log(max_cells)))),
             int(rng.integers(1, 4)), int(rng.integers(2, 5))) for i in range(count)]
    order = sorted(range(count), key=lambda i: -jobs[i][1])     # largest first
    if workers <= 1:
        built = {i: _uneven_item(jobs[i]) for i in order}
    else:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context('spawn')) as pool:
            built = dict(zip(order, pool.map(_uneven_item, [jobs[i] for i in order])))
    return [built[i] for i in range(count)]


def scenario_variants(base: Scenario, count: int, seed: int = 0, peak_range=(20., 150.),
                      manning_sigma: float = 0.2, pond: float = 0.3) -> list[Scenario]:
    """`count` scenarios on the mesh of `base`: rain peak uniform in `peak_range` (mm/h, same
    triangular shape), every land class's Manning n times one lognormal(0, manning_sigma) factor,
    and a dry start or a `pond` m pond on the lowest quarter of the cells, each drawn per scenario."""
    rng = np.random.default_rng(seed)
    out = []
    low = base.mesh.z <= np.quantile(base.mesh.z[1:], 0.25)
    for i in range(count):
        peak = float(rng.uniform(*peak_range))
        factor = float(np.exp(rng.normal(0., manning_sigma)))
        wet = bool(rng.integers(0, 2))
        starts, ends, mm = triangular_hyetograph(peak, base.duration / 3, base.duration,
                                                 float(base.rain_ends[0] - base.rain_starts[0]))
        land = {k: dict(v, manning_n=float(np.float32(v['manning_n'] * factor)))
                for k, v in base.land_parameters.items()}
        depth = None
        if wet:
            depth = np.where(low, np.float32(pond), np.float32(0.)).astype(np.float32)
            depth[0] = 0.
        out.append(Scenario(f'{base.name}_v{i}', base.mesh, land, starts, ends, mm[:, None], base.rain_pixel,
                            list(base.boundaries), base.duration, base.afa, base.sita, base.min_h, depth))
    return out


def serving_load(seed: int, count: int, rate: float = 2.0, min_cells: float = 100_000,
                 max_cells: float = 8_000_000, min_hours: float = 1.0, max_hours: float = 6.0) -> list:
    """The serving load: `count` solves arriving as a Poisson process of `rate` per second, each a
    seeded two-level mesh with a cell count log-uniform in [min_cells, max_cells], a simulated
    duration uniform in [min_hours, max_hours] (whole 300 s), a triangular hyetograph with peak
    U[20, 150] mm/h and, for about half of them, a sinusoidal tide boundary. Returns
    (index, scenario, arrival seconds) in arrival order."""
    rng = np.random.default_rng(seed + 11)
    out, arrival = [], 0.
    for i in range(count):
        cells = float(np.exp(rng.uniform(np.log(min_cells), np.log(max_cells))))
        blocks = max(8, int(round(np.sqrt(cells / 1.6))))
        duration = float(np.round(rng.uniform(min_hours, max_hours) * 3600. / 300.) * 300.)
        base = two_level_fractal(seed + 100 + i, blocks=blocks, duration=duration,
                                 tide=1.0 if rng.uniform() < 0.5 else 0.)
        arrival += float(rng.exponential(1. / rate)) if i else 0.
        out.append((i, scenario_variants(base, 1, seed=seed + 200 + i)[0], arrival))
    return out


def simulation(sc: Scenario, engine_factory=None, end_time: float | None = None, yield_step: float = 300.,
               batch_size: int = 32, base_layout=None, outputs=None, coupling=None):
    """A Simulation of a synthetic scenario (no files); `base_layout` reuses another scenario's
    layout of the same mesh."""
    from .simulation import RunSettings, Simulation
    settings = RunSettings(afa=sc.afa, sita=sc.sita, min_h=sc.min_h,
                           end_time=sc.duration if end_time is None else end_time,
                           yield_step=yield_step, batch_size=batch_size)
    return Simulation(sc.mesh, sc.land_parameters, scenario_forcing(sc), settings, engine_factory=engine_factory,
                      initial_depth=sc.initial_depth, base_layout=base_layout, outputs=outputs,
                      coupling=coupling)
