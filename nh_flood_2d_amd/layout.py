"""Device layout: Morton-ordered structure-of-arrays for the HIP kernels.

The host keeps every array in the reference's (original) numbering; a :class:`Layout` holds the
permutation to the device order and the packed arrays the kernels read:

* cells sorted by the Morton (Z-order) code of their lower-left corner on the finest grid, so the
  cells of one quadtree block are contiguous and a wavefront's neighbours are a few cache lines
  away across all refinement levels;
* sides sorted by their owner cell in that order (the lower cell, or the interior cell of an
  exterior side), so the face pass reads cell data almost sequentially and the depth pass writes
  each side from one cell;
* a compact side list per cell: ``csr_start`` plus four 8-bit face counts packed in one word
  (8 B of header per cell instead of the reference's 32 B of pointers and counts);
* ``side_cells`` as interleaved (lower, higher) pairs and a 1-byte orient/exterior word.

Within each face the reference's side order is kept, so every per-cell sum adds the same values
in the same order as ``solver_compact.py`` and results do not depend on the renumbering.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

from .hydraulics import (class_tables, hanging_geometry, land_parameters_of, open_face_fractions,
                         side_centre_distance)
from .mesh import Mesh


def morton_codes(ix: np.ndarray, iy: np.ndarray) -> np.ndarray:
    """Interleave the bits of two non-negative integer arrays (x in even bits)."""
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0xFFFFFFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x0000FFFF0000FFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x3333333333333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x5555555555555555)
        return v
    return spread(ix) | (spread(iy) << np.uint64(1))


def cell_order(mesh: Mesh, mode: str = 'morton') -> np.ndarray:
    """Permutation new -> original cell id (index 0 stays the sentinel)."""
    n = mesh.e_num
    if mode == 'identity' or n <= 2:
        return np.arange(n, dtype=np.int64)
    bounds, widths, _, _ = mesh.geometry()
    fine = widths[1:].min()
    x0 = bounds[1:, 0] - bounds[1:, 0].min()
    y0 = bounds[1:, 1] - bounds[1:, 1].min()
    ix = np.floor(x0 / fine + 0.5).astype(np.int64)
    iy = np.floor(y0 / fine + 0.5).astype(np.int64)
    code = morton_codes(ix, iy)
    order = np.argsort(code, kind='stable') + 1
    return np.concatenate([[0], order]).astype(np.int64)


@dataclass
class Forcing:
    """Per-cell forcing geometry in original numbering (loaded or synthetic)."""
    rain_pixel: np.ndarray               # [e_num] cell -> pixel
    n_pixels: int
    boundary_id: np.ndarray              # [e_num] -1 or boundary index
    bweight: np.ndarray                  # [e_num] F32 discharge share
    kinds: np.ndarray                    # [b] 0 level, 1 discharge
    selected_sides: np.ndarray           # [s_num] bool: side belongs to a prescribed boundary


@dataclass
class Layout:
    e_num: int
    s_num: int
    cell_perm: np.ndarray      # new -> original
    cell_inv: np.ndarray       # original -> new
    side_perm: np.ndarray
    side_inv: np.ndarray
    arrays: dict = field(default_factory=dict)   # static device arrays (numpy, device order)
    n_pixels: int = 1
    b_num: int = 0
    n_grad: int = 0
    n_hang: int = 0
    theta: float = 1.0
    cell_f0: np.ndarray | None = None
    size_base: float = 0.0     # > 0: every cell's size is size_base * 2**level (level in cell_meta)

    # --- permutations between the reference numbering and the device order
    def cells_to_device(self, a: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(a[self.cell_perm])

    def cells_from_device(self, a: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(a[self.cell_inv])

    def sides_to_device(self, a: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(a[self.side_perm])

    def sides_from_device(self, a: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(a[self.side_inv])


def horton_table(land: np.ndarray, f0: np.ndarray, fc: np.ndarray, hk: np.ndarray) -> np.ndarray:
    """Horton parameters are per land class (map_land_parameters assigns class constants), so the
    cells carry their class and the kernels read a 256-entry table; cell 0 keeps zeros."""
    horton = np.zeros((256, 3), dtype=np.float32)
    for cls in np.unique(land[1:]):
        first = 1 + int(np.flatnonzero(land[1:] == cls)[0])
        horton[cls] = (f0[first], fc[first], hk[first])
    if np.any(horton[land[1:], 0] != f0[1:]) or np.any(horton[land[1:], 2] != hk[1:]):
        raise ValueError('Horton parameters must be constant per land class')
    return horton


def relayout(base: Layout, mesh: Mesh, land_parameters: dict, forcing: Forcing) -> Layout:
    """The layout of another scenario on the mesh and forcing geometry of `base`: only the arrays
    that depend on the land table (side Manning n, the Horton class table) are recomputed."""
    if (mesh.e_num, mesh.s_num) != (base.e_num, base.s_num):
        raise ValueError('relayout needs the mesh of the base layout')
    bid = forcing.boundary_id.astype(np.int64)
    same = (np.array_equal(meta_boundary(base.arrays['cell_meta'])[1:], bid[base.cell_perm][1:])
            and np.array_equal(base.arrays['bweight'], forcing.bweight.astype(np.float32)[base.cell_perm])
            and np.array_equal(base.arrays['rain_index'], forcing.rain_pixel.astype(np.int32)[base.cell_perm])
            and base.n_pixels == forcing.n_pixels and base.b_num == len(forcing.kinds))
    if not same:
        raise ValueError('relayout needs the forcing geometry of the base layout')
    rows = class_tables(land_parameters, mesh.land)
    side_n, f0, fc, hk = land_parameters_of(rows, mesh.land, mesh.topo)
    arrays = dict(base.arrays)
    arrays['manning'] = side_n[base.side_perm]
    arrays['_class_manning'] = rows[0].copy()
    arrays['class_horton'] = horton_table(mesh.land.astype(np.int64), f0, fc, hk).ravel()
    layout = copy.copy(base)
    layout.arrays = arrays
    layout.cell_f0 = f0.astype(np.float32)
    return layout


META_LEVEL_SHIFT = 16   # cell_meta bits 16..19: size level (size = size_base * 2**level)


def meta_boundary(meta: np.ndarray) -> np.ndarray:
    """Boundary index (-1 for none) from cell_meta."""
    return ((meta.astype(np.int64) >> 8) & 255) - 1


def size_levels(size: np.ndarray):
    """(base, levels) with size == base * 2**levels exactly in F32 for every real cell (index 0 is
    the sentinel), levels in 0..15; (0.0, None) when the sizes are not of that form."""
    s = size[1:].astype(np.float32)
    if len(s) == 0 or not np.all(np.isfinite(s)) or s.min() <= 0:
        return 0.0, None
    base = np.float32(s.min())
    levels = np.rint(np.log2(s.astype(np.float64) / float(base))).astype(np.int64)
    if levels.max() > 15 or not np.array_equal(np.ldexp(base, levels).astype(np.float32), s):
        return 0.0, None
    return float(base), np.concatenate([[0], levels])


def cell_records(counts, start, packed, isl, length, size, area) -> np.ndarray:
    """One 16-byte record per cell for the cell passes (int32 [e_num*4]).

    A regular cell has exactly one side per face (nearly every cell of a 2:1 quadtree: a face
    with a finer neighbour has two sides) and each of its side lengths equals its size, and its
    area equals size*size, all in F32: its record holds its left, right, bottom and top side ids,
    and the kernels take the lengths and area from the size. Any other cell holds (-1, first list
    entry, packed face counts, list length) and reads its list, lengths and area as they are
    stored. Every real cell must list a side (the kernels read a list's first entry in place of
    the entries past its end).
    """
    e_num = len(counts)
    totals = counts.sum(axis=1)
    if e_num > 1 and np.any(totals[1:] == 0):
        raise ValueError('a cell lists no side')
    regular = np.all(counts == 1, axis=1)
    regular[0] = False
    quad = np.zeros((e_num, 4), dtype=np.int32)
    quad[:, 0] = -1
    quad[:, 1] = start
    quad[:, 2] = packed.view(np.int32)
    quad[:, 3] = totals
    cand = np.flatnonzero(regular)
    if len(cand):
        ids = isl[start[cand][:, None] + np.arange(4)[None, :]]
        s = size[cand].astype(np.float32)
        same = np.all(length[ids] == s[:, None], axis=1) & (s * s == area[cand])
        quad[cand[same]] = ids[same]
    return quad.ravel()


def build_layout(mesh: Mesh, land_parameters: dict, forcing: Forcing, theta: float,
                 order: str = 'morton') -> Layout:
    """Derive every static kernel array from the mesh, land table and forcing geometry."""
    e_num, s_num = mesh.e_num, mesh.s_num
    topo = mesh.topo
    _, widths, heights, area = mesh.geometry()
    rows = class_tables(land_parameters, mesh.land)
    side_n, f0, fc, hk = land_parameters_of(rows, mesh.land, topo)
    sdc = side_centre_distance(mesh.x, mesh.y, topo)

    cperm = cell_order(mesh, order)
    cinv = np.empty_like(cperm)
    cinv[cperm] = np.arange(e_num)
    # sides ordered by owner cell in the new order, then by original id
    owner = np.where(topo[:, 1] != 0, topo[:, 1], topo[:, 2])
    owner_new = cinv[owner]
    owner_new[0] = -1
    sperm = np.lexsort((np.arange(s_num), owner_new)).astype(np.int64)
    if order == 'identity':
        sperm = np.arange(s_num, dtype=np.int64)
    sinv = np.empty_like(sperm)
    sinv[sperm] = np.arange(s_num)

    # compact side lists in the new numbering, keeping each face's side order
    counts = mesh.face_count[cperm]
    totals = counts.sum(axis=1).astype(np.int64)
    start = np.concatenate([[0], np.cumsum(totals)[:-1]]).astype(np.int64)
    src_start = mesh.face_ptr[cperm, 0].astype(np.int64)
    r_len = int(totals.sum())
    gather = np.repeat(src_start - start, totals) + np.arange(r_len)
    isl = sinv[mesh.isl[gather]].astype(np.int32) if r_len else np.zeros(1, np.int32)
    if np.any(counts > 255):
        raise ValueError('a cell face has more than 255 sides')
    packed = (counts[:, 0].astype(np.uint32) | (counts[:, 1].astype(np.uint32) << 8) |
              (counts[:, 2].astype(np.uint32) << 16) | (counts[:, 3].astype(np.uint32) << 24))

    # the cell across each listed side: a cell is the higher cell of its left/bottom sides and the
    # lower cell of its right/top sides (the reference's face convention, checked here)
    owner_cell = np.repeat(np.arange(e_num), totals)
    face_dir = np.concatenate([np.repeat(np.arange(4), c) for c in counts]) if r_len else np.zeros(0, int)
    ref_sides = mesh.isl[gather] if r_len else np.zeros(0, np.int64)
    lower_of, higher_of = topo[ref_sides, 1], topo[ref_sides, 2]
    upper_face = (face_dir & 1) == 1
    expected_self = np.where(upper_face, lower_of, higher_of)
    if np.any((ref_sides != 0) & (cperm[owner_cell] != expected_self)):
        raise ValueError('a side is listed on a face whose cell is not its left/bottom or right/top neighbour')
    isl_cell = cinv[np.where(upper_face, higher_of, lower_of)].astype(np.int32) if r_len else np.zeros(1, np.int32)

    side_cells = np.empty((s_num, 2), dtype=np.int32)
    side_cells[:, 0] = cinv[topo[sperm, 1]]
    side_cells[:, 1] = cinv[topo[sperm, 2]]
    meta = topo[sperm, 0].astype(np.uint8) | (mesh.sbf[sperm] * 4).astype(np.uint8)

    land = mesh.land.astype(np.int64)
    horton = horton_table(land, f0, fc, hk)
    bid = forcing.boundary_id.astype(np.int64)
    if np.any(bid > 253):
        raise ValueError('at most 254 prescribed boundaries are supported')
    meta_cells = (land | ((bid + 1) << 8)).astype(np.uint32)
    meta_cells[0] = 0
    arrays = dict(
        area=area.astype(np.float32)[cperm], size=widths.astype(np.float32)[cperm],
        cell_meta=meta_cells[cperm], class_horton=horton.ravel(),
        rain_index=forcing.rain_pixel.astype(np.int32)[cperm],
        bweight=forcing.bweight.astype(np.float32)[cperm],
        csr_start=start.astype(np.int32), csr_count=packed.astype(np.uint32), isl=isl, isl_cell=isl_cell,
        side_cells=side_cells.ravel(), side_meta=meta,
        length=mesh.length[sperm].astype(np.float32), sdc=sdc[sperm],
        manning=side_n[sperm], side_key=sperm.astype(np.uint32), _class_manning=rows[0].copy(),
        kind=(forcing.kinds.astype(np.int32) if len(forcing.kinds) else np.zeros(1, np.int32)),
    )
    arrays['cell_quad'] = cell_records(counts, start, packed, isl, arrays['length'], arrays['size'],
                                       arrays['area'])
    size_base, levels = size_levels(arrays['size'])
    if levels is not None:
        arrays['cell_meta'] = arrays['cell_meta'] | (levels.astype(np.uint32) << META_LEVEL_SHIFT)
        arrays['cell_meta'][0] = 0
    layout = Layout(e_num, s_num, cperm, cinv, sperm, sinv, arrays, forcing.n_pixels,
                    len(forcing.kinds), 0, 0, float(theta), size_base=size_base)
    layout.cell_f0 = f0.astype(np.float32)
    if theta != 1.:
        fr = open_face_fractions(forcing.selected_sides, mesh.length, topo, mesh.face_ptr,
                                 mesh.face_count, mesh.isl)
        arrays['open_face'] = np.ascontiguousarray(fr[cperm]).ravel()
    geometry = hanging_geometry(mesh.x, mesh.y, mesh.sx, mesh.sy, topo)
    if geometry is not None:
        arrays['x'] = mesh.x[cperm].astype(np.float32)
        arrays['y'] = mesh.y[cperm].astype(np.float32)
        arrays['edge_index'] = geometry.edge_index[sperm].astype(np.int32)
        arrays['hang_offset'] = geometry.offsets.astype(np.float32).ravel()
        arrays['hang_slot'] = geometry.slots.astype(np.int32).ravel()
        arrays['grad_cell'] = cinv[geometry.cells].astype(np.int32)
        arrays['grad_axes'] = (geometry.axes[:, 0] | (geometry.axes[:, 1] << 1)).astype(np.int32)
        layout.n_grad = len(geometry.cells)
        layout.n_hang = len(geometry.offsets)
    layout.arrays['_r_len'] = np.array([r_len])
    return layout


def initial_state(mesh: Mesh, layout: Layout) -> dict:
    """Cold-start state in original numbering (``initialize``, solver_compact.py:398-410)."""
    e_num, s_num = mesh.e_num, mesh.s_num
    f0 = layout.cell_f0.copy()
    f0[0] = 0.
    return dict(depth=np.zeros(e_num, np.float32), bed=mesh.z.astype(np.float32).copy(),
                rain_time=np.zeros(e_num, np.float32), capacity=f0,
                max_depth=np.zeros(e_num, np.float32), max_speed=np.zeros(e_num, np.float32),
                peak_tick=np.zeros(e_num, np.uint64), q=np.zeros(s_num, np.float32))
