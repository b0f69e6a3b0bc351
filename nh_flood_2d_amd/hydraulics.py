"""Host-side hydraulic set-up of a mesh, computed once before the first step.

* the land table as per-class lookup tables, indexed by each cell's class code to give cell
  Horton parameters and, averaged over a side's two cells, the side's Manning n
  (the reference's ``map_land_parameters``, preprocess/hydraulics.py:63-93);
* the open fraction of every cell face that holds sides of a prescribed boundary line
  (``open_boundary_face_fractions``, preprocess/hydraulics.py:50-61);
* the water bodies a level boundary fills at the start: every cell connected to the boundary
  through sides whose crest lies below the level (``connected_still_water``,
  preprocess/hydraulics.py:7-47);
* the index tables of the hanging-edge pressure correction (``hanging_geometry``,
  core/pressure.py:19-45) and the centre distance of every side (``initialize``,
  core/solver_compact.py:404-410).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from .config import HortonInfiltration, LandParameters

MM_PER_H_TO_M_PER_S = np.float32(1. / 3_600_000.)
CLASS_CODES = 256          # land class codes are one byte


# ------------------------------------------------------------------------------ land table

def class_tables(table: dict, land_types: np.ndarray) -> np.ndarray:
    """[4, 256] F32 rows (Manning n, Horton f0 m/s, fc m/s, k 1/h) by land class code, filled for
    the classes the real cells use; a class used by a cell but absent from the table is an
    error (the reference's message)."""
    spec = {int(code): value if isinstance(value, LandParameters) else LandParameters.model_validate(value)
            for code, value in table.items()}
    used = [int(code) for code in np.unique(land_types[1:])]
    absent = sorted(code for code in used if code not in spec)
    if absent:
        raise ValueError(f'land_parameters missing land types: {absent}')
    rows = np.zeros((4, CLASS_CODES), dtype=np.float32)
    for code in used:
        entry = spec[code]
        rows[0, code] = entry.manning_n
        horton = entry.infiltration
        if isinstance(horton, HortonInfiltration):
            rows[1, code] = np.float32(horton.f0_mm_h) * MM_PER_H_TO_M_PER_S
            rows[2, code] = np.float32(horton.fc_mm_h) * MM_PER_H_TO_M_PER_S
            rows[3, code] = horton.k_per_h
    return rows


def map_land_parameters(table: dict, land_types: np.ndarray, topo: np.ndarray):
    """Side Manning n and cell Horton f0, fc (m/s), k (1/h), all F32.

    Cell 0, the exterior, carries zeros whatever its class code. A side's n is the F32 mean of
    its two cells' n; on an exterior side the exterior's zero drops out and the interior cell's
    n is taken as it is.
    """
    return land_parameters_of(class_tables(table, land_types), land_types, topo)


def land_parameters_of(rows: np.ndarray, land_types: np.ndarray, topo: np.ndarray):
    """map_land_parameters from the class rows of class_tables."""
    per_cell = rows[:, np.asarray(land_types, dtype=np.intp)]
    per_cell[:, 0] = 0.
    cell_n, f0, fc, k = (np.ascontiguousarray(row) for row in per_cell)
    return side_manning(cell_n, np.asarray(topo)[:, 1:]), f0, fc, k


def side_manning(cell_n: np.ndarray, ends: np.ndarray) -> np.ndarray:
    """F32 side n from its two cells' n (`ends`: [sides, 2] cell numbers, 0 the exterior, whose n
    is zero): the mean, or on an exterior side the interior cell's n (n + 0)."""
    both = cell_n[ends[:, 0]] + cell_n[ends[:, 1]]
    return np.where(np.any(ends == 0, axis=1), both, both * np.float32(.5)).astype(np.float32)


# ------------------------------------------------------------------------------ open faces

def face_entries(face_ptr: np.ndarray, face_count: np.ndarray, isl: np.ndarray):
    """(face number, side id) of every side-list entry, face number = 4 * cell + direction."""
    counts = np.asarray(face_count, dtype=np.int64).ravel()
    firsts = np.asarray(face_ptr, dtype=np.int64).ravel()
    face = np.repeat(np.arange(counts.size), counts)
    within = np.arange(face.size) - np.repeat(np.cumsum(counts) - counts, counts)
    return face, np.asarray(isl)[np.repeat(firsts, counts) + within]


def open_face_fractions(selected: np.ndarray, lengths: np.ndarray, topo: np.ndarray,
                        face_ptr: np.ndarray, face_count: np.ndarray, isl: np.ndarray) -> np.ndarray:
    """[e_num, 4] F32: for each face of a cell that owns a selected (prescribed, exterior) side,
    the selected length over the face's whole length, both summed in F64 in list order; zero
    on every face of the other cells."""
    e_num = len(face_ptr)
    face, side = face_entries(face_ptr, face_count, isl)
    length = np.asarray(lengths)[side].astype(np.float64)
    whole = np.bincount(face, weights=length, minlength=4 * e_num)
    chosen = np.bincount(face, weights=np.where(np.asarray(selected)[side], length, 0.), minlength=4 * e_num)
    owners = np.zeros(e_num, dtype=bool)
    owners[np.asarray(topo)[selected, 1:].sum(axis=1)] = True      # the interior cell of each side
    with np.errstate(invalid='ignore', divide='ignore'):
        ratio = (chosen / whole).reshape(e_num, 4)
    return np.where(owners[:, None], ratio, 0.).astype(np.float32)


def side_centre_distance(x: np.ndarray, y: np.ndarray, topo: np.ndarray) -> np.ndarray:
    """F32 ``max(|dx|, 0.01)`` for vertical sides, ``max(|dy|, 0.01)`` for horizontal ones."""
    lower, higher = topo[:, 1], topo[:, 2]
    dx = np.abs(x[higher] - x[lower])
    dy = np.abs(y[higher] - y[lower])
    d = np.where(topo[:, 0] == 2, dx, dy).astype(np.float32)
    out = np.maximum(d, np.float32(0.01))
    out[0] = 0.
    return out


# ------------------------------------------------------------------------------ hanging edges

@dataclass
class HangingGeometry:
    edge_index: np.ndarray    # [s_num] -1 or edge slot
    offsets: np.ndarray       # [E, 2] F32 tangential offset of the lower / higher centre
    slots: np.ndarray         # [E, 2] gradient slot or -1
    cells: np.ndarray         # [G] gradient cells (sorted)
    axes: np.ndarray          # [G, 2] 1 where the x / y gradient is needed


def tangential_offsets(x, y, side_x, side_y, topo):
    """[s_num, 2] position of the side's midpoint along the side, minus that of its lower and
    higher cell centres, and the side's tangential axis (1 = y for a vertical side)."""
    ends = np.asarray(topo)[:, 1:]
    axis = (np.asarray(topo)[:, 0] == 2).astype(np.intp)
    midpoint = np.where(axis == 1, side_y, side_x)
    centres = np.stack([x[ends], y[ends]])                       # [2 axes, s_num, 2 ends]
    along = np.take_along_axis(centres, axis[None, :, None], axis=0)[0]
    return midpoint[:, None] - along, axis


def hanging_geometry(x, y, side_x, side_y, topo) -> HangingGeometry | None:
    """Index tables of the pressure correction: every interior side with a cell centre off its
    midpoint (a hanging edge of a 2:1 interface), the cells such edges need a tangential level
    gradient of (sorted, with the axes needed), and for each edge end its cell's gradient slot.
    None when every centre lies on its sides' midpoints (a mesh with no hanging edge)."""
    offsets, axis = tangential_offsets(x, y, side_x, side_y, topo)
    ends = np.asarray(topo)[:, 1:]
    edges = np.flatnonzero(np.all(ends != 0, axis=1) & np.any(offsets != 0., axis=1))
    if edges.size == 0:
        return None
    offsets = offsets[edges].astype(np.float32)
    shifted = offsets != 0.
    cells, slot_of = np.unique(ends[edges][shifted], return_inverse=True)
    slots = np.full(offsets.shape, -1, dtype=np.int32)
    slots[shifted] = slot_of
    axes = np.zeros((cells.size, 2), dtype=np.int32)
    axes[slot_of, np.broadcast_to(axis[edges, None], offsets.shape)[shifted]] = 1
    edge_index = np.full(len(ends), -1, dtype=np.int32)
    edge_index[edges] = np.arange(edges.size, dtype=np.int32)
    return HangingGeometry(edge_index, offsets, slots, cells.astype(np.int32), axes)


# ------------------------------------------------------------------------------ still water

def level_groups(boundaries) -> dict:
    """Level boundaries grouped by their level at t = 0, in order of first appearance."""
    groups: dict[float, list] = {}
    for b in boundaries:
        if b.kind == 'level':
            groups.setdefault(b.series.at(0.), []).append(b)
    return groups


def filled_by(level: float, members, bed, crest, open_side, ends):
    """Cells a level fills: the connected pieces of the mesh graph over open sides with crest
    strictly below the level that hold a member boundary cell lying below it. Returns (mask,
    number of pieces) or (None, 0) when no boundary cell lies below the level."""
    seeds = np.unique(np.concatenate([np.asarray(b.cells) for b in members]))
    seeds = seeds[bed[seeds] < level]
    if seeds.size == 0:
        return None, 0
    links = ends[open_side & (crest < level)]
    n = len(bed)
    graph = csr_matrix((np.ones(len(links), dtype=np.int8), (links[:, 0], links[:, 1])), shape=(n, n))
    _, piece = connected_components(graph, directed=False)
    pieces = np.unique(piece[seeds])
    return np.isin(piece, pieces), int(pieces.size)


def connected_still_water(bed, side_crest, closed_sides, topo, boundaries, area):
    """Initial depth of the water bodies level boundaries fill at their t = 0 level.

    Levels are compared in F64. A body reached by two different levels is an error (the
    reference's message). Returns the F32 depth and a record of the bodies.
    """
    z = np.asarray(bed).astype(np.float64)
    ends = np.asarray(topo)[:, 1:]
    crest = np.maximum(z[ends].max(axis=1), np.asarray(side_crest, dtype=np.float64))
    open_side = np.all(ends != 0, axis=1) & ~np.asarray(closed_sides).astype(bool)
    groups = level_groups(boundaries)
    depth = np.zeros(len(z), dtype=np.float32)
    bodies = []
    for level, members in groups.items():
        wet, pieces = filled_by(level, members, z, crest, open_side, ends)
        if wet is None:
            continue
        clash = next((b for other, others in groups.items() if other != level for b in others
                      if wet[np.asarray(b.cells)].any()), None)
        if clash is not None:
            raise ValueError('Conflicting initial levels on connected water body: '
                             f"{', '.join(b.name for b in members)} ({level}) and "
                             f'{clash.name} ({clash.series.at(0.)})')
        depth[wet] = level - z[wet]
        bodies.append(dict(level_m=float(level), boundaries=[b.name for b in members],
                           connected_components=pieces, wet_cells=int(np.count_nonzero(wet))))
    return depth, dict(wet_cells=int(np.count_nonzero(depth)),
                       storage_m3=float(np.dot(depth.astype(np.float64), area)), levels=bodies)
