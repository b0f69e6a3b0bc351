"""Host-side hydraulic set-up: land parameters, open-face fractions, still water, hanging edges.

Restatements (own code, same results) of:
  map_land_parameters            preprocess/hydraulics.py:63-93
  open_boundary_face_fractions   preprocess/hydraulics.py:50-61
  connected_still_water          preprocess/hydraulics.py:7-47
  hanging_geometry               core/pressure.py:19-45
  side centre distance (sdc)     core/solver_compact.py:404-410 (``initialize``)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from .config import HortonInfiltration, LandParameters

MM_PER_H_TO_M_PER_S = np.float32(1. / 3_600_000.)


def map_land_parameters(table: dict, land_types: np.ndarray, topo: np.ndarray):
    """Side Manning n and cell Horton f0/fc (m/s), k (1/h), all F32.

    A side takes the F32 mean of its two cells; an exterior side takes its interior cell's n.
    Cell 0 (the exterior) carries zeros whatever its land class.
    """
    classes = np.unique(land_types[1:])
    table = {int(k): (v if isinstance(v, LandParameters) else LandParameters.model_validate(v))
             for k, v in table.items()}
    missing = sorted(set(int(c) for c in classes) - table.keys())
    if missing:
        raise ValueError(f'land_parameters missing land types: {missing}')
    cell_n = np.zeros(len(land_types), dtype=np.float32)
    f0 = np.zeros_like(cell_n)
    fc = np.zeros_like(cell_n)
    k = np.zeros_like(cell_n)
    for land in classes:
        params = table[int(land)]
        mask = land_types == land
        cell_n[mask] = params.manning_n
        if isinstance(params.infiltration, HortonInfiltration):
            f0[mask] = np.float32(params.infiltration.f0_mm_h) * MM_PER_H_TO_M_PER_S
            fc[mask] = np.float32(params.infiltration.fc_mm_h) * MM_PER_H_TO_M_PER_S
            k[mask] = params.infiltration.k_per_h
    cell_n[0] = f0[0] = fc[0] = k[0] = 0.
    lower, higher = topo[:, 1], topo[:, 2]
    side_n = (cell_n[lower] + cell_n[higher]) * np.float32(.5)
    exterior = (lower == 0) | (higher == 0)
    side_n[exterior] = cell_n[lower[exterior]] + cell_n[higher[exterior]]
    return side_n.astype(np.float32), f0, fc, k


def open_face_fractions(selected: np.ndarray, lengths: np.ndarray, topo: np.ndarray,
                        face_ptr: np.ndarray, face_count: np.ndarray, isl: np.ndarray) -> np.ndarray:
    """Per cell face: selected exterior side length / total face length (F64 -> F32)."""
    out = np.zeros((len(face_ptr), 4), dtype=np.float32)
    cells = np.unique(topo[selected, 1:].sum(axis=1))
    for cell in cells:
        for d in range(4):
            ids = isl[face_ptr[cell, d]:face_ptr[cell, d] + face_count[cell, d]]
            l64 = lengths[ids].astype(np.float64)
            out[cell, d] = l64[selected[ids]].sum() / l64.sum()
    return out


def side_centre_distance(x: np.ndarray, y: np.ndarray, topo: np.ndarray) -> np.ndarray:
    """F32 ``max(|dx|, 0.01)`` for vertical sides, ``max(|dy|, 0.01)`` for horizontal ones."""
    lower, higher = topo[:, 1], topo[:, 2]
    dx = np.abs(x[higher] - x[lower])
    dy = np.abs(y[higher] - y[lower])
    d = np.where(topo[:, 0] == 2, dx, dy).astype(np.float32)
    out = np.maximum(d, np.float32(0.01))
    out[0] = 0.
    return out


@dataclass
class HangingGeometry:
    edge_index: np.ndarray    # [s_num] -1 or edge slot
    offsets: np.ndarray       # [E, 2] F32 tangential offset of the lower / higher centre
    slots: np.ndarray         # [E, 2] gradient slot or -1
    cells: np.ndarray         # [G] gradient cells (sorted)
    axes: np.ndarray          # [G, 2] 1 where the x / y gradient is needed


def hanging_geometry(x, y, side_x, side_y, topo) -> HangingGeometry | None:
    """Sparse reconstruction indices; None on a mesh whose centres are aligned with every side."""
    lower, higher = topo[:, 1], topo[:, 2]
    interior = (lower != 0) & (higher != 0)
    vertical = topo[:, 0] == 2
    tangent = np.where(vertical, side_y, side_x)
    centre = np.where(vertical[:, None], y[topo[:, 1:]], x[topo[:, 1:]])
    offsets = tangent[:, None] - centre
    edges = np.flatnonzero(interior & np.any(offsets != 0., axis=1))
    if not len(edges):
        return None
    offsets = offsets[edges].astype(np.float32)
    nb = topo[edges, 1:]
    displaced = offsets != 0.
    cells = np.unique(nb[displaced]).astype(np.int32)
    slots = np.full(nb.shape, -1, dtype=np.int32)
    slots[displaced] = np.searchsorted(cells, nb[displaced])
    axes = np.zeros((len(cells), 2), dtype=np.int32)
    edge_axis = np.broadcast_to(vertical[edges, None], nb.shape).astype(np.int32)
    axes[slots[displaced], edge_axis[displaced]] = 1
    edge_index = np.full(len(topo), -1, dtype=np.int32)
    edge_index[edges] = np.arange(len(edges), dtype=np.int32)
    return HangingGeometry(edge_index, offsets, slots, cells, axes)


def connected_still_water(bed, side_crest, closed_sides, topo, boundaries, area):
    """Fill water bodies reachable from level boundaries at their t = 0 level.

    Comparisons are made in F64 against the input level; cells are connected across interior,
    open sides whose crest (max of both beds and the side crest) is strictly below the level.
    ``boundaries`` are objects with ``name``, ``kind``, ``cells`` and ``series.at``.
    """
    bed = bed.astype(np.float64)
    side_crest = np.asarray(side_crest, dtype=np.float64)
    levels: dict[float, list] = {}
    for b in boundaries:
        if b.kind == 'level':
            levels.setdefault(b.series.at(0.), []).append(b)
    depth = np.zeros(len(bed), dtype=np.float32)
    lower, higher = topo[:, 1], topo[:, 2]
    crest = np.maximum(np.maximum(bed[lower], bed[higher]), side_crest)
    interior = (lower != 0) & (higher != 0) & ~np.asarray(closed_sides).astype(bool)
    bodies = []
    n = len(bed)
    for level, group in levels.items():
        seeds = np.unique(np.concatenate([np.asarray(b.cells) for b in group]))
        seeds = seeds[bed[seeds] < level]
        if not len(seeds):
            continue
        edges = interior & (crest < level)
        graph = coo_matrix((np.ones(int(edges.sum())), (lower[edges], higher[edges])), shape=(n, n))
        _, labels = connected_components(graph, directed=False)
        selected = np.unique(labels[seeds])
        wet = np.isin(labels, selected)
        for other_level, other_group in levels.items():
            if other_level != level:
                for b in other_group:
                    if np.any(wet[np.asarray(b.cells)]):
                        names = ', '.join(x.name for x in group)
                        raise ValueError('Conflicting initial levels on connected water body: '
                                         f'{names} ({level}) and {b.name} ({other_level})')
        depth[wet] = level - bed[wet]
        bodies.append(dict(level_m=float(level), boundaries=[b.name for b in group],
                           connected_components=len(selected), wet_cells=int(np.count_nonzero(wet))))
    return depth, dict(wet_cells=int(np.count_nonzero(depth)),
                       storage_m3=float(np.dot(depth.astype(np.float64), area)), levels=bodies)
