"""NE/NS multi-resolution mesh ingestion.

Text formats and the filtered, renumbered, compact-CSR result follow the reference's
preprocessing (``src/nh_flood_2d/preprocess/domain.py``):

* NE row ``id, n_l, n_r, n_b, n_t, side_ids..., x, y, z, type`` (``domain.py:243-253``);
* NS row ``id, orient, left, right, bottom, top, length, x, y, z, attr`` (``domain.py:268-281``);
* cells with ``z == -9999`` are dropped, cells and sides renumbered in file order and dangling
  sides removed (``_filter_ne_ns``, ``domain.py:221-347``);
* the compact side list ``isl_data`` holds each cell's l, r, b, t side ids in that order, with
  per-face pointers and counts (``create_ne_fdb_compact``, ``domain.py:83-174``);
* side topology ``(orient, lower, higher)`` uses bottom/top for horizontal sides (orient 1) and
  left/right for vertical sides (orient 2) (``_batch_ns_worker``, ``domain.py:439-470``);
* the exterior flag ``sbf`` is ``(lower == 0) XOR (higher == 0)`` (``build_boundary_fdb``,
  ``domain.py:472-479``);
* cell bounds come from the first side of each face and cells must be square
  (``MeshGeometry.from_fdb``, ``preprocess/force.py:35-80``).

Index 0 is the exterior sentinel for both cells and sides, exactly as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

LEFT, RIGHT, BOTTOM, TOP = 0, 1, 2, 3


@dataclass
class Mesh:
    """Filtered compact mesh in the reference's (original) cell and side numbering."""
    x: np.ndarray            # f32 [e_num] cell centres
    y: np.ndarray
    z: np.ndarray            # f32 [e_num] bed
    land: np.ndarray         # u8  [e_num] land class
    face_count: np.ndarray   # i32 [e_num, 4] sides per face (l, r, b, t)
    face_ptr: np.ndarray     # i32 [e_num, 4] start of each face in isl
    isl: np.ndarray          # i32 [R] side ids, per cell l, r, b, t
    topo: np.ndarray         # i32 [s_num, 3] orient, lower, higher
    length: np.ndarray       # f32 [s_num]
    sx: np.ndarray           # f32 [s_num] side midpoints
    sy: np.ndarray
    sz: np.ndarray           # f32 [s_num] (loaded, never used by the step)
    attr: np.ndarray         # u8  [s_num]
    _geometry: dict = field(default_factory=dict, repr=False)

    @property
    def e_num(self) -> int:
        return len(self.x)

    @property
    def s_num(self) -> int:
        return len(self.length)

    @property
    def sbf(self) -> np.ndarray:
        """1 on exterior sides (one missing neighbour), else 0."""
        return ((self.topo[:, 1] == 0) ^ (self.topo[:, 2] == 0)).astype(np.uint8)

    def geometry(self):
        """Cell bounds (l, b, r, t), widths, heights and areas in F64."""
        if not self._geometry:
            first = [self.isl[self.face_ptr[:, d]] if len(self.isl) else
                     np.zeros(self.e_num, dtype=np.int64) for d in range(4)]
            bounds = np.column_stack([
                self.sx[first[LEFT]].astype(float), self.sy[first[BOTTOM]].astype(float),
                self.sx[first[RIGHT]].astype(float), self.sy[first[TOP]].astype(float)])
            widths = bounds[:, 2] - bounds[:, 0]
            heights = bounds[:, 3] - bounds[:, 1]
            area = widths * heights
            area[0] = 0.
            if np.any(widths[1:] <= 0) or np.any(heights[1:] <= 0):
                raise ValueError('Mesh cells must have positive rectangular areas')
            if np.any(widths[1:] != heights[1:]):
                raise ValueError('Mesh cells must be square for the edge-length CFL scale')
            self._geometry.update(bounds=bounds, widths=widths, heights=heights, area=area)
        g = self._geometry
        return g['bounds'], g['widths'], g['heights'], g['area']

    def exterior_edges(self):
        """Exterior sides with their cell, face direction, length and segment endpoints.

        Mirrors ``MeshGeometry.from_fdb`` (``preprocess/force.py:54-79``)."""
        bounds = self.geometry()[0]
        topo = self.topo
        outer = (topo[:, 1] == 0) ^ (topo[:, 2] == 0)
        side_ids = np.flatnonzero(outer)
        cells = topo[side_ids, 1] + topo[side_ids, 2]
        orient = topo[side_ids, 0]
        vertical, horizontal = orient == 2, orient == 1
        if np.any(~(vertical | horizontal)):
            raise ValueError('Exterior sides must be horizontal or vertical')
        sx = self.sx[side_ids].astype(float)
        sy = self.sy[side_ids].astype(float)
        lengths = self.length[side_ids].astype(float)
        on_faces = np.column_stack((
            vertical & (sx == bounds[cells, 0]), vertical & (sx == bounds[cells, 2]),
            horizontal & (sy == bounds[cells, 1]), horizontal & (sy == bounds[cells, 3])))
        if np.any(on_faces.sum(axis=1) != 1):
            raise ValueError('Exterior side does not lie on one cell face')
        segments = np.empty((len(side_ids), 2, 2), dtype=float)
        segments[:, :, 0] = sx[:, None]
        segments[:, :, 1] = sy[:, None]
        segments[vertical, 0, 1] -= lengths[vertical] / 2
        segments[vertical, 1, 1] += lengths[vertical] / 2
        segments[horizontal, 0, 0] -= lengths[horizontal] / 2
        segments[horizontal, 1, 0] += lengths[horizontal] / 2
        return side_ids, cells, on_faces.argmax(axis=1), lengths, segments

    def face_sides(self, cell: int, direction: int) -> np.ndarray:
        start = self.face_ptr[cell, direction]
        return self.isl[start:start + self.face_count[cell, direction]]


def csr_pointers(counts: np.ndarray) -> np.ndarray:
    """Per-face start offsets of a compact l, r, b, t side list (cell 0 owns nothing)."""
    totals = counts.sum(axis=1).astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(totals)[:-1]])
    ptr = starts[:, None] + np.concatenate(
        [np.zeros((len(counts), 1), dtype=np.int64), np.cumsum(counts, axis=1)[:, :3]], axis=1)
    return ptr.astype(np.int32)


def _split(line: str) -> list[str]:
    return line.strip().split(',')


def filter_rows(ne_lines: list[str], ns_lines: list[str]):
    """Drop ``z == -9999`` cells and dangling sides, renumber both in file order.

    Restates ``_filter_ne_ns`` (``preprocess/domain.py:221-347``). Returns lists of split rows.
    """
    ne_map: dict[int, int] = {}
    kept_ne = []
    for line in ne_lines:
        parts = _split(line)
        if not parts or parts == ['']:
            continue
        if float(parts[-2]) != -9999:
            ne_map[int(parts[0])] = len(ne_map) + 1
            kept_ne.append(parts)
    ns_map: dict[int, int] = {}
    kept_ns = []
    for line in ns_lines:
        parts = _split(line)
        if len(parts) < 6:
            continue
        orient = int(parts[1])
        left, right, bottom, top = (ne_map.get(int(v), 0) for v in parts[2:6])
        keep = (bottom != 0 or top != 0) if orient == 1 else (left != 0 or right != 0)
        if keep:
            ns_map[int(parts[0])] = len(ns_map) + 1
            parts[0] = str(len(ns_map))
            parts[2:6] = [str(left), str(right), str(bottom), str(top)]
            kept_ns.append(parts)
    out_ne = []
    for parts in kept_ne:
        parts[0] = str(ne_map[int(parts[0])])
        total = sum(int(v) for v in parts[1:5])
        for i in range(5, 5 + total):
            parts[i] = str(ns_map.get(int(parts[i]), 0))
        out_ne.append(parts)
    return out_ne, kept_ns


def compact_mesh(ne_rows: list[list[str]], ns_rows: list[list[str]]) -> Mesh:
    """Build the compact CSR mesh from filtered rows (``create_ne_fdb_compact`` and the NS worker)."""
    e_num, s_num = len(ne_rows) + 1, len(ns_rows) + 1
    counts = np.zeros((e_num, 4), dtype=np.int32)
    x = np.zeros(e_num, dtype=np.float32)
    y = np.zeros_like(x)
    z = np.zeros_like(x)
    land = np.zeros(e_num, dtype=np.uint8)
    ids: list[int] = []
    for i, parts in enumerate(ne_rows, start=1):
        c = [int(v) for v in parts[1:5]]
        counts[i] = c
        total = sum(c)
        ids.extend(int(v) for v in parts[5:5 + total])
        x[i], y[i], z[i] = float(parts[-4]), float(parts[-3]), float(parts[-2])
        land[i] = int(parts[-1])
    isl = np.asarray(ids, dtype=np.int32)
    ptr = csr_pointers(counts)
    topo = np.zeros((s_num, 3), dtype=np.int32)
    length = np.zeros(s_num, dtype=np.float32)
    sx, sy, sz = (np.zeros(s_num, dtype=np.float32) for _ in range(3))
    attr = np.zeros(s_num, dtype=np.uint8)
    for i, parts in enumerate(ns_rows, start=1):
        orient = int(parts[1])
        nb = [int(v) for v in parts[2:6]]
        topo[i] = (orient, nb[2], nb[3]) if orient == 1 else (orient, nb[0], nb[1])
        length[i], sx[i], sy[i], sz[i] = (float(v) for v in parts[6:10])
        attr[i] = int(float(parts[10]))
    return Mesh(x, y, z, land, counts, ptr, isl, topo, length, sx, sy, sz, attr)


def load_text_mesh(ne_path: str | Path, ns_path: str | Path) -> Mesh:
    """Read NE/NS text, filter, renumber and compact (``prepare_domain``, ``domain.py:590-606``)."""
    ne_lines = Path(ne_path).read_text(encoding='utf-8').splitlines()
    ns_lines = Path(ns_path).read_text(encoding='utf-8').splitlines()
    ne_rows, ns_rows = filter_rows(ne_lines, ns_lines)
    return compact_mesh(ne_rows, ns_rows)


def write_text_mesh(mesh: Mesh, ne_path: str | Path, ns_path: str | Path) -> None:
    """Write a mesh back to NE/NS text (the reference's input format)."""
    rows = []
    for i in range(1, mesh.e_num):
        c = mesh.face_count[i]
        sides = mesh.isl[mesh.face_ptr[i, 0]:mesh.face_ptr[i, 0] + int(c.sum())]
        rows.append(','.join(map(str, [i, *c.tolist(), *sides.tolist(),
                                       repr(float(mesh.x[i])), repr(float(mesh.y[i])),
                                       repr(float(mesh.z[i])), int(mesh.land[i])])))
    Path(ne_path).write_text('\n'.join(rows) + '\n')
    rows = []
    for i in range(1, mesh.s_num):
        orient, lower, higher = mesh.topo[i].tolist()
        nb = [0, 0, lower, higher] if orient == 1 else [lower, higher, 0, 0]
        rows.append(','.join(map(str, [i, orient, *nb, repr(float(mesh.length[i])),
                                       repr(float(mesh.sx[i])), repr(float(mesh.sy[i])),
                                       repr(float(mesh.sz[i])), int(mesh.attr[i])])))
    Path(ns_path).write_text('\n'.join(rows) + '\n')
