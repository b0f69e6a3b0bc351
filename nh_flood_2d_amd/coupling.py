"""Host-side pipe-network coupling: surface/pipe exchange windows around the device step.

Restates the reference's coupled mode (``core/coupled/exchange.py``, ``core/coupled/runtime.py``,
``preprocess/pipe.py``, the window logic of ``solver_compact.py:921-929,986-988`` and the
compensated per-cell volume of ``:535-539``):

* :func:`prepare_pipe` maps the nodes of a SWMM ``.inp`` onto the surface cells (primary cell and
  the weak neighbourhood) and keeps the junction rims;
* every coupling window starts with :func:`compute_exchange`: orifice drainage from the cells and
  overflow from surcharged junctions give one F32 source rate per cell, uploaded to the device's
  ``pipe_q`` through a pinned buffer; the device step applies it (donor-limited) and sums each
  cell's applied volume in F32 with a compensation term;
* at the window end :meth:`PipeCoupling.end_window` pairs the measured surface volumes with the
  junction inflows, so the surface and the pipe model exchange exactly the same volume, and
  advances the pipe model.

The 1D model runs behind an actor interface (``start() -> heads``, ``advance(boundaries,
seconds) -> (heads, seconds)``, ``close()``). :class:`SwmmActor` drives SWMM through pyswmm when it
is installed; :class:`StorageActor` is a simple stand-in (prismatic junction storage draining
to the outfalls) for tests and for runs without SWMM.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from pydantic import BaseModel, Field

# orifice law q = C h^1.5 with C = 0.85 pi 0.8 (core/coupled/exchange.py:13-15)
ORIFICE_COEFFICIENT = 0.85 * math.pi * 0.8
MAX_OVERFLOW_HEAD_M = 3.0
OVERFLOW_RELAXATION = 0.5
IMPERVIOUS_BUILDING_CLASS = 8   # cells of land class 8 never exchange (exchange.py:52,79)


class PipeConfig(BaseModel):
    """input/pipe.py: the SWMM input, the pipe working directory and the window length."""
    inp: str
    pipe_dir: str
    coupling_interval: int = Field(default=600, gt=0)   # whole seconds
    exchange_timeout: float = Field(default=600.0, gt=0)
    weak_dist_thresh: float = 50.0

    @property
    def topology_file(self) -> Path:
        return Path(self.pipe_dir) / 'pipe_topology.npz'

    @property
    def inp_runtime(self) -> Path:
        return Path(self.pipe_dir) / Path(self.inp).name


@dataclass
class PipeTopology:
    """Nodes of the pipe network and the surface cells each one exchanges with."""
    node_names: list
    is_outfall: np.ndarray          # [nodes] bool
    x: np.ndarray
    y: np.ndarray
    primary_ei: np.ndarray          # [nodes] cell whose box holds the node, 0 = none
    topo_ptr: np.ndarray            # [nodes + 1] weak-neighbourhood ranges
    topo_ei: np.ndarray             # weak-neighbourhood cells
    rims: np.ndarray                # [nodes] junction rim elevation, 0 for outfalls

    def save(self, path) -> None:
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, node_names=np.array(self.node_names, dtype=str), is_outfall=self.is_outfall,
                 x=self.x, y=self.y, primary_ei=self.primary_ei, topo_ptr=self.topo_ptr,
                 topo_ei=self.topo_ei, rims=self.rims)

    @classmethod
    def load(cls, path) -> 'PipeTopology':
        with np.load(path, allow_pickle=False) as z:
            t = cls([str(s) for s in z['node_names']], z['is_outfall'].astype(bool), z['x'], z['y'],
                    z['primary_ei'].astype(np.int64), z['topo_ptr'].astype(np.int64),
                    z['topo_ei'].astype(np.int64), z['rims'].astype(np.float64))
        t.validate()
        return t

    def validate(self, e_num: int | None = None) -> None:
        n = len(self.node_names)
        if len(self.primary_ei) != n or len(self.rims) != n or len(self.topo_ptr) != n + 1:
            raise ValueError('pipe topology and node arrays have inconsistent sizes')
        if self.topo_ptr[0] != 0 or np.any(np.diff(self.topo_ptr) < 0) or self.topo_ptr[-1] > len(self.topo_ei):
            raise ValueError('pipe topology pointer is invalid')
        if len(set(self.node_names)) != n:
            raise ValueError('pipe topology contains duplicate node names')
        if e_num is not None:
            linked = np.concatenate([self.primary_ei, self.topo_ei[:int(self.topo_ptr[-1])]])
            if np.any(linked >= e_num) or np.any(linked < 0):
                raise ValueError('pipe topology cell index is outside the surface mesh')


# ------------------------------------------------------------------------------ .inp and mapping

def _sections(inp_path) -> dict:
    """Rows of each [SECTION] of a SWMM input, comments and blank lines dropped."""
    out, name = {}, None
    for line in Path(inp_path).read_text(encoding='utf-8').splitlines():
        s = line.strip()
        if not s or s.startswith(';'):
            continue
        if s.startswith('['):
            name = s.lower()
            out.setdefault(name, [])
            continue
        if name is not None:
            out[name].append(s.split())
    return out


def parse_inp(inp_path):
    """Node coordinates (file order of [COORDINATES]), outfall flags, and junction rims
    (invert elevation + maximum depth) of a SWMM input (preprocess/pipe.py:33-102)."""
    sec = _sections(inp_path)
    outfalls = {row[0] for row in sec.get('[outfalls]', []) if row}
    nodes = [(row[0], float(row[1]), float(row[2]), row[0] in outfalls)
             for row in sec.get('[coordinates]', []) if len(row) >= 3]
    rims = {row[0]: float(row[1]) + float(row[2]) for row in sec.get('[junctions]', []) if len(row) >= 3}
    return nodes, rims


def map_nodes(nodes, x: np.ndarray, y: np.ndarray, size: np.ndarray, dist_thresh: float):
    """Primary cell of each node: the first cell (in index order) whose square holds it, else the
    nearest centre within 2 x dist_thresh (0 when none); and each junction's weak neighbourhood:
    cells whose square holds it or whose centre is within dist_thresh, primary excluded
    (preprocess/pipe.py:105-186). Coordinates compare in F32, as the reference's fields."""
    cx, cy = x[1:].astype(np.float32), y[1:].astype(np.float32)
    half = (size[1:].astype(np.float32) * np.float32(0.5)).astype(np.float32)
    primary = np.zeros(len(nodes), dtype=np.int64)
    rows, ptr = [], [0]
    limit2 = np.float32(dist_thresh * 2)
    for i, (name, nx, ny, outfall) in enumerate(nodes):
        px, py = np.float32(nx), np.float32(ny)
        inside = (px >= cx - half) & (px <= cx + half) & (py >= cy - half) & (py <= cy + half)
        hits = np.flatnonzero(inside)
        if len(hits):
            primary[i] = hits[0] + 1
        else:
            d = np.sqrt((cx - px) * (cx - px) + (cy - py) * (cy - py)).astype(np.float32)
            k = int(np.argmin(d)) if len(d) else -1
            primary[i] = k + 1 if k >= 0 and d[k] <= limit2 else 0
        if outfall:
            ptr.append(ptr[-1])
            continue
        dx, dy = cx - nx, cy - ny
        mask = ((np.abs(dx) <= half) & (np.abs(dy) <= half)) | ((dx * dx + dy * dy) <= dist_thresh ** 2)
        if primary[i] > 0:
            mask[primary[i] - 1] = False
        row = np.flatnonzero(mask) + 1
        rows.append(row)
        ptr.append(ptr[-1] + len(row))
    topo = np.concatenate(rows).astype(np.int64) if rows else np.zeros(0, np.int64)
    return primary, np.array(ptr, dtype=np.int64), topo


def prepare_pipe(pipe_cfg: PipeConfig, mesh) -> PipeTopology:
    """Build and save the pipe topology of `pipe_cfg.inp` on a surface mesh (preprocess/pipe.py:190)."""
    nodes, rims = parse_inp(pipe_cfg.inp)
    if not nodes:
        raise ValueError(f'No nodes found in SWMM .inp: {pipe_cfg.inp}')
    _, widths, _, _ = mesh.geometry()
    primary, ptr, topo = map_nodes(nodes, mesh.x, mesh.y, widths, pipe_cfg.weak_dist_thresh)
    t = PipeTopology([n[0] for n in nodes], np.array([n[3] for n in nodes], bool),
                     np.array([n[1] for n in nodes]), np.array([n[2] for n in nodes]), primary, ptr, topo,
                     np.array([0. if n[3] else float(np.float32(rims.get(n[0], 0.))) for n in nodes]))
    t.validate(mesh.e_num)
    t.save(pipe_cfg.topology_file)
    return t


# ------------------------------------------------------------------------------ one window

@dataclass
class Links:
    """The (cell, junction) pairs of one mesh that exchange water, sorted by cell then junction:
    a junction drains its primary cell and its weak neighbourhood, except cells of land class 8
    and the exterior; outfalls exchange nothing (only their stage is read). `count` is the number
    of junctions of each cell."""
    cell: np.ndarray
    node: np.ndarray
    count: np.ndarray

    @classmethod
    def of(cls, topology: PipeTopology, land_types: np.ndarray, n_cells: int) -> 'Links':
        n_nodes = len(topology.node_names)
        sizes = np.diff(np.asarray(topology.topo_ptr, dtype=np.int64))
        node = np.concatenate([np.arange(n_nodes), np.repeat(np.arange(n_nodes), sizes)]).astype(np.int64)
        cell = np.concatenate([np.asarray(topology.primary_ei, dtype=np.int64),
                               np.asarray(topology.topo_ei, dtype=np.int64)])
        keep = ~np.asarray(topology.is_outfall, dtype=bool)[node] & (cell != 0)
        keep[keep] = np.asarray(land_types)[cell[keep]] != IMPERVIOUS_BUILDING_CLASS
        pair = np.unique(cell[keep] * n_nodes + node[keep])
        cell, node = pair // n_nodes, pair % n_nodes
        return cls(cell, node, np.bincount(cell, minlength=n_cells))

    def nodes_of(self, c: int) -> tuple:
        lo, hi = np.searchsorted(self.cell, [c, c + 1])
        return tuple(int(n) for n in self.node[lo:hi])

    __getitem__ = nodes_of


def node_cells(topology: PipeTopology, land_types: np.ndarray, n_cells: int) -> list:
    """For each cell, the sorted junctions it exchanges with (see Links)."""
    links = Links.of(topology, land_types, n_cells)
    out = [()] * n_cells
    for c in np.unique(links.cell):
        out[int(c)] = links.nodes_of(int(c))
    return out


def orifice(head: float) -> float:
    """Orifice rate C h^1.5 (m^3/s) for a head h (m), evaluated by the C library's pow."""
    return ORIFICE_COEFFICIENT * head ** 1.5


def compute_exchange(window_dt: float, depth, bed, area, land_types, topology: PipeTopology, heads: dict):
    """Source rates (F32, m^3/s, + into the surface), junction boundaries, diagnostics and the
    cell -> junction links for one window (core/coupled/exchange.py:18-122).

    * Drainage: each linked cell loses min(C h^1.5, h A / window), split equally over its
      junctions (the water present at the window start bounds it).
    * Overflow: a surcharged junction (head above rim + h of its primary cell) returns
      min(C dh^1.5, dh A / window / k) x 0.5 to its primary cell, dh capped at 3 m, k the
      number of junctions sharing that primary cell.
    * The surface receives the F32 rounding of each cell's source; that rounding is charged back
      to the cell's junctions, so surface and pipe see the same volume.
    Every per-junction sum adds its cells in increasing cell order.
    """
    depth = np.asarray(depth, dtype=np.float64)
    bed = np.asarray(bed, dtype=np.float64)
    area = np.asarray(area, dtype=np.float64)
    n_cells = len(depth)
    names = topology.node_names
    links = Links.of(topology, land_types, n_cells)
    linked = np.flatnonzero(links.count)
    source = np.zeros(n_cells, dtype=np.float64)
    node_flow = np.zeros(len(names), dtype=np.float64)
    # drainage of every linked cell, then its equal shares to the junctions
    drain = np.array([min(orifice(float(depth[c])), float(depth[c]) * float(area[c]) / window_dt) for c in linked])
    source[linked] -= drain
    drainage = float(np.add.accumulate(np.concatenate([[0.], drain]))[-1])
    rate = np.zeros(n_cells)
    rate[linked] = drain
    np.add.at(node_flow, links.node, rate[links.cell] / links.count[links.cell])
    # overflow of surcharged junctions into their primary cell
    outfall = np.asarray(topology.is_outfall, dtype=bool)
    primary = np.asarray(topology.primary_ei, dtype=np.int64)
    sharing = np.bincount(primary[~outfall & (primary > 0)], minlength=n_cells)
    overflow, active = 0.0, 0
    for node, name in enumerate(names):
        c = int(primary[node])
        if outfall[node] or c == 0 or land_types[c] == IMPERVIOUS_BUILDING_CLASS:
            continue
        rise = float(heads[name]) - (float(topology.rims[node]) + float(depth[c]))
        if rise > 0:
            rise = min(rise, MAX_OVERFLOW_HEAD_M)
            q = min(orifice(rise), rise * float(area[c]) / window_dt / int(sharing[c])) * OVERFLOW_RELAXATION
            source[c] += q
            node_flow[node] -= q
            overflow += q
            active += 1
    # the F32 rounding of each linked cell's source, charged back in equal shares
    source_f32 = source.astype(np.float32)
    rounding = np.zeros(n_cells)
    rounding[linked] = (source_f32[linked].astype(np.float64) - source[linked]) / links.count[linked]
    np.subtract.at(node_flow, links.node, rounding[links.cell])
    has_cell = primary > 0
    stage = np.where(has_cell, depth[primary] + bed[primary], 0.)
    boundaries = {name: ({'level': float(stage[k]), 'flow': 0.0} if outfall[k] else
                         {'level': 0.0, 'flow': float(node_flow[k])}) for k, name in enumerate(names)}
    source_f32[0] = 0.
    surface = float(source_f32.sum(dtype=np.float64))
    diagnostics = {'drainage_m3_s': drainage, 'overflow_m3_s': overflow, 'net_pipe_source_m3_s': surface,
                   'net_swmm_inflow_m3_s': float(node_flow.sum()), 'overflow_nodes': active,
                   'surface_pipe_balance_m3_s': float(surface + node_flow.sum())}
    return source_f32, boundaries, diagnostics, links


# ------------------------------------------------------------------------------ pipe models

class StorageActor:
    """Stand-in pipe model: each junction a prismatic store (plan area `area` m^2, invert at rim
    minus `depth_below_rim`) that empties into the outfalls at `release` m^3/s per metre of
    stored head; outfalls take the stage they are given. Heads are invert + volume / area."""

    def __init__(self, topology: PipeTopology, area: float = 10.0, depth_below_rim: float = 2.0,
                 release: float = 0.05):
        self.t = topology
        self.area, self.release = area, release
        self.invert = np.asarray(topology.rims, float) - depth_below_rim
        self.volume = np.zeros(len(topology.node_names))
        self.seconds = 0.0

    def heads(self) -> dict:
        h = self.invert + self.volume / self.area
        return {name: float(h[i]) for i, name in enumerate(self.t.node_names)}

    def start(self) -> dict:
        return self.heads()

    def advance(self, boundaries: dict, seconds: int):
        for i, name in enumerate(self.t.node_names):
            if self.t.is_outfall[i]:
                continue
            self.volume[i] += boundaries[name]['flow'] * seconds
            out = min(max(self.volume[i], 0.), self.release * self.volume[i] / self.area * seconds)
            self.volume[i] -= out
        self.seconds += seconds
        return self.heads(), self.seconds

    def close(self) -> None:
        pass


class SwmmActor:
    """SWMM through pyswmm (not installed in every environment): junction inflows and outfall
    stages set per window, then one stride of the window's whole seconds."""

    def __init__(self, pipe_cfg: PipeConfig, topology: PipeTopology, start, end):
        try:
            import pyswmm  # noqa: F401
        except ImportError as exc:
            raise ImportError('pipe coupling with SWMM needs pyswmm; pass another actor to run '
                              'without it') from exc
        self.cfg, self.t, self.begin, self.end = pipe_cfg, topology, start, end
        self.sim = None

    def start(self) -> dict:
        from pyswmm import Nodes, Simulation
        self.sim = Simulation(str(self.cfg.inp))
        self.sim.start_time = self.begin.replace(tzinfo=None)
        self.sim.end_time = self.end.replace(tzinfo=None)
        self.nodes = {n.nodeid: n for n in Nodes(self.sim)}
        missing = set(self.t.node_names) - self.nodes.keys()
        if missing:
            raise ValueError(f'pipe nodes missing from SWMM: {sorted(missing)}')
        self.sim.start()
        self.seconds = 0.0
        return self.heads()

    def heads(self) -> dict:
        return {name: float(self.nodes[name].head) for name in self.t.node_names}

    def advance(self, boundaries: dict, seconds: int):
        for i, name in enumerate(self.t.node_names):
            node = self.nodes[name]
            if self.t.is_outfall[i]:
                node.outfall_stage = boundaries[name]['level']
            else:
                node.generated_inflow(boundaries[name]['flow'])
        self.sim._model.swmm_stride(int(seconds))
        self.seconds += seconds
        return self.heads(), self.seconds

    def close(self) -> None:
        if self.sim is not None:
            self.sim.close()


# ------------------------------------------------------------------------------ windows

class PipeCoupling:
    """Window protocol of one coupled run (core/coupled/runtime.py): begin_window plans the
    sources from the current heads; end_window pairs the measured surface volumes with junction
    inflows, advances the pipe model and refreshes the heads."""

    def __init__(self, topology: PipeTopology, interval: int, warmup: float, duration: float, actor,
                 start_iso: str = ''):
        self.t = topology
        self.interval = int(interval)
        self.duration = float(duration)
        self.actor = actor
        self.start_iso = start_iso
        self._elapsed_next = 0.0 - float(warmup)
        self._pending = None
        self._plan = None
        self.windows = []
        self._seconds = 0.0
        self._heads = actor.start()

    def begin_window(self, elapsed: float, window_dt: float, depth, bed, area, land_types) -> np.ndarray:
        if self._pending is not None:
            raise RuntimeError('end_window must finish before begin_window')
        self.t.validate(len(depth))
        if abs(elapsed - self._elapsed_next) > 1e-5:
            raise ValueError(f'Pipe window starts at {elapsed}, expected {self._elapsed_next}')
        seconds = round(float(window_dt))
        if seconds <= 0 or not math.isclose(float(window_dt), seconds, abs_tol=1e-5):
            raise ValueError('pipe exchange windows must contain positive whole seconds')
        if elapsed + seconds > self.duration + 1e-5:
            raise ValueError('Pipe window extends past simulation end')
        if elapsed < 0 < elapsed + seconds:
            raise ValueError('Pipe window must end at zero before official simulation')
        source, boundaries, diag, links = compute_exchange(seconds, depth, bed, area, land_types, self.t, self._heads)
        self._plan = (source, boundaries, links)
        self._pending = dict(elapsed_start_s=elapsed, elapsed_end_s=elapsed + seconds, window_seconds=seconds,
                             swmm_before_s=self._seconds, **diag)
        return source

    def end_window(self, actual_volumes) -> None:
        """Pair the volume each cell actually exchanged with its junctions, then advance the pipe
        model (core/coupled/runtime.py:152-202): a cell whose volume differs from its request
        (the limiter held back drainage, or nothing was there to drain) moves its junctions'
        inflow by the difference, in equal shares; drainage given back that way is recorded."""
        if self._pending is None:
            raise RuntimeError('No open pipe exchange window')
        source, planned, links = self._plan
        actual = np.asarray(actual_volumes, dtype=np.float64)
        if actual.shape != source.shape or not np.all(np.isfinite(actual)):
            raise ValueError('actual surface volumes must be finite and match the surface mesh')
        seconds = self._pending['window_seconds']
        requested = source.astype(np.float64) * seconds
        change = actual - requested
        links_of_node = np.bincount(links.node, minlength=len(self.t.node_names)) > 0
        moved = np.flatnonzero(change)
        orphans = moved[links.count[moved] == 0]
        if orphans.size:
            raise ValueError(f'Unpaired surface/pipe exchange at cell {int(orphans[0])}')
        per_cell = np.zeros(len(change))
        per_cell[moved] = change[moved] / (seconds * links.count[moved])
        names = self.t.node_names
        flows = np.array([planned[name]['flow'] for name in names], dtype=np.float64)
        np.subtract.at(flows, links.node, per_cell[links.cell])     # junction by junction, cell order
        boundaries = {name: dict(planned[name], flow=float(flows[k]) if links_of_node[k] else planned[name]['flow'])
                      for k, name in enumerate(names)}
        given_back = moved[(requested[moved] < 0.) & (change[moved] > 0.)]
        returned = float(np.add.accumulate(np.concatenate([[0.], change[given_back]]))[-1])
        requested_pipe = sum(v['flow'] for v in planned.values()) * seconds
        applied_pipe = sum(v['flow'] for v in boundaries.values()) * seconds
        balance = float(actual.sum(dtype=np.float64) + applied_pipe)
        if abs(balance) > 1e-9 + 1e-12 * float(np.abs(actual).sum(dtype=np.float64)):
            raise ArithmeticError(f'Surface/pipe exchange mismatch {balance} m3')
        self._pending.update(requested_surface_exchange_m3=float(requested.sum(dtype=np.float64)),
                             applied_surface_exchange_m3=float(actual.sum(dtype=np.float64)),
                             requested_swmm_inflow_m3=float(requested_pipe),
                             applied_swmm_inflow_m3=float(applied_pipe),
                             limited_drainage_return_m3=returned, surface_swmm_balance_error_m3=balance)
        self._heads, total = self.actor.advance(boundaries, seconds)
        if abs(total - (self._seconds + seconds)) > 0.05:
            raise RuntimeError(f'pipe model clock {total}s differs from {self._seconds + seconds}s')
        self._seconds = total
        self._pending['swmm_after_s'] = total
        self.windows.append(self._pending)
        self.last_boundaries = boundaries
        self._elapsed_next = self._pending['elapsed_end_s']
        self._pending = None
        self._plan = None

    @property
    def pending(self) -> bool:
        return self._pending is not None

    def metadata(self) -> dict:
        return dict(swmm_cumulative_seconds=self._seconds,
                    warmup_completed_seconds=sum(w['window_seconds'] for w in self.windows if w['elapsed_end_s'] <= 0),
                    official_completed_seconds=sum(w['window_seconds'] for w in self.windows
                                                   if w['elapsed_start_s'] >= 0),
                    window_count=len(self.windows), pending_window=self.pending,
                    windows=[dict(w) for w in self.windows])

    def close(self) -> None:
        self.actor.close()
