"""Gate control: an irreversible per-gate command, moving gate-cell beds, exchange budgets.

The reference runs all of it inside its step graph (``core/gate.py:35-268``, layout
``preprocess/gate.py:8-96``). Here the device does only what has to happen inside a step: the
trigger test at each step start (one thread per gate), the bed of the gate cells at step start
and end, and each step's post-limiter exchange volumes and cell sources summed per term
(``nhf_set_gates``). Everything else is host code in this module: the layout and its checks, the
lists the device walks, the snapshot a checkpoint carries, the exact budget totals, the report
and ``gate_balance.csv``. The device state (flags and clock ticks per gate) lives in
caller-owned arrays the host reads and restores directly.
"""
from __future__ import annotations

import csv
from dataclasses import dataclass
from fractions import Fraction
from pathlib import Path

import numpy as np

from .clock import seconds_to_ticks

CHUNK = 256     # accounting items per device workgroup
SIDE_GROUP, CELL_GROUP = 0, 1
NEGATIVE = np.int64(1) << 31   # side item flag: the recorded volume is the side's volume negated

BALANCE_FIELDS = ['time', 'elapsed_seconds', 'region', 'pond_in_m3', 'pond_out_m3', 'pond_net_m3', 'river_in_m3',
                  'river_out_m3', 'channel_storage_m3', 'initial_storage_m3', 'storage_m3', 'overtopping_in_m3',
                  'overtopping_out_m3', 'rain_m3', 'infiltration_m3', 'positivity_m3', 'pipe_m3', 'boundary_m3',
                  'rounding_m3', 'residual_m3', 'river_net_m3']


@dataclass
class GateLayout:
    """Gate components in the reference's cell numbering (GateLayout, preprocess/gate.py:8-13)."""
    cell_gate: np.ndarray        # [e_num] gate index or -1
    controls: np.ndarray         # [count] control cell
    cells: list                  # per gate: sorted gate cells
    pond: np.ndarray             # [e_num] bool


def _components(selected: np.ndarray, topo: np.ndarray) -> np.ndarray:
    """Connected components of the selected cells over shared sides; numbered in order of their
    smallest cell (-1 for cells not selected)."""
    n = len(selected)
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for left, right in topo[1:, 1:3]:
        if selected[left] and selected[right]:
            ra, rb = find(int(left)), find(int(right))
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    component = np.full(n, -1, dtype=np.int32)
    roots = {}
    for cell in np.flatnonzero(selected):
        root = find(int(cell))
        if root not in roots:
            roots[root] = len(roots)
        component[cell] = roots[root]
    return component


def load_gate_layout(config, mesh, boundaries, bed: np.ndarray) -> GateLayout:
    """Resolve and check the gate components (GateLayout.load, preprocess/gate.py:15-96)."""
    land = np.asarray(mesh.land)
    selected = land == config.land_type
    selected[0] = False
    pond = land == config.pond_land_type
    pond[0] = False
    if np.any(selected & pond):
        raise ValueError('Gate and pond cells must not overlap')
    if not np.any(pond):
        raise ValueError('pond_land_type selects no pond cells')
    topo = np.asarray(mesh.topo)
    component = _components(selected, topo)
    groups = [np.flatnonzero(component == k).astype(np.int32) for k in range(component.max() + 1)]
    bounds = mesh.geometry()[0]

    def locate(point, label):
        x, y = point
        hits = np.flatnonzero((bounds[:, 0] <= x) & (x <= bounds[:, 2]) & (bounds[:, 1] <= y) & (y <= bounds[:, 3]))
        hits = hits[hits != 0]
        if not len(hits):
            raise ValueError(f'{label}: point lies outside the mesh')
        return int(hits[0])        # the lowest cell id, as station sampling does

    interior = topo[1:][(topo[1:, 1] != 0) & (topo[1:, 2] != 0)]
    cell_gate = np.full(len(land), -1, dtype=np.int32)
    controls, cells, claimed = [], [], set()
    for index, gate in enumerate(config.gates):
        inside = locate(gate.inside_point, f'{gate.name}.inside_point')
        group = int(component[inside])
        if group < 0:
            raise ValueError(f'{gate.name}: inside_point is not on a gate cell')
        if group in claimed:
            raise ValueError('Each gate component must correspond to exactly one gate')
        claimed.add(group)
        control = locate(gate.control_point, f'{gate.name}.control_point')
        if selected[control]:
            raise ValueError(f'{gate.name}: control_point is on a gate cell')
        # compared as the solver stores beds, in F32
        if not np.all(np.float32(config.open_elevation) <= np.asarray(bed, np.float32)[groups[group]]):
            raise ValueError(f'{gate.name}: open_elevation exceeds a closed gate bed')
        lin = component[interior[:, 1]] == group
        rin = component[interior[:, 2]] == group
        across = np.where(lin & ~rin, interior[:, 2], np.where(rin & ~lin, interior[:, 1], -1))
        across = across[across >= 0]
        if not (np.any(pond[across]) and np.any(~pond[across] & ~selected[across])):
            raise ValueError(f'{gate.name}: gate needs internal edges to pond and non-pond non-gate cells')
        controls.append(control)
        cells.append(groups[group])
        cell_gate[groups[group]] = index
    if len(claimed) != len(groups):
        raise ValueError('Each gate component must correspond to exactly one gate')
    for b in boundaries:
        if b.kind == 'level' and np.any(selected[np.asarray(b.cells)]):
            raise ValueError('Gate cells must not overlap a level boundary')
    return GateLayout(cell_gate, np.array(controls, dtype=np.int32), cells, pond)


# ------------------------------------------------------------------------------ budget terms

def term_count(count: int) -> int:
    """Per gate: pond in/out, river in/out; then overtopping in/out and the pond's 6 source terms;
    then 6 source terms per gate (GateController.__init__, core/gate.py:56-59)."""
    return 10 * count + 8


def accounting_items(layout: GateLayout, device_layout, count: int):
    """Device items of the per-step accounting, in groups of one term base.

    Sides (record_flux, core/gate.py:119-139), in the library's side order: a gate cell beside a
    non-gate cell records the volume into the gate (pond side: terms 4g/4g+1, river side: 4g+2/
    4g+3, directed); a side between pond and non-pond cells records overtopping into the pond
    (terms 4G/4G+1). Cells (record_cell, :149-163): pond cells (terms 4G+2..4G+7) and each gate's
    cells (4G+8+6g..). Returns (items int32, chunks int32 [n, 4]: first item, items, kind | (first-
    flow gate + 1) << 8, first term)."""
    e_num = device_layout.e_num
    gate_dev = np.full(e_num, -1, dtype=np.int64)
    gate_dev[:] = layout.cell_gate[device_layout.cell_perm]
    pond_dev = layout.pond[device_layout.cell_perm].copy()
    gate_dev[0], pond_dev[0] = -1, False
    sc = device_layout.arrays['side_cells'].reshape(-1, 2).astype(np.int64)
    gl, gr = gate_dev[sc[:, 0]], gate_dev[sc[:, 1]]
    pl, pr = pond_dev[sc[:, 0]], pond_dev[sc[:, 1]]
    side = np.arange(len(sc))
    groups = {}

    def add(term, flow_gate, ids, negative):
        if len(ids):
            key = (SIDE_GROUP, term, flow_gate)
            groups.setdefault(key, []).append(ids.astype(np.int64) | np.where(negative, NEGATIVE, 0))

    left_gate = (gl >= 0) & (gr < 0)
    right_gate = (gr >= 0) & (gl < 0)
    over = ~left_gate & ~right_gate & (pl != pr)
    valid = side >= 1
    for g in range(count):
        m = valid & left_gate & (gl == g)
        add(4 * g, g, side[m & pr], np.zeros(int(np.sum(m & pr)), bool))
        add(4 * g + 2, -1, side[m & ~pr], np.ones(int(np.sum(m & ~pr)), bool))
        m = valid & right_gate & (gr == g)
        add(4 * g, g, side[m & pl], np.ones(int(np.sum(m & pl)), bool))
        add(4 * g + 2, -1, side[m & ~pl], np.zeros(int(np.sum(m & ~pl)), bool))
    m = valid & over
    add(4 * count, -1, side[m], ~pr[m])
    cells = np.arange(e_num)
    if np.any(pond_dev):
        groups[(CELL_GROUP, 4 * count + 2, -1)] = [cells[pond_dev]]
    for g in range(count):
        ids = cells[(gate_dev == g) & ~pond_dev]
        if len(ids):
            groups[(CELL_GROUP, 4 * count + 8 + 6 * g, -1)] = [ids]
    items, chunks, placed = [], [], 0
    for (kind, term, flow_gate) in sorted(groups):
        ids = np.concatenate(groups[(kind, term, flow_gate)])
        ids = ids[np.argsort(ids & ~NEGATIVE, kind='stable')]
        for start in range(0, len(ids), CHUNK):
            part = ids[start:start + CHUNK]
            chunks.append((placed, len(part), kind | ((flow_gate + 1) << 8), term))
            items.append(part)
            placed += len(part)
    flat = (np.concatenate(items) if items else np.zeros(0, np.int64)).astype(np.int64)
    return flat.astype(np.uint32).view(np.int32), np.array(chunks, dtype=np.int32).reshape(-1, 4)


def pair_totals(pairs: np.ndarray):
    """Exact per-term values of one flush (double-double pairs [terms, 2])."""
    return [Fraction(float(hi)) + Fraction(float(lo)) for hi, lo in np.asarray(pairs).reshape(-1, 2)]


# ------------------------------------------------------------------------------ controller

class GateController:
    """Host side of gate control for one solve; the engine holds the device state."""

    def __init__(self, config, layout: GateLayout, origin_seconds: float):
        self.config, self.layout = config, layout
        self.count = len(config.gates)
        self.formal_tick = seconds_to_ticks(origin_seconds)
        self.totals = [Fraction(0)] * term_count(self.count)
        self.initial_storage = None
        self.initial_channel_storage = None
        self.engine = None

    def bind(self, engine, device_layout, bed_reference: np.ndarray, min_h: float) -> None:
        """Give the engine the gate lists (device order) and fresh device state."""
        items, chunks = accounting_items(self.layout, device_layout, self.count)
        perm = device_layout.cell_perm
        cells = np.sort(device_layout.cell_inv[np.concatenate(self.layout.cells)]).astype(np.int32)
        spec = dict(count=self.count, trigger_level=self.config.trigger_level,
                    opening_seconds=self.config.opening_seconds, open_elevation=self.config.open_elevation,
                    min_h=min_h, formal_tick=self.formal_tick,
                    control=device_layout.cell_inv[self.layout.controls].astype(np.int32),
                    cells=cells,
                    cell_gate=self.layout.cell_gate[perm].astype(np.int32),
                    pond=self.layout.pond[perm].astype(np.uint8),
                    closed=np.asarray(bed_reference, np.float32)[perm],
                    items=items, chunks=chunks)
        engine.set_gates(spec)
        self.engine = engine

    # device state: flags [count, 3] (commanded, flow seen, opened) and ticks [count, 3]
    def flags(self):
        return self.engine.gate_state()[0]

    def ticks(self):
        return self.engine.gate_state()[1]

    def any_commanded(self) -> bool:
        return bool(np.any(self.flags()[:, 0]))

    def flush(self) -> np.ndarray:
        """Add the device's per-chunk sums since the last flush to the exact totals."""
        for t, v in enumerate(pair_totals(self.engine.flush_gate_budget())):
            self.totals[t] += v
        return np.array([float(t) for t in self.totals])

    def initialize_inventory(self, depth: np.ndarray, area: np.ndarray) -> None:
        pond = self.layout.pond
        self.initial_storage = float(np.sum(depth[pond].astype(float) * area[pond]))
        self.initial_channel_storage = np.array([np.sum(depth[c].astype(float) * area[c]) for c in self.layout.cells])

    def snapshot(self) -> dict:
        flags, ticks = self.flags(), self.ticks()

        def words(t):
            return [[int(v) >> 32, int(v) & 0xffffffff] for v in t]
        return dict(commanded=flags[:, 0].tolist(), command_ticks=words(ticks[:, 0]),
                    flow_seen=flags[:, 1].tolist(), first_flow_ticks=words(ticks[:, 1]),
                    opened=flags[:, 2].tolist(), opened_ticks=words(ticks[:, 2]),
                    initial_storage_m3=self.initial_storage,
                    initial_channel_storage_m3=self.initial_channel_storage.tolist(),
                    budget_units=[int(round(t * 2 ** 149)) for t in self.totals],
                    budget_terms_exact=[[t.numerator, t.denominator] for t in self.totals])

    def restore(self, saved: dict) -> None:
        flags = np.stack([saved['commanded'], saved['flow_seen'], saved['opened']], axis=1).astype(np.int32)

        def ticks(words):
            return [(int(hi) << 32) | int(lo) for hi, lo in words]
        t = np.stack([ticks(saved['command_ticks']), ticks(saved['first_flow_ticks']),
                      ticks(saved['opened_ticks'])], axis=1).astype(np.uint64)
        self.engine.set_gate_state(flags, t)
        self.initial_storage = saved['initial_storage_m3']
        self.initial_channel_storage = np.array(saved['initial_channel_storage_m3'], dtype=float)
        if 'budget_terms_exact' in saved:
            self.totals = [Fraction(int(n), int(d)) for n, d in saved['budget_terms_exact']]
        else:
            self.totals = [Fraction(int(u), 2 ** 149) for u in saved['budget_units']]

    def report(self, depth: np.ndarray, area: np.ndarray, clock_seconds: float, closed_bed: np.ndarray) -> dict:
        """GateController.report (core/gate.py:221-268): per-gate times and exchanges, per-gate
        channel budget and the pond control-volume budget."""
        totals = self.flush()
        flags, ticks = self.flags(), self.ticks()
        formal = self.formal_tick / 2 ** 40
        opening = self.config.opening_seconds

        def seconds(tick):
            return int(tick) / 2 ** 40 - formal
        gates = []
        for i, gate in enumerate(self.config.gates):
            command = seconds(ticks[i, 0]) if flags[i, 0] else None
            end = command + opening if command is not None else None
            cells = self.layout.cells[i]
            storage = float(np.sum(depth[cells].astype(float) * area[cells]))
            s = totals[4 * self.count + 8 + 6 * i:4 * self.count + 14 + 6 * i]
            river_net = float(totals[4 * i + 2] - totals[4 * i + 3])
            pond_net = float(totals[4 * i] - totals[4 * i + 1])
            channel_net = river_net - pond_net + s[0] - s[1] + s[2] + s[3] + s[4] + s[5]
            channel = dict(initial_storage_m3=float(self.initial_channel_storage[i]), storage_m3=storage,
                           river_net_m3=river_net, pond_net_m3=pond_net, rain_m3=float(s[0]),
                           infiltration_m3=float(s[1]), positivity_m3=float(s[2]), pipe_m3=float(s[3]),
                           boundary_m3=float(s[4]), rounding_m3=float(s[5]),
                           residual_m3=float(storage - self.initial_channel_storage[i] - channel_net))
            tick_words = [int(ticks[i, 0]) >> 32, int(ticks[i, 0]) & 0xffffffff]
            gates.append(dict(
                name=gate.name, cell_ids=cells.tolist(), closed_elevation=np.asarray(closed_bed)[cells].tolist(),
                command_clock_ticks=tick_words if flags[i, 0] else None, control_cell=int(self.layout.controls[i]),
                command_seconds=command,
                first_effective_flow_seconds=seconds(ticks[i, 1]) if flags[i, 1] else None,
                fully_open_seconds=end if end is not None and clock_seconds >= end else None,
                first_fully_open_sample_seconds=seconds(ticks[i, 2]) if flags[i, 2] else None,
                scheduled_fully_open_seconds=end,
                pond_in_m3=float(totals[4 * i]), pond_out_m3=float(totals[4 * i + 1]), pond_net_m3=pond_net,
                river_in_m3=float(totals[4 * i + 2]), river_out_m3=float(totals[4 * i + 3]),
                channel_storage_m3=storage, channel_budget=channel))
        t = totals[4 * self.count:4 * self.count + 8]
        pond = self.layout.pond
        storage = float(np.sum(depth[pond].astype(float) * area[pond]))
        net = sum(g['pond_net_m3'] for g in gates) + t[0] - t[1] + t[2] - t[3] + t[4] + t[5] + t[6] + t[7]
        return dict(
            enabled=True, trigger_level=self.config.trigger_level, pond_land_type=self.config.pond_land_type,
            opening_seconds=opening, open_elevation=self.config.open_elevation,
            clock_origin_before_start_s=formal, clock_quantum_s=2. ** -40,
            first_effective_flow_definition='End of first completed step with nonzero post-limiter gate-to-pond '
                                            'subedge volume in either direction',
            gates=gates, pond_budget=dict(
                initial_storage_m3=self.initial_storage, storage_m3=storage,
                overtopping_in_m3=float(t[0]), overtopping_out_m3=float(t[1]), rain_m3=float(t[2]),
                infiltration_m3=float(t[3]), positivity_m3=float(t[4]), pipe_m3=float(t[5]),
                boundary_m3=float(t[6]), rounding_m3=float(t[7]),
                residual_m3=float(storage - self.initial_storage - net)))


def append_gate_balance(path, date, seconds, report) -> None:
    """One row per gate and one for the pond at an output time (append_gate_balance,
    core/gate.py:12-32)."""
    path = Path(path)
    header = not path.exists()
    with path.open('a', newline='') as stream:
        writer = csv.DictWriter(stream, fieldnames=BALANCE_FIELDS)
        if header:
            writer.writeheader()
        for gate in report['gates']:
            row = dict(time=date.isoformat(), elapsed_seconds=seconds, region=gate['name'])
            row.update({k: gate[k] for k in BALANCE_FIELDS[3:9]})
            row.update({k: v for k, v in gate['channel_budget'].items() if k != 'pond_net_m3'})
            writer.writerow(row)
        writer.writerow(dict(time=date.isoformat(), elapsed_seconds=seconds, region='pond_control_volume',
                             **report['pond_budget']))
