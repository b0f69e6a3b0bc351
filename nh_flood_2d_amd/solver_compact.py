"""File-based entry points: ``run_simulation``, ``solver``, ``warmup_solver``, ``solver_coupled``.

Same signatures, return value (the final mass-balance row) and side effects as
``src/nh_flood_2d/core/solver_compact.py:152-176,1011-1016``: a manifest in ``domain_dir``, the
preprocessed mesh under ``domain_dir/preprocessed``, and under ``domain_dir/output`` the depth
frames (every 900 s and at start/end), ``bed.fdb``, ``extrema.fdb``, ``stations.csv`` and the
checkpoints ``state_start.fdb`` / ``state_final.fdb`` / ``state_latest.fdb`` (every frame once
30 min of wall time have passed, written atomically); ``mass_balance.csv`` and
``step_statistics.csv`` in ``domain_dir``. Restart identity checks follow
``solver_compact.py:698-724``. The device step always runs on the HIP library.
"""
from __future__ import annotations

import csv
import hashlib
import json
import os
import platform
from datetime import datetime, timedelta, timezone
from pathlib import Path
from time import perf_counter

import numpy as np

from . import __version__
from .clock import seconds_to_ticks, TICKS_PER_SECOND
from .config import DomainConfig, ForceConfig
from .forcing import load_forcing
from .gate import GateController, append_gate_balance, load_gate_layout
from .mesh import Mesh, filter_rows, compact_mesh
from .output import (append_station_samples, read_checkpoint, resolve_station_cells, save_tables,
                     load_tables, write_bed, write_checkpoint, write_depth_frame, write_extrema,
                     write_manifest)
from .simulation import RunSettings, Simulation, velocities

CHECKPOINT_INTERVAL_S = 30 * 60
# Engine constructor used by _solve; None selects the HIP engine. Tests substitute the oracle.
ENGINE_FACTORY = None
# pipe model used by coupled runs: (pipe_cfg, topology, begin, end) -> actor; None = SWMM (pyswmm)
PIPE_ACTOR_FACTORY = None
STATS_FIELDS = ['time', 'elapsed_seconds', 'steps', 'standard_dt_min_s', 'standard_dt_max_s',
                'standard_dt_mean_s', 'standard_steps', 'event_short_steps', 'event_dt_min_s',
                'event_dt_sum_s', 'cfl_min_s', 'limited_cell_steps', 'limited_face_steps',
                'minimum_donor_scale', 'limiting_edge', 'lower_cell', 'higher_cell', 'lower_depth_m',
                'higher_depth_m', 'face_depth_m', 'face_velocity_m_s']


def sha256(path) -> str:
    digest = hashlib.sha256()
    with open(path, 'rb') as f:
        for block in iter(lambda: f.read(1 << 20), b''):
            digest.update(block)
    return digest.hexdigest()


def source_tree_sha256(root: Path) -> str:
    digest = hashlib.sha256(b'nh-flood-2d-amd/src-sha256/v1\0')
    files = sorted(list(root.rglob('*.py')) + list(root.rglob('*.hip')), key=lambda p: p.relative_to(root).as_posix())
    for path in files:
        content = path.read_bytes()
        digest.update(path.relative_to(root).as_posix().encode() + b'\0')
        digest.update(len(content).to_bytes(8, 'big'))
        digest.update(content)
    return digest.hexdigest()


class RunRecord:
    """Provenance manifest (core/run_record.py:36-104); refuses a used output directory."""

    def __init__(self, domain: DomainConfig, force: ForceConfig, pipe=None):
        self.path = Path(domain.domain_dir) / 'manifest.json'
        output = Path(domain.domain_dir) / 'output'
        if (self.path.exists() or (self.path.parent / 'mass_balance.csv').exists()
                or (output.exists() and any(output.iterdir()))):
            raise FileExistsError(f'Use a new output directory: {domain.domain_dir}')
        self.clock = perf_counter()
        package = Path(__file__).resolve().parent
        self.data = {
            'source_sha256': source_tree_sha256(package), 'source_hash_rule': 'nh-flood-2d-amd/src-sha256/v1',
            'version': __version__,
            'config': {'domain': domain.model_dump(mode='json'), 'force': force.model_dump(mode='json'),
                       'pipe': None},
            'input_sha256': {}, 'simulation_start': domain.start.isoformat(),
            'warmup_start': (domain.start - timedelta(seconds=domain.warmup_duration)).isoformat(),
            'simulation_end': (domain.start + timedelta(seconds=domain.duration)).isoformat(),
            'started_at': datetime.now(timezone.utc).isoformat(),
            'steps': 0, 'warmup_steps': 0, 'backend': 'hip-gfx950', 'device_name': None,
            'status': 'running',
            'budget_accumulator': {
                'device_step_reduction': 'F32 per-cell terms summed in FP64 per wavefront',
                'device_time_reduction': 'FP64 compensated (Neumaier) accumulator per wavefront and term',
                'host_reduction': 'exact rational merge of double-double interval totals, then F64',
            },
        }
        self.write()

    def __enter__(self):
        return self

    def __exit__(self, error_type, error, traceback):
        if error is not None:
            self.data['error'] = str(error)
        self.finish('failed' if error_type else 'completed')

    def hash_inputs(self, domain, force):
        paths = [domain.ne, domain.ns, *force.input_paths()] + ([domain.restart_state] if domain.restart_state else [])
        for path in paths:
            self.data['input_sha256'][str(Path(path).resolve())] = sha256(path)
        self.write()

    def device(self, name: str) -> None:
        self.data['device_name'] = name

    def write(self):
        write_manifest(self.path, self.data)

    def finish(self, status):
        self.data.update(status=status, finished_at=datetime.now(timezone.utc).isoformat(),
                         wall_seconds=perf_counter() - self.clock)
        self.write()


# ------------------------------------------------------------------------------- preprocessing

def prepare_domain(cfg: DomainConfig) -> Mesh:
    """Filter and compact NE/NS text into the preprocessed tables (preprocess/domain.py:590-606)."""
    ne_lines = Path(cfg.ne).read_text(encoding='utf-8').splitlines()
    ns_lines = Path(cfg.ns).read_text(encoding='utf-8').splitlines()
    ne_rows, ns_rows = filter_rows(ne_lines, ns_lines)
    mesh = compact_mesh(ne_rows, ns_rows)
    write_preprocessed(cfg, mesh)
    return mesh


def preprocess(domain_cfg: DomainConfig, force_cfg: ForceConfig | None = None) -> None:
    prepare_domain(domain_cfg)


def write_preprocessed(cfg: DomainConfig, mesh: Mesh) -> None:
    e = np.arange(mesh.e_num, dtype=np.uint32)
    ne = [('IndexLike', 'Ne.index', e), ('F32Value', 'Ne.x', mesh.x), ('F32Value', 'Ne.y', mesh.y),
          ('F32Value', 'Ne.z', mesh.z)]
    ne += [('IndexLike', f'Ne.{d}_side_num', mesh.face_count[:, i].astype(np.uint32)) for i, d in enumerate('lrbt')]
    ne += [('U8Value', 'Ne.type', mesh.land), ('IndexLike', 'isl_data', mesh.isl.astype(np.uint32))]
    ne += [('IndexLike', f'isl_ptr_{d}', mesh.face_ptr[:, i].astype(np.uint32)) for i, d in enumerate('lrbt')]
    save_tables(cfg.ne_fdb, ne)
    s = np.arange(mesh.s_num, dtype=np.uint32)
    save_tables(cfg.ns_fdb, [('IndexLike', 'Ns.index', s), ('F32Value', 'Ns.length', mesh.length),
                             ('F32Value', 'Ns.x', mesh.sx), ('F32Value', 'Ns.y', mesh.sy),
                             ('F32Value', 'Ns.z', mesh.sz), ('U8Value', 'Ns.attr', mesh.attr),
                             ('IndexLike', 'SideTopoInfo.info', mesh.topo.astype(np.uint32).ravel())])
    save_tables(cfg.boundary_fdb, [('U8Value', 'sbf', mesh.sbf)])


def load_preprocessed(cfg: DomainConfig) -> Mesh:
    ne, ns = load_tables(cfg.ne_fdb), load_tables(cfg.ns_fdb)
    counts = np.column_stack([ne[f'Ne.{d}_side_num'] for d in 'lrbt']).astype(np.int32)
    ptr = np.column_stack([ne[f'isl_ptr_{d}'] for d in 'lrbt']).astype(np.int32)
    return Mesh(ne['Ne.x'], ne['Ne.y'], ne['Ne.z'], ne['Ne.type'], counts, ptr,
                ne['isl_data'].astype(np.int32), ns['SideTopoInfo.info'].astype(np.int32).reshape(-1, 3),
                ns['Ns.length'], ns['Ns.x'], ns['Ns.y'], ns['Ns.z'], ns['Ns.attr'])


# ------------------------------------------------------------------------------- outputs

class FileOutputs:
    """Writes every output instant of a Simulation (solver_compact.py:792-855, 986-1005)."""

    def __init__(self, cfg: DomainConfig, record: RunRecord, mesh: Mesh, forcing, identity: dict,
                 boundary_identity: list, stations: dict, initial_bed: np.ndarray):
        self.cfg, self.record, self.mesh = cfg, record, mesh
        self.root = Path(cfg.output_dir)
        self.identity, self.boundary_identity = identity, boundary_identity
        self.stations = stations
        self.budget_file = (Path(cfg.domain_dir) / 'mass_balance.csv').open('w', newline='')
        self.budget_writer = None
        self.stats_file = (Path(cfg.domain_dir) / 'step_statistics.csv').open('w', newline='')
        self.stats_writer = csv.DictWriter(self.stats_file, fieldnames=STATS_FIELDS)
        self.stats_writer.writeheader()
        self.last_checkpoint_wall = perf_counter()
        write_bed(self.root / 'bed.fdb', initial_bed)

    def close(self):
        self.budget_file.close()
        self.stats_file.close()

    def emit(self, sim: Simulation, seconds, phase, frame, final, depth, row, stats):
        if self.budget_writer is None:
            self.budget_writer = csv.DictWriter(self.budget_file, sim.budget.fields())
            self.budget_writer.writeheader()
        self.budget_writer.writerow(row)
        self.budget_file.flush()
        if sim.gates is not None:
            append_gate_balance(Path(self.cfg.domain_dir) / 'gate_balance.csv',
                                self.cfg.start + timedelta(seconds=seconds), seconds, sim.record['gate_control'])
        if self.stations:
            q = sim.engine.read_state('q')
            bed = sim.engine.read_state('bed')
            u, v = velocities(self.mesh, depth, q, sim.settings.min_h)
            samples = {name: dict(cell_id=cell, depth=float(depth[cell]),
                                  level=float(np.float32(bed[cell] + depth[cell])),
                                  u=float(u[cell]), v=float(v[cell]))
                       for name, cell in self.stations.items()}
            append_station_samples(self.root / 'stations.csv', elapsed_seconds=seconds,
                                   time=sim.time_text(seconds), samples=samples)
        if frame:
            write_depth_frame(Path(self.cfg.frame_dir) / f'depth_{int(seconds)}.fdb', depth)
        if seconds == 0. or final:
            self.checkpoint(sim, self.root / ('state_final.fdb' if final else 'state_start.fdb'), depth)
        elif frame and perf_counter() - self.last_checkpoint_wall >= CHECKPOINT_INTERVAL_S:
            self.checkpoint(sim, self.root / 'state_latest.fdb', depth)
        self.stats_writer.writerow(stats)
        self.stats_file.flush()
        print(f't={seconds:.6f}s phase={phase} storage={row["storage_m3"]:.9g} '
              f'residual={row["residual_m3"]:.9g} m3', flush=True)

    def checkpoint(self, sim: Simulation, path, depth):
        cfg = self.cfg
        units = sim.device_units
        metadata = dict(self.identity, run_state=True, boundary_identity=self.boundary_identity,
                        clock_origin_before_start_s=sim.origin_before_start,
                        physics_config={'start': cfg.start.isoformat(), 'afa': cfg.afa, 'sita': cfg.sita,
                                        'min_h': cfg.min_h,
                                        'land_parameters': cfg.model_dump(mode='json')['land_parameters']},
                        total_steps=sim.previous_steps + sim.steps,
                        total_warmup_steps=sim.previous_warmup_steps + sim.warmup_steps,
                        elapsed_seconds=sim.elapsed,
                        budget={'initial_storage_m3': sim.budget.initial, 'rainfall_m3': float(sim.budget.rain),
                                'requested_discharge_m3': sim.budget.discharge.tolist(),
                                'device_units': [int(round(t * 2 ** 149)) for t in units.total],
                                'device_terms_exact': units.state()})
        if sim.gates is not None:
            metadata['gate_controller'] = sim.gates.snapshot()
        state = sim.state()
        tick = seconds_to_ticks(sim.host_now_seconds)
        peak = state['peak_tick']
        write_checkpoint(path, depth=depth, bed=state['bed'], q=state['q'], rain_time=state['rain_time'],
                         infiltration=state['capacity'], max_depth=state['max_depth'],
                         first_time_hi=(peak >> np.uint64(32)).astype(np.uint32),
                         first_time_lo=(peak & np.uint64(0xffffffff)).astype(np.uint32),
                         max_speed=state['max_speed'], clock_hi=tick >> 32, clock_lo=tick & 0xffffffff,
                         gate_open=sim.gates is not None and sim.gates.any_commanded(),
                         metadata_json=json.dumps(metadata))
        self.last_checkpoint_wall = perf_counter()

    def finish(self, sim: Simulation):
        state = sim.state()
        peak = state['peak_tick']
        write_extrema(self.root / 'extrema.fdb', state['max_depth'], (peak >> np.uint64(32)).astype(np.uint32),
                      (peak & np.uint64(0xffffffff)).astype(np.uint32), state['max_speed'])
        depth = state['depth']
        self.record.data.update(
            gate_open=sim.gates is not None and sim.gates.any_commanded(),
            storage_depth_one_ulp_volume_m3=float(np.sum(np.abs(np.spacing(depth[1:])) * sim.area[1:],
                                                         dtype=np.float64)))


# ------------------------------------------------------------------------------- entry points

def run_simulation(domain_cfg: DomainConfig, force_cfg: ForceConfig, pipe_cfg=None):
    """Preprocess and simulate, recording failures from either stage."""
    with RunRecord(domain_cfg, force_cfg, pipe_cfg) as record:
        record.hash_inputs(domain_cfg, force_cfg)
        preprocess(domain_cfg, force_cfg)
        if pipe_cfg is not None:
            from .coupling import prepare_pipe
            prepare_pipe(pipe_cfg, load_preprocessed(domain_cfg))
        return _solve(domain_cfg, force_cfg, record, pipe_cfg)


def solver(domain_cfg: DomainConfig, force_cfg: ForceConfig, pipe_cfg=None):
    """Simulate an already preprocessed domain, including configured warmup."""
    with RunRecord(domain_cfg, force_cfg, pipe_cfg) as record:
        record.hash_inputs(domain_cfg, force_cfg)
        return _solve(domain_cfg, force_cfg, record, pipe_cfg)


def warmup_solver(domain_cfg: DomainConfig, force_cfg: ForceConfig, pipe_cfg=None):
    """Run an already preprocessed domain through warmup, stopping at start."""
    with RunRecord(domain_cfg, force_cfg, pipe_cfg) as record:
        record.data.update(run_kind='warmup', simulation_end=domain_cfg.start.isoformat())
        record.hash_inputs(domain_cfg, force_cfg)
        return _solve(domain_cfg, force_cfg, record, pipe_cfg, end_time=0.)


def solver_coupled(domain_cfg: DomainConfig, force_cfg: ForceConfig, pipe_cfg=None):
    """Run an already preprocessed domain with optional pipe coupling.

    The reference keeps this name as an alias of ``solver`` (core/coupled/driver.py:12-20): the
    pipe exchange is part of the one evolution loop, selected by ``pipe_cfg``.
    """
    return solver(domain_cfg, force_cfg, pipe_cfg)


def _device_name(factory) -> str:
    if factory is not None:
        return getattr(factory, 'device_name', 'injected engine')
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return platform.machine()


def _solve(domain_cfg: DomainConfig, force_cfg: ForceConfig, record: RunRecord, pipe_cfg=None,
           end_time=None):
    end_time = float(domain_cfg.duration if end_time is None else end_time)
    factory = ENGINE_FACTORY
    record.device(_device_name(factory))
    mesh = load_preprocessed(domain_cfg)
    forcing = load_forcing(domain_cfg, force_cfg, mesh)
    state = read_checkpoint(domain_cfg.restart_state) if domain_cfg.restart_state else None
    meta = state.meta if state else {}
    continuation = bool(meta.get('run_state', False))
    identity = {'ne_sha256': record.data['input_sha256'][str(Path(domain_cfg.ne).resolve())],
                'ns_sha256': record.data['input_sha256'][str(Path(domain_cfg.ns).resolve())],
                'gate_config': force_cfg.gate.model_dump_json() if force_cfg.gate else 'null'}
    boundary_identity = [dict(name=b.name, kind=b.kind, selected_edges_sha256=hashlib.sha256(
        np.sort(b.side_ids).astype('<i8').tobytes()).hexdigest()) for b in forcing.boundaries]
    if state:
        if len(state.depth) != mesh.e_num or len(state.q) != mesh.s_num:
            raise ValueError('Checkpoint size differs from mesh')
        for key in identity:
            if meta[key] != identity[key]:
                raise ValueError(f'Checkpoint {key} differs from the current mesh/configuration')
        if continuation:
            if meta.get('boundary_identity') != boundary_identity:
                raise ValueError('Checkpoint boundary identity or order differs from configuration')
            expected = {'start': domain_cfg.start.isoformat(), 'sita': domain_cfg.sita, 'min_h': domain_cfg.min_h,
                        'land_parameters': domain_cfg.model_dump(mode='json')['land_parameters']}
            physics = {k: v for k, v in meta['physics_config'].items() if k != 'afa'}
            if physics != expected:
                raise ValueError('Checkpoint clock or physical parameters differ from configuration')
            if meta.get('budget', {}).get('device_terms_exact') is None:
                raise ValueError('Checkpoint budget was not written by this solver')
    origin_before_start = meta['clock_origin_before_start_s'] if continuation else domain_cfg.warmup_duration
    if state is not None and ((state.clock_hi << 32) | state.clock_lo) > seconds_to_ticks(end_time + origin_before_start):
        raise ValueError('Checkpoint time is later than the requested simulation end')
    gate_layout = (load_gate_layout(force_cfg.gate, mesh, forcing.boundaries, mesh.z)
                   if force_cfg.gate is not None else None)   # checked even when disabled
    record.data['gate_cell_count'] = int(np.sum(gate_layout.cell_gate >= 0)) if gate_layout else 0
    gates = (GateController(force_cfg.gate, gate_layout, origin_before_start)
             if gate_layout is not None and force_cfg.gate.enabled else None)
    if gates is not None and pipe_cfg is not None:
        raise ValueError('gate control together with pipe coupling is not supported')
    stations = resolve_station_cells(domain_cfg.stations, mesh)
    settings = RunSettings(afa=domain_cfg.afa, sita=domain_cfg.sita, min_h=domain_cfg.min_h, end_time=end_time,
                           warmup_duration=float(domain_cfg.warmup_duration), yield_step=float(domain_cfg.yield_step),
                           start_text=domain_cfg.start.isoformat(), stations=stations)
    record.data['output'] = {
        'root': 'output', 'frame_interval_s': 900, 'station_budget_interval_s': domain_cfg.yield_step,
        'frames': 'F32 depth only; bed.fdb reconstructs water level',
        'extrema_time_origin': (domain_cfg.start - timedelta(seconds=origin_before_start)).isoformat(),
        'clock_quantum_s': 2. ** -40, 'extrema_window': 'simulation' if end_time > 0 else 'warmup'}
    record.data['station_cells'] = stations
    initial_bed = state.bed if state is not None else mesh.z
    outputs = FileOutputs(domain_cfg, record, mesh, forcing, identity, boundary_identity, stations,
                          np.asarray(initial_bed, np.float32))
    started = perf_counter()
    coupling = None
    try:
        if pipe_cfg is not None:
            if domain_cfg.start.microsecond:
                raise ValueError('pipe coupling requires domain.start at a whole second')
            from .coupling import PipeCoupling, PipeTopology
            topology = PipeTopology.load(pipe_cfg.topology_file)
            begin = domain_cfg.start - timedelta(seconds=domain_cfg.warmup_duration)
            end = domain_cfg.start + timedelta(seconds=domain_cfg.duration)
            actor = (PIPE_ACTOR_FACTORY or _swmm_actor)(pipe_cfg, topology, begin, end)
            coupling = PipeCoupling(topology, pipe_cfg.coupling_interval, float(domain_cfg.warmup_duration),
                                    float(domain_cfg.duration), actor, domain_cfg.start.isoformat())
        sim = Simulation(mesh, domain_cfg.land_parameters, forcing, settings, checkpoint=state,
                         engine_factory=factory, outputs=outputs, record=record.data, coupling=coupling,
                         gates=gates)
        record.data['budget_accumulator']['interval_batches'] = settings.batch_size
        setup = perf_counter() - started
        result = sim.run()
        wall = perf_counter() - started - setup
        formal = sim.steps - sim.warmup_steps
        record.data['timing_seconds'] = dict(preprocess_and_setup=perf_counter() - record.clock - wall,
                                             device_batches=sim.device_seconds, loop=wall,
                                             simulation_per_step=wall / sim.steps if sim.steps else None,
                                             formal_steps=formal)
        record.data['pressure_gradient'] = dict(method='tangential_subedge_reconstruction',
                                                hanging_edges=sim.layout.n_hang, gradient_cells=sim.layout.n_grad)
        return result
    finally:
        if coupling is not None:
            coupling.close()
        outputs.close()
        record.write()


def _swmm_actor(pipe_cfg, topology, begin, end):
    from .coupling import SwmmActor
    return SwmmActor(pipe_cfg, topology, begin, end)
