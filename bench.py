"""Benchmark: cell-updates/s of the HIP local-inertial step on one MI355X.

Workload: the single large multi-resolution mesh (about 1.58 M cells over 2/4/8 m levels around a
seeded channel network, about 3.17 M sides, 2:1 balanced, diamond-square DEM amplitude 40 m and
roughness 0.6, four land classes with Horton infiltration, triangular hyetograph peaking at
80 mm/h, tide boundary 1.0 m sin(2 pi t / 44640 s)). The timed state is the configuration's own
storm: the solver runs it on the GPU from the dry start to t = --develop-s (3600 s, the rain
peak) with its full host protocol, then `--warmup` steps are run and `--steps` steps are timed
from there. A step is one pass of the device time step over the whole mesh, clock and CFL
included; batches of 32 steps run as one captured hipGraph with one trace readback each, exactly
the host protocol of the solver. Mesh, state and forcing are resident in HBM before timing.

Prints one JSON line (rank 0, N = 1 only):
  value        whole-job cell-updates/s = cells x steps / wall seconds of the timed region
  roofline     the dominant phase's algorithmic bytes / its mean duration (HIP events), with
               its measured HBM bytes (rocprofv3 counters) as `traffic`
  cpu_baseline the CPU oracle (C, OpenMP) on the host cores, on a bounded sample of steps, and on
               one thread
  batched      64 scenarios on one shared 1 M-cell mesh (nhb_*), against one of them alone
  batched_small the same on the 256 x 256 grid of the CPU-runnable configuration
  uneven       32 meshes of 50 k-4 M cells in one union (nhu_*), against one mesh at a time
  served       Poisson arrivals of 100 k-8 M-cell solves, p50/p99 latency, against one at a time
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = 'cell-updates/s + %HBM roofline, batched/served vs plain; p50/p99 solve latency'
HBM_PEAK_GBS = 8000.0
KERNELS = ['prepare', 'gradients', 'face_avg', 'flux', 'donor', 'depth', 'cfl', 'finish']


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def build_workload(seed: int):
    from nh_flood_2d_amd import synth
    sc = synth.multires_channel(seed)
    return sc, synth.scenario_forcing(sc)


def develop_state(sc, device, until: float, batch_size: int = 32) -> dict:
    """The scenario's own forcing run by the solver (Simulation: forcing events, budgets, extrema)
    on the HIP engine, graph-captured batches, from its initial state to `until` seconds. Returns
    every state array in the mesh's numbering, with the step count and wall time."""
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.engine import Engine

    def factory(layout, params, state, **kw):
        eng = Engine(layout, params, state, device=device, **kw)
        eng.set_graph(True)
        return eng
    w0 = time.perf_counter()
    sim = synth.simulation(sc, engine_factory=factory, end_time=until, yield_step=until, batch_size=batch_size)
    sim.run()
    state = sim.state()
    info = dict(until_s=until, steps=int(sim.steps), wall_s=time.perf_counter() - w0,
                storage_m3=float(sim.budget.last_row['storage_m3']),
                rainfall_m3=float(sim.budget.last_row['rainfall_m3']),
                infiltration_m3=float(sim.budget.last_row['infiltration_m3']),
                dt_sequence_sha256=sim.record.get('dt_sequence_sha256'))
    sim.engine.close()
    return state, info


def wet_fractions(mesh, state, min_h: float) -> dict:
    """Cells at or above min_h (the wet test of the velocity), and interior faces whose flow depth
    max(eta_L, eta_H) - max(z_L, z_H) is at least min_h (the flux pass's wet test)."""
    depth, bed = state['depth'], state['bed']
    lo, hi = mesh.topo[1:, 1], mesh.topo[1:, 2]
    interior = (lo > 0) & (hi > 0)
    lo, hi = lo[interior], hi[interior]
    eta_l, eta_h = bed[lo] + depth[lo], bed[hi] + depth[hi]
    wh = np.maximum(eta_l, eta_h) - np.maximum(bed[lo], bed[hi])
    return dict(wet_cell_frac=float(np.mean(depth[1:] >= np.float32(min_h))),
                wet_face_frac=float(np.mean(wh >= np.float32(min_h))))


def algorithmic_bytes(layout, e_num: int, s_num: int, r_len: int) -> dict:
    """Algorithmic bytes per step and per phase: every per-step input read once, every output
    written once, with these per-unit figures:
      per side       q r+w 8, lower/higher 8, length 4, centre distance 4, Manning 4, crest 4,
                     orient/flag 1                                                   = 33 B
      per cell       depth r+w 8, bed 4, area 4, Horton 12, rain time r+w 8, capacity r+w 8,
                     rain index 4, boundary id 4, side-list header 8, donor w+r 8, max depth
                     r+w 8, max speed r+w 8                                           = 84 B
      per side-list entry                                                              4 B
      2:1 meshes     edge index 4 per side, 24 per hanging edge, 24 per gradient cell
    Phase (a), the face flux, is 33 Ns + 8 Nc (+ the 2:1 terms); phase (b), the cell update, is
    the rest, 76 Nc + 4 R; the CFL reads only inputs already counted."""
    nc, ns, r = e_num - 1, s_num - 1, r_len
    hang = (4 * ns + 24 * layout.n_hang + 24 * layout.n_grad) if layout.n_hang else 0
    a = 33 * ns + 8 * nc + hang
    b = 76 * nc + 4 * r
    return {'step': a + b, 'flux': a, 'cell_update': b}


def launch_bytes(layout, e_num: int, s_num: int, r_len: int, uniform_rain: bool = False) -> dict:
    """Bytes each launch itself must move in this layout (its inputs read once, its outputs
    written once, conditional extrema writes excluded): Nc cells, Ns sides, R side-list entries,
    E hanging edges, G gradient cells. In the uniform rain state (one rain pixel, one rain time)
    the depth pass holds the rain time and capacities per land class instead of per cell."""
    nc, ns, r = e_num - 1, s_num - 1, r_len
    e_hang, g = layout.n_hang, layout.n_grad
    pixel = 4 if layout.n_pixels > 1 else 0
    return {
        # per side: side_cells 8, meta 1, q r+w 8, sdc 4, manning 4 (+ edge_index 4); per cell:
        # bed + depth 8; per hanging edge: offsets + slots 16; per gradient cell 8
        'flux': ns * (25 + (4 if e_hang else 0)) + 8 * nc + 16 * e_hang + 8 * g,
        # the donor certificate leaves the donor pass one flag byte per cell to read; the cells
        # it looks at (flagged, boundary) and their rewrites are excluded
        'donor': nc,
        # per cell: meta 4, side record 16, depth r+w 8, max speed 4, max depth 4 (+ rain pixel),
        # capacity r+w 8 and rain time r+w 8 unless the uniform rain state holds them; per side:
        # q 4 (cell sizes come from the meta word)
        'depth': (36 + (0 if uniform_rain else 16) + pixel) * nc + 4 * ns,
        # per side: side_cells 8, length 4, q 4 (reference ids only on ties); per cell: bed +
        # depth 8
        'cfl': 16 * ns + 8 * nc,
        # per gradient cell: its face neighbours' bed, depth, centre (about 5 x 12 B)
        'gradients': g * 60,
        'face_avg': 0, 'prepare': 0, 'finish': 0,
    }


TRAFFIC_SUMMARY = ROOT / 'profiles' / 'hbm_traffic_multires_channel.txt'
PMC_KERNELS = {'k_gradients': 'gradients', 'k_face_avg': 'face_avg', 'k_flux': 'flux', 'k_donor': 'donor',
               'k_depth': 'depth', 'k_cfl': 'cfl', 'k_finish': 'finish',
               'kb_gradients': 'batch_gradients', 'kb_face_avg': 'batch_face_avg', 'kb_flux': 'batch_flux',
               'kb_donor': 'batch_donor', 'kb_donor_wide': 'batch_donor', 'kb_depth': 'batch_depth',
               'kb_cfl': 'batch_cfl', 'kb_finish': 'batch_finish'}


def kernel_base_name(name: str) -> str:
    """'void kb_flux<true>(BDev)' -> 'kb_flux'."""
    m = re.match(r'\s*(?:void\s+)?([A-Za-z_]\w*)', name)
    return m.group(1) if m else name


def pmc_summary(fetch_csv: str, write_csv: str, launches: dict | None = None) -> str:
    """Per-kernel HBM bytes per dispatch from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE in
    KiB), scaled by what each counter reads for the 256 MiB device copy of --calibrate-copy.
    `launches` (summary name -> count) keeps only each kernel's last that many dispatches of its
    section: the per-kernel timing pass, after the storm that develops the state.
    Lines: 'calibration <counter> ... factor F' and 'kernel <name> ... hbm_bytes B'."""
    import csv
    from collections import defaultdict
    lines, scaled = [], {}
    copy_bytes = float(64 << 20) * 4
    for counter, path in (('FETCH_SIZE', fetch_csv), ('WRITE_SIZE', write_csv)):
        per_kernel, copies = defaultdict(list), []
        rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r['Dispatch_Id']))
        # the single-mesh section runs before any batch kernel: later k_* dispatches belong to
        # other sections (a plain engine beside a batch, the parts of a union)
        first_batch = min((int(r['Dispatch_Id']) for r in rows if kernel_base_name(r['Kernel_Name']).startswith('kb_')),
                          default=None)
        for row in rows:
            name = kernel_base_name(row['Kernel_Name'])
            if (name.startswith('k_') and first_batch is not None and int(row['Dispatch_Id']) > first_batch):
                continue
            value = float(row['Counter_Value']) * 1024.0
            if 'copyBuffer' in name and int(row['Grid_Size']) >= 131072:
                copies.append(value)
            elif name in PMC_KERNELS:
                per_kernel[PMC_KERNELS[name]].append(value)
        if not copies:
            raise SystemExit(f'{path}: no calibration copy found (run bench.py with --calibrate-copy)')
        if launches:
            for k in per_kernel:
                n = launches.get(k)
                if n:
                    per_kernel[k] = per_kernel[k][-n:]
        factor = copy_bytes / (sum(copies) / len(copies))
        lines.append(f'calibration {counter} copy_bytes {copy_bytes:.0f} counter_bytes '
                     f'{sum(copies) / len(copies):.0f} factor {factor:.4f}')
        for k, v in per_kernel.items():
            scaled.setdefault(k, {})[counter] = (sum(v) / len(v) * factor, len(v))
    for k in sorted(scaled, key=list(PMC_KERNELS.values()).index):
        f, nf = scaled[k].get('FETCH_SIZE', (0.0, 0))
        w, _ = scaled[k].get('WRITE_SIZE', (0.0, 0))
        lines.append(f'kernel {k} dispatches {nf} fetch_bytes {f:.0f} write_bytes {w:.0f} hbm_bytes {f + w:.0f}')
    return '\n'.join(lines) + '\n'


def l2_summary(counter_csv: str, launches: dict | None = None) -> str:
    """Per-kernel L2 hits and misses per dispatch from a rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum
    run of this script, over the same dispatches as pmc_summary (single-mesh kernels before the
    first batch kernel; each kernel's last `launches` dispatches: the per-kernel timing pass)."""
    import csv
    from collections import defaultdict
    rows = sorted(csv.DictReader(open(counter_csv)), key=lambda r: int(r['Dispatch_Id']))
    first_batch = min((int(r['Dispatch_Id']) for r in rows if kernel_base_name(r['Kernel_Name']).startswith('kb_')),
                      default=None)
    per = defaultdict(lambda: defaultdict(dict))
    for row in rows:
        name, i = kernel_base_name(row['Kernel_Name']), int(row['Dispatch_Id'])
        if name.startswith('k_') and first_batch is not None and i > first_batch:
            continue
        if name in PMC_KERNELS:
            per[PMC_KERNELS[name]][i][row['Counter_Name']] = float(row['Counter_Value'])
    lines = []
    for k in sorted(per, key=list(PMC_KERNELS.values()).index):
        ids = sorted(per[k])
        n = (launches or {}).get(k)
        ids = ids[-n:] if n else ids
        hit = sum(per[k][i].get('TCC_HIT_sum', 0.) for i in ids) / len(ids)
        miss = sum(per[k][i].get('TCC_MISS_sum', 0.) for i in ids) / len(ids)
        rate = hit / (hit + miss) if hit + miss else 0.
        lines.append(f'kernel {k} dispatches {len(ids)} l2_hits {hit:.0f} l2_misses {miss:.0f} hit_rate {rate:.3f}')
    return '\n'.join(lines) + '\n'


def trace_summary(trace_csv: str, bench_json: str) -> str:
    """Per-kernel durations from a rocprofv3 kernel trace of this script: every dispatch, and the
    dispatches of the per-kernel HIP-event pass of the single-mesh section (the figures the JSON
    line's roofline uses). That pass is the last --kernel-steps dispatches of each step kernel on
    the single-mesh workload's grid; CFL and finish dispatches are taken from the same interval."""
    import csv
    from collections import defaultdict
    with open(bench_json) as f:
        res = json.loads([ln for ln in f if ln.startswith('{')][-1])
    cfg = res['config']
    cells, sides = cfg['cells'], cfg['sides']
    steps = res['kernels']['flux']['launches']
    tuning = cfg.get('tuning') or {}
    block, cpt = tuning.get('cell_block', 256), tuning.get('cells_per_thread', 1)
    grid = lambda n, per, threads: ((n + per - 1) // per + 7) // 8 * 8 * threads
    # threads of each pass's grid: the flux pass takes two sides per thread, the cell passes the
    # tuned workgroup size and cells per thread
    unique = {'k_flux': grid(sides, 512, 256), 'k_donor': grid(cells, block * cpt, block),
              'k_depth': grid(cells, block * cpt, block)}
    rows = list(csv.DictReader(open(trace_csv)))
    every = defaultdict(list)
    for r in rows:
        every[kernel_base_name(r['Kernel_Name'])].append(int(r['End_Timestamp']) - int(r['Start_Timestamp']))
    picked = {}
    for name, g in unique.items():
        sel = [r for r in rows if kernel_base_name(r['Kernel_Name']) == name and int(r['Grid_Size_X']) == g]
        picked[name] = sel[-steps:]
    # the timing pass launches each kernel `steps` times in a row: gradients, flux, donor, depth,
    # then CFL and finish, so the CFL and finish repeats are the first ones after the depth repeats
    t1 = max(int(r['End_Timestamp']) for r in picked['k_depth'])
    for name in ('k_cfl', 'k_finish'):
        picked[name] = [r for r in rows if kernel_base_name(r['Kernel_Name']) == name
                        and int(r['Start_Timestamp']) > t1][:steps]
    t0 = min(int(r['Start_Timestamp']) for r in picked['k_flux'])
    grads = [r for r in rows if kernel_base_name(r['Kernel_Name']) == 'k_gradients' and int(r['End_Timestamp']) <= t0]
    picked['k_gradients'] = grads[-steps:]
    lines = ['| Kernel | Dispatches (all) | Average all (us) | Dispatches (timing pass) | Average timing pass (us) | HIP-event average (us) |',
             '| --- | ---: | ---: | ---: | ---: | ---: |']
    short = {'k_gradients': 'gradients', 'k_flux': 'flux', 'k_donor': 'donor', 'k_depth': 'depth', 'k_cfl': 'cfl',
             'k_finish': 'finish'}
    for name in ('k_gradients', 'k_flux', 'k_donor', 'k_depth', 'k_cfl', 'k_finish'):
        d = [int(r['End_Timestamp']) - int(r['Start_Timestamp']) for r in picked[name]]
        ev = res['kernels'].get(short[name], {}).get('avg_us')
        lines.append(f'| {name} | {len(every[name])} | {np.mean(every[name]) / 1e3:.1f} | {len(d)} | '
                     f'{np.mean(d) / 1e3:.1f} | {ev:.1f} |' if d else f'| {name} | {len(every[name])} | - | 0 | - | - |')
    batch = sorted((k for k in every if k.startswith('kb_')), key=lambda k: -sum(every[k]))
    bk = res.get('batched', {}).get('kernels', {})
    lines += ['', '| Batch kernel | Dispatches (all) | Average all (us) | Dispatches (timing pass) | '
              'Average timing pass (us) | HIP-event average (us) |', '| --- | ---: | ---: | ---: | ---: | ---: |']
    for name in batch:
        d = np.array(every[name]) / 1e3
        short = name[3:].replace('_wide', '')   # kb_donor_wide: the 64-slot donor pass
        n = bk.get(short, {}).get('launches', 0)
        tail = d[-n:] if n else d[:0]
        ev = bk.get(short, {}).get('avg_us')
        lines.append(f'| {name} | {len(d)} | {d.mean():.1f} | {len(tail)} | '
                     + (f'{tail.mean():.1f}' if len(tail) else '-') + ' | ' + (f'{ev:.1f}' if ev else '-') + ' |')
    # the parts instantiations of the single-mesh kernels (meshes of different sizes in one union)
    parts = defaultdict(list)
    for r in rows:
        full = r['Kernel_Name']
        name = kernel_base_name(full)
        if name.startswith('k_') and re.search(r'<[^>]*\btrue\b', full):
            parts[name].append(int(r['End_Timestamp']) - int(r['Start_Timestamp']))
    if parts:
        # the timed union run is the union's last activity: its steps are each kernel's last dispatches
        timed = res.get('uneven', {}).get('steps', 0)
        lines += ['', '| Union kernel (parts) | Dispatches (all) | Average all (us) | Dispatches (timed run) | '
                  'Average timed run (us) | Min timed (us) | Max timed (us) |',
                  '| --- | ---: | ---: | ---: | ---: | ---: | ---: |']
        for name in sorted(parts, key=lambda k: -sum(parts[k])):
            d = np.array(parts[name]) / 1e3
            tail = d[-timed:] if timed and name not in ('k_gradients',) else d[:0]
            stats = (f'{len(tail)} | {tail.mean():.1f} | {tail.min():.1f} | {tail.max():.1f}' if len(tail)
                     else '0 | - | - | -')
            lines.append(f'| {name} | {len(d)} | {d.mean():.1f} | {stats} |')
    return '\n'.join(lines) + '\n'


def read_traffic(cells: int) -> dict | None:
    """HBM bytes per dispatch from the committed counter summary, if it was made on this workload."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    text = TRAFFIC_SUMMARY.read_text()
    m = re.search(r'^workload cells (\d+)', text, re.M)
    if not m or int(m.group(1)) != cells:
        return None
    return {k: float(v) for k, v in re.findall(r'^kernel (\w+) .* hbm_bytes (\d+)', text, re.M)}


def _cpu_model() -> str:
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(sc, forcing, developed: dict, t_start: float, seconds_budget: float = 10.0):
    """The CPU restatement of the reference's passes (the C oracle, OpenMP) timed on bounded
    samples of the same workload: with all the host threads it is given, then with one thread
    (the reference's own single-thread protocol)."""
    threads = int(os.environ.get('OMP_NUM_THREADS', '16'))
    os.environ['OMP_NUM_THREADS'] = str(threads)
    import ctypes
    import oracle
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, make_params
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout, initial_state
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
    state = initial_state(sc.mesh, layout)
    state.update({k: v.copy() for k, v in developed.items()})
    eng = Engine(layout, make_params(sc.afa, sc.sita, sc.min_h, 32), state, device=oracle.HostDevice('f32'))
    _set_forcing(eng, forcing, t_start)
    eng.reset(seconds_to_ticks(t_start))
    eng.compute_cfl()
    eng.begin(seconds_to_ticks(t_start + 1e5), seconds_to_ticks(t_start))
    eng.run_batch(1)  # first touch
    cells = sc.mesh.e_num - 1

    def sample(budget):
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            rows = eng.run_batch(4)
            steps += int(np.count_nonzero(rows[:, 3] == 1))
        return steps, time.perf_counter() - t0
    steps, wall = sample(seconds_budget)
    gomp = ctypes.CDLL('libgomp.so.1')      # the oracle's OpenMP runtime, already loaded
    gomp.omp_set_num_threads(1)
    steps1, wall1 = sample(seconds_budget)
    gomp.omp_set_num_threads(threads)
    eng.close()
    return dict(value=cells * steps / wall, unit='cell-updates/s', cores=threads, kind='port',
                label='CPU restatement of the reference passes (C oracle, OpenMP), not Taichi',
                cpu_model=_cpu_model(), nproc=len(os.sched_getaffinity(0)),
                sample=f'{steps} steps of the same {cells}-cell mesh and state, '
                       f'{wall:.1f} s wall, C oracle with {threads} OpenMP threads',
                single_thread_value=cells * steps1 / wall1,
                single_thread_sample=f'{steps1} steps, {wall1:.1f} s wall, 1 thread')


def fp64_error(sc, forcing, developed: dict, t_start: float, device, window: float = 10.0) -> dict:
    """Error of the F32 HIP step against an FP64 run of the same passes (the C oracle built with
    double arithmetic, as the checker), over `window` simulated seconds from the developed state,
    adaptive dt on each side and the forcing set at each batch's start: max |dh|, max |dq|, every
    water-budget term and the stored volume, relative to the volume scale (storage at the start
    plus the window's rain). Beside it, the F32 step against itself started from depths one ulp
    deeper: the flow's own sensitivity over the same window, which bounds what any arithmetic can
    be held to from this state."""
    import oracle
    from nh_flood_2d_amd import abi
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, make_params
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout, initial_state
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
    params = make_params(sc.afa, sc.sita, sc.min_h, 32)
    now, target = seconds_to_ticks(t_start), seconds_to_ticks(t_start + window)
    bumped = {k: v.copy() for k, v in developed.items()}
    wet = bumped['depth'] > 0
    bumped['depth'][wet] = np.nextafter(bumped['depth'][wet], np.float32(1))

    def run(dev, start):
        st = initial_state(sc.mesh, layout)
        st.update({k: v.copy() for k, v in start.items()})
        e = Engine(layout, params, st, device=dev)
        e.reset(now)
        _set_forcing(e, forcing, t_start)
        e.compute_cfl()
        t, steps = now, 0
        while t < target:
            _set_forcing(e, forcing, t / 2 ** 40)
            e.begin(target, t)
            rows = e.run_batch(32)
            done = rows[rows[:, 3] == abi.COMPLETED]
            if not len(done):
                raise RuntimeError(f'no step completed: status {rows[0, 3]}')
            t = (int(done[-1, 1]) << 32) | int(done[-1, 2])
            steps += len(done)
        out = dict(depth=e.read_state('depth').astype(np.float64), q=e.read_state('q').astype(np.float64),
                   budget=e.flush_budget().sum(axis=1), steps=steps)
        e.close()
        return out
    f32, f64, ulp = run(device, developed), run(oracle.HostDevice('f64'), developed), run(device, bumped)
    _, _, _, area = sc.mesh.geometry()
    rain = forcing.rain
    scale = float(np.dot(developed['depth'].astype(np.float64), area)) + float(
        np.dot(rain.rates(rain.index_at(t_start))[rain.pixel_index][1:].astype(np.float64), area[1:])) * window
    terms = np.abs(f32['budget'] - f64['budget']) / scale
    names = ['infiltration', 'floor_nonlevel', 'floor_level', 'pipe']
    for b in range(layout.b_num):
        names += [f'level{b}_{c}' for c in ('storage', 'faces', 'pipe', 'rain', 'infiltration')]
    names += [f'discharge{b}' for b in range(layout.b_num)]
    worst = int(np.argmax(terms))
    volume = abs(float(np.dot(f32['depth'] - f64['depth'], area))) / scale
    return dict(window_s=window, start_s=t_start, steps_f32=f32['steps'], steps_fp64=f64['steps'],
                max_abs_dh_m=float(np.max(np.abs(f32['depth'] - f64['depth']))),
                max_abs_dq_m2_s=float(np.max(np.abs(f32['q'] - f64['q']))),
                budget_terms_rel=dict(zip(names, (float(x) for x in terms))), worst_budget_term=names[worst],
                worst_budget_rel=float(terms[worst]), stored_volume_rel=volume, volume_scale_m3=scale,
                one_ulp_max_abs_dh_m=float(np.max(np.abs(f32['depth'] - ulp['depth']))),
                one_ulp_max_abs_dq_m2_s=float(np.max(np.abs(f32['q'] - ulp['q']))),
                note='F32 HIP against the FP64 C oracle from the developed state; one_ulp_* is the F32 step '
                     'against itself from depths one ulp deeper (the flow\'s sensitivity over the window)')


def _set_forcing(eng, forcing, t: float):
    c = np.zeros((max(1, len(forcing.boundaries)), 2), dtype=np.float32)
    for bi, b in enumerate(forcing.boundaries):
        times = b.series.times
        i = min(int(np.searchsorted(times, t, side='right')) - 1, len(times) - 2)
        c[bi] = (b.series.at(t), (b.series.values[i + 1] - b.series.values[i]) / (times[i + 1] - times[i]))
    rain = forcing.rain
    eng.set_forcing(c.ravel() if forcing.boundaries else np.zeros(0, np.float32),
                    rain.rates(rain.index_at(t)).astype(np.float32), True)


def _chunks(steps: int, batch: int) -> list:
    """Batch sizes the protocol below uses for `steps` steps."""
    return [min(batch, steps - done) for done in range(0, steps, batch)]


def _prime(engines, run_batch, steps: int, now: list, batch: int = 32) -> None:
    """Capture, outside any timed region, the graph of every batch size a run of `steps` steps will
    launch: each is launched once towards a target equal to the clock, so every step is inactive
    and the state does not change (the kernels return at once)."""
    for n in sorted(set(_chunks(steps, batch))):
        for i, eng in enumerate(engines):
            eng.begin(now[i], now[i])
        run_batch(n)


def _run_protocol(engines, run_batch, steps: int, horizon: int, now: list, batch: int = 32) -> int:
    """Advance every engine `steps` steps with the solver's batch protocol (begin, run, read the
    trace); returns the completed steps summed over the engines."""
    done, total = 0, 0
    while done < steps:
        n = min(batch, steps - done)
        for i, eng in enumerate(engines):
            eng.begin(horizon, now[i])
        trace = run_batch(n)
        for i in range(len(engines)):
            rows = trace[i]
            ok = rows[:, 3] == 1
            if not ok.all():
                raise RuntimeError(f'scenario {i}: batch stopped early, status {rows[~ok][0, 3]}')
            now[i] = (int(rows[-1, 1]) << 32) | int(rows[-1, 2])
            total += n
        done += n
    return total


# The reference's only published step time: a 3,850-cell 4 m mesh, gates disabled, 0.212545 ms per
# step of synchronized device-graph wall time excluding the first two batches, on an Apple M1 Max
# with Taichi 1.7.4 Metal (docs/gate-validation/metal-performance-v2.json:92-95).
REFERENCE_SMALL_MS = 0.212545


def flood_map_bench(sc, developed: dict, device, repeats: int = 20) -> dict:
    """The flood-map rasteriser on the benchmark mesh's developed state: every cell's masked
    depth written to the pixels of the mesh's 2 m grid (one 2400 x 2400 tile), timed with HIP
    events on the rasteriser's stream over `repeats` back-to-back passes. Algorithmic bytes per
    pass: 20 B per cell (centre, half sizes, value) read and 4 B per pixel written."""
    from nh_flood_2d_amd import maps
    meta = maps.area_meta(sc.mesh)
    grid = meta.grid()
    raster = maps.Rasterizer(meta, device)
    image = raster.raster(maps.masked_depth(developed['depth'][1:], sc.min_h, maps.NODATA), grid)
    painted = int(np.sum((2 * meta.half_w / meta.hr) * (2 * meta.half_h / meta.vr)))
    torch = device.torch
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    raster.enqueue(grid)
    start.record(device.stream)
    for _ in range(repeats):
        raster.enqueue(grid)
    end.record(device.stream)
    end.synchronize()
    us = start.elapsed_time(end) * 1e3 / repeats
    cells = len(meta.x)
    nbytes = 20 * cells + 4 * painted
    gbs = nbytes / us / 1e3
    log(f'flood map: {cells} cells on {grid.width} x {grid.height} px ({painted} painted), {us:.1f} us per map, '
        f'{gbs:.0f} GB/s of its bytes')
    return dict(workload=f'masked depth of the developed benchmark state on its {float(meta.hr):g} m grid',
                cells=cells, width=grid.width, height=grid.height, painted_pixels=painted,
                wet_pixels=int(np.count_nonzero(image != maps.NODATA)), us_per_map=us,
                cells_per_s=cells / us * 1e6, pixels_per_s=painted / us * 1e6, algorithmic_bytes=nbytes,
                achieved_gbs=gbs, frac_of_hbm_peak=gbs / 8000., launches=repeats)


def small_mesh_bench(args, device) -> dict:
    """Time per step on a mesh of the size of the reference's published timing: 62 x 62 cells of
    4 m (3,844 cells), seeded DEM with hills and a channel, rain, wet start, plain path with one
    graph per 32-step batch; the first two batches are excluded as in the reference's definition."""
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, make_params
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout, initial_state
    sc = synth.plain_uniform(args.seed, n=62, dx=4.)
    forcing = synth.scenario_forcing(sc)
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
    state = initial_state(sc.mesh, layout)
    state['depth'][1:] = np.float32(0.05)
    eng = Engine(layout, make_params(sc.afa, sc.sita, sc.min_h, 32), state, device=device)
    t_start = 600.
    _set_forcing(eng, forcing, t_start)
    eng.reset(seconds_to_ticks(t_start))
    eng.compute_cfl()
    eng.set_graph(True)
    now, horizon = [seconds_to_ticks(t_start)], seconds_to_ticks(t_start + 1e5)
    one = lambda n: eng.run_batch(n)[None]
    _run_protocol([eng], one, 64, horizon, now)          # the two excluded batches
    eng.synchronize()
    steps = 32 * 40
    w0 = time.perf_counter()
    done = _run_protocol([eng], one, steps, horizon, now)
    eng.synchronize()
    ms = (time.perf_counter() - w0) / done * 1e3
    cells = sc.mesh.e_num - 1
    log(f'small mesh: {cells} cells of 4 m, {ms * 1e3:.1f} us per step ({cells / ms / 1e6:.3f} G cell-updates/s); '
        f'the reference on an M1 Max with Metal: {REFERENCE_SMALL_MS * 1e3:.1f} us')
    eng.close()
    return {'workload': f'{cells} cells, uniform 4 m grid (62 x 62), seeded hills and channel, rain, wet start, '
                        'plain path, one graph per 32-step batch, first two batches excluded',
            'cells': cells, 'steps': done, 'ms_per_step': ms, 'value': cells / (ms * 1e-3), 'unit': 'cell-updates/s',
            'reference_ms_per_step': REFERENCE_SMALL_MS,
            'reference': 'published: 3,850-cell 4 m mesh, gates disabled, Apple M1 Max, Taichi 1.7.4 Metal',
            'ratio_to_reference': REFERENCE_SMALL_MS / ms}


def develop_batch(variants, layout0, params, device, until: float, per_graph: int = 64) -> list:
    """Every scenario's own storm run on the same-mesh batch (BatchRunner: one slot each, own dt)
    from its initial state to `until` seconds; returns each scenario's state."""
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.batch import BatchEngine, BatchRunner
    batch = BatchEngine(len(variants), device=device)
    batch.set_graph(True)
    runner = BatchRunner(batch, batch_size=per_graph)
    sims, states = {}, {}

    class Final:   # the state at the last output, before the runner frees the slot
        def __init__(self, key):
            self.key = key

        def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
            if final:
                states[self.key] = sim.state()

        def finish(self, sim):
            pass
    for i, v in enumerate(variants):
        def make(factory, i=i, v=v):
            sims[i] = synth.simulation(v, engine_factory=factory, end_time=until, yield_step=until,
                                       batch_size=per_graph, base_layout=layout0, outputs=Final(i))
            return sims[i]
        runner.submit(i, make)
    runner.run()
    batch.close()
    return [states[i] for i in range(len(variants))], max(sim.steps for sim in sims.values())


def batched_bench(args, device, base=None, label: str = 'batch') -> dict:
    """Same-mesh scenario batch: `--batch-slots` scenarios (rain peak, Manning factor and start
    state drawn per scenario) on one shared mesh (by default the two-level 1 M-cell mesh), each
    developed by its own storm to the rain peak (a third of its duration) in the batch, then
    timed from there; against the first of them alone on the plain path from the same state."""
    import torch
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.batch import BatchEngine
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, make_params
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout, initial_state, relayout
    t0 = time.time()
    if base is None:
        base = synth.two_level_fractal(args.seed, blocks=args.batch_blocks)
    variants = synth.scenario_variants(base, args.batch_slots, seed=args.seed + 1)
    mesh = base.mesh
    cells = mesh.e_num - 1
    cf = cell_forcing_arrays(mesh, synth.scenario_forcing(variants[0]))
    layout0 = build_layout(mesh, variants[0].land_parameters, cf, base.sita)
    params = make_params(base.afa, base.sita, base.min_h, 64)
    t_start = base.duration / 3
    log(f'{label} workload: {len(variants)} scenarios x {cells} cells, {mesh.s_num - 1} sides ({time.time() - t0:.1f} s)')
    w0 = time.perf_counter()
    developed, develop_steps = develop_batch(variants, layout0, params, device, t_start)
    wet = float(np.mean([np.mean(st['depth'][1:] >= base.min_h) for st in developed]))
    log(f'{label} developed: every scenario to {t_start:.0f} s (the rain peak), at most {develop_steps} steps, '
        f'{time.perf_counter() - w0:.1f} s wall; wet cells {wet:.3f} on average')
    start = seconds_to_ticks(t_start)
    horizon = seconds_to_ticks(t_start + 1e5)

    def states():
        for v, dev_state in zip(variants, developed):
            lay = relayout(layout0, mesh, v.land_parameters, cf)
            st = initial_state(mesh, lay)
            st.update({k: x.copy() for k, x in dev_state.items()})
            yield v, lay, st

    # plain single-scenario path on the first scenario
    v, lay, st = next(states())
    plain = Engine(lay, params, st, device=device)
    _set_forcing(plain, synth.scenario_forcing(v), t_start)
    plain.reset(start)
    plain.compute_cfl()
    plain.set_graph(True)
    now = [start]
    one = lambda n: plain.run_batch(n)[None]
    _run_protocol([plain], one, args.warmup, horizon, now)
    _prime([plain], one, args.steps, now)
    plain.synchronize()
    w0 = time.perf_counter()
    steps_plain = _run_protocol([plain], one, args.steps, horizon, now)
    plain.synchronize()
    plain_wall = time.perf_counter() - w0
    plain_value = cells * steps_plain / plain_wall
    del plain
    torch.cuda.synchronize()

    batch = BatchEngine(len(variants), device=device)
    views = []
    for v, lay, st in states():
        view = batch.view()(lay, params, st)
        _set_forcing(view, synth.scenario_forcing(v), t_start)
        view.reset(start)
        views.append(view)
    batch.compute_cfl()
    batch.set_graph(True)
    btune, per_graph = None, 32
    if args.tune:
        from nh_flood_2d_amd.tune import tune_batch
        btune = tune_batch(batch, views, steps=32, rounds=1)
        per_graph = btune['steps_per_graph']
        log(f"{label} tuned: depth grid {btune['depth_blocks']}, CFL grid {btune['cfl_blocks']}, "
            f"{per_graph} steps per graph, {btune['us_per_step']:.0f} us/step")
    now = [start] * len(views)   # forcing origins; the slots' clocks continue from the tuning steps
    _run_protocol(views, batch.run_batch, args.warmup, horizon, now, per_graph)
    _prime(views, batch.run_batch, args.steps, now, per_graph)
    batch.synchronize()
    w0 = time.perf_counter()
    steps_batch = _run_protocol(views, batch.run_batch, args.steps, horizon, now, per_graph)
    batch.synchronize()
    batch_wall = time.perf_counter() - w0
    batch_value = cells * steps_batch / batch_wall
    # per-kernel times of the batched step (eager pass, HIP events)
    batch.set_graph(False)
    batch.set_timing(True)
    batch.kernel_times(reset=True)
    _run_protocol(views, batch.run_batch, args.kernel_steps, horizon, now, per_graph)
    ms, count = batch.kernel_times(reset=True)
    batch.set_timing(False)
    kernels = {name: dict(avg_us=ms[i] / count[i] * 1e3, launches=int(count[i]))
               for i, name in enumerate(KERNELS) if count[i]}
    sim_seconds = [(n - start) / 2 ** 40 for n in now]
    log(f'{label}ed: {steps_batch} scenario-steps in {batch_wall:.3f} s, {batch_value / 1e9:.2f} G cell-updates/s; '
        f'plain: {plain_value / 1e9:.2f} G; ratio {batch_value / plain_value:.2f}')
    log(f'{label} kernels: ' + ', '.join(f"{k} {v['avg_us']:.1f} us" for k, v in kernels.items()))
    batch.close()
    return {
        'workload': f'{len(variants)} scenarios on one shared mesh ({base.name}), {cells} cells, '
                    f'{mesh.s_num - 1} sides; rain peak U[20,150] mm/h, Manning x lognormal(0.2), dry or 0.3 m '
                    f'pond start, own dt per scenario; each developed by its own storm to the rain peak '
                    f'({t_start:.0f} s) in the batch, then timed',
        'scenarios': len(variants), 'cells': cells, 'value': batch_value, 'unit': 'cell-updates/s',
        'plain_value': plain_value, 'batched_vs_plain': batch_value / plain_value,
        'scenario_steps': steps_batch, 'wall_s': batch_wall, 'plain_steps': steps_plain,
        'simulated_s_min': min(sim_seconds), 'simulated_s_max': max(sim_seconds), 'kernels': kernels,
        'tuning': btune, 'batch_steps': per_graph, 'wet_cell_frac': wet,
        'developed': dict(until_s=t_start, max_steps=int(develop_steps)),
    }


def uneven_bench(args, device) -> dict:
    """Uneven-size batch: --uneven-meshes meshes with cell counts log-uniform in
    [--uneven-min-cells, --uneven-max-cells], 1-3 refinement levels (finest cells 5 m), seeded
    fractal DEMs, 2-4 land classes with Horton infiltration, a tide on half of them and a 0.3 m
    pond on half, each developed by its own storm to the rain peak in one union. From there all
    advance together in one union (one launch sequence per step
    for every mesh, each with its own dt; autotuned cell-pass shape and steps per graph, one
    hipGraph per batch); the same steps are then timed one mesh at a time on single-mesh engines."""
    import torch
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, make_params
    from nh_flood_2d_amd.layout import initial_state
    from nh_flood_2d_amd.union import UnionEngine
    t0 = time.time()
    items = synth.uneven_batch(args.seed + 21, args.uneven_meshes, args.uneven_min_cells, args.uneven_max_cells,
                               workers=args.uneven_workers)
    cells = np.array([sc.mesh.e_num - 1 for sc, _ in items], dtype=np.int64)
    levels = [int(np.max((lay.arrays['cell_meta'][1:] >> 16) & 15)) + 1 for _, lay in items]
    log(f'uneven workload: {len(items)} meshes, {int(cells.sum())} cells, from {int(cells.min())} to '
        f'{int(cells.max())}, levels {levels} ({time.time() - t0:.1f} s)')
    sc0 = items[0][0]
    params = make_params(sc0.afa, sc0.sita, sc0.min_h, 64)

    t_peak = sc0.duration / 3
    # every mesh's own storm to its rain peak, all in one union (UnionRunner: own dt per mesh)
    from nh_flood_2d_amd.union import UnionRunner
    w0 = time.perf_counter()
    developer = UnionEngine([lay for _, lay in items], params, device=device)
    developer.set_graph(True)
    sims = {p: synth.simulation(sc, engine_factory=developer.view(p), end_time=t_peak, yield_step=t_peak,
                                base_layout=lay) for p, (sc, lay) in enumerate(items)}
    UnionRunner(developer, sims).run()
    developed = [sims[p].state() for p in range(len(items))]
    develop_steps = max(sim.steps for sim in sims.values())
    developer.close()
    del developer, sims
    wet = float(np.sum([np.sum(st['depth'][1:] >= sc0.min_h) for st in developed]) / cells.sum())
    log(f'uneven developed: every mesh to {t_peak:.0f} s (the rain peak), at most {develop_steps} steps, '
        f'{time.perf_counter() - w0:.1f} s wall; wet cells {wet:.3f}')

    def state_of(p, sc, layout):
        st = initial_state(sc.mesh, layout)
        st.update({k: x.copy() for k, x in developed[p].items()})
        return st

    start = seconds_to_ticks(t_peak)
    horizon = seconds_to_ticks(t_peak + 1e6)
    union = UnionEngine([lay for _, lay in items], params, device=device)
    views = []
    for p, (sc, lay) in enumerate(items):
        v = union.view(p)(lay, params, state_of(p, sc, lay))
        _set_forcing(v, synth.scenario_forcing(sc), t_peak)
        v.reset(start)
        views.append(v)
    union.compute_cfl()
    union.set_graph(True)
    utune, per_graph = None, 32
    if args.tune:
        from nh_flood_2d_amd.tune import tune_union
        utune = tune_union(union, views, steps=32, rounds=1)
        per_graph = utune['steps_per_graph']
        log(f"uneven tuned: cell workgroup {utune['cell_block']} x {utune['cells_per_thread']}, "
            f"{per_graph} steps per graph, {utune['us_per_step']:.0f} us/step")
    now = [start] * len(views)    # forcing origins; the parts' clocks continue from the tuning steps
    _run_protocol(views, union.run_batch, args.uneven_warmup, horizon, now, per_graph)
    _prime(views, union.run_batch, args.uneven_steps, now, per_graph)
    union.synchronize()
    w0 = time.perf_counter()
    _run_protocol(views, union.run_batch, args.uneven_steps, horizon, now, per_graph)
    union.synchronize()
    union_wall = time.perf_counter() - w0
    updates = int(cells.sum()) * args.uneven_steps
    union_value = updates / union_wall
    union.close()
    del union, views
    torch.cuda.synchronize()
    # the same steps one mesh at a time (tuned like the union's cell passes)
    plain_wall = 0.0
    for p, (sc, lay) in enumerate(items):
        eng = Engine(lay, params, state_of(p, sc, lay), device=device)
        if utune:
            eng.set_tuning(eng.tuning_range()['cfl_blocks'], utune['cell_block'], utune['cells_per_thread'])
        _set_forcing(eng, synth.scenario_forcing(sc), t_peak)
        eng.reset(start)
        eng.compute_cfl()
        eng.set_graph(True)
        one = lambda n, eng=eng: eng.run_batch(n)[None]
        mine = [start]
        _run_protocol([eng], one, args.uneven_warmup, horizon, mine, per_graph)
        _prime([eng], one, args.uneven_steps, mine, per_graph)
        eng.synchronize()
        w0 = time.perf_counter()
        _run_protocol([eng], one, args.uneven_steps, horizon, mine, per_graph)
        eng.synchronize()
        plain_wall += time.perf_counter() - w0
        eng.close()
        del eng
    plain_value = updates / plain_wall
    log(f'uneven: {len(items)} meshes x {args.uneven_steps} steps in {union_wall:.3f} s, '
        f'{union_value / 1e9:.2f} G cell-updates/s; one mesh at a time {plain_wall:.3f} s, '
        f'{plain_value / 1e9:.2f} G; ratio {union_value / plain_value:.2f}')
    return {
        'workload': f'{len(items)} meshes, {int(cells.sum())} cells in all ({int(cells.min())} to {int(cells.max())}, '
                    f'log-uniform {args.uneven_min_cells}-{args.uneven_max_cells}), 1-3 levels (finest 5 m), '
                    'fractal DEMs, 2-4 land classes with Horton, tide on half, pond on half, own dt per mesh; '
                    f'each developed by its own storm to the rain peak ({t_peak:.0f} s) in the union, then timed',
        'wet_cell_frac': wet, 'developed': dict(until_s=t_peak, max_steps=int(develop_steps)),
        'meshes': len(items), 'cells': int(cells.sum()), 'value': union_value, 'unit': 'cell-updates/s',
        'plain_value': plain_value, 'uneven_vs_plain': union_value / plain_value, 'steps': args.uneven_steps,
        'wall_s': union_wall, 'plain_wall_s': plain_wall, 'tuning': utune, 'batch_steps': per_graph,
    }


def served_bench(args) -> dict:
    """Serving load: Poisson arrivals (rate --serve-rate per second) of --serve-requests solves, each a
    seeded two-level mesh with log-uniform cell count in [--serve-min-cells, --serve-max-cells],
    simulated duration uniform in [--serve-min-h, --serve-max-h] hours, a triangular hyetograph
    with peak U[20, 150] mm/h and, for half of them, a sinusoidal tide boundary; outputs every
    300 s leave through the frame stream. Run once served (continuous batching, per-solve streams)
    and once one solve at a time on the plain path."""
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout
    from nh_flood_2d_amd.serve import FloodServer, FrameStream, Request, footprint_bytes
    t0 = time.time()
    specs = []
    for i, sc, arrival in synth.serving_load(args.seed, args.serve_requests, args.serve_rate, args.serve_min_cells,
                                             args.serve_max_cells, args.serve_min_h, args.serve_max_h):
        layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, synth.scenario_forcing(sc)),
                              sc.sita)
        specs.append((i, sc, layout, arrival))
    log(f'serving workload: {len(specs)} solves, cells {[s[1].mesh.e_num - 1 for s in specs]}, '
        f'hours {[round(s[1].duration / 3600, 2) for s in specs]} ({time.time() - t0:.1f} s)')

    def requests(stream):
        out = []
        for i, sc, layout, arr in specs:
            def build(factory, sc=sc, layout=layout, i=i):
                return synth.simulation(sc, engine_factory=factory, yield_step=300., base_layout=layout,
                                        batch_size=64, outputs=stream.sink(i))
            out.append(Request(i, build, arr, sc.mesh.e_num - 1, footprint_bytes(layout), layout=layout,
                               params=(sc.afa, sc.sita, sc.min_h)))
        return out

    stream = FrameStream(None, every_output=True)
    server = FloodServer(batch_size=64)
    served = server.serve(requests(stream), realtime=True)
    stream.close()
    s = served.summary()
    import torch
    device_bytes = torch.cuda.get_device_properties(0).total_memory
    log(f"served: {s['requests']} solves, p50 {s['p50_latency_s']:.2f} s, p99 {s['p99_latency_s']:.2f} s, "
        f"{s['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s, peak live {s['peak_live']}, peak footprint "
        f"{s['peak_device_bytes'] / 1e9:.2f} GB of a {server.budget / 1e9:.0f} GB admission budget "
        f"({device_bytes / 1e9:.0f} GB device)")
    u = None
    if args.serve_union_window > 0:
        stream = FrameStream(None, every_output=True)
        grouped = FloodServer(batch_size=64, union_window=args.serve_union_window,
                              union_parts=args.serve_union_parts).serve(requests(stream), realtime=True)
        stream.close()
        u = grouped.summary()
        log(f"served through unions (window {args.serve_union_window} s, up to {args.serve_union_parts} solves): "
            f"{u['unions']} unions, p50 {u['p50_latency_s']:.2f} s, p99 {u['p99_latency_s']:.2f} s, "
            f"{u['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s")
    stream = FrameStream(None, every_output=True)
    alone = FloodServer(batch_size=64, max_live=1).serve(requests(stream), realtime=False)
    stream.close()
    a = alone.summary()
    log(f"one at a time: {a['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s")
    return {
        'workload': f'{len(specs)} solves, Poisson arrivals {args.serve_rate}/s, two-level meshes '
                    f'{args.serve_min_cells}-{args.serve_max_cells} cells (log-uniform), '
                    f'{args.serve_min_h}-{args.serve_max_h} h simulated, rain peak U[20,150] mm/h, tide on half, '
                    'outputs every 300 s',
        'value': s['cell_updates_per_s'], 'unit': 'cell-updates/s', 'p50_latency_s': s['p50_latency_s'],
        'p99_latency_s': s['p99_latency_s'], 'plain_value': a['cell_updates_per_s'],
        'served_vs_plain': s['cell_updates_per_s'] / a['cell_updates_per_s'], 'served': s, 'one_at_a_time': a,
        'admission_budget_bytes': server.budget, 'device_bytes': device_bytes,
        'through_unions': None if u is None else dict(
            value=u['cell_updates_per_s'], vs_plain=u['cell_updates_per_s'] / a['cell_updates_per_s'],
            window_s=args.serve_union_window, max_parts=args.serve_union_parts, summary=u),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=640)
    ap.add_argument('--warmup', type=int, default=128)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--develop-s', type=float, default=3600.,
                    help='simulated seconds of the storm run from the dry start before timing')
    ap.add_argument('--kernel-steps', type=int, default=64, help='steps of the batch engine\'s per-kernel timing pass')
    ap.add_argument('--kernel-repeats', type=int, default=50, help='back-to-back launches per kernel when timing kernels')
    ap.add_argument('--cpu-seconds', type=float, default=10.0, help='per CPU sample (all threads, then one)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='also copy a 256 MiB buffer 3 times (reference dispatch for HBM counters)')
    ap.add_argument('--no-tune', dest='tune', action='store_false',
                    help='skip the autotuner (default launch grids, 32 steps per graph)')
    ap.add_argument('--batch-slots', type=int, default=64, help='scenarios of the same-mesh batch (0: skip)')
    ap.add_argument('--batch-blocks', type=int, default=790, help='10 m blocks per side of the batch mesh')
    ap.add_argument('--no-small-batch', dest='small_batch', action='store_false',
                    help='skip the batch on the 256 x 256 grid of the CPU-runnable configuration')
    ap.add_argument('--no-flood-map', dest='flood_map', action='store_false',
                    help='skip the flood-map rasteriser leg')
    ap.add_argument('--no-small-mesh', dest='small_mesh', action='store_false',
                    help='skip the time per step on a mesh of the size of the reference\'s published timing')
    ap.add_argument('--uneven-meshes', type=int, default=32, help='meshes of the uneven-size batch (0: skip)')
    ap.add_argument('--uneven-min-cells', type=int, default=50_000)
    ap.add_argument('--uneven-max-cells', type=int, default=4_000_000)
    ap.add_argument('--uneven-steps', type=int, default=256)
    ap.add_argument('--uneven-warmup', type=int, default=64)
    ap.add_argument('--uneven-workers', type=int, default=16, help='processes that build the uneven meshes')
    ap.add_argument('--serve-requests', type=int, default=12, help='solves of the serving load (0: skip)')
    ap.add_argument('--serve-rate', type=float, default=2.0)
    ap.add_argument('--serve-min-cells', type=int, default=100_000)
    ap.add_argument('--serve-max-cells', type=int, default=8_000_000)
    ap.add_argument('--serve-union-window', type=float, default=0.0,
                    help='also serve the load through unions: seconds an arrival may wait for others (0: skip)')
    ap.add_argument('--serve-union-parts', type=int, default=8)
    ap.add_argument('--serve-min-h', type=float, default=1.0)
    ap.add_argument('--serve-max-h', type=float, default=6.0)
    ap.add_argument('--trace-summary', nargs=2, metavar=('KERNEL_TRACE_CSV', 'BENCH_JSON'),
                    help='summarise a rocprofv3 --kernel-trace run of this script (no GPU needed) and exit')
    ap.add_argument('--l2-summary', nargs=2, metavar=('COUNTER_CSV', 'BENCH_JSON'),
                    help='summarise a rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum run of this script (no GPU needed) and exit')
    ap.add_argument('--pmc-summary', nargs=3, metavar=('FETCH_CSV', 'WRITE_CSV', 'BENCH_JSON'),
                    help='summarise two rocprofv3 --pmc runs of this script (no GPU needed) and exit')
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(*args.trace_summary), end='')
        return
    if args.l2_summary:
        counter_csv, bench_json = args.l2_summary
        with open(bench_json) as f:
            res = json.loads([ln for ln in f if ln.startswith('{')][-1])
        launches = {k: v['launches'] for k, v in res.get('kernels', {}).items()}
        launches.update({f'batch_{k}': v['launches'] for k, v in res.get('batched', {}).get('kernels', {}).items()})
        print(l2_summary(counter_csv, launches), end='')
        return
    if args.pmc_summary:
        fetch_csv, write_csv, bench_json = args.pmc_summary
        with open(bench_json) as f:
            line = [ln for ln in f if ln.startswith('{')][-1]
        cfg = json.loads(line)['config']
        res = json.loads(line)
        launches = {k: v['launches'] for k, v in res.get('kernels', {}).items()}
        launches.update({f'batch_{k}': v['launches'] for k, v in res.get('batched', {}).get('kernels', {}).items()})
        print(f"workload cells {cfg['cells']} sides {cfg['sides']} side_list {cfg['side_list']}")
        print(pmc_summary(fetch_csv, write_csv, launches), end='')
        return
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py runs on one GPU only: there is no multi-GPU path '
              '(independent scenarios would shard whole scenarios per GPU)', file=sys.stderr)
        sys.exit(2)

    import torch
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, TorchDevice, make_params
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout, initial_state

    t0 = time.time()
    sc, forcing = build_workload(args.seed)
    mesh = sc.mesh
    layout = build_layout(mesh, sc.land_parameters, cell_forcing_arrays(mesh, forcing), sc.sita)
    r_len = int(layout.arrays['_r_len'][0])
    log(f'workload: {mesh.e_num - 1} cells, {mesh.s_num - 1} sides, R={r_len}, '
        f'hanging edges {layout.n_hang}, gradient cells {layout.n_grad} ({time.time() - t0:.1f} s)')
    device = TorchDevice()
    t_start = float(args.develop_s)
    developed, develop = develop_state(sc, device, t_start)
    wet = wet_fractions(mesh, developed, sc.min_h)
    log(f"developed: the storm from the dry start to {t_start:.0f} s in {develop['steps']} steps "
        f"({develop['wall_s']:.1f} s wall); wet cells {wet['wet_cell_frac']:.3f}, wet faces {wet['wet_face_frac']:.3f}")
    state = initial_state(mesh, layout)
    state.update({k: v.copy() for k, v in developed.items()})
    params = make_params(sc.afa, sc.sita, sc.min_h, 64)
    tuning = None
    per_graph = 32
    if args.tune:
        from nh_flood_2d_amd.tune import tuned_engine_parameters
        c = np.zeros((max(1, len(forcing.boundaries)), 2), dtype=np.float32)
        tuning = tuned_engine_parameters(layout, params, state, device,
                                         forcing=(c.ravel() if forcing.boundaries else np.zeros(0, np.float32),
                                                  forcing.rain.rates(forcing.rain.index_at(t_start)).astype(np.float32),
                                                  True), steps=96)
        per_graph = tuning['steps_per_graph']
        log(f"tuned: cell workgroup {tuning['cell_block']} x {tuning.get('cells_per_thread', 1)} cells per thread, "
            f"CFL grid {tuning['cfl_blocks']}, {per_graph} steps per graph, "
            f"{tuning['us_per_step']:.1f} us/step (candidates {tuning['table']})")
    eng = Engine(layout, params, state, device=device)
    if tuning:
        eng.set_tuning(tuning['cfl_blocks'], tuning['cell_block'], tuning.get('cells_per_thread', 1))
    start = seconds_to_ticks(t_start)
    _set_forcing(eng, forcing, t_start)
    eng.reset(start)
    eng.compute_cfl()
    eng.set_graph(True)
    horizon = seconds_to_ticks(t_start + 1e5)
    now = start

    def run(steps):
        nonlocal now
        done = 0
        while done < steps:
            n = min(per_graph, steps - done)
            eng.begin(horizon, now)
            rows = eng.run_batch(n)
            ok = rows[:, 3] == 1
            if not ok.all():
                raise RuntimeError(f'batch stopped early: status {rows[~ok][0, 3]}')
            last = rows[-1]
            now = (int(last[1]) << 32) | int(last[2])
            done += n
        return done

    run(args.warmup)
    _prime([eng], lambda n: eng.run_batch(n)[None], args.steps, [now], per_graph)
    eng.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(device.stream):
        ev0.record()
    w0 = time.perf_counter()
    steps = run(args.steps)
    with torch.cuda.stream(device.stream):
        ev1.record()
    eng.synchronize()
    wall = time.perf_counter() - w0
    event_ms = ev0.elapsed_time(ev1)
    cells = mesh.e_num - 1
    value = cells * steps / wall
    sim_seconds = (now - start) / 2 ** 40
    log(f'timed: {steps} steps in {wall:.4f} s wall ({event_ms:.2f} ms by events), '
        f'{value / 1e9:.3f} G cell-updates/s, simulated {sim_seconds:.2f} s')

    # per-kernel timing: each step kernel launched --kernel-repeats times back to back on the
    # current inputs, bracketed by HIP events on the engine's stream (the state is discarded after)
    eng.begin(horizon, now)
    us = eng.time_kernels(args.kernel_repeats)
    ms = us * 1e-3 * args.kernel_repeats
    count = np.where(us > 0, args.kernel_repeats, 0)
    nbytes = launch_bytes(layout, mesh.e_num, mesh.s_num, r_len, uniform_rain=eng.rain_uniform)
    alg = algorithmic_bytes(layout, mesh.e_num, mesh.s_num, r_len)
    traffic = read_traffic(cells)
    kernels = {}
    for i, name in enumerate(KERNELS):
        if count[i]:
            avg_ms = ms[i] / count[i]
            kernels[name] = dict(avg_us=avg_ms * 1e3, launches=int(count[i]),
                                 launch_bytes=int(nbytes[name]),
                                 launch_gbs=nbytes[name] / (avg_ms * 1e-3) / 1e9 if nbytes[name] else None,
                                 hbm_bytes=traffic.get(name) if traffic else None)
    # phases of the step: (a) face flux = k_flux; (b) cell update = k_donor + k_depth (the donor
    # limiter needs every face of a cell before any face is scaled, hence two launches). Each
    # phase's rate is taken over the bytes its kernels must move in this layout (model_bytes);
    # the algorithmic figure (every per-step input read once, as the reference's passes do) is kept
    # beside it.
    phases = {'flux': ['flux'], 'cell_update': ['donor', 'depth']}
    # Phase (a) needs the friction and slope coefficients (centre distance, Manning n: 8 B) and
    # the hanging-edge data only on a wet face: a face below min_h carries nothing and reads
    # neither. Its rate is therefore taken over the bytes the wet fraction requires.
    wet_faces = wet['wet_face_frac'] * (mesh.s_num - 1)
    hang = (4 * (mesh.s_num - 1) + 24 * layout.n_hang + 24 * layout.n_grad) if layout.n_hang else 0
    flux_wet_bytes = int(17 * (mesh.s_num - 1) + 8 * wet_faces + 8 * cells + hang)
    phase = {}
    for p, ks in phases.items():
        us = sum(kernels[k]['avg_us'] for k in ks)
        hbm = sum(traffic[k] for k in ks) if traffic and all(k in traffic for k in ks) else None
        model = flux_wet_bytes if p == 'flux' else sum(nbytes[k] for k in ks)
        phase[p] = dict(kernels=ks, avg_us=us, algorithmic_bytes=alg[p], model_bytes=model,
                        achieved_gbs=model / (us * 1e-6) / 1e9, hbm_bytes=hbm)
    # the roofline's unit is the longer of the two phases with algorithmic bytes of their own
    # ((a) the face flux, (b) the cell update: k_donor + k_depth, two launches because the donor
    # limiter needs every face of a cell before any face is scaled); `achieved` is the phase's
    # algorithmic bytes over the sum of its launches' average durations, beside the rates over
    # the bytes its kernels must move in this state (model) and over the counters' bytes
    dominant = max(phase, key=lambda p: phase[p]['avg_us'])
    dp = dict(phase[dominant])
    dp['achieved_gbs'] = dp['algorithmic_bytes'] / (dp['avg_us'] * 1e-6) / 1e9
    log('kernels: ' + ', '.join(f"{k} {v['avg_us']:.1f} us" for k, v in kernels.items()))
    log('phases: ' + ', '.join(f"{p} {v['avg_us']:.1f} us {v['achieved_gbs']:.0f} GB/s"
                               for p, v in phase.items()))
    result = {
        'metric': METRIC, 'value': value, 'unit': 'cell-updates/s', 'n_gpus': 1, 'steps': steps,
        'warmup': args.warmup, 'ms_per_step': event_ms / steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': f'multi-resolution 2/4/8 m mesh, {cells} cells, {mesh.s_num - 1} sides, '
                               'seeded channel network + diamond-square DEM, 4 Horton land classes, '
                               'tide boundary, triangular hyetograph peaking at 80 mm/h at 1 h; state developed '
                               f'by that storm from the dry start to t = {t_start:.0f} s, then timed at the '
                               'rain peak, plain single-scenario path',
                   'cells': cells, 'sides': mesh.s_num - 1, 'side_list': r_len,
                   'developed': develop, **wet,
                   'hanging_edges': layout.n_hang, 'batch_steps': per_graph,
                   'launch': f'hipGraph per {per_graph}-step batch', 'tuning': tuning},
        'roofline': {'bound': 'hbm', 'kernel': '+'.join(f'k_{k}' for k in dp['kernels']), 'phase': dominant,
                     'achieved': dp['achieved_gbs'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': dp['achieved_gbs'] / HBM_PEAK_GBS, 'traffic': dp['hbm_bytes'],
                     'algorithmic_bytes': dp['algorithmic_bytes'], 'avg_us': dp['avg_us'],
                     'model_bytes': dp['model_bytes'],
                     'frac_model': dp['model_bytes'] / (dp['avg_us'] * 1e-6) / 1e9 / HBM_PEAK_GBS,
                     'frac_traffic': (dp['hbm_bytes'] / (dp['avg_us'] * 1e-6) / 1e9 / HBM_PEAK_GBS
                                      if dp['hbm_bytes'] else None)},
        'phases': phase,
        'phase_note': 'achieved_gbs is taken over model_bytes, the bytes the phase\'s kernels must move: for (a) '
                      '17 B per side (cell ids 8, flags 1, q read and write 8), 8 B more per wet face (centre '
                      'distance, Manning n), 8 B per cell (bed, depth); for (b) the depth pass\'s launch bytes and '
                      'one donor flag byte per cell (the donor certificate leaves the donor pass only flagged and '
                      'boundary cells to read); algorithmic_bytes is the algorithmic figure, every input read once as '
                      'the reference does; hbm_bytes is what rocprofv3 counters measured',
        'kernels': kernels,
        'step': {'algorithmic_bytes': alg['step'],
                 'achieved_gbs': alg['step'] / (event_ms / steps * 1e-3) / 1e9,
                 'hbm_bytes': sum(traffic[k] for k in kernels if k in traffic) if traffic else None},
    }
    # what a plain 256 MiB device copy reaches on this part (read + write bytes over its average
    # duration by HIP events on the engine's stream), beside the roofline's published peak
    with torch.cuda.stream(device.stream):
        src = torch.ones(64 << 20, dtype=torch.float32, device=device.device)
        dst = torch.empty_like(src)
        for _ in range(3 if args.calibrate_copy else 2):
            dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            dst.copy_(src)
        e1.record()
    device.stream.synchronize()
    copy_gbs = 20 * 2 * src.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    # the phase's physical rate (counter bytes when the committed summary matches this workload,
    # else the bytes its kernels must move) against what the copy reaches
    moved = dp['hbm_bytes'] or dp['model_bytes']
    moved_gbs = moved / (dp['avg_us'] * 1e-6) / 1e9
    result['roofline']['copy_gbs'] = copy_gbs
    result['roofline']['frac_of_copy'] = moved_gbs / copy_gbs
    result['roofline']['frac_of_copy_bytes'] = 'traffic' if dp['hbm_bytes'] else 'model'
    log(f'device copy: {copy_gbs:.0f} GB/s; {dominant} moves {moved_gbs:.0f} GB/s, {moved_gbs / copy_gbs:.2f} of it; '
        f'algorithmic {dp["achieved_gbs"]:.0f} GB/s, {dp["achieved_gbs"] / HBM_PEAK_GBS:.2f} of the peak')
    if args.calibrate_copy:
        result['copy_calibration_bytes'] = 2 * src.numel() * 4
    del src, dst
    if args.flood_map:
        result['flood_map'] = flood_map_bench(sc, developed, device)
    if args.small_mesh:
        result['small_mesh'] = small_mesh_bench(args, device)
    if args.batch_slots:
        result['batched'] = batched_bench(args, device)
        if args.small_batch:
            # the same batch on the CPU-runnable configuration's 256 x 256 grid of 10 m cells, where
            # the plain path is bound by launches rather than bytes
            from nh_flood_2d_amd import synth
            result['batched_small'] = batched_bench(args, device, base=synth.plain_uniform(args.seed, n=256, dx=10.),
                                                    label='small batch')
    if args.uneven_meshes:
        result['uneven'] = uneven_bench(args, device)
    if args.serve_requests:
        result['served'] = served_bench(args)
    if not args.no_cpu:
        del eng
        torch.cuda.synchronize()
        result['cpu_baseline'] = cpu_baseline(sc, forcing, developed, t_start, args.cpu_seconds)
        result['precision'] = fp64_error(sc, forcing, developed, t_start, device)
        p = result['precision']
        log(f"precision over {p['window_s']:.0f} s from the developed state: max|dh| {p['max_abs_dh_m']:.3g} m, "
            f"max|dq| {p['max_abs_dq_m2_s']:.3g} m2/s, worst budget term {p['worst_budget_rel']:.3g}, stored volume "
            f"{p['stored_volume_rel']:.3g} against FP64; one ulp deeper: max|dh| {p['one_ulp_max_abs_dh_m']:.3g} m")
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
