"""Scenario batches on the GPU: each scenario alone and in the mix (needs an MI355X).

Every scenario runs once on the single-scenario HIP engine and once inside a batch on the same
mesh (more scenarios than slots, so a slot is reused). The dt sequence, final state and extrema
must be identical bit for bit; the mass-balance rows may differ only by the grouping of the FP64
budget sums (relative 1e-9, far inside the 1e-7 volume-budget tolerance).
"""
import dataclasses
import math

import numpy as np
import pytest

from nh_flood_2d_amd import synth
from nh_flood_2d_amd.batch import BatchEngine, BatchRunner

pytestmark = pytest.mark.gpu


def _say(text):
    print(text, flush=True)


def _close_rows(a: dict, b: dict, rel=1e-9, atol=1e-9):
    assert a.keys() == b.keys()
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, float) and isinstance(y, float):
            if math.isfinite(x) or math.isfinite(y):
                assert abs(x - y) <= atol + rel * max(abs(x), abs(y)), (k, x, y)
        else:
            assert x == y, (k, x, y)


def _run_alone(v, device, yield_step):
    from nh_flood_2d_amd.engine import Engine

    def factory(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=device, pipe_q=pipe_q)
    sim = synth.simulation(v, engine_factory=factory, yield_step=yield_step)
    row = sim.run()
    out = row, sim.record['dt_sequence_sha256'], sim.state(), sim.steps
    sim.engine.close()
    return out


def _two_pixels(variants):
    """The variants with a two-pixel rain raster: cells west of the median centre get the
    hyetograph, the others 0.6 of it, so no slot can keep one rain time (the per-cell rain state)."""
    out = []
    for v in variants:
        x = v.mesh.geometry()[0][:, 0]
        pixel = (x > np.median(x[1:])).astype(np.int32)
        pixel[0] = 0
        mm = np.asarray(v.rain_mm, float)[:, :1]
        out.append(dataclasses.replace(v, rain_mm=np.concatenate([mm, 0.6 * mm], axis=1), rain_pixel=pixel))
    return out


def _compare(variants, alone, results, sims):
    for i, (row, sha, state, steps) in enumerate(alone):
        assert sims[i].record['dt_sequence_sha256'] == sha, f'scenario {i}: dt sequence differs'
        assert sims[i].steps == steps
        _close_rows(results[i], row)


# 64 slots take the wide kernels (several sides or cells per wavefront, one wavefront holding
# the 64 slots of each); fewer slots take the narrow ones. The wide cases bind more scenarios than
# slots (every lane busy, slots reused) or few scenarios with boundaries, hanging edges and sita < 1.
# One rain pixel keeps every slot in the uniform rain state (one rain time, one capacity per class);
# the two-pixel cases keep the per-cell rain time and capacity arrays.
@pytest.mark.parametrize('case', ['two_level', 'multires_tide_river', 'multires_theta', 'two_level_wide',
                                  'multires_tide_river_wide', 'multires_theta_wide', 'two_pixel', 'two_pixel_wide'])
def test_batch_matches_each_scenario_alone(gpu_device, case):
    if case == 'two_level':
        base = synth.two_level_fractal(5, blocks=48, duration=600.)
        variants, slots = synth.scenario_variants(base, 6, seed=7), 4
    elif case == 'multires_tide_river':
        base = synth.small_multires(3, duration=600., discharge=1.5)
        variants, slots = synth.scenario_variants(base, 3, seed=8), 2
    elif case == 'multires_theta':
        base = synth.small_multires(4, duration=480., theta=0.7, pond=0.2)
        variants, slots = synth.scenario_variants(base, 3, seed=9), 4
    elif case == 'two_level_wide':
        base = synth.two_level_fractal(12, blocks=20, duration=300.)
        variants, slots = synth.scenario_variants(base, 70, seed=13), 64
    elif case == 'multires_tide_river_wide':
        base = synth.small_multires(14, duration=480., discharge=1.5)
        variants, slots = synth.scenario_variants(base, 5, seed=15), 64
    elif case == 'multires_theta_wide':
        base = synth.small_multires(16, duration=420., theta=0.7, pond=0.2)
        variants, slots = synth.scenario_variants(base, 5, seed=17), 64
    elif case == 'two_pixel':
        base = synth.small_multires(18, duration=480., discharge=1.0)
        variants, slots = _two_pixels(synth.scenario_variants(base, 3, seed=19)), 4
    else:
        base = synth.two_level_fractal(20, blocks=20, duration=300.)
        variants, slots = _two_pixels(synth.scenario_variants(base, 6, seed=21)), 64
    _say(f'[batch] {case}: {base.mesh.e_num - 1} cells, {len(variants)} scenarios, {slots} slots; alone runs')
    alone = [_run_alone(v, gpu_device, 120.) for v in variants]
    _say(f'[batch] {case}: batched run')
    states = {}
    batch = BatchEngine(slots, device=gpu_device)
    sims = {}
    runner = BatchRunner(batch, batch_size=32)
    for i, v in enumerate(variants):
        def make(factory, i=i, v=v):
            sims[i] = synth.simulation(v, engine_factory=factory, yield_step=120.)
            return sims[i]
        runner.submit(i, make)
    # the wide cases change the grids after the first batch: the smallest (every workgroup strides
    # over many cells and sides) or the largest the context allows; every grid gives the same bits
    regrid = case.endswith('_wide')
    # read each scenario's final state before its slot is reused: run until it leaves
    while runner.step():
        if regrid:
            rng = batch.tuning_range()
            if case == 'two_level_wide':
                batch.set_tuning(8, 8)
            else:
                batch.set_tuning(rng['depth_max'] // 8 * 8, rng['cfl_max'] // 8 * 8)
            regrid = False
        for i, sim in sims.items():
            if i not in states and sim.elapsed >= sim.end_time:
                states[i] = sim.state()
    for i in sims:
        if i not in states:
            states[i] = sims[i].state()
    _compare(variants, alone, runner.results, sims)
    for i, (_, _, state, _) in enumerate(alone):
        for name, values in state.items():
            np.testing.assert_array_equal(states[i][name], values, err_msg=f'{case} scenario {i} {name}')
    batch.close()


def test_batch_donor_limiting_matches_each_scenario_alone(gpu_device):
    """Thin sheets of water on the multi-resolution mesh's slopes drain faster than some cells
    hold, in the 64-slot batch: its donor pass scans the flags of 16 cells per wavefront and takes
    the flagged ones two at a time, in several rounds where flagged cells cluster. Each scenario
    (sheet depth, sheet extent and Manning factor drawn per scenario) must take the dt sequence,
    final state and mass balance of its run alone, limiter counts included, and the limiter must
    have acted in every scenario."""
    base = synth.small_multires(11, duration=240., discharge=1.0)
    rng = np.random.default_rng(23)
    variants = []
    for i in range(6):
        sheet = np.float32(rng.uniform(0.021, 0.032))
        below = np.float32(rng.uniform(2.5, 4.0))
        depth = np.where(base.mesh.z < below, sheet, np.float32(0.)).astype(np.float32)
        depth[0] = 0.
        factor = float(np.exp(rng.normal(0., 0.3)))
        land = {k: dict(v, manning_n=float(np.float32(v['manning_n'] * factor)))
                for k, v in base.land_parameters.items()}
        variants.append(dataclasses.replace(base, name=f'sheet_v{i}', land_parameters=land, initial_depth=depth))
    _say(f'[batch] thin sheets: {base.mesh.e_num - 1} cells, {len(variants)} scenarios, 64 slots; alone runs')
    alone = [_run_alone(v, gpu_device, 60.) for v in variants]
    _say('[batch] thin sheets: batched run')
    batch = BatchEngine(64, device=gpu_device)
    sims = {}
    runner = BatchRunner(batch, batch_size=32)
    for i, v in enumerate(variants):
        def make(factory, i=i, v=v):
            sims[i] = synth.simulation(v, engine_factory=factory, yield_step=60.)
            return sims[i]
        runner.submit(i, make)
    while runner.step():
        pass
    _compare(variants, alone, runner.results, sims)
    for i, (_, _, state, _) in enumerate(alone):
        got = sims[i].state()
        for name, values in state.items():
            np.testing.assert_array_equal(got[name], values, err_msg=f'thin sheet scenario {i} {name}')
    cells = [sims[i].record['limiter_activity']['limited_cell_steps'] for i in range(len(variants))]
    _say(f'[batch] thin sheets: limited cell-steps per scenario {cells}')
    assert min(cells) > 0
    batch.close()


def test_batch_graph_replays_match_eager(gpu_device):
    base = synth.small_multires(6, duration=300., discharge=1.0)
    variants = synth.scenario_variants(base, 2, seed=10)
    _say('[batch] graph vs eager, small_multires seed 6')
    out = []
    for graph in (False, True):
        batch = BatchEngine(2, device=gpu_device)
        sims = {}
        runner = BatchRunner(batch, batch_size=16)
        for i, v in enumerate(variants):
            def make(factory, i=i, v=v):
                sims[i] = synth.simulation(v, engine_factory=factory, yield_step=60., batch_size=16)
                sims[i].engine.set_graph(graph)
                return sims[i]
            runner.submit(i, make)
        finals = {}
        while runner.step():
            for i, sim in sims.items():
                if i not in finals and sim.elapsed >= sim.end_time:
                    finals[i] = sim.state()
        out.append((finals, {i: s.record['dt_sequence_sha256'] for i, s in sims.items()}))
        batch.close()
    (fa, sa), (fb, sb) = out
    assert sa == sb
    for i in fa:
        for name in fa[i]:
            np.testing.assert_array_equal(fa[i][name], fb[i][name], err_msg=f'{i} {name}')


def test_batch_rain_state_follows_the_written_arrays(gpu_device):
    """A slot's rain state follows what the host writes: uniform at the cold start (one rain pixel),
    per cell after a rain time that differs between cells is written, uniform again after one rain
    time and one capacity per class are written. In every state the slot steps as the plain engine
    given the same writes, bit for bit, and reads back the arrays as written."""
    from nh_flood_2d_amd.clock import seconds_to_ticks
    from nh_flood_2d_amd.engine import Engine, make_params
    from nh_flood_2d_amd.forcing import cell_forcing_arrays
    from nh_flood_2d_amd.layout import build_layout, initial_state
    from parity import coefficients_at

    sc = synth.small_multires(22, duration=900., discharge=1.0, pond=0.25)
    forcing = synth.scenario_forcing(sc)
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
    params = make_params(sc.afa, sc.sita, sc.min_h, 32)
    state = initial_state(sc.mesh, layout)
    state['depth'] = sc.initial_depth.astype(np.float32).copy()
    plain = Engine(layout, params, state, device=gpu_device)
    batch = BatchEngine(2, device=gpu_device)
    view = batch.view(1)(layout, params, state)
    engines = (plain, view)
    for e in engines:
        e.reset(0)
        e.compute_cfl()
    now = 0
    cls = layout.arrays['cell_meta'].astype(np.int64) & 255
    f0 = layout.cell_f0

    def advance(batches):
        nonlocal now
        for _ in range(batches):
            t = now / 2 ** 40
            rates = forcing.rain.rates(forcing.rain.index_at(t)).astype(np.float32)
            coeff = coefficients_at(forcing, t).ravel()[:2 * len(forcing.boundaries)]
            for e in engines:
                e.set_forcing(coeff, rates, True)
                e.begin(seconds_to_ticks(t + 30.), now)
            rows = [e.run_batch(32) for e in engines]
            np.testing.assert_array_equal(rows[0], rows[1])
            done = rows[0][rows[0][:, 3] == 1]
            now = (int(done[-1, 1]) << 32) | int(done[-1, 2])
        for name in ('depth', 'q', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick'):
            np.testing.assert_array_equal(plain.read_state(name), view.read_state(name), err_msg=name)

    def write_both(name, values):
        for e in engines:
            e.write_state(name, values)
        np.testing.assert_array_equal(view.read_state(name), values, err_msg=f'{name} read back')

    _say('[batch] rain state: uniform from the cold start')
    advance(4)
    rt = plain.read_state('rain_time')
    assert rt[1:].min() == rt[1:].max() > 0
    _say('[batch] rain state: per cell after a rain time that differs between cells')
    uneven = rt.copy()
    uneven[1::3] += np.float32(7.5)
    write_both('rain_time', uneven)
    advance(4)
    _say('[batch] rain state: uniform again after one rain time and per-class capacities')
    write_both('rain_time', np.full_like(rt, np.float32(40.)))
    cap = f0.astype(np.float32) * np.float32(0.5)
    cap[0] = 0.
    write_both('capacity', cap)
    advance(4)
    plain.close()
    batch.close()
