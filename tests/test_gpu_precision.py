"""Error of the F32 HIP step against an FP64 run of the same kernels (the oracle built with double
arithmetic), and the benchmark mesh against the F32 oracle bit for bit.

Each result states max |dh|, max |dq| and the error of every water-budget term against the FP64
run, relative to the budget's volume scale (initial storage plus every input); the stored volume
Sum h*A is stated separately, because it is the F32 depths themselves summed (their rounding, not
an accounting error: every budget sum is FP64 on both sides).

Bars: each budget term <= 1e-7 relative (the volume-budget tolerance); the stored volume and the depths to
the envelopes each test states. F32 and FP64 part where a cell's depth meets the wet/dry gate
(min_h) in a different step; away from such crossings they differ by rounding only.

* The CPU-runnable configuration in full: 256 x 256 cells of 10 m, 50 mm/h for 1800 s on a dry
  start, closed boundaries, 3600 s, an output every 300 s, adaptive dt on both sides.
* The benchmark mesh (1.58 M cells, 2/4/8 m, Horton land classes, a tide boundary), its own storm
  from the dry start to 1800 s, outputs every 300 s.
* From that storm's developed state at 3600 s, where the flow amplifies rounding-size
  perturbations, F32 against FP64 beside F32 against F32 with one ulp more depth.
"""
import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import abi, synth
from nh_flood_2d_amd.clock import seconds_to_ticks
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, initial_state
from nh_flood_2d_amd.simulation import RunSettings, Simulation
from parity import coefficients_at

pytestmark = pytest.mark.gpu

H_BAR = 1e-4
TERM_BAR = 1e-7


def _say(*args):
    print(*args, flush=True)


@pytest.fixture(scope='module')
def developed(gpu_device):
    """The benchmark configuration's storm run by the solver on the GPU from the dry start to
    3600 s; returns the scenario, its layout and the state in the mesh's numbering."""
    sc = synth.multires_channel(0)

    def factory(layout, params, state, **kw):
        e = Engine(layout, params, state, device=gpu_device, **kw)
        e.set_graph(True)
        return e
    sim = synth.simulation(sc, engine_factory=factory, end_time=3600., yield_step=3600.)
    sim.run()
    state = sim.state()
    _say(f'[developed] {sc.mesh.e_num - 1} cells to 3600 s in {sim.steps} steps, '
         f'wet cells {np.mean(state["depth"][1:] >= sc.min_h):.3f}')
    layout = sim.layout
    sim.engine.close()
    return sc, layout, state


def _engines(gpu_device, developed, precision):
    sc, layout, state = developed
    forcing = synth.scenario_forcing(sc)
    params = make_params(sc.afa, sc.sita, sc.min_h, 32)
    st = initial_state(sc.mesh, layout)
    st.update({k: v.copy() for k, v in state.items()})
    gpu = Engine(layout, params, st, device=gpu_device)
    ref = Engine(layout, params, st, device=oracle.HostDevice(precision))
    now = seconds_to_ticks(3600.)
    for e in (gpu, ref):
        e.reset(now)
        _force(e, forcing, now)
        e.compute_cfl()
    return sc, forcing, gpu, ref, now


def _force(engine, forcing, now):
    t = now / 2 ** 40
    coeff = coefficients_at(forcing, t)[:len(forcing.boundaries)].ravel()
    rates = forcing.rain.rates(forcing.rain.index_at(t)).astype(np.float32)
    engine.set_forcing(coeff, rates, True)


def _advance(engine, forcing, now, target):
    """Batches of adaptive steps, each with the forcing at its start, until the clock is at target."""
    steps = 0
    while now < target:
        _force(engine, forcing, now)
        engine.begin(target, now)
        rows = engine.run_batch(32)
        done = rows[rows[:, 3] == abi.COMPLETED]
        assert len(done) > 0 and np.all(rows[len(done):, 3] == abi.DONE), rows[len(done)]
        now = (int(done[-1, 1]) << 32) | int(done[-1, 2])
        steps += len(done)
    return now, steps


def test_benchmark_mesh_matches_the_f32_oracle(gpu_device, developed):
    """The benchmark's whole step at its full size on its storm-developed state, 32 adaptive steps
    in one batch: the same trace (dt bits, clock words, limiter counts) and the same state bits as
    the F32 oracle."""
    sc, _, gpu, f32, now = _engines(gpu_device, developed, 'f32')
    _say(f'[full size] {sc.mesh.e_num - 1} cells against the F32 oracle')
    rows = []
    for e in (gpu, f32):
        e.begin(now + seconds_to_ticks(600.), now)
        rows.append(e.run_batch(32))
    np.testing.assert_array_equal(rows[0], rows[1])
    assert np.all(rows[0][:, 3] == 1)
    for name in ('depth', 'q', 'max_depth', 'max_speed', 'capacity', 'rain_time', 'peak_tick'):
        a, b = gpu.read_state(name), f32.read_state(name)
        assert a.tobytes() == b.tobytes(), name
    np.testing.assert_allclose(gpu.flush_budget().sum(axis=1), f32.flush_budget().sum(axis=1), rtol=1e-9, atol=1e-9)
    gpu.close()
    f32.close()


def _term_names(b_num):
    names = ['infiltration', 'floor_nonlevel', 'floor_level', 'pipe']
    for b in range(b_num):
        names += [f'level{b}_{c}' for c in ('storage', 'faces', 'pipe', 'rain', 'infiltration')]
    names += [f'discharge{b}' for b in range(b_num)]
    return names


def test_benchmark_storm_first_half_hour_against_fp64(gpu_device):
    """The benchmark configuration's storm from its dry start to 1800 s, adaptive dt on both sides,
    outputs every 300 s: every budget term within 1e-7 of the FP64 run, the depths within 1e-4 m
    (rounding only: nothing crosses the wet/dry gate differently yet)."""
    sc = synth.multires_channel(0)
    _say(f'[precision] {sc.mesh.e_num - 1} cells, 0-1800 s from the dry start, against FP64')

    def gpu_factory(layout, params, state):
        return Engine(layout, params, state, device=gpu_device)
    sim_g, fg = _simulate(sc, gpu_factory, 1800., 300.)
    sim_c, fc = _simulate(sc, oracle.engine_factory('f64'), 1800., 300.)
    assert [f[0] for f in fg] == [f[0] for f in fc] and len(fg) == 7
    for (t, hg, qg, rg), (_, hc, qc, rc) in zip(fg, fc):
        scale = max(rc['initial_storage_m3'] + rc['rainfall_m3'] +
                    sum(abs(v) for k, v in rc.items() if k.startswith('level:') or k.startswith('discharge:')), 1.)
        terms = {k: abs(rg[k] - rc[k]) / scale for k, v in rc.items() if k.endswith('_m3') and
                 k not in ('initial_storage_m3', 'storage_m3', 'residual_m3') and isinstance(v, float)}
        worst = max(terms, key=terms.get)
        dh, dq = float(np.max(np.abs(hg - hc))), float(np.max(np.abs(qg - qc)))
        volume = abs(rg['storage_m3'] - rc['storage_m3']) / scale
        _say(f'[precision] t={t:.0f} s: max|dh| {dh:.3g} m, max|dq| {dq:.3g} m2/s, worst budget term {worst} '
             f'{terms[worst]:.3g}, stored volume {volume:.3g} (relative to {scale:.6g} m3)')
        assert dh <= H_BAR, (t, dh)
        assert terms[worst] <= TERM_BAR, (t, terms)
        assert volume <= 5e-6, (t, volume)   # F32 depths summed: up to 2.8e-6 over the first 2400 s
    assert fc[-1][3]['infiltration_m3'] > 0 and fc[-1][3]['rainfall_m3'] > 0
    _say(f'[precision] 0-1800 s: steps gpu={sim_g.steps} fp64={sim_c.steps}')


def test_developed_storm_is_as_sensitive_to_one_ulp_as_to_fp64(gpu_device, developed):
    """From the storm-developed state at 3600 s (channels filling from the tide, sheet flow at the
    wet/dry gate on impervious ground, sita 1) the flow amplifies any perturbation: a run whose
    depths are moved by one F32 ulp parts from the unperturbed F32 run as far as the FP64 run
    does. Past this point the F32/FP64 difference measures the flow's sensitivity, not the
    arithmetic. Both runs' divergence after 10 s is printed and compared."""
    sc, forcing, gpu, f64, now = _engines(gpu_device, developed, 'f64')
    _, layout, state = developed
    bumped = {k: v.copy() for k, v in state.items()}
    wet = bumped['depth'] > 0
    bumped['depth'][wet] = np.nextafter(bumped['depth'][wet], np.float32(1))
    st = initial_state(sc.mesh, layout)
    st.update(bumped)
    ulp = Engine(layout, make_params(sc.afa, sc.sita, sc.min_h, 32), st, device=gpu_device)
    ulp.reset(now)
    _force(ulp, forcing, now)
    ulp.compute_cfl()
    _, _, _, area = sc.mesh.geometry()
    target = seconds_to_ticks(3610.)
    for e in (gpu, f64, ulp):
        _advance(e, forcing, now, target)
    h = [e.read_state('depth').astype(np.float64) for e in (gpu, f64, ulp)]
    flooded = [float(area[x >= sc.min_h].sum()) for x in h]
    d64, d1 = float(np.max(np.abs(h[0] - h[1]))), float(np.max(np.abs(h[0] - h[2])))
    a64, a1 = abs(flooded[0] - flooded[1]) / flooded[1], abs(flooded[0] - flooded[2]) / flooded[2]
    _say(f'[sensitivity] 3600-3610 s: F32 vs FP64 max|dh| {d64:.3g} m, flooded area {a64:.3g}; '
         f'F32 vs F32 one ulp deeper max|dh| {d1:.3g} m, flooded area {a1:.3g}')
    assert d1 > 1e-3 and d64 > 1e-3            # the flow is ill-conditioned here
    assert d64 <= 3 * d1 and a64 <= 3 * a1     # FP64 parts from F32 no further than one ulp does
    for e in (gpu, f64, ulp):
        e.close()


def _simulate(sc, factory, end, yield_step):
    forcing = synth.scenario_forcing(sc)
    settings = RunSettings(afa=sc.afa, sita=sc.sita, min_h=sc.min_h, end_time=end, yield_step=yield_step)
    frames = []

    class Grab:
        def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
            frames.append((seconds, depth.astype(np.float64), sim.engine.read_state('q').astype(np.float64),
                           dict(row)))

        def finish(self, sim):
            pass

    sim = Simulation(sc.mesh, sc.land_parameters, forcing, settings, engine_factory=factory, outputs=Grab())
    sim.run()
    return sim, frames


def test_plain_config_full_run_against_fp64(gpu_device):
    """The CPU-runnable configuration at full size and duration, adaptive dt on both sides.

    Uniform rain fills every cell alike, so the whole grid reaches the wet/dry gate (min_h, below
    which a face carries nothing) within one step; F32 rounding of the depth (about 1e-8 m by
    then) decides which step each cell crosses it in. Before that (outputs up to 1200 s) the two
    runs differ by rounding only and are held to 1e-6 m; after it, individual cells differ by
    millimetres where the crossing step differed, and the test holds the frame's maximum depth
    and flooded area to measured envelopes (1e-2 m, 1 %) and prints the figures. The budget
    terms are held to the 1e-7 bar throughout; the stored volume (F32 depths summed) to 2e-5."""
    sc = synth.plain_uniform(0, n=256, dx=10.)
    _say('[precision] plain 256 x 256, 3600 s, against FP64')

    def gpu_factory(layout, params, state):
        return Engine(layout, params, state, device=gpu_device)
    sim_g, fg = _simulate(sc, gpu_factory, 3600., 300.)
    sim_c, fc = _simulate(sc, oracle.engine_factory('f64'), 3600., 300.)
    assert [f[0] for f in fg] == [f[0] for f in fc] and len(fg) == 13
    _, _, _, area = sc.mesh.geometry()
    for (t, hg, qg, rg), (_, hc, qc, rc) in zip(fg, fc):
        scale = rc['initial_storage_m3'] + rc['rainfall_m3'] + sum(abs(v) for k, v in rc.items()
                                                                    if k.startswith('discharge:'))
        scale = max(scale, 1.)
        terms = [k for k, v in rc.items() if k.endswith('_m3') and k not in
                 ('initial_storage_m3', 'storage_m3', 'residual_m3') and isinstance(v, float)]
        term = max(abs(rg[k] - rc[k]) / scale for k in terms)
        d = np.abs(hg - hc)
        dh, dq = float(d.max()), float(np.max(np.abs(qg - qc)))
        peak = abs(float(hg.max()) - float(hc.max()))
        wet_g, wet_c = float(area[hg > sc.min_h].sum()), float(area[hc > sc.min_h].sum())
        flooded = abs(wet_g - wet_c) / max(wet_c, 1.)
        volume = abs(rg['storage_m3'] - rc['storage_m3']) / scale
        _say(f'[precision] t={t:.0f} s: max|dh| {dh:.3g} m (99.9 % of cells within '
             f'{float(np.quantile(d[1:], 0.999)):.3g}), max depth {float(hc.max()):.4f} m differs by '
             f'{peak:.3g}, flooded area by {flooded:.3g}, max|dq| {dq:.3g} m2/s, worst budget term {term:.3g}, '
             f'stored volume {volume:.3g} (relative to {scale:.6g} m3)')
        assert term <= TERM_BAR, (t, term)
        assert volume <= 2e-5, (t, volume)
        if t <= 1200.:
            assert dh <= 1e-6, (t, dh)
        else:
            assert peak <= 1e-2 and flooded <= 1e-2, (t, peak, flooded)
    assert fc[-1][3]['rainfall_m3'] > 0
    _say(f'[precision] plain 256 x 256: steps gpu={sim_g.steps} fp64={sim_c.steps}')
