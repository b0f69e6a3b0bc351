"""HIP step vs the CPU oracle on the same seeded inputs (needs an MI355X).

Tolerances are the ones the step is specified to meet: per step in fixed-dt mode
|dh| <= 1e-5 m and |dq| <= 1e-5 m^2/s; in adaptive-dt mode at every output 1e-4 m max depth and
0.1 % flooded area. The clock is integer arithmetic and must match bit for bit.
"""
import numpy as np
import pytest

from nh_flood_2d_amd import abi, synth
from nh_flood_2d_amd.simulation import RunSettings, Simulation

import oracle
from parity import Pair

pytestmark = pytest.mark.gpu

H_TOL = 1e-5
Q_TOL = 1e-5


def _scenarios():
    return {
        'plain64': lambda: synth.plain_uniform(3, n=64, dx=10.),
        'multires_theta1': lambda: synth.small_multires(1, pond=0.3, discharge=2.0),
        'multires_theta07': lambda: synth.small_multires(2, pond=0.3, discharge=2.0, theta=0.7),
    }


def _say(*args):
    print(*args, flush=True)


def test_clock_probe_matches_oracle(gpu_device):
    rng = np.random.default_rng(1234)
    n = 4096
    now = rng.integers(0, 1 << 62, size=n, dtype=np.uint64)
    now[:8] = [0, 169199 * 2 ** 40 + 879609302220, (1 << 32) - 219902326, 47 * 3600 * 2 ** 40,
               47 * 3600 * 2 ** 40 - 2 ** 38, 5, 169199 * 2 ** 40, 1 << 40]
    gap = rng.integers(0, 1 << 42, size=n, dtype=np.uint64)
    gap[:8] = [2 ** 40, 219902325555, 549755814, 0, 2 ** 38, 2 ** 40, 2 ** 40 - 1, 7]
    target = now + gap
    origin = now - rng.integers(0, 1 << 40, size=n, dtype=np.uint64) % np.maximum(now, 1)
    cand = rng.uniform(0, 2, size=n).astype(np.float32)
    cand[:8] = [0.2, 0.25, 0.0005, 0.25, 5e-5, np.nextafter(np.float32(2 ** -20), np.float32(1)), 0.2, 1.0]
    cand[100:200] = np.float32(np.inf)
    cand[200:300] = (rng.uniform(0, 1, 100) * 1e-6).astype(np.float32)
    out = {}
    for name, lib in (('gpu', gpu_device.lib), ('oracle', oracle.load('f32'))):
        words = np.zeros((n, 4), np.uint32)
        ticks = np.zeros((n, 2), np.uint64)
        times = np.zeros((n, 2), np.float32)
        rc = lib.nhf_clock_probe(now.ctypes.data, target.ctypes.data, origin.ctypes.data, cand.ctypes.data,
                                 np.float32(1e-4), n, words.ctypes.data, ticks.ctypes.data, times.ctypes.data)
        assert rc == 0, lib.nhf_last_error()
        out[name] = (words, ticks, times)
    for a, b in zip(out['gpu'], out['oracle']):
        np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize('name', list(_scenarios()))
def test_fixed_dt_steps_match_oracle(gpu_device, name):
    sc = _scenarios()[name]()
    _say(f'[fixed-dt] {name}: {sc.mesh.e_num - 1} cells, {sc.mesh.s_num - 1} sides')
    pair = Pair(sc, gpu_device)
    dt = 0.05 if 'multires' in name else 0.5
    worst_h = worst_q = 0.
    for step in range(60):
        g, c = pair.fixed_step(dt)
        assert g[0, 3] == c[0, 3] == abi.COMPLETED, (step, g, c)
        np.testing.assert_array_equal(g[0, :4], c[0, :4])
        np.testing.assert_array_equal(g[0, 6:8], c[0, 6:8])
        dh, dq = pair.diff('depth'), pair.diff('q')
        worst_h, worst_q = max(worst_h, dh), max(worst_q, dq)
        assert dh <= H_TOL and dq <= Q_TOL, (name, step, dh, dq)
        # every operation is IEEE and in the same order: the bits agree, not just the values
        for state in ('depth', 'q'):
            np.testing.assert_array_equal(pair.gpu.read_state(state).view(np.uint32),
                                          pair.cpu.read_state(state).view(np.uint32), err_msg=f'{state} step {step}')
        if step % 15 == 7:   # the device defers rising cells' extrema: reading settles them mid-run
            np.testing.assert_array_equal(pair.gpu.read_state('max_depth').view(np.uint32),
                                          pair.cpu.read_state('max_depth').view(np.uint32), err_msg=f'step {step}')
            np.testing.assert_array_equal(pair.gpu.read_state('peak_tick'), pair.cpu.read_state('peak_tick'),
                                          err_msg=f'peak_tick step {step}')
    np.testing.assert_array_equal(pair.gpu.read_state('peak_tick'), pair.cpu.read_state('peak_tick'))
    for state, tol in (('capacity', 1e-9), ('rain_time', 1e-6), ('max_depth', H_TOL), ('max_speed', 1e-5)):
        assert pair.diff(state) <= tol, (state, pair.diff(state))
        # the same operations in the same order: equal bits (max_speed's formula is pinned on
        # the CPU against the host velocities and the uniform-flow speed)
        np.testing.assert_array_equal(pair.gpu.read_state(state).view(np.uint32),
                                      pair.cpu.read_state(state).view(np.uint32), err_msg=state)
    # the clock fields and the water budget
    bg, bc = pair.gpu.flush_budget(), pair.cpu.flush_budget()
    tg, tc = bg.sum(axis=1), bc.sum(axis=1)
    scale = max(np.abs(tc).max(), 1.)
    assert np.max(np.abs(tg - tc)) <= 1e-5 * scale, (tg, tc)
    _say(f'[fixed-dt] {name}: max|dh|={worst_h:.3g} max|dq|={worst_q:.3g}')


def _simulate(sc, factory, end=600., yield_step=150.):
    forcing = synth.scenario_forcing(sc)
    from nh_flood_2d_amd.output import seed_checkpoint
    seed = None
    if sc.initial_depth is not None:
        seed = seed_checkpoint(sc.initial_depth, sc.mesh.z, np.zeros(sc.mesh.s_num, np.float32))
    settings = RunSettings(afa=sc.afa, sita=sc.sita, min_h=sc.min_h, end_time=end, yield_step=yield_step)
    frames = []

    class Grab:
        def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
            frames.append((seconds, depth.copy(), dict(row)))

        def finish(self, sim):
            pass

    sim = Simulation(sc.mesh, sc.land_parameters, forcing, settings, checkpoint=seed,
                     engine_factory=factory, outputs=Grab())
    sim.run()
    return sim, frames


@pytest.mark.parametrize('name', ['plain64', 'multires_theta1'])
def test_adaptive_run_frames_match_oracle(gpu_device, name):
    sc = _scenarios()[name]()
    _say(f'[adaptive] {name}')
    from nh_flood_2d_amd.engine import Engine

    def gpu_factory(layout, params, state):
        return Engine(layout, params, state, device=gpu_device)

    sim_g, fg = _simulate(sc, gpu_factory)
    sim_c, fc = _simulate(sc, oracle.engine_factory('f32'))
    assert [f[0] for f in fg] == [f[0] for f in fc]
    assert sim_g.steps == sim_c.steps
    assert sim_g.record['dt_sequence_sha256'] == sim_c.record['dt_sequence_sha256']
    _, _, _, area = sc.mesh.geometry()
    for (t, dg, rg), (_, dc, rc) in zip(fg, fc):
        assert np.max(np.abs(dg - dc)) <= 1e-4, (t, np.max(np.abs(dg - dc)))
        wet_g = area[dg > 0.02].sum()
        wet_c = area[dc > 0.02].sum()
        assert abs(wet_g - wet_c) <= 1e-3 * max(wet_c, 1.), (t, wet_g, wet_c)
        assert abs(rg['storage_m3'] - rc['storage_m3']) <= 1e-5 * max(rc['storage_m3'], 1.)
        np.testing.assert_array_equal(dg.view(np.uint32), dc.view(np.uint32))
        # budget terms: FP64 block sums on the GPU vs exact per-cell sums in the oracle
        for key in ('infiltration_m3', 'bed_clamp_nonlevel_m3', 'residual_m3'):
            assert abs(rg[key] - rc[key]) <= 1e-9 * max(abs(rc['storage_m3']), 1.), (key, rg[key], rc[key])
    _say(f'[adaptive] {name}: steps gpu={sim_g.steps} oracle={sim_c.steps}')


def test_plain_config_full_run_matches_oracle(gpu_device):
    """The CPU-runnable configuration at full size and duration: 256 x 256 cells of 10 m, 50 mm/h
    for 1800 s on a dry start, closed boundaries, 3600 s with a frame every 300 s. The adaptive-dt
    bar is 1e-4 m and 0.1 % flooded area per frame; the HIP step gives the oracle's bits."""
    from nh_flood_2d_amd.engine import Engine
    sc = synth.plain_uniform(0, n=256, dx=10.)
    _say('[adaptive] plain 256 x 256, 3600 s')

    def gpu_factory(layout, params, state):
        return Engine(layout, params, state, device=gpu_device)
    sim_g, fg = _simulate(sc, gpu_factory, end=3600., yield_step=300.)
    sim_c, fc = _simulate(sc, oracle.engine_factory('f32'), end=3600., yield_step=300.)
    assert [f[0] for f in fg] == [f[0] for f in fc] and len(fg) == 13
    assert sim_g.record['dt_sequence_sha256'] == sim_c.record['dt_sequence_sha256']
    _, _, _, area = sc.mesh.geometry()
    for (t, dg, rg), (_, dc, rc) in zip(fg, fc):
        assert np.max(np.abs(dg - dc)) <= 1e-4, t
        assert abs(area[dg > 0.02].sum() - area[dc > 0.02].sum()) <= 1e-3 * max(area[dc > 0.02].sum(), 1.), t
        np.testing.assert_array_equal(dg.view(np.uint32), dc.view(np.uint32))
        assert abs(rg['residual_m3'] - rc['residual_m3']) <= 1e-9 * max(abs(rc['storage_m3']), 1.), t
    assert fg[-1][1].max() > 0.1      # the rain has ponded in the channel and behind the hills
    _say(f'[adaptive] plain 256 x 256: {sim_g.steps} steps, wet area {area[fg[-1][1] > 0.02].sum():.0f} m2')


def test_graph_replays_match_eager(gpu_device):
    """A captured batch replayed many times gives the eager launches' bits at every replay."""
    sc = synth.small_multires(5, pond=0.3, discharge=1.0)
    _say('[graph] small_multires seed 5')
    eager, graph = Pair(sc, gpu_device), Pair(sc, gpu_device)
    graph.gpu.set_graph(True)
    for k in range(8):
        re = eager.adaptive_batch(15. * (k + 1), engines=(eager.gpu,))[0]
        rg = graph.adaptive_batch(15. * (k + 1), engines=(graph.gpu,))[0]
        np.testing.assert_array_equal(re, rg)
        for name in ('depth', 'q', 'max_depth', 'capacity'):
            a, b = eager.gpu.read_state(name), graph.gpu.read_state(name)
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f'{name} replay {k}')


@pytest.mark.parametrize('cfl_blocks,cell_block,cells_per_thread', [(512, 128, 1), (4096, 512, 1), (2048, 256, 2),
                                                                    (1024, 128, 2)])
def test_launch_tuning_keeps_the_bits(gpu_device, cfl_blocks, cell_block, cells_per_thread):
    """Every workgroup size and CFL grid the autotuner may pick computes the same state, trace
    and budget as the default launch shapes."""
    from nh_flood_2d_amd.engine import Engine
    sc = synth.small_multires(9, pond=0.3, discharge=1.0)
    _say(f'[tuning] small_multires seed 9, CFL grid {cfl_blocks}, cell workgroup {cell_block} x {cells_per_thread}')
    runs = []
    for tuned in (False, True):
        pair = Pair(sc, gpu_device)
        if tuned:
            top = pair.gpu.tuning_range()['cfl_max']
            pair.gpu.set_tuning(min(cfl_blocks, top - top % 8), cell_block, cells_per_thread)
        rows = [pair.fixed_step(0.05)[0] for _ in range(12)]
        runs.append((np.array(rows), pair.gpu.read_state('depth'), pair.gpu.read_state('q'),
                     pair.gpu.flush_budget()))
    (ra, da, qa, ba), (rb, db, qb, bb) = runs
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(qa, qb)
    np.testing.assert_allclose(ba.sum(axis=1), bb.sum(axis=1), rtol=1e-12, atol=1e-12)


def test_stored_and_computed_cell_sizes_give_the_same_bits(gpu_device):
    """A mesh whose sizes are not base * 2^level makes the kernels read the size array; on a mesh
    where both work, both give the same state, trace and budget."""
    sc = synth.small_multires(5, pond=0.3, discharge=1.0)
    _say('[sizes] small_multires seed 5: computed vs stored cell sizes')
    runs = []
    for stored in (False, True):
        pair = Pair(sc, gpu_device, stored_sizes=stored)
        assert (pair.layout.size_base == 0.0) == stored
        rows = [pair.fixed_step(0.05)[0] for _ in range(12)]
        assert pair.diff('depth') <= H_TOL and pair.diff('q') <= Q_TOL
        runs.append((np.array(rows), pair.gpu.read_state('depth'), pair.gpu.read_state('q'),
                     pair.gpu.read_state('max_speed'), pair.gpu.flush_budget()))
    (ra, da, qa, sa, ba), (rb, db, qb, sb, bb) = runs
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))
    np.testing.assert_array_equal(qa.view(np.uint32), qb.view(np.uint32))
    np.testing.assert_array_equal(sa.view(np.uint32), sb.view(np.uint32))
    np.testing.assert_allclose(ba.sum(axis=1), bb.sum(axis=1), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('theta', [1.0, 0.7])
def test_donor_limiting_matches_oracle(gpu_device, theta):
    """A thin sheet of water on the multi-resolution mesh's slopes drains faster than some cells
    hold, so the donor limiter scales their outflows. The device looks only at cells the face pass
    flags (one of whose outflows exceeds its certified share of the cell's volume) or that the
    certificate does not cover; the trace's limiter columns, depth and discharge must still be the
    oracle's bits at every step, and the limiter must have acted."""
    sc = synth.small_multires(11, pond=0.04, discharge=1.0, theta=theta)
    sc.initial_depth[1:] = np.where(sc.mesh.z[1:] < 2.0, np.float32(0.03), np.float32(0.)).astype(np.float32)
    pair = Pair(sc, gpu_device)
    limited = faces = 0
    for step in range(50):
        g, c = pair.fixed_step(1.0)
        assert g[0, 3] == c[0, 3] == abi.COMPLETED, (step, g, c)
        np.testing.assert_array_equal(g[0, :9], c[0, :9])
        for state in ('depth', 'q'):
            np.testing.assert_array_equal(pair.gpu.read_state(state).view(np.uint32),
                                          pair.cpu.read_state(state).view(np.uint32), err_msg=f'{state} step {step}')
        limited += int(g[0, 6])
        faces += int(g[0, 7])
    _say(f'[limiter] theta {theta}: {limited} limited cells, {faces} limited faces over 50 steps')
    assert limited > 0 and faces > 0


def test_rain_state_follows_the_written_arrays(gpu_device):
    """With one rain pixel the device keeps one rain time and one capacity per land class while
    every cell agrees (the uniform rain state), and the per-cell arrays otherwise; the oracle
    always keeps the per-cell arrays. Uniform from the cold start, per cell after a rain time that
    differs between cells is written, uniform again after one rain time and one capacity per class
    are written: in every state each batch's trace and every state array, rain time and capacity
    included, are the oracle's bits."""
    sc = synth.small_multires(22, duration=900., discharge=1.0, pond=0.25)
    pair = Pair(sc, gpu_device)
    assert pair.layout.n_pixels == 1

    def advance(batches):
        for _ in range(batches):
            g, c = pair.adaptive_batch(pair.now / 2 ** 40 + 30.)
            np.testing.assert_array_equal(g, c)
        for name in ('depth', 'q', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick'):
            a, b = pair.gpu.read_state(name), pair.cpu.read_state(name)
            np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8), err_msg=name)

    def write_both(name, values):
        for e in pair.engines:
            e.write_state(name, values)
        np.testing.assert_array_equal(pair.gpu.read_state(name), values, err_msg=f'{name} read back')

    _say('[rain state] uniform from the cold start')
    advance(4)
    rt = pair.gpu.read_state('rain_time')
    assert rt[1:].min() == rt[1:].max() > 0
    _say('[rain state] per cell after a rain time that differs between cells')
    uneven = rt.copy()
    uneven[1::3] += np.float32(7.5)
    write_both('rain_time', uneven)
    advance(4)
    assert pair.gpu.read_state('rain_time')[1:].max() > pair.gpu.read_state('rain_time')[1:].min()
    _say('[rain state] uniform again after one rain time and per-class capacities')
    write_both('rain_time', np.full_like(rt, np.float32(40.)))
    cap = pair.layout.cell_f0.astype(np.float32) * np.float32(0.5)
    cap[0] = 0.
    write_both('capacity', cap)
    advance(4)
