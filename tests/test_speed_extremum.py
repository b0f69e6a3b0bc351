"""Pins of the oracle's velocity extremum (``max_speed``).

The reference updates ``max_speed = max(max_speed, sqrt(u*u + v*v))`` at the end of every step
(``solver_compact.py:606-607``) with ``u = (eq_l + eq_r) / height / max(h, 0.01) / 2`` and ``v``
likewise over the bottom/top face totals, both zero below ``min_h`` (``:355-362``). Two
independent checks hold the oracle's C code to that:

* step by step, on meshes with hanging faces, ponds, rain and boundaries, the oracle's
  ``max_speed`` equals the running maximum of the host restatement ``simulation.velocities``
  (numpy, its own code), bit for bit; the same comparison with the ``/2`` dropped or with the
  division by the floored depth dropped must fail, so the check tells those mistakes apart;
* in the reference's steady coarse/fine channel (``tests/test_face_flux.py:245-271``), the
  extremum over a short formal window after a long warmup equals the closed-form uniform-flow
  speed ``q/H = H^(2/3) S^(1/2) / n`` with ``H = 1 m``.
"""
import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import abi, synth
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.clock import seconds_to_ticks
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, initial_state
from nh_flood_2d_amd.output import read_checkpoint
from nh_flood_2d_amd.simulation import velocities
from nh_flood_2d_amd.synth import scenario_forcing

import cases
from parity import coefficients_at


class OracleRun:
    """One scenario on the F32 oracle, stepped with a forced dt."""

    def __init__(self, sc):
        self.sc = sc
        self.forcing = scenario_forcing(sc)
        self.layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, self.forcing),
                                   sc.sita, 'morton')
        state = initial_state(sc.mesh, self.layout)
        if sc.initial_depth is not None:
            state['depth'] = sc.initial_depth.astype(np.float32).copy()
        self.engine = Engine(self.layout, make_params(sc.afa, sc.sita, sc.min_h, 32), state,
                             device=oracle.HostDevice('f32'))
        self.engine.reset(0)
        self.engine.compute_cfl()
        self.now = 0

    def step(self, dt):
        e, t0 = self.engine, self.now / 2 ** 40
        rain = self.forcing.rain
        rates = rain.rates(rain.index_at(t0)).astype(np.float32)
        coeff = (coefficients_at(self.forcing, t0).ravel()[:2 * len(self.forcing.boundaries)]
                 if self.forcing.boundaries else np.zeros(0, np.float32))
        e.set_forcing(coeff, rates, True)
        step = min(dt, 0.5 * e.cfl_state()[0])
        e.begin(seconds_to_ticks(t0 + step), self.now)
        row = e.run_batch(1)[0]
        assert row[3] == abi.COMPLETED, row
        self.now = (int(row[1]) << 32) | int(row[2])


def _host_speed(mesh, depth, q, min_h, variant='reference'):
    u, v = velocities(mesh, depth, q, min_h)
    if variant == 'no_half':
        u, v = u * np.float32(2), v * np.float32(2)
    elif variant == 'no_depth':
        h = np.maximum(depth, np.float32(0.01))
        u, v = (u * h).astype(np.float32), (v * h).astype(np.float32)
    speed = np.sqrt(u * u + v * v).astype(np.float32)
    speed[0] = 0
    return speed


@pytest.mark.parametrize('make', [lambda: synth.small_multires(1, pond=0.3, discharge=2.0),
                                  lambda: synth.small_multires(2, pond=0.3, discharge=2.0, theta=0.7),
                                  lambda: synth.plain_uniform(3, n=32, dx=10.)],
                         ids=['multires_theta1', 'multires_theta07', 'plain32'])
def test_oracle_max_speed_is_running_max_of_host_velocities(make):
    sc = make()
    if sc.initial_depth is None:      # the plain grid starts dry: pond its lowest quarter
        depth = np.zeros(sc.mesh.e_num, np.float32)
        low = sc.mesh.z[1:] <= np.quantile(sc.mesh.z[1:], 0.25)
        depth[1:][low] = 0.5
        sc.initial_depth = depth
    run = OracleRun(sc)
    e = run.engine
    previous = e.read_state('max_speed')
    seen_moving = 0
    mistakes = {'no_half': 0, 'no_depth': 0}
    for k in range(40):
        run.step(0.05 if 'multires' in sc.name else 0.5)
        depth, q, got = e.read_state('depth'), e.read_state('q'), e.read_state('max_speed')
        speed = _host_speed(sc.mesh, depth, q, sc.min_h)
        expected = np.maximum(previous, speed)
        np.testing.assert_array_equal(got.view(np.uint32), expected.view(np.uint32), err_msg=f'step {k}')
        seen_moving += int(np.count_nonzero(speed > 0))
        for variant in mistakes:
            wrong = np.maximum(previous, _host_speed(sc.mesh, depth, q, sc.min_h, variant))
            mistakes[variant] += int(np.count_nonzero(wrong != got))
        previous = got
    assert seen_moving > 0
    assert mistakes['no_half'] > 0 and mistakes['no_depth'] > 0   # both mistakes would be seen


def test_steady_channel_extremum_equals_uniform_flow_speed(tmp_path, monkeypatch):
    """The reference's coarse/fine channel: after a 4000 s warmup the flow is steady and uniform,
    so the extremum over the 300 s formal window is q/H = sqrt(S)/n at H = 1 m everywhere away
    from the two boundary cells."""
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))
    slope, warmup, duration = 0.002, 4000, 300
    length = 1024
    half = length // 2
    cells = [(x, x + 2, 0, 2) for x in range(0, half, 2)]
    cells += [(x, x + 1, y, y + 1) for x in range(half, length) for y in (0, 1)]
    mesh = cases.boxes_mesh(cells, bed=lambda x, y, i: 10 - slope * x)
    expected_q = np.sqrt(slope) / 0.03
    total = warmup + duration
    rain = cases.write_rain(tmp_path, np.zeros((1, 1, 1), np.float32), 0., 2., 1024., [(-warmup, duration)])
    bnd = [cases.write_boundary(tmp_path, 'upstream', 'discharge', [[0, .1], [0, 1.9]],
                                [(-warmup, 2 * expected_q), (duration, 2 * expected_q)]),
           cases.write_boundary(tmp_path, 'downstream', 'level', [[1024, .1], [1024, 1.9]],
                                [(-warmup, float(mesh.z[-1] + 1)), (duration, float(mesh.z[-1] + 1))])]
    depth = np.ones(mesh.e_num, np.float32)
    depth[0] = 0
    seed = cases.seed(tmp_path, mesh, depth)
    _, _, dc = cases.run(tmp_path, mesh, duration=duration, warmup=warmup, rain=rain, boundaries=bnd,
                         restart=seed, yield_step=duration)
    assert total == 4300
    state = read_checkpoint(dc.output_dir + '/state_final.fdb')
    x = mesh.x[1:]
    interior = (x > 100) & (x < 924)
    speed = state.max_speed[1:][interior]
    h = state.depth[1:][interior]
    assert np.all(np.abs(h - 1) < 0.02)
    np.testing.assert_allclose(speed, expected_q, rtol=3e-3)
    # the maximum over the window is at least the final speed, and the flow was steady in it
    final = _host_speed(mesh, state.depth, state.q, 0.02)[1:][interior]
    assert np.all(speed >= final)
    assert np.max(speed - final) < 1e-4 * expected_q
