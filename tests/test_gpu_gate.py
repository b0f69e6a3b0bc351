"""Gate control with the HIP step (needs an MI355X) against the CPU oracle through the solver.

The device runs the trigger test, the gate-bed motion and the per-step accounting; the host runs
the rest. On the same inputs the two runs take the same steps (dt sequence), end in the same state
bit for bit (depth, q and the moving bed included), report the same command and opening times,
and their gate budgets agree to the FP64 summation grouping (1e-9 relative).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.engine import Engine
import test_gate_cases as gc

pytestmark = pytest.mark.gpu


def _close(a, b, path=''):
    if isinstance(a, dict):
        assert set(a) == set(b), path
        for k in a:
            _close(a[k], b[k], f'{path}.{k}')
    elif isinstance(a, list):
        assert len(a) == len(b), path
        for i, (x, y) in enumerate(zip(a, b)):
            _close(x, y, f'{path}[{i}]')
    elif isinstance(a, float) and (math.isfinite(a) or math.isfinite(b)):
        assert abs(a - b) <= 1e-9 * max(1., abs(a), abs(b)), (path, a, b)
    else:
        assert a == b, (path, a, b)


@pytest.mark.parametrize('case', ['channel_and_pond', 'short_pulse'])
def test_gate_run_matches_oracle(gpu_device, tmp_path, monkeypatch, case):
    print(f'[gate] {case}', flush=True)
    if case == 'channel_and_pond':
        domain, force = gc.channel_and_pond(tmp_path / 'in')
        seconds = 160
    else:
        (tmp_path / 'in').mkdir()
        boundaries = gc.level_boundary(tmp_path / 'in', [(0, 1), (.125, 2), (.25, 1), (3600, 1)])
        domain, force, *_ = gc.setup(tmp_path / 'in', bed=[0, 10, 0], depth=[1, 0, 0],
                                     gate=dict(trigger_level=1.5, opening_seconds=1.875, open_elevation=0),
                                     boundaries=boundaries)
        seconds = 2

    def hip(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
    runs = {}
    for name, factory in (('gpu', hip), ('cpu', oracle.engine_factory('f32'))):
        monkeypatch.setattr(solver, 'ENGINE_FACTORY', factory)
        runs[name] = gc.run(domain, force, seconds, output=tmp_path / name)
    (gs, gm), (cs, cm) = runs['gpu'], runs['cpu']
    assert gm['dt_sequence_sha256'] == cm['dt_sequence_sha256']
    for field in ('depth', 'bed', 'q', 'max_depth', 'first_time_hi', 'first_time_lo'):
        assert getattr(gs, field).tobytes() == getattr(cs, field).tobytes(), field
    g, c = gm['gate_control'], cm['gate_control']
    for key in ('command_seconds', 'first_effective_flow_seconds', 'fully_open_seconds',
                'first_fully_open_sample_seconds', 'command_clock_ticks'):
        assert [x[key] for x in g['gates']] == [x[key] for x in c['gates']], key
    _close(g, c)
    assert json.loads(gs.metadata_json)['gate_controller']['commanded'] == [1]
    rows = [(tmp_path / n / 'gate_balance.csv').read_text().splitlines() for n in ('gpu', 'cpu')]
    assert len(rows[0]) == len(rows[1]) > 2
