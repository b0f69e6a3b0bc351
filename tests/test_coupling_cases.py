"""Pipe coupling on the host: exchange, window pairing, node mapping and coupled runs (no GPU).

Known answers are the reference's (``tests/test_coupling.py:62-103,258-300``); the coupled runs
restate ``tests/test_coupling.py:180-256`` with the storage stand-in in place of SWMM (pyswmm is
not installed here), checking the surface/pipe pairing, the budget closure and the donor limit.
"""
import json
import math

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import solver_compact as solver
from nh_flood_2d_amd.coupling import (ORIFICE_COEFFICIENT, PipeConfig, PipeCoupling, PipeTopology, StorageActor,
                                      compute_exchange, map_nodes, parse_inp)
import cases


@pytest.fixture(autouse=True)
def oracle_engine(monkeypatch):
    monkeypatch.setattr(solver, 'ENGINE_FACTORY', oracle.engine_factory('f32'))


def topo(names, outfall, primary, topo_ei, topo_ptr, rims):
    return PipeTopology(list(names), np.array(outfall, bool), np.zeros(len(names)), np.zeros(len(names)),
                        np.array(primary), np.array(topo_ptr), np.array(topo_ei), np.array(rims, float))


def test_shared_cells_and_outfall_conserve_exchange():
    t = topo(['J1', 'J2', 'O1'], [False, False, True], [1, 1, 2], [2, 2, 2], [0, 2, 3, 3], [6., 6., 0.])
    source, boundaries, diag, rel = compute_exchange(
        5., np.array([0., .25, .25]), np.array([0., 5., 5.]), np.array([0., 100., 100.]), np.array([0, 10, 10]),
        t, {'J1': 8., 'J2': 0., 'O1': 0.})
    assert source.dtype == np.float32
    assert rel[1] == (0, 1) and rel[2] == (0, 1)
    assert boundaries['O1'] == {'level': 5.25, 'flow': 0.0}
    assert diag['surface_pipe_balance_m3_s'] == pytest.approx(0., abs=1e-12)
    assert float(source.sum()) == pytest.approx(-(boundaries['J1']['flow'] + boundaries['J2']['flow']), abs=1e-6)
    # closed forms: drainage C h^1.5 from each linked cell (h A / dt is larger), overflow at J1
    # (head 8 over rim 6 + 0.25 m: C dh^1.5, shared by the two junctions whose primary cell it is
    # only through the h A / dt / 2 cap, which does not bind here) at half strength
    drain = ORIFICE_COEFFICIENT * .25 ** 1.5
    over = ORIFICE_COEFFICIENT * 1.75 ** 1.5 * .5
    assert diag['drainage_m3_s'] == pytest.approx(2 * drain, rel=1e-12)
    assert diag['overflow_m3_s'] == pytest.approx(over, rel=1e-12)
    assert source[1] == np.float32(-drain + over) and source[2] == np.float32(-drain) < 0


def test_limited_drainage_is_returned_equally_to_unique_junctions():
    class Recorder(StorageActor):
        def advance(self, boundaries, seconds):
            self.sent = (boundaries, seconds)
            return super().advance(boundaries, seconds)
    t = topo(['J1', 'J2'], [False, False], [1, 1], [1, 1], [0, 1, 2], [2., 2.])
    actor = Recorder(t, depth_below_rim=2.)
    c = PipeCoupling(t, 5, 0., 5., actor)
    rate = c.begin_window(0., 5., np.array([0., .25]), np.zeros(2), np.array([0., 100.]), np.array([0, 10]))
    assert not hasattr(actor, 'sent')            # the pipe model has not advanced yet
    requested = float(rate[1]) * 5.
    actual = np.array([0., requested * .4])
    c.end_window(actual)
    boundaries, seconds = actor.sent
    assert seconds == 5
    flows = [boundaries[n]['flow'] for n in ('J1', 'J2')]
    assert flows[0] == pytest.approx(flows[1], abs=1e-12)
    assert sum(flows) * 5 == pytest.approx(-actual.sum(), abs=1e-9)
    w = c.metadata()['windows'][0]
    assert w['limited_drainage_return_m3'] == pytest.approx(actual[1] - requested)
    assert w['requested_surface_exchange_m3'] == pytest.approx(requested)
    assert w['applied_surface_exchange_m3'] == pytest.approx(actual[1])
    assert abs(w['surface_swmm_balance_error_m3']) <= 1e-9


def test_window_protocol_rejects_gaps_and_fractions():
    t = topo(['J1'], [False], [1], [], [0, 0], [2.])
    c = PipeCoupling(t, 5, 0., 10., StorageActor(t))
    args = (np.array([0., .1]), np.zeros(2), np.array([0., 100.]), np.array([0, 10]))
    with pytest.raises(ValueError, match='expected 0'):
        c.begin_window(5., 5., *args)
    with pytest.raises(ValueError, match='whole seconds'):
        c.begin_window(0., 2.5, *args)
    with pytest.raises(ValueError, match='past simulation end'):
        c.begin_window(0., 20., *args)
    c.begin_window(0., 5., *args)
    with pytest.raises(RuntimeError, match='end_window must finish'):
        c.begin_window(0., 5., *args)
    with pytest.raises(ValueError, match='Unpaired'):
        c.end_window(np.array([1., 0.]))


INP = """[OPTIONS]
FLOW_UNITS CMS
[JUNCTIONS]
;name elev maxdepth
J1 0 2 0 0 0
J2 -1 2.5 0 0 0
[OUTFALLS]
O1 -2 FREE
[COORDINATES]
J1 5 5
J2 24 5
O1 45 5
"""


def test_inp_parsing_and_node_mapping(tmp_path):
    (tmp_path / 'p.inp').write_text(INP)
    nodes, rims = parse_inp(tmp_path / 'p.inp')
    assert [n[0] for n in nodes] == ['J1', 'J2', 'O1'] and [n[3] for n in nodes] == [False, False, True]
    assert rims == {'J1': 2.0, 'J2': 1.5}
    mesh = cases.grid(3, 1, 10.)                  # centres (5, 5), (15, 5), (25, 5)
    _, widths, _, _ = mesh.geometry()
    primary, ptr, topo_ei = map_nodes(nodes, mesh.x, mesh.y, widths, 10.)
    c1, c3 = cases.cell_at(mesh, 5, 5), cases.cell_at(mesh, 25, 5)
    # O1 lies 20 m beyond the last centre: nearest centre within 2 x 10 m
    assert list(primary) == [c1, c3, c3]
    # weak neighbourhood: centres within 10 m (the primary excluded); none for the outfall
    assert list(ptr) == [0, 1, 2, 2]
    assert list(topo_ei) == [cases.cell_at(mesh, 15, 5), cases.cell_at(mesh, 15, 5)]


def _coupled(tmp_path, depth0, q0=None, duration=10, interval=5, area=10.):
    mesh = cases.grid(3, 3, 10.)
    centre = cases.cell_at(mesh, 15, 15)
    (tmp_path / 'p.inp').write_text('[JUNCTIONS]\nJ1 -2 2 0 0 0\n[OUTFALLS]\nO1 -3 FREE\n'
                                    '[COORDINATES]\nJ1 15 15\nO1 25 25\n')
    pipe = PipeConfig(inp=str(tmp_path / 'p.inp'), pipe_dir=str(tmp_path / 'pipe'), coupling_interval=interval)
    depth = np.zeros(mesh.e_num, np.float32)
    depth[centre] = depth0
    q = np.zeros(mesh.s_num, np.float32) if q0 is None else q0(mesh)
    restart = cases.seed(tmp_path, mesh, depth, q=q)
    rain = cases.write_rain(tmp_path, np.zeros((1, 1, 1), np.float32), 0., 30., 30., [(0, duration)])
    sent = []

    class Actor(StorageActor):
        def advance(self, boundaries, seconds):
            sent.append((boundaries, seconds))
            return super().advance(boundaries, seconds)
    import nh_flood_2d_amd.solver_compact as s
    old = s.PIPE_ACTOR_FACTORY
    s.PIPE_ACTOR_FACTORY = lambda cfg, t, begin, end: Actor(t, area=area)
    try:
        row, manifest, _ = cases.run(tmp_path, mesh, duration=duration, rain=rain, restart=restart,
                                     land={10: dict(manning_n=1e-6, infiltration='none')}, pipe=pipe)
    finally:
        s.PIPE_ACTOR_FACTORY = old
    return row, manifest, sent


def test_coupled_run_pairs_surface_and_pipe_volumes(tmp_path):
    row, manifest, sent = _coupled(tmp_path, .5)
    windows = manifest['coupling']['windows']
    assert manifest['coupling']['window_count'] == 2 and manifest['coupling']['swmm_cumulative_seconds'] == 10
    applied = sum(w['applied_surface_exchange_m3'] for w in windows)
    assert applied < 0                                        # the junction drained the centre cell
    assert row['pipe_exchange_m3'] == pytest.approx(applied, abs=1e-6)
    pipe_in = sum(sum(v['flow'] for v in b.values()) * s for b, s in sent)
    assert pipe_in == pytest.approx(-row['pipe_exchange_m3'], abs=1e-6)
    assert abs(row['residual_m3']) < 1e-5
    for w in windows:
        assert abs(w['surface_swmm_balance_error_m3']) <= 1e-9


def test_donor_limited_drainage_returns_to_the_pipe_model(tmp_path):
    """A shallow cell draining to the junction and into all four neighbours at once: the donor
    limit scales the applied drainage below the planned one, and the unapplied part goes back
    to the junction inflow (restates tests/test_coupling.py:180-256)."""
    def q0(mesh):
        q = np.zeros(mesh.s_num, np.float32)
        for x, y, v in ((10, 15, -1.), (20, 15, 1.), (15, 10, -1.), (15, 20, 1.)):
            q[cases.side_at(mesh, x, y)] = v
        return q
    row, manifest, sent = _coupled(tmp_path, .1, q0=q0, duration=5, interval=5)
    w = manifest['coupling']['windows'][0]
    assert w['limited_drainage_return_m3'] > 0
    assert w['requested_surface_exchange_m3'] < w['applied_surface_exchange_m3'] < 0
    assert row['pipe_exchange_m3'] == pytest.approx(w['applied_surface_exchange_m3'], abs=1e-6)
    b, s = sent[0]
    assert sum(v['flow'] for v in b.values()) * s == pytest.approx(-row['pipe_exchange_m3'], abs=1e-6)
    assert abs(row['residual_m3']) < 1e-5
