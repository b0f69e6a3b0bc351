"""Pipe coupling with the HIP step (needs an MI355X): same windows, sources and volumes as the
CPU oracle. The per-cell compensated pipe volume is device code; the exchange is host code fed
from the device state, so the two runs agree bit for bit in state and window volumes, and in the
budget rows up to the FP64 summation grouping."""
import math

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import synth
from nh_flood_2d_amd.coupling import PipeCoupling, PipeTopology, StorageActor, map_nodes
from nh_flood_2d_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _run(sc, topology, factory):
    coupling = PipeCoupling(topology, 60, 0., sc.duration, StorageActor(topology, area=20.))
    sim = synth.simulation(sc, engine_factory=factory, yield_step=120., coupling=coupling)
    row = sim.run()
    return row, sim.state(), sim.record['coupling'], sim.record['dt_sequence_sha256']


def test_coupled_run_matches_oracle(gpu_device):
    sc = synth.small_multires(11, duration=360., pond=0.4)
    _, widths, _, _ = sc.mesh.geometry()
    low = np.argsort(sc.mesh.z[1:])[:3] + 1
    nodes = [(f'J{i}', float(sc.mesh.x[c]), float(sc.mesh.y[c]), False) for i, c in enumerate(low)]
    nodes.append(('O1', float(sc.mesh.x[low[0]]) + 30., float(sc.mesh.y[low[0]]), True))
    primary, ptr, topo_ei = map_nodes(nodes, sc.mesh.x, sc.mesh.y, widths, 8.)
    topology = PipeTopology([n[0] for n in nodes], np.array([n[3] for n in nodes]), np.zeros(4), np.zeros(4),
                            primary, ptr, topo_ei, np.array([float(sc.mesh.z[c]) + .1 for c in low] + [0.]))
    print(f'[coupling] small_multires seed 11, junction cells {low.tolist()}', flush=True)

    def hip(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=gpu_device, pipe_q=pipe_q)
    g = _run(sc, topology, hip)
    c = _run(sc, topology, oracle.engine_factory('f32'))
    assert g[3] == c[3]
    for name in g[1]:
        np.testing.assert_array_equal(g[1][name], c[1][name], err_msg=name)
    assert g[2]['window_count'] == c[2]['window_count'] == 6
    for wg, wc in zip(g[2]['windows'], c[2]['windows']):
        assert wg['applied_surface_exchange_m3'] == wc['applied_surface_exchange_m3']
        assert wg['applied_surface_exchange_m3'] < 0
    for k, x in c[0].items():
        y = g[0][k]
        if isinstance(x, float) and (math.isfinite(x) or math.isfinite(y)):
            assert abs(x - y) <= 1e-9 + 1e-9 * max(abs(x), abs(y)), (k, x, y)
