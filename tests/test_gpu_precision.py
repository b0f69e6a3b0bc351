"""The benchmark's 1.58 M-cell multi-resolution mesh: the HIP step against the F32 oracle (bit for
bit) and its error against an FP64 run of the same kernels (the oracle built with double
arithmetic).

Both start from the benchmark's wet state at the peak of the rain and take 8 steps, each forced to
the same dt (half the smaller CFL bound). Bars: max |dh| 1e-4 m, the flushed infiltration budget
within 1e-7 relative. The stored volume is the F32 depths summed, so its difference (F32 rounding
of about 2 m of water per cell) is only held to 1e-6.
"""
import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import synth
from nh_flood_2d_amd.clock import seconds_to_ticks
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import build_layout, initial_state
from parity import coefficients_at

pytestmark = pytest.mark.gpu


def _benchmark_engines(gpu_device, precision):
    """The benchmark mesh's wet state at the rain peak (3600 s) on the HIP engine and on the
    oracle with `precision` arithmetic, forced and with their first CFL bound computed."""
    sc = synth.multires_channel(0)
    forcing = synth.scenario_forcing(sc)
    layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
    state = initial_state(sc.mesh, layout)
    state['depth'][1:] = np.maximum(np.float32(2.0) - sc.mesh.z[1:], np.float32(0.01))
    params = make_params(sc.afa, sc.sita, sc.min_h, 32)
    gpu = Engine(layout, params, state, device=gpu_device)
    cpu = Engine(layout, params, state, device=oracle.HostDevice(precision))
    now = seconds_to_ticks(3600.)
    coeff = coefficients_at(forcing, 3600.)[:layout.b_num].ravel()
    rates = forcing.rain.rates(forcing.rain.index_at(3600.)).astype(np.float32)
    for e in (gpu, cpu):
        e.set_forcing(coeff, rates, True)
        e.reset(now)
        e.compute_cfl()
    return sc, gpu, cpu, now


def test_benchmark_mesh_matches_the_f32_oracle(gpu_device):
    """The benchmark's whole step at its full size, 32 adaptive steps in one batch: the same trace
    (dt bits, clock words, limiter counts) and the same state bits as the F32 oracle."""
    sc, gpu, f32, now = _benchmark_engines(gpu_device, 'f32')
    print(f'[full size] {sc.mesh.e_num - 1} cells against the F32 oracle', flush=True)
    rows = []
    for e in (gpu, f32):
        e.begin(now + seconds_to_ticks(600.), now)
        rows.append(e.run_batch(32))
    np.testing.assert_array_equal(rows[0], rows[1])
    assert np.all(rows[0][:, 3] == 1)
    for name in ('depth', 'q', 'max_depth', 'max_speed', 'capacity', 'rain_time'):
        a, b = gpu.read_state(name), f32.read_state(name)
        assert a.tobytes() == b.tobytes(), name
    np.testing.assert_allclose(gpu.flush_budget().sum(axis=1), f32.flush_budget().sum(axis=1), rtol=1e-9, atol=1e-9)
    gpu.close()
    f32.close()


def test_benchmark_mesh_against_fp64(gpu_device):
    sc, gpu, f64, now = _benchmark_engines(gpu_device, 'f64')
    print(f'[precision] {sc.mesh.e_num - 1} cells against FP64', flush=True)
    for _ in range(8):
        target = now + seconds_to_ticks(0.5 * min(e.cfl_state()[0] for e in (gpu, f64)))
        rows = []
        for e in (gpu, f64):
            e.begin(target, now)
            rows.append(e.run_batch(1))
        assert all(r[0, 3] == 1 for r in rows)
        now = (int(rows[0][0, 1]) << 32) | int(rows[0][0, 2])
    infil = [float(e.flush_budget()[0].sum()) for e in (gpu, f64)]
    _, _, _, area = sc.mesh.geometry()
    hg, hc = (e.read_state('depth').astype(np.float64) for e in (gpu, f64))
    qg, qc = (e.read_state('q').astype(np.float64) for e in (gpu, f64))
    dh, dq = float(np.max(np.abs(hg - hc))), float(np.max(np.abs(qg - qc)))
    volume = abs(np.sum(hg * area) - np.sum(hc * area)) / np.sum(hc * area)
    budget = abs(infil[0] - infil[1]) / abs(infil[1])
    print(f'[precision] max|dh| {dh:.3g} m, max|dq| {dq:.3g} m2/s, volume {volume:.3g}, '
          f'infiltration {budget:.3g} (relative)', flush=True)
    assert dh <= 1e-4 and budget <= 1e-7 and volume <= 1e-6 and infil[1] > 0
    gpu.close()
    f64.close()
