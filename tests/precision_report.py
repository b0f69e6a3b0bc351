"""One simulated hour of the benchmark configuration from its dry start: the F32 HIP step against
an FP64 run of the same kernels (the oracle built with double arithmetic, OpenMP on the host).

The configuration is the 1.58 M-cell 2/4/8 m mesh with four Horton land classes, the triangular
hyetograph peaking at 80 mm/h at 1 h and the tide boundary. Both runs go through the solver's host
protocol (Simulation) with adaptive dt, outputs every --every seconds. At each output the report
gives max |dh| (and the 99.9th percentile over cells), the difference in the frame's maximum
depth and flooded area (cells at or above min_h), max |dq|, the worst budget term's error and the
stored volume's, both relative to the budget's volume scale (initial storage plus inputs).

The FP64 run takes several minutes of host time; it is a measurement, not a test (pytest does not
collect this file). Run on a machine with an MI355X and the built library:
    python tests/precision_report.py --out precision_c2_hour.md
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / 'tests'))


def run(sc, factory, end, every, label):
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.simulation import RunSettings, Simulation
    frames = []
    t0 = time.perf_counter()

    class Grab:
        def emit(self, sim, seconds, phase, frame, final, depth, row, stats):
            frames.append((seconds, depth.astype(np.float32).copy(), sim.engine.read_state('q'), dict(row)))
            print(f'[{label}] t={seconds:.0f} s, {sim.steps} steps, {time.perf_counter() - t0:.1f} s wall',
                  flush=True)

        def finish(self, sim):
            pass

    settings = RunSettings(afa=sc.afa, sita=sc.sita, min_h=sc.min_h, end_time=end, yield_step=every)
    sim = Simulation(sc.mesh, sc.land_parameters, synth.scenario_forcing(sc), settings, engine_factory=factory,
                     outputs=Grab())
    sim.run()
    return sim, frames


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--end', type=float, default=3600.)
    ap.add_argument('--every', type=float, default=300.)
    ap.add_argument('--out', default='')
    args = ap.parse_args()
    import oracle
    from nh_flood_2d_amd import synth
    from nh_flood_2d_amd.engine import Engine, TorchDevice
    sc = synth.multires_channel(0)
    device = TorchDevice()

    def gpu(layout, params, state, **kw):
        e = Engine(layout, params, state, device=device, **kw)
        e.set_graph(True)
        return e
    sim_g, fg = run(sc, gpu, args.end, args.every, 'f32 gpu')
    sim_c, fc = run(sc, oracle.engine_factory('f64'), args.end, args.every, 'fp64 oracle')
    _, _, _, area = sc.mesh.geometry()
    lines = [f'# F32 HIP step against FP64, {sc.mesh.e_num - 1} cells, 0-{args.end:.0f} s from the dry start',
             '',
             f'Steps: F32 HIP {sim_g.steps}, FP64 oracle {sim_c.steps}. Relative figures are over the budget\'s '
             'volume scale (initial storage plus rain and boundary inflow).',
             '',
             '| t (s) | max depth (m) | max abs dh (m) | 99.9 % of cells within (m) | max depth diff (m) | '
             'flooded area diff | max abs dq (m2/s) | worst budget term | its error | stored volume error |',
             '| ---: | ---: | ---: | ---: | ---: | ---: | ---: | --- | ---: | ---: |']
    for (t, hg, qg, rg), (_, hc, qc, rc) in zip(fg, fc):
        hg, hc = hg.astype(np.float64), hc.astype(np.float64)
        d = np.abs(hg - hc)
        scale = max(rc['initial_storage_m3'] + rc['rainfall_m3'] +
                    sum(abs(v) for k, v in rc.items() if k.startswith('level:') or k.startswith('discharge:')), 1.)
        terms = {k: abs(rg[k] - rc[k]) / scale for k, v in rc.items() if k.endswith('_m3') and
                 k not in ('initial_storage_m3', 'storage_m3', 'residual_m3') and isinstance(v, float)}
        worst = max(terms, key=terms.get)
        wet_g, wet_c = float(area[hg >= sc.min_h].sum()), float(area[hc >= sc.min_h].sum())
        lines.append(f'| {t:.0f} | {hc.max():.4f} | {d.max():.3g} | {np.quantile(d[1:], 0.999):.3g} | '
                     f'{abs(hg.max() - hc.max()):.3g} | {abs(wet_g - wet_c) / max(wet_c, 1.):.3g} | '
                     f'{np.max(np.abs(qg.astype(np.float64) - qc.astype(np.float64))):.3g} | {worst} | '
                     f'{terms[worst]:.3g} | {abs(rg["storage_m3"] - rc["storage_m3"]) / scale:.3g} |')
    text = '\n'.join(lines) + '\n'
    print(text, flush=True)
    if args.out:
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        Path(args.out).write_text(text)


if __name__ == '__main__':
    main()
