"""bench.py's contract on the CPU: one GPU only, and the counter summary it commits to profiles/."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused_with_a_message():
    run = subprocess.run([sys.executable, str(ROOT / 'bench.py'), '--gpus', '2'], capture_output=True, text=True,
                         timeout=120)
    assert run.returncode == 2
    assert 'one GPU only' in run.stderr
    assert run.stdout == ''


def test_pmc_summary_scales_counters_by_the_copy_calibration(tmp_path):
    """FETCH_SIZE reads half the bytes of the 256 MiB calibration copy here, so fetch counts scale
    by 2; WRITE_SIZE reads them all (factor 1); kernels are averaged over their dispatches, the
    single-mesh ones over those before the first batch kernel (later ones are other sections')."""
    header = 'Dispatch_Id,Kernel_Name,Grid_Size,Counter_Value\n'
    copy = 256 << 20
    fetch = tmp_path / 'fetch.csv'
    fetch.write_text(header + f'1,copyBuffer,1048576,{copy / 2 / 1024}\n'
                     '3,"void k_flux<2, false>(Dev)",1024,3000\n2,"void k_flux<2, false>(Dev)",1024,1000\n'
                     '4,"void kb_depth<true, 1>(BDev)",2048,500\n5,"void k_flux<2, false>(Dev)",512,9000\n')
    write = tmp_path / 'write.csv'
    write.write_text(header + f'1,copyBuffer,1048576,{copy / 1024}\n2,"void k_flux<2, false>(Dev)",1024,100\n'
                     '3,"void kb_depth<true, 1>(BDev)",2048,50\n4,"void k_flux<2, false>(Dev)",512,700\n')
    result = tmp_path / 'bench.json'
    result.write_text('log line\n' + json.dumps({'config': {'cells': 10, 'sides': 20, 'side_list': 40}}) + '\n')
    run = subprocess.run([sys.executable, str(ROOT / 'bench.py'), '--pmc-summary', str(fetch), str(write),
                          str(result)], capture_output=True, text=True, timeout=120, check=True)
    lines = run.stdout.splitlines()
    assert lines[0] == 'workload cells 10 sides 20 side_list 40'
    assert 'factor 2.0000' in lines[1] and 'factor 1.0000' in lines[2]
    flux = next(ln for ln in lines if ln.startswith('kernel flux '))
    assert f'fetch_bytes {2 * 2000 * 1024} write_bytes {100 * 1024} ' in flux      # mean of 1000 and 3000 KiB
    depth = next(ln for ln in lines if ln.startswith('kernel batch_depth '))
    assert f'hbm_bytes {2 * 500 * 1024 + 50 * 1024}' in depth


def test_l2_summary_averages_the_timing_pass_dispatches(tmp_path):
    """L2 hits and misses per dispatch: each kernel over its last `launches` dispatches (the JSON
    line's kernel counts), single-mesh kernels only before the first batch kernel."""
    rows = [(1, 'void k_depth<128, 1, false>(Dev)', 10, 30), (2, 'void k_depth<128, 1, false>(Dev)', 20, 20),
            (3, 'void kb_depth<true, 1>(BDev)', 5, 15), (4, 'void k_depth<128, 1, false>(Dev)', 99, 1)]
    counters = tmp_path / 'l2.csv'
    counters.write_text('Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n' + ''.join(
        f'{i},"{name}",TCC_HIT_sum,{hit}\n{i},"{name}",TCC_MISS_sum,{miss}\n' for i, name, hit, miss in rows))
    result = tmp_path / 'bench.json'
    result.write_text(json.dumps({'kernels': {'depth': {'launches': 1}},
                                  'batched': {'kernels': {'depth': {'launches': 1}}}}) + '\n')
    run = subprocess.run([sys.executable, str(ROOT / 'bench.py'), '--l2-summary', str(counters), str(result)],
                         capture_output=True, text=True, timeout=120, check=True)
    assert run.stdout.splitlines() == ['kernel depth dispatches 1 l2_hits 20 l2_misses 20 hit_rate 0.500',
                                       'kernel batch_depth dispatches 1 l2_hits 5 l2_misses 15 hit_rate 0.250']


def test_fp64_error_reports_both_runs_and_the_one_ulp_sensitivity():
    """The precision object of the bench line, with the F32 oracle standing in for the HIP engine
    (the same passes): over 10 s of a developed 2:1 mesh with tide and river, F32 against FP64
    differs by rounding only, and every budget term is named."""
    sys.path.insert(0, str(ROOT))
    import bench
    import oracle
    from nh_flood_2d_amd import synth
    sc = synth.small_multires(2, discharge=1.0)
    forcing = synth.scenario_forcing(sc)
    sim = synth.simulation(sc, engine_factory=oracle.engine_factory('f32'), end_time=600., yield_step=600.)
    sim.run()
    p = bench.fp64_error(sc, forcing, sim.state(), 600., oracle.HostDevice('f32'), window=10.)
    assert p['steps_f32'] > 10 and abs(p['steps_f32'] - p['steps_fp64']) <= 1
    assert 0 < p['max_abs_dh_m'] < 1e-5 and p['max_abs_dq_m2_s'] < 1e-5
    assert set(p['budget_terms_rel']) >= {'infiltration', 'floor_nonlevel', 'level0_storage', 'discharge0'}
    assert p['worst_budget_rel'] < 1e-7 and p['stored_volume_rel'] < 1e-6
    assert p['one_ulp_max_abs_dh_m'] > 0
