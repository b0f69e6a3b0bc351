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
