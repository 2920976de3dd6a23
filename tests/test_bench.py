"""bench.py's command-line contract: one GPU only, and one JSON line with the metric, the roofline of the
dominant kernel and the CPU baseline."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None, timeout=600):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py')] + list(args), capture_output=True,
                          text=True, timeout=timeout, env=e, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench('--gpus', '2', timeout=120)
    assert r.returncode == 2 and 'one GPU' in r.stderr
    r = _bench(env={'WORLD_SIZE': '4'}, timeout=120)
    assert r.returncode == 2


@pytest.mark.gpu
def test_bench_prints_one_json_line_with_roofline_and_cpu_baseline():
    r = _bench('--workload', 'C2', '--steps', '1', '--warmup', '0', '--no-autotune', '--plain-tilts', '0',
               '--cpu-positions', '4', '--object-only-steps', '1')
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ('metric', 'value', 'unit', 'n_gpus', 'steps', 'warmup', 'ms_per_step', 'higher_is_better',
                'scaling', 'vs_baseline', 'dtype', 'data', 'config', 'roofline', 'cpu_baseline'):
        assert key in d, key
    assert d['n_gpus'] == 1 and d['steps'] == 1 and d['value'] > 0
    assert d['config']['positions_per_step'] == 11 * 1024 and d['config']['planes'] == 16
    rf = d['roofline']
    assert rf['bound'] == 'hbm' and rf['peak'] == 8000.0 and 0 < rf['frac'] < 1
    assert abs(rf['achieved'] / rf['peak'] - rf['frac']) < 1e-9
    cb = d['cpu_baseline']
    assert cb['kind'] == 'port' and cb['cores'] >= 1 and cb['value'] > 0 and cb['nproc'] >= cb['cores']
    assert cb['single_thread']['cores'] == 1 and cb['single_thread']['value'] > 0
    # all three gradient terms are formed in the timed pass; the object-only pass is reported beside it
    assert all(s != 0 for s in d['config']['step_size'])
    assert d['object_only']['step_size'][1:] == [0.0, 0.0] and d['object_only']['value'] > 0
