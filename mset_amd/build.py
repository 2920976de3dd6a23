"""Builds mset_amd/libmset_amd.so from the HIP sources for gfx950 (hipcc, no CMake).

    python -m mset_amd.build          # or __graft_entry__.build()
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ['batch_store.hip', 'batch_variants.hip', 'multislice.hip', 'forward.hip', 'elementwise.hip', 'rotate.hip',
           'tv.hip', 'capi.hip']
LIB = os.path.join(HERE, 'libmset_amd.so')
# stand-alone counter-calibration program (known-size copies for rocprofv3's FETCH_SIZE / WRITE_SIZE)
CALIBRATE = os.path.join(HERE, 'mset_calibrate')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('MSET_OFFLOAD_ARCH', 'gfx950')
# packed FP32 instructions are disabled: some of their forms have produced wrong results on this hardware
FLAGS = ['--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-Xclang', '-target-feature', '-Xclang',
         '-packed-fp32-ops', '-Wno-unused-command-line-argument']


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(HERE, 'csrc', src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = '\n'.join(l for l in (r.stdout + r.stderr).splitlines() if 'packed-fp32-ops' not in l)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, out))
    return out


def build(verbose=True):
    bdir = os.path.join(HERE, '_build')
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, os.path.splitext(s)[0] + '.o') for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        futs = [ex.submit(_compile, s, o) for s, o in zip(SOURCES, objs)]
        for f in futs:
            msg = f.result()
            if verbose and msg.strip():
                print(msg)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    tmp = CALIBRATE + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-o', tmp, os.path.join(HERE, 'csrc', 'calibrate.hip')]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for calibrate.hip:\n' + r.stdout + r.stderr)
    os.replace(tmp, CALIBRATE)
    if verbose:
        print('built', LIB, 'and', CALIBRATE)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
