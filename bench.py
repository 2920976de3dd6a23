"""Benchmark of the multislice forward + adjoint path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload C2|C3] [--batch B] [--no-autotune] [--gpus 1]

Workload: the GPU configuration of the reference's plain method (64^3 volume at 0.4 A in a 128 x 128
field, probe 128 x 128 at 300 kV / 21 mrad, 32 x 32 scan with sub-pixel jitter, 11 tilts from -50 to
+50 deg, 1.6 A slices -> 16 propagation planes, Poisson data at 5e4 e/A^2), synthetic and seeded.
One step = one full update pass of the reconstruction over all 11 tilts: for each tilt the inverse
Fourier-shear rotation of the volume, forward + adjoint for all 1024 positions with the object update,
the forward rotation and positivity (Run_MSET_recon.m:119-187).  Inputs are resident in HBM before
timing starts.  The headline path runs positions in mini-batches of B per launch (one workgroup per
position, up to one per CU in flight) replayed from a captured hipGraph; B and the persistent grid are
picked by the autotuner (mset_amd/autotune.py) before timing.  The plain path (one position per launch,
sequential updates, the reference's semantics) is measured on the same workload and reported beside it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'probe-position·slices/sec per fwd+adjoint iteration; HBM GB/s; p50/p99 latency'
UNIT = 'probe-position·slices/s'
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    'C2': 'C2: 64^3 volume at 0.4 A padded to 128x128x64, probe 128x128 300 kV 21 mrad, 32x32 scan '
          '(2 px + U[-0.5,0.5]), 11 tilts -50..50 deg, 16 planes (1.6 A), one update pass over all tilts per step',
    'C3': 'C3: 128^3 volume at 0.4 A padded to 256x256x128, probe 256x256 300 kV 21 mrad, 21 tilts -60..60 deg, '
          'each tilt a square scan of side U{24..48} (2 px + U[-0.5,0.5]), 32 planes (1.6 A), one update pass '
          'over all tilts per step',
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, positions, threads):
    """The oracle's single-precision restatement of the MATLAB CPU path (Run_MSET_recon.m:134-158 with
    STEP02_FORWARD_4D / STEP03_BACK: sequential positions, whole-volume transmission per position),
    timed on a bounded sample of the same workload."""
    from oracle import oracle as O
    used = O.set_threads(threads)
    bprop = np.conj(wl.prop)
    V0 = wl.truth * 0.0
    t0 = time.perf_counter()
    O.run_tilt(V0, wl.probe, wl.prop, bprop, wl.xlists[0][:positions], wl.ylists[0][:positions],
               wl.measured[0][:positions], positions, 1, wl.binning, wl.sigma, wl.dx, (100.0, 0.0, 0.0), single=True)
    dt = time.perf_counter() - t0
    return dict(value=positions * wl.L / dt, unit=UNIT, cores=used, kind='port',
                sample='%d positions of tilt 1 (%d planes each), sequential updates, FP32, %.1f s' %
                       (positions, wl.L, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', default='C2', choices=['C2', 'C3'],
                    help='C2 (default): the configuration the metric is quoted on; C3: 256^2 probe, uneven scans')
    ap.add_argument('--batch', type=int, default=None,
                    help='positions per mini-batch (default: picked by the autotuner from 256..1024 for C2, '
                         '64..512 for C3)')
    ap.add_argument('--no-autotune', action='store_true',
                    help='skip the autotuner: batch 1024 (C2) / 512 (C3) on every resident workgroup')
    ap.add_argument('--no-graph', action='store_true')
    ap.add_argument('--plain-tilts', type=int, default=None, help='tilts of the plain path to time (0 = skip; '
                    'default 1 for C2, 0 for C3)')
    ap.add_argument('--cpu-positions', type=int, default=None,
                    help='positions of the CPU sample (0 = skip; default one full tilt for C2, 32 for C3)')
    ap.add_argument('--cpu-threads', type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (each box has one GPU); '
              'run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    from mset_amd import lib, synth
    big = args.workload == 'C3'
    batches = (64, 128, 256, 512) if big else (256, 512, 1024)
    if args.batch is not None:
        batches = (args.batch,)
    elif args.no_autotune:
        batches = (max(batches),)
    if args.plain_tilts is None:
        args.plain_tilts = 0 if big else 1
    if args.cpu_positions is None:
        args.cpu_positions = 32 if big else 1024

    steps = (100.0, 0.0, 0.0)     # object step only, as the reference's example configuration (main.m:42)
    log('generating workload %s ...' % args.workload)
    wl = synth.Workload(args.workload)
    ctx = lib.Context(0)
    wl.simulate(ctx)
    ctx.close()
    ntilt = len(wl.angles)
    rec0 = np.zeros_like(wl.truth)
    mean_I = sum(float(m.sum()) for m in wl.measured) / wl.positions()
    from mset_amd import physics
    probe0 = physics.normalise_probe(wl.probe.astype(np.complex128), mean_I).astype(np.complex64)

    def make_recon():
        r = lib.Recon(rec0, wl.prop, ntilt, wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps)
        for j, ang in enumerate(wl.angles):
            r.set_tilt(j, ang, probe0, wl.xlists[j], wl.ylists[j], wl.measured[j], n_scan=wl.n_scans[j])
        return r

    units = wl.pos_slices()
    use_graph = not args.no_graph
    r = make_recon()
    tuned = None
    if not args.no_autotune:
        from mset_amd import autotune
        tuned = autotune.tune_recon(r, batches=batches, log=log)
        args.batch = tuned['batch']
        tuned = dict(batch=tuned['batch'], slots=tuned['slots'],
                     pass_ms={'%d/%d' % (sl, b): round(dt, 2) for sl, b, dt, _ in tuned['table']})
    else:
        args.batch = batches[0]
    log('batched path: batch %d, graph %s' % (args.batch, use_graph))
    for _ in range(args.warmup):
        r.iterate(lib.PATH_BATCHED, args.batch, use_graph=use_graph, want_error=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.iterate(lib.PATH_BATCHED, args.batch, use_graph=use_graph, want_error=False)
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = units * args.steps / dt
    err_after = float(np.mean(r.error_pass()))

    # dominant kernel timed with HIP events on the reconstruction's stream (no graph while profiling)
    r.profile(True)
    r.iterate(lib.PATH_BATCHED, args.batch, use_graph=False, want_error=False)
    nb, ms_b = r.profile_read('position_batch')
    nr, ms_r = r.profile_read('rotate')
    r.profile(False)
    r.close()
    per_launch_units = units / nb
    avg_ms = ms_b / nb
    b_alg = 32.0 * wl.P * wl.P            # bytes per probe-position-slice (store-exit-wave formulation)
    achieved = b_alg * per_launch_units / (avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, 'profiles', 'traffic_position_batch_%s.json' % args.workload.lower())
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get('bytes_per_launch')

    plain = None
    if args.plain_tilts > 0:
        log('plain path: %d tilt(s)' % args.plain_tilts)
        rp = lib.Recon(rec0, wl.prop, args.plain_tilts, wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps)
        for j in range(args.plain_tilts):
            rp.set_tilt(j, wl.angles[j], probe0, wl.xlists[j], wl.ylists[j], wl.measured[j], n_scan=wl.n_scans[j])
        rp.profile(True)
        t0 = time.perf_counter()
        rp.iterate(lib.PATH_PLAIN, 1, use_graph=False, want_error=False)
        dtp = time.perf_counter() - t0
        npl, ms_pl = rp.profile_read('position_plain')
        rp.close()
        pvalue = sum(wl.npos_of(j) for j in range(args.plain_tilts)) * wl.L / dtp
        plain = dict(value=pvalue, unit=UNIT, tilts_timed=args.plain_tilts,
                     kernel_avg_ms=ms_pl / max(npl, 1), speedup_of_batched=value / pvalue)

    cpu = None
    if args.cpu_positions > 0:
        log('cpu baseline: %d positions' % args.cpu_positions)
        cpu = cpu_baseline(wl, args.cpu_positions, args.cpu_threads)

    out = dict(
        metric=METRIC, value=value, unit=UNIT, n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms_step, higher_is_better=True, scaling='weak', vs_baseline=None, dtype='f32 (complex64 waves)',
        data='synthetic: seeded FCC Gaussian-atom potential, device-simulated Poisson patterns',
        config=dict(workload=WORKLOADS[args.workload],
                    path='batched', batch=args.batch, hip_graph=use_graph, positions_per_step=wl.positions(),
                    planes=wl.L, pos_slices_per_step=units, step_size=list(steps), autotune=tuned,
                    error_after=err_after, rotate_ms_per_step=ms_r),
        roofline=dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                      traffic=traffic, kernel='position_batch', kernel_avg_ms=avg_ms,
                      units_per_launch=per_launch_units, bytes_per_unit=b_alg),
        cpu_baseline=cpu, plain=plain,
    )
    print(json.dumps(out))


if __name__ == '__main__':
    main()
