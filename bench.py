"""Benchmark of the multislice forward + adjoint path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload C3|C2|C1] [--batch B] [--no-autotune] [--gpus 1]

Default workload (C3, the largest single-GPU configuration): a 128^3 Gaussian-atom volume at 0.4 A in a
256 x 256 field (256 x 256 x 128 voxels), probe 256 x 256 at 300 kV / 21 mrad, 21 tilts from -60 to +60
deg, each tilt with its own square scan of side drawn from 24..48 (2 px steps + U[-0.5, 0.5] px jitter),
1.6 A slices -> 32 propagation planes, Poisson data at 5e4 e/A^2; synthetic and seeded.  `--workload C2`
runs the 128 x 128 configuration (64^3 volume, 32 x 32 scans, 11 tilts, 16 planes), `--workload C1` the small
CPU-runnable case (24^3 volume in a 64 x 64 field, 8 x 8 scans, 3 tilts, 6 planes; its CPU sample is a whole tilt).

One step = one full update pass of the reconstruction over every tilt (Run_MSET_recon.m:119-187): for
each tilt the inverse Fourier-shear rotation of the volume, forward + adjoint for all of the tilt's
positions with the object, probe and position updates, the forward rotation and positivity.  The step
sizes form all three gradient terms on every position, as the reference does (STEP03_BACK.m:26-38); the
object-only pass of the reference's example (main.m:42, step [100, 0, 0]) is timed and reported beside
it.  Inputs are resident in HBM before timing starts.  The headline path runs positions in mini-batches
of B per launch on a persistent grid, replayed from a captured hipGraph; B and the grid are picked by the
autotuner (mset_amd/autotune.py) before timing.  The plain path (one position per launch, sequential
updates, the reference's semantics) is timed on one tilt of the same workload and reported beside it.

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
    'C1': 'C1: 24^3 volume at 0.5 A padded to 64x64x24, probe 64x64 300 kV 25 mrad, 8x8 scan at 0.8 A, 3 tilts '
          '-30/0/30 deg, 6 planes (2 A), one update pass over all tilts per step',
    'C2': 'C2: 64^3 volume at 0.4 A padded to 128x128x64, probe 128x128 300 kV 21 mrad, 32x32 scan '
          '(2 px + U[-0.5,0.5]), 11 tilts -50..50 deg, 16 planes (1.6 A), one update pass over all tilts per step',
    'C3': 'C3: 128^3 volume at 0.4 A padded to 256x256x128, probe 256x256 300 kV 21 mrad, 21 tilts -60..60 deg, '
          'each tilt a square scan of side U{24..48} (2 px + U[-0.5,0.5]), 32 planes (1.6 A), one update pass '
          'over all tilts per step',
}
# object, probe and position step sizes of the timed pass.  The object step is the example's (main.m:42);
# the probe and position steps are the largest of a sweep on five C3 tilts (probe 1e-3 / 1e-2, positions
# 1e6 / 1e7) that kept the error finite and falling (1e-2 diverged; 1e6 moved positions by thousands of
# pixels), the position step cut to keep moves well under a pixel.  Nonzero, they make the kernels form
# the probe and position terms on every position, as the reference does.
STEPS_ALL = (100.0, 1e-3, 10.0)
STEPS_OBJECT = (100.0, 0.0, 0.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Cores this process may use: its CPU affinity, capped by OMP_NUM_THREADS when that is set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = os.environ.get('OMP_NUM_THREADS')
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(wl, tilt, positions, threads, steps):
    """The oracle's single-precision restatement of the MATLAB CPU path (Run_MSET_recon.m:134-158 with
    STEP02_FORWARD_4D / STEP03_BACK: sequential positions, whole-volume transmission per position), timed
    on `positions` positions of one tilt with `threads` OpenMP threads."""
    from oracle import oracle as O
    used = O.set_threads(threads)
    bprop = np.conj(wl.prop)
    V0 = wl.truth * 0.0
    t0 = time.perf_counter()
    O.run_tilt(V0, wl.probe, wl.prop, bprop, wl.xlists[tilt][:positions], wl.ylists[tilt][:positions],
               wl.measured[tilt][:positions], positions, 1, wl.binning, wl.sigma, wl.dx, steps, single=True)
    dt = time.perf_counter() - t0
    return dict(value=positions * wl.L / dt, unit=UNIT, cores=used, seconds=dt,
                sample='%d positions of tilt %d (%d planes each), sequential updates, FP32, %d thread(s), %.1f s'
                       % (positions, tilt + 1, wl.L, used, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='C3', choices=['C1', 'C2', 'C3'],
                    help='C3 (default): 256^2 probe, 21 tilts with uneven scans; C2: 128^2 probe, 11 tilts; C1: the '
                         'small CPU-runnable case (64^2 probe, 3 tilts of 64 positions)')
    ap.add_argument('--batch', type=int, default=None,
                    help='positions per mini-batch (default: picked by the autotuner from 256 and 512 for C3, '
                         'the range of batch sizes that workload names, and from 512 and 1024 for C2)')
    ap.add_argument('--no-autotune', action='store_true',
                    help='skip the autotuner: batch 512 (C3) / 1024 (C2) on every resident workgroup')
    ap.add_argument('--no-graph', action='store_true')
    ap.add_argument('--object-only-steps', type=int, default=3,
                    help='timed passes with the object step only (0 = skip)')
    ap.add_argument('--plain-tilts', type=int, default=1, help='tilts of the plain path to time (0 = skip); the '
                    'smallest tilts are taken')
    ap.add_argument('--cpu-positions', type=int, default=None,
                    help='positions of the all-cores CPU sample (0 = skip both CPU samples; default 120 for C3, '
                         '1024 for C2, about 10 s each); the one-thread sample takes a third of them')
    ap.add_argument('--cpu-threads', type=int, default=None, help='threads of the all-cores sample (default: the '
                    'cores this process may use)')
    args = ap.parse_args()
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (each box has one GPU); '
              'run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    from mset_amd import lib, physics, synth
    big = args.workload == 'C3'
    small = args.workload == 'C1'
    batches = (256, 512) if big else (16, 64) if small else (512, 1024)
    if args.batch is not None:
        batches = (args.batch,)
    elif args.no_autotune:
        batches = (batches[-1],)
    if args.cpu_positions is None:
        # C1 in full (a whole tilt of 64 positions); C2 and C3 a sample of about 10 s
        args.cpu_positions = 120 if big else 64 if small else 1024
    if args.cpu_threads is None:
        args.cpu_threads = host_threads()

    t_setup = time.perf_counter()
    log('generating workload %s ...' % args.workload)
    wl = synth.Workload(args.workload)
    ctx = lib.Context(0)
    wl.simulate(ctx)
    ctx.close()
    ntilt = len(wl.angles)
    rec0 = np.zeros_like(wl.truth)
    mean_I = sum(float(m.sum()) for m in wl.measured) / wl.positions()
    probe0 = physics.normalise_probe(wl.probe.astype(np.complex128), mean_I).astype(np.complex64)
    log('workload ready: %d tilts, %d positions, %.1f s' % (ntilt, wl.positions(), time.perf_counter() - t_setup))

    def make_recon(tilts, steps):
        r = lib.Recon(rec0, wl.prop, len(tilts), wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps)
        for i, j in enumerate(tilts):
            r.set_tilt(i, wl.angles[j], probe0, wl.xlists[j], wl.ylists[j], wl.measured[j], n_scan=wl.n_scans[j])
        return r

    units = wl.pos_slices()
    use_graph = not args.no_graph
    r = make_recon(range(ntilt), STEPS_ALL)
    tuned = None
    engine = lib.ENGINE_AUTO
    if not args.no_autotune:
        from mset_amd import autotune
        full = lib.batch_slots(wl.P)
        tuned = autotune.tune_recon(r, batches=batches, slots=[0, full // 2], log=log)
        args.batch = tuned['batch']
        engine = tuned['engine']

        def slots_name(sl):
            return 'all resident' if sl == 0 else str(sl)
        tuned = dict(engine=autotune.engine_name(tuned['engine']), batch=tuned['batch'],
                     slots=slots_name(tuned['slots']),
                     pass_ms={'%s / slots %s / batch %d' % (autotune.engine_name(e), slots_name(sl), b): round(dt, 2)
                              for e, sl, b, dt, _ in tuned['table']})
    else:
        args.batch = batches[0]
    log('batched path: batch %d, graph %s, steps %s' % (args.batch, use_graph, STEPS_ALL))

    def timed(nsteps, nwarm, marks=None):
        # every iterate call returns after its pass has finished on the device (it synchronises the stream),
        # so the time between two returns is one step
        for _ in range(nwarm):
            r.iterate(lib.PATH_BATCHED, args.batch, use_graph=use_graph, want_error=False)
        t0 = time.perf_counter()
        if marks is not None:
            marks.append(t0)
        for _ in range(nsteps):
            r.iterate(lib.PATH_BATCHED, args.batch, use_graph=use_graph, want_error=False)
            if marks is not None:
                marks.append(time.perf_counter())
        return time.perf_counter() - t0

    marks = []
    dt = timed(args.steps, args.warmup, marks)
    ms_step = dt / args.steps * 1e3
    step_ms = np.diff(marks) * 1e3
    value = units * args.steps / dt
    err_after = float(np.mean(r.error_pass()))
    log('timed: %.2f ms per step, %.0f pos-slices/s' % (ms_step, value))

    # dominant kernel timed with HIP events on the reconstruction's stream (no graph while profiling)
    r.profile(True)
    r.iterate(lib.PATH_BATCHED, args.batch, use_graph=False, want_error=False)
    nb, ms_b = r.profile_read('position_batch')
    nr, ms_r = r.profile_read('rotate')
    r.profile(False)

    object_only = None
    if args.object_only_steps > 0:
        r.set_steps(STEPS_OBJECT)
        dto = timed(args.object_only_steps, 1)
        object_only = dict(step_size=list(STEPS_OBJECT), steps=args.object_only_steps,
                           ms_per_step=dto / args.object_only_steps * 1e3, value=units * args.object_only_steps / dto)
    r.close()

    per_launch_units = units / nb
    avg_ms = ms_b / nb
    b_alg = 32.0 * wl.P * wl.P            # bytes per probe-position-slice (store-exit-wave formulation)
    achieved = b_alg * per_launch_units / (avg_ms * 1e-3) / 1e9
    # A 256 x 256 wave does not fit on chip: every pass of the plane engine over it reads and writes the plane,
    # 16 P^2 bytes.  Passes per position: 2L + 4 with the separable-propagator routine (one per propagation
    # step each way), 4L + 3 with the general 2D routine.
    off_chip = None
    if wl.P == 256:
        passes = (4 * wl.L + 3) if (engine & lib.ENGINE_GENERAL) else (2 * wl.L + 4)
        b_off = b_alg + 16.0 * wl.P * wl.P * passes / wl.L
        a_off = b_off * per_launch_units / (avg_ms * 1e-3) / 1e9
        off_chip = dict(plane_passes_per_unit=passes / wl.L, bytes_per_unit=b_off, achieved=a_off,
                        frac=a_off / HBM_PEAK_GBS)
    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE summary of this kernel and workload:
    # its measured bytes per probe-position-slice times this run's average units per launch (the profiled
    # launches have the same batch size when the summary's command is the bench's own choice)
    traffic = None
    tpath = os.path.join(ROOT, 'profiles', 'traffic_position_batch_%s.json' % args.workload.lower())
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        if tj.get('bytes_per_unit'):
            traffic = tj['bytes_per_unit'] * per_launch_units

    plain = None
    if args.plain_tilts > 0:
        small = sorted(range(ntilt), key=lambda j: wl.npos_of(j))[:args.plain_tilts]
        log('plain path: tilt(s) %s' % [j + 1 for j in small])
        rp = make_recon(small, STEPS_ALL)
        rp.profile(True)
        t0 = time.perf_counter()
        rp.iterate(lib.PATH_PLAIN, 1, use_graph=False, want_error=False)
        dtp = time.perf_counter() - t0
        npl, ms_pl = rp.profile_read('position_plain')
        rp.close()
        pvalue = sum(wl.npos_of(j) for j in small) * wl.L / dtp
        plain = dict(value=pvalue, unit=UNIT, tilts_timed=[j + 1 for j in small], step_size=list(STEPS_ALL),
                     kernel_avg_ms=ms_pl / max(npl, 1), speedup_of_batched=value / pvalue)

    cpu = None
    if args.cpu_positions > 0:
        tilt = ntilt // 2
        one = max(1, args.cpu_positions // 3)
        log('cpu baseline: %d positions on 1 thread, %d on %d threads' % (one, args.cpu_positions, args.cpu_threads))
        c1 = cpu_baseline(wl, tilt, one, 1, STEPS_ALL)
        ca = cpu_baseline(wl, tilt, args.cpu_positions, args.cpu_threads, STEPS_ALL)
        cpu = dict(value=ca['value'], unit=UNIT, cores=ca['cores'], kind='port', sample=ca['sample'],
                   nproc=os.cpu_count(), single_thread=dict(value=c1['value'], cores=1, sample=c1['sample']))

    out = dict(
        metric=METRIC, value=value, unit=UNIT, n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms_step, ms_per_step_median=float(np.median(step_ms)),
        ms_per_step_range=[float(step_ms.min()), float(step_ms.max())], higher_is_better=True, scaling='weak', vs_baseline=None, dtype='f32',
        data='synthetic: seeded FCC Gaussian-atom potential, device-simulated Poisson patterns',
        config=dict(workload=WORKLOADS[args.workload],
                    path='batched', precision='complex64 waves, f64 loss and position-term sums', batch=args.batch, hip_graph=use_graph, positions_per_step=wl.positions(),
                    tilts=ntilt, planes=wl.L, probe=wl.P, pos_slices_per_step=units, step_size=list(STEPS_ALL),
                    autotune=tuned, error_after=err_after, rotate_ms_per_step=ms_r),
        roofline=dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                      traffic=traffic, kernel='position_batch (k_batch<%d>)' % wl.P, kernel_avg_ms=avg_ms,
                      units_per_launch=per_launch_units, bytes_per_unit=b_alg,
                      with_plane_passes=off_chip),
        cpu_baseline=cpu, plain=plain, object_only=object_only,
    )
    print(json.dumps(out))


if __name__ == '__main__':
    main()
