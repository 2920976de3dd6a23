"""Autotuner of the batched path: picks the engine, the persistent grid and the mini-batch size by timing
them on the workload itself.

Tunables (both leave the kernels' arithmetic unchanged):
  slots  the persistent grid of the batched kernels (mset_set_slot_limit; 0 = every resident workgroup).
         Fewer resident positions hold less scratch (saved waves and, at 256 x 256, the plane buffers) and
         keep more of it in L2 / MALL; more of them hide memory latency better.  Gradient slabs are
         assigned by position, not by workgroup, and a capped grid divides the slab count, so for one
         engine the sums are bitwise the same for every grid: tuning the grid never changes a result.
  batch  positions per launch of an update pass (the mini-batch size), from the range the caller
         allows: it sets how many launches, reductions and updates a pass makes and how well the last
         wave of workgroups fills the device.
  engine mset_set_engine flags: the plane routine at 256 x 256 (the separable-propagator routine, one pass
         over the plane per propagation step, or the general 2D routine, two passes per step) and the
         workgroup shape of the batched kernel at every size (the default shape, or the alternative one:
         threads per position of the LDS engine, two register-engine workgroups per CU at 128,
         16- instead of 32-line tiles at 256).  Engines sum in different orders: results agree to FP32
         rounding.

Each candidate runs one update pass from a snapshot of the tilt series, which is restored afterwards,
so tuning leaves the reconstruction unchanged.  The kernels are timed with HIP events on the series'
stream; the first pass at each grid size (which sizes the scratch) is not timed.  Results are cached per
shape in memory and, if a path is given, in a JSON file.
"""
import json
import os
import time

from . import lib

_CACHE = {}


def slot_candidates(n, fracs=(1.0, 0.5, 0.25), device=0):
    """Grid sizes to try: 0 (every resident workgroup of the engine's shape), then fractions of the default
    shape's resident workgroups (a capped grid is reduced to a divisor of the slab count)."""
    full = lib.batch_slots(n, device)
    out = []
    for f in fracs:
        s = 0 if f >= 1.0 else max(1, int(round(full * f)))
        if s not in out:
            out.append(s)
    return out


def engine_candidates(n):
    """Engines worth timing at plane size n: the default and the alternative workgroup shape, and at 256 the
    general routine (the separable one is the default)."""
    if n == 256:
        return (lib.ENGINE_AUTO, lib.ENGINE_ALT_SHAPE, lib.ENGINE_GENERAL)
    return (lib.ENGINE_AUTO, lib.ENGINE_ALT_SHAPE)


def engine_name(e):
    """readable name of mset_set_engine flags"""
    parts = ['general routine' if e & lib.ENGINE_GENERAL else 'default routine',
             'alternative shape' if e & lib.ENGINE_ALT_SHAPE else 'default shape']
    return ', '.join(parts)


def _key(recon, batches, slots, engines):
    return json.dumps(dict(shape=list(recon.shape), npos=list(recon.npos), batches=list(batches),
                           slots=list(slots), engines=list(engines)), sort_keys=True)


def _apply(recon, best):
    recon.set_engine(best.get('engine', lib.ENGINE_AUTO))
    recon.set_slot_limit(best['slots'])


def tune_recon(recon, batches=(64, 128, 256, 512), slots=None, engines=None, cache_path=None, log=None):
    """Time one batched update pass per (engine, slots, batch) candidate on `recon` and leave the fastest
    engine and grid set on it.  Returns dict(engine, batch, slots, ms, table=[(engine, slots, batch, pass ms,
    position-kernel ms)])."""
    if slots is None:
        slots = slot_candidates(recon.shape[1])
    if engines is None:
        engines = engine_candidates(recon.shape[1])
    key = _key(recon, batches, slots, engines)
    if key in _CACHE:
        best = _CACHE[key]
        _apply(recon, best)
        return best
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            disk = json.load(fh)
        if key in disk:
            _CACHE[key] = disk[key]
            _apply(recon, disk[key])
            return disk[key]
    recon.snapshot()
    table = []
    try:
        for e in engines:
            recon.set_engine(e)
            for s in slots:
                recon.set_slot_limit(s)
                recon.iterate(lib.PATH_BATCHED, max(batches), use_graph=False, want_error=False)   # sizes scratch
                recon.restore()
                for b in batches:
                    recon.profile(True)
                    t0 = time.perf_counter()
                    recon.iterate(lib.PATH_BATCHED, b, use_graph=False, want_error=False)
                    dt = (time.perf_counter() - t0) * 1e3
                    _, kms = recon.profile_read('position_batch')
                    recon.profile_read('rotate')                # consumed here, not left for the caller
                    recon.profile(False)
                    recon.restore()
                    table.append((e, s, b, dt, kms))
                    if log:
                        log('autotune: engine %d slots %d batch %d: pass %.2f ms, position kernels %.2f ms'
                            % (e, s, b, dt, kms))
    finally:
        recon.profile(False)
        recon.restore()
    e, s, b, dt, kms = min(table, key=lambda r: r[3])
    best = dict(engine=e, batch=b, slots=s, ms=dt, table=table)
    _apply(recon, best)
    _CACHE[key] = best
    if cache_path:
        disk = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                disk = json.load(fh)
        disk[key] = best
        with open(cache_path, 'w') as fh:
            json.dump(disk, fh)
    return best


def tune_grad(ctx, n, slots=None, first=0, count=None, repeats=2, log=None):
    """Grid size of the fixed-state gradient (mset_grad) on `ctx`; leaves the fastest set.  Returns
    dict(slots, ms, table=[(slots, kernel ms)])."""
    if slots is None:
        slots = slot_candidates(n)
    table = []
    for s in slots:
        ctx.set_slot_limit(s)
        ctx.grad(first, count)                              # sizes the scratch
        ctx.profile(True)
        for _ in range(repeats):
            ctx.grad(first, count)
        k, ms = ctx.profile_read('position_batch')
        ctx.profile(False)
        table.append((s, ms / max(k, 1)))
        if log:
            log('autotune: slots %d: %.3f ms per launch' % (s, ms / max(k, 1)))
    s, ms = min(table, key=lambda r: r[1])
    ctx.set_slot_limit(s)
    return dict(slots=s, ms=ms, table=table)
