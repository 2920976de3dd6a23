"""Stored, recomputed and checkpointed exit waves at 256^3 (the memory case of the adjoint verification config).

    python -m mset_amd.exitwaves [--positions 256] [--repeat 2]

Volume 256^3 at 0.4 A (FCC Gaussian atoms, 8 % occupancy), probe 256 x 256 at 300 kV / 21 mrad, slices
of one voxel (0.4 A) so 256 propagation planes, a 16 x 16 scan at 0.8 A with sub-pixel jitter, Poisson
data simulated on the device.  For each variant the batched forward + gradient at the fixed state runs
over every position; the line reports the per-slot and total scratch (waves and gradient slabs of the
resident positions), the kernel time and the gradient difference between the variants.  At 256 x 256 the
stored and the checkpointed variants run the separable-propagator routine, the recomputing one the general
routine (it inverts the 2D step); `store_general` is the stored variant on the general routine.  The
checkpointed waves are the forward passes rerun, so they are compared with the stored ones of the same
routine.  Prints one JSON line.
"""
import argparse
import json
import sys

import numpy as np

from . import lib, synth


def make_workload(positions=256, seed=4):
    """The 256^3 workload with a square scan of about `positions` positions and its simulated data."""
    side = int(round(np.sqrt(positions)))
    synth.CONFIGS['C4-256'] = dict(vol=256, P=256, dx=0.4, kV=300.0, alpha=21.0, defocus=0.0, scan=(side, side),
                                   step_A=0.8, jitter=0.5, tilts=[0.0], slice_A=0.4, dose=1e5, occupancy=0.08)
    wl = synth.Workload('C4-256', seed=seed)
    ctx = lib.Context(0)
    wl.simulate(ctx)
    ctx.close()
    return wl


def run(positions=256, repeat=2, seed=4, wl=None):
    if wl is None:
        wl = make_workload(positions, seed)
    xl, yl, M = wl.xlists[0], wl.ylists[0], wl.measured[0]
    V0 = 0.5 * wl.truth
    out = dict(workload='256^3 volume at 0.4 A, probe 256^2, %d planes, %d positions, 1 tilt' % (wl.L, len(xl)),
               variants={})
    grads = {}
    planes = {lib.EXIT_WAVES_STORE: wl.L + 2, lib.EXIT_WAVES_RECOMPUTE: 5,
              lib.EXIT_WAVES_CHECKPOINT: checkpoint_planes(wl.L)}
    for name, mode, engine in (('store', lib.EXIT_WAVES_STORE, lib.ENGINE_AUTO),
                               ('store_general', lib.EXIT_WAVES_STORE, lib.ENGINE_GENERAL),
                               ('recompute', lib.EXIT_WAVES_RECOMPUTE, lib.ENGINE_AUTO),
                               ('checkpoint', lib.EXIT_WAVES_CHECKPOINT, lib.ENGINE_AUTO)):
        ctx = lib.Context(0)
        try:
            ctx.upload(V0, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning,
                       n_scan=wl.n_scans[0])
            ctx.set_exit_waves(mode)
            ctx.set_engine(engine)
            g = ctx.grad()                                  # sizes the scratch
            ctx.profile(True)
            for _ in range(repeat):
                g = ctx.grad()
            n, ms = ctx.profile_read('position_batch')
            resolved, slot_bytes = ctx.exit_waves()
            slots = lib.batch_slots(wl.P)
        finally:
            ctx.close()
        grads[name] = g
        out['variants'][name] = dict(
            resolved=resolved, slots=min(slots, len(xl)), scratch_bytes_per_slot=slot_bytes,
            wave_bytes_per_slot=planes[resolved] * wl.P * wl.P * 8,
            scratch_gb=min(slots, len(xl)) * slot_bytes / 1e9,
            kernel_ms=ms / max(n, 1), pos_slices_per_s=len(xl) * wl.L / (ms / max(n, 1) * 1e-3))
    def diff(r, s):
        return dict(dV_rel=float(np.linalg.norm(r[0] - s[0]) / np.linalg.norm(s[0])),
                    dprobe_rel=float(np.linalg.norm(r[1] - s[1]) / np.linalg.norm(s[1])),
                    loss_rel=float(abs(r[3] / s[3] - 1)), dV_bitwise=bool(np.array_equal(r[0], s[0])))
    out['recompute_vs_store'] = diff(grads['recompute'], grads['store'])
    out['checkpoint_vs_store'] = diff(grads['checkpoint'], grads['store'])
    vs, vr, vc = out['variants']['store'], out['variants']['recompute'], out['variants']['checkpoint']
    out['wave_memory_ratio'] = vs['wave_bytes_per_slot'] / vr['wave_bytes_per_slot']
    out['scratch_memory_ratio'] = vs['scratch_bytes_per_slot'] / vr['scratch_bytes_per_slot']
    out['time_ratio'] = vr['kernel_ms'] / vs['kernel_ms']
    out['checkpoint_wave_memory_ratio'] = vs['wave_bytes_per_slot'] / vc['wave_bytes_per_slot']
    out['checkpoint_time_ratio'] = vc['kernel_ms'] / vs['kernel_ms']
    out['recompute_time_vs_store_general'] = vr['kernel_ms'] / out['variants']['store_general']['kernel_ms']
    return out


def checkpoint_planes(L):
    """Per-slot wave planes of the checkpoint variant: psi0, Fourier psi0, L // K checkpoints, the K - 1 other
    waves of a segment and the parked adjoint wave, K = ceil(sqrt(L + 1))."""
    k = 2
    while k * k < L + 1:
        k += 1
    return L // k + k + 3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--positions', type=int, default=256)
    ap.add_argument('--repeat', type=int, default=2)
    a = ap.parse_args()
    print(json.dumps(run(a.positions, a.repeat)))
    return 0


if __name__ == '__main__':
    sys.exit(main())
