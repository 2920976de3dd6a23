"""GPU parity at the named configurations: C1 (the reference's CPU-runnable case, one update pass over all
tilts), C2 (the plain path and a slot-reused batched tilt against the oracle; full-size repeatability,
additivity and graph replay), C3 (the full series' graph-captured pass against the eager one; FP32 device
gradient against the FP64 oracle at the 256^2, 32-plane shape) and C4 (device gradients against central
finite differences of the FP64 loss along 50 seeded directions, h = 1e-4; stored, recomputed and
checkpointed waves at 256^3)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def _workload(name, ctx, tilts=None):
    from mset_amd import synth
    wl = synth.Workload(name)
    if tilts is not None:
        wl.angles = [wl.angles[j] for j in tilts]
        wl.xlists = [wl.xlists[j] for j in tilts]
        wl.ylists = [wl.ylists[j] for j in tilts]
        wl.n_scans = [wl.n_scans[j] for j in tilts]
    wl.simulate(ctx)
    return wl


def test_c1_update_pass_plain_vs_oracle():
    from mset_amd import lib
    ctx = lib.Context(0)
    wl = _workload('C1', ctx)
    ctx.close()
    steps = (100.0, 0.0, 0.0)
    rec0 = (0.5 * wl.truth).astype(np.float32)
    r = lib.Recon(rec0, wl.prop, len(wl.angles), wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps)
    for j, ang in enumerate(wl.angles):
        r.set_tilt(j, ang, wl.probe, wl.xlists[j], wl.ylists[j], wl.measured[j], n_scan=wl.n_scans[j])
    e = r.iterate(lib.PATH_PLAIN, 1, want_error=True)
    rec = r.get()[0]
    r.close()
    # the oracle composition of Run_MSET_recon.m:119-187
    orec, errs = rec0.astype(np.float64), []
    bprop = np.conj(wl.prop)
    for j, ang in enumerate(wl.angles):
        RV = O.rot_3dvol(orec, ang, inverse=True)
        RV, _, _, _, ej, _ = O.run_tilt(RV, wl.probe, wl.prop, bprop, wl.xlists[j], wl.ylists[j], wl.measured[j],
                                        *wl.n_scans[j], wl.binning, wl.sigma, wl.dx, steps)
        errs.append(ej)
        orec = O.rot_3dvol(RV, ang)
        orec[orec < 0] = 0
    print('C1 rec rel', rel(rec, orec), 'error', e, np.mean(errs), flush=True)
    assert abs(e / np.mean(errs) - 1) < 1e-4
    assert rel(rec - rec0, orec - rec0) < 1e-3


def test_c2_tilt_batched_with_slot_reuse_vs_oracle():
    """One C2 tilt (128 x 128, 16 planes) at a state away from the data (half the true potential): the batched
    gradient over 32 of its positions on 8 slots, four positions per slot, against the FP64 oracle.  An FP32
    chain of single-precision FFTs puts the position terms within ~1.5e-4 of FP64 at this state."""
    from mset_amd import lib
    ctx = lib.Context(0)
    wl = _workload('C2', ctx, tilts=[2])
    k = np.arange(0, wl.npos_of(0), 32)
    V = (0.5 * wl.truth).astype(np.float64)
    xl, yl, M = wl.xlists[0][k], wl.ylists[0][k], wl.measured[0][k]
    ctx.upload(V, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning, n_scan=(len(k), 1))
    ctx.set_slot_limit(8)
    dV, dp, dpos, err = ctx.grad()
    w = ctx.waves(5)
    ctx.close()
    odV, odp, odpos, oep = O.grad_fixed(V, wl.probe, wl.prop, np.conj(wl.prop), xl, yl, M, wl.binning, wl.sigma,
                                        wl.dx)
    psi0 = O.subpixel_shift(wl.probe, xl[5], yl[5])
    save, _, _, _ = O.forward(V, psi0, wl.prop, M[5], wl.binning, wl.sigma, wl.dx)
    r = dict(waves=max(rel(w[a + 1], save[a]) for a in range(wl.L)), loss=abs(err / (oep.sum() / len(k)) - 1),
             dV=rel(dV, odV), dprobe=rel(dp, odp), dpos=rel(dpos, odpos))
    print('C2 (32 positions on 8 slots) FP32 vs FP64 relative errors', r, flush=True)
    assert r['waves'] < 1e-4 and r['loss'] < 1e-4
    assert r['dV'] < 1e-3 and r['dprobe'] < 1e-3 and r['dpos'] < 1e-3


def test_c2_tilt_plain_vs_oracle():
    """The C2 configuration's own path: one position per launch with the volume, probe and positions updated
    after every position (Run_MSET_recon.m:134-158), at the C2 shape (128 x 128, 16 planes), over 24 positions
    spread across a C2 tilt, from half the true potential, with the bench's three nonzero steps, against the
    FP64 oracle's sequential loop."""
    from mset_amd import lib
    from tests.test_gpu_parity import assert_update_close
    ctx = lib.Context(0)
    wl = _workload('C2', ctx, tilts=[2])
    k = np.arange(0, wl.npos_of(0), wl.npos_of(0) // 24)[:24]
    V = (0.5 * wl.truth).astype(np.float64)
    xl, yl, M = wl.xlists[0][k], wl.ylists[0][k], wl.measured[0][k]
    steps = (100.0, 1e-3, 10.0)
    print('case: C2 plain, %d positions, steps %s' % (len(k), steps), flush=True)
    ctx.upload(V, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps,
               n_scan=(len(k), 1))
    nV, npr, nxl, nyl, err = ctx.run()
    ctx.close()
    oV, opr, oxl, oyl, oerr, _ = O.run_tilt(V, wl.probe, wl.prop, np.conj(wl.prop), xl, yl, M, len(k), 1,
                                            wl.binning, wl.sigma, wl.dx, steps)
    d, od = np.c_[nxl - xl, nyl - yl], np.c_[oxl - xl, oyl - yl]
    print('C2 plain: loss %.3e, V update %.2e, probe update %.2e, position moves %.2e (largest %.3f px)'
          % (abs(err / oerr - 1), rel(nV - V, oV - V), rel(npr - wl.probe, opr - wl.probe), rel(d, od),
             np.abs(od).max()), flush=True)
    assert abs(err / oerr - 1) < 1e-4
    # from half the potential the 24 object steps move V by about three times its norm, so the volume is
    # compared through its update (measured 3.4e-5); the probe moves by 0.2 % and is compared both ways
    assert rel(nV - V, oV - V) < 1e-3
    assert_update_close(npr, opr, wl.probe, len(k))
    ulp = np.spacing(np.float32(max(np.abs(xl).max(), np.abs(yl).max())))
    assert np.abs(od).max() > 10 * ulp                      # the position step moved the positions
    assert np.abs(d - od).max() <= max(1e-3 * np.abs(od).max(), 4 * ulp)


def test_c2_full_size_properties():
    """At the full C2 size (11 tilts x 1024 positions, 16 planes): the batched gradient of a tilt is bitwise
    repeatable and additive over position ranges (per-position terms bitwise, sums to FP32 reordering), and
    a graph-captured update pass of the 11-tilt series replays to the eager pass's state, bitwise, with
    eager work and synchronisations between the replays."""
    from mset_amd import lib
    ctx = lib.Context(0)
    wl = _workload('C2', ctx)
    j = 4
    V = (0.5 * wl.truth).astype(np.float32)
    ctx.upload(V, wl.probe, wl.prop, wl.xlists[j], wl.ylists[j], wl.measured[j], wl.sigma, wl.dx,
               slice_binning=wl.binning, n_scan=wl.n_scans[j])
    a = ctx.grad()
    b = ctx.grad()
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3] == b[3]
    lo, hi = ctx.grad(0, 400), ctx.grad(400, 624)
    assert np.array_equal(np.concatenate([lo[2], hi[2]]), a[2])
    assert rel(lo[0] + hi[0], a[0]) < 1e-5 and rel(lo[1] + hi[1], a[1]) < 1e-5
    assert abs((lo[3] + hi[3]) / a[3] - 1) < 1e-12
    ctx.close()

    steps = (100.0, 1e-3, 10.0)
    r = lib.Recon(V, wl.prop, len(wl.angles), wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps)
    try:
        for i, ang in enumerate(wl.angles):
            r.set_tilt(i, ang, wl.probe, wl.xlists[i], wl.ylists[i], wl.measured[i], n_scan=wl.n_scans[i])
        r.snapshot()
        r.iterate(lib.PATH_BATCHED, 1024, use_graph=False, want_error=False)
        eager = r.get()
        r.restore()
        r.iterate(lib.PATH_BATCHED, 1024, use_graph=True, want_error=False)      # eager pass, then capture
        for rep in range(3):
            r.restore()
            r.error_pass()                                                       # eager work between replays
            r.iterate(lib.PATH_BATCHED, 1024, use_graph=True, want_error=False)  # replay
            got = r.get()
            assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1]), rep
            for gx, ex in zip(got[2] + got[3], eager[2] + eager[3]):
                assert np.array_equal(gx, ex), rep
    finally:
        r.close()


def test_c3_full_size_graph_replay_equals_eager():
    """The headline configuration end to end: the full C3 series (21 tilts, 25,767 positions on uneven scans,
    256^2 probes, 32 planes), batches of 512 on the separable routine with all three gradient terms.  A
    graph-captured update pass replays to the eager pass's volume, probes and positions bit for bit, twice,
    with the one-launch error pass (eager work) between the replays; the error after the pass is finite and
    below the error before it."""
    from mset_amd import lib
    ctx = lib.Context(0)
    wl = _workload('C3', ctx)
    ctx.close()
    V = np.zeros_like(wl.truth, dtype=np.float32)
    steps = (100.0, 1e-3, 10.0)
    r = lib.Recon(V, wl.prop, len(wl.angles), wl.sigma, wl.dx, slice_binning=wl.binning, step_size=steps)
    try:
        for i, ang in enumerate(wl.angles):
            r.set_tilt(i, ang, wl.probe, wl.xlists[i], wl.ylists[i], wl.measured[i], n_scan=wl.n_scans[i])
        before = float(np.mean(r.error_pass()))
        r.snapshot()
        print('case: C3 eager pass', flush=True)
        r.iterate(lib.PATH_BATCHED, 512, use_graph=False, want_error=False)
        eager = r.get()
        after = float(np.mean(r.error_pass()))
        r.restore()
        r.iterate(lib.PATH_BATCHED, 512, use_graph=True, want_error=False)       # eager pass, then capture
        for rep in range(2):
            print('case: C3 graph replay', rep, flush=True)
            r.restore()
            r.error_pass()                                                       # eager work between replays
            r.iterate(lib.PATH_BATCHED, 512, use_graph=True, want_error=False)   # replay
            got = r.get()
            assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1]), rep
            for gx, ex in zip(got[2] + got[3], eager[2] + eager[3]):
                assert np.array_equal(gx, ex), rep
    finally:
        r.close()
    print('C3 error before %.5f after one pass %.5f' % (before, after), flush=True)
    assert np.isfinite(after) and after < before


def test_c3_fp32_gradient_vs_fp64_oracle():
    from mset_amd import lib
    ctx = lib.Context(0)
    wl = _workload('C3', ctx, tilts=[3])
    k = np.arange(0, wl.npos_of(0), max(1, wl.npos_of(0) // 4))[:4]
    V = wl.truth
    xl, yl, M = wl.xlists[0][k], wl.ylists[0][k], wl.measured[0][k]
    ctx.upload(V, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning, n_scan=(len(k), 1))
    dV, dp, dpos, err = ctx.grad()
    w = ctx.waves(0)
    one = [ctx.grad(i, 1) for i in range(len(k))]           # per-position terms
    ctx.close()
    bprop = np.conj(wl.prop)
    odV, odp, odpos, oep = O.grad_fixed(V, wl.probe, wl.prop, bprop, xl, yl, M, wl.binning, wl.sigma, wl.dx)
    psi0 = O.subpixel_shift(wl.probe, xl[0], yl[0])
    save, _, _, _ = O.forward(V, psi0, wl.prop, M[0], wl.binning, wl.sigma, wl.dx)
    r = dict(exit_wave=rel(w[-1], save[-1]), loss=abs(err / (oep.sum() / len(k)) - 1), dV=rel(dV, odV),
             dprobe=rel(dp, odp), dpos=rel(dpos, odpos))
    print('C3 FP32 vs FP64 relative errors', r, flush=True)
    assert r['exit_wave'] < 1e-4 and r['loss'] < 1e-4
    assert r['dV'] < 1e-3 and r['dprobe'] < 1e-3
    # The position terms d = Re sum(g . bw), g = MATLAB gradient(conj psi0) (numpy's edge_order=1 gradient),
    # cancel over the 65536 pixels by a factor of ~1e4 on this weak-phase, noisy state, so FP32 errors of the
    # back wave bw are amplified in d: an independent FP32 chain (single-precision FFTs) shows 1.5e-3 to 8.5e-3,
    # and forming the products in FP64 does not change it.  The bound used here follows from Cauchy-Schwarz:
    # |d_dev - d| <= (eps_bw + eps_g + 2 eps_32) ||g|| ||bw||, where eps_bw is the device's own relative error of
    # bw, read from its one-position probe term (a unitary shift of bw), and 1e-6 covers eps_g + 2 eps_32.
    for i in range(len(k)):
        p0 = O.subpixel_shift(wl.probe, xl[i], yl[i])
        sv, _, resi, _ = O.forward(V, p0, wl.prop, M[i], wl.binning, wl.sigma, wl.dx)
        _, bw, _ = O.back(V, sv, p0, resi, bprop, wl.binning, wl.sigma, wl.dx, xl[i], yl[i], O.MODE_REF)
        _, odp1, odpos1, _ = O.grad_fixed(V, wl.probe, wl.prop, bprop, xl[i:i + 1], yl[i:i + 1], M[i:i + 1],
                                          wl.binning, wl.sigma, wl.dx)
        eps_bw = rel(one[i][1], odp1)
        gy, gx = np.gradient(np.conj(p0))
        for ax, gg in ((0, gx), (1, gy)):
            bound = (eps_bw + 1e-6) * np.linalg.norm(gg) * np.linalg.norm(bw)
            e_all, e_one = abs(dpos[i, ax] - odpos[i, ax]), abs(one[i][2][0, ax] - odpos1[0, ax])
            print('  position %d axis %d: rel %.2e, error / bound %.3f (eps_bw %.1e)'
                  % (i, ax, e_all / abs(odpos[i, ax]), e_all / bound, eps_bw), flush=True)
            assert e_all <= bound and e_one <= bound, (i, ax, e_all, e_one, bound)


def test_c4_gradients_vs_central_differences():
    """Exact-mode device gradients along 50 seeded directions against (L(x + h d) - L(x - h d)) / 2h of
    the FP64 loss, whose own derivatives the oracle tests pin.  The reference-mode direction
    (STEP03_BACK) is compared with the oracle's restatement of it."""
    from mset_amd import lib
    ctx = lib.Context(0)
    wl = _workload('C4', ctx)
    V = (0.5 * wl.truth + 0.25 * wl.truth.max()).astype(np.float64)
    xl, yl, M = wl.xlists[0], wl.ylists[0], wl.measured[0]
    n = wl.P
    ctx.upload(V, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning, n_scan=wl.n_scans[0],
               grad_mode=lib.GRAD_EXACT)
    dV, dp, dpos, _ = ctx.grad()
    loss = lambda V_=V, p_=wl.probe, x_=xl, y_=yl: O.loss(V_, p_, wl.prop, np.float64(x_), np.float64(y_), M,
                                                        wl.binning, wl.sigma, wl.dx)
    og = O.grad_fixed(V, wl.probe, wl.prop, np.conj(wl.prop), xl, yl, M, wl.binning, wl.sigma, wl.dx,
                      mode=O.MODE_EXACT)
    rng = np.random.default_rng(4)
    h = 1e-4
    c = -2 * n * n * wl.sigma * wl.dx
    g = np.repeat(dV, wl.binning, axis=0)[:V.shape[0]].astype(np.float64)
    go = np.repeat(og[0], wl.binning, axis=0)[:V.shape[0]]
    fd, an, ao = [], [], []
    for _ in range(50):
        d = rng.standard_normal(V.shape)
        fd.append((loss(V_=V + h * d) - loss(V_=V - h * d)) / (2 * h))
        an.append(c * np.sum(g * d))
        ao.append(c * np.sum(go * d))
    fd, an, ao = map(np.array, (fd, an, ao))
    # the FD is exact to FP64 precision here (checked against the FP64 adjoint); the device gradient carries
    # FP32 rounding, so single directional derivatives that nearly cancel are compared as one 50-vector
    print('C4: FD vs FP64 adjoint %.2e, FD vs device (FP32) %.2e over 50 directions'
          % (np.linalg.norm(fd - ao) / np.linalg.norm(fd), np.linalg.norm(fd - an) / np.linalg.norm(fd)), flush=True)
    assert np.linalg.norm(fd - ao) <= 1e-6 * np.linalg.norm(fd)
    assert np.linalg.norm(fd - an) <= 1e-3 * np.linalg.norm(fd)
    # probe (complex directions): the probe carries the dose (|probe| ~ sqrt(counts)), so the loss curves
    # strongly along it and the central difference needs h ~ 1e-6 max|probe| to be exact to ~1e-7
    odp, odpos = og[1], og[2]
    hp = 1e-6 * float(np.abs(wl.probe).max())
    fp, ap, aq = [], [], []
    for _ in range(10):
        d = rng.standard_normal(dp.shape) + 1j * rng.standard_normal(dp.shape)
        fp.append((loss(p_=wl.probe + hp * d) - loss(p_=wl.probe - hp * d)) / (2 * hp))
        ap.append(2 * n * n * np.real(np.sum(np.conj(dp.astype(np.complex128)) * d)))
        aq.append(2 * n * n * np.real(np.sum(np.conj(odp) * d)))
    fp, ap, aq = np.array(fp), np.array(ap), np.array(aq)
    print('C4 probe: FD vs FP64 adjoint %.2e, FD vs device %.2e'
          % (np.linalg.norm(fp - aq) / np.linalg.norm(fp), np.linalg.norm(fp - ap) / np.linalg.norm(fp)), flush=True)
    assert np.linalg.norm(fp - aq) <= 1e-5 * np.linalg.norm(fp)
    assert np.linalg.norm(fp - ap) <= 1e-3 * np.linalg.norm(fp), (fp, ap)
    # positions: dL/dx_k = -2 n^2 dpos[k, 0] (the spectral position term); the FD with h = 1e-3 px is
    # exact to ~1e-4 against the FP64 term, and the FP32 device term is bounded by its summands as in C3
    hx = 1e-3
    for kk in range(3):
        xp, xm = np.float64(xl).copy(), np.float64(xl).copy()
        xp[kk] += hx
        xm[kk] -= hx
        fd = (loss(x_=xp) - loss(x_=xm)) / (2 * hx)
        assert abs(fd + 2 * n * n * odpos[kk, 0]) <= 1e-3 * abs(fd), (fd, odpos[kk, 0])
        p0 = O.subpixel_shift(wl.probe, xl[kk], yl[kk])
        sv, _, resi, _ = O.forward(V, p0, wl.prop, M[kk], wl.binning, wl.sigma, wl.dx)
        _, bw, _ = O.back(V, sv, p0, resi, np.conj(wl.prop), wl.binning, wl.sigma, wl.dx, xl[kk], yl[kk],
                          O.MODE_EXACT)
        scale = np.sum(np.abs(np.gradient(np.conj(p0))[1]) * np.abs(bw))
        assert abs(dpos[kk, 0] - odpos[kk, 0]) <= 1e-4 * scale, (kk, dpos[kk, 0], odpos[kk, 0], scale)
    # the reference (STEP03_BACK) direction against its oracle restatement
    ctx.upload(V, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning, n_scan=wl.n_scans[0],
               grad_mode=lib.GRAD_REF)
    rdV, rdp, rdpos, _ = ctx.grad()
    ctx.close()
    odV, odp, odpos, _ = O.grad_fixed(V, wl.probe, wl.prop, np.conj(wl.prop), xl, yl, M, wl.binning, wl.sigma,
                                      wl.dx, mode=O.MODE_REF)
    assert rel(rdV, odV) < 1e-3 and rel(rdp, odp) < 1e-3 and rel(rdpos, odpos) < 1e-3


def test_c4_store_vs_recompute_at_256_cubed():
    """The memory case of the adjoint-verification configuration: a 256^3 volume with one-voxel slices (256
    propagation planes), probe 256^2, 64 positions.  Recomputing the exit waves keeps 5 wave planes per
    resident position instead of 258, checkpointing 35 (K = 17); both give the stored-wave gradient, and the
    three variants match the FP64 oracle on two positions."""
    from mset_amd import exitwaves, lib
    wl = exitwaves.make_workload(64)
    out = exitwaves.run(64, repeat=1, wl=wl)
    print('256^3 store vs recompute vs checkpoint:',
          {k: out[k] for k in ('wave_memory_ratio', 'scratch_memory_ratio', 'time_ratio', 'recompute_vs_store',
                               'checkpoint_wave_memory_ratio', 'checkpoint_time_ratio', 'checkpoint_vs_store')},
          flush=True)
    assert wl.L == 256 and out['wave_memory_ratio'] == (wl.L + 2) / 5
    assert out['checkpoint_wave_memory_ratio'] == (wl.L + 2) / 35
    assert out['variants']['store']['resolved'] == lib.EXIT_WAVES_STORE
    assert out['variants']['recompute']['resolved'] == lib.EXIT_WAVES_RECOMPUTE
    assert out['variants']['checkpoint']['resolved'] == lib.EXIT_WAVES_CHECKPOINT
    # checkpointed waves are the stored ones, rebuilt by the same forward passes (separable routine)
    ck = out['checkpoint_vs_store']
    assert ck['dV_rel'] < 1e-5 and ck['dprobe_rel'] < 1e-5 and ck['loss_rel'] < 1e-6
    # the stored-wave variant runs the separable-propagator routine, the recomputing one the general routine:
    # their forward sweeps round differently (~1e-8 per plane over 256 planes), both far inside 1e-4 of FP64
    assert out['recompute_vs_store']['dV_rel'] < 1e-3 and out['recompute_vs_store']['loss_rel'] < 1e-5
    V0 = (0.5 * wl.truth).astype(np.float64)
    xl, yl, M = wl.xlists[0][:2], wl.ylists[0][:2], wl.measured[0][:2]
    odV, odp, _, oep = O.grad_fixed(V0, wl.probe, wl.prop, np.conj(wl.prop), xl, yl, M, wl.binning, wl.sigma, wl.dx)
    for mode in (lib.EXIT_WAVES_STORE, lib.EXIT_WAVES_RECOMPUTE, lib.EXIT_WAVES_CHECKPOINT):
        with lib.Context(0) as ctx:
            ctx.upload(V0, wl.probe, wl.prop, xl, yl, M, wl.sigma, wl.dx, slice_binning=wl.binning, n_scan=(2, 1))
            ctx.set_exit_waves(mode)
            dV, dp, _, err = ctx.grad()
        print('  exit waves %d vs FP64: dV %.2e dprobe %.2e loss %.2e'
              % (mode, rel(dV, odV), rel(dp, odp), abs(err / (oep.sum() / 2) - 1)), flush=True)
        assert rel(dV, odV) < 1e-3 and rel(dp, odp) < 1e-3 and abs(err / (oep.sum() / 2) - 1) < 1e-4
