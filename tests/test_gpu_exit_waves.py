"""Recomputed and checkpointed exit waves (mset_set_exit_waves).  The batched adjoint that rebuilds psi_a by
inverting the forward sweep gives the stored-wave gradient and the FP64 oracle's, with 5 scratch planes per
resident position instead of L + 2.  The checkpointed variant keeps every K-th wave and one segment and
rebuilds the other segments with the forward sweep's own steps: the stored variant's waves, for any
propagator.

Tolerances: the forward pass is the same arithmetic in both variants (the two kernel instantiations
may order a few roundings differently), so the losses agree to 1e-6; the rebuilt waves carry the FP32
rounding of the inverted planes (about L ulp), so the gradients agree with the stored-wave ones to 1e-4
relative and with the FP64 oracle to the usual 1e-3.  The adjoint recursion itself does not read the
waves, so the probe and reference-mode position terms match the stored variant's closely too.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import small_case
from tests.conftest import rel

pytestmark = pytest.mark.gpu

REBUILD_TOL = 1e-4
GRAD_TOL = 1e-3
# checkpointed vs stored waves: the same forward steps, in another kernel instantiation.  Some engines give
# the stored variant's gradient bit for bit (the N = 32 and 64 LDS engines); elsewhere the compiler contracts
# a complex product into FMAs differently in the rebuild than in the forward sweep (up to 3.5e-6 after 8-18
# planes at N = 48 and 128), so the bound is a few ulps per plane
CKPT_TOL = 1e-5


def ck_planes(L):
    """per-slot wave planes of the checkpoint variant, K = ceil(sqrt(L + 1))"""
    k = 2
    while k * k < L + 1:
        k += 1
    return L // k + k + 3


def _upload(ctx, c, mode, prop=None, steps=(0.0, 0.0, 0.0)):
    ctx.upload(c['V'], c['probe'], c['prop'] if prop is None else prop, c['xl'], c['yl'], c['M'], c['sig'],
               c['dx'], slice_binning=c['bin'], step_size=steps, n_scan=(c['npos'], 1), grad_mode=mode)


@pytest.mark.parametrize('n,nz,binning,mode', [(64, 16, 2, O.MODE_REF), (64, 24, 1, O.MODE_EXACT),
                                               (32, 40, 1, O.MODE_EXACT), (128, 16, 2, O.MODE_EXACT),
                                               (128, 32, 4, O.MODE_REF), (256, 12, 2, O.MODE_REF),
                                               (256, 10, 2, O.MODE_EXACT)])
def test_recomputed_waves_give_the_stored_gradient(n, nz, binning, mode):
    from mset_amd import lib
    c = small_case(n=n, nz=nz, binning=binning, npos=6, seed=3 * n + nz, spread=n / 8)
    L = -(-nz // binning)
    print('case: exit waves', n, nz, binning, mode, flush=True)
    ctx = lib.Context(0)
    try:
        _upload(ctx, c, mode)
        ctx.set_exit_waves(lib.EXIT_WAVES_STORE)
        dV, dp, dpos, err = ctx.grad()
        ms, bs = ctx.exit_waves()
        ctx.set_exit_waves(lib.EXIT_WAVES_RECOMPUTE)
        rV, rp, rpos, rerr = ctx.grad()
        mr, br = ctx.exit_waves()
    finally:
        ctx.close()
    assert (ms, mr) == (lib.EXIT_WAVES_STORE, lib.EXIT_WAVES_RECOMPUTE)
    assert bs == n * n * ((L + 2) * 8 + L * 4) and br == n * n * (5 * 8 + L * 4)
    # the same forward pass, except at 256 with the reference pairing, where the stored-wave variant runs the
    # separable-propagator routine and the recomputing one the general routine (rounding differs slightly)
    assert abs(rerr / err - 1) < (1e-5 if n == 256 and mode == O.MODE_REF else 1e-6)
    print('rebuilt vs stored: dV %.2e dprobe %.2e' % (rel(rV, dV), rel(rp, dp)), flush=True)
    assert rel(rV, dV) < REBUILD_TOL
    assert rel(rp, dp) < REBUILD_TOL
    # position terms cancel over the plane; bound them by the largest term
    assert np.max(np.abs(rpos - dpos)) <= REBUILD_TOL * np.max(np.abs(dpos))
    odV, odp, odpos, _ = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                      c['bin'], c['sig'], c['dx'], mode=mode)
    assert rel(rV, odV) < GRAD_TOL
    assert rel(rp, odp) < GRAD_TOL
    assert rel(rpos, odpos) < GRAD_TOL


@pytest.mark.parametrize('n,nz,binning,mode', [(64, 16, 2, O.MODE_REF), (64, 12, 1, O.MODE_REF),
                                               (64, 12, 1, O.MODE_EXACT), (32, 40, 1, O.MODE_EXACT),
                                               (48, 15, 1, O.MODE_REF), (128, 16, 2, O.MODE_EXACT),
                                               (128, 36, 2, O.MODE_REF), (256, 12, 2, O.MODE_REF),
                                               (256, 10, 1, O.MODE_EXACT), (256, 13, 1, O.MODE_REF),
                                               (256, 18, 2, O.MODE_REF)])
def test_checkpointed_waves_give_the_stored_gradient(n, nz, binning, mode):
    """Segments that end on the last plane (L a multiple of K) and ones that do not, both pairings, every
    engine; at 256 with the reference pairing both plane routines (the separable one reruns its one-axis
    passes from checkpointed pass inputs), each against the stored variant on the same routine."""
    from mset_amd import lib
    c = small_case(n=n, nz=nz, binning=binning, npos=6, seed=7 * n + nz, spread=n / 8)
    L = -(-nz // binning)
    print('case: checkpointed exit waves', n, nz, binning, mode, 'L', L, flush=True)
    odV, odp, odpos, _ = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                      c['bin'], c['sig'], c['dx'], mode=mode)
    engines = [lib.ENGINE_GENERAL] + ([lib.ENGINE_AUTO] if n == 256 and mode == O.MODE_REF else [])
    ctx = lib.Context(0)
    try:
        _upload(ctx, c, mode)
        for e in engines:
            ctx.set_engine(e)
            ctx.set_exit_waves(lib.EXIT_WAVES_STORE)
            dV, dp, dpos, err = ctx.grad()
            ctx.set_exit_waves(lib.EXIT_WAVES_CHECKPOINT)
            cV, cp, cpos, cerr = ctx.grad()
            mc, bc = ctx.exit_waves()
            assert mc == lib.EXIT_WAVES_CHECKPOINT and bc == n * n * (ck_planes(L) * 8 + L * 4)
            print('  engine %d, checkpointed vs stored: dV %.2e dprobe %.2e dpos %.2e loss %.2e bitwise dV %s'
                  % (e, rel(cV, dV), rel(cp, dp), rel(cpos, dpos), abs(cerr / err - 1), np.array_equal(cV, dV)),
                  flush=True)
            assert abs(cerr / err - 1) < CKPT_TOL
            assert rel(cV, dV) < CKPT_TOL and rel(cp, dp) < CKPT_TOL
            assert np.max(np.abs(cpos - dpos)) <= CKPT_TOL * np.max(np.abs(dpos))
            assert rel(cV, odV) < GRAD_TOL and rel(cp, odp) < GRAD_TOL and rel(cpos, odpos) < GRAD_TOL
    finally:
        ctx.close()


def test_checkpoints_serve_a_propagator_that_recompute_refuses():
    """A non-unit-modulus propagator (0.9 of the Fresnel one): the inverted step is refused, the checkpointed
    variant gives the stored-wave gradient."""
    from mset_amd import lib
    c = small_case(n=64, nz=24, binning=1, npos=5, seed=21, spread=6.0)
    prop = 0.9 * c['prop']
    ctx = lib.Context(0)
    try:
        _upload(ctx, c, O.MODE_REF, prop=prop)
        ctx.set_exit_waves(lib.EXIT_WAVES_RECOMPUTE)
        with pytest.raises(lib.MsetError) as e:
            ctx.grad()
        assert e.value.code == lib.MSET_EINVAL
        ctx.set_exit_waves(lib.EXIT_WAVES_STORE)
        dV, dp, dpos, err = ctx.grad()
        ctx.set_exit_waves(lib.EXIT_WAVES_CHECKPOINT)
        cV, cp, cpos, cerr = ctx.grad()
        assert ctx.exit_waves()[0] == lib.EXIT_WAVES_CHECKPOINT
    finally:
        ctx.close()
    print('non-unit propagator, checkpointed vs stored: dV %.2e dprobe %.2e' % (rel(cV, dV), rel(cp, dp)), flush=True)
    assert abs(cerr / err - 1) < CKPT_TOL and rel(cV, dV) < CKPT_TOL and rel(cp, dp) < CKPT_TOL
    assert np.max(np.abs(cpos - dpos)) <= CKPT_TOL * np.max(np.abs(dpos))


def test_recompute_needs_a_unit_modulus_propagator_and_auto_stores():
    from mset_amd import lib
    c = small_case(n=32, nz=12, binning=1, npos=3, seed=5, spread=4.0)
    ctx = lib.Context(0)
    try:
        _upload(ctx, c, O.MODE_REF, prop=0.9 * c['prop'])
        ctx.set_exit_waves(lib.EXIT_WAVES_RECOMPUTE)
        with pytest.raises(lib.MsetError) as e:
            ctx.grad()
        assert e.value.code == lib.MSET_EINVAL
        ctx.set_exit_waves(lib.EXIT_WAVES_AUTO)
        ctx.grad()
        assert ctx.exit_waves()[0] == lib.EXIT_WAVES_STORE
        # a unit-modulus propagator at a small size: the store scratch is far below a quarter of the device
        _upload(ctx, c, O.MODE_REF)
        ctx.grad()
        assert ctx.exit_waves()[0] == lib.EXIT_WAVES_STORE
    finally:
        ctx.close()


def test_recon_pass_with_recomputed_waves():
    """A graph-captured batched update pass of a two-tilt series: rebuilt and checkpointed waves give the
    stored-wave volume, and switching the variant drops the captured graph."""
    from mset_amd import lib
    c = small_case(n=64, nz=16, binning=2, npos=8, seed=11, spread=6.0)
    out = {}
    for mode in (lib.EXIT_WAVES_STORE, lib.EXIT_WAVES_RECOMPUTE, lib.EXIT_WAVES_CHECKPOINT):
        r = lib.Recon(c['V'], c['prop'], 2, c['sig'], c['dx'], slice_binning=c['bin'], step_size=(50.0, 0.0, 0.0))
        try:
            r.set_exit_waves(mode)
            for j, ang in enumerate(((0.0, -20.0, 0.0), (0.0, 25.0, 0.0))):
                r.set_tilt(j, ang, c['probe'], c['xl'], c['yl'], c['M'], n_scan=(c['npos'], 1))
            for _ in range(2):
                r.iterate(lib.PATH_BATCHED, 4, use_graph=True, want_error=False)
            out[mode] = r.get()[0]
        finally:
            r.close()
    ds = out[lib.EXIT_WAVES_STORE] - c['V']
    for mode, tol in ((lib.EXIT_WAVES_RECOMPUTE, REBUILD_TOL), (lib.EXIT_WAVES_CHECKPOINT, CKPT_TOL)):
        d = out[mode] - c['V']
        assert np.linalg.norm(d - ds) <= tol * np.linalg.norm(ds) + 1e-6 * np.linalg.norm(c['V']), mode


@pytest.mark.parametrize('n,nz,binning,mode', [(128, 16, 2, O.MODE_REF), (128, 12, 3, O.MODE_EXACT),
                                               (256, 8, 2, O.MODE_REF), (256, 6, 1, O.MODE_EXACT)])
def test_slots_reused_by_several_positions_stored_and_recomputed(n, nz, binning, mode):
    """Two resident slots for eight positions: each slot runs four positions in one launch, so its later
    positions add to the slabs its first one wrote, and its wave and plane scratch are reused.  The three
    exit-wave variants against the FP64 oracle, and the per-position terms against the unlimited grid (bitwise: they do
    not depend on which slot ran the position)."""
    from mset_amd import lib
    c = small_case(n=n, nz=nz, binning=binning, npos=8, seed=5 * n + nz, spread=n / 8)
    print('case: slot reuse', n, nz, binning, mode, flush=True)
    odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                        c['bin'], c['sig'], c['dx'], mode=mode)
    ctx = lib.Context(0)
    try:
        _upload(ctx, c, mode)
        full = ctx.grad()
        for ew in (lib.EXIT_WAVES_STORE, lib.EXIT_WAVES_RECOMPUTE, lib.EXIT_WAVES_CHECKPOINT):
            ctx.set_exit_waves(ew)
            ctx.set_slot_limit(2)
            dV, dp, dpos, err = ctx.grad()
            assert ctx.exit_waves()[0] == ew
            ctx.set_slot_limit(0)
            print('  exit waves %d: dV %.2e dprobe %.2e dpos %.2e' % (ew, rel(dV, odV), rel(dp, odp), rel(dpos, odpos)),
                  flush=True)
            assert abs(err / (oep.sum() / c['npos']) - 1) < 1e-4
            assert rel(dV, odV) < GRAD_TOL and rel(dp, odp) < GRAD_TOL and rel(dpos, odpos) < GRAD_TOL
            if ew == lib.EXIT_WAVES_STORE:
                assert np.array_equal(dpos, full[2])
                assert rel(dV, full[0]) < 1e-5 and rel(dp, full[1]) < 1e-5
    finally:
        ctx.close()
