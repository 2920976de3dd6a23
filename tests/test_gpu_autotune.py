"""The autotuner and the knobs it turns: a capped persistent grid gives the same gradient, bitwise (the
gradient slabs are assigned by position, not by workgroup); the alternative workgroup shape of every plane
size gives the FP64 oracle's gradient; tuning an update pass leaves the tilt series unchanged."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import small_case
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def test_slot_limit_keeps_the_gradient_bitwise():
    """More positions than slabs (each slab sums several positions): a grid capped at 7 (rounded down to a
    divisor of the slab count), at half the slabs or at 1 gives the full grid's sums bit for bit."""
    from mset_amd import lib
    S = lib.batch_slots(32)
    c = small_case(n=32, nz=6, binning=2, npos=2 * S + 37, seed=21, spread=6.0)
    print('case: slot limit, %d positions on %d slabs' % (c['npos'], S), flush=True)
    ctx = lib.Context(0)
    try:
        ctx.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'],
                   slice_binning=c['bin'], n_scan=(c['npos'], 1))
        full = ctx.grad()
        out = []
        for s in (7, S // 2, 1):
            ctx.set_slot_limit(s)
            out.append(ctx.grad())
        ctx.set_slot_limit(0)
        again = ctx.grad()
    finally:
        ctx.close()
    for g in out + [again]:
        for a, b in zip(g[:3], full[:3]):
            assert np.array_equal(a, b)
        assert g[3] == full[3]
    odV = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'][:64], c['yl'][:64], c['M'][:64],
                       c['bin'], c['sig'], c['dx'])[0]
    sub = lib.Context(0)
    try:
        sub.upload(c['V'], c['probe'], c['prop'], c['xl'][:64], c['yl'][:64], c['M'][:64], c['sig'], c['dx'],
                   slice_binning=c['bin'], n_scan=(64, 1))
        assert rel(sub.grad()[0], odV) < 1e-3
    finally:
        sub.close()


@pytest.mark.parametrize('n,nz,binning,mode', [(16, 6, 1, O.MODE_REF), (24, 8, 2, O.MODE_EXACT),
                                               (32, 8, 2, O.MODE_REF), (48, 9, 3, O.MODE_REF),
                                               (64, 12, 2, O.MODE_EXACT), (96, 8, 2, O.MODE_REF),
                                               (128, 16, 2, O.MODE_REF), (128, 12, 2, O.MODE_EXACT),
                                               (256, 8, 2, O.MODE_REF), (256, 6, 1, O.MODE_EXACT)])
def test_alternative_workgroup_shape_gives_the_oracle_gradient(n, nz, binning, mode):
    """MSET_ENGINE_ALT_SHAPE at every plane size (at 256 with the separable and the general routine), with
    two slots for eight positions so that slots and slabs are reused: the FP64 oracle's gradient, and the
    default shape's to FP32 rounding (the shapes sum in different orders)."""
    from mset_amd import lib
    c = small_case(n=n, nz=nz, binning=binning, npos=8, seed=11 * n + nz, spread=n / 8)
    print('case: alternative shape', n, nz, binning, mode, flush=True)
    odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                        c['bin'], c['sig'], c['dx'], mode=mode)
    engines = [lib.ENGINE_ALT_SHAPE] + ([lib.ENGINE_ALT_SHAPE | lib.ENGINE_GENERAL] if n == 256 else [])
    ctx = lib.Context(0)
    try:
        ctx.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'],
                   slice_binning=c['bin'], n_scan=(c['npos'], 1), grad_mode=mode)
        base = ctx.grad()
        for e in engines:
            ctx.set_engine(e)
            for limit in (0, 2):
                ctx.set_slot_limit(limit)
                dV, dp, dpos, err = ctx.grad()
                print('  engine %d slots %d: dV %.2e dprobe %.2e dpos %.2e vs FP64; dV %.2e vs default shape'
                      % (e, limit, rel(dV, odV), rel(dp, odp), rel(dpos, odpos), rel(dV, base[0])), flush=True)
                assert abs(err / (oep.sum() / c['npos']) - 1) < 1e-4
                assert rel(dV, odV) < 1e-3 and rel(dp, odp) < 1e-3 and rel(dpos, odpos) < 1e-3
                assert rel(dV, base[0]) < 1e-4 and rel(dp, base[1]) < 1e-4
            ctx.set_slot_limit(0)
    finally:
        ctx.close()


def test_tune_recon_leaves_the_series_unchanged_and_sets_the_best_grid():
    from mset_amd import autotune, lib
    c = small_case(n=64, nz=16, binning=2, npos=24, seed=22, spread=8.0)
    r = lib.Recon(c['V'], c['prop'], 2, c['sig'], c['dx'], slice_binning=c['bin'], step_size=(50.0, 0.0, 0.0))
    try:
        for j, ang in enumerate(((0.0, -15.0, 0.0), (0.0, 15.0, 0.0))):
            r.set_tilt(j, ang, c['probe'], c['xl'], c['yl'], c['M'], n_scan=(c['npos'], 1))
        before = r.get()[0]
        best = autotune.tune_recon(r, batches=(8, 24), slots=[24, 12])
        after = r.get()[0]
        assert np.array_equal(before, after)
        # two workgroup shapes x two grids x two batch sizes
        assert best['batch'] in (8, 24) and best['slots'] in (24, 12) and len(best['table']) == 8
        assert best['engine'] in (lib.ENGINE_AUTO, lib.ENGINE_ALT_SHAPE)
        # the tuned grid is in place and a pass runs with it
        r.iterate(lib.PATH_BATCHED, best['batch'], use_graph=True, want_error=False)
        assert not np.array_equal(r.get()[0], before)
        # cached: a second call returns the same choice without timing again
        assert autotune.tune_recon(r, batches=(8, 24), slots=[24, 12]) is best
    finally:
        r.close()


def test_tune_grad_picks_a_candidate():
    from mset_amd import autotune, lib
    c = small_case(n=32, nz=6, binning=2, npos=30, seed=23, spread=4.0)
    ctx = lib.Context(0)
    try:
        ctx.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'],
                   slice_binning=c['bin'], n_scan=(c['npos'], 1))
        best = autotune.tune_grad(ctx, 32, slots=[30, 10], repeats=1)
        assert best['slots'] in (30, 10) and len(best['table']) == 2
        odV = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], c['bin'],
                           c['sig'], c['dx'])[0]
        assert rel(ctx.grad()[0], odV) < 1e-3
    finally:
        ctx.close()
