"""The device-resident tilt series (mset_recon_*) against the oracle composition of Run_MSET_recon's
update pass (:119-187: inverse rotation, sequential positions, forward rotation, positivity) and
error pass (:189-239)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import small_case
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def _series(n=32, nz=16, binning=2, npos=6, tilts=(-30.0, 0.0, 25.0), seed=4):
    base = small_case(n=n, nz=nz, binning=binning, npos=npos, seed=seed, spread=n / 8)
    rng = np.random.default_rng(seed + 1)
    T = []
    for j, t in enumerate(tilts):
        c = small_case(n=n, nz=nz, binning=binning, npos=npos, seed=seed + 10 + j, spread=n / 8)
        T.append(dict(ang=(0.0, t, 0.0), xl=c['xl'], yl=c['yl'], M=c['M'], probe=base['probe'] * (1 + 0.1 * j)))
    rec = np.abs(rng.standard_normal((nz, n, n))) * base['V'].mean()
    return base, T, rec


def _oracle_pass(base, T, rec, steps, positivity=True, mode=O.MODE_REF):
    rec = rec.astype(np.float64).copy()
    probes, xs, ys = [], [], []
    for t in T:
        RV = O.rot_3dvol(rec, t['ang'], inverse=True)
        RV, pr, xl, yl, _, _ = O.run_tilt(RV, t['probe'], base['prop'], base['bprop'], t['xl'], t['yl'], t['M'],
                                          len(t['xl']), 1, base['bin'], base['sig'], base['dx'], steps, mode=mode)
        rec = O.rot_3dvol(RV, t['ang'])
        if positivity:
            rec[rec < 0] = 0
        probes.append(pr)
        xs.append(xl)
        ys.append(yl)
    return rec, probes, xs, ys


def _oracle_errors(base, T, rec):
    out = []
    for t in T:
        RV = O.rot_3dvol(rec, t['ang'], inverse=True)
        e, _ = O.error_tilt(RV, t['probe'], base['prop'], t['xl'], t['yl'], t['M'], len(t['xl']), 1, base['bin'],
                            base['sig'], base['dx'])
        out.append(e)
    return np.array(out)


def _recon(base, T, rec, steps, mode=0):
    from mset_amd import lib
    r = lib.Recon(rec, base['prop'], len(T), base['sig'], base['dx'], slice_binning=base['bin'], step_size=steps,
                  grad_mode=mode)
    for j, t in enumerate(T):
        r.set_tilt(j, t['ang'], t['probe'], t['xl'], t['yl'], t['M'])
    return r


def test_update_pass_plain_matches_oracle():
    from mset_amd import lib
    base, T, rec = _series()
    steps = (20.0, 1e-3, 1e-2)
    print('case: recon plain update pass, 3 tilts, N=32', flush=True)
    with _recon(base, T, rec, steps) as r:
        e0 = r.error_pass()
        np.testing.assert_allclose(e0, _oracle_errors(base, T, rec), rtol=1e-4)
        r.iterate(lib.PATH_PLAIN, 1)
        grec, gpr, gx, gy = r.get()
    orec, opr, ox, oy = _oracle_pass(base, T, rec, steps)
    d, od = grec - rec, orec - rec
    assert np.linalg.norm(d - od) <= 2e-3 * np.linalg.norm(od) + 1e-5 * np.linalg.norm(rec)
    for j in range(len(T)):
        assert rel(gpr[j] - T[j]['probe'], opr[j] - T[j]['probe']) < 2e-3
        assert rel(np.c_[gx[j], gy[j]] - np.c_[T[j]['xl'], T[j]['yl']],
                   np.c_[ox[j], oy[j]] - np.c_[T[j]['xl'], T[j]['yl']]) < 5e-3


def test_graph_replay_equals_eager_and_snapshot_restores():
    from mset_amd import lib
    base, T, rec = _series(n=64, nz=16, binning=4, npos=9)
    steps = (20.0, 1e-3, 1e-2)
    print('case: recon batched, graph vs eager, N=64', flush=True)
    with _recon(base, T, rec, steps) as a, _recon(base, T, rec, steps) as b:
        for _ in range(3):
            a.iterate(lib.PATH_BATCHED, 4, use_graph=False)
            b.iterate(lib.PATH_BATCHED, 4, use_graph=True)
        ra, pa, xa, ya = a.get()
        rb, pb, xb, yb = b.get()
        assert np.array_equal(ra, rb) and np.array_equal(pa, pb)
        assert all(np.array_equal(u, v) for u, v in zip(xa, xb))
        b.snapshot()
        b.iterate(lib.PATH_BATCHED, 4, use_graph=True)
        assert not np.array_equal(b.get()[0], rb)
        b.restore()
        rc, pc, xc, yc = b.get()
        assert np.array_equal(rc, rb) and np.array_equal(pc, pb) and all(np.array_equal(u, v) for u, v in zip(xc, xb))


def test_batched_with_zero_steps_is_forward_plus_gradient_only():
    from mset_amd import lib
    base, T, rec = _series()
    with _recon(base, T, rec, (0.0, 0.0, 0.0)) as r:
        e = r.iterate(lib.PATH_BATCHED, 1024, positivity=False)
        grec, _, _, _ = r.get()
    # rec only passes through inverse + forward rotation per tilt
    ref = rec.astype(np.float64)
    for t in T:
        ref = O.rot_3dvol(O.rot_3dvol(ref, t['ang'], inverse=True), t['ang'])
    assert rel(grec, ref) < 1e-5
    assert np.isfinite(e) and e > 0


@pytest.mark.parametrize('n,nz,binning', [(64, 16, 2), (256, 16, 4)])
def test_error_pass_is_one_launch_over_uneven_tilts(n, nz, binning):
    """Run_MSET_recon.m:189-239 over four tilts with 5, 11, 3 and 8 positions: one persistent launch, whose
    per-tilt errors match one launch per tilt on the same rotated volumes and the oracle's error pass."""
    from mset_amd import lib
    base = small_case(n=n, nz=nz, binning=binning, npos=4, seed=n + 3, spread=n / 8)
    rng = np.random.default_rng(n + 4)
    T = []
    for j, (t, npos) in enumerate(((-40.0, 5), (-10.0, 11), (15.0, 3), (35.0, 8))):
        c = small_case(n=n, nz=nz, binning=binning, npos=npos, seed=n + 20 + j, spread=n / 8)
        T.append(dict(ang=(0.0, t, 0.0), xl=c['xl'], yl=c['yl'], M=c['M'], probe=base['probe'] * (1 + 0.05 * j)))
    rec = (np.abs(rng.standard_normal((nz, n, n))) * base['V'].mean()).astype(np.float32)
    print('case: one-launch error pass, N=%d, tilts of %s positions' % (n, [len(t['xl']) for t in T]), flush=True)
    with _recon(base, T, rec, (0.0, 0.0, 0.0)) as r:
        r.profile(True)
        e = r.error_pass()
        launches, _ = r.profile_read('position_fwd')
        r.profile(False)
    assert launches == 1
    per_tilt = []
    for t in T:
        with lib.Context(0) as ctx:
            ctx.upload(rec, t['probe'], base['prop'], t['xl'], t['yl'], t['M'], base['sig'], base['dx'],
                       slice_binning=binning, n_scan=(len(t['xl']), 1))
            ctx.rotate_resident(t['ang'], inverse=True)
            per_tilt.append(ctx.error())
    np.testing.assert_allclose(e, per_tilt, rtol=1e-6)
    np.testing.assert_allclose(e, _oracle_errors(base, T, rec), rtol=1e-4)
