"""GPU: the TV regulariser (TV_method.m) and the reconstruction driver (Run_MSET_recon.m:102-292) against
the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel
from tests.test_gpu_recon import _series, _recon

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    from mset_amd import lib
    c = lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize('shape,lam', [((16, 24, 32), 0.05), ((64, 128, 128), 0.02), ((5, 7, 9), 0.3)])
def test_tv_matches_oracle(ctx, shape, lam):
    rng = np.random.default_rng(sum(shape))
    f = rng.uniform(0, 1, shape).astype(np.float32)
    print('case: tv', shape, lam, flush=True)
    out, err = ctx.tv(f, lam, 300)
    of, ef = O.tv(f, lam, 300, single=True)
    od, ed = O.tv(f, lam, 300)
    assert rel(out, of) < 1e-5 and rel(out, od) < 2e-5
    # the error sums |grad u|; where u is nearly flat its terms are float rounding of near-equal neighbours
    floor = lam * f.size * 4 * np.finfo(np.float32).eps * np.abs(f).max()
    assert abs(err - ed) <= 2e-5 * abs(ed) + floor
    assert abs(out.sum() / f.astype(np.float64).sum() - 1) < 1e-5      # the mean is preserved


def test_recon_tv_normalises_by_the_maximum():
    base, T, rec = _series(n=32, nz=16, npos=2, tilts=(0.0,))
    r = _recon(base, T, rec, (0.0, 0.0, 0.0))
    err = r.tv(0.05, 300)
    got = r.get()[0]
    m = np.float32(rec.astype(np.float32).max())
    of, ef = O.tv(rec.astype(np.float32) / m, 0.05, 300, single=True)
    assert rel(got, m * of) < 1e-5 and abs(err / ef - 1) < 2e-5
    r.close()


def _oracle_state_pass(base, T, state, steps, positivity=True):
    rec, probes, xs, ys = state
    rec = rec.copy()
    probes, xs, ys = list(probes), list(xs), list(ys)
    for j, t in enumerate(T):
        RV = O.rot_3dvol(rec, t['ang'], inverse=True)
        RV, probes[j], xs[j], ys[j], _, _ = O.run_tilt(RV, probes[j], base['prop'], base['bprop'], xs[j], ys[j], t['M'],
                                                       len(xs[j]), 1, base['bin'], base['sig'], base['dx'], steps)
        rec = O.rot_3dvol(RV, t['ang'])
        if positivity:
            rec[rec < 0] = 0
    return rec, probes, xs, ys


def _oracle_state_errors(base, T, state):
    rec, probes, xs, ys = state
    return np.array([O.error_tilt(O.rot_3dvol(rec, t['ang'], inverse=True), probes[j], base['prop'], xs[j], ys[j], t['M'],
                                  len(xs[j]), 1, base['bin'], base['sig'], base['dx'])[0] for j, t in enumerate(T)])


def test_driver_with_line_search_and_tv_matches_oracle_loop():
    from mset_amd import lib
    from mset_amd.recon import Driver
    base, T, rec = _series(n=32, nz=16, npos=4, tilts=(-20.0, 15.0))
    # an object step far too large: the first attempt of iteration 2 is rejected by the line search
    steps = np.array([3e4, 1e-3, 0.0])
    r = _recon(base, T, rec, tuple(steps))
    d = Driver(r, steps, len(T), rec.shape, bls_parameter=0.1, use_tv=True, tv_lambda=0.01, path=lib.PATH_PLAIN)
    h = d.run(3)
    # the same loop on the oracle
    state = (rec.astype(np.float64), [t['probe'].astype(np.complex128) for t in T],
             [t['xl'].copy() for t in T], [t['yl'].copy() for t in T])
    old, st, hist, bls = 1e31, steps.copy(), [], []
    for i in range(3):
        while True:
            new = _oracle_state_pass(base, T, state, st)
            errs = _oracle_state_errors(base, T, new)
            if old > errs.mean():
                state = new
                break
            bls.append(i)
            st = 0.1 * st
        m = np.float64(np.float32(state[0].max()))
        out, _ = O.tv(state[0] / m, 0.01, 300)
        state = (m * out,) + state[1:]
        hist.append(errs.mean())
        old = errs.mean()
    print('bls events', h.bls, bls, flush=True)
    assert [b[0] for b in h.bls] == bls and len(bls) >= 1
    assert np.allclose(h.mean_error, hist, rtol=1e-3)
    got = r.get()[0]
    assert rel(got, state[0]) < 2e-3
    r.close()
