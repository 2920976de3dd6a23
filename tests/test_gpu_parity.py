"""GPU parity: the HIP path through the C ABI against the FP64 oracle and the reference's golden vectors.

Tolerances (FP32 device arithmetic against FP64): 1e-4 relative on exit waves and losses, 1e-3 on
gradients and on the updates they produce; rotation 1e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import small_case
from tests.conftest import rel

pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-4
LOSS_TOL = 1e-4
GRAD_TOL = 1e-3


@pytest.fixture(scope='module')
def mset():
    from mset_amd import lib
    ctx = lib.Context(0)
    yield ctx
    ctx.close()


def _phys(man):
    E0 = man['E0_kV'] * 1e3
    return O.wavelength(E0), O.interaction_sigma(E0), man['potential_pixelsize']


def assert_update_close(new, onew, old, nupd, tol=GRAD_TOL):
    """new - old must match onew - old to `tol` relative, above the floor of FP32 storage rounding."""
    d, od = np.asarray(new) - old, np.asarray(onew) - old
    floor = nupd * np.finfo(np.float32).eps * np.linalg.norm(np.ravel(old))
    err = np.linalg.norm(np.ravel(d - od))
    assert err <= tol * np.linalg.norm(np.ravel(od)) + floor, (err, np.linalg.norm(np.ravel(od)), floor)
    assert rel(new, onew) < 1e-5


def _upload_case(ctx, c, steps=(0.0, 0.0, 0.0), mode=0, V=None):
    ctx.upload(c['V'] if V is None else V, c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'],
               slice_binning=c['bin'], step_size=steps, n_scan=(c['npos'], 1), grad_mode=mode)


# ------------------------------------------------------------------ reference golden vectors
def test_waves_match_saved_multislice_fixture(mset, cuau):
    """STEP02_FORWARD_4D save_wave_f of position 576, tilt 8 (MSET_recon.mat) and the oracle."""
    man, s = cuau['man'], cuau['snap']
    lam, sig, dx = _phys(man)
    j = man['Nth_angle'] - 1
    print('case: CuAu N=24 tilt', j + 1, 'position', man['k'], flush=True)
    mset.upload(s['RVol'], s['probe_wfn'][j], s['prop'], s['scan_pos'][j, 0], s['scan_pos'][j, 1], cuau['meas'][j],
                sig, dx, slice_binning=1, n_scan=(24, 24))
    w = mset.waves(man['k'] - 1)
    assert rel(w[0], s['tmp_probe_wfn']) < 1e-5
    for a in range(24):
        assert rel(w[a + 1], s['save_wave_f'][a]) < 2e-5, a
    save, _, _, _ = O.forward(s['RVol'], O.subpixel_shift(s['probe_wfn'][j], man['row'], man['col']), s['prop'],
                              cuau['meas'][j][man['k'] - 1], 1, sig, dx)
    for a in range(24):
        assert rel(w[a + 1], save[a]) < WAVE_TOL, a


def test_error_pass_matches_saved_per_tilt_errors(mset, cuau):
    """Fixture 7: error pass from the saved rec (device rotation + device forward/loss), all 8 tilts."""
    man, s = cuau['man'], cuau['snap']
    lam, sig, dx = _phys(man)
    errs = []
    for j in range(8):
        print('case: CuAu error pass tilt', j + 1, flush=True)
        mset.upload(s['rec'], s['probe_wfn'][j], s['prop'], s['scan_pos'][j, 0], s['scan_pos'][j, 1],
                    cuau['meas'][j], sig, dx, slice_binning=1, n_scan=(24, 24))
        mset.rotate_resident(s['tilt_angles'][j], inverse=True)
        errs.append(mset.error_resident())
    errs = np.array(errs)
    assert np.abs(errs / s['error'] - 1).max() < 5e-5, errs / s['error'] - 1


def test_rotation_matches_saved_rotated_volume(mset, cuau):
    """Fixture 6: RVol = Func_inv_rot_3Dvol_FourierShear(rec, tilt 8)."""
    s = cuau['snap']
    R, _ = O.rot_matrix(s['tilt_angles'][7], inverse=True)
    out = mset.rotate(s['rec'], R)
    assert rel(out, s['RVol']) < 5e-6


@pytest.mark.parametrize('shape,ang', [((24, 32, 32), (0.0, 37.0, 0.0)), ((64, 64, 64), (12.0, -50.0, 7.0)),
                                       ((48, 96, 96), (-30.0, 0.0, 0.0)), ((32, 128, 128), (0.0, 0.0, 22.0)),
                                       ((96, 48, 64), (5.0, 60.0, -15.0)), ((16, 256, 256), (0.0, 20.0, 0.0))])
def test_rotation_vs_oracle(mset, shape, ang):
    rng = np.random.default_rng(sum(shape))
    v = rng.standard_normal(shape).astype(np.float32)
    R, _ = O.rot_matrix(ang)
    print('case: rotation', shape, ang, flush=True)
    assert rel(mset.rotate(v, R), O.rotate(v, R)) < 1e-5


# ------------------------------------------------------------------ operator parity on synthetic states
@pytest.mark.parametrize('n,nz,binning,mode', [(24, 6, 2, O.MODE_REF), (32, 12, 3, O.MODE_EXACT),
                                               (64, 8, 4, O.MODE_REF), (128, 16, 4, O.MODE_EXACT),
                                               (128, 7, 2, O.MODE_REF), (48, 5, 2, O.MODE_EXACT),
                                               (96, 4, 1, O.MODE_REF), (64, 32, 1, O.MODE_EXACT),
                                               (64, 6, 2, O.MODE_EXACT), (16, 4, 1, O.MODE_EXACT),
                                               (256, 8, 4, O.MODE_REF),
                                               (256, 6, 2, O.MODE_EXACT)])
def test_grad_at_fixed_state_vs_oracle(mset, n, nz, binning, mode):
    c = small_case(n=n, nz=nz, binning=binning, npos=5, seed=n + nz, spread=n / 8)
    print('case: grad', n, nz, binning, mode, flush=True)
    _upload_case(mset, c, mode=mode)
    dV, dp, dpos, err = mset.grad()
    odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                        c['bin'], c['sig'], c['dx'], mode=mode)
    oerr = oep.sum() / c['npos']          # oracle terms are per-position means
    assert abs(err / oerr - 1) < LOSS_TOL
    assert rel(dV, odV) < GRAD_TOL
    assert rel(dp, odp) < GRAD_TOL
    assert rel(dpos, odpos) < GRAD_TOL
    w = mset.waves(2)
    psi0 = O.subpixel_shift(c['probe'], c['xl'][2], c['yl'][2])
    save, _, _, _ = O.forward(c['V'], psi0, c['prop'], c['M'][2], c['bin'], c['sig'], c['dx'])
    assert rel(w[0], psi0) < WAVE_TOL
    assert rel(w[-1], save[-1]) < WAVE_TOL


@pytest.mark.parametrize('n,nz,binning,mode', [(24, 6, 2, O.MODE_REF), (32, 9, 3, O.MODE_EXACT),
                                               (64, 8, 4, O.MODE_REF), (48, 7, 3, O.MODE_EXACT),
                                               (128, 8, 4, O.MODE_REF), (128, 5, 2, O.MODE_EXACT),
                                               (256, 4, 2, O.MODE_REF)])
def test_sequential_run_vs_oracle(mset, n, nz, binning, mode):
    """'run_mset' semantics: per-position updates of V, probe and position (Run_MSET_recon.m:134-158)."""
    c = small_case(n=n, nz=nz, binning=binning, npos=6, seed=3 * n + nz, spread=n / 8)
    steps = (30.0, 2e-3, 2e-2)
    print('case: run', n, nz, binning, mode, flush=True)
    _upload_case(mset, c, steps=steps, mode=mode)
    V, pr, xl, yl, err = mset.run()
    oV, opr, oxl, oyl, oerr, _ = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                            c['npos'], 1, c['bin'], c['sig'], c['dx'], steps, mode=mode)
    assert abs(err / oerr - 1) < LOSS_TOL
    # the state is stored in single precision (as the reference stores RVol and the probe): after npos
    # updates each element carries up to ~npos ulp of rounding on top of the gradient tolerance
    assert_update_close(V, oV, c['V'], c['npos'])
    assert_update_close(pr, opr, c['probe'], c['npos'])
    d = np.c_[xl - c['xl'], yl - c['yl']]
    od = np.c_[oxl - c['xl'], oyl - c['yl']]
    # positions are stored in single precision (MATLAB scan_pos is single): a step can round away
    ulp = np.spacing(np.float32(max(np.abs(c['xl']).max(), np.abs(c['yl']).max())))
    assert np.abs(d - od).max() <= max(5e-3 * np.abs(od).max(), 4 * ulp)


def test_error_pass_vs_oracle_and_zero_step_paths(mset):
    c = small_case(n=32, nz=8, binning=2, npos=7, seed=11)
    _upload_case(mset, c)
    e_fwd = mset.error_resident()
    oerr, _ = O.error_tilt(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['npos'], 1, c['bin'],
                           c['sig'], c['dx'])
    assert abs(e_fwd / oerr - 1) < LOSS_TOL
    e_run = mset.run_resident()            # steps are zero: the state must not move
    V, pr, xl, yl = mset.download()
    assert np.array_equal(V, c['V'].astype(np.float32)) and np.array_equal(xl, c['xl'])
    assert abs(e_run / e_fwd - 1) < 1e-6
    e_b = mset.run_batched(3)
    assert abs(e_b / e_fwd - 1) < 1e-6


def test_batched_grad_is_bitwise_repeatable_and_additive(mset):
    c = small_case(n=64, nz=8, binning=2, npos=9, seed=5)
    _upload_case(mset, c)
    g1 = mset.grad()
    g2 = mset.grad()
    assert np.array_equal(g1[0], g2[0]) and np.array_equal(g1[1], g2[1])
    a = mset.grad(0, 4)
    b = mset.grad(4, 5)
    assert rel(a[0] + b[0], g1[0]) < 1e-5
    assert rel(np.r_[a[2], b[2]], g1[2]) < 1e-6


def test_mini_batch_of_one_equals_sequential(mset):
    c = small_case(n=32, nz=6, binning=2, npos=5, seed=8)
    steps = (30.0, 2e-3, 2e-2)
    _upload_case(mset, c, steps=steps)
    V1, p1, x1, y1, e1 = mset.run()
    _upload_case(mset, c, steps=steps)
    e2 = mset.run_batched(1)
    V2, p2, x2, y2 = mset.download()
    assert rel(V2 - c['V'], V1 - c['V']) < 1e-4
    assert rel(p2 - c['probe'], p1 - c['probe']) < 1e-4
    assert abs(e2 / e1 - 1) < 1e-5


def _oracle_batches(c, steps, sizes):
    """Mini-batches of the given sizes in order: each batch's gradient at the state the previous batches
    left, then V += s1 dV (replicated over each bin), probe -= s2 dprobe, the batch's positions += s3 dpos."""
    V, pr = c['V'].copy(), c['probe'].copy()
    xl, yl = c['xl'].astype(np.float64), c['yl'].astype(np.float64)
    first = 0
    for cnt in sizes:
        s = slice(first, first + cnt)
        odV, odp, odpos, _ = O.grad_fixed(V, pr, c['prop'], c['bprop'], xl[s], yl[s], c['M'][s], c['bin'], c['sig'],
                                          c['dx'])
        V = V + steps[0] * np.repeat(odV, c['bin'], axis=0)[:V.shape[0]]
        pr = pr - steps[1] * odp
        xl[s] += steps[2] * odpos[:, 0]
        yl[s] += steps[2] * odpos[:, 1]
        first += cnt
    return V, pr, xl, yl


def test_mini_batches_vs_oracle_composition(mset):
    """7 positions in batches of 3 run 3 + 3 + 1, each batch's gradient at the state the previous batches
    left.  With these steps the partition 3 + 2 + 2 would end 0.7 % away in V and 1.3 % in the probe, so
    the batch boundaries and the update order are pinned by the FP64 composition."""
    c = small_case(n=32, nz=6, binning=2, npos=7, seed=13, spread=4.0)
    steps = (300.0, 2e-2, 2e-2)
    _upload_case(mset, c, steps=steps)
    mset.run_batched(3)
    V, pr, xl, yl = mset.download()
    oV, opr, oxl, oyl = _oracle_batches(c, steps, [3, 3, 1])
    print('case: batches 3 + 3 + 1: V', rel(V - c['V'], oV - c['V']), 'probe', rel(pr - c['probe'], opr - c['probe']),
          flush=True)
    assert_update_close(V, oV, c['V'], 3)
    assert_update_close(pr, opr, c['probe'], 3)
    ulp = np.spacing(np.float32(max(np.abs(c['xl']).max(), np.abs(c['yl']).max())))
    assert np.abs(np.c_[xl - oxl, yl - oyl]).max() <= 4 * ulp


def test_bad_method_and_shapes_raise(mset):
    from mset_amd import lib
    c = small_case(n=32, nz=4, binning=1, npos=2, seed=1)
    _upload_case(mset, c)
    with pytest.raises(lib.MsetError) as e:
        mset.run(method=7)
    assert e.value.code == lib.MSET_EINVAL
    with pytest.raises(lib.MsetError):
        mset.upload(np.zeros((4, 32, 30), np.float32), c['probe'], c['prop'], c['xl'], c['yl'],
                    np.zeros((2, 32, 30), np.float32), c['sig'], c['dx'])
    c40 = small_case(n=40, nz=2, binning=1, npos=1, seed=1)
    with pytest.raises(lib.MsetError) as e:
        _upload_case(mset, c40)
    assert e.value.code == lib.MSET_EUNSUPPORTED


# ------------------------------------------------------------------ SSET ('run_sset' / 'error_sset')
@pytest.mark.parametrize('n,nz,binning,mode', [(32, 6, 2, O.MODE_REF), (64, 5, 1, O.MODE_EXACT),
                                               (128, 8, 4, O.MODE_REF), (256, 4, 2, O.MODE_EXACT)])
def test_sset_sequential_and_error_vs_oracle(mset, n, nz, binning, mode):
    """STEP02_FORWARD_4D_SSET / STEP03_BACK_SSET with the sequential updates of Run_MSET_recon.m:134-158."""
    from mset_amd import lib
    c = small_case(n=n, nz=nz, binning=binning, npos=5, seed=7 * n + nz, spread=n / 8)
    print('case: sset', n, nz, binning, mode, flush=True)
    _upload_case(mset, c, mode=mode)
    e = mset.error(method=lib.METHOD_SSET)
    oe, _ = O.error_tilt(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['npos'], 1, c['bin'], c['sig'],
                         c['dx'], mode=O.MODE_SSET)
    assert abs(e / oe - 1) < LOSS_TOL
    em = mset.error(method=lib.METHOD_MSET)
    assert abs(em / e - 1) > 2 * LOSS_TOL             # the two operators differ on this state
    steps = (30.0, 2e-3, 2e-2)
    _upload_case(mset, c, steps=steps, mode=mode)
    V, pr, xl, yl, err = mset.run(method=lib.METHOD_SSET)
    oV, opr, oxl, oyl, oerr, _ = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                            c['npos'], 1, c['bin'], c['sig'], c['dx'], steps, mode=mode | O.MODE_SSET)
    assert abs(err / oerr - 1) < LOSS_TOL
    assert_update_close(V, oV, c['V'], c['npos'])
    assert_update_close(pr, opr, c['probe'], c['npos'])
    dV = V - c['V']
    assert np.abs(dV - dV[0]).max() <= 1e-5 * np.abs(dV).max() + 1e-5 * np.abs(c['V']).max()   # every slice
    ulp = np.spacing(np.float32(max(np.abs(c['xl']).max(), np.abs(c['yl']).max())))
    d, od = np.c_[xl - c['xl'], yl - c['yl']], np.c_[oxl - c['xl'], oyl - c['yl']]
    assert np.abs(d - od).max() <= max(5e-3 * np.abs(od).max(), 4 * ulp)


@pytest.mark.parametrize('n,nz,binning,npos', [(64, 6, 3, 7), (256, 4, 2, 5)])
def test_sset_batched_step_is_the_fixed_state_gradient(mset, n, nz, binning, npos):
    """One mini-batch over every position: the object step is s1 * sum_k Re(grad_k) on every slice."""
    from mset_amd import lib
    c = small_case(n=n, nz=nz, binning=binning, npos=npos, seed=21 + n, spread=n / 8)
    steps = (30.0, 0.0, 0.0)
    _upload_case(mset, c, steps=steps)
    e = mset.run_batched(c['npos'], method=lib.METHOD_SSET)
    V, _, _, _ = mset.download()
    odV, _, _, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], c['bin'],
                                  c['sig'], c['dx'], mode=O.MODE_REF | O.MODE_SSET)
    assert abs(e / (oep.sum() / c['npos']) - 1) < LOSS_TOL
    assert_update_close(V, c['V'] + steps[0] * odV[0][None], c['V'], 1)


@pytest.mark.parametrize('n,nz,binning,npos,mode', [(64, 3, 5, 1, O.MODE_REF), (32, 1, 1, 2, O.MODE_EXACT),
                                                    (128, 2, 1, 1, O.MODE_REF), (256, 3, 4, 1, O.MODE_EXACT)])
def test_single_plane_and_single_position_edges(mset, n, nz, binning, npos, mode):
    """One plane (binning >= nz, clamped as STEP00_INIT.m:127-129) and one position: the batched grid
    shrinks to the positions and the sweeps have a single step."""
    c = small_case(n=n, nz=nz, binning=binning, npos=npos, seed=7 * n + nz, spread=n / 8)
    print('case: edge', n, nz, binning, npos, mode, flush=True)
    _upload_case(mset, c, mode=mode)
    dV, dp, dpos, err = mset.grad()
    odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                        c['bin'], c['sig'], c['dx'], mode=mode)
    assert dV.shape[0] == -(-nz // min(binning, nz))
    assert abs(err / (oep.sum() / npos) - 1) < LOSS_TOL
    assert rel(dV, odV) < GRAD_TOL
    assert rel(dp, odp) < GRAD_TOL
    assert rel(dpos, odpos) < GRAD_TOL
    # the sequential path agrees with its oracle on the same state
    steps = (30.0, 2e-3, 2e-2)
    _upload_case(mset, c, steps=steps, mode=mode)
    V2, p2, x2, y2, e2 = mset.run()
    oV, op, ox, oy, oe, _ = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], npos, 1,
                                       c['bin'], c['sig'], c['dx'], steps, mode=mode)
    assert abs(e2 / oe - 1) < LOSS_TOL
    assert_update_close(V2, oV, c['V'], npos)
    assert_update_close(p2, op, c['probe'], npos)


@pytest.mark.parametrize('n', [64, 128, 256])
def test_positions_on_and_beyond_the_field_border(mset, n):
    """Scan positions on the field border and up to a field width outside it.  The sub-pixel shift is a
    circular Fourier shift (Func_generate_probe_wave_scan_sub_pixel.m:17-25), so the probe wraps around the
    field; the device's phase ramp, its position terms and the waves must follow the oracle for any shift."""
    c = small_case(n=n, nz=4, binning=2, npos=6, seed=5 * n, spread=1.0)
    c['xl'] = np.array([1.0, n, n + 5.3, -3.7, 0.5, 2 * n - 1.25], np.float32)
    c['yl'] = np.array([n, 1.0, -2.6, n + 7.9, n / 2 + 0.25, 3.5 - n], np.float32)
    print('case: border positions', n, flush=True)
    _upload_case(mset, c)
    dV, dp, dpos, err = mset.grad()
    odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                        c['bin'], c['sig'], c['dx'])
    assert abs(err / (oep.sum() / c['npos']) - 1) < LOSS_TOL
    assert rel(dV, odV) < GRAD_TOL and rel(dp, odp) < GRAD_TOL and rel(dpos, odpos) < GRAD_TOL
    for k in (2, 3, 5):
        w = mset.waves(k)
        psi0 = O.subpixel_shift(c['probe'], c['xl'][k], c['yl'][k])
        save, _, _, _ = O.forward(c['V'], psi0, c['prop'], c['M'][k], c['bin'], c['sig'], c['dx'])
        assert rel(w[0], psi0) < WAVE_TOL and rel(w[-1], save[-1]) < WAVE_TOL, k


def test_separable_256_launch_above_one_position_per_cu(mset):
    """At 256 x 256 a launch of at most one position per CU runs the separable routine's 1024-thread shape,
    a longer one the default 512-thread shape.  The other 256 cases are short; this one (264 positions)
    runs the default shape in one launch and the 1024-thread shape in two halves, against the FP64 oracle."""
    c = small_case(n=256, nz=4, binning=2, npos=264, seed=17, spread=40.0)
    _upload_case(mset, c)
    dV, dp, dpos, err = mset.grad()
    a, b = mset.grad(0, 132), mset.grad(132, 132)
    odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'],
                                        c['bin'], c['sig'], c['dx'])
    print('case: 264 positions at 256: dV', rel(dV, odV), 'dprobe', rel(dp, odp), 'dpos', rel(dpos, odpos), flush=True)
    assert abs(err / (oep.sum() / c['npos']) - 1) < LOSS_TOL
    assert rel(dV, odV) < GRAD_TOL and rel(dp, odp) < GRAD_TOL and rel(dpos, odpos) < GRAD_TOL
    assert rel(a[0] + b[0], dV) < 1e-5 and rel(a[1] + b[1], dp) < 1e-5
    assert rel(np.r_[a[2], b[2]], dpos) < 1e-6


def test_upload_from_page_locked_host_memory(mset):
    """Patterns staged in mset_host_alloc memory upload to the same device state as pageable ones."""
    from mset_amd import lib
    c = small_case(n=64, nz=8, binning=2, npos=9, seed=31, spread=8.0)
    _upload_case(mset, c)
    ref = mset.grad()
    pin = lib.PinnedArray(c['M'].shape, np.float32)
    try:
        pin.array[...] = c['M']
        mset.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], pin.array, c['sig'], c['dx'], slice_binning=c['bin'],
                    n_scan=(c['npos'], 1))
        got = mset.grad()
    finally:
        pin.free()
    for a, b in zip(got[:3], ref[:3]):
        assert np.array_equal(a, b)
    assert got[3] == ref[3]


def test_diffraction_transpose_at_ingest(mset):
    """STEP00_INIT.m:58-73: uploading the patterns with diffraction_transpose set equals uploading their
    transposes, bit for bit."""
    c = small_case(n=64, nz=8, binning=2, npos=5, seed=33, spread=6.0)
    MT = np.ascontiguousarray(np.transpose(c['M'], (0, 2, 1)))
    mset.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], MT, c['sig'], c['dx'], slice_binning=c['bin'],
                n_scan=(c['npos'], 1))
    ref = mset.grad()
    mset.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'], slice_binning=c['bin'],
                n_scan=(c['npos'], 1), diffraction_transpose=True)
    got = mset.grad()
    for a, b in zip(got[:3], ref[:3]):
        assert np.array_equal(a, b)
    assert got[3] == ref[3]
    # and it does change the result for patterns that are not symmetric
    mset.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'], slice_binning=c['bin'],
                n_scan=(c['npos'], 1))
    assert not np.array_equal(mset.grad()[0], ref[0])


@pytest.mark.parametrize('mode', [O.MODE_REF, O.MODE_EXACT])
def test_256_separable_and_general_propagator_paths_vs_oracle(mset, mode):
    """At 256 x 256 the Fresnel propagator is separable and the reference-pairing gradient and the forward
    kernels take the separable routine (one pass per plane); a unit-modulus propagator that does not factor
    sends the same work through the general routine.  Both against the FP64 oracle with their own
    propagator."""
    c = small_case(n=256, nz=10, binning=2, npos=6, seed=77, spread=24.0)
    n = 256
    i = np.arange(n)
    bump = np.exp(0.3j * np.sin(2 * np.pi * np.outer(i, i) / n ** 2 * 7.0))     # does not factor in kx, ky
    for name, prop in (('separable', c['prop']), ('general', c['prop'] * bump)):
        bprop = np.conj(prop)
        print('case: 256 %s propagator, mode %d' % (name, mode), flush=True)
        mset.upload(c['V'], c['probe'], prop, c['xl'], c['yl'], c['M'], c['sig'], c['dx'], slice_binning=c['bin'],
                    n_scan=(c['npos'], 1), grad_mode=mode)
        dV, dp, dpos, err = mset.grad()
        e1 = mset.error()
        odV, odp, odpos, oep = O.grad_fixed(c['V'], c['probe'], prop, bprop, c['xl'], c['yl'], c['M'], c['bin'],
                                            c['sig'], c['dx'], mode=mode)
        oerr = oep.sum() / c['npos']
        assert abs(err / oerr - 1) < LOSS_TOL and abs(e1 / oerr - 1) < LOSS_TOL
        assert rel(dV, odV) < GRAD_TOL and rel(dp, odp) < GRAD_TOL and rel(dpos, odpos) < GRAD_TOL
