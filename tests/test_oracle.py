"""Pins the CPU oracle (oracle/mset_oracle.cpp) before anything is checked against it.

Every oracle function is checked against something it does not itself compute:
  * golden vectors saved by the reference's own run (examples/CuAu_24size_dataset/MSET_recon.mat),
  * a brute-force O(N^2) DFT and numpy's FFT,
  * closed forms of the physics (free propagation, unitarity, global phase, Parseval),
  * central finite differences of the amplitude loss (exact-gradient mode) and an adjoint identity,
  * rotation invariants (round trip, 90-degree index permutation, identity below 0.5 degree).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel

O.set_threads(8)


def _phys(man):
    E0 = man['E0_kV'] * 1e3
    return O.wavelength(E0), O.interaction_sigma(E0), man['potential_pixelsize']


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize('n', [1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 27, 32, 48, 64, 96, 100])
def test_fft1_vs_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert np.abs(O.fft1(x) - W @ x).max() <= 1e-12 * max(1.0, np.abs(W @ x).max())
    assert np.abs(O.fft1(x, inverse=True) - (W.conj() @ x) / n).max() <= 1e-12 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize('n', [8, 24, 64, 128])
def test_fft2_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    assert rel(O.fft2(x), np.fft.fft2(x)) < 1e-14
    assert rel(O.fft2(x, inverse=True), np.fft.ifft2(x)) < 1e-14
    assert rel(O.fft2c(x), np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x)))) < 1e-14
    assert rel(O.fft2_f(x), np.fft.fft2(x)) < 1e-6


def test_fft_axis_vs_numpy():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 10, 12)) + 1j * rng.standard_normal((6, 10, 12))
    for ax in range(3):
        assert rel(O.fft_axis(x, ax), np.fft.fft(x, axis=ax)) < 1e-14
        assert rel(O.fft_axis(x, ax, inverse=True), np.fft.ifft(x, axis=ax)) < 1e-14


# ------------------------------------------------------------------ reference golden vectors
def test_constants_match_saved_state(cuau):
    man = cuau['man']
    lam, sig, _ = _phys(man)
    assert abs(lam / man['lambda_A'] - 1) < 1e-14       # electron_wavelength.m
    assert abs(sig / man['sigma'] - 1) < 1e-14          # interaction_constant.m


def test_fixture_probe_generation(cuau):
    """Func_generate_probe_wave.m:14-56 with C10 = -probeDefocus reproduces the saved probe up to the
    real positive factor of the normalisation (STEP00_INIT.m:214-218, pinned separately by Parseval)."""
    man, s = cuau['man'], cuau['snap']
    lam, _, dx = _phys(man)
    p = O.probe(24, dx, lam, man['alpha_mrad'], C10=-man['probeDefocus_A'])
    for j in range(s['probe_wfn'].shape[0]):
        q = s['probe_wfn'][j].astype(np.complex128)
        a = np.vdot(p, q) / np.vdot(p, p)
        assert abs(a.imag) < 1e-6 * abs(a) and a.real > 0
        assert rel(a * p, q) < 1e-6
    # the defocus sign matters: the opposite sign is far off
    pw = O.probe(24, dx, lam, man['alpha_mrad'], C10=man['probeDefocus_A'])
    a = np.vdot(pw, s['probe_wfn'][0]) / np.vdot(pw, pw)
    assert rel(a * pw, s['probe_wfn'][0]) > 0.5


def test_fixture1_propagator(cuau):
    man, s = cuau['man'], cuau['snap']
    lam, _, dx = _phys(man)
    assert rel(O.propagator(24, dx, lam, man['slice_binning']), s['prop']) < 1e-7
    assert rel(O.propagator(24, dx, lam, man['slice_binning'], back=True), s['back_prop']) < 1e-7


def test_fixture2_transmission(cuau):
    man, s = cuau['man'], cuau['snap']
    _, sig, dx = _phys(man)
    assert rel(O.transmission(s['RVol'], 1, sig, dx), s['trans']) < 1e-6


def test_fixture3_subpixel_shift(cuau):
    man, s = cuau['man'], cuau['snap']
    psi0 = O.subpixel_shift(s['probe_wfn'][man['Nth_angle'] - 1], man['row'], man['col'])
    assert rel(psi0, s['tmp_probe_wfn']) < 5e-6
    psi0f = O.subpixel_shift(s['probe_wfn'][man['Nth_angle'] - 1], man['row'], man['col'], single=True)
    assert rel(psi0f, s['tmp_probe_wfn']) < 5e-6


def test_fixture4_5_forward_and_residual(cuau):
    man, s = cuau['man'], cuau['snap']
    _, sig, dx = _phys(man)
    M = cuau['meas'][man['Nth_angle'] - 1][man['k'] - 1]
    save, ex, resi, _ = O.forward(s['RVol'], s['tmp_probe_wfn'], s['prop'], M, 1, sig, dx)
    for a in range(save.shape[0]):
        assert rel(save[a], s['save_wave_f'][a]) < 5e-6, a
    # resi_vec divides by sqrt(I) computed in single precision by the reference: ~1e-4 near zeros of I
    assert rel(resi, s['resi_vec']) < 5e-4


def test_fixture6_inverse_rotation(cuau):
    s = cuau['snap']
    RV = O.rot_3dvol(s['rec'], s['tilt_angles'][7], inverse=True)
    assert rel(RV, s['RVol']) < 5e-6
    ns, dims, _ = O.shear_plan(O.rot_matrix(s['tilt_angles'][7], inverse=True)[0])
    assert ns == 5      # floor(52.5/11)+1


def _error_pass(cuau, rec, single=False):
    man, s = cuau['man'], cuau['snap']
    _, sig, dx = _phys(man)
    errs = []
    for j in range(8):
        RV = O.rot_3dvol(rec, s['tilt_angles'][j], inverse=True, single=single)
        e, _ = O.error_tilt(RV, s['probe_wfn'][j], s['prop'], s['scan_pos'][j, 0], s['scan_pos'][j, 1],
                            cuau['meas'][j], 24, 24, 1, sig, dx, single=single)
        errs.append(e)
    return np.array(errs)


def test_fixture7_per_tilt_error(cuau):
    errs = _error_pass(cuau, cuau['snap']['rec'])
    assert np.abs(errs / cuau['snap']['error'] - 1).max() < 2e-5
    errs_f = _error_pass(cuau, cuau['snap']['rec'], single=True)
    assert np.abs(errs_f / cuau['snap']['error'] - 1).max() < 5e-5


@pytest.mark.slow
def test_fixture8_mean_error_history(cuau):
    rs = cuau['recon']['rec_save']
    got = np.array([_error_pass(cuau, rs[i]).mean() for i in range(rs.shape[0])])
    assert np.abs(got / cuau['recon']['mean_error'] - 1).max() < 2e-5


# ------------------------------------------------------------------ closed forms
def _small_case(n=16, nz=6, binning=2, npos=3, strength=0.3, seed=0):
    """A small strong-phase state for derivative checks (phases ~strength rad per plane)."""
    rng = np.random.default_rng(seed)
    lam, sig, dx = O.wavelength(300e3), O.interaction_sigma(300e3), 0.5
    V = rng.uniform(0, 1, (nz, n, n)) * strength / (sig * dx * binning)
    p = O.probe(n, dx, lam, 60.0, C10=20.0)
    p = O.probe_normalise(p, 1e3)
    xl = (n / 2 + 1 + rng.uniform(-2, 2, npos)).astype(np.float32)
    yl = (n / 2 + 1 + rng.uniform(-2, 2, npos)).astype(np.float32)
    M = rng.uniform(0.5, 1.5, (npos, n, n)) * 1e3 / n ** 2
    prop = O.propagator(n, dx, lam, binning)
    bprop = O.propagator(n, dx, lam, binning, back=True)
    return dict(V=V, probe=p, xl=xl, yl=yl, M=M.astype(np.float32), prop=prop, bprop=bprop, bin=binning,
                sig=sig, dx=dx, n=n, lam=lam)


def test_free_space_exit_intensity():
    """t = 1 (V = 0) and |prop| = 1: |Psi|^2 equals |fftshift(fft2(ifftshift(psi0)))|^2."""
    c = _small_case()
    V0 = np.zeros_like(c['V'])
    psi0 = O.subpixel_shift(c['probe'], 7.3, 9.6)
    save, ex, _, _ = O.forward(V0, psi0, c['prop'], c['M'][0], c['bin'], c['sig'], c['dx'])
    # the propagator is rounded to single precision (Func_generate_propagator.m:36), |p| = 1 +- 6e-8
    assert rel(np.abs(ex) ** 2, np.abs(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(psi0)))) ** 2) < 1e-6
    # unitarity: every plane keeps the norm of the incident wave
    for a in range(save.shape[0]):
        assert abs(np.linalg.norm(save[a]) / np.linalg.norm(psi0) - 1) < 1e-6
    # with an exactly unimodular propagator both identities hold to rounding
    pu = np.exp(1j * np.angle(c['prop']))
    save, ex, _, _ = O.forward(V0, psi0, pu, c['M'][0], c['bin'], c['sig'], c['dx'])
    assert rel(np.abs(ex) ** 2, np.abs(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(psi0)))) ** 2) < 1e-13


def test_constant_potential_is_global_phase():
    c = _small_case()
    psi0 = O.subpixel_shift(c['probe'], 8.0, 8.0)
    _, ex0, _, e0 = O.forward(np.zeros_like(c['V']), psi0, c['prop'], c['M'][0], c['bin'], c['sig'], c['dx'])
    _, ex1, _, e1 = O.forward(np.full_like(c['V'], 123.0), psi0, c['prop'], c['M'][0], c['bin'], c['sig'], c['dx'])
    assert rel(np.abs(ex1), np.abs(ex0)) < 1e-12 and abs(e1 / e0 - 1) < 1e-12
    ph = c['sig'] * c['dx'] * 123.0 * c['V'].shape[0]
    assert rel(ex1, ex0 * np.exp(1j * ph)) < 1e-12


def test_subpixel_shift_integer_is_roll_and_fraction_is_bandlimited_translation():
    c = _small_case(n=32)
    p = c['probe']
    n = 32
    cen = n // 2 + 1                        # round((n+1)/2) for even n
    s = O.subpixel_shift(p, cen + 3, cen - 5)
    assert rel(s, np.roll(np.roll(p, 3, axis=1), -5, axis=0)) < 1e-12
    # fractional shift of a band-limited Gaussian against its analytic translate
    x = np.arange(n) - n / 2
    X, Y = np.meshgrid(x, x)
    g = lambda x0, y0: np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2 * 2.5 ** 2))
    s = O.subpixel_shift(g(0, 0), cen + 0.37, cen - 1.61)
    assert rel(s, g(0.37, -1.61)) < 1e-6


def test_probe_normalisation_parseval():
    c = _small_case()
    p = O.probe_normalise(c['probe'], 777.0)
    assert abs(np.sum(np.abs(np.fft.fft2(p)) ** 2) / 777.0 - 1) < 1e-12


# ------------------------------------------------------------------ gradients
def _grad(c, mode):
    return O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], c['bin'], c['sig'],
                        c['dx'], mode=mode)


def _loss(c, V=None, probe=None, xl=None, yl=None):
    return O.loss(c['V'] if V is None else V, c['probe'] if probe is None else probe, c['prop'],
                  c['xl'] if xl is None else xl, c['yl'] if yl is None else yl, c['M'], c['bin'], c['sig'], c['dx'])


def test_exact_object_gradient_vs_central_fd():
    c = _small_case()
    n = c['n']
    dV, _, _, _ = _grad(c, O.MODE_EXACT)
    # dL/dV_v = -2 n^2 sigma dx Re(g_exact) for every voxel of the bin
    rng = np.random.default_rng(1)
    h = 1e-2
    for _ in range(6):
        d = rng.standard_normal(c['V'].shape)
        fd = (_loss(c, V=c['V'] + h * d) - _loss(c, V=c['V'] - h * d)) / (2 * h)
        g = np.repeat(dV, c['bin'], axis=0)[:c['V'].shape[0]]
        an = -2 * n * n * c['sig'] * c['dx'] * np.sum(g * d)
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-9, (fd, an)


def test_reference_object_gradient_pairs_the_wave_after_each_plane():
    """STEP03_BACK.m:14-15 uses conj(save_wave_f(:,:,a)), the wave AFTER plane a.

    With the propagator set to 1 the wave after plane a is t_a times the wave entering it, so the
    reference term must equal conj(t_a) times the exact term (which the FD test pins), plane by plane.
    With real propagation the two differ, i.e. the reference direction is not the loss gradient."""
    c = _small_case(nz=6, binning=2)
    one = np.ones_like(c['prop'])
    psi0 = O.subpixel_shift(c['probe'], c['xl'][0], c['yl'][0])
    save, _, resi, _ = O.forward(c['V'], psi0, one, c['M'][0], c['bin'], c['sig'], c['dx'])
    gr, bwr, _ = O.back(c['V'], save, psi0, resi, one, c['bin'], c['sig'], c['dx'], c['xl'][0], c['yl'][0], O.MODE_REF)
    ge, bwe, _ = O.back(c['V'], save, psi0, resi, one, c['bin'], c['sig'], c['dx'], c['xl'][0], c['yl'][0], O.MODE_EXACT)
    t = np.repeat(O.transmission(c['V'], c['bin'], c['sig'], c['dx']), c['bin'], axis=0)
    assert rel(gr, np.conj(t) * ge) < 1e-12
    assert rel(bwr, bwe) < 1e-15                       # the back-propagated wave is mode independent
    dVr, _, _, _ = _grad(c, O.MODE_REF)
    dVe, _, _, _ = _grad(c, O.MODE_EXACT)
    assert rel(dVr, dVe) > 0.05


def test_probe_gradient_vs_central_fd():
    c = _small_case()
    n = c['n']
    _, dp, _, _ = _grad(c, O.MODE_REF)
    rng = np.random.default_rng(2)
    h = 1e-6 * np.abs(c['probe']).max()
    for _ in range(4):
        d = rng.standard_normal(dp.shape) + 1j * rng.standard_normal(dp.shape)
        fd = (_loss(c, probe=c['probe'] + h * d) - _loss(c, probe=c['probe'] - h * d)) / (2 * h)
        an = 2 * n * n * np.real(np.sum(np.conj(dp) * d))
        assert abs(fd - an) <= 1e-7 * abs(fd) + 1e-9, (fd, an)


def test_exact_position_gradient_vs_central_fd():
    c = _small_case()
    n = c['n']
    _, _, dpos, _ = _grad(c, O.MODE_EXACT)
    h = 1e-5
    for k in range(len(c['xl'])):
        for ax in range(2):
            lp = [np.array(c['xl'], np.float64), np.array(c['yl'], np.float64)]
            lm = [a.copy() for a in lp]
            lp[ax][k] += h
            lm[ax][k] -= h
            fd = (_loss(c, xl=lp[0], yl=lp[1]) - _loss(c, xl=lm[0], yl=lm[1])) / (2 * h)
            # row += s3 * dpos is a descent step: dpos = -(dL/drow) / (2 n^2)
            an = -2 * n * n * dpos[k, ax]
            assert abs(fd - an) <= 2e-5 * abs(fd) + 1e-7, (k, ax, fd, an)


def test_reference_position_gradient_approximates_exact():
    """STEP03_BACK.m:89-92 uses MATLAB gradient() (a finite difference).  For a smooth band-limited
    probe away from the field edges the finite difference must agree with the spectral derivative;
    a swapped axis or a sign error would not."""
    c = _small_case(n=64, npos=4)
    p = O.probe(64, c['dx'], c['lam'], 4.0, C12=300.0, Phi12=0.6)    # k_max dx ~ 0.1: FD error ~ 5 %
    c['probe'] = O.probe_normalise(p, 1e3)
    _, _, dr, _ = _grad(c, O.MODE_REF)
    _, _, de, _ = _grad(c, O.MODE_EXACT)
    assert rel(dr, de) < 0.06, (dr, de)
    assert rel(dr[:, ::-1], de) > 0.3 and rel(-dr, de) > 1.5       # discriminates axis swap and sign


# ------------------------------------------------------------------ SSET (single-slice operator)
def test_sset_is_one_plane_of_the_multislice_without_propagation():
    """STEP02_FORWARD_4D_SSET / STEP03_BACK_SSET: the whole depth as one transmission and no propagator.
    That is the multislice operator with slice_binning = nz and p = 1; the object term pairs the incident
    wave (the exact-adjoint pairing) and the position terms use the same formula as STEP03_BACK."""
    c = _small_case(nz=5, binning=2)
    one = np.ones_like(c['prop'])
    nz = c['V'].shape[0]
    for mode in (O.MODE_REF, O.MODE_EXACT):
        s = O.grad_fixed(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], c['bin'], c['sig'],
                         c['dx'], mode=mode | O.MODE_SSET)
        m_obj = O.grad_fixed(c['V'], c['probe'], one, one, c['xl'], c['yl'], c['M'], nz, c['sig'], c['dx'],
                             mode=O.MODE_EXACT)
        m_pos = O.grad_fixed(c['V'], c['probe'], one, one, c['xl'], c['yl'], c['M'], nz, c['sig'], c['dx'], mode=mode)
        assert s[0].shape == (1,) + c['V'].shape[1:]
        assert rel(s[0], m_obj[0]) < 1e-10 and rel(s[1], m_obj[1]) < 1e-10
        assert rel(s[2], m_pos[2]) < 1e-10 and rel(s[3], m_obj[3]) < 1e-10
    # SSET pairs the incident wave even in reference mode: the multislice reference pairing (the wave
    # after the plane) gives a different object term
    m_ref = O.grad_fixed(c['V'], c['probe'], one, one, c['xl'], c['yl'], c['M'], nz, c['sig'], c['dx'],
                         mode=O.MODE_REF)
    assert rel(s[0], m_ref[0]) > 1e-2


def test_sset_object_gradient_vs_central_fd():
    c = _small_case(nz=4, binning=1)
    n = c['n']
    dV, _, _, _ = _grad(c, O.MODE_REF | O.MODE_SSET)
    rng = np.random.default_rng(4)
    h = 1e-2
    lossf = lambda V: O.loss(V, c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['bin'], c['sig'], c['dx'],
                             mode=O.MODE_SSET)
    for _ in range(4):
        d = rng.standard_normal(c['V'].shape)
        fd = (lossf(c['V'] + h * d) - lossf(c['V'] - h * d)) / (2 * h)
        an = -2 * n * n * c['sig'] * c['dx'] * np.sum(dV[0] * d)      # one plane, every slice
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-9, (fd, an)


def test_sset_run_tilt_single_matches_double_and_updates_every_slice():
    c = _small_case(nz=4, binning=2, npos=3)
    steps = (30.0, 2e-3, 2e-2)
    out_d = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], 3, 1, c['bin'], c['sig'],
                       c['dx'], steps, mode=O.MODE_SSET)
    out_f = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], 3, 1, c['bin'], c['sig'],
                       c['dx'], steps, mode=O.MODE_SSET, single=True)
    dV = out_d[0] - c['V']
    assert np.abs(dV - dV[0]).max() <= 1e-12 * np.abs(dV).max()       # grad is replicated over all slices
    assert rel(out_f[0] - c['V'], dV) < 1e-4 and abs(out_f[4] / out_d[4] - 1) < 1e-5


def test_adjoint_identity_of_one_plane():
    """<A u, y> = <u, A^H y> for A u = ifft2(p . fft2(t . u)) (the forward plane operator)."""
    c = _small_case(nz=1, binning=1)
    n = c['n']
    rng = np.random.default_rng(5)
    t = O.transmission(c['V'], 1, c['sig'], c['dx'])[0]
    p = np.fft.ifftshift(c['prop'])
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    y = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Au = np.fft.ifft2(p * np.fft.fft2(t * u))
    AHy = np.conj(t) * np.fft.ifft2(np.conj(p) * np.fft.fft2(y))
    assert abs(np.vdot(y, Au) - np.vdot(AHy, u)) < 1e-12 * abs(np.vdot(y, Au))


# ------------------------------------------------------------------ sequential loop
def test_run_tilt_single_matches_double():
    c = _small_case(npos=4)
    args = (c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], 2, 2, c['bin'], c['sig'], c['dx'],
            [5.0, 1e-3, 1e-2])
    Vd, pd, xd, yd, ed, _ = O.run_tilt(*args)
    Vf, pf, xf, yf, ef, _ = O.run_tilt(*args, single=True)
    assert rel(Vf, Vd) < 1e-5 and rel(pf, pd) < 1e-5 and abs(ef / ed - 1) < 1e-5
    assert np.abs(xf - xd).max() < 1e-4 and np.abs(yf - yd).max() < 1e-4


def _sequential_by_hand(c, steps, mode, stale=False):
    """Run_MSET_recon.m:134-158 spelled out as a loop of one-position gradients: for each position k the
    gradient is taken at the state left by positions 0..k-1, then V += s1 Re(grad) on every voxel of each
    bin (STEP03_BACK.m:17-21), probe -= s2 shift(bw, -d) (STEP03_BACK.m:29-31) and (row, col) += s3 terms
    (STEP03_BACK.m:35-38).  stale: every gradient at the initial state (the state never advances)."""
    V = np.array(c['V'], np.float64)
    pr = np.array(c['probe'], np.complex128)
    xl = np.array(c['xl'], np.float32)
    yl = np.array(c['yl'], np.float32)
    V0, p0 = V.copy(), pr.copy()
    zbin = np.arange(V.shape[0]) // c['bin']
    for k in range(len(xl)):
        dV, dp, dpos, _ = O.grad_fixed(V0 if stale else V, p0 if stale else pr, c['prop'], c['bprop'],
                                       xl[k:k + 1].astype(np.float64), yl[k:k + 1].astype(np.float64),
                                       c['M'][k:k + 1], c['bin'], c['sig'], c['dx'], mode=mode)
        V = V + steps[0] * dV[zbin]
        pr = pr - steps[1] * dp
        xl[k] = np.float32(np.float64(xl[k]) + steps[2] * dpos[0, 0])
        yl[k] = np.float32(np.float64(yl[k]) + steps[2] * dpos[0, 1])
    return V, pr, xl, yl


@pytest.mark.parametrize('mode', [O.MODE_REF, O.MODE_EXACT])
def test_run_tilt_is_the_sequential_composition_of_one_position_gradients(mode):
    """Pins the update order of oracle run_tilt: after each position the volume, the probe and that
    position are advanced before the next position's forward pass.  A loop that did not advance the state,
    or applied a step before taking the gradient, fails the first comparison; the stale loop shows the
    steps are large enough for that to be visible."""
    c = _small_case(npos=5, seed=11)
    steps = [40.0, 2e-3, 5e-2]
    V, pr, xl, yl, _, _ = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], 5, 1,
                                     c['bin'], c['sig'], c['dx'], steps, mode=mode)
    hV, hp, hx, hy = _sequential_by_hand(c, steps, mode)
    assert rel(V - c['V'], hV - c['V']) < 1e-10 and rel(pr - c['probe'], hp - c['probe']) < 1e-10
    assert np.array_equal(xl, hx) and np.array_equal(yl, hy)
    sV, sp, sx, _ = _sequential_by_hand(c, steps, mode, stale=True)
    assert rel(sV - c['V'], V - c['V']) > 1e-3 and rel(sp - c['probe'], pr - c['probe']) > 1e-3
    x0 = np.asarray(c['xl'], np.float64)
    assert rel(sx - x0, xl - x0) > 1e-3


def test_reference_position_term_uses_matlab_gradient_edges():
    """STEP03_BACK.m:36-38: [de_probey, de_probex] = gradient(conj(psi0)); row += Re sum(de_probex .* bw),
    col += Re sum(de_probey .* bw).  numpy.gradient with edge_order=1 is MATLAB's gradient (central
    difference inside, one-sided first difference on the first and last sample).  A white-spectrum
    incident wave puts energy on the field border, so a wrap-around or zeroed edge rule fails."""
    c = _small_case(n=16, nz=4, binning=2, npos=1, seed=4)
    n = c['n']
    rng = np.random.default_rng(21)
    psi0 = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 3.0
    save, _, resi, _ = O.forward(c['V'], psi0, c['prop'], c['M'][0], c['bin'], c['sig'], c['dx'])
    _, bw, dpos = O.back(c['V'], save, psi0, resi, c['bprop'], c['bin'], c['sig'], c['dx'], 9.0, 9.0,
                         mode=O.MODE_REF)
    cp = np.conj(psi0)                                  # [y][x]: x (MATLAB dim 1, row) is the last axis
    gx = np.gradient(cp, axis=1, edge_order=1)
    gy = np.gradient(cp, axis=0, edge_order=1)
    want = np.array([np.real(np.sum(gx * bw)), np.real(np.sum(gy * bw))])
    assert np.abs(dpos - want).max() <= 1e-12 * np.abs(want).max(), (dpos, want)
    # the same sums with a periodic central difference differ at this level
    px = (np.roll(cp, -1, axis=1) - np.roll(cp, 1, axis=1)) / 2
    py = (np.roll(cp, -1, axis=0) - np.roll(cp, 1, axis=0)) / 2
    wrap = np.array([np.real(np.sum(px * bw)), np.real(np.sum(py * bw))])
    assert np.abs(wrap - want).max() > 0.05 * np.abs(want).max()


def test_run_tilt_with_zero_steps_is_the_error_pass():
    c = _small_case(npos=4)
    V, p, xl, yl, e, ep = O.run_tilt(c['V'], c['probe'], c['prop'], c['bprop'], c['xl'], c['yl'], c['M'], 2, 2,
                                     c['bin'], c['sig'], c['dx'], [0, 0, 0])
    e2, ep2 = O.error_tilt(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], 2, 2, c['bin'], c['sig'], c['dx'])
    assert np.array_equal(V, c['V']) and abs(e / e2 - 1) < 1e-14 and np.allclose(ep, ep2, rtol=1e-14)


# ------------------------------------------------------------------ rotation
def test_rotation_round_trip_and_small_angle_identity():
    rng = np.random.default_rng(7)
    v = rng.standard_normal((12, 16, 20))
    ang = (10.0, -33.0, 5.0)
    r = O.rot_3dvol(v, ang, inverse=True)
    assert rel(O.rot_3dvol(r, ang), v) < 1e-12
    assert np.array_equal(O.rot_3dvol(v, (0.2, -0.4, 0.5)), v)


def test_rotation_by_90_degrees_is_an_index_permutation():
    """Quarter turns map the odd-size grid onto itself; a compact blob must land on np.rot90 of itself."""
    n = 25
    x = np.arange(n) - n // 2
    Z, Y, X = np.meshgrid(x, x, x, indexing='ij')
    blob = np.exp(-((X - 3) ** 2 + (Y + 2) ** 2 * 0.5 + (Z - 1) ** 2 * 2) / 3.0)
    assert rel(O.rot_3dvol(blob, (90.0, 0.0, 0.0)), np.rot90(blob, 3, axes=(1, 2))) < 1e-3
    assert rel(O.rot_3dvol(blob, (0.0, 90.0, 0.0)), np.rot90(blob, 1, axes=(0, 2))) < 3e-2
    assert rel(O.rot_3dvol(blob, (0.0, 0.0, 90.0)), np.rot90(blob, 3, axes=(0, 1))) < 2e-2


@pytest.mark.parametrize('ang', [(0.0, 30.0, 0.0), (20.0, 0.0, 0.0), (0.0, 0.0, 25.0), (10.0, -20.0, 30.0)])
def test_rotation_of_gaussian_blob_matches_analytic(ang):
    """A well-sampled Gaussian at c is carried to (R1 R2 R3) c, i.e. R' c with R = (R1 R2 R3)'."""
    n = 32
    x = np.arange(n) - n // 2
    Z, Y, X = np.meshgrid(x, x, x, indexing='ij')
    blob = lambda c: np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * 2.0 ** 2))
    c0 = np.array([4.0, -3.0, 2.0])   # (x, y, z) in voxels from the centre voxel n/2
    R, _ = O.rot_matrix(ang)
    assert rel(O.rot_3dvol(blob(c0), ang), blob(R.T @ c0)) < 1e-6
    assert rel(O.rot_3dvol(blob(c0), ang, inverse=True), blob(R @ c0)) < 1e-6


# ------------------------------------------------------------------ TV regulariser (TV_method.m)
def test_tv_constant_volume_is_a_fixed_point():
    f = np.full((6, 5, 7), 0.37)
    out, err = O.tv(f, 0.1, 50)
    assert np.array_equal(out, f) and err == 0.0


def test_tv_preserves_the_mean_and_denoises():
    rng = np.random.default_rng(9)
    f = rng.uniform(0, 1, (8, 10, 12))
    out, err = O.tv(f, 0.05, 100)
    assert abs(out.sum() / f.sum() - 1) < 1e-12          # sum of a periodic divergence is zero
    tv_of = lambda u: sum(np.sqrt((np.roll(u, -1, 2) - u) ** 2 + (np.roll(u, -1, 1) - u) ** 2
                              + (np.roll(u, -1, 0) - u) ** 2).ravel())
    assert tv_of(out) < 0.7 * tv_of(f)
    assert abs(err / (0.05 * tv_of(out)) - 1) < 1e-12     # error = lambda * TV(output)


def test_tv_periodic_step_converges_to_the_rof_solution():
    """A step along z with two periodic jumps: the ROF minimiser keeps the jumps and moves each plateau of
    length n/2 by 2 lambda/(n/2) towards the other (0.5 lambda^-1 |u - f|^2 + TV).  A wrong shift direction
    or a div that is not the adjoint of the difference would not converge to it."""
    nz, lam, a, b = 16, 0.4, 1.0, 0.0
    f = np.zeros((nz, 3, 4))
    f[: nz // 2] = a
    f[nz // 2:] = b
    out, err = O.tv(f, lam, 4000)
    alpha, beta = a - 2 * lam / (nz / 2), b + 2 * lam / (nz / 2)
    assert np.abs(out[: nz // 2] - alpha).max() < 1e-3 and np.abs(out[nz // 2:] - beta).max() < 1e-3
    assert abs(err / (lam * 3 * 4 * 2 * (alpha - beta)) - 1) < 1e-3
    # the same step along x and along y (axes are not interchanged)
    for ax in (1, 2):
        g = np.moveaxis(f, 0, ax)
        o2, _ = O.tv(np.ascontiguousarray(g), lam, 4000)
        assert np.abs(np.moveaxis(o2, ax, 0) - out).max() < 1e-3


def test_tv_single_matches_double():
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 1, (6, 8, 8))
    od, ed = O.tv(f, 0.1, 300)
    of, ef = O.tv(f, 0.1, 300, single=True)
    assert rel(of, od) < 1e-5 and abs(ef / ed - 1) < 1e-5
