"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product package.  Arrays follow the C-order, MATLAB-dims-reversed convention: a MATLAB
volume V(x, y, z) is a numpy array v[z, y, x]; a plane A(x, y) is a[y, x]; complex arrays are numpy
complex128 (double oracle) or complex64 (float restatement).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODE_REF = 0     # STEP03_BACK.m:14-15 pairing (wave after plane a)
MODE_EXACT = 1   # true adjoint (wave entering plane a), spectral position derivative
MODE_SSET = 2    # flag: single-slice operator (STEP02_FORWARD_4D_SSET.m / STEP03_BACK_SSET.m)


def build():
    import subprocess
    subprocess.run(['make', '-C', _HERE, '-s'], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        sig = {
            'oracle_set_threads': ([I], I),
            'oracle_fft1_d': ([P, I, I], None), 'oracle_fft2_d': ([P, I, I], None),
            'oracle_fft2_f': ([P, I, I], None), 'oracle_fft2c_d': ([P, I, I], None),
            'oracle_fft_axis_d': ([P, I, I, I, I, I], None),
            'oracle_wavelength': ([D], D), 'oracle_sigma': ([D], D),
            'oracle_propagator': ([P, I, D, D, I, I], None),
            'oracle_probe': ([P, I, D, D, D, P], None),
            'oracle_probe_normalise': ([P, I, D], None),
            'oracle_subpixel_d': ([P, P, I, D, D], None), 'oracle_subpixel_f': ([P, P, I, D, D], None),
            'oracle_transmission_d': ([I, I, I, D, D, P, P], None),
            'oracle_forward_d': ([I, I, I, D, D, P, P, P, P, P, P, P], D),
            'oracle_back_d': ([I, I, I, D, D, I, P, P, P, P, P, P, P, D, D, P], None),
            'oracle_run_tilt_d': ([I, I, I, D, D, I, P, P, P, P, P, P, P, I, I, I, P, P, P], I),
            'oracle_run_tilt_f': ([I, I, I, D, D, I, P, P, P, P, P, P, P, I, I, I, P, P, P], I),
            'oracle_error_tilt_d': ([I, I, I, D, D, I, P, P, P, P, P, P, I, I, I, P, P], I),
            'oracle_error_tilt_f': ([I, I, I, D, D, I, P, P, P, P, P, P, I, I, I, P, P], I),
            'oracle_grad_fixed_d': ([I, I, I, D, D, I, P, P, P, P, P, P, P, I, P, P, P, P], I),
            'oracle_loss_d': ([I, I, I, D, D, I, P, P, P, P, P, P, I], D),
            'oracle_tv_d': ([P, P, I, I, I, D, I], D), 'oracle_tv_f': ([P, P, I, I, I, D, I], D),
            'oracle_rot_matrix': ([P, P, P, D, D, D, I, P], I),
            'oracle_shear_plan': ([P, D, P, P, P], None),
            'oracle_rotate_d': ([P, P, I, I, I, P, D], None),
            'oracle_rotate_f': ([P, P, I, I, I, P, D], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n):
    return lib().oracle_set_threads(int(n))


# ---------------------------------------------------------------- transforms
def fft1(x, inverse=False):
    a = _c128(x).copy()
    lib().oracle_fft1_d(_p(a), a.shape[0], 1 if inverse else -1)
    return a


def fft2(x, inverse=False):
    a = _c128(x).copy()
    lib().oracle_fft2_d(_p(a), a.shape[0], 1 if inverse else -1)
    return a


def fft2_f(x, inverse=False):
    a = _c64(x).copy()
    lib().oracle_fft2_f(_p(a), a.shape[0], 1 if inverse else -1)
    return a


def fft2c(x, inverse=False):
    """fftshift(fftn(ifftshift(x))) (or the ifftn twin), as written throughout STEP02/STEP03."""
    a = _c128(x).copy()
    lib().oracle_fft2c_d(_p(a), a.shape[0], 1 if inverse else -1)
    return a


def fft_axis(x, axis, inverse=False):
    """1D transform along numpy axis (0 = slowest) of a 3D complex array."""
    a = _c128(x).copy()
    n2, n1, n0 = a.shape
    lib().oracle_fft_axis_d(_p(a), n0, n1, n2, 2 - axis, 1 if inverse else -1)
    return a


# ---------------------------------------------------------------- physics inputs
def wavelength(E0_eV):
    return lib().oracle_wavelength(float(E0_eV))


def interaction_sigma(E0_eV):
    return lib().oracle_sigma(float(E0_eV))


def propagator(n, dx, lam, binning, back=False):
    """Func_generate_propagator.m, MATLAB (centred) layout, values rounded to single."""
    out = np.empty((n, n), np.complex128)
    lib().oracle_propagator(_p(out), n, dx, lam, int(binning), int(back))
    return out


def probe(n, dx, lam, alpha_mrad, C10=0.0, C30=0.0, C50=0.0, C12=0.0, Phi12=0.0, C21=0.0, Phi21=0.0,
          C23=0.0, Phi23=0.0):
    """Func_generate_probe_wave.m for one tilt (C10 = -probeDefocus)."""
    out = np.empty((n, n), np.complex128)
    ab = np.array([C10, C30, C50, C12, Phi12, C21, Phi21, C23, Phi23], np.float64)
    lib().oracle_probe(_p(out), n, dx, lam, alpha_mrad, _p(ab))
    return out


def probe_normalise(p, mean_intensity):
    a = _c128(p).copy()
    lib().oracle_probe_normalise(_p(a), a.shape[0], float(mean_intensity))
    return a


def subpixel_shift(p, row, col, single=False):
    """Func_generate_probe_wave_scan_sub_pixel(p, row, col)."""
    if single:
        a = _c64(p)
        out = np.empty_like(a)
        lib().oracle_subpixel_f(_p(a), _p(out), a.shape[0], float(row), float(col))
        return out
    a = _c128(p)
    out = np.empty_like(a)
    lib().oracle_subpixel_d(_p(a), _p(out), a.shape[0], float(row), float(col))
    return out


def transmission(V, binning, sigma, dx):
    V = _f64(V)
    nz, n, _ = V.shape
    L = -(-nz // binning)
    out = np.empty((L, n, n), np.complex128)
    lib().oracle_transmission_d(n, nz, int(binning), sigma, dx, _p(V), _p(out))
    return out


# ---------------------------------------------------------------- operator
def forward(V, psi0, prop, M, binning, sigma, dx):
    """STEP02_FORWARD_4D from the shifted incident wave. Returns (save_waves, exit_centred, resi, err_mean)."""
    V = _f64(V)
    nz, n, _ = V.shape
    L = -(-nz // binning)
    psi0, prop, M = _c128(psi0), _c128(prop), _f64(M)
    save = np.empty((L, n, n), np.complex128)
    ex = np.empty((n, n), np.complex128)
    resi = np.empty((n, n), np.complex128)
    e = lib().oracle_forward_d(n, nz, int(binning), sigma, dx, _p(V), _p(psi0), _p(prop), _p(M), _p(save), _p(ex), _p(resi))
    return save, ex, resi, e


def back(V, save, psi0, resi, back_prop, binning, sigma, dx, row, col, mode=MODE_REF):
    """STEP03_BACK object gradient (complex, nz planes), back wave, position terms (drow, dcol)."""
    V = _f64(V)
    nz, n, _ = V.shape
    grad = np.empty((nz, n, n), np.complex128)
    bw = np.empty((n, n), np.complex128)
    dpos = np.empty(2, np.float64)
    lib().oracle_back_d(n, nz, int(binning), sigma, dx, int(mode), _p(V), _p(_c128(save)), _p(_c128(psi0)),
                        _p(_c128(resi)), _p(_c128(back_prop)), _p(grad), _p(bw), float(row), float(col), _p(dpos))
    return grad, bw, dpos


def run_tilt(V, probe_, prop, back_prop, xlist, ylist, M, nsx, nsy, binning, sigma, dx, steps, mode=MODE_REF,
             single=False):
    """Run_MSET_recon.m:133-158 for one tilt: sequential per-position forward/back/updates.

    Returns (V, probe, xlist, ylist, err_acc, err_pos); inputs are not modified.
    """
    if single:
        V = _f32(V).copy(); pr = _c64(probe_).copy(); pp = _c64(prop); bp = _c64(back_prop)
        fn = lib().oracle_run_tilt_f
    else:
        V = _f64(V).copy(); pr = _c128(probe_).copy(); pp = _c128(prop); bp = _c128(back_prop)
        fn = lib().oracle_run_tilt_d
    nz, n, _ = V.shape
    xl = _f32(xlist).copy(); yl = _f32(ylist).copy()
    M = _f32(M)
    npos = xl.shape[0]
    steps = _f64(steps)
    acc = np.zeros(1, np.float64)
    ep = np.zeros(npos, np.float64)
    fn(n, nz, int(binning), sigma, dx, int(mode), _p(V), _p(pr), _p(pp), _p(bp), _p(xl), _p(yl), _p(M), npos,
       int(nsx), int(nsy), _p(steps), _p(acc), _p(ep))
    return V, pr, xl, yl, float(acc[0]), ep


def error_tilt(V, probe_, prop, xlist, ylist, M, nsx, nsy, binning, sigma, dx, single=False, mode=MODE_REF):
    """Run_MSET_recon.m:207-222: forward + loss for every position of one tilt. Returns (err, err_pos)."""
    if single:
        V = _f32(V); pr = _c64(probe_); pp = _c64(prop); fn = lib().oracle_error_tilt_f
    else:
        V = _f64(V); pr = _c128(probe_); pp = _c128(prop); fn = lib().oracle_error_tilt_d
    nz, n, _ = V.shape
    xl = _f32(xlist); yl = _f32(ylist); M = _f32(M)
    npos = xl.shape[0]
    acc = np.zeros(1, np.float64)
    ep = np.zeros(npos, np.float64)
    fn(n, nz, int(binning), sigma, dx, int(mode), _p(V), _p(pr), _p(pp), _p(xl), _p(yl), _p(M), npos, int(nsx),
       int(nsy), _p(acc), _p(ep))
    return float(acc[0]), ep


def grad_fixed(V, probe_, prop, back_prop, xlist, ylist, M, binning, sigma, dx, mode=MODE_REF):
    """Gradient sums at fixed (V, probe) over all given positions (no updates).

    Returns (dV[L, n, n] = sum_k Re(grad_k) per plane, dprobe[n, n] = sum_k shift(bw_k, -d_k),
             dpos[npos, 2] = position update terms, err_pos[npos] = mean((sqrt M - sqrt I)^2)).
    """
    V = _f64(V)
    nz, n, _ = V.shape
    L = 1 if mode & MODE_SSET else -(-nz // binning)
    xl = _f64(xlist); yl = _f64(ylist); M = _f32(M)
    npos = xl.shape[0]
    dV = np.empty((L, n, n), np.float64)
    dp = np.empty((n, n), np.complex128)
    dpos = np.empty((npos, 2), np.float64)
    ep = np.empty(npos, np.float64)
    lib().oracle_grad_fixed_d(n, nz, int(binning), sigma, dx, int(mode), _p(V), _p(_c128(probe_)), _p(_c128(prop)),
                              _p(_c128(back_prop)), _p(xl), _p(yl), _p(M), npos, _p(dV), _p(dp), _p(dpos), _p(ep))
    return dV, dp, dpos, ep


def loss(V, probe_, prop, xlist, ylist, M, binning, sigma, dx, mode=MODE_REF):
    """L = sum_k sum_q (|Psi_k(q)| - sqrt(M_k(q)))^2 at a fixed state."""
    V = _f64(V)
    nz, n, _ = V.shape
    xl = _f64(xlist); yl = _f64(ylist)
    return lib().oracle_loss_d(n, nz, int(binning), sigma, dx, int(mode), _p(V), _p(_c128(probe_)), _p(_c128(prop)), _p(xl),
                               _p(yl), _p(_f32(M)), xl.shape[0])


# ---------------------------------------------------------------- total variation
def tv(vol, lam, n_iter=300, single=False):
    """TV_method.m(input, lambda, Niter) on a volume v[z, y, x] -> (output, error)."""
    if single:
        v = _f32(vol); out = np.empty_like(v); fn = lib().oracle_tv_f
    else:
        v = _f64(vol); out = np.empty_like(v); fn = lib().oracle_tv_d
    nz, ny, nx = v.shape
    e = fn(_p(v), _p(out), nx, ny, nz, float(lam), int(n_iter))
    return out, float(e)


# ---------------------------------------------------------------- rotation
def rot_matrix(angles, vec1=(0, 0, 1), vec2=(0, 1, 0), vec3=(1, 0, 0), inverse=False):
    """R of Func_rot_3Dvol_FourierShear (inverse=False) / Func_inv_rot_3Dvol_FourierShear (inverse=True).

    Returns (R as 3x3 row-major numpy array, rotate_flag)."""
    R = np.empty(9, np.float64)
    phi, theta, psi = [float(a) for a in angles]
    flag = lib().oracle_rot_matrix(_p(_f64(vec1)), _p(_f64(vec2)), _p(_f64(vec3)), phi, theta, psi, int(inverse), _p(R))
    return R.reshape(3, 3), bool(flag)


def shear_plan(R, rot_step_init=11.0):
    ns = np.zeros(1, np.int32)
    dims = np.zeros(4, np.int32)
    ab = np.zeros(8, np.float64)
    lib().oracle_shear_plan(_p(_f64(R).ravel()), rot_step_init, _p(ns), _p(dims), _p(ab))
    return int(ns[0]), dims.tolist(), ab.reshape(4, 2)


def rotate(vol, R, rot_step_init=11.0, single=False):
    """Func_rotate_3Dvol_ROTmat_FourierShear(vol, R, rot_step_init) on a [z, y, x] volume."""
    if single:
        v = _f32(vol); out = np.empty_like(v); fn = lib().oracle_rotate_f
    else:
        v = _f64(vol); out = np.empty_like(v); fn = lib().oracle_rotate_d
    nz, ny, nx = v.shape
    fn(_p(v), _p(out), nx, ny, nz, _p(_f64(R).ravel()), rot_step_init)
    return out


def rot_3dvol(vol, angles, inverse=False, vec=((0, 0, 1), (0, 1, 0), (1, 0, 0)), single=False):
    """Func_rot_3Dvol_FourierShear / Func_inv_rot_3Dvol_FourierShear (identity when all |angles| <= 0.5)."""
    R, flag = rot_matrix(angles, *vec, inverse=inverse)
    if not flag:
        return np.array(vol, dtype=np.float32 if single else np.float64, copy=True)
    return rotate(vol, R, 11.0, single=single)
