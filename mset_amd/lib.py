"""ctypes binding of the C ABI in include/mset.h (libmset_amd.so).

There is no CPU fallback: if the library is missing or no HIP device is present, the calls raise
MsetError.  Array layout is the MEX's (MATLAB column-major = C order with dims reversed): a volume
RVol(x, y, z) is a numpy array [nz, ny, nx] float32; complex planes are complex64 [ny, nx]; the
measured data is float32 [num_scan_pos, ny, nx] in MATLAB's linear cell order.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmset_amd.so')

MSET_OK, MSET_EINVAL, MSET_ENOMEM, MSET_EDEVICE, MSET_EUNSUPPORTED = 0, 1, 2, 3, 4
METHOD_MSET, METHOD_SSET = 1, 2
GRAD_REF, GRAD_EXACT = 0, 1

# every symbol declared in include/mset.h
EXPORTS = ['mset_create', 'mset_destroy', 'mset_last_error', 'mset_stream', 'mset_upload', 'mset_run', 'mset_error',
           'mset_run_resident', 'mset_error_resident', 'mset_download', 'mset_set_state', 'mset_grad', 'mset_waves',
           'mset_run_batched', 'mset_rotate', 'mset_rotate_resident', 'mset_profile_enable', 'mset_profile_read',
           'mset_synchronize', 'mset_simulate', 'mset_recon_create', 'mset_recon_destroy', 'mset_recon_set_tilt',
           'mset_recon_set_steps', 'mset_recon_iterate', 'mset_recon_error_pass', 'mset_recon_get',
           'mset_recon_snapshot', 'mset_recon_restore', 'mset_recon_profile_enable', 'mset_recon_profile_read',
           'mset_recon_stream', 'mset_tv', 'mset_recon_tv', 'mset_device_memory', 'mset_batch_slots',
           'mset_set_exit_waves', 'mset_exit_waves', 'mset_recon_set_exit_waves', 'mset_set_slot_limit',
           'mset_recon_set_slot_limit', 'mset_host_alloc', 'mset_host_free', 'mset_grad_noalloc', 'mset_set_engine',
           'mset_recon_set_engine']
PATH_PLAIN, PATH_BATCHED = 0, 1
EXIT_WAVES_AUTO, EXIT_WAVES_STORE, EXIT_WAVES_RECOMPUTE, EXIT_WAVES_CHECKPOINT = 0, 1, 2, 3
ENGINE_AUTO, ENGINE_GENERAL, ENGINE_ALT_SHAPE = 0, 1, 2


class MsetError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('mset error %d: %s' % (code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [('nx', C.c_int32), ('ny', C.c_int32), ('nz', C.c_int32),
                ('n_scan_x', C.c_int32), ('n_scan_y', C.c_int32), ('num_scan_pos', C.c_int32),
                ('slice_binning', C.c_int32),
                ('alpha', C.c_float), ('sigma', C.c_float), ('E0', C.c_float), ('potential_pixelsize', C.c_float),
                ('step_size', C.c_float * 3), ('grad_mode', C.c_int32), ('diffraction_transpose', C.c_int32)]


_LIB = None


def load():
    """Load libmset_amd.so (raises MsetError if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise MsetError(MSET_EINVAL, 'HIP library %s not built: run `python -m mset_amd.build`' % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P, I, D, F = C.c_void_p, C.c_int, C.c_double, C.c_float
    PP = C.POINTER(C.c_void_p)
    sig = {
        'mset_create': ([PP, I], I), 'mset_destroy': ([P], I), 'mset_last_error': ([], C.c_char_p),
        'mset_stream': ([P], P),
        'mset_upload': ([P, C.POINTER(Params), P, P, P, P, P, P], I),
        'mset_run': ([P, I, P, P, P, P, P], I), 'mset_error': ([P, I, P], I),
        'mset_run_resident': ([P, I, P], I), 'mset_error_resident': ([P, I, P], I),
        'mset_download': ([P, P, P, P, P], I), 'mset_set_state': ([P, P, P, P, P], I),
        'mset_grad': ([P, I, I, P, P, P, P], I), 'mset_run_batched': ([P, I, I, P], I),
        'mset_waves': ([P, I, P], I),
        'mset_rotate': ([P, P, P, I, I, I, P, D], I),
        'mset_rotate_resident': ([P, P, P, P, D, D, D, I], I),
        'mset_profile_enable': ([P, I], I), 'mset_profile_read': ([P, C.c_char_p, P, P], I),
        'mset_synchronize': ([P], I), 'mset_simulate': ([P, P], I),
        'mset_recon_create': ([PP, I, C.POINTER(Params), I, P, P, P], I), 'mset_recon_destroy': ([P], I),
        'mset_recon_set_tilt': ([P, I, P, P, I, I, I, P, P, P], I), 'mset_recon_set_steps': ([P, P], I),
        'mset_recon_iterate': ([P, I, I, I, I, P], I), 'mset_recon_error_pass': ([P, P], I),
        'mset_recon_get': ([P, P, P, P, P], I), 'mset_recon_snapshot': ([P], I), 'mset_recon_restore': ([P], I),
        'mset_recon_profile_enable': ([P, I], I), 'mset_recon_profile_read': ([P, C.c_char_p, P, P], I),
        'mset_recon_stream': ([P], P),
        'mset_tv': ([P, P, P, I, I, I, F, I, P], I), 'mset_recon_tv': ([P, F, I, P], I),
        'mset_device_memory': ([I, P, P], I), 'mset_batch_slots': ([I, I, P], I),
        'mset_set_exit_waves': ([P, I], I), 'mset_exit_waves': ([P, P, P], I),
        'mset_recon_set_exit_waves': ([P, I], I), 'mset_set_slot_limit': ([P, I], I),
        'mset_recon_set_slot_limit': ([P, I], I), 'mset_host_alloc': ([C.c_size_t, PP], I),
        'mset_host_free': ([P], I),
        'mset_grad_noalloc': ([P, I, I, P, P, P, P], I),
        'mset_set_engine': ([P, I], I), 'mset_recon_set_engine': ([P, I], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def device_memory(device=0):
    """(free, total) device memory in bytes."""
    L = load()
    f, t = C.c_size_t(0), C.c_size_t(0)
    _chk(L.mset_device_memory(int(device), C.addressof(f), C.addressof(t)))
    return int(f.value), int(t.value)


def batch_slots(n, device=0):
    """Resident workgroups of the batched position kernel for plane size n."""
    L = load()
    s = C.c_int(0)
    _chk(L.mset_batch_slots(int(device), int(n), C.addressof(s)))
    return int(s.value)


class PinnedArray:
    """A numpy array over page-locked host memory (mset_host_alloc), released as soon as neither this object
    nor any view of `.array` is reachable, or at once by free() (after which no view may be used).  Upload
    from `.array` like from any array: the copy to the device then runs at the full link rate."""

    def __init__(self, shape, dtype=np.float32):
        L = load()
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        _chk(L.mset_host_alloc(C.c_size_t(max(n, 1)), C.byref(p)))
        buf = (C.c_char * max(n, 1)).from_address(p.value)
        # the release hangs off the ctypes buffer that every view of the array keeps alive through its
        # base chain; nothing on that chain refers back to this object, so reference counting frees the
        # memory when the last view goes (no wait for the cycle collector)
        self._fin = weakref.finalize(buf, L.mset_host_free, C.c_void_p(p.value))
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    @property
    def alive(self):
        """True until the page-locked memory has been released."""
        return self._fin.alive

    def free(self):
        self.array = None
        self._fin()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(code):
    if code != MSET_OK:
        raise MsetError(code, load().mset_last_error().decode(errors='replace'))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


class Context:
    """One device context: the resident state of one tilt (volume, probe, positions, patterns)."""

    def __init__(self, device=0):
        self._L = load()
        h = C.c_void_p()
        _chk(self._L.mset_create(C.byref(h), int(device)))
        self._h = h
        self.params = None

    def close(self):
        if getattr(self, '_h', None):
            self._L.mset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return self._L.mset_stream(self._h)

    # ---------------------------------------------------------------- MEX-compatible
    def upload(self, RVol, probe, prop2D, xlist, ylist, measured, sigma, dx, slice_binning=1,
               step_size=(0.0, 0.0, 0.0), n_scan=None, grad_mode=GRAD_REF, alpha=0.0, E0=0.0,
               diffraction_transpose=False):
        RVol = _f32(RVol)
        nz, ny, nx = RVol.shape
        xl, yl = _f32(xlist).ravel(), _f32(ylist).ravel()
        M = _f32(measured)
        npos = xl.shape[0]
        if M.shape != (npos, ny, nx):
            raise MsetError(MSET_EINVAL, 'measured must be (num_scan_pos, ny, nx) = %s, got %s' % ((npos, ny, nx), M.shape))
        if n_scan is None:
            n_scan = (npos, 1)
        p = Params(nx=nx, ny=ny, nz=nz, n_scan_x=int(n_scan[0]), n_scan_y=int(n_scan[1]), num_scan_pos=npos,
                   slice_binning=int(slice_binning), alpha=float(alpha), sigma=float(sigma), E0=float(E0),
                   potential_pixelsize=float(dx), step_size=(C.c_float * 3)(*[float(s) for s in step_size]),
                   grad_mode=int(grad_mode), diffraction_transpose=int(bool(diffraction_transpose)))
        pr, pp = _c64(probe), _c64(prop2D)
        _chk(self._L.mset_upload(self._h, C.byref(p), _ptr(RVol), _ptr(pr), _ptr(pp), _ptr(xl), _ptr(yl), _ptr(M)))
        self.params = p
        self.shape = (nz, ny, nx)
        self.npos = npos
        self.L = -(-nz // min(int(slice_binning), nz))

    def _outs(self):
        nz, ny, nx = self.shape
        return (np.empty((nz, ny, nx), np.float32), np.empty((ny, nx), np.complex64),
                np.empty(self.npos, np.float32), np.empty(self.npos, np.float32))

    def run(self, method=METHOD_MSET):
        """'run_mset': sequential updates over all positions -> (RVol, probe, xlist, ylist, error)."""
        V, pr, xl, yl = self._outs()
        e = np.zeros(1, np.float32)
        _chk(self._L.mset_run(self._h, int(method), _ptr(V), _ptr(pr), _ptr(xl), _ptr(yl), _ptr(e)))
        return V, pr, xl, yl, float(e[0])

    def error(self, method=METHOD_MSET):
        """'error_mset': forward + loss at the resident state."""
        e = np.zeros(1, np.float32)
        _chk(self._L.mset_error(self._h, int(method), _ptr(e)))
        return float(e[0])

    # ---------------------------------------------------------------- device-resident
    def run_resident(self, method=METHOD_MSET):
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_run_resident(self._h, int(method), _ptr(e)))
        return float(e[0])

    def error_resident(self, method=METHOD_MSET):
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_error_resident(self._h, int(method), _ptr(e)))
        return float(e[0])

    def download(self):
        V, pr, xl, yl = self._outs()
        _chk(self._L.mset_download(self._h, _ptr(V), _ptr(pr), _ptr(xl), _ptr(yl)))
        return V, pr, xl, yl

    def set_state(self, RVol=None, probe=None, xlist=None, ylist=None):
        a = [None if RVol is None else _f32(RVol), None if probe is None else _c64(probe),
             None if xlist is None else _f32(xlist), None if ylist is None else _f32(ylist)]
        _chk(self._L.mset_set_state(self._h, *[_ptr(x) for x in a]))

    def grad(self, first=0, count=None):
        """Forward + gradient sums at the fixed resident state -> (dV[L], dprobe, dpos[count, 2], error)."""
        if count is None:
            count = self.npos - first
        nz, ny, nx = self.shape
        dV = np.empty((self.L, ny, nx), np.float32)
        dp = np.empty((ny, nx), np.complex64)
        dpos = np.empty((count, 2), np.float32)
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_grad(self._h, int(first), int(count), _ptr(dV), _ptr(dp), _ptr(dpos), _ptr(e)))
        return dV, dp, dpos, float(e[0])

    def grad_noalloc(self, first, count, dV_ptr, dprobe_ptr, dpos_ptr, err_ptr):
        """mset_grad_noalloc: issues the gradient into caller-owned DEVICE buffers (raw addresses) using the
        library's static scratch; call synchronize() before reading them."""
        _chk(self._L.mset_grad_noalloc(self._h, int(first), int(count), C.c_void_p(dV_ptr), C.c_void_p(dprobe_ptr),
                                       C.c_void_p(dpos_ptr), C.c_void_p(err_ptr)))

    def waves(self, k):
        """Incident wave and the wave after each plane for position k -> complex64 [L + 1, ny, nx]."""
        nz, ny, nx = self.shape
        w = np.empty((self.L + 1, ny, nx), np.complex64)
        _chk(self._L.mset_waves(self._h, int(k), _ptr(w)))
        return w

    def run_batched(self, batch, method=METHOD_MSET):
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_run_batched(self._h, int(method), int(batch), _ptr(e)))
        return float(e[0])

    def rotate_resident(self, angles, inverse, vec1=(0, 0, 1), vec2=(0, 1, 0), vec3=(1, 0, 0)):
        v = [np.ascontiguousarray(x, np.float64) for x in (vec1, vec2, vec3)]
        phi, theta, psi = [float(a) for a in angles]
        _chk(self._L.mset_rotate_resident(self._h, _ptr(v[0]), _ptr(v[1]), _ptr(v[2]), phi, theta, psi, int(inverse)))

    def rotate(self, vol, R, rot_step_init=11.0):
        """Func_rotate_3Dvol_ROTmat_FourierShear(vol, R, rot_step_init) of a host volume [nz, ny, nx]."""
        v = _f32(vol)
        out = np.empty_like(v)
        nz, ny, nx = v.shape
        Rm = np.ascontiguousarray(R, np.float64).ravel()
        _chk(self._L.mset_rotate(self._h, _ptr(v), _ptr(out), nx, ny, nz, _ptr(Rm), float(rot_step_init)))
        return out

    def tv(self, vol, lam, n_iter=300):
        """TV_method(vol, lambda, n_iter) of a host volume [nz, ny, nx] -> (output, error)."""
        v = _f32(vol)
        out = np.empty_like(v)
        nz, ny, nx = v.shape
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_tv(self._h, _ptr(v), _ptr(out), nx, ny, nz, float(lam), int(n_iter), _ptr(e)))
        return out, float(e[0])

    def set_exit_waves(self, mode):
        """EXIT_WAVES_AUTO / _STORE / _RECOMPUTE / _CHECKPOINT for the batched calls (mset_set_exit_waves)."""
        _chk(self._L.mset_set_exit_waves(self._h, int(mode)))

    def set_slot_limit(self, slots):
        """cap on the batched kernels' persistent grid (0 = all resident workgroups)"""
        _chk(self._L.mset_set_slot_limit(self._h, int(slots)))

    def set_engine(self, engine):
        """ENGINE_AUTO (separable routine at 256 when the propagator factors) or ENGINE_GENERAL"""
        _chk(self._L.mset_set_engine(self._h, int(engine)))

    def exit_waves(self):
        """(variant, per-slot scratch bytes) of the last batched call."""
        m = C.c_int(0)
        b = C.c_size_t(0)
        _chk(self._L.mset_exit_waves(self._h, C.byref(m), C.byref(b)))
        return m.value, int(b.value)

    # ---------------------------------------------------------------- instrumentation
    def profile(self, enable=True):
        _chk(self._L.mset_profile_enable(self._h, int(bool(enable))))

    def profile_read(self, name):
        n = C.c_int(0)
        ms = C.c_double(0.0)
        _chk(self._L.mset_profile_read(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def synchronize(self):
        _chk(self._L.mset_synchronize(self._h))


def _simulate(self):
    """|fftn(exit wave)|^2 per position at the resident state, DC-centred -> float32 [npos, ny, nx]."""
    nz, ny, nx = self.shape
    out = np.empty((self.npos, ny, nx), np.float32)
    _chk(self._L.mset_simulate(self._h, _ptr(out)))
    return out


Context.simulate = _simulate


class Recon:
    """A device-resident tilt series: Run_MSET_recon's update and error passes on the GPU."""

    def __init__(self, rec, prop2D, n_tilts, sigma, dx, slice_binning=1, step_size=(0.0, 0.0, 0.0),
                 grad_mode=GRAD_REF, vecs=((0, 0, 1), (0, 1, 0), (1, 0, 0)), device=0, diffraction_transpose=False):
        self._L = load()
        rec = _f32(rec)
        nz, ny, nx = rec.shape
        p = Params(nx=nx, ny=ny, nz=nz, n_scan_x=1, n_scan_y=1, num_scan_pos=1, slice_binning=int(slice_binning),
                   sigma=float(sigma), potential_pixelsize=float(dx),
                   step_size=(C.c_float * 3)(*[float(s) for s in step_size]), grad_mode=int(grad_mode),
                   diffraction_transpose=int(bool(diffraction_transpose)))
        v = np.ascontiguousarray(vecs, np.float64).ravel()
        h = C.c_void_p()
        _chk(self._L.mset_recon_create(C.byref(h), int(device), C.byref(p), int(n_tilts), _ptr(v), _ptr(rec),
                                       _ptr(_c64(prop2D))))
        self._h = h
        self.shape = (nz, ny, nx)
        self.n_tilts = n_tilts
        self.npos = [0] * n_tilts

    def close(self):
        if getattr(self, '_h', None):
            self._L.mset_recon_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tilt(self, j, angles, probe, xlist, ylist, measured, n_scan=None):
        xl, yl = _f32(xlist).ravel(), _f32(ylist).ravel()
        M = _f32(measured)
        npos = xl.shape[0]
        if n_scan is None:
            n_scan = (npos, 1)
        a = np.ascontiguousarray(angles, np.float64)
        _chk(self._L.mset_recon_set_tilt(self._h, int(j), _ptr(a), _ptr(_c64(probe)), int(n_scan[0]), int(n_scan[1]),
                                         npos, _ptr(xl), _ptr(yl), _ptr(M)))
        self.npos[j] = npos

    def set_steps(self, steps):
        s = np.ascontiguousarray(steps, np.float32)
        _chk(self._L.mset_recon_set_steps(self._h, _ptr(s)))

    def iterate(self, path=PATH_BATCHED, batch=256, use_graph=False, positivity=True, want_error=True):
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_recon_iterate(self._h, int(path), int(batch), int(bool(use_graph)), int(bool(positivity)),
                                        _ptr(e) if want_error else None))
        return float(e[0])

    def error_pass(self):
        e = np.zeros(self.n_tilts, np.float64)
        _chk(self._L.mset_recon_error_pass(self._h, _ptr(e)))
        return e

    def get(self):
        nz, ny, nx = self.shape
        rec = np.empty((nz, ny, nx), np.float32)
        probes = np.empty((self.n_tilts, ny, nx), np.complex64)
        tot = sum(self.npos)
        xl = np.empty(tot, np.float32)
        yl = np.empty(tot, np.float32)
        _chk(self._L.mset_recon_get(self._h, _ptr(rec), _ptr(probes), _ptr(xl), _ptr(yl)))
        split = np.cumsum(self.npos)[:-1]
        return rec, probes, np.split(xl, split), np.split(yl, split)

    def tv(self, lam, n_iter=300):
        """rec = max(rec) * TV_method(rec / max(rec), lambda, n_iter); returns the TV error."""
        e = np.zeros(1, np.float64)
        _chk(self._L.mset_recon_tv(self._h, float(lam), int(n_iter), _ptr(e)))
        return float(e[0])

    def set_exit_waves(self, mode):
        """exit-wave variant of the batched passes (EXIT_WAVES_*)"""
        _chk(self._L.mset_recon_set_exit_waves(self._h, int(mode)))

    def set_slot_limit(self, slots):
        """cap on the batched passes' persistent grid (0 = all resident workgroups)"""
        _chk(self._L.mset_recon_set_slot_limit(self._h, int(slots)))

    def set_engine(self, engine):
        """ENGINE_AUTO or ENGINE_GENERAL for the series' passes"""
        _chk(self._L.mset_recon_set_engine(self._h, int(engine)))

    def snapshot(self):
        _chk(self._L.mset_recon_snapshot(self._h))

    def restore(self):
        _chk(self._L.mset_recon_restore(self._h))

    def profile(self, enable=True):
        _chk(self._L.mset_recon_profile_enable(self._h, int(bool(enable))))

    def profile_read(self, name):
        n = C.c_int(0)
        ms = C.c_double(0.0)
        _chk(self._L.mset_recon_profile_read(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value
