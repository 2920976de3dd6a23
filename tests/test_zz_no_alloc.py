"""The allocation-free gradient entry point (mset_grad_noalloc) against mset_grad, through the C ABI.

Both run the same kernels on the same resident state with the same persistent grid (the slot cap is
set on the context), so the outputs must agree bitwise.  Two calls in a row share the static scratch
array: the second, on a smaller problem, must not depend on what the first left there.
"""
import numpy as np
import pytest

from tests.cases import small_case

pytestmark = pytest.mark.gpu


class _DeviceBuffers:
    """Device buffers for the outputs, from the HIP runtime the library itself runs on (libamdhip64.so.7,
    already loaded by libmset_amd.so), each filled with NaN first so that an element the call does not
    write shows up."""

    def __init__(self, shapes):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL('libamdhip64.so.7')
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.bufs = {}
        for name, (shape, dtype) in shapes.items():
            host = np.full(shape, np.nan, dtype)
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), host.nbytes) == 0
            assert self.hip.hipMemcpy(p, host.ctypes.data, host.nbytes, 1) == 0      # host to device
            self.bufs[name] = (p, shape, dtype, host.nbytes)

    def ptr(self, name):
        return self.bufs[name][0].value

    def read(self, name):
        p, shape, dtype, nb = self.bufs[name]
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, nb, 2) == 0                    # device to host
        return out

    def free(self):
        for p, _, _, _ in self.bufs.values():
            self.hip.hipFree(p)
        self.bufs = {}


def _device_outputs(L, n, count):
    return _DeviceBuffers(dict(dV=((L, n, n), np.float32), dp=((n, n), np.complex64), dpos=((count, 2), np.float32),
                               err=((count,), np.float64)))


def _run_noalloc(ctx, first, count):
    o = _device_outputs(ctx.L, ctx.shape[1], count)
    try:
        ctx.grad_noalloc(first, count, o.ptr('dV'), o.ptr('dp'), o.ptr('dpos'), o.ptr('err'))
        ctx.synchronize()
        return {k: o.read(k) for k in ('dV', 'dp', 'dpos', 'err')}
    finally:
        o.free()


def _check_case(ctx, c, first, count, slots):
    print('case: noalloc n=%d nz=%d bin=%d positions %d..%d slots %d' %
          (c['n'], c['nz'], c['bin'], first, first + count, slots), flush=True)
    ctx.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'], slice_binning=c['bin'],
               n_scan=(c['npos'], 1))
    ctx.set_slot_limit(slots)
    dV, dp, dpos, e = ctx.grad(first, count)
    got = _run_noalloc(ctx, first, count)
    assert np.array_equal(got['dV'], dV)
    assert np.array_equal(got['dp'], dp)
    assert np.array_equal(got['dpos'], dpos)
    s = 0.0
    for v in got['err']:
        s += float(v)
    assert s / (c['npos'] * 1.0) / (c['n'] * c['n']) == e
    ctx.set_slot_limit(0)


def test_noalloc_gradient_matches_mset_grad_on_two_calls():
    from mset_amd import lib
    with lib.Context(0) as ctx:
        # more positions than slots, so slots accumulate into their slabs; a range that does not start at 0
        _check_case(ctx, small_case(n=128, nz=8, binning=2, npos=12, seed=41, spread=8.0), 2, 9, 4)
        # then a smaller problem on the same scratch
        _check_case(ctx, small_case(n=64, nz=6, binning=2, npos=5, seed=42, spread=6.0), 0, 5, 2)


def test_noalloc_rejects_a_problem_beyond_the_scratch_maximum():
    from mset_amd import lib
    c = small_case(n=256, nz=64, binning=1, npos=1, seed=43, spread=4.0)
    with lib.Context(0) as ctx:
        ctx.upload(c['V'], c['probe'], c['prop'], c['xl'], c['yl'], c['M'], c['sig'], c['dx'], slice_binning=1,
                   n_scan=(1, 1))
        o = _device_outputs(ctx.L, 256, 1)
        try:
            with pytest.raises(lib.MsetError) as ei:
                ctx.grad_noalloc(0, 1, o.ptr('dV'), o.ptr('dp'), o.ptr('dpos'), o.ptr('err'))
            assert ei.value.code == lib.MSET_EINVAL and 'maximum' in str(ei.value)
            # the rejected call issued nothing: the outputs are untouched
            ctx.synchronize()
            assert np.isnan(o.read('dV')).all()
        finally:
            o.free()
