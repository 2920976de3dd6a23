"""CPU-side checks of the C ABI: the library loads, exports every symbol include/mset.h declares,
the ctypes structure matches the C struct, and argument errors are reported without a device."""
import ctypes as C
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'mset.h')


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mset_[a-z_]+)\s*\(', text)))


def test_header_declares_the_mex_modes():
    fns = header_functions()
    for name in ('mset_upload', 'mset_run', 'mset_error', 'mset_destroy', 'mset_create'):
        assert name in fns


def test_library_exports_every_declared_symbol():
    from mset_amd import lib
    L = lib.load()
    missing = [f for f in header_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert sorted(lib.EXPORTS) == header_functions()


def test_params_struct_layout():
    from mset_amd.lib import Params
    # 7 int32 + 4 float + float[3] + 2 int32 = 16 * 4 bytes, no padding
    assert C.sizeof(Params) == 16 * 4
    assert Params.step_size.offset == 11 * 4 and Params.grad_mode.offset == 14 * 4
    assert Params.diffraction_transpose.offset == 15 * 4


def test_null_context_is_an_error_not_a_crash():
    from mset_amd import lib
    L = lib.load()
    assert L.mset_synchronize(None) == lib.MSET_EINVAL
    assert b'null context' in L.mset_last_error()
    assert L.mset_create(None, 0) == lib.MSET_EINVAL


def test_build_script_flags():
    from mset_amd import build
    assert '--offload-arch=gfx950' in build.FLAGS
    assert '-packed-fp32-ops' in build.FLAGS


class _FakeHostAlloc:
    """Stands in for mset_host_alloc / mset_host_free (which need a device) to test PinnedArray's lifetime."""

    def __init__(self):
        self.bufs, self.freed = {}, []

    def mset_host_alloc(self, size, out):
        b = C.create_string_buffer(size.value)
        C.cast(out, C.POINTER(C.c_void_p))[0] = C.addressof(b)
        self.bufs[C.addressof(b)] = b
        return 0

    def mset_host_free(self, ptr):
        self.freed.append(ptr.value)
        return 0


def test_pinned_array_is_released_with_its_last_view(monkeypatch):
    import gc

    import numpy as np

    from mset_amd import lib
    fake = _FakeHostAlloc()
    monkeypatch.setattr(lib, 'load', lambda: fake)
    gc.disable()                        # the release must not depend on the cycle collector
    try:
        pa = lib.PinnedArray((4, 8), np.float32)
        addr = pa.array.ctypes.data
        view = pa.array[1:]
        view[...] = 3.0
        del pa
        assert fake.freed == []         # a view still refers to the memory
        assert float(view.sum()) == 3.0 * 24
        del view
        assert fake.freed == [addr]
        pb = lib.PinnedArray((2,), np.complex64)
        b = pb.array.ctypes.data
        pb.free()                       # explicit release, exactly once
        assert not pb.alive and fake.freed == [addr, b]
        del pb
        assert fake.freed == [addr, b]
    finally:
        gc.enable()
