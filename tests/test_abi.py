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
    # 7 int32 + 4 float + float[3] + int32 = 15 * 4 bytes, no padding
    assert C.sizeof(Params) == 15 * 4
    assert Params.step_size.offset == 11 * 4 and Params.grad_mode.offset == 14 * 4


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
