"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/taurus_gpu.h (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared():
    hdr = open(os.path.join(ROOT, "include", "taurus_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tg_\w+)\s*\(", hdr, re.M)))


def test_header_matches_binding_list():
    from taurus_vap_amd import gpu
    assert _declared() == sorted(gpu.EXPORTS)


def test_library_exports_every_symbol():
    from taurus_vap_amd import gpu
    if not os.path.exists(gpu.LIBPATH):
        pytest.skip("libtaurus_gpu.so not built")
    lib = ctypes.CDLL(gpu.LIBPATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_error_convention_without_device():
    """Argument errors are reported before any device work."""
    from taurus_vap_amd import gpu
    if not os.path.exists(gpu.LIBPATH):
        pytest.skip("libtaurus_gpu.so not built")
    lib = gpu.load_library()
    h = ctypes.c_void_p()
    assert lib.tg_init(ctypes.byref(h), 3, None, None, None, None, 0.0, 0.0, None, 0, None, None,
                       4, None, 0) == -1
    assert b"invalid" in lib.tg_last_error()
    assert lib.tg_zskpfa(-1, 1, None, 1, None, None) == -1
