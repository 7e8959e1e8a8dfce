"""C ABI checks that need no GPU: the library loads, exports every entry
point include/omdg.h declares, and refuses to run without a device."""
import ctypes
import os
import re

import pytest

import objectmd_amd as om
from tests.common import GOLDEN  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "omdg.h")).read()
    return sorted(set(re.findall(r"\b(omdg_[a-z_]+)\s*\(", src)))


def test_header_and_python_lists_agree():
    assert _declared() == sorted(om.ABI_SYMBOLS)


def test_library_exports_every_symbol():
    lib = om.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_no_device_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = om.load_library()
    h = ctypes.c_void_p()
    rc = lib.omdg_create(ctypes.byref(h), 0)
    assert rc == om.ENODEV
    assert b"no HIP device" in lib.omdg_last_error()
    with pytest.raises(om.OmdgError):
        om.Engine()


def test_null_arguments_rejected():
    lib = om.load_library()
    assert lib.omdg_create(None, 0) == om.EINVAL
    assert lib.omdg_step(None, 1, ctypes.c_double(1e-3), None, None) == om.EINVAL
    assert lib.omdg_destroy(None) == om.OK
