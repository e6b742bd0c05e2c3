"""Device ceiling entry points (include/gss_ceilings.h): they run, report a
plausible device, and reject bad arguments before touching the GPU."""
import ctypes

import pytest

from gpusharesat_amd import _lib
from gpusharesat_amd._lib import GssError

pytestmark = pytest.mark.gpu


def test_device_query_and_small_ceilings():
    buf = ctypes.create_string_buffer(256)
    _lib.call("gss_device_query", 0, buf)
    arch = buf.raw[64:128].split(b"\0")[0].decode()
    cus = int.from_bytes(buf.raw[128:132], "little", signed=True)
    print("arch", arch, "cus", cus, flush=True)
    assert arch.startswith("gfx950") and cus > 0
    gbps, ms = ctypes.c_double(), ctypes.c_double()
    _lib.call("gss_ceiling_copy", 0, 64 << 20, 3, ctypes.byref(gbps), ctypes.byref(ms))
    assert gbps.value > 0 and ms.value > 0
    d2h, h2d = ctypes.c_double(), ctypes.c_double()
    _lib.call("gss_ceiling_pcie", 0, 16 << 20, 2, ctypes.byref(d2h), ctypes.byref(h2d))
    assert d2h.value > 0 and h2d.value > 0
    for unroll in (1, 2, 4, 8):
        _lib.call("gss_ceiling_gather", 0, 2 << 20, 1 << 20, unroll, 2, ctypes.byref(gbps),
                  ctypes.byref(ms))
        assert gbps.value > 0 and ms.value > 0


def test_ceilings_reject_bad_arguments():
    gbps, ms = ctypes.c_double(), ctypes.c_double()
    with pytest.raises(GssError):  # not a multiple of 16
        _lib.call("gss_ceiling_copy", 0, 1000, 1, ctypes.byref(gbps), ctypes.byref(ms))
    with pytest.raises(GssError):  # unroll 3
        _lib.call("gss_ceiling_gather", 0, 1 << 20, 1000, 3, 1, ctypes.byref(gbps),
                  ctypes.byref(ms))
    with pytest.raises(GssError):  # table not a multiple of 128
        _lib.call("gss_ceiling_gather", 0, 1000, 1000, 4, 1, ctypes.byref(gbps),
                  ctypes.byref(ms))
    with pytest.raises(GssError):  # no such device
        _lib.call("gss_device_query", 99, ctypes.create_string_buffer(256))
