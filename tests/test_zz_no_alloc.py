"""The no-allocation content analysis entry point against the allocating
one, through the C ABI: two calls in a row (the second on a smaller image,
so a stale count would show), gray and RGB, plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from jp2lam_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _stats(s):
    return (s.bright_share, s.entropy, s.colorfulness, s.content_class, s.quality)


def _to_device(lib, a):
    a = np.ascontiguousarray(a)
    ptr = lib.jp2lam_device_alloc(a.nbytes)
    assert ptr
    assert lib.jp2lam_device_copy(ptr, a.ctypes.data, a.nbytes, 0) == 0
    return ptr


@pytest.mark.parametrize("channels", [1, 3])
def test_noalloc_matches_allocating_entry(gpu, channels):
    lib = gpu
    big = synth.document_page(1200, 900, seed=5) if channels == 1 else synth.photo_page(1000, 700, seed=5)
    small = synth.gradient_page(301, 77, seed=6)
    if channels == 3:
        small = np.ascontiguousarray(np.stack([small, small[::-1], 255 - small], axis=-1))
    ptrs = []
    try:
        for img in (big, small):
            h, w = img.shape[:2]
            ptr = _to_device(lib, img)
            ptrs.append(ptr)
            want, got = _lib.ContentStatsC(), _lib.ContentStatsC()
            assert lib.jp2lam_analyze_content_u8(ptr, w, h, channels, 1, C.byref(want)) == 0
            assert lib.jp2lam_analyze_content_u8_noalloc(ptr, w, h, channels, None, C.byref(got)) == 0
            assert _stats(got) == _stats(want), (w, h)
            # and again, with the first call's counts still in the scratch
            again = _lib.ContentStatsC()
            assert lib.jp2lam_analyze_content_u8_noalloc(ptr, w, h, channels, None, C.byref(again)) == 0
            assert _stats(again) == _stats(want)
    finally:
        for p in ptrs:
            lib.jp2lam_device_free(p)


def test_noalloc_limits(gpu):
    lib = gpu
    s = _lib.ContentStatsC()
    one = _to_device(lib, np.zeros(16, np.uint8))
    try:
        assert lib.jp2lam_analyze_content_u8_noalloc(one, 65536, 1, 1, None, C.byref(s)) == 1
        assert b"maximum" in lib.jp2lam_last_error()
        assert lib.jp2lam_analyze_content_u8_noalloc(one, 4, 4, 2, None, C.byref(s)) == 1
        assert lib.jp2lam_analyze_content_u8_noalloc(None, 4, 4, 1, None, C.byref(s)) == 1
        # an empty image gives zero statistics, classified Graphics, as the reference
        assert lib.jp2lam_analyze_content_u8_noalloc(one, 0, 4, 1, None, C.byref(s)) == 0
        assert _stats(s) == (0.0, 0.0, 0.0, 1, 55)
    finally:
        lib.jp2lam_device_free(one)
