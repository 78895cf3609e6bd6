"""Content analysis (src/content.rs:119-174) on the device against the
oracle: the histogram is integer work and the f32 statistics are formed in
the reference's order, so every figure is bit-identical."""
import numpy as np
import pytest

import jp2lam_amd as j
from jp2lam_amd import synth

pytestmark = pytest.mark.gpu


def _gray(px):
    return j.Image.from_gray_bytes(px.shape[1], px.shape[0], px)


def _same(st, want):
    return (np.float32(st.bright_share) == want[0] and np.float32(st.entropy) == want[1]
            and np.float32(st.colorfulness) == want[2])


def test_reference_cases(gpu):
    # content.rs:190-235
    px = np.full(1600, 255, np.uint8)
    px[:200] = 0
    img = _gray(px.reshape(40, 40))
    st = j.analyze_content(img)
    assert st.bright_share > 0.45 and st.classify() is j.ContentClass.GRAPHICS
    assert j.auto_quality(img) == j.GRAPHICS_QUALITY == 55
    ramp = np.tile(np.arange(256, dtype=np.uint8), (64, 1))
    st = j.analyze_content(_gray(ramp))
    assert st.entropy > 5.0 and st.classify() is j.ContentClass.PHOTO
    assert j.auto_class_and_quality(_gray(ramp)) == (j.ContentClass.PHOTO, 78)
    assert j.analyze_content(_gray(np.full((8, 8), 128, np.uint8))).colorfulness == 0.0


@pytest.mark.parametrize("kind,w,h", [("gray", 1, 1), ("gray", 300, 217), ("doc", 2550, 3300), ("rgb", 129, 77),
                                      ("rgb", 4000, 3000)])
def test_matches_oracle(gpu, ora, kind, w, h):
    if kind == "rgb":
        px = synth.photo_page(w, h, seed=w)
        planes = np.moveaxis(px, -1, 0).astype(np.int32)
        img = j.Image.from_rgb_bytes(w, h, px)
        cs = 1
    else:
        px = synth.document_page(w, h, seed=2) if kind == "doc" else synth.gradient_page(w, h, seed=w)
        planes = px[None].astype(np.int32)
        img = _gray(px)
        cs = 0
    want, cls = ora.analyze_content(planes, cs)
    st = j.analyze_content(img)
    assert _same(st, want), (st, want)
    assert st.classify() is (j.ContentClass.GRAPHICS if cls else j.ContentClass.PHOTO)
    assert _same(j.analyze_content_u8(px), want)


@pytest.mark.parametrize("cs", [j.ColorSpace.RGB, j.ColorSpace.YUV, j.ColorSpace.YCBCR])
def test_i32_views_clamp_and_rgb_family(gpu, ora, cs):
    # samples outside 0..255 are clamped (content.rs:110-117, :126-147);
    # every non-gray colour space is read as R, G, B
    rng = np.random.default_rng(int(cs))
    planes = rng.integers(-40, 300, size=(3, 45, 61)).astype(np.int32)
    img = j.Image(61, 45, [j.Component(planes[c].reshape(-1), 61, 45) for c in range(3)], cs)
    want, _ = ora.analyze_content(planes, int(cs))
    assert _same(j.analyze_content(img), want)
    g = j.Image(61, 45, [j.Component(planes[0].reshape(-1), 61, 45)], j.ColorSpace.GRAY)
    assert _same(j.analyze_content(g), ora.analyze_content(planes[:1], 0)[0])


def test_too_few_components_give_zeros(gpu):
    # content.rs:132-148, :155-157: an RGB-family image without three planes counts nothing
    px = np.full((4, 4), 7, np.int32)
    img = j.Image(4, 4, [j.Component(px.reshape(-1), 4, 4)], j.ColorSpace.SRGB)
    st = j.analyze_content(img)
    assert (st.bright_share, st.entropy, st.colorfulness) == (0.0, 0.0, 0.0)
    assert st.classify() is j.ContentClass.GRAPHICS
    assert j.auto_quality(img) == 55


def test_device_resident_pixels(gpu, lib, ora):
    import ctypes as C
    px = synth.photo_page(200, 150, seed=4)
    d = lib.jp2lam_device_alloc(px.size)
    assert d
    try:
        assert lib.jp2lam_device_copy(d, px.ctypes.data, px.size, 0) == 0
        st = j._lib.ContentStatsC()
        assert lib.jp2lam_analyze_content_u8(d, 200, 150, 3, 1, C.byref(st)) == 0
        want, cls = ora.analyze_content(np.moveaxis(px, -1, 0).astype(np.int32), 1)
        assert (np.float32(st.bright_share), np.float32(st.entropy), np.float32(st.colorfulness)) == tuple(want)
        assert st.content_class == cls and st.quality == (55 if cls else 78)
    finally:
        lib.jp2lam_device_free(d)
