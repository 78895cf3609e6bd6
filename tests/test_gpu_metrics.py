"""encode_with_psnr (src/encode/mod.rs:124-135) on the device against the
oracle's restatement of compute_quality_metrics (backend.rs:297-433).

The reconstruction is bit-exact (same quantisation, truncation and f32
inverse 9/7) and each 8x8 cell's SSIM follows the reference's order; only
the order of the image-wide f64 sums differs, hence the 1e-12 relative
tolerance on both figures."""
import numpy as np
import pytest

import jp2lam_amd as j
from jp2lam_amd import synth

pytestmark = pytest.mark.gpu

REL = 1e-12


def _image(px):
    h, w = px.shape[:2]
    return j.Image.from_gray_bytes(w, h, px) if px.ndim == 2 else j.Image.from_rgb_bytes(w, h, px)


def _close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("q", [1, 20, 50, 85, 95, 99])
@pytest.mark.parametrize("kind,w,h", [("gray", 257, 193), ("rgb", 130, 97), ("gray", 64, 64), ("rgb", 33, 17)])
def test_metrics_match_oracle(gpu, ora, q, kind, w, h):
    px = synth.gradient_page(w, h, seed=q) if kind == "gray" else synth.photo_page(w, h, seed=q)
    data, m = j.encode_with_psnr(_image(px), j.EncodeOptions(quality=q))
    assert data == j.encode(_image(px), j.EncodeOptions(quality=q))
    want = ora.encode_metrics(px, q)
    assert _close(m.psnr_db, want[0]), (m, want)
    assert _close(m.ssim, want[1]), (m, want)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (9, 3), (8, 8), (3, 40)])
def test_metrics_small_and_ragged(gpu, ora, w, h):
    # no full 8x8 block -> SSIM 1.0; partial cells still count for PSNR
    px = synth.gradient_page(w, h, seed=w * h)
    _, m = j.encode_with_psnr(_image(px), j.EncodeOptions(quality=60))
    want = ora.encode_metrics(px, 60)
    assert _close(m.psnr_db, want[0]) and _close(m.ssim, want[1]), (m, want)


def test_metrics_document_page_full_size(gpu, ora):
    px = synth.document_page(2550, 3300, seed=2)
    data, m = j.encode_with_psnr(_image(px), j.EncodeOptions(quality=85))
    want = ora.encode_metrics(px, 85)
    assert _close(m.psnr_db, want[0]) and _close(m.ssim, want[1]), (m, want)
    assert data == j.encode_u8(px, 85)


@pytest.mark.parametrize("kind", ["gray", "rgb"])
def test_metrics_lossless(gpu, kind):
    px = synth.gradient_page(50, 40, seed=1) if kind == "gray" else synth.photo_page(50, 40, seed=1)
    data, m = j.encode_with_psnr(_image(px), j.EncodeOptions(quality=100))
    assert m.psnr_db == float("inf") and m.ssim == 1.0
    assert data == j.encode(_image(px), j.EncodeOptions(quality=100))
    if kind == "gray":  # (sRGB lossless streams are rejected by the reference decoder too)
        assert np.array_equal(j.decode_jp2(data).to_numpy(), px)


@pytest.mark.parametrize("cs", [j.ColorSpace.RGB, j.ColorSpace.YUV, j.ColorSpace.YCBCR])
def test_metrics_other_colour_spaces(gpu, cs):
    # backend.rs:417-421: lossy metrics exist for gray and sRGB only; the
    # lossless shortcut comes first
    px = synth.photo_page(40, 24, seed=2)
    img = _image(px)
    img.colorspace = cs
    with pytest.raises(j.EncodeFailed):
        j.encode_with_psnr(img, j.EncodeOptions(quality=80))
    _, m = j.encode_with_psnr(img, j.EncodeOptions(quality=100))
    assert m.psnr_db == float("inf") and m.ssim == 1.0


def test_metrics_consecutive_calls(gpu, ora):
    # two calls in a row reuse the engine's arena (larger, then smaller)
    a = synth.document_page(300, 200, seed=6)
    b = synth.photo_page(90, 70, seed=5)
    ma = j.encode_with_psnr(_image(a), j.EncodeOptions(quality=40))[1]
    mb = j.encode_with_psnr(_image(b), j.EncodeOptions(quality=70))[1]
    wa, wb = ora.encode_metrics(a, 40), ora.encode_metrics(b, 70)
    assert _close(ma.psnr_db, wa[0]) and _close(ma.ssim, wa[1])
    assert _close(mb.psnr_db, wb[0]) and _close(mb.ssim, wb[1])
