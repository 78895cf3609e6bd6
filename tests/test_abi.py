"""C-ABI checks that need no GPU: the library loads, exports every symbol
the header declares, carries the reference's tables, and the host-side
parsing and validation behave like the reference."""
import ctypes as C
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

import jp2lam_amd as j
from jp2lam_amd import _lib, synth
from test_oracle import REF_MQ, REF_SC, REF_SPB, REF_ZC, mq_digest

HEADER = Path(__file__).resolve().parents[1] / "include" / "jp2lam_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(jp2(?:lam|g)_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(lib):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound


def test_product_tables_match_reference(lib, ora):
    qe = np.zeros(94, np.uint16)
    mps, nmps, nlps = (np.zeros(94, np.uint8) for _ in range(3))
    zc = np.zeros(2048, np.uint8)
    sc, spb = np.zeros(256, np.uint8), np.zeros(256, np.uint8)
    lib.jp2g_tables(*(a.ctypes.data for a in (qe, mps, nmps, nlps, zc, sc, spb)))
    assert hashlib.sha256(zc.tobytes()).hexdigest() == REF_ZC
    assert hashlib.sha256(sc.tobytes()).hexdigest() == REF_SC
    assert hashlib.sha256(spb.tobytes()).hexdigest() == REF_SPB
    assert mq_digest(qe, mps, nmps, nlps) == REF_MQ


@pytest.mark.parametrize("q,fmt", [(100, 0), (85, 0), (20, 0)])
def test_inspect_matches_plan(ora, q, fmt):
    px = synth.gradient_page(70, 45, seed=q)
    data = ora.encode(px, q, fmt)
    m = j.inspect_jp2(data)
    assert (m.width, m.height, m.num_components) == (70, 45, 1)
    assert m.colorspace == j.ColorSpace.GRAY and m.bits_per_component == 8
    assert m.transform == (1 if q == 100 else 0)
    assert m.code_block_width == (32 if q < 30 else 64)
    assert m.decomposition_levels == ora.L.ora_max_decompositions(70, 45)
    assert m.layers == 1 and m.tile_part_count == 1 and m.num_quant_steps == 1 + 3 * m.decomposition_levels


def test_inspect_rejects_bad_streams(ora):
    data = ora.encode(synth.gradient_page(16, 16), 85)
    with pytest.raises(j.DecodeFailed):
        j.inspect_jp2(b"not a jp2 file at all")
    with pytest.raises(j.DecodeFailed):
        j.inspect_jp2(data[:40])
    cs = data[data.find(b"jp2c") + 4:]
    with pytest.raises(j.DecodeFailed):  # raw J2K is rejected (jp2_parse.rs:42-49)
        j.inspect_jp2(cs)
    bad = bytearray(data)
    i = bad.find(b"colr") + 4 + 3
    bad[i:i + 4] = (99).to_bytes(4, "big")  # EnumCS 99
    with pytest.raises(j.UnsupportedFeature):
        j.inspect_jp2(bytes(bad))


def test_validation_errors_before_device():
    # plan/validate.rs:4-62 -- raised before any device work
    comp = j.Component(np.zeros(4, np.int32), 2, 2)
    with pytest.raises(j.InvalidInput, match="non-zero"):
        j.encode(j.Image(0, 2, [j.Component(np.zeros(0, np.int32), 0, 2)]))
    with pytest.raises(j.InvalidInput, match="component"):
        j.encode(j.Image(2, 2, [comp, comp], j.ColorSpace.GRAY))
    with pytest.raises(j.InvalidInput, match="precision"):
        j.encode(j.Image(2, 2, [j.Component(np.zeros(4, np.int32), 2, 2, precision=12)]))
    with pytest.raises(j.InvalidInput, match="unsigned"):
        j.encode(j.Image(2, 2, [j.Component(np.zeros(4, np.int32), 2, 2, signed=True)]))
    with pytest.raises(j.InvalidInput, match="subsampling"):
        j.encode(j.Image(2, 2, [j.Component(np.zeros(4, np.int32), 2, 2, dx=2)]))
    with pytest.raises(j.InvalidInput):
        j.Image.from_gray_bytes(3, 3, b"\x00" * 8)
    with pytest.raises(j.InvalidInput):
        j.Image.from_rgb_bytes(2, 2, b"\x00" * 11)


def test_error_classes():
    e = j.DecodeFailed("x")
    assert e.is_decode_failure() and not e.is_unsupported_feature()
    assert str(e) == "decode failed: x"
    u = j.UnsupportedFeature("y")
    assert u.is_decode_failure() and u.is_unsupported_feature()
    assert j.EncodeOptions().quality == 85 and j.EncodeOptions().format == j.OutputFormat.JP2
    assert j.EncodeOptions.from_preset(j.Preset.DOCUMENT_LOW).quality == 30


def test_colorspace_numbering_follows_model_rs():
    """ColorSpace discriminants in src/model.rs:1-12 declaration order, in
    the header, the Python enum and the library's own view validation."""
    text = HEADER.read_text()
    vals = dict(re.findall(r"JP2LAM_(GRAY|RGB|SRGB|YUV|YCBCR) = (\d)", text))
    assert vals == {"GRAY": "0", "RGB": "1", "SRGB": "2", "YUV": "3", "YCBCR": "4"}
    assert [c.name for c in sorted(j.ColorSpace)] == ["GRAY", "RGB", "SRGB", "YUV", "YCBCR"]
    comp = j.Component(np.zeros(4, np.int32), 2, 2)
    # a gray view with three components is rejected with the reference's
    # name for colour space 0; colour space 2 (Srgb) with one component too
    with pytest.raises(j.InvalidInput, match="Gray images must have exactly 1"):
        j.encode(j.Image(2, 2, [comp] * 3, j.ColorSpace.GRAY))
    with pytest.raises(j.InvalidInput, match="Srgb images must have exactly 3"):
        j.encode(j.Image(2, 2, [comp], j.ColorSpace.SRGB))
    with pytest.raises(j.InvalidInput, match="Rgb images must have exactly 3"):
        j.encode(j.Image(2, 2, [comp], j.ColorSpace.RGB))


def test_batch_profile_getters_without_device(lib):
    enc = j.BatchEncoder(j.EncodeOptions(50, j.OutputFormat.J2K))
    assert enc.profile() is None and not enc.has_profile
    dec = j.BatchDecoder()
    assert dec.profile() is None
    p = _lib.BatchProfileC()
    assert lib.jp2lam_batch_encoder_profile(None, C.byref(p), None, 0) == -1
    assert lib.jp2lam_batch_decoder_profile(None, C.byref(p), None, 0) == -1
