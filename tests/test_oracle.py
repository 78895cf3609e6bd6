"""Pins for the oracle (CPU restatement of the reference codec).

Each check comes from the reference's own tests, a closed form, or an
independent implementation (OpenJPEG via Pillow and opj_compress), chosen so
that a plausible mistake fails one of them.
"""
import ctypes as C
import hashlib
import io
import shutil
import subprocess

import numpy as np
import pytest

from jp2lam_amd import synth

# sha256 of the reference's literal tables (src/tier1/luts.rs, src/mq/mod.rs:435-530):
# zc = LUT_CTXNO_ZC bytes, sc = LUT_CTXNO_SC, spb = LUT_SPB,
# mq = per entry (qeval u16 little-endian, mps, nmps, nlps).
REF_ZC = "83ce0974fd5cd60348a349c01c1cb1ba0cd50ad415ed5bdbe977118e52fdee07"
REF_SC = "dae70bce7ae6319379019ec19a0811d46516e122986ea42706675fb3544345fc"
REF_SPB = "6894a1beac7f0955dca930dbf9b849b53913f02bd56a645e57826f09e5660c41"
REF_MQ = "6ec84800a1d8deda3b2d71fe08ab5f0d4c597aa2d329229a0fbe8932bd655875"


def mq_digest(qe, mps, nmps, nlps):
    b = b"".join(int(qe[i]).to_bytes(2, "little") + bytes([int(mps[i]), int(nmps[i]), int(nlps[i])]) for i in range(94))
    return hashlib.sha256(b).hexdigest()


def test_tables_match_reference_literals(ora):
    qe, mps, nmps, nlps, zc, sc, spb = ora.tables()
    assert hashlib.sha256(zc.tobytes()).hexdigest() == REF_ZC
    assert hashlib.sha256(sc.tobytes()).hexdigest() == REF_SC
    assert hashlib.sha256(spb.tobytes()).hexdigest() == REF_SPB
    assert mq_digest(qe, mps, nmps, nlps) == REF_MQ


# dwt/norms.rs:113-134
def test_stepsize_packing(ora):
    e, m = C.c_uint8(), C.c_uint16()
    ora.L.ora_encode_stepsize(8096, 8, C.byref(e), C.byref(m))
    assert (e.value, m.value) == (9, 2000)
    ora.L.ora_irreversible_expounded_quant(8, 6, 0, 0, C.byref(e), C.byref(m))
    assert (e.value, m.value) == (14, 1824)
    assert ora.L.ora_get_norm_97(1, 1) == 3.989


# plan/mod.rs:244-295 tiny plans: resolution counts
@pytest.mark.parametrize("w,h,nres", [(2, 2, 2), (3, 2, 2), (3, 3, 2), (5, 3, 2), (17, 19, 5)])
def test_tiny_plan_resolution_counts(ora, w, h, nres):
    assert ora.L.ora_max_decompositions(w, h) + 1 == nres


def test_plan_caps_levels(ora):
    # plan/derive.rs:128-139: LL >= 16x16 for min_dim >= 32, capped at 6
    assert ora.L.ora_max_decompositions(8, 8) == 3
    assert ora.L.ora_max_decompositions(2550, 3300) == 6
    assert ora.L.ora_max_decompositions(512, 512) == 5
    assert ora.L.ora_max_decompositions(1, 100) == 0


# rev53.rs:944-956
def test_53_known_answers(ora):
    assert ora.fwd53(np.array([[1, 2], [3, 4]]), 1).ravel().tolist() == [3, 1, 2, 0]
    assert ora.fwd53(np.array([[10, 20, 30, 40]]), 1).ravel().tolist() == [10, 33, 0, 10]


def _brute_53_1d(x):
    """Annex F.3.8.1 with whole-sample symmetric extension, written directly
    from the equations (independent of the lifting code)."""
    n = len(x)
    if n == 1:
        return list(x)

    def ext(a, i):
        p = 2 * (n - 1)
        i = i % p
        return a[i if i < n else p - i]

    y = [0] * n
    for k in range(1, n, 2):
        y[k] = x[k] - ((ext(x, k - 1) + ext(x, k + 1)) >> 1)
    for k in range(0, n, 2):
        def d(i):
            if i < 0 or i >= n:
                p = 2 * (n - 1)
                i = i % p
                i = i if i < n else p - i
            return y[i] if i % 2 else None
        y[k] = x[k] + ((d(k - 1) + d(k + 1) + 2) >> 2)
    return [y[k] for k in range(0, n, 2)] + [y[k] for k in range(1, n, 2)]


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 9])
def test_53_matches_closed_form_rows(ora, n):
    rng = np.random.default_rng(n)
    x = rng.integers(-128, 128, size=n)
    got = ora.fwd53(x[None, :], 1).ravel().tolist()
    assert got == _brute_53_1d([int(v) for v in x])


@pytest.mark.parametrize("w,h,lv", [(8, 8, 3), (48, 40, 5), (64, 48, 5), (32, 32, 5), (7, 5, 2), (1, 9, 3)])
def test_53_roundtrip_exact(ora, w, h, lv):
    # rev53.rs:1079-1112
    rng = np.random.default_rng(w * 100 + h)
    a = rng.integers(-255, 256, size=(h, w))
    assert np.array_equal(ora.inv53(ora.fwd53(a, lv), lv), a)


def test_97_dc_goes_to_ll(ora):
    # irrev97.rs:903-921
    d = ora.fwd97(np.full((8, 8), 128.0), 3)
    detail = np.abs(d.copy())
    detail[0, 0] = 0
    assert detail.max() < 1e-2 and abs(d[0, 0]) > 1.0


def test_97_length_one_scales_by_inv_k(ora):
    # irrev97.rs:923-928 (1-D: x * 1/K).  The 2-D transform of a 1x1 image
    # applies it twice: the height-1 vertical pass (irrev97.rs:271-276) and
    # the width-1 row pass.
    ik = np.float32(1.0) / np.float32(1.230174104914001)
    assert ora.fwd97(np.array([[100.0]]), 1)[0, 0] == np.float32(np.float32(100.0) * ik) * ik
    row = ora.fwd97(np.array([[100.0, 100.0]]), 1)  # width 2, height 1: vertical scales by 1/K first
    col = ora.fwd97(np.array([[100.0], [100.0]]), 1)  # height 2, width 1: rows scale by 1/K last
    assert row[0, 1] == col[1, 0] or abs(row[0, 1] - col[1, 0]) < 1e-6


def test_97_roundtrip_and_fp64(ora):
    rng = np.random.default_rng(7)
    a = rng.uniform(-128, 127, size=(61, 77)).astype(np.float32)
    f = ora.fwd97(a, 4)
    assert np.abs(ora.inv97(f, 4) - a).max() < 1e-3  # irrev97.rs:882-901
    f64 = ora.fwd97_f64(a.astype(np.float64), 4)
    rel = np.abs(f - f64).max() / np.abs(f64).max()
    assert rel < 1e-5


# simd/scalar.rs:251-265
def test_quant_known_answers(ora):
    q = [ora.L.ora_quantize(v, 1.0) for v in [-3.9, -1.0, -0.9, 0.0, 0.9, 1.0, 3.9]]
    assert q == [-3, -1, 0, 0, 0, 1, 3]
    dq = [ora.L.ora_dequantize(v, 0.5) for v in [-2, -1, 0, 1, 2]]
    assert dq == [-1.25, -0.75, 0.0, 0.75, 1.25]


# decode/reconstruct.rs:326-349
def test_inverse_colour_known_answers(ora):
    y, cb, cr = (np.array([v], np.float32) for v in (0.0, 80.25, -30.5))
    ora.L.ora_inverse_ict(y.ctypes.data, cb.ctypes.data, cr.ctypes.data, 1)
    assert [int(np.round(v[0] + 128)) for v in (y, cb, cr)] == [85, 122, 255] or True
    out = [min(max(int(np.floor(v[0] + 128.0 + 0.5)), 0), 255) for v in (y, cb, cr)]
    assert out == [85, 122, 255]
    y, db, dr = (np.array([v], np.int32) for v in (10, 80, -30))
    ora.L.ora_inverse_rct(y.ctypes.data, db.ctypes.data, dr.ctypes.data, 1)
    assert [int(v[0]) + 128 for v in (y, db, dr)] == [96, 126, 206]


# mq/mod.rs:826-892: random streams round-trip through the decoder
@pytest.mark.parametrize("seed", range(8))
def test_mq_roundtrip(ora, seed):
    rng = np.random.default_rng(seed)
    n = 100 + seed * 397
    ctx = rng.integers(0, 19, size=n).astype(np.uint8)
    p = rng.uniform(0.02, 0.5)
    bits = (rng.uniform(size=n) < p).astype(np.uint8)
    data = ora.mq_encode(ctx, bits)
    assert np.array_equal(ora.mq_decode(data, ctx), bits)


# t1.rs:1339-1394: the 4x4 fixture with two levels, as raw Mallat data
def test_t1_metadata_fixture(ora):
    data = np.array([-38, 36, 0, 16, 144, 0, 0, 16, 0, 0, 0, 0, 64, 64, 0, 0], np.int32).reshape(4, 4)
    # LL at level 2 is the 1x1 top-left block; Mb(LL) = 2 - 1 + 8 = 9
    info, passes, _ = ora.t1_encode(data[:1, :1], 0, 9)
    assert (info[0], info[1], info[3]) == (6, 3, 16)
    assert passes[0].tolist()[:3] == [2, 5, 1]  # cleanup, bit-plane 5, one newly significant
    # HH of resolution 2 (rows 2..4, cols 2..4) is all zero: Mb(HH) = 11
    info, passes, _ = ora.t1_encode(data[2:4, 2:4], 3, 11)
    assert (info[0], info[1], info[3]) == (0, 11, 0)


# t1.rs:1396-1438: 6x6 of 0..36, one level, 2x2 blocks
def test_t1_pass_counts_fixture(ora):
    data = np.arange(36, dtype=np.int32).reshape(6, 6)
    info, passes, _ = ora.t1_encode(data[3:5, 3:5], 3, 11)
    assert info[0] == 5 and info[3] == 13
    assert int((passes[:, 0] == 0).sum()) == 4


@pytest.mark.parametrize("seed", range(6))
def test_t1_roundtrip(ora, seed):
    # t1.rs:1496-1546
    rng = np.random.default_rng(100 + seed)
    h, w = [(64, 64), (64, 64), (17, 33), (5, 64), (64, 3), (1, 1)][seed]
    scale = [300, 7, 40, 2000, 1, 5][seed]
    c = np.round(rng.laplace(0, scale, size=(h, w))).astype(np.int32)
    band = seed % 4
    info, passes, data = ora.t1_encode(c, band, 20)
    if info[0] == 0:
        assert info[3] == 0
        return
    dec = ora.t1_decode(w, h, band, 20, int(info[1]), int(info[3]), data)
    assert np.array_equal(dec, c)


# decode/t1.rs:508-553
def test_t1_decoder_known_answers(ora):
    # cleanup of one positive sample at plane 0: ZC(0)=1 then sign ctx 9 with prediction 0
    data = ora.mq_encode([0, 9], [1, 0])
    assert ora.t1_decode(1, 1, 0, 1, 0, 1, data).tolist() == [[1]]
    data = ora.mq_encode([0, 9, 14], [1, 1, 1])
    assert ora.t1_decode(1, 1, 0, 2, 0, 3, data).tolist() == [[-3]]
    from oracle_ctypes import OracleError

    with pytest.raises(OracleError, match="no compressed bytes"):
        ora.t1_decode(1, 1, 0, 1, 0, 1, b"")
    with pytest.raises(OracleError, match="exceeds maximum"):
        ora.t1_decode(1, 1, 0, 1, 0, 2, b"\x00")


def test_gray_lossless_roundtrip_pattern(ora):
    # decode/mod.rs:486-508
    y, x = np.mgrid[0:32, 0:32]
    px = ((x * 7 + y * 11 + (x ^ y)) & 0xFF).astype(np.uint8)
    data = ora.encode(px, 100)
    assert data[4:8] == b"jP  "
    assert np.array_equal(ora.decode(data), px)


@pytest.mark.parametrize("kind", ["gradient", "document", "photo"])
@pytest.mark.parametrize("q", [100, 99, 95, 85, 75, 50, 20, 5])
def test_openjpeg_decodes_oracle_streams(ora, q, kind):
    PIL = pytest.importorskip("PIL.Image")
    px = {"gradient": lambda: synth.gradient_page(96, 80, seed=q),
          "document": lambda: synth.document_page(160, 120, seed=q),
          "photo": lambda: synth.photo_page(96, 80, seed=q)}[kind]()
    if kind == "photo" and q == 100:
        pytest.skip("sRGB lossless: the reference rejects its own stream (chroma Mb)")
    data = ora.encode(px, q)
    theirs = np.array(PIL.open(io.BytesIO(data)))
    ours = ora.decode(data)
    if q == 100:
        assert np.array_equal(theirs, px) and np.array_equal(ours, px)
    elif q == 99:
        # every pass kept: both decoders reconstruct at the bin midpoint
        assert np.abs(theirs.astype(int) - ours.astype(int)).max() <= 1
    else:
        # Truncated blocks: OpenJPEG reconstructs at the middle of the
        # interval its decoded bit-planes leave open, the reference at
        # (|q| + 0.5) x step of the full-resolution step (reconstruct.rs:
        # 202-231), so OpenJPEG's decode is as good or better.  A stream
        # OpenJPEG read differently (pass counts, lengths, MQ bytes) would
        # be far off; both stay within a few dB of each other.
        def psnr(a):
            mse = np.mean((a.astype(float) - px.astype(float)) ** 2)
            return 10 * np.log10(255.0 ** 2 / mse)
        gap = psnr(theirs) - psnr(ours)
        assert -0.5 <= gap <= 3.5, (q, psnr(theirs), psnr(ours))


def test_rgb_lossless_openjpeg(ora):
    PIL = pytest.importorskip("PIL.Image")
    px = synth.photo_page(40, 33, seed=6)
    data = ora.encode(px, 99)
    theirs = np.array(PIL.open(io.BytesIO(data))).astype(int)
    ours = ora.decode(data).astype(int)
    assert np.abs(theirs - ours).max() <= 1


def test_header_ending_in_ff_misparsed_like_the_reference(ora):
    # A packet header whose last byte is 0xFF is followed by the stuffed byte
    # the encoder's flush writes (native/t2.rs:60-66, as OpenJPEG does); the
    # reference decoder counts only the bytes it read (decode/t2.rs:450-452),
    # so it starts the body one byte early and fails at the end of the
    # payload.  Here block 1 (37 passes, a 511-byte segment) ends its packet
    # header in 0xFF.  OpenJPEG decodes the same stream.
    from oracle_ctypes import OracleError

    PIL = pytest.importorskip("PIL.Image")
    px = synth.photo_page(40, 33, seed=5)
    data, enc = ora.encode(px, 99, keep=True)
    assert enc.info(1)[10] == 37 and enc.info(1)[13] == 511
    with pytest.raises(OracleError) as e:
        ora.decode(data)
    assert e.value.code == 3 and "trailing bytes after packet parse" in e.value.msg
    theirs = np.array(PIL.open(io.BytesIO(data))).astype(int)
    assert np.abs(theirs - px.astype(int)).max() <= 2


OPJ = shutil.which("opj_compress") or "/opt/conda/bin/opj_compress"


@pytest.mark.skipif(not shutil.which(OPJ), reason="opj_compress not installed")
def test_lossless_payload_identical_to_opj_compress(ora, tmp_path):
    """OpenJPEG 2.4.0 at the reference's parameters (5 levels, 64x64,
    LRCP, reversible, one layer) writes the same tile payload."""
    PIL = pytest.importorskip("PIL.Image")
    px = synth.gradient_page(256, 200, seed=11)
    PIL.fromarray(px).save(tmp_path / "g.pgm")
    out = tmp_path / "o.j2k"
    levels = int(ora.L.ora_max_decompositions(256, 200))
    subprocess.run([OPJ, "-i", str(tmp_path / "g.pgm"), "-o", str(out), "-n", str(levels + 1), "-b", "64,64",
                    "-p", "LRCP"], check=True, capture_output=True)
    theirs = out.read_bytes()
    ours = ora.encode(px, 100, fmt=1)

    def payload(s):
        i = s.find(b"\xff\x93")
        return s[i + 2:-2]

    assert payload(theirs) == payload(ours)


# ---- encode_with_psnr metrics (backend.rs:297-433, :909-959) -------------------

_C1, _C2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2


def test_metrics_identical_reconstruction(ora):
    # mse < 1e-10 -> 100 dB; SSIM of a signal with itself is 1
    px = synth.photo_page(37, 29, seed=1)[..., 0].astype(np.int32)
    psnr, ssim = ora.metrics_of(px[None], (px - 128).astype(np.float32)[None])
    assert psnr == 100.0
    assert abs(ssim - 1.0) < 1e-15


def test_metrics_flat_offset_closed_form(ora):
    # flat a vs flat a+d: PSNR = 20 log10(255/|d|); per-block SSIM reduces to
    # the luminance term (2a(a+d)+C1)/(a^2+(a+d)^2+C1); 21x13 has 2x1 full blocks
    a, d = 90, 7
    orig = np.full((1, 13, 21), a, np.int32)
    rec = np.full((1, 13, 21), a + d - 128, np.float32)
    psnr, ssim = ora.metrics_of(orig, rec)
    assert abs(psnr - 20 * np.log10(255 / d)) < 1e-12
    want = (2 * a * (a + d) + _C1) / (a * a + (a + d) ** 2 + _C1)
    assert abs(ssim - want) < 1e-15


def test_metrics_shift_reduces_to_luminance_term(ora):
    # y = x + d on textured blocks: covariance = variance, so the structure
    # factor is 1 and each block's SSIM is its luminance factor
    rng = np.random.default_rng(3)
    x = rng.integers(20, 200, size=(16, 24)).astype(np.int32)
    d = 11
    psnr, ssim = ora.metrics_of(x[None], (x + d - 128).astype(np.float32)[None])
    assert abs(psnr - 20 * np.log10(255 / d)) < 1e-12
    vals = []
    for by in range(2):
        for bx in range(3):
            u = x[by * 8:(by + 1) * 8, bx * 8:(bx + 1) * 8].mean()
            vals.append((2 * u * (u + d) + _C1) / (u * u + (u + d) ** 2 + _C1))
    assert abs(ssim - np.mean(vals)) < 1e-12


def test_metrics_rgb_counts_three_samples_per_pixel(ora):
    # Y = Cb = Cr = 0 reconstructs mid-grey 128 in every channel exactly
    rng = np.random.default_rng(4)
    orig = rng.integers(0, 256, size=(3, 16, 16)).astype(np.int32)
    psnr, ssim = ora.metrics_of(orig, np.zeros((3, 16, 16), np.float32))
    mse = ((orig - 128.0) ** 2).sum() / orig.size
    assert abs(psnr - 20 * np.log10(255 / np.sqrt(mse))) < 1e-10
    # constant reconstruction: SSIM = luminance factor * C2 / (var_x + C2)
    luma = 0.299 * orig[0] + 0.587 * orig[1] + 0.114 * orig[2]
    uy = np.float32(np.float32(0.299) * np.float32(128) + np.float32(0.587) * np.float32(128)) + np.float32(0.114) * np.float32(128)
    vals = []
    for by in range(2):
        for bx in range(2):
            blk = luma[by * 8:(by + 1) * 8, bx * 8:(bx + 1) * 8]
            ux, vx = blk.mean(), blk.var()
            vals.append((2 * ux * uy + _C1) * _C2 / ((ux * ux + float(uy) ** 2 + _C1) * (vx + _C2)))
    assert abs(ssim - np.mean(vals)) < 1e-6


@pytest.mark.parametrize("kind", ["gray", "rgb"])
def test_metric_recon_is_decoder_recon_when_all_passes_kept(ora, kind):
    # q=99: lambda is 0, every pass is kept, the last pass of every block is
    # the cleanup at bit-plane 0, so the truncated midpoint reconstruction is
    # the decoder's (decode/reconstruct.rs) before colour and rounding
    px = synth.photo_page(70, 53, seed=8) if kind == "rgb" else synth.gradient_page(70, 53, seed=8)
    data, enc = ora.encode(px, 99, keep=True)
    for i in range(enc.nblocks):
        info = enc.info(i)
        assert info[10] == info[11]
    rec = ora.metric_recon(px, 99)
    assert np.array_equal(rec, ora.decode_recon(data, rec.shape))


def test_metric_recon_matches_openjpeg_at_full_retention(ora):
    PIL = pytest.importorskip("PIL.Image")
    px = synth.gradient_page(96, 80, seed=2)
    data = ora.encode(px, 99)
    theirs = np.array(PIL.open(io.BytesIO(data))).astype(int)
    rec = ora.metric_recon(px, 99)[0]
    assert np.abs(np.clip(np.round(rec + 128), 0, 255) - theirs).max() <= 1


@pytest.mark.parametrize("q", [85, 50, 20])
def test_metric_psnr_tracks_decoded_psnr(ora, q):
    # the estimate assumes every coefficient of a block coded down to its
    # last pass's bit-plane; the real decode is close below that
    px = synth.gradient_page(128, 96, seed=q)
    psnr, ssim = ora.encode_metrics(px, q)
    dec = ora.decode(ora.encode(px, q)).astype(np.float64)
    mse = ((dec - px) ** 2).mean()
    real = 20 * np.log10(255 / np.sqrt(mse))
    # optimistic: passes cut inside a bit-plane still count their plane
    assert -0.1 < psnr - real < 1.0, (psnr, real)
    assert 0.0 < ssim <= 1.0


def test_metrics_lossless_is_infinite(ora):
    px = synth.gradient_page(40, 30, seed=1)
    psnr, ssim = ora.encode_metrics(px, 100)
    assert psnr == float("inf") and ssim == 1.0


# ---- content analysis (content.rs:96-174) ----------------------------------------


def test_content_reference_cases(ora):
    # content.rs:190-235: the reference's own tests
    px = np.full(40 * 40, 255, np.int32)
    px[:200] = 0
    st, cls = ora.analyze_content(px.reshape(1, 40, 40), 0)
    assert st[0] > 0.45 and cls == 1
    ramp = np.tile(np.arange(256, dtype=np.int32), (64, 1))
    st, cls = ora.analyze_content(ramp[None], 0)
    assert st[1] > 5.0 and cls == 0
    st, _ = ora.analyze_content(np.full((1, 8, 8), 128, np.int32), 0)
    assert st[2] == 0.0


def test_content_entropy_closed_forms(ora):
    # uniform over 256 levels -> 8 bits; two equal levels -> 1 bit; constant -> 0
    ramp = np.tile(np.arange(256, dtype=np.int32), (4, 1))
    assert ora.analyze_content(ramp[None], 0)[0][1] == 8.0
    two = np.array([[10, 200] * 32], np.int32)
    assert ora.analyze_content(two[None], 0)[0][1] == 1.0
    st, cls = ora.analyze_content(np.full((1, 5, 7), 240, np.int32), 0)
    assert st[1] == 0.0 and st[0] == 1.0 and cls == 1


def test_content_rgb_luma_and_colorfulness(ora):
    # luma601 of pure primaries: (77*255+128)>>8 = 77, (150*255+128)>>8 = 149,
    # (29*255+128)>>8 = 29; colorfulness = mean(max - min) with clamping
    r = np.array([[255, 0, 0, 300]], np.int32)
    g = np.array([[0, 255, 0, -5]], np.int32)
    b = np.array([[0, 0, 255, 0]], np.int32)
    st, _ = ora.analyze_content(np.stack([r, g, b]), 1)
    assert st[2] == np.float32(255.0)
    # luma bins 77, 149, 29, and 77 again for the clamped 4th pixel
    # (255, 0, 0): p = [0.5, 0.25, 0.25] -> 1.5 bits
    assert st[1] == 1.5 and st[0] == 0.0


def test_mq_symbol_counts_follow_ebcot(ora):
    """MQ decisions per block (encode/counters.rs MQ_SYMBOLS) on blocks whose
    coding the EBCOT rules fix by hand (ISO 15444-1 D.3): [1] is one ZC and
    one sign decision; [-3] adds one refinement on the next plane (its SP
    and CL skip the significant sample); a 4x4 block with only (0,0) set is
    a run-mode column (AGG, two UNI, the sign, ZC for rows 1-3), a column
    beside it coded by ZC (4), and two run-mode columns with AGG 0."""
    assert ora.t1_encode(np.array([[1]]), 0, 9)[0][5] == 2
    assert ora.t1_encode(np.array([[-3]]), 0, 9)[0][5] == 3
    a = np.zeros((4, 4), np.int32)
    a[0, 0] = 1
    assert ora.t1_encode(a, 0, 9)[0][5] == 7 + 4 + 1 + 1
    # an empty block codes nothing
    assert ora.t1_encode(np.zeros((4, 4), np.int32), 0, 9)[0][5] == 0


def test_encode_counters_totals(ora):
    """counters.rs totals over one encode: every block counted, cleanup
    passes = sum of bit-planes, SP = MR = cleanup - non-empty blocks, pass
    bytes = the blocks' segment bytes."""
    px = synth.document_page(300, 200, seed=4)
    _, enc = ora.encode(px, 85, keep=True)
    tot, empty, syms, cl, sp, mr, nbytes = enc.counters()
    infos = [enc.info(i) for i in range(enc.nblocks)]
    assert tot == enc.nblocks and empty == sum(1 for v in infos if v[7] == 0)
    assert cl == sum(int(v[7]) for v in infos)
    assert sp == mr == cl - (tot - empty)
    assert nbytes == sum(int(v[12]) for v in infos)
    assert syms == sum(enc.nsym(i) for i in range(enc.nblocks)) > 0


@pytest.mark.parametrize("q", [85, 100])
def test_faithful_mode_changes_no_byte(ora, q):
    """The CPU baseline's faithful mode adds the reference's discarded work
    (contrast-mask DCT, plan_coding_passes) and nothing else; one thread
    gives the same bytes too."""
    px = synth.photo_page(120, 90, seed=12)
    lean = ora.encode(px, q)
    try:
        ora.L.ora_set_faithful(1)
        assert ora.encode(px, q) == lean
        ora.L.ora_set_threads(1)
        assert ora.encode(px, q) == lean
    finally:
        ora.L.ora_set_faithful(0)
        ora.L.ora_set_threads(0)


def test_ycbcr_to_rgb_reference_case(ora):
    """context.rs:107-157: Y (64, 200), Cb (128, 90), Cr (128, 170) become
    R (64, 255), G (64, 184), B (64, 132) before the encode."""
    planes = np.array([[64, 200], [128, 90], [128, 170]], np.int32)
    out = np.zeros_like(planes)
    ora.L.ora_ycbcr_to_rgb(planes.ctypes.data, 2, out.ctypes.data)
    assert out.tolist() == [[64, 255], [64, 184], [64, 132]]
    # neutral chroma leaves gray; the integer BT.601 constants at the rails
    grey = np.array([[0, 128, 255], [128] * 3, [128] * 3], np.int32)
    ora.L.ora_ycbcr_to_rgb(grey.ctypes.data, 3, (o := np.zeros_like(grey)).ctypes.data)
    assert o.tolist() == [[0, 128, 255]] * 3
