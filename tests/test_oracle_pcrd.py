"""Pins for the oracle's rate allocation (PCRD) and lossy step scaling.

The lossy codestream bytes depend on the distortion model, the convex hull,
the quality-to-lambda curve, the lambda calibration and the step scaler.
The product equals the oracle there byte for byte, so these tests hold the
oracle itself to the reference, without running it:

- the reference's own unit tests of dwt/pcrd.rs:1020-1361 and
  encode/backend/native/rate.rs:286-324, restated with the same inputs and
  assertions;
- the calibration the reference documents for the step scaler
  (plan/derive.rs:72-73) and the lambda curve (dwt/pcrd.rs:1417-1424),
  checked both on the scaler and on the (exponent, mantissa) pairs the
  encoder writes;
- a brute-force Lagrangian check: on small images each code-block's kept
  pass count maximises cumulative distortion reduction minus lambda times
  bytes over every truncation point of its raw curve, and every raw point
  lies on or below the hull;
- lambda calibration against its documented closed form.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from jp2lam_amd import synth
from oracle_ctypes import RdPoint

SP, MR, CL = 0, 1, 2
LL, HL, LH, HH = 0, 1, 2, 3
FLAT, GRADIENT, TEXTURE = 0, 1, 2

NONMONOTONE_CUMULATIVE, INCONSISTENT_INCREMENT, INVALID_DELTA, INVALID_ORIGIN = 1, 2, 3, 4


def sample_curve(ora):
    # pcrd.rs:1021-1036: incremental bytes 10 each, distortion 100, 50, 20, 5
    return ora.hull_curve([(0, 10, 10, 100.0), (1, 10, 20, 50.0), (2, 10, 30, 20.0), (3, 10, 40, 5.0)])


def total_bytes(ora, curves, lam):
    tot = 0
    for c in curves:
        rc, p = ora.choose(c, lam)
        assert rc == 0
        tot += p.bytes
    return tot


# ---- dwt/pcrd.rs unit tests ------------------------------------------------------
def test_build_raw_curve_includes_omitted_origin(ora):
    rc, pts = ora.raw_curve([(0, 4, 4, 9.0), (1, 3, 7, 4.0)])
    assert rc == 0 and len(pts) == 3
    assert [(p.passes, p.bytes) for p in pts] == [(0, 0), (1, 4), (2, 7)]
    assert abs(pts[2].dist - 13.0) < 1e-9


def test_hull_prunes_non_convex_middle_points(ora):
    rc, raw = ora.raw_curve([(0, 10, 10, 100.0), (1, 10, 20, 120.0), (2, 10, 30, 10.0)])
    hull = ora.hull(raw)
    assert len(hull) < len(raw)
    for i in range(2, len(hull)):
        assert hull[i].slope < hull[i - 1].slope


def test_choose_block_point_obeys_lambda_threshold(ora):
    curve = sample_curve(ora)
    got = [ora.choose(curve, lam)[1].passes for lam in (20.0, 6.0, 3.0, 0.75, 0.1)]
    assert got == [0, 1, 2, 3, 4]


def test_choose_block_point_keeps_equal_slopes_and_stops_at_first_below(ora):
    """slope >= lambda keeps a point (pcrd.rs:414), and the walk stops at
    the first point below lambda even when a later one is steeper
    (pcrd.rs:417-420; only hulls reach it on the path)."""
    curve = sample_curve(ora)  # slopes 10, 5, 2, 0.5
    assert [ora.choose(curve, lam)[1].passes for lam in (10.0, 5.0, 2.0, 0.5)] == [1, 2, 3, 4]
    rc, raw = ora.raw_curve([(0, 10, 10, 100.0), (1, 10, 20, 10.0), (2, 10, 30, 120.0)])  # 10, 1, 12
    assert ora.choose(raw, 5.0)[1].passes == 1


def test_evaluate_lambda_is_monotone_in_total_bytes(ora):
    curves = [sample_curve(ora), sample_curve(ora),
              ora.hull_curve([(0, 8, 8, 64.0), (1, 8, 16, 16.0), (2, 8, 24, 4.0)])]
    hi, mid, lo = (total_bytes(ora, curves, lam) for lam in (100.0, 4.0, 0.1))
    assert hi <= mid <= lo


def test_evaluate_lambda_monotone_in_quality(ora):
    curves = [sample_curve(ora), sample_curve(ora)]
    assert total_bytes(ora, curves, 50.0) <= total_bytes(ora, curves, 1.0)


def test_build_raw_curve_rejects_nonmonotone_cumulative_bytes(ora):
    rc, _ = ora.raw_curve([(0, 10, 10, 5.0), (1, 5, 8, 3.0)])
    assert rc == NONMONOTONE_CUMULATIVE


def test_build_raw_curve_rejects_inconsistent_incremental_bytes(ora):
    rc, _ = ora.raw_curve([(0, 10, 10, 5.0), (1, 3, 20, 3.0)])
    assert rc == INCONSISTENT_INCREMENT


def test_build_raw_curve_rejects_negative_distortion_delta(ora):
    assert ora.raw_curve([(0, 10, 10, -1.0)])[0] == INVALID_DELTA
    assert ora.raw_curve([(0, 10, 10, math.nan)])[0] == INVALID_DELTA
    assert ora.raw_curve([(0, 10, 10, math.inf)])[0] == INVALID_DELTA


def test_validate_curve_rejects_bad_origin(ora):
    bad = [RdPoint(1, 0, 0.0, math.inf), RdPoint(2, 10, 5.0, 0.5)]
    assert ora.choose(bad, 0.1)[0] == INVALID_ORIGIN


def test_hull_dominates_equal_slope_middle_points(ora):
    rc, raw = ora.raw_curve([(0, 10, 10, 100.0), (1, 10, 20, 100.0), (2, 10, 30, 20.0)])
    hull = ora.hull(raw)
    assert len(hull) < len(raw)
    for a, b in zip(hull[1:], hull[2:]):
        assert b.slope < a.slope


def test_distortion_prefers_higher_bitplanes(ora):
    f = ora.L.ora_pass_delta_pass_kind_aware
    low = f(SP, 2, 3, 0, 1.0, 1.0, LL, 75, TEXTURE)
    high = f(SP, 5, 3, 0, 1.0, 1.0, LL, 75, TEXTURE)
    assert high > low


def test_distortion_models_agree_on_sp_passes(ora):
    # pcrd.rs:1314-1337: the pass-kind-aware model adds the LL band bias
    pka = ora.L.ora_pass_delta_pass_kind_aware(SP, 4, 8, 0, 1.0, 1.5, LL, 75, TEXTURE)
    base = ora.L.ora_pass_delta_baseline(4, 8, 0, 1.0, 1.5)
    assert pka > base


def test_distortion_model_reduces_to_baseline_where_weights_are_one(ora):
    """Where band bias, class weight and masking are all 1 (HL/LH, texture
    class, q >= 50) the pass-kind-aware model is the baseline model for SP
    and MR passes, and cleanup is 0.85 of a significance pass
    (pcrd.rs:652-696, :728-748, profile.rs:141-154)."""
    pka, base = ora.L.ora_pass_delta_pass_kind_aware, ora.L.ora_pass_delta_baseline
    for band in (HL, LH):
        for bp in (0, 3, 11, 30, 35):
            for step in (0.37, 1.0, 2.5):
                sp = pka(SP, bp, 7, 0, step, 1.7, band, 80, TEXTURE)
                assert sp == pytest.approx(base(bp, 7, 0, step, 1.7), rel=1e-15)
                mr = pka(MR, bp, 0, 9, step, 1.7, band, 80, TEXTURE)
                assert mr == pytest.approx(base(bp, 0, 9, step, 1.7), rel=1e-15)
                assert pka(CL, bp, 7, 0, step, 1.7, band, 80, TEXTURE) == pytest.approx(0.85 * sp, rel=1e-15)
    # bit-planes above 30 are capped (pcrd.rs:653)
    assert base(31, 1, 0, 1.0, 1.0) == base(30, 1, 0, 1.0, 1.0) == float(1 << 60)
    # LL / HH and the sparser classes weigh by the documented factors
    b = base(4, 8, 0, 1.0, 1.0)
    assert pka(SP, 4, 8, 0, 1.0, 1.0, LL, 75, TEXTURE) == pytest.approx(1.2 * 1.1 * b)
    assert pka(SP, 4, 8, 0, 1.0, 1.0, HH, 75, TEXTURE) == pytest.approx(0.82 * b)
    assert pka(SP, 4, 8, 0, 1.0, 1.0, HL, 75, FLAT) == pytest.approx(2.5 * b)
    assert pka(SP, 4, 8, 0, 1.0, 1.0, HL, 75, GRADIENT) == pytest.approx(1.5 * b)
    assert pka(SP, 4, 8, 0, 1.0, 1.0, LL, 0, TEXTURE) == pytest.approx(3.0 * 1.1 * b)
    assert pka(SP, 4, 8, 0, 1.0, 1.0, HH, 25, TEXTURE) == pytest.approx((0.20 + 0.62 * 0.5) * b)
    assert pka(SP, 4, 0, 0, 1.0, 1.0, HL, 75, TEXTURE) == 0.0


# ---- lambda curve and calibration ---------------------------------------------------
def test_quality_to_lambda_documented_points(ora):
    lam = ora.L.ora_quality_to_lambda
    # pcrd.rs:1417-1424: q=50 -> about 1.7e4 at the 12 Mpx reference size
    assert lam(50, 12_000_000) == pytest.approx(1.7e4, rel=0.01)
    assert lam(100, 12_000_000) == 0.0 and lam(255, 5) == 0.0
    assert lam(0, 12_000_000) == np.finfo(np.float64).max
    # the q90..q99 tail: factor 1 at q90, 0 at q99
    t = (5.68 - 2.618 * (89.0 / 98.0) ** 0.85)
    assert lam(90, 12_000_000) == pytest.approx(10 ** t, rel=1e-12)
    assert lam(99, 12_000_000) == 0.0
    # resolution factor (P / 12e6)^0.35, exactly multiplicative
    for q in (1, 30, 50, 85):
        assert lam(q, 4 * 3_000_000) / lam(q, 3_000_000) == pytest.approx(4 ** 0.35, rel=1e-12)
    seq = [lam(q, 8_415_000) for q in range(1, 100)]
    assert all(a > b for a, b in zip(seq, seq[1:]))


CALIB_CASES = [("doc", lambda q: synth.document_page(200, 160, seed=q)),
               ("gradient", lambda q: synth.gradient_page(200, 160, seed=q)),
               ("flat", lambda q: np.clip(128 + np.random.default_rng(q).normal(0, 1.5, (160, 200)), 0, 255)
                .astype(np.uint8))]
def _max_hull_slope(enc):
    """Steepest finite positive hull slope over the component's blocks,
    recomputed from the raw curves."""
    slopes = []
    for i in range(enc.nblocks):
        rb, rd, hp = enc.rd(i)
        for a, b in zip(hp, hp[1:]):
            db, dd = float(rb[b]) - float(rb[a]), rd[b] - rd[a]
            s = math.inf if db == 0 and dd > 0 else (0.0 if db == 0 else dd / db)
            if math.isfinite(s) and s > 0:
                slopes.append(s)
    return max(slopes)


@pytest.mark.parametrize("kind,make", CALIB_CASES, ids=[c[0] for c in CALIB_CASES])
@pytest.mark.parametrize("q", [20, 60, 85, 95])
def test_lambda_calibration_closed_form(ora, q, kind, make):
    """calibrate_lambda: raw x min(1, 1.2 max_slope / lambda(q=1))
    (pcrd.rs:1488-1514), with max_slope the steepest finite positive hull
    slope of the component, recomputed here from the raw curves."""
    px = make(q)
    data, enc = ora.encode(px, q, keep=True)
    pixels = 200 * 160
    mx = _max_hull_slope(enc)
    assert enc.max_slope(0) == pytest.approx(mx, rel=1e-12)
    raw = ora.L.ora_quality_to_lambda(q, pixels)
    floor = ora.L.ora_quality_to_lambda(1, pixels)
    want = raw * min(1.0, 1.2 * mx / floor)
    assert enc.lam(0) == pytest.approx(want, rel=1e-12)


def test_lambda_calibration_was_exercised(ora):
    """The cases of the closed-form test include images whose steepest
    slope sits below lambda(q=1), where the scaling is active."""
    active = []
    for kind, make in CALIB_CASES:
        for q in (20, 60, 85, 95):
            _, enc = ora.encode(make(q), q, keep=True)
            if 1.2 * _max_hull_slope(enc) < ora.L.ora_quality_to_lambda(1, 200 * 160):
                active.append((kind, q))
    assert len(active) >= 4, active


# ---- brute-force Lagrangian selection -----------------------------------------------
LAGRANGE_CASES = [
    ("doc q85", lambda: synth.document_page(160, 120, seed=21), 85),
    ("doc q40", lambda: synth.document_page(130, 97, seed=22), 40),
    ("doc q20 (32x32 blocks)", lambda: synth.document_page(150, 110, seed=23), 20),
    ("photo q75", lambda: synth.photo_page(96, 80, seed=24), 75),
    ("photo q95", lambda: synth.photo_page(64, 72, seed=25), 95),
    ("gradient q60", lambda: synth.gradient_page(128, 128, seed=26), 60),
]


@pytest.mark.parametrize("name,make,q", LAGRANGE_CASES, ids=[c[0] for c in LAGRANGE_CASES])
def test_kept_point_maximises_lagrangian(ora, name, make, q):
    px = make()
    data, enc = ora.encode(px, q, keep=True)
    ncomp = 1 if px.ndim == 2 else 3
    words = np.zeros(5 + 2 * 19, np.int32)
    ora.L.ora_enc_plan(enc.h, words.ctypes.data)
    levels = int(words[0])
    steps = _steps(words, int(words[2]))
    checked = 0
    for i in range(enc.nblocks):
        info = enc.info(i)
        comp, kept, full = int(info[0]), int(info[10]), int(info[11])
        rb, rd, hp = enc.rd(i)
        assert len(rb) == full + 1 and rb[0] == 0 and rd[0] == 0.0
        # rate: cumulative pass lengths plus 2 bytes of header overhead
        # (rate.rs:31, :168-170)
        ps = enc.passes(i)
        assert rb[1:].tolist() == (np.cumsum(ps[:, 4]) + 2).tolist()
        # distortion: each pass's delta is the model's value at unit step and
        # weight times step^2 x norm^2 x component weight (rate.rs:43-180,
        # :232-252, backend.rs:696-718)
        res, band, cls = int(info[1]), int(info[2]), int(info[9])
        bi = 0 if res == 0 else 1 + 3 * (res - 1) + (band - 1)
        step = steps[bi]
        level = (levels - res) if band != LL else levels
        norm = ora.L.ora_get_norm_97(max(level, 0), band)
        cw = 1.0 if comp == 0 else 1.20 + 1.80 * (1.0 - min(q, 99) / 99.0) ** 0.70
        for k in range(full):
            kind, bp, ns, sb = (int(v) for v in ps[k, :4])
            unit = ora.L.ora_pass_delta_pass_kind_aware(kind, bp, ns, sb if kind == MR else 0, 1.0, 1.0, band, q, cls)
            dd = rd[k + 1] - rd[k]
            assert dd == pytest.approx(unit * step * step * norm * norm * cw, rel=1e-9, abs=1e-300), (i, k)
        lam = enc.lam(comp)
        obj = rd - lam * rb.astype(np.float64)
        best = obj.max()
        tol = 1e-9 * max(1.0, abs(best), lam * float(rb[-1]))
        # the kept truncation point attains the maximum over every raw point
        assert obj[kept] >= best - tol, (i, kept, int(np.argmax(obj)))
        # and is the deepest maximiser on the hull
        on_hull = [k for k in hp if obj[k] >= best - tol]
        assert kept == max(on_hull), (i, kept, on_hull)
        # every raw point lies on or below the hull (upper concave envelope)
        for k in range(len(rb)):
            j = int(np.searchsorted(hp, k))
            if j < len(hp) and hp[j] == k:
                continue
            a, b = hp[j - 1], hp[j]
            x0, x1 = float(rb[a]), float(rb[b])
            y = rd[a] if x1 == x0 else rd[a] + (rd[b] - rd[a]) * (float(rb[k]) - x0) / (x1 - x0)
            assert rd[k] <= y + 1e-9 * max(1.0, abs(y)), (i, k)
        checked += 1
    assert checked == enc.nblocks and ncomp in (1, 3)


def test_tiny_image_hulls_have_origin_and_decreasing_slopes(ora):
    """rate.rs:286-324: the 4x4 ramp at DocumentHigh (q85), J2K: every
    block's hull starts at the origin with strictly decreasing slopes."""
    px = np.arange(0, 256, 16, dtype=np.uint8).reshape(4, 4)
    data, enc = ora.encode(px, 85, fmt=1, keep=True)
    assert enc.nblocks > 0
    for i in range(enc.nblocks):
        rb, rd, hp = enc.rd(i)
        assert hp[0] == 0 and rb[hp[0]] == 0
        slopes = []
        for a, b in zip(hp, hp[1:]):
            db, dd = float(rb[b]) - float(rb[a]), rd[b] - rd[a]
            slopes.append(math.inf if db == 0 and dd > 0 else (0.0 if db == 0 else dd / db))
        for s0, s1 in zip(slopes, slopes[1:]):
            assert s1 < s0


# ---- step scaler (plan/derive.rs:60-107) --------------------------------------------
@pytest.mark.parametrize("q,want,rel", [(42, 1.32, 0.005), (30, 2.0, 1e-15), (10, 4.0, 1e-15), (1, 5.5, 0.01),
                                        (75, 0.42, 0.005), (90, 0.25, 1e-15), (99, 0.031, 0.01), (50, 1.0, 0)])
def test_step_scaler_documented_calibration(ora, q, want, rel):
    assert ora.L.ora_quality_step_scaler(q) == pytest.approx(want, rel=rel, abs=0)


def _steps(plan_words, nbands):
    """Decode the written (exponent, mantissa) pairs into step sizes
    (1 + mu/2048) * 2^(numbps - eps), numbps = 8 + band gain."""
    out = []
    for b in range(nbands):
        e, m = plan_words[5 + 2 * b], plan_words[6 + 2 * b]
        band = 0 if b == 0 else (b - 1) % 3 + 1
        gain = 0 if band == 0 else (2 if band == 3 else 1)
        out.append((1 + m / 2048.0) * 2.0 ** (8 + gain - e))
    return out


def _plan(ora, px, q):
    _, enc = ora.encode(px, q, keep=True)
    words = np.zeros(5 + 2 * 19, np.int32)
    ora.L.ora_enc_plan(enc.h, words.ctypes.data)
    return words


@pytest.mark.parametrize("q", [1, 10, 30, 42, 60, 75, 85, 90, 95, 99])
def test_written_steps_scale_by_the_scaler(ora, q):
    """Each band's written step at quality q is its q=50 (unscaled, ISO
    derived) step times quality_step_scaler(q), up to the 11-bit mantissa's
    rounding (relative 2^-12), with neither clamp reached at these sizes."""
    px = synth.gradient_page(512, 512, seed=3)
    w50, wq = _plan(ora, px, 50), _plan(ora, px, q)
    nb = int(w50[2])
    s50, sq = _steps(w50, nb), _steps(wq, nb)
    scale = ora.L.ora_quality_step_scaler(q)
    for a, b in zip(s50, sq):
        assert b / a == pytest.approx(scale, rel=2.0 ** -12 + 1e-15)


def test_scale_one_quant_edges(ora):
    from ctypes import byref, c_uint8, c_uint16

    def scale(e, m, s):
        eo, mo = c_uint8(), c_uint16()
        ora.L.ora_scale_one_quant(e, m, s, byref(eo), byref(mo))
        return eo.value, mo.value

    assert scale(10, 0, 1.0) == (10, 0)
    assert scale(10, 0, 2.0) == (9, 0)            # doubling: one less exponent
    assert scale(10, 0, 0.5) == (11, 0)
    assert scale(10, 1024, 1.0) == (10, 1024)     # 1.5 stays 1.5
    assert scale(10, 1024, 2.0) == (9, 1024)
    assert scale(10, 0, 1.5) == (10, 1024)
    assert scale(1, 0, 8.0) == (0, 0)             # exponent clamps at 0
    assert scale(20, 0, 1e-9) == (33, 0)          # scaled step floors at 2^-13
