/* Oracle (test infrastructure only): PCRD rate/distortion bookkeeping,
 * convex hull, quality-to-lambda and truncation.  Restates
 * encode/backend/native/rate.rs:43-252, dwt/pcrd.rs:151-252, :348-424,
 * :527-554, :669-802, :1425-1514, profile.rs:141-154 and
 * encode/backend/native/backend.rs:641-718, :834-897.  f64 arithmetic is
 * evaluated in the reference's order. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

typedef ora_rd_point pt;

/* pcrd.rs:541-554 (the NegativeByteDelta case cannot arise on validated
 * curves, whose bytes never decrease) */
static double rd_slope(const pt *a, const pt *b) {
    double db = (double)(b->bytes - a->bytes);
    double dd = b->dist - a->dist;
    if (db == 0.0) return dd > 0.0 ? INFINITY : 0.0;
    return dd / db;
}

/* pcrd.rs:527-539 */
static void fill_slopes(pt *p, int n) {
    if (n <= 0) return;
    p[0].slope = INFINITY;
    for (int i = 1; i < n; i++) p[i].slope = rd_slope(&p[i - 1], &p[i]);
}

/* pcrd.rs:151-202: origin, then one point per pass record, with the
 * record validation.  out holds n + 1 points. */
int ora_pcrd_raw_curve(const ora_raw_record *recs, int n, pt *out) {
    out[0] = (pt){0, 0, 0.0, INFINITY};
    uint32_t prev = 0;
    double cum = 0.0;
    for (int i = 0; i < n; i++) {
        const ora_raw_record *r = &recs[i];
        if (r->cumulative_bytes < prev) return ORA_PCRD_NONMONOTONE_CUMULATIVE_BYTES;
        if (r->bytes != r->cumulative_bytes - prev) return ORA_PCRD_INCONSISTENT_INCREMENTAL_BYTES;
        if (!isfinite(r->distortion_delta) || r->distortion_delta < 0.0) return ORA_PCRD_INVALID_DISTORTION_DELTA;
        cum += r->distortion_delta;
        out[i + 1] = (pt){r->pass_index + 1, r->cumulative_bytes, cum, 0.0};
        prev = r->cumulative_bytes;
    }
    fill_slopes(out, n + 1);
    return ORA_PCRD_OK;
}

/* pcrd.rs:209-243: push each point, then drop the middle of the last three
 * while its slopes do not strictly decrease. */
int ora_pcrd_prune_hull(const pt *raw, int n, pt *out, int *m_out) {
    if (n <= 0) return ORA_PCRD_EMPTY_CURVE;
    int m = 0;
    for (int i = 0; i < n; i++) {
        out[m++] = raw[i];
        while (m >= 3) {
            double sab = rd_slope(&out[m - 3], &out[m - 2]);
            double sbc = rd_slope(&out[m - 2], &out[m - 1]);
            if (sbc >= sab) {
                out[m - 2] = out[m - 1];
                m--;
            } else {
                break;
            }
        }
    }
    fill_slopes(out, m);
    *m_out = m;
    return ORA_PCRD_OK;
}

/* pcrd.rs:488-525 */
static int validate_curve(const pt *p, int n) {
    if (n <= 0) return ORA_PCRD_EMPTY_CURVE;
    if (p[0].passes != 0 || p[0].bytes != 0 || p[0].dist != 0.0) return ORA_PCRD_INVALID_ORIGIN;
    for (int i = 1; i < n; i++) {
        if (p[i].passes <= p[i - 1].passes) return ORA_PCRD_NONMONOTONE_PASSES;
        if (p[i].bytes < p[i - 1].bytes) return ORA_PCRD_NONMONOTONE_POINT_BYTES;
        if (p[i].dist < p[i - 1].dist) return ORA_PCRD_NONMONOTONE_DISTORTION;
    }
    return ORA_PCRD_OK;
}

/* pcrd.rs:406-424: the deepest point reached while slopes stay >= lambda */
int ora_pcrd_choose_point(const pt *p, int n, double lambda, int *index) {
    int rc = validate_curve(p, n);
    if (rc) return rc;
    int best = 0;
    for (int i = 1; i < n; i++) {
        if (p[i].slope >= lambda) best = i;
        else break;
    }
    *index = best;
    return ORA_PCRD_OK;
}

/* pcrd.rs:728-748 */
static double band_bias(int band, int q) {
    if (q < 50) {
        double t = (double)q / 50.0;
        if (band == ORA_LL) return 3.0 - 1.8 * t;
        if (band == ORA_HH) return 0.20 + 0.62 * t;
        return 1.0;
    }
    if (band == ORA_LL) return 1.20;
    if (band == ORA_HH) return 0.82;
    return 1.00;
}

/* profile.rs:141-154 (EdgeText/BackgroundNoise never produced) */
static double class_weight(int cls, int is_ll) {
    double base = cls == 0 ? 2.5 : (cls == 1 ? 1.5 : 1.0);
    return is_ll ? base * 1.1 : base;
}

/* pcrd.rs:764-802 with contrast_visibility_weight fixed at 1.0
 * (rate.rs:157-158): the blend is 1 + s * (1 - 1) = 1, clamped to [0.2, 1]. */
static double masking(double raw, int cls, int q) {
    if (raw <= 0.0) return 0.0;
    double qq = (double)q / 100.0;
    double low = 1.0 - qq;
    if (low < 0.0) low = 0.0;
    if (low > 1.0) low = 1.0;
    double cs = cls == 0 ? 0.10 : (cls == 1 ? 0.20 : 1.00);
    double eff = low * cs;
    double blended = 1.0 + eff * (1.0 - 1.0);
    if (blended < 0.20) blended = 0.20;
    if (blended > 1.0) blended = 1.0;
    return raw * blended;
}

/* pcrd.rs:652-658 */
double ora_pass_delta_baseline(int bitplane, int newsig, int refine, double step, double sw) {
    int bp = bitplane < 30 ? bitplane : 30;
    double pw = (double)(1ull << bp) * step;
    double energy = pw * pw;
    double sig = (double)newsig * energy;
    double ref = (double)refine * energy * 0.25;
    return sw * (sig + ref);
}

/* pcrd.rs:669-696 */
double ora_pass_delta_pass_kind_aware(int kind, int bitplane, int newsig, int refine, double step,
                                      double sw, int band, int q, int cls) {
    int bp = bitplane < 30 ? bitplane : 30;
    double pw = (double)(1ull << bp) * step;
    double energy = pw * pw;
    double contrib;
    if (kind == ORA_SP) contrib = (double)newsig * energy;
    else if (kind == ORA_MR) contrib = (double)refine * energy * 0.25;
    else contrib = (double)newsig * energy * 0.85;
    double raw = sw * band_bias(band, q) * class_weight(cls, band == ORA_LL) * contrib;
    return masking(raw, cls, q);
}

/* pcrd.rs:1425-1456 */
double ora_quality_to_lambda(int q, uint32_t pixels) {
    if (q >= 100) return 0.0;
    if (q == 0) return 1.7976931348623157e308;
    double t = ((double)q - 1.0) / 98.0;
    double log_lambda = 5.68 - 2.618 * pow(t, 0.85);
    double base = pow(10.0, log_lambda);
    if (base < 1e-3) base = 1e-3;
    double tail = 1.0;
    if (q >= 90) {
        double tt = (99.0 - (double)(q < 99 ? q : 99)) / 9.0;
        tail = tt * tt;
    }
    double rf = pow((double)pixels / 12000000.0, 0.35);
    return base * rf * tail;
}

/* rate.rs:248-252 */
static double quant_step_f64(const ora_sq *sq) {
    int gain = sq->band == ORA_LL ? 0 : (sq->band == ORA_HH ? 2 : 1);
    int numbps = 8 + gain;
    double mf = (double)sq->mantissa / 2048.0;
    return (1.0 + mf) * ldexp(1.0, numbps - (int)sq->exponent);
}

/* rate.rs:232-241 */
static double subband_weight(int nres, int res, int band) {
    int lv = nres - 1;
    if (lv < 0) lv = 0;
    if (band != ORA_LL) {
        lv -= res;
        if (lv < 0) lv = 0;
    }
    double n = ora_get_norm_97((uint32_t)lv, band);
    return n * n;
}

/* backend.rs:696-718 */
static double component_weight(const ora_plan *p, int c) {
    if (!p->use_mct || p->domain != ORA_SRGB) return 1.0;
    if (c == 0) return 1.0;
    double q = (double)(p->quality < 99 ? p->quality : 99);
    return 1.20 + 1.80 * pow(1.0 - q / 99.0, 0.70);
}

/* pcrd.rs:1493-1514: scale the quality's lambda down when even q=1's lambda
 * lies above 1.2 x the steepest finite positive slope. */
double ora_calibrate_lambda(double max_slope, double raw_lambda, int quality, uint32_t pixels) {
    (void)quality;
    if (max_slope <= 0.0) return raw_lambda;
    double floor_l = ora_quality_to_lambda(1, pixels);
    double target = max_slope * 1.2;
    if (target >= floor_l) return raw_lambda;
    return raw_lambda * (target / floor_l);
}

/* Raw pass records of one block (rate.rs:128-180), its raw curve and the
 * pruned hull (pcrd.rs:246-252).  The raw curve is kept in the block for the
 * tests.  Returns the number of hull points in h[]. */
static int build_hull(ora_block *b, const ora_plan *p, double cw, pt *h) {
    const ora_sq *sq = NULL;
    for (int i = 0; i < p->nbands; i++)
        if (p->sq[i].resolution == b->res && p->sq[i].band == b->band) sq = &p->sq[i];
    double step = sq ? quant_step_f64(sq) : 1.0;
    double sw = subband_weight(p->levels + 1, b->res, b->band);
    ora_raw_record recs[ORA_MAX_PASSES];
    uint32_t prev = 0;
    for (int i = 0; i < b->npasses; i++) {
        const ora_pass *ps = &b->passes[i];
        int refine = ps->kind == ORA_MR ? ps->significant_before : 0;
        double dd = ora_pass_delta_pass_kind_aware(ps->kind, ps->bitplane, ps->newly_significant, refine, step, sw,
                                                   b->band, p->quality, b->cls) * cw;
        uint32_t cum = (uint32_t)ps->cumulative_length + 2u; /* HEADER_OVERHEAD_BYTES, rate.rs:31 */
        recs[i] = (ora_raw_record){i, cum - prev, cum, dd};
        prev = cum;
    }
    pt raw[ORA_MAX_PASSES + 1];
    int n = b->npasses + 1, m = 0;
    /* the encoder's records always validate; a failure here is a bug */
    if (ora_pcrd_raw_curve(recs, b->npasses, raw) != ORA_PCRD_OK) abort();
    ora_pcrd_prune_hull(raw, n, h, &m);
    b->raw_n = n;
    for (int i = 0; i < n; i++) {
        b->raw_bytes[i] = raw[i].bytes;
        b->raw_dist[i] = raw[i].dist;
    }
    b->hull_n = m;
    for (int i = 0; i < m; i++) b->hull_passes[i] = h[i].passes;
    return m;
}

/* backend.rs:834-897: keep n passes; the last kept pass carries the prefix
 * of the single terminated segment. */
static void truncate_block(ora_block *b, int keep) {
    if (keep > b->npasses) keep = b->npasses;
    if (keep == 0) {
        b->npasses = 0;
        b->body_len = 0;
        return;
    }
    int cum = 0;
    for (int i = 0; i < keep; i++) {
        cum += b->passes[i].length;
        b->passes[i].cumulative_length = cum;
        b->passes[i].is_terminated = i + 1 == keep;
    }
    b->npasses = keep;
    b->body_len = cum < b->nbytes ? cum : b->nbytes;
}

/* backend.rs:641-694 + pcrd.rs:1465-1514: per component, hulls, calibrated
 * lambda, deepest hull point with slope >= lambda, then truncation. */
int ora_pcrd_apply(ora_encoding *enc, const ora_plan *p) {
    for (int i = 0; i < enc->nblocks; i++) enc->blocks[i].body_len = enc->blocks[i].nbytes;
    if (p->quality >= 100) return ORA_OK;
    uint32_t pixels = p->width * p->height;
    pt *hulls = malloc(sizeof(pt) * (ORA_MAX_PASSES + 1) * (size_t)enc->nblocks);
    int *hn = malloc(sizeof(int) * (size_t)enc->nblocks);
    for (int c = 0; c < p->ncomp; c++) {
        double cw = component_weight(p, c);
        double mx = 0.0;
        for (int i = 0; i < enc->nblocks; i++) {
            ora_block *b = &enc->blocks[i];
            if (b->comp != c) continue;
            pt *h = hulls + (size_t)i * (ORA_MAX_PASSES + 1);
            hn[i] = build_hull(b, p, cw, h);
            for (int j = 1; j < hn[i]; j++)
                if (isfinite(h[j].slope) && h[j].slope > 0.0 && h[j].slope > mx) mx = h[j].slope;
        }
        double raw = ora_quality_to_lambda(p->quality, pixels);
        double lambda = ora_calibrate_lambda(mx, raw, p->quality, pixels);
        enc->lambda[c] = lambda;
        enc->max_slope[c] = mx;
        for (int i = 0; i < enc->nblocks; i++) {
            ora_block *b = &enc->blocks[i];
            if (b->comp != c) continue;
            pt *h = hulls + (size_t)i * (ORA_MAX_PASSES + 1);
            int k = 0;
            if (ora_pcrd_choose_point(h, hn[i], lambda, &k) != ORA_PCRD_OK) abort();
            int keep = h[k].passes;
            truncate_block(b, keep);
        }
    }
    free(hulls);
    free(hn);
    return ORA_OK;
}
