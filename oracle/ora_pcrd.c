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

typedef struct {
    int passes;
    uint32_t bytes;
    double dist;
    double slope;
} pt;

/* pcrd.rs:541-554 */
static double rd_slope(const pt *a, const pt *b) {
    double db = (double)(b->bytes - a->bytes);
    double dd = b->dist - a->dist;
    if (db == 0.0) return dd > 0.0 ? INFINITY : 0.0;
    return dd / db;
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

/* pcrd.rs:669-696 */
static double pass_delta(int kind, int bitplane, int newsig, int refine, double step,
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
static double quality_to_lambda(int q, uint32_t pixels) {
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

/* Build the pruned hull for one block (rate.rs:128-180, pcrd.rs:151-252).
 * Returns the number of hull points in h[]. */
static int build_hull(const ora_block *b, const ora_plan *p, double cw, pt *h) {
    const ora_sq *sq = NULL;
    for (int i = 0; i < p->nbands; i++)
        if (p->sq[i].resolution == b->res && p->sq[i].band == b->band) sq = &p->sq[i];
    double step = sq ? quant_step_f64(sq) : 1.0;
    double sw = subband_weight(p->levels + 1, b->res, b->band);
    pt raw[ORA_MAX_PASSES + 1];
    raw[0] = (pt){0, 0, 0.0, INFINITY};
    double cum = 0.0;
    for (int i = 0; i < b->npasses; i++) {
        const ora_pass *ps = &b->passes[i];
        int refine = ps->kind == ORA_MR ? ps->significant_before : 0;
        double dd = pass_delta(ps->kind, ps->bitplane, ps->newly_significant, refine, step, sw,
                               b->band, p->quality, b->cls) * cw;
        cum += dd;
        raw[i + 1] = (pt){i + 1, (uint32_t)ps->cumulative_length + 2u, cum, 0.0};
    }
    int n = b->npasses + 1;
    for (int i = 1; i < n; i++) raw[i].slope = rd_slope(&raw[i - 1], &raw[i]);
    int m = 0;
    for (int i = 0; i < n; i++) {
        h[m++] = raw[i];
        while (m >= 3) {
            double sab = rd_slope(&h[m - 3], &h[m - 2]);
            double sbc = rd_slope(&h[m - 2], &h[m - 1]);
            if (sbc >= sab) {
                h[m - 2] = h[m - 1];
                m--;
            } else {
                break;
            }
        }
    }
    h[0].slope = INFINITY;
    for (int i = 1; i < m; i++) h[i].slope = rd_slope(&h[i - 1], &h[i]);
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
        double raw = quality_to_lambda(p->quality, pixels);
        double lambda = raw;
        if (mx > 0.0) {
            double floor_l = quality_to_lambda(1, pixels);
            double target = mx * 1.2;
            if (target < floor_l) lambda = raw * (target / floor_l);
        }
        enc->lambda[c] = lambda;
        enc->max_slope[c] = mx;
        for (int i = 0; i < enc->nblocks; i++) {
            ora_block *b = &enc->blocks[i];
            if (b->comp != c) continue;
            pt *h = hulls + (size_t)i * (ORA_MAX_PASSES + 1);
            int keep = 0;
            for (int j = 1; j < hn[i]; j++) {
                if (h[j].slope >= lambda) keep = h[j].passes;
                else break;
            }
            truncate_block(b, keep);
        }
    }
    free(hulls);
    free(hn);
    return ORA_OK;
}
