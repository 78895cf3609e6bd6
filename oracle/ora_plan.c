/* Oracle (test infrastructure only): encoding plan and step sizes.
 * Restates plan/mod.rs:116-177, plan/derive.rs, plan/validate.rs:4-62 and
 * dwt/norms.rs.  Host f64 arithmetic uses glibc pow/log2/floor/round like
 * Rust's f64::powf/log2/floor/round on Linux. */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "ora.h"

/* Synthesis norms of the 9/7 filter bank (dwt/norms.rs:24-37).  This is the
 * OpenJPEG/ISO table the step sizes are derived from; it has to be the same
 * data for the QCD bytes to match. */
static const double NORMS_97[4][10] = {
    {1.000, 1.965, 4.177, 8.403, 16.90, 33.84, 67.69, 135.3, 270.6, 540.9},
    {2.022, 3.989, 8.355, 17.04, 34.27, 68.63, 137.3, 274.6, 549.0, 0.0},
    {2.022, 3.989, 8.355, 17.04, 34.27, 68.63, 137.3, 274.6, 549.0, 0.0},
    {2.080, 3.865, 8.307, 17.18, 34.71, 69.59, 139.3, 278.6, 557.2, 0.0},
};

static int band_gain(int band) { return band == ORA_LL ? 0 : (band == ORA_HH ? 2 : 1); }

/* dwt/norms.rs:49-57 */
double ora_get_norm_97(uint32_t level, int band) {
    uint32_t cl = band == ORA_LL ? (level < 9 ? level : 9) : (level < 8 ? level : 8);
    return NORMS_97[band][cl];
}

static int floor_log2_i32(int32_t v) {
    if (v < 1) v = 1;
    return 31 - __builtin_clz((uint32_t)v);
}

/* dwt/norms.rs:84-90 */
void ora_encode_stepsize(int32_t stepsize, uint32_t numbps, uint8_t *exp, uint16_t *mant) {
    int p = floor_log2_i32(stepsize) - 13;
    int n = 11 - floor_log2_i32(stepsize);
    int32_t m = (n < 0 ? (stepsize >> -n) : (stepsize << n)) & 0x7ff;
    int e = (int)numbps - p;
    if (e < 0) e = 0;
    *exp = (uint8_t)e;
    *mant = (uint16_t)m;
}

/* dwt/norms.rs:68-82 */
void ora_irreversible_expounded_quant(uint32_t prec, uint8_t nres, uint8_t res, int band,
                                      uint8_t *exp, uint16_t *mant) {
    uint32_t gain = (uint32_t)band_gain(band);
    int lv = (int)nres - 1;
    if (lv < 0) lv = 0;
    lv -= res;
    if (lv < 0) lv = 0;
    double norm = ora_get_norm_97((uint32_t)lv, band);
    double stepsize = (double)(1u << gain) / norm;
    ora_encode_stepsize((int32_t)floor(stepsize * 8192.0), prec + gain, exp, mant);
}

/* plan/derive.rs:128-139 */
uint32_t ora_max_decompositions(uint32_t w, uint32_t h) {
    uint32_t m = w < h ? w : h;
    if (m <= 1) return 0;
    if (m < 32) return 31 - __builtin_clz(m);
    uint32_t ratio = m / 32;
    uint32_t lv = 1 + (31 - __builtin_clz(ratio));
    return lv < 6 ? lv : 6;
}

/* plan/derive.rs:74-86 */
double ora_quality_step_scaler(int q) {
    if (q == 0) return 8.0;
    if (q <= 50) return pow(2.0, (50.0 - (double)q) / 20.0);
    if (q <= 90) return pow(2.0, -(((double)q - 50.0) / 20.0));
    double t = ((double)q - 90.0) / 9.0;
    return 0.25 * pow(2.0, -3.0 * t);
}

/* plan/derive.rs:94-107 */
static void scale_one_quant(ora_sq *sq, double scale) {
    double base = 1.0 + (double)sq->mantissa / 2048.0;
    double scaled = base * scale;
    if (scaled < 1.0 / 8192.0) scaled = 1.0 / 8192.0;
    int k = (int)floor(log2(scaled));
    double base_new = scaled / ldexp(1.0, k);
    int e = (int)sq->exponent - k;
    if (e < 0) e = 0;
    double m = round((base_new - 1.0) * 2048.0);
    if (m < 0.0) m = 0.0;
    if (m > 2047.0) m = 2047.0;
    sq->exponent = (uint8_t)e;
    sq->mantissa = (uint16_t)m;
}

/* scale_one_quant on one (exponent, mantissa) pair, for the tests */
void ora_scale_one_quant(uint8_t exponent, uint16_t mantissa, double scale, uint8_t *exp_out, uint16_t *mant_out) {
    ora_sq sq = {0, 0, exponent, mantissa};
    scale_one_quant(&sq, scale);
    *exp_out = sq.exponent;
    *mant_out = sq.mantissa;
}

/* plan/validate.rs:4-62 plus model.rs colour-space component counts */
static int validate(const int32_t *const *planes, uint32_t w, uint32_t h, int ncomp, int cs, char *err) {
    if (w == 0 || h == 0) {
        snprintf(err, 256, "image dimensions must be non-zero");
        return ORA_INVALID_INPUT;
    }
    int want = cs == ORA_GRAY ? 1 : 3;
    if (ncomp != want) {
        snprintf(err, 256, "images must have exactly %d component(s)", want);
        return ORA_INVALID_INPUT;
    }
    size_t n = (size_t)w * h;
    for (int c = 0; c < ncomp; c++) {
        for (size_t i = 0; i < n; i++) {
            if (planes[c][i] < 0 || planes[c][i] > 255) {
                snprintf(err, 256, "component %d contains samples outside 0..=255", c);
                return ORA_INVALID_INPUT;
            }
        }
    }
    return ORA_OK;
}

/* plan/mod.rs:116-177 with derive.rs:23-53, :109-165 */
int ora_build_plan(const int32_t *const *planes, uint32_t w, uint32_t h, int ncomp,
                   int colorspace, int quality, int format, ora_plan *p, char *err) {
    int rc = validate(planes, w, h, ncomp, colorspace, err);
    if (rc) return rc;
    memset(p, 0, sizeof(*p));
    p->width = w;
    p->height = h;
    p->ncomp = ncomp;
    p->domain = colorspace == ORA_GRAY ? ORA_GRAY : ORA_SRGB;
    p->quality = quality;
    p->format = format;
    p->lossless = quality >= 100;
    uint32_t cap = p->domain == ORA_GRAY ? 5 : 6;
    uint32_t md = ora_max_decompositions(w, h);
    p->levels = (int)(md < cap ? md : cap);
    p->use_mct = p->domain == ORA_SRGB;
    p->cbw = p->cbh = quality < 30 ? 32 : 64;
    p->guard = 2;
    uint32_t prec = 8;
    uint8_t nres = (uint8_t)(p->levels + 1);
    int nb = 0;
    for (int r = 0; r <= p->levels; r++) {
        int first = r == 0 ? ORA_LL : ORA_HL, last = r == 0 ? ORA_LL : ORA_HH;
        for (int b = first; b <= last; b++) {
            ora_sq *sq = &p->sq[nb++];
            sq->resolution = (uint8_t)r;
            sq->band = (uint8_t)b;
            if (p->lossless) {
                sq->exponent = (uint8_t)(prec + band_gain(b)); /* norms.rs:63-66 */
                sq->mantissa = 0;
            } else {
                ora_irreversible_expounded_quant(prec, nres, (uint8_t)r, b, &sq->exponent, &sq->mantissa);
            }
        }
    }
    p->nbands = nb;
    if (!p->lossless) {
        double s = ora_quality_step_scaler(quality);
        for (int i = 0; i < nb; i++) scale_one_quant(&p->sq[i], s);
    }
    return ORA_OK;
}

/* encode/backend/native/backend.rs:602-620 and decode/reconstruct.rs:251-256:
 * (1 + m/2048) * 2^(prec+gain-exp) in f32, floored at 1e-6. */
float ora_quant_step_f32(uint32_t precision, int band, uint8_t exponent, uint16_t mantissa) {
    int numbps = (int)precision + band_gain(band);
    float base = 1.0f + (float)mantissa / 2048.0f;
    float v = base * ldexpf(1.0f, numbps - (int)exponent);
    return v > 1e-6f ? v : 1e-6f;
}
