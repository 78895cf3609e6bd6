/* Oracle (test infrastructure only): colour transforms, 5/3 and 9/7 DWT,
 * quantisation.  Restates simd/scalar.rs:141-343, dwt/rev53.rs and
 * dwt/irrev97.rs.  All f32 arithmetic is evaluated left to right with no FMA
 * (build flag -ffp-contract=off), matching the reference's scalar and `wide`
 * kernels (simd/wide.rs:372-438 asserts they are bit-equal). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

/* irrev97.rs:10-15 */
static const float ALPHA = -1.586134342059924f;
static const float BETA = -0.052980118572961f;
static const float GAMMA = 0.882911075530934f;
static const float DELTA = 0.443506852043971f;
static const float K = 1.230174104914001f;
#define INV_K (1.0f / K)
/* irrev97.rs:16 PARALLEL_ROW_THRESHOLD */
#define ROW_PARALLEL_THRESHOLD ((size_t)2 * 1024 * 1024)

/* scalar.rs:242-247 */
void ora_level_shift_f32(const int32_t *in, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) out[i] = (float)(in[i] - 128);
}
/* scalar.rs:249-254 */
void ora_level_shift_i32(const int32_t *in, size_t n, int32_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = in[i] - 128;
}
/* scalar.rs:256-278 (wide.rs:232-256 has the same association) */
void ora_forward_ict(const int32_t *r, const int32_t *g, const int32_t *b, size_t n, int comp, float *out) {
    for (size_t i = 0; i < n; i++) {
        float rf = (float)(r[i] - 128), gf = (float)(g[i] - 128), bf = (float)(b[i] - 128);
        float v;
        if (comp == 0) v = (0.299f * rf + 0.587f * gf) + 0.114f * bf;
        else if (comp == 1) v = (-0.16875f * rf + -0.33126f * gf) + 0.5f * bf;
        else v = (0.5f * rf - 0.41869f * gf) - 0.08131f * bf;
        out[i] = v;
    }
}
/* scalar.rs:280-302 */
void ora_forward_rct(const int32_t *r, const int32_t *g, const int32_t *b, size_t n, int comp, int32_t *out) {
    for (size_t i = 0; i < n; i++) {
        int32_t rv = r[i] - 128, gv = g[i] - 128, bv = b[i] - 128;
        if (comp == 0) out[i] = (rv + 2 * gv + bv) >> 2;
        else if (comp == 1) out[i] = bv - gv;
        else out[i] = rv - gv;
    }
}
/* scalar.rs:304-315 */
void ora_inverse_ict(float *y, float *cb, float *cr, size_t n) {
    for (size_t i = 0; i < n; i++) {
        float yy = y[i], cbb = cb[i], crr = cr[i];
        y[i] = yy + 1.402f * crr;
        cb[i] = (yy - 0.34413f * cbb) - 0.71414f * crr;
        cr[i] = yy + 1.772f * cbb;
    }
}
/* scalar.rs:317-329 */
void ora_inverse_rct(int32_t *y, int32_t *db, int32_t *dr, size_t n) {
    for (size_t i = 0; i < n; i++) {
        int32_t yy = y[i], dbv = db[i], drv = dr[i];
        int32_t g = yy - ((dbv + drv) >> 2);
        y[i] = drv + g;
        db[i] = g;
        dr[i] = dbv + g;
    }
}
/* scalar.rs:331-343: f32::round is half away from zero (C roundf) */
void ora_finalize_f32(const float *in, size_t n, int32_t *out) {
    for (size_t i = 0; i < n; i++) {
        float v = roundf(in[i] + 128.0f);
        if (v < 0.0f) v = 0.0f;
        if (v > 255.0f) v = 255.0f;
        out[i] = (int32_t)v;
    }
}
void ora_finalize_i32(const int32_t *in, size_t n, int32_t *out) {
    for (size_t i = 0; i < n; i++) {
        int32_t v = in[i] + 128;
        out[i] = v < 0 ? 0 : (v > 255 ? 255 : v);
    }
}

/* irrev97.rs:580-593 / rev53 encode_resolutions: resolution ladder,
 * rw[0..levels] from coarsest (index 0) to full size (index levels). */
void ora_resolutions(size_t w, size_t h, int levels, size_t *rw, size_t *rh) {
    rw[levels] = w;
    rh[levels] = h;
    for (int l = levels - 1; l >= 0; l--) {
        rw[l] = (rw[l + 1] + 1) / 2;
        rh[l] = (rh[l + 1] + 1) / 2;
    }
}

/* irrev97.rs:589-605 sym_index: whole-sample symmetric extension */
static size_t sym(size_t n, long i) {
    if (n <= 1) return 0;
    if (i >= 0 && i < (long)n) return (size_t)i;
    long period = 2 * ((long)n - 1);
    long k = i % period;
    if (k < 0) k += period;
    return (size_t)(k >= (long)n ? period - k : k);
}

/* ---- 5/3 (rev53.rs) ------------------------------------------------------ */

/* One forward 5/3 lift on a strided signal of n samples (rev53.rs:143-265):
 * predict odd x[2i+1] -= (x[2i] + x[2i+2]) >> 1, update even
 * x[2i] += (x[2i-1] + x[2i+1] + 2) >> 2 with symmetric extension, then
 * deinterleave to [low | high].  n == 1 is left unchanged (even origin). */
static void fwd53_1d(int32_t *x, size_t n, size_t stride, int32_t *tmp) {
    if (n <= 1) return;
    for (size_t i = 0; i < n; i++) tmp[i] = x[i * stride];
    for (size_t j = 1; j < n; j += 2) tmp[j] -= (tmp[sym(n, (long)j - 1)] + tmp[sym(n, (long)j + 1)]) >> 1;
    for (size_t j = 0; j < n; j += 2) tmp[j] += (tmp[sym(n, (long)j - 1)] + tmp[sym(n, (long)j + 1)] + 2) >> 2;
    size_t sn = (n + 1) / 2;
    for (size_t i = 0; i < n; i++) x[(i % 2 ? sn + i / 2 : i / 2) * stride] = tmp[i];
}

/* rev53.rs:272-417: interleave, undo update on even, undo predict on odd. */
static void inv53_1d(int32_t *x, size_t n, size_t stride, int32_t *tmp) {
    if (n <= 1) return;
    size_t sn = (n + 1) / 2;
    for (size_t i = 0; i < n; i++) tmp[i] = x[(i % 2 ? sn + i / 2 : i / 2) * stride];
    for (size_t j = 0; j < n; j += 2) tmp[j] -= (tmp[sym(n, (long)j - 1)] + tmp[sym(n, (long)j + 1)] + 2) >> 2;
    for (size_t j = 1; j < n; j += 2) tmp[j] += (tmp[sym(n, (long)j - 1)] + tmp[sym(n, (long)j + 1)]) >> 1;
    for (size_t i = 0; i < n; i++) x[i * stride] = tmp[i];
}

/* rev53.rs:35-75: per level finest first, vertical (height > 1) then every
 * row of the active region. */
void ora_fwd53(int32_t *d, size_t w, size_t h, int levels) {
    if (!w || !h || !levels) return;
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    int32_t *tmp = malloc(sizeof(int32_t) * (w > h ? w : h));
    for (int l = levels; l >= 1; l--) {
        size_t aw = rw[l], ah = rh[l];
        if (ah > 1)
            for (size_t x = 0; x < aw; x++) fwd53_1d(d + x, ah, w, tmp);
        for (size_t y = 0; y < ah; y++) fwd53_1d(d + y * w, aw, 1, tmp);
    }
    free(tmp);
}

/* rev53.rs:95-136: coarsest first, rows then vertical. */
void ora_inv53(int32_t *d, size_t w, size_t h, int levels) {
    if (!w || !h || !levels) return;
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    int32_t *tmp = malloc(sizeof(int32_t) * (w > h ? w : h));
    for (int l = 1; l <= levels; l++) {
        size_t aw = rw[l], ah = rh[l];
        for (size_t y = 0; y < ah; y++) inv53_1d(d + y * w, aw, 1, tmp);
        if (ah > 1)
            for (size_t x = 0; x < aw; x++) inv53_1d(d + x, ah, w, tmp);
    }
    free(tmp);
}

/* ---- 9/7 (irrev97.rs) ---------------------------------------------------- */

/* irrev97.rs:92-129 (rows) and :261-336 (columns): four lifts
 * x[j] += c * (x[j-1] + x[j+1]) over one parity, symmetric extension, then
 * even * (1/K), odd * K, deinterleave.  n == 1 scales by 1/K. */
#define DEF_FWD97(NAME, T, CA, CB, CG, CD, CK, CIK)                                   \
    static void NAME(T *x, size_t n, size_t stride, T *tmp) {                          \
        if (n < 2) {                                                                   \
            if (n == 1) x[0] = x[0] * CIK;                                             \
            return;                                                                    \
        }                                                                              \
        for (size_t i = 0; i < n; i++) tmp[i] = x[i * stride];                         \
        const T cs[4] = {CA, CB, CG, CD};                                              \
        for (int s = 0; s < 4; s++) {                                                  \
            for (size_t j = (s & 1) ? 0 : 1; j < n; j += 2) {                          \
                T l = tmp[sym(n, (long)j - 1)], r = tmp[sym(n, (long)j + 1)];          \
                T t = l + r;                                                           \
                t = cs[s] * t;                                                         \
                tmp[j] = tmp[j] + t;                                                   \
            }                                                                          \
        }                                                                              \
        for (size_t j = 0; j < n; j += 2) tmp[j] = tmp[j] * CIK;                       \
        for (size_t j = 1; j < n; j += 2) tmp[j] = tmp[j] * CK;                        \
        size_t sn = (n + 1) / 2;                                                       \
        for (size_t i = 0; i < n; i++) x[(i % 2 ? sn + i / 2 : i / 2) * stride] = tmp[i]; \
    }

DEF_FWD97(fwd97_1d, float, ALPHA, BETA, GAMMA, DELTA, K, INV_K)
DEF_FWD97(fwd97_1d_f64, double, -1.586134342059924, -0.052980118572961, 0.882911075530934,
          0.443506852043971, 1.230174104914001, 1.0 / 1.230174104914001)

/* irrev97.rs:661-693 (rows) and :707-757 (columns): interleave, even * K,
 * odd * (1/K), then lifts -delta (even), -gamma (odd), -beta (even),
 * -alpha (odd).  n == 1 scales by K. */
static void inv97_1d(float *x, size_t n, size_t stride, float *tmp) {
    if (n < 2) {
        if (n == 1) x[0] = x[0] * K;
        return;
    }
    size_t sn = (n + 1) / 2;
    for (size_t i = 0; i < n; i++) tmp[i] = x[(i % 2 ? sn + i / 2 : i / 2) * stride];
    for (size_t j = 0; j < n; j += 2) tmp[j] = tmp[j] * K;
    for (size_t j = 1; j < n; j += 2) tmp[j] = tmp[j] * INV_K;
    const float cs[4] = {-DELTA, -GAMMA, -BETA, -ALPHA};
    for (int s = 0; s < 4; s++) {
        for (size_t j = (s & 1) ? 1 : 0; j < n; j += 2) {
            float l = tmp[sym(n, (long)j - 1)], r = tmp[sym(n, (long)j + 1)];
            float t = l + r;
            t = cs[s] * t;
            tmp[j] = tmp[j] + t;
        }
    }
    for (size_t i = 0; i < n; i++) x[i * stride] = tmp[i];
}

/* irrev97.rs:51-82: per level finest first, vertical then rows. */
void ora_fwd97(float *d, size_t w, size_t h, int levels) {
    if (!w || !h || !levels) return;
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    float *tmp = malloc(sizeof(float) * (w > h ? w : h));
    for (int l = levels; l >= 1; l--) {
        size_t aw = rw[l], ah = rh[l];
        for (size_t x = 0; x < aw; x++) fwd97_1d(d + x, ah, w, tmp);
        if (aw * ah >= ROW_PARALLEL_THRESHOLD) {
            /* irrev97.rs:139-145: rows run in parallel for large levels */
#pragma omp parallel
            {
                float *t2 = malloc(sizeof(float) * aw);
#pragma omp for schedule(static)
                for (size_t y = 0; y < ah; y++) fwd97_1d(d + y * w, aw, 1, t2);
                free(t2);
            }
        } else {
            for (size_t y = 0; y < ah; y++) fwd97_1d(d + y * w, aw, 1, tmp);
        }
    }
    free(tmp);
}

/* FP64 lifting oracle with the same steps (error reporting only). */
void ora_fwd97_f64(double *d, size_t w, size_t h, int levels) {
    if (!w || !h || !levels) return;
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    double *tmp = malloc(sizeof(double) * (w > h ? w : h));
    for (int l = levels; l >= 1; l--) {
        size_t aw = rw[l], ah = rh[l];
        for (size_t x = 0; x < aw; x++) fwd97_1d_f64(d + x, ah, w, tmp);
        for (size_t y = 0; y < ah; y++) fwd97_1d_f64(d + y * w, aw, 1, tmp);
    }
    free(tmp);
}

/* irrev97.rs:636-656: coarsest first, rows then vertical. */
void ora_inv97(float *d, size_t w, size_t h, int levels) {
    if (!w || !h || !levels) return;
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    float *tmp = malloc(sizeof(float) * (w > h ? w : h));
    for (int l = 1; l <= levels; l++) {
        size_t aw = rw[l], ah = rh[l];
        if (aw * ah >= ROW_PARALLEL_THRESHOLD) {
#pragma omp parallel
            {
                float *t2 = malloc(sizeof(float) * aw);
#pragma omp for schedule(static)
                for (size_t y = 0; y < ah; y++) inv97_1d(d + y * w, aw, 1, t2);
                free(t2);
            }
        } else {
            for (size_t y = 0; y < ah; y++) inv97_1d(d + y * w, aw, 1, tmp);
        }
        for (size_t x = 0; x < aw; x++) inv97_1d(d + x, ah, w, tmp);
    }
    free(tmp);
}

/* scalar.rs:207-213: trunc(v / step), non-finite or step <= 0 -> 0 */
int32_t ora_quantize(float v, float step) {
    if (step <= 0.0f || !isfinite(v)) return 0;
    float q = truncf(v / step);
    if (q >= 2147483648.0f) return INT32_MAX;
    if (q < -2147483648.0f) return INT32_MIN;
    return (int32_t)q;
}

/* scalar.rs:233-240 / wide.rs:141-152: (sign * (|q| + 0.5)) * step */
float ora_dequantize(int32_t q, float step) {
    if (q == 0) return 0.0f;
    float sign = q >= 0 ? 1.0f : -1.0f;
    float mag = (float)(q < 0 ? -(int64_t)q : (int64_t)q);
    return (sign * (mag + 0.5f)) * step;
}

/* backend.rs:519-587: quantise every band rect of the Mallat plane. */
void ora_quantize_plane(const float *d, size_t w, size_t h, int levels, const ora_plan *plan, int32_t *out) {
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    for (int b = 0; b < plan->nbands; b++) {
        const ora_sq *sq = &plan->sq[b];
        int r = sq->resolution;
        size_t x0, y0, x1, y1;
        if (r == 0) {
            x0 = 0; y0 = 0; x1 = rw[0]; y1 = rh[0];
        } else if (sq->band == ORA_HL) {
            x0 = rw[r - 1]; y0 = 0; x1 = rw[r]; y1 = rh[r - 1];
        } else if (sq->band == ORA_LH) {
            x0 = 0; y0 = rh[r - 1]; x1 = rw[r - 1]; y1 = rh[r];
        } else {
            x0 = rw[r - 1]; y0 = rh[r - 1]; x1 = rw[r]; y1 = rh[r];
        }
        float step = ora_quant_step_f32(8, sq->band, sq->exponent, sq->mantissa);
        for (size_t y = y0; y < y1; y++)
            for (size_t x = x0; x < x1; x++) out[y * w + x] = ora_quantize(d[y * w + x], step);
    }
}
