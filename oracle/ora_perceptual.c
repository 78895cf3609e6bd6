/* Oracle (test infrastructure only), timing-faithful CPU baseline: the work
 * the reference performs on every lossy encode and then discards.
 *
 *  - The full-image contrast-mask map (encode/backend/native/backend.rs:
 *    650-652, :724-769 -> perceptual/contrast_mask.rs:225-291): an 8x8 naive
 *    DCT per block (:78-101), its CSF-weighted energy (:105-118), the
 *    quadrant-variance edge factor (:120-172) and the 75th-percentile
 *    normalisation (:313-321).  rate.rs:60 ignores the map.
 *  - plan_coding_passes (encode/backend/native/t1.rs:1226-1322), called
 *    from analyze_codeblock (:234) and never read by the coder.
 *
 * Nothing here changes a byte of output; ora_encode runs it only in the
 * faithful mode (ora_set_faithful), so that the CPU figure includes it.
 * The DCT is evaluated as the reference writes it, cos() inside the loops;
 * this file is built with -fno-math-errno so that, as for Rust's cos, the
 * compiler may hoist the calls that do not depend on the inner index. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

/* contrast_mask.rs:48-57 (PSNR_HVS_M_CSF_WEIGHTS, Ponomarenko et al.) */
static const double CSF[8][8] = {
    {0.0000, 0.8264, 1.0000, 0.3906, 0.1736, 0.0625, 0.0384, 0.0269},
    {0.6944, 0.6944, 0.5102, 0.2770, 0.1479, 0.0297, 0.0278, 0.0331},
    {0.5102, 0.5917, 0.3906, 0.1736, 0.0625, 0.0308, 0.0210, 0.0319},
    {0.5102, 0.3460, 0.2066, 0.1189, 0.0384, 0.0132, 0.0156, 0.0260},
    {0.3086, 0.2066, 0.0730, 0.0319, 0.0216, 0.0084, 0.0094, 0.0169},
    {0.1736, 0.0816, 0.0331, 0.0244, 0.0152, 0.0092, 0.0078, 0.0118},
    {0.0416, 0.0244, 0.0164, 0.0132, 0.0094, 0.0068, 0.0069, 0.0098},
    {0.0193, 0.0118, 0.0111, 0.0104, 0.0080, 0.0100, 0.0094, 0.0102},
};

static int g_faithful;
void ora_set_faithful(int on) { g_faithful = on != 0; }
int ora_faithful(void) { return g_faithful; }

/* contrast_mask.rs:78-101 */
static void dct8x8(const double *block, double out[8][8]) {
    const double pi = 3.14159265358979323846;
    for (int u = 0; u < 8; u++)
        for (int v = 0; v < 8; v++) {
            double au = u == 0 ? 1.0 / sqrt(2.0) : 1.0;
            double av = v == 0 ? 1.0 / sqrt(2.0) : 1.0;
            double sum = 0.0;
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    double sx = block[y * 8 + x];
                    double cx = cos(((double)(2 * x + 1) * (double)u * pi) / 16.0);
                    double cy = cos(((double)(2 * y + 1) * (double)v * pi) / 16.0);
                    sum += sx * cx * cy;
                }
            out[u][v] = 0.25 * au * av * sum;
        }
}

/* contrast_mask.rs:120-133 */
static double variance(const double *v, int n) {
    double mean = 0.0;
    for (int i = 0; i < n; i++) mean += v[i];
    mean /= n;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += (v[i] - mean) * (v[i] - mean);
    return s / n;
}

/* contrast_mask.rs:141-172 */
static double quadrant_delta(const double *block, double eps) {
    double whole = variance(block, 64);
    if (whole <= eps) return 0.0;
    double q[4][16];
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) q[(y >= 4) * 2 + (x >= 4)][(y & 3) * 4 + (x & 3)] = block[y * 8 + x];
    double avg = (variance(q[0], 16) + variance(q[1], 16) + variance(q[2], 16) + variance(q[3], 16)) / 4.0;
    double d = avg / whole;
    return d < 0.0 ? 0.0 : d > 1.0 ? 1.0 : d;
}

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

/* backend.rs:724-769 + contrast_mask.rs:225-291: returns a checksum of the
 * visibility weights (the map itself is discarded, as by the reference). */
double ora_contrast_mask(const int32_t *const *planes, int ncomp, int colorspace, uint32_t w, uint32_t h) {
    size_t n = (size_t)w * h;
    uint8_t *luma = malloc(n ? n : 1);
    for (size_t i = 0; i < n; i++) {
        if (colorspace == ORA_GRAY || ncomp < 3) {
            int32_t v = planes[0][i];
            luma[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
        } else {
            double y = 0.299 * planes[0][i] + 0.587 * planes[1][i] + 0.114 * planes[2][i];
            y = y < 0.0 ? 0.0 : y > 255.0 ? 255.0 : y;
            luma[i] = (uint8_t)y;
        }
    }
    size_t bx = (w + 7) / 8, by = (h + 7) / 8, nb = bx * by;
    double *energy = malloc(sizeof(double) * (nb ? nb : 1));
    double *sorted = malloc(sizeof(double) * (nb ? nb : 1));
    double *mean_l = malloc(sizeof(double) * (nb ? nb : 1));
    for (size_t j = 0; j < by; j++)
        for (size_t i = 0; i < bx; i++) {
            double block[64], coef[8][8];
            double s = 0.0;
            for (int y = 0; y < 8; y++) {
                size_t sy = j * 8 + y < h ? j * 8 + y : h - 1;
                for (int x = 0; x < 8; x++) {
                    size_t sx = i * 8 + x < w ? i * 8 + x : w - 1;
                    block[y * 8 + x] = luma[sy * w + sx];
                    s += block[y * 8 + x];
                }
            }
            dct8x8(block, coef);
            double ew = 0.0;
            for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++) ew += coef[a][b] * coef[a][b] * CSF[a][b];
            double delta = quadrant_delta(block, 1e-9);
            energy[j * bx + i] = ew * delta / 16.0;
            mean_l[j * bx + i] = s / 64.0;
        }
    memcpy(sorted, energy, sizeof(double) * nb);
    qsort(sorted, nb, sizeof(double), cmp_double);
    double norm = nb ? sorted[(size_t)llround((double)(nb - 1) * 0.75)] : 1.0;
    if (norm < 1.0) norm = 1.0;
    double sum = 0.0;
    for (size_t k = 0; k < nb; k++) {
        double nm = energy[k] / (energy[k] + norm);
        double vis = 1.0 / (1.0 + 0.75 * nm * 4.0);
        vis = vis < 0.25 ? 0.25 : vis > 1.0 ? 1.0 : vis;
        if (mean_l[k] > 210.0) {
            double fade = (mean_l[k] - 210.0) / 40.0;
            fade = fade < 0.0 ? 0.0 : fade > 1.0 ? 1.0 : fade;
            double ceil_w = 0.25 * (1.0 - fade) + 0.05 * fade;
            if (vis > ceil_w) vis = ceil_w;
        }
        sum += vis;
    }
    free(luma);
    free(energy);
    free(sorted);
    free(mean_l);
    return sum;
}

/* t1.rs:1226-1322: per bit-plane, the newly significant count of every
 * sample; returns the total (the reference keeps the pass list unused). */
long ora_plan_coding_passes(const int32_t *coef, int n, int mbp) {
    if (mbp == 0) return 0;
    unsigned char *sig = calloc((size_t)n, 1);
    long total = 0;
    for (int bp = mbp - 1; bp >= 0; bp--) {
        for (int i = 0; i < n; i++) {
            if (!sig[i]) {
                uint32_t m = coef[i] < 0 ? (uint32_t)(-(int64_t)coef[i]) : (uint32_t)coef[i];
                if ((m >> bp) & 1u) {
                    sig[i] = 1;
                    total++;
                }
            }
        }
    }
    free(sig);
    return total;
}
