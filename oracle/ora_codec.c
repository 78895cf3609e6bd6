/* Oracle (test infrastructure only): encode/decode orchestration and the
 * flat C API the tests and the CPU-baseline leg load through ctypes.
 * Restates encode/mod.rs:84-101, encode/context.rs:27-70,
 * encode/backend/native/backend.rs:52-242, :436-517, layout.rs:37-173,
 * decode/mod.rs:56-129, decode/t1.rs:32-73 and decode/reconstruct.rs. */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ora_t2.h"

static int band_gain(int b) { return b == ORA_LL ? 0 : (b == ORA_HH ? 2 : 1); }

/* subband rectangle of (res, band) in the Mallat plane (layout.rs:37-97) */
static void band_rect(size_t w, size_t h, int levels, int res, int band, int *x0, int *y0, int *x1, int *y1) {
    size_t rw[33], rh[33];
    ora_resolutions(w, h, levels, rw, rh);
    if (res == 0) {
        *x0 = 0; *y0 = 0; *x1 = (int)rw[0]; *y1 = (int)rh[0];
    } else if (band == ORA_HL) {
        *x0 = (int)rw[res - 1]; *y0 = 0; *x1 = (int)rw[res]; *y1 = (int)rh[res - 1];
    } else if (band == ORA_LH) {
        *x0 = 0; *y0 = (int)rh[res - 1]; *x1 = (int)rw[res - 1]; *y1 = (int)rh[res];
    } else {
        *x0 = (int)rw[res - 1]; *y0 = (int)rh[res - 1]; *x1 = (int)rw[res]; *y1 = (int)rh[res];
    }
}

/* context.rs:39-70: integer BT.601 YCbCr -> RGB for the hidden colour spaces */
static void yuv_to_rgb(const int32_t *const *in, size_t n, int32_t **out) {
    for (size_t i = 0; i < n; i++) {
        int32_t y = in[0][i], d = in[1][i] - 128, e = in[2][i] - 128;
        int32_t r = y + ((91881 * e) >> 16);
        int32_t g = y - ((22554 * d + 46802 * e) >> 16);
        int32_t b = y + ((116130 * d) >> 16);
        out[0][i] = r < 0 ? 0 : (r > 255 ? 255 : r);
        out[1][i] = g < 0 ? 0 : (g > 255 ? 255 : g);
        out[2][i] = b < 0 ? 0 : (b > 255 ? 255 : b);
    }
}

/* yuv_to_rgb on flat planes (3 x n), for the context.rs:107-157 test */
void ora_ycbcr_to_rgb(const int32_t *planes_flat, size_t n, int32_t *out_flat) {
    const int32_t *in[3] = {planes_flat, planes_flat + n, planes_flat + 2 * n};
    int32_t *out[3] = {out_flat, out_flat + n, out_flat + 2 * n};
    yuv_to_rgb(in, n, out);
}

/* Quantised Mallat coefficient plane of one component
 * (backend.rs:52-84 lossy, :118-159 lossless). */
static void component_coefficients(const ora_plan *p, const int32_t *const *src, int c, int32_t *out) {
    size_t n = (size_t)p->width * p->height;
    if (p->lossless) {
        if (p->use_mct) ora_forward_rct(src[0], src[1], src[2], n, c, out);
        else ora_level_shift_i32(src[c], n, out);
        ora_fwd53(out, p->width, p->height, p->levels);
    } else {
        float *f = malloc(sizeof(float) * n);
        if (p->use_mct) ora_forward_ict(src[0], src[1], src[2], n, c, f);
        else ora_level_shift_f32(src[c], n, f);
        ora_fwd97(f, p->width, p->height, p->levels);
        ora_quantize_plane(f, p->width, p->height, p->levels, p, out);
        free(f);
    }
}

/* Mb of a subband (t1.rs:19-22 lossless with backend.rs:179-189 effective
 * precision; backend.rs:190-205 lossy: G - 1 + exponent, u8 saturating). */
static int subband_mb(const ora_plan *p, int c, int res, int band) {
    if (p->lossless) {
        int prec = (p->use_mct && c > 0) ? 9 : 8;
        int v = (p->guard - 1) + prec + band_gain(band);
        return v > 255 ? 255 : v;
    }
    for (int i = 0; i < p->nbands; i++)
        if (p->sq[i].resolution == res && p->sq[i].band == band) {
            int v = (p->guard - 1) + p->sq[i].exponent;
            return v > 255 ? 255 : v;
        }
    return (p->guard - 1) + 8 + band_gain(band);
}

static int count_blocks(int x0, int x1, int cb) { return x1 > x0 ? (x1 - x0 + cb - 1) / cb : 0; }

/* Results of the faithful-mode dead work land here, so it is not elided. */
static volatile long faithful_sink;

void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }

int ora_encode(const int32_t *const *planes, uint32_t w, uint32_t h, int ncomp, int colorspace,
               int quality, int format, ora_encoding *enc) {
    memset(enc, 0, sizeof(*enc));
    ora_plan *p = &enc->plan;
    int rc = ora_build_plan(planes, w, h, ncomp, colorspace, quality, format, p, enc->err);
    if (rc) return rc;
    size_t n = (size_t)w * h;
    const int32_t *src[3] = {planes[0], ncomp > 1 ? planes[1] : NULL, ncomp > 2 ? planes[2] : NULL};
    int32_t *conv[3] = {NULL, NULL, NULL};
    if (colorspace == ORA_YUV || colorspace == ORA_YCBCR) {
        for (int c = 0; c < 3; c++) conv[c] = malloc(sizeof(int32_t) * n);
        yuv_to_rgb(planes, n, conv);
        for (int c = 0; c < 3; c++) src[c] = conv[c];
    }
    int32_t **coef = malloc(sizeof(int32_t *) * (size_t)ncomp);
    for (int c = 0; c < ncomp; c++) {
        coef[c] = malloc(sizeof(int32_t) * n);
        component_coefficients(p, src, c, coef[c]);
    }
    /* blocks in codestream order: resolution, component, band, raster */
    int total = 0;
    for (int r = 0; r <= p->levels; r++)
        for (int c = 0; c < ncomp; c++)
            for (int b = (r ? ORA_HL : ORA_LL); b <= (r ? ORA_HH : ORA_LL); b++) {
                int x0, y0, x1, y1;
                band_rect(w, h, p->levels, r, b, &x0, &y0, &x1, &y1);
                total += count_blocks(x0, x1, p->cbw) * count_blocks(y0, y1, p->cbh);
            }
    enc->nblocks = total;
    enc->blocks = calloc((size_t)(total ? total : 1), sizeof(ora_block));
    int k = 0;
    for (int r = 0; r <= p->levels; r++)
        for (int c = 0; c < ncomp; c++)
            for (int b = (r ? ORA_HL : ORA_LL); b <= (r ? ORA_HH : ORA_LL); b++) {
                int x0, y0, x1, y1;
                band_rect(w, h, p->levels, r, b, &x0, &y0, &x1, &y1);
                for (int by = y0; by < y1; by += p->cbh)
                    for (int bx = x0; bx < x1; bx += p->cbw) {
                        ora_block *blk = &enc->blocks[k++];
                        blk->comp = c;
                        blk->res = r;
                        blk->band = b;
                        blk->x0 = bx;
                        blk->y0 = by;
                        blk->x1 = bx + p->cbw < x1 ? bx + p->cbw : x1;
                        blk->y1 = by + p->cbh < y1 ? by + p->cbh : y1;
                    }
            }
    /* Tier-1, parallel over code-blocks like the reference's rayon loops
     * (t1.rs:191-196, :287-290) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < total; i++) {
        ora_block *blk = &enc->blocks[i];
        int bw = blk->x1 - blk->x0, bh = blk->y1 - blk->y0;
        int32_t *cb = malloc(sizeof(int32_t) * (size_t)(bw * bh));
        for (int y = 0; y < bh; y++)
            memcpy(cb + y * bw, coef[blk->comp] + (size_t)(blk->y0 + y) * w + blk->x0, sizeof(int32_t) * (size_t)bw);
        ora_t1_encode_block(cb, bw, bh, blk->band, subband_mb(p, blk->comp, blk->res, blk->band), blk);
        if (ora_faithful()) faithful_sink += ora_plan_coding_passes(cb, bw * bh, blk->mbp);
        free(cb);
    }
    for (int c = 0; c < ncomp; c++) free(coef[c]);
    free(coef);
    for (int c = 0; c < 3; c++) free(conv[c]);
    /* backend.rs:650-652: the lossy path builds the contrast mask that
     * rate.rs:60 then ignores (faithful mode only) */
    if (ora_faithful() && !p->lossless) faithful_sink += (long)ora_contrast_mask(src, ncomp, p->domain, w, h);
    ora_pcrd_apply(enc, p);
    /* Tier-2: one packet per (resolution, component) (t2.rs:409-470) */
    obuf pay = {0};
    int bstart = 0;
    for (int r = 0; r <= p->levels; r++)
        for (int c = 0; c < ncomp; c++) {
            ora_band_ref refs[3];
            int nb = 0;
            for (int b = (r ? ORA_HL : ORA_LL); b <= (r ? ORA_HH : ORA_LL); b++) {
                int x0, y0, x1, y1;
                band_rect(w, h, p->levels, r, b, &x0, &y0, &x1, &y1);
                int cx = count_blocks(x0, x1, p->cbw), cy = count_blocks(y0, y1, p->cbh);
                refs[nb].nblocks = cx * cy;
                refs[nb].cx = cx;
                refs[nb].cy = cy;
                refs[nb].blocks = malloc(sizeof(ora_block *) * (size_t)(cx * cy > 0 ? cx * cy : 1));
                for (int i = 0; i < cx * cy; i++) refs[nb].blocks[i] = &enc->blocks[bstart + i];
                bstart += cx * cy;
                nb++;
            }
            ora_packet_header(refs, nb, &pay);
            for (int i = 0; i < nb; i++) {
                for (int j = 0; j < refs[i].nblocks; j++) {
                    ora_block *blk = refs[i].blocks[j];
                    if (blk->npasses) obuf_put(&pay, blk->bytes, (size_t)blk->body_len);
                }
                free(refs[i].blocks);
            }
        }
    obuf cs = {0};
    ora_write_codestream(p, pay.data, pay.len, &cs);
    free(pay.data);
    enc->codestream = cs.data;
    enc->codestream_len = cs.len;
    if (format == 0) {
        obuf jp = {0};
        ora_wrap_jp2(p, cs.data, cs.len, &jp);
        enc->file = jp.data;
        enc->file_len = jp.len;
    } else {
        enc->file = malloc(cs.len);
        memcpy(enc->file, cs.data, cs.len);
        enc->file_len = cs.len;
    }
    return ORA_OK;
}

void ora_encoding_free(ora_encoding *enc) {
    for (int i = 0; i < enc->nblocks; i++) free(enc->blocks[i].bytes);
    free(enc->blocks);
    free(enc->codestream);
    free(enc->file);
    memset(enc, 0, sizeof(*enc));
}

/* decode/mod.rs:56-129 + decode/t1.rs:32-73 + decode/reconstruct.rs:24-302.
 * recon (optional, 9/7 only): the f32 component planes after the inverse
 * DWT, before the inverse colour transform and finalisation. */
static int decode_impl(const uint8_t *bytes, size_t len, uint32_t *ow, uint32_t *oh, int *onc, int *ocs,
                       int32_t **oplanes, char *err, float **recon) {
    ora_jp2_header jh;
    const uint8_t *cs;
    size_t cslen;
    memset(&jh, 0, sizeof(jh));
    int rc = ora_parse_jp2(bytes, len, &jh, &cs, &cslen, err);
    if (rc) return rc;
    ora_cs_header hd;
    rc = ora_parse_codestream(cs, cslen, &hd, err);
    if (rc) return rc;
    if (jh.width != hd.width || jh.height != hd.height) {
        snprintf(err, 256, "JP2 ihdr dimensions do not match SIZ dimensions");
        return ORA_DECODE_FAILED;
    }
    if (jh.bpc != 8) {
        snprintf(err, 256, "unsupported JP2 bit depth: decoder currently supports 8-bit components, found %d bits", jh.bpc);
        return ORA_UNSUPPORTED;
    }
    if (jh.ncomp != hd.ncomp) {
        snprintf(err, 256, "JP2 ihdr component count does not match SIZ component count");
        return ORA_DECODE_FAILED;
    }
    if (jh.colorspace == ORA_GRAY && jh.ncomp != 1) {
        snprintf(err, 256, "unsupported JP2 component count");
        return ORA_UNSUPPORTED;
    }
    if (jh.colorspace == ORA_SRGB && jh.ncomp != 3) {
        snprintf(err, 256, "unsupported JP2 component count");
        return ORA_UNSUPPORTED;
    }
    uint32_t w = hd.width, h = hd.height;
    int L = hd.levels, nc = hd.ncomp;
    /* band states in component-major order (decode/t2.rs:266-340) */
    int nbands = nc * (1 + 3 * L);
    ora_dec_band *bands = calloc((size_t)nbands, sizeof(ora_dec_band));
    int nblocks = 0, bi = 0;
    for (int c = 0; c < nc; c++)
        for (int r = 0; r <= L; r++)
            for (int b = (r ? ORA_HL : ORA_LL); b <= (r ? ORA_HH : ORA_LL); b++) {
                int x0, y0, x1, y1;
                band_rect(w, h, L, r, b, &x0, &y0, &x1, &y1);
                int cx = count_blocks(x0, x1, hd.cbw), cy = count_blocks(y0, y1, hd.cbh);
                bands[bi] = (ora_dec_band){c, r, b, nblocks, cx * cy, cx, cy};
                nblocks += cx * cy;
                bi++;
            }
    ora_dec_block *blocks = calloc((size_t)(nblocks ? nblocks : 1), sizeof(ora_dec_block));
    for (int i = 0; i < nbands; i++) {
        int x0, y0, x1, y1;
        band_rect(w, h, L, bands[i].res, bands[i].band, &x0, &y0, &x1, &y1);
        int k = bands[i].first;
        for (int by = y0; by < y1; by += hd.cbh)
            for (int bx = x0; bx < x1; bx += hd.cbw) {
                ora_dec_block *blk = &blocks[k++];
                blk->comp = bands[i].comp;
                blk->res = bands[i].res;
                blk->band = bands[i].band;
                blk->band_index = i;
                blk->x0 = bx;
                blk->y0 = by;
                blk->x1 = bx + hd.cbw < x1 ? bx + hd.cbw : x1;
                blk->y1 = by + hd.cbh < y1 ? by + hd.cbh : y1;
                blk->zbp = -1;
                blk->numlenbits = 3;
            }
    }
    rc = ora_parse_packets(&hd, blocks, nblocks, bands, nbands, err);
    if (rc) {
        free(bands);
        free(blocks);
        return rc;
    }
    size_t n = (size_t)w * h;
    int32_t *coef = calloc(n * (size_t)nc, sizeof(int32_t));
    /* the reference decodes code-blocks serially (decode/t1.rs:50) */
    for (int i = 0; i < nblocks && !rc; i++) {
        ora_dec_block *blk = &blocks[i];
        if (!blk->included) continue;
        int step_i = blk->band_index % hd.nsteps;
        int mb = (hd.guard > 0 ? hd.guard - 1 : 0) + hd.exp[step_i];
        if (mb > 255) mb &= 255;
        int bw = blk->x1 - blk->x0, bh = blk->y1 - blk->y0;
        int32_t *out = malloc(sizeof(int32_t) * (size_t)(bw * bh));
        rc = ora_t1_decode_block(bw, bh, blk->band, mb, blk->zbp, blk->passes, blk->data, blk->len, out, err);
        if (!rc)
            for (int y = 0; y < bh; y++)
                memcpy(coef + (size_t)blk->comp * n + (size_t)(blk->y0 + y) * w + blk->x0, out + y * bw,
                       sizeof(int32_t) * (size_t)bw);
        free(out);
    }
    free(blocks);
    if (rc) {
        free(bands);
        free(coef);
        return rc;
    }
    int32_t *planes = malloc(sizeof(int32_t) * n * (size_t)nc);
    if (hd.transform == 1) {
        for (int c = 0; c < nc; c++) ora_inv53(coef + (size_t)c * n, w, h, L);
        if (nc == 3 && hd.mct) ora_inverse_rct(coef, coef + n, coef + 2 * n, n);
        for (int c = 0; c < nc; c++) ora_finalize_i32(coef + (size_t)c * n, n, planes + (size_t)c * n);
    } else {
        float *f = malloc(sizeof(float) * n * (size_t)nc);
        for (int c = 0; c < nc; c++) {
            float *fc = f + (size_t)c * n;
            for (int b = 0; b < 1 + 3 * L; b++) {
                int res = b == 0 ? 0 : (b - 1) / 3 + 1;
                int band = b == 0 ? ORA_LL : (b - 1) % 3 + 1;
                int x0, y0, x1, y1;
                band_rect(w, h, L, res, band, &x0, &y0, &x1, &y1);
                float step = ora_quant_step_f32((uint32_t)hd.prec[c], band, hd.exp[b], hd.mant[b]);
                for (int y = y0; y < y1; y++)
                    for (int x = x0; x < x1; x++)
                        fc[(size_t)y * w + x] = ora_dequantize(coef[(size_t)c * n + (size_t)y * w + x], step);
            }
            ora_inv97(fc, w, h, L);
        }
        if (recon) {
            *recon = malloc(sizeof(float) * n * (size_t)nc);
            memcpy(*recon, f, sizeof(float) * n * (size_t)nc);
        }
        if (nc == 3 && hd.mct) ora_inverse_ict(f, f + n, f + 2 * n, n);
        for (int c = 0; c < nc; c++) ora_finalize_f32(f + (size_t)c * n, n, planes + (size_t)c * n);
        free(f);
    }
    free(coef);
    free(bands);
    *ow = w;
    *oh = h;
    *onc = nc;
    *ocs = jh.colorspace;
    *oplanes = planes;
    return ORA_OK;
}

int ora_decode(const uint8_t *bytes, size_t len, uint32_t *ow, uint32_t *oh, int *onc, int *ocs,
               int32_t **oplanes, char *err) {
    return decode_impl(bytes, len, ow, oh, onc, ocs, oplanes, err, NULL);
}

/* ---- encode_with_psnr metrics (encode/mod.rs:114-135) ------------------------ */

/* backend.rs:626-639: midpoint reconstruction of a coefficient whose block
 * kept its passes down to bit-plane last_bp (-1: no pass kept). */
static float dequantize_truncated(int32_t q, int last_bp, float step) {
    if (last_bp < 0) return 0.0f;
    uint32_t aq = q < 0 ? (uint32_t)(-(int64_t)q) : (uint32_t)q;
    uint32_t at = (aq >> last_bp) << last_bp;
    if (at == 0) return 0.0f;
    float sign = q >= 0 ? 1.0f : -1.0f;
    return (sign * ((float)at + 0.5f)) * step;
}

/* backend.rs:331-368: the f32 component planes the encoder's PCRD choice
 * implies, after the inverse 9/7 (level-shifted domain).  planes are the
 * encoder's input samples (RGB for sRGB); recon holds ncomp planes. */
int ora_metric_recon(const int32_t *const *planes, const ora_encoding *enc, float *recon) {
    const ora_plan *p = &enc->plan;
    if (p->lossless) return ORA_ENCODE_FAILED;
    size_t n = (size_t)p->width * p->height;
    int32_t *coef = malloc(sizeof(int32_t) * n);
    for (int c = 0; c < p->ncomp; c++) {
        float *rc = recon + (size_t)c * n;
        component_coefficients(p, planes, c, coef);
        memset(rc, 0, sizeof(float) * n);
        for (int i = 0; i < enc->nblocks; i++) {
            const ora_block *b = &enc->blocks[i];
            if (b->comp != c) continue;
            float step = 0.0f;
            for (int k = 0; k < p->nbands; k++)
                if (p->sq[k].resolution == b->res && p->sq[k].band == b->band)
                    step = ora_quant_step_f32(8, b->band, p->sq[k].exponent, p->sq[k].mantissa);
            int last_bp = b->npasses ? b->passes[b->npasses - 1].bitplane : -1;
            for (int y = b->y0; y < b->y1; y++)
                for (int x = b->x0; x < b->x1; x++) {
                    size_t k = (size_t)y * p->width + (size_t)x;
                    rc[k] = dequantize_truncated(coef[k], last_bp, step);
                }
        }
        ora_inv97(rc, p->width, p->height, p->levels);
    }
    free(coef);
    return ORA_OK;
}

/* backend.rs:909-959: mean SSIM over the non-overlapping 8x8 blocks. */
static double mssim_8x8(const float *orig, const float *rec, size_t w, size_t h) {
    const double C1 = 6.5025, C2 = 58.5225;
    size_t rows = h / 8, cols = w / 8;
    if (rows == 0 || cols == 0) return 1.0;
    double total = 0.0;
    for (size_t br = 0; br < rows; br++)
        for (size_t bc = 0; bc < cols; bc++) {
            double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
            for (size_t dy = 0; dy < 8; dy++)
                for (size_t dx = 0; dx < 8; dx++) {
                    size_t i = (br * 8 + dy) * w + bc * 8 + dx;
                    double x = orig[i], y = rec[i];
                    sx += x;
                    sy += y;
                    sxx += x * x;
                    syy += y * y;
                    sxy += x * y;
                }
            double ux = sx / 64.0, uy = sy / 64.0;
            double vx = sxx / 64.0 - ux * ux, vy = syy / 64.0 - uy * uy;
            if (vx < 0.0) vx = 0.0;
            if (vy < 0.0) vy = 0.0;
            double cxy = sxy / 64.0 - ux * uy;
            total += ((2.0 * ux * uy + C1) * (2.0 * cxy + C2)) / ((ux * ux + uy * uy + C1) * (vx + vy + C2));
        }
    return total / (double)(rows * cols);
}

static float clamp255(float v) { return v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v); }

/* backend.rs:370-432: PSNR over every sample and SSIM over BT.601 luma of
 * the reconstruction recon (gray: 1 plane; sRGB: Y, Cb, Cr planes). */
void ora_metrics_from_recon(const int32_t *const *orig, const float *recon, uint32_t w, uint32_t h, int ncomp,
                            double *psnr, double *ssim) {
    size_t n = (size_t)w * h;
    float *dl = malloc(sizeof(float) * n), *ol = malloc(sizeof(float) * n);
    double sse = 0.0;
    size_t samples;
    if (ncomp == 1) {
        samples = n;
        for (size_t i = 0; i < n; i++) {
            float d = clamp255(recon[i] + 128.0f);
            float diff = d - (float)orig[0][i];
            sse += (double)(diff * diff);
            dl[i] = d;
            ol[i] = (float)orig[0][i];
        }
    } else {
        samples = n * 3;
        const float *yc = recon, *cb = recon + n, *cr = recon + 2 * n;
        for (size_t i = 0; i < n; i++) {
            float r = clamp255((yc[i] + 1.402f * cr[i]) + 128.0f);
            float g = clamp255(((yc[i] - 0.34413f * cb[i]) - 0.71414f * cr[i]) + 128.0f);
            float b = clamp255((yc[i] + 1.772f * cb[i]) + 128.0f);
            float dr = r - (float)orig[0][i], dg = g - (float)orig[1][i], db = b - (float)orig[2][i];
            sse += (double)((dr * dr + dg * dg) + db * db);
            dl[i] = (0.299f * r + 0.587f * g) + 0.114f * b;
            ol[i] = (0.299f * (float)orig[0][i] + 0.587f * (float)orig[1][i]) + 0.114f * (float)orig[2][i];
        }
    }
    double mse = sse / (double)samples;
    *psnr = mse < 1e-10 ? 100.0 : 20.0 * log10(255.0 / sqrt(mse));
    *ssim = mssim_8x8(ol, dl, w, h);
    free(dl);
    free(ol);
}

/* backend.rs:305-433 (compute_quality_metrics) after a successful encode
 * of planes (the caller's image components, colour space as given). */
int ora_quality_metrics(const int32_t *const *planes, int colorspace, const ora_encoding *enc, double *psnr,
                        double *ssim, char *err) {
    const ora_plan *p = &enc->plan;
    if (p->quality >= 100) {
        *psnr = INFINITY;
        *ssim = 1.0;
        return ORA_OK;
    }
    if (!(colorspace == ORA_GRAY || (colorspace == ORA_SRGB && p->ncomp == 3 && p->use_mct))) {
        snprintf(err, 256, "quality metrics not implemented for this colorspace configuration");
        return ORA_ENCODE_FAILED;
    }
    size_t n = (size_t)p->width * p->height;
    float *recon = malloc(sizeof(float) * n * (size_t)p->ncomp);
    int rc = ora_metric_recon(planes, enc, recon);
    if (!rc) ora_metrics_from_recon(planes, recon, p->width, p->height, p->ncomp, psnr, ssim);
    free(recon);
    return rc;
}

/* ---- flat API for ctypes ---------------------------------------------------- */
static void u8_to_planes(const uint8_t *pix, size_t n, int ch, int32_t **planes) {
    for (int c = 0; c < ch; c++) {
        planes[c] = malloc(sizeof(int32_t) * n);
        for (size_t i = 0; i < n; i++) planes[c][i] = pix[i * (size_t)ch + (size_t)c];
    }
}

void ora_free(void *p) { free(p); }

/* Encode interleaved u8 pixels (gray: 1 channel, sRGB: 3) like
 * Image::from_gray_bytes / from_rgb_bytes + encode (model.rs:87-140). */
ora_encoding *ora_encode_u8(const uint8_t *pix, uint32_t w, uint32_t h, int channels, int quality, int format, int *status) {
    ora_encoding *e = calloc(1, sizeof(*e));
    int32_t *planes[3] = {0};
    u8_to_planes(pix, (size_t)w * h, channels, planes);
    *status = ora_encode((const int32_t *const *)planes, w, h, channels, channels == 1 ? ORA_GRAY : ORA_SRGB,
                         quality, format, e);
    for (int c = 0; c < channels; c++) free(planes[c]);
    return e;
}
ora_encoding *ora_encode_planes(const int32_t *planes_flat, uint32_t w, uint32_t h, int ncomp, int colorspace,
                                int quality, int format, int *status) {
    ora_encoding *e = calloc(1, sizeof(*e));
    const int32_t *planes[3] = {0};
    for (int c = 0; c < ncomp && c < 3; c++) planes[c] = planes_flat + (size_t)c * w * h;
    *status = ora_encode(planes, w, h, ncomp, colorspace, quality, format, e);
    return e;
}
void ora_enc_release(ora_encoding *e) {
    if (!e) return;
    ora_encoding_free(e);
    free(e);
}
const char *ora_enc_error(const ora_encoding *e) { return e->err; }
size_t ora_enc_file(const ora_encoding *e, const uint8_t **data) {
    *data = e->file;
    return e->file_len;
}
size_t ora_enc_codestream(const ora_encoding *e, const uint8_t **data) {
    *data = e->codestream;
    return e->codestream_len;
}
int ora_enc_nblocks(const ora_encoding *e) { return e->nblocks; }
double ora_enc_lambda(const ora_encoding *e, int c) { return e->lambda[c]; }
void ora_enc_plan(const ora_encoding *e, int32_t *out) {
    const ora_plan *p = &e->plan;
    out[0] = p->levels;
    out[1] = p->cbw;
    out[2] = p->nbands;
    out[3] = p->use_mct;
    out[4] = p->lossless;
    for (int i = 0; i < p->nbands; i++) {
        out[5 + 2 * i] = p->sq[i].exponent;
        out[6 + 2 * i] = p->sq[i].mantissa;
    }
}
/* info[16] = comp res band x0 y0 x1 y1 mbp zbp cls npasses full_npasses nbytes body_len max nnz */
void ora_enc_block_info(const ora_encoding *e, int i, int32_t *info) {
    const ora_block *b = &e->blocks[i];
    int32_t v[16] = {b->comp, b->res, b->band, b->x0, b->y0, b->x1, b->y1, b->mbp, b->zbp, b->cls,
                     b->npasses, b->full_npasses, b->nbytes, b->body_len, (int32_t)b->max_magnitude, b->nnz};
    memcpy(info, v, sizeof(v));
}
/* per full pass: kind bitplane newsig sigbefore length */
void ora_enc_block_passes(const ora_encoding *e, int i, int32_t *out) {
    const ora_block *b = &e->blocks[i];
    for (int k = 0; k < b->full_npasses; k++) {
        out[5 * k + 0] = b->passes[k].kind;
        out[5 * k + 1] = b->passes[k].bitplane;
        out[5 * k + 2] = b->passes[k].newly_significant;
        out[5 * k + 3] = b->passes[k].significant_before;
        out[5 * k + 4] = b->full_lengths[k];
    }
}
/* PCRD record of block i (lossy encodes): the raw curve, origin first, as
 * (bytes, cumulative distortion) pairs, and the hull's pass counts.
 * Returns raw_n; *hull_n receives the hull length. */
int ora_enc_block_rd(const ora_encoding *e, int i, uint32_t *raw_bytes, double *raw_dist, int32_t *hull_passes,
                     int32_t *hull_n) {
    const ora_block *b = &e->blocks[i];
    for (int k = 0; k < b->raw_n; k++) {
        raw_bytes[k] = b->raw_bytes[k];
        raw_dist[k] = b->raw_dist[k];
    }
    for (int k = 0; k < b->hull_n; k++) hull_passes[k] = b->hull_passes[k];
    *hull_n = b->hull_n;
    return b->raw_n;
}
double ora_enc_max_slope(const ora_encoding *e, int c) { return e->max_slope[c]; }
/* encode/counters.rs:3-9 for this encode: total blocks, empty blocks, MQ
 * symbols, cleanup / SP / MR passes (as coded, before truncation) and the
 * coded bytes of every pass. */
void ora_enc_counters(const ora_encoding *e, uint64_t *out) {
    memset(out, 0, 7 * sizeof(uint64_t));
    for (int i = 0; i < e->nblocks; i++) {
        const ora_block *b = &e->blocks[i];
        out[0]++;
        if (b->mbp == 0) {
            out[1]++;
            continue;
        }
        out[2] += b->nsym;
        for (int k = 0; k < b->full_npasses; k++) {
            int kind = b->passes[k].kind;
            out[kind == ORA_CL ? 3 : kind == ORA_SP ? 4 : 5]++;
            out[6] += (uint64_t)b->full_lengths[k];
        }
    }
}
uint32_t ora_enc_block_nsym(const ora_encoding *e, int i) { return e->blocks[i].nsym; }
void ora_enc_block_bytes(const ora_encoding *e, int i, uint8_t *out) {
    const ora_block *b = &e->blocks[i];
    if (b->nbytes) memcpy(out, b->bytes, (size_t)b->nbytes);
}

/* Quantised coefficient planes (Mallat layout) as Tier-1 sees them. */
int ora_coefficients_u8(const uint8_t *pix, uint32_t w, uint32_t h, int channels, int quality, int32_t *out) {
    int32_t *planes[3] = {0};
    size_t n = (size_t)w * h;
    u8_to_planes(pix, n, channels, planes);
    ora_plan p;
    char err[256];
    int rc = ora_build_plan((const int32_t *const *)planes, w, h, channels, channels == 1 ? ORA_GRAY : ORA_SRGB,
                            quality, 0, &p, err);
    if (!rc)
        for (int c = 0; c < channels; c++) component_coefficients(&p, (const int32_t *const *)planes, c, out + (size_t)c * n);
    for (int c = 0; c < channels; c++) free(planes[c]);
    return rc;
}

/* Decode to interleaved u8 (caller buffer w*h*ncomp, sized via inspect). */
int ora_decode_u8(const uint8_t *bytes, size_t len, uint32_t *w, uint32_t *h, int *nc, uint8_t *out, size_t cap,
                  char *err) {
    int32_t *planes = NULL;
    int cs;
    int rc = ora_decode(bytes, len, w, h, nc, &cs, &planes, err);
    if (rc) return rc;
    size_t n = (size_t)*w * *h;
    if (n * (size_t)*nc > cap) {
        free(planes);
        snprintf(err, 256, "output buffer too small");
        return ORA_IO;
    }
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < *nc; c++) out[i * (size_t)*nc + (size_t)c] = (uint8_t)planes[(size_t)c * n + i];
    free(planes);
    return ORA_OK;
}

/* Tier-1 block coder entry points for the block-level parity tests. */
int ora_t1_encode_flat(const int32_t *coef, int w, int h, int band, int mb, int32_t *info, int32_t *passes,
                       uint8_t *bytes, int cap) {
    ora_block b;
    memset(&b, 0, sizeof(b));
    ora_t1_encode_block(coef, w, h, band, mb, &b);
    info[0] = b.mbp;
    info[1] = b.zbp;
    info[2] = b.cls;
    info[3] = b.npasses;
    info[4] = b.nbytes;
    info[5] = (int32_t)b.nsym;
    for (int k = 0; k < b.npasses; k++) {
        passes[5 * k + 0] = b.passes[k].kind;
        passes[5 * k + 1] = b.passes[k].bitplane;
        passes[5 * k + 2] = b.passes[k].newly_significant;
        passes[5 * k + 3] = b.passes[k].significant_before;
        passes[5 * k + 4] = b.passes[k].length;
    }
    int rc = b.nbytes <= cap ? ORA_OK : ORA_ENCODE_FAILED;
    if (!rc && b.nbytes) memcpy(bytes, b.bytes, (size_t)b.nbytes);
    free(b.bytes);
    return rc;
}

/* MQ streams for the coder KATs (mq/mod.rs:213-242, :61-81): encode n
 * (ctx, bit) symbols, return the terminated length (or -1 if cap is too
 * small); decode n symbols of the given contexts. */
int ora_mq_encode_stream(const uint8_t *ctx, const uint8_t *bits, int n, uint8_t *out, int cap) {
    size_t bcap = (size_t)n * 2 + 64;
    uint8_t *buf = malloc(bcap);
    ora_mqc m;
    ora_mqc_init(&m, buf, bcap);
    for (int i = 0; i < n; i++) ora_mqc_encode(&m, ctx[i], bits[i]);
    uint8_t *seg;
    size_t len = ora_mqc_flush(&m, &seg);
    int rc = (int)len <= cap ? (int)len : -1;
    if (rc >= 0) memcpy(out, seg, len);
    free(buf);
    return rc;
}
void ora_mq_decode_stream(const uint8_t *data, int len, const uint8_t *ctx, int n, uint8_t *bits) {
    ora_mqd d;
    ora_mqd_init(&d, data, (size_t)len);
    for (int i = 0; i < n; i++) bits[i] = (uint8_t)ora_mqd_decode(&d, ctx[i]);
}
void ora_tables(uint16_t *qe, uint8_t *mps, uint8_t *nmps, uint8_t *nlps, uint8_t *zc, uint8_t *sc, uint8_t *spb) {
    ora_mq_state_table(qe, mps, nmps, nlps);
    ora_t1_luts((uint8_t(*)[512])zc, sc, spb);
}

/* encode_with_psnr on interleaved u8 pixels: out2 = {psnr_db, ssim}. */
int ora_encode_metrics_u8(const uint8_t *pix, uint32_t w, uint32_t h, int channels, int quality, double *out2) {
    int32_t *planes[3] = {0};
    u8_to_planes(pix, (size_t)w * h, channels, planes);
    ora_encoding e;
    int rc = ora_encode((const int32_t *const *)planes, w, h, channels, channels == 1 ? ORA_GRAY : ORA_SRGB, quality, 0, &e);
    if (!rc) rc = ora_quality_metrics((const int32_t *const *)planes, channels == 1 ? ORA_GRAY : ORA_SRGB, &e,
                                      &out2[0], &out2[1], e.err);
    ora_encoding_free(&e);
    for (int c = 0; c < channels; c++) free(planes[c]);
    return rc;
}
/* The truncated reconstruction the metrics use (channels planes of w*h). */
int ora_metric_recon_u8(const uint8_t *pix, uint32_t w, uint32_t h, int channels, int quality, float *out) {
    int32_t *planes[3] = {0};
    u8_to_planes(pix, (size_t)w * h, channels, planes);
    ora_encoding e;
    int rc = ora_encode((const int32_t *const *)planes, w, h, channels, channels == 1 ? ORA_GRAY : ORA_SRGB, quality, 0, &e);
    if (!rc) rc = ora_metric_recon((const int32_t *const *)planes, &e, out);
    ora_encoding_free(&e);
    for (int c = 0; c < channels; c++) free(planes[c]);
    return rc;
}
/* PSNR/SSIM of a given reconstruction against i32 planes: out2 = {psnr, ssim}. */
void ora_metrics_flat(const int32_t *orig_flat, const float *recon, uint32_t w, uint32_t h, int ncomp, double *out2) {
    const int32_t *planes[3] = {0};
    for (int c = 0; c < ncomp && c < 3; c++) planes[c] = orig_flat + (size_t)c * w * h;
    ora_metrics_from_recon(planes, recon, w, h, ncomp, &out2[0], &out2[1]);
}
/* The decoder's f32 planes after the inverse 9/7 (before colour and
 * rounding); cap in floats.  Returns the status. */
int ora_decode_recon_f32(const uint8_t *bytes, size_t len, float *out, size_t cap) {
    uint32_t w, h;
    int nc, cs;
    int32_t *planes = NULL;
    float *rec = NULL;
    char err[256];
    int rc = decode_impl(bytes, len, &w, &h, &nc, &cs, &planes, err, &rec);
    if (!rc && (!rec || (size_t)w * h * (size_t)nc > cap)) rc = ORA_IO;
    if (!rc) memcpy(out, rec, sizeof(float) * (size_t)w * h * (size_t)nc);
    free(planes);
    free(rec);
    return rc;
}

/* ---- content analysis (content.rs:119-174) ---------------------------------- */

/* content.rs:110-117: integer BT.601 luma of clamped 8-bit R, G, B */
static int luma601(int32_t r, int32_t g, int32_t b) {
    r = r < 0 ? 0 : (r > 255 ? 255 : r);
    g = g < 0 ? 0 : (g > 255 ? 255 : g);
    b = b < 0 ? 0 : (b > 255 ? 255 : b);
    int v = (77 * r + 150 * g + 29 * b + 128) >> 8;
    return v < 0 ? 0 : (v > 255 ? 255 : v);
}

/* content.rs:120-166: out3 = {bright_share, entropy, colorfulness}.  Gray
 * reads component 0; every other colour space reads components 0..2 as
 * R, G, B.  n = samples per component. */
void ora_analyze_content(const int32_t *const *planes, size_t n, int ncomp, int colorspace, float *out3) {
    uint64_t hist[256] = {0}, colorful = 0, count = 0;
    if (colorspace == ORA_GRAY) {
        if (ncomp >= 1)
            for (size_t i = 0; i < n; i++) {
                int32_t v = planes[0][i];
                hist[v < 0 ? 0 : (v > 255 ? 255 : v)]++;
                count++;
            }
    } else if (ncomp >= 3) {
        for (size_t i = 0; i < n; i++) {
            int32_t r = planes[0][i], g = planes[1][i], b = planes[2][i];
            hist[luma601(r, g, b)]++;
            int32_t mx = r > g ? r : g, mn = r < g ? r : g;
            mx = mx > b ? mx : b;
            mn = mn < b ? mn : b;
            mx = mx < 0 ? 0 : (mx > 255 ? 255 : mx);
            mn = mn < 0 ? 0 : (mn > 255 ? 255 : mn);
            colorful += (uint64_t)(int64_t)(mx - mn);
            count++;
        }
    }
    if (count == 0) {
        out3[0] = out3[1] = out3[2] = 0.0f;
        return;
    }
    float total = (float)count;
    uint64_t bright = 0;
    for (int v = 225; v < 256; v++) bright += hist[v];
    float entropy = 0.0f;
    for (int v = 0; v < 256; v++)
        if (hist[v]) {
            float p = (float)hist[v] / total;
            entropy -= p * log2f(p);
        }
    out3[0] = (float)bright / total;
    out3[1] = entropy;
    out3[2] = (float)colorful / total;
}

/* content.rs:96-105: 1 = Graphics, 0 = Photo */
int ora_classify_content(const float *stats3) {
    return stats3[1] < 3.0f || (stats3[0] > 0.45f && stats3[1] < 5.0f);
}

void ora_analyze_content_flat(const int32_t *planes_flat, size_t n, int ncomp, int colorspace, float *out3) {
    const int32_t *planes[3] = {0};
    for (int c = 0; c < ncomp && c < 3; c++) planes[c] = planes_flat + (size_t)c * n;
    ora_analyze_content(planes, n, ncomp, colorspace, out3);
}
