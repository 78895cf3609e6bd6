/*
 * Oracle: a plain-C CPU restatement of the jp2lam 0.1.0 codec path, used as
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline.  The product (jp2lam_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference crate root (src/...).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, IEEE f32/f64, no
 * FMA contraction, glibc libm for the host-side pow/log2 of the planner).
 */
#ifndef ORA_H
#define ORA_H
#include <stddef.h>
#include <stdint.h>

#define ORA_OK 0
#define ORA_INVALID_INPUT 1
#define ORA_ENCODE_FAILED 2
#define ORA_DECODE_FAILED 3
#define ORA_UNSUPPORTED 4
#define ORA_IO 5

#define ORA_MAX_LEVELS 6
#define ORA_MAX_BANDS (1 + 3 * ORA_MAX_LEVELS)
#define ORA_MAX_PASSES 96

/* band orientation (plan/mod.rs:45-50; tier1/helpers.rs:13-20) */
enum { ORA_LL = 0, ORA_HL = 1, ORA_LH = 2, ORA_HH = 3 };
/* colour spaces (model.rs:1-12) */
enum { ORA_GRAY = 0, ORA_RGB = 1, ORA_SRGB = 2, ORA_YUV = 3, ORA_YCBCR = 4 }; /* model.rs:1-12 order */
/* pass kinds (t1.rs:99-104) */
enum { ORA_SP = 0, ORA_MR = 1, ORA_CL = 2 };

typedef struct {
    uint8_t resolution, band, exponent;
    uint16_t mantissa;
} ora_sq;

typedef struct {
    uint32_t width, height;
    int ncomp;
    int domain;   /* ORA_GRAY or ORA_SRGB (model.rs:14-19) */
    int quality;
    int format;   /* 0 = JP2, 1 = J2K */
    int lossless; /* quality >= 100 (plan/mod.rs:121) */
    int levels;
    int cbw, cbh;
    int guard;
    int use_mct;
    int nbands;
    ora_sq sq[ORA_MAX_BANDS];
} ora_plan;

typedef struct {
    int kind, bitplane;
    int newly_significant, significant_before;
    int length, cumulative_length;
    int is_terminated;
} ora_pass;

typedef struct {
    int comp, res, band;
    int x0, y0, x1, y1;
    uint32_t max_magnitude;
    int nnz;
    int mbp, zbp, cls;
    int npasses;
    ora_pass passes[ORA_MAX_PASSES];
    int full_npasses;          /* passes produced by Tier-1 before PCRD */
    int full_lengths[ORA_MAX_PASSES];
    uint8_t *bytes;            /* single terminated MQ segment */
    int nbytes;                /* full segment length */
    int body_len;              /* bytes emitted into the packet body */
    /* PCRD record of the last lossy encode (kept for the tests): the raw
     * curve, origin first (rate.rs:128-180), and the hull's pass counts. */
    uint32_t nsym;             /* MQ symbols coded for the block, all passes */
    int raw_n, hull_n;
    uint32_t raw_bytes[ORA_MAX_PASSES + 1];
    double raw_dist[ORA_MAX_PASSES + 1];
    int hull_passes[ORA_MAX_PASSES + 1];
} ora_block;

typedef struct {
    ora_plan plan;
    int nblocks;
    ora_block *blocks;          /* codestream order: res, comp, band, raster */
    double lambda[3];
    double max_slope[3];
    uint8_t *codestream;
    size_t codestream_len;
    uint8_t *file;              /* J2K or JP2 bytes as returned by encode */
    size_t file_len;
    char err[256];
} ora_encoding;

/* plan */
int ora_build_plan(const int32_t *const *planes, uint32_t w, uint32_t h, int ncomp,
                   int colorspace, int quality, int format, ora_plan *plan, char *err);
uint32_t ora_max_decompositions(uint32_t w, uint32_t h);
void ora_encode_stepsize(int32_t stepsize, uint32_t numbps, uint8_t *exp, uint16_t *mant);
void ora_irreversible_expounded_quant(uint32_t prec, uint8_t nres, uint8_t res, int band,
                                      uint8_t *exp, uint16_t *mant);
double ora_get_norm_97(uint32_t level, int band);
float ora_quant_step_f32(uint32_t precision, int band, uint8_t exponent, uint16_t mantissa);
double ora_quality_step_scaler(int quality);
void ora_scale_one_quant(uint8_t exponent, uint16_t mantissa, double scale, uint8_t *exp_out, uint16_t *mant_out);

/* colour + DWT + quantisation */
void ora_level_shift_f32(const int32_t *in, size_t n, float *out);
void ora_level_shift_i32(const int32_t *in, size_t n, int32_t *out);
void ora_forward_ict(const int32_t *r, const int32_t *g, const int32_t *b, size_t n, int comp, float *out);
void ora_forward_rct(const int32_t *r, const int32_t *g, const int32_t *b, size_t n, int comp, int32_t *out);
void ora_inverse_ict(float *y, float *cb, float *cr, size_t n);
void ora_inverse_rct(int32_t *y, int32_t *db, int32_t *dr, size_t n);
void ora_finalize_f32(const float *in, size_t n, int32_t *out);
void ora_finalize_i32(const int32_t *in, size_t n, int32_t *out);
void ora_fwd53(int32_t *d, size_t w, size_t h, int levels);
void ora_inv53(int32_t *d, size_t w, size_t h, int levels);
void ora_fwd97(float *d, size_t w, size_t h, int levels);
void ora_inv97(float *d, size_t w, size_t h, int levels);
void ora_fwd97_f64(double *d, size_t w, size_t h, int levels);
int32_t ora_quantize(float v, float step);
float ora_dequantize(int32_t q, float step);
void ora_quantize_plane(const float *d, size_t w, size_t h, int levels, const ora_plan *plan, int32_t *out);
void ora_resolutions(size_t w, size_t h, int levels, size_t *rw, size_t *rh);

/* MQ */
typedef struct {
    uint32_t a, c, ct;
    uint8_t *out;
    size_t cap, pos; /* out[0] is the dummy pre-byte */
    uint8_t ctx[19];
    uint64_t nsym;   /* symbols coded (counters.rs MQ_SYMBOLS, mq/mod.rs:213-216) */
} ora_mqc;
void ora_mqc_init(ora_mqc *m, uint8_t *buf, size_t cap);
void ora_mqc_encode(ora_mqc *m, int ctx, int bit);
size_t ora_mqc_flush(ora_mqc *m, uint8_t **bytes);
typedef struct {
    const uint8_t *bytes;
    size_t len, bp;
    uint32_t a, c, ct;
    uint8_t ctx[19];
} ora_mqd;
void ora_mqd_init(ora_mqd *d, const uint8_t *bytes, size_t len);
int ora_mqd_decode(ora_mqd *d, int ctx);
int ora_mq_state_table(uint16_t *qe, uint8_t *mps, uint8_t *nmps, uint8_t *nlps);

/* Tier-1 */
void ora_t1_luts(uint8_t zc[4][512], uint8_t sc[256], uint8_t spb[256]);
int ora_t1_encode_block(const int32_t *coeffs, int w, int h, int band, int mb, ora_block *blk);
int ora_t1_decode_block(int w, int h, int band, int mb, int zbp, int passes,
                        const uint8_t *data, size_t len, int32_t *out, char *err);

/* PCRD (dwt/pcrd.rs) */
typedef struct {
    int passes;
    uint32_t bytes;
    double dist;   /* cumulative distortion reduction */
    double slope;
} ora_rd_point;
/* RawPassRecord (pcrd.rs:33-50) */
typedef struct {
    int pass_index;
    uint32_t bytes, cumulative_bytes;
    double distortion_delta;
} ora_raw_record;
/* PcrdError variants the path can raise (pcrd.rs:872-935) */
enum {
    ORA_PCRD_OK = 0,
    ORA_PCRD_NONMONOTONE_CUMULATIVE_BYTES = 1,
    ORA_PCRD_INCONSISTENT_INCREMENTAL_BYTES = 2,
    ORA_PCRD_INVALID_DISTORTION_DELTA = 3,
    ORA_PCRD_INVALID_ORIGIN = 4,
    ORA_PCRD_NONMONOTONE_PASSES = 5,
    ORA_PCRD_NONMONOTONE_POINT_BYTES = 6,
    ORA_PCRD_NONMONOTONE_DISTORTION = 7,
    ORA_PCRD_EMPTY_CURVE = 8,
};
int ora_pcrd_raw_curve(const ora_raw_record *recs, int n, ora_rd_point *out);
int ora_pcrd_prune_hull(const ora_rd_point *raw, int n, ora_rd_point *out, int *m);
int ora_pcrd_choose_point(const ora_rd_point *pts, int n, double lambda, int *index);
double ora_pass_delta_pass_kind_aware(int kind, int bitplane, int newsig, int refine, double step, double sw,
                                      int band, int quality, int cls);
double ora_pass_delta_baseline(int bitplane, int newsig, int refine, double step, double sw);
double ora_quality_to_lambda(int quality, uint32_t pixels);
double ora_calibrate_lambda(double max_slope, double raw_lambda, int quality, uint32_t pixels);
int ora_pcrd_apply(ora_encoding *enc, const ora_plan *plan);

/* timing-faithful dead work (ora_perceptual.c): the contrast-mask map and
 * plan_coding_passes, run by ora_encode when faithful mode is on */
void ora_set_faithful(int on);
int ora_faithful(void);
double ora_contrast_mask(const int32_t *const *planes, int ncomp, int colorspace, uint32_t w, uint32_t h);
long ora_plan_coding_passes(const int32_t *coef, int n, int mbp);
/* host threads of the OpenMP regions (0 = the runtime default) */
void ora_set_threads(int n);

/* codec */
int ora_encode(const int32_t *const *planes, uint32_t w, uint32_t h, int ncomp, int colorspace,
               int quality, int format, ora_encoding *enc);
void ora_encoding_free(ora_encoding *enc);
/* encode_with_psnr metrics (backend.rs:297-433, :909-959) */
int ora_metric_recon(const int32_t *const *planes, const ora_encoding *enc, float *recon);
void ora_metrics_from_recon(const int32_t *const *orig, const float *recon, uint32_t w, uint32_t h, int ncomp,
                            double *psnr, double *ssim);
/* content analysis (content.rs:96-174) */
void ora_analyze_content(const int32_t *const *planes, size_t n, int ncomp, int colorspace, float *out3);
int ora_classify_content(const float *stats3);
int ora_quality_metrics(const int32_t *const *planes, int colorspace, const ora_encoding *enc, double *psnr,
                        double *ssim, char *err);
int ora_decode(const uint8_t *bytes, size_t len, uint32_t *w, uint32_t *h, int *ncomp,
               int *colorspace, int32_t **planes, char *err);

#endif
