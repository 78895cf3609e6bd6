/* Oracle (test infrastructure only): MQ coder/decoder and EBCOT Tier-1.
 * Restates mq/mod.rs:30-432, tier1/flags.rs:33-296, tier1/helpers.rs,
 * encode/backend/native/t1.rs:166-258, :508-658, :805-1197 and
 * decode/t1.rs:87-466. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

/* ---- MQ state machine ------------------------------------------------------
 * ISO/IEC 15444-1 Table C.2 (Qe, NMPS, NLPS, SWITCH).  The reference stores
 * the same machine expanded to 94 (state, mps) entries (mq/mod.rs:435-530);
 * ora_mq_state_table() produces that expansion for the checksum test. */
static const uint16_t QE[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801, 0x3001, 0x2401,
    0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401, 0x3001, 0x2801, 0x2401, 0x2201,
    0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101, 0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1,
    0x0221, 0x0141, 0x0111, 0x0085, 0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
static const uint8_t NMPS[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
static const uint8_t NLPS[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                 15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
static const uint8_t SWITCH[47] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                   0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

static uint16_t st_qe(uint8_t s) { return QE[s >> 1]; }
static uint8_t st_mps(uint8_t s) { return s & 1; }
static uint8_t st_nmps(uint8_t s) { return (uint8_t)(2 * NMPS[s >> 1] + (s & 1)); }
static uint8_t st_nlps(uint8_t s) { return (uint8_t)(2 * NLPS[s >> 1] + ((s & 1) ^ SWITCH[s >> 1])); }

int ora_mq_state_table(uint16_t *qe, uint8_t *mps, uint8_t *nmps, uint8_t *nlps) {
    for (int s = 0; s < 94; s++) {
        qe[s] = st_qe((uint8_t)s);
        mps[s] = st_mps((uint8_t)s);
        nmps[s] = st_nmps((uint8_t)s);
        nlps[s] = st_nlps((uint8_t)s);
    }
    return 94;
}

/* context numbers (mq/mod.rs:5-11) */
#define CTX_ZC 0
#define CTX_SC 9
#define CTX_MAG 14
#define CTX_AGG 17
#define CTX_UNI 18

/* mq/mod.rs:195-200 reset: UNI -> state 46, AGG -> 3, ZC -> 4 (mps 0) */
static void reset_ctx(uint8_t *ctx) {
    memset(ctx, 0, 19);
    ctx[CTX_UNI] = 2 * 46;
    ctx[CTX_AGG] = 2 * 3;
    ctx[CTX_ZC] = 2 * 4;
}

/* mq/mod.rs:171-183 */
void ora_mqc_init(ora_mqc *m, uint8_t *buf, size_t cap) {
    m->a = 0x8000;
    m->c = 0;
    m->ct = 12;
    m->out = buf;
    m->cap = cap;
    m->pos = 0;
    m->out[0] = 0;
    m->nsym = 0;
    reset_ctx(m->ctx);
}

static void write_next(ora_mqc *m, uint8_t b) {
    m->pos++;
    if (m->pos >= m->cap) {
        fprintf(stderr, "oracle MQ buffer overflow\n");
        abort();
    }
    m->out[m->pos] = b;
}

/* mq/mod.rs:381-402 */
static void byteout(ora_mqc *m) {
    if (m->out[m->pos] == 0xff) {
        write_next(m, (uint8_t)(m->c >> 20));
        m->c &= 0xfffff;
        m->ct = 7;
    } else if ((m->c & 0x8000000) == 0) {
        write_next(m, (uint8_t)(m->c >> 19));
        m->c &= 0x7ffff;
        m->ct = 8;
    } else {
        m->out[m->pos]++;
        if (m->out[m->pos] == 0xff) {
            m->c &= 0x7ffffff;
            write_next(m, (uint8_t)(m->c >> 20));
            m->c &= 0xfffff;
            m->ct = 7;
        } else {
            write_next(m, (uint8_t)(m->c >> 19));
            m->c &= 0x7ffff;
            m->ct = 8;
        }
    }
}

/* mq/mod.rs:362-371 */
static void renorme(ora_mqc *m) {
    while ((m->a & 0x8000) == 0) {
        m->a <<= 1;
        m->c <<= 1;
        m->ct--;
        if (m->ct == 0) byteout(m);
    }
}

/* mq/mod.rs:213-242 */
void ora_mqc_encode(ora_mqc *m, int ctx, int bit) {
    m->nsym++;
    uint8_t s = m->ctx[ctx];
    uint32_t qe = st_qe(s);
    if (bit == st_mps(s)) {
        m->a -= qe;
        if ((m->a & 0x8000) == 0) {
            if (m->a < qe) m->a = qe;
            else m->c += qe;
            m->ctx[ctx] = st_nmps(s);
            renorme(m);
        } else {
            m->c += qe;
        }
    } else {
        m->a -= qe;
        if (m->a < qe) m->c += qe;
        else m->a = qe;
        m->ctx[ctx] = st_nlps(s);
        renorme(m);
    }
}

/* mq/mod.rs:244-250 (flush), :373-379 (setbits), :416-423 (take_bytes:
 * a trailing 0xFF is dropped).  Returns the terminated segment length. */
size_t ora_mqc_flush(ora_mqc *m, uint8_t **bytes) {
    uint32_t temp = m->c + m->a;
    m->c |= 0xffff;
    if (m->c >= temp) m->c -= 0x8000;
    m->c <<= m->ct;
    byteout(m);
    m->c <<= m->ct;
    byteout(m);
    size_t end = m->out[m->pos] == 0xff ? m->pos : m->pos + 1;
    *bytes = m->out + 1;
    return end - 1;
}

/* mq/mod.rs:44-58 */
static uint8_t byte_at(const ora_mqd *d, size_t i) { return i < d->len ? d->bytes[i] : 0xff; }
static void bytein(ora_mqd *d) {
    uint8_t next = byte_at(d, d->bp + 1);
    if (byte_at(d, d->bp) == 0xff) {
        if (next > 0x8f) {
            d->c += 0xff00;
            d->ct = 8;
        } else {
            d->bp++;
            d->c += (uint32_t)next << 9;
            d->ct = 7;
        }
    } else {
        d->bp++;
        d->c += (uint32_t)next << 8;
        d->ct = 8;
    }
}
void ora_mqd_init(ora_mqd *d, const uint8_t *bytes, size_t len) {
    d->bytes = bytes;
    d->len = len;
    d->bp = 0;
    d->a = 0x8000;
    d->c = len == 0 ? (0xffu << 16) : ((uint32_t)bytes[0] << 16);
    d->ct = 0;
    reset_ctx(d->ctx);
    bytein(d);
    d->c <<= 7;
    d->ct = d->ct >= 7 ? d->ct - 7 : 0;
}
/* mq/mod.rs:61-81, :97-131 */
int ora_mqd_decode(ora_mqd *d, int ctx) {
    uint8_t s = d->ctx[ctx];
    uint32_t qe = st_qe(s);
    int bit;
    d->a -= qe;
    if ((d->c >> 16) < qe) {
        uint32_t post = d->a;
        d->a = qe;
        if (post < qe) {
            d->ctx[ctx] = st_nmps(s);
            bit = st_mps(s);
        } else {
            d->ctx[ctx] = st_nlps(s);
            bit = 1 ^ st_mps(s);
        }
    } else {
        d->c -= qe << 16;
        if ((d->a & 0x8000) == 0) {
            if (d->a < qe) {
                d->ctx[ctx] = st_nlps(s);
                bit = 1 ^ st_mps(s);
            } else {
                d->ctx[ctx] = st_nmps(s);
                bit = st_mps(s);
            }
        } else {
            bit = st_mps(s);
        }
    }
    if ((d->a & 0x8000) == 0) {
        while (d->a < 0x8000) {
            if (d->ct == 0) bytein(d);
            d->a <<= 1;
            d->c <<= 1;
            d->ct--;
        }
    }
    return bit;
}

/* ---- context lookup tables ------------------------------------------------
 * Generated from ISO/IEC 15444-1 Tables D.1 (zero coding), D.3 (sign
 * coding) in the neighbour-bit layout of tier1/consts.rs:1-39:
 *   sigma bits NW=0 N=1 NE=2 W=3 (THIS=4) E=5 SW=6 S=7 SE=8,
 *   sign index bits SGN_W=0 SIG_N=1 SGN_E=2 SIG_W=3 SGN_N=4 SIG_E=5
 *   SGN_S=6 SIG_S=7.
 * The reference ships the same tables as literals (tier1/luts.rs); the
 * checksum test compares the two. */
static int zc_ctx(int orient, int h, int v, int d) {
    if (orient == ORA_HL) { int t = h; h = v; v = t; }
    if (orient == ORA_HH) {
        int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv >= 1 ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv == 1 ? 4 : 3);
        return hv >= 2 ? 2 : (hv == 1 ? 1 : 0);
    }
    if (h == 2) return 8;
    if (h == 1) return v >= 1 ? 7 : (d >= 1 ? 6 : 5);
    if (v == 2) return 4;
    if (v == 1) return 3;
    return d >= 2 ? 2 : (d == 1 ? 1 : 0);
}

void ora_t1_luts(uint8_t zc[4][512], uint8_t sc[256], uint8_t spb[256]) {
    for (int o = 0; o < 4; o++)
        for (int i = 0; i < 512; i++) {
            int nw = i & 1, n = (i >> 1) & 1, ne = (i >> 2) & 1, w = (i >> 3) & 1;
            int e = (i >> 5) & 1, sw = (i >> 6) & 1, s = (i >> 7) & 1, se = (i >> 8) & 1;
            zc[o][i] = (uint8_t)zc_ctx(o, w + e, n + s, nw + ne + sw + se);
        }
    for (int i = 0; i < 256; i++) {
        int sgn_w = i & 1, sig_n = (i >> 1) & 1, sgn_e = (i >> 2) & 1, sig_w = (i >> 3) & 1;
        int sgn_n = (i >> 4) & 1, sig_e = (i >> 5) & 1, sgn_s = (i >> 6) & 1, sig_s = (i >> 7) & 1;
        int hc = (sig_w ? (sgn_w ? -1 : 1) : 0) + (sig_e ? (sgn_e ? -1 : 1) : 0);
        int vc = (sig_n ? (sgn_n ? -1 : 1) : 0) + (sig_s ? (sgn_s ? -1 : 1) : 0);
        int h = hc > 0 ? 1 : (hc < 0 ? -1 : 0), v = vc > 0 ? 1 : (vc < 0 ? -1 : 0);
        int ctx, xr = 0;
        if (h == 0 && v == 0) ctx = 9;
        else if (h == 0) { ctx = 10; xr = v < 0; }
        else {
            int hv = h * v; /* H=1: V=1,0,-1 -> 13,12,11 ; H=-1 mirrors with xor */
            ctx = hv > 0 ? 13 : (v == 0 ? 12 : 11);
            xr = h < 0;
        }
        sc[i] = (uint8_t)ctx;
        spb[i] = (uint8_t)xr;
    }
}

static uint8_t LUT_ZC[4][512], LUT_SC[256], LUT_SPB[256];
/* Filled once when the library loads, before any thread can code a block. */
__attribute__((constructor)) static void init_luts(void) { ora_t1_luts(LUT_ZC, LUT_SC, LUT_SPB); }
static void ensure_luts(void) {}

/* ---- flag grid (tier1/flags.rs:11-296) ----------------------------------- */
#define F_SIG (1u << 0)
#define F_VIS (1u << 1)
#define F_REF (1u << 2)
#define F_SGN (1u << 3)
#define N_W (1u << 4)
#define N_E (1u << 5)
#define N_N (1u << 6)
#define N_S (1u << 7)
#define N_NW (1u << 8)
#define N_NE (1u << 9)
#define N_SW (1u << 10)
#define N_SE (1u << 11)
#define N_ANY (N_W | N_E | N_N | N_S | N_NW | N_NE | N_SW | N_SE)

typedef struct {
    int w, h;
    uint16_t *f;
} grid;

static void mark_sig(grid *g, int x, int y, int sgn) {
    int w = g->w, h = g->h, i = y * w + x;
    g->f[i] |= F_SIG;
    if (sgn) g->f[i] |= F_SGN;
    else g->f[i] &= (uint16_t)~F_SGN;
    if (x > 0) g->f[i - 1] |= N_E;
    if (x + 1 < w) g->f[i + 1] |= N_W;
    if (y > 0) g->f[i - w] |= N_S;
    if (y + 1 < h) g->f[i + w] |= N_N;
    if (x > 0 && y > 0) g->f[i - w - 1] |= N_SE;
    if (x + 1 < w && y > 0) g->f[i - w + 1] |= N_SW;
    if (x > 0 && y + 1 < h) g->f[i + w - 1] |= N_NE;
    if (x + 1 < w && y + 1 < h) g->f[i + w + 1] |= N_NW;
}

/* flags.rs:178-191: internal neighbour bits -> T1_SIGMA layout */
static uint32_t nmask(const grid *g, int x, int y) {
    uint16_t f = g->f[y * g->w + x];
    uint32_t m = 0;
    if (f & N_NW) m |= 1u << 0;
    if (f & N_N) m |= 1u << 1;
    if (f & N_NE) m |= 1u << 2;
    if (f & N_W) m |= 1u << 3;
    if (f & N_E) m |= 1u << 5;
    if (f & N_SW) m |= 1u << 6;
    if (f & N_S) m |= 1u << 7;
    if (f & N_SE) m |= 1u << 8;
    return m;
}

/* flags.rs:202-215 */
static int stripe_clean(const grid *g, int x, int k) {
    for (int ci = 0; ci < 4; ci++) {
        int y = k + ci;
        if (y >= g->h) break;
        if (g->f[y * g->w + x] & (F_SIG | F_VIS | N_ANY)) return 0;
    }
    return 1;
}

/* flags.rs:217-296 + helpers.rs:53-89: cardinal sign-context index */
static uint32_t sign_index(const grid *g, int x, int y) {
    uint32_t lu = 0;
    int w = g->w, h = g->h;
    if (x > 0) {
        uint16_t f = g->f[y * w + x - 1];
        if (f & F_SIG) { lu |= 1u << 3; if (f & F_SGN) lu |= 1u << 0; }
    }
    if (y > 0) {
        uint16_t f = g->f[(y - 1) * w + x];
        if (f & F_SIG) { lu |= 1u << 1; if (f & F_SGN) lu |= 1u << 4; }
    }
    if (x + 1 < w) {
        uint16_t f = g->f[y * w + x + 1];
        if (f & F_SIG) { lu |= 1u << 5; if (f & F_SGN) lu |= 1u << 2; }
    }
    if (y + 1 < h) {
        uint16_t f = g->f[(y + 1) * w + x];
        if (f & F_SIG) { lu |= 1u << 7; if (f & F_SGN) lu |= 1u << 6; }
    }
    return lu;
}

static void clear_visited_all(grid *g) {
    for (int i = 0; i < g->w * g->h; i++) g->f[i] &= (uint16_t)~F_VIS;
}

/* helpers.rs:36-50 */
static int mag_ctx(int has_nb, int refined) { return refined ? CTX_MAG + 2 : (has_nb ? CTX_MAG + 1 : CTX_MAG); }

/* ---- encoder ---------------------------------------------------------------- */
typedef struct {
    ora_mqc *mq;
    grid *g;
    const int32_t *coef;
    int band;
    uint32_t one;
    int newsig;
} encst;

static uint32_t absmag(int32_t v) { return v < 0 ? (uint32_t)(-(int64_t)v) : (uint32_t)v; }

/* t1.rs:861-905 */
static void sp_step(encst *e, int x, int y) {
    grid *g = e->g;
    int i = y * g->w + x;
    if (g->f[i] & F_SIG) return;
    uint32_t nb = nmask(g, x, y);
    if (!nb) return;
    int32_t c = e->coef[i];
    int bit = (absmag(c) & e->one) != 0;
    ora_mqc_encode(e->mq, LUT_ZC[e->band][nb], bit);
    g->f[i] |= F_VIS;
    if (bit) {
        int sgn = c < 0;
        uint32_t si = sign_index(g, x, y);
        ora_mqc_encode(e->mq, LUT_SC[si], sgn ^ LUT_SPB[si]);
        mark_sig(g, x, y, sgn);
        e->newsig++;
    }
}

/* t1.rs:914-965 */
static void mr_step(encst *e, int x, int y) {
    grid *g = e->g;
    int i = y * g->w + x;
    if (!(g->f[i] & F_SIG) || (g->f[i] & F_VIS)) return;
    int ctx = mag_ctx(nmask(g, x, y) != 0, (g->f[i] & F_REF) != 0);
    int bit = (absmag(e->coef[i]) & e->one) != 0;
    ora_mqc_encode(e->mq, ctx, bit);
    g->f[i] |= F_REF;
}

/* t1.rs:1159-1197 */
static void cl_regular(encst *e, int x, int y) {
    grid *g = e->g;
    int i = y * g->w + x;
    if ((g->f[i] & F_SIG) || (g->f[i] & F_VIS)) {
        g->f[i] &= (uint16_t)~F_VIS;
        return;
    }
    uint32_t nb = nmask(g, x, y);
    int ctx = nb ? LUT_ZC[e->band][nb] : CTX_ZC;
    int32_t c = e->coef[i];
    int bit = (absmag(c) & e->one) != 0;
    ora_mqc_encode(e->mq, ctx, bit);
    g->f[i] &= (uint16_t)~F_VIS;
    if (bit) {
        int sgn = c < 0;
        uint32_t si = sign_index(g, x, y);
        ora_mqc_encode(e->mq, LUT_SC[si], sgn ^ LUT_SPB[si]);
        mark_sig(g, x, y, sgn);
        e->newsig++;
    }
}

static void sp_pass(encst *e) {
    grid *g = e->g;
    for (int k = 0; k < g->h; k += 4)
        for (int x = 0; x < g->w; x++)
            for (int ci = 0; ci < 4 && k + ci < g->h; ci++) sp_step(e, x, k + ci);
}
static void mr_pass(encst *e) {
    grid *g = e->g;
    for (int k = 0; k < g->h; k += 4)
        for (int x = 0; x < g->w; x++)
            for (int ci = 0; ci < 4 && k + ci < g->h; ci++) mr_step(e, x, k + ci);
}
/* t1.rs:977-1157: full stripes may use run mode, the partial stripe never */
static void cl_pass(encst *e) {
    grid *g = e->g;
    int full = (g->h / 4) * 4;
    for (int k = 0; k < full; k += 4) {
        for (int x = 0; x < g->w; x++) {
            if (stripe_clean(g, x, k)) {
                int rl = -1;
                for (int ci = 0; ci < 4; ci++)
                    if (absmag(e->coef[(k + ci) * g->w + x]) & e->one) { rl = ci; break; }
                if (rl >= 0) {
                    ora_mqc_encode(e->mq, CTX_AGG, 1);
                    ora_mqc_encode(e->mq, CTX_UNI, (rl >> 1) & 1);
                    ora_mqc_encode(e->mq, CTX_UNI, rl & 1);
                    int y = k + rl;
                    int32_t c = e->coef[y * g->w + x];
                    int sgn = c < 0;
                    uint32_t si = sign_index(g, x, y);
                    ora_mqc_encode(e->mq, LUT_SC[si], sgn ^ LUT_SPB[si]);
                    mark_sig(g, x, y, sgn);
                    e->newsig++;
                    g->f[y * g->w + x] &= (uint16_t)~F_VIS;
                    for (int ci = rl + 1; ci < 4; ci++) cl_regular(e, x, k + ci);
                } else {
                    ora_mqc_encode(e->mq, CTX_AGG, 0);
                    for (int ci = 0; ci < 4; ci++) g->f[(k + ci) * g->w + x] &= (uint16_t)~F_VIS;
                }
            } else {
                for (int ci = 0; ci < 4; ci++) cl_regular(e, x, k + ci);
            }
        }
    }
    if (full < g->h)
        for (int x = 0; x < g->w; x++)
            for (int y = full; y < g->h; y++) cl_regular(e, x, y);
}

/* t1.rs:215-258 (analysis) + t1.rs:508-658 (single-termination encode).
 * Fills blk->mbp, nnz, max_magnitude, zbp, cls, passes, bytes.  `mb` is the
 * subband's Mb (t1.rs:19-22, backend.rs:190-205). */
int ora_t1_encode_block(const int32_t *coef, int w, int h, int band, int mb, ora_block *blk) {
    ensure_luts();
    int n = w * h;
    uint32_t mx = 0;
    int nnz = 0;
    for (int i = 0; i < n; i++) {
        uint32_t m = absmag(coef[i]);
        if (m > mx) mx = m;
        if (coef[i]) nnz++;
    }
    blk->max_magnitude = mx;
    blk->nnz = nnz;
    blk->mbp = mx ? 32 - __builtin_clz(mx) : 0;
    blk->zbp = mb > blk->mbp ? mb - blk->mbp : 0;
    float frac = n ? (float)nnz / (float)n : 0.0f; /* profile.rs:122-130 */
    blk->cls = frac <= 0.05f ? 0 : (frac <= 0.30f ? 1 : 2);
    blk->npasses = 0;
    blk->full_npasses = 0;
    blk->bytes = NULL;
    blk->nbytes = 0;
    if (blk->mbp == 0) {
        blk->cls = 0; /* t1.rs:320-335: empty blocks report Flat */
        return ORA_OK;
    }
    size_t cap = (size_t)n * 4 + 1024;
    uint8_t *buf = malloc(cap);
    grid g = {w, h, calloc((size_t)n, sizeof(uint16_t))};
    ora_mqc mq;
    ora_mqc_init(&mq, buf, cap);
    encst e = {&mq, &g, coef, band, 0, 0};
    int np = 0, sig_so_far = 0;
    size_t snaps[ORA_MAX_PASSES];
    int top = blk->mbp - 1;
    for (int bp = top; bp >= 0; bp--) {
        e.one = 1u << bp;
        int before_sp = sig_so_far;
        if (bp != top) {
            e.newsig = 0;
            sp_pass(&e);
            blk->passes[np] = (ora_pass){ORA_SP, bp, e.newsig, before_sp, 0, 0, 0};
            snaps[np++] = mq.pos;
            sig_so_far += e.newsig;
            mr_pass(&e);
            blk->passes[np] = (ora_pass){ORA_MR, bp, 0, before_sp, 0, 0, 0};
            snaps[np++] = mq.pos;
        }
        e.newsig = 0;
        cl_pass(&e);
        blk->passes[np] = (ora_pass){ORA_CL, bp, e.newsig, bp == top ? 0 : sig_so_far, 0, 0, 0};
        snaps[np++] = mq.pos;
        sig_so_far += e.newsig;
        clear_visited_all(&g);
    }
    uint8_t *seg;
    blk->nsym = (uint32_t)mq.nsym;
    size_t total = ora_mqc_flush(&mq, &seg);
    /* t1.rs:586-600: lengths from pos snapshots, last pass absorbs the rest */
    size_t prev = 0, sum = 0, cum = 0;
    for (int i = 0; i < np; i++) {
        size_t len = snaps[i] > prev ? snaps[i] - prev : 0;
        prev = snaps[i];
        if (i == np - 1) len = total > sum ? total - sum : 0;
        sum += len;
        cum += len;
        blk->passes[i].length = (int)len;
        blk->passes[i].cumulative_length = (int)cum;
        blk->passes[i].is_terminated = i == np - 1;
        blk->full_lengths[i] = (int)len;
    }
    blk->npasses = np;
    blk->full_npasses = np;
    blk->bytes = malloc(total ? total : 1);
    memcpy(blk->bytes, seg, total);
    blk->nbytes = (int)total;
    free(buf);
    free(g.f);
    return ORA_OK;
}

/* ---- decoder (decode/t1.rs:87-466) ------------------------------------------ */
typedef struct {
    ora_mqd *mq;
    grid *g;
    uint32_t *mag;
    uint8_t *sgn;
    int band, bp;
} decst;

static void dec_newsig(decst *d, int x, int y) {
    uint32_t si = sign_index(d->g, x, y);
    int s = ora_mqd_decode(d->mq, LUT_SC[si]) ^ LUT_SPB[si];
    int i = y * d->g->w + x;
    d->mag[i] |= 1u << d->bp;
    d->sgn[i] = (uint8_t)s;
    mark_sig(d->g, x, y, s);
}
static void dec_sp_step(decst *d, int x, int y) {
    grid *g = d->g;
    int i = y * g->w + x;
    if (g->f[i] & F_SIG) return;
    uint32_t nb = nmask(g, x, y);
    if (!nb) return;
    int bit = ora_mqd_decode(d->mq, LUT_ZC[d->band][nb]);
    g->f[i] |= F_VIS;
    if (bit) dec_newsig(d, x, y);
}
static void dec_mr_step(decst *d, int x, int y) {
    grid *g = d->g;
    int i = y * g->w + x;
    if (!(g->f[i] & F_SIG) || (g->f[i] & F_VIS)) return;
    int ctx = mag_ctx(nmask(g, x, y) != 0, (g->f[i] & F_REF) != 0);
    if (ora_mqd_decode(d->mq, ctx)) d->mag[i] |= 1u << d->bp;
    g->f[i] |= F_REF;
}
static void dec_cl_regular(decst *d, int x, int y) {
    grid *g = d->g;
    int i = y * g->w + x;
    if ((g->f[i] & F_SIG) || (g->f[i] & F_VIS)) {
        g->f[i] &= (uint16_t)~F_VIS;
        return;
    }
    uint32_t nb = nmask(g, x, y);
    int bit = ora_mqd_decode(d->mq, nb ? LUT_ZC[d->band][nb] : CTX_ZC);
    g->f[i] &= (uint16_t)~F_VIS;
    if (bit) dec_newsig(d, x, y);
}
static void dec_sp(decst *d) {
    grid *g = d->g;
    for (int k = 0; k < g->h; k += 4)
        for (int x = 0; x < g->w; x++)
            for (int ci = 0; ci < 4 && k + ci < g->h; ci++) dec_sp_step(d, x, k + ci);
}
static void dec_mr(decst *d) {
    grid *g = d->g;
    for (int k = 0; k < g->h; k += 4)
        for (int x = 0; x < g->w; x++)
            for (int ci = 0; ci < 4 && k + ci < g->h; ci++) dec_mr_step(d, x, k + ci);
}
static void dec_cl(decst *d) {
    grid *g = d->g;
    int full = (g->h / 4) * 4;
    for (int k = 0; k < full; k += 4) {
        for (int x = 0; x < g->w; x++) {
            if (stripe_clean(g, x, k)) {
                if (ora_mqd_decode(d->mq, CTX_AGG) == 0) {
                    for (int ci = 0; ci < 4; ci++) g->f[(k + ci) * g->w + x] &= (uint16_t)~F_VIS;
                    continue;
                }
                int hi = ora_mqd_decode(d->mq, CTX_UNI);
                int lo = ora_mqd_decode(d->mq, CTX_UNI);
                int rl = (hi << 1) | lo;
                dec_newsig(d, x, k + rl);
                g->f[(k + rl) * g->w + x] &= (uint16_t)~F_VIS;
                for (int ci = rl + 1; ci < 4; ci++) dec_cl_regular(d, x, k + ci);
            } else {
                for (int ci = 0; ci < 4; ci++) dec_cl_regular(d, x, k + ci);
            }
        }
    }
    if (full < g->h)
        for (int x = 0; x < g->w; x++)
            for (int y = full; y < g->h; y++) dec_cl_regular(d, x, y);
}

int ora_t1_decode_block(int w, int h, int band, int mb, int zbp, int passes,
                        const uint8_t *data, size_t len, int32_t *out, char *err) {
    ensure_luts();
    int n = w * h;
    if (passes == 0) {
        memset(out, 0, sizeof(int32_t) * (size_t)n);
        return ORA_OK;
    }
    if (len == 0) {
        snprintf(err, 256, "non-empty code-block has no compressed bytes");
        return ORA_DECODE_FAILED;
    }
    if (zbp > 255) {
        snprintf(err, 256, "zero-bitplane count exceeds supported range");
        return ORA_DECODE_FAILED;
    }
    if (zbp > mb) {
        snprintf(err, 256, "zero-bitplane count exceeds subband bit-plane count");
        return ORA_DECODE_FAILED;
    }
    int mbp = mb - zbp;
    if (mbp == 0) {
        snprintf(err, 256, "non-empty code-block has no magnitude bit-planes");
        return ORA_DECODE_FAILED;
    }
    int maxp = 1 + 3 * (mbp - 1);
    if (passes > maxp) {
        snprintf(err, 256, "code-block pass count %d exceeds maximum %d for %d magnitude bit-planes", passes, maxp, mbp);
        return ORA_DECODE_FAILED;
    }
    ora_mqd mq;
    ora_mqd_init(&mq, data, len);
    grid g = {w, h, calloc((size_t)n, sizeof(uint16_t))};
    uint32_t *mag = calloc((size_t)n, sizeof(uint32_t));
    uint8_t *sg = calloc((size_t)n, 1);
    decst d = {&mq, &g, mag, sg, band, mbp - 1};
    int rem = passes;
    dec_cl(&d);
    rem--;
    clear_visited_all(&g);
    for (int bp = mbp - 2; bp >= 0 && rem > 0; bp--) {
        d.bp = bp;
        dec_sp(&d);
        if (--rem == 0) break;
        dec_mr(&d);
        if (--rem == 0) break;
        dec_cl(&d);
        rem--;
        clear_visited_all(&g);
    }
    for (int i = 0; i < n; i++) out[i] = sg[i] ? -(int32_t)mag[i] : (int32_t)mag[i];
    free(g.f);
    free(mag);
    free(sg);
    return ORA_OK;
}
