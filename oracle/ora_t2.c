/* Oracle (test infrastructure only): Tier-2 packet headers (encode and
 * decode), tag trees, codestream markers, JP2 boxes.  Restates
 * encode/backend/native/t2.rs:20-470, j2k/markers.rs, j2k/write.rs:9-118,
 * jp2.rs:13-71, decode/t2.rs:75-660, decode/jp2_parse.rs,
 * decode/codestream.rs and j2k/decode_markers.rs:139-501. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ora_t2.h"

/* ---- growable byte buffer ------------------------------------------------ */
void obuf_put(obuf *b, const void *p, size_t n) {
    if (b->len + n > b->cap) {
        size_t nc = b->cap ? b->cap * 2 : 256;
        while (nc < b->len + n) nc *= 2;
        b->data = realloc(b->data, nc);
        b->cap = nc;
    }
    memcpy(b->data + b->len, p, n);
    b->len += n;
}
void obuf_u8(obuf *b, uint8_t v) { obuf_put(b, &v, 1); }
void obuf_u16(obuf *b, uint16_t v) {
    uint8_t t[2] = {(uint8_t)(v >> 8), (uint8_t)v};
    obuf_put(b, t, 2);
}
void obuf_u32(obuf *b, uint32_t v) {
    uint8_t t[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
    obuf_put(b, t, 4);
}

/* ---- bit writer (t2.rs:27-68) -------------------------------------------- */
typedef struct {
    obuf *out;
    uint32_t reg, ct;
} bio;
static void bio_byteout(bio *b) {
    b->reg = (b->reg << 8) & 0xffff;
    b->ct = b->reg == 0xff00 ? 7 : 8;
    obuf_u8(b->out, (uint8_t)(b->reg >> 8));
}
static void bio_putbit(bio *b, uint32_t bit) {
    if (b->ct == 0) bio_byteout(b);
    b->ct--;
    b->reg |= bit << b->ct;
}
static void bio_bits(bio *b, uint32_t v, int n) {
    for (int i = n - 1; i >= 0; i--) bio_putbit(b, (v >> i) & 1);
}
static void bio_flush(bio *b) {
    bio_byteout(b);
    if (b->ct == 7) bio_byteout(b);
}

/* ---- tag tree (ISO 15444-1 B.10.2; t2.rs:73-214, decode/t2.rs:480-570) ----
 * Quad tree stored level by level in raster order; the parent of node (x,y)
 * at level l is (x/2, y/2) at level l+1. */
typedef struct {
    int nnodes;
    int *parent;
    int *value, *low, *known;
} tagtree;

static void tgt_init(tagtree *t, int w, int h, int init_value) {
    int lw[32], lh[32], off[33], nl = 0, total = 0;
    lw[0] = w;
    lh[0] = h;
    for (;;) {
        int n = lw[nl] * lh[nl];
        off[nl] = total;
        total += n;
        lw[nl + 1] = (lw[nl] + 1) / 2;
        lh[nl + 1] = (lh[nl] + 1) / 2;
        nl++;
        if (n <= 1) break;
    }
    t->nnodes = total;
    t->parent = malloc(sizeof(int) * (size_t)total);
    t->value = malloc(sizeof(int) * (size_t)total);
    t->low = calloc((size_t)total, sizeof(int));
    t->known = calloc((size_t)total, sizeof(int));
    for (int l = 0; l < nl; l++)
        for (int y = 0; y < lh[l]; y++)
            for (int x = 0; x < lw[l]; x++) {
                int i = off[l] + y * lw[l] + x;
                t->parent[i] = l + 1 < nl ? off[l + 1] + (y / 2) * lw[l + 1] + x / 2 : -1;
                t->value[i] = init_value;
            }
}
static void tgt_free(tagtree *t) {
    free(t->parent);
    free(t->value);
    free(t->low);
    free(t->known);
}
static void tgt_set(tagtree *t, int leaf, int v) {
    for (int i = leaf; i >= 0; i = t->parent[i]) {
        if (t->value[i] <= v) break;
        t->value[i] = v;
    }
}
static void tgt_encode(tagtree *t, bio *b, int leaf, int threshold) {
    int stk[32], depth = 0, n = leaf;
    while (t->parent[n] >= 0) {
        stk[depth++] = n;
        n = t->parent[n];
    }
    int low = 0;
    for (;;) {
        if (low > t->low[n]) t->low[n] = low;
        else low = t->low[n];
        while (low < threshold) {
            if (low >= t->value[n]) {
                if (!t->known[n]) {
                    bio_putbit(b, 1);
                    t->known[n] = 1;
                }
                break;
            }
            bio_putbit(b, 0);
            low++;
        }
        t->low[n] = low;
        if (depth == 0) break;
        n = stk[--depth];
    }
}

/* t2.rs:223-243, Table B.4 */
static void put_numpasses(bio *b, uint32_t n) {
    if (n == 1) bio_bits(b, 0, 1);
    else if (n == 2) bio_bits(b, 2, 2);
    else if (n <= 5) bio_bits(b, 0xc | (n - 3), 4);
    else if (n <= 36) bio_bits(b, 0x1e0 | (n - 6), 9);
    else if (n <= 164) bio_bits(b, 0xff80 | (n - 37), 16);
    else bio_bits(b, 0xff80 | (164 - 37), 16);
}
static uint32_t floor_log2u(uint32_t v) { return v ? 31u - (uint32_t)__builtin_clz(v) : 0u; }

/* t2.rs:345-396: one packet header for the bands of (resolution, component);
 * blocks[] are the band's code-blocks in raster order with a grid of
 * cx x cy. */
void ora_packet_header(const ora_band_ref *bands, int nbands, obuf *out) {
    bio b = {out, 0, 8};
    bio_putbit(&b, 1);
    for (int bi = 0; bi < nbands; bi++) {
        const ora_band_ref *bd = &bands[bi];
        if (bd->nblocks == 0) continue;
        tagtree incl, imsb;
        tgt_init(&incl, bd->cx, bd->cy, 999);
        tgt_init(&imsb, bd->cx, bd->cy, 999);
        for (int k = 0; k < bd->nblocks; k++) {
            const ora_block *blk = bd->blocks[k];
            tgt_set(&imsb, k, blk->zbp);
            if (blk->npasses) tgt_set(&incl, k, 0);
        }
        for (int k = 0; k < bd->nblocks; k++) {
            const ora_block *blk = bd->blocks[k];
            tgt_encode(&incl, &b, k, 1);
            if (!blk->npasses) continue;
            tgt_encode(&imsb, &b, k, 999);
            put_numpasses(&b, (uint32_t)blk->npasses);
            /* t2.rs:245-276: one segment (single termination); Lblock starts at 3 */
            uint32_t seg_len = (uint32_t)blk->passes[blk->npasses - 1].cumulative_length;
            uint32_t seg_n = (uint32_t)blk->npasses;
            uint32_t inc = 0;
            if (seg_len) {
                uint32_t need = floor_log2u(seg_len) + 1, have = 3 + floor_log2u(seg_n);
                inc = need > have ? need - have : 0;
            }
            for (uint32_t i = 0; i < inc; i++) bio_putbit(&b, 1);
            bio_putbit(&b, 0);
            bio_bits(&b, seg_len, (int)(3 + inc + floor_log2u(seg_n)));
        }
        tgt_free(&incl);
        tgt_free(&imsb);
    }
    bio_flush(&b);
}

/* ---- markers (j2k/markers.rs) + framing (j2k/write.rs) ------------------- */
void ora_write_codestream(const ora_plan *p, const uint8_t *payload, size_t plen, obuf *o) {
    obuf_u16(o, 0xff4f); /* SOC */
    obuf_u16(o, 0xff51); /* SIZ */
    obuf_u16(o, (uint16_t)(38 + 3 * p->ncomp));
    obuf_u16(o, 0);
    obuf_u32(o, p->width);
    obuf_u32(o, p->height);
    obuf_u32(o, 0);
    obuf_u32(o, 0);
    obuf_u32(o, p->width);
    obuf_u32(o, p->height);
    obuf_u32(o, 0);
    obuf_u32(o, 0);
    obuf_u16(o, (uint16_t)p->ncomp);
    for (int c = 0; c < p->ncomp; c++) {
        obuf_u8(o, 7);
        obuf_u8(o, 1);
        obuf_u8(o, 1);
    }
    obuf_u16(o, 0xff52); /* COD */
    obuf_u16(o, 12);
    obuf_u8(o, 0);
    obuf_u8(o, 0);
    obuf_u16(o, 1);
    obuf_u8(o, (uint8_t)(p->use_mct ? 1 : 0));
    obuf_u8(o, (uint8_t)p->levels);
    obuf_u8(o, (uint8_t)(__builtin_ctz((unsigned)p->cbw) - 2));
    obuf_u8(o, (uint8_t)(__builtin_ctz((unsigned)p->cbh) - 2));
    obuf_u8(o, 0);
    obuf_u8(o, (uint8_t)(p->lossless ? 1 : 0));
    obuf_u16(o, 0xff5c); /* QCD */
    if (p->lossless) {
        obuf_u16(o, (uint16_t)(3 + p->nbands));
        obuf_u8(o, (uint8_t)(p->guard << 5));
        for (int i = 0; i < p->nbands; i++) obuf_u8(o, (uint8_t)(p->sq[i].exponent << 3));
    } else {
        obuf_u16(o, (uint16_t)(3 + 2 * p->nbands));
        obuf_u8(o, (uint8_t)((p->guard << 5) | 2));
        for (int i = 0; i < p->nbands; i++)
            obuf_u16(o, (uint16_t)(((uint16_t)p->sq[i].exponent << 11) | (p->sq[i].mantissa & 0x7ff)));
    }
    obuf_u16(o, 0xff90); /* SOT */
    obuf_u16(o, 10);
    obuf_u16(o, 0);
    obuf_u32(o, (uint32_t)(12 + 2 + plen));
    obuf_u8(o, 0);
    obuf_u8(o, 1);
    obuf_u16(o, 0xff93); /* SOD */
    obuf_put(o, payload, plen);
    obuf_u16(o, 0xffd9); /* EOC */
}

/* jp2.rs:13-50 */
void ora_wrap_jp2(const ora_plan *p, const uint8_t *cs, size_t len, obuf *o) {
    obuf_u32(o, 12);
    obuf_put(o, "jP  ", 4);
    obuf_u32(o, 0x0d0a870a);
    obuf_u32(o, 20);
    obuf_put(o, "ftyp", 4);
    obuf_put(o, "jp2 ", 4);
    obuf_u32(o, 0);
    obuf_put(o, "jp2 ", 4);
    obuf_u32(o, 45);
    obuf_put(o, "jp2h", 4);
    obuf_u32(o, 22);
    obuf_put(o, "ihdr", 4);
    obuf_u32(o, p->height);
    obuf_u32(o, p->width);
    obuf_u16(o, (uint16_t)p->ncomp);
    obuf_u8(o, 7);
    obuf_u8(o, 7);
    obuf_u8(o, 0);
    obuf_u8(o, 0);
    obuf_u32(o, 15);
    obuf_put(o, "colr", 4);
    obuf_u8(o, 1);
    obuf_u8(o, 0);
    obuf_u8(o, 0);
    obuf_u32(o, p->domain == ORA_GRAY ? 17 : 16);
    obuf_u32(o, (uint32_t)(len + 8));
    obuf_put(o, "jp2c", 4);
    obuf_put(o, cs, len);
}

/* ---- parsing ---------------------------------------------------------------- */
static uint32_t rd32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
static uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] << 8 | p[1]); }

#define FAIL(code, ...)                      \
    do {                                     \
        snprintf(err, 256, __VA_ARGS__);     \
        return code;                         \
    } while (0)

/* decode/jp2_parse.rs:42-258 (box walk, ihdr, colr) */
int ora_parse_jp2(const uint8_t *b, size_t n, ora_jp2_header *h, const uint8_t **cs, size_t *cslen, char *err) {
    size_t pos = 0;
    int idx = 0, have_hdr = 0, have_colr = 0;
    *cs = NULL;
    while (pos < n) {
        if (n - pos < 8) FAIL(ORA_DECODE_FAILED, "truncated JP2 box header");
        size_t start = pos;
        uint64_t lbox = rd32(b + pos);
        const uint8_t *type = b + pos + 4;
        pos += 8;
        size_t end;
        if (lbox == 0) end = n;
        else if (lbox == 1) {
            if (n - pos < 8) FAIL(ORA_DECODE_FAILED, "unexpected end of JP2 u64");
            uint64_t xl = ((uint64_t)rd32(b + pos) << 32) | rd32(b + pos + 4);
            pos += 8;
            if (xl < 16) FAIL(ORA_DECODE_FAILED, "invalid extended JP2 box length");
            end = start + (size_t)xl;
        } else if (lbox < 8) FAIL(ORA_DECODE_FAILED, "invalid JP2 box length below header size");
        else end = start + (size_t)lbox;
        if (end > n || end < pos) FAIL(ORA_DECODE_FAILED, "JP2 box extends past input");
        const uint8_t *pl = b + pos;
        size_t plen = end - pos;
        pos = end;
        if (idx == 0) {
            if (memcmp(type, "jP  ", 4) || plen != 4 || rd32(pl) != 0x0d0a870a) FAIL(ORA_DECODE_FAILED, "invalid JP2 signature box");
        } else if (idx == 1) {
            if (memcmp(type, "ftyp", 4)) FAIL(ORA_DECODE_FAILED, "JP2 file type box must follow signature box");
            if (plen < 12) FAIL(ORA_DECODE_FAILED, "JP2 file type box is too short");
            if (memcmp(pl, "jp2 ", 4)) FAIL(ORA_DECODE_FAILED, "JP2 brand is not jp2");
            int ok = 0;
            for (size_t k = 8; k + 4 <= plen; k += 4)
                if (!memcmp(pl + k, "jp2 ", 4)) ok = 1;
            if (!ok) FAIL(ORA_DECODE_FAILED, "JP2 file type lacks jp2 compatibility brand");
        } else if (!memcmp(type, "jp2h", 4)) {
            size_t q = 0;
            have_hdr = 0;
            have_colr = 0;
            while (q < plen) {
                if (plen - q < 8) FAIL(ORA_DECODE_FAILED, "truncated JP2 box header");
                size_t s2 = q;
                uint64_t l2 = rd32(pl + q);
                const uint8_t *t2 = pl + q + 4;
                q += 8;
                size_t e2;
                if (l2 == 0) e2 = plen;
                else if (l2 == 1) {
                    if (plen - q < 8) FAIL(ORA_DECODE_FAILED, "unexpected end of JP2 u64");
                    uint64_t xl = ((uint64_t)rd32(pl + q) << 32) | rd32(pl + q + 4);
                    q += 8;
                    if (xl < 16) FAIL(ORA_DECODE_FAILED, "invalid extended JP2 box length");
                    e2 = s2 + (size_t)xl;
                } else if (l2 < 8) FAIL(ORA_DECODE_FAILED, "invalid JP2 box length below header size");
                else e2 = s2 + (size_t)l2;
                if (e2 > plen || e2 < q) FAIL(ORA_DECODE_FAILED, "JP2 box extends past input");
                const uint8_t *p2 = pl + q;
                size_t n2 = e2 - q;
                q = e2;
                if (!memcmp(t2, "ihdr", 4)) {
                    if (n2 != 14) FAIL(ORA_DECODE_FAILED, "JP2 ihdr box must be 14 bytes");
                    if (p2[11] != 7) FAIL(ORA_DECODE_FAILED, "JP2 ihdr compression type is not JPEG 2000");
                    if (p2[10] == 0xff) FAIL(ORA_UNSUPPORTED, "unsupported JP2 feature: per-component bit depths require bpcc");
                    h->height = rd32(p2);
                    h->width = rd32(p2 + 4);
                    h->ncomp = rd16(p2 + 8);
                    h->bpc = (p2[10] & 0x7f) + 1;
                    have_hdr = 1;
                } else if (!memcmp(t2, "bpcc", 4)) {
                    FAIL(ORA_UNSUPPORTED, "unsupported JP2 feature: per-component bit-depth box (bpcc)");
                } else if (!memcmp(t2, "colr", 4) && !have_colr) {
                    if (n2 < 3) FAIL(ORA_DECODE_FAILED, "JP2 colr box is too short");
                    if (p2[0] != 1) FAIL(ORA_UNSUPPORTED, "only enumerated JP2 colorspaces are supported");
                    if (n2 != 7) FAIL(ORA_DECODE_FAILED, "enumerated JP2 colr box must be 7 bytes");
                    uint32_t e = rd32(p2 + 3);
                    if (e == 17) h->colorspace = ORA_GRAY;
                    else if (e == 16) h->colorspace = ORA_SRGB;
                    else FAIL(ORA_UNSUPPORTED, "unsupported JP2 EnumCS value %u", e);
                    have_colr = 1;
                } else if (!memcmp(t2, "pclr", 4)) {
                    FAIL(ORA_UNSUPPORTED, "unsupported JP2 feature: palette box (pclr)");
                } else if (!memcmp(t2, "cmap", 4)) {
                    FAIL(ORA_UNSUPPORTED, "unsupported JP2 feature: component mapping box (cmap)");
                } else if (!memcmp(t2, "cdef", 4)) {
                    FAIL(ORA_UNSUPPORTED, "unsupported JP2 feature: channel definition box (cdef)");
                }
            }
            if (!have_hdr) FAIL(ORA_DECODE_FAILED, "JP2 header lacks ihdr box");
            if (!have_colr) FAIL(ORA_DECODE_FAILED, "JP2 header lacks colr box");
        } else if (!memcmp(type, "jp2c", 4)) {
            if (*cs) FAIL(ORA_DECODE_FAILED, "unsupported JP2 layout: multiple contiguous codestream boxes");
            *cs = pl;
            *cslen = plen;
        }
        idx++;
    }
    if (idx == 0) FAIL(ORA_DECODE_FAILED, "missing JP2 signature box");
    if (idx == 1) FAIL(ORA_DECODE_FAILED, "missing JP2 file type box");
    if (!have_hdr) FAIL(ORA_DECODE_FAILED, "missing JP2 header box");
    if (!*cs) FAIL(ORA_DECODE_FAILED, "missing contiguous codestream box");
    return ORA_OK;
}

/* decode/codestream.rs:27-198 + j2k/decode_markers.rs:139-501 */
int ora_parse_codestream(const uint8_t *b, size_t n, ora_cs_header *hd, char *err) {
    memset(hd, 0, sizeof(*hd));
    size_t pos = 0;
    if (n < 2 || rd16(b) != 0xff4f) FAIL(ORA_DECODE_FAILED, "codestream did not start with SOC");
    pos = 2;
    int have_siz = 0, have_cod = 0, have_qcd = 0, tiles = 0;
    for (;;) {
        if (n - pos < 2) FAIL(ORA_DECODE_FAILED, "codestream ended before EOC");
        uint16_t m = rd16(b + pos);
        if (m == 0xff90) break;
        if (n - pos < 4) FAIL(ORA_DECODE_FAILED, "truncated marker segment");
        size_t L = rd16(b + pos + 2);
        if (L < 2) FAIL(ORA_DECODE_FAILED, "invalid marker length %zu for marker 0x%04x", L, m);
        if (pos + 2 + L > n) FAIL(ORA_DECODE_FAILED, "marker 0x%04x extended past end of codestream", m);
        const uint8_t *body = b + pos + 4;
        size_t blen = L - 2;
        if (m == 0xff51) {
            if (blen < 38) FAIL(ORA_DECODE_FAILED, "SIZ marker body is too short");
            hd->rsiz = rd16(body);
            hd->width = rd32(body + 2);
            hd->height = rd32(body + 6);
            hd->xo = rd32(body + 10);
            hd->yo = rd32(body + 14);
            hd->tw = rd32(body + 18);
            hd->th = rd32(body + 22);
            hd->txo = rd32(body + 26);
            hd->tyo = rd32(body + 30);
            hd->ncomp = rd16(body + 34);
            if (blen != 36 + 3 * (size_t)hd->ncomp || hd->ncomp > 16) FAIL(ORA_DECODE_FAILED, "SIZ component count mismatch");
            for (int c = 0; c < hd->ncomp; c++) {
                hd->prec[c] = (body[36 + 3 * c] & 0x7f) + 1;
                hd->sgnd[c] = body[36 + 3 * c] >> 7;
                hd->dx[c] = body[37 + 3 * c];
                hd->dy[c] = body[38 + 3 * c];
            }
            have_siz = 1;
        } else if (m == 0xff52) {
            if (blen < 10) FAIL(ORA_DECODE_FAILED, "COD marker body is too short");
            hd->scod = body[0];
            hd->prog = body[1];
            hd->layers = rd16(body + 2);
            hd->mct = body[4];
            hd->levels = body[5];
            if (body[6] > 8 || body[7] > 8) FAIL(ORA_DECODE_FAILED, "unsupported code-block size");
            hd->cbw = 1 << (body[6] + 2);
            hd->cbh = 1 << (body[7] + 2);
            hd->cbstyle = body[8];
            hd->transform = body[9];
            have_cod = 1;
        } else if (m == 0xff5c) {
            if (blen < 1) FAIL(ORA_DECODE_FAILED, "QCD marker body is too short");
            hd->qstyle = body[0] & 0x1f;
            hd->guard = body[0] >> 5;
            if (hd->qstyle == 0) {
                hd->nsteps = (int)(blen - 1);
                for (int i = 0; i < hd->nsteps && i < 128; i++) { hd->exp[i] = body[1 + i] >> 3; hd->mant[i] = 0; }
            } else {
                hd->nsteps = (int)((blen - 1) / 2);
                for (int i = 0; i < hd->nsteps && i < 128; i++) {
                    uint16_t v = rd16(body + 1 + 2 * i);
                    hd->exp[i] = (uint8_t)(v >> 11);
                    hd->mant[i] = v & 0x7ff;
                }
            }
            have_qcd = 1;
        } else if (m == 0xff64) {
            /* COM */
        } else {
            FAIL(ORA_UNSUPPORTED, "unsupported codestream marker 0x%04x in main header", m);
        }
        pos += 2 + L;
    }
    /* single tile-part */
    for (;;) {
        if (n - pos < 12) FAIL(ORA_DECODE_FAILED, "missing Psot in SOT segment");
        size_t sot = pos;
        size_t L = rd16(b + pos + 2);
        if (L != 10) FAIL(ORA_DECODE_FAILED, "invalid SOT length");
        hd->tile_index = rd16(b + pos + 4);
        uint32_t psot = rd32(b + pos + 6);
        hd->part_index = b[pos + 10];
        hd->total_parts = b[pos + 11];
        pos += 12;
        while (n - pos >= 2 && rd16(b + pos) != 0xff93) {
            uint16_t m = rd16(b + pos);
            if (m != 0xff64) FAIL(ORA_UNSUPPORTED, "unsupported codestream marker 0x%04x in tile-part header", m);
            size_t L2 = rd16(b + pos + 2);
            pos += 2 + L2;
        }
        if (n - pos < 2) FAIL(ORA_DECODE_FAILED, "missing SOD");
        pos += 2;
        size_t end = psot ? sot + psot : n - 2;
        if (end > n || end < pos) FAIL(ORA_DECODE_FAILED, "tile-part extends past end of codestream");
        if (tiles == 0) {
            hd->payload = b + pos;
            hd->payload_len = end - pos;
        }
        tiles++;
        pos = end;
        if (n - pos >= 2 && rd16(b + pos) == 0xff90) continue;
        break;
    }
    if (n - pos != 2 || rd16(b + pos) != 0xffd9) FAIL(ORA_DECODE_FAILED, "expected EOC at end of codestream");
    if (!have_siz) FAIL(ORA_DECODE_FAILED, "codestream main header lacks SIZ");
    if (!have_cod) FAIL(ORA_DECODE_FAILED, "codestream main header lacks COD");
    if (!have_qcd) FAIL(ORA_DECODE_FAILED, "codestream main header lacks QCD");
    if (tiles != 1) FAIL(ORA_DECODE_FAILED, "unsupported tiling: decoder currently supports one full-image tile in one tile-part");
    /* decode_markers.rs:348-501 */
    if (hd->tile_index != 0) FAIL(ORA_DECODE_FAILED, "only tile index 0 is supported");
    if (hd->part_index != 0) FAIL(ORA_DECODE_FAILED, "only first tile-part is supported");
    if (hd->total_parts > 1) FAIL(ORA_DECODE_FAILED, "unsupported tiling: multiple tile-parts are not implemented");
    if (hd->rsiz != 0) FAIL(ORA_DECODE_FAILED, "only Part 1 Rsiz=0 codestreams are supported");
    if (hd->xo || hd->yo || hd->txo || hd->tyo) FAIL(ORA_DECODE_FAILED, "non-zero image or tile origins are unsupported");
    if (hd->tw != hd->width || hd->th != hd->height) FAIL(ORA_DECODE_FAILED, "unsupported tiling: decoder currently supports only a single full-image tile");
    if (hd->ncomp != 1 && hd->ncomp != 3) FAIL(ORA_DECODE_FAILED, "unsupported component layout");
    for (int c = 0; c < hd->ncomp; c++) {
        if (hd->prec[c] != 8 || hd->sgnd[c]) FAIL(ORA_DECODE_FAILED, "unsupported sample precision");
        if (hd->dx[c] != 1 || hd->dy[c] != 1) FAIL(ORA_DECODE_FAILED, "subsampled components are unsupported");
    }
    if (hd->prog != 0) FAIL(ORA_DECODE_FAILED, "only LRCP progression is supported");
    uint32_t md = hd->width < hd->height ? hd->width : hd->height;
    uint32_t maxd = md <= 1 ? 0 : 31 - (uint32_t)__builtin_clz(md);
    if (hd->levels > maxd) FAIL(ORA_DECODE_FAILED, "COD decomposition levels exceed DWT limit");
    if (hd->levels > ORA_MAX_LEVELS * 5) FAIL(ORA_DECODE_FAILED, "COD decomposition levels exceed DWT limit");
    if (hd->layers == 0) FAIL(ORA_DECODE_FAILED, "COD must signal at least one layer");
    if (hd->layers != 1) FAIL(ORA_DECODE_FAILED, "unsupported quality layers: decoder currently supports exactly one layer");
    if (hd->ncomp == 1 && hd->mct) FAIL(ORA_DECODE_FAILED, "MCT is unsupported for grayscale decoder slice");
    if (hd->ncomp == 3 && !hd->mct) FAIL(ORA_DECODE_FAILED, "unsupported RGB codestream: expected MCT=1 for Archive.org sRGB profile");
    if (hd->scod & 7) FAIL(ORA_DECODE_FAILED, "unsupported packet syntax: precinct partitioning, SOP markers, and EPH markers are not implemented");
    if (hd->cbstyle) FAIL(ORA_DECODE_FAILED, "unsupported code-block style");
    if (hd->transform == 1 && hd->qstyle != 0) FAIL(ORA_DECODE_FAILED, "unsupported quantization: reversible 5/3 currently supports no-quantization QCD only");
    if (hd->transform == 0 && hd->qstyle == 0) FAIL(ORA_DECODE_FAILED, "unsupported quantization: irreversible 9/7 requires scalar quantization");
    if (hd->transform == 0 && hd->qstyle == 1) FAIL(ORA_DECODE_FAILED, "unsupported quantization: scalar-derived QCD is not implemented");
    if (hd->cbw * hd->cbh > 4096) FAIL(ORA_DECODE_FAILED, "COD code-block size exceeds Part 1 limits");
    if (hd->nsteps != 1 + 3 * hd->levels) FAIL(ORA_DECODE_FAILED, "QCD step count does not match decomposition levels");
    return ORA_OK;
}

/* ---- packet-header reading (decode/t2.rs:75-660) ------------------------- */
typedef struct {
    const uint8_t *b;
    size_t n, pos;
    uint32_t reg, ct;
    int prev_ff;
} bior;
static int bior_bit(bior *r, uint32_t *bit) {
    if (r->ct == 0) {
        if (r->pos >= r->n) return -1;
        uint8_t byte = r->b[r->pos++];
        r->reg = byte;
        r->ct = r->prev_ff ? 7 : 8;
        r->prev_ff = byte == 0xff;
    }
    r->ct--;
    *bit = (r->reg >> r->ct) & 1;
    return 0;
}
static int bior_bits(bior *r, int n, uint32_t *v) {
    *v = 0;
    for (int i = 0; i < n; i++) {
        uint32_t bit;
        if (bior_bit(r, &bit)) return -1;
        *v = (*v << 1) | bit;
    }
    return 0;
}
/* decode/t2.rs:518-566 */
static int tgt_decode(tagtree *t, bior *r, int leaf, int threshold, int *res) {
    int stk[32], depth = 0, n = leaf;
    while (t->parent[n] >= 0) {
        stk[depth++] = n;
        n = t->parent[n];
    }
    int low = 0;
    for (;;) {
        if (low > t->low[n]) t->low[n] = low;
        else low = t->low[n];
        while (low < threshold && !t->known[n]) {
            uint32_t bit;
            if (bior_bit(r, &bit)) return -1;
            if (bit) {
                t->value[n] = low;
                t->known[n] = 1;
            } else {
                low++;
            }
        }
        if (t->known[n] && t->value[n] > low) low = t->value[n];
        t->low[n] = low;
        if (depth == 0) break;
        n = stk[--depth];
    }
    *res = t->known[leaf] && t->value[leaf] < threshold;
    return 0;
}

int ora_parse_packets(const ora_cs_header *hd, ora_dec_block *blocks, int nblocks,
                      const ora_dec_band *bands, int nbands, char *err) {
    const uint8_t *pay = hd->payload;
    size_t plen = hd->payload_len, pos = 0;
    tagtree *incl = calloc((size_t)nbands, sizeof(tagtree));
    tagtree *imsb = calloc((size_t)nbands, sizeof(tagtree));
    /* a packet's contributions in header order (decode/t2.rs:96-152 collects
       them in a Vec): at most one per code-block */
    int *order = malloc((size_t)(nblocks > 0 ? nblocks : 1) * sizeof(int));
    for (int i = 0; i < nbands; i++) {
        tgt_init(&incl[i], bands[i].cx > 0 ? bands[i].cx : 1, bands[i].cy > 0 ? bands[i].cy : 1, 0x7fffffff);
        tgt_init(&imsb[i], bands[i].cx > 0 ? bands[i].cx : 1, bands[i].cy > 0 ? bands[i].cy : 1, 0x7fffffff);
    }
    int rc = ORA_OK;
    for (int r = 0; r <= hd->levels && !rc; r++) {
        for (int c = 0; c < hd->ncomp && !rc; c++) {
            bior br = {pay + pos, plen - pos, 0, 0, 0, 0};
            uint32_t present;
            size_t body_total = 0;
            int contrib_first = -1, contrib_count = 0;
            if (bior_bit(&br, &present)) { rc = ORA_DECODE_FAILED; snprintf(err, 256, "packet header ended before requested bit"); break; }
            if (present) {
                for (int bi = 0; bi < nbands && !rc; bi++) {
                    if (bands[bi].comp != c || bands[bi].res != r) continue;
                    for (int k = 0; k < bands[bi].nblocks; k++) {
                        ora_dec_block *blk = &blocks[bands[bi].first + k];
                        int inc;
                        uint32_t bit;
                        if (blk->included) {
                            if (bior_bit(&br, &bit)) goto trunc;
                            inc = (int)bit;
                        } else {
                            if (tgt_decode(&incl[bi], &br, k, 1, &inc)) goto trunc;
                        }
                        if (!inc) continue;
                        if (blk->zbp < 0) {
                            int dummy;
                            if (tgt_decode(&imsb[bi], &br, k, 999, &dummy)) goto trunc;
                            if (!imsb[bi].known[k]) { rc = ORA_DECODE_FAILED; snprintf(err, 256, "zero-bitplane tag tree did not terminate"); break; }
                            blk->zbp = imsb[bi].value[k];
                        }
                        uint32_t np, v;
                        if (bior_bit(&br, &bit)) goto trunc;
                        if (!bit) np = 1;
                        else {
                            if (bior_bit(&br, &bit)) goto trunc;
                            if (!bit) np = 2;
                            else {
                                if (bior_bits(&br, 2, &v)) goto trunc;
                                if (v != 3) np = 3 + v;
                                else {
                                    if (bior_bits(&br, 5, &v)) goto trunc;
                                    if (v != 31) np = 6 + v;
                                    else {
                                        if (bior_bits(&br, 7, &v)) goto trunc;
                                        np = 37 + v;
                                    }
                                }
                            }
                        }
                        uint32_t incr = 0;
                        for (;;) {
                            if (bior_bit(&br, &bit)) goto trunc;
                            if (!bit) break;
                            incr++;
                        }
                        blk->numlenbits += (int)incr;
                        uint32_t len;
                        if (bior_bits(&br, blk->numlenbits + (int)floor_log2u(np), &len)) goto trunc;
                        if (len == 0) { rc = ORA_DECODE_FAILED; snprintf(err, 256, "non-empty code-block contribution has zero byte length"); break; }
                        blk->included = 1;
                        blk->passes = (int)np;
                        blk->len = len;
                        order[contrib_count] = bands[bi].first + k;
                        if (contrib_first < 0) contrib_first = bands[bi].first + k;
                        contrib_count++;
                        body_total += len;
                    }
                }
            }
            if (rc) break;
            pos += br.pos;
            if (pos + body_total > plen) { rc = ORA_DECODE_FAILED; snprintf(err, 256, "packet body extends past tile payload"); break; }
            for (int i = 0; i < contrib_count; i++) {
                blocks[order[i]].data = pay + pos;
                pos += blocks[order[i]].len;
            }
            continue;
        trunc:
            rc = ORA_DECODE_FAILED;
            snprintf(err, 256, "packet header ended before requested bit");
        }
    }
    if (!rc && pos != plen) {
        rc = ORA_DECODE_FAILED;
        snprintf(err, 256, "tile payload has %zu trailing bytes after packet parse", plen - pos);
    }
    for (int i = 0; i < nbands; i++) {
        tgt_free(&incl[i]);
        tgt_free(&imsb[i]);
    }
    free(incl);
    free(imsb);
    free(order);
    return rc;
}
