/* Oracle (test infrastructure only): Tier-2 / container types. */
#ifndef ORA_T2_H
#define ORA_T2_H
#include "ora.h"

typedef struct {
    uint8_t *data;
    size_t len, cap;
} obuf;
void obuf_put(obuf *b, const void *p, size_t n);
void obuf_u8(obuf *b, uint8_t v);
void obuf_u16(obuf *b, uint16_t v);
void obuf_u32(obuf *b, uint32_t v);

typedef struct {
    int nblocks, cx, cy;
    ora_block **blocks;
} ora_band_ref;

void ora_packet_header(const ora_band_ref *bands, int nbands, obuf *out);
void ora_write_codestream(const ora_plan *p, const uint8_t *payload, size_t plen, obuf *o);
void ora_wrap_jp2(const ora_plan *p, const uint8_t *cs, size_t len, obuf *o);

typedef struct {
    uint32_t width, height;
    int ncomp, bpc, colorspace;
} ora_jp2_header;

typedef struct {
    uint16_t rsiz;
    uint32_t width, height, xo, yo, tw, th, txo, tyo;
    int ncomp;
    int prec[16], sgnd[16], dx[16], dy[16];
    int scod, prog, layers, mct, levels, cbw, cbh, cbstyle, transform;
    int qstyle, guard, nsteps;
    uint8_t exp[128];
    uint16_t mant[128];
    int tile_index, part_index, total_parts;
    const uint8_t *payload;
    size_t payload_len;
} ora_cs_header;

typedef struct {
    int comp, res, band, band_index;
    int x0, y0, x1, y1;
    int included, zbp, passes, numlenbits;
    uint32_t len;
    const uint8_t *data;
} ora_dec_block;

typedef struct {
    int comp, res, band;
    int first, nblocks, cx, cy;
} ora_dec_band;

int ora_parse_jp2(const uint8_t *b, size_t n, ora_jp2_header *h, const uint8_t **cs, size_t *cslen, char *err);
int ora_parse_codestream(const uint8_t *b, size_t n, ora_cs_header *hd, char *err);
int ora_parse_packets(const ora_cs_header *hd, ora_dec_block *blocks, int nblocks,
                      const ora_dec_band *bands, int nbands, char *err);
#endif
