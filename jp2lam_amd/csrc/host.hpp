// Host-side stages of the codec: planning, Tier-2 packets, codestream
// markers and JP2 boxes (encode and decode).  These stay on the CPU; the
// data-parallel stages run in the HIP kernels.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "types.hpp"

namespace j2l {

// ColorSpace in the order of src/model.rs:1-12.
enum ColorSpace : int { GRAY = 0, RGB = 1, SRGB = 2, YUV = 3, YCBCR = 4 };

struct SubbandQuant {
    uint8_t res, band, exponent;
    uint16_t mantissa;
};

// plan/mod.rs:116-177
struct Plan {
    uint32_t width = 0, height = 0;
    int ncomp = 1;
    int domain = GRAY;  // encoding domain: GRAY or SRGB
    int quality = 85;
    int format = 0;     // 0 JP2, 1 J2K
    bool lossless = false;
    int levels = 0;
    int cbw = 64, cbh = 64;
    int guard = 2;
    bool use_mct = false;
    std::vector<SubbandQuant> quants;  // LL, then (HL, LH, HH) per resolution
};

Plan build_plan(uint32_t w, uint32_t h, int colorspace, int quality, int format);
uint32_t max_decompositions(uint32_t w, uint32_t h);
float quant_step_f32(int band, uint8_t exponent, uint16_t mantissa, uint32_t precision = 8);
double quant_step_f64(int band, uint8_t exponent, uint16_t mantissa, uint32_t precision = 8);
double norm97(uint32_t level, int band);
double band_bias(int band, int quality);
double component_weight(const Plan& p, int comp);
double quality_to_lambda(int quality, uint32_t pixels);
int subband_mb(const Plan& p, int comp, int res, int band);

// Geometry (layout.rs:37-173): Mallat rectangle of (res, band).
struct Rect {
    int x0, y0, x1, y1;
};
void resolution_ladder(uint32_t w, uint32_t h, int levels, std::vector<uint32_t>& rw, std::vector<uint32_t>& rh);
Rect band_rect(uint32_t w, uint32_t h, int levels, int res, int band);
// Storage of a component's Mallat plane on the device.  The bands keep
// their Mallat rows; the HL and HH bands of resolution r start at column
// hx[r].  Aligned (the encoder): hx[r] is a multiple of 32 samples (128
// bytes) past the end of the lower resolutions' columns and the row stride
// a multiple of 32 too, so every code-block row starts on a 128-byte line:
// with the plain Mallat columns a 64-sample block row spans three lines and
// the Tier-1 reads fetch 1.5x their bytes.  Plain (the decoder): the Mallat
// columns, stride w; the inverse DWT reads each band row from two samples
// before a tile's start, which on an aligned band spans one more line.  The
// aligned form falls back to the plain one when its rows would not fit the
// 16-bit block offsets.
struct PlaneLayout {
    uint32_t stride = 0;
    uint32_t hx[33] = {};
};
PlaneLayout plane_layout(uint32_t w, int levels, bool aligned);
// Storage column of the sample at Mallat column x of band `band` at
// resolution `res` (band_rect's coordinates).
inline uint32_t plane_x(const PlaneLayout& pl, const Rect& rc, int res, int band, int x) {
    return band == HL || band == HH ? pl.hx[res] + (uint32_t)(x - rc.x0) : (uint32_t)x;
}
// Row stride of an LL buffer of width w: a multiple of 32 samples.
inline uint32_t ll_stride(uint32_t w) { return (w + 31u) & ~31u; }

// Kernel lookup tables (MQ states, ZC/SC contexts) in the per-lane layout
// the Tier-1 kernels read: 64 MQ words, 4 x 64 ZC words, 64 SC words.
std::array<uint32_t, 448> make_t1_tables();
// Plain tables for tests: MQ expanded to 94 states, ZC [4][512], SC/SPB [256].
void expanded_tables(uint16_t* qe, uint8_t* mps, uint8_t* nmps, uint8_t* nlps, uint8_t* zc, uint8_t* sc,
                     uint8_t* spb);

// Code-block view used by the packet-header writer.
struct CodedBlock {
    uint8_t keep;      // passes included (0 = not included)
    uint8_t zbp;
    uint32_t seg_len;  // bytes of its single segment
};
struct BandBlocks {
    int cx, cy;
    const CodedBlock* blocks;  // cx * cy in raster order
};
// One LRCP packet header for (resolution, component) (t2.rs:337-396).
void packet_header(const BandBlocks* bands, int nbands, std::vector<uint8_t>& out);
// SOC..EOC around one tile-part payload (j2k/markers.rs, j2k/write.rs).
void write_codestream_head(const Plan& p, size_t payload_len, std::vector<uint8_t>& out);
void write_codestream_tail(std::vector<uint8_t>& out);
// JP2 boxes before the codestream (jp2.rs:13-50).
void write_jp2_head(const Plan& p, size_t codestream_len, std::vector<uint8_t>& out);

// ---- content analysis (content.rs:46-174) ----------------------------------------
struct ContentOut {
    float bright_share = 0.0f, entropy = 0.0f, colorfulness = 0.0f;
    int graphics = 1;  // 1 Graphics, 0 Photo (all-zero statistics classify as Graphics)
};
// Statistics and class from the device histogram: hist[0..255] luma bins,
// hist[256] the sum of chroma spreads, n samples (content.rs:120-174).
ContentOut content_from_hist(const uint64_t* hist, uint64_t n);

// ---- decoding -----------------------------------------------------------------
struct Jp2Info {
    uint32_t width = 0, height = 0;
    int ncomp = 0, bpc = 0, colorspace = GRAY;
    bool ipr = false;
};
struct CodestreamInfo {
    uint32_t width = 0, height = 0;
    int ncomp = 0;
    int levels = 0, cbw = 64, cbh = 64, layers = 1, mct = 0, transform = 0, guard = 2, qstyle = 0;
    int progression = 0, tile_parts = 0;
    std::vector<uint8_t> exps;
    std::vector<uint16_t> mants;
    std::vector<int> prec;
    const uint8_t* payload = nullptr;
    size_t payload_len = 0;
    size_t first_payload_len = 0;
};
struct DecBlock {
    int comp, res, band, band_index;
    Rect r;
    int zbp = -1;
    int passes = 0;
    int numlenbits = 3;
    bool included = false;
    uint32_t off = 0, len = 0;  // within the tile payload
};
// decode/jp2_parse.rs:42-258 -> codestream slice
void parse_jp2(const uint8_t* b, size_t n, Jp2Info& info, const uint8_t*& cs, size_t& cslen);
// decode/codestream.rs + j2k/decode_markers.rs:139-501
void parse_codestream(const uint8_t* b, size_t n, CodestreamInfo& ci);
// decode/t2.rs:75-256: walk the LRCP packets; fills blocks (component-major
// band order, raster within band).
void parse_packets(const CodestreamInfo& ci, std::vector<DecBlock>& blocks);
// The same walk in steps: the constructor lays out every band's blocks,
// run(rmax) parses the packets of the resolutions up to rmax not parsed yet
// (one layer, so LRCP order is resolution-major), finish() checks that the
// payload is used up.  parse_packets is run(levels) then finish().
class PacketParser {
public:
    PacketParser(const CodestreamInfo& ci, std::vector<DecBlock>& blocks);
    ~PacketParser();
    void run(int rmax);
    void finish();
    // the header and block list moved (their contents unchanged)
    void rebind(const CodestreamInfo& ci, std::vector<DecBlock>& blocks) {
        ci_ = &ci;
        blocks_ = &blocks;
    }

private:
    struct Band;
    const CodestreamInfo* ci_;
    std::vector<DecBlock>* blocks_;
    std::vector<Band> bands_;
    int layer_ = 0, res_ = 0;
    size_t pos_ = 0;
};
// decode/t1.rs:96-127 per-block checks (zero bit-planes, pass counts).
void check_decode_blocks(const CodestreamInfo& ci, const std::vector<DecBlock>& blocks);
// The container and codestream part of parse_for_decode: JP2 boxes,
// markers and the header checks, no packets.
void parse_decode_header(const uint8_t* bytes, size_t len, Jp2Info& jp, CodestreamInfo& ci);
// Every host-side step of decode_jp2 before Tier-1 (decode/mod.rs:56-129):
// JP2 boxes, codestream markers, header consistency, packets, block checks.
// Throws Error(DECODE_FAILED / UNSUPPORTED_FEATURE) on malformed input.
void parse_for_decode(const uint8_t* bytes, size_t len, Jp2Info& jp, CodestreamInfo& ci, std::vector<DecBlock>& blocks);

}  // namespace j2l
