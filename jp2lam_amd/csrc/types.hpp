// Types and constants shared by the host stages and the device engine.
// Plain C++: the host-side parsers and writers build from this header
// alone, with any C++17 compiler.
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace j2l {

// Band orientation, same numbering as the reference's BandOrientation
// (plan/mod.rs:45-50) and LUT orientation index (tier1/helpers.rs:13-20).
enum Band : int { LL = 0, HL = 1, LH = 2, HH = 3 };

// Status codes of the C ABI, 1:1 with Jp2LamError's variants (error.rs:6-13).
enum Status : int {
    OK = 0,
    INVALID_INPUT = 1,
    ENCODE_FAILED = 2,
    DECODE_FAILED = 3,
    UNSUPPORTED_FEATURE = 4,
    IO = 5,
};

struct Error : std::runtime_error {
    Status status;
    Error(Status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

constexpr int kMaxLevels = 6;      // encoder caps: 5 gray / 6 sRGB (plan/derive.rs:109-115)
// Pixels per workgroup of the colour kernels (prepare, finalize): 256
// threads, four coalesced passes each.
constexpr uint32_t kPixPerBlock = 1024;
inline uint32_t pix_blocks(size_t n) { return (uint32_t)((n + kPixPerBlock - 1) / kPixPerBlock); }
constexpr int kMaxPasses = 96;     // 1 + 3 * (31 - 1) rounded up
constexpr int kMaxDecodeLevels = 32;

// One code-block job for the Tier-1 kernels.  Coordinates are in the
// component's coefficient plane: the Mallat layout of layout.rs:37-173, with
// the detail bands' columns placed as host.hpp's plane_layout says.
struct BlockJob {
    uint64_t plane;     // device address of the coefficient plane (f32 lossy / i32 lossless)
    uint32_t stride;    // plane row stride in elements
    uint16_t x0, y0;    // top-left in the plane (planes are <= 65535 wide here)
    uint8_t w, h;       // <= 64
    uint8_t band;       // orientation
    uint8_t mb;         // Mb of the subband
    float step;         // quantisation step (lossy), 0 for lossless
    uint32_t out_off;   // byte offset of this block's output slot
    uint32_t out_cap;   // capacity of the slot
};

// Per-block Tier-1 encode result.
struct BlockResult {
    uint32_t nbytes;    // terminated segment length
    uint32_t max_mag;
    uint16_t nnz;       // <= 4096
    uint8_t mbp, zbp, cls, npasses, overflow, pad;
};

// Per-block decode job.
struct DecodeJob {
    uint64_t plane;     // f32 (9/7) or i32 (5/3) Mallat plane of the component
    uint32_t stride;
    uint16_t x0, y0;
    uint8_t w, h, band, mb;
    float step;         // dequantisation step (9/7), 0 for 5/3
    uint32_t data_off;  // offset of the block's bytes in the payload buffer
    uint32_t data_len;
    int16_t zbp;        // -1: not included
    uint8_t passes;
    uint8_t pad;        // issue priority 0..3 (heaviest blocks highest)
};

// Per-block PCRD inputs precomputed on the host.
struct PcrdJob {
    double step;     // quantisation step (rate.rs:248-252)
    double swb;      // subband weight * band bias (rate.rs:232-241, pcrd.rs:728-748)
    double cw;       // component weight (backend.rs:696-718)
    uint32_t group;  // (image, component) index for the lambda reduction
    uint32_t is_ll;
};

// Per-block truncation decision.
struct Selection {
    uint32_t body_len;  // bytes sent for the block
    uint8_t keep;       // passes kept
    uint8_t zbp;
    uint8_t pad[2];
};

struct HullPt {
    double slope;
    int32_t passes;
    int32_t pad;
};

// DWT launch segment: one image component at one decomposition level.  A
// launch covers every segment; workgroups find theirs by tile_start.
struct DwtSeg {
    uint64_t in;       // forward: level input; inverse: LL input
    uint64_t plane;    // Mallat plane holding HL/LH/HH
    uint64_t out;      // forward: LL output; inverse: interleaved output
    int32_t in_stride, pstride, out_stride;
    int32_t rw, rh, tiles_x;
    uint32_t tile_start;
    uint32_t hx;       // plane column of this level's HL and HH bands
};

// Pixel launch segment: one image component (prepare) or one image
// (finalize), covering 256-element blocks [blk_start, next blk_start).
struct PixSeg {
    uint64_t src;         // prepare: u8 interleaved or i32 planar; finalize: comp 0
    uint64_t src1, src2;  // finalize: comp 1 and 2
    uint64_t dst;         // prepare: f32/i32 plane; finalize: u8 interleaved (or 0)
    uint64_t dst_i32;     // finalize: i32 planar (or 0)
    uint64_t n;           // pixels
    int32_t channels, comp, planar_i32, irrev;
    uint32_t blk_start, pad;
};

// Quality-metric launch segment: one image.  8x8 cells tile the image
// (partial cells on the right and bottom edges count for PSNR only).
struct MetricSeg {
    uint64_t rec[3];    // f32 reconstruction planes (level-shifted; Y, Cb, Cr for sRGB)
    uint64_t orig;      // encoder input samples: u8 interleaved or i32 planar
    uint32_t width, height;
    uint32_t cells_x, cells_y;
    uint32_t cell_start;
    int32_t ncomp, orig_i32;
    uint32_t pad;
};

}  // namespace j2l
