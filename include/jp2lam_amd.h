/*
 * jp2lam_amd: C ABI of the MI355X JPEG 2000 Part 1 codec core.
 *
 * The entry points mirror the public API of the jp2lam crate so that a
 * Rust `build.rs` + `extern "C"` shim can bind them one for one (see
 * INTEGRATION.md).  Reference locations are given as crate-relative paths.
 *
 * Conventions
 *  - Every function returning jp2lam_status sets a thread-local message
 *    readable with jp2lam_last_error() when it fails.
 *  - Status codes map 1:1 onto Jp2LamError's variants (src/error.rs:6-13).
 *  - Buffers returned through `uint8_t** out` and jp2lam_image are owned by
 *    the library; release them with jp2lam_free / jp2lam_image_free.
 *  - Images hold planar i32 samples like the crate's Image (src/model.rs:64-81).
 */
#ifndef JP2LAM_AMD_H
#define JP2LAM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Jp2LamError variants (src/error.rs:6-13). */
typedef enum {
    JP2LAM_OK = 0,
    JP2LAM_INVALID_INPUT = 1,
    JP2LAM_ENCODE_FAILED = 2,
    JP2LAM_DECODE_FAILED = 3,
    JP2LAM_UNSUPPORTED_FEATURE = 4,
    JP2LAM_IO = 5
} jp2lam_status;

/* ColorSpace (src/model.rs:1-12), numbered in the enum's declaration order
 * so that `ColorSpace as u32` on the Rust side is the value passed here. */
typedef enum {
    JP2LAM_GRAY = 0,
    JP2LAM_RGB = 1,
    JP2LAM_SRGB = 2,
    JP2LAM_YUV = 3,
    JP2LAM_YCBCR = 4
} jp2lam_colorspace;

/* OutputFormat (src/model.rs:29-33). */
typedef enum { JP2LAM_FORMAT_JP2 = 0, JP2LAM_FORMAT_J2K = 1 } jp2lam_format;

/* EncodeOptions{quality, format} (src/model.rs:155-177); quality >= 100 is
 * lossless (src/plan/mod.rs:121). */
typedef struct {
    uint8_t quality;
    uint8_t format;
} jp2lam_encode_options;

/* Component (src/model.rs:64-73), borrowed. */
typedef struct {
    const int32_t* data;
    uint32_t width, height, precision;
    uint32_t is_signed;
    uint32_t dx, dy;
} jp2lam_component;

/* Image (src/model.rs:75-81), borrowed view. */
typedef struct {
    uint32_t width, height;
    uint32_t colorspace;
    uint32_t num_components;
    const jp2lam_component* components;
} jp2lam_image_view;

/* Decoded Image, library-owned: planes holds num_components planes of
 * width*height samples (precision 8, unsigned, dx = dy = 1). */
typedef struct {
    uint32_t width, height, colorspace, num_components, precision;
    int32_t* planes;
} jp2lam_image;

/* DecodeMetadata (src/decode/mod.rs:23-31) flattened. */
typedef struct {
    uint32_t width, height, colorspace, num_components, bits_per_component, has_ipr_metadata;
    uint32_t decomposition_levels, code_block_width, code_block_height, layers, transform, use_mct;
    uint32_t guard_bits, num_quant_steps;
    uint64_t tile_part_count, first_tile_payload_len;
} jp2lam_metadata;

/* ---- crate API ------------------------------------------------------------ */

/* encode(&Image, &EncodeOptions) (src/encode/mod.rs:84-101). */
jp2lam_status jp2lam_encode(const jp2lam_image_view* image, const jp2lam_encode_options* opts, uint8_t** out,
                            size_t* out_len);
/* EncodeMetrics (src/encode/mod.rs:114-122). */
typedef struct {
    double psnr_db; /* INFINITY for lossless encodes */
    double ssim;    /* mean SSIM over 8x8 luma blocks; 1.0 for lossless */
} jp2lam_encode_metrics;
/* encode_with_psnr(&Image, &EncodeOptions) (src/encode/mod.rs:124-135):
 * the encode plus compute_quality_metrics (encode/backend/native/
 * backend.rs:297-433), evaluated on the device from the encode's own
 * coefficients.  Gray and sRGB only for quality < 100 (ENCODE_FAILED
 * otherwise, as the reference); on failure nothing is returned. */
jp2lam_status jp2lam_encode_with_psnr(const jp2lam_image_view* image, const jp2lam_encode_options* opts, uint8_t** out,
                                      size_t* out_len, jp2lam_encode_metrics* metrics);
/* ContentStats + ContentClass (src/content.rs:46-105). */
typedef struct {
    float bright_share;     /* share of pixels with luma > 224 */
    float entropy;          /* bits, 256-bin luma histogram */
    float colorfulness;     /* mean max(r,g,b) - min(r,g,b); 0 for gray */
    uint32_t content_class; /* 0 Photo, 1 Graphics (ContentStats::classify) */
    uint32_t quality;       /* ContentClass::quality: 78 Photo, 55 Graphics */
} jp2lam_content_stats;
/* analyze_content(&Image) / auto_quality (src/content.rs:119-174): one
 * histogram pass on the device.  Gray reads component 0, every other
 * colour space components 0..2 as R, G, B; too few components give zero
 * statistics (classified Graphics), as the reference. */
jp2lam_status jp2lam_analyze_content(const jp2lam_image_view* image, jp2lam_content_stats* out);
/* The same on interleaved u8 pixels (channels 1 gray / 3 RGB), host or
 * device (on_device) resident. */
jp2lam_status jp2lam_analyze_content_u8(const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                                        int on_device, jp2lam_content_stats* out);
/* No counterpart in the reference.  jp2lam_analyze_content_u8 for pixels
 * already in device memory, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * it calls no hipMalloc/hipFree and touches no engine.  Its counters live in
 * a fixed 257 x 8-byte device array per GPU (256 luma bins + the chroma
 * spread sum), cleared by a kernel at the start of every call; calls that
 * overlap on one GPU share that array and race, so serialise them.  The
 * work runs on `stream` (a hipStream_t; NULL = the default stream), which
 * the call synchronises.  width and height are at most
 * JP2LAM_NOALLOC_MAX_DIM (INVALID_INPUT beyond).  Results equal
 * jp2lam_analyze_content_u8 with on_device = 1. */
#define JP2LAM_NOALLOC_MAX_DIM 65535u
jp2lam_status jp2lam_analyze_content_u8_noalloc(const uint8_t* dev_pixels, uint32_t width, uint32_t height,
                                                uint32_t channels, void* stream, jp2lam_content_stats* out);
/* Image::from_gray_bytes / from_rgb_bytes (src/model.rs:87-140) + encode:
 * interleaved u8 pixels, channels 1 (gray) or 3 (sRGB). */
jp2lam_status jp2lam_encode_u8(const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                               const jp2lam_encode_options* opts, uint8_t** out, size_t* out_len);
/* decode_jp2(&[u8]) (src/decode/mod.rs:56-61). */
jp2lam_status jp2lam_decode_jp2(const uint8_t* bytes, size_t len, jp2lam_image* out);
/* decode_jp2 into caller-owned interleaved u8 pixels (cap bytes). */
jp2lam_status jp2lam_decode_jp2_u8(const uint8_t* bytes, size_t len, uint8_t* pixels, size_t cap, uint32_t* width,
                                   uint32_t* height, uint32_t* channels);
/* inspect_jp2(&[u8]) (src/decode/mod.rs:38-49). */
jp2lam_status jp2lam_inspect_jp2(const uint8_t* bytes, size_t len, jp2lam_metadata* out);
void jp2lam_free(void* p);
void jp2lam_image_free(jp2lam_image* image);
const char* jp2lam_last_error(void);

/* ---- batch handles (src/batch.rs:32-130) ------------------------------------
 * Same contract as BatchEncoder / BatchDecoder: the first item fixes the
 * profile, later items with another profile fail with INVALID_INPUT.  One
 * thread per handle. */
typedef struct jp2lam_batch_encoder jp2lam_batch_encoder;
typedef struct jp2lam_batch_decoder jp2lam_batch_decoder;
jp2lam_batch_encoder* jp2lam_batch_encoder_new(const jp2lam_encode_options* opts);
jp2lam_status jp2lam_batch_encoder_encode_one(jp2lam_batch_encoder* enc, const jp2lam_image_view* image, uint8_t** out,
                                              size_t* out_len);
size_t jp2lam_batch_encoder_encoded_count(const jp2lam_batch_encoder* enc);
int jp2lam_batch_encoder_has_profile(const jp2lam_batch_encoder* enc);
void jp2lam_batch_encoder_free(jp2lam_batch_encoder* enc);
/* BatchComponentProfile (src/batch.rs:15-23). */
typedef struct {
    uint32_t width, height, precision, is_signed, dx, dy;
} jp2lam_component_profile;
/* BatchProfile (src/batch.rs:5-13) without its component list, which the
 * getters below write to a caller array. */
typedef struct {
    uint32_t width, height, colorspace, num_components;
    uint8_t quality, format, pad[2];
} jp2lam_batch_profile;
/* BatchEncoder::profile (src/batch.rs:47-49): 1 and *out filled (plus the
 * first min(cap, num_components) component profiles into comps, which may
 * be NULL when cap is 0) once the first item has fixed the profile; 0 while
 * it is None; -1 for a null handle or out. */
int jp2lam_batch_encoder_profile(const jp2lam_batch_encoder* enc, jp2lam_batch_profile* out,
                                 jp2lam_component_profile* comps, size_t cap);
jp2lam_batch_decoder* jp2lam_batch_decoder_new(void);
jp2lam_status jp2lam_batch_decoder_decode_one(jp2lam_batch_decoder* dec, const uint8_t* bytes, size_t len,
                                              jp2lam_image* out);
size_t jp2lam_batch_decoder_decoded_count(const jp2lam_batch_decoder* dec);
void jp2lam_batch_decoder_free(jp2lam_batch_decoder* dec);
/* BatchDecoder::profile (src/batch.rs:91-93); same contract as
 * jp2lam_batch_encoder_profile.  A decoder's profile carries quality 0 and
 * format JP2 (BatchProfile::for_decode, src/batch.rs:152-154). */
int jp2lam_batch_decoder_profile(const jp2lam_batch_decoder* dec, jp2lam_batch_profile* out,
                                 jp2lam_component_profile* comps, size_t cap);

/* ---- mixed batches (no counterpart in the reference) ------------------------
 * Pages of different sizes and qualities in one device batch; each result
 * equals jp2lam_encode_u8 / jp2lam_decode_jp2_u8 of that item alone.
 * `on_device` marks pixel pointers that are device (HBM) addresses. */
typedef struct {
    const uint8_t* pixels;
    uint32_t width, height, channels;
    uint8_t quality, format, on_device, pad;
    uint8_t* out;      /* set by the library, free with jp2lam_free */
    size_t out_len;
    int32_t status;
} jp2lam_encode_item;
typedef struct {
    const uint8_t* bytes;
    size_t len;
    uint8_t* pixels;   /* caller buffer, host or device (on_device) */
    size_t cap;
    uint32_t width, height, channels;
    uint8_t on_device, pad[3];
    int32_t status;
} jp2lam_decode_item;
jp2lam_status jp2lam_encode_many(jp2lam_encode_item* items, size_t n);
jp2lam_status jp2lam_decode_many(jp2lam_decode_item* items, size_t n);
/* Queue form of the two calls above: the list is cut, in order, into device
 * batches of at most `chunk_blocks` code-blocks (decodes weighed by their
 * header size), which run on `engines` engines at once, so one batch's host
 * work overlaps another's kernels.  0 restores the default for either value,
 * which jp2lam_amd/autotune.py measured (at most 262144 blocks, 8 engines). */
void jp2g_set_queue(uint32_t chunk_blocks, int engines);
void jp2g_get_queue(uint32_t* chunk_blocks, int* engines);

/* ---- serving (no counterpart in the reference) ---------------------------
 * Continuous batching over `slots` device engines: a free slot takes the
 * oldest request and every later one that fits its share of the memory
 * budget (at most `max_batch`), runs them as one batch, and publishes the
 * results; batches on different slots overlap on the GPU.  Each result is
 * the same bytes/pixels as the request alone.  Host pixels and streams are
 * copied at submission; device pixels (on_device) must stay valid and
 * complete until the request is done. */
typedef struct jp2lam_server jp2lam_server;
typedef struct {
    int32_t status;        /* the request's own status */
    uint8_t* data;         /* encode: codestream; decode: interleaved u8; free with jp2lam_free */
    size_t len;
    uint32_t width, height, channels, batch_size;
    double queue_ms, run_ms, total_ms;  /* submit->start, start->done, submit->done */
} jp2lam_served;
jp2lam_server* jp2lam_server_new(int slots, size_t mem_budget_mib, int max_batch);
/* The same request loop on a simulated executor, for testing admission,
 * cancellation and fault handling without a GPU: no device work; an
 * encode's result is a 24-byte digest of the request (width, height,
 * colour space, quality as u32, FNV-1a of the pixels), a decode's the
 * stream's bytes as a len x 1 gray image, each batch after delay_us. */
jp2lam_server* jp2lam_server_new_simulated(int slots, size_t mem_budget_mib, int max_batch, uint32_t delay_us);
/* Return a request id, or 0 with jp2lam_last_error() set.  A request whose
 * device-memory estimate exceeds one slot's share of the budget is refused
 * here with INVALID_INPUT.  A device error while a batch runs fails that
 * batch's requests with ENCODE_FAILED / DECODE_FAILED ("device error: ...")
 * and the server goes on serving; JP2LAM_FAULT_INJECT="encode[:N]" or
 * "decode[:N]", read when the server starts, injects such an error into the
 * N-th batch call of that kind (default the first), for tests. */
uint64_t jp2lam_server_submit_encode(jp2lam_server* srv, const uint8_t* pixels, uint32_t width, uint32_t height,
                                     uint32_t channels, uint8_t quality, uint8_t format, uint8_t on_device);
uint64_t jp2lam_server_submit_decode(jp2lam_server* srv, const uint8_t* bytes, size_t len);
/* 1: done (out filled), 0: timed out (timeout_ms >= 0), -1: bad arguments. */
int jp2lam_server_wait(jp2lam_server* srv, uint64_t id, int timeout_ms, jp2lam_served* out);
/* {submitted, completed, batches, largest batch} */
void jp2lam_server_stats(jp2lam_server* srv, uint64_t* out4);
/* Cancel a request that no slot has taken yet: 1 done (its wait then gives
 * status IO, "request cancelled"), 0 already running or finished, -1 for
 * an unknown id or null server. */
int jp2lam_server_cancel(jp2lam_server* srv, uint64_t id);
/* {requests refused at admission, cancelled, batch calls failed by a device
 * error} */
void jp2lam_server_faults(jp2lam_server* srv, uint64_t* out3);
void jp2lam_server_free(jp2lam_server* srv);

/* ---- device buffers (no counterpart in the reference) ----------------------
 * For callers without their own HIP allocator: HBM buffers to hold pixels
 * for on_device encode items and decode targets.  kind: 0 host->device,
 * 1 device->host, 2 device->device; copies are synchronous. */
void* jp2lam_device_alloc(size_t bytes);
jp2lam_status jp2lam_device_copy(void* dst, const void* src, size_t bytes, int kind);
void jp2lam_device_free(void* ptr);

/* ---- stage-level entry points for parity tests ------------------------------- */
/* Forward colour + DWT: coefficient planes (Mallat layout) as the Tier-1
 * stage sees them before quantisation: f32 (quality < 100) or i32.  `out`
 * holds channels * width * height elements. */
jp2lam_status jp2g_forward_dwt(const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                               uint32_t quality, void* out);
/* Inverse DWT + inverse colour + finalisation of coefficient planes
 * (f32 dequantised when irreversible, else i32) to interleaved u8. */
jp2lam_status jp2g_inverse_dwt(const void* planes, uint32_t width, uint32_t height, uint32_t channels,
                               uint32_t levels, uint32_t irreversible, uint8_t* out);
/* Tier-1 coding of one i32 code-block (<= 64x64): info = {mbp, zbp, cls,
 * npasses, nbytes}, passes = 5 ints per pass {kind, bitplane, newly
 * significant, significant before, length}. */
jp2lam_status jp2g_t1_encode_block(const int32_t* coeffs, uint32_t w, uint32_t h, uint32_t band, uint32_t mb,
                                   int32_t* info, int32_t* passes, uint8_t* bytes, size_t cap);
jp2lam_status jp2g_t1_decode_block(uint32_t w, uint32_t h, uint32_t band, uint32_t mb, uint32_t zbp, uint32_t passes,
                                   const uint8_t* data, size_t len, int32_t* out);
/* Tier-1 counters (src/encode/counters.rs:3-41, the crate's "counters"
 * feature): while on, every encode adds {total blocks, empty blocks, MQ
 * symbols, cleanup passes, significance passes, refinement passes, pass
 * bytes} to process-wide totals, counted on the device (the MQ decisions of
 * every pass formed again after coding; the coding kernels are unchanged).
 * jp2g_counters copies the 7 totals, jp2g_reset_counters clears them
 * (counters.rs:11-19). */
void jp2g_set_counters(int on);
void jp2g_counters(uint64_t* out7);
void jp2g_reset_counters(void);
/* MQ decisions of block `index` of the last encode (debug capture and
 * counting on): all its passes, or only the passes PCRD kept (the ones the
 * decoder reads back). */
uint32_t jp2g_last_block_symbols(int index, int kept_only);
/* Per-block record of the last encode while debug capture is on. */
void jp2g_set_debug(int on);
int jp2g_last_block_count(void);
void jp2g_last_block_info(int index, int32_t* info16);
void jp2g_last_block_passes(int index, int32_t* out);
void jp2g_last_block_bytes(int index, uint8_t* out);
double jp2g_last_lambda(int group);
/* Tier-1 encoder selection (no reference counterpart; the reference codes
 * every block the same way, t1.rs:287-290): mode 0 = tuned: the coder and
 * split point come from a table that jp2lam_amd/autotune.py measured
 * offline on an idle MI355X, looked up by the batch's size class (bit
 * length of its code-block count); nothing is probed or timed while
 * encoding, so a batch's cost does not depend on earlier calls; 1 = one
 * wavefront per code-block; 2 = one lane per code-block; 3 = one lane per
 * code-block from `lane_min_blocks` code-blocks up; 4 = hybrid: blocks whose
 * formed symbol stream is longer than the split fraction (see
 * jp2g_set_hybrid_frac) x the batch's longest are coded by a wavefront
 * each, the rest by lanes, in one launch.  Output is the same in every
 * mode. */
void jp2g_set_t1_mode(int mode, uint32_t lane_min_blocks);
/* No counterpart in the reference (a CPU library).  Each engine runs its
 * device launch sequence (colour transform, DWT levels, Tier-1, PCRD and
 * packing for an encode; Tier-1, inverse DWT and pixel output for a decode)
 * as a HIP graph once the same sequence (same arena offsets, counts and
 * grids, i.e. a batch of the same shape) has been seen twice; other batches
 * run eagerly.  Output is the same either way.  Off by default (measured no
 * faster than eager launches); jp2g_set_graphs(1) turns the graphs on for
 * every engine, jp2g_set_graphs(0) off; jp2g_graph_stats reports how many
 * sequences were captured and how many graph launches ran. */
void jp2g_set_graphs(int on);
void jp2g_graph_stats(uint64_t* captures, uint64_t* launches);
/* 1 when the last encode ran the lane-per-block coder. */
int jp2g_last_t1_lanes(void);
/* Tier-1 coder of the last encode: 1 wave per block, 2 lane per block,
 * 4 hybrid (long streams on the scalar coder, short ones on lanes). */
int jp2g_last_t1_coder(void);
/* Hybrid split point for mode 4: streams longer than frac x the batch's
 * longest go to the scalar coder (0 < frac < 1; default 0.55). */
void jp2g_set_hybrid_frac(float frac);
/* Stage timings of the last call in ms: {dwt, tier1, pcrd+pack, host tier2}. */
void jp2g_last_timings(double* out4);
/* Context/MQ tables in the reference's layouts (for table checks). */
void jp2g_tables(uint16_t* qe, uint8_t* mps, uint8_t* nmps, uint8_t* nlps, uint8_t* zc, uint8_t* sc, uint8_t* spb);
/* 1 if a HIP device is usable. */
int jp2g_device_available(void);

#ifdef __cplusplus
}
#endif
#endif
