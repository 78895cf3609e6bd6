// C ABI of jp2lam_amd (include/jp2lam_amd.h).
#include "../../include/jp2lam_amd.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <optional>
#include <sstream>

#include "engine.hpp"
#include "queue.hpp"
#include "server.hpp"

using namespace j2l;

namespace {

thread_local std::string g_last_error;

jp2lam_status fail(Status s, const std::string& m) {
    g_last_error = m;
    return (jp2lam_status)s;
}

template <class F>
jp2lam_status guarded(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc&) {
        return fail(IO, "out of host memory");
    } catch (const std::exception& e) {
        return fail(IO, e.what());
    }
}

uint8_t* copy_out(const std::vector<uint8_t>& v) {
    uint8_t* p = (uint8_t*)std::malloc(v.size() ? v.size() : 1);
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), v.size());
    return p;
}

const char* cs_name(uint32_t cs) {
    switch (cs) {
        case JP2LAM_GRAY: return "Gray";
        case JP2LAM_SRGB: return "Srgb";
        case JP2LAM_RGB: return "Rgb";
        case JP2LAM_YUV: return "Yuv";
        case JP2LAM_YCBCR: return "YCbCr";
    }
    return "Unknown";
}

// plan/validate.rs:4-62 (the sample range scan runs in the engine)
void validate_view(const jp2lam_image_view* im, std::vector<int32_t>& planar) {
    if (!im) throw Error(INVALID_INPUT, "null image");
    if (im->colorspace > JP2LAM_YCBCR) throw Error(INVALID_INPUT, "unknown colour space");
    if (im->width == 0 || im->height == 0) throw Error(INVALID_INPUT, "image dimensions must be non-zero");
    uint32_t want = im->colorspace == JP2LAM_GRAY ? 1 : 3;
    if (im->num_components != want)
        throw Error(INVALID_INPUT, std::string(cs_name(im->colorspace)) + " images must have exactly " + std::to_string(want) +
                                       " component(s)");
    size_t n = (size_t)im->width * im->height;
    planar.resize(n * want);
    for (uint32_t c = 0; c < want; c++) {
        const jp2lam_component& k = im->components[c];
        const std::string idx = std::to_string(c);
        if (k.width != im->width || k.height != im->height)
            throw Error(INVALID_INPUT, "component " + idx + " dimensions " + std::to_string(k.width) + "x" +
                                           std::to_string(k.height) + " do not match image " + std::to_string(im->width) +
                                           "x" + std::to_string(im->height));
        if (k.precision != 8)
            throw Error(INVALID_INPUT, "component " + idx + " precision " + std::to_string(k.precision) +
                                           " is unsupported; only 8-bit is supported");
        if (k.is_signed) throw Error(INVALID_INPUT, "component " + idx + " must be unsigned");
        if (k.dx != 1 || k.dy != 1)
            throw Error(INVALID_INPUT, "component " + idx + " subsampling " + std::to_string(k.dx) + "x" +
                                           std::to_string(k.dy) + " is unsupported");
        if (!k.data) throw Error(INVALID_INPUT, "component " + idx + " has no samples");
        std::memcpy(planar.data() + c * n, k.data, n * 4);
    }
}

EncOut encode_one(const EncItem& it) {
    Engine& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    return std::move(e.encode({it})[0]);
}

DecOut decode_one(const DecItem& it) {
    Engine& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    return std::move(e.decode({it})[0]);
}

jp2lam_status encode_view(const jp2lam_image_view* image, const jp2lam_encode_options* opts, uint8_t** out,
                          size_t* out_len, jp2lam_encode_metrics* metrics = nullptr) {
    if (!opts || !out || !out_len) return fail(INVALID_INPUT, "null argument");
    std::vector<int32_t> planar;
    validate_view(image, planar);
    EncItem it;
    it.width = image->width;
    it.height = image->height;
    it.colorspace = (int)image->colorspace;
    it.quality = opts->quality;
    it.format = opts->format;
    it.src_kind = SRC_HOST_I32;
    it.src = planar.data();
    it.metrics = metrics != nullptr;
    EncOut r = encode_one(it);
    if (r.status != OK) return fail(r.status, r.message);
    *out = copy_out(r.bytes);
    *out_len = r.bytes.size();
    if (metrics) {
        metrics->psnr_db = r.psnr_db;
        metrics->ssim = r.ssim;
    }
    return JP2LAM_OK;
}

void fill_image(DecOut& r, jp2lam_image* out) {
    out->width = r.width;
    out->height = r.height;
    out->colorspace = (uint32_t)r.colorspace;
    out->num_components = (uint32_t)r.ncomp;
    out->precision = 8;
    out->planes = (int32_t*)std::malloc(sizeof(int32_t) * (r.planes.empty() ? 1 : r.planes.size()));
    if (!out->planes) throw std::bad_alloc();
    std::memcpy(out->planes, r.planes.data(), sizeof(int32_t) * r.planes.size());
}

// batch.rs:6-30
struct Profile {
    uint32_t width, height, colorspace;
    std::vector<std::array<uint32_t, 6>> comps;
    int quality, format;
    bool operator==(const Profile& o) const {
        return width == o.width && height == o.height && colorspace == o.colorspace && comps == o.comps &&
               quality == o.quality && format == o.format;
    }
    bool same_image(const Profile& o) const {
        return width == o.width && height == o.height && colorspace == o.colorspace && comps == o.comps;
    }
    std::string describe() const {
        std::ostringstream s;
        s << "BatchProfile { width: " << width << ", height: " << height << ", colorspace: " << cs_name(colorspace)
          << ", components: " << comps.size() << ", quality: " << quality << ", format: " << (format ? "J2k" : "Jp2")
          << " }";
        return s.str();
    }
};

Profile profile_of(const jp2lam_image_view* im, int quality, int format) {
    Profile p{im->width, im->height, im->colorspace, {}, quality, format};
    for (uint32_t c = 0; c < im->num_components; c++) {
        const auto& k = im->components[c];
        p.comps.push_back({k.width, k.height, k.precision, k.is_signed, k.dx, k.dy});
    }
    return p;
}

int export_profile(const std::optional<Profile>& p, jp2lam_batch_profile* out, jp2lam_component_profile* comps,
                   size_t cap) {
    std::memset(out, 0, sizeof(*out));
    if (!p) return 0;
    out->width = p->width;
    out->height = p->height;
    out->colorspace = p->colorspace;
    out->num_components = (uint32_t)p->comps.size();
    out->quality = (uint8_t)p->quality;
    out->format = (uint8_t)p->format;
    for (size_t c = 0; c < p->comps.size() && c < cap && comps; c++) {
        const auto& k = p->comps[c];
        comps[c] = {k[0], k[1], k[2], k[3], k[4], k[5]};
    }
    return 1;
}

}  // namespace

struct jp2lam_batch_encoder {
    jp2lam_encode_options opts;
    std::optional<Profile> profile;
    size_t count = 0;
};
struct jp2lam_batch_decoder {
    std::optional<Profile> profile;
    size_t count = 0;
};

extern "C" {

jp2lam_status jp2lam_encode(const jp2lam_image_view* image, const jp2lam_encode_options* opts, uint8_t** out,
                            size_t* out_len) {
    return guarded([&] { return encode_view(image, opts, out, out_len); });
}

jp2lam_status jp2lam_encode_with_psnr(const jp2lam_image_view* image, const jp2lam_encode_options* opts, uint8_t** out,
                                     size_t* out_len, jp2lam_encode_metrics* metrics) {
    return guarded([&] {
        if (!metrics) return fail(INVALID_INPUT, "null argument");
        return encode_view(image, opts, out, out_len, metrics);
    });
}

jp2lam_status jp2lam_analyze_content(const jp2lam_image_view* image, jp2lam_content_stats* out) {
    return guarded([&] {
        if (!image || !out) return fail(INVALID_INPUT, "null argument");
        std::memset(out, 0, sizeof(*out));
        const bool gray = image->colorspace == JP2LAM_GRAY;
        const uint32_t need = gray ? 1 : 3;
        ContentOut r;
        if (image->num_components >= need) {
            size_t n = SIZE_MAX;
            for (uint32_t c = 0; c < need; c++) {
                const jp2lam_component& k = image->components[c];
                if (!k.data && (size_t)k.width * k.height) return fail(INVALID_INPUT, "component has no samples");
                n = std::min(n, (size_t)k.width * k.height);
            }
            std::vector<int32_t> planar(n * need);
            for (uint32_t c = 0; c < need; c++)
                if (n) std::memcpy(planar.data() + c * n, image->components[c].data, n * 4);
            Engine& e = engine();
            std::lock_guard<std::mutex> lk(e.mu);
            r = e.analyze_content(planar.data(), SRC_HOST_I32, n, gray);
        }
        out->bright_share = r.bright_share;
        out->entropy = r.entropy;
        out->colorfulness = r.colorfulness;
        out->content_class = (uint32_t)r.graphics;
        out->quality = r.graphics ? 55u : 78u;
        return JP2LAM_OK;
    });
}

jp2lam_status jp2lam_analyze_content_u8(const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                                        int on_device, jp2lam_content_stats* out) {
    return guarded([&] {
        if (!pixels || !out) return fail(INVALID_INPUT, "null argument");
        if (channels != 1 && channels != 3) return fail(INVALID_INPUT, "channels must be 1 (gray) or 3 (sRGB)");
        Engine& e = engine();
        std::lock_guard<std::mutex> lk(e.mu);
        ContentOut r = e.analyze_content(pixels, on_device ? SRC_DEV_U8 : SRC_HOST_U8, (size_t)width * height, channels == 1);
        out->bright_share = r.bright_share;
        out->entropy = r.entropy;
        out->colorfulness = r.colorfulness;
        out->content_class = (uint32_t)r.graphics;
        out->quality = r.graphics ? 55u : 78u;
        return JP2LAM_OK;
    });
}

jp2lam_status jp2lam_analyze_content_u8_noalloc(const uint8_t* dev_pixels, uint32_t width, uint32_t height,
                                                uint32_t channels, void* stream, jp2lam_content_stats* out) {
    return guarded([&] {
        if (!dev_pixels || !out) return fail(INVALID_INPUT, "null argument");
        if (channels != 1 && channels != 3) return fail(INVALID_INPUT, "channels must be 1 (gray) or 3 (sRGB)");
        if (width > JP2LAM_NOALLOC_MAX_DIM || height > JP2LAM_NOALLOC_MAX_DIM)
            return fail(INVALID_INPUT, "image dimensions " + std::to_string(width) + "x" + std::to_string(height) +
                                           " exceed the no-allocation maximum of " +
                                           std::to_string(JP2LAM_NOALLOC_MAX_DIM) + " per side");
        std::memset(out, 0, sizeof(*out));
        const uint64_t n = (uint64_t)width * height;
        uint64_t hist[257] = {};
        if (n) content_hist_noalloc(dev_pixels, n, (int)channels, (hipStream_t)stream, hist);
        const ContentOut r = content_from_hist(hist, n);
        out->bright_share = r.bright_share;
        out->entropy = r.entropy;
        out->colorfulness = r.colorfulness;
        out->content_class = (uint32_t)r.graphics;
        out->quality = r.graphics ? 55u : 78u;
        return JP2LAM_OK;
    });
}

jp2lam_status jp2lam_encode_u8(const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                               const jp2lam_encode_options* opts, uint8_t** out, size_t* out_len) {
    return guarded([&] {
        if (!pixels || !opts || !out || !out_len) return fail(INVALID_INPUT, "null argument");
        if (channels != 1 && channels != 3) return fail(INVALID_INPUT, "channels must be 1 (gray) or 3 (sRGB)");
        EncItem it;
        it.width = width;
        it.height = height;
        it.colorspace = channels == 1 ? GRAY : SRGB;
        it.quality = opts->quality;
        it.format = opts->format;
        it.src_kind = SRC_HOST_U8;
        it.src = pixels;
        EncOut r = encode_one(it);
        if (r.status != OK) return fail(r.status, r.message);
        *out = copy_out(r.bytes);
        *out_len = r.bytes.size();
        return JP2LAM_OK;
    });
}

jp2lam_status jp2lam_decode_jp2(const uint8_t* bytes, size_t len, jp2lam_image* out) {
    return guarded([&] {
        if (!bytes || !out) return fail(INVALID_INPUT, "null argument");
        std::memset(out, 0, sizeof(*out));
        DecItem it;
        it.bytes = bytes;
        it.len = len;
        it.target = DST_HOST_I32;
        DecOut r = decode_one(it);
        if (r.status != OK) return fail(r.status, r.message);
        fill_image(r, out);
        return JP2LAM_OK;
    });
}

jp2lam_status jp2lam_decode_jp2_u8(const uint8_t* bytes, size_t len, uint8_t* pixels, size_t cap, uint32_t* width,
                                   uint32_t* height, uint32_t* channels) {
    return guarded([&] {
        if (!bytes || !pixels) return fail(INVALID_INPUT, "null argument");
        DecItem it;
        it.bytes = bytes;
        it.len = len;
        it.target = DST_HOST_U8;
        DecOut r = decode_one(it);
        if (r.status != OK) return fail(r.status, r.message);
        if (r.pixels.size() > cap) return fail(IO, "output buffer too small");
        std::memcpy(pixels, r.pixels.data(), r.pixels.size());
        if (width) *width = r.width;
        if (height) *height = r.height;
        if (channels) *channels = (uint32_t)r.ncomp;
        return JP2LAM_OK;
    });
}

jp2lam_status jp2lam_inspect_jp2(const uint8_t* bytes, size_t len, jp2lam_metadata* out) {
    return guarded([&] {
        if (!bytes || !out) return fail(INVALID_INPUT, "null argument");
        Jp2Info jp;
        CodestreamInfo ci;
        const uint8_t* cs;
        size_t cslen;
        parse_jp2(bytes, len, jp, cs, cslen);
        parse_codestream(cs, cslen, ci);
        if (jp.width != ci.width || jp.height != ci.height)
            return fail(DECODE_FAILED, "JP2 ihdr dimensions do not match SIZ dimensions");
        if (jp.bpc != 8) return fail(UNSUPPORTED_FEATURE, "unsupported JP2 bit depth");
        if (jp.ncomp != ci.ncomp) return fail(DECODE_FAILED, "JP2 ihdr component count does not match SIZ component count");
        std::memset(out, 0, sizeof(*out));
        out->width = ci.width;
        out->height = ci.height;
        out->colorspace = (uint32_t)jp.colorspace;
        out->num_components = (uint32_t)ci.ncomp;
        out->bits_per_component = (uint32_t)jp.bpc;
        out->has_ipr_metadata = jp.ipr ? 1 : 0;
        out->decomposition_levels = (uint32_t)ci.levels;
        out->code_block_width = (uint32_t)ci.cbw;
        out->code_block_height = (uint32_t)ci.cbh;
        out->layers = (uint32_t)ci.layers;
        out->transform = (uint32_t)ci.transform;
        out->use_mct = (uint32_t)ci.mct;
        out->guard_bits = (uint32_t)ci.guard;
        out->num_quant_steps = (uint32_t)ci.exps.size();
        out->tile_part_count = (uint64_t)ci.tile_parts;
        out->first_tile_payload_len = (uint64_t)ci.payload_len;
        return JP2LAM_OK;
    });
}

void jp2lam_free(void* p) { std::free(p); }

void jp2lam_image_free(jp2lam_image* image) {
    if (!image) return;
    std::free(image->planes);
    image->planes = nullptr;
}

const char* jp2lam_last_error(void) { return g_last_error.c_str(); }

jp2lam_batch_encoder* jp2lam_batch_encoder_new(const jp2lam_encode_options* opts) {
    auto* e = new (std::nothrow) jp2lam_batch_encoder();
    if (e && opts) e->opts = *opts;
    else if (e) e->opts = {85, 0};
    return e;
}

jp2lam_status jp2lam_batch_encoder_encode_one(jp2lam_batch_encoder* enc, const jp2lam_image_view* image, uint8_t** out,
                                              size_t* out_len) {
    return guarded([&] {
        if (!enc || !image) return fail(INVALID_INPUT, "null argument");
        Profile p = profile_of(image, enc->opts.quality, enc->opts.format);
        if (enc->profile && !(*enc->profile == p))
            return fail(INVALID_INPUT, "batch item profile mismatch: expected " + enc->profile->describe() + ", found " +
                                           p.describe());
        if (!enc->profile) enc->profile = p;
        jp2lam_status s = encode_view(image, &enc->opts, out, out_len);
        if (s == JP2LAM_OK) enc->count++;
        return s;
    });
}

size_t jp2lam_batch_encoder_encoded_count(const jp2lam_batch_encoder* enc) { return enc ? enc->count : 0; }
int jp2lam_batch_encoder_has_profile(const jp2lam_batch_encoder* enc) { return enc && enc->profile ? 1 : 0; }
void jp2lam_batch_encoder_free(jp2lam_batch_encoder* enc) { delete enc; }
int jp2lam_batch_encoder_profile(const jp2lam_batch_encoder* enc, jp2lam_batch_profile* out,
                                 jp2lam_component_profile* comps, size_t cap) {
    if (!enc || !out) return -1;
    return export_profile(enc->profile, out, comps, cap);
}

jp2lam_batch_decoder* jp2lam_batch_decoder_new(void) { return new (std::nothrow) jp2lam_batch_decoder(); }

jp2lam_status jp2lam_batch_decoder_decode_one(jp2lam_batch_decoder* dec, const uint8_t* bytes, size_t len,
                                              jp2lam_image* out) {
    return guarded([&] {
        if (!dec || !bytes || !out) return fail(INVALID_INPUT, "null argument");
        std::memset(out, 0, sizeof(*out));
        DecItem it;
        it.bytes = bytes;
        it.len = len;
        it.target = DST_HOST_I32;
        DecOut r = decode_one(it);
        if (r.status != OK) return fail(r.status, r.message);
        // batch.rs:99-130: decode first, then check the image profile
        Profile p{r.width, r.height, (uint32_t)r.colorspace, {}, 0, 0};
        for (int c = 0; c < r.ncomp; c++) p.comps.push_back({r.width, r.height, 8u, 0u, 1u, 1u});
        if (dec->profile && !dec->profile->same_image(p))
            return fail(INVALID_INPUT, "batch item profile mismatch: expected " + dec->profile->describe() + ", found " +
                                           p.describe());
        if (!dec->profile) dec->profile = p;
        fill_image(r, out);
        dec->count++;
        return JP2LAM_OK;
    });
}

size_t jp2lam_batch_decoder_decoded_count(const jp2lam_batch_decoder* dec) { return dec ? dec->count : 0; }
void jp2lam_batch_decoder_free(jp2lam_batch_decoder* dec) { delete dec; }
int jp2lam_batch_decoder_profile(const jp2lam_batch_decoder* dec, jp2lam_batch_profile* out,
                                 jp2lam_component_profile* comps, size_t cap) {
    if (!dec || !out) return -1;
    return export_profile(dec->profile, out, comps, cap);
}

jp2lam_status jp2lam_encode_many(jp2lam_encode_item* items, size_t n) {
    return guarded([&] {
        if (!items && n) return fail(INVALID_INPUT, "null argument");
        std::vector<EncItem> v(n);
        for (size_t i = 0; i < n; i++) {
            const auto& s = items[i];
            if (s.channels != 1 && s.channels != 3) return fail(INVALID_INPUT, "channels must be 1 (gray) or 3 (sRGB)");
            v[i].width = s.width;
            v[i].height = s.height;
            v[i].colorspace = s.channels == 1 ? GRAY : SRGB;
            v[i].quality = s.quality;
            v[i].format = s.format;
            v[i].src_kind = s.on_device ? SRC_DEV_U8 : SRC_HOST_U8;
            v[i].src = s.pixels;
        }
        std::vector<EncOut> r = encode_queue(v);
        jp2lam_status first = JP2LAM_OK;
        for (size_t i = 0; i < n; i++) {
            items[i].status = r[i].status;
            items[i].out = nullptr;
            items[i].out_len = 0;
            if (r[i].status == OK) {
                items[i].out = copy_out(r[i].bytes);
                items[i].out_len = r[i].bytes.size();
            } else if (first == JP2LAM_OK) {
                first = fail(r[i].status, r[i].message);
            }
        }
        return first;
    });
}

jp2lam_status jp2lam_decode_many(jp2lam_decode_item* items, size_t n) {
    return guarded([&] {
        if (!items && n) return fail(INVALID_INPUT, "null argument");
        std::vector<DecItem> v(n);
        for (size_t i = 0; i < n; i++) {
            v[i].bytes = items[i].bytes;
            v[i].len = items[i].len;
            v[i].target = items[i].on_device ? DST_DEV_U8 : DST_HOST_U8;
            v[i].dev_out = items[i].pixels;
            v[i].dev_cap = items[i].cap;
        }
        std::vector<DecOut> r = decode_queue(v);
        jp2lam_status first = JP2LAM_OK;
        for (size_t i = 0; i < n; i++) {
            items[i].status = r[i].status;
            items[i].width = r[i].width;
            items[i].height = r[i].height;
            items[i].channels = (uint32_t)r[i].ncomp;
            if (r[i].status == OK && !items[i].on_device) {
                if (r[i].pixels.size() > items[i].cap) {
                    items[i].status = IO;
                    if (first == JP2LAM_OK) first = fail(IO, "output buffer too small");
                    continue;
                }
                std::memcpy(items[i].pixels, r[i].pixels.data(), r[i].pixels.size());
            }
            if (r[i].status != OK && first == JP2LAM_OK) first = fail(r[i].status, r[i].message);
        }
        return first;
    });
}

// ---- stage-level entry points -----------------------------------------------------

jp2lam_status jp2g_forward_dwt(const uint8_t* pixels, uint32_t width, uint32_t height, uint32_t channels,
                               uint32_t quality, void* out) {
    return guarded([&] {
        if (channels != 1 && channels != 3) return fail(INVALID_INPUT, "channels must be 1 or 3");
        Plan p = build_plan(width, height, channels == 1 ? GRAY : SRGB, (int)quality, 0);
        Engine& e = engine();
        std::lock_guard<std::mutex> lk(e.mu);
        hipStream_t st = e.stream();
        const size_t n = (size_t)width * height;
        std::vector<uint32_t> rw, rh;
        resolution_ladder(width, height, p.levels, rw, rh);
        uint8_t* din;
        char* planes;
        char* ll;
        J2L_HIP(hipMalloc((void**)&din, n * channels));
        J2L_HIP(hipMalloc((void**)&planes, n * 4 * channels * 2));
        J2L_HIP(hipMalloc((void**)&ll, ((width + 1) / 2) * ((height + 1) / 2) * 4 * 2 + 256));
        J2L_HIP(hipMemcpyAsync(din, pixels, n * channels, hipMemcpyHostToDevice, st));
        const bool irrev = quality < 100;
        PixSeg* dps;
        DwtSeg* dds;
        J2L_HIP(hipMalloc((void**)&dps, sizeof(PixSeg)));
        J2L_HIP(hipMalloc((void**)&dds, sizeof(DwtSeg)));
        const size_t llhalf = ((width + 1) / 2) * ((height + 1) / 2) * 4;
        for (uint32_t c = 0; c < channels; c++) {
            char* X = planes + (size_t)c * n * 4 * 2;
            char* P = X + n * 4;
            PixSeg ps{};
            ps.src = (uint64_t)(uintptr_t)din;
            ps.dst = (uint64_t)(uintptr_t)X;
            ps.n = n;
            ps.channels = (int)channels;
            ps.comp = (int)c;
            ps.irrev = irrev;
            J2L_HIP(hipMemcpyAsync(dps, &ps, sizeof(ps), hipMemcpyHostToDevice, st));
            launch_prepare(dps, 1, pix_blocks(n), st);
            if (p.levels == 0) J2L_HIP(hipMemcpyAsync(P, X, n * 4, hipMemcpyDeviceToDevice, st));
            for (int k = 0; k < p.levels; k++) {
                const int Lv = p.levels;
                DwtSeg s{};
                s.in = (uint64_t)(uintptr_t)(k == 0 ? X : ll + ((k - 1) & 1) * llhalf);
                s.in_stride = k == 0 ? (int)width : (int)rw[Lv - k];
                s.plane = (uint64_t)(uintptr_t)P;
                s.pstride = (int)width;
                s.hx = rw[Lv - k - 1];  // the plain Mallat layout
                const bool last = k == Lv - 1;
                s.out = (uint64_t)(uintptr_t)(last ? P : ll + (k & 1) * llhalf);
                s.out_stride = last ? (int)width : (int)rw[Lv - k - 1];
                s.rw = (int)rw[Lv - k];
                s.rh = (int)rh[Lv - k];
                s.tiles_x = (s.rw + 63) / 64;
                J2L_HIP(hipMemcpyAsync(dds, &s, sizeof(s), hipMemcpyHostToDevice, st));
                launch_fwd_level(irrev, dds, 1, (uint32_t)s.tiles_x * (uint32_t)((s.rh + 63) / 64), st);
            }
            J2L_HIP(hipMemcpyAsync((char*)out + (size_t)c * n * 4, P, n * 4, hipMemcpyDeviceToHost, st));
            J2L_HIP(hipStreamSynchronize(st));
        }
        (void)hipFree(din);
        (void)hipFree(planes);
        (void)hipFree(ll);
        (void)hipFree(dps);
        (void)hipFree(dds);
        return JP2LAM_OK;
    });
}

jp2lam_status jp2g_inverse_dwt(const void* planes, uint32_t width, uint32_t height, uint32_t channels, uint32_t levels,
                               uint32_t irreversible, uint8_t* out) {
    return guarded([&] {
        if (channels != 1 && channels != 3) return fail(INVALID_INPUT, "channels must be 1 or 3");
        Engine& e = engine();
        std::lock_guard<std::mutex> lk(e.mu);
        hipStream_t st = e.stream();
        const size_t n = (size_t)width * height;
        std::vector<uint32_t> rw, rh;
        resolution_ladder(width, height, (int)levels, rw, rh);
        char* P;
        char* O;
        char* ll;
        uint8_t* d8;
        const size_t llhalf = ((width + 1) / 2) * ((height + 1) / 2) * 4;
        J2L_HIP(hipMalloc((void**)&P, n * 4 * channels));
        J2L_HIP(hipMalloc((void**)&O, n * 4 * channels));
        J2L_HIP(hipMalloc((void**)&ll, llhalf * 2 + 256));
        J2L_HIP(hipMalloc((void**)&d8, n * channels));
        J2L_HIP(hipMemcpyAsync(P, planes, n * 4 * channels, hipMemcpyHostToDevice, st));
        DwtSeg* dds;
        PixSeg* dps;
        J2L_HIP(hipMalloc((void**)&dds, sizeof(DwtSeg)));
        J2L_HIP(hipMalloc((void**)&dps, sizeof(PixSeg)));
        for (uint32_t c = 0; c < channels; c++) {
            char* Pc = P + (size_t)c * n * 4;
            char* Oc = O + (size_t)c * n * 4;
            if (levels == 0) J2L_HIP(hipMemcpyAsync(Oc, Pc, n * 4, hipMemcpyDeviceToDevice, st));
            for (uint32_t k = 0; k < levels; k++) {
                const uint32_t r = k + 1;
                DwtSeg s{};
                s.in = (uint64_t)(uintptr_t)(k == 0 ? Pc : ll + ((k - 1) & 1) * llhalf);
                s.in_stride = k == 0 ? (int)width : (int)rw[r - 1];
                s.plane = (uint64_t)(uintptr_t)Pc;
                s.pstride = (int)width;
                s.hx = rw[r - 1];  // the plain Mallat layout
                const bool last = k == levels - 1;
                s.out = (uint64_t)(uintptr_t)(last ? Oc : ll + (k & 1) * llhalf);
                s.out_stride = last ? (int)width : (int)rw[r];
                s.rw = (int)rw[r];
                s.rh = (int)rh[r];
                s.tiles_x = (s.rw + 63) / 64;
                J2L_HIP(hipMemcpyAsync(dds, &s, sizeof(s), hipMemcpyHostToDevice, st));
                launch_inv_level(irreversible != 0, dds, 1, (uint32_t)s.tiles_x * (uint32_t)((s.rh + 63) / 64), st);
                J2L_HIP(hipStreamSynchronize(st));
            }
        }
        PixSeg f{};
        f.src = (uint64_t)(uintptr_t)O;
        f.src1 = channels == 3 ? (uint64_t)(uintptr_t)(O + n * 4) : 0;
        f.src2 = channels == 3 ? (uint64_t)(uintptr_t)(O + 2 * n * 4) : 0;
        f.dst = (uint64_t)(uintptr_t)d8;
        f.n = n;
        f.channels = (int)channels;
        f.irrev = irreversible != 0;
        J2L_HIP(hipMemcpyAsync(dps, &f, sizeof(f), hipMemcpyHostToDevice, st));
        launch_finalize(dps, 1, pix_blocks(n), st);
        J2L_HIP(hipMemcpyAsync(out, d8, n * channels, hipMemcpyDeviceToHost, st));
        J2L_HIP(hipStreamSynchronize(st));
        (void)hipFree(P);
        (void)hipFree(O);
        (void)hipFree(ll);
        (void)hipFree(d8);
        (void)hipFree(dds);
        (void)hipFree(dps);
        return JP2LAM_OK;
    });
}

jp2lam_status jp2g_t1_encode_block(const int32_t* coeffs, uint32_t w, uint32_t h, uint32_t band, uint32_t mb,
                                   int32_t* info, int32_t* passes, uint8_t* bytes, size_t cap) {
    return guarded([&] {
        if (w == 0 || h == 0 || w > 64 || h > 64 || band > 3) return fail(INVALID_INPUT, "block must be 1..64 x 1..64");
        Engine& e = engine();
        std::lock_guard<std::mutex> lk(e.mu);
        hipStream_t st = e.stream();
        auto tabs = make_t1_tables();
        const uint32_t slot = w * h * 4 + 1024;
        int32_t* dplane;
        uint8_t* dout;
        BlockJob* djob;
        BlockResult* dres;
        uint32_t* dsnap;
        uint16_t* dns;
        uint32_t* dtab;
        uint32_t* dsched;
        J2L_HIP(hipMalloc((void**)&dplane, 4 * w * h));
        J2L_HIP(hipMalloc((void**)&dout, slot));
        J2L_HIP(hipMalloc((void**)&djob, sizeof(BlockJob)));
        J2L_HIP(hipMalloc((void**)&dres, sizeof(BlockResult)));
        J2L_HIP(hipMalloc((void**)&dsnap, 4 * kMaxPasses));
        J2L_HIP(hipMalloc((void**)&dns, 2 * kMaxPasses));
        J2L_HIP(hipMalloc((void**)&dtab, 4 * tabs.size()));
        J2L_HIP(hipMalloc((void**)&dsched, 4 * t1_sched_words(1)));
        BlockJob j{};
        j.plane = (uint64_t)(uintptr_t)dplane;
        j.stride = w;
        j.w = (uint8_t)w;
        j.h = (uint8_t)h;
        j.band = (uint8_t)band;
        j.mb = (uint8_t)mb;
        j.step = 0.0f;
        j.out_off = 0;
        j.out_cap = slot;
        J2L_HIP(hipMemcpyAsync(dplane, coeffs, 4 * w * h, hipMemcpyHostToDevice, st));
        J2L_HIP(hipMemcpyAsync(djob, &j, sizeof(j), hipMemcpyHostToDevice, st));
        J2L_HIP(hipMemcpyAsync(dtab, tabs.data(), 4 * tabs.size(), hipMemcpyHostToDevice, st));
        launch_t1_encode(djob, 1, dtab, dout, dres, dsnap, dns, dsched, st);
        BlockResult r;
        uint32_t snaps[kMaxPasses];
        uint16_t ns[kMaxPasses];
        J2L_HIP(hipMemcpyAsync(&r, dres, sizeof(r), hipMemcpyDeviceToHost, st));
        J2L_HIP(hipMemcpyAsync(snaps, dsnap, sizeof(snaps), hipMemcpyDeviceToHost, st));
        J2L_HIP(hipMemcpyAsync(ns, dns, sizeof(ns), hipMemcpyDeviceToHost, st));
        J2L_HIP(hipStreamSynchronize(st));
        std::vector<uint8_t> seg(r.nbytes);
        if (r.nbytes) J2L_HIP(hipMemcpy(seg.data(), dout, r.nbytes, hipMemcpyDeviceToHost));
        (void)hipFree(dplane);
        (void)hipFree(dout);
        (void)hipFree(djob);
        (void)hipFree(dres);
        (void)hipFree(dsnap);
        (void)hipFree(dns);
        (void)hipFree(dtab);
        (void)hipFree(dsched);
        if (r.overflow) return fail(ENCODE_FAILED, "code-block segment exceeded its device slot");
        info[0] = r.mbp;
        info[1] = r.zbp;
        info[2] = r.cls;
        info[3] = r.npasses;
        info[4] = (int32_t)r.nbytes;
        uint32_t prev = 0, sum = 0, sig = 0, sig_sp = 0;
        const int top = r.mbp - 1;
        for (int i = 0; i < r.npasses; i++) {
            uint32_t len = snaps[i] > prev ? snaps[i] - prev : 0;
            prev = snaps[i];
            if (i == r.npasses - 1) len = r.nbytes > sum ? r.nbytes - sum : 0;
            sum += len;
            int kind = i == 0 ? 2 : (i - 1) % 3;
            if (kind == 0) sig_sp = sig;
            passes[5 * i + 0] = kind;
            passes[5 * i + 1] = i == 0 ? top : top - 1 - (i - 1) / 3;
            passes[5 * i + 2] = ns[i];
            passes[5 * i + 3] = kind == 2 ? (i == 0 ? 0 : (int)sig) : (int)sig_sp;
            passes[5 * i + 4] = (int)len;
            sig += ns[i];
        }
        if (r.nbytes > cap) return fail(ENCODE_FAILED, "output buffer too small");
        if (r.nbytes) std::memcpy(bytes, seg.data(), r.nbytes);
        return JP2LAM_OK;
    });
}

jp2lam_status jp2g_t1_decode_block(uint32_t w, uint32_t h, uint32_t band, uint32_t mb, uint32_t zbp, uint32_t passes,
                                   const uint8_t* data, size_t len, int32_t* out) {
    return guarded([&] {
        if (w == 0 || h == 0 || w > 64 || h > 64 || band > 3) return fail(INVALID_INPUT, "block must be 1..64 x 1..64");
        if (passes > 0 && (zbp > mb || mb - zbp > 31 || mb - zbp == 0 || passes > 1 + 3 * (mb - zbp - 1)))
            return fail(DECODE_FAILED, "invalid block parameters");
        Engine& e = engine();
        std::lock_guard<std::mutex> lk(e.mu);
        hipStream_t st = e.stream();
        auto tabs = make_t1_tables();
        int32_t* dplane;
        uint8_t* ddata;
        DecodeJob* djob;
        uint32_t* dtab;
        J2L_HIP(hipMalloc((void**)&dplane, 4 * w * h));
        J2L_HIP(hipMalloc((void**)&ddata, len + 64));
        J2L_HIP(hipMalloc((void**)&djob, sizeof(DecodeJob)));
        J2L_HIP(hipMalloc((void**)&dtab, 4 * tabs.size()));
        DecodeJob j{};
        j.plane = (uint64_t)(uintptr_t)dplane;
        j.stride = w;
        j.w = (uint8_t)w;
        j.h = (uint8_t)h;
        j.band = (uint8_t)band;
        j.mb = (uint8_t)mb;
        j.step = 0.0f;
        j.data_off = 0;
        j.data_len = (uint32_t)len;
        j.zbp = (int16_t)(passes ? (int)zbp : -1);
        j.passes = (uint8_t)passes;
        if (len) J2L_HIP(hipMemcpyAsync(ddata, data, len, hipMemcpyHostToDevice, st));
        J2L_HIP(hipMemcpyAsync(djob, &j, sizeof(j), hipMemcpyHostToDevice, st));
        J2L_HIP(hipMemcpyAsync(dtab, tabs.data(), 4 * tabs.size(), hipMemcpyHostToDevice, st));
        launch_t1_decode(djob, 1, dtab, ddata, st);
        J2L_HIP(hipMemcpyAsync(out, dplane, 4 * w * h, hipMemcpyDeviceToHost, st));
        J2L_HIP(hipStreamSynchronize(st));
        (void)hipFree(dplane);
        (void)hipFree(ddata);
        (void)hipFree(djob);
        (void)hipFree(dtab);
        return JP2LAM_OK;
    });
}

void jp2g_set_debug(int on) { engine().keep_debug = on != 0; }
void jp2g_set_t1_mode(int mode, uint32_t lane_min_blocks) {
    Engine& e = engine();
    std::lock_guard<std::mutex> g(e.mu);
    e.t1_mode = mode < 0 || mode > 4 ? 0 : mode;
    if (lane_min_blocks) e.lane_min_blocks = lane_min_blocks;
}
int jp2g_last_t1_lanes(void) { return engine().last_t1_lanes ? 1 : 0; }
int jp2g_last_t1_coder(void) { return engine().last_t1_coder; }
void jp2g_set_queue(uint32_t chunk_blocks, int engines) { set_queue_config(chunk_blocks, engines); }
void jp2g_get_queue(uint32_t* chunk_blocks, int* engines) {
    const QueueConfig c = queue_config();
    if (chunk_blocks) *chunk_blocks = c.chunk_blocks;
    if (engines) *engines = c.engines;
}
void jp2g_set_hybrid_frac(float frac) {
    Engine& e = engine();
    std::lock_guard<std::mutex> g(e.mu);
    e.hybrid_frac = frac > 0.0f && frac < 1.0f ? frac : 0.55f;
}
void jp2g_set_graphs(int on) { set_graphs(on != 0); }
void jp2g_graph_stats(uint64_t* captures, uint64_t* launches) { graph_stats(captures, launches); }
void jp2g_set_counters(int on) { set_t1_counting(on != 0); }
void jp2g_counters(uint64_t* out7) {
    if (out7) t1_counters(out7);
}
void jp2g_reset_counters(void) { reset_t1_counters(); }
uint32_t jp2g_last_block_symbols(int i, int kept_only) {
    const auto& v = engine().last_blocks();
    if (i < 0 || (size_t)i >= v.size()) return 0;
    return kept_only ? v[i].nsym_kept : v[i].nsym;
}
int jp2g_last_block_count(void) { return (int)engine().last_blocks().size(); }
void jp2g_last_block_info(int i, int32_t* info) {
    const BlockStat& b = engine().last_blocks()[i];
    int32_t v[16] = {b.comp, b.res, b.band, b.x0, b.y0, b.x1, b.y1, b.mbp, b.zbp, b.cls,
                     b.keep, b.npasses, (int32_t)b.nbytes, (int32_t)b.body_len, (int32_t)b.max_mag, (int32_t)b.nnz};
    std::memcpy(info, v, sizeof(v));
}
void jp2g_last_block_passes(int i, int32_t* out) {
    Engine& e = engine();
    const BlockStat& b = e.last_blocks()[i];
    const uint32_t* sp = e.last_snaps().data() + (size_t)i * kMaxPasses;
    const uint16_t* ns = e.last_newsigs().data() + (size_t)i * kMaxPasses;
    uint32_t prev = 0, sum = 0, sig = 0, sig_sp = 0;
    const int top = b.mbp - 1;
    for (int k = 0; k < b.npasses; k++) {
        uint32_t len = sp[k] > prev ? sp[k] - prev : 0;
        prev = sp[k];
        if (k == b.npasses - 1) len = b.nbytes > sum ? b.nbytes - sum : 0;
        sum += len;
        int kind = k == 0 ? 2 : (k - 1) % 3;
        if (kind == 0) sig_sp = sig;
        out[5 * k + 0] = kind;
        out[5 * k + 1] = k == 0 ? top : top - 1 - (k - 1) / 3;
        out[5 * k + 2] = ns[k];
        out[5 * k + 3] = kind == 2 ? (k == 0 ? 0 : (int)sig) : (int)sig_sp;
        out[5 * k + 4] = (int)len;
        sig += ns[k];
    }
}
void jp2g_last_block_bytes(int i, uint8_t* out) {
    Engine& e = engine();
    const BlockStat& b = e.last_blocks()[i];
    std::memcpy(out, e.last_segments().data() + e.last_segment_offsets()[i], b.nbytes);
}
double jp2g_last_lambda(int g) {
    const auto& l = engine().last_lambdas();
    return g >= 0 && (size_t)g < l.size() ? l[g] : 0.0;
}
void jp2g_last_timings(double* out4) {
    Engine& e = engine();
    out4[0] = e.t_dwt_ms;
    out4[1] = e.t_t1_ms;
    out4[2] = e.t_pcrd_ms;
    out4[3] = e.t_host_ms;
}
void jp2g_tables(uint16_t* qe, uint8_t* mps, uint8_t* nmps, uint8_t* nlps, uint8_t* zc, uint8_t* sc, uint8_t* spb) {
    expanded_tables(qe, mps, nmps, nlps, zc, sc, spb);
}
int jp2g_device_available(void) { return device_available(nullptr) ? 1 : 0; }

// ---- device buffers ------------------------------------------------------------
void* jp2lam_device_alloc(size_t bytes) {
    void* p = nullptr;
    guarded([&] {
        if (!device_available(nullptr)) return fail(IO, "no HIP device");
        J2L_HIP(hipMalloc(&p, bytes ? bytes : 1));
        return JP2LAM_OK;
    });
    return p;
}
jp2lam_status jp2lam_device_copy(void* dst, const void* src, size_t bytes, int kind) {
    return guarded([&] {
        if ((!dst || !src) && bytes) return fail(INVALID_INPUT, "null argument");
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                : kind == 1 ? hipMemcpyDeviceToHost
                                : kind == 2 ? hipMemcpyDeviceToDevice
                                            : hipMemcpyDefault;
        if (kind < 0 || kind > 2) return fail(INVALID_INPUT, "copy kind must be 0, 1 or 2");
        if (bytes) J2L_HIP(hipMemcpy(dst, src, bytes, k));
        return JP2LAM_OK;
    });
}
void jp2lam_device_free(void* ptr) {
    if (ptr) (void)hipFree(ptr);
}

// ---- serving -------------------------------------------------------------------
struct jp2lam_server {
    std::unique_ptr<j2l::Server> s;
};

jp2lam_server* jp2lam_server_new(int slots, size_t mem_budget_mib, int max_batch) {
    jp2lam_server* out = nullptr;
    guarded([&] {
        out = new jp2lam_server{std::make_unique<Server>(slots, mem_budget_mib << 20, max_batch)};
        return JP2LAM_OK;
    });
    return out;
}
jp2lam_server* jp2lam_server_new_simulated(int slots, size_t mem_budget_mib, int max_batch, uint32_t delay_us) {
    jp2lam_server* out = nullptr;
    guarded([&] {
        out = new jp2lam_server{std::make_unique<Server>(slots, mem_budget_mib << 20, max_batch, true, delay_us)};
        return JP2LAM_OK;
    });
    return out;
}
int jp2lam_server_cancel(jp2lam_server* srv, uint64_t id) {
    int rc = -1;
    guarded([&] {
        if (!srv) return fail(INVALID_INPUT, "null argument");
        rc = srv->s->cancel(id);
        return JP2LAM_OK;
    });
    return rc;
}
uint64_t jp2lam_server_submit_encode(jp2lam_server* srv, const uint8_t* pixels, uint32_t width, uint32_t height,
                                     uint32_t channels, uint8_t quality, uint8_t format, uint8_t on_device) {
    uint64_t id = 0;
    guarded([&] {
        if (!srv) return fail(INVALID_INPUT, "null argument");
        id = srv->s->submit_encode(pixels, width, height, channels, quality, format, on_device != 0);
        return JP2LAM_OK;
    });
    return id;
}
uint64_t jp2lam_server_submit_decode(jp2lam_server* srv, const uint8_t* bytes, size_t len) {
    uint64_t id = 0;
    guarded([&] {
        if (!srv) return fail(INVALID_INPUT, "null argument");
        id = srv->s->submit_decode(bytes, len);
        return JP2LAM_OK;
    });
    return id;
}
int jp2lam_server_wait(jp2lam_server* srv, uint64_t id, int timeout_ms, jp2lam_served* out) {
    int rc = -1;
    guarded([&] {
        if (!srv || !out) return fail(INVALID_INPUT, "null argument");
        std::shared_ptr<Request> r = srv->s->wait(id, timeout_ms);
        if (!r) {
            rc = 0;
            return JP2LAM_OK;
        }
        std::memset(out, 0, sizeof(*out));
        using ms = std::chrono::duration<double, std::milli>;
        out->queue_ms = ms(r->t_start - r->t_submit).count();
        out->run_ms = ms(r->t_done - r->t_start).count();
        out->total_ms = ms(r->t_done - r->t_submit).count();
        out->batch_size = (uint32_t)r->batch_size;
        if (r->is_encode) {
            out->status = r->eout.status;
            if (r->eout.status == OK) {
                out->data = copy_out(r->eout.bytes);
                out->len = r->eout.bytes.size();
            } else {
                fail(r->eout.status, r->eout.message);
            }
            out->width = r->enc.width;
            out->height = r->enc.height;
            out->channels = r->enc.colorspace == GRAY ? 1 : 3;
        } else {
            out->status = r->dout.status;
            if (r->dout.status == OK) {
                out->data = copy_out(r->dout.pixels);
                out->len = r->dout.pixels.size();
            } else {
                fail(r->dout.status, r->dout.message);
            }
            out->width = r->dout.width;
            out->height = r->dout.height;
            out->channels = (uint32_t)r->dout.ncomp;
        }
        rc = 1;
        return JP2LAM_OK;
    });
    return rc;
}
void jp2lam_server_stats(jp2lam_server* srv, uint64_t* out4) {
    if (!srv || !out4) return;
    ServerStats st = srv->s->stats();
    out4[0] = st.submitted;
    out4[1] = st.completed;
    out4[2] = st.batches;
    out4[3] = st.max_batch;
}
void jp2lam_server_faults(jp2lam_server* srv, uint64_t* out3) {
    if (!srv || !out3) return;
    ServerStats st = srv->s->stats();
    out3[0] = st.rejected;
    out3[1] = st.cancelled;
    out3[2] = st.device_faults;
}
void jp2lam_server_free(jp2lam_server* srv) { delete srv; }

}  // extern "C"
