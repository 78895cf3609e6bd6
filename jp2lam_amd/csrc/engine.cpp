// Device engine: plans a batch of images, lays out one device arena for it,
// launches every stage once for the whole batch (per DWT level, Tier-1,
// PCRD, packing) and assembles the codestreams on the host.
#include "engine.hpp"

#include <functional>
#include <memory>

#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>

namespace j2l {

namespace {

constexpr int kTile = 64;

struct Layout {
    size_t used = 0;
    size_t add(size_t bytes, size_t align = 256) {
        used = (used + align - 1) & ~(align - 1);
        size_t o = used;
        used += bytes;
        return o;
    }
};

// A table bound for device offset dst, staged at hoff in the page-locked
// arena.
struct Staged {
    size_t dst, hoff, n;
    const void* src;
};
// Fill the arena and upload: consecutive tables whose device and host
// offsets advance alike go in one copy (the gaps between them are
// alignment padding on both sides).
void copy_staged(const std::vector<Staged>& staged, char* D, char* P, hipStream_t st) {
    for (size_t i = 0; i < staged.size();) {
        size_t j = i + 1;
        while (j < staged.size() && staged[j].dst >= staged[i].dst &&
               staged[j].dst - staged[i].dst == staged[j].hoff - staged[i].hoff)
            j++;
        for (size_t k = i; k < j; k++) std::memcpy(P + staged[k].hoff, staged[k].src, staged[k].n);
        const size_t n = staged[j - 1].hoff + staged[j - 1].n - staged[i].hoff;
        J2L_HIP(hipMemcpyAsync(D + staged[i].dst, P + staged[i].hoff, n, hipMemcpyHostToDevice, st));
        i = j;
    }
}


// context.rs:39-70: integer BT.601 YCbCr -> RGB for the hidden colour spaces
void yuv_to_rgb(const int32_t* src, size_t n, std::vector<int32_t>& out) {
    out.resize(3 * n);
    auto cl = [](int v) { return std::min(std::max(v, 0), 255); };
    for (size_t i = 0; i < n; i++) {
        int y = src[i], d = src[n + i] - 128, e = src[2 * n + i] - 128;
        out[i] = cl(y + ((91881 * e) >> 16));
        out[n + i] = cl(y - ((22554 * d + 46802 * e) >> 16));
        out[2 * n + i] = cl(y + ((116130 * d) >> 16));
    }
}

}  // namespace

void Arena::reserve(size_t bytes) {
    if (bytes <= cap) return;
    size_t want = std::max(bytes, cap + cap / 2);
    if (base) {
        if (pinned) J2L_HIP(hipHostFree(base));
        else J2L_HIP(hipFree(base));
        base = nullptr;
        cap = 0;
    }
    if (pinned) J2L_HIP(hipHostMalloc((void**)&base, want, hipHostMallocDefault));
    else J2L_HIP(hipMalloc((void**)&base, want));
    cap = want;
}

Arena::~Arena() {
    if (base) {
        if (pinned) (void)hipHostFree(base);
        else (void)hipFree(base);
    }
}

bool device_available(std::string* why) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        if (why) *why = e != hipSuccess ? hipGetErrorString(e) : "no HIP device";
        return false;
    }
    return true;
}

Engine::Engine() {
    std::string why;
    if (!device_available(&why)) throw Error(IO, "jp2lam_amd needs a HIP device (gfx950): " + why);
    J2L_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    pin_.pinned = true;
    auto t = make_t1_tables();
    J2L_HIP(hipMalloc((void**)&tables_, sizeof(uint32_t) * t.size()));
    J2L_HIP(hipMemcpy(tables_, t.data(), sizeof(uint32_t) * t.size(), hipMemcpyHostToDevice));
}

namespace {
enum T1Coder : int { kT1Wave = 1, kT1Lanes = 2, kT1Hybrid = 4 };
struct T1TableEntry {
    int bucket;  // bit length of the batch's code-block count
    int coder;
    float frac;
};
#include "t1_tuning.inc"
}  // namespace

Engine::T1Choice Engine::tuned_t1_choice(int nblocks) {
    const int bucket = nblocks > 0 ? 32 - __builtin_clz((unsigned)nblocks) : 0;
    T1Choice c{kT1Wave, 0.55f};
    for (const auto& e : kT1Table)
        if (e.bucket <= bucket) c = {e.coder, e.frac};
    return c;
}

// ---- launch sequences as HIP graphs -------------------------------------------------
// Off by default: on the MI355X the graphs measured no faster than the
// eager launches (a 64x64 page through a 2550x3300 one, within 2 %), and the
// capture of each new shape added latency under mixed-shape serving.
static std::atomic<bool> g_graphs{false};
static std::atomic<uint64_t> g_captures{0}, g_replays{0};
void set_graphs(bool on) { g_graphs.store(on); }
void graph_stats(uint64_t* captures, uint64_t* replays) {
    if (captures) *captures = g_captures.load();
    if (replays) *replays = g_replays.load();
}

GraphCache::~GraphCache() { clear(); }

void GraphCache::clear() {
    for (auto& kv : exec_) (void)hipGraphExecDestroy(kv.second);
    exec_.clear();
    seen_.clear();
}

void GraphCache::run(hipStream_t st, const std::function<void(bool, std::vector<uint64_t>&)>& seq) {
    std::vector<uint64_t> sig;
    seq(false, sig);
    if (g_graphs.load()) {
        auto it = exec_.find(sig);
        if (it != exec_.end()) {
            J2L_HIP(hipGraphLaunch(it->second, st));
            g_replays++;
            return;
        }
        // capture a sequence on its second occurrence; one-off batch shapes
        // (and the first call of each) run eagerly
        if (++seen_[sig] >= 2 && exec_.size() < kMaxGraphs) {
            J2L_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            std::vector<uint64_t> again;
            hipGraph_t g = nullptr;
            try {
                seq(true, again);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            J2L_HIP(hipStreamEndCapture(st, &g));
            hipGraphExec_t x = nullptr;
            const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            J2L_HIP(e);
            exec_.emplace(std::move(sig), x);
            g_captures++;
            g_replays++;
            J2L_HIP(hipGraphLaunch(x, st));
            return;
        }
        if (seen_.size() > 4 * kMaxGraphs) seen_.clear();
    }
    std::vector<uint64_t> unused;
    seq(true, unused);
}

Engine::~Engine() {
    graphs_.clear();
    if (tables_) (void)hipFree(tables_);
    if (ev_up_) (void)hipEventDestroy(ev_up_);
    if (ev_side_) (void)hipEventDestroy(ev_side_);
    if (side_) (void)hipStreamDestroy(side_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::ensure_side_stream() {
    if (side_) return;
    J2L_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    J2L_HIP(hipEventCreateWithFlags(&ev_up_, hipEventDisableTiming));
    J2L_HIP(hipEventCreateWithFlags(&ev_side_, hipEventDisableTiming));
}

namespace {
std::atomic<bool> g_counting{false};
std::mutex g_counter_mu;
uint64_t g_counters[7] = {0, 0, 0, 0, 0, 0, 0};
void add_t1_counters(const uint64_t* c) {
    std::lock_guard<std::mutex> g(g_counter_mu);
    for (int i = 0; i < 7; i++) g_counters[i] += c[i];
}
}  // namespace

void set_t1_counting(bool on) { g_counting.store(on); }
bool t1_counting() { return g_counting.load(); }
void t1_counters(uint64_t* out7) {
    std::lock_guard<std::mutex> g(g_counter_mu);
    for (int i = 0; i < 7; i++) out7[i] = g_counters[i];
}
void reset_t1_counters() {
    std::lock_guard<std::mutex> g(g_counter_mu);
    for (int i = 0; i < 7; i++) g_counters[i] = 0;
}

Engine& engine() {
    static Engine* e = new Engine();
    return *e;
}

// ============================================================================
// Encode
// ============================================================================
std::vector<EncOut> Engine::encode(const std::vector<EncItem>& items) {
    std::vector<EncOut> outs(items.size());
    struct Img {
        int idx;
        Plan plan;
        std::vector<uint32_t> rw, rh;
        PlaneLayout pl;  // coefficient plane columns and stride
        size_t n;
        int kind;
        const void* src;
        std::vector<int32_t> conv;
        size_t in_off = 0, in_bytes = 0;
        size_t x_off[3], p_off[3], ll_off[3][2];
        uint32_t blk_first = 0, blk_count = 0;
        size_t pack_cap = 0;
        uint32_t group0 = 0;
        int mseg = -1;  // index of its quality-metric segment
    };
    std::vector<Img> imgs;
    for (size_t i = 0; i < items.size(); i++) {
        const EncItem& it = items[i];
        try {
            if (it.width == 0 || it.height == 0) throw Error(INVALID_INPUT, "image dimensions must be non-zero");
            if (it.width > 65535 || it.height > 65535)
                throw Error(UNSUPPORTED_FEATURE, "image dimensions above 65535 are not supported by the device path");
            Img im;
            im.idx = (int)i;
            im.plan = build_plan(it.width, it.height, it.colorspace, it.quality, it.format);
            im.n = (size_t)it.width * it.height;
            im.kind = it.src_kind;
            im.src = it.src;
            const int nc = im.plan.ncomp;
            if (it.src_kind == SRC_HOST_I32) {
                const int32_t* s = (const int32_t*)it.src;
                for (int c = 0; c < nc; c++)
                    for (size_t k = 0; k < im.n; k++)
                        if (s[(size_t)c * im.n + k] < 0 || s[(size_t)c * im.n + k] > 255)
                            throw Error(INVALID_INPUT, "component " + std::to_string(c) + " contains samples outside 0..=255");
                if (it.colorspace == YUV || it.colorspace == YCBCR) {
                    yuv_to_rgb(s, im.n, im.conv);
                    im.src = im.conv.data();
                }
            }
            resolution_ladder(it.width, it.height, im.plan.levels, im.rw, im.rh);
            im.pl = plane_layout(it.width, im.plan.levels, true);
            // backend.rs:371-421: metrics cover gray and sRGB only
            if (it.metrics && !im.plan.lossless && it.colorspace != GRAY && it.colorspace != SRGB)
                throw Error(ENCODE_FAILED, "quality metrics not implemented for this colorspace configuration");
            imgs.push_back(std::move(im));
        } catch (const Error& e) {
            outs[i].status = e.status;
            outs[i].message = e.what();
        }
    }
    if (imgs.empty()) return outs;

    // ---- host descriptors --------------------------------------------------------
    std::vector<BlockJob> jobs;
    std::vector<PcrdJob> pjobs;
    std::vector<double> raw_floor;
    std::vector<uint32_t> img_first, blk_img;
    std::vector<uint8_t> blk_comp, blk_res;
    Layout L;
    uint32_t ngroups = 0;
    for (auto& im : imgs) {
        const Plan& p = im.plan;
        const int nc = p.ncomp;
        const size_t ch = nc;
        if (im.kind == SRC_HOST_U8) im.in_bytes = im.n * ch;
        else if (im.kind == SRC_HOST_I32) im.in_bytes = im.n * ch * 4;
        if (im.in_bytes) im.in_off = L.add(im.in_bytes);
        const size_t llw = ll_stride((p.width + 1) / 2), llh = (p.height + 1) / 2;
        for (int c = 0; c < nc; c++) {
            im.x_off[c] = L.add(im.n * 4);
            im.p_off[c] = L.add((size_t)im.pl.stride * p.height * 4);
            im.ll_off[c][0] = L.add(llw * llh * 4);
            im.ll_off[c][1] = L.add(llw * llh * 4);
        }
        im.group0 = ngroups;
        double lraw = quality_to_lambda(p.quality, p.width * p.height);
        double lfloor = quality_to_lambda(1, p.width * p.height);
        for (int c = 0; c < nc; c++) {
            raw_floor.push_back(lraw);
            raw_floor.push_back(lfloor);
        }
        ngroups += nc;
        im.blk_first = (uint32_t)jobs.size();
        img_first.push_back(im.blk_first);
        for (int r = 0; r <= p.levels; r++)
            for (int c = 0; c < nc; c++)
                for (int b = r ? HL : LL; b <= (r ? HH : LL); b++) {
                    Rect rc = band_rect(p.width, p.height, p.levels, r, b);
                    const SubbandQuant* sq = nullptr;
                    for (auto& q : p.quants)
                        if (q.res == r && q.band == b) sq = &q;
                    const int mb = subband_mb(p, c, r, b);
                    const float step = p.lossless ? 0.0f : quant_step_f32(b, sq->exponent, sq->mantissa);
                    const int lv = b == LL ? p.levels : std::max(p.levels - r, 0);
                    const double nrm = norm97((uint32_t)lv, b);
                    const double swb = (nrm * nrm) * band_bias(b, p.quality);
                    const double cw = component_weight(p, c);
                    const double dstep = quant_step_f64(b, sq->exponent, sq->mantissa);
                    for (int by = rc.y0; by < rc.y1; by += p.cbh)
                        for (int bx = rc.x0; bx < rc.x1; bx += p.cbw) {
                            BlockJob j;
                            j.plane = 0;  // patched after the arena is placed
                            j.stride = im.pl.stride;
                            j.x0 = (uint16_t)plane_x(im.pl, rc, r, b, bx);
                            j.y0 = (uint16_t)by;
                            j.w = (uint8_t)(std::min(bx + p.cbw, rc.x1) - bx);
                            j.h = (uint8_t)(std::min(by + p.cbh, rc.y1) - by);
                            j.band = (uint8_t)b;
                            j.mb = (uint8_t)mb;
                            j.step = step;
                            j.out_cap = ((uint32_t)j.w * j.h * 3 + 256 + 63) & ~63u;
                            j.out_off = 0;
                            jobs.push_back(j);
                            pjobs.push_back({dstep, swb, cw, im.group0 + (uint32_t)c, b == LL ? 1u : 0u});
                            blk_img.push_back((uint32_t)(&im - imgs.data()));
                            blk_comp.push_back((uint8_t)c);
                            blk_res.push_back((uint8_t)r);
                        }
                }
        im.blk_count = (uint32_t)jobs.size() - im.blk_first;
    }
    img_first.push_back((uint32_t)jobs.size());
    const int nb = (int)jobs.size();
    // slots
    size_t slot_total = 0;
    for (auto& j : jobs) {
        j.out_off = (uint32_t)slot_total;
        slot_total += j.out_cap;
    }
    std::vector<uint64_t> pack_base(imgs.size());
    size_t pack_total = 0;
    for (size_t k = 0; k < imgs.size(); k++) {
        size_t cap = 0;
        for (uint32_t b = imgs[k].blk_first; b < imgs[k].blk_first + imgs[k].blk_count; b++) cap += jobs[b].out_cap;
        pack_base[k] = pack_total;
        imgs[k].pack_cap = cap;
        pack_total += (cap + 255) & ~(size_t)255;
    }
    if (slot_total > 0xffffffffull) throw Error(ENCODE_FAILED, "batch too large for one launch");

    // The readback group (selection, results, totals, lambdas) is laid out
    // in the order of its host buffers, so one copy returns it.
    const size_t o_sel = L.add(sizeof(Selection) * nb);
    const size_t o_res = L.add(sizeof(BlockResult) * nb);
    const size_t o_totals = L.add(4 * imgs.size());
    const size_t o_lambda = L.add(8 * ngroups);
    const size_t o_snaps = L.add(sizeof(uint32_t) * kMaxPasses * nb);
    const size_t o_ns = L.add(sizeof(uint16_t) * kMaxPasses * nb);
    const size_t o_slots = L.add(slot_total);
    const size_t o_hull = L.add(sizeof(HullPt) * (kMaxPasses + 1) * nb);
    const size_t o_hulln = L.add(nb);
    const size_t o_offs = L.add(4 * nb);
    const size_t o_packed = L.add(pack_total);
    const size_t o_sched = L.add(4 * t1_sched_words(nb));
    const bool counting = t1_counting();
    const size_t o_counts = counting ? L.add(sizeof(uint32_t) * kMaxPasses * nb) : 0;
    // Lane-per-block and hybrid Tier-1 (large batches) code formed symbol
    // streams, bounded per block by the band's Mb + 2 bit-planes; an
    // overflow falls back to the wave-per-block kernel below.  The coder is
    // forced (t1_mode 1 / 2 / 4), a fixed threshold (3), or taken from the
    // offline-tuned table by the batch's size class (0).  Nothing is timed
    // or probed here, so a batch's cost never depends on what ran before.
    int coder = t1_mode == 1 || t1_mode == 2 || t1_mode == 4 ? t1_mode : 1;
    float frac = hybrid_frac;
    if (t1_mode == 3) {
        coder = (uint32_t)nb >= lane_min_blocks ? 2 : 1;
    } else if (t1_mode == 0) {
        const T1Choice c = tuned_t1_choice(nb);
        coder = c.coder;
        frac = c.frac;
    }
    const bool lanes = coder != 1;  // lane (2) and hybrid (4) code formed streams
    size_t o_pend = 0, o_soff = 0, o_pused = 0, o_pool = 0, o_lists = 0;
    unsigned long long pool_cap = 0;
    if (lanes) {
        for (auto& j : jobs) pool_cap += t1_stream_bound((uint32_t)j.mb + 2u);
        o_pend = L.add(sizeof(uint32_t) * kMaxPasses * nb);
        o_soff = L.add(sizeof(uint32_t) * nb);
        o_pused = L.add(8);
        o_pool = L.add(pool_cap + 64);
        o_lists = L.add(4 * (2 + 2 * (size_t)nb));
    }

    // Segment tables, split by sample type (f32 for 9/7, i32 for 5/3) so
    // every launch is homogeneous.  Addresses are arena offsets until the
    // arena is placed.
    int maxL = 0;
    for (auto& im : imgs) maxL = std::max(maxL, im.plan.levels);
    std::vector<PixSeg> pseg_k[2];
    uint32_t pblk_k[2] = {0, 0};
    std::vector<DwtSeg> fseg_k[2][kMaxLevels];
    uint32_t ftile_k[2][kMaxLevels] = {};
    for (auto& im : imgs) {
        const int kind = im.plan.lossless ? 0 : 1;
        for (int c = 0; c < im.plan.ncomp; c++) {
            PixSeg s{};
            s.src = im.in_off;
            s.dst = im.x_off[c];
            s.n = im.n;
            s.channels = im.plan.ncomp;
            s.comp = c;
            s.planar_i32 = im.kind == SRC_HOST_I32;
            s.irrev = kind;
            s.blk_start = pblk_k[kind];
            pblk_k[kind] += pix_blocks(im.n);
            pseg_k[kind].push_back(s);
        }
    }
    for (int k = 0; k < maxL; k++) {
        for (auto& im : imgs) {
            const Plan& p = im.plan;
            if (k >= p.levels) continue;
            const int kind = p.lossless ? 0 : 1;
            const int Lv = p.levels;
            const uint32_t rw = im.rw[Lv - k], rh = im.rh[Lv - k];
            for (int c = 0; c < p.ncomp; c++) {
                DwtSeg s{};
                s.in = k == 0 ? im.x_off[c] : im.ll_off[c][(k - 1) & 1];
                s.in_stride = k == 0 ? (int)p.width : (int)ll_stride(rw);
                s.plane = im.p_off[c];
                s.pstride = (int)im.pl.stride;
                s.hx = im.pl.hx[Lv - k];
                const bool last = k == Lv - 1;
                s.out = last ? im.p_off[c] : im.ll_off[c][k & 1];
                s.out_stride = last ? (int)im.pl.stride : (int)ll_stride(im.rw[Lv - k - 1]);
                s.rw = (int)rw;
                s.rh = (int)rh;
                s.tiles_x = (int)((rw + kTile - 1) / kTile);
                s.tile_start = ftile_k[kind][k];
                ftile_k[kind][k] += (uint32_t)s.tiles_x * ((rh + kTile - 1) / kTile);
                fseg_k[kind][k].push_back(s);
            }
        }
    }
    // The uploaded tables follow each other in the order they are staged in
    // host memory (the same sizes and alignment on both sides), so their
    // copies merge into one.
    const size_t o_jobs = L.add(sizeof(BlockJob) * nb);
    const size_t o_pjobs = L.add(sizeof(PcrdJob) * nb);
    const size_t o_gmax = L.add(8 * ngroups);
    const size_t o_rawf = L.add(16 * ngroups);
    const size_t o_imgfirst = L.add(4 * img_first.size());
    const size_t o_blkimg = L.add(4 * nb);
    const size_t o_packbase = L.add(8 * imgs.size());
    size_t o_pseg_k[2], o_fseg_k[2][kMaxLevels];
    for (int kind = 0; kind < 2; kind++) {
        // (an empty table takes no space, as in the host staging; its
        // launch is skipped)
        o_pseg_k[kind] = L.add(sizeof(PixSeg) * pseg_k[kind].size());
        for (int k = 0; k < maxL; k++) o_fseg_k[kind][k] = L.add(sizeof(DwtSeg) * fseg_k[kind][k].size());
    }
    // encode_with_psnr (backend.rs:305-433): the lossy images' blocks are
    // dequantised as truncated in place, inverted into their x planes
    // (free after the forward DWT) and compared cell by cell with the input.
    std::vector<uint32_t> m_list;
    std::vector<DwtSeg> m_iseg[kMaxLevels];
    uint32_t m_itiles[kMaxLevels] = {};
    std::vector<MetricSeg> m_seg;
    std::vector<char> m_orig_dev;
    uint32_t m_cells = 0;
    for (auto& im : imgs) {
        if (!items[im.idx].metrics || im.plan.lossless) continue;
        const Plan& p = im.plan;
        const int Lv = p.levels;
        for (uint32_t b = im.blk_first; b < im.blk_first + im.blk_count; b++) m_list.push_back(b);
        for (int k = 0; k < Lv; k++) {
            const int r = k + 1;
            for (int c = 0; c < p.ncomp; c++) {
                DwtSeg s{};
                s.in = k == 0 ? im.p_off[c] : im.ll_off[c][(k - 1) & 1];
                s.in_stride = k == 0 ? (int)im.pl.stride : (int)ll_stride(im.rw[r - 1]);
                s.plane = im.p_off[c];
                s.pstride = (int)im.pl.stride;
                s.hx = im.pl.hx[r];
                const bool last = k == Lv - 1;
                s.out = last ? im.x_off[c] : im.ll_off[c][k & 1];
                s.out_stride = last ? (int)p.width : (int)ll_stride(im.rw[r]);
                s.rw = (int)im.rw[r];
                s.rh = (int)im.rh[r];
                s.tiles_x = (s.rw + kTile - 1) / kTile;
                s.tile_start = m_itiles[k];
                m_itiles[k] += (uint32_t)s.tiles_x * (uint32_t)((s.rh + kTile - 1) / kTile);
                m_iseg[k].push_back(s);
            }
        }
        MetricSeg ms{};
        for (int c = 0; c < 3; c++) ms.rec[c] = Lv ? im.x_off[std::min(c, p.ncomp - 1)] : im.p_off[std::min(c, p.ncomp - 1)];
        const bool dev_src = im.kind == SRC_DEV_U8;
        ms.orig = dev_src ? (uint64_t)(uintptr_t)im.src : im.in_off;
        ms.width = p.width;
        ms.height = p.height;
        ms.cells_x = (p.width + 7) / 8;
        ms.cells_y = (p.height + 7) / 8;
        ms.cell_start = m_cells;
        ms.ncomp = p.ncomp;
        ms.orig_i32 = im.kind == SRC_HOST_I32;
        m_cells += ms.cells_x * ms.cells_y;
        im.mseg = (int)m_seg.size();
        m_seg.push_back(ms);
        m_orig_dev.push_back(dev_src);
    }
    size_t o_mlist = 0, o_mseg = 0, o_mcells = 0, o_mout = 0, o_miseg[kMaxLevels] = {};
    if (!m_seg.empty()) {
        o_mlist = L.add(4 * m_list.size());
        for (int k = 0; k < kMaxLevels; k++)
            if (!m_iseg[k].empty()) o_miseg[k] = L.add(sizeof(DwtSeg) * m_iseg[k].size());
        o_mseg = L.add(sizeof(MetricSeg) * m_seg.size());
        o_mcells = L.add(16 * (size_t)m_cells);
        o_mout = L.add(16 * m_seg.size());
    }

    dev_.reserve(L.used);
    char* D = dev_.base;
    auto A = [&](size_t off) { return (uint64_t)(uintptr_t)(D + off); };
    for (size_t b = 0; b < jobs.size(); b++) jobs[b].plane = A(imgs[blk_img[b]].p_off[blk_comp[b]]);
    for (int kind = 0; kind < 2; kind++) {
        for (auto& s : pseg_k[kind]) s.dst = A(s.dst);
        for (int k = 0; k < maxL; k++)
            for (auto& s : fseg_k[kind][k]) {
                s.in = A(s.in);
                s.plane = A(s.plane);
                s.out = A(s.out);
            }
    }
    for (int k = 0; k < kMaxLevels; k++)
        for (auto& s : m_iseg[k]) {
            s.in = A(s.in);
            s.plane = A(s.plane);
            s.out = A(s.out);
        }
    for (size_t q = 0; q < m_seg.size(); q++) {
        for (auto& r : m_seg[q].rec) r = A(r);
        if (!m_orig_dev[q]) m_seg[q].orig = A(m_seg[q].orig);
    }
    {
        size_t q[2] = {0, 0};
        for (auto& im : imgs) {
            const int kind = im.plan.lossless ? 0 : 1;
            for (int c = 0; c < im.plan.ncomp; c++)
                pseg_k[kind][q[kind]++].src = im.kind == SRC_DEV_U8 ? (uint64_t)(uintptr_t)im.src : A(im.in_off);
        }
    }

    // ---- uploads ---------------------------------------------------------------------
    // The descriptor tables go through the page-locked arena, so each copy is
    // one asynchronous DMA (a pageable source makes the runtime stage it
    // synchronously); the same arena then receives the readback.
    hipStream_t st = stream_;
    auto up = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) J2L_HIP(hipMemcpyAsync(D + off, src, bytes, hipMemcpyHostToDevice, st));
    };
    Layout H;
    std::vector<Staged> staged;
    auto stage = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) staged.push_back({off, H.add(bytes), bytes, src});
    };
    std::vector<unsigned long long> gmax0(ngroups, 0ull);
    stage(o_jobs, jobs.data(), sizeof(BlockJob) * nb);
    stage(o_pjobs, pjobs.data(), sizeof(PcrdJob) * nb);
    stage(o_gmax, gmax0.data(), 8 * ngroups);
    stage(o_rawf, raw_floor.data(), 16 * ngroups);
    stage(o_imgfirst, img_first.data(), 4 * img_first.size());
    stage(o_blkimg, blk_img.data(), 4 * nb);
    stage(o_packbase, pack_base.data(), 8 * imgs.size());
    for (int kind = 0; kind < 2; kind++) {
        stage(o_pseg_k[kind], pseg_k[kind].data(), sizeof(PixSeg) * pseg_k[kind].size());
        for (int k = 0; k < maxL; k++) stage(o_fseg_k[kind][k], fseg_k[kind][k].data(), sizeof(DwtSeg) * fseg_k[kind][k].size());
    }
    const size_t h_sel = H.add(sizeof(Selection) * nb);
    const size_t h_res = H.add(sizeof(BlockResult) * nb);
    const size_t h_tot = H.add(4 * imgs.size());
    const size_t h_lam = H.add(8 * ngroups);
    pin_.reserve(H.used);  // once per call: DMAs from it stay in flight
    char* P = pin_.base;
    copy_staged(staged, D, P, st);
    for (auto& im : imgs)
        if (im.in_bytes) up(im.in_off, im.src, im.in_bytes);

    // ---- device pipeline -------------------------------------------------------------
    // The launch sequence is written once and run either eagerly or through
    // the engine's graph cache: each step appends what identifies it (kind,
    // arena offsets, counts, grid sizes) to `sig`, and issues it when `go`.
    Events& ev = ev_;
    auto dwt_seq = [&](bool go, std::vector<uint64_t>& sig) {
        sig.insert(sig.end(), {1, (uint64_t)(uintptr_t)D, (uint64_t)(uintptr_t)tables_});
        for (int kind = 0; kind < 2; kind++) {
            if (pseg_k[kind].empty()) continue;
            sig.insert(sig.end(), {2, (uint64_t)kind, o_pseg_k[kind], pseg_k[kind].size(), pblk_k[kind]});
            if (go) launch_prepare((const PixSeg*)(D + o_pseg_k[kind]), (int)pseg_k[kind].size(), pblk_k[kind], st);
            for (int k = 0; k < maxL; k++)
                if (!fseg_k[kind][k].empty()) {
                    sig.insert(sig.end(), {3, (uint64_t)kind, o_fseg_k[kind][k], fseg_k[kind][k].size(), ftile_k[kind][k]});
                    if (go)
                        launch_fwd_level(kind == 1, (const DwtSeg*)(D + o_fseg_k[kind][k]), (int)fseg_k[kind][k].size(),
                                         ftile_k[kind][k], st);
                }
        }
        // images without any decomposition level: the coefficient plane is the
        // prepared plane itself
        for (auto& im : imgs)
            if (im.plan.levels == 0)
                for (int c = 0; c < im.plan.ncomp; c++) {
                    sig.insert(sig.end(), {4, im.p_off[c], im.x_off[c], im.n});
                    if (go) J2L_HIP(hipMemcpyAsync(D + im.p_off[c], D + im.x_off[c], im.n * 4, hipMemcpyDeviceToDevice, st));
                }
    };
    bool any_lossy = false, any_lossless = false;
    for (auto& im : imgs) (im.plan.lossless ? any_lossless : any_lossy) = true;
    auto t1_seq = [&](bool go, std::vector<uint64_t>& sig, int which) {
        sig.insert(sig.end(), {5, (uint64_t)which, (uint64_t)nb, pool_cap, (uint64_t)(frac * 1e6f)});
        if (which != 1) {
            // formed-stream coders: 2 = lane per block, 4 = hybrid; the pool
            // counter is cleared by a kernel (no host copy or memset node)
            if (go) {
                launch_zero_u64((unsigned long long*)(D + o_pused), 1, st);
                launch_t1_form((const BlockJob*)(D + o_jobs), nb, tables_, (BlockResult*)(D + o_res),
                               (uint32_t*)(D + o_pend), (uint16_t*)(D + o_ns), (uint32_t*)(D + o_soff),
                               (uint8_t*)(D + o_pool), (unsigned long long*)(D + o_pused), pool_cap, st);
                if (which == 2)
                    launch_t1_mq((const BlockJob*)(D + o_jobs), nb, tables_, (const uint8_t*)(D + o_pool),
                                 (const uint32_t*)(D + o_soff), (const uint32_t*)(D + o_pend), (BlockResult*)(D + o_res),
                                 (uint32_t*)(D + o_snaps), (uint8_t*)(D + o_slots), st);
                else
                    launch_t1_hybrid((const BlockJob*)(D + o_jobs), nb, tables_, (const uint8_t*)(D + o_pool),
                                     (const uint32_t*)(D + o_soff), (const uint32_t*)(D + o_pend),
                                     (BlockResult*)(D + o_res), (uint32_t*)(D + o_snaps), (uint8_t*)(D + o_slots),
                                     (uint32_t*)(D + o_lists), frac, st);
            }
        } else if (go) {
            launch_t1_encode((const BlockJob*)(D + o_jobs), nb, tables_, (uint8_t*)(D + o_slots),
                             (BlockResult*)(D + o_res), (uint32_t*)(D + o_snaps), (uint16_t*)(D + o_ns),
                             (uint32_t*)(D + o_sched), st);
        }
    };
    auto pcrd_seq = [&](bool go, std::vector<uint64_t>& sig) {
        // PCRD: lossless images keep every pass; lossy ones go through the hull.
        sig.insert(sig.end(), {6, (uint64_t)any_lossy, (uint64_t)any_lossless, ngroups, imgs.size()});
        if (any_lossy && go)
            launch_pcrd(nb, false, (int)ngroups, (const BlockResult*)(D + o_res), (const uint32_t*)(D + o_snaps),
                        (const uint16_t*)(D + o_ns), (const PcrdJob*)(D + o_pjobs), (HullPt*)(D + o_hull),
                        (uint8_t*)(D + o_hulln), (unsigned long long*)(D + o_gmax), (const double*)(D + o_rawf),
                        (double*)(D + o_lambda), (Selection*)(D + o_sel), st);
        if (any_lossless) {
            // lossless blocks: select "all passes" for the lossless images only
            for (auto& im : imgs) {
                if (!im.plan.lossless) continue;
                sig.insert(sig.end(), {7, im.blk_first, im.blk_count});
                if (go)
                    launch_pcrd((int)im.blk_count, true, 0, (const BlockResult*)(D + o_res) + im.blk_first,
                                (const uint32_t*)(D + o_snaps) + (size_t)im.blk_first * kMaxPasses, nullptr,
                                (const PcrdJob*)(D + o_pjobs) + im.blk_first, nullptr, nullptr, nullptr, nullptr,
                                nullptr, (Selection*)(D + o_sel) + im.blk_first, st);
            }
        }
        if (go)
            for (int pass = 0; pass < 2; pass++)
                launch_pack(nb, (int)imgs.size(), (const uint32_t*)(D + o_imgfirst), (const uint32_t*)(D + o_blkimg),
                            (const uint64_t*)(D + o_packbase), (const BlockJob*)(D + o_jobs),
                            (const Selection*)(D + o_sel), (uint32_t*)(D + o_offs), (uint32_t*)(D + o_totals),
                            (const uint8_t*)(D + o_slots), (uint8_t*)(D + o_packed), st, pass == 0);
    };
    // every arena offset the steps use, once: the signature then names the
    // whole launch sequence exactly
    const std::vector<uint64_t> layout_sig = {
        L.used, o_jobs, o_pjobs, o_gmax, o_rawf, o_imgfirst, o_blkimg, o_packbase, o_sel, o_res, o_totals, o_lambda,
        o_snaps, o_ns, o_slots, o_hull, o_hulln, o_offs, o_packed, o_sched, o_pend, o_soff, o_pused, o_pool, o_lists};
    // Each stage is one graph (or its launches), with the stage events
    // recorded between them on the stream.
    auto staged_run = [&](const std::function<void(bool, std::vector<uint64_t>&)>& seq, bool graph) {
        if (!graph) {
            std::vector<uint64_t> unused;
            seq(true, unused);
            return;
        }
        graphs_.run(st, [&](bool go, std::vector<uint64_t>& sig) {
            sig.insert(sig.end(), layout_sig.begin(), layout_sig.end());
            seq(go, sig);
        });
    };
    int run = coder;
    auto coding = [&](int which, bool graph) {
        J2L_HIP(hipEventRecord(ev.e[1], st));
        staged_run([&](bool go, std::vector<uint64_t>& sig) { t1_seq(go, sig, which); }, graph);
        J2L_HIP(hipEventRecord(ev.e[2], st));
        staged_run(pcrd_seq, graph);
        J2L_HIP(hipEventRecord(ev.e[3], st));
    };
    J2L_HIP(hipEventRecord(ev.e[0], st));
    staged_run(dwt_seq, true);
    coding(run, true);
    std::vector<uint32_t> pass_syms, pass_snaps;
    const Selection* sel = nullptr;
    const BlockResult* res = nullptr;
    // after the pipeline: counters (their own launches, outside the
    // sequence), readback, and the stage times from the engine's events
    auto finish = [&] {
        // counters: MQ decisions per pass, formed again from the coefficients
        // after the stages timed above (the coding kernels are untouched)
        if (counting) {
            J2L_HIP(hipMemsetAsync(D + o_counts, 0, sizeof(uint32_t) * kMaxPasses * nb, st));
            launch_t1_count((const BlockJob*)(D + o_jobs), nb, tables_, (uint32_t*)(D + o_counts), st);
            pass_syms.resize((size_t)kMaxPasses * nb);
            pass_snaps.resize((size_t)kMaxPasses * nb);
            J2L_HIP(hipMemcpyAsync(pass_syms.data(), D + o_counts, 4 * pass_syms.size(), hipMemcpyDeviceToHost, st));
            J2L_HIP(hipMemcpyAsync(pass_snaps.data(), D + o_snaps, 4 * pass_snaps.size(), hipMemcpyDeviceToHost, st));
        }
        // ---- readback ----------------------------------------------------------------------
        if (o_res - o_sel == h_res - h_sel && o_totals - o_sel == h_tot - h_sel && o_lambda - o_sel == h_lam - h_sel) {
            J2L_HIP(hipMemcpyAsync(P + h_sel, D + o_sel, o_lambda + 8 * ngroups - o_sel, hipMemcpyDeviceToHost, st));
        } else {
            J2L_HIP(hipMemcpyAsync(P + h_sel, D + o_sel, sizeof(Selection) * nb, hipMemcpyDeviceToHost, st));
            J2L_HIP(hipMemcpyAsync(P + h_res, D + o_res, sizeof(BlockResult) * nb, hipMemcpyDeviceToHost, st));
            J2L_HIP(hipMemcpyAsync(P + h_tot, D + o_totals, 4 * imgs.size(), hipMemcpyDeviceToHost, st));
            J2L_HIP(hipMemcpyAsync(P + h_lam, D + o_lambda, 8 * ngroups, hipMemcpyDeviceToHost, st));
        }
        J2L_HIP(hipStreamSynchronize(st));
        t_dwt_ms = ev.ms(0, 1);
        t_t1_ms = ev.ms(1, 2);
        t_pcrd_ms = ev.ms(2, 3);
        sel = (const Selection*)(P + h_sel);
        res = (const BlockResult*)(P + h_res);
    };
    finish();
    if (run != 1) {
        bool pool_full = false;
        for (int b = 0; b < nb && !pool_full; b++) pool_full = res[b].overflow == 2;
        if (pool_full) {
            // the formed streams overflowed their pool: code with the wave
            // coder instead (eagerly; the DWT output is still in place)
            run = 1;
            coding(1, false);
            finish();
        }
    }
    last_t1_lanes = run == 2;
    last_t1_coder = run;
    const uint32_t* totals = (const uint32_t*)(P + h_tot);
    std::vector<Selection> selv(sel, sel + nb);
    std::vector<BlockResult> resv(res, res + nb);
    if (counting) {
        // t1.rs:270-336, :593-641: blocks, empty blocks, MQ decisions,
        // passes by kind (cleanup first, then SP, MR, CL per plane) and pass
        // bytes (the per-pass lengths of the single terminated segment)
        uint64_t c[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int b = 0; b < nb; b++) {
            const BlockResult& r = resv[b];
            c[0]++;
            if (r.mbp == 0) {
                c[1]++;
                continue;
            }
            const uint32_t* sy = pass_syms.data() + (size_t)b * kMaxPasses;
            const uint32_t* sn = pass_snaps.data() + (size_t)b * kMaxPasses;
            uint32_t prev = 0, sum = 0;
            for (int k = 0; k < r.npasses; k++) {
                c[2] += sy[k];
                const int kind = k == 0 ? 2 : (k - 1) % 3;  // 0 SP, 1 MR, 2 CL
                c[kind == 2 ? 3 : kind == 0 ? 4 : 5]++;
                uint32_t len = sn[k] > prev ? sn[k] - prev : 0;
                prev = sn[k];
                if (k == r.npasses - 1) len = r.nbytes > sum ? r.nbytes - sum : 0;
                sum += len;
                c[6] += len;
            }
        }
        add_t1_counters(c);
    }
    std::vector<uint32_t> totv(totals, totals + imgs.size());
    last_lambdas_.assign((const double*)(P + h_lam), (const double*)(P + h_lam) + ngroups);
    std::vector<double> m_out(2 * m_seg.size());
    if (!m_seg.empty()) {
        up(o_mlist, m_list.data(), 4 * m_list.size());
        for (int k = 0; k < kMaxLevels; k++) up(o_miseg[k], m_iseg[k].data(), sizeof(DwtSeg) * m_iseg[k].size());
        up(o_mseg, m_seg.data(), sizeof(MetricSeg) * m_seg.size());
        J2L_HIP(hipEventRecord(ev.e[4], st));
        launch_metric_dequant((const BlockJob*)(D + o_jobs), (const BlockResult*)(D + o_res), (const Selection*)(D + o_sel),
                              (const uint32_t*)(D + o_mlist), (int)m_list.size(), st);
        for (int k = 0; k < kMaxLevels; k++)
            if (!m_iseg[k].empty())
                launch_inv_level(true, (const DwtSeg*)(D + o_miseg[k]), (int)m_iseg[k].size(), m_itiles[k], st);
        launch_metric_cells((const MetricSeg*)(D + o_mseg), (int)m_seg.size(), m_cells, (double*)(D + o_mcells), st);
        launch_metric_reduce((const MetricSeg*)(D + o_mseg), (int)m_seg.size(), (const double*)(D + o_mcells),
                             (double*)(D + o_mout), st);
        J2L_HIP(hipEventRecord(ev.e[5], st));
        J2L_HIP(hipMemcpyAsync(m_out.data(), D + o_mout, 8 * m_out.size(), hipMemcpyDeviceToHost, st));
    }
    size_t packed_bytes = 0;
    for (size_t k = 0; k < imgs.size(); k++) packed_bytes += totv[k];
    // The selected bytes come back into the page-locked arena (one DMA per
    // image, no staging).  The synchronisation above has retired every copy
    // from the arena, and its readback is already copied out.
    pin_.reserve(packed_bytes);
    const uint8_t* packed_host = (const uint8_t*)pin_.base;
    {
        size_t o = 0;
        for (size_t k = 0; k < imgs.size(); k++) {
            if (totv[k])
                J2L_HIP(hipMemcpyAsync(pin_.base + o, D + o_packed + pack_base[k], totv[k], hipMemcpyDeviceToHost, st));
            o += totv[k];
        }
    }
    if (keep_debug) {
        last_snaps_.resize((size_t)kMaxPasses * nb);
        last_newsigs_.resize((size_t)kMaxPasses * nb);
        last_segments_.resize(slot_total);
        J2L_HIP(hipMemcpyAsync(last_snaps_.data(), D + o_snaps, 4 * last_snaps_.size(), hipMemcpyDeviceToHost, st));
        J2L_HIP(hipMemcpyAsync(last_newsigs_.data(), D + o_ns, 2 * last_newsigs_.size(), hipMemcpyDeviceToHost, st));
        J2L_HIP(hipMemcpyAsync(last_segments_.data(), D + o_slots, slot_total, hipMemcpyDeviceToHost, st));
        last_seg_offs_.resize(nb);
        for (int b = 0; b < nb; b++) last_seg_offs_[b] = jobs[b].out_off;
    }
    // The packet headers need only the selection: they are formed while the
    // bodies come back, and the first image to write its bodies waits for
    // the copies (once, for every image and thread).
    std::once_flag landed;
    auto bodies_landed = [&] { std::call_once(landed, [&] { J2L_HIP(hipStreamSynchronize(st)); }); };
    struct LandGuard {  // an exception out of the host work still waits for the copies into the arena
        hipStream_t s;
        std::once_flag* f;
        ~LandGuard() { std::call_once(*f, [this] { (void)hipStreamSynchronize(s); }); }
    } land_guard{st, &landed};

    // ---- host Tier-2 + container --------------------------------------------------------
    auto th0 = std::chrono::steady_clock::now();
    last_blocks_.clear();
    std::vector<size_t> pkoff(imgs.size());
    {
        size_t pk = 0;
        for (size_t k = 0; k < imgs.size(); k++) {
            pkoff[k] = pk;
            pk += totv[k];
        }
    }
    if (keep_debug) {
        for (size_t k = 0; k < imgs.size(); k++) {
            const Img& im = imgs[k];
            bool overflow = false;
            for (uint32_t b = im.blk_first; b < im.blk_first + im.blk_count; b++) overflow |= resv[b].overflow != 0;
            if (overflow) continue;  // the image fails below; no statistics
            for (uint32_t b = 0; b < im.blk_count; b++) {
                const BlockJob& j = jobs[im.blk_first + b];
                const BlockResult& r = resv[im.blk_first + b];
                const Selection& sl = selv[im.blk_first + b];
                BlockStat bs{};
                bs.mbp = r.mbp;
                bs.zbp = r.zbp;
                bs.cls = r.cls;
                bs.npasses = r.npasses;
                bs.keep = sl.keep;
                bs.band = j.band;
                bs.comp = blk_comp[im.blk_first + b];
                bs.res = blk_res[im.blk_first + b];
                // Mallat column (the plane stores the detail bands at
                // im.pl.hx[res], aligned)
                bs.x0 = j.band == HL || j.band == HH ? j.x0 - im.pl.hx[bs.res] + im.rw[bs.res - 1] : j.x0;
                bs.y0 = j.y0;
                bs.x1 = (uint16_t)(bs.x0 + j.w);
                bs.y1 = (uint16_t)(j.y0 + j.h);
                bs.nbytes = r.nbytes;
                bs.body_len = sl.body_len;
                bs.max_mag = r.max_mag;
                bs.nnz = r.nnz;
                if (counting) {
                    const uint32_t* sy = pass_syms.data() + (size_t)(im.blk_first + b) * kMaxPasses;
                    for (int k = 0; k < r.npasses; k++) {
                        bs.nsym += sy[k];
                        if (k < sl.keep) bs.nsym_kept += sy[k];
                    }
                }
                last_blocks_.push_back(bs);
            }
        }
    }
    // Tier-2 and containers of one image, written once into its output: the
    // packet headers go to a side buffer first, which fixes the codestream
    // length the JP2 header needs, then header, packets and bodies are
    // appended in order.
    auto assemble = [&](size_t k) {
        Img& im = imgs[k];
        EncOut& out = outs[im.idx];
        const Plan& p = im.plan;
        const uint8_t* body = packed_host + pkoff[k];
        bool overflow = false;
        for (uint32_t b = im.blk_first; b < im.blk_first + im.blk_count; b++) overflow |= resv[b].overflow != 0;
        if (overflow) {
            out.status = ENCODE_FAILED;
            out.message = "code-block segment exceeded its device slot";
            return;
        }
        std::vector<CodedBlock> cbs(im.blk_count);
        for (uint32_t b = 0; b < im.blk_count; b++) {
            const Selection& sl = selv[im.blk_first + b];
            cbs[b] = {sl.keep, sl.zbp, sl.body_len};
        }
        // packets in LRCP order: resolution, component (t2.rs:409-470)
        struct Pkt {
            size_t h0, h1, b0, blen;
        };
        std::vector<Pkt> pkts;
        pkts.reserve((size_t)(p.levels + 1) * p.ncomp);
        std::vector<uint8_t> hdrs;
        hdrs.reserve(64 * (size_t)(p.levels + 1) * p.ncomp + 1024);
        uint32_t bi = 0;
        size_t bo = 0;
        for (int r = 0; r <= p.levels; r++) {
            for (int c = 0; c < p.ncomp; c++) {
                BandBlocks bands[3];
                int nbands = 0;
                uint32_t first = bi;
                for (int b = r ? HL : LL; b <= (r ? HH : LL); b++) {
                    Rect rc = band_rect(p.width, p.height, p.levels, r, b);
                    int cx = rc.x1 > rc.x0 ? (rc.x1 - rc.x0 + p.cbw - 1) / p.cbw : 0;
                    int cy = rc.y1 > rc.y0 ? (rc.y1 - rc.y0 + p.cbh - 1) / p.cbh : 0;
                    bands[nbands++] = {cx, cy, cbs.data() + bi};
                    bi += (uint32_t)(cx * cy);
                }
                const size_t h0 = hdrs.size();
                packet_header(bands, nbands, hdrs);
                size_t blen = 0;
                for (uint32_t b = first; b < bi; b++) blen += cbs[b].seg_len;
                pkts.push_back({h0, hdrs.size(), bo, blen});
                bo += blen;
            }
        }
        std::vector<uint8_t> head;
        write_codestream_head(p, hdrs.size() + bo, head);
        const size_t cs_len = head.size() + hdrs.size() + bo + 2;  // + EOC
        bodies_landed();
        std::vector<uint8_t>& o = out.bytes;
        o.clear();
        o.reserve(cs_len + (p.format == 0 ? 128 : 0));
        if (p.format == 0) write_jp2_head(p, cs_len, o);
        o.insert(o.end(), head.begin(), head.end());
        for (const Pkt& q : pkts) {
            o.insert(o.end(), hdrs.begin() + (ptrdiff_t)q.h0, hdrs.begin() + (ptrdiff_t)q.h1);
            o.insert(o.end(), body + q.b0, body + q.b0 + q.blen);
        }
        write_codestream_tail(o);
        if (items[im.idx].metrics) {
            if (im.mseg < 0) {
                // backend.rs:308-314: lossless encodes
                out.psnr_db = HUGE_VAL;
                out.ssim = 1.0;
            } else {
                // backend.rs:423-430
                const MetricSeg& ms = m_seg[im.mseg];
                const double mse = m_out[2 * im.mseg] / (double)((size_t)ms.width * ms.height * ms.ncomp);
                out.psnr_db = mse < 1e-10 ? 100.0 : 20.0 * std::log10(255.0 / std::sqrt(mse));
                const size_t full = (size_t)(ms.width / 8) * (ms.height / 8);
                out.ssim = full ? m_out[2 * im.mseg + 1] / (double)full : 1.0;
            }
        }
    };
    // Images are independent: a batch of several spreads them over host
    // threads (each writes only its own output).
    const size_t nthr = imgs.size() < 4 ? 1 : std::min<size_t>({imgs.size(), 8, std::max(1u, std::thread::hardware_concurrency())});
    if (nthr <= 1) {
        for (size_t k = 0; k < imgs.size(); k++) assemble(k);
    } else {
        std::atomic<size_t> next{0};
        std::vector<std::string> errs(imgs.size());
        auto work = [&] {
            for (size_t k; (k = next.fetch_add(1)) < imgs.size();) {
                try {
                    assemble(k);
                } catch (const std::exception& e) {
                    errs[k] = e.what();
                }
            }
        };
        std::vector<std::thread> ts;
        for (size_t t = 1; t < nthr; t++) ts.emplace_back(work);
        work();
        for (auto& t : ts) t.join();
        for (size_t k = 0; k < imgs.size(); k++)
            if (!errs[k].empty()) {
                outs[imgs[k].idx].status = ENCODE_FAILED;
                outs[imgs[k].idx].message = errs[k];
            }
    }
    bodies_landed();  // also when no image got that far; the debug copies and events are read later
    t_metrics_ms = m_seg.empty() ? 0.0 : ev.ms(4, 5);
    t_host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count();
    return outs;
}

// ============================================================================
// Content analysis
// ============================================================================
// content.rs:120-174: the histogram pass runs on the device; the statistics
// are formed on the host from the 257 counters in the reference's f32 order.
ContentOut Engine::analyze_content(const void* src, int src_kind, size_t n, bool gray) {
    ContentOut o;
    if (n == 0) return o;  // content.rs:155-157
    const int ch = gray ? 1 : 3;
    Layout L;
    size_t in_bytes = src_kind == SRC_HOST_U8 ? n * ch : src_kind == SRC_HOST_I32 ? n * ch * 4 : 0;
    const size_t in_off = in_bytes ? L.add(in_bytes) : 0;
    const size_t o_hist = L.add(8 * 257);
    dev_.reserve(L.used);
    pin_.reserve(8 * 257);
    char* D = dev_.base;
    hipStream_t st = stream_;
    if (in_bytes) J2L_HIP(hipMemcpyAsync(D + in_off, src, in_bytes, hipMemcpyHostToDevice, st));
    J2L_HIP(hipMemsetAsync(D + o_hist, 0, 8 * 257, st));
    launch_content_hist(src_kind == SRC_DEV_U8 ? src : (const void*)(D + in_off), n, ch, src_kind == SRC_HOST_I32, gray,
                        (unsigned long long*)(D + o_hist), st);
    J2L_HIP(hipMemcpyAsync(pin_.base, D + o_hist, 8 * 257, hipMemcpyDeviceToHost, st));
    J2L_HIP(hipStreamSynchronize(st));
    o = content_from_hist((const uint64_t*)pin_.base, n);
    return o;
}

// ============================================================================
// Decode
// ============================================================================
std::vector<DecOut> Engine::decode(const std::vector<DecItem>& items) {
    std::vector<DecOut> outs(items.size());
    struct Img {
        int idx;
        Jp2Info jp;
        CodestreamInfo ci;
        std::vector<DecBlock> blocks;
        std::vector<uint32_t> rw, rh;
        PlaneLayout pl;  // coefficient plane columns and stride
        size_t n;
        size_t pay_off, p_off[3], ll_off[3][2], o_off[3], u8_off;
        uint32_t job_first;
    };
    // Container, codestream and packet headers: each stream on its own, so a
    // batch of several spreads them over host threads (decode/t2.rs parses
    // one stream at a time).
    std::vector<Img> parsed(items.size());
    std::vector<char> ok(items.size(), 0);
    // A single image of four or more resolutions is decoded in two parts:
    // the packets of all but the two finest resolutions are parsed first and
    // their blocks start on the device (they hold the longest segments);
    // the rest are parsed meanwhile and coded on the side stream.  If the
    // first part's blocks fail a check, the image is parsed and checked
    // whole, as below, so the error is the one the whole parse reports.
    std::unique_ptr<PacketParser> split_pp;
    int split_res = -1;
    auto parse_one = [&](size_t i) {
        try {
            Img& im = parsed[i];
            im.idx = (int)i;
            if (items.size() == 1) {
                parse_decode_header(items[i].bytes, items[i].len, im.jp, im.ci);
                auto pp = std::make_unique<PacketParser>(im.ci, im.blocks);
                bool split = im.ci.levels >= 3;
                if (split) {
                    pp->run(im.ci.levels - 2);
                    try {
                        std::vector<DecBlock> first;
                        for (const auto& b : im.blocks)
                            if (b.res <= im.ci.levels - 2) first.push_back(b);
                        check_decode_blocks(im.ci, first);
                    } catch (const Error&) {
                        split = false;
                    }
                }
                if (split) {
                    split_pp = std::move(pp);
                    split_res = im.ci.levels - 2;
                } else {
                    pp->run(im.ci.levels);
                    pp->finish();
                    check_decode_blocks(im.ci, im.blocks);
                }
            } else {
                parse_for_decode(items[i].bytes, items[i].len, im.jp, im.ci, im.blocks);
            }
            im.n = (size_t)im.ci.width * im.ci.height;
            resolution_ladder(im.ci.width, im.ci.height, im.ci.levels, im.rw, im.rh);
            im.pl = plane_layout(im.ci.width, im.ci.levels, false);
            if (items[i].target == DST_DEV_U8 && items[i].dev_cap < im.n * im.ci.ncomp)
                throw Error(IO, "device output buffer too small");
            ok[i] = 1;
        } catch (const Error& e) {
            outs[i].status = e.status;
            outs[i].message = e.what();
        }
    };
    const size_t nthr = items.size() < 4 ? 1 : std::min<size_t>({items.size(), 8, std::max(1u, std::thread::hardware_concurrency())});
    if (nthr <= 1) {
        for (size_t i = 0; i < items.size(); i++) parse_one(i);
    } else {
        // other exceptions (out of memory) fail the whole call, as they do
        // when the streams are parsed one after another
        std::atomic<size_t> next{0};
        std::exception_ptr fatal;
        std::mutex fatal_mu;
        auto work = [&] {
            for (size_t i; (i = next.fetch_add(1)) < items.size();) {
                try {
                    parse_one(i);
                } catch (...) {
                    std::lock_guard<std::mutex> g(fatal_mu);
                    if (!fatal) fatal = std::current_exception();
                }
            }
        };
        std::vector<std::thread> ts;
        for (size_t t = 1; t < nthr; t++) ts.emplace_back(work);
        work();
        for (auto& t : ts) t.join();
        if (fatal) std::rethrow_exception(fatal);
    }
    std::vector<Img> imgs;
    for (size_t i = 0; i < items.size(); i++)
        if (ok[i]) imgs.push_back(std::move(parsed[i]));
    if (imgs.empty()) return outs;
    if (split_pp) split_pp->rebind(imgs[0].ci, imgs[0].blocks);

    // one block's Tier-1 job; its plane is the component index until the
    // arena is placed
    auto make_job = [](const Img& im, const DecBlock& b) {
        DecodeJob j{};
        j.plane = (uint64_t)b.comp;
        j.stride = im.pl.stride;
        const bool detail_x = b.band == HL || b.band == HH;
        j.x0 = (uint16_t)(detail_x ? im.pl.hx[b.res] + (uint32_t)b.r.x0 - im.rw[b.res - 1] : (uint32_t)b.r.x0);
        j.y0 = (uint16_t)b.r.y0;
        j.w = (uint8_t)(b.r.x1 - b.r.x0);
        j.h = (uint8_t)(b.r.y1 - b.r.y0);
        j.band = (uint8_t)b.band;
        const int si = b.band_index % (int)im.ci.exps.size();
        j.mb = (uint8_t)((std::max(im.ci.guard - 1, 0) + im.ci.exps[si]) & 255);
        // decode/reconstruct.rs:202-231: step from the component precision
        j.step = im.ci.transform == 0
                     ? quant_step_f32(b.band, im.ci.exps[si], im.ci.mants[si], (uint32_t)im.ci.prec[b.comp])
                     : 0.0f;
        j.data_off = (uint32_t)(im.pay_off + b.off);
        j.data_len = b.len;
        j.zbp = (int16_t)(b.included && b.passes > 0 ? b.zbp : -1);
        j.passes = (uint8_t)std::min(b.passes, 255);
        return j;
    };
    // issue priority by segment length rank: the kernel ends with its
    // heaviest block, so those blocks win their SIMD's issue slots.  Levels
    // 3 / 2 / 1 go to the longest 2% / 10% / 30% of a table (f = 1); the
    // first part of a split decode holds the coarse resolutions, where the
    // longest segments are, and ranks with f = 15 (30% / 100% / all).
    auto rank_jobs = [](DecodeJob* jb, int n, size_t f) {
        if (n <= 0) return;
        std::vector<uint32_t> sorted(n);
        for (int k = 0; k < n; k++) sorted[k] = jb[k].data_len;
        std::sort(sorted.begin(), sorted.end(), std::greater<uint32_t>());
        const size_t m = (size_t)n - 1;
        const uint32_t t3 = sorted[std::min(m, (size_t)n * f / 50)];
        const uint32_t t2 = sorted[std::min(m, (size_t)n * f / 10)];
        const uint32_t t1 = sorted[std::min(m, (size_t)n * f * 3 / 10)];
        for (int k = 0; k < n; k++) {
            const uint32_t l = jb[k].data_len;
            jb[k].pad = (uint8_t)(l >= t3 && l ? 3 : l >= t2 && l ? 2 : l >= t1 && l ? 1 : 0);
        }
    };
    Layout L;
    std::vector<DecodeJob> jobs;
    size_t pay_total = 0;
    for (auto& im : imgs) {
        im.pay_off = pay_total;
        pay_total += im.ci.payload_len;
    }
    // The payload region is rounded up to a granule of an eighth of its
    // size (a power of two, at least 4 KiB), so that pages of one shape whose
    // code-streams differ a little in length share an arena layout, and with
    // it a captured launch sequence.
    size_t pay_gran = 4096;
    while (pay_gran * 8 < pay_total) pay_gran *= 2;
    const size_t o_pay = L.add((pay_total + 64 + pay_gran - 1) / pay_gran * pay_gran);
    int maxL = 0;
    for (auto& im : imgs) {
        const int nc = im.ci.ncomp;
        const size_t llw = (im.ci.width + 1) / 2, llh = (im.ci.height + 1) / 2;
        for (int c = 0; c < nc; c++) {
            im.p_off[c] = L.add((size_t)im.pl.stride * im.ci.height * 4);
            im.ll_off[c][0] = L.add(llw * llh * 4);
            im.ll_off[c][1] = L.add(llw * llh * 4);
            im.o_off[c] = L.add(im.n * 4);
        }
        im.u8_off = items[im.idx].target == DST_DEV_U8 ? 0 : L.add(im.n * nc);
        im.job_first = (uint32_t)jobs.size();
        maxL = std::max(maxL, im.ci.levels);
        for (auto& b : im.blocks)
            if (split_res < 0 || b.res <= split_res) jobs.push_back(make_job(im, b));  // the rest: parsed later
    }
    const int nj = (int)jobs.size();
    // a split decode's table holds every block: its second part follows
    const int nj_all = split_res >= 0 ? (int)imgs[0].blocks.size() : nj;
    // inverse segment tables per step k (coarsest first)
    std::vector<std::vector<DwtSeg>> isegs(2 * std::max(maxL, 1));
    std::vector<PixSeg> fsegs[2];
    for (auto& im : imgs) {
        const int Lv = im.ci.levels;
        const bool irrev = im.ci.transform == 0;
        for (int k = 0; k < Lv; k++) {
            const int r = k + 1;
            for (int c = 0; c < im.ci.ncomp; c++) {
                DwtSeg s{};
                s.in = k == 0 ? im.p_off[c] : im.ll_off[c][(k - 1) & 1];
                s.in_stride = k == 0 ? (int)im.pl.stride : (int)im.rw[r - 1];
                s.plane = im.p_off[c];
                s.pstride = (int)im.pl.stride;
                s.hx = im.pl.hx[r];
                const bool last = k == Lv - 1;
                s.out = last ? im.o_off[c] : im.ll_off[c][k & 1];
                s.out_stride = last ? (int)im.ci.width : (int)im.rw[r];
                s.rw = (int)im.rw[r];
                s.rh = (int)im.rh[r];
                s.tiles_x = (s.rw + kTile - 1) / kTile;
                isegs[2 * k + (irrev ? 1 : 0)].push_back(s);
            }
        }
        PixSeg f{};
        const size_t* src = Lv ? im.o_off : im.p_off;
        f.src = src[0];
        f.src1 = im.ci.ncomp == 3 ? src[1] : 0;
        f.src2 = im.ci.ncomp == 3 ? src[2] : 0;
        f.n = im.n;
        f.channels = im.ci.ncomp;
        f.irrev = irrev;
        f.dst = 0;
        f.dst_i32 = 0;
        fsegs[irrev ? 1 : 0].push_back(f);
    }
    // i32 host outputs are produced on the device too
    std::vector<size_t> o_i32(imgs.size(), 0);
    for (size_t k = 0; k < imgs.size(); k++)
        if (items[imgs[k].idx].target == DST_HOST_I32) o_i32[k] = L.add(imgs[k].n * 4 * imgs[k].ci.ncomp);
    // jobs and segment tables follow each other in their staging order, so
    // their uploads merge into one copy (an empty table takes no space; its
    // launch is skipped)
    const size_t o_jobs = L.add(sizeof(DecodeJob) * nj_all);
    std::vector<size_t> o_iseg(isegs.size());
    for (size_t q = 0; q < isegs.size(); q++) o_iseg[q] = L.add(sizeof(DwtSeg) * isegs[q].size());
    const size_t o_fseg0 = L.add(sizeof(PixSeg) * std::max<size_t>(fsegs[0].size(), 1));
    const size_t o_fseg1 = L.add(sizeof(PixSeg) * std::max<size_t>(fsegs[1].size(), 1));

    dev_.reserve(L.used);
    char* D = dev_.base;
    auto A = [&](size_t off) { return (uint64_t)(uintptr_t)(D + off); };
    for (auto& im : imgs)
        for (uint32_t j = im.job_first; j < std::min<size_t>(im.job_first + im.blocks.size(), jobs.size()); j++)
            jobs[j].plane = A(im.p_off[jobs[j].plane]);
    std::vector<uint32_t> itiles(isegs.size(), 0);
    for (size_t q = 0; q < isegs.size(); q++)
        for (auto& s : isegs[q]) {
            s.in = A(s.in);
            s.plane = A(s.plane);
            s.out = A(s.out);
            s.tile_start = itiles[q];
            itiles[q] += (uint32_t)s.tiles_x * (uint32_t)((s.rh + kTile - 1) / kTile);
        }
    uint32_t fblocks[2] = {0, 0};
    {
        size_t cnt[2] = {0, 0};
        for (size_t k = 0; k < imgs.size(); k++) {
            auto& im = imgs[k];
            int kind = im.ci.transform == 0 ? 1 : 0;
            PixSeg& f = fsegs[kind][cnt[kind]++];
            f.src = A(f.src);
            if (f.src1) f.src1 = A(f.src1);
            if (f.src2) f.src2 = A(f.src2);
            const DecItem& it = items[im.idx];
            if (it.target == DST_DEV_U8) f.dst = (uint64_t)(uintptr_t)it.dev_out;
            else if (it.target == DST_HOST_U8) f.dst = A(im.u8_off);
            if (it.target == DST_HOST_I32) f.dst_i32 = A(o_i32[k]);
            f.blk_start = fblocks[kind];
            fblocks[kind] += pix_blocks(im.n);
        }
    }

    hipStream_t st = stream_;
    // payloads and tables go through the page-locked arena (one asynchronous
    // DMA each; a pageable source makes the runtime stage it synchronously)
    Layout H;
    std::vector<Staged> staged;
    auto stage = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) staged.push_back({off, H.add(bytes), bytes, src});
    };
    // A split decode uploads only the payload bytes its first part reads
    // before that part's launch (the coarse resolutions' packets come first
    // in the codestream); the tail goes up on the side stream while the
    // first part runs.
    size_t pay_head = 0;
    if (split_res >= 0) {
        for (int k = 0; k < nj; k++) pay_head = std::max(pay_head, (size_t)jobs[k].data_off + jobs[k].data_len);
        pay_head = std::min(pay_head, imgs[0].ci.payload_len);
    }
    for (auto& im : imgs) stage(o_pay + im.pay_off, im.ci.payload, split_res >= 0 ? pay_head : im.ci.payload_len);
    if (nj) {
        rank_jobs(jobs.data(), nj, split_res >= 0 ? 15 : 1);
        stage(o_jobs, jobs.data(), sizeof(DecodeJob) * nj);
    }
    for (size_t q = 0; q < isegs.size(); q++) stage(o_iseg[q], isegs[q].data(), sizeof(DwtSeg) * isegs[q].size());
    stage(o_fseg0, fsegs[0].data(), sizeof(PixSeg) * fsegs[0].size());
    stage(o_fseg1, fsegs[1].data(), sizeof(PixSeg) * fsegs[1].size());
    // a split decode stages its second job table after these (the first
    // uploads may still be reading the arena then)
    const size_t h_rest = H.add(sizeof(DecodeJob) * (size_t)(nj_all - nj));
    const size_t pay_tail = split_res >= 0 ? imgs[0].ci.payload_len - pay_head : 0;
    const size_t h_pay_tail = H.add(pay_tail);
    pin_.reserve(H.used);
    copy_staged(staged, D, pin_.base, st);

    Events& ev = ev_;
    // Tier-1 is one launch (a split decode: one per part, the second on the
    // side stream); the inverse DWT levels and the pixel output run as one
    // graph (or their launches), between the stage events.
    if (split_res >= 0) {
        ensure_side_stream();
        J2L_HIP(hipEventRecord(ev_up_, st));  // the payload is on the device
    }
    J2L_HIP(hipEventRecord(ev.e[0], st));
    launch_t1_decode((const DecodeJob*)(D + o_jobs), nj, tables_, (const uint8_t*)(D + o_pay), st);
    if (split_res >= 0) {
        // the finest resolutions' packets, parsed while the first part runs
        Img& im = imgs[0];
        try {
            split_pp->run(im.ci.levels);
            split_pp->finish();
            check_decode_blocks(im.ci, im.blocks);
        } catch (const Error& e) {
            J2L_HIP(hipStreamSynchronize(st));
            outs[im.idx].status = e.status;
            outs[im.idx].message = e.what();
            return outs;
        } catch (...) {
            // the first part still runs on the arena: let it finish first
            (void)hipStreamSynchronize(st);
            throw;
        }
        std::vector<DecodeJob> rest;
        rest.reserve(nj_all - nj);
        for (auto& b : im.blocks)
            if (b.res > split_res) {
                DecodeJob j = make_job(im, b);
                j.plane = A(im.p_off[j.plane]);
                rest.push_back(j);
            }
        const int nr = (int)rest.size();
        if (nr) {
            rank_jobs(rest.data(), nr, 1);
            std::memcpy(pin_.base + h_rest, rest.data(), sizeof(DecodeJob) * nr);
            if (pay_tail) std::memcpy(pin_.base + h_pay_tail, im.ci.payload + pay_head, pay_tail);
            J2L_HIP(hipStreamWaitEvent(side_, ev_up_, 0));
            if (pay_tail)
                J2L_HIP(hipMemcpyAsync(D + o_pay + pay_head, pin_.base + h_pay_tail, pay_tail, hipMemcpyHostToDevice,
                                       side_));
            J2L_HIP(hipMemcpyAsync(D + o_jobs + sizeof(DecodeJob) * nj, pin_.base + h_rest, sizeof(DecodeJob) * nr,
                                   hipMemcpyHostToDevice, side_));
            launch_t1_decode((const DecodeJob*)(D + o_jobs) + nj, nr, tables_, (const uint8_t*)(D + o_pay), side_);
            J2L_HIP(hipEventRecord(ev_side_, side_));
            J2L_HIP(hipStreamWaitEvent(st, ev_side_, 0));
        }
    }
    J2L_HIP(hipEventRecord(ev.e[1], st));
    graphs_.run(st, [&](bool go, std::vector<uint64_t>& sig) {
        sig.insert(sig.end(), {9, (uint64_t)(uintptr_t)D, (uint64_t)L.used, o_fseg0, o_fseg1, fsegs[0].size(),
                               fsegs[1].size(), (uint64_t)fblocks[0], (uint64_t)fblocks[1]});
        for (int k = 0; k < maxL; k++)
            for (int irrev = 0; irrev <= 1; irrev++) {
                auto& v = isegs[2 * k + irrev];
                if (v.empty()) continue;
                sig.insert(sig.end(), {10, (uint64_t)(2 * k + irrev), o_iseg[2 * k + irrev], v.size(),
                                       (uint64_t)itiles[2 * k + irrev]});
                if (go)
                    launch_inv_level(irrev != 0, (const DwtSeg*)(D + o_iseg[2 * k + irrev]), (int)v.size(),
                                     itiles[2 * k + irrev], st);
            }
        if (go) {
            launch_finalize((const PixSeg*)(D + o_fseg0), (int)fsegs[0].size(), fblocks[0], st);
            launch_finalize((const PixSeg*)(D + o_fseg1), (int)fsegs[1].size(), fblocks[1], st);
        }
    });
    J2L_HIP(hipEventRecord(ev.e[2], st));
    for (size_t k = 0; k < imgs.size(); k++) {
        auto& im = imgs[k];
        DecOut& o = outs[im.idx];
        o.width = im.ci.width;
        o.height = im.ci.height;
        o.ncomp = im.ci.ncomp;
        o.colorspace = im.jp.colorspace;
        const DecItem& it = items[im.idx];
        if (it.target == DST_HOST_I32) {
            o.planes.resize(im.n * im.ci.ncomp);
            J2L_HIP(hipMemcpyAsync(o.planes.data(), D + o_i32[k], 4 * o.planes.size(), hipMemcpyDeviceToHost, st));
        } else if (it.target == DST_HOST_U8) {
            o.pixels.resize(im.n * im.ci.ncomp);
            J2L_HIP(hipMemcpyAsync(o.pixels.data(), D + im.u8_off, o.pixels.size(), hipMemcpyDeviceToHost, st));
        }
    }
    J2L_HIP(hipStreamSynchronize(st));
    t_t1_ms = ev.ms(0, 1);
    t_dwt_ms = ev.ms(1, 2);
    return outs;
}

}  // namespace j2l
