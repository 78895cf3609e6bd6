// Device engine: owns the HIP stream and the device/pinned arenas and runs
// whole encode/decode batches (one image = a batch of one).
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace j2l {

enum SourceKind : int { SRC_HOST_U8 = 0, SRC_HOST_I32 = 1, SRC_DEV_U8 = 2 };

struct EncItem {
    uint32_t width = 0, height = 0;
    int colorspace = GRAY;  // ColorSpace
    int quality = 85;
    int format = 0;         // 0 JP2, 1 J2K
    int src_kind = SRC_HOST_U8;
    const void* src = nullptr;  // u8 interleaved (gray/RGB) or i32 planar (ncomp planes)
    bool metrics = false;       // also compute encode_with_psnr's PSNR / SSIM
};

struct EncOut {
    Status status = OK;
    std::string message;
    std::vector<uint8_t> bytes;
    double psnr_db = 0.0, ssim = 0.0;  // EncItem::metrics only
};

enum DecodeTarget : int { DST_HOST_I32 = 0, DST_HOST_U8 = 1, DST_DEV_U8 = 2 };

struct DecItem {
    const uint8_t* bytes = nullptr;
    size_t len = 0;
    int target = DST_HOST_I32;
    uint8_t* dev_out = nullptr;  // DST_DEV_U8: caller's device buffer (w*h*ncomp bytes)
    size_t dev_cap = 0;
};

struct DecOut {
    Status status = OK;
    std::string message;
    uint32_t width = 0, height = 0;
    int ncomp = 0, colorspace = GRAY;
    std::vector<int32_t> planes;  // DST_HOST_I32: ncomp planes of width*height
    std::vector<uint8_t> pixels;  // DST_HOST_U8: interleaved
};

// ContentStats and ContentClass (content.rs:80-105).

// Tier-1 block statistics of the last encode (tests and counters).
struct BlockStat {
    uint8_t comp, res, band, mbp, zbp, cls, npasses, keep;
    uint16_t x0, y0, x1, y1;
    uint32_t nbytes, body_len, max_mag, nnz;
    uint32_t nsym, nsym_kept;  // MQ decisions, all passes / kept passes (counting on)
};

struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    bool pinned = false;
    void reserve(size_t bytes);
    ~Arena();
};

struct Events {
    hipEvent_t e[8];
    Events() {
        for (auto& x : e) J2L_HIP(hipEventCreate(&x));
    }
    ~Events() {
        for (auto& x : e) (void)hipEventDestroy(x);
    }
    float ms(int a, int b) {
        float t = 0;
        J2L_HIP(hipEventElapsedTime(&t, e[a], e[b]));
        return t;
    }
};

// Launch sequences (kernels and device-to-device copies of one stage) run
// as HIP graphs; the stage events are recorded on the stream between them,
// since event records captured into a graph did not give usable times.  A sequence is named by its exact signature: every step's
// kind, arena offsets, counts and grid sizes, and the arena base.  It is
// captured on its second occurrence and replayed from then on; a batch of a
// new shape runs eagerly.  Inputs and descriptor tables are copied into the
// arena before the launch, so a replay codes the new data.
class GraphCache {
public:
    ~GraphCache();
    void run(hipStream_t st, const std::function<void(bool go, std::vector<uint64_t>& sig)>& seq);
    void clear();

private:
    static constexpr size_t kMaxGraphs = 32;
    std::map<std::vector<uint64_t>, hipGraphExec_t> exec_;
    std::map<std::vector<uint64_t>, int> seen_;
};

// Process-wide switch for the graph caches of every engine (default off) and
// their totals: sequences captured, graph launches.
void set_graphs(bool on);
void graph_stats(uint64_t* captures, uint64_t* replays);

class Engine {
public:
    Engine();
    ~Engine();
    std::vector<EncOut> encode(const std::vector<EncItem>& items);
    std::vector<DecOut> decode(const std::vector<DecItem>& items);
    // analyze (content.rs:120-166) of n samples per component: gray reads one
    // plane / channel, otherwise R, G, B.  src: host u8 interleaved, host i32
    // planar (ncomp planes of n) or device u8 interleaved (SourceKind).
    ContentOut analyze_content(const void* src, int src_kind, size_t n, bool gray);
    // Debug/test access to the last encode: per-block statistics, pass
    // snapshots and the full Tier-1 segments (before truncation).
    const std::vector<BlockStat>& last_blocks() const { return last_blocks_; }
    const std::vector<uint32_t>& last_snaps() const { return last_snaps_; }
    const std::vector<uint16_t>& last_newsigs() const { return last_newsigs_; }
    const std::vector<uint8_t>& last_segments() const { return last_segments_; }
    const std::vector<uint32_t>& last_segment_offsets() const { return last_seg_offs_; }
    const std::vector<double>& last_lambdas() const { return last_lambdas_; }
    bool keep_debug = false;
    // Tier-1 encoder choice: 0 = the offline-tuned table by batch-size class
    // (tuned_t1_choice), 1 = wave per block, 2 = lane per block, 3 = lane per block from lane_min_blocks up,
    // 4 = hybrid (streams longer than hybrid_frac x the longest on the
    // scalar coder, the rest on lanes).
    int t1_mode = 0;
    uint32_t lane_min_blocks = 8192;
    float hybrid_frac = 0.55f;
    bool last_t1_lanes = false;
    int last_t1_coder = 1;
    // Coder (1 wave, 2 lane, 4 hybrid) and hybrid split point for a batch of
    // `nblocks` code-blocks in mode 0: the entry of the offline-tuned table
    // (jp2lam_amd/autotune.py) for the batch's size class.
    struct T1Choice {
        int coder;
        float frac;
    };
    static T1Choice tuned_t1_choice(int nblocks);
    // Stage timings of the last call (ms, from HIP events): prepare+DWT,
    // Tier-1, PCRD+pack, host Tier-2.
    double t_dwt_ms = 0, t_t1_ms = 0, t_pcrd_ms = 0, t_host_ms = 0;
    double t_metrics_ms = 0;  // encode_with_psnr reconstruction + reductions
    hipStream_t stream() const { return stream_; }
    GraphCache& graphs() { return graphs_; }
    std::mutex mu;

private:
    // A second stream and two events (no timing) for a single-image decode:
    // the packets of its finest resolutions are parsed while Tier-1 decodes
    // the coarser ones, whose blocks are then coded on this stream beside
    // them (created on first use).
    void ensure_side_stream();
    hipStream_t stream_ = nullptr, side_ = nullptr;
    hipEvent_t ev_up_ = nullptr, ev_side_ = nullptr;
    Arena dev_, pin_;
    Events ev_;
    GraphCache graphs_;
    uint32_t* tables_ = nullptr;
    std::vector<BlockStat> last_blocks_;
    std::vector<uint32_t> last_snaps_;
    std::vector<uint16_t> last_newsigs_;
    std::vector<uint8_t> last_segments_;
    std::vector<uint32_t> last_seg_offs_;
    std::vector<double> last_lambdas_;
};

Engine& engine();
// Tier-1 counters (encode/counters.rs:3-9), summed over every encode of
// every engine while counting is on: {total blocks, empty blocks, MQ
// symbols, cleanup passes, SP passes, MR passes, pass bytes}.
void set_t1_counting(bool on);
bool t1_counting();
void t1_counters(uint64_t* out7);
void reset_t1_counters();
bool device_available(std::string* why = nullptr);

}  // namespace j2l
