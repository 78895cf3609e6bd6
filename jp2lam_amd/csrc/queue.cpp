// Queue execution of long encode/decode lists (jp2lam_encode_many /
// jp2lam_decode_many): the list is cut, in order, into device batches
// ("chunks") of bounded size, and the chunks run on several engines from
// host threads.  Two engines keep one chunk's host Tier-2, packing and
// copies under another chunk's kernels, as the serving slots do.  Every
// stage is per image, so an item's result is the same in any chunk.
#include "queue.hpp"

#include <algorithm>
#include <atomic>
#include <memory>
#include <thread>

namespace j2l {

namespace {
#include "t1_tuning_queue.inc"

std::mutex g_cfg_mu;
QueueConfig g_cfg{kQueueChunkBlocks, kQueueEngines};

// Engines beyond the process-wide one, created on first use and kept.
std::mutex g_pool_mu;
std::vector<std::unique_ptr<Engine>> g_pool;

Engine& queue_engine(int k) {
    if (k == 0) return engine();
    std::lock_guard<std::mutex> g(g_pool_mu);
    while ((int)g_pool.size() < k) g_pool.emplace_back(new Engine());
    return *g_pool[k - 1];
}

// Cut [0, n) into consecutive chunks whose weights stay within `limit`
// (a single heavier item forms a chunk of its own).
std::vector<std::pair<size_t, size_t>> cut(const std::vector<uint64_t>& weight, uint64_t limit) {
    std::vector<std::pair<size_t, size_t>> chunks;
    size_t start = 0;
    uint64_t acc = 0;
    for (size_t i = 0; i < weight.size(); i++) {
        if (i > start && acc + weight[i] > limit) {
            chunks.push_back({start, i});
            start = i;
            acc = 0;
        }
        acc += weight[i];
    }
    if (start < weight.size()) chunks.push_back({start, weight.size()});
    return chunks;
}

template <class Item, class Out, class Run>
std::vector<Out> run_chunks(const std::vector<Item>& items, const std::vector<std::pair<size_t, size_t>>& chunks,
                            int engines, Run run) {
    std::vector<Out> outs(items.size());
    std::atomic<size_t> next{0};
    auto worker = [&](int k) {
        for (;;) {
            const size_t c = next.fetch_add(1);
            if (c >= chunks.size()) return;
            const auto [a, b] = chunks[c];
            std::vector<Item> part(items.begin() + (ptrdiff_t)a, items.begin() + (ptrdiff_t)b);
            std::vector<Out> r;
            try {
                Engine& e = queue_engine(k);
                std::lock_guard<std::mutex> lk(e.mu);
                r = run(e, part);
            } catch (const Error& err) {
                r.assign(part.size(), Out{});
                for (auto& o : r) o.status = err.status, o.message = err.what();
            } catch (const std::exception& err) {
                r.assign(part.size(), Out{});
                for (auto& o : r) o.status = IO, o.message = err.what();
            }
            for (size_t i = a; i < b; i++) outs[i] = std::move(r[i - a]);
        }
    };
    const int nthreads = std::max(1, std::min<int>(engines, (int)chunks.size()));
    if (nthreads == 1) {
        worker(0);
        return outs;
    }
    std::vector<std::thread> th;
    for (int k = 1; k < nthreads; k++) th.emplace_back(worker, k);
    worker(0);
    for (auto& t : th) t.join();
    return outs;
}
}  // namespace

QueueConfig queue_config() {
    std::lock_guard<std::mutex> g(g_cfg_mu);
    return g_cfg;
}

void set_queue_config(uint32_t chunk_blocks, int engines) {
    std::lock_guard<std::mutex> g(g_cfg_mu);
    g_cfg.chunk_blocks = chunk_blocks ? std::min(chunk_blocks, kMaxChunkBlocks) : kQueueChunkBlocks;
    g_cfg.engines = engines > 0 ? std::min(engines, 8) : kQueueEngines;
}

uint64_t estimate_blocks(uint32_t w, uint32_t h, int ncomp, int quality) {
    if (!w || !h) return 1;
    const Plan p = build_plan(w, h, ncomp == 1 ? GRAY : SRGB, quality, 0);
    uint64_t n = 0;
    for (int r = 0; r <= p.levels; r++)
        for (int b = r ? HL : LL; b <= (r ? HH : LL); b++) {
            const Rect rc = band_rect(w, h, p.levels, r, b);
            if (rc.x1 <= rc.x0 || rc.y1 <= rc.y0) continue;
            n += (uint64_t)((rc.x1 - rc.x0 + p.cbw - 1) / p.cbw) * ((rc.y1 - rc.y0 + p.cbh - 1) / p.cbh);
        }
    return std::max<uint64_t>(n * (uint64_t)ncomp, 1);
}

std::vector<EncOut> encode_queue(const std::vector<EncItem>& items) {
    const QueueConfig cfg = queue_config();
    std::vector<uint64_t> weight(items.size());
    for (size_t i = 0; i < items.size(); i++) {
        const auto& it = items[i];
        weight[i] = estimate_blocks(it.width, it.height, it.colorspace == GRAY ? 1 : 3, it.quality);
    }
    const auto chunks = cut(weight, cfg.chunk_blocks);
    return run_chunks<EncItem, EncOut>(items, chunks, cfg.engines,
                                       [](Engine& e, const std::vector<EncItem>& part) { return e.encode(part); });
}

std::vector<DecOut> decode_queue(const std::vector<DecItem>& items) {
    const QueueConfig cfg = queue_config();
    // weight by the decoded size, read from the headers; a stream whose
    // headers do not parse weighs one block (the engine reports its error)
    std::vector<uint64_t> weight(items.size(), 1);
    for (size_t i = 0; i < items.size(); i++) {
        try {
            Jp2Info jp;
            CodestreamInfo ci;
            const uint8_t* cs;
            size_t cslen;
            parse_jp2(items[i].bytes, items[i].len, jp, cs, cslen);
            parse_codestream(cs, cslen, ci);
            weight[i] = estimate_blocks(ci.width, ci.height, ci.ncomp, 85);
        } catch (const Error&) {
        }
    }
    const auto chunks = cut(weight, cfg.chunk_blocks);
    return run_chunks<DecItem, DecOut>(items, chunks, cfg.engines,
                                       [](Engine& e, const std::vector<DecItem>& part) { return e.decode(part); });
}

}  // namespace j2l
