#include "server.hpp"

#include <algorithm>
#include <cstring>

namespace j2l {

namespace {
// Device bytes an encode holds per sample: input, level-shifted plane,
// coefficient plane, LL ping-pong, segment slots, Tier-1 symbol streams.
constexpr size_t kEncBytesPerSample = 48;
// Decode: payload, coefficient plane, LL ping-pong, output.
constexpr size_t kDecBytesPerSample = 16;
}  // namespace

PinnedPool::~PinnedPool() {
    for (auto& kv : free_) (void)hipHostFree(kv.second);
}

uint8_t* PinnedPool::get(size_t n) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.lower_bound(n);
    if (it != free_.end() && it->first <= 2 * n) {  // reuse a buffer of a similar size
        uint8_t* p = it->second;
        free_.erase(it);
        return p;
    }
    if (held_ + n > cap_) return nullptr;
    uint8_t* p = nullptr;
    if (hipHostMalloc((void**)&p, n, hipHostMallocDefault) != hipSuccess) return nullptr;
    held_ += n;
    return p;
}

void PinnedPool::put(uint8_t* p, size_t n) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_.emplace(n, p);
}

Server::Server(int slots, size_t mem_budget_bytes, int max_batch) {
    if (slots < 1 || slots > 8) throw Error(INVALID_INPUT, "server slots must be 1..8");
    if (max_batch < 1) throw Error(INVALID_INPUT, "server batch cap must be positive");
    budget_per_slot_ = std::max<size_t>(mem_budget_bytes / (size_t)slots, 1);
    max_batch_ = max_batch;
    for (int i = 0; i < slots; i++) engines_.emplace_back(new Engine());
    for (int i = 0; i < slots; i++) threads_.emplace_back(&Server::worker, this, i);
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
}

uint64_t Server::enqueue(std::shared_ptr<Request> r) {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error(IO, "server is shutting down");
    r->id = next_id_++;
    r->t_submit = std::chrono::steady_clock::now();
    live_[r->id] = r;
    queue_.push_back(r);
    stats_.submitted++;
    cv_work_.notify_one();
    return r->id;
}

uint64_t Server::submit_encode(const uint8_t* pixels, uint32_t w, uint32_t h, uint32_t channels, int quality,
                               int format, bool on_device) {
    if (!pixels) throw Error(INVALID_INPUT, "null argument");
    if (channels != 1 && channels != 3) throw Error(INVALID_INPUT, "channels must be 1 (gray) or 3 (sRGB)");
    auto r = std::make_shared<Request>();
    r->is_encode = true;
    r->enc.width = w;
    r->enc.height = h;
    r->enc.colorspace = channels == 1 ? GRAY : SRGB;
    r->enc.quality = quality;
    r->enc.format = format;
    const size_t n = (size_t)w * h * channels;
    if (on_device) {
        r->enc.src_kind = SRC_DEV_U8;
        r->enc.src = pixels;
    } else {
        r->enc.src_kind = SRC_HOST_U8;
        r->pinned = pinned_.get(n);
        if (r->pinned) {
            // a pooled buffer can be up to twice n; the request records n
            r->pinned_size = n;
            std::memcpy(r->pinned, pixels, n);
            r->enc.src = r->pinned;
        } else {
            r->pixels.assign(pixels, pixels + n);
            r->enc.src = r->pixels.data();
        }
    }
    r->est_bytes = n * kEncBytesPerSample;
    return enqueue(r);
}

uint64_t Server::submit_decode(const uint8_t* bytes, size_t len) {
    if (!bytes) throw Error(INVALID_INPUT, "null argument");
    auto r = std::make_shared<Request>();
    r->is_encode = false;
    r->stream.assign(bytes, bytes + len);
    // size from the header when it parses; a bad stream fails in its batch
    size_t samples = len * 8;
    try {
        Jp2Info jp;
        const uint8_t* cs;
        size_t cslen;
        parse_jp2(r->stream.data(), r->stream.size(), jp, cs, cslen);
        samples = (size_t)jp.width * jp.height * std::max(jp.ncomp, 1);
    } catch (const Error&) {
    }
    r->est_bytes = samples * kDecBytesPerSample;
    return enqueue(r);
}

std::shared_ptr<Request> Server::wait(uint64_t id, int timeout_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = live_.find(id);
    if (it == live_.end()) throw Error(INVALID_INPUT, "unknown request id");
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) cv_done_.wait(lk, ready);
    else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return nullptr;
    live_.erase(id);
    return r;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

void Server::worker(int slot) {
    Engine& eng = *engines_[slot];
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;  // stop requested and nothing left
            // admission: the head always, then in order while the budget holds
            size_t mem = 0;
            while (!queue_.empty() && (int)batch.size() < max_batch_) {
                const auto& r = queue_.front();
                if (!batch.empty() && mem + r->est_bytes > budget_per_slot_) break;
                mem += r->est_bytes;
                batch.push_back(r);
                queue_.pop_front();
            }
            stats_.batches++;
            stats_.max_batch = std::max<uint64_t>(stats_.max_batch, batch.size());
        }
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<EncItem> ei;
        std::vector<DecItem> di;
        std::vector<Request*> er, dr;
        for (auto& r : batch) {
            r->t_start = t0;
            r->batch_size = (int)batch.size();
            if (r->is_encode) {
                ei.push_back(r->enc);
                er.push_back(r.get());
            } else {
                DecItem d;
                d.bytes = r->stream.data();
                d.len = r->stream.size();
                d.target = DST_HOST_U8;
                di.push_back(d);
                dr.push_back(r.get());
            }
        }
        // a failure of the whole encode (or decode) call is every such item's result
        try {
            if (!ei.empty()) {
                auto out = eng.encode(ei);
                for (size_t i = 0; i < er.size(); i++) er[i]->eout = std::move(out[i]);
            }
        } catch (const Error& e) {
            for (auto* r : er) r->eout = EncOut{e.status, e.what(), {}};
        } catch (const std::exception& e) {
            for (auto* r : er) r->eout = EncOut{IO, e.what(), {}};
        }
        try {
            if (!di.empty()) {
                auto out = eng.decode(di);
                for (size_t i = 0; i < dr.size(); i++) dr[i]->dout = std::move(out[i]);
            }
        } catch (const Error& e) {
            for (auto* r : dr) r->dout.status = e.status, r->dout.message = e.what();
        } catch (const std::exception& e) {
            for (auto* r : dr) r->dout.status = IO, r->dout.message = e.what();
        }
        const auto t1 = std::chrono::steady_clock::now();
        for (auto& r : batch) {  // the batch has finished with its inputs
            if (r->pinned) {
                pinned_.put(r->pinned, r->pinned_size);
                r->pinned = nullptr;
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto& r : batch) {
                r->t_done = t1;
                r->done = true;
                r->pixels.clear();
                r->pixels.shrink_to_fit();
            }
            stats_.completed += batch.size();
        }
        cv_done_.notify_all();
    }
}

}  // namespace j2l
