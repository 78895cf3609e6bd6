#include "server.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace j2l {

namespace {
// Device bytes an encode holds per sample: input, level-shifted plane,
// coefficient plane, LL ping-pong, segment slots, Tier-1 symbol streams.
constexpr size_t kEncBytesPerSample = 48;
// Decode: payload, coefficient plane, LL ping-pong, output.
constexpr size_t kDecBytesPerSample = 16;
}  // namespace

namespace {
class EngineRunner : public Runner {
public:
    std::vector<EncOut> encode(const std::vector<EncItem>& items) override { return eng_.encode(items); }
    std::vector<DecOut> decode(const std::vector<DecItem>& items) override { return eng_.decode(items); }

private:
    Engine eng_;
};

class SimulatedRunner : public Runner {
public:
    explicit SimulatedRunner(uint32_t us) : us_(us) {}
    std::vector<EncOut> encode(const std::vector<EncItem>& items) override {
        pause();
        std::vector<EncOut> out(items.size());
        for (size_t i = 0; i < items.size(); i++) {
            const auto& it = items[i];
            const size_t n = (size_t)it.width * it.height * (it.colorspace == GRAY ? 1 : 3);
            uint64_t h = 1469598103934665603ull;  // FNV-1a over the pixels
            const uint8_t* p = (const uint8_t*)it.src;
            if (it.src_kind == SRC_HOST_U8)
                for (size_t k = 0; k < n; k++) h = (h ^ p[k]) * 1099511628211ull;
            auto& b = out[i].bytes;
            const uint32_t v[4] = {it.width, it.height, (uint32_t)it.colorspace, (uint32_t)it.quality};
            b.resize(24);
            std::memcpy(b.data(), v, 16);
            std::memcpy(b.data() + 16, &h, 8);
        }
        return out;
    }
    std::vector<DecOut> decode(const std::vector<DecItem>& items) override {
        pause();
        std::vector<DecOut> out(items.size());
        for (size_t i = 0; i < items.size(); i++) {
            out[i].width = (uint32_t)items[i].len;
            out[i].height = 1;
            out[i].ncomp = 1;
            out[i].pixels.assign(items[i].bytes, items[i].bytes + items[i].len);
        }
        return out;
    }

private:
    void pause() {
        if (us_) std::this_thread::sleep_for(std::chrono::microseconds(us_));
    }
    uint32_t us_;
};
}  // namespace

std::unique_ptr<Runner> make_engine_runner() { return std::make_unique<EngineRunner>(); }
std::unique_ptr<Runner> make_simulated_runner(uint32_t delay_us) { return std::make_unique<SimulatedRunner>(delay_us); }

FaultPlan FaultPlan::from_env() {
    FaultPlan f;
    const char* v = std::getenv("JP2LAM_FAULT_INJECT");
    if (!v) return f;
    const std::string s(v);
    const size_t colon = s.find(':');
    const std::string kind = s.substr(0, colon);
    int at = 1;
    if (colon != std::string::npos) at = std::max(1, std::atoi(s.c_str() + colon + 1));
    if (kind == "encode") f.encode_at = at;
    else if (kind == "decode") f.decode_at = at;
    return f;
}

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

Server::Server(int slots, size_t mem_budget_bytes, int max_batch, bool simulated, uint32_t sim_delay_us) {
    if (slots < 1 || slots > 8) throw Error(INVALID_INPUT, "server slots must be 1..8");
    if (max_batch < 1) throw Error(INVALID_INPUT, "server batch cap must be positive");
    budget_per_slot_ = std::max<size_t>(mem_budget_bytes / (size_t)slots, 1);
    max_batch_ = max_batch;
    faults_ = FaultPlan::from_env();
    simulated_ = simulated;
    for (int i = 0; i < slots; i++)
        runners_.push_back(simulated ? make_simulated_runner(sim_delay_us) : make_engine_runner());
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
    // admission by memory budget: a request that could never fit one
    // slot's share is refused at once instead of running over the budget
    if (r->est_bytes > budget_per_slot_) {
        stats_.rejected++;
        throw Error(INVALID_INPUT, "request needs about " + std::to_string((r->est_bytes + (1u << 20) - 1) >> 20) +
                                       " MiB of device memory, above the per-slot budget of " +
                                       std::to_string(budget_per_slot_ >> 20) + " MiB");
    }
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
        r->pinned = simulated_ ? nullptr : pinned_.get(n);
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

int Server::cancel(uint64_t id) {
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = live_.find(id);
        if (it == live_.end()) throw Error(INVALID_INPUT, "unknown request id");
        r = it->second;
        auto q = std::find(queue_.begin(), queue_.end(), r);
        if (q == queue_.end()) return 0;  // a slot has taken it, or it is done
        queue_.erase(q);
        const auto now = std::chrono::steady_clock::now();
        r->t_start = r->t_done = now;
        if (r->is_encode) r->eout = EncOut{IO, "request cancelled", {}};
        else r->dout.status = IO, r->dout.message = "request cancelled";
        r->done = true;
        stats_.cancelled++;
        stats_.completed++;
    }
    if (r->pinned) {
        pinned_.put(r->pinned, r->pinned_size);
        r->pinned = nullptr;
    }
    cv_done_.notify_all();
    return 1;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

void Server::worker(int slot) {
    Runner& eng = *runners_[slot];
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
        // which calls of this batch carry an injected device fault
        bool fault_enc = false, fault_dec = false;
        {
            std::lock_guard<std::mutex> lk(mu_);
            bool has_enc = false, has_dec = false;
            for (auto& r : batch) (r->is_encode ? has_enc : has_dec) = true;
            if (has_enc) fault_enc = ++enc_calls_ == faults_.encode_at;
            if (has_dec) fault_dec = ++dec_calls_ == faults_.decode_at;
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
        // A failure of the whole encode (or decode) call is every such
        // item's result.  Device errors (HIP failures, reported as IO by the
        // engine) become ENCODE_FAILED / DECODE_FAILED for the requests that
        // met them; the slot goes on serving.
        uint64_t faults = 0;
        try {
            if (!ei.empty()) {
                if (fault_enc) throw Error(IO, "injected device fault (JP2LAM_FAULT_INJECT)");
                auto out = eng.encode(ei);
                for (size_t i = 0; i < er.size(); i++) er[i]->eout = std::move(out[i]);
            }
        } catch (const Error& e) {
            const bool dev = e.status == IO;
            faults += dev;
            for (auto* r : er)
                r->eout = EncOut{dev ? ENCODE_FAILED : e.status, dev ? std::string("device error: ") + e.what() : e.what(), {}};
        } catch (const std::exception& e) {
            for (auto* r : er) r->eout = EncOut{ENCODE_FAILED, e.what(), {}};
        }
        try {
            if (!di.empty()) {
                if (fault_dec) throw Error(IO, "injected device fault (JP2LAM_FAULT_INJECT)");
                auto out = eng.decode(di);
                for (size_t i = 0; i < dr.size(); i++) dr[i]->dout = std::move(out[i]);
            }
        } catch (const Error& e) {
            const bool dev = e.status == IO;
            faults += dev;
            for (auto* r : dr) {
                r->dout.status = dev ? DECODE_FAILED : e.status;
                r->dout.message = dev ? std::string("device error: ") + e.what() : e.what();
            }
        } catch (const std::exception& e) {
            for (auto* r : dr) r->dout.status = DECODE_FAILED, r->dout.message = e.what();
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
            stats_.device_faults += faults;
        }
        cv_done_.notify_all();
    }
}

}  // namespace j2l
