// Serving loop: continuous batching of encode and decode requests.
//
// `slots` workers each own an Engine (HIP stream, device and pinned
// arenas).  A free worker takes the queue's head and then, in arrival
// order, every further request that fits its memory budget (and the batch
// cap), runs them as one batch, and hands results back.  Requests that
// arrive while batches are on the GPU wait only until a slot frees; batches
// on different slots overlap on the device, and one slot's host Tier-2
// overlaps another's kernels.  A result is the same bytes or pixels as the
// request alone (every stage is per image; batching only shares launches).
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace j2l {

// Pinned (page-locked) host buffers for submitted pixels, recycled across
// requests, so an encode's upload is one DMA instead of a staged copy.
// Beyond `cap` bytes in use, requests fall back to pageable memory.
class PinnedPool {
public:
    explicit PinnedPool(size_t cap) : cap_(cap) {}
    ~PinnedPool();
    uint8_t* get(size_t n);  // nullptr when over the cap or out of memory
    void put(uint8_t* p, size_t n);

private:
    std::mutex mu_;
    std::multimap<size_t, uint8_t*> free_;
    size_t held_ = 0, cap_;
};

struct Request {
    uint64_t id = 0;
    bool is_encode = true;
    EncItem enc;
    uint8_t* pinned = nullptr;    // host pixels copied at submission (pinned pool) ...
    size_t pinned_size = 0;
    std::vector<uint8_t> pixels;  // ... or pageable when the pool is full
    std::vector<uint8_t> stream;  // decode input copied at submission
    size_t est_bytes = 0;         // device bytes the request is expected to hold
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    int batch_size = 0;
    bool done = false;
    EncOut eout;
    DecOut dout;
};

// What a serving slot runs its batches on: a device engine, or, for tests
// of the request loop on a machine without a GPU, a simulated executor
// (no device work: an encode's result is a 24-byte digest of the request,
// a decode's the stream's bytes as a 1-row gray image, after a fixed delay).
class Runner {
public:
    virtual ~Runner() = default;
    virtual std::vector<EncOut> encode(const std::vector<EncItem>& items) = 0;
    virtual std::vector<DecOut> decode(const std::vector<DecItem>& items) = 0;
};
std::unique_ptr<Runner> make_engine_runner();
std::unique_ptr<Runner> make_simulated_runner(uint32_t delay_us);

// Injected device faults (JP2LAM_FAULT_INJECT="encode" | "decode" |
// "encode:N" | "decode:N", read when the server starts): the N-th (default
// first) batch call of that kind fails as a device error would.
struct FaultPlan {
    int encode_at = 0, decode_at = 0;  // 1-based call number, 0 = never
    static FaultPlan from_env();
};

struct ServerStats {
    uint64_t submitted = 0, completed = 0, batches = 0;
    uint64_t max_batch = 0;
    uint64_t rejected = 0, cancelled = 0, device_faults = 0;
};

class Server {
public:
    // simulated: slots run the simulated executor with `sim_delay_us` per
    // batch instead of device engines (tests of the request loop only).
    Server(int slots, size_t mem_budget_bytes, int max_batch, bool simulated = false, uint32_t sim_delay_us = 0);
    ~Server();
    // Encode request: host pixels (u8, interleaved, channels 1 or 3) are
    // copied; device pixels (on_device) must stay valid until completion.
    uint64_t submit_encode(const uint8_t* pixels, uint32_t w, uint32_t h, uint32_t channels, int quality, int format,
                           bool on_device);
    uint64_t submit_decode(const uint8_t* bytes, size_t len);
    // Waits up to timeout_ms (< 0: no limit) and moves the result out.
    // Returns nullptr on timeout; throws Error for an unknown id.
    std::shared_ptr<Request> wait(uint64_t id, int timeout_ms);
    // 1: removed before it started (it completes with status IO, "request
    // cancelled"); 0: already running or done; throws for an unknown id.
    int cancel(uint64_t id);
    ServerStats stats();

private:
    void worker(int slot);
    uint64_t enqueue(std::shared_ptr<Request> r);

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Request>> queue_;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live_;
    std::vector<std::unique_ptr<Runner>> runners_;
    FaultPlan faults_;
    bool simulated_ = false;
    int enc_calls_ = 0, dec_calls_ = 0;
    std::vector<std::thread> threads_;
    bool stop_ = false;
    uint64_t next_id_ = 1;
    PinnedPool pinned_{size_t(4) << 30};
    size_t budget_per_slot_;
    int max_batch_;
    ServerStats stats_;
};

}  // namespace j2l
