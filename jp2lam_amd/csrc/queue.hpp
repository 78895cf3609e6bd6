// Queue execution of long encode/decode lists: bounded device batches run
// on several engines at once (queue.cpp).
#pragma once
#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace j2l {

// Largest chunk, in code-blocks: keeps a chunk's Tier-1 output slots
// (at most 3 bytes per sample plus 256 per block) within 32-bit offsets.
constexpr uint32_t kMaxChunkBlocks = 1u << 18;

struct QueueConfig {
    uint32_t chunk_blocks;  // code-blocks per device batch
    int engines;            // engines (host threads) running chunks at once
};
QueueConfig queue_config();
// 0 restores the tuned default for either value.
void set_queue_config(uint32_t chunk_blocks, int engines);
// Code-blocks of one image under the plan's geometry (layout.rs:37-173).
uint64_t estimate_blocks(uint32_t w, uint32_t h, int ncomp, int quality);
std::vector<EncOut> encode_queue(const std::vector<EncItem>& items);
std::vector<DecOut> decode_queue(const std::vector<DecItem>& items);

}  // namespace j2l
