// Mutation fuzz of the host-side decode parsers (JP2 boxes, codestream
// markers, packet headers, per-block checks) for sanitizer builds on the CPU.
//
// Usage: parse_fuzz <iterations> <seed> <stream.jp2>...
// Every seed stream must parse; each mutant must either parse or throw
// j2l::Error.  Any other outcome (a sanitizer report, another exception)
// ends the process with a non-zero status.  Prints one summary line.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <memory>
#include <vector>

#include "../../jp2lam_amd/csrc/host.hpp"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    size_t below(size_t n) { return n ? (size_t)(next() % n) : 0; }
};

// Offsets of the 0xFF-prefixed marker codes and of the JP2 box headers:
// mutations aimed at length fields find the interesting bytes there.
std::vector<size_t> marker_offsets(const std::vector<uint8_t>& b) {
    std::vector<size_t> m;
    for (size_t i = 0; i + 1 < b.size(); i++)
        if (b[i] == 0xff && b[i + 1] >= 0x4f && b[i + 1] <= 0x93) m.push_back(i);
    return m;
}

void mutate(std::vector<uint8_t>& b, Rng& r, const std::vector<size_t>& marks) {
    const int kinds = 10;
    int rounds = 1 + (int)r.below(4);
    for (int k = 0; k < rounds && !b.empty(); k++) {
        switch (r.below(kinds)) {
            case 0:  // flip one bit
                b[r.below(b.size())] ^= (uint8_t)(1u << r.below(8));
                break;
            case 1:  // random byte
                b[r.below(b.size())] = (uint8_t)r.next();
                break;
            case 2:  // truncate
                b.resize(r.below(b.size() + 1));
                break;
            case 3: {  // insert random bytes
                size_t at = r.below(b.size() + 1), n = 1 + r.below(16);
                std::vector<uint8_t> ins(n);
                for (auto& x : ins) x = (uint8_t)r.next();
                b.insert(b.begin() + (ptrdiff_t)at, ins.begin(), ins.end());
                break;
            }
            case 4: {  // delete a range
                size_t at = r.below(b.size()), n = 1 + r.below(32);
                b.erase(b.begin() + (ptrdiff_t)at, b.begin() + (ptrdiff_t)std::min(b.size(), at + n));
                break;
            }
            case 5:  // extreme marker-segment length after a marker code
                if (!marks.empty()) {
                    size_t at = marks[r.below(marks.size())] + 2;
                    static const uint16_t lens[] = {0, 1, 2, 3, 0xfff0, 0xffff, 0x8000, 10, 38};
                    uint16_t L = lens[r.below(sizeof lens / sizeof lens[0])];
                    if (at + 1 < b.size()) b[at] = (uint8_t)(L >> 8), b[at + 1] = (uint8_t)L;
                }
                break;
            case 6: {  // extreme 32-bit field (box lengths, Psot, SIZ sizes)
                size_t at = r.below(b.size());
                static const uint32_t vals[] = {0, 1, 7, 8, 16, 0xffffffffu, 0x7fffffffu, 0x10000u, 65536u * 2};
                uint32_t v = vals[r.below(sizeof vals / sizeof vals[0])];
                for (int i = 0; i < 4 && at + i < b.size(); i++) b[at + i] = (uint8_t)(v >> (24 - 8 * i));
                break;
            }
            case 7:  // 0xFF into the packet data (stuffing rules)
                b[r.below(b.size())] = 0xff;
                break;
            case 8: {  // duplicate a range
                size_t at = r.below(b.size()), n = 1 + r.below(64);
                n = std::min(n, b.size() - at);
                std::vector<uint8_t> dup(b.begin() + (ptrdiff_t)at, b.begin() + (ptrdiff_t)(at + n));
                b.insert(b.begin() + (ptrdiff_t)r.below(b.size() + 1), dup.begin(), dup.end());
                break;
            }
            case 9:  // a marker segment inserted at a marker boundary
                if (!marks.empty()) {
                    static const uint8_t codes[] = {0x64, 0x64, 0x90, 0x93, 0x52, 0x5c, 0x51, 0xd9};
                    static const uint16_t lens[] = {0, 1, 2, 4, 0xfff0, 0xffff, 10};
                    size_t at = std::min(marks[r.below(marks.size())], b.size());
                    uint16_t L = lens[r.below(sizeof lens / sizeof lens[0])];
                    std::vector<uint8_t> seg = {0xff, codes[r.below(sizeof codes)], (uint8_t)(L >> 8), (uint8_t)L};
                    for (size_t i = 0, n = r.below(8); i < n; i++) seg.push_back((uint8_t)r.next());
                    b.insert(b.begin() + (ptrdiff_t)at, seg.begin(), seg.end());
                }
                break;
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s iterations seed stream...\n", argv[0]);
        return 2;
    }
    const long iters = std::atol(argv[1]);
    Rng rng{std::strtoull(argv[2], nullptr, 10)};
    std::vector<std::vector<uint8_t>> seeds;
    for (int i = 3; i < argc; i++) {
        std::ifstream f(argv[i], std::ios::binary);
        seeds.emplace_back(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        j2l::Jp2Info jp;
        j2l::CodestreamInfo ci;
        std::vector<j2l::DecBlock> blocks;
        try {
            j2l::parse_for_decode(seeds.back().data(), seeds.back().size(), jp, ci, blocks);
        } catch (const j2l::Error& e) {
            std::fprintf(stderr, "seed %s does not parse: %s\n", argv[i], e.what());
            return 3;
        }
    }
    long parsed = 0, rejected = 0;
    for (long it = 0; it < iters; it++) {
        const auto& s = seeds[(size_t)it % seeds.size()];
        std::vector<uint8_t> b = s;
        mutate(b, rng, marker_offsets(s));
        // An exact-size heap copy, so that any read past the end is caught.
        std::unique_ptr<uint8_t[]> buf(new uint8_t[b.size() ? b.size() : 1]);
        if (!b.empty()) std::memcpy(buf.get(), b.data(), b.size());
        j2l::Jp2Info jp;
        j2l::CodestreamInfo ci;
        std::vector<j2l::DecBlock> blocks;
        try {
            j2l::parse_for_decode(buf.get(), b.size(), jp, ci, blocks);
            parsed++;
        } catch (const j2l::Error& e) {
            if (e.status != j2l::DECODE_FAILED && e.status != j2l::UNSUPPORTED_FEATURE) {
                std::fprintf(stderr, "iteration %ld: unexpected status %d: %s\n", it, (int)e.status, e.what());
                return 4;
            }
            rejected++;
        }
    }
    std::printf("fuzz ok: %ld mutants, %ld parsed, %ld rejected\n", iters, parsed, rejected);
    return 0;
}
