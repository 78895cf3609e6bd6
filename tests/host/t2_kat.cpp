// Known-answer harness for the host Tier-2 code (packet headers, packet
// parsing, tile-part framing), built with any C++17 compiler from
// jp2lam_amd/csrc/host.cpp.  Driven by tests/test_host_t2.py.
//
//   t2_kat hdr <numpasses> <zbp> <seg_len>   one band, one block, first
//                                            inclusion: header bytes (hex)
//   t2_kat parse <hex payload>               1x1 gray image, 0 levels, one
//                                            64x64 block: parse result
//   t2_kat frame <payload_len>               SOT..EOC framing of a 1x1
//                                            q85 J2K around a payload
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../jp2lam_amd/csrc/host.hpp"

static void hex(const std::vector<uint8_t>& v) {
    for (uint8_t b : v) std::printf("%02x", b);
    std::printf("\n");
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "hdr" && argc == 5) {
            j2l::CodedBlock cb{(uint8_t)std::atoi(argv[2]), (uint8_t)std::atoi(argv[3]), (uint32_t)std::atol(argv[4])};
            j2l::BandBlocks band{1, 1, &cb};
            std::vector<uint8_t> out;
            j2l::packet_header(&band, 1, out);
            hex(out);
            return 0;
        }
        if (mode == "parse" && argc == 3) {
            const std::string h = argv[2];
            std::vector<uint8_t> pay;
            for (size_t i = 0; i + 1 < h.size(); i += 2) pay.push_back((uint8_t)std::strtoul(h.substr(i, 2).c_str(), nullptr, 16));
            j2l::CodestreamInfo ci;
            ci.width = ci.height = 1;
            ci.ncomp = 1;
            ci.levels = 0;
            ci.cbw = ci.cbh = 64;
            ci.layers = 1;
            ci.guard = 1;
            ci.qstyle = 2;
            ci.exps = {8};
            ci.mants = {0};
            ci.prec = {8};
            ci.payload = pay.data();
            ci.payload_len = pay.size();
            std::vector<j2l::DecBlock> blocks;
            j2l::parse_packets(ci, blocks);
            const auto& b = blocks.at(0);
            std::printf("ok included=%d passes=%d zbp=%d len=%u off=%u\n", (int)b.included, b.passes, b.zbp, b.len, b.off);
            return 0;
        }
        if (mode == "frame" && argc == 3) {
            const j2l::Plan p = j2l::build_plan(1, 1, j2l::GRAY, 85, 1);
            std::vector<uint8_t> head, tail;
            j2l::write_codestream_head(p, (size_t)std::atol(argv[2]), head);
            j2l::write_codestream_tail(tail);
            hex(head);
            hex(tail);
            return 0;
        }
    } catch (const j2l::Error& e) {
        std::printf("error %d %s\n", (int)e.status, e.what());
        return 0;
    }
    return 2;
}
