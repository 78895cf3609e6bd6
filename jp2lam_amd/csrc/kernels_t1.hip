// EBCOT Tier-1 encoder and decoder for CDNA4 (gfx950).
//
// One 64-lane wavefront codes one code-block.  Lane x owns column x of the
// block; its significance, sign, visited and refined state are 64-bit
// row masks held in VGPRs.  For each 4-row stripe the wave packs, in
// parallel, a 32-bit neighbourhood word per lane (rows 4s-1..4s+4 of its
// column), then a wave-uniform loop walks the columns in the reference's
// scan order (tier1 passes: t1.rs:805-1197, decode/t1.rs:211-466).  The loop
// carries the west column's new significance from this pass, reads the
// other columns' words with v_readlane and keeps all coder state in SGPRs,
// so it runs on the scalar unit while other waves on the SIMD use the VALU.
// Context and MQ probability tables live one entry per lane in VGPRs and are
// read with readlane, not memory loads.  Columns with nothing to code are
// skipped with a ballot mask.
//
// The MQ coder/decoder follows mq/mod.rs:30-432 bit for bit, including the
// byte-position snapshots used as pass lengths (t1.rs:583-600).
#include "common.hpp"

#include <algorithm>
#include <mutex>

namespace j2l {

namespace {

constexpr int CTX_ZC = 0, CTX_MAG = 14, CTX_AGG = 17, CTX_UNI = 18;

// Lane index within the wavefront (the encoder runs two waves per block).
__device__ __forceinline__ uint32_t lid() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
// Write a wave-uniform value into one lane of a VGPR (v_cmp + v_cndmask).
__device__ __forceinline__ uint32_t wl(uint32_t v, uint32_t val, uint32_t lane) {
    return lid() == lane ? val : v;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// Wave issue priority within its SIMD (s_setprio takes an immediate).
__device__ __forceinline__ void set_priority(int p) {
    if (p >= 3) __builtin_amdgcn_s_setprio(3);
    else if (p == 2) __builtin_amdgcn_s_setprio(2);
    else if (p == 1) __builtin_amdgcn_s_setprio(1);
}

// The decoder's context states live in s[80:98] (context k in s[80 + k])
// for the whole kernel.  The kernel caps the SGPRs the compiler allocates
// below 80 (kStateSgprCap), so no compiler code touches the block, and a
// decision reaches state M0 with s_movrels / s_movreld: no VALU-to-SALU hop
// and no writelane-to-readlane wait on its critical path.  Listing the
// block as clobbered here makes the kernel descriptor reserve it.  (An SALU
// write of M0 needs one wait state before s_movrels.)
#define kStateSgprCap 80
__device__ __forceinline__ void states_init(uint32_t cx) {
    asm volatile(
        "v_readlane_b32 s80, %0, 0\n\tv_readlane_b32 s81, %0, 1\n\tv_readlane_b32 s82, %0, 2\n\t"
        "v_readlane_b32 s83, %0, 3\n\tv_readlane_b32 s84, %0, 4\n\tv_readlane_b32 s85, %0, 5\n\t"
        "v_readlane_b32 s86, %0, 6\n\tv_readlane_b32 s87, %0, 7\n\tv_readlane_b32 s88, %0, 8\n\t"
        "v_readlane_b32 s89, %0, 9\n\tv_readlane_b32 s90, %0, 10\n\tv_readlane_b32 s91, %0, 11\n\t"
        "v_readlane_b32 s92, %0, 12\n\tv_readlane_b32 s93, %0, 13\n\tv_readlane_b32 s94, %0, 14\n\t"
        "v_readlane_b32 s95, %0, 15\n\tv_readlane_b32 s96, %0, 16\n\tv_readlane_b32 s97, %0, 17\n\t"
        "v_readlane_b32 s98, %0, 18"
        :
        : "v"(cx)
        : "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", "s91", "s92", "s93", "s94",
          "s95", "s96", "s97", "s98");
}

// Rows 4s-1..4s+4 of a column mask as bits 0..5 (missing rows read 0).
__device__ __forceinline__ uint32_t win6(uint64_t m, int s) {
    return s ? (uint32_t)(m >> (4 * s - 1)) & 63u : (uint32_t)(m << 1) & 63u;
}

struct Tables {
    uint32_t mq;  // lane i < 50: qe | nmps << 16 | nlps << 22 | flip-MPS << 31
    uint32_t zc;  // lane i: ZC contexts of neighbour masks 8i..8i+7 (nibbles)
    uint32_t sc;  // lane i: sign entries 4i..4i+3 (ctx | spb << 4, bytes)
};

__device__ __forceinline__ uint32_t zc_ctx(const Tables& t, uint32_t nb) { return (rl(t.zc, nb >> 3) >> ((nb & 7) * 4)) & 15u; }
__device__ __forceinline__ uint32_t sc_entry(const Tables& t, uint32_t si) { return (rl(t.sc, si >> 2) >> ((si & 3) * 8)) & 255u; }


// ---- MQ decoder (mq/mod.rs:30-155) -----------------------------------------
// Coder registers in SGPRs; context k's state is its table entry with the
// MPS in bit 31 (as in the encoder), kept in s[80 + k].  The segment is read through a
// 128-byte window of two VGPRs (lane i of w0/w1 = byte dbase + i / + 64 + i,
// 0xFF past the end, which the decoder reads as the reference does,
// mq/mod.rs:152-154); the scan slides it before each column or group of at
// most 16 decisions, which read at most 32 bytes (two byte-ins per decision),
// and a slide leaves at least 40 bytes ahead, so a decision never leaves it.  The next 64 bytes
// (w2) are loaded one slide ahead, so a slide never waits on memory.
struct MqDec {
    uint32_t a, c, ct, bp, cur;  // cur = byte at bp
    uint32_t tab;                // VGPR: transition table
    uint32_t w0, w1, w2, dbase, lim;  // lim = dbase + 88: slide when bp reaches it
    const uint8_t* data;
    uint32_t len;
};
__device__ __forceinline__ uint32_t dec_load64(const MqDec& d, uint32_t base) {
    const uint32_t i = base + lid();
    return i < d.len ? (uint32_t)d.data[i] : 0xffu;
}
__device__ __forceinline__ void dec_window_init(MqDec& d) {
    d.dbase = 0;
    d.lim = 88;
    d.w0 = dec_load64(d, 0);
    d.w1 = dec_load64(d, 64);
    d.w2 = dec_load64(d, 128);
}
__device__ __forceinline__ void dec_slide(MqDec& d) {
    if (__builtin_expect(d.bp >= d.lim, 0)) {
        d.lim += 64u;
        // w0 = w1, w1 = w2 in place (written as plain copies, the rotation
        // made the compiler copy the window registers on every column)
        asm volatile("v_mov_b32 %0, %1\n\tv_mov_b32 %1, %2" : "+v"(d.w0), "+v"(d.w1) : "v"(d.w2));
        d.dbase += 64u;
        d.w2 = dec_load64(d, d.dbase + 128u);
    }
}
// Start of decoding (mq/mod.rs:46-60): c from the first byte, one byte-in,
// then c <<= 7.
__device__ __forceinline__ void dec_start(MqDec& d) {
    dec_window_init(d);
    d.bp = 0;
    d.a = 0x80000000u;  // 0x8000 << 16
    d.cur = rl(d.w0, 0);
    d.c = d.cur << 16;
    const uint32_t next = rl(d.w0, 1);
    if (d.cur == 0xffu) {
        if (next > 0x8fu) {
            d.c += 0xff00u;
            d.ct = 8;
        } else {
            d.bp = 1;
            d.cur = next;
            d.c += next << 9;
            d.ct = 7;
        }
    } else {
        d.bp = 1;
        d.cur = next;
        d.c += next << 8;
        d.ct = 8;
    }
    d.c <<= 7;
    d.ct = d.ct >= 7 ? d.ct - 7 : 0;
}
// One decision as one scalar-unit sequence (mq/mod.rs:62-155): the MPS
// case that keeps a >= 0x8000 exits early; otherwise the LPS flag becomes a
// mask and one common tail updates the state, forms the decision and shifts
// (byte-in only when ct runs out).  M0 holds the context index on entry.
// The interval a is held shifted up by 16 bits, and the state entries of
// the decoder carry qe << 16 (dec_entry), so the comparisons with c's high
// half and the subtraction from it need no shifts, the a >= 0x8000 test is
// a sign test, and the renormalising shift is the leading-zero count of a.
__device__ __forceinline__ uint32_t mq_decode(MqDec& d, const Tables&, uint32_t ctx) {
    uint32_t bit, e, q, t, k, n0, n1;
    asm volatile(
        "s_nop 0\n\t"                 // the M0 wait state
        "s_movrels_b32 %[e], s80\n\t"
        "s_and_b32 %[q], %[e], 0x7fff0000\n\t"
        "s_sub_u32 %[a], %[a], %[q]\n\t"
        "s_cmp_lt_u32 %[c], %[q]\n\t"   // c >> 16 < qe
        "s_cbranch_scc1 1f\n\t"
        "s_sub_u32 %[c], %[c], %[q]\n\t"
        "s_cmp_lt_i32 %[a], 0\n\t"
        "s_cbranch_scc0 7f\n\t"         // renormalising upper path: the one taken branch
        "s_lshr_b32 %[bit], %[e], 31\n\t"   // MPS, a stays >= 0x8000
        "s_branch 9f\n"
        "1:\n\t"                     // lower sub-interval: a = qe; LPS iff a >= qe
        "s_cmp_ge_u32 %[a], %[q]\n\t"
        "s_cselect_b32 %[k], -1, 0\n\t"
        "s_mov_b32 %[a], %[q]\n\t"
        "s_branch 8f\n"
        "2:\n\t"                    // t > ct: shift by ct, then a byte-in
        "s_lshl_b32 %[a], %[a], %[ct]\n\t"
        "s_lshl_b32 %[c], %[c], %[ct]\n\t"
        "s_sub_u32 %[t], %[t], %[ct]\n\t"
        // byte-in, ct is 0 (mq/mod.rs:133-155): the usual byte (the held
        // one not 0xFF) runs straight through
        "s_sub_u32 %[k], %[bp], %[db]\n\t"
        "s_add_u32 %[k], %[k], 1\n\t"
        "v_readlane_b32 %[n0], %[w0], %[k]\n\t"
        "v_readlane_b32 %[n1], %[w1], %[k]\n\t"
        "s_cmp_lt_u32 %[k], 64\n\t"
        "s_cselect_b32 %[n0], %[n0], %[n1]\n\t"
        "s_cmp_eq_u32 %[cur], 0xff\n\t"
        "s_cbranch_scc1 5f\n\t"
        "s_add_u32 %[bp], %[bp], 1\n\t"
        "s_mov_b32 %[cur], %[n0]\n\t"
        "s_lshl_b32 %[n0], %[n0], 8\n\t"
        "s_add_u32 %[c], %[c], %[n0]\n\t"
        "s_mov_b32 %[ct], 8\n"
        "3:\n\t"                    // the rest of the shift
        "s_cmp_gt_u32 %[t], %[ct]\n\t"
        "s_cbranch_scc1 2b\n\t"
        "s_lshl_b32 %[a], %[a], %[t]\n\t"
        "s_lshl_b32 %[c], %[c], %[t]\n\t"
        "s_sub_u32 %[ct], %[ct], %[t]\n\t"
        "s_branch 9f\n"
        "5:\n\t"                    // after 0xFF: a 7-bit byte, or a marker
        "s_cmp_gt_u32 %[n0], 0x8f\n\t"
        "s_cbranch_scc1 6f\n\t"
        "s_add_u32 %[bp], %[bp], 1\n\t"
        "s_mov_b32 %[cur], %[n0]\n\t"
        "s_lshl_b32 %[n0], %[n0], 9\n\t"
        "s_add_u32 %[c], %[c], %[n0]\n\t"
        "s_mov_b32 %[ct], 7\n\t"
        "s_branch 3b\n"
        "6:\n\t"
        "s_add_u32 %[c], %[c], 0xff00\n\t"
        "s_mov_b32 %[ct], 8\n\t"
        "s_branch 3b\n"
        "7:\n\t"                        // upper sub-interval, a < 0x8000: LPS iff a < qe
        "s_cmp_lt_u32 %[a], %[q]\n\t"
        "s_cselect_b32 %[k], -1, 0\n"
        "8:\n\t"                    // k = all ones on LPS, and so is SCC
        "s_cselect_b32 %[t], 0x60006, %[f0]\n\t"  // NLPS / NMPS field: width 6 at 6 / 0
        "s_bfe_u32 %[t], %[e], %[t]\n\t"
        "v_readlane_b32 %[n1], %[tab], %[t]\n\t"  // used four instructions later
        "s_xor_b32 %[bit], %[e], %[k]\n\t"   // bit 31: MPS ^ LPS = the decision
        "s_lshr_b32 %[bit], %[bit], 31\n\t"
        "s_and_b32 %[n0], %[e], 0x80000000\n\t"
        "s_flbit_i32_b32 %[t], %[a]\n\t"
        "s_xor_b32 %[n1], %[n1], %[n0]\n\t"   // bit 31: MPS, flipped by a switching LPS
        "s_movreld_b32 s80, %[n1]\n\t"
        "s_cmp_gt_u32 %[t], %[ct]\n\t"
        "s_cbranch_scc1 2b\n\t"
        "s_lshl_b32 %[a], %[a], %[t]\n\t"
        "s_lshl_b32 %[c], %[c], %[t]\n\t"
        "s_sub_u32 %[ct], %[ct], %[t]\n"
        "9:"
        : [a] "+s"(d.a), [c] "+s"(d.c), [ct] "+s"(d.ct), [bp] "+s"(d.bp), [cur] "+s"(d.cur), [m0] "+{m0}"(ctx), [bit] "=&s"(bit), [e] "=&s"(e), [q] "=&s"(q), [t] "=&s"(t), [k] "=&s"(k),
          [n0] "=&s"(n0), [n1] "=&s"(n1)
        : [tab] "v"(d.tab), [w0] "v"(d.w0), [w1] "v"(d.w1), [db] "s"(d.dbase), [f0] "s"(0x60000u)
        : "scc");
    return bit;
}
// A state-table lane (qe | nmps << 16 | nlps << 22 | flip-MPS << 31) in the
// decoder's layout: qe << 16 | nmps | nlps << 6 | flip-MPS << 31.
__device__ __forceinline__ uint32_t dec_entry(uint32_t v) {
    return ((v & 0xffffu) << 16) | ((v >> 16) & 63u) | (((v >> 22) & 63u) << 6) | (v & 0x80000000u);
}

// ---- shared scan machinery ------------------------------------------------------
// Per-lane column state.
struct Col {
    uint64_t sig, sgn, vis, ref;
};

// Neighbour mask in the T1_SIGMA layout (tier1/consts.rs:1-27) of row ci of
// the current column, given effective 6-row windows of W, C (own) and E.
__device__ __forceinline__ uint32_t nbmask(uint32_t w6, uint32_t c6, uint32_t e6, uint32_t r) {
    return ((w6 >> (r - 1)) & 1u) | (((c6 >> (r - 1)) & 1u) << 1) | (((e6 >> (r - 1)) & 1u) << 2) |
           (((w6 >> r) & 1u) << 3) | (((e6 >> r) & 1u) << 5) | (((w6 >> (r + 1)) & 1u) << 6) |
           (((c6 >> (r + 1)) & 1u) << 7) | (((e6 >> (r + 1)) & 1u) << 8);
}
// Sign index (tier1/helpers.rs:53-89 bit layout).
__device__ __forceinline__ uint32_t signidx(uint32_t w6, uint32_t ws6, uint32_t c6, uint32_t cs6, uint32_t e6,
                                            uint32_t es6, uint32_t r) {
    uint32_t lu = 0;
    if ((w6 >> r) & 1u) lu |= 8u | ((ws6 >> r) & 1u);
    if ((c6 >> (r - 1)) & 1u) lu |= 2u | (((cs6 >> (r - 1)) & 1u) << 4);
    if ((e6 >> r) & 1u) lu |= 32u | (((es6 >> r) & 1u) << 2);
    if ((c6 >> (r + 1)) & 1u) lu |= 128u | (((cs6 >> (r + 1)) & 1u) << 6);
    return lu;
}

template <bool DEC>
struct Coder;

template <>
struct Coder<true> {
    MqDec m;
    __device__ __forceinline__ uint32_t code(const Tables& t, uint32_t ctx, uint32_t) { return mq_decode(m, t, ctx); }
    __device__ __forceinline__ void column() { dec_slide(m); }
};

__device__ __forceinline__ void load_tables(Tables& t, const uint32_t* __restrict__ tabs, int band) {
    const int lane = lid();
    t.mq = tabs[lane];
    t.zc = tabs[64 + band * 64 + lane];
    t.sc = tabs[320 + lane];
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

// ---- encoder: lane-parallel context formation + scalar MQ coder ---------------
//
// Each coding pass is formed for the whole block first: for every 4-row
// stripe, lane x decides which of its four samples are coded, with which
// context and decision, writes those (context, decision) bytes into LDS at
// an offset given by a wave prefix sum, and then the scalar MQ loop codes the
// bytes in order.  The scan order of the reference (stripe, then column,
// then row) is the byte order in LDS.
//
// Symbol byte: context in bits 0..4, decision in bit 7.

constexpr uint32_t kSymCap = 10240;  // 1024 stripe-columns x at most 10 symbols

// v_writelane: write a wave-uniform value into one lane of a VGPR.  The
// lane select goes through M0 (one SGPR operand per VALU instruction), which
// is always written by the scalar unit here.
__device__ __forceinline__ uint32_t wlane(uint32_t v, uint32_t val, uint32_t lane) {
    // val and lane are wave-uniform by contract; readfirstlane states it
    // for the compiler (free when they already sit in SGPRs).
    val = (uint32_t)__builtin_amdgcn_readfirstlane((int)val);
    lane = (uint32_t)__builtin_amdgcn_readfirstlane((int)lane);
    asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(val), "{m0}"(lane));
    return v;
}
// Lane i receives lane i-1's value (west neighbour); lane 0 receives 0.
__device__ __forceinline__ uint32_t from_west(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}
// Lane i receives lane i+1's value (east neighbour); lane 63 receives 0.
__device__ __forceinline__ uint32_t from_east(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}
// Number of set bits of m in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Exclusive prefix sum over lanes of a count < 16, and the wave total.
__device__ __forceinline__ uint32_t lane_offsets(uint32_t cnt, uint32_t& total) {
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t m = ballot((cnt >> k) & 1u);
        off += lanes_below(m) << k;
        tot += (uint32_t)__builtin_popcountll(m) << k;
    }
    total = tot;
    return off;
}
__device__ __forceinline__ uint32_t wave_popc4(uint32_t v4) {
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) n += (uint32_t)__builtin_popcountll(ballot((v4 >> k) & 1u));
    return n;
}
// Per-lane table lookups (each lane its own index) through ds_bpermute.
__device__ __forceinline__ uint32_t zc_lane(const Tables& t, uint32_t nb) {
    return ((uint32_t)__shfl((int)t.zc, (int)(nb >> 3), 64) >> ((nb & 7u) * 4u)) & 15u;
}
__device__ __forceinline__ uint32_t sc_lane(const Tables& t, uint32_t si) {
    return ((uint32_t)__shfl((int)t.sc, (int)(si >> 2), 64) >> ((si & 3u) * 8u)) & 255u;
}
__device__ __forceinline__ uint32_t zc_sym(const Tables& t, uint32_t nb, uint32_t bit) { return zc_lane(t, nb) | (bit << 7); }
// Sign symbol: context and (sign XOR predicted sign) (t1.rs:884-893).
__device__ __forceinline__ uint32_t sc_sym(const Tables& t, uint32_t si, uint32_t neg) {
    const uint32_t e = sc_lane(t, si);
    return (e & 15u) | ((neg ^ (e >> 4)) << 7);
}

// Significance propagation pass, context formation (t1.rs:805-905).  A
// sample is coded when it is insignificant and some neighbour is
// significant at the time it is reached: neighbours earlier in scan order
// (the whole west column of the stripe, rows above in its own column) count
// with this pass's new significance, later ones without.  The new
// significance of a stripe is the least fixpoint of that rule, found by
// iterating lane-parallel; each round moves information one column east.
__device__ __forceinline__ uint32_t form_spp(const Tables& t, Col& st, uint64_t bits, uint64_t negs, int w, int h, uint8_t* symb,
                             uint32_t& newsig) {
    const bool valid = (int)lid() < w;
    const int nstripes = (h + 3) >> 2;
    uint32_t total = 0, ns = 0;
    for (int s = 0; s < nstripes; s++) {
        const int nrows = __builtin_amdgcn_readfirstlane(min(4, h - 4 * s));
        const uint32_t rmask = valid ? (1u << nrows) - 1u : 0u;
        const uint32_t sig6 = win6(st.sig, s), sgn6 = win6(st.sgn, s);
        const uint32_t b4 = (uint32_t)(bits >> (4 * s)) & 15u, n4 = (uint32_t)(negs >> (4 * s)) & 15u;
        const uint32_t w6 = from_west(sig6), e6 = from_east(sig6);
        const uint32_t ws6 = from_west(sgn6), es6 = from_east(sgn6);
        const uint32_t x6 = w6 | e6;
        const uint32_t nbs = (x6 | (x6 >> 1) | (x6 >> 2) | sig6 | (sig6 >> 2)) & 15u;
        const uint32_t uns = ~(sig6 >> 1) & rmask;
        uint32_t nsig = 0, coded;
        for (;;) {
            const uint32_t wn = from_west(nsig);
            uint32_t c = uns & (nbs | ((wn | (wn << 1) | (wn >> 1)) & 15u));
            uint32_t n = c & b4;
            c |= (n << 1) & uns;
            n = c & b4;
            c |= (n << 1) & uns;
            n = c & b4;
            c |= (n << 1) & uns;
            n = c & b4;
            coded = c;
            const bool grew = n != nsig;
            nsig = n;
            if (!ballot(grew)) break;
        }
        const uint32_t nsgn = nsig & n4;
        const uint32_t wE = w6 | (from_west(nsig) << 1), wsE = ws6 | (from_west(nsgn) << 1);
        const uint32_t cnt = (uint32_t)__builtin_popcount(coded) + (uint32_t)__builtin_popcount(nsig);
        uint32_t tot;
        uint32_t pos = total + lane_offsets(cnt, tot);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t above = (1u << r) - 1u;
            const uint32_t cE = sig6 | ((nsig & above) << 1), csE = sgn6 | ((nsgn & above) << 1);
            const uint32_t zs = zc_sym(t, nbmask(wE, cE, e6, r + 1), (b4 >> r) & 1u);
            const uint32_t ss = sc_sym(t, signidx(wE, wsE, cE, csE, e6, es6, r + 1), (n4 >> r) & 1u);
            if ((coded >> r) & 1u) {
                symb[pos++] = (uint8_t)zs;
                if ((nsig >> r) & 1u) symb[pos++] = (uint8_t)ss;
            }
        }
        total += tot;
        ns += wave_popc4(nsig);
        st.sig |= (uint64_t)nsig << (4 * s);
        st.sgn |= (uint64_t)nsgn << (4 * s);
        st.vis |= (uint64_t)coded << (4 * s);
    }
    newsig = ns;
    return total;
}

// Magnitude refinement pass, context formation (t1.rs:914-965): members are
// significant samples not coded in this plane's SP pass; the context is 16
// once refined, else 15/14 by any significant 8-neighbour.
__device__ __forceinline__ uint32_t form_mrp(Col& st, uint64_t bits, int w, int h, uint8_t* symb) {
    const bool valid = (int)lid() < w;
    const int nstripes = (h + 3) >> 2;
    uint32_t total = 0;
    for (int s = 0; s < nstripes; s++) {
        const int nrows = __builtin_amdgcn_readfirstlane(min(4, h - 4 * s));
        const uint32_t rmask = valid ? (1u << nrows) - 1u : 0u;
        const uint32_t sig6 = win6(st.sig, s);
        const uint32_t vis4 = (uint32_t)(st.vis >> (4 * s)) & 15u, ref4 = (uint32_t)(st.ref >> (4 * s)) & 15u;
        const uint32_t b4 = (uint32_t)(bits >> (4 * s)) & 15u;
        const uint32_t mem = (sig6 >> 1) & ~vis4 & rmask;
        const uint32_t w6 = from_west(sig6), e6 = from_east(sig6);
        uint32_t tot;
        uint32_t pos = total + lane_offsets((uint32_t)__builtin_popcount(mem), tot);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if ((mem >> r) & 1u) {
                const uint32_t ctx = ((ref4 >> r) & 1u) ? CTX_MAG + 2u : (nbmask(w6, sig6, e6, r + 1) ? CTX_MAG + 1u : CTX_MAG);
                symb[pos++] = (uint8_t)(ctx | (((b4 >> r) & 1u) << 7));
            }
        }
        total += tot;
    }
    st.ref |= st.sig & ~st.vis;
    return total;
}

// Cleanup pass, context formation (t1.rs:977-1197): every insignificant,
// unvisited sample is coded, so the pass's new significance is simply
// bit & todo.  Full stripes that are clean (flags.rs:202-215) and have no
// new significance in the west column use run mode: AGG, then if any bit is
// set the run length as two UNI decisions and the first sample's sign.
__device__ __forceinline__ uint32_t form_cup(const Tables& t, Col& st, uint64_t bits, uint64_t negs, int w, int h, uint8_t* symb,
                             uint32_t& newsig) {
    const bool valid = (int)lid() < w;
    const int nstripes = (h + 3) >> 2;
    uint32_t total = 0, ns = 0;
    for (int s = 0; s < nstripes; s++) {
        const int nrows = __builtin_amdgcn_readfirstlane(min(4, h - 4 * s));
        const uint32_t rmask = valid ? (1u << nrows) - 1u : 0u;
        const uint32_t sig6 = win6(st.sig, s), sgn6 = win6(st.sgn, s);
        const uint32_t vis4 = (uint32_t)(st.vis >> (4 * s)) & 15u;
        const uint32_t b4 = (uint32_t)(bits >> (4 * s)) & 15u, n4 = (uint32_t)(negs >> (4 * s)) & 15u;
        const uint32_t w6 = from_west(sig6), e6 = from_east(sig6);
        const uint32_t ws6 = from_west(sgn6), es6 = from_east(sgn6);
        const uint32_t todo = ~((sig6 >> 1) | vis4) & rmask;
        const uint32_t nsig = todo & b4, nsgn = nsig & n4;
        const uint32_t wn = from_west(nsig);
        const uint32_t wE = w6 | (wn << 1), wsE = ws6 | (from_west(nsgn) << 1);
        const bool run = valid && nrows == 4 && sig6 == 0 && vis4 == 0 && w6 == 0 && e6 == 0 && wn == 0;
        const uint32_t rl0 = b4 ? (uint32_t)__builtin_ctz(b4) : 4u;
        const uint32_t reg = run ? (b4 ? todo & ~((2u << rl0) - 1u) : 0u) : todo;
        const uint32_t hdr = run ? (b4 ? 4u : 1u) : 0u;
        const uint32_t cnt = hdr + (uint32_t)__builtin_popcount(reg) + (uint32_t)__builtin_popcount(nsig & reg);
        uint32_t tot;
        uint32_t pos = total + lane_offsets(cnt, tot);
        // run-mode sign of the first set row: nothing above it is significant
        const uint32_t rs = sc_sym(t, signidx(wE, wsE, sig6, sgn6, e6, es6, rl0 + 1), (n4 >> (rl0 & 3u)) & 1u);
        if (run) {
            symb[pos] = (uint8_t)(CTX_AGG | (b4 ? 128u : 0u));
            if (b4) {
                symb[pos + 1] = (uint8_t)(CTX_UNI | (((rl0 >> 1) & 1u) << 7));
                symb[pos + 2] = (uint8_t)(CTX_UNI | ((rl0 & 1u) << 7));
                symb[pos + 3] = (uint8_t)rs;
            }
            pos += hdr;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t above = (1u << r) - 1u;
            const uint32_t cE = sig6 | ((nsig & above) << 1), csE = sgn6 | ((nsgn & above) << 1);
            const uint32_t zs = zc_sym(t, nbmask(wE, cE, e6, r + 1), (b4 >> r) & 1u);
            const uint32_t ss = sc_sym(t, signidx(wE, wsE, cE, csE, e6, es6, r + 1), (n4 >> r) & 1u);
            if ((reg >> r) & 1u) {
                symb[pos++] = (uint8_t)zs;
                if ((nsig >> r) & 1u) symb[pos++] = (uint8_t)ss;
            }
        }
        total += tot;
        ns += wave_popc4(nsig);
        st.sig |= (uint64_t)nsig << (4 * s);
        st.sgn |= (uint64_t)nsgn << (4 * s);
    }
    newsig = ns;
    return total;
}

// ---- MQ encoder (mq/mod.rs:163-432) --------------------------------------------
// Coder registers in SGPRs.  cx lane k holds context k's current state as
// its table entry (qe | nmps << 16 | nlps << 22) with the MPS in bit 31, so
// a decision needs one readlane; the next state is the table entry at the
// NMPS or NLPS field XOR the MPS bit (entries 47..49, the LPS targets of the
// switching states, carry bit 31 to flip it).  The pre-formed heaviest
// blocks keep the states in the decoder's SGPR block instead (mqe_group16s,
// mqe_symbol_s): faster for a lone heavy wave, but for every block it
// measured slower, as the two extra scalar instructions per decision then
// compete with the waves sharing the SIMD.  Output bytes collect in
// a 64-byte ring in one VGPR: byte b sits in lane (b + 1) & 63, so the byte
// committed when pos moves on goes to lane pos & 63 with no range test (the
// first commit, of no byte, lands in lane 0 as "byte -1", never stored).
// Bytes [wbase, wbase + 32) are stored with one write once final; between
// spill checks at most 32 bytes are added, so no unstored byte is reused.
struct MqE {
    uint32_t a, c, ct, pos, last;
    uint32_t cx, tab;       // VGPRs
    uint32_t win, wbase;    // byte ring (VGPR) and the first unstored byte
    uint8_t* out;
    uint32_t cap, overflow;
};

// Byte-out (mq/mod.rs:373-414) after the shift by ct: the usual case (held
// byte not 0xFF, no carry) runs straight to the commit; then back to the
// renormalisation test (BACK: label reference with its direction).  The
// commit writes the held byte to ring lane pos & 63 (M0 is free here: the
// state write is done).  L8, L7, LCOM, LCARRY are label numbers local to the
// caller: 8-bit byte, 7-bit byte, commit, carry into the held byte.
#define MQE_BYTEOUT(L8, L7, LCOM, LCARRY, BACK)                                    \
    "s_lshl_b32 %[a], %[a], %[ct]\n\t"                                          \
    "s_lshl_b32 %[c], %[c], %[ct]\n\t"                                          \
    "s_sub_u32 %[t], %[t], %[ct]\n\t"                                           \
    "s_cmp_eq_u32 %[last], 0xff\n\t"                                            \
    "s_cbranch_scc1 " L7 "f\n\t"                                                \
    "s_bitcmp1_b32 %[c], 27\n\t"                                                \
    "s_cbranch_scc1 " LCARRY "f\n" L8 ":\n\t"                                  \
    "s_lshr_b32 %[q], %[c], 19\n\t"                                             \
    "s_and_b32 %[c], %[c], 0x7ffff\n\t"                                         \
    "s_mov_b32 %[ct], 8\n" LCOM ":\n\t"                                        \
    "s_mov_b32 m0, %[pos]\n\t"                                                  \
    "s_add_u32 %[pos], %[pos], 1\n\t"                                           \
    "v_writelane_b32 %[win], %[last], m0\n\t"                                   \
    "s_and_b32 %[last], %[q], 0xff\n\t"                                         \
    "s_branch " BACK "\n" LCARRY ":\n\t"                                       \
    "s_add_u32 %[last], %[last], 1\n\t"                                         \
    "s_cmp_eq_u32 %[last], 0xff\n\t"                                            \
    "s_cbranch_scc0 " L8 "b\n\t"                                                \
    "s_bitset0_b32 %[c], 27\n" L7 ":\n\t"                                      \
    "s_lshr_b32 %[q], %[c], 20\n\t"                                             \
    "s_and_b32 %[c], %[c], 0xfffff\n\t"                                         \
    "s_mov_b32 %[ct], 7\n\t"                                                    \
    "s_branch " LCOM "b\n"

// One decision, written as one scalar-unit sequence (mq/mod.rs:213-242 with
// renormalisation :362-371 and byte-out :373-414): the MPS case that keeps
// a >= 0x8000 takes 9 instructions; the MPS case that renormalises shifts by
// exactly one bit; the LPS case shifts by the leading zeros of a.  M0 holds
// the context index on entry (the window write reuses it).
__device__ __forceinline__ void mqe_symbol(MqE& m, uint32_t ctx, uint32_t d31) {
    uint32_t e, q, u, t;
    asm volatile(
        "v_readlane_b32 %[e], %[cx], m0\n\t"
        "s_and_b32 %[q], %[e], 0xffff\n\t"
        "s_sub_u32 %[a], %[a], %[q]\n\t"
        "s_xor_b32 %[u], %[e], %[d]\n\t"
        "s_lshl_b32 %[t], %[a], 16\n\t"
        "s_andn2_b32 %[t], %[t], %[u]\n\t"
        "s_cmp_lt_i32 %[t], 0\n\t"
        "s_cbranch_scc0 7f\n\t"       // renormalising path: the one taken branch
        "s_add_u32 %[c], %[c], %[q]\n\t"
        "s_branch 9f\n"
        "1:\n\t"                      // t >= ct: shift by ct and move a byte out
        MQE_BYTEOUT("2", "3", "5", "6", "8f")
        "7:\n\t"
        // renormalising decision, MPS or LPS, branch-free: conditional
        // exchange to_qe = (a < qe) XOR lps; MPS: a = qe if a < qe else
        // c += qe; LPS: c += qe if a < qe else a = qe; state NMPS / NLPS,
        // MPS ^= SWITCH on LPS
        // (a, qe < 2^16, so bit 31 of a - qe is a < qe)
        "s_sub_u32 %[t], %[a], %[q]\n\t"
        "s_xor_b32 %[t], %[t], %[u]\n\t"
        "s_cmp_lt_i32 %[t], 0\n\t"
        "s_cselect_b32 %[a], %[q], %[a]\n\t"
        "s_cselect_b32 %[t], 0, %[q]\n\t"
        "s_add_u32 %[c], %[c], %[t]\n\t"
        "s_cmp_lt_i32 %[u], 0\n\t"
        "s_cselect_b32 %[t], 0x60016, %[f16]\n\t"  // NLPS / NMPS field: width 6 at 22 / 16
        "s_bfe_u32 %[t], %[e], %[t]\n\t"
        "v_readlane_b32 %[t], %[tab], %[t]\n\t"
        "s_and_b32 %[q], %[e], 0x80000000\n\t"
        "s_xor_b32 %[t], %[t], %[q]\n\t"             // bit 31: MPS, flipped by a switching LPS
        "v_writelane_b32 %[cx], %[t], m0\n\t"
        "s_flbit_i32_b32 %[t], %[a]\n\t"
        "s_sub_u32 %[t], %[t], 16\n"
        // renormalise by t bits, a byte out whenever ct reaches zero
        "8:\n\t"
        "s_cmp_lt_u32 %[t], %[ct]\n\t"
        "s_cbranch_scc0 1b\n\t"
        "s_lshl_b32 %[a], %[a], %[t]\n\t"
        "s_lshl_b32 %[c], %[c], %[t]\n\t"
        "s_sub_u32 %[ct], %[ct], %[t]\n"
        "9:"
        : [a] "+s"(m.a), [c] "+s"(m.c), [ct] "+s"(m.ct), [pos] "+s"(m.pos), [last] "+s"(m.last),
          [cx] "+v"(m.cx), [win] "+v"(m.win), [m0] "+{m0}"(ctx), [e] "=&s"(e), [q] "=&s"(q),
          [u] "=&s"(u), [t] "=&s"(t)
        : [d] "s"(d31), [tab] "v"(m.tab), [f16] "s"(0x60010u)
        : "scc");
}
// mqe_symbol with the context states in s[80:98].
__device__ __forceinline__ void mqe_symbol_s(MqE& m, uint32_t ctx, uint32_t d31) {
    uint32_t e, q, u, t;
    asm volatile(
        "s_nop 0\n\t"                 // the M0 wait state
        "s_movrels_b32 %[e], s80\n\t"
        "s_and_b32 %[q], %[e], 0xffff\n\t"
        "s_sub_u32 %[a], %[a], %[q]\n\t"
        "s_xor_b32 %[u], %[e], %[d]\n\t"
        "s_lshl_b32 %[t], %[a], 16\n\t"
        "s_andn2_b32 %[t], %[t], %[u]\n\t"
        "s_cmp_lt_i32 %[t], 0\n\t"
        "s_cbranch_scc0 7f\n\t"       // renormalising path: the one taken branch
        "s_add_u32 %[c], %[c], %[q]\n\t"
        "s_branch 9f\n"
        "1:\n\t"                      // t >= ct: shift by ct and move a byte out
        MQE_BYTEOUT("2", "3", "5", "6", "8f")
        "7:\n\t"
        // renormalising decision, MPS or LPS, branch-free: conditional
        // exchange to_qe = (a < qe) XOR lps; MPS: a = qe if a < qe else
        // c += qe; LPS: c += qe if a < qe else a = qe; state NMPS / NLPS,
        // MPS ^= SWITCH on LPS
        // (a, qe < 2^16, so bit 31 of a - qe is a < qe)
        "s_sub_u32 %[t], %[a], %[q]\n\t"
        "s_xor_b32 %[t], %[t], %[u]\n\t"
        "s_cmp_lt_i32 %[t], 0\n\t"
        "s_cselect_b32 %[a], %[q], %[a]\n\t"
        "s_cselect_b32 %[t], 0, %[q]\n\t"
        "s_add_u32 %[c], %[c], %[t]\n\t"
        "s_cmp_lt_i32 %[u], 0\n\t"
        "s_cselect_b32 %[t], 0x60016, %[f16]\n\t"  // NLPS / NMPS field: width 6 at 22 / 16
        "s_bfe_u32 %[t], %[e], %[t]\n\t"
        "v_readlane_b32 %[t], %[tab], %[t]\n\t"
        "s_and_b32 %[q], %[e], 0x80000000\n\t"
        "s_xor_b32 %[t], %[t], %[q]\n\t"             // bit 31: MPS, flipped by a switching LPS
        "s_movreld_b32 s80, %[t]\n\t"
        "s_flbit_i32_b32 %[t], %[a]\n\t"
        "s_sub_u32 %[t], %[t], 16\n"
        // renormalise by t bits, a byte out whenever ct reaches zero
        "8:\n\t"
        "s_cmp_lt_u32 %[t], %[ct]\n\t"
        "s_cbranch_scc0 1b\n\t"
        "s_lshl_b32 %[a], %[a], %[t]\n\t"
        "s_lshl_b32 %[c], %[c], %[t]\n\t"
        "s_sub_u32 %[ct], %[ct], %[t]\n"
        "9:"
        : [a] "+s"(m.a), [c] "+s"(m.c), [ct] "+s"(m.ct), [pos] "+s"(m.pos), [last] "+s"(m.last),
          [win] "+v"(m.win), [m0] "+{m0}"(ctx), [e] "=&s"(e), [q] "=&s"(q),
          [u] "=&s"(u), [t] "=&s"(t)
        : [d] "s"(d31), [tab] "v"(m.tab), [f16] "s"(0x60010u)
        : "scc");
}
// Coder registers at the start of a block (mq/mod.rs:183-200).  Their
// initial values come from one asm that writes SGPRs: a constant the
// compiler materialises with v_mov (0 is shared with the ring, a VGPR) would
// reach the asm operands through the pass loop's phis as a VGPR, which the
// backend cannot turn back into an SGPR.
__device__ __forceinline__ void mqe_init(MqE& m, uint32_t mq, uint8_t* out, uint32_t cap) {
    uint32_t z, a0, ct0;
    asm volatile("s_mov_b32 %0, 0\n\ts_mov_b32 %1, 0x8000\n\ts_mov_b32 %2, 12" : "=s"(z), "=s"(a0), "=s"(ct0));
    m.a = a0;
    m.c = z;
    m.ct = ct0;
    m.pos = z;
    m.last = z;
    m.wbase = z;
    m.tab = mq;
    m.win = 0;
    m.out = out;
    m.cap = cap;
    m.overflow = 0;
}
// One decision of a 16-decision group (mqe_group16), labels 1<ID><n>.  X is
// the symbol word, OFF the s_bfe field of its context, DSH the shift putting
// the decision into bit 31 of DREG.  c += qe is done before the test, so the
// MPS case that keeps a >= 0x8000 is one taken branch to the next decision;
// the renormalising path falls through (and takes qe back from c when it
// exchanges), with the byte-out out of line after the group.
#define MQE_G_SYM(ID, X, OFF, DSH, DREG) MQE_G_SYM_AT(ID, X, OFF, DSH, DREG, MQE_V_READ, MQE_V_WRITE)
#define MQE_V_READ "v_readlane_b32 %[e], %[cx], m0\n\t"
#define MQE_V_WRITE "v_writelane_b32 %[cx], %[t], m0\n\t"
#define MQE_G_SYM_AT(ID, X, OFF, DSH, DREG, SREAD, SWRITE)                         \
    "s_bfe_u32 m0, %[" X "], " OFF "\n\t" DSH                                     \
    SREAD                                                                        \
    "s_and_b32 %[q], %[e], 0xffff\n\t"                                           \
    "s_sub_u32 %[a], %[a], %[q]\n\t"                                             \
    "s_xor_b32 %[u], %[e], " DREG "\n\t"                                         \
    "s_lshl_b32 %[t], %[a], 16\n\t"                                              \
    "s_andn2_b32 %[t], %[t], %[u]\n\t"                                           \
    "s_add_u32 %[c], %[c], %[q]\n\t"                                             \
    "s_cmp_lt_i32 %[t], 0\n\t"                                                   \
    "s_cbranch_scc1 1" ID "1f\n\t"                                               \
    "s_sub_u32 %[t], %[a], %[q]\n\t"                                             \
    "s_xor_b32 %[t], %[t], %[u]\n\t"                                             \
    "s_cmp_lt_i32 %[t], 0\n\t"                                                   \
    "s_cselect_b32 %[a], %[q], %[a]\n\t"                                         \
    "s_cselect_b32 %[t], %[q], 0\n\t"                                            \
    "s_sub_u32 %[c], %[c], %[t]\n\t"                                             \
    "s_cmp_lt_i32 %[u], 0\n\t"                                                   \
    "s_cselect_b32 %[t], 0x60016, %[f16]\n\t"                                    \
    "s_bfe_u32 %[t], %[e], %[t]\n\t"                                             \
    "v_readlane_b32 %[t], %[tab], %[t]\n\t"                                      \
    "s_and_b32 %[q], %[e], 0x80000000\n\t"                                       \
    "s_xor_b32 %[t], %[t], %[q]\n\t"                                             \
    SWRITE                                                                       \
    "s_flbit_i32_b32 %[t], %[a]\n\t"                                             \
    "s_sub_u32 %[t], %[t], 16\n"                                                  \
    "1" ID "2:\n\t"                                                              \
    "s_cmp_lt_u32 %[t], %[ct]\n\t"                                               \
    "s_cbranch_scc0 1" ID "3f\n\t"                                               \
    "s_lshl_b32 %[a], %[a], %[t]\n\t"                                            \
    "s_lshl_b32 %[c], %[c], %[t]\n\t"                                            \
    "s_sub_u32 %[ct], %[ct], %[t]\n"                                              \
    "1" ID "1:\n\t"
#define MQE_G_BO(ID) "1" ID "3:\n\t" MQE_BYTEOUT("1" ID "4", "1" ID "5", "1" ID "6", "1" ID "7", "1" ID "2b")
#define MQE_G_WORD(I0, I1, I2, I3, X)                                              \
    MQE_G_SYM(I0, X, "0x50000", "s_lshl_b32 %[d], %[" X "], 24\n\t", "%[d]")     \
    MQE_G_SYM(I1, X, "0x50008", "s_lshl_b32 %[d], %[" X "], 16\n\t", "%[d]")     \
    MQE_G_SYM(I2, X, "0x50010", "s_lshl_b32 %[d], %[" X "], 8\n\t", "%[d]")      \
    MQE_G_SYM(I3, X, "0x50018", "", "%[" X "]")

// The same with the context states in s[80:98] (the decoder's block, see
// states_init): s_movrels / s_movreld instead of the lane read and write.
// An SALU write of M0 needs one wait state before s_movrels; the shift
// forming the decision fills it for the first three bytes of a word.
#define MQE_S_READ "s_movrels_b32 %[e], s80\n\t"
#define MQE_S_WRITE "s_movreld_b32 s80, %[t]\n\t"
#define MQE_GS_WORD(I0, I1, I2, I3, X)                                                                   \
    MQE_G_SYM_AT(I0, X, "0x50000", "s_lshl_b32 %[d], %[" X "], 24\n\t", "%[d]", MQE_S_READ, MQE_S_WRITE) \
    MQE_G_SYM_AT(I1, X, "0x50008", "s_lshl_b32 %[d], %[" X "], 16\n\t", "%[d]", MQE_S_READ, MQE_S_WRITE) \
    MQE_G_SYM_AT(I2, X, "0x50010", "s_lshl_b32 %[d], %[" X "], 8\n\t", "%[d]", MQE_S_READ, MQE_S_WRITE)  \
    MQE_G_SYM_AT(I3, X, "0x50018", "s_nop 0\n\t", "%[" X "]", MQE_S_READ, MQE_S_WRITE)

// 16 decisions, the symbol bytes of words x0..x3 in order (at most 32
// bytes move out).
__device__ __forceinline__ void mqe_group16(MqE& m, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
    uint32_t e, q, u, t, d;
    asm volatile(
        MQE_G_WORD("00", "01", "02", "03", "x0")
        MQE_G_WORD("04", "05", "06", "07", "x1")
        MQE_G_WORD("08", "09", "10", "11", "x2")
        MQE_G_WORD("12", "13", "14", "15", "x3")
        "s_branch 9f\n"
        MQE_G_BO("00") MQE_G_BO("01") MQE_G_BO("02") MQE_G_BO("03")
        MQE_G_BO("04") MQE_G_BO("05") MQE_G_BO("06") MQE_G_BO("07")
        MQE_G_BO("08") MQE_G_BO("09") MQE_G_BO("10") MQE_G_BO("11")
        MQE_G_BO("12") MQE_G_BO("13") MQE_G_BO("14") MQE_G_BO("15")
        "9:"
        : [a] "+s"(m.a), [c] "+s"(m.c), [ct] "+s"(m.ct), [pos] "+s"(m.pos), [last] "+s"(m.last),
          [cx] "+v"(m.cx), [win] "+v"(m.win), [e] "=&s"(e), [q] "=&s"(q), [u] "=&s"(u),
          [t] "=&s"(t), [d] "=&s"(d)
        : [x0] "s"(x0), [x1] "s"(x1), [x2] "s"(x2), [x3] "s"(x3), [tab] "v"(m.tab), [f16] "s"(0x60010u)
        : "scc", "m0");
}
// mqe_group16 with the context states in s[80:98].
__device__ __forceinline__ void mqe_group16s(MqE& m, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
    uint32_t e, q, u, t, d;
    asm volatile(
        MQE_GS_WORD("00", "01", "02", "03", "x0")
        MQE_GS_WORD("04", "05", "06", "07", "x1")
        MQE_GS_WORD("08", "09", "10", "11", "x2")
        MQE_GS_WORD("12", "13", "14", "15", "x3")
        "s_branch 9f\n"
        MQE_G_BO("00") MQE_G_BO("01") MQE_G_BO("02") MQE_G_BO("03")
        MQE_G_BO("04") MQE_G_BO("05") MQE_G_BO("06") MQE_G_BO("07")
        MQE_G_BO("08") MQE_G_BO("09") MQE_G_BO("10") MQE_G_BO("11")
        MQE_G_BO("12") MQE_G_BO("13") MQE_G_BO("14") MQE_G_BO("15")
        "9:"
        : [a] "+s"(m.a), [c] "+s"(m.c), [ct] "+s"(m.ct), [pos] "+s"(m.pos), [last] "+s"(m.last),
          [win] "+v"(m.win), [e] "=&s"(e), [q] "=&s"(q), [u] "=&s"(u),
          [t] "=&s"(t), [d] "=&s"(d)
        : [x0] "s"(x0), [x1] "s"(x1), [x2] "s"(x2), [x3] "s"(x3), [tab] "v"(m.tab), [f16] "s"(0x60010u)
        : "scc", "m0");
}
// Store ring bytes [wbase, wbase + 32) once they are all committed (byte
// pos - 1 is still held).
__device__ __forceinline__ void mqe_spill(MqE& m) {
    if (m.pos > m.wbase + 32u) {
        const uint32_t k = (lid() - 1u - m.wbase) & 63u;
        if (m.wbase + 32u <= m.cap) {
            if (k < 32u) m.out[m.wbase + k] = (uint8_t)m.win;
        } else {
            m.overflow = 1u;
        }
        m.wbase += 32u;
    }
}
// Code the first cnt (<= 256) symbol bytes held in the wave's lanes (lane i:
// bytes 4i..4i+3).
__device__ __forceinline__ void mqe_chunk(MqE& m, uint32_t sv, uint32_t cnt) {
    uint32_t q = 0;
    // 16 symbols move at most 32 bytes out, so one spill check per 16
    // keeps the ring from reusing an unstored byte (at most 64 pending)
    for (; q + 16u <= cnt; q += 16u) {
        mqe_group16(m, rl(sv, q >> 2), rl(sv, (q >> 2) + 1u), rl(sv, (q >> 2) + 2u), rl(sv, (q >> 2) + 3u));
        mqe_spill(m);
    }
    for (; q + 4u <= cnt; q += 4u) {
        const uint32_t w = rl(sv, q >> 2);
        mqe_symbol(m, w & 31u, w << 24);
        mqe_symbol(m, (w >> 8) & 31u, w << 16);
        mqe_symbol(m, (w >> 16) & 31u, w << 8);
        mqe_symbol(m, (w >> 24) & 31u, w);
        mqe_spill(m);
    }
    if (q < cnt) {
        uint32_t w = rl(sv, q >> 2);
        for (; q < cnt; q++, w >>= 8) mqe_symbol(m, w & 31u, w << 24);
        mqe_spill(m);
    }
}
// mqe_chunk with the context states in s[80:98].
__device__ __forceinline__ void mqe_chunk_s(MqE& m, uint32_t sv, uint32_t cnt) {
    uint32_t q = 0;
    // 16 symbols move at most 32 bytes out, so one spill check per 16
    // keeps the ring from reusing an unstored byte (at most 64 pending)
    for (; q + 16u <= cnt; q += 16u) {
        mqe_group16s(m, rl(sv, q >> 2), rl(sv, (q >> 2) + 1u), rl(sv, (q >> 2) + 2u), rl(sv, (q >> 2) + 3u));
        mqe_spill(m);
    }
    for (; q + 4u <= cnt; q += 4u) {
        const uint32_t w = rl(sv, q >> 2);
        mqe_symbol_s(m, w & 31u, w << 24);
        mqe_symbol_s(m, (w >> 8) & 31u, w << 16);
        mqe_symbol_s(m, (w >> 16) & 31u, w << 8);
        mqe_symbol_s(m, (w >> 24) & 31u, w);
        mqe_spill(m);
    }
    if (q < cnt) {
        uint32_t w = rl(sv, q >> 2);
        for (; q < cnt; q++, w >>= 8) mqe_symbol_s(m, w & 31u, w << 24);
        mqe_spill(m);
    }
}
// Code the n symbol bytes of one pass from LDS.
__device__ __forceinline__ void mqe_pass(MqE& m, const uint32_t* symw, uint32_t n) {
    for (uint32_t base = 0; base < n; base += 256u) mqe_chunk(m, symw[(base >> 2) + lid()], min(256u, n - base));
}
// Byte-out for the flush (same as the asm sequence above).
__device__ __forceinline__ void mqe_byteout(MqE& m) {
    uint32_t b;
    if (m.last == 0xffu) {
        b = m.c >> 20;
        m.c &= 0xfffffu;
        m.ct = 7;
    } else if ((m.c & 0x8000000u) == 0) {
        b = m.c >> 19;
        m.c &= 0x7ffffu;
        m.ct = 8;
    } else {
        m.last = (m.last + 1u) & 0xffu;
        if (m.last == 0xffu) {
            m.c &= 0x7ffffffu;
            b = m.c >> 20;
            m.c &= 0xfffffu;
            m.ct = 7;
        } else {
            b = m.c >> 19;
            m.c &= 0x7ffffu;
            m.ct = 8;
        }
    }
    m.win = wlane(m.win, m.last, m.pos & 63u);
    m.pos++;
    m.last = b & 0xffu;
}
// flush + take_bytes (mq/mod.rs:244-250, :373-379, :416-423); returns the
// terminated length.
__device__ __forceinline__ uint32_t mqe_flush(MqE& m) {
    mqe_spill(m);
    uint32_t temp = m.c + m.a;
    m.c |= 0xffffu;
    if (m.c >= temp) m.c -= 0x8000u;
    m.c <<= m.ct;
    mqe_byteout(m);
    m.c <<= m.ct;
    mqe_byteout(m);
    const uint32_t total = m.last == 0xffu ? m.pos - 1u : m.pos;
    m.win = wlane(m.win, m.last, m.pos & 63u);  // commit the held byte
    const uint32_t nwin = m.pos - m.wbase;    // <= 34 bytes pending
    const uint32_t k = (lid() - 1u - m.wbase) & 63u;
    if (k < nwin) {
        const uint32_t o = m.wbase + k;
        if (o < m.cap) m.out[o] = (uint8_t)m.win;
    }
    if (m.wbase + nwin > m.cap) m.overflow = 1u;
    return total;
}

}  // namespace

// Load one code-block, quantise it (9/7) and form the block statistics of
// t1.rs:215-258: lane x gets column x as row magnitudes and a negative mask.
// Returns mbp (0 for an all-zero block).
__device__ __forceinline__ int load_block(const BlockJob& J, uint32_t (&mag)[64], uint64_t& neg, BlockResult& R) {
    const int lane = lid();
    const int w = J.w, h = J.h;
    const bool lossy = J.step > 0.0f;
    neg = 0;
    uint32_t mx = 0, nz = 0;
#pragma unroll
    for (int y = 0; y < 64; y++) {
        uint32_t m = 0;
        if (y < h && lane < w) {
            size_t idx = (size_t)(J.y0 + y) * J.stride + J.x0 + lane;
            int32_t q = lossy ? quantize(((const float*)J.plane)[idx], J.step) : ((const int32_t*)J.plane)[idx];
            m = q < 0 ? (uint32_t)(-(int64_t)q) : (uint32_t)q;
            if (q < 0) neg |= 1ull << y;
            if (q) nz++;
        }
        mag[y] = m;
        mx = max(mx, m);
    }
    // The butterfly leaves the same value in every lane; readfirstlane tells
    // the compiler so, keeping the bit-plane loop and the coder in SGPRs.
    mx = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_max(mx));
    nz = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_sum(nz));
    const int mbp = mx ? 32 - __builtin_clz(mx) : 0;
    R.max_mag = mx;
    R.nnz = (uint16_t)nz;
    R.mbp = (uint8_t)mbp;
    R.zbp = (uint8_t)(J.mb > mbp ? J.mb - mbp : 0);
    float frac = (float)nz / (float)(w * h);  // profile.rs:122-130
    R.cls = mbp == 0 ? 0 : (frac <= 0.05f ? 0 : (frac <= 0.30f ? 1 : 2));
    R.overflow = 0;
    R.pad = 0;
    R.nbytes = 0;
    R.npasses = 0;
    return mbp;
}

// ---- wave-coder schedule ----------------------------------------------------
// A SIMD runs about two coder waves at once, and the kernel lasts as long as
// its most loaded SIMD (a block with many bit-planes runs at raised issue
// priority and starves the wave beside it).  So blocks are handed out in
// descending estimated cost: the first wave to start on a SIMD takes the next
// heavy block, any later wave there a light one.  When every wave of a
// launch is resident and the waves do not divide evenly, some SIMDs start
// with three, and a heavy block beside two others ends late (measured on a
// 2550x3300 page: the launch ended 0.3 ms after its heaviest block).  For
// such launches (`three`) a SIMD's first wave waits until the launch's waves
// have arrived (at most kPickWaitTicks) and takes a light block too if its
// SIMD runs three.  The order is laid out for the two ends: the `split`
// heaviest blocks in descending cost from the front, where the heavy takers
// read, and the rest in ascending cost from there to the back, where the
// light takers read from the end inward.  So the light takers get the
// heavier of the light blocks first and the lightest last: the waves that
// arrive last (third waves and the first waves of SIMDs with three, which
// decide after their wait) get the lightest, and a launch whose waves are
// not all resident hands out its light end longest first.  sched holds
// cost[n] | order[n] | heavy count, light count, arrivals | arrivals per SIMD.
constexpr uint32_t kSimdSlots = 1u << 14;  // xcc(4) se(3) sh(1) cu(4) simd(2)
constexpr uint64_t kPickWaitTicks = 2500;   // 25 us of the 100 MHz real-time clock

// Estimated cost of a block in MQ decisions: fitted on document and photo
// pages (correlation 0.98 with the decisions coded), plus about 470
// decisions' worth of context formation per bit-plane.
__global__ __launch_bounds__(64) void t1_cost_kernel(const BlockJob* __restrict__ jobs, uint32_t* __restrict__ sched) {
    const BlockJob J = jobs[blockIdx.x];
    uint32_t mag[64];
    uint64_t neg;
    BlockResult R;
    const int mbp = load_block(J, mag, neg, R);
    uint32_t bits = 0, z = 0;
#pragma unroll
    for (int y = 0; y < 64; y++) {
        const uint32_t bl = mag[y] ? 32u - (uint32_t)__builtin_clz(mag[y]) : 0u;
        if (y < J.h && (int)lid() < J.w) {
            bits += bl;
            z += (uint32_t)mbp - bl;
        }
    }
    bits = wave_sum(bits);
    z = wave_sum(z);
    if (lid() == 0) sched[blockIdx.x] = mbp ? (9u * bits + 4u * z + 12u * R.nnz) / 8u + 470u * (uint32_t)mbp : 0u;
}
// Counting sort of the blocks by descending cost (512-decision buckets), the
// ranks from `split` on placed back to front, and reset of the hand-out
// counters.  One workgroup.
__global__ __launch_bounds__(1024) void t1_order_kernel(int n, uint32_t split, uint32_t* __restrict__ sched) {
    __shared__ uint32_t hist[256];
    const int tid = threadIdx.x;
    if (tid < 256) hist[tid] = 0;
    uint32_t* ctrl = sched + 2 * (size_t)n;
    for (uint32_t i = tid; i < 3u + kSimdSlots; i += 1024u) ctrl[i] = 0;
    __syncthreads();
    for (int b = tid; b < n; b += 1024) atomicAdd(&hist[255u - min(255u, sched[b] >> 9)], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int k = 0; k < 256; k++) {
            const uint32_t c = hist[k];
            hist[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (int b = tid; b < n; b += 1024) {
        const uint32_t k = atomicAdd(&hist[255u - min(255u, sched[b] >> 9)], 1u);
        sched[n + (k < split ? k : split + ((uint32_t)n - 1u - k))] = (uint32_t)b;
    }
}
// The block this wave codes.  The n waves of the launch take n distinct
// positions of the order (heavy ones from the front, light ones from the
// back), wherever they are placed and however the wait ends, so every block
// is coded exactly once.  The wait is bounded by the clock: waves beyond the
// resident ones start only as others end, and then the total never reaches n
// within it.
__device__ __forceinline__ uint32_t t1_pick_block(uint32_t n, uint32_t three, uint32_t* sched, uint32_t& pos) {
    uint32_t b = 0, at = 0;
    if (lid() == 0) {
        uint32_t hw, xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)\n\ts_getreg_b32 %1, hwreg(HW_REG_XCC_ID)" : "=s"(hw), "=s"(xcc));
        const uint32_t simd = ((xcc & 15u) << 10) | (((hw >> 13) & 7u) << 7) | (((hw >> 12) & 1u) << 6) |
                              (((hw >> 8) & 15u) << 2) | ((hw >> 4) & 3u);
        uint32_t* ctrl = sched + 2 * (size_t)n;
        uint32_t* const here = &ctrl[3 + simd];
        bool heavy = atomicAdd(here, 1u) == 0u;
        atomicAdd(&ctrl[2], 1u);
        if (heavy && three) {
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            while (__hip_atomic_load(&ctrl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n &&
                   __builtin_amdgcn_s_memrealtime() - t0 < kPickWaitTicks)
                __builtin_amdgcn_s_sleep(4);
            heavy = __hip_atomic_load(here, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 3u;
        }
        const uint32_t k = atomicAdd(&ctrl[heavy ? 0 : 1], 1u);
        at = heavy ? k : n - 1u - k;
        b = sched[n + at];
    }
    pos = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
}

// ---- heaviest blocks formed ahead, one wave per bit-plane --------------------
// A wave codes its block's bit-planes one after another, forming each pass
// (lane-parallel) and then coding its decisions (scalar), so the heaviest
// block of a launch, which sets the launch's length, also pays for its own
// formation.  The coder state at the start of plane p depends on the
// magnitudes alone: significant = magnitude >= 2^(p+1), refined =
// magnitude >= 2^(p+2), nothing visited.  So t1_preform_kernel forms the
// planes of the kPre heaviest blocks (order positions 0..kPre-1 with at least
// half the cost of position 0) in parallel, one wave per plane, and the
// coder wave that takes such a position only codes.
//
// Per position a slot after the hand-out counters in sched: pass k's start
// (in words from the slot's data) at [k], its symbol count | new
// significance << 16 at [kMaxPasses + k], then kPrePlanes plane regions of
// kPlaneWords.  In a plane region the passes follow each other, each
// rounded up to 64 bytes (the coder's read unit); a plane has at most 10
// symbols per stripe-column, 10240 in all.
constexpr uint32_t kPre = 32;
constexpr uint32_t kPrePlanes = 32;
constexpr uint32_t kPlaneWords = (kSymCap + 3u * 64u) / 4u;
constexpr uint32_t kPreSlotWords = 2u * kMaxPasses + kPrePlanes * kPlaneWords;
__host__ __device__ __forceinline__ size_t pre_base(uint32_t n) {
    return (2 * (size_t)n + 3 + kSimdSlots + 63) & ~(size_t)63;
}
__device__ __forceinline__ bool pre_formed(const uint32_t* sched, uint32_t n, uint32_t pos) {
    if (pos >= kPre || pos >= n) return false;
    const uint32_t c0 = sched[sched[n]], c = sched[sched[n + pos]];
    return c > 0u && 2u * c >= c0;
}

// One wave per (position, plane): the plane's passes of a pre-formed
// position, in the coder's pass order (plane mbp-1: cleanup only, pass 0;
// plane p below: passes 1 + 3 (mbp - 2 - p) + 0, 1, 2).
__global__ __launch_bounds__(64) void t1_preform_kernel(const BlockJob* __restrict__ jobs, const uint32_t* __restrict__ tabs,
                                                        uint32_t n, uint32_t* __restrict__ sched) {
    __shared__ uint32_t symw[kSymCap / 4];
    uint8_t* symb = reinterpret_cast<uint8_t*>(symw);
    const uint32_t lane = lid();
    const uint32_t slot = blockIdx.x / kPrePlanes, p = blockIdx.x % kPrePlanes;
    if (!pre_formed(sched, n, slot)) return;
    const BlockJob J = jobs[sched[n + slot]];
    const int w = J.w, h = J.h;
    uint32_t mag[64];
    uint64_t neg;
    BlockResult R;
    const int mbp = load_block(J, mag, neg, R);
    if ((int)p >= mbp) return;
    Tables t;
    load_tables(t, tabs, J.band);
    uint64_t bits = 0, sig = 0, ref = 0;
#pragma unroll
    for (int y = 0; y < 64; y++) {
        const uint64_t v = mag[y];
        bits |= ((v >> p) & 1u) << y;
        sig |= (uint64_t)((v >> (p + 1u)) != 0u) << y;
        ref |= (uint64_t)((v >> (p + 2u)) != 0u) << y;
    }
    Col st = {sig, sig & neg, 0, ref};
    uint32_t* slotp = sched + pre_base(n) + (size_t)slot * kPreSlotWords;
    uint32_t* data = slotp + 2u * kMaxPasses + p * kPlaneWords;
    const uint32_t top = (uint32_t)mbp - 1u;
    uint32_t k = p == top ? 0u : 1u + 3u * (top - 1u - p), off = 0;
    auto put = [&](uint32_t cnt, uint32_t ns) {
        __syncthreads();
        for (uint32_t i = lane; i < (cnt + 3u) >> 2; i += 64u) data[off + i] = symw[i];
        if (lane == 0) {
            slotp[k] = p * kPlaneWords + off;
            slotp[kMaxPasses + k] = cnt | ns << 16;
        }
        off += ((cnt + 63u) >> 6) << 4;
        k++;
        __syncthreads();
    };
    uint32_t ns, cnt;
    if (p != top) {
        cnt = form_spp(t, st, bits, neg, w, h, symb, ns);
        put(cnt, ns);
        cnt = form_mrp(st, bits, w, h, symb);
        put(cnt, 0);
    }
    cnt = form_cup(t, st, bits, neg, w, h, symb, ns);
    put(cnt, ns);
}

// 64 symbol bytes, read with scalar loads (the read unit of pre-formed
// passes; 64-byte aligned).  The region is written by an earlier launch only,
// so it is read through the constant address space: the coder's asm
// statements would otherwise count as possible stores to it and turn the
// loads into vector loads.
struct W16 {
    uint32_t w[16];
};
typedef const __attribute__((address_space(4))) uint32_t const_u32;
__device__ __forceinline__ W16 load16(const uint32_t* p) {
    const const_u32* q = (const const_u32*)p;
    W16 r;
#pragma unroll
    for (int i = 0; i < 16; i++) r.w[i] = q[i];
    return r;
}

// Tier-1 encoder: one wavefront per code-block (the block from
// t1_pick_block).  Outputs the terminated segment into out + job.out_off,
// the per-pass byte snapshots and newly-significant counts, and the block
// statistics.
__global__ __launch_bounds__(64) __attribute__((amdgpu_num_sgpr(kStateSgprCap), amdgpu_waves_per_eu(1, 8))) void t1_encode_kernel(const BlockJob* __restrict__ jobs, const uint32_t* __restrict__ tabs,
                                                       uint8_t* __restrict__ out, BlockResult* __restrict__ results,
                                                       uint32_t* __restrict__ snaps, uint16_t* __restrict__ newsigs,
                                                       uint32_t* __restrict__ sched, const uint32_t* __restrict__ pre,
                                                       uint32_t three) {
    extern __shared__ uint32_t symw[];
    uint8_t* symb = reinterpret_cast<uint8_t*>(symw);
    const int lane = lid();
    uint32_t pos;
    const uint32_t blk = t1_pick_block(gridDim.x, three, sched, pos);
    const BlockJob J = jobs[blk];
    const int w = J.w, h = J.h;
    uint32_t mag[64];
    uint64_t neg;
    BlockResult R;
    const int mbp = load_block(J, mag, neg, R);
    if (mbp == 0) {
        if (lane == 0) results[blk] = R;
        return;
    }
    // The kernel ends with its heaviest block; blocks with many bit-planes
    // get issue priority over the light block sharing their SIMD.
    set_priority(mbp >= 12 ? 3 : mbp >= 10 ? 2 : mbp >= 8 ? 1 : 0);
    Tables t;
    load_tables(t, tabs, J.band);
    MqE m;
    mqe_init(m, t.mq, out + J.out_off, J.out_cap);
    // initial states (mq/mod.rs:195-200): UNI 46, AGG 3, ZC 4, others 0; MPS 0
    m.cx = (uint32_t)__shfl((int)t.mq, lane == CTX_UNI ? 46 : lane == CTX_AGG ? 3 : lane == CTX_ZC ? 4 : 0, 64);
    Col st = {0, 0, 0, 0};
    uint32_t snap0 = 0, snap1 = 0, ns0 = 0, ns1 = 0;
    uint32_t np = 0;
    auto record = [&](uint32_t n) {
        if (np < 64u) {
            snap0 = wlane(snap0, m.pos, np);
            ns0 = wlane(ns0, n, np);
        } else {
            snap1 = wlane(snap1, m.pos, np - 64u);
            ns1 = wlane(ns1, n, np - 64u);
        }
        np++;
    };
    const int top = mbp - 1;
    if (pre_formed(sched, gridDim.x, pos)) {
        // Code the pre-formed passes: 64 symbols per step from SGPRs, the
        // next 64 (or the next pass's first) loaded while these are coded.
        // The context states move to s[80:98] for the block: the wave coding
        // the heaviest block sets the launch's length, and the scalar reads
        // and writes shorten its decisions (the light blocks keep the states
        // in a VGPR, which leaves the scalar unit to the waves beside them).
        states_init(m.cx);
        const uint32_t* slotp = pre + (size_t)pos * kPreSlotWords;
        const uint32_t* data = slotp + 2u * kMaxPasses;
        const uint32_t hs0 = slotp[lane], hn0 = slotp[kMaxPasses + lane];
        const uint32_t hs1 = lane < kMaxPasses - 64 ? slotp[64 + lane] : 0u;
        const uint32_t hn1 = lane < kMaxPasses - 64 ? slotp[kMaxPasses + 64 + lane] : 0u;
        const uint32_t npass = 3u * (uint32_t)mbp - 2u;
        uint32_t at = rl(hs0, 0);
        W16 cur = load16(data + at);
        for (uint32_t k = 0; k < npass; k++) {
            const uint32_t S = k < 64u ? rl(hs0, k) : rl(hs1, k - 64u);
            const uint32_t nn = k < 64u ? rl(hn0, k) : rl(hn1, k - 64u);
            const uint32_t cnt = nn & 0xffffu;
            if (S != at) {
                at = S;
                cur = load16(data + at);
            }
            const uint32_t k1 = k + 1u;
            const uint32_t after = k1 >= npass ? at : k1 < 64u ? rl(hs0, k1) : rl(hs1, k1 - 64u);
            uint32_t b = 0;
            for (; b + 64u <= cnt; b += 64u) {
                const uint32_t nxt = b + 64u < cnt ? at + 16u : after;
                const W16 nx = load16(data + nxt);
                mqe_group16s(m, cur.w[0], cur.w[1], cur.w[2], cur.w[3]);
                mqe_spill(m);
                mqe_group16s(m, cur.w[4], cur.w[5], cur.w[6], cur.w[7]);
                mqe_spill(m);
                mqe_group16s(m, cur.w[8], cur.w[9], cur.w[10], cur.w[11]);
                mqe_spill(m);
                mqe_group16s(m, cur.w[12], cur.w[13], cur.w[14], cur.w[15]);
                mqe_spill(m);
                cur = nx;
                at = nxt;
            }
            if (b < cnt) {
                // last partial 64: into a VGPR for the generic chunk coder
                const W16 nx = load16(data + after);
                uint32_t sv = 0;
#pragma unroll
                for (int i = 0; i < 16; i++) sv = wlane(sv, cur.w[i], (uint32_t)i);
                mqe_chunk_s(m, sv, cnt - b);
                cur = nx;
                at = after;
            }
            record(nn >> 16);
        }
    } else
    for (int p = top; p >= 0; p--) {
        uint64_t bits = 0;
#pragma unroll
        for (int y = 0; y < 64; y++) bits |= (uint64_t)((mag[y] >> p) & 1u) << y;
        uint32_t ns, n;
        if (p != top) {
            n = form_spp(t, st, bits, neg, w, h, symb, ns);
            __syncthreads();
            mqe_pass(m, symw, n);
            record(ns);
            __syncthreads();
            n = form_mrp(st, bits, w, h, symb);
            __syncthreads();
            mqe_pass(m, symw, n);
            record(0);
            __syncthreads();
        }
        n = form_cup(t, st, bits, neg, w, h, symb, ns);
        __syncthreads();
        mqe_pass(m, symw, n);
        record(ns);
        __syncthreads();
        st.vis = 0;
    }
    const uint32_t total = mqe_flush(m);
    R.nbytes = total;
    R.npasses = (uint8_t)np;
    R.overflow = (uint8_t)(m.overflow ? 1 : 0);
    uint32_t* sp = snaps + (size_t)blk * kMaxPasses;
    uint16_t* nsp = newsigs + (size_t)blk * kMaxPasses;
    if ((uint32_t)lane < np) {
        sp[lane] = snap0;
        nsp[lane] = (uint16_t)ns0;
    }
    if ((uint32_t)lane + 64u < np && lane + 64 < kMaxPasses) {
        sp[lane + 64] = snap1;
        nsp[lane + 64] = (uint16_t)ns1;
    }
    if (lane == 0) results[blk] = R;
}

// ---- Tier-1 counters (encode/counters.rs:3-9) ----------------------------------
// MQ decisions per coding pass of every block: the same context formation as
// the coders (every formed symbol is one MQ decision, mq/mod.rs:213-216),
// with no coding.  Runs only when counting is on, after the coder, so the
// coding kernels stay as they are.  counts[b * kMaxPasses + k].
__global__ __launch_bounds__(64) void t1_count_kernel(const BlockJob* __restrict__ jobs,
                                                      const uint32_t* __restrict__ tabs,
                                                      uint32_t* __restrict__ counts) {
    __shared__ uint32_t symw[kSymCap / 4];
    uint8_t* symb = reinterpret_cast<uint8_t*>(symw);
    const int lane = lid();
    const BlockJob J = jobs[blockIdx.x];
    const int w = J.w, h = J.h;
    uint32_t mag[64];
    uint64_t neg;
    BlockResult R;
    const int mbp = load_block(J, mag, neg, R);
    if (mbp == 0) return;
    Tables t;
    load_tables(t, tabs, J.band);
    Col st = {0, 0, 0, 0};
    uint32_t c0 = 0, c1 = 0, np = 0;
    auto count = [&](uint32_t n) {
        __syncthreads();
        if (np < 64u) c0 = wlane(c0, n, np);
        else c1 = wlane(c1, n, np - 64u);
        np++;
    };
    const int top = mbp - 1;
    for (int p = top; p >= 0; p--) {
        uint64_t bits = 0;
#pragma unroll
        for (int y = 0; y < 64; y++) bits |= (uint64_t)((mag[y] >> p) & 1u) << y;
        uint32_t ns;
        if (p != top) {
            count(form_spp(t, st, bits, neg, w, h, symb, ns));
            __syncthreads();
            count(form_mrp(st, bits, w, h, symb));
            __syncthreads();
        }
        count(form_cup(t, st, bits, neg, w, h, symb, ns));
        __syncthreads();
        st.vis = 0;
    }
    uint32_t* cp = counts + (size_t)blockIdx.x * kMaxPasses;
    if ((uint32_t)lane < np) cp[lane] = c0;
    if ((uint32_t)lane + 64u < np && lane + 64 < kMaxPasses) cp[lane + 64] = c1;
}

// ---- batched Tier-1 encoder: context formation, then one lane per block ------
//
// For large batches the MQ coder runs one code-block per lane: a VALU
// instruction then advances 64 coders, where the wave-per-block coder above
// spends one scalar instruction on one.  t1_form_kernel writes each block's
// (context, decision) bytes, all passes in order, into one contiguous
// stream; t1_mq_kernel codes the streams.  Every active lane consumes one
// symbol per step, so the lanes of a wave stay in step and their 16-byte
// stream reads are issued together.

// Stream region of a block with mbp bit-planes: per plane at most one
// symbol per sample plus two extra per run-mode stripe-column (AGG=1 adds
// AGG and two UNI and drops the samples before the run length), plus one
// sign per sample, per pass one marker and up to three padding bytes, and
// the reader's 32-byte look-ahead; 16-byte aligned.
__host__ __device__ __forceinline__ uint64_t stream_bound(uint32_t mbp) {
    return ((uint64_t)mbp * 6144u + 4096u + 4u * kMaxPasses + 32u + 15u) & ~(uint64_t)15u;
}
// Pass marker: context 31, whose entry (Qe = 0, MPS 0) makes coding it a
// no-op (a stays >= 0x8000, c += 0, no renormalisation).  Each pass's
// region (symbols, padding, marker) is a multiple of 4 bytes, padded with
// context 30, another no-op entry that records nothing, so a reader may take
// a pass as whole words.
constexpr uint32_t kMarker = 31u;
constexpr uint32_t kPad = 30u;

// Context formation for the lane-per-block coder: same passes as the fused
// kernel, symbols copied from LDS to the block's stream in global memory,
// each pass followed by a marker.  pend[b * kMaxPasses + k] = stream length
// up to and including pass k's marker.
__global__ __launch_bounds__(64) void t1_form_kernel(const BlockJob* __restrict__ jobs, const uint32_t* __restrict__ tabs,
                                                     BlockResult* __restrict__ results, uint32_t* __restrict__ pend,
                                                     uint16_t* __restrict__ newsigs, uint32_t* __restrict__ soff16,
                                                     uint8_t* __restrict__ pool, unsigned long long* __restrict__ pool_used,
                                                     unsigned long long pool_cap) {
    __shared__ uint32_t symw[kSymCap / 4];
    uint8_t* symb = reinterpret_cast<uint8_t*>(symw);
    const int lane = lid();
    const BlockJob J = jobs[blockIdx.x];
    const int w = J.w, h = J.h;
    uint32_t mag[64];
    uint64_t neg;
    BlockResult R;
    const int mbp = load_block(J, mag, neg, R);
    if (mbp == 0) {
        if (lane == 0) {
            results[blockIdx.x] = R;
            soff16[blockIdx.x] = 0;
        }
        return;
    }
    const uint64_t bound = stream_bound((uint32_t)mbp);
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(pool_used, (unsigned long long)bound);
    base = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(base >> 32), 0) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, 0);
    bool fits = base + bound <= pool_cap;
    uint8_t* dst = pool + base;
    Tables t;
    load_tables(t, tabs, J.band);
    Col st = {0, 0, 0, 0};
    uint32_t end0 = 0, end1 = 0, ns0 = 0, ns1 = 0;
    uint32_t np = 0, cum = 0;
    auto emit_pass = [&](uint32_t n, uint32_t ns) {
        __syncthreads();
        const uint32_t len = (n + 4u) & ~3u;  // symbols, padding, marker
        fits = fits && cum + len + 32u <= bound;  // never past the region, whatever the bound's margin
        if (fits) {
            for (uint32_t i = lane; i < len - 1u; i += 64u) dst[cum + i] = i < n ? symb[i] : (uint8_t)kPad;
            if (lane == 0) dst[cum + len - 1u] = (uint8_t)kMarker;
        }
        cum += len;
        if (np < 64u) {
            end0 = wlane(end0, cum, np);
            ns0 = wlane(ns0, ns, np);
        } else {
            end1 = wlane(end1, cum, np - 64u);
            ns1 = wlane(ns1, ns, np - 64u);
        }
        np++;
        __syncthreads();
    };
    const int top = mbp - 1;
    for (int p = top; p >= 0; p--) {
        uint64_t bits = 0;
#pragma unroll
        for (int y = 0; y < 64; y++) bits |= (uint64_t)((mag[y] >> p) & 1u) << y;
        uint32_t ns, n;
        if (p != top) {
            n = form_spp(t, st, bits, neg, w, h, symb, ns);
            emit_pass(n, ns);
            n = form_mrp(st, bits, w, h, symb);
            emit_pass(n, 0);
        }
        n = form_cup(t, st, bits, neg, w, h, symb, ns);
        emit_pass(n, ns);
        st.vis = 0;
    }
    R.npasses = (uint8_t)np;
    R.overflow = fits ? 0 : 2;
    uint32_t* pe = pend + (size_t)blockIdx.x * kMaxPasses;
    uint16_t* nsp = newsigs + (size_t)blockIdx.x * kMaxPasses;
    if ((uint32_t)lane < np) {
        pe[lane] = end0;
        nsp[lane] = (uint16_t)ns0;
    }
    if ((uint32_t)lane + 64u < np && lane + 64 < kMaxPasses) {
        pe[lane + 64] = end1;
        nsp[lane + 64] = (uint16_t)ns1;
    }
    if (lane == 0) {
        results[blockIdx.x] = R;
        soff16[blockIdx.x] = (uint32_t)(base >> 4);
    }
}

namespace {
// One lane's MQ encoder state (mq/mod.rs:163-432); context entries live in
// LDS, cst[ctx * 64 + lane], in the table-entry form used by MqE.
struct LaneMq {
    uint32_t a, c, ct, pos, last, ovf;
    uint8_t* o;
    uint32_t cap;
};
// Byte-out (mq/mod.rs:373-414) as selects: the carry into the held byte,
// the 7-bit byte after 0xFF, then commit the held byte and hold the new one.
__device__ __forceinline__ void lmq_byteout(LaneMq& m) {
    const bool lff = m.last == 0xffu;
    const bool carry = !lff && (m.c & 0x8000000u);
    const uint32_t held = carry ? (m.last + 1u) & 0xffu : m.last;
    const bool clr = carry && held == 0xffu;
    const uint32_t c1 = clr ? m.c & 0x7ffffffu : m.c;
    const bool seven = lff || clr;
    const uint32_t byte = seven ? c1 >> 20 : c1 >> 19;
    m.c = c1 & (seven ? 0xfffffu : 0x7ffffu);
    m.ct = seven ? 7u : 8u;
    if (m.pos > 0) {
        if (m.pos - 1u < m.cap) m.o[m.pos - 1u] = (uint8_t)held;
        else m.ovf = 1u;
    }
    m.pos++;
    m.last = byte & 0xffu;
}
}  // namespace

// MQ coding of the streams, one code-block per lane (block = blockIdx.x * 64
// + lane).  Writes the segment bytes, the per-pass byte snapshots and the
// final length into the block's BlockResult.
//
// The step counter is wave-uniform: lane b codes symbol s of its stream at
// step s while s < its stream length.  Each step is branch-free apart from
// the byte output: the MPS-without-renormalisation case is the exchange
// formula with the state written back unchanged and a zero shift.
// Lane coder: this lane codes block b (b < 0: none); cst is 32 x 64 words
// and tab 64 words of LDS.
__device__ __forceinline__ void lane_mq(int b, const BlockJob* __restrict__ jobs, const uint32_t* __restrict__ tabs,
                                        const uint8_t* __restrict__ pool, const uint32_t* __restrict__ soff16,
                                        const uint32_t* __restrict__ pend, BlockResult* __restrict__ results,
                                        uint32_t* __restrict__ snaps, uint8_t* __restrict__ out, uint32_t* cst,
                                        uint32_t* tab) {
    const int lane = lid();
    tab[lane] = tabs[lane];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 19; k++)
        cst[k * 64 + lane] = tab[k == CTX_UNI ? 46 : k == CTX_AGG ? 3 : k == CTX_ZC ? 4 : 0];
    uint32_t np = 0, total = 0;
    if (b >= 0 && results[b].overflow == 0) np = results[b].npasses;
    const uint32_t* pe = pend + (size_t)b * kMaxPasses;
    if (np) total = pe[np - 1];
    const uint32_t smax = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_max(total));
    if (smax == 0) return;
    uint32_t* cx = cst + lane;
    const uint4* src = reinterpret_cast<const uint4*>(pool + (np ? (uint64_t)soff16[b] << 4 : 0));
    uint32_t* sn = snaps + (size_t)b * kMaxPasses;
    LaneMq m;
    m.a = 0x8000u;
    m.c = 0;
    m.ct = 12;
    m.pos = 0;
    m.last = 0;
    m.ovf = 0;
    m.o = nullptr;
    m.cap = 0;
    if (np) {
        const BlockJob& J = jobs[b];
        m.o = out + J.out_off;
        m.cap = J.out_cap;
    }
    uint32_t k = 0;
    uint4 cur = make_uint4(0, 0, 0, 0), nxt = cur;
    if (np) {
        cur = src[0];
        nxt = src[1];
    }
    // rows kMarker and kPad: Qe 0, MPS 0 (no-op symbols)
    cst[kMarker * 64 + lane] = 0;
    cst[kPad * 64 + lane] = 0;
    auto step = [&](uint32_t si, uint32_t sym) {
        sym = si < total ? sym & 0xffu : kMarker;  // past the end: no-op symbols
        if (si < total && (sym & 31u) == kMarker) sn[k++] = m.pos;  // end of a pass
        uint32_t* cp = cx + (sym & 31u) * 64u;
        const uint32_t e = *cp;
        const uint32_t qe = e & 0xffffu;
        const uint32_t a2 = m.a - qe;
        const uint32_t x = e ^ (sym << 24);  // bit 31: decision differs from the MPS
        const bool lps = (int32_t)x < 0;
        const bool keep = !lps && (a2 & 0x8000u);
        const bool to_qe = (a2 < qe) != lps;
        m.a = to_qe ? qe : a2;
        m.c += to_qe ? 0u : qe;
        const uint32_t next = (e >> (lps ? 22 : 16)) & 63u;
        const uint32_t ne = tab[next] ^ (e & 0x80000000u);
        *cp = keep ? e : ne;
        uint32_t sh = (uint32_t)__builtin_clz(m.a) - 16u;
        while (sh >= m.ct) {
            m.a <<= m.ct;
            m.c <<= m.ct;
            sh -= m.ct;
            lmq_byteout(m);
        }
        m.a <<= sh;
        m.c <<= sh;
        m.ct -= sh;
    };
    for (uint32_t s = 0; s < smax; s += 16u) {
        if (s) {
            cur = nxt;
            // look-ahead only within this lane's region: its stream plus the
            // 32 bytes of slack the region bound provides
            if (np && (s >> 4) + 1u <= (total + 15u) >> 4) nxt = src[(s >> 4) + 1];
        }
        const uint32_t w4[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t wd = w4[q];
            step(s + 4 * q, wd);
            step(s + 4 * q + 1, wd >> 8);
            step(s + 4 * q + 2, wd >> 16);
            step(s + 4 * q + 3, wd >> 24);
        }
    }
    if (!np) return;
    // flush + take_bytes (mq/mod.rs:244-250, :373-379, :416-423)
    const uint32_t temp = m.c + m.a;
    m.c |= 0xffffu;
    if (m.c >= temp) m.c -= 0x8000u;
    m.c <<= m.ct;
    lmq_byteout(m);
    m.c <<= m.ct;
    lmq_byteout(m);
    const uint32_t fin = m.last == 0xffu ? m.pos - 1u : m.pos;
    if (m.pos > 0) {
        if (m.pos - 1u < m.cap) m.o[m.pos - 1u] = (uint8_t)m.last;
        else m.ovf = 1u;
    }
    results[b].nbytes = fin;
    results[b].overflow = m.ovf ? 1 : 0;
}

__global__ __launch_bounds__(64) void t1_mq_kernel(int nb, const BlockJob* __restrict__ jobs,
                                                   const uint32_t* __restrict__ tabs, const uint8_t* __restrict__ pool,
                                                   const uint32_t* __restrict__ soff16, const uint32_t* __restrict__ pend,
                                                   BlockResult* __restrict__ results, uint32_t* __restrict__ snaps,
                                                   uint8_t* __restrict__ out) {
    __shared__ uint32_t cst[32 * 64];  // 19 contexts used; sized for any 5-bit context code
    __shared__ uint32_t tab[64];
    const int b = blockIdx.x * 64 + lid();
    lane_mq(b < nb ? b : -1, jobs, tabs, pool, soff16, pend, results, snaps, out, cst, tab);
}

// ---- hybrid Tier-1 coder: long streams on the scalar coder, the rest on lanes --
//
// The lane coder's time is its longest stream at one vector step per symbol,
// the scalar coder's is its block's chain at one scalar sequence per
// symbol (about three times faster per symbol) but all its blocks share the
// CUs' scalar units.  When a batch's streams are long but few, the longest
// go to the scalar coder and the rest to lanes, in one launch, so both units
// work.  The formation is t1_form_kernel's for every block.

// Scalar coder over a formed stream (wave per block): the pass regions are
// whole words, read 64 words (256 symbols) per load with the next load in
// flight; markers and padding code as no-ops (their context lanes hold a
// zero entry).
__device__ __forceinline__ void stream_mq(int b, const BlockJob* __restrict__ jobs, const uint32_t* __restrict__ tabs,
                                          const uint8_t* __restrict__ pool, const uint32_t* __restrict__ soff16,
                                          const uint32_t* __restrict__ pend, BlockResult* __restrict__ results,
                                          uint32_t* __restrict__ snaps, uint8_t* __restrict__ out) {
    const uint32_t lane = lid();
    const BlockJob J = jobs[b];
    const uint32_t np = results[b].npasses;
    Tables t;
    load_tables(t, tabs, J.band);
    MqE m;
    mqe_init(m, t.mq, out + J.out_off, J.out_cap);
    // initial states (mq/mod.rs:195-200): UNI 46, AGG 3, ZC 4, others 0; MPS 0
    m.cx = (uint32_t)__shfl((int)t.mq, lane == CTX_UNI ? 46 : lane == CTX_AGG ? 3 : lane == CTX_ZC ? 4 : 0, 64);
    if (lane == kPad || lane == kMarker) m.cx = 0;
    const uint32_t* words = reinterpret_cast<const uint32_t*>(pool + ((uint64_t)soff16[b] << 4));
    const uint32_t* pe = pend + (size_t)b * kMaxPasses;
    const uint32_t pe0 = lane < np ? pe[lane] >> 2 : 0u;  // pass ends in words
    const uint32_t pe1 = lane + 64u < np && lane + 64u < kMaxPasses ? pe[lane + 64] >> 2 : 0u;
    const uint32_t total_w = np <= 64u ? rl(pe0, np - 1u) : rl(pe1, np - 65u);
    uint32_t k = 0, kend = rl(pe0, 0);
    uint32_t snap0 = 0, snap1 = 0;
    uint32_t nxt = lane < total_w ? words[lane] : 0u;
    for (uint32_t base = 0; base < total_w; base += 64u) {
        const uint32_t cur = nxt;
        nxt = base + 64u + lane < total_w ? words[base + 64u + lane] : 0u;
        const uint32_t cnt = min(64u, total_w - base);
        for (uint32_t q = 0; q < cnt;) {
            // words up to the end of pass k or of the chunk: groups of four,
            // then single words
            const uint32_t lim = min(cnt, kend - base);
            for (; q + 4u <= lim; q += 4u) {
                mqe_group16(m, rl(cur, q), rl(cur, q + 1u), rl(cur, q + 2u), rl(cur, q + 3u));
                mqe_spill(m);
            }
            for (; q < lim; q++) {
                const uint32_t w = rl(cur, q);
                mqe_symbol(m, w & 31u, w << 24);
                mqe_symbol(m, (w >> 8) & 31u, w << 16);
                mqe_symbol(m, (w >> 16) & 31u, w << 8);
                mqe_symbol(m, (w >> 24) & 31u, w);
                mqe_spill(m);
            }
            if (base + q == kend) {  // end of pass k
                if (k < 64u) snap0 = wlane(snap0, m.pos, k);
                else snap1 = wlane(snap1, m.pos, k - 64u);
                k++;
                kend = k < np ? (k < 64u ? rl(pe0, k) : rl(pe1, k - 64u)) : 0xffffffffu;
            }
        }
    }
    const uint32_t total = mqe_flush(m);
    uint32_t* sp = snaps + (size_t)b * kMaxPasses;
    if (lane < np) sp[lane] = snap0;
    if (lane + 64u < np && lane + 64u < kMaxPasses) sp[lane + 64] = snap1;
    if (lane == 0) {
        results[b].nbytes = total;
        results[b].overflow = (uint8_t)(m.overflow ? 1 : 0);
    }
}

// Split the formed blocks: streams longer than frac x the longest go to the
// scalar coder.  lists = {n_long, n_short, long[nb], short[nb]}; blocks with
// nothing to code or whose stream did not fit are in neither.  One
// workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void t1_split_kernel(int nb, const BlockResult* __restrict__ results,
                                                        const uint32_t* __restrict__ pend, float frac,
                                                        uint32_t* __restrict__ lists) {
    __shared__ uint32_t smax, nlong, nshort;
    if (threadIdx.x == 0) {
        smax = 0;
        nlong = 0;
        nshort = 0;
    }
    __syncthreads();
    uint32_t mx = 0;
    for (int b = threadIdx.x; b < nb; b += 1024) {
        const BlockResult r = results[b];
        if (r.overflow == 0 && r.npasses) mx = max(mx, pend[(size_t)b * kMaxPasses + r.npasses - 1]);
    }
    atomicMax(&smax, mx);
    __syncthreads();
    const uint32_t thr = (uint32_t)(frac * (float)smax);
    for (int b = threadIdx.x; b < nb; b += 1024) {
        const BlockResult r = results[b];
        if (r.overflow != 0 || r.npasses == 0) continue;
        const uint32_t len = pend[(size_t)b * kMaxPasses + r.npasses - 1];
        if (len > thr) lists[2 + atomicAdd(&nlong, 1u)] = (uint32_t)b;
        else lists[2 + nb + atomicAdd(&nshort, 1u)] = (uint32_t)b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        lists[0] = nlong;
        lists[1] = nshort;
    }
}

// Workgroups [0, ceil(nb / 64)): one lane-coder wave per 64 short streams
// (if there are that many); then [.., + nb): the scalar coder on long
// stream i (if there is one).  The grid is sized for any split, and
// workgroups start in index order, so the lane waves and the long streams
// come first and the surplus workgroups, which exit at once, last.
__global__ __launch_bounds__(64) void t1_hybrid_kernel(int nb, const BlockJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ tabs, const uint8_t* __restrict__ pool,
                                                       const uint32_t* __restrict__ soff16,
                                                       const uint32_t* __restrict__ pend, BlockResult* __restrict__ results,
                                                       uint32_t* __restrict__ snaps, uint8_t* __restrict__ out,
                                                       const uint32_t* __restrict__ lists) {
    __shared__ uint32_t cst[32 * 64];
    __shared__ uint32_t tab[64];
    const uint32_t nlane = ((uint32_t)nb + 63u) / 64u;
    if (blockIdx.x < nlane) {
        const uint32_t g0 = blockIdx.x * 64u, nshort = lists[1];
        if (g0 >= nshort) return;
        const uint32_t j = g0 + lid();
        lane_mq(j < nshort ? (int)lists[2 + nb + j] : -1, jobs, tabs, pool, soff16, pend, results, snaps, out, cst, tab);
    } else {
        const uint32_t i = blockIdx.x - nlane;
        if (i >= lists[0]) return;
        stream_mq((int)lists[2 + i], jobs, tabs, pool, soff16, pend, results, snaps, out);
    }
}

// ---- decoder significance propagation and cleanup passes ----------------------
// At the start of a stripe, lane x gets X: the significance of columns
// x - 1, x and x + 1 interleaved, bits 3k, 3k + 1, 3k + 2 for window row k
// (row 4s - 1 + k), and Y: their signs in the same places.  Row ci's
// neighbourhood is then (X >> 3ci) & 0x1ff, bits NW N NE W C E SW S SE: the
// ZC table index of tier1/consts.rs:1-27 with the own bit (C) in the
// middle; its N W E S significance (bits 1 3 5 7) and the matching sign bits
// of Y shifted down by one form the index of the second sign table.

// spr lane v = v (six bits) with bit k moved to bit 3k.
__device__ __forceinline__ uint32_t spread_lut() {
    const uint32_t v = lid();
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6) | ((v & 16u) << 8) | ((v & 32u) << 10);
}
// Interleaved windows of one stripe: lane x gets column x's word.
__device__ __forceinline__ uint32_t interleave3(uint32_t spr, uint32_t own6) {
    const uint32_t sp = (uint32_t)__shfl((int)spr, (int)own6, 64);
    return from_west(sp) | (sp << 1) | (from_east(sp) << 2);
}
__device__ __forceinline__ uint32_t sc2_entry(uint32_t sc2, uint32_t si) { return (rl(sc2, si >> 2) >> ((si & 3u) * 8u)) & 255u; }

// Significance propagation and cleanup scans.  The scan of a stripe keeps,
// in two scalar words, the neighbourhood of the column it is in: W holds
// the significance of the west, own and east columns for the six window
// rows (bit 3k + j, window row k = stripe row k - 1, j = 0 west, 1 own,
// 2 east) and S their signs in the same places.  Row ci's neighbourhood is
// then (W >> 3ci) & 0x1ff (the ZC table index, NW N NE W C E SW S SE), so
// the contexts are looked up when a row is coded, from words that already
// hold this pass's decisions: a row that becomes significant sets one bit
// in each word, and the next column's words take the west and own columns
// of the current ones (W >> 1) when it follows a column with new
// significance.  Nothing else is updated on the significant path.
constexpr uint32_t kWinWestOwn = 0x1B6DBu;  // bits 3k and 3k + 1, k = 0..5

// ZC context of row ci from the window word (the row's neighbourhood in
// bits 3ci..3ci+8).
__device__ __forceinline__ uint32_t win_zc(const Tables& t, uint32_t nb) {
    return (rl(t.zc, (nb >> 3) & 63u) >> ((nb & 7u) * 4u)) & 15u;
}
// Sign decision of a row that became significant (decode/t1.rs sign coding
// with the context of tier1/helpers.rs:53-89): nb and sb are the window
// words shifted to the row; returns the sign.
__device__ __forceinline__ uint32_t win_sign(Coder<true>& cd, const Tables& t, uint32_t sc2, uint32_t nb,
                                             uint32_t sb) {
    const uint32_t si = (nb & 0xAAu) | ((sb & 0xAAu) >> 1);
    const uint32_t e = sc2_entry(sc2, si);
    return cd.code(t, e & 15u, 0) ^ (e >> 4);
}
// Scan registers start from SGPR constants written by one asm: constants
// the compiler materialises with v_mov could reach the MQ asm operands
// through the loop phis as VGPRs.
struct WinScan {
    uint32_t W, Sx, fx, pnew;
};
__device__ __forceinline__ WinScan win_init() {
    uint32_t z, none;
    asm volatile("s_mov_b32 %0, 0\n\ts_mov_b32 %1, -1" : "=s"(z), "=s"(none));
    return WinScan{z, z, none, z};
}
// Column x leaves the 64-bit candidate mask (one instruction).
__device__ __forceinline__ void drop_column(uint64_t& m, uint32_t x) {
    asm volatile("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(x));
}
// The column after x joins the scan (if the block has one).  cols is the
// block's column mask without column 0: s_bitset1_b64 takes the bit index
// mod 64, so for x = 63 it sets bit 0, which the mask then clears.
__device__ __forceinline__ void next_column(uint64_t& m, uint32_t x, uint64_t cols) {
    asm volatile("s_bitset1_b64 %0, %1\n\ts_and_b64 %0, %0, %2" : "+s"(m) : "s"(x + 1u), "s"(cols) : "scc");
}
// Column x's words: the significance word from the stripe-start words X,
// plus, when the previous column (x - 1) had new significance, the west
// and own columns of that column's word shifted into place (a real branch:
// the empty asm keeps the compiler from turning it into selects on every
// column).  The column's signs are the stripe-start word Y plus Sx (those
// merged bits and the signs decoded in the column); Y is read only when a
// row turns significant.
// Row ci (constant after unrolling) of column x turned significant: decode
// its sign and record both in the column words.
__device__ __forceinline__ uint32_t win_newsig(Coder<true>& cd, const Tables& t, uint32_t sc2, WinScan& z, uint32_t Y,
                                               uint32_t x, uint32_t ci) {
    const uint32_t sh = ci * 3u;
    const uint32_t sg = win_sign(cd, t, sc2, z.W >> sh, (rl(Y, x) | z.Sx) >> sh);
    z.W |= 1u << (sh + 4u);
    z.Sx |= sg << (sh + 4u);
    z.fx = x + 1u;
    return sg;
}

// Significance propagation pass (t1.rs:805-905, decode/t1.rs:211-266).
// Coded: insignificant rows with a significant neighbour, in stripe-column
// order.  Lane-parallel at the start of each stripe, every column gets one
// word: those rows (bits 0..3) and the other insignificant rows (bits
// 4..7), which become coded when a neighbour becomes significant during
// the scan: the row below a new significance, and, in the next column, the
// rows next to it.
__device__ void sp_pass_dec(Coder<true>& cd, const Tables& t, uint32_t sc2, uint32_t spr, Col& st, uint64_t& plane,
                            int w, int h) {
    const int lane = lid();
    const int nstripes = (h + 3) >> 2;
    const uint64_t cols = ballot(lane < w) & ~1ull;  // the columns another column can add
    for (int s = 0; s < nstripes; s++) {
        const int nrows = __builtin_amdgcn_readfirstlane(min(4, h - 4 * s));
        const uint32_t rmask = (1u << nrows) - 1u;
        const uint32_t sig6 = win6(st.sig, s), sgn6 = win6(st.sgn, s);
        const uint32_t X = interleave3(spr, sig6), Y = interleave3(spr, sgn6);
        const uint32_t o6 = from_west(sig6) | sig6 | from_east(sig6);
        const uint32_t insig4 = (~(sig6 >> 1)) & rmask;
        const uint32_t cand = insig4 & (o6 | (o6 >> 1) | (o6 >> 2));
        const uint32_t PL = cand | ((insig4 & ~cand) << 4);
        uint64_t M = ballot(lane < w && cand != 0);
        uint32_t upd = 0;
        WinScan z = win_init();
        while (M) {
            const uint32_t x = (uint32_t)__builtin_ctzll(M);
            drop_column(M, x);
            cd.column();
            const uint32_t P = rl(PL, x);
            const uint32_t ix = P >> 4;
            // acc bits 0..3: new significance, 4..7: its signs, 8..11: rows
            // coded beyond the word's
            const uint32_t Wp = z.W, Sp = z.Sx;
            z.W = rl(X, x);
            z.Sx = 0;
            uint32_t visit = P & 15u, acc = 0;
            if (__builtin_expect(z.fx == x, 0)) {
                asm volatile("");
                z.W |= (Wp >> 1) & kWinWestOwn;
                z.Sx = ((rl(Y, x - 1u) | Sp) >> 1) & kWinWestOwn;
                // rows next to the west column's new significance
                const uint32_t nw = z.pnew;
                const uint32_t vadd = (nw | (nw << 1) | (nw >> 1)) & ix;
                visit |= vadd;
                acc = vadd << 8;
            }
#pragma unroll
            for (uint32_t ci = 0; ci < 4u; ci++) {
                if (visit & (1u << ci)) {
                    if (__builtin_expect(cd.code(t, win_zc(t, z.W >> (3u * ci)), 0), 0)) {
                        const uint32_t sg = win_newsig(cd, t, sc2, z, Y, x, ci);
                        const uint32_t below = (2u << ci) & ix;  // its south neighbour
                        visit |= below;
                        acc |= (1u << ci) | (sg << (ci + 4u)) | (below << 8);
                        next_column(M, x, cols);
                    }
                }
            }
            upd = wlane(upd, acc, x);
            z.pnew = acc & 15u;
        }
        st.sig |= (uint64_t)(upd & 15u) << (4 * s);
        st.sgn |= (uint64_t)((upd >> 4) & 15u) << (4 * s);
        st.vis |= (uint64_t)(((lane < w ? cand : 0u) | (upd >> 8)) & 15u) << (4 * s);
        plane |= (uint64_t)(upd & 15u) << (4 * s);
    }
}

// Cleanup pass with run-length mode on full clean stripes (t1.rs:977-1197,
// decode/t1.rs:317-446).  Every row neither significant nor visited is
// coded.  The column word holds those rows (bits 0..3) and whether the
// column starts in run mode (bit 4: four rows, nothing significant around
// them, nothing visited); a column after one with new significance loses
// run mode.
__device__ void cl_pass_dec(Coder<true>& cd, const Tables& t, uint32_t sc2, uint32_t spr, Col& st, uint64_t& plane,
                            int w, int h) {
    const int lane = lid();
    const int nstripes = (h + 3) >> 2;
    for (int s = 0; s < nstripes; s++) {
        const int nrows = __builtin_amdgcn_readfirstlane(min(4, h - 4 * s));
        const uint32_t rmask = (1u << nrows) - 1u;
        const uint32_t sig6 = win6(st.sig, s), sgn6 = win6(st.sgn, s);
        const uint32_t vis4 = (uint32_t)(st.vis >> (4 * s)) & 15u;
        const uint32_t X = interleave3(spr, sig6), Y = interleave3(spr, sgn6);
        // static part of stripe_is_clean (flags.rs:202-215)
        const uint32_t sclean = (nrows == 4 && X == 0 && vis4 == 0) ? 1u : 0u;
        const uint32_t todo = ~((sig6 >> 1) | vis4) & rmask;
        const uint32_t PL = todo | (sclean << 4);
        uint64_t M = ballot(lane < w && todo != 0);
        uint32_t upd = 0;
        WinScan z = win_init();
        while (M) {
            const uint32_t x = (uint32_t)__builtin_ctzll(M);
            drop_column(M, x);
            cd.column();
            const uint32_t P = rl(PL, x);
            const uint32_t Wp = z.W, Sp = z.Sx;
            z.W = rl(X, x);
            z.Sx = 0;
            uint32_t visit = P & 15u;
            uint32_t acc = 0;
            if (__builtin_expect(z.fx == x, 0)) {
                // the west column had new significance (no run mode then)
                asm volatile("");
                z.W |= (Wp >> 1) & kWinWestOwn;
                z.Sx = ((rl(Y, x - 1u) | Sp) >> 1) & kWinWestOwn;
            } else if (P & 16u) {  // run mode
                if (__builtin_expect(!cd.code(t, CTX_AGG, 0), 1)) {
                    visit = 0;
                } else {
                    const uint32_t hi = cd.code(t, CTX_UNI, 0);
                    const uint32_t lo = cd.code(t, CTX_UNI, 0);
                    const uint32_t ci = (hi << 1) | lo;
                    const uint32_t sh = ci * 3u;
                    // a clean stripe has no significant neighbour: sign
                    // index 0 (its contexts are 0 as well)
                    const uint32_t sg = win_sign(cd, t, sc2, z.W >> sh, z.Sx >> sh);
                    z.W |= 1u << (sh + 4u);
                    z.Sx |= sg << (sh + 4u);
                    acc = (1u << ci) | (sg << (ci + 4u));
                    visit &= ~((2u << ci) - 1u);  // rows after it
                    z.fx = x + 1u;
                }
            }
#pragma unroll
            for (uint32_t ci = 0; ci < 4u; ci++) {
                if (visit & (1u << ci)) {
                    if (__builtin_expect(cd.code(t, win_zc(t, z.W >> (3u * ci)), 0), 0)) {
                        const uint32_t sg = win_newsig(cd, t, sc2, z, Y, x, ci);
                        acc |= (1u << ci) | (sg << (ci + 4u));
                    }
                }
            }
            upd = wlane(upd, acc, x);
        }
        st.sig |= (uint64_t)(upd & 15u) << (4 * s);
        st.sgn |= (uint64_t)((upd >> 4) & 15u) << (4 * s);
        plane |= (uint64_t)(upd & 15u) << (4 * s);
    }
}

// Magnitude refinement pass of the decoder (t1.rs:914-965, decode/t1.rs:268-315).
// Members and contexts do not depend on the pass's own decisions, so the
// pass is formed lane-parallel into LDS in scan order (as the encoder forms
// it), decoded by one scalar loop that appends each decision to a bit
// string (one bit-field extract into M0 and one shift-add per decision),
// and the bits go back to their columns lane-parallel.  symw: 1024 words
// (at most 4096 members), bitw: 128 words.
__device__ __forceinline__ void mr_pass_dec(Coder<true>& cd, const Tables& t, Col& st, uint64_t& plane, int w, int h,
                                            uint32_t* symw, uint32_t* bitw) {
    const uint32_t lane = lid();
    const bool valid = (int)lane < w;
    const int nstripes = (h + 3) >> 2;
    uint8_t* symb = reinterpret_cast<uint8_t*>(symw);
    auto members = [&](int s, uint32_t& sig6, uint32_t& ref4) {
        const int nrows = __builtin_amdgcn_readfirstlane(min(4, h - 4 * s));
        const uint32_t rmask = valid ? (1u << nrows) - 1u : 0u;
        sig6 = win6(st.sig, s);
        ref4 = (uint32_t)(st.ref >> (4 * s)) & 15u;
        const uint32_t vis4 = (uint32_t)(st.vis >> (4 * s)) & 15u;
        return (sig6 >> 1) & ~vis4 & rmask;
    };
    uint32_t total = 0;
    for (int s = 0; s < nstripes; s++) {
        uint32_t sig6, ref4;
        const uint32_t mem = members(s, sig6, ref4);
        const uint32_t w6 = from_west(sig6), e6 = from_east(sig6);
        uint32_t tot;
        uint32_t pos = total + lane_offsets((uint32_t)__builtin_popcount(mem), tot);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if ((mem >> r) & 1u) {
                const uint32_t ctx = ((ref4 >> r) & 1u) ? CTX_MAG + 2u : (nbmask(w6, sig6, e6, r + 1) ? CTX_MAG + 1u : CTX_MAG);
                symb[pos++] = (uint8_t)ctx;
            }
        }
        total += tot;
    }
    __syncthreads();
    // the scan slides the segment window at least every 16 decisions
    uint32_t acc = 0;
    for (uint32_t base = 0; base < total; base += 256u) {
        const uint32_t sv = symw[(base >> 2) + lane];
        const uint32_t cnt = min(256u, total - base);
        uint32_t q = 0;
        for (; q + 16u <= cnt; q += 16u) {
            cd.column();
#pragma unroll
            for (uint32_t g = 0; g < 4u; g++) {
                const uint32_t wd = rl(sv, (q >> 2) + g);
                acc = acc * 2u + cd.code(t, wd & 31u, 0);
                acc = acc * 2u + cd.code(t, (wd >> 8) & 31u, 0);
                acc = acc * 2u + cd.code(t, (wd >> 16) & 31u, 0);
                acc = acc * 2u + cd.code(t, (wd >> 24) & 31u, 0);
            }
            const uint32_t done = base + q + 16u;
            if ((done & 31u) == 0u) {
                if (lane == 0) bitw[(done >> 5) - 1u] = acc;
                acc = 0;
            }
        }
        for (; q + 4u <= cnt; q += 4u) {
            cd.column();
            const uint32_t wd = rl(sv, q >> 2);
            acc = acc * 2u + cd.code(t, wd & 31u, 0);
            acc = acc * 2u + cd.code(t, (wd >> 8) & 31u, 0);
            acc = acc * 2u + cd.code(t, (wd >> 16) & 31u, 0);
            acc = acc * 2u + cd.code(t, (wd >> 24) & 31u, 0);
            const uint32_t done = base + q + 4u;
            if ((done & 31u) == 0u) {
                if (lane == 0) bitw[(done >> 5) - 1u] = acc;
                acc = 0;
            }
        }
        if (q < cnt) {
            cd.column();
            uint32_t wd = rl(sv, q >> 2);
            for (; q < cnt; q++, wd >>= 8) acc = acc * 2u + cd.code(t, wd & 31u, 0);
        }
    }
    if (total & 31u) {  // the last, partial word, first decision in bit 31
        acc <<= 32u - (total & 31u);
        if (lane == 0) bitw[total >> 5] = acc;
    }
    __syncthreads();
    total = 0;
    for (int s = 0; s < nstripes; s++) {
        uint32_t sig6, ref4;
        const uint32_t mem = members(s, sig6, ref4);
        uint32_t tot;
        uint32_t pos = total + lane_offsets((uint32_t)__builtin_popcount(mem), tot);
        uint32_t ones = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if ((mem >> r) & 1u) {
                ones |= ((bitw[pos >> 5] >> (31u - (pos & 31u))) & 1u) << r;
                pos++;
            }
        }
        plane |= (uint64_t)ones << (4 * s);
        total += tot;
    }
    st.ref |= st.sig & ~st.vis;
    __syncthreads();  // symw and bitw are reused by the next pass
}

// Tier-1 decoder: one wavefront per code-block.  Decodes `passes` coding
// passes and writes the reconstructed coefficients into the plane:
// dequantised f32 (sign * (|q| + 0.5)) * step for 9/7 (scalar.rs:233-240),
// i32 for 5/3.  Blocks that are not included are written as zeros.
__global__ __launch_bounds__(64) __attribute__((amdgpu_num_sgpr(kStateSgprCap))) void t1_decode_kernel(const DecodeJob* __restrict__ jobs, const uint32_t* __restrict__ tabs,
                                                       const uint8_t* __restrict__ payload) {
    __shared__ uint32_t symw[1024];
    __shared__ uint32_t bitw[128];
    const int lane = lid();
    const DecodeJob J = jobs[blockIdx.x];
    const int w = __builtin_amdgcn_readfirstlane(J.w), h = __builtin_amdgcn_readfirstlane(J.h);
    const bool lossy = J.step > 0.0f;
    uint32_t mag[64];
#pragma unroll
    for (int y = 0; y < 64; y++) mag[y] = 0;
    uint64_t sgnmask = 0;
    set_priority(J.pad);  // the host ranks blocks by segment length
    if (J.zbp >= 0 && J.passes > 0) {
        Tables t;
        load_tables(t, tabs, J.band);
        const uint32_t sc2 = tabs[384 + lane];
        const uint32_t spr = spread_lut();
        Coder<true> cd;
        cd.m.data = payload + J.data_off;
        cd.m.len = J.data_len;
        cd.m.tab = dec_entry(t.mq);
        // initial states (mq/mod.rs:195-200): UNI 46, AGG 3, ZC 4, others 0; MPS 0
        states_init((uint32_t)__shfl((int)cd.m.tab, lane == CTX_UNI ? 46 : lane == CTX_AGG ? 3 : lane == CTX_ZC ? 4 : 0, 64));
        dec_start(cd.m);
        Col st = {0, 0, 0, 0};
        const int mbp = (int)J.mb - (int)J.zbp;
        int remaining = J.passes;
        for (int p = mbp - 1; p >= 0 && remaining > 0; p--) {
            uint64_t plane = 0;
            if (p != mbp - 1) {
                sp_pass_dec(cd, t, sc2, spr, st, plane, w, h);
                if (--remaining == 0) goto plane_done;
                mr_pass_dec(cd, t, st, plane, w, h, symw, bitw);
                if (--remaining == 0) goto plane_done;
            }
            cl_pass_dec(cd, t, sc2, spr, st, plane, w, h);
            remaining--;
        plane_done:
#pragma unroll
            for (int y = 0; y < 64; y++) mag[y] |= (uint32_t)((plane >> y) & 1ull) << p;
            st.vis = 0;
        }
        sgnmask = st.sgn;
    }
#pragma unroll
    for (int y = 0; y < 64; y++) {
        if (y >= h || lane >= w) continue;
        uint32_t m = mag[y];
        bool negv = (sgnmask >> y) & 1ull;
        size_t idx = (size_t)(J.y0 + y) * J.stride + J.x0 + lane;
        if (lossy) {
            float v = 0.0f;
            if (m) v = __fmul_rn(negv ? -__fadd_rn((float)m, 0.5f) : __fadd_rn((float)m, 0.5f), J.step);
            ((float*)J.plane)[idx] = v;
        } else {
            ((int32_t*)J.plane)[idx] = negv ? -(int32_t)m : (int32_t)m;
        }
    }
}

size_t t1_sched_words(int n) { return pre_base((uint32_t)n) + kPre * (size_t)kPreSlotWords; }

namespace {
// SIMDs of the current device and encode waves resident on it at once
// (zeros if the queries fail), queried once per device.
struct EncodeShape {
    uint32_t simds = 0, resident = 0;
};
EncodeShape encode_shape() {
    static std::mutex mu;
    static EncodeShape cache[64];
    static bool known[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return EncodeShape{};
    std::lock_guard<std::mutex> g(mu);
    if (!known[dev]) {
        int cus = 0, per_cu = 0;
        EncodeShape e;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, t1_encode_kernel, 64, kSymCap) == hipSuccess &&
            cus > 0 && per_cu > 0) {
            e.simds = 4u * (uint32_t)cus;
            e.resident = (uint32_t)cus * (uint32_t)per_cu;
        }
        cache[dev] = e;
        known[dev] = true;
    }
    return cache[dev];
}
}  // namespace

void launch_t1_encode(const BlockJob* jobs, int n, const uint32_t* tabs, uint8_t* out, BlockResult* res, uint32_t* snaps,
                      uint16_t* newsigs, uint32_t* sched, hipStream_t st) {
    if (n <= 0) return;
    // Heavy takers: one per SIMD, none on SIMDs that run three waves when
    // every wave is resident (all SIMDs hold two, 3 * simds - n hold only
    // two).  The rule stays off where fewer than kPre SIMDs would take heavy
    // blocks, so the pre-formed positions are always heavy takers' ones.
    const EncodeShape e = encode_shape();
    const uint32_t un = (uint32_t)n;
    const bool three = e.simds && un > 2u * e.simds && un <= e.resident && un + kPre <= 3u * e.simds;
    const uint32_t split = !e.simds ? un : three ? 3u * e.simds - un : std::min(un, e.simds);
    t1_cost_kernel<<<n, 64, 0, st>>>(jobs, sched);
    t1_order_kernel<<<1, 1024, 0, st>>>(n, split, sched);
    t1_preform_kernel<<<kPre * kPrePlanes, 64, 0, st>>>(jobs, tabs, un, sched);
    t1_encode_kernel<<<n, 64, kSymCap, st>>>(jobs, tabs, out, res, snaps, newsigs, sched, sched + pre_base(un),
                                             three ? 1u : 0u);
}
void launch_t1_form(const BlockJob* jobs, int n, const uint32_t* tabs, BlockResult* res, uint32_t* pend, uint16_t* newsigs,
                    uint32_t* soff16, uint8_t* pool, unsigned long long* pool_used, unsigned long long pool_cap,
                    hipStream_t st) {
    if (n > 0) t1_form_kernel<<<n, 64, 0, st>>>(jobs, tabs, res, pend, newsigs, soff16, pool, pool_used, pool_cap);
}
void launch_t1_mq(const BlockJob* jobs, int n, const uint32_t* tabs, const uint8_t* pool, const uint32_t* soff16,
                  const uint32_t* pend, BlockResult* res, uint32_t* snaps, uint8_t* out, hipStream_t st) {
    if (n > 0) t1_mq_kernel<<<(n + 63) / 64, 64, 0, st>>>(n, jobs, tabs, pool, soff16, pend, res, snaps, out);
}
void launch_t1_hybrid(const BlockJob* jobs, int n, const uint32_t* tabs, const uint8_t* pool, const uint32_t* soff16,
                      const uint32_t* pend, BlockResult* res, uint32_t* snaps, uint8_t* out, uint32_t* lists, float frac,
                      hipStream_t st) {
    if (n <= 0) return;
    t1_split_kernel<<<1, 1024, 0, st>>>(n, res, pend, frac, lists);
    t1_hybrid_kernel<<<n + (n + 63) / 64, 64, 0, st>>>(n, jobs, tabs, pool, soff16, pend, res, snaps, out, lists);
}
void launch_t1_count(const BlockJob* jobs, int n, const uint32_t* tabs, uint32_t* counts, hipStream_t st) {
    if (n > 0) t1_count_kernel<<<n, 64, 0, st>>>(jobs, tabs, counts);
}
__global__ __launch_bounds__(64) void zero_u64_kernel(unsigned long long* __restrict__ p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st) {
    if (n > 0) zero_u64_kernel<<<(n + 63) / 64, 64, 0, st>>>(p, n);
}
uint64_t t1_stream_bound(uint32_t mbp) { return stream_bound(mbp); }
void launch_t1_decode(const DecodeJob* jobs, int n, const uint32_t* tabs, const uint8_t* payload, hipStream_t st) {
    if (n > 0) t1_decode_kernel<<<n, 64, 0, st>>>(jobs, tabs, payload);
}

}  // namespace j2l
