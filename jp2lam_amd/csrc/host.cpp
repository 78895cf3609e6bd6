// Host-side codec stages: planning, Tier-2, markers, JP2 boxes, parsing.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace j2l {

namespace {

int gain_of(int band) { return band == LL ? 0 : (band == HH ? 2 : 1); }

// 9/7 synthesis norms per orientation and level (dwt/norms.rs:24-37).  The
// step sizes are derived from these, so they must be the same numbers.
constexpr double kNorm97[4][10] = {
    {1.000, 1.965, 4.177, 8.403, 16.90, 33.84, 67.69, 135.3, 270.6, 540.9},
    {2.022, 3.989, 8.355, 17.04, 34.27, 68.63, 137.3, 274.6, 549.0, 0.0},
    {2.022, 3.989, 8.355, 17.04, 34.27, 68.63, 137.3, 274.6, 549.0, 0.0},
    {2.080, 3.865, 8.307, 17.18, 34.71, 69.59, 139.3, 278.6, 557.2, 0.0},
};

int ilog2(int32_t v) { return 31 - __builtin_clz((uint32_t)std::max<int32_t>(v, 1)); }

// dwt/norms.rs:84-90
void pack_step(int32_t stepsize, uint32_t numbps, uint8_t& e, uint16_t& m) {
    int lg = ilog2(stepsize);
    int p = lg - 13, n = 11 - lg;
    int32_t mant = (n < 0 ? (stepsize >> -n) : (stepsize << n)) & 0x7ff;
    e = (uint8_t)std::max<int>((int)numbps - p, 0);
    m = (uint16_t)mant;
}

// plan/derive.rs:74-86
double step_scale(int q) {
    if (q == 0) return 8.0;
    if (q <= 50) return std::pow(2.0, (50.0 - q) / 20.0);
    if (q <= 90) return std::pow(2.0, -((q - 50.0) / 20.0));
    double t = (q - 90.0) / 9.0;
    return 0.25 * std::pow(2.0, -3.0 * t);
}

// plan/derive.rs:94-107
void rescale(SubbandQuant& sq, double scale) {
    double base = 1.0 + sq.mantissa / 2048.0;
    double s = std::max(base * scale, 1.0 / 8192.0);
    int k = (int)std::floor(std::log2(s));
    double nb = s / std::ldexp(1.0, k);
    sq.exponent = (uint8_t)std::max<int>((int)sq.exponent - k, 0);
    sq.mantissa = (uint16_t)std::clamp(std::round((nb - 1.0) * 2048.0), 0.0, 2047.0);
}

}  // namespace

double norm97(uint32_t level, int band) {
    uint32_t cap = band == LL ? 9 : 8;
    return kNorm97[band][std::min(level, cap)];
}

uint32_t max_decompositions(uint32_t w, uint32_t h) {
    uint32_t m = std::min(w, h);
    if (m <= 1) return 0;
    if (m < 32) return 31 - __builtin_clz(m);
    return std::min<uint32_t>(1 + (31 - __builtin_clz(m / 32)), 6);
}

Plan build_plan(uint32_t w, uint32_t h, int colorspace, int quality, int format) {
    Plan p;
    p.width = w;
    p.height = h;
    p.domain = colorspace == GRAY ? GRAY : SRGB;
    p.ncomp = p.domain == GRAY ? 1 : 3;
    p.quality = quality;
    p.format = format;
    p.lossless = quality >= 100;
    p.levels = (int)std::min<uint32_t>(max_decompositions(w, h), p.domain == GRAY ? 5 : 6);
    p.use_mct = p.domain == SRGB;
    p.cbw = p.cbh = quality < 30 ? 32 : 64;
    p.guard = 2;
    const uint32_t prec = 8;
    const int nres = p.levels + 1;
    for (int r = 0; r <= p.levels; r++) {
        for (int b = r ? HL : LL; b <= (r ? HH : LL); b++) {
            SubbandQuant sq{(uint8_t)r, (uint8_t)b, 0, 0};
            if (p.lossless) {
                sq.exponent = (uint8_t)(prec + gain_of(b));
            } else {
                // dwt/norms.rs:68-82
                int lv = std::max(std::max(nres - 1, 0) - r, 0);
                double stepsize = (double)(1u << gain_of(b)) / norm97((uint32_t)lv, b);
                pack_step((int32_t)std::floor(stepsize * 8192.0), prec + gain_of(b), sq.exponent, sq.mantissa);
            }
            p.quants.push_back(sq);
        }
    }
    if (!p.lossless) {
        double s = step_scale(quality);
        for (auto& q : p.quants) rescale(q, s);
    }
    return p;
}

float quant_step_f32(int band, uint8_t exponent, uint16_t mantissa, uint32_t precision) {
    int numbps = (int)precision + gain_of(band);
    float base = 1.0f + (float)mantissa / 2048.0f;
    float v = base * std::ldexp(1.0f, numbps - (int)exponent);
    return std::max(v, 1e-6f);
}

double quant_step_f64(int band, uint8_t exponent, uint16_t mantissa, uint32_t precision) {
    int numbps = (int)precision + gain_of(band);
    return (1.0 + (double)mantissa / 2048.0) * std::ldexp(1.0, numbps - (int)exponent);
}

double band_bias(int band, int q) {
    if (q < 50) {
        double t = q / 50.0;
        return band == LL ? 3.0 - 1.8 * t : band == HH ? 0.20 + 0.62 * t : 1.0;
    }
    return band == LL ? 1.20 : band == HH ? 0.82 : 1.00;
}

double component_weight(const Plan& p, int comp) {
    if (!p.use_mct || p.domain != SRGB || comp == 0) return 1.0;
    double q = (double)std::min(p.quality, 99);
    return 1.20 + 1.80 * std::pow(1.0 - q / 99.0, 0.70);
}

double quality_to_lambda(int q, uint32_t pixels) {
    if (q >= 100) return 0.0;
    if (q == 0) return 1.7976931348623157e308;
    double t = (q - 1.0) / 98.0;
    double base = std::max(std::pow(10.0, 5.68 - 2.618 * std::pow(t, 0.85)), 1e-3);
    double tail = 1.0;
    if (q >= 90) {
        double tt = (99.0 - std::min(q, 99)) / 9.0;
        tail = tt * tt;
    }
    return base * std::pow((double)pixels / 12000000.0, 0.35) * tail;
}

int subband_mb(const Plan& p, int comp, int res, int band) {
    int v;
    if (p.lossless) {
        int prec = (p.use_mct && comp > 0) ? 9 : 8;  // backend.rs:179-189
        v = (p.guard - 1) + prec + gain_of(band);
    } else {
        v = (p.guard - 1) + 8 + gain_of(band);
        for (auto& q : p.quants)
            if (q.res == res && q.band == band) v = (p.guard - 1) + q.exponent;
    }
    return std::min(v, 255);
}

void resolution_ladder(uint32_t w, uint32_t h, int levels, std::vector<uint32_t>& rw, std::vector<uint32_t>& rh) {
    rw.assign(levels + 1, 0);
    rh.assign(levels + 1, 0);
    rw[levels] = w;
    rh[levels] = h;
    for (int l = levels - 1; l >= 0; l--) {
        rw[l] = (rw[l + 1] + 1) / 2;
        rh[l] = (rh[l + 1] + 1) / 2;
    }
}

Rect band_rect(uint32_t w, uint32_t h, int levels, int res, int band) {
    std::vector<uint32_t> rw, rh;
    resolution_ladder(w, h, levels, rw, rh);
    if (res == 0) return {0, 0, (int)rw[0], (int)rh[0]};
    int lw = (int)rw[res - 1], lh = (int)rh[res - 1], fw = (int)rw[res], fh = (int)rh[res];
    if (band == HL) return {lw, 0, fw, lh};
    if (band == LH) return {0, lh, lw, fh};
    return {lw, lh, fw, fh};
}

PlaneLayout plane_layout(uint32_t w, int levels, bool aligned) {
    PlaneLayout pl;
    std::vector<uint32_t> rw, rh;
    resolution_ladder(w, 1, levels, rw, rh);
    uint32_t end = rw[0];  // columns used by resolutions below r
    for (int r = 1; r <= levels; r++) {
        pl.hx[r] = (end + 31u) & ~31u;
        end = pl.hx[r] + (rw[r] - rw[r - 1]);
    }
    pl.stride = levels ? (end + 31u) & ~31u : w;
    if (!aligned || pl.stride > 65535u) {
        for (int r = 1; r <= levels; r++) pl.hx[r] = rw[r - 1];
        pl.stride = w;
    }
    return pl;
}

// ---- Tier-1 tables ------------------------------------------------------------
namespace {
// ISO/IEC 15444-1 Table C.2: Qe, next-MPS, next-LPS, switch.
constexpr uint16_t kQe[47] = {0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801,
                              0x3001, 0x2401, 0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401,
                              0x3001, 0x2801, 0x2401, 0x2201, 0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101,
                              0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1, 0x0221, 0x0141, 0x0111, 0x0085,
                              0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
constexpr uint8_t kNmps[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                               17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                               33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
constexpr uint8_t kNlps[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                               15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                               30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
constexpr bool kSwitch(int i) { return i == 0 || i == 6 || i == 14; }

// Table D.1 zero-coding context from (h, v, d) neighbour counts.
int zc_context(int orient, int h, int v, int d) {
    if (orient == HL) std::swap(h, v);
    if (orient == HH) {
        int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv >= 1 ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : hv == 1 ? 4 : 3;
        return hv >= 2 ? 2 : hv;
    }
    if (h == 2) return 8;
    if (h == 1) return v >= 1 ? 7 : d >= 1 ? 6 : 5;
    if (v >= 1) return 2 + v;
    return d >= 2 ? 2 : d;
}
// Table D.3 sign context and XOR bit from the sign index bits of
// tier1/consts.rs:31-38 (SGN_W=0 SIG_N=1 SGN_E=2 SIG_W=3 SGN_N=4 SIG_E=5
// SGN_S=6 SIG_S=7).
void sign_context(int i, int& ctx, int& xr) {
    auto contrib = [&](int sig_bit, int sgn_bit) { return ((i >> sig_bit) & 1) ? (((i >> sgn_bit) & 1) ? -1 : 1) : 0; };
    int hc = contrib(3, 0) + contrib(5, 2), vc = contrib(1, 4) + contrib(7, 6);
    int hh = (hc > 0) - (hc < 0), vv = (vc > 0) - (vc < 0);
    if (hh == 0) {
        ctx = vv == 0 ? 9 : 10;
        xr = vv < 0;
    } else {
        ctx = vv == 0 ? 12 : (hh == vv ? 13 : 11);
        xr = hh < 0;
    }
}
int zc_index_context(int orient, int idx) {
    int nw = idx & 1, n = (idx >> 1) & 1, ne = (idx >> 2) & 1, w = (idx >> 3) & 1;
    int e = (idx >> 5) & 1, sw = (idx >> 6) & 1, s = (idx >> 7) & 1, se = (idx >> 8) & 1;
    return zc_context(orient, w + e, n + s, nw + ne + sw + se);
}
}  // namespace

// Words 0..63: the MQ states as the coders step them.  Word i < 47 is state
// i: qe | nmps << 16 | nlps' << 22.  The LPS targets of the three switching
// states (0 -> 1, 6 -> 6, 14 -> 14) get copies at words 47..49 with bit 31
// set, and nlps' points there, so the next state of any decision is
// word[field] XOR (MPS in bit 31): the copy's bit 31 flips the MPS.
std::array<uint32_t, 448> make_t1_tables() {
    std::array<uint32_t, 448> t{};
    auto word = [](int i) {
        const int l = kNlps[i];
        const uint32_t nl = kSwitch(i) ? (l == 1 ? 47u : l == 6 ? 48u : 49u) : (uint32_t)l;
        return kQe[i] | ((uint32_t)kNmps[i] << 16) | (nl << 22);
    };
    for (int i = 0; i < 47; i++) t[i] = word(i);
    t[47] = word(1) | 0x80000000u;
    t[48] = word(6) | 0x80000000u;
    t[49] = word(14) | 0x80000000u;
    for (int o = 0; o < 4; o++)
        for (int idx = 0; idx < 512; idx++) t[64 + o * 64 + idx / 8] |= (uint32_t)zc_index_context(o, idx) << ((idx % 8) * 4);
    for (int i = 0; i < 256; i++) {
        int ctx, xr;
        sign_context(i, ctx, xr);
        t[320 + i / 4] |= (uint32_t)(ctx | (xr << 4)) << ((i % 4) * 8);
    }
    // Words 384..447: the same sign entries indexed by the decoder's layout,
    // significance of N, W, E, S in bits 1, 3, 5, 7 (as above) and their
    // signs in bits 0, 2, 4, 6.
    for (int i = 0; i < 256; i++) {
        const int old = (i & 0xAA) | (((i >> 0) & 1) << 4) | (((i >> 2) & 1) << 0) | (((i >> 4) & 1) << 2) |
                        (((i >> 6) & 1) << 6);
        int ctx, xr;
        sign_context(old, ctx, xr);
        t[384 + i / 4] |= (uint32_t)(ctx | (xr << 4)) << ((i % 4) * 8);
    }
    return t;
}

void expanded_tables(uint16_t* qe, uint8_t* mps, uint8_t* nmps, uint8_t* nlps, uint8_t* zc, uint8_t* sc, uint8_t* spb) {
    for (int s = 0; s < 94; s++) {
        int i = s >> 1, m = s & 1;
        qe[s] = kQe[i];
        mps[s] = (uint8_t)m;
        nmps[s] = (uint8_t)(2 * kNmps[i] + m);
        nlps[s] = (uint8_t)(2 * kNlps[i] + (m ^ (kSwitch(i) ? 1 : 0)));
    }
    for (int o = 0; o < 4; o++)
        for (int idx = 0; idx < 512; idx++) zc[o * 512 + idx] = (uint8_t)zc_index_context(o, idx);
    for (int i = 0; i < 256; i++) {
        int ctx, xr;
        sign_context(i, ctx, xr);
        sc[i] = (uint8_t)ctx;
        spb[i] = (uint8_t)xr;
    }
}

// ---- Tier-2 encoding ------------------------------------------------------------
namespace {

// Packet-header bit writer with 0xFF bit stuffing (ISO 15444-1 B.10.1;
// t2.rs:27-68).
class BitWriter {
public:
    explicit BitWriter(std::vector<uint8_t>& o) : out_(o) {}
    void bit(uint32_t b) {
        if (ct_ == 0) byteout();
        ct_--;
        reg_ |= b << ct_;
    }
    void bits(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; i--) bit((v >> i) & 1u);
    }
    void flush() {
        byteout();
        if (ct_ == 7) byteout();
    }

private:
    void byteout() {
        reg_ = (reg_ << 8) & 0xffffu;
        ct_ = reg_ == 0xff00u ? 7 : 8;
        out_.push_back((uint8_t)(reg_ >> 8));
    }
    std::vector<uint8_t>& out_;
    uint32_t reg_ = 0, ct_ = 8;
};

// Tag tree (B.10.2) as a quad tree stored level by level in raster order.
class TagTree {
public:
    TagTree(int w, int h, int init) {
        int lw = w, lh = h;
        for (;;) {
            lvl_w_.push_back(lw);
            lvl_off_.push_back((int)value_.size());
            value_.resize(value_.size() + (size_t)lw * lh, init);
            if (lw * lh <= 1) break;
            lw = (lw + 1) / 2;
            lh = (lh + 1) / 2;
        }
        low_.assign(value_.size(), 0);
        known_.assign(value_.size(), 0);
        parent_.assign(value_.size(), -1);
        int nl = (int)lvl_w_.size();
        int lh2 = h;
        for (int l = 0; l + 1 < nl; l++) {
            int cw = lvl_w_[l];
            for (int y = 0; y < lh2; y++)
                for (int x = 0; x < cw; x++)
                    parent_[lvl_off_[l] + y * cw + x] = lvl_off_[l + 1] + (y / 2) * lvl_w_[l + 1] + x / 2;
            lh2 = (lh2 + 1) / 2;
        }
    }
    void set(int leaf, int v) {
        for (int i = leaf; i >= 0 && value_[i] > v; i = parent_[i]) value_[i] = v;
    }
    void encode(BitWriter& bw, int leaf, int threshold) {
        int path[40], depth = 0;
        for (int n = leaf; n >= 0; n = parent_[n]) path[depth++] = n;
        int low = 0;
        while (depth > 0) {
            int n = path[--depth];
            if (low > low_[n]) low_[n] = low;
            else low = low_[n];
            while (low < threshold) {
                if (low >= value_[n]) {
                    if (!known_[n]) {
                        bw.bit(1);
                        known_[n] = 1;
                    }
                    break;
                }
                bw.bit(0);
                low++;
            }
            low_[n] = low;
        }
    }
    template <class Reader>
    bool decode(Reader& rd, int leaf, int threshold) {
        int path[40], depth = 0;
        for (int n = leaf; n >= 0; n = parent_[n]) path[depth++] = n;
        int low = 0;
        while (depth > 0) {
            int n = path[--depth];
            if (low > low_[n]) low_[n] = low;
            else low = low_[n];
            while (low < threshold && !known_[n]) {
                if (rd.bit()) {
                    value_[n] = low;
                    known_[n] = 1;
                } else {
                    low++;
                }
            }
            if (known_[n] && value_[n] > low) low = value_[n];
            low_[n] = low;
        }
        return known_[leaf] && value_[leaf] < threshold;
    }
    bool known(int leaf) const { return known_[leaf] != 0; }
    int value(int leaf) const { return value_[leaf]; }

private:
    std::vector<int> lvl_w_, lvl_off_, value_, low_, parent_;
    std::vector<uint8_t> known_;
};

uint32_t flog2(uint32_t v) { return v ? 31u - (uint32_t)__builtin_clz(v) : 0u; }

// Table B.4 codewords for the number of coding passes.
void put_passes(BitWriter& bw, uint32_t n) {
    if (n == 1) bw.bits(0, 1);
    else if (n == 2) bw.bits(2, 2);
    else if (n <= 5) bw.bits(0xc | (n - 3), 4);
    else if (n <= 36) bw.bits(0x1e0 | (n - 6), 9);
    else bw.bits(0xff80 | (std::min<uint32_t>(n, 164) - 37), 16);
}

void be16(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}
void be32(std::vector<uint8_t>& o, uint32_t v) {
    be16(o, v >> 16);
    be16(o, v & 0xffff);
}
void tag(std::vector<uint8_t>& o, const char* t) { o.insert(o.end(), t, t + 4); }

}  // namespace

void packet_header(const BandBlocks* bands, int nbands, std::vector<uint8_t>& out) {
    BitWriter bw(out);
    bw.bit(1);  // packet present (always, like the reference: t2.rs:340-344)
    for (int bi = 0; bi < nbands; bi++) {
        const BandBlocks& bd = bands[bi];
        const int n = bd.cx * bd.cy;
        if (n == 0) continue;
        TagTree incl(bd.cx, bd.cy, 999), imsb(bd.cx, bd.cy, 999);
        for (int k = 0; k < n; k++) {
            imsb.set(k, bd.blocks[k].zbp);
            if (bd.blocks[k].keep) incl.set(k, 0);
        }
        for (int k = 0; k < n; k++) {
            const CodedBlock& cb = bd.blocks[k];
            incl.encode(bw, k, 1);
            if (!cb.keep) continue;
            imsb.encode(bw, k, 999);
            put_passes(bw, cb.keep);
            // one segment per block (single termination); Lblock starts at 3
            // and takes one increment sized for the segment (B.10.7.1)
            uint32_t inc = 0;
            if (cb.seg_len) {
                uint32_t need = flog2(cb.seg_len) + 1, have = 3 + flog2(cb.keep);
                inc = need > have ? need - have : 0;
            }
            for (uint32_t i = 0; i < inc; i++) bw.bit(1);
            bw.bit(0);
            bw.bits(cb.seg_len, (int)(3 + inc + flog2(cb.keep)));
        }
    }
    bw.flush();
}

void write_codestream_head(const Plan& p, size_t payload_len, std::vector<uint8_t>& o) {
    be16(o, 0xff4f);  // SOC
    be16(o, 0xff51);  // SIZ
    be16(o, 38 + 3 * p.ncomp);
    be16(o, 0);
    be32(o, p.width);
    be32(o, p.height);
    be32(o, 0);
    be32(o, 0);
    be32(o, p.width);
    be32(o, p.height);
    be32(o, 0);
    be32(o, 0);
    be16(o, p.ncomp);
    for (int c = 0; c < p.ncomp; c++) {
        o.push_back(7);  // 8-bit unsigned
        o.push_back(1);
        o.push_back(1);
    }
    be16(o, 0xff52);  // COD
    be16(o, 12);
    o.push_back(0);
    o.push_back(0);   // LRCP
    be16(o, 1);       // one layer
    o.push_back(p.use_mct ? 1 : 0);
    o.push_back((uint8_t)p.levels);
    o.push_back((uint8_t)(__builtin_ctz((unsigned)p.cbw) - 2));
    o.push_back((uint8_t)(__builtin_ctz((unsigned)p.cbh) - 2));
    o.push_back(0);
    o.push_back(p.lossless ? 1 : 0);
    be16(o, 0xff5c);  // QCD
    if (p.lossless) {
        be16(o, 3 + (uint32_t)p.quants.size());
        o.push_back((uint8_t)(p.guard << 5));
        for (auto& q : p.quants) o.push_back((uint8_t)(q.exponent << 3));
    } else {
        be16(o, 3 + 2 * (uint32_t)p.quants.size());
        o.push_back((uint8_t)((p.guard << 5) | 2));
        for (auto& q : p.quants) be16(o, ((uint32_t)q.exponent << 11) | (q.mantissa & 0x7ffu));
    }
    be16(o, 0xff90);  // SOT
    be16(o, 10);
    be16(o, 0);
    be32(o, (uint32_t)(14 + payload_len));
    o.push_back(0);
    o.push_back(1);
    be16(o, 0xff93);  // SOD
}

void write_codestream_tail(std::vector<uint8_t>& o) { be16(o, 0xffd9); }

void write_jp2_head(const Plan& p, size_t cslen, std::vector<uint8_t>& o) {
    be32(o, 12);
    tag(o, "jP  ");
    be32(o, 0x0d0a870a);
    be32(o, 20);
    tag(o, "ftyp");
    tag(o, "jp2 ");
    be32(o, 0);
    tag(o, "jp2 ");
    be32(o, 8 + 22 + 15);
    tag(o, "jp2h");
    be32(o, 22);
    tag(o, "ihdr");
    be32(o, p.height);
    be32(o, p.width);
    be16(o, p.ncomp);
    o.push_back(7);
    o.push_back(7);
    o.push_back(0);
    o.push_back(0);
    be32(o, 15);
    tag(o, "colr");
    o.push_back(1);
    o.push_back(0);
    o.push_back(0);
    be32(o, p.domain == GRAY ? 17 : 16);
    if (cslen + 8 > 0xffffffffull) throw Error(ENCODE_FAILED, "JP2 box length exceeds 32-bit limit");
    be32(o, (uint32_t)(cslen + 8));
    tag(o, "jp2c");
}

// ---- decoding ---------------------------------------------------------------------
namespace {
uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] << 8 | p[1]); }
[[noreturn]] void fail(const std::string& m) { throw Error(DECODE_FAILED, m); }
[[noreturn]] void unsupported(const std::string& m) { throw Error(UNSUPPORTED_FEATURE, m); }

struct Box {
    char type[4];
    const uint8_t* p;
    size_t n;
};
// One box at `pos` (decode/jp2_parse.rs:190-235)
bool next_box(const uint8_t* b, size_t n, size_t& pos, Box& bx) {
    if (pos == n) return false;
    if (n - pos < 8) fail("truncated JP2 box header");
    size_t start = pos;
    uint64_t lbox = rd32(b + pos);
    std::memcpy(bx.type, b + pos + 4, 4);
    pos += 8;
    size_t end;
    if (lbox == 0) {
        end = n;
    } else if (lbox == 1) {
        if (n - pos < 8) fail("unexpected end of JP2 u64");
        uint64_t xl = ((uint64_t)rd32(b + pos) << 32) | rd32(b + pos + 4);
        pos += 8;
        if (xl < 16) fail("invalid extended JP2 box length");
        if (xl > n - start) fail("JP2 box extends past input");
        end = start + (size_t)xl;
    } else if (lbox < 8) {
        fail("invalid JP2 box length below header size");
    } else {
        if (lbox > n - start) fail("JP2 box extends past input");
        end = start + (size_t)lbox;
    }
    if (end < pos) fail("JP2 box extends past input");
    bx.p = b + pos;
    bx.n = end - pos;
    pos = end;
    return true;
}
bool is(const Box& b, const char* t) { return std::memcmp(b.type, t, 4) == 0; }
}  // namespace

void parse_jp2(const uint8_t* b, size_t n, Jp2Info& info, const uint8_t*& cs, size_t& cslen) {
    size_t pos = 0;
    Box bx;
    if (!next_box(b, n, pos, bx)) fail("missing JP2 signature box");
    if (!is(bx, "jP  ") || bx.n != 4 || rd32(bx.p) != 0x0d0a870a) fail("invalid JP2 signature box");
    if (!next_box(b, n, pos, bx)) fail("missing JP2 file type box");
    if (!is(bx, "ftyp")) fail("JP2 file type box must follow signature box");
    if (bx.n < 12) fail("JP2 file type box is too short");
    if (std::memcmp(bx.p, "jp2 ", 4)) fail("JP2 brand is not jp2");
    bool compat = false;
    for (size_t k = 8; k + 4 <= bx.n; k += 4) compat |= std::memcmp(bx.p + k, "jp2 ", 4) == 0;
    if (!compat) fail("JP2 file type lacks jp2 compatibility brand");
    bool have_hdr = false;
    cs = nullptr;
    while (next_box(b, n, pos, bx)) {
        if (is(bx, "jp2h")) {
            size_t q = 0;
            Box sb;
            bool ihdr = false, colr = false;
            while (next_box(bx.p, bx.n, q, sb)) {
                if (is(sb, "ihdr")) {
                    if (sb.n != 14) fail("JP2 ihdr box must be 14 bytes");
                    if (sb.p[11] != 7) fail("JP2 ihdr compression type is not JPEG 2000");
                    if (sb.p[10] == 0xff) unsupported("unsupported JP2 feature: per-component bit depths require bpcc");
                    info.height = rd32(sb.p);
                    info.width = rd32(sb.p + 4);
                    info.ncomp = rd16(sb.p + 8);
                    info.bpc = (sb.p[10] & 0x7f) + 1;
                    info.ipr = sb.p[13] != 0;
                    ihdr = true;
                } else if (is(sb, "bpcc")) {
                    unsupported("unsupported JP2 feature: per-component bit-depth box (bpcc)");
                } else if (is(sb, "colr") && !colr) {
                    if (sb.n < 3) fail("JP2 colr box is too short");
                    if (sb.p[0] != 1) unsupported("only enumerated JP2 colorspaces are supported");
                    if (sb.n != 7) fail("enumerated JP2 colr box must be 7 bytes");
                    uint32_t e = rd32(sb.p + 3);
                    if (e == 17) info.colorspace = GRAY;
                    else if (e == 16) info.colorspace = SRGB;
                    else unsupported("unsupported JP2 EnumCS value " + std::to_string(e));
                    colr = true;
                } else if (is(sb, "pclr")) {
                    unsupported("unsupported JP2 feature: palette box (pclr)");
                } else if (is(sb, "cmap")) {
                    unsupported("unsupported JP2 feature: component mapping box (cmap)");
                } else if (is(sb, "cdef")) {
                    unsupported("unsupported JP2 feature: channel definition box (cdef)");
                }
            }
            if (!ihdr) fail("JP2 header lacks ihdr box");
            if (!colr) fail("JP2 header lacks colr box");
            have_hdr = true;
        } else if (is(bx, "jp2c")) {
            if (cs) fail("unsupported JP2 layout: multiple contiguous codestream boxes");
            cs = bx.p;
            cslen = bx.n;
        }
    }
    if (!have_hdr) fail("missing JP2 header box");
    if (!cs) fail("missing contiguous codestream box");
}

void parse_codestream(const uint8_t* b, size_t n, CodestreamInfo& ci) {
    if (n < 2 || rd16(b) != 0xff4f) fail("codestream did not start with SOC");
    size_t pos = 2;
    bool siz = false, cod = false, qcd = false;
    int rsiz = 0, scod = 0, cbstyle = 0;
    uint32_t xo = 0, yo = 0, tw = 0, th = 0, txo = 0, tyo = 0;
    std::vector<int> sgnd, dx, dy;
    for (;;) {
        if (n - pos < 2) fail("codestream ended before EOC");
        uint16_t m = rd16(b + pos);
        if (m == 0xff90) break;
        if (n - pos < 4) fail("truncated marker segment");
        size_t L = rd16(b + pos + 2);
        if (L < 2) fail("invalid marker length");
        if (L > n - pos - 2) fail("marker extended past end of codestream");
        const uint8_t* body = b + pos + 4;
        size_t bl = L - 2;
        switch (m) {
            case 0xff51: {
                if (bl < 38) fail("SIZ marker body is too short");
                rsiz = rd16(body);
                ci.width = rd32(body + 2);
                ci.height = rd32(body + 6);
                xo = rd32(body + 10);
                yo = rd32(body + 14);
                tw = rd32(body + 18);
                th = rd32(body + 22);
                txo = rd32(body + 26);
                tyo = rd32(body + 30);
                ci.ncomp = rd16(body + 34);
                if (bl != 36 + 3 * (size_t)ci.ncomp) fail("SIZ marker length does not match component count");
                for (int c = 0; c < ci.ncomp; c++) {
                    ci.prec.push_back((body[36 + 3 * c] & 0x7f) + 1);
                    sgnd.push_back(body[36 + 3 * c] >> 7);
                    dx.push_back(body[37 + 3 * c]);
                    dy.push_back(body[38 + 3 * c]);
                }
                siz = true;
                break;
            }
            case 0xff52: {
                if (bl < 10) fail("COD marker body is too short");
                scod = body[0];
                ci.progression = body[1];
                ci.layers = rd16(body + 2);
                ci.mct = body[4];
                ci.levels = body[5];
                if (body[6] > 8 || body[7] > 8) unsupported("unsupported code-block size exponent");
                ci.cbw = 1 << (body[6] + 2);
                ci.cbh = 1 << (body[7] + 2);
                cbstyle = body[8];
                ci.transform = body[9];
                cod = true;
                break;
            }
            case 0xff5c: {
                if (bl < 1) fail("QCD marker body is too short");
                ci.qstyle = body[0] & 0x1f;
                ci.guard = body[0] >> 5;
                ci.exps.clear();
                ci.mants.clear();
                if (ci.qstyle == 0) {
                    for (size_t i = 1; i < bl; i++) {
                        ci.exps.push_back(body[i] >> 3);
                        ci.mants.push_back(0);
                    }
                } else {
                    for (size_t i = 1; i + 2 <= bl; i += 2) {
                        uint16_t v = rd16(body + i);
                        ci.exps.push_back((uint8_t)(v >> 11));
                        ci.mants.push_back(v & 0x7ff);
                    }
                }
                qcd = true;
                break;
            }
            case 0xff64:
                break;  // COM
            case 0xff50: unsupported("unsupported JPEG 2000 feature: CAP capabilities marker");
            case 0xff53: unsupported("unsupported JPEG 2000 feature: COC component coding-style override");
            case 0xff55: unsupported("unsupported JPEG 2000 feature: TLM tile-part length marker");
            case 0xff57: unsupported("unsupported JPEG 2000 feature: PLM packet-length marker");
            case 0xff58: unsupported("unsupported JPEG 2000 feature: PLT packet-length marker");
            case 0xff5d: unsupported("unsupported JPEG 2000 feature: QCC component quantization override");
            case 0xff5e: unsupported("unsupported JPEG 2000 feature: RGN region-of-interest marker");
            case 0xff5f: unsupported("unsupported JPEG 2000 feature: POC progression-order change marker");
            case 0xff60: unsupported("unsupported JPEG 2000 feature: PPM packed packet headers");
            case 0xff61: unsupported("unsupported JPEG 2000 feature: PPT packed packet headers");
            case 0xff63: unsupported("unsupported JPEG 2000 feature: CRG component registration marker");
            default: unsupported("unsupported codestream marker in main header");
        }
        pos += 2 + L;
    }
    // tile-parts (decode/codestream.rs:41-80)
    int tiles = 0;
    int tile_index = 0, part_index = 0, total_parts = 0;
    for (;;) {
        size_t sot = pos;
        if (n - pos < 12) fail("missing Psot in SOT segment");
        if (rd16(b + pos + 2) != 10) fail("invalid SOT marker length");
        int ti = rd16(b + pos + 4);
        uint32_t psot = rd32(b + pos + 6);
        int pi = b[pos + 10], tp = b[pos + 11];
        pos += 12;
        while (n - pos >= 2 && rd16(b + pos) != 0xff93) {
            uint16_t m = rd16(b + pos);
            if (m != 0xff64) unsupported("unsupported marker in tile-part header");
            if (n - pos < 4) fail("truncated marker segment");
            size_t L = rd16(b + pos + 2);
            if (L < 2) fail("invalid marker length");
            if (L > n - pos - 2) fail("marker extended past end of codestream");
            pos += 2 + L;
        }
        if (n - pos < 2) fail("tile-part lacks SOD");
        pos += 2;
        size_t end = psot ? sot + psot : n - 2;
        if (end > n || end < pos) fail("tile-part extends past end of codestream");
        if (tiles == 0) {
            ci.payload = b + pos;
            ci.payload_len = end - pos;
            tile_index = ti;
            part_index = pi;
            total_parts = tp;
        }
        tiles++;
        pos = end;
        if (n - pos >= 2 && rd16(b + pos) == 0xff90) continue;
        break;
    }
    if (n - pos != 2 || rd16(b + pos) != 0xffd9) fail("expected EOC at end of codestream");
    if (!siz) fail("codestream main header lacks SIZ");
    if (!cod) fail("codestream main header lacks COD");
    if (!qcd) fail("codestream main header lacks QCD");
    ci.tile_parts = tiles;
    // decode_markers.rs:348-501
    if (tiles != 1) fail("unsupported tiling: decoder currently supports one full-image tile in one tile-part");
    if (tile_index != 0) fail("only tile index 0 is supported");
    if (part_index != 0) fail("only first tile-part is supported");
    if (total_parts > 1) fail("unsupported tiling: multiple tile-parts are not implemented");
    if (rsiz != 0) fail("only Part 1 Rsiz=0 codestreams are supported");
    if (xo || yo || txo || tyo) fail("non-zero image or tile origins are unsupported");
    if (tw != ci.width || th != ci.height) fail("unsupported tiling: decoder currently supports only a single full-image tile");
    if (ci.ncomp != 1 && ci.ncomp != 3) fail("unsupported component layout");
    for (int c = 0; c < ci.ncomp; c++) {
        if (ci.prec[c] != 8 || sgnd[c]) fail("unsupported sample precision: decoder currently supports unsigned 8-bit samples");
        if (dx[c] != 1 || dy[c] != 1) fail("subsampled components are unsupported");
    }
    if (ci.progression != 0) fail("only LRCP progression is supported");
    uint32_t md = std::min(ci.width, ci.height);
    uint32_t maxd = md <= 1 ? 0 : 31 - __builtin_clz(md);
    if ((uint32_t)ci.levels > maxd) fail("COD decomposition levels exceed DWT limit");
    if (ci.layers == 0) fail("COD must signal at least one layer");
    if (ci.layers != 1) fail("unsupported quality layers: decoder currently supports exactly one layer");
    if (ci.ncomp == 1 && ci.mct) fail("MCT is unsupported for grayscale decoder slice");
    if (ci.ncomp == 3 && !ci.mct) fail("unsupported RGB codestream: expected MCT=1 for Archive.org sRGB profile");
    if (scod & 7) fail("unsupported packet syntax: precinct partitioning, SOP markers, and EPH markers are not implemented");
    if (cbstyle) fail("unsupported code-block style");
    if (ci.transform == 1 && ci.qstyle != 0) fail("unsupported quantization: reversible 5/3 currently supports no-quantization QCD only");
    if (ci.transform == 0 && ci.qstyle == 0) fail("unsupported quantization: irreversible 9/7 requires scalar quantization");
    if (ci.transform == 0 && ci.qstyle == 1) fail("unsupported quantization: scalar-derived QCD is not implemented");
    if (ci.cbw * ci.cbh > 4096) fail("COD code-block size exceeds Part 1 limits");
    if (ci.exps.size() != (size_t)(1 + 3 * ci.levels)) fail("QCD step count does not match decomposition levels");
    if (ci.levels > kMaxDecodeLevels) fail("COD decomposition levels exceed DWT limit");
}

namespace {
// Packet-header bit reader (decode/t2.rs:395-445): after an 0xFF byte the
// next byte carries 7 bits.
struct BitReader {
    const uint8_t* b;
    size_t n, pos = 0;
    uint32_t reg = 0, ct = 0;
    bool prev_ff = false;
    uint32_t bit() {
        if (ct == 0) {
            if (pos >= n) fail("packet header ended before requested bit");
            uint8_t byte = b[pos++];
            reg = byte;
            ct = prev_ff ? 7 : 8;
            prev_ff = byte == 0xff;
        }
        ct--;
        return (reg >> ct) & 1u;
    }
    uint32_t bits(int k) {
        uint32_t v = 0;
        for (int i = 0; i < k; i++) v = (v << 1) | bit();
        return v;
    }
};
}  // namespace

struct PacketParser::Band {
    int comp, res, band, first, cx, cy;
    TagTree incl, imsb;
};

PacketParser::PacketParser(const CodestreamInfo& ci, std::vector<DecBlock>& blocks) : ci_(&ci), blocks_(&blocks) {
    blocks.clear();
    const int L = ci.levels;
    for (int c = 0; c < ci.ncomp; c++)
        for (int r = 0; r <= L; r++)
            for (int bd = r ? HL : LL; bd <= (r ? HH : LL); bd++) {
                Rect rc = band_rect(ci.width, ci.height, L, r, bd);
                int cx = rc.x1 > rc.x0 ? (rc.x1 - rc.x0 + ci.cbw - 1) / ci.cbw : 0;
                int cy = rc.y1 > rc.y0 ? (rc.y1 - rc.y0 + ci.cbh - 1) / ci.cbh : 0;
                int bi = (int)bands_.size();
                bands_.push_back({c, r, bd, (int)blocks.size(), cx, cy, TagTree(std::max(cx, 1), std::max(cy, 1), 0x7fffffff),
                                  TagTree(std::max(cx, 1), std::max(cy, 1), 0x7fffffff)});
                for (int by = rc.y0; by < rc.y1; by += ci.cbh)
                    for (int bx = rc.x0; bx < rc.x1; bx += ci.cbw) {
                        DecBlock db;
                        db.comp = c;
                        db.res = r;
                        db.band = bd;
                        db.band_index = bi;
                        db.r = {bx, by, std::min(bx + ci.cbw, rc.x1), std::min(by + ci.cbh, rc.y1)};
                        blocks.push_back(db);
                    }
            }
}
PacketParser::~PacketParser() = default;

void PacketParser::run(int rmax) {
    const CodestreamInfo& ci = *ci_;
    std::vector<DecBlock>& blocks = *blocks_;
    const uint8_t* pay = ci.payload;
    const size_t plen = ci.payload_len;
    std::vector<int> order;
    // one layer (checked with the header), so LRCP is resolution-major
    for (; layer_ < ci.layers; layer_++, res_ = 0)
        for (; res_ <= ci.levels; res_++) {
            if (res_ > rmax) return;
            for (int c = 0; c < ci.ncomp; c++) {
                BitReader br{pay + pos_, plen - pos_};
                order.clear();
                uint64_t body = 0;
                if (br.bit()) {
                    for (auto& bs : bands_) {
                        if (bs.comp != c || bs.res != res_) continue;
                        for (int k = 0; k < bs.cx * bs.cy; k++) {
                            DecBlock& blk = blocks[bs.first + k];
                            bool inc = blk.included ? br.bit() != 0 : bs.incl.decode(br, k, layer_ + 1);
                            if (!inc) continue;
                            if (blk.zbp < 0) {
                                bs.imsb.decode(br, k, 999);
                                if (!bs.imsb.known(k)) fail("zero-bitplane tag tree did not terminate");
                                blk.zbp = bs.imsb.value(k);
                            }
                            uint32_t np;
                            if (!br.bit()) np = 1;
                            else if (!br.bit()) np = 2;
                            else {
                                uint32_t v = br.bits(2);
                                if (v != 3) np = 3 + v;
                                else {
                                    v = br.bits(5);
                                    np = v != 31 ? 6 + v : 37 + br.bits(7);
                                }
                            }
                            uint32_t incr = 0;
                            while (br.bit()) incr++;
                            blk.numlenbits += (int)incr;
                            uint32_t len = br.bits(blk.numlenbits + (int)flog2(np));
                            if (len == 0) fail("non-empty code-block contribution has zero byte length");
                            blk.included = true;
                            blk.passes = (int)np;
                            blk.len = len;
                            order.push_back(bs.first + k);
                            body += len;
                        }
                    }
                }
                pos_ += br.pos;
                if (body > plen - pos_) fail("packet body extends past tile payload");
                for (int i : order) {
                    blocks[i].off = (uint32_t)pos_;
                    pos_ += blocks[i].len;
                }
            }
        }
}

void PacketParser::finish() {
    if (pos_ != ci_->payload_len)
        fail("tile payload has " + std::to_string(ci_->payload_len - pos_) + " trailing bytes after packet parse");
}

void parse_packets(const CodestreamInfo& ci, std::vector<DecBlock>& blocks) {
    PacketParser pp(ci, blocks);
    pp.run(ci.levels);
    pp.finish();
}


// decode/t1.rs:96-127: per-block checks, in contribution order, before any
// block reaches the device decoder.
void check_decode_blocks(const CodestreamInfo& ci, const std::vector<DecBlock>& blocks) {
    const size_t nsteps = ci.exps.size();
    if (nsteps == 0) fail("QCD step count does not match decomposition levels");
    for (const auto& b : blocks) {
        if (!b.included || b.passes == 0) continue;
        int mb = std::max(ci.guard - 1, 0) + ci.exps[(size_t)b.band_index % nsteps];
        mb &= 255;
        if (b.zbp > 255) fail("zero-bitplane count exceeds supported range");
        if (b.zbp > mb) fail("zero-bitplane count exceeds subband bit-plane count");
        int mbp = mb - b.zbp;
        if (mbp == 0) fail("non-empty code-block has no magnitude bit-planes");
        int maxp = 1 + 3 * (mbp - 1);
        if (b.passes > maxp)
            fail("code-block pass count " + std::to_string(b.passes) + " exceeds maximum " + std::to_string(maxp) +
                 " for " + std::to_string(mbp) + " magnitude bit-planes");
        if (mbp > 31) unsupported("more than 31 magnitude bit-planes");
    }
}

// decode/mod.rs:56-129 up to the Tier-1 decode: every host-side step of
// decode_jp2, with the device path's own limits.
void parse_decode_header(const uint8_t* bytes, size_t len, Jp2Info& jp, CodestreamInfo& ci) {
    const uint8_t* cs;
    size_t cslen;
    parse_jp2(bytes, len, jp, cs, cslen);
    parse_codestream(cs, cslen, ci);
    if (jp.width != ci.width || jp.height != ci.height) fail("JP2 ihdr dimensions do not match SIZ dimensions");
    if (jp.bpc != 8)
        unsupported("unsupported JP2 bit depth: decoder currently supports 8-bit components, found " +
                    std::to_string(jp.bpc) + " bits");
    if (jp.ncomp != ci.ncomp) fail("JP2 ihdr component count does not match SIZ component count");
    if (jp.colorspace == GRAY && jp.ncomp != 1)
        unsupported("unsupported JP2 component count: decoder currently supports one grayscale component");
    if (jp.colorspace == SRGB && jp.ncomp != 3)
        unsupported("unsupported JP2 component count: decoder currently supports three sRGB components");
    if (ci.width > 65535 || ci.height > 65535)
        unsupported("image dimensions above 65535 are not supported by the device path");
    if (ci.cbw > 64 || ci.cbh > 64) unsupported("code-blocks larger than 64x64 are not supported by the device path");
}
void parse_for_decode(const uint8_t* bytes, size_t len, Jp2Info& jp, CodestreamInfo& ci, std::vector<DecBlock>& blocks) {
    parse_decode_header(bytes, len, jp, ci);
    parse_packets(ci, blocks);
    check_decode_blocks(ci, blocks);
}

// content.rs:120-174: f32 statistics from the 256 luma bins and the
// chroma-spread sum, then ContentStats::classify (content.rs:96-105).
ContentOut content_from_hist(const uint64_t* h, uint64_t n) {
    ContentOut o;
    if (n == 0) return o;  // content.rs:155-157
    const float total = (float)n;
    uint64_t bright = 0;
    for (int v = 225; v < 256; v++) bright += h[v];  // luma > 224
    float entropy = 0.0f;
    for (int v = 0; v < 256; v++)
        if (h[v]) {
            const float p = (float)h[v] / total;
            entropy -= p * std::log2(p);
        }
    o.bright_share = (float)bright / total;
    o.entropy = entropy;
    o.colorfulness = (float)h[256] / total;
    o.graphics = entropy < 3.0f || (o.bright_share > 0.45f && entropy < 5.0f);
    return o;
}

}  // namespace j2l
