// PCRD rate/distortion bookkeeping, hull pruning, lambda calibration and
// truncation on the GPU, plus packing of the selected code-block bytes into
// codestream order.
//
// f64 arithmetic follows the reference's evaluation order
// (encode/backend/native/rate.rs:128-180, dwt/pcrd.rs:151-252, :348-424,
// :527-554, :669-696, :1493-1514) so the hulls and selections are identical.
#include "common.hpp"

namespace j2l {

namespace {

__device__ __forceinline__ double rd_slope(uint32_t ba, double da, uint32_t bb, double db_) {
    double db = (double)(bb - ba);
    double dd = __dsub_rn(db_, da);
    if (db == 0.0) return dd > 0.0 ? __longlong_as_double(0x7ff0000000000000ll) : 0.0;
    return __ddiv_rn(dd, db);
}

__device__ __forceinline__ double class_weight(int cls, bool ll) {
    double base = cls == 0 ? 2.5 : (cls == 1 ? 1.5 : 1.0);
    return ll ? __dmul_rn(base, 1.1) : base;
}

}  // namespace

// Per-block pass lengths (t1.rs:586-600), distortion deltas and convex hull.
// G (1 or kHullG) blocks per wave: the lanes bring the blocks' pass records
// into LDS together, then lane 8g runs block g's hull with its stack in LDS
// (a per-thread stack would sit in scratch, and every pass would wait a
// memory latency).  One block per wave is the fastest for a page; large
// batches take 8 per wave to keep the wave count down.  `hull` has
// kMaxPasses + 1 points per block.
constexpr int kHullG = 8;
__global__ __launch_bounds__(64) void pcrd_hull_kernel(int n, int G, const BlockResult* __restrict__ res,
                                                       const uint32_t* __restrict__ snaps,
                                                       const uint16_t* __restrict__ newsigs,
                                                       const PcrdJob* __restrict__ jobs, HullPt* __restrict__ hull,
                                                       uint8_t* __restrict__ hull_n,
                                                       unsigned long long* __restrict__ gmax) {
    __shared__ uint32_t s_sp[kHullG][kMaxPasses], s_hb[kHullG][kMaxPasses + 1];
    __shared__ uint16_t s_ns[kHullG][kMaxPasses];
    __shared__ double s_hd[kHullG][kMaxPasses + 1];
    __shared__ int s_hp[kHullG][kMaxPasses + 1];
    const int b0 = blockIdx.x * G;
    for (int g = 0; g < G && b0 + g < n; g++) {
        const int npg = min((int)res[b0 + g].npasses, kMaxPasses);
        for (int i = threadIdx.x; i < npg; i += blockDim.x) {
            s_sp[g][i] = snaps[(size_t)(b0 + g) * kMaxPasses + i];
            s_ns[g][i] = newsigs[(size_t)(b0 + g) * kMaxPasses + i];
        }
    }
    __syncthreads();
    const int g = (int)threadIdx.x >> 3;
    const int b = b0 + g;
    if ((threadIdx.x & 7u) != 0 || g >= G || b >= n) return;
    const uint32_t* sp = s_sp[g];
    const uint16_t* ns = s_ns[g];
    uint32_t* hb = s_hb[g];
    double* hd = s_hd[g];
    int* hp = s_hp[g];
    const BlockResult R = res[b];
    const int np = R.npasses;
    const PcrdJob J = jobs[b];
    HullPt* H = hull + (size_t)b * (kMaxPasses + 1);
    int m = 0;
    hb[0] = 0;
    hd[0] = 0.0;
    hp[0] = 0;
    m = 1;
    uint32_t prev = 0, sum = 0, cum = 0;
    uint32_t sig_total = 0, sig_before_sp = 0;
    double cumd = 0.0;
    const double bw = class_weight(R.cls, J.is_ll != 0);
    const int top = R.mbp - 1;
    for (int i = 0; i < np; i++) {
        uint32_t s = sp[i];
        uint32_t len = s > prev ? s - prev : 0u;
        prev = s;
        if (i == np - 1) len = R.nbytes > sum ? R.nbytes - sum : 0u;
        sum += len;
        cum += len;
        int kind = i == 0 ? 2 : (i - 1) % 3;
        int bp = i == 0 ? top : top - 1 - (i - 1) / 3;
        uint32_t nsig = ns[i];
        if (kind == 0) sig_before_sp = sig_total;
        int bpc = bp < 30 ? bp : 30;
        double pw = __dmul_rn((double)(1ull << bpc), J.step);
        double energy = __dmul_rn(pw, pw);
        double contrib;
        if (kind == 0) contrib = __dmul_rn((double)nsig, energy);
        else if (kind == 1) contrib = __dmul_rn(__dmul_rn((double)sig_before_sp, energy), 0.25);
        else contrib = __dmul_rn(__dmul_rn((double)nsig, energy), 0.85);
        double raw = __dmul_rn(__dmul_rn(J.swb, bw), contrib);
        double dd = raw > 0.0 ? raw : 0.0;  // pcrd.rs:764-802 with visibility weight 1
        dd = __dmul_rn(dd, J.cw);
        sig_total += nsig;
        cumd = __dadd_rn(cumd, dd);
        // push and prune (pcrd.rs:209-243)
        hb[m] = cum + 2u;  // rate.rs:31 header overhead
        hd[m] = cumd;
        hp[m] = i + 1;
        m++;
        while (m >= 3) {
            double sab = rd_slope(hb[m - 3], hd[m - 3], hb[m - 2], hd[m - 2]);
            double sbc = rd_slope(hb[m - 2], hd[m - 2], hb[m - 1], hd[m - 1]);
            if (sbc >= sab) {
                hb[m - 2] = hb[m - 1];
                hd[m - 2] = hd[m - 1];
                hp[m - 2] = hp[m - 1];
                m--;
            } else {
                break;
            }
        }
    }
    double mx = 0.0;
    H[0].slope = __longlong_as_double(0x7ff0000000000000ll);
    H[0].passes = 0;
    for (int k = 1; k < m; k++) {
        double sl = rd_slope(hb[k - 1], hd[k - 1], hb[k], hd[k]);
        H[k].slope = sl;
        H[k].passes = hp[k];
        if (isfinite(sl) && sl > 0.0 && sl > mx) mx = sl;
    }
    hull_n[b] = (uint8_t)m;
    if (mx > 0.0) atomicMax(&gmax[J.group], (unsigned long long)__double_as_longlong(mx));
}

// Lambda per (image, component) group from the group's max slope
// (pcrd.rs:1493-1514).  The hull kernel folded each block's maximum into
// gmax with an integer atomicMax: slopes are non-negative, so the IEEE bit
// patterns order like the values and the result is order-independent.
__global__ void pcrd_lambda_kernel(int ngroups, const unsigned long long* __restrict__ gmax,
                                   const double* __restrict__ raw_floor, double* __restrict__ lambda) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    double m = __longlong_as_double((long long)gmax[g]);
    double raw = raw_floor[2 * g], floor_l = raw_floor[2 * g + 1];
    double l = raw;
    if (m > 0.0) {
        double target = __dmul_rn(m, 1.2);
        if (target < floor_l) l = __dmul_rn(raw, __ddiv_rn(target, floor_l));
    }
    lambda[g] = l;
}

// Deepest hull point with slope >= lambda (pcrd.rs:404-424) and the
// truncated body length (backend.rs:834-897).  With `lossless` every pass is
// kept.
__global__ void pcrd_select_kernel(int n, int lossless, const BlockResult* __restrict__ res,
                                   const uint32_t* __restrict__ snaps, const PcrdJob* __restrict__ jobs,
                                   const HullPt* __restrict__ hull, const uint8_t* __restrict__ hull_n,
                                   const double* __restrict__ lambda, Selection* __restrict__ sel) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const BlockResult R = res[b];
    int keep = R.npasses;
    if (!lossless) {
        const double l = lambda[jobs[b].group];
        const HullPt* H = hull + (size_t)b * (kMaxPasses + 1);
        keep = 0;
        for (int k = 1; k < hull_n[b]; k++) {
            if (H[k].slope >= l) keep = H[k].passes;
            else break;
        }
    }
    uint32_t body = 0;
    if (keep > 0) {
        if (keep >= R.npasses) {
            body = R.nbytes;
        } else {
            // cumulative length of the first `keep` passes = snapshot of pass keep-1
            const uint32_t* sp = snaps + (size_t)b * kMaxPasses;
            uint32_t prev = 0, cum = 0;
            for (int i = 0; i < keep; i++) {
                uint32_t s = sp[i];
                cum += s > prev ? s - prev : 0u;
                prev = s;
            }
            body = cum < R.nbytes ? cum : R.nbytes;
        }
    }
    Selection S;
    S.body_len = body;
    S.keep = (uint8_t)keep;
    S.zbp = R.zbp;
    S.pad[0] = S.pad[1] = 0;
    sel[b] = S;
}

// Exclusive scan of the body lengths of each image's blocks (codestream
// order) -> byte offsets in that image's packed body buffer.  One workgroup
// per image; totals[i] = packed size.
__global__ void pack_offsets_kernel(const uint32_t* __restrict__ img_first, const Selection* __restrict__ sel,
                                    uint32_t* __restrict__ offs, uint32_t* __restrict__ totals) {
    __shared__ uint32_t sh[1024];
    const int img = blockIdx.x;
    const uint32_t b0 = img_first[img], b1 = img_first[img + 1];
    uint32_t carry = 0;
    for (uint32_t base = b0; base < b1; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < b1 ? sel[i].body_len : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (unsigned o = 1; o < blockDim.x; o <<= 1) {
            uint32_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < b1) offs[i] = carry + sh[threadIdx.x] - v;
        uint32_t tot = sh[blockDim.x - 1];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) totals[img] = carry;
}

// Copy each block's selected prefix into the packed buffer (one wave per
// block; image i's bytes start at pack_base[i]).
__global__ void pack_copy_kernel(int n, const uint32_t* __restrict__ blk_img, const uint64_t* __restrict__ pack_base,
                                 const BlockJob* __restrict__ jobs, const Selection* __restrict__ sel,
                                 const uint32_t* __restrict__ offs, const uint8_t* __restrict__ slots,
                                 uint8_t* __restrict__ packed) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const uint32_t len = sel[b].body_len;
    const uint8_t* src = slots + jobs[b].out_off;
    uint8_t* dst = packed + pack_base[blk_img[b]] + offs[b];
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) dst[i] = src[i];
}

void launch_pcrd(int n, bool lossless, int ngroups, const BlockResult* res, const uint32_t* snaps,
                 const uint16_t* newsigs, const PcrdJob* jobs, HullPt* hull, uint8_t* hull_n,
                 unsigned long long* gmax, const double* raw_floor, double* lambda, Selection* sel, hipStream_t st) {
    if (n <= 0) return;
    if (!lossless) {
        const int G = n <= 4096 ? 1 : kHullG;
        pcrd_hull_kernel<<<(n + G - 1) / G, 64, 0, st>>>(n, G, res, snaps, newsigs, jobs, hull, hull_n, gmax);
        pcrd_lambda_kernel<<<(ngroups + 63) / 64, 64, 0, st>>>(ngroups, gmax, raw_floor, lambda);
    }
    pcrd_select_kernel<<<(n + 127) / 128, 128, 0, st>>>(n, lossless ? 1 : 0, res, snaps, jobs, hull, hull_n, lambda, sel);
}

void launch_pack(int n, int nimg, const uint32_t* img_first, const uint32_t* blk_img, const uint64_t* pack_base,
                 const BlockJob* jobs, const Selection* sel, uint32_t* offs, uint32_t* totals, const uint8_t* slots,
                 uint8_t* packed, hipStream_t st, bool offsets_only) {
    if (n <= 0) return;
    if (offsets_only) {
        pack_offsets_kernel<<<nimg, 1024, 0, st>>>(img_first, sel, offs, totals);
        return;
    }
    pack_copy_kernel<<<n, 64, 0, st>>>(n, blk_img, pack_base, jobs, sel, offs, slots, packed);
}

}  // namespace j2l
