// Colour transforms, multi-level forward/inverse DWT (5/3 integer, 9/7
// float) and finalisation for CDNA4 (gfx950).
//
// Each DWT level is one launch of 2-D tile kernels: a workgroup loads a
// 64x64 output tile plus a 4-sample mirrored halo into LDS, runs the
// vertical then horizontal lifting passes (inverse: horizontal then
// vertical) and writes the four subbands once.  Every level therefore reads
// and writes its active region once from HBM (plus the halo).  In each pass
// a thread lifts a window of 24 samples of one column (row) in registers:
// 16 outputs with 4 samples either side, all lifting steps and the scaling
// at once, so a pass costs one LDS read and one write per sample and one
// barrier, not one per lifting step.
//
// Arithmetic follows the reference's lifting order exactly
// (dwt/irrev97.rs:92-129, :261-336, :625-820; dwt/rev53.rs:143-417):
//   x[j] = x[j] + c * (x[j-1] + x[j+1])   (f32, no contraction)
// with whole-sample symmetric extension, scaling after the four lifts and
// the vertical pass before the horizontal one.  The halo trick is exact:
// lifting preserves the symmetry of a symmetrically extended signal, so the
// mirrored halo values equal the values the reference reads through
// fetch_sym at the borders.
#include "common.hpp"

namespace j2l {

namespace {

constexpr int TX = 64, TY = 64, HALO = 4;
constexpr int TLW = TX + 2 * HALO;      // 72
constexpr int TLH = TY + 2 * HALO;      // 72
constexpr int TLP = TLW + 1;             // padded row pitch
// A pass splits the 64 outputs of a column (row) into 4 segments of 16; a
// segment's window is 16 + 2 * HALO samples.  288 threads: 72 columns x 4
// segments in the forward vertical and the inverse horizontal pass.
constexpr int SEG = 16, NSEG = TY / SEG, WIN = SEG + 2 * HALO;
constexpr int kLevelThreads = TLW * NSEG;  // 288
constexpr int kLoadsPerThread = TLH / NSEG;  // the tile load: 18 rows per thread

// irrev97.rs:10-15
__constant__ float kAlpha = -1.586134342059924f;
__constant__ float kBeta = -0.052980118572961f;
__constant__ float kGamma = 0.882911075530934f;
__constant__ float kDelta = 0.443506852043971f;
constexpr float K97 = 1.230174104914001f;
constexpr float INV_K97 = 1.0f / 1.230174104914001f;

__device__ __forceinline__ int sym(int n, int i) {
    if (n <= 1) return 0;
    if (i >= 0 && i < n) return i;
    int p = 2 * (n - 1);
    int k = i % p;
    if (k < 0) k += p;
    return k >= n ? p - k : k;
}

__device__ __forceinline__ float lift(float x, float c, float l, float r) {
    return __fadd_rn(x, __fmul_rn(c, __fadd_rn(l, r)));
}

// One lifting step on the window's samples of parity `par` in [1, WIN-2]
// (window position parity is global parity: windows start at even tile
// rows/columns and tiles at multiples of 64).  The window's outer samples
// take stale values as the steps go on, one more per step from each end,
// so after the four 9/7 steps (two 5/3 steps) positions [HALO, HALO + SEG)
// equal the lifting of the whole signal (irrev97.rs:158-228, :339-399;
// rev53.rs:207-265).
template <int PAR>
__device__ __forceinline__ void lift_step(float (&v)[WIN], float c) {
#pragma unroll
    for (int j = PAR ? 1 : 2; j <= WIN - 2; j += 2) v[j] = lift(v[j], c, v[j - 1], v[j + 1]);
}
// 5/3 predict on odd samples (x -= (l + r) >> 1) and update on even ones
// (x += (l + r + 2) >> 2); the inverse undoes them in the other order.
__device__ __forceinline__ void predict53(int32_t (&v)[WIN], int sign) {
#pragma unroll
    for (int j = 1; j <= WIN - 2; j += 2) {
        const int32_t d = (v[j - 1] + v[j + 1]) >> 1;
        v[j] = sign > 0 ? v[j] - d : v[j] + d;
    }
}
__device__ __forceinline__ void update53(int32_t (&v)[WIN], int sign) {
#pragma unroll
    for (int j = 2; j <= WIN - 2; j += 2) {
        const int32_t d = (v[j - 1] + v[j + 1] + 2) >> 2;
        v[j] = sign > 0 ? v[j] + d : v[j] - d;
    }
}
// Forward 1-D transform of a window: the four lifts, then even x 1/K, odd
// x K (9/7); predict then update (5/3).
__device__ __forceinline__ void fwd_window(float (&v)[WIN]) {
    lift_step<1>(v, kAlpha);
    lift_step<0>(v, kBeta);
    lift_step<1>(v, kGamma);
    lift_step<0>(v, kDelta);
#pragma unroll
    for (int j = 0; j < WIN; j++) v[j] = __fmul_rn(v[j], (j & 1) ? K97 : INV_K97);
}
__device__ __forceinline__ void fwd_window(int32_t (&v)[WIN]) {
    predict53(v, 1);
    update53(v, 1);
}
// Inverse 1-D transform (irrev97.rs:636-820; rev53.rs:272-417): even x K,
// odd x 1/K, then -delta (even), -gamma (odd), -beta (even), -alpha (odd);
// 5/3: undo the update, then the predict.
__device__ __forceinline__ void inv_window(float (&v)[WIN]) {
#pragma unroll
    for (int j = 0; j < WIN; j++) v[j] = __fmul_rn(v[j], (j & 1) ? INV_K97 : K97);
    lift_step<0>(v, -kDelta);
    lift_step<1>(v, -kGamma);
    lift_step<0>(v, -kBeta);
    lift_step<1>(v, -kAlpha);
}
__device__ __forceinline__ void inv_window(int32_t (&v)[WIN]) {
    update53(v, -1);
    predict53(v, -1);
}

}  // namespace


// ---- segment tables ---------------------------------------------------------
// One launch covers every image/component that has a given decomposition
// level: each workgroup finds its segment (one image component at one
// level) by binary search over the tile start indices.
__device__ __forceinline__ int find_seg(const DwtSeg* segs, int nseg, uint32_t bid) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].tile_start <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One forward decomposition level of the active region rw x rh.
// Subbands: LL -> out, HL/TLH/HH -> plane at their Mallat offsets.  Tile
// origins are multiples of 64 so tile-local parity equals global parity.
template <typename T, bool IRREV>
__global__ __launch_bounds__(kLevelThreads) void fwd_level_kernel(const DwtSeg* __restrict__ segs, int nseg) {
    __shared__ T s[TLH][TLP];
    const DwtSeg S = segs[find_seg(segs, nseg, blockIdx.x)];
    const uint32_t t = blockIdx.x - S.tile_start;
    const int tx0 = (int)(t % (uint32_t)S.tiles_x) * TX, ty0 = (int)(t / (uint32_t)S.tiles_x) * TY;
    const int rw = S.rw, rh = S.rh;
    const int tid = threadIdx.x;
    const T* __restrict__ in = (const T*)S.in;
    T* __restrict__ plane = (T*)S.plane;
    T* __restrict__ ll_out = (T*)S.out;
    // tile load: thread = column lx, rows ly0, ly0 + 4, ...; all 18 loads in
    // flight before the LDS writes
    {
        const int lx = tid % TLW, ly0 = tid / TLW;
        const int gx = sym(rw, tx0 - HALO + lx);
        T r[kLoadsPerThread];
#pragma unroll
        for (int k = 0; k < kLoadsPerThread; k++) {
            const int gy = sym(rh, ty0 - HALO + ly0 + NSEG * k);
            r[k] = in[(size_t)gy * S.in_stride + gx];
        }
#pragma unroll
        for (int k = 0; k < kLoadsPerThread; k++) s[ly0 + NSEG * k][lx] = r[k];
    }
    __syncthreads();
    // vertical pass (irrev97.rs:261-336; rev53.rs:419-497): thread = column
    // c of all 72, segment g; writes back the segment's 16 output rows
    {
        const int c = tid % TLW, g = tid / TLW;
        if (rh >= 2) {
            T v[WIN];
#pragma unroll
            for (int j = 0; j < WIN; j++) v[j] = s[g * SEG + j][c];
            fwd_window(v);
            __syncthreads();  // every window read before any write-back
#pragma unroll
            for (int j = 0; j < SEG; j++) s[g * SEG + HALO + j][c] = v[HALO + j];
        } else {
            __syncthreads();
            if constexpr (IRREV) {
                if (g == 0) s[HALO][c] = __fmul_rn(s[HALO][c], INV_K97);
            }
        }
    }
    __syncthreads();
    // horizontal pass on the output rows (irrev97.rs:92-129; rev53.rs:143-265):
    // thread = output row r, segment g (256 threads)
    if (tid < TY * NSEG) {
        const int r = HALO + (tid % TY), g = tid / TY;
        if (rw >= 2) {
            T v[WIN];
#pragma unroll
            for (int j = 0; j < WIN; j++) v[j] = s[r][g * SEG + j];
            fwd_window(v);
            // the windows of neighbouring segments overlap by 8 samples:
            // every read before any write-back
            __syncthreads();
#pragma unroll
            for (int j = 0; j < SEG; j++) s[r][g * SEG + HALO + j] = v[HALO + j];
        } else {
            __syncthreads();
            if constexpr (IRREV) {
                if (g == 0) s[r][HALO] = __fmul_rn(s[r][HALO], INV_K97);
            }
        }
    } else {
        __syncthreads();
    }
    __syncthreads();
    // HL and HH start at plane column S.hx, LH and HH at Mallat row snh
    const int snh = (rh + 1) >> 1, bhx = (int)S.hx;
    for (int idx = tid; idx < TY * TX; idx += kLevelThreads) {
        int oy = idx / TX, ox = idx % TX;
        int gy = ty0 + oy, gx = tx0 + ox;
        if (gy >= rh || gx >= rw) continue;
        T v = s[HALO + oy][HALO + ox];
        int hy = gy >> 1, hx = gx >> 1;
        if (gy & 1) {
            if (gx & 1) plane[(size_t)(snh + hy) * S.pstride + bhx + hx] = v;
            else plane[(size_t)(snh + hy) * S.pstride + hx] = v;
        } else {
            if (gx & 1) plane[(size_t)hy * S.pstride + bhx + hx] = v;
            else ll_out[(size_t)hy * S.out_stride + hx] = v;
        }
    }
}

// One inverse level: LL from `in`, HL/TLH/HH from `plane`, interleaved
// rw x rh result into `out` (irrev97.rs:636-820; rev53.rs:95-136, :272-707).
template <typename T, bool IRREV>
__global__ __launch_bounds__(kLevelThreads) void inv_level_kernel(const DwtSeg* __restrict__ segs, int nseg) {
    __shared__ T s[TLH][TLP];
    const DwtSeg S = segs[find_seg(segs, nseg, blockIdx.x)];
    const uint32_t t = blockIdx.x - S.tile_start;
    const int tx0 = (int)(t % (uint32_t)S.tiles_x) * TX, ty0 = (int)(t / (uint32_t)S.tiles_x) * TY;
    const int rw = S.rw, rh = S.rh;
    const int tid = threadIdx.x;
    const T* __restrict__ ll_in = (const T*)S.in;
    const T* __restrict__ plane = (const T*)S.plane;
    T* __restrict__ out = (T*)S.out;
    const int snh = (rh + 1) >> 1, bhx = (int)S.hx;
    // tile gather from the four bands: thread = column lx, rows ly0, ly0 + 4,
    // ...; all 18 loads in flight before the LDS writes
    {
        const int lx = tid % TLW, ly0 = tid / TLW;
        const int mx = sym(rw, tx0 - HALO + lx), hx = mx >> 1;
        const int hcol = (mx & 1) ? bhx + hx : hx;  // column in the plane
        T r[kLoadsPerThread];
#pragma unroll
        for (int k = 0; k < kLoadsPerThread; k++) {
            const int my = sym(rh, ty0 - HALO + ly0 + NSEG * k), hy = my >> 1;
            const T* src = (my & 1) ? plane + (size_t)(snh + hy) * S.pstride + hcol
                         : (mx & 1) ? plane + (size_t)hy * S.pstride + hcol
                                    : ll_in + (size_t)hy * S.in_stride + hx;
            r[k] = *src;
        }
#pragma unroll
        for (int k = 0; k < kLoadsPerThread; k++) s[ly0 + NSEG * k][lx] = r[k];
    }
    __syncthreads();
    // horizontal inverse on every tile row (the vertical pass needs them):
    // thread = row r of all 72, segment g; writes back the 16 output columns
    {
        const int r = tid % TLH, g = tid / TLH;
        if (rw >= 2) {
            T v[WIN];
#pragma unroll
            for (int j = 0; j < WIN; j++) v[j] = s[r][g * SEG + j];
            inv_window(v);
            __syncthreads();
#pragma unroll
            for (int j = 0; j < SEG; j++) s[r][g * SEG + HALO + j] = v[HALO + j];
        } else {
            __syncthreads();
            if constexpr (IRREV) {
                if (g == 0) s[r][HALO] = __fmul_rn(s[r][HALO], K97);
            }
        }
    }
    __syncthreads();
    // vertical inverse on the output columns, straight to the output:
    // thread = output column c, segment g (256 threads)
    if (tid < TX * NSEG) {
        const int c = HALO + (tid % TX), g = tid / TX;
        const int gx = tx0 + c - HALO;
        T v[WIN];
        if (rh >= 2) {
#pragma unroll
            for (int j = 0; j < WIN; j++) v[j] = s[g * SEG + j][c];
            inv_window(v);
        } else {
#pragma unroll
            for (int j = 0; j < WIN; j++) v[j] = s[g * SEG + j][c];
            if constexpr (IRREV) v[HALO] = __fmul_rn(v[HALO], K97);
        }
        if (gx < rw) {
#pragma unroll
            for (int j = 0; j < SEG; j++) {
                const int gy = ty0 + g * SEG + j;
                if (gy < rh) out[(size_t)gy * S.out_stride + gx] = v[HALO + j];
            }
        }
    }
}

// ---- colour / level shift / finalisation --------------------------------------
__device__ __forceinline__ int find_pix(const PixSeg* segs, int nseg, uint32_t bid) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].blk_start <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pixel kernels: a thread takes 4 consecutive pixels (a wave 256), read
// and written with vector accesses when the pointers allow it, else one
// element at a time.
__device__ __forceinline__ bool aligned_to(uint64_t p, uint32_t a) { return (p & (a - 1u)) == 0; }

// One transformed sample (scalar.rs:135-195): level shift, forward ICT
// evaluated left to right without contraction, or forward RCT.
__device__ __forceinline__ float ict_sample(int comp, int r, int g, int b) {
    const float rf = (float)(r - 128), gf = (float)(g - 128), bf = (float)(b - 128);
    if (comp == 0) return __fadd_rn(__fadd_rn(__fmul_rn(0.299f, rf), __fmul_rn(0.587f, gf)), __fmul_rn(0.114f, bf));
    if (comp == 1) return __fadd_rn(__fadd_rn(__fmul_rn(-0.16875f, rf), __fmul_rn(-0.33126f, gf)), __fmul_rn(0.5f, bf));
    return __fsub_rn(__fsub_rn(__fmul_rn(0.5f, rf), __fmul_rn(0.41869f, gf)), __fmul_rn(0.08131f, bf));
}
__device__ __forceinline__ int rct_sample(int comp, int r, int g, int b) {
    const int rv = r - 128, gv = g - 128, bv = b - 128;
    return comp == 0 ? ((rv + 2 * gv + bv) >> 2) : comp == 1 ? (bv - gv) : (rv - gv);
}

// Level shift (scalar.rs:242-254), forward ICT (scalar.rs:256-278) and
// forward RCT (scalar.rs:280-302) into a component plane, for ch channels
// (1 or 3) of input.
template <int ch>
__device__ __forceinline__ void prepare4(const PixSeg& S, uint64_t i) {
    const bool full = i + 4u <= S.n;
    int r[4], g[4], b[4];
    if (S.planar_i32) {
        const int32_t* p = (const int32_t*)S.src;
#pragma unroll
        for (int c = 0; c < ch; c++) {
            int* d = c == 0 ? r : c == 1 ? g : b;
            const int32_t* q = p + (uint64_t)c * S.n + i;
            if (full && aligned_to((uint64_t)q, 16)) {
                const int4 v = *(const int4*)q;
                d[0] = v.x, d[1] = v.y, d[2] = v.z, d[3] = v.w;
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) d[k] = i + k < S.n ? q[k] : 0;
            }
        }
    } else {
        const uint8_t* q = (const uint8_t*)S.src + i * (uint64_t)ch;
        uint32_t w[3] = {0, 0, 0};  // the 4 pixels' bytes, in order
        if (full && aligned_to((uint64_t)q, 4)) {
#pragma unroll
            for (int k = 0; k < ch; k++) w[k] = ((const uint32_t*)q)[k];
        } else {
            for (int k = 0; k < 4 * ch; k++)
                if (i * ch + k < S.n * ch) w[k >> 2] |= (uint32_t)q[k] << (8 * (k & 3));
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (ch == 1) {
                r[k] = (int)((w[0] >> (8 * k)) & 255u);
            } else {
                const int o = 3 * k;
                r[k] = (int)((w[o >> 2] >> (8 * (o & 3))) & 255u);
                g[k] = (int)((w[(o + 1) >> 2] >> (8 * ((o + 1) & 3))) & 255u);
                b[k] = (int)((w[(o + 2) >> 2] >> (8 * ((o + 2) & 3))) & 255u);
            }
        }
    }
    if (ch == 1) {
#pragma unroll
        for (int k = 0; k < 4; k++) g[k] = b[k] = r[k];
    }
    if (S.irrev) {
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = ch == 1 ? (float)(r[k] - 128) : ict_sample(S.comp, r[k], g[k], b[k]);
        float* d = (float*)S.dst + i;
        if (full && aligned_to((uint64_t)d, 16)) {
            *(float4*)d = make_float4(v[0], v[1], v[2], v[3]);
        } else {
            for (int k = 0; k < 4; k++)
                if (i + k < S.n) d[k] = v[k];
        }
    } else {
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = ch == 1 ? r[k] - 128 : rct_sample(S.comp, r[k], g[k], b[k]);
        int32_t* d = (int32_t*)S.dst + i;
        if (full && aligned_to((uint64_t)d, 16)) {
            *(int4*)d = make_int4(v[0], v[1], v[2], v[3]);
        } else {
            for (int k = 0; k < 4; k++)
                if (i + k < S.n) d[k] = v[k];
        }
    }
}
__global__ __launch_bounds__(256) void prepare_kernel(const PixSeg* __restrict__ segs, int nseg) {
    const PixSeg S = segs[find_pix(segs, nseg, blockIdx.x)];
    const uint64_t i = (uint64_t)(blockIdx.x - S.blk_start) * kPixPerBlock + 4u * threadIdx.x;
    if (i >= S.n) return;
    if (S.channels == 3) prepare4<3>(S, i);
    else prepare4<1>(S, i);
}

// Inverse ICT/RCT and round/clamp to 0..255 (scalar.rs:197-236; f32::round
// is half away from zero), for CH channels (1 or 3).
template <int CH>
__device__ __forceinline__ void finalize4(const PixSeg& S, uint64_t i) {
    const bool full = i + 4u <= S.n;
    // the component planes' 4 samples (f32 bits or i32)
    uint32_t x[CH][4];
    const uint64_t src[3] = {S.src, S.src1, S.src2};
#pragma unroll
    for (int c = 0; c < CH; c++) {
        const uint32_t* q = (const uint32_t*)src[c] + i;
        if (full && aligned_to((uint64_t)q, 16)) {
            const uint4 v = *(const uint4*)q;
            x[c][0] = v.x, x[c][1] = v.y, x[c][2] = v.z, x[c][3] = v.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) x[c][k] = i + k < S.n ? q[k] : 0u;
        }
    }
    int v[4][CH];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (S.irrev) {
            float f[CH];
            const float y = __uint_as_float(x[0][k]);
            if constexpr (CH == 3) {
                const float cb = __uint_as_float(x[1][k]), cr = __uint_as_float(x[2][k]);
                f[0] = __fadd_rn(y, __fmul_rn(1.402f, cr));
                f[1] = __fsub_rn(__fsub_rn(y, __fmul_rn(0.34413f, cb)), __fmul_rn(0.71414f, cr));
                f[2] = __fadd_rn(y, __fmul_rn(1.772f, cb));
            } else {
                f[0] = y;
            }
#pragma unroll
            for (int c = 0; c < CH; c++) v[k][c] = (int)fminf(fmaxf(roundf(__fadd_rn(f[c], 128.0f)), 0.0f), 255.0f);
        } else {
            const int y = (int)x[0][k];
            int t[CH];
            if constexpr (CH == 3) {
                const int db = (int)x[1][k], dr = (int)x[2][k];
                const int g = y - ((db + dr) >> 2);
                t[0] = dr + g;
                t[1] = g;
                t[2] = db + g;
            } else {
                t[0] = y;
            }
#pragma unroll
            for (int c = 0; c < CH; c++) v[k][c] = min(max(t[c] + 128, 0), 255);
        }
    }
    if (S.dst) {  // u8 interleaved: 4 * CH bytes
        uint8_t* o = (uint8_t*)S.dst + i * CH;
        uint32_t w[CH];
#pragma unroll
        for (int k = 0; k < CH; k++) w[k] = 0;
#pragma unroll
        for (int k = 0; k < 4 * CH; k++) w[k >> 2] |= (uint32_t)v[k / CH][k % CH] << (8 * (k & 3));
        if (full && aligned_to((uint64_t)o, 4)) {
#pragma unroll
            for (int k = 0; k < CH; k++) ((uint32_t*)o)[k] = w[k];
        } else {
            for (int k = 0; k < 4 * CH; k++)
                if (i * CH + k < S.n * CH) o[k] = (uint8_t)(w[k >> 2] >> (8 * (k & 3)));
        }
    }
    if (S.dst_i32) {  // i32 planar
#pragma unroll
        for (int c = 0; c < CH; c++) {
            int32_t* o = (int32_t*)S.dst_i32 + (uint64_t)c * S.n + i;
            if (full && aligned_to((uint64_t)o, 16)) {
                *(int4*)o = make_int4(v[0][c], v[1][c], v[2][c], v[3][c]);
            } else {
                for (int k = 0; k < 4; k++)
                    if (i + k < S.n) o[k] = v[k][c];
            }
        }
    }
}
__global__ __launch_bounds__(256) void finalize_kernel(const PixSeg* __restrict__ segs, int nseg) {
    const PixSeg S = segs[find_pix(segs, nseg, blockIdx.x)];
    const uint64_t i = (uint64_t)(blockIdx.x - S.blk_start) * kPixPerBlock + 4u * threadIdx.x;
    if (i >= S.n) return;
    if (S.channels == 3) finalize4<3>(S, i);
    else finalize4<1>(S, i);
}

// ---- host launchers ---------------------------------------------------------------

void launch_fwd_level(bool irrev, const DwtSeg* segs, int nseg, uint32_t ntiles, hipStream_t st) {
    if (!ntiles) return;
    if (irrev) fwd_level_kernel<float, true><<<ntiles, kLevelThreads, 0, st>>>(segs, nseg);
    else fwd_level_kernel<int32_t, false><<<ntiles, kLevelThreads, 0, st>>>(segs, nseg);
}
void launch_inv_level(bool irrev, const DwtSeg* segs, int nseg, uint32_t ntiles, hipStream_t st) {
    if (!ntiles) return;
    if (irrev) inv_level_kernel<float, true><<<ntiles, kLevelThreads, 0, st>>>(segs, nseg);
    else inv_level_kernel<int32_t, false><<<ntiles, kLevelThreads, 0, st>>>(segs, nseg);
}
void launch_prepare(const PixSeg* segs, int nseg, uint32_t nblocks, hipStream_t st) {
    if (nblocks) prepare_kernel<<<nblocks, 256, 0, st>>>(segs, nseg);
}
void launch_finalize(const PixSeg* segs, int nseg, uint32_t nblocks, hipStream_t st) {
    if (nblocks) finalize_kernel<<<nblocks, 256, 0, st>>>(segs, nseg);
}

}  // namespace j2l
