// Colour transforms, multi-level forward/inverse DWT (5/3 integer, 9/7
// float) and finalisation for CDNA4 (gfx950).
//
// Each DWT level is one launch of 2-D tile kernels: a workgroup loads a
// 64x64 output tile plus a 4-sample mirrored halo into LDS, runs the
// vertical then horizontal lifting passes there (inverse: horizontal then
// vertical) and writes the four subbands once.  Every level therefore reads
// and writes its active region exactly once from HBM.
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

// One lifting step over rows of one parity (vertical pass) for all columns
// in [c0, c1).  Rows in [1, nrows-2] of the given parity are updated.
template <bool IRREV, typename T>
__device__ __forceinline__ void vstep(T (*s)[TLP], int par, int step, int c0, int c1, int nrows) {
    const int ncols = c1 - c0;
    const int first = par ? 1 : 2;
    const int nr = (nrows - 1 - first + 1) / 2;  // rows first, first+2, ... <= nrows-2
    for (int idx = threadIdx.x; idx < nr * ncols; idx += blockDim.x) {
        int j = first + 2 * (idx / ncols);
        int c = c0 + idx % ncols;
        if constexpr (IRREV) {
            float cc = step == 0 ? kAlpha : step == 1 ? kBeta : step == 2 ? kGamma : kDelta;
            s[j][c] = lift(s[j][c], cc, s[j - 1][c], s[j + 1][c]);
        } else {
            if (step == 0) s[j][c] -= (s[j - 1][c] + s[j + 1][c]) >> 1;
            else s[j][c] += (s[j - 1][c] + s[j + 1][c] + 2) >> 2;
        }
    }
}

template <bool IRREV, typename T>
__device__ __forceinline__ void hstep(T (*s)[TLP], int par, int step, int r0, int r1, int ncols) {
    const int nrows = r1 - r0;
    const int first = par ? 1 : 2;
    const int nc = (ncols - 1 - first + 1) / 2;
    for (int idx = threadIdx.x; idx < nc * nrows; idx += blockDim.x) {
        int r = r0 + idx / nc;
        int j = first + 2 * (idx % nc);
        if constexpr (IRREV) {
            float cc = step == 0 ? kAlpha : step == 1 ? kBeta : step == 2 ? kGamma : kDelta;
            s[r][j] = lift(s[r][j], cc, s[r][j - 1], s[r][j + 1]);
        } else {
            if (step == 0) s[r][j] -= (s[r][j - 1] + s[r][j + 1]) >> 1;
            else s[r][j] += (s[r][j - 1] + s[r][j + 1] + 2) >> 2;
        }
    }
}

// Inverse steps: 9/7 uses -delta (even), -gamma (odd), -beta (even),
// -alpha (odd); 5/3 undoes update (even) then predict (odd).
template <bool IRREV, typename T>
__device__ __forceinline__ void vstep_inv(T (*s)[TLP], int par, int step, int c0, int c1, int nrows) {
    const int ncols = c1 - c0;
    const int first = par ? 1 : 2;
    const int nr = (nrows - 1 - first + 1) / 2;
    for (int idx = threadIdx.x; idx < nr * ncols; idx += blockDim.x) {
        int j = first + 2 * (idx / ncols);
        int c = c0 + idx % ncols;
        if constexpr (IRREV) {
            float cc = step == 0 ? -kDelta : step == 1 ? -kGamma : step == 2 ? -kBeta : -kAlpha;
            s[j][c] = lift(s[j][c], cc, s[j - 1][c], s[j + 1][c]);
        } else {
            if (step == 0) s[j][c] -= (s[j - 1][c] + s[j + 1][c] + 2) >> 2;
            else s[j][c] += (s[j - 1][c] + s[j + 1][c]) >> 1;
        }
    }
}

template <bool IRREV, typename T>
__device__ __forceinline__ void hstep_inv(T (*s)[TLP], int par, int step, int r0, int r1, int ncols) {
    const int nrows = r1 - r0;
    const int first = par ? 1 : 2;
    const int nc = (ncols - 1 - first + 1) / 2;
    for (int idx = threadIdx.x; idx < nc * nrows; idx += blockDim.x) {
        int r = r0 + idx / nc;
        int j = first + 2 * (idx % nc);
        if constexpr (IRREV) {
            float cc = step == 0 ? -kDelta : step == 1 ? -kGamma : step == 2 ? -kBeta : -kAlpha;
            s[r][j] = lift(s[r][j], cc, s[r][j - 1], s[r][j + 1]);
        } else {
            if (step == 0) s[r][j] -= (s[r][j - 1] + s[r][j + 1] + 2) >> 2;
            else s[r][j] += (s[r][j - 1] + s[r][j + 1]) >> 1;
        }
    }
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
__global__ __launch_bounds__(1024) void fwd_level_kernel(const DwtSeg* __restrict__ segs, int nseg) {
    __shared__ T s[TLH][TLP];
    const DwtSeg S = segs[find_seg(segs, nseg, blockIdx.x)];
    const uint32_t t = blockIdx.x - S.tile_start;
    const int tx0 = (int)(t % (uint32_t)S.tiles_x) * TX, ty0 = (int)(t / (uint32_t)S.tiles_x) * TY;
    const int rw = S.rw, rh = S.rh;
    const T* __restrict__ in = (const T*)S.in;
    T* __restrict__ plane = (T*)S.plane;
    T* __restrict__ ll_out = (T*)S.out;
    for (int idx = threadIdx.x; idx < TLH * TLW; idx += blockDim.x) {
        int ly = idx / TLW, lx = idx % TLW;
        int gy = sym(rh, ty0 - HALO + ly), gx = sym(rw, tx0 - HALO + lx);
        s[ly][lx] = in[(size_t)gy * S.in_stride + gx];
    }
    __syncthreads();
    // vertical pass (irrev97.rs:261-336; rev53.rs:419-497)
    if (rh >= 2) {
        if constexpr (IRREV) {
            for (int st = 0; st < 4; st++) {
                vstep<true>(s, (st & 1) ? 0 : 1, st, 0, TLW, TLH);
                __syncthreads();
            }
            for (int idx = threadIdx.x; idx < TLH * TLW; idx += blockDim.x) {
                int ly = idx / TLW, lx = idx % TLW;
                s[ly][lx] = __fmul_rn(s[ly][lx], (ly & 1) ? K97 : INV_K97);
            }
        } else {
            vstep<false>(s, 1, 0, 0, TLW, TLH);
            __syncthreads();
            vstep<false>(s, 0, 1, 0, TLW, TLH);
        }
    } else if constexpr (IRREV) {
        for (int lx = threadIdx.x; lx < TLW; lx += blockDim.x) s[HALO][lx] = __fmul_rn(s[HALO][lx], INV_K97);
    }
    __syncthreads();
    // horizontal pass on the output rows (irrev97.rs:92-129; rev53.rs:143-265)
    if (rw >= 2) {
        if constexpr (IRREV) {
            for (int st = 0; st < 4; st++) {
                hstep<true>(s, (st & 1) ? 0 : 1, st, HALO, HALO + TY, TLW);
                __syncthreads();
            }
            for (int idx = threadIdx.x; idx < TY * TLW; idx += blockDim.x) {
                int ly = HALO + idx / TLW, lx = idx % TLW;
                s[ly][lx] = __fmul_rn(s[ly][lx], (lx & 1) ? K97 : INV_K97);
            }
        } else {
            hstep<false>(s, 1, 0, HALO, HALO + TY, TLW);
            __syncthreads();
            hstep<false>(s, 0, 1, HALO, HALO + TY, TLW);
        }
    } else if constexpr (IRREV) {
        for (int ly = threadIdx.x; ly < TLH; ly += blockDim.x) s[ly][HALO] = __fmul_rn(s[ly][HALO], INV_K97);
    }
    __syncthreads();
    // HL and HH start at plane column S.hx, LH and HH at Mallat row snh
    const int snh = (rh + 1) >> 1, bhx = (int)S.hx;
    for (int idx = threadIdx.x; idx < TY * TX; idx += blockDim.x) {
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
__global__ __launch_bounds__(1024) void inv_level_kernel(const DwtSeg* __restrict__ segs, int nseg) {
    __shared__ T s[TLH][TLP];
    const DwtSeg S = segs[find_seg(segs, nseg, blockIdx.x)];
    const uint32_t t = blockIdx.x - S.tile_start;
    const int tx0 = (int)(t % (uint32_t)S.tiles_x) * TX, ty0 = (int)(t / (uint32_t)S.tiles_x) * TY;
    const int rw = S.rw, rh = S.rh;
    const T* __restrict__ ll_in = (const T*)S.in;
    const T* __restrict__ plane = (const T*)S.plane;
    T* __restrict__ out = (T*)S.out;
    const int snh = (rh + 1) >> 1, bhx = (int)S.hx;
    for (int idx = threadIdx.x; idx < TLH * TLW; idx += blockDim.x) {
        int ly = idx / TLW, lx = idx % TLW;
        int my = sym(rh, ty0 - HALO + ly), mx = sym(rw, tx0 - HALO + lx);
        int hy = my >> 1, hx = mx >> 1;
        T v;
        if (my & 1) v = plane[(size_t)(snh + hy) * S.pstride + ((mx & 1) ? bhx + hx : hx)];
        else if (mx & 1) v = plane[(size_t)hy * S.pstride + bhx + hx];
        else v = ll_in[(size_t)hy * S.in_stride + hx];
        s[ly][lx] = v;
    }
    __syncthreads();
    // horizontal inverse on every tile row (the vertical pass needs them)
    if (rw >= 2) {
        if constexpr (IRREV) {
            for (int idx = threadIdx.x; idx < TLH * TLW; idx += blockDim.x) {
                int ly = idx / TLW, lx = idx % TLW;
                s[ly][lx] = __fmul_rn(s[ly][lx], (lx & 1) ? INV_K97 : K97);
            }
            __syncthreads();
            for (int st = 0; st < 4; st++) {
                hstep_inv<true>(s, st & 1, st, 0, TLH, TLW);
                __syncthreads();
            }
        } else {
            hstep_inv<false>(s, 0, 0, 0, TLH, TLW);
            __syncthreads();
            hstep_inv<false>(s, 1, 1, 0, TLH, TLW);
            __syncthreads();
        }
    } else if constexpr (IRREV) {
        for (int ly = threadIdx.x; ly < TLH; ly += blockDim.x) s[ly][HALO] = __fmul_rn(s[ly][HALO], K97);
        __syncthreads();
    }
    // vertical inverse on the output columns
    if (rh >= 2) {
        if constexpr (IRREV) {
            for (int idx = threadIdx.x; idx < TLH * TX; idx += blockDim.x) {
                int ly = idx / TX, lx = HALO + idx % TX;
                s[ly][lx] = __fmul_rn(s[ly][lx], (ly & 1) ? INV_K97 : K97);
            }
            __syncthreads();
            for (int st = 0; st < 4; st++) {
                vstep_inv<true>(s, st & 1, st, HALO, HALO + TX, TLH);
                __syncthreads();
            }
        } else {
            vstep_inv<false>(s, 0, 0, HALO, HALO + TX, TLH);
            __syncthreads();
            vstep_inv<false>(s, 1, 1, HALO, HALO + TX, TLH);
            __syncthreads();
        }
    } else if constexpr (IRREV) {
        for (int lx = threadIdx.x; lx < TLW; lx += blockDim.x) s[HALO][lx] = __fmul_rn(s[HALO][lx], K97);
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < TY * TX; idx += blockDim.x) {
        int oy = idx / TX, ox = idx % TX;
        int gy = ty0 + oy, gx = tx0 + ox;
        if (gy >= rh || gx >= rw) continue;
        out[(size_t)gy * S.out_stride + gx] = s[HALO + oy][HALO + ox];
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

// Level shift (scalar.rs:242-254), forward ICT (scalar.rs:256-278) and
// forward RCT (scalar.rs:280-302) into a component plane.
__global__ __launch_bounds__(256) void prepare_kernel(const PixSeg* __restrict__ segs, int nseg) {
    const PixSeg S = segs[find_pix(segs, nseg, blockIdx.x)];
    const size_t i0 = (size_t)(blockIdx.x - S.blk_start) * kPixPerBlock + threadIdx.x;
    for (uint32_t c256 = 0; c256 < kPixPerBlock / 256u; c256++) {
    const size_t i = i0 + (size_t)c256 * 256u;  // each pass a coalesced 256
    if (i >= S.n) return;
    int r, g, b;
    if (S.planar_i32) {
        const int32_t* p = (const int32_t*)S.src;
        if (S.channels == 1) {
            r = g = b = p[i];
        } else {
            r = p[i];
            g = p[S.n + i];
            b = p[2 * S.n + i];
        }
    } else {
        const uint8_t* p = (const uint8_t*)S.src;
        if (S.channels == 1) {
            r = g = b = p[i];
        } else {
            r = p[3 * i];
            g = p[3 * i + 1];
            b = p[3 * i + 2];
        }
    }
    if (S.irrev) {
        float v;
        if (S.channels == 1) {
            v = (float)(r - 128);
        } else {
            float rf = (float)(r - 128), gf = (float)(g - 128), bf = (float)(b - 128);
            if (S.comp == 0) v = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, rf), __fmul_rn(0.587f, gf)), __fmul_rn(0.114f, bf));
            else if (S.comp == 1)
                v = __fadd_rn(__fadd_rn(__fmul_rn(-0.16875f, rf), __fmul_rn(-0.33126f, gf)), __fmul_rn(0.5f, bf));
            else v = __fsub_rn(__fsub_rn(__fmul_rn(0.5f, rf), __fmul_rn(0.41869f, gf)), __fmul_rn(0.08131f, bf));
        }
        ((float*)S.dst)[i] = v;
    } else {
        int v;
        if (S.channels == 1) {
            v = r - 128;
        } else {
            int rv = r - 128, gv = g - 128, bv = b - 128;
            v = S.comp == 0 ? ((rv + 2 * gv + bv) >> 2) : S.comp == 1 ? (bv - gv) : (rv - gv);
        }
        ((int32_t*)S.dst)[i] = v;
    }
    }
}

// Inverse ICT/RCT and round/clamp to 0..255 (scalar.rs:197-236; f32::round
// is half away from zero).
__global__ __launch_bounds__(256) void finalize_kernel(const PixSeg* __restrict__ segs, int nseg) {
    const PixSeg S = segs[find_pix(segs, nseg, blockIdx.x)];
    const size_t i0 = (size_t)(blockIdx.x - S.blk_start) * kPixPerBlock + threadIdx.x;
    for (uint32_t c256 = 0; c256 < kPixPerBlock / 256u; c256++) {
    const size_t i = i0 + (size_t)c256 * 256u;  // each pass a coalesced 256
    if (i >= S.n) return;
    const int ch = S.channels;
    int v[3];
    if (S.irrev) {
        float f[3];
        float y = ((const float*)S.src)[i];
        if (ch == 3) {
            float cb = ((const float*)S.src1)[i], cr = ((const float*)S.src2)[i];
            f[0] = __fadd_rn(y, __fmul_rn(1.402f, cr));
            f[1] = __fsub_rn(__fsub_rn(y, __fmul_rn(0.34413f, cb)), __fmul_rn(0.71414f, cr));
            f[2] = __fadd_rn(y, __fmul_rn(1.772f, cb));
        } else {
            f[0] = y;
        }
        for (int c = 0; c < ch; c++) v[c] = (int)fminf(fmaxf(roundf(__fadd_rn(f[c], 128.0f)), 0.0f), 255.0f);
    } else {
        int y = ((const int32_t*)S.src)[i];
        int t[3];
        if (ch == 3) {
            int db = ((const int32_t*)S.src1)[i], dr = ((const int32_t*)S.src2)[i];
            int g = y - ((db + dr) >> 2);
            t[0] = dr + g;
            t[1] = g;
            t[2] = db + g;
        } else {
            t[0] = y;
        }
        for (int c = 0; c < ch; c++) v[c] = min(max(t[c] + 128, 0), 255);
    }
    uint8_t* o8 = (uint8_t*)S.dst;
    int32_t* o32 = (int32_t*)S.dst_i32;
    for (int c = 0; c < ch; c++) {
        if (o8) o8[i * ch + c] = (uint8_t)v[c];
        if (o32) o32[(size_t)c * S.n + i] = v[c];
    }
    }
}

// ---- host launchers ---------------------------------------------------------------
// A level with few tiles lasts about as long as one tile's chain of lifting
// steps and barriers, so its tiles get 1024 threads, which makes each
// thread's share of a step 4x smaller.  Levels of 1024 tiles or more keep
// 256 threads: they are throughput-bound (a 16-page batch's DWT measured
// 1.24 ms with 256 there, 1.30 ms with 1024).
inline int level_threads(uint32_t ntiles) { return ntiles < 1024 ? 1024 : 256; }
void launch_fwd_level(bool irrev, const DwtSeg* segs, int nseg, uint32_t ntiles, hipStream_t st) {
    if (!ntiles) return;
    const int nt = level_threads(ntiles);
    if (irrev) fwd_level_kernel<float, true><<<ntiles, nt, 0, st>>>(segs, nseg);
    else fwd_level_kernel<int32_t, false><<<ntiles, nt, 0, st>>>(segs, nseg);
}
void launch_inv_level(bool irrev, const DwtSeg* segs, int nseg, uint32_t ntiles, hipStream_t st) {
    if (!ntiles) return;
    const int nt = level_threads(ntiles);
    if (irrev) inv_level_kernel<float, true><<<ntiles, nt, 0, st>>>(segs, nseg);
    else inv_level_kernel<int32_t, false><<<ntiles, nt, 0, st>>>(segs, nseg);
}
void launch_prepare(const PixSeg* segs, int nseg, uint32_t nblocks, hipStream_t st) {
    if (nblocks) prepare_kernel<<<nblocks, 256, 0, st>>>(segs, nseg);
}
void launch_finalize(const PixSeg* segs, int nseg, uint32_t nblocks, hipStream_t st) {
    if (nblocks) finalize_kernel<<<nblocks, 256, 0, st>>>(segs, nseg);
}

}  // namespace j2l
