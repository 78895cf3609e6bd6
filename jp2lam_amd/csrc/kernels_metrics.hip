// Quality metrics of encode_with_psnr (encode/mod.rs:114-135) from the
// coefficients still resident after an encode: the reference's simulated
// reconstruction (encode/backend/native/backend.rs:305-433) without
// re-running the pipeline.
//
//   metric_dequant_kernel  truncated midpoint dequantisation, in place in the
//                          Mallat plane (backend.rs:331-368, :626-639)
//   (inverse 9/7 levels)   inv_level_kernel, as in the decoder
//   metric_cells_kernel    per 8x8 cell: reconstructed pixels, squared
//                          error, SSIM of the luma (backend.rs:370-430, :909-959)
//   metric_reduce_kernel   per image: sum of squared errors, sum of SSIMs
//
// No FMA contraction (the build passes -ffp-contract=off): per-pixel f32
// values and per-cell f64 sums follow the reference's order, so each cell's
// SSIM equals the reference's; only the order of the image-wide sums differs.
#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace j2l {

// One workgroup per code-block of a metrics image: quantise (as the Tier-1
// load does), keep the bit-planes down to the last kept pass's, and store the
// midpoint reconstruction over the coefficient.  Pass k of a block with mbp
// magnitude bit-planes codes bit-plane mbp - 1 - (k + 2) / 3 (cleanup only at
// the top, then SP, MR, CL per plane: t1.rs:518-575).
__global__ __launch_bounds__(256) void metric_dequant_kernel(const BlockJob* __restrict__ jobs,
                                                             const BlockResult* __restrict__ res,
                                                             const Selection* __restrict__ sel,
                                                             const uint32_t* __restrict__ list) {
    const uint32_t b = list[blockIdx.x];
    const BlockJob J = jobs[b];
    const int keep = sel[b].keep, mbp = res[b].mbp;
    const int last_bp = keep ? mbp - 1 - (keep + 1) / 3 : -1;
    float* plane = (float*)J.plane;
    for (int i = threadIdx.x; i < 64 * J.h; i += 256) {
        const int x = i & 63, y = i >> 6;
        if (x >= J.w) continue;
        const size_t idx = (size_t)(J.y0 + y) * J.stride + J.x0 + x;
        float v = 0.0f;
        if (last_bp >= 0) {
            const int32_t q = quantize(plane[idx], J.step);
            const uint32_t aq = q < 0 ? (uint32_t)(-(int64_t)q) : (uint32_t)q;
            const uint32_t at = (aq >> last_bp) << last_bp;
            if (at) {
                const float m = __fmul_rn(__fadd_rn((float)at, 0.5f), J.step);
                v = q >= 0 ? m : -m;
            }
        }
        plane[idx] = v;
    }
}

namespace {

__device__ __forceinline__ float clamp255(float v) { return fminf(fmaxf(v, 0.0f), 255.0f); }

__device__ __forceinline__ int find_metric_seg(const MetricSeg* segs, int nseg, uint32_t cell) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].cell_start <= cell) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

// One thread per 8x8 cell: cells[2 * g] = squared error over the cell's
// samples, cells[2 * g + 1] = the cell's SSIM (0 for a partial cell).
__global__ __launch_bounds__(256) void metric_cells_kernel(const MetricSeg* __restrict__ segs, int nseg,
                                                           uint32_t ncells, double* __restrict__ cells) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ncells) return;
    const MetricSeg S = segs[find_metric_seg(segs, nseg, g)];
    const uint32_t k = g - S.cell_start;
    const uint32_t x0 = (k % S.cells_x) * 8, y0 = (k / S.cells_x) * 8;
    const uint32_t x1 = min(x0 + 8, S.width), y1 = min(y0 + 8, S.height);
    const bool full = x0 + 8 <= S.width && y0 + 8 <= S.height;
    const size_t n = (size_t)S.width * S.height;
    const float* r0 = (const float*)S.rec[0];
    const float* r1 = (const float*)S.rec[1];
    const float* r2 = (const float*)S.rec[2];
    const uint8_t* o8 = (const uint8_t*)S.orig;
    const int32_t* o32 = (const int32_t*)S.orig;
    double sse = 0.0, sx = 0.0, sy = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0;
    for (uint32_t y = y0; y < y1; y++)
        for (uint32_t x = x0; x < x1; x++) {
            const size_t i = (size_t)y * S.width + x;
            float ol, dl;
            if (S.ncomp == 1) {
                const float o = S.orig_i32 ? (float)o32[i] : (float)o8[i];
                const float d = clamp255(__fadd_rn(r0[i], 128.0f));
                const float diff = __fsub_rn(d, o);
                sse += (double)__fmul_rn(diff, diff);
                ol = o;
                dl = d;
            } else {
                float o[3];
                for (int c = 0; c < 3; c++) o[c] = S.orig_i32 ? (float)o32[(size_t)c * n + i] : (float)o8[3 * i + c];
                const float yc = r0[i], cb = r1[i], cr = r2[i];
                // inverse ICT (Annex G.2) and the DC shift, left to right
                const float r = clamp255(__fadd_rn(__fadd_rn(yc, __fmul_rn(1.402f, cr)), 128.0f));
                const float gg = clamp255(
                    __fadd_rn(__fsub_rn(__fsub_rn(yc, __fmul_rn(0.34413f, cb)), __fmul_rn(0.71414f, cr)), 128.0f));
                const float bb = clamp255(__fadd_rn(__fadd_rn(yc, __fmul_rn(1.772f, cb)), 128.0f));
                const float dr = __fsub_rn(r, o[0]), dg = __fsub_rn(gg, o[1]), db = __fsub_rn(bb, o[2]);
                sse += (double)__fadd_rn(__fadd_rn(__fmul_rn(dr, dr), __fmul_rn(dg, dg)), __fmul_rn(db, db));
                dl = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, r), __fmul_rn(0.587f, gg)), __fmul_rn(0.114f, bb));
                ol = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, o[0]), __fmul_rn(0.587f, o[1])), __fmul_rn(0.114f, o[2]));
            }
            const double xd = ol, yd = dl;
            sx = __dadd_rn(sx, xd);
            sy = __dadd_rn(sy, yd);
            sxx = __dadd_rn(sxx, __dmul_rn(xd, xd));
            syy = __dadd_rn(syy, __dmul_rn(yd, yd));
            sxy = __dadd_rn(sxy, __dmul_rn(xd, yd));
        }
    double ssim = 0.0;
    if (full) {
        const double C1 = 6.5025, C2 = 58.5225;
        const double ux = __ddiv_rn(sx, 64.0), uy = __ddiv_rn(sy, 64.0);
        const double vx = fmax(__dsub_rn(__ddiv_rn(sxx, 64.0), __dmul_rn(ux, ux)), 0.0);
        const double vy = fmax(__dsub_rn(__ddiv_rn(syy, 64.0), __dmul_rn(uy, uy)), 0.0);
        const double cxy = __dsub_rn(__ddiv_rn(sxy, 64.0), __dmul_rn(ux, uy));
        const double num = __dmul_rn(__dadd_rn(__dmul_rn(__dmul_rn(2.0, ux), uy), C1), __dadd_rn(__dmul_rn(2.0, cxy), C2));
        const double den =
            __dmul_rn(__dadd_rn(__dadd_rn(__dmul_rn(ux, ux), __dmul_rn(uy, uy)), C1), __dadd_rn(__dadd_rn(vx, vy), C2));
        ssim = __ddiv_rn(num, den);
    }
    cells[2 * (size_t)g] = sse;
    cells[2 * (size_t)g + 1] = ssim;
}

// One workgroup per image: out[2 * s] = sum of squared errors, out[2 * s + 1]
// = sum of the full cells' SSIMs, in a fixed order.
__global__ __launch_bounds__(256) void metric_reduce_kernel(const MetricSeg* __restrict__ segs,
                                                            const double* __restrict__ cells, double* __restrict__ out) {
    __shared__ double sa[256], sb[256];
    const MetricSeg S = segs[blockIdx.x];
    const uint32_t n = S.cells_x * S.cells_y;
    double a = 0.0, b = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += 256) {
        a += cells[2 * ((size_t)S.cell_start + k)];
        b += cells[2 * ((size_t)S.cell_start + k) + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sa[0];
        out[2 * blockIdx.x + 1] = sb[0];
    }
}

// Content analysis (content.rs:120-153): the 256-bin luma histogram and the
// sum of per-pixel chroma spreads max(r,g,b) - min(r,g,b), in one pass.
// out[0..255] = bins, out[256] = spread sum.  Gray reads the samples
// clamped to 0..255; the RGB family takes integer BT.601 luma
// (content.rs:110-117).  A workgroup sees at most ceil(n / (256 G)) * 256
// samples for G workgroups, so its LDS bins stay far below 2^32 (under 2^21
// for 65535 x 65535 samples and G = 4096).
__global__ __launch_bounds__(256) void content_hist_kernel(const void* __restrict__ src, uint64_t n, int channels,
                                                           int planar_i32, int gray,
                                                           unsigned long long* __restrict__ out) {
    __shared__ uint32_t bins[256];
    __shared__ unsigned long long spread;
    bins[threadIdx.x] = 0;
    if (threadIdx.x == 0) spread = 0;
    __syncthreads();
    const uint8_t* s8 = (const uint8_t*)src;
    const int32_t* s32 = (const int32_t*)src;
    unsigned long long mine = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        int bin;
        if (gray) {
            const int v = planar_i32 ? s32[i] : (int)s8[i * channels];
            bin = min(max(v, 0), 255);
        } else {
            int r, g, b;
            if (planar_i32) {
                r = s32[i];
                g = s32[n + i];
                b = s32[2 * n + i];
            } else {
                r = s8[3 * i];
                g = s8[3 * i + 1];
                b = s8[3 * i + 2];
            }
            const int rc = min(max(r, 0), 255), gc = min(max(g, 0), 255), bc = min(max(b, 0), 255);
            bin = min(max((77 * rc + 150 * gc + 29 * bc + 128) >> 8, 0), 255);
            const int mx = min(max(max(max(r, g), b), 0), 255), mn = min(max(min(min(r, g), b), 0), 255);
            mine += (unsigned long long)(mx - mn);
        }
        atomicAdd(&bins[bin], 1u);
    }
    if (mine) atomicAdd(&spread, mine);
    __syncthreads();
    if (bins[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
    if (threadIdx.x == 0 && spread) atomicAdd(&out[256], spread);
}

void launch_content_hist(const void* src, uint64_t n, int channels, int planar_i32, int gray, unsigned long long* out,
                         hipStream_t st) {
    if (!n) return;
    const uint64_t groups = std::min<uint64_t>(std::max<uint64_t>(n / 4096, 1), 4096);
    content_hist_kernel<<<(unsigned)groups, 256, 0, st>>>(src, n, channels, planar_i32, gray, out);
}

// ---- content analysis without allocation ----------------------------------------
// Scratch of content_hist_noalloc: the 256 luma bins and the chroma-spread
// sum of the last call.  Each device holds its own copy for the life of the
// process; calls that overlap on one device share it and race.
constexpr int kContentScratchWords = 257;
static __device__ __align__(256) unsigned long long scratch[kContentScratchWords];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxScratchDevices] = {};

__global__ __launch_bounds__(256) void zero_words_kernel(unsigned long long* __restrict__ p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

// Device address of `scratch` on the current device, looked up once per
// device.  The array's own name is a host-side stand-in and never reaches a
// kernel or a copy.
unsigned long long* scratch_on_current_device() {
    int dev = 0;
    J2L_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) throw Error(IO, "device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        J2L_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error(IO, "hipGetSymbolAddress returned no address for the content scratch");
        g_scratch_ptr[dev] = (unsigned long long*)ptr;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void content_hist_noalloc(const void* dev_src, uint64_t n, int channels, hipStream_t st, uint64_t* host_hist) {
    unsigned long long* buf = scratch_on_current_device();
    zero_words_kernel<<<(kContentScratchWords + 255) / 256, 256, 0, st>>>(buf, kContentScratchWords);
    J2L_HIP(hipGetLastError());
    launch_content_hist(dev_src, n, channels, 0, channels == 1, buf, st);
    J2L_HIP(hipGetLastError());
    J2L_HIP(hipMemcpyAsync(host_hist, buf, sizeof(unsigned long long) * kContentScratchWords, hipMemcpyDeviceToHost,
                           st));
    J2L_HIP(hipStreamSynchronize(st));
}

void launch_metric_dequant(const BlockJob* jobs, const BlockResult* res, const Selection* sel, const uint32_t* list,
                           int n, hipStream_t st) {
    if (n > 0) metric_dequant_kernel<<<n, 256, 0, st>>>(jobs, res, sel, list);
}
void launch_metric_cells(const MetricSeg* segs, int nseg, uint32_t ncells, double* cells, hipStream_t st) {
    if (nseg > 0 && ncells) metric_cells_kernel<<<(ncells + 255) / 256, 256, 0, st>>>(segs, nseg, ncells, cells);
}
void launch_metric_reduce(const MetricSeg* segs, int nseg, const double* cells, double* out, hipStream_t st) {
    if (nseg > 0) metric_reduce_kernel<<<nseg, 256, 0, st>>>(segs, cells, out);
}

}  // namespace j2l
