// Device-side definitions shared by the engine and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "types.hpp"

namespace j2l {

#define J2L_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::j2l::Error(::j2l::IO, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                              " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// Quantisation of a 9/7 coefficient: trunc(v / step) with IEEE division,
// non-finite -> 0 (simd/scalar.rs:207-213).
__device__ __forceinline__ int32_t quantize(float v, float step) {
    if (!(step > 0.0f) || !isfinite(v)) return 0;
    float q = truncf(__fdiv_rn(v, step));
    if (q >= 2147483648.0f) return 0x7fffffff;
    if (q < -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)q;
}

// ---- kernel launchers (kernels_*.hip) ----
void launch_fwd_level(bool irrev, const DwtSeg* segs, int nseg, uint32_t ntiles, hipStream_t st);
void launch_inv_level(bool irrev, const DwtSeg* segs, int nseg, uint32_t ntiles, hipStream_t st);
void launch_prepare(const PixSeg* segs, int nseg, uint32_t nblocks, hipStream_t st);
void launch_finalize(const PixSeg* segs, int nseg, uint32_t nblocks, hipStream_t st);
// The wave coder: a cost estimate and a sort of the blocks, then one wave per
// block in that order.  sched: t1_sched_words(n) words of device scratch.
size_t t1_sched_words(int n);
void launch_t1_encode(const BlockJob* jobs, int n, const uint32_t* tabs, uint8_t* out, BlockResult* res, uint32_t* snaps,
                      uint16_t* newsigs, uint32_t* sched, hipStream_t st);
void launch_t1_form(const BlockJob* jobs, int n, const uint32_t* tabs, BlockResult* res, uint32_t* pend, uint16_t* newsigs,
                    uint32_t* soff16, uint8_t* pool, unsigned long long* pool_used, unsigned long long pool_cap,
                    hipStream_t st);
void launch_t1_mq(const BlockJob* jobs, int n, const uint32_t* tabs, const uint8_t* pool, const uint32_t* soff16,
                  const uint32_t* pend, BlockResult* res, uint32_t* snaps, uint8_t* out, hipStream_t st);
// Long streams on the scalar coder, short ones on lanes (after launch_t1_form);
// lists holds 2 + 2n words.
void launch_t1_hybrid(const BlockJob* jobs, int n, const uint32_t* tabs, const uint8_t* pool, const uint32_t* soff16,
                      const uint32_t* pend, BlockResult* res, uint32_t* snaps, uint8_t* out, uint32_t* lists, float frac,
                      hipStream_t st);
uint64_t t1_stream_bound(uint32_t mbp);
// Zero n 64-bit words on the stream (a kernel, so it can sit in a graph).
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
// MQ decisions per pass of every block (counts: n * kMaxPasses words).
void launch_t1_count(const BlockJob* jobs, int n, const uint32_t* tabs, uint32_t* counts, hipStream_t st);
void launch_t1_decode(const DecodeJob* jobs, int n, const uint32_t* tabs, const uint8_t* payload, hipStream_t st);
void launch_pcrd(int n, bool lossless, int ngroups, const BlockResult* res, const uint32_t* snaps,
                 const uint16_t* newsigs, const PcrdJob* jobs, HullPt* hull, uint8_t* hull_n, unsigned long long* gmax,
                 const double* raw_floor, double* lambda, Selection* sel, hipStream_t st);
void launch_metric_dequant(const BlockJob* jobs, const BlockResult* res, const Selection* sel, const uint32_t* list,
                           int n, hipStream_t st);
void launch_metric_cells(const MetricSeg* segs, int nseg, uint32_t ncells, double* cells, hipStream_t st);
void launch_metric_reduce(const MetricSeg* segs, int nseg, const double* cells, double* out, hipStream_t st);
void launch_content_hist(const void* src, uint64_t n, int channels, int planar_i32, int gray, unsigned long long* out,
                         hipStream_t st);
// Content histogram of `n` interleaved u8 pixels in device memory into the
// file-scope scratch array of kernels_metrics.hip (no allocation), read back
// into host_hist[257]; synchronises `st`.
void content_hist_noalloc(const void* dev_src, uint64_t n, int channels, hipStream_t st, uint64_t* host_hist);
void launch_pack(int n, int nimg, const uint32_t* img_first, const uint32_t* blk_img, const uint64_t* pack_base,
                 const BlockJob* jobs, const Selection* sel, uint32_t* offs, uint32_t* totals, const uint8_t* slots,
                 uint8_t* packed, hipStream_t st, bool offsets_only);

}  // namespace j2l
