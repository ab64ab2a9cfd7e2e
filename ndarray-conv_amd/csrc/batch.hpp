// Batches of independent rank-1 signal/kernel pairs of uneven lengths, run
// as one persistent launch that drains a device-side work queue.
//
// Every pair has the semantics of one rank-1 conv_fft call
// (conv_fft/mod.rs:185-292): ConvMode and PaddingMode are shared by the
// batch, lengths differ per pair.  Each pair is computed by one of
//   - direct items: a tile of outputs, FP32 multiply-adds over the taps;
//   - overlap-save FFT items: blocks of outputs, each from one real FFT of
//     length L = 2 * LC (LC = 256 .. 4096 complex points) of a window of the
//     padded signal, multiplied by the pair's kernel spectrum (computed by a
//     pre-pass at the same L) and transformed back.  Overlap-save computes the
//     same linear correlation as the reference's single FFT of the whole padded
//     signal, only blocked: window position u >= kdil - 1 of a block holds
//     full-correlation index block_start + u.
// The choice per pair comes from a cost model whose constants are measured
// on the device (the autotuner).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "launch.hpp"
#include "ndconv.h"

namespace ndc {

constexpr int kBatchThreads = 256;
constexpr int kBatchClasses = 5;             // LC = 256 << c, c = 0..4
constexpr int kBatchPoints = 4096;           // complex points per FFT item (G lines of LC)
constexpr int kBatchDirectTile = 2048;       // outputs per direct item (8 per thread)
constexpr int kBatchWindow = 6144;           // LDS window capacity of a direct item (floats)
constexpr int kBatchTapChunk = 1024;         // taps staged per round of a direct item
constexpr int kBatchMfmaTile = 4096;         // outputs per matrix-core item (1024 per wave)
constexpr int kBatchMfmaMaxKdil = 1024;      // longest dilated kernel of a matrix-core item (LDS)
constexpr int kBatchKindMfma = 1 + kBatchClasses;    // BF16x3 matrix-core tile (two-term split)
constexpr int kBatchKindMfma6 = 2 + kBatchClasses;   // BF16x6 matrix-core tile (three-term split)
// Fewest kernel taps for which the autotuned and matrix-core batch forms use
// BF16x3.  Its products carry about 16 significant bits (each within
// 3 * 2^-16 of the exact product); an output summed from a few products keeps
// that error, while with 32 or more independent products per output the
// rounding errors average to about 2^-18 of the output scale (normwise 4.7e-6
// worst in an exact-accumulation model of Uniform[-1,1) data over 1.2 M
// outputs).  Shorter kernels use the three-term split, whose products are
// within 5 * 2^-24.
constexpr int kBf16x3MinTaps = 32;

// The work queue is two ranges of the item list, each longest-first: the
// direct and FFT items [0, nfirst), then the matrix-core items [nfirst,
// nitems).  Workgroups below nfirst drain the first range, then all drain
// the second.  Each
// range is kBatchQueues sub-queues: sub-queue q of a range starting at lo
// holds items lo + q, lo + q + kBatchQueues, ...  One head per sub-queue,
// kBatchQueueStride ints (256 bytes) apart, the second range's heads after
// the first's: a single head shared by every workgroup serializes the pops
// (about 12 ns each).
constexpr int kBatchQueues = 8;
constexpr int kBatchQueueStride = 64;
constexpr int kBatchHeads = 2 * kBatchQueues;

// ceil(2^31 / d), d >= 1: for 0 <= q < 2^20 and d <= 2^11, q / d is the high
// word of 2 q * batch_dmul(d) (the error term q (2^31 mod d) / (d 2^31) stays
// below 1 / d).
constexpr uint32_t batch_dmul(int d) { return uint32_t(((uint64_t(1) << 31) + uint64_t(d) - 1) / uint64_t(d)); }

constexpr int batch_lc(int cls) { return 256 << cls; }
constexpr int batch_lines(int cls) { return kBatchPoints / batch_lc(cls); }
// Quarter-circle twiddle tables of all classes, concatenated: class c has
// LC/2 = 128 << c entries.
constexpr int batch_tw_offset(int cls) { return 128 * ((1 << cls) - 1); }

// Per-pair metadata read by the kernels (64 bytes).
struct BatchPair {
    int64_t x_off, k_off, y_off;   // element offsets into the packed x, k, y
    int64_t ks_off;                // complex offset of the kernel spectrum (FFT pairs)
    int32_t n, k;                  // signal and kernel length
    int32_t p0, pds;               // front pad, padded length n + p0 + p1
    int32_t out;                   // output length
    int32_t per;                   // outputs per FFT block
    int32_t cls;                   // FFT class, or -1 for direct
    int32_t pad_;
};

// One unit of the work queue (64 bytes).
//   kind 0: direct tile  a = pair, b = first output, c = output count
//   kind 1 + cls: FFT    a = first line, b = line count (<= batch_lines(cls))
//   kind kBatchKindMfma / kBatchKindMfma6: BF16x3 / BF16x6 matrix-core tile,
//       a = pair, b = first output, c = count
// Direct and matrix-core tiles carry their pair's metadata, so a workgroup
// starts an item after one load (which it prefetches during the item before).
struct BatchItem {
    int32_t kind, a, b, c;
    int64_t x_off, k_off, y_off;   // the pair's element offsets into x, k, y
    int32_t n, k, p0, pds;         // the pair's signal and kernel lengths, front pad, padded length
    int32_t pad_[2];
};
static_assert(sizeof(BatchItem) == 64, "work-queue records are 16 dwords");

// One overlap-save block with its pair's metadata: first output o0, outputs
// per block, and what the pair's kernel spectrum and signal need.
struct BatchLine {
    int64_t x_off, y_off, ks_off;
    int32_t n, k, p0, pds, out, per, o0, pair;
};

// Kernel-spectrum pre-pass work: up to batch_lines(cls) pairs of one class,
// klines [first, first + count).
struct BatchKItem {
    int32_t cls, first, count, pad_;
};

// What the pre-pass reads of a pair, in pre-pass order (24 bytes, one load).
struct BatchKLine {
    int64_t k_off, ks_off;         // kernel offset into k, complex offset of its spectrum
    int32_t k, pad_;               // kernel length
};

// Shared launch parameters of a batch.
struct BatchArgs {
    const float* x;
    const float* k;
    float* y;
    const BatchPair* pairs;
    const BatchItem* items;
    const BatchLine* lines;
    const BatchKLine* klines;      // pairs of the kernel-spectrum pre-pass, grouped by class
    const BatchKItem* kitems;
    float2* kspec;
    int32_t* counter;              // kBatchHeads sub-queue heads (kBatchQueueStride apart), reset by the pre-pass
    const float2* tw;              // W_L^r, r < L/4, of class c at batch_tw_offset(c) (L = 2 LC)
    const float* consts;           // {0, cfront, cback, 0}: the cells constant pad positions read
    int32_t nitems, nkitems;
    int32_t nfirst;                // items of the first range (direct and FFT)
    int32_t s, d, rev;
    uint32_t dmul;                 // batch_dmul(d)
    int32_t front, back;           // side tags of border1d.hpp
    float cfront, cback;           // constants of Const sides
};

// Device launchers (batch_conv.hip).
size_t batch_lds_bytes();
int batch_blocks_per_cu();
// Pre-pass then persistent main launch on `stream`.  With `profile` each
// launch is bracketed for ctx's per-kernel profile (not while capturing);
// the byte/flop counts are the algorithmic traffic of each launch.
ndc_status launch_batch(const BatchArgs& a, int grid, LaunchContext& ctx, hipStream_t stream, bool profile,
                        double kspec_bytes, double main_bytes, double main_flops);

// Services of the handle behind an ndc_handle (capi.cpp).
LaunchContext* handle_context(ndc_handle h);
int handle_device(ndc_handle h);
void handle_set_last(ndc_handle h, const std::string& msg);

}  // namespace ndc
