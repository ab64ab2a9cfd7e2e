// Interface between the C-ABI layer (capi.cpp) and the kernel launchers in
// the .hip translation units.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>

#include "plan.hpp"

namespace ndc {

// LDS bytes kept free for the kernels' static __shared__ arrays when sizing
// dynamic LDS (the two share one per-workgroup limit).
constexpr size_t kStaticLdsReserve = 1024;

// Services a launcher needs from the handle that owns the call.
class LaunchContext {
public:
    virtual ~LaunchContext() = default;
    virtual hipStream_t stream() = 0;
    // Device scratch of at least `bytes`, 256-byte aligned, valid for the
    // duration of the current call (slot 0..3 are independent buffers).
    virtual void* scratch(int slot, size_t bytes) = 0;
    // Device table of W_n^i = exp(-2*pi*i*I/n), i < n, generated in FP64 on
    // the host and stored as float2 (dbl = false) or double2 (dbl = true).
    virtual const void* twiddles(int64_t n, bool dbl) = 0;
    // Device copy of a constant table, uploaded once per key and kept for the
    // handle's lifetime.
    virtual const void* constant_table(const std::string& key, const void* data, size_t bytes) = 0;
    // Per-handle host memo: a byte string kept under `key` for the handle's
    // lifetime (empty the first time), for host-side plans built once.
    virtual std::string& memo(const std::string& key) = 0;
    // Device copy of the plan's border maps, concatenated per axis.
    // offsets[i] is the element offset of axis i's map; the entry at
    // offsets[kMaxRank] is a single 0 used for virtual unit axes.
    virtual const int32_t* border_maps(const Plan& p, int64_t offsets[kMaxRank + 1]) = 0;
    virtual int cu_count() = 0;
    // Largest dynamic LDS allocation one workgroup may request.
    virtual size_t max_lds() = 0;
    // Bracket one kernel launch for the optional per-kernel profile: `name`
    // identifies the kernel, `bytes` is its algorithmic HBM traffic (compulsory
    // reads + writes of that pass) and `flops` its algorithmic arithmetic.
    // The launch between them is made with NDC_LAUNCH, which asks kevents for
    // the bracket's start/stop events (false: profile off, already taken, or
    // the stream is being captured).
    virtual void kbegin(const char* name, double bytes, double flops) { (void)name; (void)bytes; (void)flops; }
    virtual bool kevents(hipEvent_t* start, hipEvent_t* stop) { (void)start; (void)stop; return false; }
    virtual void kend() {}
    // A dry run plans every launch of a call without a device: NDC_LAUNCH
    // hands the launch to record_launch instead of the runtime, and device
    // memory requests return placeholders that are never dereferenced.
    virtual bool dry_run() { return false; }
    virtual void record_launch(const char* kernel, size_t smem, unsigned grid, unsigned block) {
        (void)kernel; (void)smem; (void)grid; (void)block;
    }
    std::string err;
};

// Launches `kernel` (parenthesised when it names template arguments) on
// `stream`.  Inside a profiled kbegin/kend bracket the launch carries the
// bracket's events (hipExtLaunchKernelGGL), which then hold the dispatch's own
// start and end times rather than those of marker packets queued around it.
#define NDC_LAUNCH(ctx, stream, kernel, grid, block, smem, ...)                                           \
    do {                                                                                                  \
        hipEvent_t ndc_e0_ = nullptr, ndc_e1_ = nullptr;                                                  \
        if ((ctx).dry_run())                                                                              \
            (ctx).record_launch(#kernel, size_t(smem), dim3(grid).x, dim3(block).x);                      \
        else if ((ctx).kevents(&ndc_e0_, &ndc_e1_))                                                       \
            hipExtLaunchKernelGGL(kernel, grid, block, uint32_t(smem), stream, ndc_e0_, ndc_e1_, 0,       \
                                  __VA_ARGS__);                                                           \
        else                                                                                              \
            hipLaunchKernelGGL(kernel, grid, block, smem, stream, __VA_ARGS__);                           \
    } while (0)

// True when `kernel`'s static LDS plus `dynamic_bytes` fits one workgroup;
// a launch that would not fit is refused on the host instead of reaching the
// device.  The static size is looked up once per kernel.
inline bool lds_fits(LaunchContext& ctx, const void* kernel, size_t dynamic_bytes) {
    if (ctx.dry_run()) {
        if (dynamic_bytes <= ctx.max_lds()) return true;
        ctx.err = "kernel LDS request exceeds the workgroup limit";
        return false;
    }
    thread_local std::unordered_map<const void*, size_t> cache;
    auto it = cache.find(kernel);
    if (it == cache.end()) {
        hipFuncAttributes attr;
        if (hipFuncGetAttributes(&attr, kernel) != hipSuccess) return false;
        it = cache.emplace(kernel, size_t(attr.sharedSizeBytes)).first;
    }
    if (it->second + dynamic_bytes > ctx.max_lds() + kStaticLdsReserve) {
        ctx.err = "kernel LDS request exceeds the workgroup limit";
        return false;
    }
    return true;
}

// Direct convolution (conv/mod.rs:128-200) in the element type.
ndc_status launch_direct(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx);

// Direct 2-D f32 convolution on the matrix cores in BF16 or split BF16x3
// (p.algo), FP32 accumulation.
ndc_status launch_direct_mfma(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx);

// A kernel spectrum kept across calls (the form the reference sketches as
// "Baked", conv_fft/mod.rs:22-42,306-392): with bake set, launch_fft only
// computes the kernel's spectrum into `spec`; otherwise it reads it from there
// and transforms only the data.
struct KernelSpectrum {
    void* spec = nullptr;
    bool bake = false;
};

// FFT convolution (conv_fft/mod.rs:185-292).
ndc_status launch_fft(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx,
                      KernelSpectrum* ks = nullptr);

// Bytes of a kept kernel spectrum for the plan, or 0 when its rows are too
// long for one workgroup (those plans recompute the kernel spectrum per call).
size_t fft_kernel_spectrum_bytes(const Plan& p, size_t lds);

// Upper bound of the device memory launch_fft (without a kept spectrum)
// requests through LaunchContext::scratch, twiddles and constant_table for
// the plan, alignment included.
size_t fft_workspace_bytes(const Plan& p, size_t lds);

// The processor's plain N-d transforms (processor/real.rs, complex.rs) on
// contiguous device arrays.  dtype F32/F64: real input (forward) or output
// (backward); C64/C128: complex.  Forward: `shape` is the input's, the output
// is the spectrum with axes rotated left by one.  Backward: `shape` is the
// spectrum's; origin_len is the real last-axis length.
ndc_status launch_fft_processor(int dtype, bool forward, int N, const int64_t* shape, int64_t origin_len,
                                const void* in, void* out, LaunchContext& ctx);

}  // namespace ndc
