// extern "C" entry points declared in include/ndconv.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "launch.hpp"
#include "ndconv.h"
#include "plan.hpp"

namespace ndc {
namespace {

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

class Handle final : public LaunchContext {
public:
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t ext = nullptr;
    bool use_ext = false;
    int cus = 256;
    size_t lds = 64 * 1024;
    DevBuf slots[4];
    DevBuf io[3];
    std::map<std::pair<int64_t, bool>, DevBuf> tw;
    std::unordered_map<std::string, std::pair<DevBuf, std::vector<int64_t>>> maps;
    std::unordered_map<std::string, Plan> plans;   // keyed by the raw descriptor bytes and op
    std::string last;
    int64_t rp_origin_len = 0;   // real length of the last axis of the last real forward transform
    hipEvent_t order_ev = nullptr;   // orders a newly selected stream after the previous one

    // Makes `next` the stream of later calls.  The handle's workspace (spectra,
    // io buffers, batch queues) is reused by every call, so work already queued
    // on the stream being left must finish before `next` runs anything: an
    // event recorded on the old stream is waited for by the new one.
    ndc_status switch_stream(hipStream_t next, bool external) {
        hipStream_t prev = stream();
        if (prev != next || use_ext != external) {
            if (!order_ev && hipEventCreateWithFlags(&order_ev, hipEventDisableTiming) != hipSuccess) {
                order_ev = nullptr;
                return NDC_ERR_DEVICE;
            }
            if (hipEventRecord(order_ev, prev) != hipSuccess || hipStreamWaitEvent(next, order_ev, 0) != hipSuccess)
                return NDC_ERR_DEVICE;
        }
        ext = external ? next : nullptr;
        use_ext = external;
        return NDC_OK;
    }

    // per-kernel profile: HIP events recorded on the handle's stream around
    // every launch while enabled
    struct Rec {
        std::string name;
        double bytes, flops;
        hipEvent_t a, b;
    };
    bool prof = false;
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;

    ~Handle() override {
        for (auto& s : slots) if (s.ptr) (void)hipFree(s.ptr);
        for (auto& s : io) if (s.ptr) (void)hipFree(s.ptr);
        for (auto& t : tw) (void)hipFree(t.second.ptr);
        for (auto& m : maps) (void)hipFree(m.second.first.ptr);
        for (auto& t : tables) (void)hipFree(t.second.ptr);
        for (auto& r : recs) { pool.push_back(r.a); pool.push_back(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
        if (order_ev) (void)hipEventDestroy(order_ev);
        if (own) (void)hipStreamDestroy(own);
    }

    hipEvent_t event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }

    // A bracket's events are handed to its launch (kevents), which records
    // them at the dispatch's start and end; a bracket whose launch did not
    // take them is dropped at kend.
    bool armed = false;

    void kbegin(const char* name, double bytes, double flops) override {
        armed = false;
        if (!prof) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream(), &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
        Rec r{name, bytes, flops, event(), event()};
        if (!r.a || !r.b) {
            if (r.a) pool.push_back(r.a);
            if (r.b) pool.push_back(r.b);
            return;
        }
        recs.push_back(r);
        armed = true;
    }

    bool kevents(hipEvent_t* a, hipEvent_t* b) override {
        if (!armed) return false;
        armed = false;
        *a = recs.back().a;
        *b = recs.back().b;
        return true;
    }

    void kend() override {
        if (!armed) return;
        armed = false;
        pool.push_back(recs.back().a);
        pool.push_back(recs.back().b);
        recs.pop_back();
    }

    hipStream_t stream() override { return use_ext ? ext : own; }
    int cu_count() override { return cus; }
    size_t max_lds() override { return lds; }

    void* grow(DevBuf& b, size_t bytes) {
        if (b.bytes >= bytes && b.ptr) return b.ptr;
        if (b.ptr) {
            (void)hipStreamSynchronize(stream());
            (void)hipFree(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes < 256 ? 256 : bytes;
        if (hipMalloc(&b.ptr, want) != hipSuccess) {
            b.ptr = nullptr;
            err = "hipMalloc failed";
            return nullptr;
        }
        b.bytes = want;
        return b.ptr;
    }

    void* scratch(int slot, size_t bytes) override {
        if (slot < 0 || slot >= 4) return nullptr;
        return grow(slots[slot], bytes);
    }

    const void* twiddles(int64_t n, bool dbl) override {
        auto key = std::make_pair(n, dbl);
        auto it = tw.find(key);
        if (it != tw.end()) return it->second.ptr;
        const size_t esz = dbl ? 16 : 8;
        std::vector<unsigned char> host(size_t(n) * esz);
        for (int64_t i = 0; i < n; ++i) {
            // exp(-2*pi*i*I/n) in FP64
            const double a = -2.0 * M_PI * double(i) / double(n);
            const double c = std::cos(a), s = std::sin(a);
            if (dbl) {
                double v[2] = {c, s};
                std::memcpy(&host[size_t(i) * esz], v, esz);
            } else {
                float v[2] = {float(c), float(s)};
                std::memcpy(&host[size_t(i) * esz], v, esz);
            }
        }
        DevBuf b;
        if (hipMalloc(&b.ptr, host.size()) != hipSuccess) return nullptr;
        b.bytes = host.size();
        if (hipMemcpy(b.ptr, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(b.ptr);
            return nullptr;
        }
        tw[key] = b;
        return b.ptr;
    }

    std::unordered_map<std::string, DevBuf> tables;
    std::unordered_map<std::string, std::string> memos;

    std::string& memo(const std::string& key) override { return memos[key]; }

    const void* constant_table(const std::string& key, const void* data, size_t bytes) override {
        auto it = tables.find(key);
        if (it != tables.end()) return it->second.ptr;
        DevBuf b;
        if (hipMalloc(&b.ptr, bytes ? bytes : 16) != hipSuccess) return nullptr;
        b.bytes = bytes;
        if (bytes && hipMemcpy(b.ptr, data, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(b.ptr);
            return nullptr;
        }
        tables[key] = b;
        return b.ptr;
    }

    const int32_t* border_maps(const Plan& p, int64_t offsets[kMaxRank + 1]) override {
        const std::string key = plan_key(p);
        auto it = maps.find(key);
        if (it == maps.end()) {
            std::vector<int32_t> blob;
            std::vector<int64_t> offs(kMaxRank + 1, 0);
            for (int i = 0; i < p.N; ++i) {
                offs[i] = int64_t(blob.size());
                blob.insert(blob.end(), p.map[i].begin(), p.map[i].end());
            }
            offs[kMaxRank] = int64_t(blob.size());
            blob.push_back(0);
            DevBuf b;
            if (hipMalloc(&b.ptr, blob.size() * 4) != hipSuccess) return nullptr;
            b.bytes = blob.size() * 4;
            if (hipMemcpy(b.ptr, blob.data(), b.bytes, hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(b.ptr);
                return nullptr;
            }
            it = maps.emplace(key, std::make_pair(b, offs)).first;
        }
        for (int i = 0; i <= kMaxRank; ++i) offsets[i] = it->second.second[i];
        return static_cast<const int32_t*>(it->second.first.ptr);
    }
};

// Plans a call's launches on the host without a device (LaunchContext::dry_run):
// device memory requests get placeholders, launches are recorded.
class DryRunContext final : public LaunchContext {
public:
    DryRunContext(size_t lds, int cus) : lds_(lds), cus_(cus) {}
    struct Launch {
        std::string kernel;
        size_t smem;
        unsigned grid, block;
    };
    std::vector<Launch> launches;

    hipStream_t stream() override { return nullptr; }
    void* scratch(int slot, size_t) override { return placeholder(slot); }
    const void* twiddles(int64_t, bool) override { return placeholder(8); }
    const void* constant_table(const std::string&, const void*, size_t) override { return placeholder(9); }
    std::string& memo(const std::string& key) override { return memos_[key]; }
    const int32_t* border_maps(const Plan&, int64_t offsets[kMaxRank + 1]) override {
        for (int i = 0; i <= kMaxRank; ++i) offsets[i] = 0;
        return static_cast<const int32_t*>(placeholder(10));
    }
    int cu_count() override { return cus_; }
    size_t max_lds() override { return lds_; }
    bool dry_run() override { return true; }
    void record_launch(const char* kernel, size_t smem, unsigned grid, unsigned block) override {
        launches.push_back(Launch{kernel, smem, grid, block});
    }

private:
    // distinct, aligned, never dereferenced
    static void* placeholder(int i) { return reinterpret_cast<void*>(uintptr_t(i + 1) << 20); }
    size_t lds_;
    int cus_;
    std::unordered_map<std::string, std::string> memos_;
};

ndc_status set_err(ndc_handle h, ndc_status s, const std::string& msg) {
    if (h) reinterpret_cast<Handle*>(h)->last = msg;
    return s;
}

ndc_status run(ndc_handle hh, const ndc_conv_desc* d, int op, const void* x, const void* k, void* y) {
    if (!hh) return NDC_ERR_INVALID_ARGUMENT;
    Handle& h = *reinterpret_cast<Handle*>(hh);
    if (!d) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "null descriptor");
    std::string key(reinterpret_cast<const char*>(d), sizeof(*d));
    key.append(reinterpret_cast<const char*>(&op), sizeof(op));
    auto pit = h.plans.find(key);
    if (pit == h.plans.end()) {
        Plan fresh;
        std::string err;
        ndc_status s = make_plan(d, op, &fresh, &err, true);
        if (s != NDC_OK) return set_err(hh, s, err);
        if (h.plans.size() >= 256) h.plans.clear();
        pit = h.plans.emplace(key, std::move(fresh)).first;
    }
    const Plan& p = pit->second;
    ndc_status s = NDC_OK;
    if (!x || !k || !y) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "null data pointer");
    const bool mfma = op == NDC_OP_CONV && (p.algo == NDC_ALGO_MFMA_BF16 || p.algo == NDC_ALGO_MFMA_BF16X3);
    if (op == NDC_OP_CONV && !mfma && p.algo != NDC_ALGO_AUTO && p.algo != NDC_ALGO_DIRECT)
        return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "unknown algorithm");
    if (hipSetDevice(h.device) != hipSuccess) return set_err(hh, NDC_ERR_DEVICE, "hipSetDevice failed");

    const size_t es = dtype_size(p.dtype);
    const bool host = (p.flags & NDC_FLAG_HOST_MEMORY) != 0;
    const void* dx = x;
    const void* dk = k;
    void* dy = y;
    if (host) {
        // copy the spans the strided views cover; device pointers keep the
        // same element offsets relative to the first logical element
        const size_t xb = size_t(p.x_hi - p.x_lo + 1) * es;
        const size_t kb = size_t(p.k_hi - p.k_lo + 1) * es;
        const size_t yb = size_t(p.y_numel) * es;
        char* bx = static_cast<char*>(h.grow(h.io[0], xb));
        char* bk = static_cast<char*>(h.grow(h.io[1], kb));
        char* by = static_cast<char*>(h.grow(h.io[2], yb ? yb : es));
        if (!bx || !bk || !by) return set_err(hh, NDC_ERR_OUT_OF_MEMORY, "device buffer allocation failed");
        const char* hx = static_cast<const char*>(x) + p.x_lo * int64_t(es);
        const char* hk = static_cast<const char*>(k) + p.k_lo * int64_t(es);
        if (hipMemcpyAsync(bx, hx, xb, hipMemcpyHostToDevice, h.stream()) != hipSuccess ||
            hipMemcpyAsync(bk, hk, kb, hipMemcpyHostToDevice, h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "host to device copy failed");
        dx = bx - p.x_lo * int64_t(es);
        dk = bk - p.k_lo * int64_t(es);
        dy = by;
    }
    h.err.clear();
    if (p.y_numel > 0) {
        if (op == NDC_OP_CONV_FFT) s = launch_fft(p, dx, dk, dy, h);
        else if (mfma) s = launch_direct_mfma(p, dx, dk, dy, h);
        else s = launch_direct(p, dx, dk, dy, h);
        if (s != NDC_OK) return set_err(hh, s, h.err);
    }
    if (host) {
        if (hipMemcpyAsync(y, dy, size_t(p.y_numel) * es, hipMemcpyDeviceToHost, h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "device to host copy failed");
        if (hipStreamSynchronize(h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "stream synchronize failed");
    }
    return NDC_OK;
}

// A kept kernel spectrum (the reference's sketched Baked form,
// conv_fft/mod.rs:22-42,306-392): the plan of one data shape and one kernel.
struct Baked {
    ndc_handle h = nullptr;
    Plan p;
    DevBuf spec;    // the kernel's spectrum, or empty when the rows are too long to keep one
    DevBuf kcopy;   // the kernel itself for those plans (same element offsets as the view)
    int64_t k_lo = 0;
    ~Baked() {
        if (spec.ptr) (void)hipFree(spec.ptr);
        if (kcopy.ptr) (void)hipFree(kcopy.ptr);
    }
};

ndc_status bake(ndc_handle hh, const ndc_conv_desc* d, const void* k, Baked** out) {
    if (!hh || !d || !out) return NDC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    Handle& h = *reinterpret_cast<Handle*>(hh);
    auto b = std::make_unique<Baked>();
    b->h = hh;
    std::string err;
    ndc_status s = make_plan(d, NDC_OP_CONV_FFT, &b->p, &err, true);
    if (s != NDC_OK) return set_err(hh, s, err);
    if (!k) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "null kernel pointer");
    if (hipSetDevice(h.device) != hipSuccess) return set_err(hh, NDC_ERR_DEVICE, "hipSetDevice failed");
    const Plan& p = b->p;
    const size_t es = dtype_size(p.dtype);
    const bool host = (p.flags & NDC_FLAG_HOST_MEMORY) != 0;
    // the kernel on the device, at the view's element offsets
    const size_t kb = size_t(p.k_hi - p.k_lo + 1) * es;
    const void* dk = k;
    const size_t sb = fft_kernel_spectrum_bytes(p, h.max_lds());
    if (host || sb == 0) {
        if (hipMalloc(&b->kcopy.ptr, kb) != hipSuccess) return set_err(hh, NDC_ERR_OUT_OF_MEMORY, "kernel copy allocation failed");
        b->kcopy.bytes = kb;
        const char* src = static_cast<const char*>(k) + p.k_lo * int64_t(es);
        if (hipMemcpyAsync(b->kcopy.ptr, src, kb, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, h.stream()) !=
            hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "kernel copy failed");
        b->k_lo = p.k_lo;
        dk = static_cast<const char*>(b->kcopy.ptr) - p.k_lo * int64_t(es);
    }
    if (sb > 0) {
        if (hipMalloc(&b->spec.ptr, sb) != hipSuccess) return set_err(hh, NDC_ERR_OUT_OF_MEMORY, "spectrum allocation failed");
        b->spec.bytes = sb;
        KernelSpectrum ks;
        ks.spec = b->spec.ptr;
        ks.bake = true;
        h.err.clear();
        s = launch_fft(p, nullptr, dk, nullptr, h, &ks);
        if (s != NDC_OK) return set_err(hh, s, h.err);
        if (b->kcopy.ptr && sb > 0) {
            // the spectrum is all that later calls read
            if (hipStreamSynchronize(h.stream()) != hipSuccess) return set_err(hh, NDC_ERR_DEVICE, "synchronize failed");
            (void)hipFree(b->kcopy.ptr);
            b->kcopy.ptr = nullptr;
        }
    }
    if (host && hipStreamSynchronize(h.stream()) != hipSuccess) return set_err(hh, NDC_ERR_DEVICE, "synchronize failed");
    *out = b.release();
    return NDC_OK;
}

ndc_status run_baked(Baked& b, const void* x, void* y) {
    ndc_handle hh = b.h;
    Handle& h = *reinterpret_cast<Handle*>(hh);
    const Plan& p = b.p;
    if (!x || !y) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "null data pointer");
    if (hipSetDevice(h.device) != hipSuccess) return set_err(hh, NDC_ERR_DEVICE, "hipSetDevice failed");
    const size_t es = dtype_size(p.dtype);
    const bool host = (p.flags & NDC_FLAG_HOST_MEMORY) != 0;
    const void* dx = x;
    void* dy = y;
    if (host) {
        const size_t xb = size_t(p.x_hi - p.x_lo + 1) * es;
        const size_t yb = size_t(p.y_numel) * es;
        char* bx = static_cast<char*>(h.grow(h.io[0], xb));
        char* by = static_cast<char*>(h.grow(h.io[2], yb ? yb : es));
        if (!bx || !by) return set_err(hh, NDC_ERR_OUT_OF_MEMORY, "device buffer allocation failed");
        if (hipMemcpyAsync(bx, static_cast<const char*>(x) + p.x_lo * int64_t(es), xb, hipMemcpyHostToDevice,
                           h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "host to device copy failed");
        dx = bx - p.x_lo * int64_t(es);
        dy = by;
    }
    h.err.clear();
    ndc_status s;
    if (b.spec.ptr) {
        KernelSpectrum ks;
        ks.spec = b.spec.ptr;
        ks.bake = false;
        s = launch_fft(p, dx, nullptr, dy, h, &ks);
    } else {
        const void* dk = static_cast<const char*>(b.kcopy.ptr) - b.k_lo * int64_t(es);
        s = launch_fft(p, dx, dk, dy, h);
    }
    if (s != NDC_OK) return set_err(hh, s, h.err);
    if (host) {
        if (hipMemcpyAsync(y, dy, size_t(p.y_numel) * es, hipMemcpyDeviceToHost, h.stream()) != hipSuccess ||
            hipStreamSynchronize(h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "device to host copy failed");
    }
    return NDC_OK;
}

// The processor's forward/backward (processor/mod.rs:79-121).
ndc_status run_processor(ndc_handle hh, bool forward, int32_t dtype, int32_t rank, const int64_t* shape,
                         const void* x, void* y, uint32_t flags) {
    if (!hh) return NDC_ERR_INVALID_ARGUMENT;
    Handle& h = *reinterpret_cast<Handle*>(hh);
    if (dtype != NDC_F32 && dtype != NDC_F64 && dtype != NDC_C64 && dtype != NDC_C128)
        return set_err(hh, NDC_ERR_UNSUPPORTED, "the processor transforms f32/f64 and complex64/complex128");
    if (rank < 1 || rank > kMaxRank || !shape) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "bad rank or shape");
    if (flags & ~uint32_t(NDC_FLAG_HOST_MEMORY)) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "unknown flags");
    int64_t n_in = 1;
    for (int i = 0; i < rank; ++i) {
        if (shape[i] < 1 || shape[i] > (int64_t(1) << 30))
            return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "every axis needs 1 .. 2^30 elements");
        n_in *= shape[i];
    }
    const bool real = dtype == NDC_F32 || dtype == NDC_F64;
    const size_t rsz = (dtype == NDC_F64 || dtype == NDC_C128) ? 8 : 4;
    const size_t csz = 2 * rsz;
    size_t in_bytes, out_bytes;
    int64_t origin = h.rp_origin_len;
    if (forward) {
        const int64_t H = real ? shape[rank - 1] / 2 + 1 : shape[rank - 1];
        in_bytes = size_t(n_in) * (real ? rsz : csz);
        out_bytes = size_t(n_in / shape[rank - 1] * H) * csz;
    } else {
        if (real && origin <= 0)
            return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "backward needs a real forward transform on this processor first");
        in_bytes = size_t(n_in) * csz;
        // natural last axis: the spectrum's second-to-last axis (its first for rank 1)
        const int64_t hl = rank == 1 ? shape[0] : shape[rank - 2];
        out_bytes = real ? size_t(n_in / hl * origin) * rsz : size_t(n_in) * csz;
    }
    if (!x || !y) return set_err(hh, NDC_ERR_INVALID_ARGUMENT, "null data pointer");
    if (hipSetDevice(h.device) != hipSuccess) return set_err(hh, NDC_ERR_DEVICE, "hipSetDevice failed");
    const bool host = (flags & NDC_FLAG_HOST_MEMORY) != 0;
    const void* dx = x;
    void* dy = y;
    if (host) {
        void* bx = h.grow(h.io[0], in_bytes);
        void* by = h.grow(h.io[2], out_bytes);
        if (!bx || !by) return set_err(hh, NDC_ERR_OUT_OF_MEMORY, "device buffer allocation failed");
        if (hipMemcpyAsync(bx, x, in_bytes, hipMemcpyHostToDevice, h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "host to device copy failed");
        dx = bx;
        dy = by;
    }
    h.err.clear();
    const ndc_status s = launch_fft_processor(dtype, forward, rank, shape, origin, dx, dy, h);
    if (s != NDC_OK) return set_err(hh, s, h.err);
    if (forward && real) h.rp_origin_len = shape[rank - 1];
    if (host) {
        if (hipMemcpyAsync(y, dy, out_bytes, hipMemcpyDeviceToHost, h.stream()) != hipSuccess ||
            hipStreamSynchronize(h.stream()) != hipSuccess)
            return set_err(hh, NDC_ERR_DEVICE, "device to host copy failed");
    }
    return NDC_OK;
}

}  // namespace

LaunchContext* handle_context(ndc_handle h) { return reinterpret_cast<Handle*>(h); }
int handle_device(ndc_handle h) { return reinterpret_cast<Handle*>(h)->device; }
void handle_set_last(ndc_handle h, const std::string& msg) {
    if (h) reinterpret_cast<Handle*>(h)->last = msg;
}

}  // namespace ndc

using ndc::Handle;

extern "C" {

int32_t ndc_abi_version(void) { return NDC_ABI_VERSION; }

int64_t ndc_struct_size(int32_t which) {
    switch (which) {
        case 0: return int64_t(sizeof(ndc_conv_desc));
        case 1: return int64_t(sizeof(ndc_plan_info));
        case 2: return int64_t(sizeof(ndc_conv_mode));
        case 3: return int64_t(sizeof(ndc_padding_mode));
        case 4: return int64_t(sizeof(ndc_border));
        case 5: return int64_t(sizeof(ndc_value));
        case 6: return int64_t(sizeof(ndc_batch_desc));
        default: return -1;
    }
}

const char* ndc_status_string(ndc_status s) {
    switch (s) {
        case NDC_OK: return "ok";
        case NDC_ERR_DATA_SHAPE: return "DataShape: data shape shouldn't have ZERO";
        case NDC_ERR_KERNEL_SHAPE: return "KernelShape: kernel shape shouldn't have ZERO";
        case NDC_ERR_MISMATCH_SHAPE: return "MismatchShape: ConvMode does not match KernelWithDilation size";
        case NDC_ERR_INVALID_ARGUMENT: return "invalid argument";
        case NDC_ERR_UNSUPPORTED: return "unsupported";
        case NDC_ERR_NO_DEVICE: return "no HIP device";
        case NDC_ERR_DEVICE: return "HIP runtime error";
        case NDC_ERR_OUT_OF_MEMORY: return "out of device memory";
    }
    return "unknown status";
}

ndc_status ndc_create(ndc_handle* out) {
    if (!out) return NDC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return NDC_ERR_NO_DEVICE;
    auto h = std::make_unique<Handle>();
    if (hipGetDevice(&h->device) != hipSuccess) return NDC_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) {
        h->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        // Dynamic LDS budget: the per-workgroup limit minus a reserve for the
        // kernels' static __shared__ arrays (at most 576 bytes today), which
        // count against the same limit.
        const size_t total = prop.sharedMemPerBlock > 0 ? prop.sharedMemPerBlock : 64 * 1024;
        h->lds = total - ndc::kStaticLdsReserve;
    }
    if (hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking) != hipSuccess) return NDC_ERR_DEVICE;
    *out = reinterpret_cast<ndc_handle>(h.release());
    return NDC_OK;
}

ndc_status ndc_destroy(ndc_handle h) {
    if (!h) return NDC_ERR_INVALID_ARGUMENT;
    Handle* p = reinterpret_cast<Handle*>(h);
    (void)hipStreamSynchronize(p->stream());
    delete p;
    return NDC_OK;
}

ndc_status ndc_set_stream(ndc_handle h, void* stream) {
    if (!h) return NDC_ERR_INVALID_ARGUMENT;
    Handle* p = reinterpret_cast<Handle*>(h);
    if (hipSetDevice(p->device) != hipSuccess) return NDC_ERR_DEVICE;
    return p->switch_stream(static_cast<hipStream_t>(stream), true);
}

ndc_status ndc_reset_stream(ndc_handle h) {
    if (!h) return NDC_ERR_INVALID_ARGUMENT;
    Handle* p = reinterpret_cast<Handle*>(h);
    if (hipSetDevice(p->device) != hipSuccess) return NDC_ERR_DEVICE;
    return p->switch_stream(p->own, false);
}

ndc_status ndc_synchronize(ndc_handle h) {
    if (!h) return NDC_ERR_INVALID_ARGUMENT;
    if (hipStreamSynchronize(reinterpret_cast<Handle*>(h)->stream()) != hipSuccess) return NDC_ERR_DEVICE;
    return NDC_OK;
}

const char* ndc_last_error(ndc_handle h) {
    if (!h) return "";
    return reinterpret_cast<Handle*>(h)->last.c_str();
}

ndc_status ndc_plan(const ndc_conv_desc* desc, int32_t op, ndc_plan_info* info) {
    ndc::Plan p;
    std::string err;
    ndc_status s = ndc::make_plan(desc, op, &p, &err, false);
    if (s != NDC_OK) return s;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        for (int i = 0; i < p.N; ++i) {
            info->padding[i][0] = p.pad[i][0];
            info->padding[i][1] = p.pad[i][1];
            info->strides[i] = p.stride[i];
            info->kernel_dilated[i] = p.kdil[i];
            info->padded[i] = p.pds[i];
            info->fft_size[i] = p.fft[i];
            info->out_shape[i] = p.out[i];
        }
    }
    return NDC_OK;
}

ndc_status ndc_border_map(const ndc_conv_desc* desc, int32_t op, int32_t axis, int32_t* codes, int64_t len) {
    ndc::Plan p;
    std::string err;
    ndc_status s = ndc::make_plan(desc, op, &p, &err, true);
    if (s != NDC_OK) return s;
    if (axis < 0 || axis >= p.N || !codes) return NDC_ERR_INVALID_ARGUMENT;
    const auto& m = p.map[axis];
    if (len != int64_t(m.size())) return NDC_ERR_INVALID_ARGUMENT;
    std::memcpy(codes, m.data(), m.size() * sizeof(int32_t));
    return NDC_OK;
}

ndc_status ndc_conv_fft_launches(const ndc_conv_desc* desc, int64_t lds_bytes, int32_t cus, int64_t* smem,
                                 int64_t* grid, int64_t* block, int64_t cap, int64_t* count) {
    if (!desc || lds_bytes <= 0 || cus <= 0 || cap < 0 || !count || (cap > 0 && (!smem || !grid || !block)))
        return NDC_ERR_INVALID_ARGUMENT;
    ndc::Plan p;
    std::string err;
    ndc_status s = ndc::make_plan(desc, NDC_OP_CONV_FFT, &p, &err, false);
    if (s != NDC_OK) return s;
    ndc::DryRunContext ctx(size_t(lds_bytes), cus);
    *count = 0;
    if (p.y_numel > 0) {
        s = ndc::launch_fft(p, reinterpret_cast<const void*>(uintptr_t(1) << 24),
                            reinterpret_cast<const void*>(uintptr_t(2) << 24), reinterpret_cast<void*>(uintptr_t(3) << 24),
                            ctx);
        if (s != NDC_OK) return s;
    }
    *count = int64_t(ctx.launches.size());
    for (int64_t i = 0; i < cap && i < *count; ++i) {
        smem[i] = int64_t(ctx.launches[size_t(i)].smem);
        grid[i] = int64_t(ctx.launches[size_t(i)].grid);
        block[i] = int64_t(ctx.launches[size_t(i)].block);
    }
    return NDC_OK;
}

ndc_status ndc_conv(ndc_handle h, const ndc_conv_desc* desc, const void* x, const void* k, void* y) {
    return ndc::run(h, desc, NDC_OP_CONV, x, k, y);
}

ndc_status ndc_conv_fft(ndc_handle h, const ndc_conv_desc* desc, const void* x, const void* k, void* y) {
    return ndc::run(h, desc, NDC_OP_CONV_FFT, x, k, y);
}

ndc_status ndc_conv_fft_bake(ndc_handle h, const ndc_conv_desc* desc, const void* k, ndc_baked* out) {
    if (!out) return NDC_ERR_INVALID_ARGUMENT;
    ndc::Baked* b = nullptr;
    const ndc_status s = ndc::bake(h, desc, k, &b);
    *out = reinterpret_cast<ndc_baked>(b);
    return s;
}

ndc_status ndc_conv_fft_with_baked(ndc_baked b, const void* x, void* y) {
    if (!b) return NDC_ERR_INVALID_ARGUMENT;
    return ndc::run_baked(*reinterpret_cast<ndc::Baked*>(b), x, y);
}

ndc_status ndc_baked_destroy(ndc_baked b) {
    if (!b) return NDC_ERR_INVALID_ARGUMENT;
    ndc::Baked* p = reinterpret_cast<ndc::Baked*>(b);
    (void)hipStreamSynchronize(reinterpret_cast<Handle*>(p->h)->stream());
    delete p;
    return NDC_OK;
}

ndc_status ndc_fft_forward(ndc_handle h, int32_t dtype, int32_t rank, const int64_t* shape, const void* x, void* y,
                           uint32_t flags) {
    return ndc::run_processor(h, true, dtype, rank, shape, x, y, flags);
}

ndc_status ndc_fft_backward(ndc_handle h, int32_t dtype, int32_t rank, const int64_t* spec_shape, const void* spec,
                            void* y, uint32_t flags) {
    return ndc::run_processor(h, false, dtype, rank, spec_shape, spec, y, flags);
}

ndc_status ndc_fft_origin_len(ndc_handle h, int64_t* len) {
    if (!h || !len) return NDC_ERR_INVALID_ARGUMENT;
    *len = reinterpret_cast<Handle*>(h)->rp_origin_len;
    return NDC_OK;
}

ndc_status ndc_profile_enable(ndc_handle hh, int32_t on) {
    if (!hh) return NDC_ERR_INVALID_ARGUMENT;
    reinterpret_cast<Handle*>(hh)->prof = on != 0;
    return NDC_OK;
}

ndc_status ndc_profile_report(ndc_handle hh, char* buf, int64_t len) {
    if (!hh || !buf || len <= 0) return NDC_ERR_INVALID_ARGUMENT;
    Handle& h = *reinterpret_cast<Handle*>(hh);
    if (hipStreamSynchronize(h.stream()) != hipSuccess) return NDC_ERR_DEVICE;
    struct Acc {
        int64_t n = 0;
        double ms = 0, bytes = 0, flops = 0;
    };
    std::map<std::string, Acc> acc;
    for (auto& r : h.recs) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            Acc& a = acc[r.name];
            a.n += 1;
            a.ms += ms;
            a.bytes += r.bytes;
            a.flops += r.flops;
        }
        h.pool.push_back(r.a);
        h.pool.push_back(r.b);
    }
    h.recs.clear();
    std::string out = "{";
    bool first = true;
    char tmp[512];
    for (auto& kv : acc) {
        std::snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"launches\": %lld, \"total_ms\": %.6f, \"bytes\": %.1f, \"flops\": %.1f}",
                      first ? "" : ", ", kv.first.c_str(), static_cast<long long>(kv.second.n), kv.second.ms,
                      kv.second.bytes, kv.second.flops);
        out += tmp;
        first = false;
    }
    out += "}";
    if (int64_t(out.size()) + 1 > len) return NDC_ERR_INVALID_ARGUMENT;
    std::memcpy(buf, out.c_str(), out.size() + 1);
    return NDC_OK;
}

}  // extern "C"
