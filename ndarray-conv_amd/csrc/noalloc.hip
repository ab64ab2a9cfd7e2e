// ndc_conv_fft_noalloc: ndc_conv_fft for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The FFT convolution's workspace (the data and kernel
// spectra, the twiddle and post-processing tables) is carved from one
// file-scope device array instead of the handle's allocations; the kernels
// and their launches are those of ndc_conv_fft.
//
// The array has kNoAllocScratchBytes per GPU and lives until the process
// ends.  Host code never passes the array's name to a kernel or a copy (in
// host code it is only a stand-in for the symbol): its device address is
// looked up once per device with hipGetSymbolAddress and kept in g_scratch.
// Every call on a GPU carves the same array, so calls that overlap on one GPU
// (different streams, or different host threads) race on it; the tables are
// uploaded by every call, nothing in the array has to start at zero, and the
// array keeps the last call's contents.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "batch.hpp"
#include "launch.hpp"
#include "ndconv.h"
#include "plan.hpp"

namespace ndc {
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up
// once per device.
ndc_status scratch_of(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return NDC_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return NDC_ERR_DEVICE;
        g_scratch[device] = ptr;
    }
    *out = g_scratch[device];
    return NDC_OK;
}

// A LaunchContext whose device memory comes from the scratch array: a bump
// carve per call, tables uploaded on the call's stream.  Everything else is
// the handle's (stream, device limits); host-side plan memos live only for
// the call, since the tables they point at are rewritten by the next call.
class NoAllocContext final : public LaunchContext {
public:
    NoAllocContext(LaunchContext& h, unsigned char* base, size_t cap) : h_(h), base_(base), cap_(cap) {}

    hipStream_t stream() override { return h_.stream(); }
    int cu_count() override { return h_.cu_count(); }
    size_t max_lds() override { return h_.max_lds(); }
    std::string& memo(const std::string& key) override { return memos_[key]; }

    void* scratch(int slot, size_t bytes) override {
        if (slot < 0 || slot >= 4) return nullptr;
        if (slot_ptr_[slot] && slot_bytes_[slot] >= bytes) return slot_ptr_[slot];
        void* p = carve(bytes);
        if (p) {
            slot_ptr_[slot] = p;
            slot_bytes_[slot] = bytes;
        }
        return p;
    }

    const void* twiddles(int64_t n, bool dbl) override {
        const size_t esz = dbl ? 16 : 8;
        std::vector<unsigned char> host(size_t(n) * esz);
        for (int64_t i = 0; i < n; ++i) {
            const double a = -2.0 * M_PI * double(i) / double(n);
            if (dbl) {
                const double v[2] = {std::cos(a), std::sin(a)};
                std::memcpy(&host[size_t(i) * esz], v, esz);
            } else {
                const float v[2] = {float(std::cos(a)), float(std::sin(a))};
                std::memcpy(&host[size_t(i) * esz], v, esz);
            }
        }
        return upload(host.data(), host.size());
    }

    const void* constant_table(const std::string& key, const void* data, size_t bytes) override {
        auto it = tables_.find(key);
        if (it != tables_.end()) return it->second;
        const void* p = upload(data, bytes);
        if (p) tables_[key] = p;
        return p;
    }

    const int32_t* border_maps(const Plan& p, int64_t offsets[kMaxRank + 1]) override {
        std::vector<int32_t> blob;
        for (int i = 0; i < p.N; ++i) {
            offsets[i] = int64_t(blob.size());
            blob.insert(blob.end(), p.map[i].begin(), p.map[i].end());
        }
        for (int i = p.N; i < kMaxRank; ++i) offsets[i] = 0;
        offsets[kMaxRank] = int64_t(blob.size());
        blob.push_back(0);
        return static_cast<const int32_t*>(upload(blob.data(), blob.size() * sizeof(int32_t)));
    }

private:
    void* carve(size_t bytes) {
        const size_t off = (used_ + 255) & ~size_t(255);
        if (off > cap_ || bytes > cap_ - off) {
            err = "the allocation-free workspace (" + std::to_string(cap_) + " bytes) is exhausted";
            return nullptr;
        }
        used_ = off + bytes;
        return base_ + off;
    }

    // The host table is copied on the call's stream, ordered before the
    // kernels that read it; hipMemcpyAsync from pageable memory has taken the
    // source when it returns.
    const void* upload(const void* data, size_t bytes) {
        void* p = carve(bytes ? bytes : 16);
        if (!p) return nullptr;
        if (bytes && hipMemcpyAsync(p, data, bytes, hipMemcpyHostToDevice, stream()) != hipSuccess) {
            err = "table upload failed";
            return nullptr;
        }
        return p;
    }

    LaunchContext& h_;
    unsigned char* base_;
    size_t cap_;
    size_t used_ = 0;
    void* slot_ptr_[4] = {};
    size_t slot_bytes_[4] = {};
    std::unordered_map<std::string, const void*> tables_;
    std::unordered_map<std::string, std::string> memos_;
};

ndc_status fail(ndc_handle h, ndc_status s, const std::string& msg) {
    handle_set_last(h, msg);
    return s;
}

}  // namespace
}  // namespace ndc

extern "C" ndc_status ndc_conv_fft_noalloc(ndc_handle h, const ndc_conv_desc* desc, const void* x, const void* k,
                                          void* y) {
    using namespace ndc;
    if (!h) return NDC_ERR_INVALID_ARGUMENT;
    if (!desc) return fail(h, NDC_ERR_INVALID_ARGUMENT, "null descriptor");
    Plan p;
    std::string err;
    const ndc_status st = make_plan(desc, NDC_OP_CONV_FFT, &p, &err, false);
    if (st != NDC_OK) return fail(h, st, err);
    if (p.flags & NDC_FLAG_HOST_MEMORY)
        return fail(h, NDC_ERR_INVALID_ARGUMENT, "the allocation-free entry point takes device pointers only");
    if (!x || !k || !y) return fail(h, NDC_ERR_INVALID_ARGUMENT, "null data pointer");
    LaunchContext* hc = handle_context(h);
    const int dev = handle_device(h);
    if (hipSetDevice(dev) != hipSuccess) return fail(h, NDC_ERR_DEVICE, "hipSetDevice failed");
    const size_t need = fft_workspace_bytes(p, hc->max_lds());
    if (need > kNoAllocScratchBytes)
        return fail(h, NDC_ERR_UNSUPPORTED,
                    "this convolution needs up to " + std::to_string(need) +
                        " bytes of FFT workspace; the allocation-free entry point has " +
                        std::to_string(kNoAllocScratchBytes));
    if (p.y_numel == 0) return NDC_OK;
    void* base = nullptr;
    if (scratch_of(dev, &base) != NDC_OK)
        return fail(h, NDC_ERR_DEVICE, "hipGetSymbolAddress of the workspace array failed");
    NoAllocContext ctx(*hc, static_cast<unsigned char*>(base), kNoAllocScratchBytes);
    const ndc_status s = launch_fft(p, x, k, y, ctx);
    if (s != NDC_OK) return fail(h, s, ctx.err);
    return NDC_OK;
}
