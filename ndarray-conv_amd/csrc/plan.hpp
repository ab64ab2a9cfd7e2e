// Host-side planning for one convolution: validation with the reference's
// error precedence, ConvMode unfolding, FFT sizes and the per-axis border
// index maps that the HIP kernels use to read the padded input without ever
// materialising it.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "ndconv.h"

namespace ndc {

constexpr int kMaxRank = NDC_MAX_RANK;

// Codes stored in a border map next to non-negative source indices.
enum BorderCode : int32_t {
    kUnwritten = -1,   // pad cell never written: holds the buffer's initial value (0)
    kConstFront = -2,  // value written by this axis's front fill
    kConstBack = -3,   // value written by this axis's back fill
    kOutside = -4      // FFT buffer cell beyond the padded region (always 0)
};

struct Plan {
    int op = 0;
    int dtype = 0;
    int N = 0;
    bool reverse = true;
    int algo = 0;
    uint32_t flags = 0;
    int64_t x_shape[kMaxRank] = {};
    int64_t x_str[kMaxRank] = {};
    int64_t k_shape[kMaxRank] = {};
    int64_t k_str[kMaxRank] = {};
    int64_t dil[kMaxRank] = {};
    int64_t pad[kMaxRank][2] = {};
    int64_t stride[kMaxRank] = {};
    int64_t kdil[kMaxRank] = {};
    int64_t pds[kMaxRank] = {};
    int64_t fft[kMaxRank] = {};
    int64_t out[kMaxRank] = {};
    ndc_border side[kMaxRank][2] = {};     // resolved border of each side; Zeros becomes Const(0)
    std::vector<int32_t> map[kMaxRank];    // border map per axis (length pds, or fft for conv_fft)
    int64_t x_numel = 0, k_numel = 0, y_numel = 0;
    // Element offsets (relative to the data pointer) of the lowest and highest
    // element a strided view touches.
    int64_t x_lo = 0, x_hi = 0, k_lo = 0, k_hi = 0;
};

size_t dtype_size(int dtype);
bool dtype_valid(int dtype);
bool dtype_is_complex(int dtype);
bool dtype_is_float(int dtype);   // real or complex floating point

// good_size_cc (conv_fft/good_size.rs:6-31).
int64_t good_size(int64_t n);

// Validates `d` for `op` and fills `p`.  Returns the status the reference
// would produce (Err variant or panic) in the reference's order of checks.
// When build_maps is false the border maps are left empty (shape queries).
ndc_status make_plan(const ndc_conv_desc* d, int op, Plan* p, std::string* err, bool build_maps);

// A byte string that identifies everything the device-side metadata of a
// plan depends on (used as a cache key).
std::string plan_key(const Plan& p);

}  // namespace ndc
