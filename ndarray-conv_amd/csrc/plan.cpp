#include "plan.hpp"

#include <cstring>
#include <limits>

namespace ndc {

size_t dtype_size(int dtype) {
    switch (dtype) {
        case NDC_F32: return 4;
        case NDC_F64: return 8;
        case NDC_C64: return 8;
        case NDC_C128: return 16;
        case NDC_I8: case NDC_U8: return 1;
        case NDC_I16: case NDC_U16: return 2;
        case NDC_I32: case NDC_U32: return 4;
        case NDC_I64: case NDC_U64: return 8;
        default: return 0;
    }
}

bool dtype_valid(int dtype) { return dtype_size(dtype) != 0; }
bool dtype_is_complex(int dtype) { return dtype == NDC_C64 || dtype == NDC_C128; }
bool dtype_is_float(int dtype) {
    return dtype == NDC_F32 || dtype == NDC_F64 || dtype == NDC_C64 || dtype == NDC_C128;
}

// good_size.rs:6-31.  The truncating divisions are part of the plan: they
// produce sizes such as 425 = 5^2 * 17 or 218450 = 2 * 5^2 * 17 * 257.
int64_t good_size(int64_t n) {
    int64_t best = 1;
    while (best < n) best <<= 1;
    for (;;) {
        int64_t nf = best / 4 * 3;
        if (nf < n) break;
        if (nf == n) return n;
        best = nf;
    }
    for (;;) {
        int64_t nf = best / 6 * 5;
        if (nf < n) break;
        if (nf == n) return n;
        best = nf;
    }
    return best;
}

namespace {

bool border_tag_valid(int t) { return t >= NDC_BORDER_ZEROS && t <= NDC_BORDER_CIRCULAR; }

ndc_border make_border(int tag, const ndc_value& v) {
    ndc_border b;
    std::memset(&b, 0, sizeof(b));
    b.tag = tag;
    if (tag == NDC_BORDER_CONST) b.value = v;
    if (tag == NDC_BORDER_ZEROS) b.tag = NDC_BORDER_CONST;   // a fill with T::zero()
    return b;
}

// Builds the border map of one axis by replaying the reference's fill loops
// on indices instead of values.  `len` is the map length (>= P; positions
// beyond P are kOutside).  Returns false where the reference panics.
//   front fills: half_dim.rs constant_front :40-48, replicate_front :119-128,
//                reflect_front :198-210, circular_front :283-295
//   back fills:  constant_back :85-93, replicate_back :163-172,
//                reflect_back :245-257, circular_back :330-342
bool axis_map(int64_t n, int64_t p0, int64_t p1, const ndc_border& front, const ndc_border& back,
              int64_t len, std::vector<int32_t>* out, std::string* err) {
    const int64_t P = n + p0 + p1;
    std::vector<int32_t>& m = *out;
    m.assign(static_cast<size_t>(len), kOutside);
    for (int64_t j = 0; j < P; ++j) m[j] = kUnwritten;
    for (int64_t j = 0; j < n; ++j) m[p0 + j] = static_cast<int32_t>(j);
    // front side, increasing j
    for (int64_t j = 0; j < p0; ++j) {
        switch (front.tag) {
            case NDC_BORDER_CONST: m[j] = kConstFront; break;
            case NDC_BORDER_REPLICATE: m[j] = m[p0]; break;
            case NDC_BORDER_REFLECT: {
                int64_t src = (p0 - j) + p0;
                if (src >= P) {
                    if (err) *err = "reflect front pad reads outside the padded axis (the reference panics)";
                    return false;
                }
                m[j] = m[src];
                break;
            }
            case NDC_BORDER_CIRCULAR: m[j] = m[P - p1 - (p0 - j)]; break;
            default: break;
        }
    }
    // back side, increasing j
    const int64_t border = P - p1 - 1;
    for (int64_t j = n + p0; j < P; ++j) {
        switch (back.tag) {
            case NDC_BORDER_CONST: m[j] = kConstBack; break;
            case NDC_BORDER_REPLICATE: m[j] = m[border]; break;
            case NDC_BORDER_REFLECT: {
                int64_t dd = j - border;
                if (dd > border) {
                    if (err) *err = "reflect back pad reads outside the padded axis (the reference panics)";
                    return false;
                }
                m[j] = m[border - dd];
                break;
            }
            case NDC_BORDER_CIRCULAR: m[j] = m[p0 + (j - border - 1)]; break;
            default: break;
        }
    }
    return true;
}

void contiguous_strides(const int64_t* shape, int N, int64_t* str) {
    int64_t s = 1;
    for (int i = N - 1; i >= 0; --i) {
        str[i] = s;
        s *= shape[i] > 0 ? shape[i] : 1;
    }
}

void span(const int64_t* shape, const int64_t* str, int N, int64_t* lo, int64_t* hi) {
    *lo = 0;
    *hi = 0;
    for (int i = 0; i < N; ++i) {
        int64_t e = (shape[i] - 1) * str[i];
        if (e < 0) *lo += e; else *hi += e;
    }
}

}  // namespace

ndc_status make_plan(const ndc_conv_desc* d, int op, Plan* p, std::string* err, bool build_maps) {
    auto fail = [&](ndc_status s, const std::string& msg) {
        if (err) *err = msg;
        return s;
    };
    if (!d || !p) return fail(NDC_ERR_INVALID_ARGUMENT, "null descriptor");
    if (op != NDC_OP_CONV && op != NDC_OP_CONV_FFT) return fail(NDC_ERR_INVALID_ARGUMENT, "unknown op");
    if (d->rank < 1 || d->rank > kMaxRank) return fail(NDC_ERR_INVALID_ARGUMENT, "rank must be 1..6");
    if (!dtype_valid(d->dtype)) return fail(NDC_ERR_INVALID_ARGUMENT, "unknown dtype");
    if (op == NDC_OP_CONV_FFT && !dtype_is_float(d->dtype))
        return fail(NDC_ERR_UNSUPPORTED, "conv_fft needs f32, f64, c64 or c128 (processor/mod.rs:42-61)");
    const int N = d->rank;
    Plan& P = *p;
    P.op = op;
    P.dtype = d->dtype;
    P.N = N;
    P.reverse = d->reverse != 0;
    P.algo = d->algo;
    P.flags = d->flags;
    for (int i = 0; i < kMaxRank; ++i) P.map[i].clear();
    const int64_t kAxisMax = (int64_t(1) << 31) - 1;
    for (int i = 0; i < N; ++i) {
        if (d->x_shape[i] < 0 || d->k_shape[i] < 0) return fail(NDC_ERR_INVALID_ARGUMENT, "negative shape");
        if (d->x_shape[i] > kAxisMax || d->k_shape[i] > kAxisMax)
            return fail(NDC_ERR_UNSUPPORTED, "axis longer than 2^31-1");
        if (d->dilation[i] < 0) return fail(NDC_ERR_INVALID_ARGUMENT, "negative dilation");
        P.x_shape[i] = d->x_shape[i];
        P.k_shape[i] = d->k_shape[i];
        P.dil[i] = d->dilation[i];
    }
    if (d->flags & NDC_FLAG_X_STRIDED) {
        for (int i = 0; i < N; ++i) P.x_str[i] = d->x_strides[i];
    } else {
        contiguous_strides(P.x_shape, N, P.x_str);
    }
    if (d->flags & NDC_FLAG_K_STRIDED) {
        for (int i = 0; i < N; ++i) P.k_str[i] = d->k_strides[i];
    } else {
        contiguous_strides(P.k_shape, N, P.k_str);
    }

    // Descriptor well-formedness (not reference behaviour: the Rust types make
    // these unrepresentable).
    const ndc_conv_mode& cm = d->conv_mode;
    if (cm.tag < NDC_CONV_FULL || cm.tag > NDC_CONV_EXPLICIT) return fail(NDC_ERR_INVALID_ARGUMENT, "conv mode tag");
    const ndc_padding_mode& pm = d->padding_mode;
    if (pm.tag < NDC_PAD_ZEROS || pm.tag > NDC_PAD_EXPLICIT) return fail(NDC_ERR_INVALID_ARGUMENT, "padding mode tag");
    if (pm.tag == NDC_PAD_CUSTOM || pm.tag == NDC_PAD_EXPLICIT) {
        for (int i = 0; i < N; ++i)
            for (int s = 0; s < (pm.tag == NDC_PAD_EXPLICIT ? 2 : 1); ++s)
                if (!border_tag_valid(pm.borders[i][s].tag))
                    return fail(NDC_ERR_INVALID_ARGUMENT, "border tag");
    }
    if (cm.tag == NDC_CONV_CUSTOM || cm.tag == NDC_CONV_EXPLICIT) {
        for (int i = 0; i < N; ++i) {
            if (cm.strides[i] < 0) return fail(NDC_ERR_INVALID_ARGUMENT, "negative stride");
            for (int s = 0; s < 2; ++s)
                if (cm.padding[i][s] < 0 || cm.padding[i][s] > kAxisMax)
                    return fail(NDC_ERR_INVALID_ARGUMENT, "padding out of range");
        }
    }

    P.x_numel = 1;
    P.k_numel = 1;
    for (int i = 0; i < N; ++i) {
        P.x_numel *= P.x_shape[i];
        P.k_numel *= P.k_shape[i];
    }
    // conv/mod.rs:137-144, conv_fft/mod.rs:206-213
    if (P.x_numel == 0) return fail(NDC_ERR_DATA_SHAPE, "data shape has a zero axis");
    if (P.k_numel == 0)
        return fail(op == NDC_OP_CONV ? NDC_ERR_KERNEL_SHAPE : NDC_ERR_DATA_SHAPE,
                    "kernel shape has a zero axis");

    // kernel size with dilation (conv/mod.rs:146-147)
    for (int i = 0; i < N; ++i) P.kdil[i] = P.k_shape[i] * P.dil[i] - P.dil[i] + 1;

    // ConvMode::unfold (conv/mod.rs:28-66)
    for (int i = 0; i < N; ++i) {
        const int64_t k = P.kdil[i];
        switch (cm.tag) {
            case NDC_CONV_FULL: P.pad[i][0] = P.pad[i][1] = k - 1; P.stride[i] = 1; break;
            case NDC_CONV_SAME:
                if (k % 2 == 0) { P.pad[i][0] = (k - 1) / 2 + 1; P.pad[i][1] = (k - 1) / 2; }
                else { P.pad[i][0] = P.pad[i][1] = (k - 1) / 2; }
                P.stride[i] = 1;
                break;
            case NDC_CONV_VALID: P.pad[i][0] = P.pad[i][1] = 0; P.stride[i] = 1; break;
            case NDC_CONV_CUSTOM: P.pad[i][0] = P.pad[i][1] = cm.padding[i][0]; P.stride[i] = cm.strides[i]; break;
            case NDC_CONV_EXPLICIT:
                P.pad[i][0] = cm.padding[i][0]; P.pad[i][1] = cm.padding[i][1]; P.stride[i] = cm.strides[i];
                break;
        }
        P.pds[i] = P.x_shape[i] + P.pad[i][0] + P.pad[i][1];
        if (P.pds[i] > kAxisMax) return fail(NDC_ERR_UNSUPPORTED, "padded axis longer than 2^31-1");
    }

    // Border of each side (padding/mod.rs:84-153, 346-452)
    for (int i = 0; i < N; ++i) {
        for (int s = 0; s < 2; ++s) {
            switch (pm.tag) {
                case NDC_PAD_ZEROS: P.side[i][s] = make_border(NDC_BORDER_ZEROS, pm.value); break;
                case NDC_PAD_CONST: P.side[i][s] = make_border(NDC_BORDER_CONST, pm.value); break;
                case NDC_PAD_REFLECT: P.side[i][s] = make_border(NDC_BORDER_REFLECT, pm.value); break;
                case NDC_PAD_REPLICATE: P.side[i][s] = make_border(NDC_BORDER_REPLICATE, pm.value); break;
                case NDC_PAD_CIRCULAR: P.side[i][s] = make_border(NDC_BORDER_CIRCULAR, pm.value); break;
                case NDC_PAD_CUSTOM: P.side[i][s] = make_border(pm.borders[i][0].tag, pm.borders[i][0].value); break;
                case NDC_PAD_EXPLICIT: P.side[i][s] = make_border(pm.borders[i][s].tag, pm.borders[i][s].value); break;
            }
        }
    }

    auto maps = [&](bool fft) -> ndc_status {
        for (int i = 0; i < N; ++i) {
            std::vector<int32_t> tmp;
            std::string e;
            const int64_t len = fft ? P.fft[i] : P.pds[i];
            if (!axis_map(P.x_shape[i], P.pad[i][0], P.pad[i][1], P.side[i][0], P.side[i][1], len, &tmp, &e))
                return fail(NDC_ERR_INVALID_ARGUMENT, "axis " + std::to_string(i) + ": " + e);
            if (build_maps) P.map[i].swap(tmp);
        }
        return NDC_OK;
    };

    if (op == NDC_OP_CONV) {
        // conv/mod.rs:150 pads before the shape check, so a reflect panic comes first
        ndc_status s = maps(false);
        if (s != NDC_OK) return s;
        for (int i = 0; i < N; ++i)
            if (P.kdil[i] > P.pds[i]) return fail(NDC_ERR_MISMATCH_SHAPE, "kernel with dilation larger than the padded input");
        for (int i = 0; i < N; ++i)
            if (P.stride[i] == 0) return fail(NDC_ERR_INVALID_ARGUMENT, "stride 0 (the reference divides by zero)");
        for (int i = 0; i < N; ++i) {
            P.out[i] = (P.pad[i][0] + P.pad[i][1] + P.x_shape[i] - P.kdil[i]) / P.stride[i] + 1;
            P.fft[i] = 0;
        }
    } else {
        // conv_fft/mod.rs:220-231: shape check first, then plan, then padding
        for (int i = 0; i < N; ++i)
            if (P.kdil[i] > P.pds[i]) return fail(NDC_ERR_MISMATCH_SHAPE, "kernel with dilation larger than the padded input");
        for (int i = 0; i < N; ++i) {
            const int64_t need = P.pds[i] > P.kdil[i] ? P.pds[i] : P.kdil[i];
            P.fft[i] = good_size(need);
        }
        ndc_status s = maps(true);
        if (s != NDC_OK) return s;
        for (int i = 0; i < N; ++i)
            if (P.dil[i] == 0) return fail(NDC_ERR_INVALID_ARGUMENT, "dilation 0 (the reference's kernel slice step is 0)");
        for (int i = 0; i < N; ++i)
            if (P.stride[i] == 0) return fail(NDC_ERR_INVALID_ARGUMENT, "stride 0 (the reference's crop step is 0)");
        for (int i = 0; i < N; ++i) P.out[i] = (P.pds[i] - P.kdil[i]) / P.stride[i] + 1;
    }
    P.y_numel = 1;
    for (int i = 0; i < N; ++i) P.y_numel *= P.out[i];
    span(P.x_shape, P.x_str, N, &P.x_lo, &P.x_hi);
    span(P.k_shape, P.k_str, N, &P.k_lo, &P.k_hi);
    return NDC_OK;
}

std::string plan_key(const Plan& p) {
    std::string k;
    auto put = [&](const void* ptr, size_t n) { k.append(static_cast<const char*>(ptr), n); };
    put(&p.op, sizeof(p.op));
    put(&p.N, sizeof(p.N));
    for (int i = 0; i < p.N; ++i) {
        put(&p.x_shape[i], 8);
        put(&p.pad[i][0], 16);
        put(&p.fft[i], 8);
        for (int s = 0; s < 2; ++s) put(&p.side[i][s].tag, 4);
    }
    return k;
}

}  // namespace ndc
