// C++ mirror of ndarray-conv's convolution API over the C ABI (ndconv.h).
//
//   ndc::Array<T> x = ...;                      // shape + strides + data (host)
//   auto y = ndc::conv(x, ndc::with_dilation(k, 2).no_reverse(),
//                      ndc::ConvMode::Same(), ndc::PaddingMode::Custom({ndc::BorderType::Reflect(),
//                                                                        ndc::BorderType::Circular()}));
//   auto z = ndc::conv_fft(x, k, ndc::ConvMode::Full(), ndc::PaddingMode::Zeros());
//
// Names and argument meaning follow the crate: ConvMode (lib.rs:80-105),
// PaddingMode (lib.rs:106-127), BorderType (lib.rs:129-143),
// KernelWithDilation / with_dilation / reverse / no_reverse
// (dilation/mod.rs:8-12,103-189), ConvExt::conv (conv/mod.rs:110-115),
// ConvFFTExt::conv_fft / conv_fft_with_processor (conv_fft/mod.rs:113-157),
// Processor::forward / backward (processor/mod.rs:79-121), the sketched
// conv_fft_bake / conv_fft_with_baked (conv_fft/mod.rs:175-182), and Batch for
// many rank-1 pairs (one conv_fft per pair).
// Errors are thrown as DataShape / KernelShape / MismatchShape (lib.rs:145-159);
// inputs on which the crate panics throw InvalidArgument.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "ndconv.h"

namespace ndc {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};
struct DataShape : Error { using Error::Error; };
struct KernelShape : Error { using Error::Error; };
struct MismatchShape : Error { using Error::Error; };
struct InvalidArgument : Error { using Error::Error; };
struct Unsupported : Error { using Error::Error; };
struct DeviceError : Error { using Error::Error; };

inline void check(ndc_status s, ndc_handle h = nullptr) {
    if (s == NDC_OK) return;
    std::string msg = ndc_status_string(s);
    if (h) {
        const char* d = ndc_last_error(h);
        if (d && *d) msg += std::string(": ") + d;
    }
    switch (s) {
        case NDC_ERR_DATA_SHAPE: throw DataShape(msg);
        case NDC_ERR_KERNEL_SHAPE: throw KernelShape(msg);
        case NDC_ERR_MISMATCH_SHAPE: throw MismatchShape(msg);
        case NDC_ERR_INVALID_ARGUMENT: throw InvalidArgument(msg);
        case NDC_ERR_UNSUPPORTED: throw Unsupported(msg);
        default: throw DeviceError(msg);
    }
}

template <typename T> struct DType;
template <> struct DType<float> { static constexpr int value = NDC_F32; };
template <> struct DType<double> { static constexpr int value = NDC_F64; };
template <> struct DType<std::complex<float>> { static constexpr int value = NDC_C64; };
template <> struct DType<std::complex<double>> { static constexpr int value = NDC_C128; };
template <> struct DType<int8_t> { static constexpr int value = NDC_I8; };
template <> struct DType<int16_t> { static constexpr int value = NDC_I16; };
template <> struct DType<int32_t> { static constexpr int value = NDC_I32; };
template <> struct DType<int64_t> { static constexpr int value = NDC_I64; };
template <> struct DType<uint8_t> { static constexpr int value = NDC_U8; };
template <> struct DType<uint16_t> { static constexpr int value = NDC_U16; };
template <> struct DType<uint32_t> { static constexpr int value = NDC_U32; };
template <> struct DType<uint64_t> { static constexpr int value = NDC_U64; };

template <typename T>
ndc_value to_value(const T& v) {
    ndc_value out{0, 0, 0};
    if constexpr (std::is_same<T, std::complex<float>>::value || std::is_same<T, std::complex<double>>::value) {
        out.re = double(v.real());
        out.im = double(v.imag());
    } else if constexpr (std::is_floating_point<T>::value) {
        out.re = double(v);
    } else {
        out.i = int64_t(v);
    }
    return out;
}

// An owned or borrowed N-D array: shape, element strides (any sign) and a
// pointer to the first logical element.
template <typename T>
struct Array {
    std::vector<int64_t> shape;
    std::vector<int64_t> strides;
    const T* ptr = nullptr;
    std::vector<T> owned;

    static Array from_vec(std::vector<int64_t> shape, std::vector<T> data) {
        Array a;
        a.shape = std::move(shape);
        a.owned = std::move(data);
        a.strides.assign(a.shape.size(), 1);
        for (int i = int(a.shape.size()) - 2; i >= 0; --i) a.strides[i] = a.strides[i + 1] * a.shape[i + 1];
        a.ptr = a.owned.data();
        return a;
    }
    static Array view(std::vector<int64_t> shape, std::vector<int64_t> strides, const T* first) {
        Array a;
        a.shape = std::move(shape);
        a.strides = std::move(strides);
        a.ptr = first;
        return a;
    }
    size_t ndim() const { return shape.size(); }
    const std::vector<T>& data() const { return owned; }
};

class BorderType {
public:
    static BorderType Zeros() { return BorderType(NDC_BORDER_ZEROS); }
    template <typename T> static BorderType Const(T c) { BorderType b(NDC_BORDER_CONST); b.v_ = to_value(c); return b; }
    static BorderType Reflect() { return BorderType(NDC_BORDER_REFLECT); }
    static BorderType Replicate() { return BorderType(NDC_BORDER_REPLICATE); }
    static BorderType Circular() { return BorderType(NDC_BORDER_CIRCULAR); }
    ndc_border raw() const { ndc_border b; std::memset(&b, 0, sizeof(b)); b.tag = tag_; b.value = v_; return b; }
private:
    explicit BorderType(int tag) : tag_(tag), v_{0, 0, 0} {}
    int tag_;
    ndc_value v_;
};

class ConvMode {
public:
    static ConvMode Full() { return ConvMode(NDC_CONV_FULL); }
    static ConvMode Same() { return ConvMode(NDC_CONV_SAME); }
    static ConvMode Valid() { return ConvMode(NDC_CONV_VALID); }
    static ConvMode Custom(std::vector<int64_t> padding, std::vector<int64_t> strides) {
        ConvMode m(NDC_CONV_CUSTOM);
        for (size_t i = 0; i < padding.size() && i < NDC_MAX_RANK; ++i) m.raw_.padding[i][0] = m.raw_.padding[i][1] = padding[i];
        for (size_t i = 0; i < strides.size() && i < NDC_MAX_RANK; ++i) m.raw_.strides[i] = strides[i];
        return m;
    }
    static ConvMode Explicit(std::vector<std::array<int64_t, 2>> padding, std::vector<int64_t> strides) {
        ConvMode m(NDC_CONV_EXPLICIT);
        for (size_t i = 0; i < padding.size() && i < NDC_MAX_RANK; ++i) {
            m.raw_.padding[i][0] = padding[i][0];
            m.raw_.padding[i][1] = padding[i][1];
        }
        for (size_t i = 0; i < strides.size() && i < NDC_MAX_RANK; ++i) m.raw_.strides[i] = strides[i];
        return m;
    }
    const ndc_conv_mode& raw() const { return raw_; }
private:
    explicit ConvMode(int tag) { std::memset(&raw_, 0, sizeof(raw_)); raw_.tag = tag; }
    ndc_conv_mode raw_;
};

class PaddingMode {
public:
    static PaddingMode Zeros() { return PaddingMode(NDC_PAD_ZEROS); }
    template <typename T> static PaddingMode Const(T c) { PaddingMode p(NDC_PAD_CONST); p.raw_.value = to_value(c); return p; }
    static PaddingMode Reflect() { return PaddingMode(NDC_PAD_REFLECT); }
    static PaddingMode Replicate() { return PaddingMode(NDC_PAD_REPLICATE); }
    static PaddingMode Circular() { return PaddingMode(NDC_PAD_CIRCULAR); }
    static PaddingMode Custom(std::vector<BorderType> b) {
        PaddingMode p(NDC_PAD_CUSTOM);
        for (size_t i = 0; i < b.size() && i < NDC_MAX_RANK; ++i) p.raw_.borders[i][0] = b[i].raw();
        return p;
    }
    static PaddingMode Explicit(std::vector<std::array<BorderType, 2>> b) {
        PaddingMode p(NDC_PAD_EXPLICIT);
        for (size_t i = 0; i < b.size() && i < NDC_MAX_RANK; ++i) {
            p.raw_.borders[i][0] = b[i][0].raw();
            p.raw_.borders[i][1] = b[i][1].raw();
        }
        return p;
    }
    const ndc_padding_mode& raw() const { return raw_; }
private:
    explicit PaddingMode(int tag) { std::memset(&raw_, 0, sizeof(raw_)); raw_.tag = tag; }
    ndc_padding_mode raw_;
};

template <typename T>
struct KernelWithDilation {
    const Array<T>* kernel;
    std::vector<int64_t> dilation;
    bool reverse_ = true;
    KernelWithDilation reverse() const { KernelWithDilation k = *this; k.reverse_ = true; return k; }
    KernelWithDilation no_reverse() const { KernelWithDilation k = *this; k.reverse_ = false; return k; }
};

template <typename T>
KernelWithDilation<T> with_dilation(const Array<T>& k, int64_t d) {
    return KernelWithDilation<T>{&k, std::vector<int64_t>(k.ndim(), d), true};
}
template <typename T>
KernelWithDilation<T> with_dilation(const Array<T>& k, std::vector<int64_t> d) {
    return KernelWithDilation<T>{&k, std::move(d), true};
}

template <typename T> struct RealOf { using type = T; };
template <typename R> struct RealOf<std::complex<R>> { using type = R; };
template <typename T> struct IsComplex : std::false_type {};
template <typename R> struct IsComplex<std::complex<R>> : std::true_type {};

// The elements of `a` in C order (a copy that honours any strides).
template <typename T>
std::vector<T> contiguous(const Array<T>& a) {
    int64_t n = 1;
    for (int64_t s : a.shape) n *= s;
    std::vector<T> out(static_cast<size_t>(n));
    std::vector<int64_t> idx(a.ndim(), 0);
    for (int64_t i = 0; i < n; ++i) {
        int64_t off = 0;
        for (size_t d = 0; d < a.ndim(); ++d) off += idx[d] * a.strides[d];
        out[static_cast<size_t>(i)] = a.ptr[off];
        for (int d = int(a.ndim()) - 1; d >= 0; --d) {
            if (++idx[d] < a.shape[d]) break;
            idx[d] = 0;
        }
    }
    return out;
}

// The reusable processor: a library handle (stream, workspace, plan caches),
// and Processor::forward / backward (processor/mod.rs:79-121).
class Processor {
public:
    Processor() { check(ndc_create(&h_)); }
    ~Processor() { if (h_) ndc_destroy(h_); }
    Processor(const Processor&) = delete;
    Processor& operator=(const Processor&) = delete;
    ndc_handle handle() const { return h_; }

    // The unnormalised N-d DFT (e^{-2 pi i}) of x, with the axes rotated left by
    // one: [F1, ..., F(N-1)', F0], where a real input keeps F(N-1)/2 + 1 bins of
    // its last axis (real.rs:105-155, complex.rs:33-88).
    template <typename In>
    Array<std::complex<typename RealOf<In>::type>> forward(const Array<In>& x) {
        static_assert(std::is_floating_point<typename RealOf<In>::type>::value, "FFT of a float type");
        using C = std::complex<typename RealOf<In>::type>;
        const size_t n = x.ndim();
        if (n == 0 || n > NDC_MAX_RANK) throw InvalidArgument("rank out of range");
        std::vector<In> src = contiguous(x);
        std::vector<int64_t> shape = x.shape;
        const int64_t last = IsComplex<In>::value ? shape[n - 1] : shape[n - 1] / 2 + 1;
        std::vector<int64_t> out_shape;
        for (size_t i = 1; i + 1 < n; ++i) out_shape.push_back(shape[i]);
        out_shape.push_back(last);
        if (n > 1) out_shape.push_back(shape[0]);
        int64_t total = 1;
        for (int64_t s : out_shape) total *= s;
        std::vector<C> out(static_cast<size_t>(total));
        check(ndc_fft_forward(h_, DType<In>::value, int32_t(n), shape.data(), src.data(), out.data(),
                              NDC_FLAG_HOST_MEMORY), h_);
        return Array<C>::from_vec(std::move(out_shape), std::move(out));
    }

    // The inverse of forward, divided by the number of output elements, in
    // natural axis order; a real Out restores the last axis to the length the
    // latest real forward recorded (real.rs:233-281, complex.rs:90-145).
    template <typename Out>
    Array<Out> backward(const Array<std::complex<typename RealOf<Out>::type>>& spec) {
        using C = std::complex<typename RealOf<Out>::type>;
        const size_t n = spec.ndim();
        if (n == 0 || n > NDC_MAX_RANK) throw InvalidArgument("rank out of range");
        std::vector<C> src = contiguous(spec);
        std::vector<int64_t> shape = spec.shape;
        std::vector<int64_t> out_shape;
        if (n > 1) out_shape.push_back(shape[n - 1]);
        for (size_t i = 0; i + 1 < n; ++i) out_shape.push_back(shape[i]);
        if (n == 1) out_shape.push_back(shape[0]);
        if (!IsComplex<Out>::value) {
            int64_t len = 0;
            check(ndc_fft_origin_len(h_, &len), h_);
            out_shape.back() = len;
        }
        int64_t total = 1;
        for (int64_t s : out_shape) total *= s;
        std::vector<Out> out(static_cast<size_t>(total));
        check(ndc_fft_backward(h_, DType<Out>::value, int32_t(n), shape.data(), src.data(), out.data(),
                               NDC_FLAG_HOST_MEMORY), h_);
        return Array<Out>::from_vec(std::move(out_shape), std::move(out));
    }

private:
    ndc_handle h_ = nullptr;
};

template <typename T>
ndc_conv_desc make_desc(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm,
                        int algo = NDC_ALGO_AUTO) {
    if (x.ndim() != k.kernel->ndim() || x.ndim() == 0 || x.ndim() > NDC_MAX_RANK || k.dilation.size() != x.ndim())
        throw InvalidArgument("rank mismatch");
    ndc_conv_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dtype = DType<T>::value;
    d.rank = int32_t(x.ndim());
    for (size_t i = 0; i < x.ndim(); ++i) {
        d.x_shape[i] = x.shape[i];
        d.x_strides[i] = x.strides[i];
        d.k_shape[i] = k.kernel->shape[i];
        d.k_strides[i] = k.kernel->strides[i];
        d.dilation[i] = k.dilation[i];
    }
    d.reverse = k.reverse_ ? 1 : 0;
    d.algo = algo;
    d.conv_mode = cm.raw();
    d.padding_mode = pm.raw();
    d.flags = NDC_FLAG_HOST_MEMORY | NDC_FLAG_X_STRIDED | NDC_FLAG_K_STRIDED;
    return d;
}

template <typename T>
Array<T> run(int op, const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm,
             Processor* proc, int algo) {
    ndc_conv_desc d = make_desc(x, k, cm, pm, algo);
    ndc_plan_info info;
    check(ndc_plan(&d, op, &info));
    std::vector<int64_t> shape(x.ndim());
    int64_t n = 1;
    for (size_t i = 0; i < x.ndim(); ++i) {
        shape[i] = info.out_shape[i];
        n *= shape[i];
    }
    std::vector<T> out(static_cast<size_t>(n));
    ndc_handle h = proc->handle();
    check(op == NDC_OP_CONV ? ndc_conv(h, &d, x.ptr, k.kernel->ptr, out.data())
                            : ndc_conv_fft(h, &d, x.ptr, k.kernel->ptr, out.data()), h);
    return Array<T>::from_vec(std::move(shape), std::move(out));
}

// ConvExt::conv
template <typename T>
Array<T> conv(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm,
              Processor& proc, int algo = NDC_ALGO_AUTO) {
    return run(NDC_OP_CONV, x, k, cm, pm, &proc, algo);
}
template <typename T>
Array<T> conv(const Array<T>& x, const Array<T>& k, const ConvMode& cm, const PaddingMode& pm, Processor& proc) {
    return conv(x, with_dilation(k, 1), cm, pm, proc);
}

// ConvFFTExt::conv_fft_with_processor (conv_fft and conv_fft_par are the same
// call with a processor the caller does not keep)
template <typename T>
Array<T> conv_fft_with_processor(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm,
                                 const PaddingMode& pm, Processor& proc) {
    return run(NDC_OP_CONV_FFT, x, k, cm, pm, &proc, NDC_ALGO_AUTO);
}
template <typename T>
Array<T> conv_fft(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm) {
    Processor proc;
    return conv_fft_with_processor(x, k, cm, pm, proc);
}
template <typename T>
Array<T> conv_fft(const Array<T>& x, const Array<T>& k, const ConvMode& cm, const PaddingMode& pm) {
    return conv_fft(x, with_dilation(k, 1), cm, pm);
}

// Shape arithmetic only (no device needed).
template <typename T>
ndc_plan_info plan(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm,
                   bool fft) {
    ndc_conv_desc d = make_desc(x, k, cm, pm);
    ndc_plan_info info;
    check(ndc_plan(&d, fft ? NDC_OP_CONV_FFT : NDC_OP_CONV, &info));
    return info;
}

// A kernel spectrum kept for repeated conv_fft of data shaped like `x` with one
// kernel: the sketched Baked / conv_fft_bake / conv_fft_with_baked
// (conv_fft/mod.rs:22-42,175-182,306-392).  The processor must outlive it.
template <typename T>
class Baked {
public:
    Baked(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm,
          Processor& proc)
        : proc_(&proc), desc_(make_desc(x, k, cm, pm)) {
        ndc_plan_info info;
        check(ndc_plan(&desc_, NDC_OP_CONV_FFT, &info));
        out_shape_.assign(info.out_shape, info.out_shape + x.ndim());
        check(ndc_conv_fft_bake(proc.handle(), &desc_, k.kernel->ptr, &b_), proc.handle());
    }
    Baked(Baked&& o) noexcept : proc_(o.proc_), desc_(o.desc_), out_shape_(std::move(o.out_shape_)), b_(o.b_) {
        o.b_ = nullptr;
    }
    Baked(const Baked&) = delete;
    Baked& operator=(const Baked&) = delete;
    Baked& operator=(Baked&&) = delete;
    ~Baked() { if (b_) ndc_baked_destroy(b_); }

    Array<T> conv(const Array<T>& x) {
        bool same = int32_t(x.ndim()) == desc_.rank;
        for (int32_t i = 0; same && i < desc_.rank; ++i)
            same = x.shape[i] == desc_.x_shape[i] && x.strides[i] == desc_.x_strides[i];
        if (!same) throw InvalidArgument("data must have the shape and strides the kernel was baked for");
        int64_t total = 1;
        for (int64_t s : out_shape_) total *= s;
        std::vector<T> out(static_cast<size_t>(total));
        check(ndc_conv_fft_with_baked(b_, x.ptr, out.data()), proc_->handle());
        return Array<T>::from_vec(out_shape_, std::move(out));
    }

private:
    Processor* proc_;
    ndc_conv_desc desc_;
    std::vector<int64_t> out_shape_;
    ndc_baked b_ = nullptr;
};

template <typename T>
Baked<T> conv_fft_bake(const Array<T>& x, const KernelWithDilation<T>& k, const ConvMode& cm, const PaddingMode& pm,
                       Processor& proc) {
    return Baked<T>(x, k, cm, pm, proc);
}
template <typename T>
Array<T> conv_fft_with_baked(const Array<T>& x, Baked<T>& baked) {
    return baked.conv(x);
}

// Many independent rank-1 f32 signal/kernel pairs of uneven lengths, each with
// the semantics of one conv_fft call (conv_fft/mod.rs:113-171) under the shared
// ConvMode, PaddingMode, dilation and reverse flag; signals, kernels and outputs
// are packed back to back in pair order.  The processor must outlive it.
class Batch {
public:
    Batch(Processor& proc, std::vector<int64_t> x_len, std::vector<int64_t> k_len, const ConvMode& cm,
          const PaddingMode& pm, int64_t dilation = 1, bool reverse = true, int algo = NDC_BATCH_AUTO)
        : proc_(&proc), x_len_(std::move(x_len)), k_len_(std::move(k_len)) {
        if (x_len_.size() != k_len_.size()) throw InvalidArgument("x_len and k_len differ in length");
        ndc_batch_desc d;
        std::memset(&d, 0, sizeof(d));
        d.dtype = NDC_F32;
        d.algo = algo;
        d.count = int64_t(x_len_.size());
        d.x_len = x_len_.data();
        d.k_len = k_len_.data();
        d.dilation = dilation;
        d.reverse = reverse ? 1 : 0;
        d.flags = NDC_FLAG_HOST_MEMORY;
        d.conv_mode = cm.raw();
        d.padding_mode = pm.raw();
        check(ndc_batch_create(proc.handle(), &d, &b_), proc.handle());
        y_len_.resize(x_len_.size());
        check(ndc_batch_out_lens(b_, y_len_.data(), &x_total_, &k_total_, &y_total_), proc.handle());
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    ~Batch() { if (b_) ndc_batch_destroy(b_); }

    const std::vector<int64_t>& out_lens() const { return y_len_; }
    int64_t x_total() const { return x_total_; }
    int64_t k_total() const { return k_total_; }
    int64_t y_total() const { return y_total_; }

    // The packed outputs of the packed signals x and kernels k.
    std::vector<float> run(const std::vector<float>& x, const std::vector<float>& k, bool graph = false) {
        if (int64_t(x.size()) != x_total_ || int64_t(k.size()) != k_total_)
            throw InvalidArgument("packed sizes differ from the batch");
        std::vector<float> y(static_cast<size_t>(y_total_));
        check(graph ? ndc_batch_run_graph(b_, x.data(), k.data(), y.data())
                    : ndc_batch_run(b_, x.data(), k.data(), y.data()), proc_->handle());
        return y;
    }

    // JSON: cost constants, the choice per length bucket, item counts.
    std::string report() const {
        std::vector<char> buf(1 << 16);
        check(ndc_batch_report(b_, buf.data(), int64_t(buf.size())), proc_->handle());
        return std::string(buf.data());
    }

private:
    Processor* proc_;
    std::vector<int64_t> x_len_, k_len_, y_len_;
    int64_t x_total_ = 0, k_total_ = 0, y_total_ = 0;
    ndc_batch b_ = nullptr;
};

}  // namespace ndc
