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
// ConvFFTExt::conv_fft / conv_fft_with_processor (conv_fft/mod.rs:113-157).
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

// The reusable processor: a library handle (stream, workspace, plan caches).
class Processor {
public:
    Processor() { check(ndc_create(&h_)); }
    ~Processor() { if (h_) ndc_destroy(h_); }
    Processor(const Processor&) = delete;
    Processor& operator=(const Processor&) = delete;
    ndc_handle handle() const { return h_; }
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

}  // namespace ndc
