// Exercises the C++ mirror of the crate's API (ndarray_conv.hpp) through the
// C ABI.  Plan and error checks run anywhere; with a GPU present the
// convolutions are compared with a brute-force loop (integers: exact), the
// processor's transforms with the DFT's definition and a round trip, and the
// baked and batch calls with conv_fft.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ndarray_conv.hpp"

static int failures = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            ++failures;                                                  \
        }                                                                \
    } while (0)

template <typename E, typename F>
bool throws(F&& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

// brute force of out[o] = sum_t pad(o*s + t*d) * k_rev[t] for 1-D zero padding
static std::vector<int32_t> brute_1d(const std::vector<int32_t>& x, const std::vector<int32_t>& k, int p0, int p1,
                                     int s, int d) {
    const int n = int(x.size()), m = int(k.size());
    const int kd = m * d - d + 1;
    const int P = n + p0 + p1;
    const int out = (P - kd) / s + 1;
    std::vector<int32_t> y(out, 0);
    for (int o = 0; o < out; ++o)
        for (int t = 0; t < m; ++t) {
            const int p = o * s + t * d - p0;
            if (p >= 0 && p < n) y[o] += x[p] * k[m - 1 - t];
        }
    return y;
}

int main() {
    using namespace ndc;
    auto x = Array<int32_t>::from_vec({5}, {1, 2, 3, 4, 5});
    auto k = Array<int32_t>::from_vec({3}, {1, 2, 1});
    // shape arithmetic (conv/mod.rs:28-66, 162-167)
    auto info = plan(x, with_dilation(k, 1), ConvMode::Same(), PaddingMode::Zeros(), false);
    EXPECT(info.out_shape[0] == 5 && info.padding[0][0] == 1 && info.padding[0][1] == 1);
    auto xf = Array<double>::from_vec({5}, {1, 2, 3, 4, 5});
    auto kf = Array<double>::from_vec({3}, {1, 2, 1});
    info = plan(xf, with_dilation(kf, 2), ConvMode::Full(), PaddingMode::Zeros(), true);
    EXPECT(info.kernel_dilated[0] == 5 && info.out_shape[0] == 9 && info.fft_size[0] == 16);
    // integer element types have no FFT path (processor/mod.rs:46-52)
    EXPECT(throws<Unsupported>([&] { plan(x, with_dilation(k, 1), ConvMode::Full(), PaddingMode::Zeros(), true); }));
    // errors (lib.rs:145-159) and panics
    auto empty = Array<int32_t>::from_vec({0}, {});
    EXPECT(throws<DataShape>([&] { plan(empty, with_dilation(k, 1), ConvMode::Same(), PaddingMode::Zeros(), false); }));
    EXPECT(throws<KernelShape>([&] { plan(x, with_dilation(empty, 1), ConvMode::Same(), PaddingMode::Zeros(), false); }));
    auto emptyf = Array<double>::from_vec({0}, {});
    EXPECT(throws<DataShape>([&] { plan(xf, with_dilation(emptyf, 1), ConvMode::Same(), PaddingMode::Zeros(), true); }));
    auto big = Array<int32_t>::from_vec({7}, {1, 1, 1, 1, 1, 1, 1});
    EXPECT(throws<MismatchShape>([&] { plan(x, with_dilation(big, 1), ConvMode::Valid(), PaddingMode::Zeros(), false); }));
    EXPECT(throws<InvalidArgument>([&] {
        plan(x, with_dilation(k, 1), ConvMode::Custom({1}, {0}), PaddingMode::Zeros(), false);
    }));
    EXPECT(throws<InvalidArgument>([&] {
        plan(x, with_dilation(k, 1), ConvMode::Explicit({{9, 0}}, {1}), PaddingMode::Reflect(), false);
    }));

    // strided copy (a reversed view of a 2x3 array)
    std::vector<int32_t> buf{1, 2, 3, 4, 5, 6};
    auto rv = Array<int32_t>::view({2, 3}, {-3, -1}, buf.data() + 5);
    EXPECT((contiguous(rv) == std::vector<int32_t>{6, 5, 4, 3, 2, 1}));

    ndc_handle probe = nullptr;
    if (ndc_create(&probe) != NDC_OK) {
        std::printf("no device: plan and error checks only (%d failures)\n", failures);
        return failures ? 1 : 0;
    }
    ndc_destroy(probe);
    Processor proc;
    // conv/tests.rs:123-143 (Same, torch-checked) and the edge cases :556-583
    auto y = conv(x, k, ConvMode::Same(), PaddingMode::Zeros(), proc);
    EXPECT((y.data() == std::vector<int32_t>{4, 8, 12, 16, 14}));
    auto s42 = Array<int32_t>::from_vec({1}, {42});
    auto s2 = Array<int32_t>::from_vec({1}, {2});
    EXPECT((conv(s42, s2, ConvMode::Same(), PaddingMode::Zeros(), proc).data() == std::vector<int32_t>{84}));
    // strided / dilated / no_reverse against brute force
    std::srand(7);
    for (int trial = 0; trial < 50; ++trial) {
        const int n = 1 + std::rand() % 60, m = 1 + std::rand() % 9;
        const int d = 1 + std::rand() % 3, s = 1 + std::rand() % 3;
        const int p0 = std::rand() % 6, p1 = std::rand() % 6;
        std::vector<int32_t> xv(n), kv(m);
        for (auto& v : xv) v = std::rand() % 17 - 8;
        for (auto& v : kv) v = std::rand() % 17 - 8;
        if (n + p0 + p1 < m * d - d + 1) continue;
        auto xa = Array<int32_t>::from_vec({n}, xv);
        auto ka = Array<int32_t>::from_vec({m}, kv);
        auto got = conv(xa, with_dilation(ka, d), ConvMode::Explicit({{p0, p1}}, {s}), PaddingMode::Zeros(), proc);
        EXPECT(got.data() == brute_1d(xv, kv, p0, p1, s, d));
        auto gf = conv_fft_with_processor(Array<double>::from_vec({n}, std::vector<double>(xv.begin(), xv.end())),
                                          with_dilation(Array<double>::from_vec({m}, std::vector<double>(kv.begin(), kv.end())), d),
                                          ConvMode::Explicit({{p0, p1}}, {s}), PaddingMode::Zeros(), proc);
        auto ref = brute_1d(xv, kv, p0, p1, s, d);
        for (size_t i = 0; i < ref.size(); ++i) EXPECT(std::llround(gf.data()[i]) == ref[i]);
    }

    // Processor::forward of a real 3x5 array against the DFT's definition; the
    // spectrum's axes come rotated left ([5/2+1, 3]); backward restores x.
    {
        const double pi = std::acos(-1.0);
        std::vector<double> xv(15);
        for (int i = 0; i < 15; ++i) xv[i] = std::sin(1.7 * i) + 0.1 * i;
        auto xr = Array<double>::from_vec({3, 5}, xv);
        Processor p;
        auto spec = p.forward(xr);
        EXPECT((spec.shape == std::vector<int64_t>{3, 3}));
        double err = 0;
        for (int k1 = 0; k1 < 3; ++k1)
            for (int k0 = 0; k0 < 3; ++k0) {
                std::complex<double> acc = 0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 5; ++b)
                        acc += xv[a * 5 + b] * std::polar(1.0, -2 * pi * (double(k0 * a) / 3 + double(k1 * b) / 5));
                err = std::max(err, std::abs(spec.data()[k1 * 3 + k0] - acc));
            }
        EXPECT(err < 1e-9);
        auto back = p.backward<double>(spec);
        EXPECT((back.shape == std::vector<int64_t>{3, 5}));
        err = 0;
        for (int i = 0; i < 15; ++i) err = std::max(err, std::abs(back.data()[i] - xv[i]));
        EXPECT(err < 1e-12);
        // complex round trip, rank 1, length 12
        std::vector<std::complex<float>> cv(12);
        for (int i = 0; i < 12; ++i) cv[i] = {float(i % 5) - 2.0f, float(i % 3)};
        auto cs = p.forward(Array<std::complex<float>>::from_vec({12}, cv));
        auto cb = p.backward<std::complex<float>>(cs);
        float cerr = 0;
        for (int i = 0; i < 12; ++i) cerr = std::max(cerr, std::abs(cb.data()[i] - cv[i]));
        EXPECT(cb.shape == std::vector<int64_t>{12} && cerr < 1e-5f);
    }

    // conv_fft_with_baked equals conv_fft on the same data, twice in a row;
    // data of another shape is refused.
    {
        std::vector<float> xv(7 * 9), kv(3 * 4);
        for (auto& v : xv) v = float(std::rand() % 200) / 100.0f - 1.0f;
        for (auto& v : kv) v = float(std::rand() % 200) / 100.0f - 1.0f;
        auto xa = Array<float>::from_vec({7, 9}, xv);
        auto ka = Array<float>::from_vec({3, 4}, kv);
        auto kd = with_dilation(ka, 2);
        auto baked = conv_fft_bake(xa, kd, ConvMode::Same(), PaddingMode::Reflect(), proc);
        for (int rep = 0; rep < 2; ++rep) {
            for (auto& v : xv) v = float(std::rand() % 200) / 100.0f - 1.0f;
            auto xb = Array<float>::from_vec({7, 9}, xv);
            auto got = conv_fft_with_baked(xb, baked);
            auto ref = conv_fft_with_processor(xb, kd, ConvMode::Same(), PaddingMode::Reflect(), proc);
            EXPECT(got.shape == ref.shape);
            float e = 0;
            for (size_t i = 0; i < ref.data().size(); ++i) e = std::max(e, std::abs(got.data()[i] - ref.data()[i]));
            EXPECT(e < 1e-5f);
        }
        auto other = Array<float>::from_vec({9, 7}, xv);
        EXPECT(throws<InvalidArgument>([&] { conv_fft_with_baked(other, baked); }));
    }

    // a batch of uneven rank-1 pairs equals one conv_fft per pair (direct FP32
    // and the FFT form, eager and through a graph)
    {
        std::vector<int64_t> xl{1, 37, 300, 5000, 64}, kl{1, 5, 31, 129, 64};
        std::vector<float> xp, kp;
        for (int64_t n : xl)
            for (int64_t i = 0; i < n; ++i) xp.push_back(float(std::rand() % 200) / 100.0f - 1.0f);
        for (int64_t n : kl)
            for (int64_t i = 0; i < n; ++i) kp.push_back(float(std::rand() % 200) / 100.0f - 1.0f);
        for (int algo : {int(NDC_BATCH_DIRECT), int(NDC_BATCH_FFT)}) {
            Batch b(proc, xl, kl, ConvMode::Same(), PaddingMode::Circular(), 1, true, algo);
            EXPECT(b.x_total() == int64_t(xp.size()) && b.k_total() == int64_t(kp.size()));
            auto y = b.run(xp, kp);
            auto yg = b.run(xp, kp, true);
            EXPECT(y == yg);
            int64_t xo = 0, ko = 0, yo = 0;
            for (size_t i = 0; i < xl.size(); ++i) {
                auto xi = Array<float>::from_vec({xl[i]}, std::vector<float>(xp.begin() + xo, xp.begin() + xo + xl[i]));
                auto ki = Array<float>::from_vec({kl[i]}, std::vector<float>(kp.begin() + ko, kp.begin() + ko + kl[i]));
                auto ref = conv_fft_with_processor(xi, with_dilation(ki, 1), ConvMode::Same(), PaddingMode::Circular(),
                                                   proc);
                EXPECT(int64_t(ref.data().size()) == b.out_lens()[i]);
                float e = 0, m = 1e-30f;
                for (int64_t j = 0; j < b.out_lens()[i]; ++j) {
                    e = std::max(e, std::abs(y[size_t(yo + j)] - ref.data()[size_t(j)]));
                    m = std::max(m, std::abs(ref.data()[size_t(j)]));
                }
                EXPECT(e <= 1e-5f * m + 1e-6f);
                xo += xl[i];
                ko += kl[i];
                yo += b.out_lens()[i];
            }
        }
        EXPECT(throws<InvalidArgument>([&] { Batch(proc, {3}, {1, 2}, ConvMode::Same(), PaddingMode::Zeros()); }));
        // conv_fft reports an empty kernel as DataShape (conv_fft/mod.rs:210-213)
        EXPECT(throws<DataShape>([&] { Batch(proc, {3, 4}, {1, 0}, ConvMode::Same(), PaddingMode::Zeros()); }));
    }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
