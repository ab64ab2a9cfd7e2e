// Exercises the C++ mirror of the crate's API (ndarray_conv.hpp) through the
// C ABI.  Plan and error checks run anywhere; with a GPU present the
// convolutions are compared with a brute-force loop (integers: exact).
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
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
