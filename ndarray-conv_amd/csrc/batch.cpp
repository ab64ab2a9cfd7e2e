// Host side of the batched rank-1 path: per-pair validation through the
// single-problem planner, the cost model and its calibration on the device,
// work-queue construction (longest items first), graph capture, and the
// extern "C" entry points declared in ndconv.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "batch.hpp"
#include "border1d.hpp"
#include "plan.hpp"

namespace ndc {
namespace {

struct DevMem {
    void* p = nullptr;
    size_t n = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    bool ensure(size_t bytes) {
        if (bytes == 0) bytes = 256;
        if (p && n >= bytes) return true;
        release();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = bytes;
        return true;
    }
    bool upload(const void* src, size_t bytes) {
        if (!ensure(bytes)) return false;
        return bytes == 0 || hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    }
};

// Throughput costs (ms per item with the device full of workgroups).
struct CostModel {
    double fft_item[kBatchClasses];
    double dir_a, dir_b;     // FP32 direct item: dir_a + dir_b * outputs * taps
    double mfma_a, mfma_b;   // BF16x3 matrix-core item: mfma_a + mfma_b * K-blocks
    double mfma6_a, mfma6_b; // BF16x6 matrix-core item, likewise
    bool measured = false;
};

// Starting point before calibration, and what the calibration falls back to.
CostModel default_costs() {
    CostModel c;
    for (int i = 0; i < kBatchClasses; ++i) c.fft_item[i] = 1.2e-5;
    c.dir_a = 2e-6;
    c.dir_b = 6e-11;
    c.mfma_a = 4e-6;
    c.mfma_b = 4e-7;
    c.mfma6_a = 4e-6;
    c.mfma6_b = 8e-7;
    return c;
}

// Shape arithmetic of one pair (from the single-problem planner).
struct PairShape {
    int64_t n, k, p0, pds, out, kdil;
};

int direct_tile(int s) {
    const int cap = (kBatchWindow / 2 - 1) / s + 1;
    return std::min(kBatchDirectTile, std::max(cap, 1));
}

int64_t fft_per(int cls, int64_t kdil, int s) {
    const int64_t L = 2 * int64_t(batch_lc(cls));
    if (L < kdil) return 0;
    return (L - kdil) / s + 1;
}

double direct_cost(const CostModel& cm, const PairShape& ps, int s) {
    const int64_t tile = direct_tile(s);
    const int64_t nt = (ps.out + tile - 1) / tile;
    return double(nt) * cm.dir_a + cm.dir_b * double(ps.out) * double(ps.k);
}

double fft_cost(const CostModel& cm, const PairShape& ps, int s, int cls) {
    const int64_t per = fft_per(cls, ps.kdil, s);
    if (per <= 0) return INFINITY;
    const int64_t lines = (ps.out + per - 1) / per;
    // one more line's worth (forward only) for the kernel spectrum
    return (double(lines) + 0.5) / double(batch_lines(cls)) * cm.fft_item[cls];
}

constexpr int kMethodDirect = -1;   // FP32 direct items
constexpr int kMethodMfma = -2;     // BF16x3 matrix-core items
constexpr int kMethodMfma6 = -3;    // BF16x6 matrix-core items

int mfma_nkb(const PairShape& ps) { return int((ps.kdil + 31 + 15) / 16); }

bool mfma_eligible(const PairShape& ps, int s) { return s == 1 && ps.kdil <= kBatchMfmaMaxKdil; }

// BF16x3 meets the batch's accuracy only where every output sums enough
// products of distinct samples for their rounding errors to average (see
// kBf16x3MinTaps): the kernel needs that many taps, and no window may read one
// sample many times over, which a signal shorter than the dilated kernel (a
// circular pad wraps it) or a border repeating one value (Replicate, a
// nonzero constant) would do.  The three-term split has no such limit.
bool bf16x3_admitted(const PairShape& ps, bool repeating_border) {
    return ps.k >= kBf16x3MinTaps && ps.n >= ps.kdil && !repeating_border;
}

double mfma_item_cost(const CostModel& cm, const PairShape& ps, bool six) {
    return six ? cm.mfma6_a + cm.mfma6_b * double(mfma_nkb(ps)) : cm.mfma_a + cm.mfma_b * double(mfma_nkb(ps));
}

double mfma_cost(const CostModel& cm, const PairShape& ps, int s, bool six) {
    if (!mfma_eligible(ps, s)) return INFINITY;
    const int64_t nt = (ps.out + kBatchMfmaTile - 1) / kBatchMfmaTile;
    return double(nt) * mfma_item_cost(cm, ps, six);
}

// Method per pair: kMethodDirect, kMethodMfma, kMethodMfma6, or an FFT class.
int choose(const CostModel& cm, const PairShape& ps, int s, int algo, bool repeating_border) {
    int best_cls = kMethodDirect;
    double best = INFINITY;
    for (int c = 0; c < kBatchClasses; ++c) {
        const double v = fft_cost(cm, ps, s, c);
        if (v < best) {
            best = v;
            best_cls = c;
        }
    }
    if (algo == NDC_BATCH_DIRECT) return kMethodDirect;
    if (algo == NDC_BATCH_FFT) return best_cls;   // direct when no class holds the dilated kernel
    if (algo == NDC_BATCH_MFMA && mfma_eligible(ps, s))
        return bf16x3_admitted(ps, repeating_border) ? kMethodMfma : kMethodMfma6;
    if (algo == NDC_BATCH_BF16X6 && mfma_eligible(ps, s)) return kMethodMfma6;
    const double vd = direct_cost(cm, ps, s);
    int pick = vd <= best ? kMethodDirect : best_cls;
    if (algo == NDC_BATCH_AUTO_BF16X3) {
        const bool x3 = bf16x3_admitted(ps, repeating_border);
        const double vm = mfma_cost(cm, ps, s, !x3);
        if (vm < std::min(best, vd)) pick = x3 ? kMethodMfma : kMethodMfma6;
    }
    return pick;
}

struct Schedule {
    std::vector<BatchPair> pairs;
    std::vector<BatchItem> items;
    std::vector<BatchLine> lines;
    std::vector<int32_t> kpairs;
    std::vector<BatchKItem> kitems;
    int64_t ks_total = 0;
    int32_t nfirst = 0;            // direct and FFT items, ahead of the matrix-core items
    double est_ms = 0;
    double kspec_bytes = 0, main_bytes = 0, main_flops = 0;
};

// Builds the work queue.  offs[i] = {x, k, y} element offsets of pair i.
Schedule build_schedule(const std::vector<PairShape>& sh, const std::vector<std::array<int64_t, 3>>& offs,
                        const std::vector<int>& method, int s, const CostModel& cm) {
    Schedule S;
    const size_t P = sh.size();
    S.pairs.resize(P);
    std::vector<std::vector<BatchLine>> cls_lines(kBatchClasses);
    std::vector<std::vector<int32_t>> cls_kpairs(kBatchClasses);
    std::vector<std::pair<double, BatchItem>> work;
    const int tile = direct_tile(s);
    // a direct or matrix-core tile record with its pair's metadata
    auto tile_item = [&](int kind, size_t i, int64_t o, int cnt) {
        const PairShape& ps = sh[i];
        BatchItem it{};
        it.kind = kind;
        it.a = int32_t(i);
        it.b = int32_t(o);
        it.c = cnt;
        it.x_off = offs[i][0];
        it.k_off = offs[i][1];
        it.y_off = offs[i][2];
        it.n = int32_t(ps.n);
        it.k = int32_t(ps.k);
        it.p0 = int32_t(ps.p0);
        it.pds = int32_t(ps.pds);
        return it;
    };
    for (size_t i = 0; i < P; ++i) {
        const PairShape& ps = sh[i];
        BatchPair& bp = S.pairs[i];
        bp.x_off = offs[i][0];
        bp.k_off = offs[i][1];
        bp.y_off = offs[i][2];
        bp.n = int32_t(ps.n);
        bp.k = int32_t(ps.k);
        bp.p0 = int32_t(ps.p0);
        bp.pds = int32_t(ps.pds);
        bp.out = int32_t(ps.out);
        bp.cls = method[i];
        bp.ks_off = 0;
        bp.per = 0;
        S.main_bytes += 4.0 * double(ps.n + ps.out);
        if (ps.out <= 0) continue;
        if (method[i] == kMethodDirect) {
            S.main_bytes += 4.0 * double(ps.k);
            S.main_flops += 2.0 * double(ps.out) * double(ps.k);
            for (int64_t o = 0; o < ps.out; o += tile) {
                const int cnt = int(std::min<int64_t>(tile, ps.out - o));
                work.emplace_back(cm.dir_a + cm.dir_b * double(cnt) * double(ps.k), tile_item(0, i, o, cnt));
            }
        } else if (method[i] == kMethodMfma || method[i] == kMethodMfma6) {
            const bool six = method[i] == kMethodMfma6;
            S.main_bytes += 4.0 * double(ps.k);
            S.main_flops += 2.0 * double(ps.out) * double(ps.k);
            const double c = mfma_item_cost(cm, ps, six);
            const int kind = six ? kBatchKindMfma6 : kBatchKindMfma;
            for (int64_t o = 0; o < ps.out; o += kBatchMfmaTile) {
                const int cnt = int(std::min<int64_t>(kBatchMfmaTile, ps.out - o));
                work.emplace_back(c, tile_item(kind, i, o, cnt));
            }
        } else {
            const int c = method[i];
            const int LC = batch_lc(c);
            bp.per = int32_t(fft_per(c, ps.kdil, s));
            bp.ks_off = S.ks_total;
            S.ks_total += LC + 1;
            cls_kpairs[c].push_back(int32_t(i));
            S.kspec_bytes += 4.0 * double(ps.k) + 8.0 * double(LC + 1);
            S.main_bytes += 8.0 * double(LC + 1);   // the kernel spectrum, read once (its blocks share it in cache)
            for (int64_t o = 0; o < ps.out; o += bp.per) {
                BatchLine ln{};
                ln.x_off = bp.x_off;
                ln.y_off = bp.y_off;
                ln.ks_off = bp.ks_off;
                ln.n = bp.n;
                ln.k = bp.k;
                ln.p0 = bp.p0;
                ln.pds = bp.pds;
                ln.out = bp.out;
                ln.per = bp.per;
                ln.o0 = int32_t(o);
                ln.pair = int32_t(i);
                cls_lines[c].push_back(ln);
            }
        }
    }
    for (int c = 0; c < kBatchClasses; ++c) {
        const int G = batch_lines(c);
        const int32_t base = int32_t(S.lines.size());
        const int32_t nl = int32_t(cls_lines[c].size());
        S.lines.insert(S.lines.end(), cls_lines[c].begin(), cls_lines[c].end());
        for (int32_t f = 0; f < nl; f += G) {
            BatchItem it{};
            it.kind = 1 + c;
            it.a = base + f;
            it.b = std::min(G, nl - f);
            work.emplace_back(cm.fft_item[c], it);
        }
        const int32_t kb = int32_t(S.kpairs.size());
        const int32_t nk = int32_t(cls_kpairs[c].size());
        S.kpairs.insert(S.kpairs.end(), cls_kpairs[c].begin(), cls_kpairs[c].end());
        for (int32_t f = 0; f < nk; f += G) S.kitems.push_back(BatchKItem{c, kb + f, std::min(G, nk - f), 0});
    }
    // longest items first: the queue hands them out in this order, so the
    // tail of the launch is made of the shortest items.  The direct and FFT
    // items form the first range, the matrix-core items the second.
    std::stable_sort(work.begin(), work.end(),
                     [](const std::pair<double, BatchItem>& a, const std::pair<double, BatchItem>& b) {
                         return a.first > b.first;
                     });
    const auto first_end = std::stable_partition(work.begin(), work.end(), [](const std::pair<double, BatchItem>& w) {
        return w.second.kind != kBatchKindMfma && w.second.kind != kBatchKindMfma6;
    });
    S.nfirst = int32_t(first_end - work.begin());
    S.items.reserve(work.size());
    for (auto& w : work) {
        S.items.push_back(w.second);
        S.est_ms += w.first;
    }
    return S;
}

// Device-resident copy of a schedule plus the launch arguments.
struct DevSchedule {
    DevMem pairs, items, lines, klines, kitems, kspec, counter, consts;
    BatchArgs args{};
    int grid = 0;
    // Border sides of every pair, and the constant cells the kernels read for
    // pad positions that hold a constant.
    bool set_borders(int32_t front, int32_t back, float cfront, float cback) {
        const float cells[4] = {0.f, cfront, cback, 0.f};
        if (!consts.upload(cells, sizeof(cells))) return false;
        args.consts = static_cast<const float*>(consts.p);
        args.front = front;
        args.back = back;
        args.cfront = cfront;
        args.cback = cback;
        return true;
    }
    bool upload(const Schedule& S, const float2* tw, int blocks) {
        std::vector<BatchKLine> kl(S.kpairs.size());
        for (size_t i = 0; i < kl.size(); ++i) {
            const BatchPair& p = S.pairs[size_t(S.kpairs[i])];
            kl[i] = BatchKLine{p.k_off, p.ks_off, p.k, 0};
        }
        if (!pairs.upload(S.pairs.data(), S.pairs.size() * sizeof(BatchPair)) ||
            !items.upload(S.items.data(), S.items.size() * sizeof(BatchItem)) ||
            !lines.upload(S.lines.data(), S.lines.size() * sizeof(BatchLine)) ||
            !klines.upload(kl.data(), kl.size() * sizeof(BatchKLine)) ||
            !kitems.upload(S.kitems.data(), S.kitems.size() * sizeof(BatchKItem)) ||
            !kspec.ensure(size_t(S.ks_total) * sizeof(float2)) || !counter.ensure(size_t(kBatchHeads) * kBatchQueueStride * sizeof(int32_t)))
            return false;
        args.pairs = static_cast<const BatchPair*>(pairs.p);
        args.items = static_cast<const BatchItem*>(items.p);
        args.lines = static_cast<const BatchLine*>(lines.p);
        args.klines = static_cast<const BatchKLine*>(klines.p);
        args.kitems = static_cast<const BatchKItem*>(kitems.p);
        args.kspec = static_cast<float2*>(kspec.p);
        args.counter = static_cast<int32_t*>(counter.p);
        args.tw = tw;
        args.nitems = int32_t(S.items.size());
        args.nfirst = S.nfirst;
        args.nkitems = int32_t(S.kitems.size());
        grid = int(std::min<int64_t>(int64_t(S.items.size()), int64_t(blocks)));
        return true;
    }
};

// Quarter-circle tables W_L^r (r < L/4, L = 2 LC) of every class, FP64 on the
// host, concatenated at batch_tw_offset(c).
const float2* batch_twiddles(LaunchContext& ctx) {
    std::vector<float> tab;
    for (int c = 0; c < kBatchClasses; ++c) {
        const int LC = batch_lc(c);
        const double L = 2.0 * LC;
        for (int r = 0; r < LC / 2; ++r) {
            const double ang = -2.0 * M_PI * double(r) / L;
            tab.push_back(float(std::cos(ang)));
            tab.push_back(float(std::sin(ang)));
        }
    }
    return static_cast<const float2*>(ctx.constant_table("batch_tw", tab.data(), tab.size() * sizeof(float)));
}

int resident_blocks(LaunchContext& ctx) { return batch_blocks_per_cu() * ctx.cu_count(); }

// Times `reps` runs of a schedule on synthetic data (ms per run).
bool time_schedule(LaunchContext& ctx, const Schedule& S, const float* x, const float* k, float* y, int s, int d,
                   double* ms) {
    DevSchedule D;
    const float2* tw = batch_twiddles(ctx);
    if (!tw || !D.upload(S, tw, resident_blocks(ctx))) return false;
    D.args.x = x;
    D.args.k = k;
    D.args.y = y;
    D.args.s = s;
    D.args.d = d;
    D.args.dmul = batch_dmul(d);
    D.args.rev = 1;
    if (!D.set_borders(kS1Circular, kS1Circular, 0.f, 0.f)) return false;
    hipStream_t st = ctx.stream();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return false;
    bool ok = launch_batch(D.args, D.grid, ctx, st, false, 0, 0, 0) == NDC_OK;
    constexpr int reps = 3;
    ok = ok && hipEventRecord(e0, st) == hipSuccess;
    for (int r = 0; ok && r < reps; ++r) ok = launch_batch(D.args, D.grid, ctx, st, false, 0, 0, 0) == NDC_OK;
    ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float t = 0;
    ok = ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = double(t) / reps;
    return ok;
}

// Measures the cost model on the device: synthetic batches of 96 signals of
// 2^18 samples, one method each (every FFT class; direct at two tap counts).
bool calibrate(LaunchContext& ctx, int s, int d, CostModel* cm) {
    constexpr int P = 96;
    constexpr int64_t N = 1 << 18;
    float* x = static_cast<float*>(ctx.scratch(0, size_t(P) * N * sizeof(float)));
    float* y = static_cast<float*>(ctx.scratch(1, size_t(P) * N * sizeof(float)));
    float* k = static_cast<float*>(ctx.scratch(2, size_t(P) * 1024 * sizeof(float)));
    if (!x || !y || !k) return false;
    if (hipMemsetAsync(x, 0, size_t(P) * N * sizeof(float), ctx.stream()) != hipSuccess ||
        hipMemsetAsync(k, 0, size_t(P) * 1024 * sizeof(float), ctx.stream()) != hipSuccess)
        return false;
    auto shapes = [&](int64_t kl) {
        std::vector<PairShape> sh(P);
        std::vector<std::array<int64_t, 3>> offs(P);
        for (int i = 0; i < P; ++i) {
            const int64_t kdil = (kl - 1) * d + 1;
            const int64_t p0 = (kdil - 1) / 2 + ((kdil - 1) % 2), p1 = (kdil - 1) / 2;   // Same
            const int64_t pds = N + p0 + p1;
            sh[i] = PairShape{N, kl, p0, pds, (pds - kdil) / s + 1, kdil};
            offs[i] = {i * N, i * 1024, i * N};
        }
        return std::make_pair(sh, offs);
    };
    CostModel out = *cm;
    for (int c = 0; c < kBatchClasses; ++c) {
        const int64_t kl = std::max<int64_t>(1, std::min<int64_t>(33, (2 * batch_lc(c) - 1) / (2 * d) + 1));
        auto so = shapes(kl);
        std::vector<int> method(P, c);
        Schedule S = build_schedule(so.first, so.second, method, s, out);
        double ms = 0;
        if (!time_schedule(ctx, S, x, k, y, s, d, &ms) || S.items.empty()) return false;
        out.fft_item[c] = ms / double(S.items.size());
    }
    double t[2] = {0, 0}, work[2] = {0, 0}, nitems[2] = {0, 0};
    const int64_t taps[2] = {16, 256};
    for (int j = 0; j < 2; ++j) {
        auto so = shapes(taps[j]);
        std::vector<int> method(P, -1);
        Schedule S = build_schedule(so.first, so.second, method, s, out);
        if (!time_schedule(ctx, S, x, k, y, s, d, &t[j])) return false;
        nitems[j] = double(S.items.size());
        for (auto& ps : so.first) work[j] += double(ps.out) * double(ps.k);
    }
    // t = nitems a + work b, from the two measurements
    const double det = nitems[0] * work[1] - nitems[1] * work[0];
    if (det != 0) {
        out.dir_b = std::max(1e-14, (nitems[0] * t[1] - nitems[1] * t[0]) / det);
        out.dir_a = std::max(0.0, (t[0] - work[0] * out.dir_b) / nitems[0]);
    }
    // matrix-core items (stride 1 only): t = nitems (a + b nkb) at two kernel
    // lengths, for each split
    if (s == 1) {
        const int64_t kcap = (kBatchMfmaMaxKdil - 1) / d + 1;
        const int64_t mt[2] = {std::min<int64_t>(8, kcap), std::min<int64_t>(512, kcap)};
        for (int six = 0; six < 2; ++six) {
            double tm[2] = {0, 0}, nm[2] = {0, 0}, kbm[2] = {0, 0};
            for (int j = 0; j < 2; ++j) {
                auto so = shapes(mt[j]);
                std::vector<int> method(P, six ? kMethodMfma6 : kMethodMfma);
                Schedule S = build_schedule(so.first, so.second, method, s, out);
                if (!time_schedule(ctx, S, x, k, y, s, d, &tm[j]) || S.items.empty()) return false;
                nm[j] = double(S.items.size());
                kbm[j] = double(mfma_nkb(so.first[0]));
            }
            const double pi0 = tm[0] / nm[0], pi1 = tm[1] / nm[1];   // ms per item
            if (kbm[1] > kbm[0]) {
                const double b = std::max(1e-12, (pi1 - pi0) / (kbm[1] - kbm[0]));
                const double a = std::max(0.0, pi0 - b * kbm[0]);
                (six ? out.mfma6_b : out.mfma_b) = b;
                (six ? out.mfma6_a : out.mfma_a) = a;
            }
        }
    }
    out.measured = true;
    *cm = out;
    return true;
}

// The cost model of this handle for strides/dilation (s, d), measured once.
CostModel cost_model(LaunchContext& ctx, int s, int d, bool measure) {
    const std::string key = "batch_costs3:" + std::to_string(s) + ":" + std::to_string(d);
    std::string& m = ctx.memo(key);
    CostModel cm = default_costs();
    if (m.size() == sizeof(CostModel)) {
        std::memcpy(&cm, m.data(), sizeof(cm));
        return cm;
    }
    if (!measure) return cm;
    if (!calibrate(ctx, s, d, &cm)) cm = default_costs();
    m.assign(reinterpret_cast<const char*>(&cm), sizeof(cm));
    return cm;
}

// The rank-1 single-problem descriptor of pair i.
ndc_conv_desc pair_desc(const ndc_batch_desc* d, int64_t i) {
    ndc_conv_desc c;
    std::memset(&c, 0, sizeof(c));
    c.dtype = d->dtype;
    c.rank = 1;
    c.x_shape[0] = d->x_len[i];
    c.k_shape[0] = d->k_len[i];
    c.dilation[0] = d->dilation;
    c.reverse = d->reverse;
    c.conv_mode = d->conv_mode;
    c.padding_mode = d->padding_mode;
    return c;
}

struct Batch {
    ndc_handle h = nullptr;
    LaunchContext* ctx = nullptr;
    int algo = 0;
    bool host = false;
    int s = 1, d = 1;
    std::vector<PairShape> shapes;
    std::vector<std::array<int64_t, 3>> offs;
    std::vector<int> method;
    int64_t x_total = 0, k_total = 0, y_total = 0;
    CostModel cm;
    Schedule S;
    DevSchedule D;
    std::vector<Plan> plain;
    DevMem io[3];
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* gx = nullptr;
    const void* gk = nullptr;
    void* gy = nullptr;

    ~Batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
    }

    ndc_status enqueue(const float* x, const float* k, float* y, hipStream_t st, bool profile) {
        if (algo == NDC_BATCH_PLAIN) {
            for (size_t i = 0; i < plain.size(); ++i) {
                if (shapes[i].out <= 0) continue;
                const ndc_status r = launch_fft(plain[i], x + offs[i][0], k + offs[i][1], y + offs[i][2], *ctx);
                if (r != NDC_OK) return r;
            }
            return NDC_OK;
        }
        BatchArgs a = D.args;
        a.x = x;
        a.k = k;
        a.y = y;
        D.args = a;
        return launch_batch(a, D.grid, *ctx, st, profile, S.kspec_bytes, S.main_bytes, S.main_flops);
    }
};

ndc_status fail(ndc_handle h, ndc_status s, const std::string& msg) {
    handle_set_last(h, msg);
    return s;
}

ndc_status batch_create(ndc_handle h, const ndc_batch_desc* d, Batch** out) {
    if (!h || !d || !out) return NDC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (d->dtype != NDC_F32) return fail(h, NDC_ERR_UNSUPPORTED, "batches are f32");
    if (d->count < 0 || (d->count > 0 && (!d->x_len || !d->k_len)))
        return fail(h, NDC_ERR_INVALID_ARGUMENT, "bad pair count or length arrays");
    if (d->algo < NDC_BATCH_AUTO || d->algo > NDC_BATCH_BF16X6)
        return fail(h, NDC_ERR_INVALID_ARGUMENT, "unknown batch algorithm");
    if (d->flags & ~uint32_t(NDC_FLAG_HOST_MEMORY)) return fail(h, NDC_ERR_INVALID_ARGUMENT, "unknown batch flags");
    LaunchContext* ctx = handle_context(h);
    if (hipSetDevice(handle_device(h)) != hipSuccess) return fail(h, NDC_ERR_DEVICE, "hipSetDevice failed");
    auto b = std::make_unique<Batch>();
    b->h = h;
    b->ctx = ctx;
    b->algo = d->algo;
    b->host = (d->flags & NDC_FLAG_HOST_MEMORY) != 0;
    const int64_t P = d->count;
    b->shapes.resize(size_t(P));
    b->offs.resize(size_t(P));
    ndc_border front{}, back{};
    for (int64_t i = 0; i < P; ++i) {
        ndc_conv_desc cd = pair_desc(d, i);
        Plan p;
        std::string err;
        const bool maps = d->algo == NDC_BATCH_PLAIN;
        const ndc_status st = make_plan(&cd, NDC_OP_CONV_FFT, &p, &err, maps);
        if (st != NDC_OK) return fail(h, st, "pair " + std::to_string(i) + ": " + err);
        if (p.x_shape[0] >= (int64_t(1) << 30) || p.pds[0] >= (int64_t(1) << 30))
            return fail(h, NDC_ERR_UNSUPPORTED, "pair " + std::to_string(i) + ": longer than 2^30 samples");
        b->shapes[size_t(i)] = PairShape{p.x_shape[0], p.k_shape[0], p.pad[0][0], p.pds[0], p.out[0], p.kdil[0]};
        b->offs[size_t(i)] = {b->x_total, b->k_total, b->y_total};
        b->x_total += p.x_shape[0];
        b->k_total += p.k_shape[0];
        b->y_total += p.out[0];
        if (i == 0) {
            b->s = int(p.stride[0]);
            b->d = int(p.dil[0]);
            front = p.side[0][0];
            back = p.side[0][1];
        }
        if (maps) b->plain.push_back(std::move(p));
    }
    if (b->algo != NDC_BATCH_PLAIN) {
        if (int64_t(b->s) >= (int64_t(1) << 20) || int64_t(b->d) >= (int64_t(1) << 20))
            return fail(h, NDC_ERR_UNSUPPORTED, "stride or dilation too large for the batch kernels");
        b->cm = cost_model(*ctx, b->s, b->d, b->algo == NDC_BATCH_AUTO || b->algo == NDC_BATCH_AUTO_BF16X3);
        b->method.resize(size_t(P));
        auto repeats = [](const ndc_border& side) {
            return side.tag == NDC_BORDER_REPLICATE || (side.tag == NDC_BORDER_CONST && side.value.re != 0.0);
        };
        const bool rb = repeats(front) || repeats(back);
        for (int64_t i = 0; i < P; ++i)
            b->method[size_t(i)] = choose(b->cm, b->shapes[size_t(i)], b->s, b->algo, rb);
        b->S = build_schedule(b->shapes, b->offs, b->method, b->s, b->cm);
        const float2* tw = batch_twiddles(*ctx);
        if (!tw || !b->D.upload(b->S, tw, resident_blocks(*ctx)))
            return fail(h, NDC_ERR_OUT_OF_MEMORY, "batch metadata upload failed");
        BatchArgs& a = b->D.args;
        a.s = b->s;
        a.d = b->d;
        a.dmul = batch_dmul(b->d);
        a.rev = d->reverse ? 1 : 0;
        if (!b->D.set_borders(side1d(front.tag), side1d(back.tag), float(front.value.re), float(back.value.re)))
            return fail(h, NDC_ERR_OUT_OF_MEMORY, "batch metadata upload failed");
    }
    *out = b.release();
    return NDC_OK;
}

ndc_status batch_run(Batch& b, const void* x, const void* k, void* y, bool use_graph) {
    if (!x || !k || !y) {
        if (b.y_total > 0 || b.x_total > 0) return fail(b.h, NDC_ERR_INVALID_ARGUMENT, "null data pointer");
    }
    if (hipSetDevice(handle_device(b.h)) != hipSuccess) return fail(b.h, NDC_ERR_DEVICE, "hipSetDevice failed");
    LaunchContext& ctx = *b.ctx;
    ctx.err.clear();
    hipStream_t st = ctx.stream();
    const float* dx = static_cast<const float*>(x);
    const float* dk = static_cast<const float*>(k);
    float* dy = static_cast<float*>(y);
    if (b.host) {
        if (!b.io[0].ensure(size_t(b.x_total) * 4) || !b.io[1].ensure(size_t(b.k_total) * 4) ||
            !b.io[2].ensure(size_t(b.y_total) * 4))
            return fail(b.h, NDC_ERR_OUT_OF_MEMORY, "device buffer allocation failed");
        if (hipMemcpyAsync(b.io[0].p, x, size_t(b.x_total) * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(b.io[1].p, k, size_t(b.k_total) * 4, hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(b.h, NDC_ERR_DEVICE, "host to device copy failed");
        dx = static_cast<const float*>(b.io[0].p);
        dk = static_cast<const float*>(b.io[1].p);
        dy = static_cast<float*>(b.io[2].p);
    }
    ndc_status r = NDC_OK;
    if (use_graph && b.algo != NDC_BATCH_PLAIN) {
        if (!b.exec || b.gx != dx || b.gk != dk || b.gy != dy) {
            if (b.exec) (void)hipGraphExecDestroy(b.exec);
            if (b.graph) (void)hipGraphDestroy(b.graph);
            b.exec = nullptr;
            b.graph = nullptr;
            if (!b.cap && hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking) != hipSuccess)
                return fail(b.h, NDC_ERR_DEVICE, "capture stream creation failed");
            if (hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return fail(b.h, NDC_ERR_DEVICE, "graph capture failed to start");
            r = b.enqueue(dx, dk, dy, b.cap, false);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b.cap, &g);
            if (r != NDC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return fail(b.h, r, ctx.err);
            }
            if (e != hipSuccess || !g || hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0) != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                b.exec = nullptr;
                return fail(b.h, NDC_ERR_DEVICE, "graph capture failed");
            }
            b.graph = g;
            b.gx = dx;
            b.gk = dk;
            b.gy = dy;
        }
        if (hipGraphLaunch(b.exec, st) != hipSuccess) return fail(b.h, NDC_ERR_DEVICE, "graph launch failed");
    } else {
        r = b.enqueue(dx, dk, dy, st, true);
        if (r != NDC_OK) return fail(b.h, r, ctx.err);
    }
    if (b.host) {
        if (hipMemcpyAsync(y, dy, size_t(b.y_total) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(b.h, NDC_ERR_DEVICE, "device to host copy failed");
    }
    return NDC_OK;
}

void append(std::string& s, const char* fmt, double v) {
    char tmp[64];
    std::snprintf(tmp, sizeof(tmp), fmt, v);
    s += tmp;
}

std::string batch_report(const Batch& b) {
    std::string r = "{\"algo\": " + std::to_string(b.algo) + ", \"pairs\": " + std::to_string(b.shapes.size());
    r += ", \"outputs\": " + std::to_string(b.y_total);
    r += ", \"items\": " + std::to_string(b.S.items.size()) + ", \"lines\": " + std::to_string(b.S.lines.size());
    r += ", \"kernel_spectra\": " + std::to_string(b.S.kpairs.size());
    r += ", \"grid\": " + std::to_string(b.D.grid);
    r += ", \"blocks_per_cu\": " + std::to_string(batch_blocks_per_cu());
    append(r, ", \"est_ms\": %.6f", b.S.est_ms);
    append(r, ", \"main_bytes\": %.1f", b.S.main_bytes);
    append(r, ", \"kspec_bytes\": %.1f", b.S.kspec_bytes);
    append(r, ", \"main_flops\": %.1f", b.S.main_flops);
    r += ", \"costs\": {\"measured\": ";
    r += b.cm.measured ? "true" : "false";
    r += ", \"fft_item_ms\": [";
    for (int c = 0; c < kBatchClasses; ++c) append(r, c ? ", %.4e" : "%.4e", b.cm.fft_item[c]);
    r += "]";
    append(r, ", \"direct_item_ms\": %.4e", b.cm.dir_a);
    append(r, ", \"direct_mac_ms\": %.4e", b.cm.dir_b);
    append(r, ", \"mfma_item_ms\": %.4e", b.cm.mfma_a);
    append(r, ", \"mfma_kblock_ms\": %.4e", b.cm.mfma_b);
    append(r, ", \"mfma6_item_ms\": %.4e", b.cm.mfma6_a);
    append(r, ", \"mfma6_kblock_ms\": %.4e}", b.cm.mfma6_b);
    // choice per bucket: kernel-length octave x signal-length octave
    std::map<std::pair<int, int>, std::map<std::string, int64_t>> bk;
    for (size_t i = 0; i < b.shapes.size() && i < b.method.size(); ++i) {
        const int kb = int(std::floor(std::log2(double(std::max<int64_t>(1, b.shapes[i].k)))));
        const int nb = int(std::floor(std::log2(double(std::max<int64_t>(1, b.shapes[i].n)))));
        const int m = b.method[i];
        const std::string name = m == kMethodDirect ? "direct_fp32"
                                 : m == kMethodMfma  ? "direct_bf16x3"
                                 : m == kMethodMfma6 ? "direct_bf16x6"
                                                     : "fft" + std::to_string(2 * batch_lc(m));
        bk[{kb, nb}][name] += 1;
    }
    r += ", \"buckets\": [";
    bool first = true;
    for (auto& e : bk) {
        r += first ? "" : ", ";
        first = false;
        r += "{\"k_log2\": " + std::to_string(e.first.first) + ", \"n_log2\": " + std::to_string(e.first.second) +
             ", \"choice\": {";
        bool f2 = true;
        for (auto& c : e.second) {
            r += (f2 ? "\"" : ", \"") + c.first + "\": " + std::to_string(c.second);
            f2 = false;
        }
        r += "}}";
    }
    r += "]}";
    return r;
}

}  // namespace
}  // namespace ndc

using ndc::Batch;

extern "C" {

ndc_status ndc_batch_create(ndc_handle h, const ndc_batch_desc* desc, ndc_batch* out) {
    if (!out) return NDC_ERR_INVALID_ARGUMENT;
    Batch* b = nullptr;
    const ndc_status s = ndc::batch_create(h, desc, &b);
    *out = reinterpret_cast<ndc_batch>(b);
    return s;
}

ndc_status ndc_batch_destroy(ndc_batch bb) {
    if (!bb) return NDC_ERR_INVALID_ARGUMENT;
    Batch* b = reinterpret_cast<Batch*>(bb);
    (void)hipStreamSynchronize(b->ctx->stream());
    delete b;
    return NDC_OK;
}

ndc_status ndc_batch_out_lens(ndc_batch bb, int64_t* y_len, int64_t* x_total, int64_t* k_total, int64_t* y_total) {
    if (!bb) return NDC_ERR_INVALID_ARGUMENT;
    const Batch* b = reinterpret_cast<const Batch*>(bb);
    if (y_len)
        for (size_t i = 0; i < b->shapes.size(); ++i) y_len[i] = b->shapes[i].out;
    if (x_total) *x_total = b->x_total;
    if (k_total) *k_total = b->k_total;
    if (y_total) *y_total = b->y_total;
    return NDC_OK;
}

ndc_status ndc_batch_run(ndc_batch bb, const void* x, const void* k, void* y) {
    if (!bb) return NDC_ERR_INVALID_ARGUMENT;
    return ndc::batch_run(*reinterpret_cast<Batch*>(bb), x, k, y, false);
}

ndc_status ndc_batch_run_graph(ndc_batch bb, const void* x, const void* k, void* y) {
    if (!bb) return NDC_ERR_INVALID_ARGUMENT;
    return ndc::batch_run(*reinterpret_cast<Batch*>(bb), x, k, y, true);
}

ndc_status ndc_batch_report(ndc_batch bb, char* buf, int64_t len) {
    if (!bb || !buf || len <= 0) return NDC_ERR_INVALID_ARGUMENT;
    const std::string r = ndc::batch_report(*reinterpret_cast<const Batch*>(bb));
    if (int64_t(r.size()) + 1 > len) return NDC_ERR_INVALID_ARGUMENT;
    std::memcpy(buf, r.c_str(), r.size() + 1);
    return NDC_OK;
}

ndc_status ndc_batch_border_map(const ndc_batch_desc* desc, int64_t pair, int32_t* codes, int64_t len) {
    if (!desc || !codes || pair < 0 || pair >= desc->count || !desc->x_len || !desc->k_len)
        return NDC_ERR_INVALID_ARGUMENT;
    ndc_conv_desc cd = ndc::pair_desc(desc, pair);
    cd.dtype = NDC_F32;
    ndc::Plan p;
    std::string err;
    const ndc_status s = ndc::make_plan(&cd, NDC_OP_CONV, &p, &err, false);
    if (s != NDC_OK) return s;
    if (len != p.pds[0]) return NDC_ERR_INVALID_ARGUMENT;
    const int32_t front = ndc::side1d(p.side[0][0].tag), back = ndc::side1d(p.side[0][1].tag);
    for (int64_t q = 0; q < len; ++q)
        codes[q] = int32_t(ndc::resolve1d<int64_t>(q, p.x_shape[0], p.pad[0][0], front, back));
    return NDC_OK;
}

}  // extern "C"
