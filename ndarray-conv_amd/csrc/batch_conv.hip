// Kernels of the batched rank-1 path: a kernel-spectrum pre-pass and one
// persistent launch that drains the work queue (direct tiles and
// overlap-save FFT blocks of all pairs).
//
// FFT items hold kBatchPoints complex points in LDS: G = kBatchPoints / LC
// lines of LC points, each line the half-length complex view
// z[n] = w[2n] + i w[2n+1] of a real window of L = 2 LC samples.  A line is
// transformed by radix-16 Stockham passes with the 16 points of a butterfly
// in registers (the last pass is radix 2/4/8/16 to reach LC), LDS only being
// used to exchange between passes.  The padded layout i + i/16 keeps the
// stride-16 accesses of the first pass free of bank conflicts.  The window is
// read from global memory with the border remap fused into the load, and only
// the cropped, strided outputs are written back.  Twiddles W_L^r come from a
// quarter-circle table generated in FP64 on the host and staged in LDS.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "batch.hpp"
#include "bf16_split.hpp"
#include "border1d.hpp"
#include "fft_bfly.hpp"

namespace ndc {
namespace {

using namespace fftb;
using C = float2;

constexpr int T = kBatchThreads;
constexpr int kTwSlots = 2048;                                   // L/4 for the largest class
constexpr int kDataSlots = kBatchPoints + kBatchPoints / 16;     // padded FFT data
constexpr int kWinSlots = kBatchWindow + kBatchWindow / 8;       // padded direct window
constexpr size_t kFftLdsBytes = size_t(kTwSlots + kDataSlots) * sizeof(C);

__device__ __forceinline__ int padh(int e) { return e + 8 * (e >> 5); }   // 16 bytes of padding per 64

// Word 3 of a raw (unswizzled, stride 0) buffer resource on gfx950: 32-bit
// data format.  A raw buffer's accesses at or past num_records bytes read
// zero and write nothing.
constexpr int kRawBufferWord3 = 0x00020000;

// s_waitcnt immediates (vmcnt in bits 3:0 and 15:14, expcnt 6:4 and lgkmcnt
// 11:8 left at their maxima, which wait for nothing).
constexpr int kWaitVmcnt0 = 0x0F70;
constexpr int kWaitVmcnt16 = 0x4F70;

// Raw buffer over [base, base + bytes); base and bytes are the same in every
// lane (the item's record), taken from the first so the resource is scalar.
// readfirstlane returns int: each half goes through uint32_t before widening,
// or a low half with bit 31 set would sign-extend over the high half.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_buffer(const float* base, int bytes) {
    const uint64_t b = reinterpret_cast<uint64_t>(base);
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(b >> 32))));
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(b))));
    const uint64_t u = (uint64_t(hi) << 32) | uint64_t(lo);
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(u), 0, __builtin_amdgcn_readfirstlane(bytes),
                                             kRawBufferWord3);
}

template <int P>
struct MfmaLds {
    static constexpr int nkb = (kBatchMfmaMaxKdil + 31 + 15) / 16;
    static constexpr int win = kBatchMfmaTile + 16 * nkb + 32;         // window samples
    static constexpr int win_slots = win + 8 * ((win - 1) / 32);       // padded plane length: padh(win - 1) + 1
    static constexpr int taps = 16 * nkb + 48;                         // tap copy length (even)
    // Copy 1 of a term starts copy1 entries after copy 0, 17 banks further
    // on: in one B read the even lanes take 16 consecutive dwords of copy 0
    // ending at some bank b and the odd lanes 16 of copy 1 ending at b + 16
    // (the shift less one), so the two sets share no bank.  A term's two
    // copies take term entries.
    static constexpr int copy1 = taps + 18;
    static constexpr int term = taps + copy1;
    static_assert((copy1 / 2) % 32 == 17, "tap copies 17 banks apart");
    static constexpr size_t bytes = size_t(P * win_slots + P * term) * 2;
};

// dynamic LDS of both kernels: the larger of the FFT layout and BF16x6 staging
constexpr size_t kLdsBytes = ((kFftLdsBytes > MfmaLds<3>::bytes ? kFftLdsBytes : MfmaLds<3>::bytes) + 255) & ~size_t(255);
static_assert(size_t(kWinSlots + kBatchTapChunk) * sizeof(float) <= size_t(kDataSlots) * sizeof(C),
              "direct-item staging must fit the FFT data region");

// threadIdx.x behind an empty volatile asm: inside the persistent loop it is
// re-read per item, so the compiler cannot hoist the address arithmetic of
// every item type out of the loop and keep it live in registers.
__device__ __forceinline__ int tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__device__ __forceinline__ int padi(int i) { return i + (i >> 4); }
__device__ __forceinline__ int padw(int i) { return i + (i >> 3); }

// st(e, f(e)) for e < n, with U values of f (global loads) in flight per
// thread before any is used: item stages are otherwise bound by load latency.
template <int U, typename F, typename S>
__device__ __forceinline__ void staged(int n, F&& f, S&& st) {
    for (int base = tid(); base < n; base += U * T) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * T;
            v[u] = e < n ? f(e) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * T;
            if (e < n) st(e, v[u]);
        }
    }
}

// W_L^e from the quarter table q[r] = W_L^r, r < L/4 (L a power of two).
__device__ __forceinline__ C twl(const C* q, int e, int L) {
    e &= L - 1;
    const int q4 = L >> 2;
    const int quad = e / q4;
    const C v = q[e & (q4 - 1)];
    switch (quad) {
        case 0: return v;
        case 1: return cmake<C>(v.y, -v.x);     // * -i
        case 2: return cmake<C>(-v.x, -v.y);    // * -1
        default: return cmake<C>(-v.y, v.x);    // * i
    }
}

// v[r] *= W_L^(r E) (conjugated for the inverse), r = 1 .. R-1; powers
// beyond 1, 2, 4, 8 are products of at most three looked-up factors.
template <int R>
__device__ __forceinline__ void twiddle(C* v, int E, const C* q, int L, bool inv) {
    C w[16];
    w[1] = twl(q, E, L);
    if constexpr (R > 2) {
        w[2] = twl(q, 2 * E, L);
        w[3] = cmul(w[2], w[1]);
    }
    if constexpr (R > 4) {
        w[4] = twl(q, 4 * E, L);
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R > 8) {
        w[8] = twl(q, 8 * E, L);
#pragma unroll
        for (int r = 9; r < 16; ++r) w[r] = cmul(w[8], w[r - 8]);
    }
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmulc(v[r], w[r], inv);
}

// One radix-R Stockham pass, in place, over the G lines of an item in LDS:
// butterfly j of a line reads points j + r LC/R, applies W_(Ns R)^(r k)
// (k = j mod Ns), and writes (j - k) R + k + r Ns.  Ends with a barrier.
template <int LC, int R>
__device__ __forceinline__ void pass(C* buf, const C* q, int Ns, bool inv) {
    constexpr int NB = 16 / R;
    constexpr int BPL = LC / R;
    constexpr int LS = LC + LC / 16;
    constexpr int L = 2 * LC;
    C v[16];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int beta = tid() + b * T;
        const int g = beta / BPL, j = beta % BPL;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = buf[g * LS + padi(j + r * BPL)];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int beta = tid() + b * T;
        const int g = beta / BPL, j = beta % BPL;
        const int k = j & (Ns - 1);
        if (Ns > 1) twiddle<R>(v + b * R, k * (L / (Ns * R)), q, L, inv);
        bfly<R>(v + b * R, inv);
        const int ob = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[g * LS + padi(ob + r * Ns)] = v[b * R + r];
    }
    __syncthreads();
}

template <int LC> struct LastRadix { static constexpr int value = LC / 256; };
template <> struct LastRadix<256> { static constexpr int value = 16; };

// Transform of the G lines in LDS (natural order in and out): radix 16,
// radix 16, then the radix that completes LC.
template <int LC>
__device__ __forceinline__ void fft_lds(C* buf, const C* q, bool inv) {
    pass<LC, 16>(buf, q, 1, inv);
    pass<LC, 16>(buf, q, 16, inv);
    if constexpr (LC > 256) pass<LC, LastRadix<LC>::value>(buf, q, 256, inv);
}

// Real sample u of line g in the padded layout (the complex view's point
// u/2, real part for even u).
template <int LC>
__device__ __forceinline__ int real_slot(int g, int u) {
    constexpr int LS = LC + LC / 16;
    return 2 * (g * LS + padi(u >> 1)) + (u & 1);
}

// X[m] of a real line from its half-length spectrum: zk = Z[m], zo = Z[LC-m].
__device__ __forceinline__ C r2c(C zk, C zo, C w) {
    const C zc = cconj(zo);
    const C e = cscale(cadd(zk, zc), 0.5f);
    const C o = cmul_negi(cscale(csub(zk, zc), 0.5f));
    return cadd(e, cmul(o, w));
}

// Half-length inverse input from A = P[m], B = conj P[LC-m].
__device__ __forceinline__ C c2r(C A, C B, C w) {
    return cadd(cadd(A, B), cmul_i(cmul(csub(A, B), cconj(w))));
}

// Per-line context of an item, staged in LDS.
struct LineCtx {
    int64_t x_off, y_off, ks_off, k_off;
    int64_t src;                   // x offset of window sample 0 when the window is all interior
    int32_t n, k, p0, pds, out, per, o0, valid;
    int32_t interior, pad_;
};

// Address of padded position P: an element of x, or the cell of a.consts
// holding the position's constant (zero past the padded end and for cells the
// fill never writes).  Address arithmetic only, so that a window's loads can
// all be issued before the first one returns: a load inside the border
// branches makes the compiler wait for it before the next position.
__device__ __forceinline__ const float* pad_ptr(const BatchArgs& a, const LineCtx& c, int64_t P) {
    if (P >= c.pds) return a.consts;
    const int64_t s = resolve1d<int64_t>(P, c.n, c.p0, a.front, a.back);
    if (s >= 0) return a.x + c.x_off + s;
    return a.consts + (s == kConstFront ? 1 : s == kConstBack ? 2 : 0);
}

__device__ __forceinline__ float pad_value(const BatchArgs& a, const LineCtx& c, int64_t P) { return *pad_ptr(a, c, P); }

// Tap t of the correlation form out[o] = sum_t pad[o s + t d] kc[t]:
// kc[t] = k[K-1-t] for a convolution (reverse), k[t] for no_reverse.
__device__ __forceinline__ float corr_tap(const BatchArgs& a, int64_t k_off, int K, int t) {
    return a.k[k_off + (a.rev ? K - 1 - t : t)];
}

// span: window length in samples (the interior test covers [o0 s, o0 s + span)).
__device__ __forceinline__ void load_line_ctx(const BatchArgs& a, LineCtx* lc, int G, int first, int count, int span) {
    const int g = tid();
    if (g < G) {
        LineCtx c{};
        c.valid = g < count;
        if (c.valid) {
            // a data line's record carries its pair's metadata: one load
            const BatchLine ln = a.lines[first + g];
            c.o0 = ln.o0;
            c.x_off = ln.x_off;
            c.y_off = ln.y_off;
            c.ks_off = ln.ks_off;
            c.n = ln.n;
            c.k = ln.k;
            c.p0 = ln.p0;
            c.pds = ln.pds;
            c.out = ln.out;
            c.per = ln.per;
            const int64_t P0 = int64_t(c.o0) * a.s;
            c.interior = P0 >= c.p0 && P0 + span <= int64_t(c.p0) + c.n;
            c.src = c.x_off + P0 - c.p0;
        }
        lc[g] = c;
    }
}

// Writes the real samples of the G lines: sample u of line g is *addr(c, u)
// (addr returns a readable address for every line, valid or not).  Each
// thread loads its L / T samples of every line, all in flight at once.
template <int LC, typename F>
__device__ __forceinline__ void load_reals(float* fb, const LineCtx* lc, F&& addr) {
    constexpr int G = kBatchPoints / LC;
    constexpr int L = 2 * LC;
    constexpr int J = L / T;
    static_assert(L % T == 0, "a line is whole rounds of the workgroup");
    const int t = tid();
    float v[G * J];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const LineCtx& c = lc[g];
#pragma unroll
        for (int j = 0; j < J; ++j) v[g * J + j] = *addr(c, t + j * T);
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int j = 0; j < J; ++j) fb[real_slot<LC>(g, t + j * T)] = v[g * J + j];
    __syncthreads();
}

template <int LC>
__device__ __forceinline__ void fft_item(const BatchArgs& a, const BatchItem& it, const C* q, C* buf, LineCtx* lc) {
    constexpr int G = kBatchPoints / LC;
    constexpr int LS = LC + LC / 16;
    constexpr int L = 2 * LC;
    constexpr int H = LC / 2 + 1;
    load_line_ctx(a, lc, G, it.a, it.b, L);
    __syncthreads();
    const int s = a.s;
    float* fb = reinterpret_cast<float*>(buf);
    // window of padded positions [o0 s, o0 s + L), border remap fused in
    load_reals<LC>(fb, lc, [&](const LineCtx& c, int u) -> const float* {
        if (!c.valid) return a.consts;
        return c.interior ? a.x + c.src + u : pad_ptr(a, c, int64_t(c.o0) * s + u);
    });
    // the kernel-spectrum bins this thread multiplies, loaded while the
    // forward transform runs
    constexpr int NI = (G * H + T - 1) / T;
    C kmv[NI], kpv[NI];
#pragma unroll
    for (int r = 0; r < NI; ++r) {
        const int idx = tid() + r * T;
        kmv[r] = kpv[r] = cmake<C>(0.f, 0.f);
        if (idx < G * H) {
            const int g = idx / H, m = idx - g * H;
            const LineCtx& c = lc[g];
            if (c.valid) {
                kmv[r] = a.kspec[c.ks_off + m];
                kpv[r] = a.kspec[c.ks_off + LC - m];
            }
        }
    }
    fft_lds<LC>(buf, q, false);
    // spectrum product, bins m and LC - m together
#pragma unroll
    for (int r = 0; r < NI; ++r) {
        const int idx = tid() + r * T;
        if (idx >= G * H) break;
        const int g = idx / H, m = idx - g * H;
        const LineCtx& c = lc[g];
        if (!c.valid) continue;
        C* Z = buf + g * LS;
        const int mp = LC - m;
        const C zm = Z[padi(m)];
        const C zp = Z[padi(mp == LC ? 0 : mp)];
        const C wm = twl(q, m, L), wp = twl(q, mp, L);
        C Pm = cmul(r2c(zm, zp, wm), kmv[r]);
        C Pp = cmul(r2c(zp, zm, wp), kpv[r]);
        if (m == 0) {   // imaginary parts of DC and Nyquist are ignored by c2r
            Pm.y = 0.f;
            Pp.y = 0.f;
        }
        Z[padi(m)] = c2r(Pm, cconj(Pp), wm);
        if (mp != m && mp < LC) Z[padi(mp)] = c2r(Pp, cconj(Pm), wp);
    }
    __syncthreads();
    fft_lds<LC>(buf, q, true);
    // crop and stride: output o0 + j is window sample kdil - 1 + j s (< L);
    // every thread reads its L / T candidates of each line first (clamped
    // into the line), then stores the ones that are outputs
    constexpr int J = L / T;
    const int t = tid();
    float v[G * J];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int kdil = lc[g].valid ? (lc[g].k - 1) * a.d + 1 : 1;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int u = kdil - 1 + (t + j * T) * s;
            v[g * J + j] = fb[real_slot<LC>(g, u < L ? u : L - 1)];
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const LineCtx& c = lc[g];
        const int cnt = !c.valid ? 0 : c.per < c.out - c.o0 ? c.per : c.out - c.o0;
        float* y = a.y + c.y_off + c.o0;
#pragma unroll
        for (int j = 0; j < J; ++j)
            if (t + j * T < cnt) y[t + j * T] = v[g * J + j];
    }
}

// The lines' context (valid, k, k_off, ks_off) is in lc, and the class's
// twiddle table in q.
template <int LC>
__device__ __forceinline__ void kspec_item(const BatchArgs& a, const C* q, C* buf, const LineCtx* lc) {
    constexpr int G = kBatchPoints / LC;
    constexpr int LS = LC + LC / 16;
    constexpr int L = 2 * LC;
    const int d = a.d;
    float* fb = reinterpret_cast<float*>(buf);
    // scattered kernel of the reference's FFT path (conv_fft/padding.rs:78-111):
    // k[i] at i d (reverse) or (K-1-i) d (no_reverse)
    // u / d as the high word of 2u dmul (exact: u < 2 LC <= 8192, and a pair
    // whose dilation passes 2^18 has one tap, whose only position is u = 0)
    load_reals<LC>(fb, lc, [&](const LineCtx& c, int u) -> const float* {
        const int i = int(__umulhi(uint32_t(u) << 1, a.dmul));
        if (!c.valid || u != i * d || i >= c.k) return a.consts;
        return a.k + c.k_off + (a.rev ? i : c.k - 1 - i);
    });
    fft_lds<LC>(buf, q, false);
    const float scale = 1.0f / float(L);   // the inverse normalisation, folded in
    for (int idx = tid(); idx < G * (LC + 1); idx += T) {
        const int g = idx / (LC + 1), m = idx - g * (LC + 1);
        const LineCtx& c = lc[g];
        if (!c.valid) continue;
        const C* Z = buf + g * LS;
        const C zm = Z[padi(m == LC ? 0 : m)];
        const C zp = Z[padi(m == 0 ? 0 : LC - m)];
        a.kspec[c.ks_off + m] = cscale(r2c(zm, zp, twl(q, m, L)), scale);
    }
}

// Direct tile: outputs [o0, o0 + cnt) of one pair, taps staged in chunks.
__device__ __forceinline__ void direct_item(const BatchArgs& a, const BatchItem& p, float* win, float* taps) {
    const int o0 = p.b, cnt = p.c;
    const int s = a.s, d = a.d;
    const int K = p.k;
    const bool fast = s == 1 && d == 1;
    const int span = (cnt - 1) * s + 1;
    if (span > kBatchWindow / 2) return;   // the planner sizes tiles below this
    int chunk = (kBatchWindow - span) / d + 1;
    if (chunk > kBatchTapChunk) chunk = kBatchTapChunk;
    float acc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = 0.f;
    LineCtx c{};
    c.x_off = p.x_off;
    c.n = p.n;
    c.p0 = p.p0;
    c.pds = p.pds;
    for (int t0 = 0; t0 < K; t0 += chunk) {
        const int nt = K - t0 < chunk ? K - t0 : chunk;
        __builtin_amdgcn_s_setprio(2);   // staging, which the barrier below waits for, first
        const int64_t wbase = int64_t(o0) * s + int64_t(t0) * d;
        const int wlen = span + (nt - 1) * d;
        auto put = [&](int e, float v) { win[padw(e)] = v; };
        if (wbase >= c.p0 && wbase + wlen <= int64_t(c.p0) + c.n) {
            const float* src = a.x + c.x_off + (wbase - c.p0);   // all interior: no border logic
            staged<24>(wlen, [&](int e) { return src[e]; }, put);
        } else {
            staged<24>(wlen, [&](int e) { return pad_value(a, c, wbase + e); }, put);
        }
        staged<4>(nt, [&](int t) { return corr_tap(a, p.k_off, K, t0 + t); }, [&](int t, float v) { taps[t] = v; });
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        if (fast) {
            // 8 consecutive outputs per thread, a sliding register window
            const int ob = 8 * tid();
            float r[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) r[m] = win[padw(ob + m)];
            for (int t = 0; t < nt; ++t) {
                const float kv = taps[t];
#pragma unroll
                for (int m = 0; m < 8; ++m) acc[m] += r[m] * kv;
#pragma unroll
                for (int m = 0; m < 7; ++m) r[m] = r[m + 1];
                r[7] = win[padw(ob + t + 8)];
            }
        } else {
            for (int t = 0; t < nt; ++t) {
                const float kv = taps[t];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int o = tid() + m * T;
                    if (o < cnt) acc[m] += win[padw(o * s + t * d)] * kv;
                }
            }
        }
        __syncthreads();
    }
    float* y = a.y + p.y_off + o0;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int o = fast ? 8 * tid() + m : tid() + m * T;
        if (o < cnt) y[o] = acc[m];
    }
}

// Matrix-core tile (stride 1): outputs [o0, o0 + cnt), cnt <= 4096, 1024 per
// wave as one 32x32 tile D[n][m] = out[32 n + m] of the product
//   D = A B,  A[n][j] = pad[32 n + j],  B[j][m] = kc[j - m]   (Toeplitz),
// K = j running over kdil + 31 values in blocks of 16.  The window and the
// taps are split into P bf16 terms once, when staged, each term in a plane of
// its own, so an A operand is one aligned 16-byte read per plane; the taps are
// kept in two copies shifted by one element so every B read is 4-byte aligned.
// P = 2 (BF16x3) issues t0*u0, t0*u1, t1*u0 per block; P = 3 (BF16x6) also
// t1*u1, t2*u0, t0*u2.
// Tap q of the dilated correlation kernel kc_dil (zero outside [0, kdil) and
// between dilated taps) through a raw buffer kr over the pair's k taps: an
// index that is no tap is clamped to k, the first offset past the buffer.
// q / d is mulhi(2 q, dmul), dmul = ceil(2^31 / d): exact for q < 2^20 when
// d <= 2^11, and matrix-core items have (k - 1) d < 1024 (q = 0 is exact for
// any d).
__device__ __forceinline__ float dil_tap(const BatchArgs& a, __amdgpu_buffer_rsrc_t kr, int k, int kdil, int q) {
    const uint32_t uq = uint32_t(q);
    const uint32_t t = __umulhi(uq << 1, a.dmul);
    const bool tap = uq < uint32_t(kdil) && t * uint32_t(a.d) == uq;
    const uint32_t i = tap ? (a.rev ? uint32_t(k - 1) - t : t) : uint32_t(k);
    return __builtin_bit_cast(float, uint32_t(__builtin_amdgcn_raw_buffer_load_b32(kr, int(i * 4), 0, 0)));
}

// Taps q0, q0 + 1, q0 + 2 of the dilated correlation kernel.
__device__ __forceinline__ void taps3(const BatchArgs& a, const BatchItem& p, int kdil, int q0, float& f0, float& f1,
                                      float& f2) {
    const __amdgpu_buffer_rsrc_t kr = raw_buffer(a.k + p.k_off, p.k * int(sizeof(float)));
    f0 = dil_tap(a, kr, p.k, kdil, q0);
    f1 = dil_tap(a, kr, p.k, kdil, q0 + 1);
    f2 = dil_tap(a, kr, p.k, kdil, q0 + 2);
}

// A matrix-core item's window samples and first round of taps, in registers:
// thread t holds window quads e = 4 stage_quad(t) + u 4 T (four consecutive
// samples each) and the taps of tap-plane pair e = 2 t.  Loaded either at
// the item's start or during the item before it.
struct MfmaStage {
    static constexpr int U = (MfmaLds<3>::win + 4 * T - 1) / (4 * T);
    float w[U][4];
    float f0, f1, f2;
};
static_assert(MfmaLds<2>::win == MfmaLds<3>::win, "both splits read the same window");

// Window quad of thread t in each round of 4 T samples.  A quad's two pairs
// are two consecutive words of a BF16 plane (padh keeps 4-aligned samples'
// pairs together), stored as one 8-byte write.  padh adds 4 words per 16, so
// a run of 8 quads fills 16 words and runs 4 apart are 80 words (16 banks)
// apart: each 16-lane group takes two such runs and writes every bank once.
__device__ __forceinline__ int stage_quad(int t) {
    const int g = t >> 4, l = t & 15;
    return 8 * ((g & 3) + 8 * (g >> 2) + 4 * (l >> 3)) + (l & 7);
}
static_assert(T == 256, "stage_quad maps 256 threads onto 32 runs of 8 quads");

// Issues the loads of item p's stage (no waits).  The window covers padded
// positions o0 + e for every sample an A operand reads.
__device__ __forceinline__ void mfma_stage_load(const BatchArgs& a, const BatchItem& p, MfmaStage& st) {
    const int kdil = (p.k - 1) * a.d + 1;
    const int nkb = (kdil + 31 + 15) / 16;
    const int W = kBatchMfmaTile + 16 * nkb + 32;
    const int64_t o0 = p.b;
    LineCtx c{};
    c.x_off = p.x_off;
    c.n = p.n;
    c.p0 = p.p0;
    c.pds = p.pds;
    const int t2 = 2 * tid();
    const int e4 = 4 * stage_quad(tid());   // the thread's first window quad
    if (o0 >= c.p0 && o0 + W <= int64_t(c.p0) + c.n) {
        // All interior: a raw buffer over the window's W samples, each quad
        // one 16-byte load at the lane's offset plus a scalar one (dword
        // alignment is all a multi-dword load needs).  Lanes past the window
        // read zero, which the stage store never uses.
        const __amdgpu_buffer_rsrc_t wr = raw_buffer(a.x + c.x_off + (o0 - c.p0), W * int(sizeof(float)));
#pragma unroll
        for (int u = 0; u < MfmaStage::U; ++u) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(wr, e4 * int(sizeof(float)),
                                                                 u * 4 * T * int(sizeof(float)), 0);
            // copied out before the bit cast: clang's bit cast of a vector
            // element reads element 0 whatever the index
            const uint32_t v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
            st.w[u][0] = __builtin_bit_cast(float, v0);
            st.w[u][1] = __builtin_bit_cast(float, v1);
            st.w[u][2] = __builtin_bit_cast(float, v2);
            st.w[u][3] = __builtin_bit_cast(float, v3);
        }
    } else {
#pragma unroll
        for (int u = 0; u < MfmaStage::U; ++u) {
            const int e = e4 + 4 * u * T;
#pragma unroll
            for (int i = 0; i < 4; ++i) st.w[u][i] = *(e < W ? pad_ptr(a, c, o0 + e + i) : a.consts);
        }
    }
    taps3(a, p, kdil, t2 - 32, st.f0, st.f1, st.f2);
}

// Splits item p's stage into the P bf16 planes in LDS: window plane t at
// wp + t * win_slots (pairs stored as one 32-bit word; padh pads 16 bytes per
// 64), tap term t, copy c at tp + t * term + c * copy1, copy 0 holding tap q at
// 32 + q and copy 1 at 31 + q, zero outside [0, kdil).  Tap rounds past the
// first are loaded here.  No barrier.
template <int P>
__device__ __forceinline__ void mfma_stage_store(const BatchArgs& a, const BatchItem& p, const MfmaStage& st,
                                                 uint16_t* lds) {
    using L = MfmaLds<P>;
    uint16_t* wp = lds;
    uint16_t* tp = lds + P * L::win_slots;
    const int kdil = (p.k - 1) * a.d + 1;
    const int nkb = (kdil + 31 + 15) / 16;
    const int W = kBatchMfmaTile + 16 * nkb + 32;
    const int ntp = 16 * nkb + 48;
    const int t2 = 2 * tid();
    const int e4 = 4 * stage_quad(tid());
    // Quad u sits at e = e4 + 4 u T, padded position padh(e4) + u (4 T) 5/4
    // (8-byte aligned); quads u < 4 are inside every window (W > 4 (4 T)), and
    // W is a multiple of 4, so a quad is inside or outside whole.
    static_assert(4 * T % 32 == 0 && 4 * 4 * T <= kBatchMfmaTile, "quad positions of the stage store");
    const int h0 = padh(e4);
#pragma unroll
    for (int u = 0; u < MfmaStage::U; ++u) {
        if (u < 4 || e4 + 4 * u * T < W) {
            uint32_t lo[P], hi[P];
            split_pair<P>(st.w[u][0], st.w[u][1], lo);
            split_pair<P>(st.w[u][2], st.w[u][3], hi);
#pragma unroll
            for (int q = 0; q < P; ++q)
                *reinterpret_cast<uint2*>(wp + q * L::win_slots + h0 + u * (4 * T) * 5 / 4) = make_uint2(lo[q], hi[q]);
        }
    }
    // pairs of both tap copies from three consecutive taps
    auto put_taps = [&](int e, float t0, float t1, float t2_) {
        uint32_t c0[P], c1[P];
        split_pair<P>(t0, t1, c0);
        split_pair<P>(t1, t2_, c1);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            reinterpret_cast<uint32_t*>(tp + q * L::term + e)[0] = c0[q];
            reinterpret_cast<uint32_t*>(tp + q * L::term + L::copy1 + e)[0] = c1[q];
        }
    };
    if (t2 < ntp) put_taps(t2, st.f0, st.f1, st.f2);
    for (int e = t2 + 2 * T; e < ntp; e += 2 * T) {
        float f0, f1, f2;
        taps3(a, p, kdil, e - 32, f0, f1, f2);
        put_taps(e, f0, f1, f2);
    }
}

// Matrix-core tile: outputs [o0, o0 + cnt) of one pair with stride 1.  With
// `have`, st holds this item's stage (loaded during the item before).  When
// `next` is a matrix-core item, its stage is loaded into st during this
// item's products and `have` set.  `advance` issues the work queue's own
// reads for the items ahead after those loads: the wait for the queue's
// reads at the item's end then covers loads issued before the products, and
// never the item's own output stores.
template <int P, typename Adv>
__device__ __forceinline__ void mfma_item(const BatchArgs& a, const BatchItem& p, uint16_t* lds, MfmaStage& st,
                                          bool& have, const BatchItem* next, Adv&& advance) {
    using L = MfmaLds<P>;
    const int o0 = p.b, cnt = p.c;
    const int kdil = (p.k - 1) * a.d + 1;
    const int nkb = (kdil + 31 + 15) / 16;
    uint16_t* wp = lds;
    uint16_t* tp = lds + P * L::win_slots;
    if (kdil > kBatchMfmaMaxKdil) {   // the planner never queues these
        advance();
        have = false;
        __syncthreads();
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);   // the queue's reads, on every lane
        return;
    }
    // The stage and the prefetch run at a raised wave priority: the item's
    // barrier waits for them, while the products of the other workgroups on
    // the CU keep the matrix cores busy meanwhile.
    __builtin_amdgcn_s_setprio(2);
    // two copies of the split: after fresh loads it waits for all of them,
    // after a prefetch only for loads older than the last item's stores
    if (have) {
        mfma_stage_store<P>(a, p, st, lds);
    } else {
        mfma_stage_load(a, p, st);
        mfma_stage_store<P>(a, p, st, lds);
    }
    have = next != nullptr;
    if (have) mfma_stage_load(a, *next, st);
    advance();
    __syncthreads();
    __builtin_amdgcn_s_setprio(0);
    const int t = tid();
    const int lane = t & 63, wave = t >> 6;
    const int half = lane >> 5, m = lane & 31;
    const int wb = wave * 1024;
    // operands of K-block kb: A lane holds row m (a 32-output segment), B lane
    // column m (an output offset); both K entries 16 kb + 8 half + i.  The A
    // operand is window entry padh(wb + 32 m + 8 half + 16 kb), which is
    // pa + 40 (kb / 2) + 16 (kb % 2) (8 half + 16 (kb % 2) < 32: a pair of
    // blocks spans one padding step); the B operand is taps
    // 16 kb + 8 half - m + i, i < 8, from the copy where they start even, at
    // pb + 16 kb.  Both advance by constants, with no index arithmetic.
    const int par = (32 + 8 * half - m) & 1;   // parity of this lane's first tap index
    const uint16_t* pa = wp + padh(wb + 32 * m + 8 * half);
    const uint16_t* pb = tp + par * L::copy1 + (32 + 8 * half - m - par);
    auto load = [&](int j, int odd, u32x4* av, u32x4* bv) {   // block kb = 2 j + odd
        const uint16_t* qa = pa + 40 * j + 16 * odd;
        const uint16_t* qb = pb + 32 * j + 16 * odd;
#pragma unroll
        for (int s = 0; s < P; ++s) {
            av[s] = *reinterpret_cast<const u32x4*>(qa + s * L::win_slots);
            const uint32_t* b = reinterpret_cast<const uint32_t*>(qb + s * L::term);
#pragma unroll
            for (int i = 0; i < 4; ++i) bv[s][i] = b[i];
        }
    };
    auto mfma = [](const u32x4& x, const u32x4& y, floatx16 acc) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, x), __builtin_bit_cast(bf16x8, y),
                                                       acc, 0, 0, 0);
    };
    // Two accumulators, written strictly in turn: each matrix instruction
    // has one independent instruction between it and the one it depends on.
    // P = 2 takes three products per block, so the block's parity picks the
    // accumulator it starts with; P = 3 takes six.  Every product lands in
    // one of them and the output is their sum (FP32 accumulation, as an FP32
    // dot product; the split's own error is 2^-16 per product).
    floatx16 acc0, acc1;
    // block 0's products start from the instruction's inline zero (no
    // register clears)
    const floatx16 zero = {};
    auto first = [&](const u32x4* av, const u32x4* bv) {
        acc0 = mfma(av[0], bv[0], zero);
        acc1 = mfma(av[0], bv[1], zero);
        acc0 = mfma(av[1], bv[0], acc0);
        if constexpr (P == 3) {
            acc1 = mfma(av[1], bv[1], acc1);
            acc0 = mfma(av[2], bv[0], acc0);
            acc1 = mfma(av[0], bv[2], acc1);
        }
    };
    auto products = [&](const u32x4* av, const u32x4* bv, auto odd) {
        floatx16& x = (P == 2 && odd.value) ? acc1 : acc0;
        floatx16& y = (P == 2 && odd.value) ? acc0 : acc1;
        x = mfma(av[0], bv[0], x);
        y = mfma(av[0], bv[1], y);
        x = mfma(av[1], bv[0], x);
        if constexpr (P == 3) {
            y = mfma(av[1], bv[1], y);
            x = mfma(av[2], bv[0], x);
            y = mfma(av[0], bv[2], y);
        }
    };
    using Even = std::integral_constant<bool, false>;
    using Odd = std::integral_constant<bool, true>;
    // two operand sets in turn: the next block's reads are in flight while
    // the current block's products issue, with no register copies between
    u32x4 a0[P], b0[P], a1[P], b1[P];
    // block 0, then odd blocks in the second set and even ones in the first
    load(0, 0, a0, b0);
    first(a0, b0);
    load(0, 1, a1, b1);   // nkb >= 2: kdil + 31 >= 32
    int kb = 1;
    for (; kb + 1 < nkb; kb += 2) {
        load((kb + 1) >> 1, 0, a0, b0);
        products(a1, b1, Odd{});
        if (kb + 2 < nkb) load((kb + 2) >> 1, 1, a1, b1);
        products(a0, b0, Even{});
    }
    if (kb < nkb) products(a1, b1, Odd{});
    // D lane holds D[8 (i/4) + 4 half + i%4][m].  Every wave issues all 16
    // stores: the tile's outputs are a raw buffer of cnt floats, and the
    // hardware drops the stores of lanes past its end.  With no store
    // skipped, the next item's wait for its prefetched stage can let these
    // stores run on instead of waiting for all of them.
    const __amdgpu_buffer_rsrc_t yr = raw_buffer(a.y + p.y_off + o0, cnt * int(sizeof(float)));
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int o = wb + 32 * (8 * (i / 4) + 4 * half + (i % 4)) + m;
        const float v = acc0[i] + acc1[i];
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), yr, o * int(sizeof(float)), 0, 0);
    }
    // With at most 16 memory operations outstanding (the stores above),
    // every load issued before them has returned: the prefetched stage and
    // the queue's reads.  Said here, on every lane, the compiler knows those
    // registers are written when the loop comes round, and does not wait for
    // the stores to copy them.
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt16);
}

// BF16x3 staging fits the FFT data region (the twiddle table stays); BF16x6
// staging takes the whole dynamic allocation, twiddles included.
static_assert(MfmaLds<2>::bytes <= size_t(kDataSlots) * sizeof(C), "BF16x3 staging must fit the FFT data region");
static_assert(MfmaLds<3>::bytes <= kLdsBytes, "BF16x6 staging must fit the dynamic LDS");

// A class's quarter table (n = LC/2 <= kTwSlots points): every load of a
// thread is issued before its first LDS store.  A plain copy loop waits for
// each load before the next, eight serial round trips for the largest class.
__device__ __forceinline__ void stage_tw(C* dst, const C* src, int n) {
    constexpr int U = kTwSlots / T;
    static_assert(kTwSlots % T == 0 && batch_lc(kBatchClasses - 1) / 2 <= kTwSlots, "one round per thread slot");
    C v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = tid() + u * T;
        if (i < n) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = tid() + u * T;
        if (i < n) dst[i] = v[u];
    }
}

__global__ void __launch_bounds__(kBatchThreads) batch_kspec_kernel(BatchArgs a) {
    extern __shared__ C smem[];
    __shared__ LineCtx lc[16];
    if (blockIdx.x == 0 && threadIdx.x < kBatchHeads) a.counter[threadIdx.x * kBatchQueueStride] = 0;   // sub-queue heads
    if (int(blockIdx.x) >= a.nkitems) return;
    const BatchKItem it = a.kitems[blockIdx.x];
    C* q = smem;
    C* buf = smem + kTwSlots;
    const int LC = batch_lc(it.cls);
    const int G = kBatchPoints / LC;
    // the lines' records and the twiddle table in flight together
    const int g = tid();
    const bool valid = g < G && g < it.count;
    BatchKLine kl{};
    if (valid) kl = a.klines[it.first + g];
    stage_tw(q, a.tw + batch_tw_offset(it.cls), LC / 2);
    if (g < G) {
        LineCtx c{};
        c.valid = valid;
        c.k_off = kl.k_off;
        c.ks_off = kl.ks_off;
        c.k = kl.k;
        lc[g] = c;
    }
    __syncthreads();
    switch (it.cls) {
        case 0: kspec_item<256>(a, q, buf, lc); break;
        case 1: kspec_item<512>(a, q, buf, lc); break;
        case 2: kspec_item<1024>(a, q, buf, lc); break;
        case 3: kspec_item<2048>(a, q, buf, lc); break;
        default: kspec_item<4096>(a, q, buf, lc); break;
    }
}

// LDS of the persistent loop beside the dynamic allocation: the FFT lines'
// context and the queue ring.
__shared__ LineCtx g_lc[16];
__shared__ int g_item[3];
__shared__ BatchItem g_rec[3];

// Persistent: the workgroup takes items of one range of the queue until the
// range is empty (the only exit, reached by every workgroup): the direct and
// FFT items (M false) or the matrix-core items (M true).  The matrix-core
// range's loop is a function of its own, called once per workgroup, so the
// FFT code's registers do not crowd it: in one loop with the FFT code the
// matrix-core items' scalar values spilled to vector lanes (about a fifth of
// their vector instructions).  The first range's loop stays in the kernel,
// which has the larger scalar register budget (no call convention).
typedef const __attribute__((address_space(4))) BatchArgs* KernargArgs;

// The launch arguments, field by field from the kernel argument segment
// (where the kernel's one parameter lies, at ap): scalar loads, whose
// registers the loop never waits for.  A function argument arrives in vector
// registers, so the pointer is first made uniform; loads through a vector
// address, or a copy through a generic pointer, land in vector registers,
// and the compiler, unable to tell those loads finished, waits for every
// load of the wave at their first use inside the loop.  (The kernel argument
// segment's own intrinsic is valid in kernels only: in a called function it
// yields a null pointer.)
__device__ __forceinline__ BatchArgs kernel_args(KernargArgs arg) {
    const uint64_t u = reinterpret_cast<uint64_t>(arg);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(u))));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(u >> 32))));
    const KernargArgs ap = reinterpret_cast<KernargArgs>((uint64_t(hi) << 32) | uint64_t(lo));
    BatchArgs a;
    a.x = ap->x;
    a.k = ap->k;
    a.y = ap->y;
    a.pairs = ap->pairs;
    a.items = ap->items;
    a.lines = ap->lines;
    a.klines = ap->klines;
    a.kitems = ap->kitems;
    a.kspec = ap->kspec;
    a.counter = ap->counter;
    a.tw = ap->tw;
    a.consts = ap->consts;
    a.nitems = ap->nitems;
    a.nkitems = ap->nkitems;
    a.nfirst = ap->nfirst;
    a.s = ap->s;
    a.d = ap->d;
    a.rev = ap->rev;
    a.dmul = ap->dmul;
    a.front = ap->front;
    a.back = ap->back;
    a.cfront = ap->cfront;
    a.cback = ap->cback;
    return a;
}

template <bool M>
__device__ __forceinline__ void drain(KernargArgs ap) {
    const BatchArgs a = kernel_args(ap);
    extern __shared__ C smem[];
    LineCtx* lc = g_lc;
    int* s_item = g_item;
    BatchItem* s_rec = g_rec;
    C* q = smem;
    C* buf = smem + kTwSlots;
    float* win = reinterpret_cast<float*>(buf);
    float* taps = win + kWinSlots;
    int cur = -1;   // class whose twiddle table is in LDS
    const int lo = M ? a.nfirst : 0, hi = M ? a.nitems : a.nfirst;
    if (lo >= hi) return;   // uniform: an empty range costs no queue pops
    int32_t* heads = a.counter + (M ? kBatchQueues * kBatchQueueStride : 0);
    // The queue runs three items ahead of the work.  While item i (ring slot
    // r0) is processed, the records of items i and i+1 are in s_rec[r0] and
    // s_rec[r1], item i+2's record (16 dwords, one per lane of the first 16
    // lanes) is in flight, and a ticket is drawn for item i+3.  Writes for
    // later items go to slots no thread reads in the same iteration.
    //
    // Thread 0 draws tickets from sub-queue cq, starting with blockIdx.x mod
    // kBatchQueues; ticket t of sub-queue q is item lo + q + t kBatchQueues.
    // An empty sub-queue stays empty, so once cq is empty thread 0 moves on
    // through the others in turn, and an index >= hi after visiting all of
    // them means every sub-queue is empty.
    int cq = blockIdx.x % kBatchQueues, moved = 0;
    auto ticket = [&]() { return atomicAdd(heads + cq * kBatchQueueStride, 1); };
    auto resolve = [&](int t) {
        int g = lo + cq + t * kBatchQueues;
        while (g >= hi && moved < kBatchQueues - 1) {
            cq = cq + 1 == kBatchQueues ? 0 : cq + 1;
            ++moved;
            g = lo + cq + ticket() * kBatchQueues;
        }
        return g;
    };
    if (threadIdx.x == 0) {
        s_item[0] = resolve(ticket());
        s_item[1] = resolve(ticket());
        s_item[2] = resolve(ticket());
    }
    __syncthreads();
    {
        const int r = threadIdx.x >> 4, i = s_item[r & 1];
        if (threadIdx.x < 32 && i < hi)
            reinterpret_cast<int32_t*>(&s_rec[r])[threadIdx.x & 15] =
                reinterpret_cast<const int32_t*>(a.items + i)[threadIdx.x & 15];
    }
    __syncthreads();
    MfmaStage st;
    bool have = false;   // st holds the current item's stage
    // The queue's reads for items i+2 and i+3 (the ticket is resolved at the
    // item's end), carried across iterations and written only by their
    // loads: a register write over one that a load may still be filling
    // makes the compiler wait for every load of the wave (the prefetched
    // stage included).  Every lane loads pre (dword tid mod 16 of item i+2's
    // record, or of item 0 past the end of the range); thread 0 draws next3,
    // and the other lanes keep whatever the register held.
    int32_t pre = 0;
    int next3 = 0;
    for (int r0 = 0;;) {
        const int r1 = r0 == 2 ? 0 : r0 + 1, r2 = r1 == 2 ? 0 : r1 + 1;
        const int idx = s_item[r0];
        if (idx >= hi) break;   // uniform: every workgroup leaves once the range is empty
        auto advance = [&]() {
            const int i2 = s_item[r2];
            pre = reinterpret_cast<const int32_t*>(a.items + (i2 < hi ? i2 : 0))[threadIdx.x & 15];
            if (threadIdx.x == 0) next3 = ticket();
        };
        const BatchItem& it = s_rec[r0];
        const int kind = it.kind;
        if constexpr (M) {
            // every item of this range is a matrix-core item
            const BatchItem* next = s_item[r1] < hi ? &s_rec[r1] : nullptr;
            if (kind == kBatchKindMfma)
                mfma_item<2>(a, it, reinterpret_cast<uint16_t*>(buf), st, have, next, advance);
            else
                mfma_item<3>(a, it, reinterpret_cast<uint16_t*>(smem), st, have, next, advance);
        } else {
            advance();
            if (kind == 0) {
                direct_item(a, it, win, taps);
            } else {
                const int cls = kind - 1;
                if (cls != cur) {
                    stage_tw(q, a.tw + batch_tw_offset(cls), batch_lc(cls) / 2);
                    cur = cls;
                    __syncthreads();
                }
                switch (cls) {
                    case 0: fft_item<256>(a, it, q, buf, lc); break;
                    case 1: fft_item<512>(a, it, q, buf, lc); break;
                    case 2: fft_item<1024>(a, it, q, buf, lc); break;
                    case 3: fft_item<2048>(a, it, q, buf, lc); break;
                    default: fft_item<4096>(a, it, q, buf, lc); break;
                }
            }
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);   // the queue's reads, on every lane (as above)
        }
        // Every item kind passes a barrier after all threads have read the
        // ring at the top, and no thread reads s_rec[r2] or writes s_item[r0]
        // before that, so the writes below need no barrier before them.
        // The slot address is formed here from an opaque copy of the thread
        // index: hoisted out of the loop it would hold a register the whole
        // loop, one more than the matrix-core items leave free.
        int lane16 = threadIdx.x;
        asm volatile("" : "+v"(lane16));
        if (lane16 < 16) reinterpret_cast<int32_t*>(&s_rec[r2])[lane16] = pre;
        if (threadIdx.x == 0) s_item[r0] = resolve(next3);
        __syncthreads();
        r0 = r1;
    }
}

// The matrix-core range's loop, called once per workgroup.
__device__ __attribute__((noinline)) void drain_matrix_core(KernargArgs ap) { drain<true>(ap); }

// The first min(grid, nfirst) workgroups drain the first range, then every
// workgroup the second: a workgroup that finds a range empty pops each of
// its sub-queues once more on the way out, so the first range's few items
// are left to that many workgroups.
__global__ void __launch_bounds__(kBatchThreads, 3) batch_main_kernel(BatchArgs a) {
    const KernargArgs ap = (KernargArgs)__builtin_amdgcn_kernarg_segment_ptr();
    if (int(blockIdx.x) < a.nfirst) {
        drain<false>(ap);
        __syncthreads();   // the ring slots are reused by the second range
    }
    drain_matrix_core(ap);
}

}  // namespace

size_t batch_lds_bytes() { return kLdsBytes; }

int batch_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_main_kernel, kBatchThreads, kLdsBytes) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

ndc_status launch_batch(const BatchArgs& a, int grid, LaunchContext& ctx, hipStream_t stream, bool profile,
                        double kspec_bytes, double main_bytes, double main_flops) {
    if (!lds_fits(ctx, reinterpret_cast<const void*>(batch_kspec_kernel), kLdsBytes) ||
        !lds_fits(ctx, reinterpret_cast<const void*>(batch_main_kernel), kLdsBytes))
        return NDC_ERR_UNSUPPORTED;
    const int kgrid = a.nkitems > 0 ? a.nkitems : 1;
    if (profile) ctx.kbegin("batch_kspec", kspec_bytes, 0);
    NDC_LAUNCH(ctx, stream, batch_kspec_kernel, dim3(kgrid), dim3(kBatchThreads), kLdsBytes, a);
    if (profile) ctx.kend();
    if (hipGetLastError() != hipSuccess) {
        ctx.err = "batch_kspec launch failed";
        return NDC_ERR_DEVICE;
    }
    if (a.nitems > 0 && grid > 0) {
        if (profile) ctx.kbegin("batch_main", main_bytes, main_flops);
        NDC_LAUNCH(ctx, stream, batch_main_kernel, dim3(grid), dim3(kBatchThreads), kLdsBytes, a);
        if (profile) ctx.kend();
        if (hipGetLastError() != hipSuccess) {
            ctx.err = "batch_main launch failed";
            return NDC_ERR_DEVICE;
        }
    }
    return NDC_OK;
}

}  // namespace ndc
