// FFT convolution on gfx950 (conv_fft/mod.rs:185-292).
//
// The padded data and the scattered kernel are never materialised: the first
// pass reads x through the per-axis border maps (conv_fft/padding.rs:30-62,
// padding/half_dim.rs), evaluated in closed form per cell (border1d.hpp),
// and k through the dilation/reversal scatter rule
// (conv_fft/padding.rs:78-111) while loading rows into LDS.  Transforms are
// mixed-radix Stockham FFTs over lines held in LDS.  Each line plan carries a
// compact twiddle table generated in FP64 on the host; a workgroup stages it
// into LDS once, so no stage touches global memory.  Axis order and layout:
//
//   pass 1  rows (last axis): r2c by the half-length trick (even F) or a full
//           complex FFT (odd F / complex input), data and kernel in one grid;
//   pass 2  axes N-2 .. 1 forward, columns in place, data and kernel together;
//   pass 3  axis 0: forward data and kernel columns, multiply, inverse, keep
//           only the rows the crop needs (conv_fft/mod.rs:268,278 and :282-289);
//   pass 4  axes 1 .. N-2 inverse, cropped rows only;
//   pass 5  rows inverse (c2r), scale 1/prod F (real.rs:278-279,
//           complex.rs:141-142), crop and stride on store.
// A rank-1 problem runs as one fused kernel.  Rows beyond the padded extent
// are known to be zero and are neither read nor written.  Lines longer than
// LDS take a four-step path (see below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "border1d.hpp"
#include "device_common.hpp"
#include "fft_bfly.hpp"
#include "launch.hpp"

namespace ndc {
namespace {

using namespace fftb;

constexpr int kMaxThreads = 1024;
constexpr int kMaxStages = 24;
constexpr int kU = 4;   // elements per thread per batched global-load round

__device__ __forceinline__ int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }

// a / d for a < 2^24 and the float reciprocal r of d: the float quotient is
// within one of the true one, and one correction step fixes it.
__device__ __forceinline__ int fdiv(int a, int d, float r) {
    int q = int(float(a) * r);
    const int rem = a - q * d;
    if (rem < 0) --q;
    else if (rem >= d) ++q;
    return q;
}

template <typename R>
struct LinePlan {
    int L;                  // complex length transformed
    int nst;                // Stockham stages
    int ntw;                // entries of the compact twiddle table
    int post;               // offset of W_F^k, k = 0..L, for the real-even post/pre-processing; -1 if none
    const Cx<R>* tw;        // compact table in global memory
    int rad[kMaxStages];
    int base[kMaxStages];   // stage twiddles W_{Ns*P}^(r*k) at base + (r-1)*Ns + k (consecutive k adjacent)
    int wp[kMaxStages];     // radix-P roots W_P^m at wp + m (stages without a fixed butterfly)
};

// LDS layout of a line: element t sits at slot pp(t).  Every line access goes
// through this pair, so the layout is decided here.  It is the plain one: a
// padded layout (one spare slot after every 16 elements) removes the bank
// conflicts of the early Stockham stages' stride-P writes, but on gfx950 the
// extra address arithmetic cost more than the conflicts (measured on the 1-D,
// 2-D and 3-D workloads: 6% slower end to end).
__host__ __device__ __forceinline__ int pp(int t) { return t; }
// LDS slots a line of L elements spans in that layout.
__host__ __device__ __forceinline__ int pp_len(int L) { return L; }

// ---------------------------------------------------------------------------
// Stockham stages over lines in LDS: element (line b, index t) at b*ls + pp(t)*es.
// LINE_FAST selects the thread->item order: line index fastest for column
// tiles (es = #lines, consecutive lines adjacent in LDS), element index
// fastest for rows (es = 1).  `tw` points at the plan's table in LDS.
// ---------------------------------------------------------------------------

// SHIFT: the item splits are powers of two (Ns, and the line count or the
// butterflies per line) or the block holds a single line, so they take shifts
// and masks instead of the reciprocal divisions; the stages are bound by the
// vector instructions of the busiest SIMD, so the index arithmetic counts.
template <typename R, int P, bool LINE_FAST, bool INV, bool SHIFT>
__device__ __forceinline__ void stage_fixed(const Cx<R>* __restrict__ src, Cx<R>* __restrict__ dst, int L, int Ns,
                                            int nlines, int ls, int es, const Cx<R>* __restrict__ twb, int T) {
    using C = Cx<R>;
    const int nbf = L / P;
    const int total = nbf * nlines;
    const int dv = LINE_FAST ? nlines : nbf;
    const float rdv = 1.0f / float(dv);
    const float rns = 1.0f / float(Ns);
    const int ldv = SHIFT ? __builtin_ctz(unsigned(dv)) : 0;
    const int lns = SHIFT ? __builtin_ctz(unsigned(Ns)) : 0;
    const bool one = !LINE_FAST && nlines == 1;
    for (int it = threadIdx.x; it < total; it += T) {
        int b, j, q, k;
        if (SHIFT) {
            if (LINE_FAST) { j = it >> ldv; b = it & (dv - 1); }
            else if (one) { b = 0; j = it; }
            else { b = it >> ldv; j = it & (dv - 1); }
            q = j >> lns;
            k = j & (Ns - 1);
        } else {
            if (LINE_FAST) { j = fdiv(it, dv, rdv); b = it - j * dv; }
            else { b = fdiv(it, dv, rdv); j = it - b * dv; }
            q = fdiv(j, Ns, rns);
            k = j - q * Ns;
        }
        const C* s = src + b * ls;
        // All LDS reads of the butterfly (data and twiddles) are issued before
        // any arithmetic: left to itself the scheduler interleaves the twiddle
        // reads with the multiplies and serialises several LDS round trips.
        C v[P], w[P];
#pragma unroll
        for (int r = 0; r < P; ++r) v[r] = s[pp(j + r * nbf) * es];
        if (Ns > 1) {
            const C* wb = twb + k - Ns;
#pragma unroll
            for (int r = 1; r < P; ++r) w[r] = wb[r * Ns];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < P; ++r) v[r] = cmulc(v[r], w[r], INV);
        }
        bfly<P>(v, INV);
        C* d = dst + b * ls;
        const int o = q * Ns * P + k;
#pragma unroll
        for (int r = 0; r < P; ++r) d[pp(o + r * Ns) * es] = v[r];
    }
}

// Any radix: one output per item, O(P) reads (used for the primes above 7
// that good_size can produce: 11, 13, 17, 31, 43, 73, 127, 257).
template <typename R, bool LINE_FAST, bool INV>
__device__ __forceinline__ void stage_generic(const Cx<R>* __restrict__ src, Cx<R>* __restrict__ dst, int L, int Ns,
                                              int P, int nlines, int ls, int es, const Cx<R>* __restrict__ twb,
                                              const Cx<R>* __restrict__ wP, int T) {
    using C = Cx<R>;
    const int nbf = L / P;
    const int total = L * nlines;
    const int dv = LINE_FAST ? nlines : L;
    const float rdv = 1.0f / float(dv), rns = 1.0f / float(Ns), rp = 1.0f / float(P);
    for (int it = threadIdx.x; it < total; it += T) {
        int b, o;
        if (LINE_FAST) { o = fdiv(it, dv, rdv); b = it - o * dv; }
        else { b = fdiv(it, dv, rdv); o = it - b * dv; }
        const int t = fdiv(o, Ns, rns);
        const int k = o - t * Ns;
        const int jq = fdiv(t, P, rp);
        const int rout = t - jq * P;
        const int j = jq * Ns + k;
        const C* s = src + b * ls;
        C acc = cmake<C>(R(0), R(0));
        int m = 0;   // (r * rout) mod P
        for (int r = 0; r < P; ++r) {
            C v = s[pp(j + r * nbf) * es];
            if (Ns > 1 && r) v = cmulc(v, twb[(r - 1) * Ns + k], INV);
            acc = cadd(acc, cmulc(v, wP[m], INV));
            m += rout;
            if (m >= P) m -= P;
        }
        dst[b * ls + pp(o) * es] = acc;
    }
}

// Runs all stages; input in b0; returns the index (0/1) of the buffer holding
// the output.  Ends with a barrier.  The direction is a template parameter of
// the stages so that no butterfly carries a runtime conjugation select.
template <typename R, bool LINE_FAST, bool INV>
__device__ __forceinline__ int fft_lines_dir(Cx<R>* b0, Cx<R>* b1, const LinePlan<R>& pl, const Cx<R>* tw,
                                             int nlines, int ls, int es, int T) {
    // The two buffers are picked by selects, not through a local pointer
    // array: an array of pointers hides that both live in LDS, and every
    // access would then be compiled as a flat (generic) access.
    int cur = 0;
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int P = pl.rad[s];
        const Cx<R>* src = cur ? b1 : b0;
        Cx<R>* dst = cur ? b0 : b1;
        const Cx<R>* twb = tw + pl.base[s];
        const int dvs = LINE_FAST ? nlines : pl.L / P;
        const bool shift = (Ns & (Ns - 1)) == 0 &&
                           (((dvs & (dvs - 1)) == 0) || (!LINE_FAST && nlines == 1));
        switch (P) {
#define NDC_STAGE(PP)                                                                                     \
    case PP:                                                                                              \
        if (shift) stage_fixed<R, PP, LINE_FAST, INV, true>(src, dst, pl.L, Ns, nlines, ls, es, twb, T);  \
        else stage_fixed<R, PP, LINE_FAST, INV, false>(src, dst, pl.L, Ns, nlines, ls, es, twb, T);       \
        break;
            NDC_STAGE(2)
            NDC_STAGE(3)
            NDC_STAGE(4)
            NDC_STAGE(5)
            NDC_STAGE(7)
            NDC_STAGE(8)
#undef NDC_STAGE
            case 16:
                if constexpr (sizeof(R) == 4) {
                    if (shift) stage_fixed<R, 16, LINE_FAST, INV, true>(src, dst, pl.L, Ns, nlines, ls, es, twb, T);
                    else stage_fixed<R, 16, LINE_FAST, INV, false>(src, dst, pl.L, Ns, nlines, ls, es, twb, T);
                }
                break;
            default:
                stage_generic<R, LINE_FAST, INV>(src, dst, pl.L, Ns, P, nlines, ls, es, twb, tw + pl.wp[s], T);
                break;
        }
        __syncthreads();
        cur ^= 1;
        Ns *= P;
    }
    return cur;
}

// T is the block's thread count, read once at kernel entry: blockDim.x is a
// load from the dispatch packet, and the barriers between stages would make
// every stage issue it again and wait for it.
template <typename R, bool LINE_FAST>
__device__ __forceinline__ int fft_lines(Cx<R>* b0, Cx<R>* b1, const LinePlan<R>& pl, const Cx<R>* tw, int nlines,
                                         int ls, int es, bool inv, int T) {
    return inv ? fft_lines_dir<R, LINE_FAST, true>(b0, b1, pl, tw, nlines, ls, es, T)
               : fft_lines_dir<R, LINE_FAST, false>(b0, b1, pl, tw, nlines, ls, es, T);
}

// Copies n entries global -> LDS with kU loads in flight per thread.
template <typename C>
__device__ __forceinline__ void stage_table(C* __restrict__ dst, const C* __restrict__ src, int n, int T) {
    for (int base = threadIdx.x; base < n; base += kU * T) {
        C v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = base + u * T;
            if (i < n) v[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = base + u * T;
            if (i < n) dst[i] = v[u];
        }
    }
}

// Twiddle staging split around a kernel's data loads: begin() issues the
// global loads of the first kTabU * blockDim entries into registers, end()
// writes them to LDS and copies any remainder, so the table's load latency
// overlaps the data's instead of preceding it.
template <typename C>
struct TableStage {
    static constexpr int kTabU = 4;
    C v[kTabU];
    __device__ __forceinline__ void begin(const C* __restrict__ src, int n, int T) {
#pragma unroll
        for (int u = 0; u < kTabU; ++u) {
            const int i = threadIdx.x + u * T;
            if (i < n) v[u] = src[i];
        }
    }
    __device__ __forceinline__ void end(C* __restrict__ dst, const C* __restrict__ src, int n, int T) {
#pragma unroll
        for (int u = 0; u < kTabU; ++u) {
            const int i = threadIdx.x + u * T;
            if (i < n) dst[i] = v[u];
        }
        for (int i = threadIdx.x + kTabU * T; i < n; i += T) dst[i] = src[i];
    }
};

// Copies a contiguous line of n entries global -> LDS in the padded layout.
template <typename C>
__device__ __forceinline__ void stage_line(C* __restrict__ dst, const C* __restrict__ src, int n, int T) {
    for (int base = threadIdx.x; base < n; base += kU * T) {
        C v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = base + u * T;
            if (i < n) v[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = base + u * T;
            if (i < n) dst[pp(i)] = v[u];
        }
    }
}

// Entries of a row plan's table that the row passes stage into LDS: the
// stage twiddles; the real-even post/pre-processing factors W_F^k (used
// once per bin) are read from global memory where they are needed.
template <typename R>
__host__ __device__ __forceinline__ int row_stage_tw(const LinePlan<R>& pl) {
    return pl.post >= 0 ? pl.post : pl.ntw;
}

// ---------------------------------------------------------------------------
// Load stages
// ---------------------------------------------------------------------------

// What the data side needs to read the padded input through the border maps.
template <typename E>
// The per-axis border map is evaluated in closed form (resolve1d), so no
// load phase waits on a map lookup.
struct PadSrc {
    const E* x;
    E cst[kMaxRank][2];
    int64_t xstr[kMaxRank];
    int32_t lo[kMaxRank];           // interior [lo, lo + n) maps to x index p - lo
    int32_t n[kMaxRank];
    int32_t pds[kMaxRank];          // padded extent; cells beyond it are zero
    int32_t side[kMaxRank][2];      // Side1D of the front and back fill
};

// Border-map code of padded cell q of axis e (plan.cpp axis_map, position by
// position): an index into x, or kUnwritten / kConstFront / kConstBack /
// kOutside.
template <typename E>
__device__ __forceinline__ int32_t axis_code(const PadSrc<E>& s, int e, int q) {
    if (q >= s.pds[e]) return kOutside;
    return resolve1d<int>(q, s.n[e], s.lo[e], s.side[e][0], s.side[e][1]);
}

// What the kernel side needs to read the scattered, dilated kernel.
template <typename E>
struct KerSrc {
    const E* k;
    int64_t kstr[kMaxRank];
    int32_t K[kMaxRank];
    int32_t d[kMaxRank];
    int32_t rev;
};

// Outer-coordinate resolution of one data row.
template <typename E>
struct RowRes {
    int64_t xoff;
    int oc_axis;    // highest outer axis whose fill wrote a constant, or -1
    E ocval;
    bool unw;       // some outer axis is an unwritten pad cell
    bool zero;      // some outer coordinate lies beyond the padded extent
};

template <typename E>
__device__ __forceinline__ E zero_of() {
    if constexpr (std::is_same<E, float2>::value || std::is_same<E, double2>::value) {
        E z;
        z.x = 0;
        z.y = 0;
        return z;
    } else {
        return E(0);
    }
}

// Value of the padded row at position p of the last axis.  Interior
// positions are computed without touching the border map.
template <typename E>
__device__ __forceinline__ E pad_value(const PadSrc<E>& s, const RowRes<E>& rr, int ax, int p) {
    if (rr.zero) return zero_of<E>();
    const int rel = p - s.lo[ax];
    if (rel >= 0 && rel < s.n[ax]) {
        if (rr.oc_axis >= 0) return rr.ocval;
        if (rr.unw) return zero_of<E>();
        return s.x[rr.xoff + int64_t(rel) * s.xstr[ax]];
    }
    const int32_t c = axis_code(s, ax, p);
    if (c == kOutside) return zero_of<E>();
    if (c <= kConstFront) return s.cst[ax][c == kConstFront ? 0 : 1];
    if (rr.oc_axis >= 0) return rr.ocval;
    if (c == kUnwritten || rr.unw) return zero_of<E>();
    return s.x[rr.xoff + int64_t(c) * s.xstr[ax]];
}

// Kernel tap at position p of the scattered, dilated line (32-bit: p is an
// index into an LDS-resident line).
template <typename E>
__device__ __forceinline__ E ker_value(const KerSrc<E>& s, int64_t koff, int ax, int p) {
    const int dd = s.d[ax];
    const int i = dd == 1 ? p : int(unsigned(p) / unsigned(dd));
    if (i * dd != p || i >= s.K[ax]) return zero_of<E>();
    return s.k[koff + int64_t(s.rev ? i : s.K[ax] - 1 - i) * s.kstr[ax]];
}

// Last-axis constants of the data border map, read once per kernel.
template <typename E>
struct LastAxisPad {
    int lo, n, pds, front, back;
    int64_t xs;
    E cfront, cback;
};

template <typename E>
__device__ __forceinline__ LastAxisPad<E> last_axis_pad(const PadSrc<E>& s, int ax) {
    LastAxisPad<E> la;
    la.lo = s.lo[ax];
    la.n = s.n[ax];
    la.pds = s.pds[ax];
    la.front = s.side[ax][0];
    la.back = s.side[ax][1];
    la.xs = s.xstr[ax];
    la.cfront = s.cst[ax][0];
    la.cback = s.cst[ax][1];
    return la;
}

// Whether padded cell p of a data row reads x: if so *off is its offset into
// x (negative for views with negative strides), else *fill holds the cell's
// value.  pad_value's rule with no memory access, so that a caller can issue
// all of a round's loads before using any of them.
template <typename E>
__device__ __forceinline__ bool pad_offset(const LastAxisPad<E>& la, const RowRes<E>& rr, int p, int64_t* off,
                                           E* fill) {
    *fill = zero_of<E>();
    *off = 0;
    if (rr.zero) return false;
    const int rel = p - la.lo;
    int c;
    if (rel >= 0 && rel < la.n) c = rel;
    else if (p >= la.pds) c = kOutside;
    else c = resolve1d<int>(p, la.n, la.lo, la.front, la.back);
    if (c == kOutside) return false;
    if (c == kConstFront || c == kConstBack) {
        *fill = c == kConstFront ? la.cfront : la.cback;
        return false;
    }
    if (rr.oc_axis >= 0) {
        *fill = rr.ocval;
        return false;
    }
    if (c == kUnwritten || rr.unw) return false;
    *off = rr.xoff + int64_t(c) * la.xs;
    return true;
}

// Last-axis constants of the kernel scatter.
struct LastAxisKer {
    int d, K, rev;
    int64_t ks;
};

template <typename E>
__device__ __forceinline__ LastAxisKer last_axis_ker(const KerSrc<E>& s, int ax) {
    LastAxisKer k;
    k.d = s.d[ax];
    k.K = s.K[ax];
    k.rev = s.rev;
    k.ks = s.kstr[ax];
    return k;
}

// Whether position p of a scattered kernel row holds a tap; if so *off is
// its offset into k (negative for views with negative strides).
__device__ __forceinline__ bool ker_offset(const LastAxisKer& k, int64_t koff, int p, int64_t* off) {
    const int i = k.d == 1 ? p : int(unsigned(p) / unsigned(k.d));
    *off = 0;
    if (i * k.d != p || i >= k.K) return false;
    *off = koff + int64_t(k.rev ? i : k.K - 1 - i) * k.ks;
    return true;
}

template <typename R, typename E>
__device__ __forceinline__ Cx<R> to_cx(E v) {
    if constexpr (std::is_same<E, Cx<R>>::value) return v;
    else return cmake<Cx<R>>(R(v), R(0));
}

// Complex sample n of a row for the FFT: the pair (x[2n], x[2n+1]) for the
// half-length real transform, x[n] otherwise.
template <typename R, typename E, typename F>
__device__ __forceinline__ Cx<R> row_sample(int mode, int n, F&& val) {
    if (mode == 0) {
        if constexpr (std::is_same<E, Cx<R>>::value) return val(n);
        else return cmake<Cx<R>>(R(val(2 * n)), R(val(2 * n + 1)));
    }
    return to_cx<R>(val(n));
}

// X[k] of a real line of length F = 2M from the half-length spectrum Z (an
// LDS line in the padded layout)
// (X = E + W_F^k O with E = (Z[k] + conj Z[M-k]) / 2, O = -i (Z[k] - conj Z[M-k]) / 2).
template <typename C>
__device__ __forceinline__ C r2c_bin(const C* Z, int M, int k, const C* wpost, int es = 1) {
    using R = decltype(C{}.x);
    const C zk = Z[pp(k == M ? 0 : k) * es];
    const C zm = cconj(Z[pp(k == 0 ? 0 : M - k) * es]);
    const C e = cscale(cadd(zk, zm), R(0.5));
    const C o = cmul_negi(cscale(csub(zk, zm), R(0.5)));
    return cadd(e, cmul(o, wpost[k]));
}

// r2c_bin with the factor W_F^k passed by value.
template <typename C>
__device__ __forceinline__ C r2c_bin_w(const C* Z, int M, int k, C w) {
    using R = decltype(C{}.x);
    const C zk = Z[pp(k == M ? 0 : k)];
    const C zm = cconj(Z[pp(k == 0 ? 0 : M - k)]);
    const C e = cscale(cadd(zk, zm), R(0.5));
    const C o = cmul_negi(cscale(csub(zk, zm), R(0.5)));
    return cadd(e, cmul(o, w));
}

// Z[k] of the half-length inverse from bins A = X[k] and B = conj X[M-k]:
// (A + B) + i W_F^-k (A - B); IFFT_M of it is F times the real line.
template <typename C>
__device__ __forceinline__ C c2r_pre(C A, C B, int k, const C* wpost) {
    const C t = cmul(csub(A, B), cconj(wpost[k]));
    return cadd(cadd(A, B), cmul_i(t));
}

// ---------------------------------------------------------------------------
// Pass 1: forward rows (last axis) for data (job 0) and kernel (job 1)
// ---------------------------------------------------------------------------

constexpr int kMaxLpb = 16;

template <typename R, typename E>
struct RowsFwdArgs {
    PadSrc<E> pad;
    KerSrc<E> ker;
    Cx<R>* spec[2];          // output spectra (data, kernel)
    int64_t sstr[kMaxRank];  // spectrum strides of the outer axes (complex elements)
    int64_t cnt[2][kMaxRank];// active extent of each outer axis per job
    int64_t nlines[2];
    int32_t nblocks0;        // blocks of job 0
    int32_t nout;            // number of outer axes (N-1)
    int32_t F;               // FFT size of the last axis
    int32_t H;               // bins written per row
    int32_t mode;            // 0: real even (half-length), 1: real odd (full complex), 2: complex input
    int32_t lpb;             // lines per block
    int32_t ls;              // LDS line stride (complex elements)
    LinePlan<R> pl;
};

template <typename R, typename E>
__global__ void __launch_bounds__(kMaxThreads) rows_fwd_kernel(RowsFwdArgs<R, E> a) {
    using C = Cx<R>;
    constexpr bool kCE = std::is_same<E, C>::value;
    extern __shared__ unsigned char smem_raw[];
    const int T = blockDim.x;
    const int ntw = row_stage_tw(a.pl);
    TableStage<C> tab;
    tab.begin(a.pl.tw, ntw, T);
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* b0 = tw + ntw;
    C* b1 = b0 + a.lpb * a.ls;
    __shared__ int64_t s_off[kMaxLpb];
    __shared__ int64_t s_xoff[kMaxLpb];
    __shared__ int s_flag[kMaxLpb];     // bit0 zero row, bit1 unwritten, oc axis + 1 in bits 8..
    __shared__ E s_oc[kMaxLpb];

    const int job = blockIdx.x < unsigned(a.nblocks0) ? 0 : 1;
    const int blk = job == 0 ? int(blockIdx.x) : int(blockIdx.x) - a.nblocks0;
    const int line0 = blk * a.lpb;
    const int nl = int(imin64(a.lpb, a.nlines[job] - line0));
    const int ax = a.nout;   // last axis index

    // post-processing factors of this thread's first store round, loaded now
    // so that they arrive with the input rather than after the transform
    // (single precision: the double kernels have no registers to spare)
    constexpr bool kPre = sizeof(R) == 4;
    const int H = a.H;
    const int nH = nl * H;
    const float rH = 1.0f / float(H);
    const C* wpost = a.pl.tw + a.pl.post;
    C wk0[kU];
    if (kPre && a.mode == 0) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = int(threadIdx.x) + u * T;
            const int itc = it < nH ? it : 0;
            const int b = fdiv(itc, H, rH), k = itc - b * H;
            wk0[u] = wpost[k];
        }
    }

    // per-line outer resolution (closed-form maps: no memory reads)
    if (threadIdx.x < nl) {
        int rem = line0 + int(threadIdx.x);
        int64_t soff = 0, xoff = 0;
        int flag = 0;
        int oc = -1;
        E ocv = zero_of<E>();
        for (int e = a.nout - 1; e >= 0; --e) {
            const int cnt = int(a.cnt[job][e]);
            const int q = rem % cnt;
            rem /= cnt;
            soff += int64_t(q) * a.sstr[e];
            if (job == 0) {
                const int32_t c = axis_code(a.pad, e, q);
                if (c >= 0) xoff += int64_t(c) * a.pad.xstr[e];
                else if (c == kOutside) flag |= 1;
                else if (c == kUnwritten) flag |= 2;
                else if (oc < e) { oc = e; ocv = a.pad.cst[e][c == kConstFront ? 0 : 1]; }
            } else {
                const int i = q / a.ker.d[e];
                if (i * a.ker.d[e] != q || i >= a.ker.K[e]) flag |= 1;
                const int ki = a.ker.rev ? i : a.ker.K[e] - 1 - i;
                xoff += int64_t(ki) * a.ker.kstr[e];
            }
        }
        s_off[threadIdx.x] = soff;
        s_xoff[threadIdx.x] = xoff;
        s_flag[threadIdx.x] = flag | ((oc + 1) << 8);
        s_oc[threadIdx.x] = ocv;
    }
    __syncthreads();

    // Batched load of the rows.  Every sample's source offset (or fill value)
    // is computed without touching memory, then all loads of the round are
    // issued back to back from clamped addresses and the fills selected
    // afterwards: no load sits behind a branch, so they overlap.
    const int L = a.pl.L;
    const int total = nl * L;
    const float rL = 1.0f / float(L);
    const LastAxisPad<E> la = last_axis_pad(a.pad, ax);
    const LastAxisKer lk = last_axis_ker(a.ker, ax);
    const E* src = job == 0 ? a.pad.x : a.ker.k;
    constexpr int kS = kCE ? 1 : 2;               // samples per complex point (half-length real rows)
    const int spt = (!kCE && a.mode == 0) ? 2 : 1;
    for (int base = threadIdx.x; base < total; base += kU * T) {
        int64_t off[kU][kS];
        bool ok[kU][kS];
        E fill[kU][kS];
        int at[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = base + u * T;
            const int itc = it < total ? it : 0;
            const int b = fdiv(itc, L, rL), n = itc - b * L;
            at[u] = b * a.ls + pp(n);
            const int flag = s_flag[b];
            const RowRes<E> rr{s_xoff[b], (flag >> 8) - 1, s_oc[b], (flag & 2) != 0, (flag & 1) != 0 || it >= total};
#pragma unroll
            for (int q = 0; q < kS; ++q) {
                fill[u][q] = zero_of<E>();
                off[u][q] = 0;
                ok[u][q] = false;
                if (q < spt) {
                    const int pos = spt * n + q;
                    if (job == 0) ok[u][q] = pad_offset(la, rr, pos, &off[u][q], &fill[u][q]);
                    else if (!rr.zero) ok[u][q] = ker_offset(lk, rr.xoff, pos, &off[u][q]);
                }
            }
        }
        E v[kU][kS];
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int q = 0; q < kS; ++q)
                if (q < spt) v[u][q] = src[off[u][q]];   // offset 0 (a valid element) where !ok
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (base + u * T >= total) continue;
            if constexpr (kCE) {
                b0[at[u]] = ok[u][0] ? v[u][0] : fill[u][0];
            } else {
                const E v0 = ok[u][0] ? v[u][0] : fill[u][0];
                const E v1 = spt == 2 ? (ok[u][1] ? v[u][1] : fill[u][1]) : E(0);
                b0[at[u]] = cmake<C>(R(v0), R(v1));
            }
        }
    }
    tab.end(tw, a.pl.tw, ntw, T);
    __syncthreads();
    const int cur = fft_lines<R, false>(b0, b1, a.pl, tw, nl, a.ls, 1, false, T);
    const C* Z = cur ? b1 : b0;

    // store H bins per row; the post-processing factors of a round are loaded
    // together before any of them is used
    Cx<R>* out = a.spec[job];
    for (int base = threadIdx.x; base < nH; base += kU * T) {
        C wk[kU];
        if (kPre && a.mode == 0 && base == int(threadIdx.x)) {
#pragma unroll
            for (int u = 0; u < kU; ++u) wk[u] = wk0[u];
        } else if (a.mode == 0) {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int it = base + u * T;
                const int itc = it < nH ? it : 0;
                const int b = fdiv(itc, H, rH), k = itc - b * H;
                wk[u] = wpost[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = base + u * T;
            if (it >= nH) continue;
            const int b = fdiv(it, H, rH), k = it - b * H;
            const C* Zb = Z + b * a.ls;
            const C X = a.mode == 0 ? r2c_bin_w(Zb, L, k, wk[u]) : Zb[pp(k)];
            out[s_off[b] + k] = X;
        }
    }
}

// ---------------------------------------------------------------------------
// Passes 2-4: column transforms along axis d of the spectrum (in place)
// ---------------------------------------------------------------------------

struct AxisSel {
    int64_t start, step, count;
};

template <typename R>
struct ColsArgs {
    Cx<R>* spec[2];
    int64_t sd;                      // element stride between consecutive points of a line
    int64_t ostr[kMaxRank];          // element stride of each outer index
    AxisSel outer[2][kMaxRank];      // selection per outer index per job
    int32_t nouter;
    int64_t nlines_outer[2];         // product of outer counts per job
    int32_t ntiles;                  // bin tiles along the last axis
    int32_t nblocks0;
    int32_t H;                       // bins
    int32_t C;                       // tile width (lines per block)
    int32_t valid[2];                // rows of axis d holding data (rest are zero)
    AxisSel outsel;                  // rows of axis d to write back
    int32_t mode;                    // 0 forward (2 jobs), 1 fused forward*kernel->inverse, 2 inverse,
                                     // 3 fused with the kernel spectrum already transformed
    int32_t nblocks;                 // grid size
    int32_t joint;                   // mode 1: kernel and data columns transformed as one batch
    // Four-step split of a column axis of length F = L1 * L2 (point j at
    // j1 L2 + j2): the L1-point pass runs along j1 with j2 the outer index
    // tw_outer.  Rows j1 with j1 L2 + j2 >= valid are zero (split_L2 = L2;
    // 0 when the line is a whole axis).  tw_dir 1 multiplies its forward
    // output k1 by W_F^(k1 j2), 2 multiplies its inverse input by the
    // conjugate; twF holds W_F^i, i < F.
    const Cx<R>* twF;
    int32_t split_L2;
    int32_t tw_outer;
    int32_t tw_dir;
    LinePlan<R> pl;
};

// Workgroups are dispatched round-robin over the eight XCDs, each with its own
// L2.  Neighbouring column tiles share cache lines (a tile holds Ct adjacent
// bins of every row), so the eight XCDs take consecutive ranges of tiles:
// workgroup g, the (g/8)-th on XCD g%8, gets slot start(g%8) + g/8.  The map
// is a permutation of [0, n) for every n.
__device__ __forceinline__ int64_t xcd_slot(int64_t g, int64_t n) {
    const int64_t per = n >> 3, rem = n & 7, x = g & 7, k = g >> 3;
    return x * per + (x < rem ? x : rem) + k;
}

template <typename R>
__global__ void __launch_bounds__(kMaxThreads) cols_kernel(ColsArgs<R> a) {
    using C = Cx<R>;
    extern __shared__ unsigned char smem_raw[];
    const int L = a.pl.L;
    const int Ct = a.C;
    const int PL = pp_len(L) * Ct;   // LDS slots of one tile buffer
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* b0 = tw + a.pl.ntw;
    C* b1 = b0 + PL;
    C* b2 = b1 + PL;
    const int T = blockDim.x;

    const int64_t g = xcd_slot(blockIdx.x, a.nblocks);
    const int job = (a.mode == 0 && g >= a.nblocks0) ? 1 : 0;
    int64_t blk = job == 0 ? g : g - a.nblocks0;
    const int tile = int(blk % a.ntiles);
    int64_t rem = blk / a.ntiles;
    int64_t base = 0;
    int j2 = 0;   // the split's outer coordinate (split passes only)
    for (int i = a.nouter - 1; i >= 0; --i) {
        const AxisSel& s = a.outer[job][i];
        const int64_t q = rem % s.count;
        rem /= s.count;
        const int64_t pos = s.start + q * s.step;
        if (i == a.tw_outer) j2 = int(pos);
        base += pos * a.ostr[i];
    }
    const int c0 = tile * Ct;
    const int nc = min(Ct, a.H - c0);
    const int64_t sd = a.sd;
    const float rC = 1.0f / float(Ct);
    // rows of this line holding data: j1 L2 + j2 < valid in a split pass
    auto valid_rows = [&](int valid) {
        if (a.split_L2 <= 0) return valid;
        return valid > j2 ? (valid - j2 + a.split_L2 - 1) / a.split_L2 : 0;
    };
    const bool pre_tw = a.tw_dir == 2;

    // tile column c of row t goes to dstbuf[pp(t) * es + c] (es = Ct, or 2 Ct
    // in the fused mode's joint kernel + data buffer)
    auto load = [&](const C* src, C* dstbuf, int valid, int es) {
        valid = valid_rows(valid);
        const int total = L * Ct;
        for (int b = threadIdx.x; b < total; b += kU * T) {
            C v[kU];
            int at[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int it = b + u * T;
                v[u] = cmake<C>(R(0), R(0));
                const int t = fdiv(it, Ct, rC), c = it - t * Ct;
                at[u] = pp(t) * es + c;
                if (it < total && t < valid && c < nc) {
                    v[u] = src[base + int64_t(t) * sd + c0 + c];
                    if (pre_tw) v[u] = cmul(v[u], cconj(a.twF[t * j2]));
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int it = b + u * T;
                if (it < total) dstbuf[at[u]] = v[u];
            }
        }
    };

    stage_table(tw, a.pl.tw, a.pl.ntw, T);
    C* res;
    int res_es = Ct;   // element stride of the result lines
    if (a.mode == 0) {
        load(a.spec[job], b0, a.valid[job], Ct);
        __syncthreads();
        res = fft_lines<R, true>(b0, b1, a.pl, tw, Ct, 1, Ct, false, T) ? b1 : b0;
    } else if (a.mode == 1 && !a.joint) {
        // kernel and data columns loaded together, then transformed in turn
        load(a.spec[1], b1, a.valid[1], Ct);
        load(a.spec[0], b0, a.valid[0], Ct);
        __syncthreads();
        const int ck = fft_lines<R, true>(b1, b2, a.pl, tw, Ct, 1, Ct, false, T);
        C* kres = ck ? b2 : b1;
        C* kfree = ck ? b1 : b2;
        const int cd = fft_lines<R, true>(b0, kfree, a.pl, tw, Ct, 1, Ct, false, T);
        C* dres = cd ? kfree : b0;
        C* dfree = cd ? b0 : kfree;
        for (int it = threadIdx.x; it < PL; it += T) dres[it] = cmul(dres[it], kres[it]);
        __syncthreads();
        res = fft_lines<R, true>(dres, dfree, a.pl, tw, Ct, 1, Ct, true, T) ? dfree : dres;
    } else if (a.mode == 1) {
        // kernel columns (lines 0..Ct-1) and data columns (lines Ct..2Ct-1) in
        // one interleaved buffer, transformed together (twice the butterflies
        // per stage, one set of stage barriers); the product lands on the data
        // lines, which alone are inverted
        C* j0 = b0;
        C* j1 = b0 + 2 * PL;
        load(a.spec[1], j0, a.valid[1], 2 * Ct);
        load(a.spec[0], j0 + Ct, a.valid[0], 2 * Ct);
        __syncthreads();
        const int cj = fft_lines<R, true>(j0, j1, a.pl, tw, 2 * Ct, 1, 2 * Ct, false, T);
        C* jr = cj ? j1 : j0;
        C* jf = cj ? j0 : j1;
        for (int it = threadIdx.x; it < PL; it += T) {
            const int t = fdiv(it, Ct, rC), c = it - t * Ct;
            const int e = t * 2 * Ct + c;   // it walks the LDS slots pp(t) of the tile (PL = pp_len(L) * Ct)
            jr[e + Ct] = cmul(jr[e + Ct], jr[e]);
        }
        __syncthreads();
        res = (fft_lines<R, true>(jr + Ct, jf + Ct, a.pl, tw, Ct, 1, 2 * Ct, true, T) ? jf : jr) + Ct;
        res_es = 2 * Ct;
    } else if (a.mode == 3) {
        // kernel columns already transformed (kept spectrum): data forward,
        // product, inverse
        load(a.spec[1], b1, L, Ct);
        load(a.spec[0], b0, a.valid[0], Ct);
        __syncthreads();
        const int cd = fft_lines<R, true>(b0, b2, a.pl, tw, Ct, 1, Ct, false, T);
        C* dres = cd ? b2 : b0;
        C* dfree = cd ? b0 : b2;
        for (int it = threadIdx.x; it < PL; it += T) dres[it] = cmul(dres[it], b1[it]);
        __syncthreads();
        res = fft_lines<R, true>(dres, dfree, a.pl, tw, Ct, 1, Ct, true, T) ? dfree : dres;
    } else {
        load(a.spec[0], b0, L, Ct);
        __syncthreads();
        res = fft_lines<R, true>(b0, b1, a.pl, tw, Ct, 1, Ct, true, T) ? b1 : b0;
    }
    // store selected rows
    C* dst = a.spec[job];
    const int nrow = int(a.outsel.count);
    const bool post_tw = a.tw_dir == 1;
    for (int it = threadIdx.x; it < nrow * Ct; it += T) {
        const int i = fdiv(it, Ct, rC), c = it - i * Ct;
        if (c >= nc) continue;
        const int t = int(a.outsel.start + int64_t(i) * a.outsel.step);
        C v = res[pp(t) * res_es + c];
        if (post_tw) v = cmul(v, a.twF[t * j2]);
        dst[base + int64_t(t) * sd + c0 + c] = v;
    }
}

// ---------------------------------------------------------------------------
// Pass 5: inverse rows (last axis), scale, crop, stride
// ---------------------------------------------------------------------------

template <typename R, typename E>
struct RowsInvArgs {
    const Cx<R>* spec;
    E* y;
    int64_t sstr[kMaxRank];
    AxisSel sel[kMaxRank];        // crop selection per outer axis
    int64_t ystr[kMaxRank];       // output strides of the outer axes
    int64_t nlines;
    int32_t nout;
    int32_t F, H, mode, lpb, ls;
    AxisSel lastsel;              // crop along the last axis
    R scale;
    LinePlan<R> pl;
};

template <typename R, typename E>
__device__ __forceinline__ E store_value(int mode, const Cx<R>* Z, int64_t p, R scale) {
    if constexpr (std::is_same<E, Cx<R>>::value) {
        return cscale(Z[pp(int(p))], scale);
    } else {
        if (mode == 0) {
            const Cx<R> zz = Z[pp(int(p >> 1))];
            return E(((p & 1) ? zz.y : zz.x) * scale);
        }
        return E(Z[pp(int(p))].x * scale);
    }
}

template <typename R, typename E>
__global__ void __launch_bounds__(kMaxThreads) rows_inv_kernel(RowsInvArgs<R, E> a) {
    using C = Cx<R>;
    extern __shared__ unsigned char smem_raw[];
    const int ntw = row_stage_tw(a.pl);
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* b0 = tw + ntw;
    C* b1 = b0 + a.lpb * a.ls;
    __shared__ int64_t s_off[kMaxLpb];
    __shared__ int64_t s_yoff[kMaxLpb];
    const int T = blockDim.x;
    const int64_t line0 = int64_t(blockIdx.x) * a.lpb;
    const int nl = int(imin64(a.lpb, a.nlines - line0));
    TableStage<C> tab;
    tab.begin(a.pl.tw, ntw, T);
    if (threadIdx.x < nl) {
        int64_t rem = line0 + threadIdx.x;
        int64_t soff = 0, yoff = 0;
        for (int e = a.nout - 1; e >= 0; --e) {
            const int64_t q = rem % a.sel[e].count;
            rem /= a.sel[e].count;
            soff += (a.sel[e].start + q * a.sel[e].step) * a.sstr[e];
            yoff += q * a.ystr[e];
        }
        s_off[threadIdx.x] = soff;
        s_yoff[threadIdx.x] = yoff;
    }
    __syncthreads();
    const int L = a.pl.L;
    const int M = L;
    const C* wpost = a.pl.tw + a.pl.post;
    const float rL = 1.0f / float(L);
    const int total = nl * L;
    // Rounds of kU bins per thread: every source index first, then all of the
    // round's loads, then the pre-processing and the LDS writes, so that no
    // load waits behind another.
    for (int base = threadIdx.x; base < total; base += kU * T) {
        int at[kU], kk[kU], i1[kU], i2[kU];
        const C* X[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = base + u * T;
            const int itc = it < total ? it : 0;
            const int b = fdiv(itc, L, rL), k = itc - b * L;
            at[u] = b * a.ls + pp(k);
            kk[u] = k;
            X[u] = a.spec + s_off[b];
            // mode 0: bins k and M - k; mode 1 (odd F): the Hermitian partner
            // of a bin beyond the stored half; mode 2: bin k
            i1[u] = (a.mode == 1 && k >= a.H) ? L - k : k;
            i2[u] = M - k;
        }
        C A[kU], B[kU], W[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) A[u] = X[u][i1[u]];
        if (a.mode == 0) {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                B[u] = X[u][i2[u]];
                W[u] = wpost[kk[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (base + u * T >= total) continue;
            const int k = kk[u];
            C z;
            if (a.mode == 0) {
                // Im X[0] and Im X[M] are ignored as a c2r transform does
                C Au = A[u], Bu = cconj(B[u]);
                if (k == 0) { Au.y = 0; Bu.y = 0; }
                const C t = cmul(csub(Au, Bu), cconj(W[u]));
                z = cadd(cadd(Au, Bu), cmul_i(t));
            } else if (a.mode == 1) {
                // odd F: Hermitian completion of the half spectrum
                z = k == 0 ? cmake<C>(A[u].x, R(0)) : (k < a.H ? A[u] : cconj(A[u]));
            } else {
                z = A[u];
            }
            b0[at[u]] = z;
        }
    }
    tab.end(tw, a.pl.tw, ntw, T);
    __syncthreads();
    const int cur = fft_lines<R, false>(b0, b1, a.pl, tw, nl, a.ls, 1, true, T);
    const C* Z = cur ? b1 : b0;
    const int nout_last = int(a.lastsel.count);
    const float rn = 1.0f / float(nout_last);
    for (int it = threadIdx.x; it < nl * nout_last; it += T) {
        const int b = fdiv(it, nout_last, rn), i = it - b * nout_last;
        const int64_t p = a.lastsel.start + int64_t(i) * a.lastsel.step;
        a.y[s_yoff[b] + i] = store_value<R, E>(a.mode, Z + b * a.ls, p, a.scale);
    }
}

// ---------------------------------------------------------------------------
// Rank 1: kernel and data transforms, multiply, inverse and crop in one block
// (the kernel spectrum is transformed in the block, or read from a previous
// pass when three line buffers do not fit LDS)
// ---------------------------------------------------------------------------

template <typename R, typename E>
struct Fused1dArgs {
    PadSrc<E> pad;
    KerSrc<E> ker;
    const Cx<R>* ks;   // kernel spectrum bins computed by a previous pass, or null
    E* y;
    int32_t F, H, mode;
    int32_t joint;     // no ks: kernel and data lines transformed as one batch (four line buffers)
    AxisSel lastsel;
    R scale;
    LinePlan<R> pl;
};

template <typename R, typename E>
__global__ void __launch_bounds__(kMaxThreads) fused1d_kernel(Fused1dArgs<R, E> a) {
    using C = Cx<R>;
    extern __shared__ unsigned char smem_raw[];
    const int L = a.pl.L;
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* const buf0 = tw + a.pl.ntw;
    C* const buf1 = buf0 + pp_len(L);
    C* const buf2 = buf1 + pp_len(L);
    const int T = blockDim.x;
    const RowRes<E> rr{0, -1, zero_of<E>(), false, false};
    auto load = [&](C* dst, bool kernel, int es) {
        for (int base = threadIdx.x; base < L; base += kU * T) {
            C z[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int n = base + u * T;
                z[u] = cmake<C>(R(0), R(0));
                if (n < L) {
                    if (kernel) z[u] = row_sample<R, E>(a.mode, n, [&](int p) { return ker_value(a.ker, 0, 0, p); });
                    else z[u] = row_sample<R, E>(a.mode, n, [&](int p) { return pad_value(a.pad, rr, 0, p); });
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int n = base + u * T;
                if (n < L) dst[pp(n) * es] = z[u];
            }
        }
    };
    stage_table(tw, a.pl.tw, a.pl.ntw, T);
    C* K = nullptr;   // kernel spectrum (LDS), element stride des
    C* D;             // data spectrum, element stride des
    C* W;             // free line for the product
    C* other;         // W's partner in the inverse
    int des = 1;
    if (!a.ks && a.joint) {
        // kernel (line 0) and data (line 1) interleaved in one buffer and
        // transformed as one batch: one set of stage barriers and twice the
        // butterflies per stage
        C* j0 = buf0;
        C* j1 = buf0 + 2 * pp_len(L);
        load(j0, true, 2);
        load(j0 + 1, false, 2);
        __syncthreads();
        const int cj = fft_lines<R, true>(j0, j1, a.pl, tw, 2, 1, 2, false, T);
        C* jr = cj ? j1 : j0;
        C* jf = cj ? j0 : j1;
        K = jr;
        D = jr + 1;
        des = 2;
        W = jf;
        other = jf + pp_len(L);
    } else {
        C* f0 = buf0;
        C* f1 = buf1;
        if (!a.ks) {
            load(buf2, true, 1);
            load(buf0, false, 1);
            __syncthreads();
            const int ck = fft_lines<R, false>(buf2, buf1, a.pl, tw, 1, pp_len(L), 1, false, T);
            K = ck ? buf1 : buf2;
            f1 = ck ? buf2 : buf1;
        } else {
            load(buf0, false, 1);
            __syncthreads();
        }
        const int cd = fft_lines<R, false>(f0, f1, a.pl, tw, 1, pp_len(L), 1, false, T);
        D = cd ? f1 : f0;
        W = cd ? f0 : f1;   // free buffer
        other = D;          // free once the product is formed
    }
    const C* wpost = tw + a.pl.post;
    // With a stored kernel spectrum, a thread's bins of it come in rounds of
    // kU, every load of a round issued before the round's first product (a
    // load per bin inside the loop would wait for each in turn).
    auto ks_rounds = [&](int n, auto&& get, auto&& use) {
        for (int base = threadIdx.x; base < n; base += kU * T) {
            C v0[kU], v1[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int k = base + u * T;
                v0[u] = v1[u] = cmake<C>(R(0), R(0));
                if (k < n) get(k, v0[u], v1[u]);
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (base + u * T < n) use(base + u * T, v0[u], v1[u]);
        }
    };
    if (a.mode == 0) {
        // bins k and M-k together: forward post-process both spectra,
        // multiply, then the c2r pre-process of the product
        const int M = L;
        auto product = [&](int k, C Kk, C Km) {
            const int km = M - k;
            C Pk = cmul(r2c_bin(D, M, k, wpost, des), Kk);
            C Pm = cmul(r2c_bin(D, M, km, wpost, des), Km);
            if (k == 0) { Pk.y = 0; Pm.y = 0; }   // Im of DC and Nyquist ignored by c2r
            W[pp(k == M ? 0 : k)] = c2r_pre(Pk, cconj(Pm), k, wpost);
            if (km != k && km < M) W[pp(km)] = c2r_pre(Pm, cconj(Pk), km, wpost);
        };
        if (a.ks) {
            ks_rounds(M / 2 + 1, [&](int k, C& k0, C& k1) { k0 = a.ks[k]; k1 = a.ks[M - k]; }, product);
        } else {
            for (int k = threadIdx.x; k <= M / 2; k += T)
                product(k, r2c_bin(K, M, k, wpost, des), r2c_bin(K, M, M - k, wpost, des));
        }
    } else if (a.mode == 1) {
        // odd F: the product of two Hermitian spectra, used whole
        if (a.ks) {
            // the stored kernel spectrum holds bins 0..F/2; the rest by symmetry
            ks_rounds(L, [&](int k, C& k0, C&) { k0 = k < a.H ? a.ks[k] : cconj(a.ks[L - k]); },
                      [&](int k, C k0, C) { W[pp(k)] = cmul(D[pp(k)], k0); });
        } else {
            for (int k = threadIdx.x; k < L; k += T) W[pp(k)] = cmul(D[pp(k) * des], K[pp(k) * des]);
        }
        if (threadIdx.x == 0) W[0].y = 0;
    } else if (a.ks) {
        // separate loops: a select between the global and the LDS pointer
        // would turn both into flat accesses
        ks_rounds(L, [&](int k, C& k0, C&) { k0 = a.ks[k]; }, [&](int k, C k0, C) { W[pp(k)] = cmul(D[pp(k)], k0); });
    } else {
        for (int k = threadIdx.x; k < L; k += T) W[pp(k)] = cmul(D[pp(k) * des], K[pp(k) * des]);
    }
    __syncthreads();
    const int ci = fft_lines<R, false>(W, other, a.pl, tw, 1, pp_len(L), 1, true, T);
    const C* Z = ci ? other : W;
    const int nout = int(a.lastsel.count);
    for (int i = threadIdx.x; i < nout; i += T) {
        const int64_t p = a.lastsel.start + int64_t(i) * a.lastsel.step;
        a.y[i] = store_value<R, E>(a.mode, Z, p, a.scale);
    }
}

// ---------------------------------------------------------------------------
// Lines longer than LDS: four-step decomposition of the last axis.
// A line of L = L1*L2 points is viewed as [n1][n2] (n = n1*L2 + n2):
//   forward  FFT along n1 (stride L2), times W_L^(k1*n2), then FFT along n2,
//            leaving X[k1 + L1*k2] at [k1][k2];
//   inverse  FFT along k2, times conj W_L^(k1*n2), FFT along k1 -> natural
//            order.
// The spectrum stays in [k1][k2] order, which the elementwise product and
// the other axes' passes do not care about.  Real input is transformed as
// complex (all L bins kept).
// ---------------------------------------------------------------------------

template <typename R, typename E>
struct BigColArgs {
    PadSrc<E> pad;
    KerSrc<E> ker;
    Cx<R>* spec[2];
    E* y;
    int64_t sstr[kMaxRank];       // spectrum strides of the outer axes
    AxisSel sel[2][kMaxRank];     // outer line selection per job
    int64_t ystr[kMaxRank];       // output strides of the outer axes (inverse)
    int64_t nlines[2];
    int32_t nout;
    int32_t L1, L2, Ct, ntiles, nblocks0;
    int32_t nblocks;              // grid size
    int32_t inv;                  // 0: forward from x / k (two jobs); 1: inverse to y (crop)
    AxisSel lastsel;              // crop along the long axis (inverse)
    R scale;
    const Cx<R>* twL;             // W_L^i, i < L
    LinePlan<R> pl;               // length-L1 plan
};

template <typename R, typename E>
__global__ void __launch_bounds__(kMaxThreads) bigcol_kernel(BigColArgs<R, E> a) {
    using C = Cx<R>;
    extern __shared__ unsigned char smem_raw[];
    const int L1 = a.pl.L, L2 = a.L2, Ct = a.Ct;
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* b0 = tw + a.pl.ntw;
    C* b1 = b0 + pp_len(L1) * Ct;
    const int T = blockDim.x;
    const int64_t g = xcd_slot(blockIdx.x, a.nblocks);   // neighbouring tiles on one XCD
    const int job = (!a.inv && g >= a.nblocks0) ? 1 : 0;
    const int64_t blk = job == 0 ? g : g - a.nblocks0;
    const int tile = int(blk % a.ntiles);
    int64_t line = blk / a.ntiles;
    // outer coordinates of this line
    int64_t soff = 0, xoff = 0, yoff = 0;
    int oc = -1;
    E ocv = zero_of<E>();
    bool zero = false, unw = false;
    {
        int64_t rem = line;
        for (int e = a.nout - 1; e >= 0; --e) {
            const AxisSel& s = a.sel[job][e];
            const int64_t i = rem % s.count;
            rem /= s.count;
            const int64_t q = s.start + i * s.step;
            soff += q * a.sstr[e];
            yoff += i * a.ystr[e];
            if (a.inv) continue;
            if (job == 0) {
                const int32_t c = axis_code(a.pad, e, int(q));
                if (c >= 0) xoff += int64_t(c) * a.pad.xstr[e];
                else if (c == kOutside) zero = true;
                else if (c == kUnwritten) unw = true;
                else if (oc < e) { oc = e; ocv = a.pad.cst[e][c == kConstFront ? 0 : 1]; }
            } else {
                if (q % a.ker.d[e] != 0) zero = true;
                const int64_t ii = q / a.ker.d[e];
                if (ii >= a.ker.K[e]) zero = true;
                xoff += (a.ker.rev ? ii : a.ker.K[e] - 1 - ii) * a.ker.kstr[e];
            }
        }
    }
    const int c0 = tile * Ct;
    const int nc = min(Ct, L2 - c0);
    const int ax = a.nout;
    const float rC = 1.0f / float(Ct);
    stage_table(tw, a.pl.tw, a.pl.ntw, T);
    const RowRes<E> rr{xoff, oc, ocv, unw, zero};
    const int total = L1 * Ct;
    for (int base = threadIdx.x; base < total; base += kU * T) {
        C z[kU];
        int at[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = base + u * T;
            z[u] = cmake<C>(R(0), R(0));
            at[u] = 0;
            if (it < total) {
                const int n1 = fdiv(it, Ct, rC), c = it - n1 * Ct;
                at[u] = pp(n1) * Ct + c;
                if (c < nc) {
                    const int n2 = c0 + c;
                    if (!a.inv) {
                        const int64_t n = int64_t(n1) * L2 + n2;
                        if (job == 0) z[u] = to_cx<R>(pad_value(a.pad, rr, ax, n));
                        else if (!zero) z[u] = to_cx<R>(ker_value(a.ker, xoff, ax, n));
                    } else {
                        z[u] = cmulc(a.spec[0][soff + int64_t(n1) * L2 + n2], a.twL[int64_t(n1) * n2], true);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = base + u * T;
            if (it < total) b0[at[u]] = z[u];
        }
    }
    __syncthreads();
    const int cur = fft_lines<R, true>(b0, b1, a.pl, tw, Ct, 1, Ct, a.inv != 0, T);
    const C* Z = cur ? b1 : b0;
    if (!a.inv) {
        C* out = a.spec[job];
        for (int it = threadIdx.x; it < total; it += T) {
            const int k1 = fdiv(it, Ct, rC), c = it - k1 * Ct;
            if (c >= nc) continue;
            const int n2 = c0 + c;
            out[soff + int64_t(k1) * L2 + n2] = cmul(Z[pp(k1) * Ct + c], a.twL[int64_t(k1) * n2]);
        }
    } else {
        for (int it = threadIdx.x; it < total; it += T) {
            const int n1 = fdiv(it, Ct, rC), c = it - n1 * Ct;
            if (c >= nc) continue;
            const int64_t n = int64_t(n1) * L2 + c0 + c;
            const int64_t rel = n - a.lastsel.start;
            if (rel < 0 || rel % a.lastsel.step) continue;
            const int64_t i = rel / a.lastsel.step;
            if (i >= a.lastsel.count) continue;
            E v;
            if constexpr (std::is_same<E, C>::value) v = cscale(Z[pp(n1) * Ct + c], a.scale);
            else v = E(Z[pp(n1) * Ct + c].x * a.scale);
            a.y[yoff + i] = v;
        }
    }
}

template <typename R>
struct BigRowArgs {
    Cx<R>* spec[2];
    int64_t sstr[kMaxRank];
    AxisSel sel[2][kMaxRank];
    int64_t nrows[2];             // outer lines * L1 per job
    int32_t nout, L1, L2, nblocks0;
    int32_t mode;                 // 0 forward (two jobs), 1 forward x kernel -> inverse (job 0), 2 inverse
    LinePlan<R> pl;               // length-L2 plan
};

template <typename R>
__global__ void __launch_bounds__(kMaxThreads) bigrow_kernel(BigRowArgs<R> a) {
    using C = Cx<R>;
    extern __shared__ unsigned char smem_raw[];
    const int L2 = a.L2;
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* b0 = tw + a.pl.ntw;
    C* b1 = b0 + pp_len(L2);
    C* b2 = b1 + pp_len(L2);
    const int T = blockDim.x;
    const int job = (a.mode == 0 && blockIdx.x >= unsigned(a.nblocks0)) ? 1 : 0;
    const int64_t row = job == 0 ? blockIdx.x : blockIdx.x - a.nblocks0;
    const int64_t k1 = row % a.L1;
    int64_t rem = row / a.L1;
    int64_t soff = k1 * L2;
    for (int e = a.nout - 1; e >= 0; --e) {
        const AxisSel& s = a.sel[job][e];
        const int64_t i = rem % s.count;
        rem /= s.count;
        soff += (s.start + i * s.step) * a.sstr[e];
    }
    stage_table(tw, a.pl.tw, a.pl.ntw, T);
    C* res;
    if (a.mode == 1) {
        stage_line(b1, a.spec[1] + soff, L2, T);
        stage_line(b0, a.spec[0] + soff, L2, T);
        __syncthreads();
        const int ck = fft_lines<R, false>(b1, b2, a.pl, tw, 1, pp_len(L2), 1, false, T);
        C* K = ck ? b2 : b1;
        C* f1 = ck ? b1 : b2;
        const int cd = fft_lines<R, false>(b0, f1, a.pl, tw, 1, pp_len(L2), 1, false, T);
        C* D = cd ? f1 : b0;
        C* Dfree = cd ? b0 : f1;
        for (int n = threadIdx.x; n < pp_len(L2); n += T) D[n] = cmul(D[n], K[n]);
        __syncthreads();
        res = fft_lines<R, false>(D, Dfree, a.pl, tw, 1, pp_len(L2), 1, true, T) ? Dfree : D;
    } else {
        stage_line(b0, a.spec[job] + soff, L2, T);
        __syncthreads();
        res = fft_lines<R, false>(b0, b1, a.pl, tw, 1, pp_len(L2), 1, a.mode == 2, T) ? b1 : b0;
    }
    for (int n = threadIdx.x; n < L2; n += T) a.spec[job][soff + n] = res[pp(n)];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Radices of a line plan.  The power-of-two part takes the fewest stages of
// radix at most 16, with the bits spread evenly over them (2^8 = 16*16,
// 2^9 = 8*8*8, 2^10 = 16*8*8): fewer stages mean fewer LDS round trips and
// barriers, and radix 8-16 butterflies cost the fewest instructions per point.
// Double-precision plans stop at radix 8: a radix-16 butterfly of complex
// doubles does not fit the registers of a 1024-thread block.
std::vector<int> factor_radices(int L, bool dbl) {
    std::vector<int> r;
    int n = L, a = 0;
    while (n % 2 == 0) { n /= 2; ++a; }
    if (a > 0) {
        const int maxb = dbl ? 3 : 4;
        const int st = (a + maxb - 1) / maxb;
        for (int i = 0; i < st; ++i) {
            const int bits = a / st + (i < a % st ? 1 : 0);
            r.push_back(1 << bits);
        }
    }
    for (int p : {3, 5, 7}) while (n % p == 0) { r.push_back(p); n /= p; }
    for (int p = 11; n > 1; p += 2) {
        while (n % p == 0) { r.push_back(p); n /= p; }
        if (int64_t(p) * p > n && n > 1) { r.push_back(n); n = 1; }
    }
    return r;
}

bool fixed_radix(int p) { return p == 2 || p == 3 || p == 4 || p == 5 || p == 7 || p == 8 || p == 16; }

// Entries of the compact twiddle table make_lineplan builds for (L, post_F).
int64_t ntw_of(int L, int post_F, bool dbl) {
    int64_t n = 0, Ns = 1;
    for (int P : factor_radices(L, dbl)) {
        n += Ns * (P - 1);
        if (!fixed_radix(P)) n += P;
        Ns *= P;
    }
    if (post_F > 0) n += L + 1;
    return n;
}

// Builds the compact FP64 twiddle table of a line plan and caches it on the
// device.  post_F > 0 appends W_F^k, k = 0..L (real-even rows, F = 2L).
template <typename R>
bool make_lineplan(LinePlan<R>* pl, int L, int post_F, LaunchContext& ctx) {
    const std::string key = "lineplan:" + std::to_string(L) + ":" + std::to_string(post_F) + ":" +
                            std::to_string(sizeof(R));
    std::string& m = ctx.memo(key);
    if (m.size() == sizeof(*pl)) {
        std::memcpy(pl, m.data(), sizeof(*pl));
        return true;
    }
    std::memset(pl, 0, sizeof(*pl));
    pl->L = L;
    auto r = factor_radices(L, sizeof(R) == 8);
    if (int(r.size()) > kMaxStages) return false;
    pl->nst = int(r.size());
    std::vector<double> tab;   // interleaved re, im
    auto push = [&](double ang) {
        tab.push_back(std::cos(ang));
        tab.push_back(std::sin(ang));
    };
    int Ns = 1;
    for (size_t s = 0; s < r.size(); ++s) {
        const int P = r[s];
        pl->rad[s] = P;
        pl->base[s] = int(tab.size() / 2);
        for (int q = 1; q < P; ++q)
            for (int k = 0; k < Ns; ++k) push(-2.0 * M_PI * double(int64_t(q) * k) / double(int64_t(Ns) * P));
        pl->wp[s] = int(tab.size() / 2);
        if (!fixed_radix(P))
            for (int m = 0; m < P; ++m) push(-2.0 * M_PI * double(m) / double(P));
        Ns *= P;
    }
    pl->post = -1;
    if (post_F > 0) {
        pl->post = int(tab.size() / 2);
        for (int k = 0; k <= L; ++k) push(-2.0 * M_PI * double(k) / double(post_F));
    }
    pl->ntw = int(tab.size() / 2);
    std::vector<R> dev(tab.begin(), tab.end());
    pl->tw = static_cast<const Cx<R>*>(ctx.constant_table(key, dev.data(), dev.size() * sizeof(R)));
    if (!pl->tw) return false;
    m.assign(reinterpret_cast<const char*>(pl), sizeof(*pl));
    return true;
}

template <typename E>
E host_value(const ndc_value& v);
template <> float host_value<float>(const ndc_value& v) { return float(v.re); }
template <> double host_value<double>(const ndc_value& v) { return v.re; }
template <> float2 host_value<float2>(const ndc_value& v) { return cmake<float2>(float(v.re), float(v.im)); }
template <> double2 host_value<double2>(const ndc_value& v) { return cmake<double2>(v.re, v.im); }

ndc_status check_launch(LaunchContext& ctx, const char* what) {
    if (ctx.dry_run()) return NDC_OK;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx.err = std::string(what) + ": " + hipGetErrorString(e);
        return NDC_ERR_DEVICE;
    }
    return NDC_OK;
}

// Threads per block: about two complex points per thread, 64..1024.
int threads_for(int64_t points) {
    int64_t t = (points / 2 + 63) / 64 * 64;
    return int(std::max<int64_t>(64, std::min<int64_t>(kMaxThreads, t)));
}

template <typename R, typename E>
ndc_status run_fft(const Plan& p, const void* xv, const void* kv, void* yv, LaunchContext& ctx, KernelSpectrum* ks) {
    using C = Cx<R>;
    constexpr bool kComplex = std::is_same<E, C>::value;
    const int N = p.N;
    const int last = N - 1;
    const int Fl = int(p.fft[last]);
    const size_t es = sizeof(C);
    const size_t lds = ctx.max_lds();
    int mode;             // 0 real even, 1 real odd, 2 complex
    int L;
    if (kComplex) { mode = 2; L = Fl; }
    else if (Fl % 2 == 0) { mode = 0; L = Fl / 2; }
    else { mode = 1; L = Fl; }
    // LDS of a row plan: the compact table (about L entries, 2L for real-even)
    // plus two line buffers.  Lines that do not fit take the four-step path,
    // as complex transforms of the full length (all Fl bins kept).
    constexpr bool kDbl = sizeof(R) == 8;
    const size_t row_tw = size_t(ntw_of(L, mode == 0 ? Fl : 0, kDbl));
    const bool big = (row_tw + size_t(2) * pp_len(L)) * es > lds;
    const int H = big ? Fl : (kComplex ? Fl : Fl / 2 + 1);
    int L1 = 0, L2 = 0, Ctb = 0;
    if (big) {
        for (int d2 = Fl; d2 >= 2; --d2) {
            if (Fl % d2 || (size_t(ntw_of(d2, 0, kDbl)) + size_t(3) * pp_len(d2)) * es > lds) continue;
            const int d1 = Fl / d2;
            const size_t t1 = size_t(ntw_of(d1, 0, kDbl));
            int ct = 16;
            while (ct > 1 && (t1 + size_t(2) * pp_len(d1) * ct) * es > lds) ct /= 2;
            if ((t1 + size_t(2) * pp_len(d1) * ct) * es > lds) continue;
            L1 = d1;
            L2 = d2;
            Ctb = ct;
            break;
        }
        if (!L2) {
            ctx.err = "FFT length of the last axis cannot be split into LDS-resident factors";
            return NDC_ERR_UNSUPPORTED;
        }
    }
    LinePlan<R> rowpl;
    if (!big && !make_lineplan<R>(&rowpl, L, mode == 0 ? Fl : 0, ctx)) {
        ctx.err = "twiddle table allocation failed";
        return NDC_ERR_OUT_OF_MEMORY;
    }

    PadSrc<E> pad;
    std::memset(&pad, 0, sizeof(pad));
    pad.x = static_cast<const E*>(xv);
    KerSrc<E> ker;
    std::memset(&ker, 0, sizeof(ker));
    ker.k = static_cast<const E*>(kv);
    ker.rev = p.reverse ? 1 : 0;
    for (int i = 0; i < N; ++i) {
        pad.pds[i] = int32_t(p.pds[i]);
        pad.side[i][0] = side1d(p.side[i][0].tag);
        pad.side[i][1] = side1d(p.side[i][1].tag);
        pad.cst[i][0] = host_value<E>(p.side[i][0].value);
        pad.cst[i][1] = host_value<E>(p.side[i][1].value);
        pad.xstr[i] = p.x_str[i];
        pad.lo[i] = int32_t(p.pad[i][0]);
        pad.n[i] = int32_t(p.x_shape[i]);
        ker.kstr[i] = p.k_str[i];
        ker.K[i] = int32_t(p.k_shape[i]);
        ker.d[i] = int32_t(p.dil[i]);
    }
    double prodF = 1;
    for (int i = 0; i < N; ++i) prodF *= double(p.fft[i]);
    const R scale = R(1.0 / prodF);
    E* y = static_cast<E*>(yv);
    const double esz = double(sizeof(E));

    if (ks && big) {
        ctx.err = "a kept kernel spectrum needs LDS-resident rows";
        return NDC_ERR_UNSUPPORTED;
    }
    // rank 1, line fits three buffers: one fused launch (four: the kernel and
    // data lines transformed as one batch)
    if (!ks && N == 1 && !big && (size_t(rowpl.ntw) + size_t(3) * pp_len(L)) * es <= lds) {
        Fused1dArgs<R, E> a;
        std::memset(&a, 0, sizeof(a));
        a.joint = (size_t(rowpl.ntw) + size_t(4) * pp_len(L)) * es <= lds ? 1 : 0;
        a.pad = pad;
        a.ker = ker;
        a.y = y;
        a.F = Fl;
        a.H = H;
        a.mode = mode;
        a.lastsel = AxisSel{p.kdil[0] - 1, p.stride[0], p.out[0]};
        a.scale = scale;
        a.pl = rowpl;
        const size_t smem = (size_t(rowpl.ntw) + size_t(a.joint ? 4 : 3) * pp_len(L)) * es;
        ctx.kbegin("fft_fused_1d", esz * double(p.x_numel + p.k_numel + p.y_numel), 15.0 * L * std::log2(double(L)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&fused1d_kernel<R, E>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (fused1d_kernel<R, E>), dim3(1), dim3(threads_for(L)), smem, a);
        ctx.kend();
        return check_launch(ctx, "fused rank-1 FFT convolution");
    }

    // spectrum layout [F0, ..., F_{N-2}, Hp]
    const int64_t Hp = big ? H : (H + 7) / 8 * 8;
    int64_t sstr[kMaxRank];
    sstr[last] = 1;
    if (N >= 2) sstr[last - 1] = Hp;
    for (int e = last - 2; e >= 0; --e) sstr[e] = sstr[e + 1] * p.fft[e + 1];
    const int64_t spec_elems = N >= 2 ? sstr[0] * p.fft[0] : Hp;
    C* Sx = static_cast<C*>(ctx.scratch(0, size_t(spec_elems) * es));
    C* Sk = ks ? static_cast<C*>(ks->spec) : static_cast<C*>(ctx.scratch(1, size_t(spec_elems) * es));
    if (!Sx || !Sk) {
        ctx.err = "spectrum workspace allocation failed";
        return NDC_ERR_OUT_OF_MEMORY;
    }
    const AxisSel none{0, 1, 1};

    // LDS of the row passes: the stage twiddles and two line buffers per line
    const size_t row_tw_lds = big ? 0 : size_t(row_stage_tw(rowpl));
    auto rows_lpb = [&]() {
        int lpb = int(std::max<int64_t>(1, std::min<int64_t>(kMaxLpb, 512 / std::max(1, L))));
        while (lpb > 1 && (row_tw_lds + size_t(2) * lpb * pp_len(L)) * es > lds) lpb /= 2;
        return lpb;
    };

    auto rows_fwd = [&](bool data_job, bool kernel_job) -> ndc_status {
        RowsFwdArgs<R, E> a;
        std::memset(&a, 0, sizeof(a));
        a.pad = pad;
        a.ker = ker;
        a.spec[0] = Sx;
        a.spec[1] = Sk;
        for (int e = 0; e < last; ++e) {
            a.sstr[e] = sstr[e];
            a.cnt[0][e] = p.pds[e];
            a.cnt[1][e] = p.kdil[e];
        }
        a.nlines[0] = data_job ? 1 : 0;
        a.nlines[1] = kernel_job ? 1 : 0;
        for (int e = 0; e < last; ++e) {
            a.nlines[0] *= a.cnt[0][e];
            a.nlines[1] *= a.cnt[1][e];
        }
        a.nout = last;
        a.F = Fl;
        a.H = H;
        a.mode = mode;
        a.ls = pp_len(L);
        const int lpb = rows_lpb();
        a.lpb = lpb;
        const size_t smem = (row_tw_lds + size_t(2) * lpb * pp_len(L)) * es;
        a.pl = rowpl;
        const int64_t nb0 = (a.nlines[0] + lpb - 1) / lpb;
        const int64_t nb1 = (a.nlines[1] + lpb - 1) / lpb;
        a.nblocks0 = int32_t(nb0);
        const double bytes = esz * double((data_job ? p.x_numel : 0) + (kernel_job ? p.k_numel : 0)) +
                             double(es) * double(a.nlines[0] + a.nlines[1]) * H;
        ctx.kbegin("fft_rows_fwd", bytes, 5.0 * double(a.nlines[0] + a.nlines[1]) * L * std::log2(double(L)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&rows_fwd_kernel<R, E>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (rows_fwd_kernel<R, E>),
            dim3(unsigned(nb0 + nb1)), dim3(threads_for(int64_t(lpb) * L)), smem, a);
        ctx.kend();
        return check_launch(ctx, "rows forward");
    };

    // rank 1, line fits two buffers: kernel spectrum first, then the fused data pass
    if (N == 1 && !big) {
        if (!ks || ks->bake) {
            ndc_status s = rows_fwd(false, true);
            if (s != NDC_OK || ks) return s;
        }
        Fused1dArgs<R, E> a;
        std::memset(&a, 0, sizeof(a));
        a.pad = pad;
        a.ker = ker;
        a.ks = Sk;
        a.y = y;
        a.F = Fl;
        a.H = H;
        a.mode = mode;
        a.lastsel = AxisSel{p.kdil[0] - 1, p.stride[0], p.out[0]};
        a.scale = scale;
        a.pl = rowpl;
        const size_t smem = (size_t(rowpl.ntw) + size_t(2) * pp_len(L)) * es;
        ctx.kbegin("fft_fused_1d", esz * double(p.x_numel + p.y_numel) + double(es) * H, 10.0 * L * std::log2(double(L)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&fused1d_kernel<R, E>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (fused1d_kernel<R, E>), dim3(1), dim3(threads_for(L)), smem, a);
        ctx.kend();
        return check_launch(ctx, "fused rank-1 FFT convolution");
    }

    // four-step passes of the long last axis
    LinePlan<R> pl1, pl2;
    const Cx<R>* twL = nullptr;
    if (big) {
        if (!make_lineplan<R>(&pl1, L1, 0, ctx) || !make_lineplan<R>(&pl2, L2, 0, ctx) ||
            !(twL = static_cast<const Cx<R>*>(ctx.twiddles(Fl, sizeof(R) == 8)))) {
            ctx.err = "twiddle table allocation failed";
            return NDC_ERR_OUT_OF_MEMORY;
        }
    }
    auto big_col = [&](bool inv) -> ndc_status {
        BigColArgs<R, E> a;
        std::memset(&a, 0, sizeof(a));
        a.pad = pad;
        a.ker = ker;
        a.spec[0] = Sx;
        a.spec[1] = Sk;
        a.y = y;
        a.nout = last;
        int64_t ys = p.out[last];
        for (int e = last - 1; e >= 0; --e) {
            a.sstr[e] = sstr[e];
            a.ystr[e] = ys;
            ys *= p.out[e];
            if (inv) a.sel[0][e] = AxisSel{p.kdil[e] - 1, p.stride[e], p.out[e]};
            else {
                a.sel[0][e] = AxisSel{0, 1, p.pds[e]};
                a.sel[1][e] = AxisSel{0, 1, p.kdil[e]};
            }
        }
        for (int e = last; e < kMaxRank; ++e) a.sel[0][e] = a.sel[1][e] = none;
        a.nlines[0] = 1;
        a.nlines[1] = inv ? 0 : 1;
        for (int e = 0; e < last; ++e) {
            a.nlines[0] *= a.sel[0][e].count;
            if (!inv) a.nlines[1] *= a.sel[1][e].count;
        }
        a.L1 = L1;
        a.L2 = L2;
        a.Ct = Ctb;
        a.ntiles = (L2 + Ctb - 1) / Ctb;
        a.inv = inv ? 1 : 0;
        a.lastsel = AxisSel{p.kdil[last] - 1, p.stride[last], p.out[last]};
        a.scale = scale;
        a.twL = twL;
        a.pl = pl1;
        const int64_t nb0 = a.nlines[0] * a.ntiles, nb1 = a.nlines[1] * a.ntiles;
        a.nblocks0 = int32_t(nb0);
        const size_t smem = (size_t(pl1.ntw) + size_t(2) * pp_len(L1) * Ctb) * es;
        const double lines = double(a.nlines[0] + a.nlines[1]);
        const double bytes = inv ? double(es) * lines * Fl + esz * double(p.y_numel)
                                 : esz * double(p.x_numel + p.k_numel) + double(es) * lines * Fl;
        a.nblocks = int32_t(nb0 + nb1);
        ctx.kbegin(inv ? "fft_bigcol_inv" : "fft_bigcol_fwd", bytes, 5.0 * lines * Fl * std::log2(double(L1)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&bigcol_kernel<R, E>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (bigcol_kernel<R, E>),
            dim3(unsigned(nb0 + nb1)), dim3(threads_for(int64_t(L1) * Ctb)), smem, a);
        ctx.kend();
        return check_launch(ctx, "long-axis column pass");
    };
    auto big_row = [&](int rmode) -> ndc_status {
        BigRowArgs<R> a;
        std::memset(&a, 0, sizeof(a));
        a.spec[0] = Sx;
        a.spec[1] = Sk;
        a.nout = last;
        a.L1 = L1;
        a.L2 = L2;
        a.mode = rmode;
        a.pl = pl2;
        for (int e = 0; e < last; ++e) {
            a.sstr[e] = sstr[e];
            if (rmode == 2) a.sel[0][e] = AxisSel{p.kdil[e] - 1, p.stride[e], p.out[e]};
            else {
                a.sel[0][e] = AxisSel{0, 1, rmode == 1 ? 1 : p.pds[e]};
                a.sel[1][e] = AxisSel{0, 1, p.kdil[e]};
            }
        }
        a.nrows[0] = L1;
        a.nrows[1] = rmode == 0 ? L1 : 0;
        for (int e = 0; e < last; ++e) {
            a.nrows[0] *= a.sel[0][e].count;
            if (rmode == 0) a.nrows[1] *= a.sel[1][e].count;
        }
        a.nblocks0 = int32_t(a.nrows[0]);
        const size_t smem = (size_t(pl2.ntw) + size_t(rmode == 1 ? 3 : 2) * pp_len(L2)) * es;
        const double rows = double(a.nrows[0] + a.nrows[1]);
        const double bytes = double(es) * L2 * (rmode == 1 ? 3.0 * a.nrows[0] : 2.0 * rows);
        static const char* names[3] = {"fft_bigrow_fwd", "fft_bigrow_fused", "fft_bigrow_inv"};
        ctx.kbegin(names[rmode], bytes, 5.0 * rows * (rmode == 1 ? 3 : 1) * L2 * std::log2(double(L2)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&bigrow_kernel<R>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (bigrow_kernel<R>),
            dim3(unsigned(a.nrows[0] + a.nrows[1])), dim3(threads_for(L2)), smem, a);
        ctx.kend();
        return check_launch(ctx, "long-axis row pass");
    };

    if (N == 1) {   // big rank-1 line
        ndc_status s = big_col(false);
        if (s == NDC_OK) s = big_row(1);
        if (s == NDC_OK) s = big_col(true);
        return s;
    }

    // ---- pass 1: last axis forward
    bool data_job = !(ks && ks->bake), kernel_job = !(ks && !ks->bake);
    if (big) {
        ndc_status s = big_col(false);
        if (s == NDC_OK) s = big_row(0);
        if (s != NDC_OK) return s;
    } else {
        ndc_status s = rows_fwd(data_job, kernel_job);
        if (s != NDC_OK) return s;
    }

    // One column-pass launch: lines of Ld points at element stride sdl, one
    // per outer position of `outer` and bin tile (see ColsArgs).
    struct OuterAx {
        int64_t stride;
        AxisSel sel[2];
    };
    const Cx<R>* twF = nullptr;   // W_F^i of the axis being split, when one is
    auto cols_launch = [&](int Ld, int64_t sdl, const std::vector<OuterAx>& outer, int cmode, int valid0,
                           int valid1, AxisSel outsel, int split_L2, int tw_outer, int tw_dir) -> ndc_status {
        ColsArgs<R> a;
        std::memset(&a, 0, sizeof(a));
        a.spec[0] = Sx;
        a.spec[1] = Sk;
        a.sd = sdl;
        a.H = H;
        a.mode = cmode;
        if (!make_lineplan<R>(&a.pl, Ld, 0, ctx)) {
            ctx.err = "twiddle table allocation failed";
            return NDC_ERR_OUT_OF_MEMORY;
        }
        a.nouter = int(outer.size());
        a.nlines_outer[0] = a.nlines_outer[1] = 1;
        for (int i = 0; i < a.nouter; ++i) {
            a.ostr[i] = outer[size_t(i)].stride;
            for (int j = 0; j < 2; ++j) {
                a.outer[j][i] = outer[size_t(i)].sel[j];
                a.nlines_outer[j] *= outer[size_t(i)].sel[j].count;
            }
        }
        a.valid[0] = valid0;
        a.valid[1] = valid1;
        a.outsel = outsel;
        a.split_L2 = split_L2;
        a.tw_outer = tw_outer;
        a.tw_dir = tw_dir;
        a.twF = twF;
        if (cmode == 0 && !data_job) a.nlines_outer[0] = 0;
        if (cmode == 0 && !kernel_job) a.nlines_outer[1] = 0;
        // fused: two joint kernel + data buffers (4 tiles) when they fit the
        // budget below at 8 columns, else kernel, data and a spare (3 tiles)
        auto smem_n = [&](int nb, int ct) { return (size_t(a.pl.ntw) + size_t(nb) * pp_len(Ld) * ct) * es; };
        a.joint = cmode == 1 && smem_n(4, 8) <= std::min<size_t>(lds, 96 * 1024) ? 1 : 0;
        const int nbuf = cmode == 1 ? (a.joint ? 4 : 3) : (cmode == 3 ? 3 : 2);
        // tile width: enough lines for coalescing, enough blocks to fill the chip
        int Ct = 16;
        auto smem_of = [&](int ct) { return smem_n(nbuf, ct); };
        while (Ct > 1 && smem_of(Ct) > std::min<size_t>(lds, 96 * 1024)) Ct /= 2;
        // 8 columns = 64-byte segments per row (f32): below that, partial-line
        // reads and writes cost more than the extra blocks gain
        while (Ct > 8 && (a.nlines_outer[0] + a.nlines_outer[1]) * ((H + Ct - 1) / Ct) < 2 * ctx.cu_count()) Ct /= 2;
        const size_t smem = smem_of(Ct);
        if (smem > lds) {
            ctx.err = "FFT length of a column axis exceeds the LDS-resident line size";
            return NDC_ERR_UNSUPPORTED;
        }
        a.C = Ct;
        a.ntiles = (H + Ct - 1) / Ct;
        const int64_t nb0 = a.nlines_outer[0] * a.ntiles;
        const int64_t nb1 = cmode == 0 ? a.nlines_outer[1] * a.ntiles : 0;
        a.nblocks0 = int32_t(nb0);
        a.nblocks = int32_t(nb0 + nb1);
        if (nb0 + nb1 == 0) return NDC_OK;
        // compulsory traffic: the rows read (valid / all) and the rows written
        double lines0 = double(a.nlines_outer[0]) * H, lines1 = double(a.nlines_outer[1]) * H;
        const double v0 = split_L2 > 0 ? double(valid0) / split_L2 : double(valid0);
        const double v1 = split_L2 > 0 ? double(valid1) / split_L2 : double(valid1);
        double bytes = 0, nfft = 0;
        if (cmode == 0) { bytes = (lines0 * (std::min<double>(v0, Ld) + Ld) + lines1 * (std::min<double>(v1, Ld) + Ld)); nfft = lines0 + lines1; }
        else if (cmode == 1) { bytes = lines0 * (v0 + v1 + double(a.outsel.count)); nfft = 3 * lines0; }
        else if (cmode == 3) { bytes = lines0 * (v0 + Ld + double(a.outsel.count)); nfft = 2 * lines0; }
        else { bytes = lines0 * (Ld + double(a.outsel.count)); nfft = lines0; }
        bytes *= double(es);
        static const char* names[4] = {"fft_cols_fwd", "fft_cols_fused", "fft_cols_inv", "fft_cols_fused_kept"};
        static const char* split_names[4] = {"fft_cols_split_fwd", "fft_cols_split_fused", "fft_cols_split_inv",
                                             "fft_cols_split_fused_kept"};
        const bool split = tw_outer >= 0;
        ctx.kbegin(split ? split_names[cmode] : names[cmode], bytes, 5.0 * nfft * Ld * std::log2(double(Ld)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&cols_kernel<R>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (cols_kernel<R>),
            dim3(unsigned(nb0 + nb1)), dim3(threads_for(int64_t(Ld) * Ct / 2)), smem, a);
        ctx.kend();
        return check_launch(ctx, "column pass");
    };

    // A column line of F points fits one workgroup when the fused pass's
    // three tile buffers of one column fit the LDS.  Narrow tiles of long
    // lines coalesce poorly, but one pass still beats the split's three: on a
    // 4096x32768 f32 convolution (axis-0 length 4608, one-column tiles) the
    // fused pass took 4.3 ms against 6.4 ms for the split passes.
    auto col_fits = [&](int64_t F) {
        return F <= (int64_t(1) << 24) && (size_t(ntw_of(int(F), 0, kDbl)) + size_t(3) * pp_len(int(F))) * es <= lds;
    };
    // F = L1 * L2 with both factors fitting, as close to sqrt(F) as possible
    // (0 when no such split exists)
    auto col_split = [&](int64_t F) -> int64_t {
        int64_t best = 0;
        for (int64_t l2 = 2; l2 * l2 <= F; ++l2) {
            if (F % l2) continue;
            const int64_t l1 = F / l2;
            if (col_fits(l1) && col_fits(l2)) best = l2;
        }
        return best;
    };

    // Axis d of the spectrum, one of the column modes 0..3 (see ColsArgs):
    // one launch when its lines fit a workgroup, otherwise the four-step
    // split F = L1 * L2 (point j1 L2 + j2, spectral index k1 + L1 k2 kept at
    // k1 L2 + k2, which the elementwise product does not care about): forward
    // L1-point passes along j1 with a W_F^(k1 j2) post-multiply, then
    // L2-point passes along j2; the inverse mirrors it and restores natural
    // order, every row written (later passes read the rows they select).
    auto col_pass = [&](int d, int cmode) -> ndc_status {
        const int64_t F = p.fft[d];
        std::vector<OuterAx> outer;
        for (int e = 0; e < last; ++e) {
            if (e == d) continue;
            OuterAx o;
            o.stride = sstr[e];
            for (int j = 0; j < 2; ++j) {
                AxisSel s{0, 1, p.fft[e]};
                if (cmode == 0 && e < d) s = AxisSel{0, 1, j == 0 ? p.pds[e] : p.kdil[e]};
                if (cmode == 2 && e < d) s = AxisSel{p.kdil[e] - 1, p.stride[e], p.out[e]};
                o.sel[j] = s;
            }
            outer.push_back(o);
        }
        const int valid0 = int(cmode == 2 ? F : p.pds[d]);
        const int valid1 = int(p.kdil[d]);
        const AxisSel crop{p.kdil[d] - 1, p.stride[d], p.out[d]};
        if (col_fits(F))
            return cols_launch(int(F), sstr[d], outer, cmode, valid0, valid1, cmode == 0 ? AxisSel{0, 1, F} : crop, 0,
                               -1, 0);
        const int64_t L2 = col_split(F);
        if (!L2) {
            ctx.err = "FFT length of a column axis cannot be split into LDS-resident factors";
            return NDC_ERR_UNSUPPORTED;
        }
        const int64_t L1 = F / L2;
        if (!(twF = static_cast<const Cx<R>*>(ctx.twiddles(F, kDbl)))) {
            ctx.err = "twiddle table allocation failed";
            return NDC_ERR_OUT_OF_MEMORY;
        }
        const int jax = int(outer.size());   // the extra outer index: j2 (A passes) or k1 (B passes)
        std::vector<OuterAx> outA = outer, outB = outer;
        outA.push_back(OuterAx{sstr[d], {AxisSel{0, 1, L2}, AxisSel{0, 1, L2}}});
        outB.push_back(OuterAx{L2 * sstr[d], {AxisSel{0, 1, L1}, AxisSel{0, 1, L1}}});
        const AxisSel allA{0, 1, L1}, allB{0, 1, L2};
        ndc_status s = NDC_OK;
        if (cmode == 2) {
            s = cols_launch(int(L2), sstr[d], outB, 2, int(L2), int(L2), allB, 0, -1, 0);
            if (s == NDC_OK) s = cols_launch(int(L1), L2 * sstr[d], outA, 2, int(L1), int(L1), allA, 0, jax, 2);
            return s;
        }
        // forward along j1 (the data and kernel rows holding values), twiddled
        const bool kjob = kernel_job && cmode != 3;
        const bool djob = data_job;
        {
            const bool sd0 = data_job, sk0 = kernel_job;
            data_job = djob;
            kernel_job = kjob;
            s = cols_launch(int(L1), L2 * sstr[d], outA, 0, valid0, valid1, allA, int(L2), jax, 1);
            data_job = sd0;
            kernel_job = sk0;
        }
        if (s != NDC_OK) return s;
        if (cmode == 0) return cols_launch(int(L2), sstr[d], outB, 0, int(L2), int(L2), allB, 0, -1, 0);
        // along j2: forward, product, inverse (cmode 1 or 3), every row kept
        s = cols_launch(int(L2), sstr[d], outB, cmode, int(L2), int(L2), allB, 0, -1, 0);
        if (s != NDC_OK) return s;
        // inverse along j1 with the conjugate twiddle, natural order again
        return cols_launch(int(L1), L2 * sstr[d], outA, 2, int(L1), int(L1), allA, 0, jax, 2);
    };

    // ---- pass 2: axes N-2 .. 1 forward
    for (int d = last - 1; d >= 1; --d) {
        ndc_status s = col_pass(d, 0);
        if (s != NDC_OK) return s;
    }
    // ---- pass 3: axis 0 fused (or, when baking, the kernel's axis-0 forward)
    if (ks && ks->bake) return col_pass(0, 0);
    {
        ndc_status s = col_pass(0, ks ? 3 : 1);
        if (s != NDC_OK) return s;
    }
    // ---- pass 4: axes 1 .. N-2 inverse
    for (int d = 1; d <= last - 1; ++d) {
        ndc_status s = col_pass(d, 2);
        if (s != NDC_OK) return s;
    }
    // ---- pass 5: last axis inverse with crop
    if (big) {
        ndc_status s = big_row(2);
        if (s == NDC_OK) s = big_col(true);
        return s;
    }
    {
        RowsInvArgs<R, E> a;
        std::memset(&a, 0, sizeof(a));
        a.spec = Sx;
        a.y = y;
        a.nout = last;
        int64_t ys = p.out[last];
        for (int e = last - 1; e >= 0; --e) {
            a.ystr[e] = ys;
            ys *= p.out[e];
        }
        a.nlines = 1;
        for (int e = 0; e < last; ++e) {
            a.sstr[e] = sstr[e];
            a.sel[e] = AxisSel{p.kdil[e] - 1, p.stride[e], p.out[e]};
            a.nlines *= p.out[e];
        }
        a.F = Fl;
        a.H = H;
        a.mode = mode;
        a.ls = pp_len(L);
        const int lpb = rows_lpb();
        a.lpb = lpb;
        a.lastsel = AxisSel{p.kdil[last] - 1, p.stride[last], p.out[last]};
        a.scale = scale;
        a.pl = rowpl;
        const size_t smem = (row_tw_lds + size_t(2) * lpb * pp_len(L)) * es;
        const int64_t nb = (a.nlines + lpb - 1) / lpb;
        const double bytes = double(es) * double(a.nlines) * H + esz * double(p.y_numel);
        ctx.kbegin("fft_rows_inv", bytes, 5.0 * double(a.nlines) * L * std::log2(double(L)));
        if (!lds_fits(ctx, reinterpret_cast<const void*>(&rows_inv_kernel<R, E>), smem)) return NDC_ERR_UNSUPPORTED;
        NDC_LAUNCH(ctx, ctx.stream(), (rows_inv_kernel<R, E>),
            dim3(unsigned(nb)), dim3(threads_for(int64_t(lpb) * L)), smem, a);
        ctx.kend();
        ndc_status s = check_launch(ctx, "rows inverse");
        if (s != NDC_OK) return s;
    }
    return NDC_OK;
}

// ---------------------------------------------------------------------------
// Plain N-d transforms of the public processor (processor/real.rs,
// processor/complex.rs): line transforms along one axis of a contiguous
// array viewed as [outer][n][inner], and a 2-D transpose for the axis
// rotation of the reference's spectral layout.
// ---------------------------------------------------------------------------

template <typename R>
struct LineFftArgs {
    const Cx<R>* in;     // complex input (c2c, c2r)
    const R* rin;        // real input (r2c)
    Cx<R>* out;          // complex output (c2c, r2c)
    R* rout;             // real output (c2r)
    int64_t outer, inner;
    int32_t n;           // elements along the axis: complex length (c2c) or real length F (r2c, c2r)
    int32_t H;           // F/2 + 1 (r2c, c2r)
    int32_t mode;        // 0 c2c, 1 r2c, 2 c2r
    int32_t Ct;          // lines per block: along inner (columns) or consecutive rows (inner == 1)
    int32_t inv;
    int32_t ntiles;      // column tiles per outer index
    int32_t nblocks;     // grid size
    R scale;
    // first pass of a long line split as n_full = n * L2 (columns only): inner
    // index I = j2 * tw_inner + i, and output k1 is multiplied by
    // W^(k1 j2) of twN (W_nfull^e, e < n_full; conjugated for the inverse)
    const Cx<R>* twN;
    int64_t tw_inner;
    // columns only: the input (in_lines) or output (out_lines) holds each
    // line contiguously, element (o, j, i) at ((o inner + i) n + j): the
    // processor's axis rotation fused into its first inverse or last forward
    // transform instead of a transpose pass
    int32_t in_lines, out_lines;
    LinePlan<R> pl;      // complex length transformed: F/2 for even real lines, else n
};

template <typename R>
__global__ void __launch_bounds__(kMaxThreads) line_fft_kernel(LineFftArgs<R> a) {
    using C = Cx<R>;
    extern __shared__ unsigned char smem_raw[];
    const int L = a.pl.L;
    const int Ct = a.Ct;
    const int PLl = pp_len(L);   // LDS slots per line
    C* tw = reinterpret_cast<C*>(smem_raw);
    C* b0 = tw + a.pl.ntw;
    C* b1 = b0 + PLl * Ct;
    const int T = blockDim.x;
    const bool inv = a.inv != 0;
    stage_table(tw, a.pl.tw, a.pl.ntw, T);
    if (a.inner > 1) {
        // columns: Ct lines at consecutive inner indices, element (c, j) at pp(j)*Ct + c
        const int64_t g = xcd_slot(blockIdx.x, a.nblocks);   // neighbouring tiles on one XCD
        const int64_t o = g / a.ntiles;
        const int64_t i0 = int64_t(g % a.ntiles) * Ct;
        const int nc = int(imin64(Ct, a.inner - i0));
        const int64_t base = o * int64_t(a.n) * a.inner + i0;
        const float rC = 1.0f / float(Ct);
        for (int b = threadIdx.x; b < L * Ct; b += kU * T) {
            C v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int it = b + u * T;
                v[u] = cmake<C>(R(0), R(0));
                if (it < L * Ct) {
                    const int j = fdiv(it, Ct, rC), c = it - j * Ct;
                    if (c < nc)
                        v[u] = a.in_lines ? a.in[(o * a.inner + i0 + c) * a.n + j] : a.in[base + int64_t(j) * a.inner + c];
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int it = b + u * T;
                if (it < L * Ct) {
                    const int j = fdiv(it, Ct, rC), c = it - j * Ct;
                    b0[pp(j) * Ct + c] = v[u];
                }
            }
        }
        __syncthreads();
        const C* res = fft_lines<R, true>(b0, b1, a.pl, tw, Ct, 1, Ct, inv, T) ? b1 : b0;
        for (int it = threadIdx.x; it < L * Ct; it += T) {
            const int j = fdiv(it, Ct, rC), c = it - j * Ct;
            if (c >= nc) continue;
            C v = cscale(res[pp(j) * Ct + c], a.scale);
            if (a.twN) {
                const int64_t j2 = (i0 + c) / a.tw_inner;
                const C w = a.twN[int64_t(j) * j2];
                v = cmul(v, inv ? cconj(w) : w);
            }
            if (a.out_lines) a.out[(o * a.inner + i0 + c) * a.n + j] = v;
            else a.out[base + int64_t(j) * a.inner + c] = v;
        }
        return;
    }
    // rows: Ct consecutive rows, element (b, j) at b*PLl + pp(j)
    const int64_t r0 = int64_t(blockIdx.x) * Ct;
    const int nr = int(imin64(Ct, a.outer - r0));
    const bool half = a.mode != 0 && L * 2 == a.n;   // even real line: half-length transform
    const int tot = L * Ct;
    const float rL = 1.0f / float(L);
    for (int b = threadIdx.x; b < tot; b += kU * T) {
        C v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = b + u * T;
            v[u] = cmake<C>(R(0), R(0));
            if (it < tot) {
                const int rb = fdiv(it, L, rL), j = it - rb * L;
                if (rb < nr) {
                    const int64_t row = r0 + rb;
                    if (a.mode == 0) {
                        v[u] = a.in[row * a.n + j];
                    } else if (a.mode == 1) {
                        const R* x = a.rin + row * a.n;
                        v[u] = half ? cmake<C>(x[2 * j], x[2 * j + 1]) : cmake<C>(x[j], R(0));
                    } else {
                        // c2r: bins of the full (odd) or half-length (even) inverse
                        const C* X = a.in + row * a.H;
                        if (half) {
                            const int M = L;
                            C A = X[j], B = cconj(X[M - j]);
                            if (j == 0) {   // imaginary parts of DC and Nyquist are ignored
                                A.y = R(0);
                                B.y = R(0);
                            }
                            v[u] = c2r_pre(A, B, j, tw + a.pl.post);
                        } else {
                            C Y = j < a.H ? X[j] : cconj(X[a.n - j]);
                            if (j == 0) Y.y = R(0);
                            v[u] = Y;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int it = b + u * T;
            if (it < tot) {
                const int rb = fdiv(it, L, rL), j = it - rb * L;
                b0[rb * PLl + pp(j)] = v[u];
            }
        }
    }
    __syncthreads();
    const C* res = fft_lines<R, false>(b0, b1, a.pl, tw, Ct, PLl, 1, inv, T) ? b1 : b0;
    if (a.mode == 0) {
        for (int it = threadIdx.x; it < tot; it += T) {
            const int rb = fdiv(it, L, rL), j = it - rb * L;
            if (rb < nr) a.out[(r0 + rb) * a.n + j] = cscale(res[rb * PLl + pp(j)], a.scale);
        }
    } else if (a.mode == 1) {
        const int Hb = a.H;
        for (int it = threadIdx.x; it < Ct * Hb; it += T) {
            const int rb = it / Hb, k = it - rb * Hb;
            if (rb >= nr) continue;
            const C* Z = res + rb * PLl;
            const C X = half ? r2c_bin(Z, L, k, tw + a.pl.post) : Z[pp(k)];
            a.out[(r0 + rb) * Hb + k] = cscale(X, a.scale);
        }
    } else {
        for (int it = threadIdx.x; it < tot; it += T) {
            const int rb = fdiv(it, L, rL), j = it - rb * L;
            if (rb >= nr) continue;
            R* y = a.rout + (r0 + rb) * a.n;
            const C zj = res[rb * PLl + pp(j)];
            if (half) {
                y[2 * j] = zj.x * a.scale;
                y[2 * j + 1] = zj.y * a.scale;
            } else {
                y[j] = zj.x * a.scale;
            }
        }
    }
}

// out[B][A] = in[A][B] (complex), 32x32 tiles through LDS.
template <typename R>
__global__ void __launch_bounds__(256) transpose_kernel(const Cx<R>* in, Cx<R>* out, int64_t A, int64_t B) {
    __shared__ Cx<R> tile[32][33];
    const int64_t a0 = int64_t(blockIdx.y) * 32, b0 = int64_t(blockIdx.x) * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int64_t ai = a0 + r, bi = b0 + threadIdx.x;
        if (ai < A && bi < B) tile[r][threadIdx.x] = in[ai * B + bi];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int64_t bi = b0 + r, ai = a0 + threadIdx.x;
        if (ai < A && bi < B) out[bi * A + ai] = tile[threadIdx.x][r];
    }
}

// Long lines split as n = L1 * L2 (point j1 L2 + j2) leave spectral index
// k1 + L1 k2 at k1 L2 + k2: per outer index an L1 x L2 matrix whose transpose
// is natural order.  out[o][k][i] = f(in[o][k1 L2 + k2][i]) for k = k1 + L1 k2
// < count: f scales (complex output) or takes the scaled real part.  Rows
// (inner == 1) go through 32x32 LDS tiles of the matrix (blockIdx.z = outer).
template <typename R>
__global__ void __launch_bounds__(256) split_order_rows_kernel(const Cx<R>* in, Cx<R>* out, R* rout, int L1, int L2,
                                                               int64_t count, R scale) {
    __shared__ Cx<R> tile[32][33];
    const int64_t o = blockIdx.z;
    const Cx<R>* src = in + o * int64_t(L1) * L2;
    const int k10 = blockIdx.y * 32, k20 = blockIdx.x * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int k1 = k10 + r, k2 = k20 + threadIdx.x;
        if (k1 < L1 && k2 < L2) tile[r][threadIdx.x] = src[int64_t(k1) * L2 + k2];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int k2 = k20 + r, k1 = k10 + threadIdx.x;
        const int64_t k = int64_t(k2) * L1 + k1;
        if (k1 < L1 && k2 < L2 && k < count) {
            const Cx<R> v = tile[threadIdx.x][r];
            if (rout) rout[o * count + k] = v.x * scale;
            else out[o * count + k] = cscale(v, scale);
        }
    }
}

// Columns (inner > 1): consecutive threads along the inner index.
template <typename R>
__global__ void __launch_bounds__(256) split_order_cols_kernel(const Cx<R>* in, Cx<R>* out, int64_t outer, int L1,
                                                               int L2, int64_t inner, R scale) {
    const int64_t n = int64_t(L1) * L2;
    const int64_t total = outer * n * inner;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t i = e % inner, q = e / inner;
        const int64_t k = q % n, o = q / n;
        const int64_t k1 = k % L1, k2 = k / L1;
        out[e] = cscale(in[(o * n + k1 * L2 + k2) * inner + i], scale);
    }
}

// Complex input of a long real line: mode 1 (r2c, odd n) x[j] + 0i; mode 2
// (c2r) the bins of the full line from the stored half: odd n completes the
// Hermitian spectrum, even n forms the half-length inverse input
// (A + B) + i W_n^-j (A - B) from A = X[j], B = conj X[M - j] (M = n / 2),
// with the imaginary parts of the DC and Nyquist bins ignored.
template <typename R>
__global__ void __launch_bounds__(256) split_real_in_kernel(const R* rin, const Cx<R>* xin, Cx<R>* out, int64_t outer,
                                                            int n, int mode, const Cx<R>* twn) {
    using C = Cx<R>;
    const int H = n / 2 + 1;
    const bool half = mode == 2 && n % 2 == 0;
    const int len = half ? n / 2 : n;
    const int64_t total = outer * len;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t o = e / len;
        const int j = int(e - o * len);
        C v;
        if (mode == 1) {
            v = cmake<C>(rin[o * n + j], R(0));
        } else if (!half) {
            const C* X = xin + o * H;
            v = j < H ? X[j] : cconj(X[n - j]);
            if (j == 0) v.y = R(0);
        } else {
            const C* X = xin + o * H;
            const int M = len;
            C A = X[j], B = cconj(X[M - j]);
            if (j == 0) {
                A.y = R(0);
                B.y = R(0);
            }
            v = c2r_pre(A, B, j, twn);
        }
        out[e] = v;
    }
}

// r2c of a long even real line from its half-length spectrum Z (natural
// order, M = n / 2 points per row): X[k], k <= M.
template <typename R>
__global__ void __launch_bounds__(256) split_r2c_post_kernel(const Cx<R>* Z, Cx<R>* out, int64_t outer, int M,
                                                             const Cx<R>* twn) {
    const int H = M + 1;
    const int64_t total = outer * H;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t o = e / H;
        const int k = int(e - o * H);
        out[e] = r2c_bin(Z + o * M, M, k, twn);
    }
}

template <typename R>
ndc_status line_fft_long(LaunchContext& ctx, int mode, bool inv, const void* in, void* out, int64_t outer, int n,
                         int64_t inner, R scale);

// Whether a line of n points (mode as in line_fft) is transformed by one
// workgroup in LDS (otherwise line_fft takes the split form).
template <typename R>
bool line_fits_lds(LaunchContext& ctx, int mode, int n) {
    const bool half = mode != 0 && n % 2 == 0;
    const int L = half ? n / 2 : n;
    return (size_t(ntw_of(L, half ? n : 0, sizeof(R) == 8)) + size_t(2) * pp_len(L)) * sizeof(Cx<R>) <= ctx.max_lds();
}

constexpr int kInLines = 1, kOutLines = 2;   // line_fft io: see LineFftArgs::in_lines / out_lines

// One line-transform launch along an axis of [outer][n][inner]; lines that do
// not fit one workgroup's LDS take the split form (line_fft_long).  twN and
// tw_inner: the first pass of such a split (see LineFftArgs).  io: kInLines /
// kOutLines for LDS-resident lines (see LineFftArgs).
template <typename R>
ndc_status line_fft(LaunchContext& ctx, int mode, bool inv, const void* in, void* out, int64_t outer, int n,
                    int64_t inner, R scale, const Cx<R>* twN = nullptr, int64_t tw_inner = 1, int io = 0) {
    using C = Cx<R>;
    LineFftArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    const bool half = mode != 0 && n % 2 == 0;
    const int L = half ? n / 2 : n;
    if (!line_fits_lds<R>(ctx, mode, n)) {
        if (io) {
            ctx.err = "line-contiguous input or output needs an LDS-resident line";
            return NDC_ERR_UNSUPPORTED;
        }
        if (twN) {
            ctx.err = "split line factor exceeds the LDS transform size";
            return NDC_ERR_UNSUPPORTED;
        }
        return line_fft_long<R>(ctx, mode, inv, in, out, outer, n, inner, scale);
    }
    if (!make_lineplan(&a.pl, L, half ? n : 0, ctx)) {
        ctx.err = "FFT plan failed";
        return NDC_ERR_UNSUPPORTED;
    }
    a.in = static_cast<const C*>(in);
    a.rin = static_cast<const R*>(in);
    a.out = static_cast<C*>(out);
    a.rout = static_cast<R*>(out);
    a.outer = outer;
    a.inner = inner;
    a.n = n;
    a.H = n / 2 + 1;
    a.mode = mode;
    a.inv = inv ? 1 : 0;
    a.scale = scale;
    a.twN = twN;
    a.tw_inner = tw_inner;
    a.in_lines = (io & kInLines) ? 1 : 0;
    a.out_lines = (io & kOutLines) ? 1 : 0;
    const size_t es = sizeof(C);
    const size_t lds = ctx.max_lds();
    int Ct = 1;
    int64_t blocks = 0;
    if (inner > 1) {
        Ct = int(std::min<int64_t>(16, inner));
        while (Ct > 1 && (size_t(a.pl.ntw) + size_t(2) * pp_len(L) * Ct) * es > lds) Ct /= 2;
        a.ntiles = int32_t((inner + Ct - 1) / Ct);
        blocks = outer * a.ntiles;
    } else {
        Ct = int(std::max<int64_t>(1, std::min<int64_t>(kMaxLpb, 512 / std::max(1, L))));
        while (Ct > 1 && (size_t(a.pl.ntw) + size_t(2) * pp_len(L) * Ct) * es > lds) Ct /= 2;
        blocks = (outer + Ct - 1) / Ct;
    }
    a.Ct = Ct;
    const size_t smem = (size_t(a.pl.ntw) + size_t(2) * pp_len(L) * Ct) * es;
    if (smem > lds || !lds_fits(ctx, reinterpret_cast<const void*>(&line_fft_kernel<R>), smem)) {
        ctx.err = "line longer than the LDS transform supports";
        return NDC_ERR_UNSUPPORTED;
    }
    if (blocks <= 0) return NDC_OK;
    a.nblocks = int32_t(blocks);
    const double bytes = double(outer) * double(inner) * double(n) * double(es) * 2.0;
    ctx.kbegin("fft_processor_lines", bytes, 5.0 * double(outer) * double(inner) * L * std::log2(double(std::max(2, L))));
    NDC_LAUNCH(ctx, ctx.stream(), (line_fft_kernel<R>),
        dim3(unsigned(blocks)), dim3(threads_for(int64_t(L) * Ct)), smem, a);
    ctx.kend();
    return check_launch(ctx, "processor line transform");
}

inline unsigned grid_for(int64_t total) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, int64_t(1) << 20)));
}

// Bluestein's chirp-z transform (L. I. Bluestein, "A linear filtering approach
// to the computation of the discrete Fourier transform", 1970).  With
// c[m] = exp(-i pi m^2 / N) (its conjugate for the inverse) and
// j k = (j^2 + k^2 - (k - j)^2) / 2,
//   X[k] = c[k] sum_j (x[j] c[j]) conj(c[k - j]),
// a circular convolution of length M >= 2N - 1, a power of two whose
// transforms split into LDS-resident factors.  Lines with no LDS-resident
// split of their own (a large prime length, or twice one) take it.  Real
// lines run as complex lines of their full length: r2c keeps bins 0..N/2;
// c2r first rebuilds the Hermitian spectrum from them (the imaginary parts
// of DC and, for even N, Nyquist ignored, as in the resident kernel).
constexpr int64_t kBluesteinMaxM = int64_t(1) << 24;

// In-place forward FFT of a power-of-two length, FP64 (iterative radix 2),
// for the chirp's spectrum.
void host_fft_pow2(std::vector<std::complex<double>>& v) {
    const size_t M = v.size();
    for (size_t i = 1, j = 0; i < M; ++i) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(v[i], v[j]);
    }
    std::vector<std::complex<double>> w(M / 2);
    for (size_t k = 0; k < M / 2; ++k) {
        const double a = -2.0 * M_PI * double(k) / double(M);
        w[k] = {std::cos(a), std::sin(a)};
    }
    for (size_t len = 2; len <= M; len <<= 1) {
        const size_t half = len / 2, step = M / len;
        for (size_t b = 0; b < M; b += len)
            for (size_t k = 0; k < half; ++k) {
                const std::complex<double> t = v[b + k + half] * w[k * step];
                v[b + k + half] = v[b + k] - t;
                v[b + k] += t;
            }
    }
}

// Line l = o * inner + i of [outer][n][inner] times the chirp, zero-extended
// to M points (M = 2^logM), as row l of dst.
template <typename R>
__global__ void __launch_bounds__(256) blue_in_kernel(const void* in, Cx<R>* dst, const Cx<R>* chirp, int64_t outer,
                                                      int n, int64_t inner, int logM, int mode) {
    using C = Cx<R>;
    const int M = 1 << logM, H = n / 2 + 1;
    const int64_t total = (outer * inner) << logM;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t l = e >> logM;
        const int j = int(e & (M - 1));
        C v = cmake<C>(R(0), R(0));
        if (j < n) {
            const int64_t o = l / inner, i = l - o * inner;
            if (mode == 0) {
                v = static_cast<const C*>(in)[(o * n + j) * inner + i];
            } else if (mode == 1) {
                v = cmake<C>(static_cast<const R*>(in)[o * n + j], R(0));
            } else {
                const C* X = static_cast<const C*>(in) + o * H;
                v = j < H ? X[j] : cconj(X[n - j]);
                if (j == 0 || 2 * j == n) v.y = R(0);
            }
            v = cmul(v, chirp[j]);
        }
        dst[e] = v;
    }
}

// Row-wise product with the chirp's spectrum (which carries the 1/M of the
// inverse transform).
template <typename R>
__global__ void __launch_bounds__(256) blue_mul_kernel(Cx<R>* buf, const Cx<R>* spec, int64_t lines, int logM) {
    const int64_t total = lines << logM;
    const int64_t mask = (int64_t(1) << logM) - 1;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x)
        buf[e] = cmul(buf[e], spec[e & mask]);
}

// The first N (c2c, c2r) or N/2 + 1 (r2c) points of each row times the
// chirp and `scale`, into the output's layout (c2r keeps real parts).
template <typename R>
__global__ void __launch_bounds__(256) blue_out_kernel(const Cx<R>* src, const Cx<R>* chirp, void* out, int64_t outer,
                                                       int n, int64_t inner, int logM, int mode, R scale) {
    using C = Cx<R>;
    const int count = mode == 1 ? n / 2 + 1 : n;
    const int64_t total = outer * inner * count;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t l = e / count;
        const int k = int(e - l * count);
        const C v = cscale(cmul(src[(l << logM) + k], chirp[k]), scale);
        if (mode == 0) {
            const int64_t o = l / inner, i = l - o * inner;
            static_cast<C*>(out)[(o * n + k) * inner + i] = v;
        } else if (mode == 1) {
            static_cast<C*>(out)[e] = v;
        } else {
            static_cast<R*>(out)[l * n + k] = v.x;
        }
    }
}

template <typename R>
ndc_status line_fft_bluestein(LaunchContext& ctx, int mode, bool inv, const void* in, void* out, int64_t outer, int n,
                              int64_t inner, R scale) {
    using C = Cx<R>;
    constexpr bool kDbl = sizeof(R) == 8;
    const size_t es = sizeof(C);
    int logM = 0;
    while ((int64_t(1) << logM) < 2 * int64_t(n) - 1) ++logM;
    const int64_t M = int64_t(1) << logM;
    if (M > kBluesteinMaxM) {
        ctx.err = "line length has no LDS-resident split and exceeds the chirp-z transform's limit";
        return NDC_ERR_UNSUPPORTED;
    }
    const bool backward = inv || mode == 2;
    // chirp (n points) then the M-point spectrum of its conjugate, scaled by
    // 1/M, made once per handle in FP64
    const std::string key = std::string("bluestein:") + std::to_string(n) + (backward ? ":inv" : ":fwd") +
                            (kDbl ? ":f64" : ":f32");
    const size_t bytes = size_t(n + M) * es;
    std::string& made = ctx.memo(key);
    const void* tab = nullptr;
    if (made.empty() && !ctx.dry_run()) {
        std::vector<std::complex<double>> c(static_cast<size_t>(n)), b(static_cast<size_t>(M));
        for (int64_t m = 0; m < n; ++m) {
            const double ang = M_PI * double((m * m) % (2 * int64_t(n))) / double(n);
            c[size_t(m)] = {std::cos(ang), backward ? std::sin(ang) : -std::sin(ang)};
            b[size_t(m)] = std::conj(c[size_t(m)]);
            if (m > 0) b[size_t(M - m)] = std::conj(c[size_t(m)]);
        }
        host_fft_pow2(b);
        std::vector<R> host(size_t(n + M) * 2);
        for (int64_t m = 0; m < n; ++m) {
            host[size_t(2 * m)] = R(c[size_t(m)].real());
            host[size_t(2 * m + 1)] = R(c[size_t(m)].imag());
        }
        for (int64_t j = 0; j < M; ++j) {
            host[size_t(2 * (n + j))] = R(b[size_t(j)].real() / double(M));
            host[size_t(2 * (n + j) + 1)] = R(b[size_t(j)].imag() / double(M));
        }
        tab = ctx.constant_table(key, host.data(), bytes);
        if (tab) made = "1";
    } else {
        tab = ctx.constant_table(key, nullptr, bytes);
    }
    const int64_t lines = outer * inner;
    C* buf = static_cast<C*>(ctx.scratch(3, size_t(lines * M) * es));
    if (!tab || !buf) {
        ctx.err = "chirp-z workspace allocation failed";
        return NDC_ERR_OUT_OF_MEMORY;
    }
    const C* chirp = static_cast<const C*>(tab);
    const C* spec = chirp + n;
    ndc_status s = NDC_OK;
    const double io_in = double(lines) * (double(n) * (mode == 1 ? sizeof(R) : es) + double(M) * es);
    ctx.kbegin("fft_processor_chirp_io", io_in, 0);
    NDC_LAUNCH(ctx, ctx.stream(), (blue_in_kernel<R>), dim3(grid_for(lines * M)), dim3(256), 0, in, buf, chirp, outer,
               n, inner, logM, mode);
    ctx.kend();
    if ((s = check_launch(ctx, "chirp-z input")) != NDC_OK) return s;
    if ((s = line_fft<R>(ctx, 0, false, buf, buf, lines, int(M), 1, R(1))) != NDC_OK) return s;
    ctx.kbegin("fft_processor_chirp_io", 2.0 * double(lines * M) * es, 0);
    NDC_LAUNCH(ctx, ctx.stream(), (blue_mul_kernel<R>), dim3(grid_for(lines * M)), dim3(256), 0, buf, spec, lines,
               logM);
    ctx.kend();
    if ((s = check_launch(ctx, "chirp-z product")) != NDC_OK) return s;
    if ((s = line_fft<R>(ctx, 0, true, buf, buf, lines, int(M), 1, R(1))) != NDC_OK) return s;
    const int64_t count = mode == 1 ? n / 2 + 1 : n;
    ctx.kbegin("fft_processor_chirp_io", double(lines * count) * (es + (mode == 2 ? sizeof(R) : es)), 0);
    NDC_LAUNCH(ctx, ctx.stream(), (blue_out_kernel<R>), dim3(grid_for(lines * count)), dim3(256), 0, buf, chirp, out,
               outer, n, inner, logM, mode, scale);
    ctx.kend();
    return check_launch(ctx, "chirp-z output");
}

// A line of n points that does not fit one workgroup: n (the complex length
// for c2c and odd real lines, n / 2 for even real ones) = L1 * L2 with both
// factors LDS-resident.  Pass A transforms L1-point lines along j1 (L2 and the
// inner index together form the inner extent) and multiplies by
// W^(k1 j2); pass B transforms L2-point lines along j2; the result, in
// k1 L2 + k2 order, is transposed back to natural order.  Real lines go
// through a complex input (odd n) or the half-length packing (even n) with
// its pre/post-processing, as in the LDS-resident kernel.
template <typename R>
ndc_status line_fft_long(LaunchContext& ctx, int mode, bool inv, const void* in, void* out, int64_t outer, int n,
                         int64_t inner, R scale) {
    using C = Cx<R>;
    constexpr bool kDbl = sizeof(R) == 8;
    const size_t es = sizeof(C);
    const size_t lds = ctx.max_lds();
    if (mode != 0 && inner != 1) {
        ctx.err = "real transforms run along rows";
        return NDC_ERR_UNSUPPORTED;
    }
    const bool half = mode != 0 && n % 2 == 0;
    const int64_t Lc = half ? n / 2 : n;
    auto fits = [&](int64_t L) {
        return L >= 2 && (size_t(ntw_of(int(L), 0, kDbl)) + size_t(2) * pp_len(int(L))) * es <= lds;
    };
    int64_t L2 = 0;
    for (int64_t l2 = 2; l2 * l2 <= Lc; ++l2)
        if (Lc % l2 == 0 && fits(l2) && fits(Lc / l2)) L2 = l2;
    if (!L2) return line_fft_bluestein<R>(ctx, mode, inv, in, out, outer, n, inner, scale);
    const int64_t L1 = Lc / L2;
    const C* twc = static_cast<const C*>(ctx.twiddles(Lc, kDbl));
    const C* twn = half ? static_cast<const C*>(ctx.twiddles(n, kDbl)) : twc;
    const int64_t elems = outer * Lc * inner;
    C* t1 = static_cast<C*>(ctx.scratch(1, size_t(elems) * es));
    C* t2 = mode != 0 ? static_cast<C*>(ctx.scratch(2, size_t(outer * std::max<int64_t>(Lc, n / 2 + 1)) * es)) : nullptr;
    if (!twc || !twn || !t1 || (mode != 0 && !t2)) {
        ctx.err = "split line workspace allocation failed";
        return NDC_ERR_OUT_OF_MEMORY;
    }
    ndc_status s = NDC_OK;
    // complex input of the c2c transform
    const C* src = static_cast<const C*>(in);
    if (mode == 1 && !half) {
        ctx.kbegin("fft_processor_split_io", double(outer) * n * (sizeof(R) + es), 0);
        NDC_LAUNCH(ctx, ctx.stream(), (split_real_in_kernel<R>), dim3(grid_for(outer * n)), dim3(256), 0,
                   static_cast<const R*>(in), static_cast<const C*>(nullptr), t2, outer, n, 1, twn);
        ctx.kend();
        if ((s = check_launch(ctx, "split line input")) != NDC_OK) return s;
        src = t2;
    } else if (mode == 2) {
        ctx.kbegin("fft_processor_split_io", double(outer) * (n / 2 + 1 + Lc) * es, 0);
        NDC_LAUNCH(ctx, ctx.stream(), (split_real_in_kernel<R>), dim3(grid_for(outer * Lc)), dim3(256), 0,
                   static_cast<const R*>(nullptr), static_cast<const C*>(in), t2, outer, n, 2, twn);
        ctx.kend();
        if ((s = check_launch(ctx, "split line input")) != NDC_OK) return s;
        src = t2;
    }
    // (an even real line's samples, read as Lc complex points, are the
    // half-length packing itself)
    s = line_fft<R>(ctx, 0, inv, src, t1, outer, int(L1), L2 * inner, R(1), twc, inner);
    if (s == NDC_OK) s = line_fft<R>(ctx, 0, inv, t1, t1, outer * L1, int(L2), inner, R(1));
    if (s != NDC_OK) return s;
    // natural order
    auto order = [&](C* cdst, R* rdst, int64_t count, R sc) -> ndc_status {
        if (inner == 1) {
            const dim3 grid(unsigned((L2 + 31) / 32), unsigned((L1 + 31) / 32), unsigned(outer));
            ctx.kbegin("fft_processor_split_order", double(outer) * (Lc + count) * es, 0);
            NDC_LAUNCH(ctx, ctx.stream(), (split_order_rows_kernel<R>), grid, dim3(32, 8), 0, t1, cdst, rdst, int(L1),
                       int(L2), count, sc);
        } else {
            ctx.kbegin("fft_processor_split_order", 2.0 * double(elems) * es, 0);
            NDC_LAUNCH(ctx, ctx.stream(), (split_order_cols_kernel<R>), dim3(grid_for(elems)), dim3(256), 0, t1, cdst,
                       outer, int(L1), int(L2), inner, sc);
        }
        ctx.kend();
        return check_launch(ctx, "split line reorder");
    };
    if (mode == 0) return order(static_cast<C*>(out), nullptr, Lc, scale);
    if (mode == 2) {
        // even: the complex result is the real line's sample pairs; odd: real parts
        if (half) return order(static_cast<C*>(out), nullptr, Lc, scale);
        return order(nullptr, static_cast<R*>(out), n, scale);
    }
    if (!half) return order(static_cast<C*>(out), nullptr, n / 2 + 1, scale);
    if ((s = order(t2, nullptr, Lc, R(1))) != NDC_OK) return s;
    ctx.kbegin("fft_processor_split_io", double(outer) * (Lc + Lc + 1) * es, 0);
    NDC_LAUNCH(ctx, ctx.stream(), (split_r2c_post_kernel<R>), dim3(grid_for(outer * (Lc + 1))), dim3(256), 0, t2,
               static_cast<C*>(out), outer, int(Lc), twn);
    ctx.kend();
    return check_launch(ctx, "split line post-processing");
}

template <typename R>
ndc_status transpose(LaunchContext& ctx, const void* in, void* out, int64_t A, int64_t B) {
    if (A <= 0 || B <= 0) return NDC_OK;
    const dim3 grid(unsigned((B + 31) / 32), unsigned((A + 31) / 32));
    ctx.kbegin("fft_processor_transpose", 2.0 * double(A) * double(B) * double(sizeof(Cx<R>)), 0);
    NDC_LAUNCH(ctx, ctx.stream(), (transpose_kernel<R>),
        grid, dim3(32, 8), 0, static_cast<const Cx<R>*>(in), static_cast<Cx<R>*>(out), A, B);
    ctx.kend();
    return check_launch(ctx, "transpose");
}

// Forward N-d transform in the reference's layout: the natural spectrum with
// its axes rotated left by one (processor/real.rs:105-155: r2c on the last
// axis, then N-1 times "rotate axes right, transform the new last axis";
// processor/complex.rs:33-115 likewise with c2c on every axis).
template <typename R>
ndc_status proc_forward(bool real, int N, const int64_t* shape, const void* in, void* out, LaunchContext& ctx) {
    using C = Cx<R>;
    const int64_t Fl = shape[N - 1];
    const int64_t H = real ? Fl / 2 + 1 : Fl;
    int64_t rows = 1;
    for (int i = 0; i < N - 1; ++i) rows *= shape[i];
    C* work = N == 1 ? static_cast<C*>(out) : static_cast<C*>(ctx.scratch(0, size_t(rows * H) * sizeof(C)));
    if (!work) return NDC_ERR_OUT_OF_MEMORY;
    ndc_status s = line_fft<R>(ctx, real ? 1 : 0, false, in, work, rows, int(Fl), 1, R(1));
    if (s != NDC_OK) return s;
    int64_t inner = H;
    // the last (axis 0) transform writes the rotated layout itself when its
    // lines fit LDS: line i of [F0][inner] lands at out[i F0 ...]
    const bool fused = N > 1 && line_fits_lds<R>(ctx, 0, int(shape[0]));
    for (int ax = N - 2; ax >= 0; --ax) {
        int64_t outer = 1;
        for (int i = 0; i < ax; ++i) outer *= shape[i];
        if (ax == 0 && fused)
            s = line_fft<R>(ctx, 0, false, work, out, 1, int(shape[0]), inner, R(1), nullptr, 1, kOutLines);
        else
            s = line_fft<R>(ctx, 0, false, work, work, outer, int(shape[ax]), inner, R(1));
        if (s != NDC_OK) return s;
        inner *= shape[ax];
    }
    if (N > 1 && !fused) return transpose<R>(ctx, work, out, shape[0], rows / shape[0] * H);
    return NDC_OK;
}

// Backward: undoes the rotation, inverse transforms, scales by 1/prod F
// (processor/real.rs:233-281, complex.rs:134-145).  `fl` is the real length
// of the last axis (the processor's rp_origin_len) for real output.
template <typename R>
ndc_status proc_backward(bool real, int N, const int64_t* sshape, int64_t fl, const void* in, void* out,
                         LaunchContext& ctx) {
    using C = Cx<R>;
    // natural shape: the spectrum's last axis first
    int64_t nat[kMaxRank];
    nat[0] = sshape[N - 1];
    for (int i = 0; i < N - 1; ++i) nat[i + 1] = sshape[i];
    if (N == 1) nat[0] = sshape[0];
    const int64_t H = nat[N - 1];
    const int64_t Fl = real ? fl : H;
    if (real && (Fl < 1 || H != Fl / 2 + 1)) {
        ctx.err = "the spectrum's last axis does not match the real length of the last forward transform";
        return NDC_ERR_INVALID_ARGUMENT;
    }
    double total = double(Fl);
    int64_t rows = 1;
    for (int i = 0; i < N - 1; ++i) {
        rows *= nat[i];
        total *= double(nat[i]);
    }
    const R scale = R(1.0 / total);
    const C* src = static_cast<const C*>(in);
    C* work = nullptr;
    if (N > 1) {
        work = static_cast<C*>(ctx.scratch(0, size_t(rows * H) * sizeof(C)));
        if (!work) return NDC_ERR_OUT_OF_MEMORY;
        // the first (axis 0) transform reads the rotated layout itself when
        // its lines fit LDS: line i of [F0][inner] is in[i F0 ...]
        const bool fused = line_fits_lds<R>(ctx, 0, int(nat[0]));
        ndc_status s = fused ? NDC_OK : transpose<R>(ctx, in, work, rows / nat[0] * H, nat[0]);
        if (s != NDC_OK) return s;
        int64_t outer = 1;
        for (int ax = 0; ax < N - 1; ++ax) {
            int64_t inner = H;
            for (int i = ax + 1; i < N - 1; ++i) inner *= nat[i];
            if (ax == 0 && fused)
                s = line_fft<R>(ctx, 0, true, in, work, 1, int(nat[0]), inner, R(1), nullptr, 1, kInLines);
            else
                s = line_fft<R>(ctx, 0, true, work, work, outer, int(nat[ax]), inner, R(1));
            if (s != NDC_OK) return s;
            outer *= nat[ax];
        }
        src = work;
    }
    return line_fft<R>(ctx, real ? 2 : 0, true, src, out, rows, int(Fl), 1, scale);
}

}  // namespace

ndc_status launch_fft(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx, KernelSpectrum* ks) {
    switch (p.dtype) {
        case NDC_F32: return run_fft<float, float>(p, x, k, y, ctx, ks);
        case NDC_F64: return run_fft<double, double>(p, x, k, y, ctx, ks);
        case NDC_C64: return run_fft<float, float2>(p, x, k, y, ctx, ks);
        case NDC_C128: return run_fft<double, double2>(p, x, k, y, ctx, ks);
        default: return NDC_ERR_UNSUPPORTED;
    }
}

size_t fft_kernel_spectrum_bytes(const Plan& p, size_t lds) {
    const bool cplx = p.dtype == NDC_C64 || p.dtype == NDC_C128;
    const size_t es = (p.dtype == NDC_F64 || p.dtype == NDC_C128) ? 16 : 8;
    const int N = p.N, last = N - 1;
    const int64_t Fl = p.fft[last];
    const int mode = cplx ? 2 : (Fl % 2 == 0 ? 0 : 1);
    const int64_t L = mode == 0 ? Fl / 2 : Fl;
    if ((size_t(ntw_of(int(L), mode == 0 ? int(Fl) : 0, es == 16)) + size_t(2) * pp_len(int(L))) * es > lds)
        return 0;   // four-step rows
    const int64_t H = cplx ? Fl : Fl / 2 + 1;
    const int64_t Hp = (H + 7) / 8 * 8;
    int64_t n = Hp;
    for (int e = 0; e < last; ++e) n *= p.fft[e];
    return size_t(n) * es;
}

size_t fft_workspace_bytes(const Plan& p, size_t lds) {
    const size_t es = (p.dtype == NDC_F64 || p.dtype == NDC_C128) ? 16 : 8;
    const int last = p.N - 1;
    const int64_t Fl = p.fft[last];
    // the two spectra: half-spectrum rows padded to 8 bins, or full rows of
    // the four-step form (all Fl bins kept)
    size_t spec = fft_kernel_spectrum_bytes(p, lds);
    const bool big = spec == 0;
    if (big) {
        int64_t n = Fl;
        for (int e = 0; e < last; ++e) n *= p.fft[e];
        spec = size_t(n) * es;
    }
    // stage and post-processing tables: at most 2 L + 257 + 64 entries per
    // axis plan (sum over stages of Ns (P - 1) is L - 1, generic radices add
    // P <= 257 each, post-processing L + 1); the four-step form adds its two
    // factor plans and an Fl-entry table
    size_t tables = 0;
    for (int e = 0; e <= last; ++e) tables += (2 * size_t(p.fft[e]) + 600) * es + 256;
    if (big) tables += (3 * size_t(Fl) + 1200) * es + 768;
    return 2 * spec + tables + 512;
}

ndc_status launch_fft_processor(int dtype, bool forward, int N, const int64_t* shape, int64_t origin_len,
                                const void* in, void* out, LaunchContext& ctx) {
    const bool real = dtype == NDC_F32 || dtype == NDC_F64;
    const bool dbl = dtype == NDC_F64 || dtype == NDC_C128;
    if (dbl) return forward ? proc_forward<double>(real, N, shape, in, out, ctx)
                            : proc_backward<double>(real, N, shape, origin_len, in, out, ctx);
    return forward ? proc_forward<float>(real, N, shape, in, out, ctx)
                   : proc_backward<float>(real, N, shape, origin_len, in, out, ctx);
}

}  // namespace ndc
