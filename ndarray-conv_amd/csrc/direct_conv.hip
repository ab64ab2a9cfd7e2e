// Direct convolution on gfx950: out[o] = sum_t pds[o*s + t*d] * kv(t) over the
// non-zero taps of the (reversed) kernel (conv/mod.rs:177-197,
// dilation/mod.rs:34-60), with the padded input never materialised: each
// workgroup stages the input region its output tile needs into LDS, resolving
// every padded coordinate through the per-axis border maps (in closed form,
// border1d.hpp) on the way in.
//
// Layout: the last two axes form the tile plane (a rank-1 problem gets a
// virtual unit axis in front); lower axes are looped over, one kernel tap at
// a time.  Arithmetic stays in the element type (integers wrap).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "border1d.hpp"
#include "device_common.hpp"
#include "launch.hpp"

namespace ndc {
namespace {

constexpr int kThreads = 256;
constexpr int kTileX = 64;           // outputs per tile row: one wavefront
// Largest tile height (a kernel template argument TY; kTileX * TY / kThreads
// outputs per thread): 64 for elements up to 4 bytes, 32 for 8-byte and 16 for
// 16-byte elements, whose wider windows would spill.  4-byte launches use 32
// rows when 64 would leave the chip underfilled or waste rows.
template <typename T> constexpr int kMaxTileY = sizeof(T) > 8 ? 16 : (sizeof(T) > 4 ? 32 : 64);

// Region cells staged per thread per round (fewer for 16-byte elements, which
// would otherwise spill).
template <typename T> constexpr int staged_per_round() { return sizeof(T) > 8 ? 2 : 8; }

// Kernel columns per register window of the sliding form (32 for 4-byte
// elements, 16 for wider ones, whose windows would spill).
template <typename T> constexpr int kCX = sizeof(T) == 4 ? 32 : 16;

// A value every lane holds identically (a broadcast LDS read), moved to
// scalar registers: the zero-tap test becomes a scalar branch and the
// multiply takes a scalar operand.
template <typename T>
__device__ __forceinline__ T wave_uniform(T v) {
    if constexpr (sizeof(T) < 4) {
        return T(__builtin_amdgcn_readfirstlane(int(v)));
    } else {
        static_assert(sizeof(T) % 4 == 0, "32-bit words");
        constexpr int n = int(sizeof(T) / 4);
        int w[n];
        __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
        for (int i = 0; i < n; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
        __builtin_memcpy(&v, w, sizeof(T));
        return v;
    }
}

template <typename T>
struct DirectArgs {
    const T* x;
    const T* k;
    T* y;
    int32_t lo[kMaxRank];       // front pad: padded cell p of the interior is x index p - lo
    int32_t n[kMaxRank];        // input extent
    int32_t pds[kMaxRank];      // padded extent
    int32_t side[kMaxRank][2];  // Side1D of the front and back fill
    T cst[kMaxRank][2];
    int64_t xstr[kMaxRank];
    int64_t kstr[kMaxRank];
    int64_t s[kMaxRank];
    int64_t d[kMaxRank];
    int64_t out[kMaxRank];
    int64_t K[kMaxRank];
    int64_t nouter;     // product of the output extents of axes 0..No-3
    int64_t ktaps_outer;// product of the kernel extents of axes 0..No-3
    int32_t No;         // rank after promotion to >= 2
    int32_t rev;
    int32_t KYc, KXc;   // kernel chunk staged per pass
    int32_t gy, gx;     // gcd(stride, dilation) per plane axis: region subsampling
    int32_t RY, RX;     // region extents in LDS
    int32_t TO;         // element offset of the taps in LDS, after the region, 16-byte aligned
    int32_t KS;         // tap row stride in LDS: KXc rounded up to kCX (zero-filled)
    int32_t ntx, nty;   // tiles per plane
};

// Elements per 16-byte LDS read of the sliding form's taps: 4 for 4-byte
// elements; 1 (scalar reads) otherwise.
template <typename T> constexpr int vec_elems() { return sizeof(T) == 4 ? 4 : 1; }

// N 4-byte elements from 16-byte aligned LDS at p, in 16-byte reads.
template <typename T, int N>
__device__ __forceinline__ void lds_vec_load(const T* p, T (&out)[N]) {
    static_assert(sizeof(T) == 4 && N % 4 == 0, "whole 16-byte reads of 4-byte elements");
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int j = 0; j < N / 4; ++j) {
        const uint4 v = q[j];
        out[4 * j + 0] = __builtin_bit_cast(T, v.x);
        out[4 * j + 1] = __builtin_bit_cast(T, v.y);
        out[4 * j + 2] = __builtin_bit_cast(T, v.z);
        out[4 * j + 3] = __builtin_bit_cast(T, v.w);
    }
}

// Border-map code of padded cell q of axis e (plan.cpp axis_map, position by
// position): an index into x, or kUnwritten / kConstFront / kConstBack.
template <typename T>
__device__ __forceinline__ int32_t axis_code(const DirectArgs<T>& a, int e, int q) {
    return resolve1d<int>(q, a.n[e], a.lo[e], a.side[e][0], a.side[e][1]);
}

// Source of a plane cell from its row and column codes: true with the offset
// into x, or false with the cell's value in *fill.  No memory access, so that
// the staging loop can issue all of a round's loads before using any.
template <typename T>
__device__ __forceinline__ bool plane_source(const DirectArgs<T>& a, int32_t cy, int32_t cx, int oc_axis, T ocval,
                                             bool ounw, int64_t xoff, int64_t* off, T* fill) {
    const int ay = a.No - 2, ax = a.No - 1;
    *off = 0;
    *fill = Arith<T>::zero();
    // the highest axis whose fill wrote a constant wins (later fills overwrite)
    if (cx <= kConstFront) {
        *fill = a.cst[ax][cx == kConstFront ? 0 : 1];
        return false;
    }
    if (cy <= kConstFront) {
        *fill = a.cst[ay][cy == kConstFront ? 0 : 1];
        return false;
    }
    if (oc_axis >= 0) {
        *fill = ocval;
        return false;
    }
    if (cx == kUnwritten || cy == kUnwritten || ounw) return false;
    *off = xoff + int64_t(cy) * a.xstr[ay] + int64_t(cx) * a.xstr[ax];
    return true;
}

// SLIDE (stride equal to dilation along the row, so consecutive outputs read
// consecutive staged columns): each thread computes kItems (16, 8 or 4)
// adjacent outputs of one tile row from a register window of
// kCX + kItems - 1 staged values per kCX kernel columns, instead of one LDS
// read per product.  The taps of an
// output are summed in the same order in both forms.
template <typename T, bool SLIDE, int TY>
__global__ void __launch_bounds__(kThreads) direct_conv_kernel(DirectArgs<T> a) {
    constexpr int kItems = TY * kTileX / kThreads;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* region = reinterpret_cast<T*>(smem_raw);
    T* taps = region + a.TO;

    const int ay = a.No - 2, ax = a.No - 1;
    int64_t bid = blockIdx.x;
    const int tx = int(bid % a.ntx);
    bid /= a.ntx;
    const int ty = int(bid % a.nty);
    int64_t outer = bid / a.nty;

    // outer output coordinates (axes 0..No-3)
    int64_t oo[kMaxRank];
    int64_t yoff = 0;
    {
        int64_t rem = outer;
        int64_t ostr = a.out[ay] * a.out[ax];
        for (int e = a.No - 3; e >= 0; --e) {
            oo[e] = rem % a.out[e];
            rem /= a.out[e];
            yoff += oo[e] * ostr;
            ostr *= a.out[e];
        }
    }
    const int64_t oy0 = int64_t(ty) * TY, ox0 = int64_t(tx) * kTileX;
    // plain form: column lx, rows ly0 + 4i; sliding form: row ly0, columns lx + i
    const int lx = SLIDE ? (threadIdx.x % (kTileX / kItems)) * kItems : threadIdx.x % kTileX;
    const int ly0 = SLIDE ? threadIdx.x / (kTileX / kItems) : threadIdx.x / kTileX;

    T acc[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) acc[i] = Arith<T>::zero();

    const int64_t sy = a.s[ay], sx = a.s[ax], dy = a.d[ay], dx = a.d[ax];
    // region offsets: gcd(s, d) divides both terms, so the subsampled index is
    // ly*(s/g) + cy*(d/g)
    const int sy_ = int(sy / a.gy), dy_ = int(dy / a.gy), sx_ = int(sx / a.gx), dx_ = int(dx / a.gx);
    int base[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) base[i] = (ly0 + i * (kThreads / kTileX)) * sy_ * a.RX + lx * sx_;

    for (int64_t to = 0; to < a.ktaps_outer; ++to) {
        // outer tap -> padded coordinate per outer axis -> border code
        int oc_axis = -1;
        T ocval = Arith<T>::zero();
        bool ounw = false;
        int64_t xoff = 0, koff = 0;
        {
            int64_t rem = to;
            for (int e = a.No - 3; e >= 0; --e) {
                const int64_t t = rem % a.K[e];
                rem /= a.K[e];
                const int64_t pe = oo[e] * a.s[e] + t * a.d[e];
                const int32_t c = axis_code(a, e, int(pe));
                if (c >= 0) xoff += int64_t(c) * a.xstr[e];
                else if (c == kUnwritten) ounw = true;
                else if (oc_axis < e) { oc_axis = e; ocval = a.cst[e][c == kConstFront ? 0 : 1]; }
                koff += (a.rev ? a.K[e] - 1 - t : t) * a.kstr[e];
            }
        }
        for (int ky0 = 0; ky0 < a.K[ay]; ky0 += a.KYc) {
            const int kyn = int(a.K[ay] - ky0 < a.KYc ? a.K[ay] - ky0 : a.KYc);
            for (int kx0 = 0; kx0 < a.K[ax]; kx0 += a.KXc) {
                const int kxn = int(a.K[ax] - kx0 < a.KXc ? a.K[ax] - kx0 : a.KXc);
                __syncthreads();
                // stage the input region: rows py = oy0*sy + ky0*dy + r*gy
                const int64_t py0 = oy0 * sy + int64_t(ky0) * dy;
                const int64_t px0 = ox0 * sx + int64_t(kx0) * dx;
                const int64_t leny = (a.out[ay] - 1) * sy + (a.K[ay] - 1) * dy + 1;  // padded extent used
                const int64_t lenx = (a.out[ax] - 1) * sx + (a.K[ax] - 1) * dx + 1;
                // Rounds of kSU cells per thread: every cell's source is worked
                // out first, then all of the round's loads are issued, then the
                // values are selected and written, so the loads overlap.
                constexpr int kSU = staged_per_round<T>();
                const int nreg = a.RY * a.RX;
                const float rRX = 1.0f / float(a.RX);
                for (int i0 = threadIdx.x; i0 < nreg; i0 += kSU * kThreads) {
                    int64_t off[kSU];
                    T fill[kSU];
                    bool ok[kSU];
#pragma unroll
                    for (int u = 0; u < kSU; ++u) {
                        const int i = i0 + u * kThreads;
                        const int ic = i < nreg ? i : 0;
                        int r = int(float(ic) * rRX);
                        int c = ic - r * a.RX;
                        if (c < 0) { --r; c += a.RX; } else if (c >= a.RX) { ++r; c -= a.RX; }
                        const int64_t py = py0 + int64_t(r) * a.gy;
                        const int64_t px = px0 + int64_t(c) * a.gx;
                        ok[u] = false;
                        off[u] = 0;
                        fill[u] = Arith<T>::zero();
                        if (i < nreg && py < leny && px < lenx)
                            ok[u] = plane_source(a, axis_code(a, ay, int(py)), axis_code(a, ax, int(px)), oc_axis,
                                                 ocval, ounw, xoff, &off[u], &fill[u]);
                    }
                    T raw[kSU];
#pragma unroll
                    for (int u = 0; u < kSU; ++u) raw[u] = a.x[off[u]];   // offset 0 (a valid element) where !ok
#pragma unroll
                    for (int u = 0; u < kSU; ++u) {
                        const int i = i0 + u * kThreads;
                        if (i < nreg) region[i] = ok[u] ? raw[u] : fill[u];
                    }
                }
                // stage the kernel taps of this chunk (value of reversed-view tap
                // t), rows KS apart, zeros past kxn
                for (int i = threadIdx.x; i < kyn * a.KS; i += kThreads) {
                    const int cy = i / a.KS, cx = i - cy * a.KS;
                    T v = Arith<T>::zero();
                    if (cx < kxn) {
                        const int64_t ty_ = ky0 + cy, tx_ = kx0 + cx;
                        const int64_t kiy = a.rev ? a.K[ay] - 1 - ty_ : ty_;
                        const int64_t kix = a.rev ? a.K[ax] - 1 - tx_ : tx_;
                        v = a.k[koff + kiy * a.kstr[ay] + kix * a.kstr[ax]];
                    }
                    taps[i] = v;
                }
                __syncthreads();
                if constexpr (SLIDE) {
                    for (int cy = 0; cy < kyn; ++cy) {
                        const T* rrow = region + (ly0 * sy_ + cy * dy_) * a.RX + lx;
                        const T* trow = taps + cy * a.KS;
                        for (int cx0 = 0; cx0 < kxn; cx0 += kCX<T>) {
                            // window values this chunk uses: up to column lx + kxn + kItems - 2 < RX
                            const int nw = min(kCX<T>, kxn - cx0) + kItems - 1;
                            T w[kCX<T> + kItems - 1];
#pragma unroll
                            for (int i = 0; i < kCX<T> + kItems - 1; ++i)
                                w[i] = i < nw ? rrow[cx0 + i] : Arith<T>::zero();
                            // the chunk's kCX taps (a broadcast read; trow + cx0 is 16-byte
                            // aligned: TO, KS and cx0 are multiples of vec_elems)
                            T tv[kCX<T>];
                            if constexpr (vec_elems<T>() > 1) {
                                lds_vec_load(trow + cx0, tv);
                            } else {
#pragma unroll
                                for (int c = 0; c < kCX<T>; ++c) tv[c] = trow[cx0 + c];
                            }
#pragma unroll
                            for (int c = 0; c < kCX<T>; ++c) {
                                // taps past kxn are the zero padding of the row: skipped below
                                const T v = wave_uniform(tv[c]);
                                if (Arith<T>::is_zero(v)) continue;   // zero taps are not in the offset list
#pragma unroll
                                for (int i = 0; i < kItems; ++i) acc[i] = Arith<T>::mac(acc[i], w[i + c], v);
                            }
                        }
                    }
                } else {
                    for (int cy = 0; cy < kyn; ++cy) {
                        const int roff = cy * dy_ * a.RX;
                        for (int cx = 0; cx < kxn; ++cx) {
                            const T v = taps[cy * a.KS + cx];
                            if (Arith<T>::is_zero(v)) continue;   // zero taps are not in the offset list
                            const int off = roff + cx * dx_;
#pragma unroll
                            for (int i = 0; i < kItems; ++i)
                                acc[i] = Arith<T>::mac(acc[i], region[base[i] + off], v);
                        }
                    }
                }
            }
        }
    }
    if constexpr (SLIDE) {
        const int64_t oy = oy0 + ly0;
        if (oy < a.out[ay]) {
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const int64_t ox = ox0 + lx + i;
                if (ox < a.out[ax]) a.y[yoff + oy * a.out[ax] + ox] = acc[i];
            }
        }
    } else {
        const int64_t ox = ox0 + lx;
        if (ox < a.out[ax]) {
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const int64_t oy = oy0 + ly0 + i * (kThreads / kTileX);
                if (oy < a.out[ay]) a.y[yoff + oy * a.out[ax] + ox] = acc[i];
            }
        }
    }
}

int64_t gcd64(int64_t a, int64_t b) {
    while (b) {
        int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// Launches the TY-row form; false when its LDS request does not fit.
template <typename T, int TY>
bool launch_tile(LaunchContext& ctx, bool slide, const DirectArgs<T>& a, int64_t nblocks, size_t smem) {
    const void* kp = slide ? reinterpret_cast<const void*>(&direct_conv_kernel<T, true, TY>)
                           : reinterpret_cast<const void*>(&direct_conv_kernel<T, false, TY>);
    if (!lds_fits(ctx, kp, smem)) return false;
    if (slide)
        NDC_LAUNCH(ctx, ctx.stream(), (direct_conv_kernel<T, true, TY>),
            dim3(unsigned(nblocks)), dim3(kThreads), smem, a);
    else
        NDC_LAUNCH(ctx, ctx.stream(), (direct_conv_kernel<T, false, TY>),
            dim3(unsigned(nblocks)), dim3(kThreads), smem, a);
    return true;
}

template <typename T>
ndc_status launch_typed(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx,
                        T (*conv_const)(const ndc_value&)) {
    DirectArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    // the closed-form border codes are evaluated in 32-bit positions
    for (int i = 0; i < p.N; ++i)
        if (p.pds[i] > INT32_MAX) {
            ctx.err = "direct convolution: a padded axis longer than 2^31 - 1 cells";
            return NDC_ERR_UNSUPPORTED;
        }
    const int shift = p.N == 1 ? 1 : 0;   // promote rank 1 to a unit-row plane
    a.No = p.N + shift;
    if (shift) {
        // a unit axis: the single cell is interior
        a.lo[0] = 0;
        a.n[0] = 1;
        a.pds[0] = 1;
        a.side[0][0] = a.side[0][1] = kS1Const;
        a.xstr[0] = 0;
        a.kstr[0] = 0;
        a.s[0] = 1;
        a.d[0] = 1;
        a.out[0] = 1;
        a.K[0] = 1;
        a.cst[0][0] = a.cst[0][1] = conv_const(ndc_value{0, 0, 0});
    }
    for (int i = 0; i < p.N; ++i) {
        const int j = i + shift;
        a.lo[j] = int32_t(p.pad[i][0]);
        a.n[j] = int32_t(p.x_shape[i]);
        a.pds[j] = int32_t(p.pds[i]);
        a.side[j][0] = side1d(p.side[i][0].tag);
        a.side[j][1] = side1d(p.side[i][1].tag);
        a.xstr[j] = p.x_str[i];
        a.kstr[j] = p.k_str[i];
        a.s[j] = p.stride[i];
        a.d[j] = p.dil[i];
        a.out[j] = p.out[i];
        a.K[j] = p.k_shape[i];
        a.cst[j][0] = conv_const(p.side[i][0].value);
        a.cst[j][1] = conv_const(p.side[i][1].value);
    }
    a.x = static_cast<const T*>(x);
    a.k = static_cast<const T*>(k);
    a.y = static_cast<T*>(y);
    a.rev = p.reverse ? 1 : 0;
    const int ay = a.No - 2, ax = a.No - 1;
    a.nouter = 1;
    a.ktaps_outer = 1;
    for (int e = 0; e < a.No - 2; ++e) {
        a.nouter *= a.out[e];
        a.ktaps_outer *= a.K[e];
    }
    a.gy = int(gcd64(a.s[ay], a.d[ay]));
    a.gx = int(gcd64(a.s[ax], a.d[ax]));
    if (a.gy <= 0) a.gy = 1;
    if (a.gx <= 0) a.gx = 1;
    // tile height: the largest unless, for 4-byte elements, 64-row tiles would
    // give under four workgroups per CU or waste over a tenth of the rows
    int ty = kMaxTileY<T>;
    if (ty == 64) {
        const int64_t ntx = (a.out[ax] + kTileX - 1) / kTileX, nty = (a.out[ay] + 63) / 64;
        if (a.nouter * ntx * nty < 4 * int64_t(ctx.cu_count()) || (nty * 64 - a.out[ay]) * 10 > a.out[ay]) ty = 32;
    }
    // Choose the kernel chunk so that region + taps fit the LDS budget.  LDS
    // layout: the region (RY rows of RX, the staged width rounded up to odd),
    // then the taps in rows of KS (a multiple of kCX, zero-filled), 16-byte
    // aligned.
    const size_t budget = 48 * 1024;
    constexpr int64_t E = vec_elems<T>();
    struct Layout { int64_t ry, rx, to, ks, total; };
    auto layout = [&](int64_t kyc, int64_t kxc) {
        Layout l;
        l.ry = ((ty - 1) * a.s[ay] + (kyc - 1) * a.d[ay]) / a.gy + 1;
        // odd row stride: a wave's eight tile rows start on distinct LDS banks
        l.rx = (((kTileX - 1) * a.s[ax] + (kxc - 1) * a.d[ax]) / a.gx + 1) | 1;
        l.to = (l.ry * l.rx + E - 1) / E * E;
        l.ks = (kxc + kCX<T> - 1) / kCX<T> * kCX<T>;
        l.total = l.to + kyc * l.ks;
        return l;
    };
    int64_t kyc = a.K[ay], kxc = a.K[ax];
    for (;;) {
        const size_t bytes = size_t(layout(kyc, kxc).total) * sizeof(T);
        if (bytes <= budget || (kyc == 1 && kxc == 1)) break;
        if (kyc >= kxc && kyc > 1) kyc = (kyc + 1) / 2;
        else if (kxc > 1) kxc = (kxc + 1) / 2;
        else kyc = (kyc + 1) / 2;
    }
    const Layout lay = layout(kyc, kxc);
    const size_t smem = size_t(lay.total) * sizeof(T);
    if (smem > ctx.max_lds()) {
        ctx.err = "direct convolution: stride/dilation too large for the LDS tile";
        return NDC_ERR_UNSUPPORTED;
    }
    a.KYc = int(kyc);
    a.KXc = int(kxc);
    a.RY = int(lay.ry);
    a.RX = int(lay.rx);
    a.TO = int(lay.to);
    a.KS = int(lay.ks);
    a.ntx = int((a.out[ax] + kTileX - 1) / kTileX);
    a.nty = int((a.out[ay] + ty - 1) / ty);
    const int64_t nblocks = a.nouter * a.ntx * a.nty;
    if (nblocks <= 0) return NDC_OK;
    if (nblocks > (int64_t(1) << 31) - 1) {
        ctx.err = "direct convolution: output too large for one launch";
        return NDC_ERR_UNSUPPORTED;
    }
    ctx.kbegin("direct_conv", double(sizeof(T)) * double(p.x_numel + p.k_numel + p.y_numel),
               2.0 * double(p.y_numel) * double(p.k_numel));
    // the sliding form needs consecutive outputs on consecutive staged columns
    const bool slide = a.s[ax] == a.d[ax];
    if (ty == kMaxTileY<T>) {
        if (!launch_tile<T, kMaxTileY<T>>(ctx, slide, a, nblocks, smem)) return NDC_ERR_UNSUPPORTED;
    } else if constexpr (kMaxTileY<T> == 64) {
        if (!launch_tile<T, 32>(ctx, slide, a, nblocks, smem)) return NDC_ERR_UNSUPPORTED;
    }
    ctx.kend();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx.err = std::string("direct convolution launch: ") + hipGetErrorString(e);
        return NDC_ERR_DEVICE;
    }
    return NDC_OK;
}

template <typename T> T cst_real(const ndc_value& v) { return T(v.re); }
template <typename T> T cst_int(const ndc_value& v) { return T(v.i); }
float2 cst_c64(const ndc_value& v) { return cmake<float2>(float(v.re), float(v.im)); }
double2 cst_c128(const ndc_value& v) { return cmake<double2>(v.re, v.im); }

}  // namespace

ndc_status launch_direct(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx) {
    switch (p.dtype) {
        case NDC_F32: return launch_typed<float>(p, x, k, y, ctx, cst_real<float>);
        case NDC_F64: return launch_typed<double>(p, x, k, y, ctx, cst_real<double>);
        case NDC_C64: return launch_typed<float2>(p, x, k, y, ctx, cst_c64);
        case NDC_C128: return launch_typed<double2>(p, x, k, y, ctx, cst_c128);
        case NDC_I8: return launch_typed<int8_t>(p, x, k, y, ctx, cst_int<int8_t>);
        case NDC_I16: return launch_typed<int16_t>(p, x, k, y, ctx, cst_int<int16_t>);
        case NDC_I32: return launch_typed<int32_t>(p, x, k, y, ctx, cst_int<int32_t>);
        case NDC_I64: return launch_typed<int64_t>(p, x, k, y, ctx, cst_int<int64_t>);
        case NDC_U8: return launch_typed<uint8_t>(p, x, k, y, ctx, cst_int<uint8_t>);
        case NDC_U16: return launch_typed<uint16_t>(p, x, k, y, ctx, cst_int<uint16_t>);
        case NDC_U32: return launch_typed<uint32_t>(p, x, k, y, ctx, cst_int<uint32_t>);
        case NDC_U64: return launch_typed<uint64_t>(p, x, k, y, ctx, cst_int<uint64_t>);
        default: return NDC_ERR_UNSUPPORTED;
    }
}

}  // namespace ndc
