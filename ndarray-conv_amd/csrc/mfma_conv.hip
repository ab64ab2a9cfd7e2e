// Direct 2-D convolution of f32 data on the gfx950 matrix cores, with BF16
// products (one, or three for the split hi/lo "BF16x3" form) and FP32
// accumulation.  Same arithmetic contract as conv/mod.rs:177-197 (sum over the
// taps of the reversed kernel), different rounding.
//
// A single-channel convolution has no channel dimension to make a GEMM of, so
// the matrix product is taken over kernel ROWS:
//   for one input row rho, P[m][r] = sum_c P(rho, (x0+m)*sx + c*dx) * kv(r, c)
// is a 32x32 tile (m: 32 outputs along x, r: 32 kernel rows, c: kernel
// columns as the K dimension), and P[.][r] belongs to output row
// y = (rho - r*dy) / sy.  Streaming rho = (y0 + t)*sy, the partial sums
// travel one lane (one kernel row) per step: the accumulator is shifted up by
// one lane (DPP wave_shr:1) and fed back as the MFMA's C operand, so the lane
// of the last kernel row holds a finished output row every step.  This needs
// sy to divide dy; Q = dy/sy accumulator sets rotate when dy > sy.
//
// MFMA lane layout (v_mfma_f32_32x32x16_bf16): A lane l holds A[l%32][8*(l/32)+j],
// B lane l holds B[8*(l/32)+j][l%32], D lane l holds D[8*(i/4)+4*(l/32)+i%4][l%32].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "bf16_split.hpp"
#include "device_common.hpp"
#include "launch.hpp"

namespace ndc {
namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kXT = 1;               // 32-output x tiles per wave (independent chains)
constexpr int kXB = 32 * kWaves * kXT;   // outputs along x per workgroup
constexpr int kMaxW = 256;           // staged row width limit (entries)
constexpr int kMaxSR = 16;            // input rows staged per batch, at most
constexpr int kPer = (kMaxSR * kMaxW + kThreads - 1) / kThreads;   // staged entries per thread per batch
constexpr int kMaxChain = 1024;      // chain rows per pass (BY + 31 * Q)
constexpr int kMfmaX3MinTaps = 32;   // fewest kernel taps BF16x3 is used for

// input rows staged per batch: a multiple of Q keeps the accumulator-set
// rotation aligned with the batches
template <int Q> constexpr int rows_per_batch() { return Q == 3 ? 15 : 16; }

// Entries of one staged BF16 plane (stride-1 columns): sr * w rounded up to
// 16 more than a multiple of 32.  A K-block read takes its even lanes from
// plane 0 and its odd lanes from the plane shifted by one element, `plane`
// 32-bit words further on; each set is 16 consecutive words, so with the
// planes 16 banks apart the two sets never share a bank.
constexpr int bf16_plane(int sr, int w) { return (sr * w + 15) / 32 * 32 + 16; }

struct MfmaArgs {
    const float* x;
    const float* k;
    float* y;
    const int32_t* map[2];   // border maps of the row and column axes
    float cst[2][2];
    int64_t xstr[2];
    int64_t kstr[2];
    int32_t OY, OX;          // output extents
    int32_t KH, KW;          // kernel extents
    int32_t sy, sx, dy, dx;
    int32_t pdy, pdx;        // padded extents (map lengths)
    int32_t rev;
    int32_t BY;              // output rows per workgroup chain
    int32_t ntx;             // x tiles
    int32_t gx;              // gcd(sx, dx): staged columns are subsampled by it
    int32_t W;               // staged row width (u32 entries)
    int32_t KC;              // kernel columns per pass
    int32_t npass_c;         // column passes
    int32_t npass_r;         // row groups of 32
};

// Source of the padded input from the border codes of its row and column
// (codes resolved once per pass into LDS): true with the offset into x, or
// false with the cell's value in *fill.  No memory access, so that a batch's
// loads can all be issued before any of them is used.
__device__ __forceinline__ bool pad_source(const MfmaArgs& a, int32_t cy, int32_t cx, int64_t* off, float* fill) {
    *off = 0;
    *fill = 0.0f;
    // the higher axis' constant wins (later fills overwrite)
    if (cx == kConstFront || cx == kConstBack) {
        *fill = a.cst[1][cx == kConstFront ? 0 : 1];
        return false;
    }
    if (cy == kConstFront || cy == kConstBack) {
        *fill = a.cst[0][cy == kConstFront ? 0 : 1];
        return false;
    }
    if (cx < 0 || cy < 0) return false;   // unwritten, or beyond the padded extent
    *off = int64_t(cy) * a.xstr[0] + int64_t(cx) * a.xstr[1];
    return true;
}

// True when v and its BF16 split stay finite (NaN and infinities fail, and
// magnitudes whose BF16 rounding would overflow).
__device__ __forceinline__ bool finite_bf16(float v) { return __builtin_fabsf(v) <= 3.38e38f; }

// D1: the subsampled column dilation is 1 (stride equal to dilation, or both
// 1), so a K-block's 8 operand columns are consecutive staged entries and
// their reads take fixed offsets from one address.
//
// Each wave runs kXT chains side by side (32 adjacent outputs each, tile xt
// of wave w at x0 + 32 (w + kWaves xt)) sharing the B operands.  One tile per
// wave is the measured best on the 4096x4096 BF16x3 case (76.5 us; two tiles
// per wave took 117-136 us: wider workgroups halve the grid, and shorter
// chains to restore it add fill steps).  Kernel columns beyond the pass hold
// zero B entries, so the A operands need no masks (their staged columns exist).
template <int NKB, int Q, bool X3, bool D1>
__global__ void __launch_bounds__(kThreads, 2) mfma_conv_kernel(MfmaArgs a) {
    constexpr int kSR = rows_per_batch<Q>();
    // D1: per buffer four BF16 planes of kSR x W (W even): hi, lo, and both
    // shifted by one element, so a K-block's 8 consecutive columns are 4
    // aligned 32-bit reads in operand order from one of them; otherwise
    // [2][kSR][W] 32-bit entries packing (hi, lo)
    extern __shared__ uint32_t rows_lds[];
    __shared__ int32_t s_row[kMaxChain];     // border code of chain row t
    __shared__ int32_t s_col[kMaxW];         // border code of staged column j
    __shared__ int32_t s_exact;              // a staged value or tap is not finite: recompute exactly
    if (threadIdx.x == 0) s_exact = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int half = lane >> 5;
    const int m = lane & 31;
    const int tx = blockIdx.x % a.ntx;
    const int ty = blockIdx.x / a.ntx;
    const int x0 = tx * kXB;
    const int y0 = ty * a.BY;
    const int sxg = a.sx / a.gx, dxg = D1 ? 1 : a.dx / a.gx;
    const int BYe = min(a.BY, a.OY - y0);

    for (int pr = 0; pr < a.npass_r; ++pr) {
        const int r0 = 32 * pr;
        const int KHg = min(32, a.KH - r0);
        for (int pc = 0; pc < a.npass_c; ++pc) {
            const int c0 = pc * a.KC;
            const int KCg = min(a.KC, a.KW - c0);
            const bool first = pr == 0 && pc == 0;
            // B operands: kernel rows r0 + (l % 32), columns c0 + 16 kb + 8 half + j
            u32x4 bh[NKB], bl[NKB];
            bool kbad = false;
#pragma unroll
            for (int kb = 0; kb < NKB; ++kb) {
                uint32_t p[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int r = r0 + m;
                    const int c = 16 * kb + 8 * half + j;
                    float v = 0.0f;
                    if (r < a.KH && c < KCg) {
                        const int ki = a.rev ? a.KH - 1 - r : r;
                        const int kj = a.rev ? a.KW - 1 - (c0 + c) : c0 + c;
                        v = a.k[int64_t(ki) * a.kstr[0] + int64_t(kj) * a.kstr[1]];
                    }
                    kbad |= !finite_bf16(v);
                    p[j] = split_pack(v);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    bh[kb][i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x05040100u);
                    bl[kb][i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x07060302u);
                }
            }
            if (kbad) s_exact = 1;
            floatx16 acc[Q][kXT];
#pragma unroll
            for (int s = 0; s < Q; ++s)
#pragma unroll
                for (int xt = 0; xt < kXT; ++xt)
#pragma unroll
                    for (int i = 0; i < 16; ++i) acc[s][xt][i] = 0.0f;

            const int nsteps = BYe + (KHg - 1) * Q;
            const int nbatch = (nsteps + kSR - 1) / kSR;
            // staging: row t of the chain is input row (y0 + t) * sy + r0 * dy,
            // columns x0*sx + c0*dx + j*gx, j < W
            const int64_t pcol0 = int64_t(x0) * a.sx + int64_t(c0) * a.dx;
            __syncthreads();   // the previous pass is done with s_row / s_col
            for (int t = threadIdx.x; t < nsteps; t += kThreads)
                s_row[t] = a.map[0][int64_t(y0 + t) * a.sy + int64_t(r0) * a.dy];
            for (int j = threadIdx.x; j < a.W; j += kThreads) {
                const int64_t px = pcol0 + int64_t(j) * a.gx;
                s_col[j] = px < a.pdx ? a.map[1][px] : int32_t(kOutside);
            }
            __syncthreads();
            const float rW = 1.0f / float(a.W);
            // Staging of a batch in two halves: fetch() computes every source
            // and issues all loads without using them; the values are selected,
            // split and written to LDS by store() after the next batch's
            // compute, so the loads overlap it instead of waiting one by one.
            float raw[kPer], fill[kPer];
            uint32_t okm = 0;
            auto fetch = [&](int b) {
                int64_t off[kPer];
                okm = 0;
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + i * kThreads;
                    off[i] = 0;
                    fill[i] = 0.0f;
                    int rr = int(float(e) * rW);
                    int cc = e - rr * a.W;
                    if (cc < 0) { --rr; cc += a.W; } else if (cc >= a.W) { ++rr; cc -= a.W; }
                    const int t = b * kSR + rr;
                    if (e < kSR * a.W && t < nsteps && pad_source(a, s_row[t], s_col[cc], &off[i], &fill[i]))
                        okm |= 1u << i;
                }
#pragma unroll
                for (int i = 0; i < kPer; ++i) raw[i] = a.x[off[i]];   // offset 0 (a valid element) where not used
            };
            const int plane = D1 ? bf16_plane(kSR, a.W) : kSR * a.W;   // entries of one BF16 plane (D1)
            auto store = [&](uint32_t* dst) {
                uint16_t* h16 = reinterpret_cast<uint16_t*>(dst);
                bool bad = false;
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + i * kThreads;
                    if (e < kSR * a.W) {
                        const float f = (okm >> i & 1u) ? raw[i] : fill[i];
                        bad |= !finite_bf16(f);
                        const uint32_t v = split_pack(f);
                        if constexpr (D1) {
                            // hi at plane 0, lo at plane 1; the shifted planes 2, 3
                            // hold entry e at e - 1
                            h16[e] = uint16_t(v);
                            if (e > 0) h16[2 * plane + e - 1] = uint16_t(v);
                            if (X3) {
                                h16[plane + e] = uint16_t(v >> 16);
                                if (e > 0) h16[3 * plane + e - 1] = uint16_t(v >> 16);
                            }
                        } else {
                            dst[e] = v;
                        }
                    }
                }
                if (bad) s_exact = 1;
            };
            const int bufw = D1 ? 2 * plane : plane;   // 32-bit words per buffer
            // host keeps W <= kMaxW, so kSR * W <= kPer * kThreads
            fetch(0);
            store(rows_lds);
            __syncthreads();
            int t = 0;
            for (int b = 0; b < nbatch; ++b) {
                if (b + 1 < nbatch) fetch(b + 1);
                const uint32_t* buf = rows_lds + (b & 1) * bufw;
                const int tend = min(nsteps, (b + 1) * kSR);
                for (; t < tend;) {
#pragma unroll
                    for (int s = 0; s < Q; ++s, ++t) {
                        if (t >= tend) break;
                        const int rowoff = (t - b * kSR) * a.W;
                        const uint32_t* row = buf + rowoff;
                        // Hankel A operands from the staged row, per tile
                        u32x4 ah[kXT][NKB], al[kXT][NKB];
#pragma unroll
                        for (int xt = 0; xt < kXT; ++xt) {
#pragma unroll
                            for (int kb = 0; kb < NKB; ++kb) {
                                const int base = (32 * (wave + kWaves * xt) + m) * sxg + (16 * kb + 8 * half) * dxg;
                                // every staged column of the K-block exists (W covers
                                // 16 * NKB kernel columns): the reads are unconditional
                                if constexpr (D1) {
                                    // columns base..base+7 of the hi / lo plane, from the
                                    // shifted planes when base is odd: even half-word
                                    // offsets, so 32-bit reads of (column 2i, 2i+1) pairs
                                    const int odd = base & 1;
                                    const uint16_t* b16 = reinterpret_cast<const uint16_t*>(buf);
                                    const int oh = (odd ? 2 * plane : 0) + rowoff + base - odd;
                                    const uint32_t* ph = reinterpret_cast<const uint32_t*>(b16 + oh);
                                    const uint32_t* pl = reinterpret_cast<const uint32_t*>(b16 + oh + plane);
#pragma unroll
                                    for (int i = 0; i < 4; ++i) {
                                        ah[xt][kb][i] = ph[i];
                                        al[xt][kb][i] = X3 ? pl[i] : 0u;
                                    }
                                } else {
                                    uint32_t p[8];
#pragma unroll
                                    for (int j = 0; j < 8; ++j) p[j] = row[base + j * dxg];
#pragma unroll
                                    for (int i = 0; i < 4; ++i) {
                                        ah[xt][kb][i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x05040100u);
                                        al[xt][kb][i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x07060302u);
                                    }
                                }
                            }
                        }
                        // shift the accumulators one kernel row up; row 0 starts fresh
                        floatx16 c[kXT];
#pragma unroll
                        for (int xt = 0; xt < kXT; ++xt) {
#pragma unroll
                            for (int i = 0; i < 16; ++i) {
                                // copy the element out first: a bit cast of the vector
                                // element lvalue reads element 0
                                const float ai = acc[s][xt][i];
                                const int v = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, ai), 0x138, 0xf,
                                                                          0xf, true);
                                c[xt][i] = m == 0 ? 0.0f : __builtin_bit_cast(float, v);
                            }
                        }
                        // products of the tiles interleaved: consecutive matrix
                        // instructions never depend on each other
#pragma unroll
                        for (int kb = 0; kb < NKB; ++kb) {
#pragma unroll
                            for (int xt = 0; xt < kXT; ++xt)
                                c[xt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                    __builtin_bit_cast(bf16x8, ah[xt][kb]), __builtin_bit_cast(bf16x8, bh[kb]), c[xt], 0,
                                    0, 0);
                            if (X3) {
#pragma unroll
                                for (int xt = 0; xt < kXT; ++xt)
                                    c[xt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                        __builtin_bit_cast(bf16x8, ah[xt][kb]), __builtin_bit_cast(bf16x8, bl[kb]),
                                        c[xt], 0, 0, 0);
#pragma unroll
                                for (int xt = 0; xt < kXT; ++xt)
                                    c[xt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                        __builtin_bit_cast(bf16x8, al[xt][kb]), __builtin_bit_cast(bf16x8, bh[kb]),
                                        c[xt], 0, 0, 0);
                            }
                        }
#pragma unroll
                        for (int xt = 0; xt < kXT; ++xt) acc[s][xt] = c[xt];
                        // the lane of the last kernel row holds output row y; the
                        // common case (first pass, tile inside the row) is plain
                        // stores at fixed offsets
                        const int yo = y0 + t - (KHg - 1) * Q;
                        if (yo >= y0 && yo < y0 + BYe && m == KHg - 1) {
#pragma unroll
                            for (int xt = 0; xt < kXT; ++xt) {
                                const int xw = x0 + 32 * (wave + kWaves * xt);
                                float* yr = a.y + int64_t(yo) * a.OX + xw + 4 * half;
                                if (first && xw + 32 <= a.OX) {
#pragma unroll
                                    for (int i = 0; i < 16; ++i) yr[8 * (i / 4) + (i % 4)] = c[xt][i];
                                } else {
#pragma unroll
                                    for (int i = 0; i < 16; ++i) {
                                        const int xo = xw + 8 * (i / 4) + 4 * half + (i % 4);
                                        if (xo < a.OX) {
                                            float* yp = yr + 8 * (i / 4) + (i % 4);
                                            *yp = first ? c[xt][i] : *yp + c[xt][i];
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
                if (b + 1 < nbatch) store(rows_lds + ((b + 1) & 1) * bufw);
                __syncthreads();
            }
        }
    }
    // A non-finite value makes a matrix product non-finite even where the
    // reference's sum skips it (zero taps, dilation gaps, neighbouring outputs
    // of a tile).  Such workgroups recompute their outputs as the reference
    // does: FP32, the view's taps in row-major order, zero taps skipped,
    // separate multiply and add (conv/mod.rs:177-197, dilation/mod.rs:34-60).
    __syncthreads();
    if (s_exact) {
        const int nout = BYe * kXB;
        for (int e = threadIdx.x; e < nout; e += kThreads) {
            const int yo = y0 + e / kXB, xo = x0 + e % kXB;
            if (xo >= a.OX) continue;
            float acc = 0.0f;
            for (int r = 0; r < a.KH; ++r) {
                const int32_t cy = a.map[0][int64_t(yo) * a.sy + int64_t(r) * a.dy];
                const int ki = a.rev ? a.KH - 1 - r : r;
                for (int c = 0; c < a.KW; ++c) {
                    const int kj = a.rev ? a.KW - 1 - c : c;
                    const float kv = a.k[int64_t(ki) * a.kstr[0] + int64_t(kj) * a.kstr[1]];
                    if (kv == 0.0f) continue;
                    const int32_t cx = a.map[1][int64_t(xo) * a.sx + int64_t(c) * a.dx];
                    int64_t off;
                    float fill;
                    const float v = pad_source(a, cy, cx, &off, &fill) ? a.x[off] : fill;
                    acc = __fadd_rn(acc, __fmul_rn(v, kv));
                }
            }
            a.y[int64_t(yo) * a.OX + xo] = acc;
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

template <int NKB, int Q, bool X3, bool D1>
const void* kernel_ptr() {
    return reinterpret_cast<const void*>(&mfma_conv_kernel<NKB, Q, X3, D1>);
}

template <int NKB, int Q, bool X3, bool D1>
void launch_one(const MfmaArgs& a, unsigned blocks, size_t smem, LaunchContext& ctx) {
    NDC_LAUNCH(ctx, ctx.stream(), (mfma_conv_kernel<NKB, Q, X3, D1>), dim3(blocks), dim3(kThreads), smem, a);
}

template <bool X3>
const void* pick(int nkb, int q, bool d1, void (**fn)(const MfmaArgs&, unsigned, size_t, LaunchContext&)) {
#define NDC_MFMA_CASE(NK, QQ, DD)                        \
    if (nkb == NK && q == QQ && d1 == DD) {              \
        *fn = &launch_one<NK, QQ, X3, DD>;               \
        return kernel_ptr<NK, QQ, X3, DD>();             \
    }
    // consecutive operand columns only for the single-set chain (row stride
    // equal to row dilation), which covers the common shapes
    NDC_MFMA_CASE(1, 1, true) NDC_MFMA_CASE(2, 1, true) NDC_MFMA_CASE(3, 1, true) NDC_MFMA_CASE(4, 1, true)
    if (d1) d1 = false;
    NDC_MFMA_CASE(1, 1, false) NDC_MFMA_CASE(2, 1, false) NDC_MFMA_CASE(3, 1, false) NDC_MFMA_CASE(4, 1, false)
    NDC_MFMA_CASE(1, 2, false) NDC_MFMA_CASE(2, 2, false) NDC_MFMA_CASE(3, 2, false) NDC_MFMA_CASE(4, 2, false)
    NDC_MFMA_CASE(1, 3, false) NDC_MFMA_CASE(2, 3, false) NDC_MFMA_CASE(3, 3, false) NDC_MFMA_CASE(4, 3, false)
    NDC_MFMA_CASE(1, 4, false) NDC_MFMA_CASE(2, 4, false) NDC_MFMA_CASE(3, 4, false) NDC_MFMA_CASE(4, 4, false)
#undef NDC_MFMA_CASE
    return nullptr;
}

}  // namespace

ndc_status launch_direct_mfma(const Plan& p, const void* x, const void* k, void* y, LaunchContext& ctx) {
    if (p.dtype != NDC_F32 || p.N != 2) {
        ctx.err = "the BF16 matrix-core forms take rank-2 f32 convolutions";
        return NDC_ERR_UNSUPPORTED;
    }
    const int64_t sy = p.stride[0], sx = p.stride[1], dy = p.dil[0], dx = p.dil[1];
    if (dy < 1 || dx < 1 || dy % sy != 0 || dy / sy > 4) {
        ctx.err = "the BF16 matrix-core forms need dilation >= 1 and the row stride dividing the row dilation "
                  "(at most 4x)";
        return NDC_ERR_UNSUPPORTED;
    }
    if (p.out[0] > (int64_t(1) << 30) || p.out[1] > (int64_t(1) << 30)) return NDC_ERR_UNSUPPORTED;
    // BF16x3 products keep about 16 bits (each within 3 * 2^-16 of the exact
    // product); an output summed from fewer than 32 of them keeps that error
    // above the f32 tolerance, and the exact FP32 kernel costs little at that
    // size, so such kernels run it instead (shapes the matrix-core forms
    // refuse are still refused).
    const bool x3 = p.algo == NDC_ALGO_MFMA_BF16X3;
    if (x3 && p.k_shape[0] * p.k_shape[1] < kMfmaX3MinTaps) {
        const int64_t g = gcd64(sx, dx);
        if (((kXB - 1) * sx + 15 * dx) / g + 1 <= kMaxW) return launch_direct(p, x, k, y, ctx);
    }
    int64_t offs[kMaxRank + 1];
    const int32_t* maps = ctx.border_maps(p, offs);
    if (!maps) return NDC_ERR_OUT_OF_MEMORY;
    MfmaArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = static_cast<const float*>(x);
    a.k = static_cast<const float*>(k);
    a.y = static_cast<float*>(y);
    for (int i = 0; i < 2; ++i) {
        a.map[i] = maps + offs[i];
        a.cst[i][0] = float(p.side[i][0].value.re);
        a.cst[i][1] = float(p.side[i][1].value.re);
        a.xstr[i] = p.x_str[i];
        a.kstr[i] = p.k_str[i];
    }
    a.OY = int32_t(p.out[0]);
    a.OX = int32_t(p.out[1]);
    a.KH = int32_t(p.k_shape[0]);
    a.KW = int32_t(p.k_shape[1]);
    a.sy = int32_t(sy);
    a.sx = int32_t(sx);
    a.dy = int32_t(dy);
    a.dx = int32_t(dx);
    a.pdy = int32_t(p.pds[0]);
    a.pdx = int32_t(p.pds[1]);
    a.rev = p.reverse ? 1 : 0;
    const int Q = int(dy / sy);
    a.gx = int32_t(gcd64(sx, dx));
    // kernel columns per pass: up to 64 (four K-blocks of 16), fewer when the
    // staged row (128 outputs plus the dilated column span) would not fit
    int nkb = 0;
    for (int cand = 4; cand >= 1; --cand) {
        const int64_t w = ((kXB - 1) * sx + (16 * cand - 1) * dx) / a.gx + 1;
        if (w <= kMaxW) {
            nkb = cand;
            a.W = int32_t(w);
            break;
        }
    }
    if (!nkb) {
        ctx.err = "stride/dilation too large for the staged rows of the matrix-core form";
        return NDC_ERR_UNSUPPORTED;
    }
    nkb = std::min(nkb, int((a.KW + 15) / 16));
    a.W = int32_t(((kXB - 1) * sx + (16 * nkb - 1) * dx) / a.gx + 1);
    const bool d1 = dx / a.gx == 1;
    if (d1) a.W += a.W & 1;   // the BF16 planes keep rows 4-byte aligned (W <= kMaxW stays: kMaxW is even)
    a.KC = std::min<int32_t>(a.KW, 16 * nkb);
    a.npass_c = (a.KW + a.KC - 1) / a.KC;
    a.npass_r = (a.KH + 31) / 32;
    const int SR = Q == 3 ? rows_per_batch<3>() : rows_per_batch<1>();
    // chain length: long chains amortise the (KH-1)*Q fill steps, short ones
    // give more workgroups; aim for two workgroups per CU
    a.ntx = int32_t((a.OX + kXB - 1) / kXB);
    const int64_t target = 2 * int64_t(ctx.cu_count());
    int64_t by = (int64_t(a.OY) * a.ntx + target - 1) / target;
    const int64_t min_by = std::max<int64_t>(16, 2 * (std::min(a.KH, 32) - 1) * Q);
    by = std::max(by, min_by);
    by = std::min<int64_t>(by, a.OY);
    by = std::min<int64_t>(by, kMaxChain - 31 * Q);   // the chain's row codes live in LDS
    a.BY = int32_t(by);
    const int64_t nty = (a.OY + a.BY - 1) / a.BY;
    const int64_t blocks = nty * a.ntx;
    // two buffers of SR x W staged entries: four 2-byte BF16 planes of
    // bf16_plane(SR, W) entries (D1) or one 32-bit (hi, lo) word per entry
    const size_t smem = d1 ? size_t(2) * 4 * bf16_plane(SR, a.W) * sizeof(uint16_t)
                           : size_t(2) * SR * a.W * sizeof(uint32_t);
    void (*fn)(const MfmaArgs&, unsigned, size_t, LaunchContext&) = nullptr;
    const void* kp = x3 ? pick<true>(nkb, Q, d1, &fn) : pick<false>(nkb, Q, d1, &fn);
    if (!kp || !fn) return NDC_ERR_UNSUPPORTED;
    if (!lds_fits(ctx, kp, smem)) return NDC_ERR_UNSUPPORTED;
    // algorithmic work: 2 flops per output and tap (the FP32-equivalent count)
    ctx.kbegin(x3 ? "mfma_conv_bf16x3" : "mfma_conv_bf16",
               4.0 * double(p.x_numel + p.k_numel + p.y_numel), 2.0 * double(p.y_numel) * double(p.k_numel));
    fn(a, unsigned(blocks), smem, ctx);
    ctx.kend();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx.err = std::string("matrix-core convolution launch: ") + hipGetErrorString(e);
        return NDC_ERR_DEVICE;
    }
    return NDC_OK;
}

}  // namespace ndc
