/*
 * ndconv.h - C ABI of the MI355X (gfx950) convolution library.
 *
 * This is the drop-in boundary for ndarray-conv's convolution hot path: the
 * Rust crate's trait methods bind to these entry points through a thin FFI
 * shim.  Every declaration names the reference interface it replaces
 * (paths relative to the ndarray-conv 0.6.1 source tree).
 *
 * Conventions
 *  - Plain C: pointers, sizes, enums and POD structs only.
 *  - Element strides are in elements and may be negative; a data pointer
 *    always points at the first logical element (ndarray's as_ptr()).
 *  - Outputs are compact C-order arrays allocated by the caller; query the
 *    shape with ndc_plan().
 *  - By default x, k and y are device (HBM) pointers and the call is
 *    asynchronous on the handle's stream.  With NDC_FLAG_HOST_MEMORY they
 *    are host pointers and the call is synchronous.
 *  - A handle is not thread-safe; use one per thread, as the reference's
 *    &mut Processor is used.
 */
#ifndef NDCONV_H
#define NDCONV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ranks 1..6: the reference's trait impls need Dim<[Ix; N]>: RemoveAxis
 * (conv/mod.rs:118-127, conv_fft/mod.rs:198-201), which ndarray provides for
 * N = 1..6. */
#define NDC_MAX_RANK 6
#define NDC_ABI_VERSION 1

/* Error<N> (lib.rs:145-159); reference panics map to NDC_ERR_INVALID_ARGUMENT. */
typedef enum ndc_status {
    NDC_OK = 0,
    NDC_ERR_DATA_SHAPE = 1,       /* Error::DataShape      (lib.rs:151-152) */
    NDC_ERR_KERNEL_SHAPE = 2,     /* Error::KernelShape    (lib.rs:154-155) */
    NDC_ERR_MISMATCH_SHAPE = 3,   /* Error::MismatchShape  (lib.rs:157-158) */
    NDC_ERR_INVALID_ARGUMENT = 4, /* inputs the reference panics on: stride 0, reflect
                                     pads leaving the axis (half_dim.rs:200,247),
                                     FFT dilation 0; or a malformed descriptor */
    NDC_ERR_UNSUPPORTED = 5,      /* dtype/algorithm combination not provided */
    NDC_ERR_NO_DEVICE = 6,        /* no HIP device visible */
    NDC_ERR_DEVICE = 7,           /* a HIP runtime call failed */
    NDC_ERR_OUT_OF_MEMORY = 8
} ndc_status;

/* Element types: T of ConvExt (any NumAssign + Copy, conv/mod.rs:118-127) and
 * InElem of ConvFFTExt (f32/f64 or Complex<f32/f64>, conv_fft/mod.rs:81-87). */
typedef enum ndc_dtype {
    NDC_F32 = 0, NDC_F64 = 1, NDC_C64 = 2, NDC_C128 = 3,
    NDC_I8 = 4, NDC_I16 = 5, NDC_I32 = 6, NDC_I64 = 7,
    NDC_U8 = 8, NDC_U16 = 9, NDC_U32 = 10, NDC_U64 = 11
} ndc_dtype;

/* ConvMode<N> (lib.rs:80-105) */
typedef enum ndc_conv_mode_tag {
    NDC_CONV_FULL = 0, NDC_CONV_SAME = 1, NDC_CONV_VALID = 2,
    NDC_CONV_CUSTOM = 3, NDC_CONV_EXPLICIT = 4
} ndc_conv_mode_tag;

typedef struct ndc_conv_mode {
    int32_t tag;                        /* ndc_conv_mode_tag */
    int32_t reserved;
    int64_t padding[NDC_MAX_RANK][2];   /* Custom: [i][0] is used for both sides; Explicit: per side */
    int64_t strides[NDC_MAX_RANK];      /* Custom / Explicit only */
} ndc_conv_mode;

/* A scalar of the element type: re (and im for complex) for floats, i for integers
 * (two's complement bits for unsigned types). */
typedef struct ndc_value {
    double re;
    double im;
    int64_t i;
} ndc_value;

/* BorderType<T> (lib.rs:129-143) */
typedef enum ndc_border_tag {
    NDC_BORDER_ZEROS = 0, NDC_BORDER_CONST = 1, NDC_BORDER_REFLECT = 2,
    NDC_BORDER_REPLICATE = 3, NDC_BORDER_CIRCULAR = 4
} ndc_border_tag;

typedef struct ndc_border {
    int32_t tag;         /* ndc_border_tag */
    int32_t reserved;
    ndc_value value;     /* Const(c) */
} ndc_border;

/* PaddingMode<N, T> (lib.rs:106-127) */
typedef enum ndc_padding_tag {
    NDC_PAD_ZEROS = 0, NDC_PAD_CONST = 1, NDC_PAD_REFLECT = 2, NDC_PAD_REPLICATE = 3,
    NDC_PAD_CIRCULAR = 4, NDC_PAD_CUSTOM = 5, NDC_PAD_EXPLICIT = 6
} ndc_padding_tag;

typedef struct ndc_padding_mode {
    int32_t tag;                               /* ndc_padding_tag */
    int32_t reserved;
    ndc_value value;                           /* Const(c) */
    ndc_border borders[NDC_MAX_RANK][2];       /* Custom: [i][0]; Explicit: [i][front/back] */
} ndc_padding_mode;

/* Algorithm selection for ndc_conv. */
typedef enum ndc_algo {
    NDC_ALGO_AUTO = 0,          /* exact direct convolution in the element type */
    NDC_ALGO_DIRECT = 1,        /* same as AUTO */
    NDC_ALGO_MFMA_BF16 = 2,     /* f32 only, rank 2: one BF16 product per tap on the
                                   matrix cores, FP32 accumulation */
    NDC_ALGO_MFMA_BF16X3 = 3    /* f32 only, rank 2: split hi/lo BF16, three products per
                                   tap, FP32 accumulation (about 16 mantissa bits per product);
                                   kernels of fewer than 32 taps, whose outputs would keep that
                                   error, run the exact FP32 direct kernel instead */
} ndc_algo;

enum {
    NDC_FLAG_HOST_MEMORY = 1u,  /* x, k, y are host pointers; the call is synchronous */
    NDC_FLAG_X_STRIDED = 2u,    /* x_strides is valid (otherwise x is C-contiguous) */
    NDC_FLAG_K_STRIDED = 4u     /* k_strides is valid (otherwise k is C-contiguous) */
};

/* One convolution: the arguments of ConvExt::conv / ConvFFTExt::conv_fft.
 * kernel + dilation + reverse stand for KernelWithDilation (dilation/mod.rs:8-12):
 * `&kernel` is dilation 1 with reverse = 1 (dilation/mod.rs:196-203); no_reverse()
 * sets reverse = 0 (dilation/mod.rs:181-188). */
typedef struct ndc_conv_desc {
    int32_t dtype;                       /* ndc_dtype of x, k and y */
    int32_t rank;                        /* N, 1..NDC_MAX_RANK */
    int64_t x_shape[NDC_MAX_RANK];
    int64_t x_strides[NDC_MAX_RANK];
    int64_t k_shape[NDC_MAX_RANK];
    int64_t k_strides[NDC_MAX_RANK];
    int64_t dilation[NDC_MAX_RANK];
    int32_t reverse;
    int32_t algo;                        /* ndc_algo (ndc_conv only) */
    ndc_conv_mode conv_mode;
    ndc_padding_mode padding_mode;
    uint32_t flags;
    uint32_t reserved;
} ndc_conv_desc;

typedef enum ndc_op {
    NDC_OP_CONV = 0,        /* direct convolution (conv/mod.rs:128-200) */
    NDC_OP_CONV_FFT = 1     /* FFT convolution (conv_fft/mod.rs:185-292) */
} ndc_op;

/* Shape arithmetic of one convolution. */
typedef struct ndc_plan_info {
    int64_t padding[NDC_MAX_RANK][2];   /* ConvMode::unfold (conv/mod.rs:28-66) */
    int64_t strides[NDC_MAX_RANK];
    int64_t kernel_dilated[NDC_MAX_RANK];/* k*d - d + 1 (conv/mod.rs:146-147) */
    int64_t padded[NDC_MAX_RANK];       /* n + p0 + p1 (conv_fft/mod.rs:220-221) */
    int64_t fft_size[NDC_MAX_RANK];     /* good_size::compute (good_size.rs:40-42); 0 for NDC_OP_CONV */
    int64_t out_shape[NDC_MAX_RANK];    /* output shape (conv/mod.rs:162-167, conv_fft/mod.rs:282-289) */
} ndc_plan_info;

typedef struct ndc_handle_s* ndc_handle;

/* Library identification.  ndc_struct_size lets an FFI binding check its
 * struct layouts: 0 ndc_conv_desc, 1 ndc_plan_info, 2 ndc_conv_mode,
 * 3 ndc_padding_mode, 4 ndc_border, 5 ndc_value, 6 ndc_batch_desc; -1 for
 * anything else.
 * (No reference counterpart: FFI hygiene for the shims.) */
int32_t ndc_abi_version(void);
int64_t ndc_struct_size(int32_t which);
const char* ndc_status_string(ndc_status s);

/* Handle lifecycle: stands in for the FFT processor (get_fft_processor,
 * processor/mod.rs:71-73; GetProcessor::get_processor, processor/mod.rs:125-143):
 * it owns the HIP stream, the device workspace, twiddle tables and plan caches,
 * which is what conv_fft_with_processor reuses across calls (conv_fft/mod.rs:151-157). */
ndc_status ndc_create(ndc_handle* out);
ndc_status ndc_destroy(ndc_handle h);
/* Use an external HIP stream (hipStream_t; NULL is the null stream) for all
 * later calls; ndc_reset_stream returns to the handle's own stream.  A switch
 * orders the new stream after the work already queued on the old one (the
 * handle's workspace is shared by its calls), so the old stream must still be
 * valid when the switch is made. */
ndc_status ndc_set_stream(ndc_handle h, void* hip_stream);
ndc_status ndc_reset_stream(ndc_handle h);
ndc_status ndc_synchronize(ndc_handle h);
/* Detail message of the last failed call on this handle. */
const char* ndc_last_error(ndc_handle h);

/* Validates a descriptor and returns its shape arithmetic, with the
 * reference's error precedence (conv/mod.rs:136-158; conv_fft/mod.rs:205-227). */
ndc_status ndc_plan(const ndc_conv_desc* desc, int32_t op, ndc_plan_info* info);

/* Resolved border map of one axis: for each padded position (length padded[axis]
 * for NDC_OP_CONV, fft_size[axis] for NDC_OP_CONV_FFT) the source index into x
 * (>= 0) or a code: -1 never written (initial value), -2 the front side's
 * constant, -3 the back side's constant, -4 outside the padded region (zero).
 * This is the sequential per-axis fill of padding/mod.rs:119-153 and
 * padding/half_dim.rs:30-343 restated as an index map. */
ndc_status ndc_border_map(const ndc_conv_desc* desc, int32_t op, int32_t axis, int32_t* codes, int64_t len);

/* The kernel launches ndc_conv_fft makes for `desc`, planned on the host
 * without a device for one with `lds_bytes` of dynamic LDS per workgroup and
 * `cus` compute units: launch i requests smem[i] bytes of dynamic LDS with
 * grid[i] workgroups of block[i] threads.  *count receives the number of
 * launches; at most `cap` are written.  Returns ndc_conv_fft's status for a
 * plan it cannot launch (NDC_ERR_UNSUPPORTED when a launch would need more
 * LDS than lds_bytes).  (No reference counterpart: launch-plan introspection
 * for tests and bindings.) */
ndc_status ndc_conv_fft_launches(const ndc_conv_desc* desc, int64_t lds_bytes, int32_t cus, int64_t* smem,
                                 int64_t* grid, int64_t* block, int64_t cap, int64_t* count);

/* ConvExt::conv (conv/mod.rs:110-115, impl :128-200). */
ndc_status ndc_conv(ndc_handle h, const ndc_conv_desc* desc, const void* x, const void* k, void* y);

/* ConvFFTExt::conv_fft / conv_fft_with_processor / conv_fft_par
 * (conv_fft/mod.rs:113-171, impls :394-423).  The three differ only in CPU
 * threading and planner reuse, which the handle subsumes. */
ndc_status ndc_conv_fft(ndc_handle h, const ndc_conv_desc* desc, const void* x, const void* k, void* y);

/* ndc_conv_fft for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  (No
 * counterpart in the reference: its CPU path has no device memory to manage.)
 * Same arguments and results as ndc_conv_fft, device pointers only
 * (NDC_FLAG_HOST_MEMORY is NDC_ERR_INVALID_ARGUMENT); the workspace (spectra
 * and transform tables) comes from one library-owned 32 MiB device array per
 * GPU instead of allocations, so a convolution whose workspace bound exceeds
 * it returns NDC_ERR_UNSUPPORTED (ndc_last_error gives both sizes; about
 * 2 x 8 x (F0 ... F(N-2)) x (F(N-1)/2 + 8) bytes for f32, so 200x5000 fits and
 * 4096x4096 does not).  Calls that overlap on one GPU (other streams or host
 * threads) share the array and race. */
ndc_status ndc_conv_fft_noalloc(ndc_handle h, const ndc_conv_desc* desc, const void* x, const void* k, void* y);

/* A kernel spectrum kept for repeated convolutions of same-shaped data with
 * one kernel: the reference's sketched ConvFFTExt::conv_fft_bake /
 * conv_fft_with_baked and Baked (conv_fft/mod.rs:22-42,175-182,306-392,
 * commented out there).  ndc_conv_fft_bake validates `desc` exactly as
 * ndc_conv_fft does, transforms the kernel once and keeps its spectrum (plans
 * whose last axis is too long for one workgroup keep a copy of the kernel
 * instead).  ndc_conv_fft_with_baked convolves x, which has desc's shape,
 * strides and flags, and writes y; it equals ndc_conv_fft on the same
 * arguments.  The baked object uses h and must be destroyed before it. */
typedef struct ndc_baked_s* ndc_baked;
ndc_status ndc_conv_fft_bake(ndc_handle h, const ndc_conv_desc* desc, const void* k, ndc_baked* out);
ndc_status ndc_conv_fft_with_baked(ndc_baked b, const void* x, void* y);
ndc_status ndc_baked_destroy(ndc_baked b);

/* The FFT processor's plain transforms (Processor::forward / backward,
 * processor/mod.rs:79-121; real.rs:105-155,233-281; complex.rs:33-145) on a
 * C-contiguous array.  The handle is the processor.
 *  - dtype NDC_F32/NDC_F64: real input (forward) or real output (backward);
 *    NDC_C64/NDC_C128: complex both ways.
 *  - forward: x has `shape` [F0, ..., F(N-1)]; y receives the N-d DFT
 *    (unnormalised, e^{-2 pi i}) with its axes rotated left by one:
 *    shape [F1, ..., F(N-1)', F0], where F(N-1)' = F(N-1)/2 + 1 for real input.
 *    A real forward records F(N-1) in the handle (the reference's rp_origin_len).
 *  - backward: `spec_shape` is such a spectrum's shape; y receives the inverse,
 *    divided by the number of output elements, in natural order; a real
 *    backward restores the last axis to the recorded length (the imaginary parts
 *    of the DC and Nyquist bins are ignored, as realfft's result is used there).
 * Lines longer than one workgroup's LDS holds are split in two LDS-resident
 * factors; lengths with no such split (e.g. a prime above about 9000 points)
 * run as Bluestein's chirp-z transform, up to 2^23 points (longer ones return
 * NDC_ERR_UNSUPPORTED). */
ndc_status ndc_fft_forward(ndc_handle h, int32_t dtype, int32_t rank, const int64_t* shape, const void* x, void* y,
                           uint32_t flags);
ndc_status ndc_fft_backward(ndc_handle h, int32_t dtype, int32_t rank, const int64_t* spec_shape, const void* spec,
                            void* y, uint32_t flags);
ndc_status ndc_fft_origin_len(ndc_handle h, int64_t* len);

/* Per-kernel profile (no reference counterpart: the reference has no
 * instrumentation).  While enabled, every kernel launch on the handle carries
 * a pair of HIP events that take its dispatch's start and end times
 * (hipExtLaunchKernel; launches into a stream being captured are not
 * recorded).  The report synchronizes
 * the stream and writes a JSON object
 *   {"<kernel>": {"launches": n, "total_ms": t, "bytes": b, "flops": f}, ...}
 * where bytes/flops are the summed algorithmic HBM traffic and arithmetic of
 * those launches; then it clears the records. */
ndc_status ndc_profile_enable(ndc_handle h, int32_t on);
ndc_status ndc_profile_report(ndc_handle h, char* buf, int64_t len);

/* ---------------------------------------------------------------------------
 * Batches of independent rank-1 f32 signal/kernel pairs of uneven lengths.
 *
 * The reference has no batch call; a caller of it loops over the pairs and
 * calls ConvFFTExt::conv_fft (conv_fft/mod.rs:113-171) once per pair.  Each
 * pair here has exactly that call's semantics and errors (ConvMode, PaddingMode,
 * dilation and reverse shared by the batch), but the whole batch runs as one
 * persistent launch draining a device-side work queue.  Signals, kernels and
 * outputs are packed back to back in pair order (x of pair i starts at
 * sum_{j<i} x_len[j], likewise k and y); query the output lengths with
 * ndc_batch_out_lens().
 * ------------------------------------------------------------------------- */

typedef enum ndc_batch_algo {
    NDC_BATCH_AUTO = 0,     /* per pair, the cheaper of FP32 direct and FP32 overlap-save FFT
                             * by the device-measured cost model (the autotuner) */
    NDC_BATCH_FFT = 1,      /* overlap-save FFT for every pair whose dilated kernel fits
                             * the largest block (8192 samples), direct otherwise */
    NDC_BATCH_DIRECT = 2,   /* direct FP32 for every pair */
    NDC_BATCH_PLAIN = 3,    /* one single-problem conv_fft launch sequence per pair (the
                             * unbatched path, for comparison); run_graph runs it eagerly */
    NDC_BATCH_MFMA = 4,     /* direct on the matrix cores for every pair with stride 1 and a dilated
                             * kernel of at most 1024 taps, FP32 accumulation: split BF16x3 when the
                             * kernel has at least 32 taps, split BF16x6 below; the others as
                             * NDC_BATCH_AUTO */
    NDC_BATCH_AUTO_BF16X3 = 5, /* the autotuner's choice among FP32 direct, FP32 FFT and the matrix-core
                               * forms.  BF16x3 (hi/lo split, three products) carries about 16
                               * significant bits per product, so it is admitted only for kernels of
                               * at least 32 taps, where the errors of the many products per output
                               * average out; shorter kernels may take BF16x6 (three-term split, six
                               * products, within 5 * 2^-24 per product) */
    NDC_BATCH_BF16X6 = 6    /* split BF16x6 on the matrix cores for every pair NDC_BATCH_MFMA takes;
                             * the others as NDC_BATCH_AUTO */
} ndc_batch_algo;

typedef struct ndc_batch_desc {
    int32_t dtype;                  /* NDC_F32 */
    int32_t algo;                   /* ndc_batch_algo */
    int64_t count;                  /* number of pairs */
    const int64_t* x_len;           /* [count] signal lengths */
    const int64_t* k_len;           /* [count] kernel lengths */
    int64_t dilation;               /* KernelWithDilation (dilation/mod.rs:64-80) */
    int32_t reverse;                /* 1 convolution, 0 no_reverse (dilation/mod.rs:117-189) */
    uint32_t flags;                 /* NDC_FLAG_HOST_MEMORY */
    ndc_conv_mode conv_mode;        /* rank-1 ConvMode, shared */
    ndc_padding_mode padding_mode;  /* rank-1 PaddingMode, shared */
} ndc_batch_desc;

typedef struct ndc_batch_s* ndc_batch;

/* Validates every pair (first failing pair: its status, and its index in
 * ndc_last_error), plans the work queue and uploads it.  AUTO measures the
 * cost model on the device the first time a handle needs it.  The batch
 * keeps using h (stream, profile) and must be destroyed before it. */
ndc_status ndc_batch_create(ndc_handle h, const ndc_batch_desc* desc, ndc_batch* out);
ndc_status ndc_batch_destroy(ndc_batch b);
/* y_len[count] (may be NULL) and the packed totals (each may be NULL). */
ndc_status ndc_batch_out_lens(ndc_batch b, int64_t* y_len, int64_t* x_total, int64_t* k_total, int64_t* y_total);
/* Enqueues the batch on the handle's stream (device pointers; with
 * NDC_FLAG_HOST_MEMORY host pointers and synchronous).  No allocation and no
 * host synchronisation in device mode, so it may be captured in a graph. */
ndc_status ndc_batch_run(ndc_batch b, const void* x, const void* k, void* y);
/* Same result through a HIP graph of the batch's launches, captured on the
 * first call and replayed while x, k, y stay the same (re-captured when they
 * change). */
ndc_status ndc_batch_run_graph(ndc_batch b, const void* x, const void* k, void* y);
/* JSON: the measured cost constants, the choice per kernel-length and
 * signal-length bucket, item counts and algorithmic bytes per launch. */
ndc_status ndc_batch_report(ndc_batch b, char* buf, int64_t len);
/* Border map of pair `pair` over its padded length (codes as ndc_border_map
 * with NDC_OP_CONV) from the closed form the batch kernels evaluate per
 * sample. */
ndc_status ndc_batch_border_map(const ndc_batch_desc* desc, int64_t pair, int32_t* codes, int64_t len);

#ifdef __cplusplus
}
#endif

#endif /* NDCONV_H */
