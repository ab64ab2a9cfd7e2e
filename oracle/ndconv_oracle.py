"""CPU restatement of ndarray-conv's convolution hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product (``ndarray-conv_amd``) never imports, links or runs
anything under ``oracle/``.

Every function restates one piece of the reference crate (ndarray-conv 0.6.1)
and cites the file:line it follows.  How each one is pinned (by the
reference's own golden vectors, by torch, or by mathematics) is listed here:

  unfold            conv/mod.rs:28-66     pinned: torch conv cases (conv/tests.rs) and
                                          padding goldens that go through unfold
  good_size         good_size.rs:6-31     pinned: invariants (F >= n, F <= next_pow2(n),
                                          exact on 2^a and 3*2^a) and the independently
                                          derived non-smooth band table; no reference test
                                          exists for it, so the bands are the strongest pin
  padding           padding/mod.rs:84-117, half_dim.rs
                                          pinned: every golden array in padding/mod.rs:464-759
                                          and torch F.pad cases (padding/mod.rs:763-861)
  padding_data      conv_fft/padding.rs:30-62
                                          pinned: golden conv_fft/padding.rs:124-153
  padding_kernel    conv_fft/padding.rs:78-111
                                          pinned: golden conv_fft/padding.rs:156-181
  gen_offset_list   dilation/mod.rs:34-60 pinned: goldens dilation/mod.rs:277-404
  conv              conv/mod.rs:128-200   pinned: torch conv1d/2d/3d cases (conv/tests.rs:36-547),
                                          goldens conv/tests.rs:552-584, int64 brute force
  conv_fft          conv_fft/mod.rs:185-292
                                          pinned: FFT == direct after rounding
                                          (conv_fft/tests.rs:148-619) and a brute-force
                                          circular-convolution DFT on tiny inputs.  Complex
                                          results are pinned by the direct complex sum.
  processor_forward / processor_backward  processor/real.rs:105-155,233-281, complex.rs:33-145
                                          pinned: the literal permute-and-transform loop
                                          (processor_forward_procedural), a brute-force DFT
                                          on tiny inputs, and the reference's roundtrip cases
                                          (real.rs:289-478, complex.rs:154-480)

Representations (kept independent of the product's own API objects):
  conv mode    "full" | "same" | "valid" | ("custom", [p]*N, [s]*N)
               | ("explicit", [[p0, p1]]*N, [s]*N)
  border       "zeros" | ("const", c) | "reflect" | "replicate" | "circular"
  padding mode "zeros" | ("const", c) | "reflect" | "replicate" | "circular"
               | ("custom", [border]*N) | ("explicit", [[border, border]]*N)
"""

import numpy as np

DATA_SHAPE = "DataShape"
KERNEL_SHAPE = "KernelShape"
MISMATCH_SHAPE = "MismatchShape"
INVALID_ARGUMENT = "InvalidArgument"   # inputs on which the reference panics


class OracleError(Exception):
    def __init__(self, kind, detail=""):
        super().__init__(f"{kind}: {detail}")
        self.kind = kind


# ----------------------------------------------------------------------------
# ConvMode::unfold (conv/mod.rs:28-66)
# ----------------------------------------------------------------------------

def kernel_dilated_shape(kshape, dilation):
    """k + (k - 1) * (d - 1) per axis (conv/mod.rs:34-37, conv_fft/mod.rs:215-216)."""
    return [k * d - d + 1 for k, d in zip(kshape, dilation)]


def unfold(mode, kshape, dilation):
    """Return (padding [[front, back]]*N, strides [s]*N) (conv/mod.rs:28-66)."""
    n = len(kshape)
    kd = kernel_dilated_shape(kshape, dilation)
    if mode == "full":
        return [[k - 1, k - 1] for k in kd], [1] * n
    if mode == "same":
        pads = []
        for k in kd:
            if k % 2 == 0:
                pads.append([(k - 1) // 2 + 1, (k - 1) // 2])
            else:
                pads.append([(k - 1) // 2, (k - 1) // 2])
        return pads, [1] * n
    if mode == "valid":
        return [[0, 0] for _ in range(n)], [1] * n
    tag = mode[0]
    if tag == "custom":
        return [[int(p), int(p)] for p in mode[1]], [int(s) for s in mode[2]]
    if tag == "explicit":
        return [[int(p[0]), int(p[1])] for p in mode[1]], [int(s) for s in mode[2]]
    raise ValueError(mode)


# ----------------------------------------------------------------------------
# good_size (conv_fft/good_size.rs:6-42)
# ----------------------------------------------------------------------------

def _next_power_of_two(n):
    return 1 if n <= 1 else 1 << (n - 1).bit_length()


def good_size(n):
    """good_size_cc (good_size.rs:6-31): next power of two, then shrink by
    x3/4 while still >= n, then by x5/6 while still >= n, with the integer
    divisions truncating exactly as the reference's usize arithmetic."""
    best = _next_power_of_two(n)
    while True:
        new = best // 4 * 3
        if new < n:
            break
        if new == n:
            return n
        best = new
    while True:
        new = best // 6 * 5
        if new < n:
            break
        if new == n:
            return n
        best = new
    return best


def fft_size(shape):
    """good_size::compute (good_size.rs:40-42)."""
    return [good_size(int(s)) for s in shape]


# ----------------------------------------------------------------------------
# Padding (padding/mod.rs, padding/dim.rs, padding/half_dim.rs)
# ----------------------------------------------------------------------------

def _plane(buf, axis, j):
    idx = [slice(None)] * buf.ndim
    idx[axis] = j
    return tuple(idx)


def _border_tag(b):
    return b if isinstance(b, str) else b[0]


def _const_front(buf, axis, pad, c):
    # half_dim.rs:30-49
    for j in range(pad[0]):
        buf[_plane(buf, axis, j)] = c


def _const_back(buf, axis, pad, n, c):
    # half_dim.rs:72-94
    for j in range(n + pad[0], buf.shape[axis]):
        buf[_plane(buf, axis, j)] = c


def _replicate_front(buf, axis, pad):
    # half_dim.rs:113-129: border = plane pad[0], copied into 0..pad[0]
    border = buf[_plane(buf, axis, pad[0])].copy()
    for j in range(pad[0]):
        buf[_plane(buf, axis, j)] = border


def _replicate_back(buf, axis, pad, n):
    # half_dim.rs:151-173: border = plane len - pad[1] - 1
    border = buf[_plane(buf, axis, buf.shape[axis] - pad[1] - 1)].copy()
    for j in range(n + pad[0], buf.shape[axis]):
        buf[_plane(buf, axis, j)] = border


def _reflect_front(buf, axis, pad):
    # half_dim.rs:192-211: src = (p0 - j) + p0, panics when out of range
    p0 = pad[0]
    size = buf.shape[axis]
    for j in range(p0):
        src = (p0 - j) + p0
        if src >= size:
            raise OracleError(INVALID_ARGUMENT, "reflect front pad reads outside the padded axis")
        buf[_plane(buf, axis, j)] = buf[_plane(buf, axis, src)].copy()


def _reflect_back(buf, axis, pad, n):
    # half_dim.rs:233-258: border = len - p1 - 1, src = border - (j - border),
    # usize underflow there is a panic in the reference
    size = buf.shape[axis]
    border = size - pad[1] - 1
    for j in range(n + pad[0], size):
        d = j - border
        if d > border:
            raise OracleError(INVALID_ARGUMENT, "reflect back pad reads outside the padded axis")
        buf[_plane(buf, axis, j)] = buf[_plane(buf, axis, border - d)].copy()


def _circular_front(buf, axis, pad):
    # half_dim.rs:277-296: src = len - p1 - (p0 - j)
    size = buf.shape[axis]
    for j in range(pad[0]):
        src = size - pad[1] - (pad[0] - j)
        buf[_plane(buf, axis, j)] = buf[_plane(buf, axis, src)].copy()


def _circular_back(buf, axis, pad, n):
    # half_dim.rs:318-343: src = p0 + (j - border - 1), border = len - p1 - 1
    size = buf.shape[axis]
    border = size - pad[1] - 1
    for j in range(n + pad[0], size):
        src = pad[0] + (j - border - 1)
        buf[_plane(buf, axis, j)] = buf[_plane(buf, axis, src)].copy()


def _front(buf, axis, pad, border):
    tag = _border_tag(border)
    if tag == "zeros":
        _const_front(buf, axis, pad, 0)
    elif tag == "const":
        _const_front(buf, axis, pad, border[1])
    elif tag == "reflect":
        _reflect_front(buf, axis, pad)
    elif tag == "replicate":
        _replicate_front(buf, axis, pad)
    elif tag == "circular":
        _circular_front(buf, axis, pad)
    else:
        raise ValueError(border)


def _back(buf, axis, pad, n, border):
    tag = _border_tag(border)
    if tag == "zeros":
        _const_back(buf, axis, pad, n, 0)
    elif tag == "const":
        _const_back(buf, axis, pad, n, border[1])
    elif tag == "reflect":
        _reflect_back(buf, axis, pad, n)
    elif tag == "replicate":
        _replicate_back(buf, axis, pad, n)
    elif tag == "circular":
        _circular_back(buf, axis, pad, n)
    else:
        raise ValueError(border)


def _fill(x, buf, mode, ep, in_place_const):
    """The mode dispatch shared by padding (padding/mod.rs:103-114) and
    padding_in (padding/mod.rs:133-152).  Axes are filled in order 0..N-1,
    front then back, each fill copying whole hyperplanes."""
    n = x.ndim
    tag = mode if isinstance(mode, str) else mode[0]
    if tag == "zeros":
        return
    if tag == "const":
        if in_place_const:
            # padding_in: dim::constant per axis (padding/mod.rs:136-143)
            for d in range(n):
                _const_front(buf, d, ep[d], mode[1])
                _const_back(buf, d, ep[d], x.shape[d], mode[1])
        return
    if tag in ("reflect", "replicate", "circular"):
        for d in range(n):
            _front(buf, d, ep[d], tag)
            _back(buf, d, ep[d], x.shape[d], tag)
        return
    if tag == "custom":
        for d in range(n):
            _front(buf, d, ep[d], mode[1][d])
            _back(buf, d, ep[d], x.shape[d], mode[1][d])
        return
    if tag == "explicit":
        for d in range(n):
            _front(buf, d, ep[d], mode[1][d][0])
            _back(buf, d, ep[d], x.shape[d], mode[1][d][1])
        return
    raise ValueError(mode)


def padding(x, mode, ep):
    """PaddingExt::padding (padding/mod.rs:84-117): buffer pre-filled with c
    for PaddingMode::Const, else zero; interior copy (padding_const,
    padding/mod.rs:175-200); then the per-axis fills."""
    x = np.asarray(x)
    shape = [x.shape[i] + ep[i][0] + ep[i][1] for i in range(x.ndim)]
    tag = mode if isinstance(mode, str) else mode[0]
    init = mode[1] if tag == "const" else 0
    buf = np.full(shape, init, dtype=x.dtype)
    buf[tuple(slice(ep[i][0], ep[i][0] + x.shape[i]) for i in range(x.ndim))] = x
    _fill(x, buf, mode, ep, in_place_const=False)
    return buf


def padding_data(x, mode, ep, fsize):
    """conv_fft::padding::data (conv_fft/padding.rs:30-62): zero buffer of the
    FFT size; padding_in writes into the [0, pds) slice (padding/mod.rs:119-153)."""
    x = np.asarray(x)
    buf = np.zeros(fsize, dtype=x.dtype)
    pds = [x.shape[i] + ep[i][0] + ep[i][1] for i in range(x.ndim)]
    sl = tuple(slice(0, p) for p in pds)
    view = buf[sl]
    view[tuple(slice(ep[i][0], ep[i][0] + x.shape[i]) for i in range(x.ndim))] = x
    _fill(x, view, mode, ep, in_place_const=True)
    return buf


def padding_kernel(k, dilation, reverse, fsize):
    """conv_fft::padding::kernel (conv_fft/padding.rs:78-111): kernel element
    i goes to i*d for reverse=True (convolution; the FFT itself flips), and
    to (K-1-i)*d for no_reverse (the negative-step slice at :103)."""
    k = np.asarray(k)
    buf = np.zeros(fsize, dtype=k.dtype)
    kd = kernel_dilated_shape(k.shape, dilation)
    sl = []
    for i in range(k.ndim):
        if reverse:
            sl.append(slice(0, kd[i], dilation[i]))
        else:
            # ndarray's negative step walks the [0, kd) range from its end
            sl.append(slice(kd[i] - 1, None, -dilation[i]) if kd[i] - 1 >= 0 else slice(0, 0))
    buf[tuple(sl)] = k
    return buf


# ----------------------------------------------------------------------------
# Offsets (dilation/mod.rs:34-60)
# ----------------------------------------------------------------------------

def gen_offset_list(k, dilation, reverse, pds_strides):
    """Row-major walk over the kernel (reversed along every axis when
    reverse), zero taps dropped; offset = sum idx_n * d_n * stride_n."""
    k = np.asarray(k)
    view = k[tuple(slice(None, None, -1) if reverse else slice(None) for _ in range(k.ndim))]
    out = []
    for idx in np.ndindex(*view.shape):
        v = view[idx]
        if v == 0:
            continue
        off = sum(int(idx[n]) * int(dilation[n]) * int(pds_strides[n]) for n in range(k.ndim))
        out.append((off, v))
    return out


def _norm_dilation(dilation, n):
    if dilation is None:
        return [1] * n
    if np.isscalar(dilation):
        return [int(dilation)] * n
    return [int(d) for d in dilation]


# ----------------------------------------------------------------------------
# Direct convolution (conv/mod.rs:128-200)
# ----------------------------------------------------------------------------

def conv(x, k, mode, padmode, dilation=None, reverse=True, acc_dtype=None):
    """ConvExt::conv.  With acc_dtype=None the arithmetic is done in x's
    dtype in offset-list order (integers wrap like the reference's release
    build; floats round after every product and every add, as the
    reference's `tmp_res += a * b` does).  acc_dtype=np.float64 (or
    complex128) gives the FP64 oracle."""
    x = np.asarray(x)
    k = np.asarray(k)
    n = x.ndim
    dil = _norm_dilation(dilation, n)
    if x.size == 0:
        raise OracleError(DATA_SHAPE, str(x.shape))
    if k.size == 0:
        raise OracleError(KERNEL_SHAPE, str(k.shape))
    kd = kernel_dilated_shape(k.shape, dil)
    ep, strides = unfold(mode, k.shape, dil)
    pds = padding(x, padmode, ep)                       # conv/mod.rs:150
    if not all(kd[i] <= pds.shape[i] for i in range(n)):
        raise OracleError(MISMATCH_SHAPE, f"kdil {kd} vs padded {pds.shape}")
    if any(s == 0 for s in strides):
        raise OracleError(INVALID_ARGUMENT, "stride 0 divides by zero (conv/mod.rs:165)")
    out_shape = [(ep[i][0] + ep[i][1] + x.shape[i] - kd[i]) // strides[i] + 1 for i in range(n)]
    dt = x.dtype if acc_dtype is None else np.dtype(acc_dtype)
    src = pds.astype(dt, copy=False)
    acc = np.zeros(out_shape, dtype=dt)
    view = k[tuple(slice(None, None, -1) if reverse else slice(None) for _ in range(n))]
    with np.errstate(over="ignore"):
        for idx in np.ndindex(*view.shape):
            v = view[idx]
            if v == 0:
                continue
            sl = tuple(slice(idx[i] * dil[i], idx[i] * dil[i] + (out_shape[i] - 1) * strides[i] + 1,
                             strides[i]) for i in range(n))
            acc += src[sl] * dt.type(v)
    return acc


def conv_int_exact(x, k, mode, padmode, dilation=None, reverse=True, bits=32):
    """Exact integer direct convolution: int64 brute force (no overflow for
    the value ranges used), then reduced to two's complement of `bits`,
    which is what wrapping arithmetic yields in any summation order."""
    acc = conv(np.asarray(x).astype(np.int64), np.asarray(k).astype(np.int64), mode, padmode,
               dilation, reverse)
    m = 1 << bits
    r = np.mod(acc, m)
    r = np.where(r >= m // 2, r - m, r)
    return r


# ----------------------------------------------------------------------------
# FFT convolution (conv_fft/mod.rs:185-292, processor/real.rs, complex.rs)
# ----------------------------------------------------------------------------

def conv_fft_plan(xshape, kshape, mode, dilation):
    """Shape arithmetic of conv_fft_proc_impl (conv_fft/mod.rs:215-231)."""
    n = len(xshape)
    dil = _norm_dilation(dilation, n)
    kd = kernel_dilated_shape(kshape, dil)
    ep, strides = unfold(mode, kshape, dil)
    pds = [xshape[i] + ep[i][0] + ep[i][1] for i in range(n)]
    fs = fft_size([max(pds[i], kd[i]) for i in range(n)])
    return dict(kdil=kd, padding=ep, strides=strides, pds=pds, fft_size=fs,
                out_shape=[(pds[i] - kd[i]) // strides[i] + 1 if strides[i] else None
                           for i in range(n)])


def conv_fft(x, k, mode, padmode, dilation=None, reverse=True, precision="f64", workers=1):
    """ConvFFTExt::conv_fft restated: pad data into the FFT buffer, scatter
    the kernel, forward N-D transform (real: r2c on the last axis then c2c on
    the others, real.rs:105-155; complex: c2c on every axis, complex.rs:33-115),
    multiply, inverse (scaled by 1/prod F, real.rs:278-279 / complex.rs:141-142),
    then crop [kdil-1, pds) with the stride step (conv_fft/mod.rs:282-289).

    precision="f64" is the FP64 oracle (numpy pocketfft in double);
    precision="f32" restates the crate's f32 arithmetic with scipy's
    single-precision pocketfft (the CPU baseline)."""
    x = np.asarray(x)
    k = np.asarray(k)
    n = x.ndim
    dil = _norm_dilation(dilation, n)
    if x.size == 0:
        raise OracleError(DATA_SHAPE, str(x.shape))
    if k.size == 0:
        raise OracleError(DATA_SHAPE, str(k.shape))   # conv_fft/mod.rs:211-213 reports DataShape
    plan = conv_fft_plan(x.shape, k.shape, mode, dil)
    kd, ep, strides, pds, fs = plan["kdil"], plan["padding"], plan["strides"], plan["pds"], plan["fft_size"]
    if not all(kd[i] <= pds[i] for i in range(n)):
        raise OracleError(MISMATCH_SHAPE, f"kdil {kd} vs padded {pds}")
    if any(d == 0 for d in dil):
        raise OracleError(INVALID_ARGUMENT, "dilation 0 makes a zero slice step (conv_fft/padding.rs:103)")
    is_complex = np.iscomplexobj(x) or np.iscomplexobj(k)
    if precision == "f64":
        rdt, cdt = np.float64, np.complex128
    else:
        rdt, cdt = np.float32, np.complex64
    wdt = cdt if is_complex else rdt
    pmode = padmode
    xd = padding_data(x.astype(wdt), pmode, ep, fs)
    kb = padding_kernel(k.astype(wdt), dil, reverse, fs)
    axes_rest = list(range(n - 2, -1, -1))
    if precision == "f64":
        fft = np.fft
        kw = {}
    else:
        import scipy.fft as fft
        kw = {"workers": workers}
    if is_complex:
        def fwd(a):
            a = fft.fft(a, axis=n - 1, **kw)
            for ax in axes_rest:
                a = fft.fft(a, axis=ax, **kw)
            return a
        X = fwd(xd)
        K = fwd(kb)
        P = X * K
        for ax in reversed(axes_rest):
            P = fft.ifft(P, axis=ax, **kw)
        full = fft.ifft(P, axis=n - 1, **kw)
    else:
        def fwd(a):
            a = fft.rfft(a, axis=n - 1, **kw)
            for ax in axes_rest:
                a = fft.fft(a, axis=ax, **kw)
            return a
        X = fwd(xd)
        K = fwd(kb)
        P = X * K
        for ax in reversed(axes_rest):
            P = fft.ifft(P, axis=ax, **kw)
        full = fft.irfft(P, n=fs[n - 1], axis=n - 1, **kw)
    full = full.astype(wdt, copy=False)
    if any(s == 0 for s in strides):
        raise OracleError(INVALID_ARGUMENT, "stride 0 slice step (conv_fft/mod.rs:286)")
    return full[tuple(slice(kd[i] - 1, pds[i], strides[i]) for i in range(n))]


def circular_conv_bruteforce(a, b):
    """Textbook N-D circular convolution by direct summation (for pinning
    conv_fft on tiny inputs): y[i] = sum_j a[j] b[(i - j) mod F]."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = np.zeros(a.shape, dtype=np.result_type(a, b, np.float64))
    shape = a.shape
    for j in np.ndindex(*shape):
        if a[j] == 0:
            continue
        rolled = b
        for ax in range(a.ndim):
            rolled = np.roll(rolled, j[ax], axis=ax)
        out += a[j] * rolled
    return out


def relerr(y, ref):
    """Normwise relative error max|y - ref| / max|ref| and max abs error."""
    y = np.asarray(y).astype(np.complex128 if np.iscomplexobj(y) or np.iscomplexobj(ref) else np.float64)
    ref = np.asarray(ref).astype(y.dtype)
    d = np.abs(y - ref)
    m = float(np.max(np.abs(ref))) if ref.size else 0.0
    mx = float(np.max(d)) if d.size else 0.0
    return (mx / m if m > 0 else mx), mx


# ----------------------------------------------------------------------------
# Processor::forward / backward (processor/real.rs:105-155,233-281,
# processor/complex.rs:33-145)
# ----------------------------------------------------------------------------

def processor_forward(x):
    """The N-d DFT (forward, unnormalised) with its axes rotated left by one,
    the layout the reference's processors return: the last axis is
    transformed first (r2c for real input, F/2+1 bins), then N-1 times the
    axes are rotated right and the new last axis is transformed
    (real.rs:117-152, complex.rs:54-110).  Pinned against the procedural
    restatement below and the reference's roundtrip tests."""
    x = np.asarray(x)
    X = np.fft.rfftn(x) if not np.iscomplexobj(x) else np.fft.fftn(x)
    return np.moveaxis(X, 0, -1) if x.ndim > 1 else X


def processor_forward_procedural(x):
    """Literal restatement of the reference's loop: transform the last axis,
    then N-1 times permute axes by rotate_right(1), copy contiguous, and
    transform the new last axis."""
    x = np.asarray(x)
    n = x.ndim
    out = np.fft.rfft(x, axis=-1) if not np.iscomplexobj(x) else np.fft.fft(x, axis=-1)
    axes = list(range(n))
    axes = axes[-1:] + axes[:-1]   # rotate_right(1)
    for _ in range(n - 1):
        out = np.ascontiguousarray(np.transpose(out, axes))
        out = np.fft.fft(out, axis=-1)
    return out


def processor_backward(spec, real_len=None):
    """Inverse of processor_forward divided by the number of output elements
    (real.rs:233-281 with rp_origin_len = real_len, complex.rs:134-145).  For
    real output the imaginary parts of the DC and Nyquist bins are ignored."""
    spec = np.asarray(spec)
    X = np.moveaxis(spec, -1, 0) if spec.ndim > 1 else spec
    if real_len is None:
        return np.fft.ifftn(X)
    X = X.copy()
    X = np.fft.ifftn(X, axes=list(range(X.ndim - 1))) if X.ndim > 1 else X
    X[..., 0] = X[..., 0].real
    if real_len % 2 == 0:
        X[..., real_len // 2] = X[..., real_len // 2].real
    return np.fft.irfft(X, n=real_len, axis=-1)
