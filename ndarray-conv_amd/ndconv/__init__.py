"""Host-side mirror of ndarray-conv's convolution API over the HIP C ABI.

The names, argument meaning and error behaviour follow the reference crate
(ndarray-conv 0.6.1):

    conv(data, kernel, conv_mode, padding_mode)          ConvExt::conv            conv/mod.rs:110-115
    conv_fft(data, kernel, conv_mode, padding_mode)      ConvFFTExt::conv_fft     conv_fft/mod.rs:113-118
    conv_fft_with_processor(..., processor)                                       conv_fft/mod.rs:151-157
    conv_fft_par(...)                                                             conv_fft/mod.rs:166-171
    ConvMode.Full/Same/Valid/Custom/Explicit                                      lib.rs:80-105
    PaddingMode.Zeros/Const/Reflect/Replicate/Circular/Custom/Explicit            lib.rs:106-127
    BorderType.Zeros/Const/Reflect/Replicate/Circular                             lib.rs:129-143
    with_dilation(kernel, d), reverse(k), no_reverse(k)                           dilation/mod.rs:103-189
    DataShape / KernelShape / MismatchShape errors                                lib.rs:145-159
    BatchPlan / conv_fft_batch: many rank-1 pairs, conv_fft semantics per pair,
    one persistent launch (the reference loops over conv_fft calls)

`data` and `kernel` are numpy arrays (host memory, copied to and from the
GPU inside the call) or torch tensors already on the GPU (HBM-resident; the
result is a new GPU tensor and the call is ordered on torch's current
stream).  Reference panics (stride 0, reflect pads leaving the axis) raise
InvalidArgument.  Every computation runs in libndconv.so; there is no CPU
fallback.
"""
import ctypes
import threading

import numpy as np

from . import _lib as C

__all__ = ["ConvMode", "PaddingMode", "BorderType", "KernelWithDilation", "with_dilation", "reverse",
           "no_reverse", "conv", "conv_fft", "conv_fft_with_processor", "conv_fft_par", "get_fft_processor",
           "FftProcessor", "plan", "border_map", "Error", "DataShape", "KernelShape", "MismatchShape",
           "InvalidArgument", "Unsupported", "DeviceError", "BatchPlan", "conv_fft_batch", "batch_border_map",
           "Baked", "conv_fft_bake", "conv_fft_with_baked", "fft_launch_plan"]


# ---------------------------------------------------------------------------
# Errors (lib.rs:145-159)
# ---------------------------------------------------------------------------

class Error(Exception):
    """Base of the convolution errors."""


class DataShape(Error):
    """Data shape shouldn't have ZERO (Error::DataShape)."""


class KernelShape(Error):
    """Kernel shape shouldn't have ZERO (Error::KernelShape)."""


class MismatchShape(Error):
    """ConvMode does not match KernelWithDilation size (Error::MismatchShape)."""


class InvalidArgument(Error):
    """An input on which the reference panics (stride 0, reflect pad out of range)."""


class Unsupported(Error):
    pass


class DeviceError(Error):
    pass


_STATUS_EXC = {C.ERR_DATA_SHAPE: DataShape, C.ERR_KERNEL_SHAPE: KernelShape,
               C.ERR_MISMATCH_SHAPE: MismatchShape, C.ERR_INVALID_ARGUMENT: InvalidArgument,
               C.ERR_UNSUPPORTED: Unsupported, C.ERR_NO_DEVICE: DeviceError, C.ERR_DEVICE: DeviceError,
               C.ERR_OUT_OF_MEMORY: DeviceError}


def _check(status, handle=None):
    if status == C.OK:
        return
    lib = C.lib()
    msg = lib.ndc_status_string(status).decode()
    if handle is not None:
        detail = lib.ndc_last_error(handle).decode()
        if detail:
            msg = f"{msg}: {detail}"
    raise _STATUS_EXC.get(status, Error)(msg)


# ---------------------------------------------------------------------------
# Modes (lib.rs:80-143)
# ---------------------------------------------------------------------------

class _ConvModeValue:
    __slots__ = ("tag", "padding", "strides", "name")

    def __init__(self, tag, name, padding=None, strides=None):
        self.tag, self.name, self.padding, self.strides = tag, name, padding, strides

    def __repr__(self):
        if self.padding is None:
            return f"ConvMode.{self.name}"
        return f"ConvMode.{self.name}(padding={self.padding}, strides={self.strides})"


class ConvMode:
    Full = _ConvModeValue(C.CONV_FULL, "Full")
    Same = _ConvModeValue(C.CONV_SAME, "Same")
    Valid = _ConvModeValue(C.CONV_VALID, "Valid")

    @staticmethod
    def Custom(padding, strides):
        return _ConvModeValue(C.CONV_CUSTOM, "Custom", [int(p) for p in padding], [int(s) for s in strides])

    @staticmethod
    def Explicit(padding, strides):
        return _ConvModeValue(C.CONV_EXPLICIT, "Explicit", [[int(a), int(b)] for a, b in padding],
                              [int(s) for s in strides])


class _BorderValue:
    __slots__ = ("tag", "value", "name")

    def __init__(self, tag, name, value=None):
        self.tag, self.name, self.value = tag, name, value

    def __repr__(self):
        return f"BorderType.{self.name}" + ("" if self.value is None else f"({self.value})")


class BorderType:
    Zeros = _BorderValue(C.BORDER_ZEROS, "Zeros")
    Reflect = _BorderValue(C.BORDER_REFLECT, "Reflect")
    Replicate = _BorderValue(C.BORDER_REPLICATE, "Replicate")
    Circular = _BorderValue(C.BORDER_CIRCULAR, "Circular")

    @staticmethod
    def Const(c):
        return _BorderValue(C.BORDER_CONST, "Const", c)


class _PaddingValue:
    __slots__ = ("tag", "value", "borders", "name")

    def __init__(self, tag, name, value=None, borders=None):
        self.tag, self.name, self.value, self.borders = tag, name, value, borders

    def __repr__(self):
        if self.borders is not None:
            return f"PaddingMode.{self.name}({self.borders})"
        return f"PaddingMode.{self.name}" + ("" if self.value is None else f"({self.value})")


class PaddingMode:
    Zeros = _PaddingValue(C.PAD_ZEROS, "Zeros")
    Reflect = _PaddingValue(C.PAD_REFLECT, "Reflect")
    Replicate = _PaddingValue(C.PAD_REPLICATE, "Replicate")
    Circular = _PaddingValue(C.PAD_CIRCULAR, "Circular")

    @staticmethod
    def Const(c):
        return _PaddingValue(C.PAD_CONST, "Const", value=c)

    @staticmethod
    def Custom(borders):
        return _PaddingValue(C.PAD_CUSTOM, "Custom", borders=list(borders))

    @staticmethod
    def Explicit(borders):
        return _PaddingValue(C.PAD_EXPLICIT, "Explicit", borders=[list(b) for b in borders])


# ---------------------------------------------------------------------------
# Kernel wrappers (dilation/mod.rs:8-12, 103-212)
# ---------------------------------------------------------------------------

class KernelWithDilation:
    """A kernel with per-axis dilation and a reverse flag (True = convolution,
    False = cross-correlation)."""
    __slots__ = ("kernel", "dilation", "reverse_flag")

    def __init__(self, kernel, dilation=1, reverse_flag=True):
        self.kernel = kernel
        nd = len(kernel.shape)
        if np.isscalar(dilation) or isinstance(dilation, (int, np.integer)):
            dilation = [int(dilation)] * nd
        self.dilation = [int(d) for d in dilation]
        self.reverse_flag = bool(reverse_flag)

    def with_dilation(self, dilation):
        return KernelWithDilation(self.kernel, dilation, self.reverse_flag)

    def reverse(self):
        return KernelWithDilation(self.kernel, self.dilation, True)

    def no_reverse(self):
        return KernelWithDilation(self.kernel, self.dilation, False)


def with_dilation(kernel, dilation):
    """WithDilation::with_dilation: a usize for every axis or one per axis; reverse stays on."""
    return KernelWithDilation(kernel, dilation, True)


def reverse(kernel):
    return _into_kwd(kernel).reverse()


def no_reverse(kernel):
    return _into_kwd(kernel).no_reverse()


def _into_kwd(k):
    if isinstance(k, KernelWithDilation):
        return k
    return KernelWithDilation(k, 1, True)


# ---------------------------------------------------------------------------
# Descriptor packing
# ---------------------------------------------------------------------------

_NP_DTYPE = {np.dtype(np.float32): C.F32, np.dtype(np.float64): C.F64, np.dtype(np.complex64): C.C64,
             np.dtype(np.complex128): C.C128, np.dtype(np.int8): C.I8, np.dtype(np.int16): C.I16,
             np.dtype(np.int32): C.I32, np.dtype(np.int64): C.I64, np.dtype(np.uint8): C.U8,
             np.dtype(np.uint16): C.U16, np.dtype(np.uint32): C.U32, np.dtype(np.uint64): C.U64}


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _np_dtype_of(a):
    if _is_torch(a):
        import torch
        m = {torch.float32: np.float32, torch.float64: np.float64, torch.complex64: np.complex64,
             torch.complex128: np.complex128, torch.int8: np.int8, torch.int16: np.int16,
             torch.int32: np.int32, torch.int64: np.int64, torch.uint8: np.uint8}
        if a.dtype not in m:
            raise Unsupported(f"dtype {a.dtype}")
        return np.dtype(m[a.dtype])
    return np.asarray(a).dtype


def _value(v, dtype):
    val = C.Value()
    if v is None:
        return val
    if dtype in (C.F32, C.F64):
        val.re = float(np.real(v))
    elif dtype in (C.C64, C.C128):
        val.re, val.im = float(np.real(v)), float(np.imag(v))
    else:
        iv = int(v)
        if iv >= 2 ** 63:
            iv -= 2 ** 64
        val.i = iv
    return val


def _fill_desc(d, x, kwd, conv_mode, padding_mode, dtype, host):
    k = kwd.kernel
    n = len(x.shape)
    if len(k.shape) != n:
        raise InvalidArgument("data and kernel ranks differ")
    if n < 1 or n > C.MAX_RANK:
        raise Unsupported(f"rank {n}")
    d.dtype = dtype
    d.rank = n
    flags = C.FLAG_HOST_MEMORY if host else 0
    for arr, shp, strd, flag in ((x, d.x_shape, d.x_strides, C.FLAG_X_STRIDED),
                                 (k, d.k_shape, d.k_strides, C.FLAG_K_STRIDED)):
        for i, s in enumerate(arr.shape):
            shp[i] = int(s)
        if _is_torch(arr):
            st = arr.stride()
        else:
            st = [s // arr.itemsize for s in arr.strides]
            if any(s % arr.itemsize for s in arr.strides):
                raise Unsupported("strides that are not a multiple of the element size")
        for i, s in enumerate(st):
            strd[i] = int(s)
        flags |= flag
    if len(kwd.dilation) != n:
        raise InvalidArgument("dilation rank differs from the kernel rank")
    for i, dl in enumerate(kwd.dilation):
        d.dilation[i] = dl
    d.reverse = 1 if kwd.reverse_flag else 0
    cm = d.conv_mode
    cm.tag = conv_mode.tag
    if conv_mode.tag == C.CONV_CUSTOM:
        if len(conv_mode.padding) != n or len(conv_mode.strides) != n:
            raise InvalidArgument("ConvMode::Custom rank differs")
        for i in range(n):
            cm.padding[i][0] = cm.padding[i][1] = conv_mode.padding[i]
            cm.strides[i] = conv_mode.strides[i]
    elif conv_mode.tag == C.CONV_EXPLICIT:
        if len(conv_mode.padding) != n or len(conv_mode.strides) != n:
            raise InvalidArgument("ConvMode::Explicit rank differs")
        for i in range(n):
            cm.padding[i][0], cm.padding[i][1] = conv_mode.padding[i]
            cm.strides[i] = conv_mode.strides[i]
    pm = d.padding_mode
    pm.tag = padding_mode.tag
    pm.value = _value(padding_mode.value, dtype)
    if padding_mode.tag == C.PAD_CUSTOM:
        if len(padding_mode.borders) != n:
            raise InvalidArgument("PaddingMode::Custom rank differs")
        for i, b in enumerate(padding_mode.borders):
            pm.borders[i][0].tag = b.tag
            pm.borders[i][0].value = _value(b.value, dtype)
    elif padding_mode.tag == C.PAD_EXPLICIT:
        if len(padding_mode.borders) != n:
            raise InvalidArgument("PaddingMode::Explicit rank differs")
        for i, (b0, b1) in enumerate(padding_mode.borders):
            pm.borders[i][0].tag = b0.tag
            pm.borders[i][0].value = _value(b0.value, dtype)
            pm.borders[i][1].tag = b1.tag
            pm.borders[i][1].value = _value(b1.value, dtype)
    d.flags = flags


def _desc_for_shapes(xshape, kshape, conv_mode, padding_mode, dilation=1, reverse_flag=True, dtype=np.float32):
    """Descriptor from shapes only (plan queries)."""
    class _S:
        def __init__(self, shape, itemsize):
            self.shape = tuple(shape)
            self.itemsize = itemsize
            st, acc = [], itemsize
            for s in reversed(self.shape):
                st.append(acc)
                acc *= max(int(s), 1)
            self.strides = tuple(reversed(st))
    dt = np.dtype(dtype)
    x = _S(xshape, dt.itemsize)
    k = _S(kshape, dt.itemsize)
    d = C.ConvDesc()
    _fill_desc(d, x, KernelWithDilation(k, dilation, reverse_flag), conv_mode, padding_mode, _NP_DTYPE[dt], True)
    return d


# ---------------------------------------------------------------------------
# Processor / handle (processor/mod.rs:71-143)
# ---------------------------------------------------------------------------

class FftProcessor:
    """Owns a library handle: HIP stream, device workspace, twiddle tables and
    plan caches.  Reusing one across calls is what conv_fft_with_processor is
    for.  Not thread-safe (use one per thread, as &mut Processor).

    forward/backward are Processor::forward/backward (processor/mod.rs:79-121):
    `elem` is the processor's InElem (a real dtype for real.rs, a complex one
    for complex.rs; by default the dtype of the first forward input)."""

    def __init__(self, elem=None):
        lib = C.lib()
        h = ctypes.c_void_p()
        _check(lib.ndc_create(ctypes.byref(h)))
        self._h = h
        self._elem = None if elem is None else np.dtype(elem)

    def _transform(self, a, forward):
        lib = C.lib()
        h = self._h
        torch_in = _is_torch(a)
        dt = _np_dtype_of(a)
        if forward:
            if self._elem is None:
                self._elem = dt
            if dt != self._elem:
                raise InvalidArgument(f"this processor transforms {self._elem}, not {dt}")
            code = _NP_DTYPE.get(dt)
        else:
            if self._elem is None:
                raise InvalidArgument("backward before any forward: the processor's element type is unknown")
            if dt != np.dtype(np.complex64 if self._elem in (np.float32, np.complex64) else np.complex128):
                raise InvalidArgument(f"a spectrum of {self._elem} is complex{64 if self._elem.itemsize <= 8 else 128}")
            code = _NP_DTYPE.get(self._elem)
        if code not in (C.F32, C.F64, C.C64, C.C128):
            raise Unsupported(f"dtype {dt}")
        shape = tuple(int(v) for v in a.shape)
        n = len(shape)
        real = self._elem.kind == "f"
        if forward:
            last = shape[-1] // 2 + 1 if real else shape[-1]
            out_shape = shape[1:-1] + (last,) + shape[:1] if n > 1 else (last,)
            out_dt = np.dtype(np.complex64 if code in (C.F32, C.C64) else np.complex128)
        else:
            nat = shape[-1:] + shape[:-1] if n > 1 else shape
            if real:
                o = ctypes.c_int64()
                _check(lib.ndc_fft_origin_len(h, ctypes.byref(o)), h)
                nat = nat[:-1] + (o.value,)
            out_shape = nat
            out_dt = self._elem
        sh = (ctypes.c_int64 * n)(*shape)
        fn = lib.ndc_fft_forward if forward else lib.ndc_fft_backward
        if torch_in:
            import torch
            if not a.is_cuda or not a.is_contiguous():
                raise InvalidArgument("torch inputs must be contiguous GPU tensors")
            tdt = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
                   np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128}[out_dt]
            y = torch.empty(out_shape, dtype=tdt, device=a.device)
            _check(lib.ndc_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)), h)
            _check(fn(h, code, n, sh, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(y.data_ptr()), 0), h)
            return y
        x = np.ascontiguousarray(a)
        y = np.empty(out_shape, dtype=out_dt)
        _check(lib.ndc_reset_stream(h), h)
        _check(fn(h, code, n, sh, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data), C.FLAG_HOST_MEMORY), h)
        return y

    def forward(self, input, parallel=False):
        """Processor::forward (processor/mod.rs:83-96): the N-d DFT with the axes
        rotated left by one (real input: half spectrum on the input's last axis)."""
        return self._transform(input, True)

    def backward(self, input, parallel=False):
        """Processor::backward (processor/mod.rs:98-119): the inverse of forward,
        divided by the number of elements, in natural axis order."""
        return self._transform(input, False)

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _check(C.lib().ndc_synchronize(self._h), self._h)

    def close(self):
        if getattr(self, "_h", None):
            C.lib().ndc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_fft_processor(elem=None):
    """get_fft_processor (processor/mod.rs:71-73); `elem` is InElem."""
    return FftProcessor(elem)


_tls = threading.local()


def _default_processor():
    p = getattr(_tls, "proc", None)
    if p is None:
        p = FftProcessor()
        _tls.proc = p
    return p


# ---------------------------------------------------------------------------
# Operations
# ---------------------------------------------------------------------------

def plan(data_shape, kernel_shape, conv_mode, padding_mode=PaddingMode.Zeros, dilation=1, reverse_flag=True,
         fft=False, dtype=np.float32):
    """Shape arithmetic of one convolution (padding, strides, kdil, padded, fft_size, out_shape)."""
    d = _desc_for_shapes(data_shape, kernel_shape, conv_mode, padding_mode, dilation, reverse_flag, dtype)
    info = C.PlanInfo()
    _check(C.lib().ndc_plan(ctypes.byref(d), C.OP_CONV_FFT if fft else C.OP_CONV, ctypes.byref(info)))
    n = len(data_shape)
    return dict(padding=[[info.padding[i][0], info.padding[i][1]] for i in range(n)],
                strides=[info.strides[i] for i in range(n)],
                kernel_dilated=[info.kernel_dilated[i] for i in range(n)],
                padded=[info.padded[i] for i in range(n)],
                fft_size=[info.fft_size[i] for i in range(n)],
                out_shape=[info.out_shape[i] for i in range(n)])


def border_map(data_shape, kernel_shape, conv_mode, padding_mode, axis, dilation=1, fft=False):
    """The resolved border map of one axis (see ndc_border_map)."""
    info = plan(data_shape, kernel_shape, conv_mode, padding_mode, dilation, True, fft)
    n = info["fft_size"][axis] if fft else info["padded"][axis]
    d = _desc_for_shapes(data_shape, kernel_shape, conv_mode, padding_mode, dilation, True)
    out = np.zeros(n, dtype=np.int32)
    _check(C.lib().ndc_border_map(ctypes.byref(d), C.OP_CONV_FFT if fft else C.OP_CONV, axis,
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n))
    return out


def fft_launch_plan(data_shape, kernel_shape, conv_mode, padding_mode=PaddingMode.Zeros, dilation=1,
                    reverse_flag=True, dtype=np.float32, lds_bytes=160 * 1024 - 1024, cus=256):
    """The kernel launches conv_fft makes for these shapes, planned on the host
    without a device (see ndc_conv_fft_launches): a list of
    (dynamic LDS bytes, grid, block) per launch, for a device with `lds_bytes`
    of dynamic LDS per workgroup and `cus` compute units."""
    d = _desc_for_shapes(data_shape, kernel_shape, conv_mode, padding_mode, dilation, reverse_flag, dtype)
    lib = C.lib()
    cap = 64
    bufs = [(ctypes.c_int64 * cap)() for _ in range(3)]
    cnt = ctypes.c_int64()
    _check(lib.ndc_conv_fft_launches(ctypes.byref(d), int(lds_bytes), int(cus), bufs[0], bufs[1], bufs[2], cap,
                                     ctypes.byref(cnt)))
    return [(bufs[0][i], bufs[1][i], bufs[2][i]) for i in range(min(cap, cnt.value))]


_ALGOS = {"auto": C.ALGO_AUTO, "direct": C.ALGO_DIRECT, "bf16": C.ALGO_MFMA_BF16, "bf16x3": C.ALGO_MFMA_BF16X3}


def _run(op, data, kernel, conv_mode, padding_mode, processor=None, algo="auto"):
    lib = C.lib()
    kwd = _into_kwd(kernel)
    proc = processor or _default_processor()
    torch_in = _is_torch(data) or _is_torch(kwd.kernel)
    if torch_in:
        import torch
        if not (_is_torch(data) and _is_torch(kwd.kernel)) or not data.is_cuda or not kwd.kernel.is_cuda:
            raise InvalidArgument("torch inputs must both be GPU tensors")
        if data.dtype != kwd.kernel.dtype:
            raise InvalidArgument("data and kernel dtypes differ")
        x, k = data, kwd.kernel
    else:
        x = np.asarray(data)
        k = np.asarray(kwd.kernel)
        if x.dtype != k.dtype:
            raise InvalidArgument("data and kernel dtypes differ")
    dt = _np_dtype_of(x)
    if dt not in _NP_DTYPE:
        raise Unsupported(f"dtype {dt}")
    dtype = _NP_DTYPE[dt]
    d = C.ConvDesc()
    _fill_desc(d, x, KernelWithDilation(k, kwd.dilation, kwd.reverse_flag), conv_mode, padding_mode, dtype,
               host=not torch_in)
    d.algo = _ALGOS[algo]
    info = C.PlanInfo()
    _check(lib.ndc_plan(ctypes.byref(d), op, ctypes.byref(info)))
    out_shape = tuple(int(info.out_shape[i]) for i in range(len(x.shape)))
    h = proc.handle
    fn = lib.ndc_conv if op == C.OP_CONV else lib.ndc_conv_fft
    if torch_in:
        y = torch.empty(out_shape, dtype=x.dtype, device=x.device)
        _check(lib.ndc_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)), h)
        _check(fn(h, ctypes.byref(d), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(k.data_ptr()),
                  ctypes.c_void_p(y.data_ptr())), h)
        return y
    y = np.empty(out_shape, dtype=dt)
    _check(lib.ndc_reset_stream(h), h)
    _check(fn(h, ctypes.byref(d), ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(k.ctypes.data),
              ctypes.c_void_p(y.ctypes.data)), h)
    return y


def conv(data, kernel, conv_mode, padding_mode, algo="auto", processor=None):
    """ConvExt::conv (conv/mod.rs:110-115).  algo: "auto"/"direct" (exact, in
    the element type), "bf16" or "bf16x3" (f32 rank 2 on the matrix cores;
    "bf16x3" runs kernels of fewer than 32 taps exactly, in FP32)."""
    return _run(C.OP_CONV, data, kernel, conv_mode, padding_mode, processor, algo)


def conv_fft(data, kernel, conv_mode, padding_mode):
    """ConvFFTExt::conv_fft (conv_fft/mod.rs:113-118)."""
    return _run(C.OP_CONV_FFT, data, kernel, conv_mode, padding_mode)


def conv_fft_with_processor(data, kernel, conv_mode, padding_mode, fft_processor):
    """ConvFFTExt::conv_fft_with_processor (conv_fft/mod.rs:151-157)."""
    return _run(C.OP_CONV_FFT, data, kernel, conv_mode, padding_mode, fft_processor)


class Baked:
    """A kernel spectrum kept for repeated conv_fft of same-shaped data with
    one kernel: the reference's sketched Baked (conv_fft/mod.rs:22-42,306-392).
    The data given to conv_fft_bake fixes the shape, strides, dtype and memory
    kind (numpy or torch GPU) that conv_fft_with_baked accepts."""

    def __init__(self, data, kernel, conv_mode, padding_mode, processor=None):
        lib = C.lib()
        self._proc = processor or _default_processor()
        kwd = _into_kwd(kernel)
        self._torch = _is_torch(data) or _is_torch(kwd.kernel)
        if self._torch:
            if not (_is_torch(data) and _is_torch(kwd.kernel)) or not data.is_cuda or not kwd.kernel.is_cuda:
                raise InvalidArgument("torch inputs must both be GPU tensors")
            x, k = data, kwd.kernel
        else:
            x, k = np.asarray(data), np.asarray(kwd.kernel)
            self._k = k   # keeps the host kernel alive for the duration of the bake
        dt = _np_dtype_of(x)
        if dt != _np_dtype_of(k):
            raise InvalidArgument("data and kernel dtypes differ")
        if dt not in _NP_DTYPE:
            raise Unsupported(f"dtype {dt}")
        d = C.ConvDesc()
        _fill_desc(d, x, KernelWithDilation(k, kwd.dilation, kwd.reverse_flag), conv_mode, padding_mode, _NP_DTYPE[dt],
                   host=not self._torch)
        info = C.PlanInfo()
        _check(lib.ndc_plan(ctypes.byref(d), C.OP_CONV_FFT, ctypes.byref(info)))
        self.out_shape = tuple(int(info.out_shape[i]) for i in range(len(x.shape)))
        self._sig = (tuple(x.shape), tuple(x.stride() if self._torch else x.strides), dt)
        self._dt = dt
        h = self._proc.handle
        b = ctypes.c_void_p()
        if self._torch:
            import torch
            _check(lib.ndc_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)), h)
            _check(lib.ndc_conv_fft_bake(h, ctypes.byref(d), ctypes.c_void_p(k.data_ptr()), ctypes.byref(b)), h)
        else:
            _check(lib.ndc_reset_stream(h), h)
            _check(lib.ndc_conv_fft_bake(h, ctypes.byref(d), ctypes.c_void_p(k.ctypes.data), ctypes.byref(b)), h)
        self._b = b

    def conv(self, data):
        lib = C.lib()
        h = self._proc.handle
        if _is_torch(data) != self._torch:
            raise InvalidArgument("data must be of the memory kind the kernel was baked for")
        x = data if self._torch else np.asarray(data)
        if (tuple(x.shape), tuple(x.stride() if self._torch else x.strides), _np_dtype_of(x)) != self._sig:
            raise InvalidArgument("data must have the shape, strides and dtype the kernel was baked for")
        if self._torch:
            import torch
            y = torch.empty(self.out_shape, dtype=x.dtype, device=x.device)
            _check(lib.ndc_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)), h)
            _check(lib.ndc_conv_fft_with_baked(self._b, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())), h)
            return y
        y = np.empty(self.out_shape, dtype=self._dt)
        _check(lib.ndc_reset_stream(h), h)
        _check(lib.ndc_conv_fft_with_baked(self._b, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data)), h)
        return y

    def close(self):
        if getattr(self, "_b", None):
            C.lib().ndc_baked_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def conv_fft_bake(data, kernel, conv_mode, padding_mode):
    """ConvFFTExt::conv_fft_bake as sketched in conv_fft/mod.rs:175-182,306-360."""
    return Baked(data, kernel, conv_mode, padding_mode)


def conv_fft_with_baked(data, baked):
    """ConvFFTExt::conv_fft_with_baked as sketched in conv_fft/mod.rs:362-392."""
    return baked.conv(data)


def conv_fft_par(data, kernel, conv_mode, padding_mode):
    """ConvFFTExt::conv_fft_par (conv_fft/mod.rs:166-171): the same transform;
    the GPU grid replaces rayon."""
    return _run(C.OP_CONV_FFT, data, kernel, conv_mode, padding_mode)


# ---------------------------------------------------------------------------
# Batches of independent rank-1 pairs (one conv_fft call per pair in the
# reference; one persistent launch here)
# ---------------------------------------------------------------------------

_BATCH_ALGOS = {"auto": C.BATCH_AUTO, "fft": C.BATCH_FFT, "direct": C.BATCH_DIRECT, "plain": C.BATCH_PLAIN,
                "bf16x3": C.BATCH_MFMA, "auto_bf16x3": C.BATCH_AUTO_BF16X3, "bf16x6": C.BATCH_BF16X6}


def _batch_desc(x_lens, k_lens, conv_mode, padding_mode, dilation, reverse_flag, algo, host):
    xl = np.ascontiguousarray(x_lens, dtype=np.int64)
    kl = np.ascontiguousarray(k_lens, dtype=np.int64)
    if xl.ndim != 1 or xl.shape != kl.shape:
        raise InvalidArgument("x_lens and k_lens must be 1-D of equal length")
    tmpl = _desc_for_shapes((1,), (1,), conv_mode, padding_mode, dilation, reverse_flag)
    d = C.BatchDesc()
    d.dtype = C.F32
    if algo not in _BATCH_ALGOS:
        raise InvalidArgument(f"unknown batch algorithm {algo!r}")
    d.algo = _BATCH_ALGOS[algo]
    d.count = int(xl.shape[0])
    d.x_len = xl.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    d.k_len = kl.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    d.dilation = int(dilation)
    d.reverse = 1 if reverse_flag else 0
    d.flags = C.FLAG_HOST_MEMORY if host else 0
    d.conv_mode = tmpl.conv_mode
    d.padding_mode = tmpl.padding_mode
    return d, xl, kl


class BatchPlan:
    """Many independent rank-1 f32 signal/kernel pairs of uneven lengths.

    Each pair has the semantics of one conv_fft call (conv_fft/mod.rs:113-118)
    with the shared ConvMode, PaddingMode, dilation and reverse flag.  Signals,
    kernels and outputs are packed back to back in pair order.  algo: "auto"
    (a device-measured cost model picks FP32 direct or FP32 overlap-save FFT per
    pair), "auto_bf16x3" (the same choice with the matrix-core direct forms:
    BF16x3 for kernels of at least 32 taps, BF16x6 below), "fft", "direct"
    (FP32), "bf16x3" (matrix cores for every eligible pair, BF16x3 or BF16x6 by
    the same tap rule), "bf16x6" (the three-term split for every eligible
    pair), or "plain" (one single-problem conv_fft per pair).
    host=True takes numpy arrays; otherwise torch GPU tensors, ordered on
    torch's current stream."""

    def __init__(self, x_lens, k_lens, conv_mode, padding_mode, dilation=1, reverse_flag=True, algo="auto",
                 host=False, processor=None):
        self._proc = processor or _default_processor()
        self.host = bool(host)
        lib = C.lib()
        h = self._proc.handle
        d, self.x_lens, self.k_lens = _batch_desc(x_lens, k_lens, conv_mode, padding_mode, dilation, reverse_flag,
                                                  algo, host)
        b = ctypes.c_void_p()
        _check(lib.ndc_batch_create(h, ctypes.byref(d), ctypes.byref(b)), h)
        self._b = b
        n = int(self.x_lens.shape[0])
        self.y_lens = np.zeros(n, dtype=np.int64)
        xt, kt, yt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib.ndc_batch_out_lens(b, self.y_lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                      ctypes.byref(xt), ctypes.byref(kt), ctypes.byref(yt)), h)
        self.x_total, self.k_total, self.y_total = xt.value, kt.value, yt.value

    def _offsets(self, lens):
        o = np.zeros(len(lens) + 1, dtype=np.int64)
        np.cumsum(lens, out=o[1:])
        return o

    def run(self, x, k, y=None, graph=False):
        """Convolves the packed batch; returns the packed outputs."""
        lib = C.lib()
        h = self._proc.handle
        fn = lib.ndc_batch_run_graph if graph else lib.ndc_batch_run
        if self.host:
            x = np.ascontiguousarray(x, dtype=np.float32)
            k = np.ascontiguousarray(k, dtype=np.float32)
            if x.size != self.x_total or k.size != self.k_total:
                raise InvalidArgument("packed sizes differ from the plan")
            if y is None:
                y = np.empty(self.y_total, dtype=np.float32)
            _check(lib.ndc_reset_stream(h), h)
            _check(fn(self._b, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(k.ctypes.data),
                      ctypes.c_void_p(y.ctypes.data)), h)
            return y
        import torch
        if not (_is_torch(x) and _is_torch(k)) or not x.is_cuda or not k.is_cuda:
            raise InvalidArgument("a device batch takes torch GPU tensors")
        if x.dtype != torch.float32 or k.dtype != torch.float32 or not x.is_contiguous() or not k.is_contiguous():
            raise InvalidArgument("packed inputs must be contiguous float32")
        if x.numel() != self.x_total or k.numel() != self.k_total:
            raise InvalidArgument("packed sizes differ from the plan")
        if y is None:
            y = torch.empty(self.y_total, dtype=torch.float32, device=x.device)
        _check(lib.ndc_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)), h)
        _check(fn(self._b, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(k.data_ptr()),
                  ctypes.c_void_p(y.data_ptr())), h)
        return y

    def split(self, y):
        """The per-pair outputs of a packed result."""
        o = self._offsets(self.y_lens)
        return [y[o[i]:o[i + 1]] for i in range(len(self.y_lens))]

    def report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        _check(C.lib().ndc_batch_report(self._b, buf, len(buf)), self._proc.handle)
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "_b", None):
            C.lib().ndc_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def conv_fft_batch(xs, ks, conv_mode, padding_mode, dilation=1, reverse_flag=True, algo="auto"):
    """conv_fft of every (xs[i], ks[i]) pair (rank-1 f32 numpy arrays), as one
    batch; returns the list of outputs."""
    xs = [np.ascontiguousarray(a, dtype=np.float32).ravel() for a in xs]
    ks = [np.ascontiguousarray(a, dtype=np.float32).ravel() for a in ks]
    bp = BatchPlan([a.size for a in xs], [a.size for a in ks], conv_mode, padding_mode, dilation, reverse_flag,
                   algo, host=True)
    try:
        x = np.concatenate(xs) if xs else np.zeros(0, np.float32)
        k = np.concatenate(ks) if ks else np.zeros(0, np.float32)
        return [a.copy() for a in bp.split(bp.run(x, k))]
    finally:
        bp.close()


def batch_border_map(x_lens, k_lens, pair, conv_mode, padding_mode, dilation=1):
    """Closed-form border map of one pair of a batch (see ndc_batch_border_map)."""
    d, xl, kl = _batch_desc(x_lens, k_lens, conv_mode, padding_mode, dilation, True, "auto", False)
    info = plan((int(xl[pair]),), (int(kl[pair]),), conv_mode, padding_mode, dilation)
    n = info["padded"][0]
    out = np.zeros(n, dtype=np.int32)
    _check(C.lib().ndc_batch_border_map(ctypes.byref(d), int(pair),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n))
    return out
