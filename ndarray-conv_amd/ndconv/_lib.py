"""ctypes binding of include/ndconv.h (the library's C ABI)."""
import ctypes
import os

MAX_RANK = 6

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("NDCONV_LIB", os.path.join(_PKG, "lib", "libndconv.so"))

# ndc_status
OK, ERR_DATA_SHAPE, ERR_KERNEL_SHAPE, ERR_MISMATCH_SHAPE, ERR_INVALID_ARGUMENT, ERR_UNSUPPORTED, \
    ERR_NO_DEVICE, ERR_DEVICE, ERR_OUT_OF_MEMORY = range(9)
# ndc_dtype
F32, F64, C64, C128, I8, I16, I32, I64, U8, U16, U32, U64 = range(12)
# ndc_conv_mode_tag
CONV_FULL, CONV_SAME, CONV_VALID, CONV_CUSTOM, CONV_EXPLICIT = range(5)
# ndc_border_tag
BORDER_ZEROS, BORDER_CONST, BORDER_REFLECT, BORDER_REPLICATE, BORDER_CIRCULAR = range(5)
# ndc_padding_tag
PAD_ZEROS, PAD_CONST, PAD_REFLECT, PAD_REPLICATE, PAD_CIRCULAR, PAD_CUSTOM, PAD_EXPLICIT = range(7)
# ndc_algo
ALGO_AUTO, ALGO_DIRECT, ALGO_MFMA_BF16, ALGO_MFMA_BF16X3 = range(4)
# flags
FLAG_HOST_MEMORY, FLAG_X_STRIDED, FLAG_K_STRIDED = 1, 2, 4
# ndc_op
OP_CONV, OP_CONV_FFT = 0, 1
# ndc_batch_algo
BATCH_AUTO, BATCH_FFT, BATCH_DIRECT, BATCH_PLAIN, BATCH_MFMA, BATCH_AUTO_BF16X3, BATCH_BF16X6 = range(7)

I64A = ctypes.c_int64 * MAX_RANK


class Value(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double), ("i", ctypes.c_int64)]


class Border(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int32), ("reserved", ctypes.c_int32), ("value", Value)]


class ConvModeC(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("padding", (ctypes.c_int64 * 2) * MAX_RANK), ("strides", I64A)]


class PaddingModeC(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int32), ("reserved", ctypes.c_int32), ("value", Value),
                ("borders", (Border * 2) * MAX_RANK)]


class ConvDesc(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("rank", ctypes.c_int32),
                ("x_shape", I64A), ("x_strides", I64A), ("k_shape", I64A), ("k_strides", I64A),
                ("dilation", I64A), ("reverse", ctypes.c_int32), ("algo", ctypes.c_int32),
                ("conv_mode", ConvModeC), ("padding_mode", PaddingModeC),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("algo", ctypes.c_int32), ("count", ctypes.c_int64),
                ("x_len", ctypes.POINTER(ctypes.c_int64)), ("k_len", ctypes.POINTER(ctypes.c_int64)),
                ("dilation", ctypes.c_int64), ("reverse", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("conv_mode", ConvModeC), ("padding_mode", PaddingModeC)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("padding", (ctypes.c_int64 * 2) * MAX_RANK), ("strides", I64A), ("kernel_dilated", I64A),
                ("padded", I64A), ("fft_size", I64A), ("out_shape", I64A)]


_lib = None


def lib():
    """Loads libndconv.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libndconv.so not found at {LIB_PATH}: build it with "
                           f"`python ndarray-conv_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    P = ctypes.c_void_p
    L.ndc_abi_version.restype = ctypes.c_int32
    L.ndc_struct_size.restype = ctypes.c_int64
    L.ndc_struct_size.argtypes = [ctypes.c_int32]
    L.ndc_status_string.restype = ctypes.c_char_p
    L.ndc_status_string.argtypes = [ctypes.c_int]
    L.ndc_create.restype = ctypes.c_int
    L.ndc_create.argtypes = [ctypes.POINTER(H)]
    L.ndc_destroy.restype = ctypes.c_int
    L.ndc_destroy.argtypes = [H]
    L.ndc_set_stream.restype = ctypes.c_int
    L.ndc_set_stream.argtypes = [H, P]
    L.ndc_reset_stream.restype = ctypes.c_int
    L.ndc_reset_stream.argtypes = [H]
    L.ndc_profile_enable.restype = ctypes.c_int
    L.ndc_profile_enable.argtypes = [H, ctypes.c_int32]
    L.ndc_profile_report.restype = ctypes.c_int
    L.ndc_profile_report.argtypes = [H, ctypes.c_char_p, ctypes.c_int64]
    L.ndc_synchronize.restype = ctypes.c_int
    L.ndc_synchronize.argtypes = [H]
    L.ndc_last_error.restype = ctypes.c_char_p
    L.ndc_last_error.argtypes = [H]
    L.ndc_plan.restype = ctypes.c_int
    L.ndc_plan.argtypes = [ctypes.POINTER(ConvDesc), ctypes.c_int32, ctypes.POINTER(PlanInfo)]
    L.ndc_border_map.restype = ctypes.c_int
    L.ndc_border_map.argtypes = [ctypes.POINTER(ConvDesc), ctypes.c_int32, ctypes.c_int32,
                                 ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
    for name in ("ndc_conv", "ndc_conv_fft", "ndc_conv_fft_noalloc"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [H, ctypes.POINTER(ConvDesc), P, P, P]
    B = ctypes.c_void_p
    I64P = ctypes.POINTER(ctypes.c_int64)
    for name in ("ndc_fft_forward", "ndc_fft_backward"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [H, ctypes.c_int32, ctypes.c_int32, I64P, P, P, ctypes.c_uint32]
    L.ndc_conv_fft_launches.restype = ctypes.c_int
    L.ndc_conv_fft_launches.argtypes = [ctypes.POINTER(ConvDesc), ctypes.c_int64, ctypes.c_int32, I64P, I64P, I64P,
                                        ctypes.c_int64, I64P]
    L.ndc_conv_fft_bake.restype = ctypes.c_int
    L.ndc_conv_fft_bake.argtypes = [H, ctypes.POINTER(ConvDesc), P, ctypes.POINTER(B)]
    L.ndc_conv_fft_with_baked.restype = ctypes.c_int
    L.ndc_conv_fft_with_baked.argtypes = [B, P, P]
    L.ndc_baked_destroy.restype = ctypes.c_int
    L.ndc_baked_destroy.argtypes = [B]
    L.ndc_fft_origin_len.restype = ctypes.c_int
    L.ndc_fft_origin_len.argtypes = [H, I64P]
    L.ndc_batch_create.restype = ctypes.c_int
    L.ndc_batch_create.argtypes = [H, ctypes.POINTER(BatchDesc), ctypes.POINTER(B)]
    L.ndc_batch_destroy.restype = ctypes.c_int
    L.ndc_batch_destroy.argtypes = [B]
    L.ndc_batch_out_lens.restype = ctypes.c_int
    L.ndc_batch_out_lens.argtypes = [B, I64P, I64P, I64P, I64P]
    for name in ("ndc_batch_run", "ndc_batch_run_graph"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [B, P, P, P]
    L.ndc_batch_report.restype = ctypes.c_int
    L.ndc_batch_report.argtypes = [B, ctypes.c_char_p, ctypes.c_int64]
    L.ndc_batch_border_map.restype = ctypes.c_int
    L.ndc_batch_border_map.argtypes = [ctypes.POINTER(BatchDesc), ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.c_int64]
    _lib = L
    return L
