"""ndc_conv_fft_noalloc through the C ABI: the allocation-free variant of
ndc_conv_fft (its workspace carved from a library-owned device array) must
give exactly ndc_conv_fft's results, also on a second, smaller call right
after the first (the array keeps the first call's contents), and refuse a
convolution whose workspace exceeds the array."""
import ctypes

import numpy as np
import pytest

import ndconv as nc
from ndconv import _lib as C
from oracle import ndconv_oracle as O

pytestmark = pytest.mark.gpu


def _run(fn, h, x, kwd, mode, pm):
    import torch
    lib = C.lib()
    d = C.ConvDesc()
    nc._fill_desc(d, x, kwd, mode, pm, nc._NP_DTYPE[nc._np_dtype_of(x)], host=False)
    info = C.PlanInfo()
    nc._check(lib.ndc_plan(ctypes.byref(d), C.OP_CONV_FFT, ctypes.byref(info)))
    shape = tuple(int(info.out_shape[i]) for i in range(len(x.shape)))
    y = torch.full(shape, float("nan"), dtype=x.dtype, device=x.device)
    nc._check(lib.ndc_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)), h)
    st = fn(h, ctypes.byref(d), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(kwd.kernel.data_ptr()),
            ctypes.c_void_p(y.data_ptr()))
    return st, y


def test_noalloc_matches_conv_fft(gpu):
    import torch
    lib = C.lib()
    proc = nc.get_fft_processor()
    h = proc.handle
    rng = np.random.default_rng(17)
    B = nc.BorderType
    cases = [
        # the 2-D workload, then a smaller call on the same array
        ((200, 5000), (11, 31), 2, nc.ConvMode.Same, nc.PaddingMode.Custom([B.Reflect, B.Circular]), np.float32),
        ((37, 300), (5, 9), 1, nc.ConvMode.Full, nc.PaddingMode.Zeros, np.float32),
        ((10, 100, 200), (5, 11, 31), 1, nc.ConvMode.Full, nc.PaddingMode.Zeros, np.complex64),
        ((5000,), (31,), 1, nc.ConvMode.Same, nc.PaddingMode.Zeros, np.float32),
        ((3, 20000), (2, 9), 1, nc.ConvMode.Same, nc.PaddingMode.Circular, np.float32),   # four-step rows
        ((64, 90), (7, 9), 3, nc.ConvMode.Custom([2, 3], [2, 1]), nc.PaddingMode.Replicate, np.float64),
    ]
    for xs, ks, dil, mode, pm, dt in cases:
        x = rng.standard_normal(xs)
        k = rng.standard_normal(ks)
        if np.dtype(dt).kind == "c":
            x = x + 1j * rng.standard_normal(xs)
            k = k + 1j * rng.standard_normal(ks)
        x, k = x.astype(dt), k.astype(dt)
        xt, kt = torch.from_numpy(x).cuda(), torch.from_numpy(k).cuda()
        kwd = nc.with_dilation(kt, dil)
        st_a, y_a = _run(lib.ndc_conv_fft, h, xt, kwd, mode, pm)
        st_b, y_b = _run(lib.ndc_conv_fft_noalloc, h, xt, kwd, mode, pm)
        torch.cuda.synchronize()
        assert st_a == C.OK and st_b == C.OK, (xs, st_a, st_b, lib.ndc_last_error(h))
        assert torch.equal(y_a, y_b), xs
        # and against the oracle, as the allocating entry point is checked elsewhere
        wide = np.complex128 if np.dtype(dt).kind == "c" else np.float64
        ref = O.conv_fft(x.astype(wide), k.astype(wide), _mode(mode), _pad(pm), dil)
        tol = 1e-5 if np.dtype(dt).itemsize <= 8 and dt != np.float64 else 1e-12
        assert O.relerr(y_b.cpu().numpy(), ref)[0] <= tol, xs
    proc.close()


def test_noalloc_refuses_oversize_and_host_memory(gpu):
    import torch
    lib = C.lib()
    proc = nc.get_fft_processor()
    h = proc.handle
    xt = torch.zeros((4096, 4096), dtype=torch.float32, device="cuda")
    kt = torch.zeros((31, 31), dtype=torch.float32, device="cuda")
    st, _ = _run(lib.ndc_conv_fft_noalloc, h, xt, nc.with_dilation(kt, 1), nc.ConvMode.Same, nc.PaddingMode.Zeros)
    assert st == C.ERR_UNSUPPORTED
    assert b"bytes" in lib.ndc_last_error(h)
    d = C.ConvDesc()
    x = np.zeros((8, 8), np.float32)
    nc._fill_desc(d, x, nc.with_dilation(x[:3, :3], 1), nc.ConvMode.Same, nc.PaddingMode.Zeros, C.F32, host=True)
    y = np.zeros((8, 8), np.float32)
    assert lib.ndc_conv_fft_noalloc(h, ctypes.byref(d), ctypes.c_void_p(x.ctypes.data),
                                    ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data)) == \
        C.ERR_INVALID_ARGUMENT
    proc.close()


def _mode(m):
    if m is nc.ConvMode.Full:
        return "full"
    if m is nc.ConvMode.Same:
        return "same"
    if m is nc.ConvMode.Valid:
        return "valid"
    return ("custom", m.padding, m.strides)


def _pad(pm):
    if pm is nc.PaddingMode.Zeros:
        return "zeros"
    if pm is nc.PaddingMode.Circular:
        return "circular"
    if pm is nc.PaddingMode.Replicate:
        return "replicate"
    names = {"Reflect": "reflect", "Circular": "circular", "Replicate": "replicate", "Zeros": "zeros"}
    return ("custom", [names[b.name] for b in pm.borders])
