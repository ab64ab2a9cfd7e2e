"""The FFT processor's forward/backward on the GPU (Processor::forward /
backward, processor/mod.rs:79-121) against the oracle's restatement of the
reference's layout: the N-d DFT with the axes rotated left by one, real input
keeping F/2+1 bins of its last axis; backward returns natural order divided by
the element count."""
import ctypes
import json

import numpy as np
import pytest

from oracle import ndconv_oracle as O

import ndconv as nc
from ndconv import _lib as C

pytestmark = pytest.mark.gpu

TOL = {np.dtype(np.float32): 2e-6, np.dtype(np.float64): 1e-12, np.dtype(np.complex64): 2e-6,
       np.dtype(np.complex128): 1e-12}

SHAPES = [(6,), (7,), (425,), (5120,), (4, 6), (9, 35), (128, 145), (18, 128, 288), (3, 5, 7), (2, 3, 4, 5),
          (256, 5120), (16, 1), (5, 1, 3), (1, 7)]


def make(rng, shape, dt):
    dt = np.dtype(dt)
    x = rng.standard_normal(shape)
    if dt.kind == "c":
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(dt)


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.complex64, np.complex128])
def test_forward_backward_layout_and_values(gpu, dt):
    rng = np.random.default_rng(5)
    for shape in SHAPES:
        x = make(rng, shape, dt)
        p = nc.get_fft_processor(dt)
        spec = p.forward(x)
        ref = O.processor_forward(x.astype(np.complex128 if np.dtype(dt).kind == "c" else np.float64))
        assert spec.shape == ref.shape, (shape, spec.shape, ref.shape)
        e = O.relerr(spec, ref)[0]
        assert e <= TOL[np.dtype(dt)] * (10 if len(shape) > 2 else 1), (shape, e)
        back = p.backward(spec)
        assert back.shape == x.shape and back.dtype == x.dtype
        e2 = O.relerr(back, x)[0]
        assert e2 <= 5 * TOL[np.dtype(dt)] * (10 if len(shape) > 2 else 1), (shape, e2)
        p.close()


def test_reference_roundtrip_cases_on_gpu(gpu):
    # real.rs roundtrip tests: 1e-10 for f64 inputs
    cases = [np.array([1.0, 2, 3, 4, 5, 6]), np.array([1.0, 2]), np.array([1.0, 2, 3]),
             np.array([1.0, 2, 3, 4, 5]), np.fromfunction(lambda i, j: (i + j) % 17, (32, 32)),
             np.fromfunction(lambda i, j, k: (i + j + k) % 13, (8, 8, 8))]
    p = nc.get_fft_processor(np.float64)
    for x in cases:
        back = p.backward(p.forward(x))
        assert np.max(np.abs(back - x)) < 1e-10


def test_backward_uses_the_last_forward_length(gpu):
    # rp_origin_len: the real length recorded by the last forward (real.rs:108,263)
    rng = np.random.default_rng(6)
    p = nc.get_fft_processor(np.float32)
    x6 = rng.standard_normal((3, 6)).astype(np.float32)
    x7 = rng.standard_normal((3, 7)).astype(np.float32)
    s6 = p.forward(x6)
    s7 = p.forward(x7)   # both have 4 bins: the last forward decides the output length
    assert s6.shape == s7.shape == (4, 3)
    assert p.backward(s6).shape == (3, 7)
    assert O.relerr(p.backward(s7), x7)[0] <= 1e-5
    with pytest.raises(nc.InvalidArgument):
        nc.get_fft_processor(np.float32).backward(s7)   # no forward on this processor yet
    bad = np.zeros((6, 3), np.complex64)                # 6 bins do not belong to length 7
    with pytest.raises(nc.InvalidArgument):
        p.backward(bad)


def test_torch_device_tensors(gpu):
    import torch
    rng = np.random.default_rng(7)
    x = rng.standard_normal((64, 300)).astype(np.float32)
    p = nc.get_fft_processor(np.float32)
    spec = p.forward(torch.from_numpy(x).cuda())
    ref = O.processor_forward(x.astype(np.float64))
    assert O.relerr(spec.cpu().numpy(), ref)[0] <= 2e-6
    back = p.backward(spec)
    assert O.relerr(back.cpu().numpy(), x)[0] <= 1e-5


# Lines longer than one workgroup's LDS holds (about 9000 complex64 or 4800
# complex128 points) are split into two LDS-resident factors; real lines go
# through the half-length packing (even) or a complex line (odd).  A prime
# length is transformed by the generic radix stage while it fits one
# workgroup (4999 here).
LONG = [((5120,), np.float64), ((256, 5120), np.float64), ((9600,), np.complex64), ((3, 12000), np.complex64),
        ((16384,), np.complex128), ((2, 40000), np.float32), ((10125,), np.float32), ((3, 10125), np.float64),
        ((20000, 3), np.complex64), ((6000, 4), np.float64), ((7, 4999), np.float32)]


@pytest.mark.parametrize("shape,dt", LONG, ids=lambda v: str(v))
def test_long_lines(gpu, shape, dt):
    rng = np.random.default_rng(sum(shape))
    x = make(rng, shape, dt)
    p = nc.get_fft_processor(dt)
    spec = p.forward(x)
    wide = np.complex128 if np.dtype(dt).kind == "c" else np.float64
    ref = O.processor_forward(x.astype(wide))
    assert spec.shape == ref.shape
    tol = TOL[np.dtype(dt)] * 10
    assert O.relerr(spec, ref)[0] <= tol, shape
    back = p.backward(spec)
    assert back.shape == x.shape and back.dtype == x.dtype
    assert O.relerr(back, x)[0] <= 5 * tol, shape
    p.close()


# Lengths with no LDS-resident split (a prime beyond one workgroup's LDS, or
# twice one for a real line's half-length packing) take the chirp-z form:
# complex lines, strided columns, odd and even real lines, both directions.
CHIRP = [((10007,), np.complex64), ((10007,), np.complex128), ((3, 10007), np.float32), ((2, 20014), np.float64),
         ((10007, 3), np.complex128), ((19946,), np.complex64), ((4, 4801), np.float64), ((65537,), np.float32)]


@pytest.mark.parametrize("shape,dt", CHIRP, ids=lambda v: str(v))
def test_lines_without_a_split(gpu, shape, dt):
    rng = np.random.default_rng(sum(shape) + 1)
    x = make(rng, shape, dt)
    p = nc.get_fft_processor(dt)
    for rep in range(2):   # the chirp tables are made on the first call and reused
        if rep == 0:
            lib = C.lib()
            nc._check(lib.ndc_profile_enable(p.handle, 1), p.handle)
            spec = p.forward(x)
            buf = ctypes.create_string_buffer(1 << 16)
            nc._check(lib.ndc_profile_report(p.handle, buf, len(buf)), p.handle)
            nc._check(lib.ndc_profile_enable(p.handle, 0), p.handle)
            assert "fft_processor_chirp_io" in json.loads(buf.value.decode()), shape   # the chirp-z form ran
        else:
            spec = p.forward(x)
        wide = np.complex128 if np.dtype(dt).kind == "c" else np.float64
        ref = O.processor_forward(x.astype(wide))
        assert spec.shape == ref.shape
        tol = TOL[np.dtype(dt)] * 10
        assert O.relerr(spec, ref)[0] <= tol, (shape, rep)
        back = p.backward(spec)
        assert back.shape == x.shape and back.dtype == x.dtype
        assert O.relerr(back, x)[0] <= 5 * tol, (shape, rep)
    p.close()


def test_chirp_length_limit(gpu):
    # a prime beyond 2^23 would need a chirp-z convolution longer than 2^24
    p = nc.get_fft_processor(np.complex64)
    with pytest.raises(nc.Unsupported):
        p.forward(np.zeros(8388617, np.complex64))
