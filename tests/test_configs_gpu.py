"""The benchmark workloads at their full sizes, seeded as the measurement
plan fixes them (seed = workload number, numpy PCG64 default_rng):

  1-D f32 conv_fft   x 5000, k 31, Same, Zeros
  2-D f32 conv_fft   x 200x5000, k 11x31, Same, Custom[Reflect, Circular], dilation 2
  2-D i32 conv       same shapes, values in -8..8, no dilation (bit-exact)
  3-D f32 conv_fft   x 10x100x200, k 5x11x31, Full, Zeros; also complex64 N(0,1)
  2-D f32 conv       x 4096x4096, k 31x31, Valid, Replicate, dilation 2, stride 2

Floating point against the FP64 oracle (normwise relative error <= 1e-5),
integers exact; plus size-independent properties (linearity, shift)."""
import numpy as np
import pytest

from oracle import ndconv_oracle as O

import ndconv as nc

pytestmark = pytest.mark.gpu

B = nc.BorderType
SAME, FULL = nc.ConvMode.Same, nc.ConvMode.Full


def workload1():
    rng = np.random.default_rng(1)
    return rng.random(5000).astype(np.float32), rng.random(31).astype(np.float32)


def workload2():
    rng = np.random.default_rng(2)
    return rng.random((200, 5000)).astype(np.float32), rng.random((11, 31)).astype(np.float32)


def workload2_int():
    rng = np.random.default_rng(2)
    return (rng.integers(-8, 9, (200, 5000)).astype(np.int32), rng.integers(-8, 9, (11, 31)).astype(np.int32))


def workload3(complex_=False):
    rng = np.random.default_rng(3)
    if complex_:
        x = (rng.standard_normal((10, 100, 200)) + 1j * rng.standard_normal((10, 100, 200))).astype(np.complex64)
        k = (rng.standard_normal((5, 11, 31)) + 1j * rng.standard_normal((5, 11, 31))).astype(np.complex64)
        return x, k
    return rng.random((10, 100, 200)).astype(np.float32), rng.random((5, 11, 31)).astype(np.float32)


def workload4():
    rng = np.random.default_rng(4)
    return (rng.standard_normal((4096, 4096)).astype(np.float32),
            rng.standard_normal((31, 31)).astype(np.float32))


def test_workload1_fft(gpu):
    x, k = workload1()
    got = nc.conv_fft(x, k, SAME, nc.PaddingMode.Zeros)
    ref = O.conv_fft(x.astype(np.float64), k.astype(np.float64), "same", "zeros")
    assert got.shape == (5000,)
    assert O.relerr(got, ref)[0] <= 1e-5


def test_workload2_fft(gpu):
    x, k = workload2()
    pm = nc.PaddingMode.Custom([B.Reflect, B.Circular])
    got = nc.conv_fft(x, nc.with_dilation(k, 2), SAME, pm)
    ref = O.conv_fft(x.astype(np.float64), k.astype(np.float64), "same", ("custom", ["reflect", "circular"]), 2)
    assert got.shape == (200, 5000)
    r, _ = O.relerr(got, ref)
    assert r <= 1e-5, r


def test_workload2_int_exact(gpu):
    x, k = workload2_int()
    pm = nc.PaddingMode.Custom([B.Reflect, B.Circular])
    got = nc.conv(x, k, SAME, pm)
    ref = O.conv(x, k, "same", ("custom", ["reflect", "circular"]))
    np.testing.assert_array_equal(got, ref)
    # wrapping is order independent: int64 exact reduced mod 2^32 agrees too
    np.testing.assert_array_equal(got.astype(np.int64), O.conv_int_exact(x, k, "same",
                                                                          ("custom", ["reflect", "circular"])))


@pytest.mark.parametrize("complex_", [False, True])
def test_workload3_fft(gpu, complex_):
    x, k = workload3(complex_)
    got = nc.conv_fft(x, k, FULL, nc.PaddingMode.Zeros)
    dt = np.complex128 if complex_ else np.float64
    ref = O.conv_fft(x.astype(dt), k.astype(dt), "full", "zeros")
    assert got.shape == (14, 110, 230)
    assert O.relerr(got, ref)[0] <= 1e-5


def test_workload4_direct_fp32(gpu):
    x, k = workload4()
    mode = nc.ConvMode.Custom([0, 0], [2, 2])
    got = nc.conv(x, nc.with_dilation(k, 2), mode, nc.PaddingMode.Replicate)
    ref = O.conv_fft(x.astype(np.float64), k.astype(np.float64), ("custom", [0, 0], [2, 2]), "replicate", 2)
    assert got.shape == (2018, 2018)
    r, _ = O.relerr(got, ref)
    assert r <= 1e-5, r


def test_linearity_full_size(gpu):
    # conv(a x1 + b x2) == a conv(x1) + b conv(x2) at the 2-D workload size
    x1, k = workload2()
    rng = np.random.default_rng(22)
    x2 = rng.random(x1.shape).astype(np.float32)
    pm = nc.PaddingMode.Custom([B.Reflect, B.Circular])
    kw = nc.with_dilation(k, 2)
    y1 = nc.conv_fft(x1, kw, SAME, pm).astype(np.float64)
    y2 = nc.conv_fft(x2, kw, SAME, pm).astype(np.float64)
    y3 = nc.conv_fft((0.5 * x1 - 2.0 * x2).astype(np.float32), kw, SAME, pm).astype(np.float64)
    assert O.relerr(y3, 0.5 * y1 - 2.0 * y2)[0] < 1e-5


def test_impulse_recovers_kernel(gpu):
    # a unit impulse convolved (Full) with k returns k itself, direct and FFT
    rng = np.random.default_rng(23)
    k = rng.standard_normal((11, 31)).astype(np.float32)
    x = np.zeros((50, 80), np.float32)
    x[20, 30] = 1
    for f in (nc.conv, nc.conv_fft):
        y = f(x, k, FULL, nc.PaddingMode.Zeros)
        np.testing.assert_allclose(y[20:31, 30:61], k, atol=1e-5)
        y[20:31, 30:61] = 0
        assert np.abs(y).max() < 1e-5
