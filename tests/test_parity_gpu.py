"""Parity of the HIP path (through the C ABI) with the oracle.

Integers must be bit-exact.  Floating point is checked against the FP64
oracle with the normwise relative error max|y - y64| / max|y64|:
    f32 / c64  <= 1e-5   (the reference's TOLERANCE_F32, conv_fft/tests.rs:15)
    f64 / c128 <= 1e-12
"""
import numpy as np
import pytest

from golden_io import load, oracle_mode
from modes import conv_mode, padding_mode, random_case, rand_array
from oracle import ndconv_oracle as O

import ndconv as nc

pytestmark = pytest.mark.gpu

G = load("reference_goldens.json")
T = load("torch_cases.json")

TOL = {np.dtype(np.float32): 1e-5, np.dtype(np.complex64): 1e-5,
       np.dtype(np.float64): 1e-12, np.dtype(np.complex128): 1e-12}


def _f64(dt):
    return np.complex128 if np.dtype(dt).kind == "c" else np.float64


def check_conv(x, k, mode, pm, dil=1, rev=True, algo="auto"):
    kw = nc.with_dilation(k, dil)
    kw = kw if rev else kw.no_reverse()
    try:
        exp_err = None
        if x.dtype.kind in "iu":
            ref = O.conv(x, k, mode, pm, dil, rev)
        else:
            ref = O.conv(x, k, mode, pm, dil, rev, acc_dtype=_f64(x.dtype))
    except O.OracleError as e:
        exp_err = e.kind
    if exp_err is not None:
        with pytest.raises(getattr(nc, exp_err)):
            nc.conv(x, kw, conv_mode(mode), padding_mode(pm), algo=algo)
        return None
    got = nc.conv(x, kw, conv_mode(mode), padding_mode(pm), algo=algo)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    if x.dtype.kind in "iu":
        np.testing.assert_array_equal(got, ref)
        return 0.0
    r, _ = O.relerr(got, ref)
    assert r <= TOL[x.dtype], f"normwise rel err {r:.3e}"
    return r


def fft_error(got, ref, x, k):
    """Normwise relative error max|y - y64| / max|y64|.  When every tap lands
    on zero padding the exact result is zero and max|y64| is FP64 roundoff, so
    the error is taken relative to the problem's scale max|x| * sum|k| (the
    largest |y| inputs of that size can give) instead."""
    r, a = O.relerr(got, ref)
    scale = float(np.max(np.abs(x))) * float(np.sum(np.abs(k))) if x.size and k.size else 0.0
    m = float(np.max(np.abs(ref))) if ref.size else 0.0
    if scale > 0 and m <= 1e-9 * scale:
        return a / scale
    return r


def check_conv_fft(x, k, mode, pm, dil=1, rev=True):
    kw = nc.with_dilation(k, dil)
    kw = kw if rev else kw.no_reverse()
    try:
        exp_err = None
        ref = O.conv_fft(x.astype(_f64(x.dtype)), k.astype(_f64(x.dtype)), mode, pm, dil, rev)
    except O.OracleError as e:
        exp_err = e.kind
    if exp_err is not None:
        with pytest.raises(getattr(nc, exp_err)):
            nc.conv_fft(x, kw, conv_mode(mode), padding_mode(pm))
        return None
    got = nc.conv_fft(x, kw, conv_mode(mode), padding_mode(pm))
    assert got.shape == ref.shape and got.dtype == x.dtype
    r = fft_error(got, ref, x, k)
    assert r <= TOL[x.dtype], f"normwise rel err {r:.3e}"
    return r


# --- the reference's own cases -------------------------------------------------------

@pytest.mark.parametrize("case", G["conv"] + G["doc_cases"], ids=lambda c: c["src"])
def test_conv_goldens_gpu(gpu, case):
    x = np.array(case["x"], dtype=np.int32)
    k = np.array(case["k"], dtype=np.int32)
    kw = nc.with_dilation(k, case.get("dilation", 1))
    if not case.get("reverse", True):
        kw = kw.no_reverse()
    got = nc.conv(x, kw, conv_mode(case["mode"]), padding_mode(case["padmode"]))
    if "expect" in case:
        np.testing.assert_array_equal(got, np.array(case["expect"]))
    np.testing.assert_array_equal(got, O.conv(x, k, case["mode"], case["padmode"], case.get("dilation", 1),
                                              case.get("reverse", True)))


@pytest.mark.parametrize("case", T["conv"], ids=lambda c: c["src"])
def test_conv_vs_torch_gpu(gpu, case):
    x = np.array(case["x"], dtype=np.int32)
    k = np.array(case["k"], dtype=np.int32)
    kw = nc.with_dilation(k, case.get("dilation", 1))
    if not case.get("reverse", True):
        kw = kw.no_reverse()
    got = nc.conv(x, kw, conv_mode(oracle_mode(case["mode"])), nc.PaddingMode.Zeros)
    np.testing.assert_array_equal(got, np.array(case["expect"]))


@pytest.mark.parametrize("case", G["fft_vs_conv"], ids=lambda c: c["src"])
def test_fft_vs_conv_gpu(gpu, case):
    # conv_fft/tests.rs: the FFT result rounded equals the direct convolution
    dt = case["dtype"]
    xi, ki = np.array(case["x"]), np.array(case["k"])
    mode, pm = oracle_mode(case["mode"]), oracle_mode(case["padmode"])
    dil, rev = case.get("dilation", 1), case.get("reverse", True)
    kw = nc.with_dilation(ki.astype(dt), dil)
    kw = kw if rev else kw.no_reverse()
    f = nc.conv_fft(xi.astype(dt), kw, conv_mode(mode), padding_mode(pm))
    if case.get("no_round"):
        kd = nc.with_dilation(ki.astype(dt), dil)
        d = nc.conv(xi.astype(dt), kd if rev else kd.no_reverse(), conv_mode(mode), padding_mode(pm))
        assert np.max(np.abs(f - d)) < case["tol"]
    else:
        pmi = ("const", int(pm[1])) if isinstance(pm, tuple) else pm
        kd = nc.with_dilation(ki.astype(np.int32), dil)
        d = nc.conv(xi.astype(np.int32), kd if rev else kd.no_reverse(), conv_mode(mode), padding_mode(pmi))
        assert f.shape == d.shape
        assert np.all(np.abs(np.round(f) - d) < 1e-5)


@pytest.mark.parametrize("case", G["complex_cases"], ids=lambda c: c["src"])
def test_complex_cases_gpu(gpu, case):
    x = (np.array(case["x"])[:, 0] + 1j * np.array(case["x"])[:, 1]).reshape(case["xshape"])
    k = (np.array(case["k"])[:, 0] + 1j * np.array(case["k"])[:, 1]).reshape(case["kshape"])
    mode, pm = oracle_mode(case["mode"]), oracle_mode(case["padmode"])
    for dt in (np.complex64, np.complex128):
        check_conv_fft(x.astype(dt), k.astype(dt), mode, pm, case.get("dilation", 1), case.get("reverse", True))
        check_conv(x.astype(dt), k.astype(dt), mode, pm, case.get("dilation", 1), case.get("reverse", True))


# --- random cases over every mode, border, rank and dtype -----------------------------

# Every element type ConvExt is used with (any NumAssign + Copy, conv/mod.rs:118-127)
# and every rank the crate's impls exist for: they need Dim<[Ix; N]>: RemoveAxis,
# which ndarray provides for N = 1..6.
CONV_DTYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64,
               np.float32, np.float64, np.complex64, np.complex128]
RANKS = [1, 2, 3, 4, 5, 6]
MAX_X = {1: 8, 2: 8, 3: 5, 4: 4, 5: 3, 6: 3}


@pytest.mark.parametrize("dtype", CONV_DTYPES)
@pytest.mark.parametrize("rank", RANKS)
def test_conv_random(gpu, dtype, rank):
    rng = np.random.default_rng(1000 * rank + np.dtype(dtype).num)
    cval = 3 if np.dtype(dtype).kind in "iu" else 1.5
    for _ in range(25 if rank <= 3 else 12):
        xs, ks, dil, rev, mode, pm = random_case(rng, rank, cval, max_x=MAX_X[rank], max_k=4 if rank <= 3 else 3,
                                                 signed=np.dtype(dtype).kind != "u")
        x = rand_array(rng, xs, dtype)
        k = rand_array(rng, ks, dtype)
        check_conv(x, k, mode, pm, dil, rev)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
@pytest.mark.parametrize("rank", RANKS)
def test_conv_fft_random(gpu, dtype, rank):
    rng = np.random.default_rng(2000 * rank + np.dtype(dtype).num)
    for _ in range(25 if rank <= 3 else 12):
        xs, ks, dil, rev, mode, pm = random_case(rng, rank, 1.5, max_x=12 if rank < 3 else MAX_X[rank] + 1,
                                                 max_k=4 if rank <= 3 else 3)
        x = rand_array(rng, xs, dtype)
        k = rand_array(rng, ks, dtype)
        check_conv_fft(x, k, mode, pm, dil, rev)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 16, 17, 100, 384, 424, 425, 1537, 1705, 3073, 5030, 6145, 8000,
                               20000, 27000, 100000, 262144])
def test_conv_fft_1d_sizes(gpu, n):
    # covers the radix-2/3/5/7 stages, odd sizes, the generic prime stages
    # (425 = 5^2*17, 1705 = 5*11*31, 6825 = 3*5^2*7*13, 27305 = 5*43*127,
    # 109225 = 5^2*17*257), the two-buffer rank-1 path (8000) and the
    # four-step long-line path (20000 and up)
    rng = np.random.default_rng(n)
    for dt in (np.float32, np.complex64, np.float64):
        x = rand_array(rng, [n], dt)
        k = rand_array(rng, [min(n, 7)], dt)
        check_conv_fft(x, k, "full", "zeros")
        check_conv_fft(x, k, "same", "circular")
    if n >= 20000:
        k = rand_array(rng, [1023], np.float32)
        x = rand_array(rng, [n], np.float32)
        check_conv_fft(x, k, "same", "circular")


@pytest.mark.parametrize("shape,kshape", [((3, 20000), (2, 9)), ((2, 3, 17000), (2, 2, 5)), ((4, 9000), (3, 3))])
def test_conv_fft_long_last_axis(gpu, shape, kshape):
    rng = np.random.default_rng(len(shape))
    for dt in (np.float32, np.complex64):
        x = rand_array(rng, list(shape), dt)
        k = rand_array(rng, list(kshape), dt)
        check_conv_fft(x, k, "same", ("custom", ["reflect"] + ["circular"] * (len(shape) - 1)))
        check_conv_fft(x, k, ("custom", [1] * len(shape), [2] * len(shape)), "replicate", 2)


@pytest.mark.parametrize("shape,kshape,dtype", [((5000, 200), (11, 31), np.float32),
                                                 ((2600, 64), (7, 5), np.float64),
                                                 ((6000, 9), (3, 4), np.complex64),
                                                 ((3, 5200, 20), (2, 9, 3), np.float32),
                                                 ((30000, 3), (17, 2), np.float32)])
def test_conv_fft_long_column_axes(gpu, shape, kshape, dtype):
    # column axes longer than one workgroup's LDS: the four-step split
    # (forward L1-point passes with twiddles, then L2-point passes; mirrored
    # for the inverse)
    rng = np.random.default_rng(sum(shape))
    x = rand_array(rng, list(shape), dtype)
    k = rand_array(rng, list(kshape), dtype)
    check_conv_fft(x, k, "same", ("custom", ["reflect"] + ["circular"] * (len(shape) - 1)))
    check_conv_fft(x, k, ("custom", [1] * len(shape), [2] * len(shape)), "replicate", 2, False)
    check_conv_fft(x, k, "full", "zeros")


@pytest.mark.parametrize("dtype", [np.float32, np.complex64, np.float64, np.complex128])
@pytest.mark.parametrize("rows", [60, 250, 500])
def test_conv_fft_fused_column_forms(gpu, dtype, rows):
    # The fused axis-0 column pass transforms the kernel and data columns of a
    # tile as one interleaved batch while four 8-column tiles fit 96 KB of LDS
    # (axis-0 FFT lengths up to 384 for f32/c64, 192 for f64/c128), and one
    # after the other beyond: 60 rows take the first form for every dtype,
    # 500 the second, 250 the first for f32/c64 and the second for f64/c128.
    rng = np.random.default_rng(rows + np.dtype(dtype).num)
    x = rand_array(rng, [rows, 23], dtype)
    k = rand_array(rng, [7, 4], dtype)
    check_conv_fft(x, k, "same", ("custom", ["reflect", "circular"]), 2)
    check_conv_fft(x, k, "full", "zeros")
    check_conv_fft(x, k, ("custom", [3, 1], [2, 1]), "replicate", 1, False)


def test_strided_views(gpu):
    rng = np.random.default_rng(5)
    base = rng.standard_normal((20, 30)).astype(np.float32)
    kb = rng.standard_normal((7, 8)).astype(np.float32)
    for x in (base[::2, ::3], base[::-1, 5:], base.T, base[3:17, ::-2]):
        for k in (kb[::-1, :], kb[1:6:2, ::3], kb.T):
            ref = O.conv(np.ascontiguousarray(x), np.ascontiguousarray(k), "same", "reflect", 1, True,
                         acc_dtype=np.float64)
            got = nc.conv(x, k, nc.ConvMode.Same, nc.PaddingMode.Reflect)
            assert O.relerr(got, ref)[0] < 1e-5
            got = nc.conv_fft(x, k, nc.ConvMode.Same, nc.PaddingMode.Reflect)
            assert O.relerr(got, ref)[0] < 1e-5


def test_torch_device_tensors(gpu):
    import torch
    rng = np.random.default_rng(9)
    x = rng.standard_normal((33, 47)).astype(np.float32)
    k = rng.standard_normal((5, 6)).astype(np.float32)
    xt, kt = torch.from_numpy(x).cuda(), torch.from_numpy(k).cuda()
    ref = O.conv(x, k, "same", "circular", 2, True, acc_dtype=np.float64)
    y1 = nc.conv(xt, nc.with_dilation(kt, 2), nc.ConvMode.Same, nc.PaddingMode.Circular)
    y2 = nc.conv_fft(xt, nc.with_dilation(kt, 2), nc.ConvMode.Same, nc.PaddingMode.Circular)
    assert y1.is_cuda and y2.is_cuda
    assert O.relerr(y1.cpu().numpy(), ref)[0] < 1e-5
    assert O.relerr(y2.cpu().numpy(), ref)[0] < 1e-5


def test_processor_reuse(gpu):
    rng = np.random.default_rng(11)
    proc = nc.get_fft_processor()
    k = rng.standard_normal((3, 3))
    for shape in [(10, 10), (17, 5), (10, 10), (64, 3)]:
        x = rng.standard_normal(shape)
        ref = O.conv_fft(x, k, "same", "zeros")
        got = nc.conv_fft_with_processor(x, k, nc.ConvMode.Same, nc.PaddingMode.Zeros, proc)
        assert O.relerr(got, ref)[0] < 1e-12


def test_stream_switch_orders_calls(gpu):
    """A device call on torch's stream followed at once by a host-memory call
    on the handle's own stream: both use the handle's spectra, so the second
    must not start before the first has finished with them."""
    import torch
    rng = np.random.default_rng(13)
    proc = nc.get_fft_processor()
    x = rng.standard_normal((1024, 4000)).astype(np.float32)
    k = rng.standard_normal((9, 17)).astype(np.float32)
    xs = rng.standard_normal((1000, 4000)).astype(np.float32)
    ks = rng.standard_normal((5, 7)).astype(np.float32)
    xt, kt = torch.from_numpy(x).cuda(), torch.from_numpy(k).cuda()
    torch.cuda.synchronize()
    for _ in range(3):
        yt = nc.conv_fft_with_processor(xt, kt, nc.ConvMode.Same, nc.PaddingMode.Zeros, proc)
        ys = nc.conv_fft_with_processor(xs, ks, nc.ConvMode.Same, nc.PaddingMode.Zeros, proc)
        torch.cuda.synchronize()
        assert O.relerr(yt.cpu().numpy(), O.conv_fft(x.astype(np.float64), k.astype(np.float64), "same",
                                                     "zeros"))[0] <= 1e-5
        assert O.relerr(ys, O.conv_fft(xs.astype(np.float64), ks.astype(np.float64), "same", "zeros"))[0] <= 1e-5
