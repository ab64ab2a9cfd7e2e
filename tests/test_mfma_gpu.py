"""The matrix-core precision forms of the 2-D f32 direct convolution.

BF16x3 (split hi/lo, three products per tap, FP32 accumulation) must stay
within the 1e-5 normwise relative error of the FP64 oracle that the plain
FP32 path meets; single BF16 products are reported with a looser bound
(about 2^-8 per product).  Shapes that the systolic form does not take
(row stride not dividing the row dilation, or a staged row segment wider
than 256 entries) must be refused, not computed."""
import math

import numpy as np
import pytest

from modes import conv_mode, padding_mode, random_border
from oracle import ndconv_oracle as O

import ndconv as nc

pytestmark = pytest.mark.gpu

TOL = {"bf16x3": 1e-5, "bf16": 1e-2}


def staged_row_fits(sx, dx):
    """The systolic form stages, per workgroup, one input row segment covering
    128 outputs and 16 dilated kernel columns, sampled every gcd(sx, dx); it
    holds at most 256 entries."""
    g = math.gcd(sx, dx)
    return (127 * sx + 15 * dx) // g + 1 <= 256


def run(x, k, mode, pm, dil, rev, algo):
    kw = nc.with_dilation(k, dil)
    kw = kw if rev else kw.no_reverse()
    ref = O.conv(x, k, mode, pm, dil, rev, acc_dtype=np.float64)   # raises first on invalid input
    got = nc.conv(x, kw, conv_mode(mode), padding_mode(pm), algo=algo)
    assert got.shape == ref.shape
    return O.relerr(got, ref)[0]


@pytest.mark.parametrize("algo", ["bf16x3", "bf16"])
def test_mfma_random_cases(gpu, algo):
    rng = np.random.default_rng(77)
    worst = 0.0
    for i in range(40):
        ys, xs = int(rng.integers(1, 70)), int(rng.integers(1, 300))
        kh, kw = int(rng.integers(1, 12)), int(rng.integers(1, 40))
        sy = int(rng.integers(1, 3))
        dy = sy * int(rng.integers(1, 4))          # sy divides dy
        sx, dx = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        mode = ("explicit", [[int(rng.integers(0, kh * dy)), int(rng.integers(0, kh * dy))],
                             [int(rng.integers(0, kw * dx)), int(rng.integers(0, kw * dx))]], [sy, sx])
        pm = ("explicit", [[random_border(rng, 0.5), random_border(rng, -0.25)] for _ in range(2)])
        x = rng.standard_normal((ys, xs)).astype(np.float32)
        k = rng.standard_normal((kh, kw)).astype(np.float32)
        rev = bool(rng.integers(0, 2))
        try:
            ref_ok = True
            O.conv(x, k, mode, pm, [dy, dx], rev, acc_dtype=np.float64)
        except O.OracleError as e:
            ref_ok = False
            with pytest.raises(getattr(nc, e.kind)):
                nc.conv(x, nc.with_dilation(k, [dy, dx]), conv_mode(mode), padding_mode(pm), algo=algo)
        if not ref_ok:
            continue
        if not staged_row_fits(sx, dx):
            with pytest.raises(nc.Unsupported):
                nc.conv(x, nc.with_dilation(k, [dy, dx]), conv_mode(mode), padding_mode(pm), algo=algo)
            continue
        r = run(x, k, mode, pm, [dy, dx], rev, algo)
        assert r <= TOL[algo], (i, r)
        worst = max(worst, r)
    print(f"{algo}: worst normwise rel err {worst:.3e}")


@pytest.mark.parametrize("kshape,dil", [((40, 7), (1, 1)), ((33, 70), (1, 2)), ((5, 130), (2, 1)),
                                        ((9, 9), (4, 3))])
def test_mfma_row_groups_and_column_passes(gpu, kshape, dil):
    # more than 32 kernel rows (row groups), more than 64 columns (column
    # passes) and dilation/stride ratios 2..4 (rotating accumulator sets)
    rng = np.random.default_rng(sum(kshape))
    x = rng.standard_normal((150, 400)).astype(np.float32)
    k = rng.standard_normal(kshape).astype(np.float32)
    for mode in ["same", "valid", "full"]:
        try:
            r = run(x, k, mode, "reflect", list(dil), True, "bf16x3")
        except O.OracleError:
            continue
        assert r <= 1e-5, (mode, r)


def test_mfma_workload4(gpu):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4096, 4096)).astype(np.float32)
    k = rng.standard_normal((31, 31)).astype(np.float32)
    mode = nc.ConvMode.Custom([0, 0], [2, 2])
    ref = O.conv_fft(x.astype(np.float64), k.astype(np.float64), ("custom", [0, 0], [2, 2]), "replicate", 2)
    errs = {}
    for algo in ["direct", "bf16x3", "bf16"]:
        got = nc.conv(x, nc.with_dilation(k, 2), mode, nc.PaddingMode.Replicate, algo=algo)
        errs[algo] = O.relerr(got, ref)[0]
    print("workload4 normwise rel err:", errs)
    assert errs["direct"] <= 1e-5
    assert errs["bf16x3"] <= 1e-5
    assert errs["bf16"] <= 1e-2


@pytest.mark.parametrize("algo", ["bf16x3", "bf16"])
def test_wide_row_cases(gpu, algo):
    """Output rows 512 to 2300 wide (many x tiles per row, ragged last tile)
    with the column stride equal to the column dilation and the row stride
    equal to the row dilation: kernel widths 1..97 (one to four column passes,
    and below 32 taps the exact FP32 kernel that BF16x3 hands over to), every
    border and both kernel orientations."""
    rng = np.random.default_rng(515)
    worst = 0.0
    for i in range(14):
        s = int(rng.integers(1, 3))
        kh, kw = int(rng.integers(1, 14)), [1, 2, 17, 31, 33, 49, 65, 81, 97, 7, 40, 3, 64, 97][i]
        ox = int(rng.integers(512, 2300))
        oy = int(rng.integers(1, 30))
        # valid-mode input of that output, plus random explicit pads
        py = [int(rng.integers(0, kh * s + 1)), int(rng.integers(0, kh * s + 1))]
        px = [int(rng.integers(0, kw * s + 1)), int(rng.integers(0, kw * s + 1))]
        xs = max(1, (ox - 1) * s + (kw - 1) * s + 1 - px[0] - px[1])
        ys = max(1, (oy - 1) * s + (kh - 1) * s + 1 - py[0] - py[1])
        mode = ("explicit", [py, px], [s, s])
        pm = ("explicit", [[random_border(rng, 0.5), random_border(rng, -0.25)] for _ in range(2)])
        x = rng.standard_normal((ys, xs)).astype(np.float32)
        k = rng.standard_normal((kh, kw)).astype(np.float32)
        rev = bool(rng.integers(0, 2))
        try:
            r = run(x, k, mode, pm, [s, s], rev, algo)
        except O.OracleError:
            continue
        assert r <= TOL[algo], (i, x.shape, k.shape, s, r)
        worst = max(worst, r)
    print(f"{algo} wide rows: worst normwise rel err {worst:.3e}")


def test_mfma_refuses_unsupported(gpu):
    x = np.ones((20, 20), np.float32)
    k = np.ones((3, 3), np.float32)
    with pytest.raises(nc.Unsupported):   # row stride 2 does not divide row dilation 1
        nc.conv(x, k, nc.ConvMode.Custom([1, 1], [2, 1]), nc.PaddingMode.Zeros, algo="bf16x3")
    with pytest.raises(nc.Unsupported):   # rank 1
        nc.conv(np.ones(20, np.float32), np.ones(3, np.float32), nc.ConvMode.Same, nc.PaddingMode.Zeros,
                algo="bf16x3")
    with pytest.raises(nc.Unsupported):   # f64
        nc.conv(x.astype(np.float64), k.astype(np.float64), nc.ConvMode.Same, nc.PaddingMode.Zeros,
                algo="bf16x3")


@pytest.mark.parametrize("algo", ["auto", "bf16x3", "bf16"])
@pytest.mark.parametrize("strides,dil,width", [([2, 2], [2, 2], 260), ([1, 1], [2, 1], 260), ([1, 1], [1, 1], 1100),
                                               ([2, 2], [2, 2], 1300)])
def test_nonfinite_inputs_follow_the_offset_list(gpu, algo, strides, dil, width):
    """Infinities and NaNs under zero taps, in dilation gaps and at tile
    edges: the reference sums only the non-zero taps of its offset list
    (dilation/mod.rs:47, conv/mod.rs:188-196), so an output is non-finite
    exactly when a non-zero tap reaches such a sample.  The matrix-core forms
    recompute tiles holding one exactly, so the non-finite outputs must match
    the oracle's float32 sum position by position."""
    rng = np.random.default_rng(21)
    x = rng.standard_normal((300, width)).astype(np.float32)   # 1100 and 1300 wide: many x tiles
    k = rng.standard_normal((7, 9)).astype(np.float32)
    k[2, :] = 0.0
    k[:, 4] = 0.0
    for i, j, v in [(50, 60, np.inf), (120, 7, -np.inf), (200, 200, np.nan), (299, width - 1, np.inf), (0, 0, np.nan),
                    (151, 33, np.inf), (151, 35, -np.inf)]:
        x[i, j] = v
    mode, pm = ("custom", [3, 4], strides), "replicate"
    with np.errstate(invalid="ignore"):
        ref = O.conv(x, k, mode, pm, dil, True)                           # float32, offset-list order
        ref64 = O.conv(x, k, mode, pm, dil, True, acc_dtype=np.float64)
    got = nc.conv(x, nc.with_dilation(k, dil), conv_mode(mode), padding_mode(pm), algo=algo)
    assert got.shape == ref.shape
    assert np.isfinite(ref).sum() < ref.size          # the case reaches some outputs
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(np.isposinf(got), np.isposinf(ref))
    np.testing.assert_array_equal(np.isneginf(got), np.isneginf(ref))
    fin = np.isfinite(ref64)
    tol = 1e-2 if algo == "bf16" else 1e-5
    assert O.relerr(got[fin], ref64[fin])[0] <= tol
