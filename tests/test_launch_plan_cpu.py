"""The FFT convolution's launch planner, run on the host without a device
(ndc_conv_fft_launches): over the rank-1 length sweep, long last axes and the
column lengths of the multi-axis passes, every planned launch must fit the
MI355X workgroup LDS (160 KiB, less the reserve the handle keeps for static
LDS).  A launch that does not fit is refused on the host before it reaches
the device; this pins that the planner never produces one for these sizes,
and that no size in the sweep is refused."""
import numpy as np
import pytest

import ndconv as nc

LDS = 160 * 1024 - 1024     # the handle's dynamic budget on MI355X
DTYPES = [np.float32, np.complex64, np.float64, np.complex128]
SIZES_1D = [1, 2, 3, 5, 7, 16, 17, 100, 384, 424, 425, 1537, 1705, 3073, 5030, 6145, 6825, 8000, 9001, 12289,
            13650, 20000, 24577, 27000, 27305, 54610, 100000, 109225, 218450, 262144, 294912]


def check(xs, ks, mode, pm, dil=1, dtype=np.float32):
    launches = nc.fft_launch_plan(xs, ks, mode, pm, dil, True, dtype, LDS)
    assert launches, (xs, ks)
    for smem, grid, block in launches:
        assert 0 <= smem <= LDS, (xs, ks, dtype, smem)
        assert grid >= 1 and 64 <= block <= 1024 and block % 64 == 0, (xs, ks, dtype, grid, block)
    return launches


@pytest.mark.parametrize("dtype", DTYPES)
def test_rank1_sweep_fits_lds(dtype):
    for n in SIZES_1D:
        for k in sorted({1, min(n, 7), min(n, 1023)}):
            check([n], [k], nc.ConvMode.Full, nc.PaddingMode.Zeros, 1, dtype)
            check([n], [k], nc.ConvMode.Same, nc.PaddingMode.Circular, 1, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_long_last_axes_fit_lds(dtype):
    for shape, kshape in [((3, 20000), (2, 9)), ((2, 3, 17000), (2, 2, 5)), ((4, 9000), (3, 3)),
                          ((200, 5000), (11, 31)), ((16, 65536), (3, 33)), ((2, 2, 2, 40000), (1, 2, 1, 3))]:
        check(list(shape), list(kshape), nc.ConvMode.Same, nc.PaddingMode.Circular, 1, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_column_lengths_fit_lds(dtype):
    # columns longer than one workgroup holds take the four-step split
    for rows in [1, 7, 60, 250, 500, 1000, 2000, 2600, 4000, 4600, 5000, 6145, 13650, 40000, 109225, 262144]:
        check([rows, 31], [5, 3], nc.ConvMode.Same, nc.PaddingMode.Reflect if rows > 4 else nc.PaddingMode.Zeros,
              1, dtype)
        check([3, rows, 17], [2, 5, 3], nc.ConvMode.Full, nc.PaddingMode.Zeros, 1, dtype)


def test_headline_shapes():
    B = nc.BorderType
    two_d = check([200, 5000], [11, 31], nc.ConvMode.Same, nc.PaddingMode.Custom([B.Reflect, B.Circular]), 2)
    assert len(two_d) == 3          # rows forward, fused column pass, rows inverse
    assert len(check([5000], [31], nc.ConvMode.Same, nc.PaddingMode.Zeros)) == 1   # one fused rank-1 launch
    check([10, 100, 200], [5, 11, 31], nc.ConvMode.Full, nc.PaddingMode.Zeros)
    check([10, 100, 200], [5, 11, 31], nc.ConvMode.Full, nc.PaddingMode.Zeros, 1, np.complex64)
