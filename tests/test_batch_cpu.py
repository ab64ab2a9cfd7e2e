"""Host-side pieces of the batched rank-1 path that need no GPU.

The batch kernels resolve borders per sample from a closed form instead of
a tabulated map (a tabulated map per pair would not fit in memory for large
batches).  The closed form is pinned here against the oracle's sequential
padding of x = [1..n] with distinguishable constants, over every side
combination, pads up to several times n (the front-pad-longer-than-n cases
of Circular and Reflect included), and against the tabulated map of the
single-problem planner."""
import itertools

import numpy as np
import pytest

from oracle import ndconv_oracle as O

import ndconv as nc
from test_boundary import BORDERS, _map_values


@pytest.mark.parametrize("fb,bb", list(itertools.product(BORDERS, BORDERS)), ids=lambda b: str(b[0]))
def test_closed_form_map_matches_oracle(fb, bb):
    for n in [1, 2, 3, 5, 8]:
        for p0 in range(0, 14):
            for p1 in range(0, 14):
                mode = nc.PaddingMode.Explicit([[fb[1], bb[1]]])
                cm = nc.ConvMode.Explicit([[p0, p1]], [1])
                fo = ("const", -1000) if fb[0] == "zeros" or isinstance(fb[0], tuple) else fb[0]
                bo = ("const", -2000) if bb[0] == "zeros" or isinstance(bb[0], tuple) else bb[0]
                try:
                    exp = O.padding(np.arange(1, n + 1, dtype=np.int64), ("explicit", [[fo, bo]]), [[p0, p1]])
                except O.OracleError as e:
                    assert e.kind == O.INVALID_ARGUMENT
                    with pytest.raises(nc.InvalidArgument):
                        nc.batch_border_map([n], [1], 0, cm, mode)
                    continue
                got = nc.batch_border_map([n], [1], 0, cm, mode)
                np.testing.assert_array_equal(_map_values(got, fb, bb), exp, err_msg=f"n={n} p=({p0},{p1})")
                np.testing.assert_array_equal(got, nc.border_map([n], [1], cm, mode, 0))


def test_closed_form_map_uneven_batch_shapes():
    # Same + Circular with kernels longer than the signal (front pad > n):
    # the front wraps once, cells before p0 - n stay 0
    rng = np.random.default_rng(11)
    xl = rng.integers(1, 400, 60)
    kl = rng.integers(1, 1024, 60)
    for i in range(60):
        for pm in [nc.PaddingMode.Circular, nc.PaddingMode.Replicate, nc.PaddingMode.Zeros]:
            got = nc.batch_border_map(xl, kl, i, nc.ConvMode.Same, pm)
            np.testing.assert_array_equal(got, nc.border_map([int(xl[i])], [int(kl[i])], nc.ConvMode.Same, pm, 0))
        if 2 * int(xl[i]) > int(kl[i]):   # reflect needs pads shorter than the signal
            got = nc.batch_border_map(xl, kl, i, nc.ConvMode.Same, nc.PaddingMode.Reflect)
            exp = nc.border_map([int(xl[i])], [int(kl[i])], nc.ConvMode.Same, nc.PaddingMode.Reflect, 0)
            np.testing.assert_array_equal(got, exp)


def test_batch_plan_needs_a_device():
    # this container has no GPU: planning a batch fails cleanly
    with pytest.raises(nc.Error):
        nc.BatchPlan([10, 20], [3, 4], nc.ConvMode.Same, nc.PaddingMode.Circular)


def test_dilation_multiplier_divides_exactly():
    """Matrix-core items find tap q / d of a dilated kernel as the high word of
    2 q * ceil(2^31 / d) (batch.hpp, batch_dmul): exhaustively equal to q // d
    for every dilation up to 2^11 and every tap position an item reads
    (q < 2^13), and for sampled dilations up to 2^18 (the kernel-spectrum
    pre-pass)."""
    q = np.arange(1 << 13, dtype=np.uint64)
    for d in range(1, (1 << 11) + 1):
        m = ((1 << 31) + d - 1) // d
        assert m < (1 << 32)
        t = ((2 * q) * np.uint64(m)) >> np.uint64(32)
        np.testing.assert_array_equal(t, q // np.uint64(d), err_msg=f"d={d}")
    # the kernel-spectrum pre-pass uses it for every FFT pair's dilation
    # (below 2^18 whenever the kernel has two taps or more) over positions
    # below 8192: sampled dilations up to 2^18 and the edges
    rng = np.random.default_rng(18)
    extra = set(rng.integers((1 << 11) + 1, (1 << 18) + 1, 3000).tolist())
    extra |= {(1 << 18), (1 << 18) - 1, 65521, 131071, 8191, 8192, 12289}
    for d in sorted(extra):
        m = ((1 << 31) + d - 1) // d
        t = ((2 * q) * np.uint64(m)) >> np.uint64(32)
        np.testing.assert_array_equal(t, q // np.uint64(d), err_msg=f"d={d}")
