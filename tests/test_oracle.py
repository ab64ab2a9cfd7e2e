"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement (a wrong
pad side, a missed chain through a pad cell, a transposed kernel, a dropped
reversal, a wrong crop start) fails at least one of them:
  * the exact arrays the reference's own tests assert (tests/golden);
  * torch's F.pad / conv{1,2,3}d on the reference's torch-checked cases;
  * mathematics: brute-force circular convolution, FFT == direct conv,
    invariants of good_size and the independently derived band table.
"""
import numpy as np
import pytest

from golden_io import load, oracle_mode
from oracle import ndconv_oracle as O

G = load("reference_goldens.json")
T = load("torch_cases.json")


@pytest.mark.parametrize("case", G["padding"], ids=lambda c: c["src"])
def test_padding_goldens(case):
    x = np.array(case["x"], dtype=case.get("dtype", "int64"))
    out = O.padding(x, oracle_mode(case["mode"]), case["ep"])
    np.testing.assert_array_equal(out, np.array(case["expect"], dtype=x.dtype))


def test_padding_large_array_zeros():
    # padding/mod.rs:723-758
    arr = np.add.outer(np.arange(100), np.arange(100)).astype(np.int32)
    p = O.padding(arr, "zeros", [[5, 5], [5, 5]])
    assert p.shape == (110, 110)
    assert (p[:5] == 0).all() and (p[105:] == 0).all()
    assert (p[5:105, :5] == 0).all() and (p[5:105, 105:] == 0).all()
    assert p[5, 5] == arr[0, 0] and p[54, 54] == arr[49, 49] and p[104, 104] == arr[99, 99]


@pytest.mark.parametrize("case", T["padding"], ids=lambda c: c["src"] + str(c["mode"]))
def test_padding_vs_torch(case):
    x = np.array(case["x"], dtype=np.int64)
    out = O.padding(x, oracle_mode(case["mode"]), case["ep"])
    np.testing.assert_array_equal(out, np.array(case["expect"]))


@pytest.mark.parametrize("case", G["fft_buffer"], ids=lambda c: c["src"])
def test_fft_buffer_goldens(case):
    if case["kind"] == "data":
        x = np.array(case["x"], dtype=np.int64)
        out = O.padding_data(x, oracle_mode(case["mode"]), case["ep"], case["fft_size"])
    else:
        k = np.array(case["k"], dtype=np.int64)
        out = O.padding_kernel(k, case["dilation"], case["reverse"], case["fft_size"])
    np.testing.assert_array_equal(out, np.array(case["expect"]))


def test_kernel_no_reverse_is_flip_of_reverse():
    # padding::kernel with no_reverse places the flipped kernel (padding.rs:103)
    k = np.arange(1, 7).reshape(2, 3)
    a = O.padding_kernel(k, [2, 3], False, [8, 8])
    b = O.padding_kernel(k[::-1, ::-1], [2, 3], True, [8, 8])
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("case", G["offsets"], ids=lambda c: c["src"])
def test_offset_goldens(case):
    offs = O.gen_offset_list(np.array(case["k"]), case["dilation"], case["reverse"], case["strides"])
    if "expect_len" in case:
        assert len(offs) == case["expect_len"]
    else:
        assert [[int(o), float(v)] for o, v in offs] == case["expect"]


@pytest.mark.parametrize("case", G["conv"], ids=lambda c: c["src"])
def test_conv_goldens(case):
    x = np.array(case["x"], dtype=np.int32)
    k = np.array(case["k"], dtype=np.int32)
    out = O.conv(x, k, case["mode"], case["padmode"])
    np.testing.assert_array_equal(out, np.array(case["expect"]))


@pytest.mark.parametrize("case", T["conv"], ids=lambda c: c["src"])
def test_conv_vs_torch(case):
    x = np.array(case["x"], dtype=np.int32)
    k = np.array(case["k"], dtype=np.int32)
    out = O.conv(x, k, oracle_mode(case["mode"]), "zeros", case.get("dilation", 1), case.get("reverse", True))
    np.testing.assert_array_equal(out, np.array(case["expect"]))


@pytest.mark.parametrize("case", G["fft_vs_conv"], ids=lambda c: c["src"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fft_vs_conv(case, prec):
    xi = np.array(case["x"])
    ki = np.array(case["k"])
    dt = case["dtype"]
    mode = oracle_mode(case["mode"])
    pm = oracle_mode(case["padmode"])
    dil = case.get("dilation", 1)
    rev = case.get("reverse", True)
    f = O.conv_fft(xi.astype(dt), ki.astype(dt), mode, pm, dil, rev, precision=prec)
    if case.get("no_round"):
        d = O.conv(xi.astype(dt), ki.astype(dt), mode, pm, dil, rev)
        assert np.max(np.abs(f - d)) < case["tol"]
    else:
        pmi = ("const", int(pm[1])) if isinstance(pm, tuple) else pm
        d = O.conv(xi.astype(np.int32), ki.astype(np.int32), mode, pmi, dil, rev)
        assert f.shape == d.shape
        assert np.all(np.abs(np.round(f) - d) < 1e-5)


@pytest.mark.parametrize("case", G["complex_cases"], ids=lambda c: c["src"])
def test_complex_fft_vs_direct(case):
    x = (np.array(case["x"])[:, 0] + 1j * np.array(case["x"])[:, 1]).reshape(case["xshape"])
    k = (np.array(case["k"])[:, 0] + 1j * np.array(case["k"])[:, 1]).reshape(case["kshape"])
    mode = oracle_mode(case["mode"])
    pm = oracle_mode(case["padmode"])
    dil = case.get("dilation", 1)
    rev = case.get("reverse", True)
    f = O.conv_fft(x, k, mode, pm, dil, rev)
    d = O.conv(x.astype(np.complex128), k.astype(np.complex128), mode, pm, dil, rev)
    assert np.max(np.abs(f - d)) < 1e-9


def test_fft_conv_is_circular_conv_then_crop():
    # The FFT path is a circular convolution of the padded buffers (size F),
    # cropped to [kdil-1, pds): checked by direct summation.
    rng = np.random.default_rng(7)
    for shape, kshape, dil in [((7,), (3,), 2), ((5, 6), (2, 3), 1), ((3, 4, 5), (2, 2, 3), 1)]:
        x = rng.standard_normal(shape)
        k = rng.standard_normal(kshape)
        plan = O.conv_fft_plan(shape, kshape, "full", [dil] * len(shape))
        fs = plan["fft_size"]
        xb = O.padding_data(x, "zeros", plan["padding"], fs)
        kb = O.padding_kernel(k, [dil] * len(shape), True, fs)
        circ = O.circular_conv_bruteforce(xb, kb)
        crop = circ[tuple(slice(plan["kdil"][i] - 1, plan["pds"][i]) for i in range(len(shape)))]
        f = O.conv_fft(x, k, "full", "zeros", dil)
        np.testing.assert_allclose(f, crop, atol=1e-10)


@pytest.mark.parametrize("case", G["doc_cases"], ids=lambda c: c["src"])
def test_doc_examples_fft_matches_direct(case):
    x = np.array(case["x"], dtype=np.int32)
    k = np.array(case["k"], dtype=np.int32)
    dil = case.get("dilation", 1)
    rev = case.get("reverse", True)
    d = O.conv(x, k, case["mode"], case["padmode"], dil, rev)
    f = O.conv_fft(x.astype(np.float64), k.astype(np.float64), case["mode"], case["padmode"], dil, rev)
    np.testing.assert_array_equal(np.round(f).astype(np.int64), d)


def test_int_wrapping_matches_exact():
    rng = np.random.default_rng(3)
    x = rng.integers(-2**31, 2**31 - 1, size=(9, 11), dtype=np.int64).astype(np.int32)
    k = rng.integers(-2**31, 2**31 - 1, size=(3, 4), dtype=np.int64).astype(np.int32)
    a = O.conv(x, k, "same", "circular")
    b = O.conv_int_exact(x, k, "same", "circular")
    np.testing.assert_array_equal(a.astype(np.int64), b)


def test_errors():
    with pytest.raises(O.OracleError) as e:
        O.conv(np.zeros((0, 3), np.int32), np.ones((2, 2), np.int32), "same", "zeros")
    assert e.value.kind == O.DATA_SHAPE
    with pytest.raises(O.OracleError) as e:
        O.conv(np.ones((3, 3), np.int32), np.ones((0, 2), np.int32), "same", "zeros")
    assert e.value.kind == O.KERNEL_SHAPE
    with pytest.raises(O.OracleError) as e:
        O.conv_fft(np.ones((3, 3)), np.ones((0, 2)), "same", "zeros")
    assert e.value.kind == O.DATA_SHAPE
    with pytest.raises(O.OracleError) as e:
        O.conv(np.ones((3, 3), np.int32), np.ones((4, 2), np.int32), "valid", "zeros")
    assert e.value.kind == O.MISMATCH_SHAPE
    with pytest.raises(O.OracleError) as e:
        # symmetric pads never leave the axis; asymmetric ones can (half_dim.rs:200,247)
        O.conv(np.ones(3, np.int32), np.ones(2, np.int32), ("explicit", [[5, 0]], [1]), "reflect")
    assert e.value.kind == O.INVALID_ARGUMENT
    with pytest.raises(O.OracleError) as e:
        O.conv(np.ones(3, np.int32), np.ones(2, np.int32), ("explicit", [[0, 3]], [1]), "reflect")
    assert e.value.kind == O.INVALID_ARGUMENT
    # the largest in-range pads still work, reading unwritten cells as 0
    np.testing.assert_array_equal(O.padding(np.array([1, 2, 3]), "reflect", [[3, 1]]), [0, 3, 2, 1, 2, 3, 2])
    np.testing.assert_array_equal(O.padding(np.array([1, 2, 3]), "reflect", [[0, 2]]), [1, 2, 3, 2, 1])
    with pytest.raises(O.OracleError) as e:
        O.conv(np.ones(3, np.int32), np.ones(2, np.int32), ("custom", [1], [0]), "zeros")
    assert e.value.kind == O.INVALID_ARGUMENT


# --- good_size ---------------------------------------------------------------

# Bands where the truncating divisions give sizes with primes above 7,
# derived independently of the restatement by following good_size.rs:6-31
# by hand for each band edge.
NON_SMOOTH_BANDS = [((385, 425), 425), ((769, 850), 850), ((1537, 1705), 1705), ((3073, 3410), 3410),
                    ((6145, 6825), 6825), ((12289, 13650), 13650), ((24577, 27305), 27305),
                    ((49153, 54610), 54610), ((98305, 109225), 109225), ((196609, 218450), 218450)]


def test_good_size_invariants():
    for n in range(1, 300001):
        g = O.good_size(n)
        p = 1 << (n - 1).bit_length() if n > 1 else 1
        assert n <= g <= p
    for a in range(0, 18):
        assert O.good_size(1 << a) == 1 << a
        if a >= 2:
            assert O.good_size(3 << (a - 2)) == 3 << (a - 2)


@pytest.mark.parametrize("band,size", NON_SMOOTH_BANDS)
def test_good_size_bands(band, size):
    for n in (band[0], (band[0] + band[1]) // 2, band[1]):
        assert O.good_size(n) == size
    assert O.good_size(band[0] - 1) != size


def test_good_size_config_plans():
    # F for the benchmark configs (shapes from the reference README/bench)
    assert O.conv_fft_plan([5000], [31], "same", [1])["fft_size"] == [5120]
    assert O.conv_fft_plan([200, 5000], [11, 31], "same", [2, 2])["fft_size"] == [256, 5120]
    assert O.conv_fft_plan([200, 5000], [11, 31], "same", [1, 1])["fft_size"] == [210, 5120]
    assert O.conv_fft_plan([10, 100, 200], [5, 11, 31], "full", [1, 1, 1])["fft_size"] == [18, 128, 288]
