"""Kept kernel spectra (the reference's sketched conv_fft_bake /
conv_fft_with_baked, conv_fft/mod.rs:22-42,306-392): one bake, several data
arrays of the same shape; every result must equal conv_fft's within the
tolerance and match the FP64 oracle."""
import numpy as np
import pytest

from modes import conv_mode, padding_mode
from oracle import ndconv_oracle as O

import ndconv as nc

pytestmark = pytest.mark.gpu

CASES = [
    # x shape, k shape, dilation, mode, padding, dtype
    ((5000,), (31,), 1, "same", "zeros", np.float32),
    ((100000,), (101,), 1, "same", "circular", np.float32),        # long rank-1 line (four-step: kernel kept as is)
    ((200, 5000), (11, 31), 2, "same", ("custom", ["reflect", "circular"]), np.float32),
    ((10, 100, 200), (5, 11, 31), 1, "full", "zeros", np.float32),
    ((10, 100, 200), (5, 11, 31), 1, "full", "zeros", np.complex64),
    ((64, 90), (7, 9), 3, ("custom", [2, 3], [2, 1]), "replicate", np.float64),
    ((33, 47, 5), (4, 6, 3), 1, "valid", "reflect", np.complex128),
    ((5000, 64), (11, 5), 1, "same", "circular", np.float32),      # long axis 0: kept spectrum in split order
]


def rand(rng, shape, dt):
    x = rng.standard_normal(shape)
    if np.dtype(dt).kind == "c":
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(dt)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[5].__name__}")
def test_bake_matches_conv_fft(gpu, case):
    xs, ks, dil, mode, pm, dt = case
    rng = np.random.default_rng(len(xs) * 7 + ks[0])
    k = rand(rng, ks, dt)
    x0 = rand(rng, xs, dt)
    kw = nc.with_dilation(k, dil)
    baked = nc.conv_fft_bake(x0, kw, conv_mode(mode), padding_mode(pm))
    tol = 1e-5 if np.dtype(dt) in (np.dtype(np.float32), np.dtype(np.complex64)) else 1e-12
    for _ in range(3):
        x = rand(rng, xs, dt)
        got = nc.conv_fft_with_baked(x, baked)
        plain = nc.conv_fft(x, kw, conv_mode(mode), padding_mode(pm))
        ref = O.conv_fft(x.astype(np.complex128 if np.dtype(dt).kind == "c" else np.float64),
                         k.astype(np.complex128 if np.dtype(dt).kind == "c" else np.float64), mode, pm, dil)
        assert got.shape == plain.shape == ref.shape
        assert O.relerr(got, ref)[0] <= tol
        assert O.relerr(got, plain)[0] <= tol
    baked.close()


def test_bake_errors_and_shape_guard(gpu):
    x = np.ones((20, 30), np.float32)
    with pytest.raises(nc.MismatchShape):
        nc.conv_fft_bake(x, np.ones((25, 3), np.float32), nc.ConvMode.Valid, nc.PaddingMode.Zeros)
    with pytest.raises(nc.DataShape):
        nc.conv_fft_bake(x, np.ones((0, 3), np.float32), nc.ConvMode.Same, nc.PaddingMode.Zeros)
    b = nc.conv_fft_bake(x, np.ones((3, 3), np.float32), nc.ConvMode.Same, nc.PaddingMode.Zeros)
    with pytest.raises(nc.InvalidArgument):
        b.conv(np.ones((20, 31), np.float32))


def test_bake_with_torch_tensors(gpu):
    import torch
    rng = np.random.default_rng(9)
    x = rng.random((200, 5000)).astype(np.float32)
    k = rng.random((11, 31)).astype(np.float32)
    xt, kt = torch.from_numpy(x).cuda(), torch.from_numpy(k).cuda()
    b = nc.conv_fft_bake(xt, nc.with_dilation(kt, 2), nc.ConvMode.Same, nc.PaddingMode.Circular)
    got = b.conv(xt).cpu().numpy()
    ref = O.conv_fft(x.astype(np.float64), k.astype(np.float64), "same", "circular", 2)
    assert O.relerr(got, ref)[0] <= 1e-5
