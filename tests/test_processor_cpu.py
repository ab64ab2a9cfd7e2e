"""The processor's spectral layout in the oracle, pinned without the GPU:
closed form vs the literal permute-and-transform loop, a brute-force DFT on
tiny inputs, and the reference's roundtrip cases (real.rs:289-478)."""
import itertools

import numpy as np
import pytest

from oracle import ndconv_oracle as O


def brute_dftn(x):
    x = np.asarray(x, dtype=np.complex128)
    X = np.zeros_like(x)
    for k in itertools.product(*[range(n) for n in x.shape]):
        acc = 0j
        for j in itertools.product(*[range(n) for n in x.shape]):
            ph = sum(kk * jj / n for kk, jj, n in zip(k, j, x.shape))
            acc += x[j] * np.exp(-2j * np.pi * ph)
        X[k] = acc
    return X


@pytest.mark.parametrize("shape", [(6,), (5,), (4, 6), (3, 5), (2, 3, 4), (3, 4, 5), (2, 2, 3, 3)])
@pytest.mark.parametrize("cplx", [False, True])
def test_closed_form_matches_the_reference_loop(shape, cplx):
    rng = np.random.default_rng(len(shape) * 10 + int(cplx))
    x = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    a = O.processor_forward(x)
    b = O.processor_forward_procedural(x)
    assert a.shape == b.shape
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-10)
    # rotated left by one; real input keeps F/2+1 bins of its last axis
    exp = tuple(shape[1:-1]) + ((shape[-1] // 2 + 1) if not cplx else shape[-1],) + tuple(shape[:1])
    if len(shape) == 1:
        exp = ((shape[0] // 2 + 1) if not cplx else shape[0],)
    assert a.shape == exp


@pytest.mark.parametrize("shape", [(5,), (3, 4), (2, 3, 3)])
def test_forward_against_brute_force_dft(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    X = brute_dftn(x)
    ref = np.moveaxis(X[..., : shape[-1] // 2 + 1], 0, -1) if len(shape) > 1 else X[: shape[0] // 2 + 1]
    np.testing.assert_allclose(O.processor_forward(x), ref, atol=1e-9)


def test_reference_roundtrip_cases():
    # real.rs roundtrip_1d::basic and different_sizes, roundtrip_2d/3d shapes
    cases = [np.array([1.0, 2, 3, 4, 5, 6]), np.array([1.0, 2]), np.array([1.0, 2, 3]), np.array([1.0, 2, 3, 4, 5]),
             np.fromfunction(lambda i, j: (i + j) % 17, (32, 32)), np.fromfunction(lambda i, j, k: (i + j + k) % 13,
                                                                                   (8, 8, 8))]
    for x in cases:
        spec = O.processor_forward(x)
        back = O.processor_backward(spec, real_len=x.shape[-1])
        np.testing.assert_allclose(back, x, atol=1e-10)
        z = x + 1j * x[::-1]
        np.testing.assert_allclose(O.processor_backward(O.processor_forward(z)), z, atol=1e-10)
